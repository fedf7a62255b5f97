"""foilbench_amd: FoilBench's per-step grid update (Stable Fluids MAC, D2Q9 TRT LBM, PIC/FLIP)
as hand-written gfx950 HIP kernels behind the reference's solver lifecycle."""

__version__ = "0.1.0"
