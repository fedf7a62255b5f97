"""CPU oracle: a NumPy/SciPy restatement of the reference's per-step grid update.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports this package; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it, and only
as the checker (or the timed CPU baseline), never as the thing shipped.

Each function cites the reference file:line it restates (paths are relative to
``implementations/python/src/`` of Gravifer/FoilBench: ``fb/`` = ``foilbench_py/``,
``kern/`` = ``foilbench_kernels/``). Operation order follows the reference so that FP64
results are reproducible to the last bit where the reference's arithmetic is plain
IEEE (no libm calls); see each module's header for what pins it.
"""
