"""Argument checks of the GPU transfer adapter (foilbench_amd/solvers/transfers.py) against the reference's
(fb/solvers/_numba_adapter.py:28-160, tests/unit/test_pic_transfers.py:101-140). Every check runs before the
library is touched, so these run without a GPU."""

from __future__ import annotations

import numpy as np
import pytest

from foilbench_amd.solvers import transfers as T


def test_rejects_inconsistent_dtypes_and_dimensions():
    # tests/unit/test_pic_transfers.py:101-124
    grid = np.zeros((4, 5, 2), dtype=np.float32)
    positions = np.zeros((3, 2), dtype=np.float64)
    with pytest.raises(TypeError, match="same dtype"):
        T.grid_to_particle(grid, positions, 0.0, 0.0, 0.2, 0.25)
    positions_3d = np.zeros((3, 3), dtype=np.float32)
    with pytest.raises((TypeError, ValueError)):
        T.grid_to_particle(grid, positions_3d, 0.0, 0.0, 0.2, 0.25)
    with pytest.raises((TypeError, ValueError)):
        T.particle_to_grid(positions_3d, np.zeros((4, 2), np.float32), 0.0, 0.0, 0.2, 0.25, 5, 4, (1.0, 0.0))


@pytest.mark.parametrize("dtype", [np.float16, np.int32])
def test_rejects_unsupported_array_dtypes(dtype):
    # tests/unit/test_pic_transfers.py:127-133
    with pytest.raises((TypeError, ValueError)):
        T.grid_to_particle(np.zeros((4, 5, 2), dtype), np.zeros((3, 2), dtype), 0.0, 0.0, 0.2, 0.25)
    with pytest.raises(TypeError):
        T.particle_to_grid(np.zeros((3, 2), dtype), np.zeros((3, 2), dtype), 0.0, 0.0, 0.2, 0.25, 5, 4, (1.0, 0.0))


def test_rejects_empty_grid_axes_and_bad_spacing():
    # tests/unit/test_pic_transfers.py:136-140 and fb/solvers/_numba_adapter.py:53-54, 92-95
    pos = np.zeros((3, 2), np.float32)
    with pytest.raises(ValueError, match="nonempty"):
        T.grid_to_particle(np.zeros((0, 5, 2), np.float32), pos, 0.0, 0.0, 0.2, 0.25)
    with pytest.raises(ValueError, match="positive"):
        T.grid_to_particle(np.zeros((4, 5, 2), np.float32), pos, 0.0, 0.0, 0.0, 0.25)
    for nx, ny, dx in ((0, 4, 0.2), (5, 0, 0.2), (5, 4, -0.2)):
        with pytest.raises(ValueError, match="positive"):
            T.particle_to_grid(pos, pos.copy(), 0.0, 0.0, dx, 0.25, nx, ny, (1.0, 0.0))
    with pytest.raises(ValueError, match="must match"):
        T.particle_to_grid(pos, np.zeros((4, 2), np.float32), 0.0, 0.0, 0.2, 0.25, 5, 4, (1.0, 0.0))


def test_face_pair_rejects_a_broken_mac_grid_and_mixed_dtypes():
    # fb/solvers/_numba_adapter.py:120-123, 149-150
    pos = np.zeros((3, 2), np.float64)
    with pytest.raises(ValueError, match="MAC grid"):
        T.particle_to_faces(pos, pos.copy(), np.zeros((4, 6)), np.zeros((4, 5)), 0.0, 0.0, 0.1, 0.1)
    with pytest.raises(ValueError, match="MAC grid"):
        T.faces_to_particle(np.zeros((4, 6)), np.zeros((4, 5)), pos, 0.0, 0.0, 0.1, 0.1)
    with pytest.raises(ValueError, match="shape"):
        T.particle_to_faces(pos, np.zeros((2, 2)), np.zeros((4, 6)), np.zeros((5, 5)), 0.0, 0.0, 0.1, 0.1)
    with pytest.raises(TypeError, match="same dtype"):
        T.faces_to_particle(np.zeros((4, 6), np.float32), np.zeros((5, 5), np.float32), pos, 0.0, 0.0, 0.1, 0.1)
