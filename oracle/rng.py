"""PCG32 (PCG-XSH-RR 64/32) restated from fb/core/rng.py:9-55 (test infrastructure only).

Pinned by spec/conformance/pcg32.json (three seed/stream cases, uint32 outputs and float32
bit patterns), copied under tests/golden/pcg32.json.
"""

from __future__ import annotations

import numpy as np

_MASK = (1 << 64) - 1
_MULT = 6364136223846793005


class PCG32:
    def __init__(self, seed: int, stream: int = 54) -> None:
        # fb/core/rng.py:19-24
        self.state = 0
        self.inc = ((stream << 1) | 1) & _MASK
        self.next_uint32()
        self.state = (self.state + seed) & _MASK
        self.next_uint32()

    def next_uint32(self) -> int:
        # fb/core/rng.py:26-31
        old = self.state
        self.state = (old * _MULT + self.inc) & _MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = (old >> 59) & 31
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def random(self, shape) -> np.ndarray:
        # fb/core/rng.py:46-52: float32(u32) * float32(2^-32), draws in row-major order
        count = int(np.prod(shape, dtype=np.int64))
        words = np.fromiter((self.next_uint32() for _ in range(count)), dtype=np.uint32, count=count)
        out = words.astype(np.float32) * np.float32(1.0 / 4294967296.0)
        return out.reshape(tuple(shape))

    def uniform(self, low: float, high: float, shape) -> np.ndarray:
        # fb/core/rng.py:54-55
        return np.asarray(low + (high - low) * self.random(shape), dtype=np.float32)
