"""PCG32 stream of the reference (fb/core/rng.py:9-55): PCG-XSH-RR 64/32 with the reference's
seeding, float32 conversion and uniform(). Draws are generated vectorised: the states of a
batch follow from the jump-ahead coefficients (O'Neill's PCG paper, section 4.3.1), so a batch
of n words costs O(log n) NumPy passes instead of n Python steps. The device keeps the same
stream for the PIC/FLIP particles (csrc/pic.hip) and this class mirrors it on the host.
"""

from __future__ import annotations

import numpy as np

_MASK = (1 << 64) - 1
_MULT = 6364136223846793005


def _jump(delta: int, inc: int) -> tuple[int, int]:
    """(a, c) with state_{k+delta} = a * state_k + c (mod 2^64)."""
    cur_mult, cur_plus, acc_mult, acc_plus = _MULT, inc, 1, 0
    while delta > 0:
        if delta & 1:
            acc_mult = (acc_mult * cur_mult) & _MASK
            acc_plus = (acc_plus * cur_mult + cur_plus) & _MASK
        cur_plus = ((cur_mult + 1) * cur_plus) & _MASK
        cur_mult = (cur_mult * cur_mult) & _MASK
        delta >>= 1
    return acc_mult, acc_plus


def _output(old: np.ndarray) -> np.ndarray:
    xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
    rot = (old >> np.uint64(59)).astype(np.uint64)
    left = (np.uint64(32) - rot) & np.uint64(31)
    return ((xs >> rot) | (xs << left)) & np.uint64(0xFFFFFFFF)


class PCG32:
    def __init__(self, seed: int, stream: int = 54) -> None:
        # fb/core/rng.py:19-24
        self.state = 0
        self.increment = ((stream << 1) | 1) & _MASK
        self.next_uint32()
        self.state = (self.state + seed) & _MASK
        self.next_uint32()

    def next_uint32(self) -> int:
        old = self.state
        self.state = (old * _MULT + self.increment) & _MASK
        return int(_output(np.asarray([old], dtype=np.uint64))[0])

    def checkpoint(self) -> tuple[int, int]:
        return self.state, self.increment

    def restore(self, words: tuple[int, int]) -> None:
        state, increment = words
        if not (0 <= state <= _MASK and 0 <= increment <= _MASK):
            raise ValueError("PCG32 checkpoint words must be unsigned 64-bit integers")
        if increment & 1 == 0:
            raise ValueError("PCG32 increment must be odd")
        self.state, self.increment = state, increment

    def advance(self, delta: int) -> None:
        a, c = _jump(delta, self.increment)
        self.state = (a * self.state + c) & _MASK

    def uint32(self, count: int) -> np.ndarray:
        """The next `count` words, advancing the stream."""
        states = np.empty(max(count, 0), dtype=np.uint64)
        if count > 0:
            states[0] = self.state
            filled = 1
            with np.errstate(over="ignore"):
                while filled < count:
                    m = min(filled, count - filled)
                    a, c = _jump(filled, self.increment)
                    states[filled:filled + m] = states[:m] * np.uint64(a) + np.uint64(c)
                    filled += m
            self.advance(count)
        return _output(states).astype(np.uint32)

    def random(self, shape) -> np.ndarray:
        # fb/core/rng.py:46-52: float32(u32) * float32(2^-32), row-major draws
        shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        count = int(np.prod(shape, dtype=np.int64))
        return (self.uint32(count).astype(np.float32) * np.float32(1.0 / 4294967296.0)).reshape(shape)

    def uniform(self, low: float, high: float, shape) -> np.ndarray:
        # fb/core/rng.py:54-55
        return np.asarray(low + (high - low) * self.random(shape), dtype=np.float32)
