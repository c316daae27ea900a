"""Seeded synthetic inputs shared by tests and the benchmark.

The reference fixes no random generator, so one is defined here and used for
every target: splitmix64 (Steele, Lea & Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014), element i drawing state seed + (i+1)*gamma,
top 24 bits scaled to a float32 in [0, 1).  oracle/himeno_ref_driver.c
restates the same generator in C.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def uniform01(n: int, seed: int = 0, start: int = 0,
              chunk: int = 1 << 22) -> np.ndarray:
    """float32[n]: elements start .. start+n-1 of the stream for `seed`.
    Chunks are filled by a thread pool (numpy releases the GIL); every
    element depends only on its index, so the result is the same for any
    chunking."""
    out = np.empty(n, dtype=np.float32)
    scale = np.float32(1.0 / 16777216.0)

    def fill(lo: int) -> None:
        hi = min(n, lo + chunk)
        with np.errstate(over="ignore"):
            idx = np.arange(start + lo + 1, start + hi + 1, dtype=np.uint64)
            z = np.uint64(seed) + idx * _GAMMA
            out[lo:hi] = (_mix(z) >> np.uint64(40)).astype(np.float32) * scale

    starts = range(0, n, chunk)
    if n <= chunk:
        for lo in starts:
            fill(lo)
    else:
        with ThreadPoolExecutor(_threads()) as ex:
            list(ex.map(fill, starts))
    return out


def uniform(n: int, lo: float, hi: float, seed: int = 0,
            start: int = 0) -> np.ndarray:
    """float32[n] in [lo, hi), computed as lo + (hi - lo) * u in float32."""
    u = uniform01(n, seed, start)
    return (np.float32(lo) + np.float32(hi - lo) * u).astype(np.float32)
