"""Synthetic workloads of the table-commitment path, shared by bench.py and the
GPU tests so that a timed table is one a test or the oracle can reproduce.

Seed rules (one SplitMix64 stream per purpose; a stream seeded s has state s,
and each draw is state += 0x9E3779B97F4A7C15 followed by the SplitMix64
finaliser):
  * field tables: row-major draws from the table's seed, draws >= p rejected,
    returned column-major (``rowmajor_table``).  Main table seed 1, aux 2,
    streaming commit 3.
  * device-drawn tables (tvh_fill_uniform_dev): element i of the column-major
    buffer is draw i + 1 of the same stream, a draw >= p lowered by p
    (``device_fill``).  Batched table t uses seed 1000 + t, served job j seed
    100000 + j.
  * batched shapes: stream 7, two draws per table: log2 height 12 + (d1 mod 9),
    width [264, 156, 379][d2 mod 3].
  * served jobs: log2 height 14 + (d mod 8) from stream 9; Poisson arrivals
    with gaps -ln(1 - u) / rate, u = d / 2^64, from stream 11; query leaves
    d mod target_len from stream 13.
"""
from __future__ import annotations

import math

import numpy as np

P = 0xFFFFFFFF00000001
GAMMA = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


def splitmix_stream(seed: int):
    """Python generator of the SplitMix64 stream with state `seed`."""
    s = seed & MASK
    while True:
        s = (s + GAMMA) & MASK
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        yield z ^ (z >> 31)


def _splitmix_block(seed: int, first: int, count: int) -> np.ndarray:
    """Draws first+1 .. first+count of the stream (numpy, vectorised)."""
    with np.errstate(over="ignore"):
        i = np.arange(first + 1, first + count + 1, dtype=np.uint64)
        z = np.uint64(seed & MASK) + i * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def splitmix_field(seed: int, count: int) -> np.ndarray:
    """The first `count` draws below p (draws >= p rejected)."""
    out = []
    have, first = 0, 0
    while have < count:
        need = count - have
        blk = _splitmix_block(seed, first, need + 16)
        first += blk.size
        blk = blk[blk < np.uint64(P)]
        out.append(blk[:need])
        have += min(need, blk.size)
    return np.concatenate(out) if out else np.zeros(0, np.uint64)


def rowmajor_table(seed: int, rows: int, cols: int) -> np.ndarray:
    """rows x cols uniform field table drawn row-major; returned column-major
    (element (r, c) at c * rows + r)."""
    flat = splitmix_field(seed, rows * cols)
    return np.ascontiguousarray(flat.reshape(rows, cols).T).reshape(-1)


def device_fill(seed: int, count: int, first: int = 0) -> np.ndarray:
    """numpy statement of tvh_fill_uniform_dev: elements first .. first+count of
    the buffer it fills from `seed`."""
    z = _splitmix_block(seed, first, count)
    return np.where(z >= np.uint64(P), z - np.uint64(P), z)


def batch_shapes(count: int = 96):
    """[(log2 height, width)] of the batched-tables workload."""
    g = splitmix_stream(7)
    out = []
    for _ in range(count):
        d1, d2 = next(g), next(g)
        out.append((12 + d1 % 9, [264, 156, 379][d2 % 3]))
    return out


def batch_seed(t: int) -> int:
    return 1000 + t


def serve_jobs(count: int = 2000, rate: float = 40.0, queries: int = 80, width: int = 379,
               expansion: int = 4, min_log: int = 14, max_log: int = 21):
    """The served job stream: dicts with i, n, m, w, seed, t (arrival, s), leaves."""
    g_h, g_t, g_q = splitmix_stream(9), splitmix_stream(11), splitmix_stream(13)
    span = max_log - min_log + 1
    jobs, t = [], 0.0
    for j in range(count):
        lg = min_log + next(g_h) % span
        n = 1 << lg
        m = expansion * n
        if j and rate > 0:
            u = next(g_t) / float(1 << 64)
            t += -math.log1p(-u) / rate
        leaves = np.array([next(g_q) % m for _ in range(queries)], dtype=np.uint64)
        jobs.append(dict(i=j, n=n, m=m, w=width, seed=100000 + j, t=t, leaves=leaves))
    return jobs
