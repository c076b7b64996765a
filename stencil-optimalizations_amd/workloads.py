"""Seeded synthetic workloads (no reference counterpart for the batched and served ones).

  single_large      16384 x 16384, RandomOnesZerosDataLoader seed 42, 10000 generations
  mostly_stable     16384 x 16384, empty with 256 patches of 64 x 64 at density 0.35, 20000 generations
  batched_uneven    4096 square grids, side log-uniform in [64, 2048] rounded to a multiple of 64,
                    density uniform in [0.2, 0.5], generations uniform in [100, 5000]
  served_stream     Poisson arrivals; side log-uniform in [128, 8192] (multiples of 64), density
                    uniform in [0.1, 0.5], generations log-uniform in [100, 50000]

Grid-level parameters come from numpy's PCG64 seeded with the workload's master
seed; each grid's cell seed is SplitMix64(master ^ index), and its cells are
golhip.fill_bernoulli(side, side, seed, density).
"""

from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import List

import numpy as np

import golhip

MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


@dataclass
class GridSpec:
    index: int
    side: int
    density: float
    generations: int
    seed: int
    arrival_s: float = 0.0

    @property
    def cells(self) -> int:
        return self.side * self.side

    @property
    def cell_updates(self) -> int:
        return self.cells * self.generations


def _round64(v: float, lo: int, hi: int) -> int:
    return int(min(hi, max(lo, 64 * round(v / 64))))


def batched_uneven(n: int = 4096, master: int = 20250101) -> List[GridSpec]:
    rng = np.random.Generator(np.random.PCG64(master))
    out = []
    for i in range(n):
        side = _round64(math.exp(rng.uniform(math.log(64), math.log(2048))), 64, 2048)
        density = float(rng.uniform(0.2, 0.5))
        gens = int(rng.integers(100, 5001))
        out.append(GridSpec(i, side, density, gens, splitmix64(master ^ i)))
    return out


def served_stream(n: int, rate_per_s: float, master: int = 20250202) -> List[GridSpec]:
    rng = np.random.Generator(np.random.PCG64(master))
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate_per_s))
        side = _round64(math.exp(rng.uniform(math.log(128), math.log(8192))), 128, 8192)
        density = float(rng.uniform(0.1, 0.5))
        gens = int(round(math.exp(rng.uniform(math.log(100), math.log(50000)))))
        out.append(GridSpec(i, side, density, gens, splitmix64(master ^ i), arrival_s=t))
    return out


def make_inputs(specs: List[GridSpec], threads: int = 16) -> List[np.ndarray]:
    """Cells of every grid (rows-64 words), generated on `threads` host threads
    (the generator releases the GIL)."""
    def one(s: GridSpec) -> np.ndarray:
        return golhip.fill_bernoulli(s.side, s.side, s.seed, s.density)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, specs))
