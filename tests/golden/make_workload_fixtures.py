"""Writes tests/golden/workload_fixtures.json: the oracle's final grid of every
grid of the batched and served workloads, at their real sizes and generation
counts.

  batched_uneven  all 4096 grids of workloads.batched_uneven(4096): sides 64..2048,
                  100..5000 generations
  served_stream   all 400 jobs of workloads.served_stream(400, .): sides 128..7936,
                  100..49500 generations (the job parameters do not depend on the
                  arrival rate), plus SERVED_EXTRA: an 8192-side job, the largest side
                  the stream allows, which its 400 draws do not reach

Every input is made by the oracle's Bernoulli loader (oracle_load_bernoulli,
cells u < p with u = (mt19937_64() >> 11) * 2^-53) and packed by the oracle's
row packer; every result comes from the oracle's bit-parallel engine
(oracle_bitpar_run), which the CPU tests pin against the naive loop. Per grid
the file keeps the first 16 hex digits of the SHA-256 of the packed input and
of the packed final grid (rows-64 words), and the final popcount; the input
digest pins each grid's density and cell seed.

Run from the repository root (oracle/liboracle.so must be built):
    python tests/golden/make_workload_fixtures.py            # both workloads
    python tests/golden/make_workload_fixtures.py batched    # one of them
About 7e12 + 3e13 cell-updates: a few minutes on 8 cores.
"""

from __future__ import annotations

import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "stencil-optimalizations_amd"))

import oracle_lib as ol  # noqa: E402
import workloads  # noqa: E402

PATH = os.path.join(HERE, "workload_fixtures.json")
BATCHED_N, SERVED_N = 4096, 400
# Jobs the served test adds to the stream: an 8192-side grid past 20000 generations.
SERVED_EXTRA = [{"side": 8192, "density": 0.3, "generations": 20480, "seed": 8192}]
DIGITS = 16


def digest(a: np.ndarray) -> str:
    return ol.sha256(a)[:DIGITS]


def one_grid(side: int, density: float, generations: int, seed: int, threads: int) -> list:
    words = ol.pack(ol.bernoulli_cells(side, side, seed, density), 64)
    out = ol.bitpar_run(words, side, side, generations, threads)
    return [digest(words), digest(out), int(np.unpackbits(out.view(np.uint8)).sum())]


def run_all(jobs: list) -> list:
    """jobs: (side, density, generations, seed). Small grids run one per host
    thread; grids of 2048 cells and more use every thread each."""
    cpus = os.cpu_count() or 1
    results = [None] * len(jobs)
    small = [i for i, j in enumerate(jobs) if j[0] < 2048]
    large = [i for i, j in enumerate(jobs) if j[0] >= 2048]
    t = time.time()
    with ThreadPoolExecutor(max_workers=cpus) as ex:
        for i, r in zip(small, ex.map(lambda i: one_grid(*jobs[i], 1), small)):
            results[i] = r
    print(f"  {len(small)} grids below 2048 cells: {time.time() - t:.1f} s", flush=True)
    for n, i in enumerate(sorted(large, key=lambda i: -jobs[i][0] ** 2 * jobs[i][2])):
        t1 = time.time()
        results[i] = one_grid(*jobs[i], cpus)
        if n < 5 or n % 20 == 0:
            print(f"  grid {i}: {jobs[i][0]}^2 x {jobs[i][2]}: {time.time() - t1:.1f} s", flush=True)
    print(f"  all {len(jobs)}: {time.time() - t:.1f} s", flush=True)
    return results


def entries(jobs: list, results: list) -> list:
    return [[s, g, *r] for (s, _, g, _), r in zip(jobs, results)]


def batched() -> dict:
    specs = workloads.batched_uneven(BATCHED_N)
    jobs = [(s.side, s.density, s.generations, s.seed) for s in specs]
    print(f"batched_uneven({BATCHED_N}): {sum(s.cell_updates for s in specs):.3e} cell-updates", flush=True)
    return {"n": BATCHED_N, "cell_updates": sum(s.cell_updates for s in specs), "grids": entries(jobs, run_all(jobs))}


def served() -> dict:
    specs = workloads.served_stream(SERVED_N, 1.0)
    jobs = [(s.side, s.density, s.generations, s.seed) for s in specs]
    jobs += [(e["side"], e["density"], e["generations"], e["seed"]) for e in SERVED_EXTRA]
    print(f"served_stream({SERVED_N}) + {len(SERVED_EXTRA)} extra: "
          f"{sum(j[0] ** 2 * j[2] for j in jobs):.3e} cell-updates", flush=True)
    return {"n": SERVED_N, "extra": SERVED_EXTRA, "cell_updates": sum(j[0] ** 2 * j[2] for j in jobs),
            "grids": entries(jobs, run_all(jobs))}


def main() -> None:
    which = set(sys.argv[1:]) or {"batched", "served"}
    out = json.load(open(PATH)) if os.path.exists(PATH) else {}
    out["format"] = ("per grid: [side, generations, sha256[:16] of the input rows-64 words, "
                     "sha256[:16] of the final rows-64 words, final popcount]")
    if "batched" in which:
        out["batched_uneven"] = batched()
    if "served" in which:
        out["served_stream"] = served()
    with open(PATH, "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
