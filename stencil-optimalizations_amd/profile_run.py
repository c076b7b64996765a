"""Fixed workloads to run under rocprofv3 (the profiles under profiles/ come from these).

  python3 profile_run.py slab     16384^2 random, 3 persistent slab launches of 10000 generations
                                  (the bench's configuration: one launch per step)
  python3 profile_run.py tile     16384^2 random, 800 generations with the temporal-tile kernel
                                  (100 launches of 8 generations, persistent launch switched off)
  python3 profile_run.py plain    16384^2 random, 100 generations with the plain kernel
  python3 profile_run.py active   16384^2 mostly-stable patches, 20000 generations, change tracking
                                  (the bench's mostly-stable configuration)
  python3 profile_run.py batch    256 mixed grids (sides 64..2048, 100..500 generations), one batch run
  python3 profile_run.py copy     calibration: a torch device copy of 256 MiB, 10 times
                                  (reads and writes exactly 256 MiB per dispatch)
"""

from __future__ import annotations

import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import golhip  # noqa: E402


def main(what: str) -> None:
    t0 = time.perf_counter()
    if what in ("tile", "plain"):
        w = golhip.fill_random(16384, 16384, 42)
        kernel = golhip.KERNEL_TEMPORAL if what == "tile" else golhip.KERNEL_PLAIN
        gens = 800 if what == "tile" else 100
        with golhip.Grid(16384, 16384, kernel=kernel, device=0, flags=golhip.FLAG_NO_PERSISTENT) as g:
            g.upload(w)
            ms = g.step_timed(gens)
            info = g.info()
        print(f"{what}: {gens} generations, {info.launches} launches, {ms:.2f} ms GPU", flush=True)
    elif what == "slab":
        w = golhip.fill_random(16384, 16384, 42)
        with golhip.Grid(16384, 16384, kernel=golhip.KERNEL_TEMPORAL, device=0) as g:
            if not g.info().persistent:
                raise SystemExit("slab: the persistent launch is not available on this device")
            for _ in range(3):
                g.upload(w)
                ms = g.step_timed(10000)
                print(f"slab: 10000 generations, {ms:.2f} ms GPU", flush=True)
            info = g.info()
        print(f"slab: {info.launches} launches", flush=True)
    elif what == "active":
        w = golhip.fill_patches(16384, 16384, 42, 256, 64, 0.35)
        with golhip.Grid(16384, 16384, kernel=golhip.KERNEL_ACTIVE, device=0) as g:
            g.upload(w)
            ms = g.step_timed(20000)
            info = g.info()
        print(f"active: 20000 generations, {info.launches} launches, {ms:.2f} ms GPU, "
              f"{info.active_tiles_total} active tile-blocks of {info.tiles * info.tracked_blocks}", flush=True)
    elif what == "batch":
        rng = np.random.default_rng(3)
        with golhip.Batch(device=0) as b:
            for i in range(256):
                side = int(64 * max(1, round(np.exp(rng.uniform(np.log(64), np.log(2048))) / 64)))
                gens = int(rng.integers(100, 501))
                b.add(golhip.fill_bernoulli(side, side, 7000 + i, float(rng.uniform(0.2, 0.5))), side, side, gens)
            ms = b.run()
            info = b.info()
        print(f"batch: {info.grids} grids, {info.rounds} rounds, {info.launches} launches, {ms:.2f} ms GPU",
              flush=True)
    elif what == "copy":
        import torch

        n = 256 * 1024 * 1024 // 4
        a = torch.ones(n, dtype=torch.int32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copy: 10 x 256 MiB device copies", flush=True)
    else:
        raise SystemExit(f"unknown workload {what}")
    print(f"done in {time.perf_counter() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "tile")
