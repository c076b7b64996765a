"""Fixed workloads to run under rocprofv3 (the profiles under profiles/ come from these).

  python3 profile_run.py slab     16384^2 random, 3 persistent slab launches of 10000 generations
                                  (the bench's configuration: one launch per step)
  python3 profile_run.py tile     16384^2 random, 2400 generations with the temporal-tile kernel at
                                  the shape and depth the engine chooses with the persistent launch
                                  switched off (tile_kernel<1, 32, 8>, 100 launches of 24 generations)
  python3 profile_run.py plain    16384^2 random, 100 generations with the plain kernel
  python3 profile_run.py active   16384^2 mostly-stable patches, 20000 generations, change tracking
                                  (the bench's mostly-stable configuration)
  python3 profile_run.py spacefiller  16384^2 empty with spacefiller[8192, 8192], 20000 generations,
                                  change tracking (the bench's spacefiller configuration)
  python3 profile_run.py spacefillers30  16384^2 with 30 spacefillers (6 x 5 positions, the
                                  reference's 66 %-busy change-tracking workload), 10000 generations,
                                  change tracking with the tile width left to the library
  python3 profile_run.py batch    the bench's batched workload: 4096 uneven grids (sides 64..2048,
                                  100..5000 generations), built once, then two graph replays
  python3 profile_run.py batch_queue  the same batch as persistent queue launches (one per kernel
                                  shape), built once, then two replays
  python3 profile_run.py served   the bench's served stream (400 jobs, sides 128..8192, 100..50000
                                  generations) submitted at once to a server (its capacity run)
  python3 profile_run.py copy     calibration: a torch device copy of 256 MiB, 10 times
                                  (reads and writes exactly 256 MiB per dispatch)
"""

from __future__ import annotations

import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import golhip  # noqa: E402
import workloads  # noqa: E402


def main(what: str) -> None:
    t0 = time.perf_counter()
    if what in ("tile", "plain"):
        w = golhip.fill_random(16384, 16384, 42)
        kernel = golhip.KERNEL_TEMPORAL if what == "tile" else golhip.KERNEL_PLAIN
        gens = 2400 if what == "tile" else 100
        with golhip.Grid(16384, 16384, kernel=kernel, device=0, flags=golhip.FLAG_NO_PERSISTENT) as g:
            g.upload(w)
            ms = g.step_timed(gens)
            info = g.info()
        print(f"{what}: {gens} generations, {info.launches} launches, {ms:.2f} ms GPU, shape "
              f"({info.lane_dwords}, {info.band_rows}, {info.waves}), {info.temporal_steps} generations per launch",
              flush=True)
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
    elif what in ("active", "spacefiller", "active_tiles", "spacefiller_tiles"):
        if what.startswith("active"):
            w = golhip.fill_patches(16384, 16384, 42, 256, 64, 0.35)
        else:
            w = golhip.fill_lexicon(16384, 16384, "spacefiller[8192, 8192]")
        with golhip.Grid(16384, 16384, kernel=golhip.KERNEL_ACTIVE, device=0) as g:
            g.upload(w)
            if what.endswith("_tiles"):
                # the tiles each block advanced: one block per call, read from
                # the handle's running total (a CSV line per block)
                print("block,tiles", flush=True)
                before = 0
                for b in range(20000 // g.info().temporal_steps):
                    g.step(g.info().temporal_steps)
                    total = g.info().active_tiles_total
                    print(f"{b},{total - before}")
                    before = total
            else:
                ms = g.step_timed(20000)
                info = g.info()
                print(f"{what}: 20000 generations, {info.launches} launches, {ms:.2f} ms GPU, "
                      f"{info.active_tiles_total} active tile-blocks of {info.tiles * info.tracked_blocks}",
                      flush=True)
    elif what == "spacefillers30":
        expr = "; ".join(f"spacefiller[{16384 // 7 * i}, {16384 // 6 * j}]" for i in range(1, 7) for j in range(1, 6))
        w = golhip.fill_lexicon(16384, 16384, expr)
        with golhip.Grid(16384, 16384, kernel=golhip.KERNEL_ACTIVE, device=0) as g:
            g.upload(w)
            ms = g.step_timed(10000)
            info = g.info()
        print(f"{what}: 10000 generations, {info.launches} launches, {ms:.2f} ms GPU, "
              f"{info.active_tiles_total} active tile-blocks, {info.tile_width_switches} tile-width switches",
              flush=True)
    elif what in ("batch", "batch_queue"):
        specs = workloads.batched_uneven(4096)
        inputs = workloads.make_inputs(specs)
        with golhip.Batch(device=0, persistent=what == "batch_queue") as b:
            for s, w in zip(specs, inputs):
                b.add(w, s.side, s.side, s.generations)
            b.run()  # plans, uploads and records the graph
            for _ in range(2):
                ms = b.run()
                print(f"batch: replay {ms:.2f} ms GPU", flush=True)
            info = b.info()
        print(f"batch: {info.grids} grids, {info.rounds} rounds, {info.launches} launches per replay", flush=True)
    elif what == "served":
        specs = workloads.served_stream(400, 1.0)
        inputs = workloads.make_inputs(specs)
        total = sum(s.cell_updates for s in specs)
        with golhip.Server(device=0, hbm_budget_bytes=32 << 30) as srv:
            t = time.perf_counter()
            ids = [srv.submit(w, s.side, s.side, s.generations) for s, w in zip(specs, inputs)]
            for i in ids:
                srv.wait(i)
            wall = time.perf_counter() - t
            st = srv.stats()
        print(f"served: {len(specs)} jobs in {wall:.3f} s, {total / wall:.3e} cell-updates/s, "
              f"{st.rounds} rounds, {st.items} items", flush=True)
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
    # the per-block tile modes print CSV on stdout
    print(f"done in {time.perf_counter() - t0:.1f} s", flush=True,
          file=sys.stderr if what.endswith("_tiles") else sys.stdout)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "tile")
