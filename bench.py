"""Benchmark: cell-updates/s of the temporal-tile kernel on a 16384x16384 random grid.

One step = one pass of the hot path over the synthetic input: restore the
seeded 16384^2 grid from a resident device copy (device-to-device, on the
engine's stream) and advance it 10000 generations. Inputs are in HBM before
the timed region starts. The last step's final grid is checked bit-exact
against the oracle fixture (tests/golden/fixtures.json).

Prints ONE JSON line (rank 0). Single GPU only: --gpus above 1 exits with a message.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "stencil-optimalizations_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import golhip  # noqa: E402

METRIC = "cell-updates/s at 16384² random; batched vs plain speedup; p50/p99 job latency"
X = Y = 16384
GENERATIONS = 10000
SEED = 42
BYTES_PER_CELL_UPDATE = 0.25  # 1 bit read + 1 bit written of packed state per cell per generation
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sample_generations: int) -> dict:
    """The oracle's naive loop (restating gol-cpu-naive) on a bounded sample of the
    reference's CPU workload (2048^2 seed-42 random grid), one host core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    cells = ol.random_cells(2048, 2048, 42)
    t0 = time.perf_counter()
    ol.naive_run(cells, sample_generations)
    dt = time.perf_counter() - t0
    return {
        "value": 2048 * 2048 * sample_generations / dt,
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"2048x2048 random seed 42, {sample_generations} of the 1000 generations of the "
                  f"reference CPU workload, naive oracle loop (char cells, dead boundary), "
                  f"{dt:.1f} s on 1 of {os.cpu_count()} host threads",
    }


def expected_hash():
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "fixtures.json")))
    for r in fx["bitpar_runs"]:
        if r["name"] == "random-16384" and r["generations"] == GENERATIONS:
            return r["sha256_input"], r["sha256"]
    return None, None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--temporal-steps", type=int, default=0)
    ap.add_argument("--kernel", default="temporal", choices=["temporal", "plain"])
    ap.add_argument("--cpu-sample-generations", type=int, default=150)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the engine has no "
              f"multi-GPU path in this build", file=sys.stderr)
        return 2

    kernel = golhip.KERNEL_TEMPORAL if args.kernel == "temporal" else golhip.KERNEL_PLAIN
    words = golhip.fill_random(X, Y, SEED)
    want_in, want_out = expected_hash()
    if want_in is not None and hashlib.sha256(words.tobytes()).hexdigest() != want_in:
        raise SystemExit("input generator does not match the fixture")

    src = golhip.Grid(X, Y, word_bits=64, kernel=golhip.KERNEL_PLAIN, device=0)
    src.upload(words)
    src.sync()
    grid = golhip.Grid(X, Y, word_bits=64, kernel=kernel, temporal_steps=args.temporal_steps, device=0)
    info = grid.info()
    log(f"kernel {info.kernel} lane_dwords {info.lane_dwords} band_rows {info.band_rows} waves "
        f"{info.waves} temporal_steps {info.temporal_steps} tiles {info.tiles}")

    def one_step() -> float:
        grid.upload_device(src.device_state())
        return grid.step_timed(GENERATIONS)

    for i in range(args.warmup):
        ms = one_step()
        log(f"warmup {i}: {ms:.1f} ms")
    grid.sync()
    launches0 = grid.info().launches
    gpu_ms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        ms = one_step()
        gpu_ms += ms
        log(f"step {i}: {ms:.1f} ms GPU")
    grid.sync()
    wall = time.perf_counter() - t0
    launches = grid.info().launches - launches0

    out = grid.download()
    got = hashlib.sha256(out.tobytes()).hexdigest()
    exact = (got == want_out) if want_out else None
    if exact is False:
        raise SystemExit(f"final grid differs from the oracle fixture ({got} != {want_out})")

    cells = X * Y
    total_updates = cells * GENERATIONS * args.steps
    value = total_updates / wall
    # roofline of the dominant kernel: algorithmic bytes per launch / average launch duration
    avg_launch_s = (gpu_ms / 1e3) / launches
    bytes_per_launch = total_updates * BYTES_PER_CELL_UPDATE / launches
    achieved_gbs = bytes_per_launch / avg_launch_s / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", "tile_kernel_16384_traffic.json")
    if os.path.exists(prof):
        traffic = json.load(open(prof)).get("hbm_bytes_per_launch")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"{X}x{Y} random (std::mt19937 seed {SEED}, p=0.5), {GENERATIONS} generations "
                        f"per step, dead boundary, rows-64 packing",
            "kernel": args.kernel,
            "lane_dwords": info.lane_dwords, "band_rows": info.band_rows, "waves": info.waves,
            "temporal_steps": info.temporal_steps, "tiles_per_launch": info.tiles,
            "launches_per_step": launches // args.steps,
            "bit_exact_vs_oracle": exact,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "gpu_ms_per_step": gpu_ms / args.steps,
        "avg_launch_us": avg_launch_s * 1e6,
    }
    if not args.no_cpu_baseline:
        log("timing the CPU baseline sample ...")
        result["cpu_baseline"] = cpu_baseline(args.cpu_sample_generations)
    grid.close()
    src.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
