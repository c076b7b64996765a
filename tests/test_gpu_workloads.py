"""GPU parity at the batched and served workloads' real sizes and generation
counts: every grid of workloads.batched_uneven(4096) (sides 64..2048,
100..5000 generations) and every job of workloads.served_stream(400, .)
(sides 128..7936, 100..49500 generations) plus an 8192-side job of 20480
generations, against the oracle's final grids in
tests/golden/workload_fixtures.json (written by make_workload_fixtures.py
from the oracle's own loader and bit-parallel engine)."""

import hashlib
import json
import os
import threading
import time

import numpy as np
import pytest

import golhip
import workloads

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "workload_fixtures.json")))
TUNE_SIDES = (64, 128, 256, 512, 1024, 2048)


def note(*a):
    print("[case]", *a, flush=True)


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def check_specs(specs, grids):
    assert len(specs) == len(grids)
    for s, g in zip(specs, grids):
        assert (s.side, s.generations) == (g[0], g[1]), s


def check_inputs(inputs, grids):
    # the engine's host generator makes the oracle loader's cells for every grid
    bad = [i for i, (w, g) in enumerate(zip(inputs, grids)) if digest(w) != g[2]]
    assert not bad, f"inputs differ from the oracle loader's for grids {bad[:10]}"


@pytest.fixture(scope="module")
def batched():
    specs = workloads.batched_uneven(FIX["batched_uneven"]["n"])
    grids = FIX["batched_uneven"]["grids"]
    check_specs(specs, grids)
    inputs = workloads.make_inputs(specs)
    check_inputs(inputs, grids)
    return specs, inputs, grids


def run_batch(specs, inputs, grids, **kw):
    with golhip.Batch(device=0, **kw) as b:
        for s, w in zip(specs, inputs):
            b.add(w, s.side, s.side, s.generations)
        b.run()
        bad = [i for i in range(len(specs)) if digest(b.fetch(i)) != grids[i][3]]
        info = b.info()
        # a second run restarts from the initial states (and, with plan
        # selection, runs the plan kept): check a spread of grids again
        b.run()
        bad += [i for i in range(0, len(specs), 97) if digest(b.fetch(i)) != grids[i][3]]
    assert not bad, f"{len(bad)} grids differ from the oracle, e.g. " + \
        ", ".join(f"#{i} {specs[i].side}^2 x {specs[i].generations}" for i in bad[:8])
    return info


@pytest.mark.parametrize("persistent", [False, True], ids=["rounds", "persistent"])
def test_batched_workload_against_oracle(batched, persistent):
    specs, inputs, grids = batched
    golhip.batch_tuning_reset()
    note("batched workload", len(specs), "grids, built-in plan, persistent", persistent)
    info = run_batch(specs, inputs, grids, persistent=persistent)
    assert info.grids == len(specs)
    assert info.cell_updates == FIX["batched_uneven"]["cell_updates"]


@pytest.mark.parametrize("persistent", [False, True], ids=["rounds", "persistent"])
def test_batched_workload_autotuned_against_oracle(batched, persistent):
    """After the autotuner: the measured per-class shapes and round depths, and
    plan selection between them and the built-in model, as the bench runs it."""
    specs, inputs, grids = batched
    note("autotune", TUNE_SIDES)
    golhip.batch_autotune([(s, s) for s in TUNE_SIDES], device=0)
    try:
        note("batched workload, autotuned + plan selection, persistent", persistent)
        run_batch(specs, inputs, grids, persistent=persistent, flags=golhip.BATCH_SELECT_PLAN)
    finally:
        golhip.batch_tuning_reset()


def test_served_stream_against_oracle():
    """The whole served stream, paced (Poisson arrivals at 300 jobs/s, about a
    second of arrivals against the server's capacity), with an 8192-side job
    arriving mid-stream; every result against the oracle."""
    fx = FIX["served_stream"]
    specs = workloads.served_stream(fx["n"], 300.0)
    mid = specs[len(specs) // 2].arrival_s
    for k, e in enumerate(fx["extra"]):
        specs.append(workloads.GridSpec(len(specs), e["side"], e["density"], e["generations"], e["seed"],
                                        arrival_s=mid + 0.01 * k))
    grids = fx["grids"]
    check_specs(specs, grids)
    inputs = workloads.make_inputs(specs)
    check_inputs(inputs, grids)
    order = sorted(range(len(specs)), key=lambda i: specs[i].arrival_s)
    ids = [None] * len(specs)
    failure = []
    with golhip.Server(device=0) as srv:
        note("served stream", len(specs), "jobs, paced")
        t0 = time.perf_counter()

        def submitter():
            try:
                for i in order:
                    delay = t0 + specs[i].arrival_s - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                    ids[i] = srv.submit(inputs[i], specs[i].side, specs[i].side, specs[i].generations)
            except Exception as e:  # reported below
                failure.append(e)

        th = threading.Thread(target=submitter)
        th.start()
        bad, odd = [], []
        try:
            for i in order:
                while ids[i] is None and not failure:
                    time.sleep(0.0005)
                if failure:
                    break
                out, st = srv.wait(ids[i])
                if digest(out) != grids[i][3]:
                    bad.append(i)
                if not st.latency_ms >= st.queue_ms >= 0:
                    odd.append(i)
        finally:
            th.join()  # the server outlives every submit
        stats = srv.stats()
    assert not failure, f"submit failed: {failure[0]}"
    assert not odd, f"jobs with latency below queue time: {odd[:8]}"
    assert not bad, f"{len(bad)} jobs differ from the oracle, e.g. " + \
        ", ".join(f"#{i} {specs[i].side}^2 x {specs[i].generations}" for i in bad[:8])
    assert stats.completed == len(specs) and stats.queued_jobs == 0 and stats.resident_jobs == 0
    assert stats.cell_updates == fx["cell_updates"]
