"""GPU tests of the continuous-batching server: every job's result equals the
same job run alone (and the oracle), whatever else is in the mix, under
concurrent submission, tight HBM budgets and arrivals during long jobs."""

import threading
import time

import numpy as np
import pytest

import golhip
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def note(*a):
    print("[case]", *a, flush=True)


def spec_list(n, seed, max_side=1024, max_gens=400):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(64 * max(1, round(np.exp(rng.uniform(np.log(64), np.log(max_side))) / 64)))
        ysz = side if i % 4 else int(rng.integers(1, 300))
        out.append((side, ysz, int(rng.integers(0, max_gens)), float(rng.uniform(0.1, 0.5)), 500 + i))
    return out


def alone(words, x, y, gens):
    with golhip.Grid(x, y) as g:
        g.upload(words)
        g.step(gens)
        return g.download()


def test_concurrent_submitters_match_oracle():
    specs = spec_list(36, 1)
    inputs = [golhip.fill_bernoulli(x, y, sd, p) for x, y, g, p, sd in specs]
    ids = [None] * len(specs)
    with golhip.Server() as srv:
        def submit(idx):
            for i in idx:
                time.sleep(0.002)
                x, y, g, p, sd = specs[i]
                ids[i] = srv.submit(inputs[i], x, y, g)
        ts = [threading.Thread(target=submit, args=(range(k, len(specs), 3),)) for k in range(3)]
        note("server concurrent submit", len(specs))
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        for i, (x, y, g, p, sd) in enumerate(specs):
            out, st = srv.wait(ids[i])
            assert np.array_equal(out, ol.bitpar_run(inputs[i], x, y, g)), (x, y, g)
            assert st.latency_ms >= st.queue_ms >= 0
        s = srv.stats()
        assert s.completed == len(specs) and s.queued_jobs == 0 and s.resident_jobs == 0


def test_budget_admission_queues_and_reuses_memory():
    # 2 MiB budget: at most a few 512^2 jobs (2 x 32 KiB) ... 1024^2 jobs (2 x 128 KiB) resident at once
    specs = [(1024, 1024, 150, 0.4, 10 + i) for i in range(12)] + [(512, 512, 300, 0.3, 40 + i) for i in range(10)]
    inputs = [golhip.fill_bernoulli(x, y, sd, p) for x, y, g, p, sd in specs]
    with golhip.Server(hbm_budget_bytes=600 * 1024) as srv:
        note("budget admission")
        ids = [srv.submit(w, x, y, g) for w, (x, y, g, p, sd) in zip(inputs, specs)]
        for i, (x, y, g, p, sd) in enumerate(specs):
            out, _ = srv.wait(ids[i])
            assert np.array_equal(out, alone(inputs[i], x, y, g))
        s = srv.stats()
        assert s.resident_bytes_peak <= 600 * 1024
        assert s.completed == len(specs)


def test_pinned_region_fills_then_reuses():
    # a 64 MiB budget gives a 2 MiB pinned host region: 32 jobs of 512^2 (2 x
    # 32 KiB of host buffers each) fill it, the rest pin their own buffers;
    # after their results are taken, later jobs take the freed blocks again
    specs = [(512, 512, 40 + i, 0.35, 70 + i) for i in range(48)]
    inputs = [golhip.fill_bernoulli(x, y, sd, p) for x, y, g, p, sd in specs]
    with golhip.Server(hbm_budget_bytes=64 << 20) as srv:
        note("pinned region")
        st = srv.stats()
        assert st.pinned_region_bytes == 2 << 20 and st.pinned_region_used == 0
        for rnd in range(2):
            ids = [srv.submit(w, x, y, g) for w, (x, y, g, p, sd) in zip(inputs, specs)]
            st = srv.stats()
            # the first 32 jobs fill the region, the other 16 pin their own buffers
            assert st.pinned_region_used == st.pinned_region_bytes == st.pinned_region_peak
            assert st.pinned_own_jobs == 16 * (rnd + 1)
            for i, (x, y, g, p, sd) in enumerate(specs):
                out, _ = srv.wait(ids[i])
                assert np.array_equal(out, alone(inputs[i], x, y, g)), (rnd, i)
            # every block went back to the region
            assert srv.stats().pinned_region_used == 0
        assert srv.stats().completed == 2 * len(specs)


def test_oversize_job_fails_alone():
    with golhip.Server(hbm_budget_bytes=256 * 1024) as srv:
        big = golhip.fill_random(2048, 2048, 1)
        small = golhip.fill_random(256, 256, 2)
        note("oversize")
        jb = srv.submit(big, 2048, 2048, 10)
        js = srv.submit(small, 256, 256, 10)
        with pytest.raises(golhip.GolError) as e:
            srv.wait(jb)
        assert e.value.code == golhip.GOL_ERR_OUT_OF_MEMORY
        out, _ = srv.wait(js)
        assert np.array_equal(out, alone(small, 256, 256, 10))


def test_arrivals_join_a_running_job():
    long_w = golhip.fill_random(4096, 4096, 3)
    shorts = [(256, 256, 50, golhip.fill_random(256, 256, 100 + i)) for i in range(8)]
    with golhip.Server() as srv:
        note("arrivals during a long job")
        jl = srv.submit(long_w, 4096, 4096, 6000)
        time.sleep(0.01)
        js = [srv.submit(w, x, y, g) for x, y, g, w in shorts]
        done_order = []
        for j, (x, y, g, w) in zip(js, shorts):
            out, st = srv.wait(j)
            assert np.array_equal(out, alone(w, x, y, g))
            done_order.append(st.latency_ms)
        assert srv.poll(jl) in (golhip.JOB_RESIDENT, golhip.JOB_DRAINING, golhip.JOB_DONE)
        out, stl = srv.wait(jl)
        assert np.array_equal(out, alone(long_w, 4096, 4096, 6000))
        # the short jobs were admitted into the running batch, not after the long job
        assert max(done_order) < stl.latency_ms


def test_zero_generations_and_word_bits_32():
    with golhip.Server() as srv:
        w = golhip.fill_random(128, 64, 5)
        j0 = srv.submit(w, 128, 64, 0)
        c = ol.bernoulli_cells(96, 33, 4, 0.5)
        j1 = srv.submit(ol.pack(c, 32), 96, 33, 21, word_bits=32)
        out0, _ = srv.wait(j0)
        assert np.array_equal(out0, w)
        out1, _ = srv.wait(j1)
        assert np.array_equal(ol.unpack(out1, 96, 33, 32), ol.naive_run(c, 21))


def test_wide_and_tall_split_grids():
    # widths beyond one 2048-cell wavefront row and heights beyond one tile:
    # every size class plans a tile shape and halo depth that fit
    specs = [(7104, 7104, 40), (4160, 600, 70), (8192, 130, 33), (2112, 3000, 45), (6144, 64, 90)]
    with golhip.Server() as srv:
        ws = [golhip.fill_bernoulli(x, y, x + y, 0.3) for x, y, g in specs]
        ids = [srv.submit(w, x, y, g) for w, (x, y, g) in zip(ws, specs)]
        for j, w, (x, y, g) in zip(ids, ws, specs):
            note("split grid", x, y, g)
            out, _ = srv.wait(j)
            assert np.array_equal(out, ol.bitpar_run(w, x, y, g)), (x, y, g)
