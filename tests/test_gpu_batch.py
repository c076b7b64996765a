"""GPU parity of the batched engine: every grid of a mixed batch is bit-exact
against the oracle and against the same grid run alone."""

import numpy as np
import pytest

import golhip
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def note(*a):
    print("[case]", *a, flush=True)


def mixed_specs(n, seed, max_side=1024, max_gens=300):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        side_x = int(64 * max(1, round(np.exp(rng.uniform(np.log(64), np.log(max_side))) / 64)))
        side_y = int(64 * max(1, round(np.exp(rng.uniform(np.log(64), np.log(max_side))) / 64)))
        if i % 5 == 0:
            side_y = int(rng.integers(1, 200))  # ragged heights
        gens = int(rng.integers(0, max_gens))
        p = float(rng.uniform(0.2, 0.5))
        specs.append((side_x, side_y, gens, p, 1000 + i))
    return specs


@pytest.mark.parametrize("use_graph", [True, False])
def test_mixed_batch_against_oracle(use_graph):
    specs = mixed_specs(40, 1)
    with golhip.Batch(use_graph=use_graph) as b:
        ids = []
        for x, y, gens, p, seed in specs:
            ids.append(b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, gens))
        note("batch of", len(specs), "graph", use_graph)
        b.run()
        info = b.info()
        assert info.grids == len(specs)
        for gid, (x, y, gens, p, seed) in zip(ids, specs):
            ref = ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
            assert np.array_equal(b.fetch(gid), ref), (x, y, gens)
        # a second run restarts from the initial states
        b.run()
        x, y, gens, p, seed = specs[3]
        assert np.array_equal(b.fetch(ids[3]), ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, gens))


def test_large_split_grids_in_batch():
    specs = [(2048, 2048, 200, 0.5, 1), (1088, 1600, 97, 0.3, 2), (64, 2048, 150, 0.4, 3), (2048, 64, 64, 0.5, 4),
             (1984, 1216, 33, 0.45, 5), (7104, 900, 41, 0.35, 6), (4160, 3000, 20, 0.4, 7)]
    with golhip.Batch(round_generations=16) as b:
        for x, y, gens, p, seed in specs:
            b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
        note("split grids")
        b.run()
        for i, (x, y, gens, p, seed) in enumerate(specs):
            ref = ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
            assert np.array_equal(b.fetch(i), ref), (x, y, gens)


def test_same_alone_and_in_the_mix():
    specs = mixed_specs(12, 7, max_side=2048, max_gens=400)
    with golhip.Batch() as b:
        for x, y, gens, p, seed in specs:
            b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
        b.run()
        for i, (x, y, gens, p, seed) in enumerate(specs):
            note("alone vs mix", i, x, y, gens)
            with golhip.Grid(x, y) as g:
                g.upload(golhip.fill_bernoulli(x, y, seed, p))
                g.step(gens)
                assert np.array_equal(g.download(), b.fetch(i))


@pytest.mark.parametrize("round_generations", [32, 5])
def test_narrow_grids_in_batch(round_generations):
    """Grids at most 512 cells wide and 512 rows tall run on the narrow kernel
    (one grid row per lane, 2/4/8/16 dwords per lane, 1..8 wavefronts): every
    width class with ragged widths, heights across wavefront edges, and long
    runs that advance hundreds of generations per item."""
    specs = []
    for i, (x, y) in enumerate([(32, 1), (64, 2), (96, 63), (128, 64), (160, 65), (256, 127), (288, 128),
                                (384, 200), (480, 255), (512, 256), (32, 512), (512, 512), (448, 449),
                                (224, 320), (352, 383), (64, 448)]):
        gens = [0, 1, 37, 500, 1500][i % 5]
        specs.append((x, y, gens, ol.bernoulli_cells(x, y, 300 + i, 0.25 + 0.05 * (i % 5))))
    with golhip.Batch(round_generations=round_generations) as b:
        for x, y, gens, c in specs:
            b.add(ol.pack(c, 32), x, y, gens, word_bits=32)
        note("narrow grids, round_generations", round_generations)
        b.run()
        for i, (x, y, gens, c) in enumerate(specs):
            assert np.array_equal(ol.unpack(b.fetch(i), x, y, 32), ol.naive_run(c, gens)), (x, y, gens)


def test_narrow_oscillators_at_every_edge():
    # blinkers across every wavefront edge and against the grid's edges: the
    # rows exchanged between wavefronts and the dead boundary
    x, y = 512, 512
    c = np.zeros((y, x), dtype=np.uint8)
    for r in list(range(0, 512, 64)) + list(range(63, 512, 64)) + [1, 510]:
        for col in (0, 31, 32, 255, 256, 479, 510):
            c[max(r - 1, 0):min(r + 2, y), col] = 1
    with golhip.Batch() as b:
        gids = [b.add(ol.pack(c, 32), x, y, g, word_bits=32) for g in (1, 2, 101)]
        b.run()
        for gid, g in zip(gids, (1, 2, 101)):
            assert np.array_equal(ol.unpack(b.fetch(gid), x, y, 32), ol.naive_run(c, g)), g


def test_word_bits_32_in_batch():
    with golhip.Batch() as b:
        c = ol.bernoulli_cells(96, 50, 3, 0.5)
        gid = b.add(ol.pack(c, 32), 96, 50, 77, word_bits=32)
        b.run()
        assert np.array_equal(ol.unpack(b.fetch(gid), 96, 50, 32), ol.naive_run(c, 77))


def test_autotune_installs_measured_choices():
    """The autotuner times every compiled shape per size class and installs the
    fastest; batches planned afterwards use it and stay bit-exact."""
    sizes = [(2048, 2048), (640, 512), (128, 128), (4096, 96)]
    try:
        note("autotune", sizes)
        res = golhip.batch_autotune(sizes, round_generations=16, generations=64, cells_per_measure=1 << 25)
        assert len(res) == len(sizes)
        for r in res:
            assert r["candidates"] >= 1
            assert r["tuned_cell_updates_per_s"] >= r["builtin_cell_updates_per_s"] > 0
            narrow = r["band_rows"] == 0
            if narrow:  # the narrow kernel: one grid row per lane
                assert r["x_cells"] <= 512 and r["y_cells"] <= 512
                assert r["x_cells"] <= 32 * r["lane_dwords"] and r["y_cells"] <= 64 * r["waves"]
            else:
                assert (r["lane_dwords"], r["band_rows"], r["waves"]) in {tuple(s) for s in golhip.TILE_SHAPES}
            one_tile = narrow or (r["x_cells"] <= 2048 * r["lane_dwords"]
                                  and r["y_cells"] <= r["band_rows"] * r["waves"])
            assert (r["generations_per_round"] == 0) == one_tile
            if not one_tile:
                assert 1 <= r["generations_per_round"] <= 16
        specs = [(2048, 2048, 70, 0.4, 11), (640, 512, 150, 0.3, 12), (128, 128, 400, 0.5, 13),
                 (4096, 96, 33, 0.45, 14), (1984, 1600, 50, 0.35, 15)]
        with golhip.Batch(round_generations=16) as b:
            for x, y, gens, p, seed in specs:
                b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
            note("batch after tuning")
            b.run()
            for i, (x, y, gens, p, seed) in enumerate(specs):
                ref = ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
                assert np.array_equal(b.fetch(i), ref), (x, y, gens)
    finally:
        golhip.batch_tuning_reset()


# ---- persistent queue launch (GOL_BATCH_PERSISTENT) ----

@pytest.mark.parametrize("use_graph", [True, False])
def test_persistent_batch_against_oracle(use_graph):
    """One persistent launch per kernel shape drains a device work queue; every
    grid, including split grids whose rounds wait for each other on the device,
    is bit-exact against the oracle, and a re-run restarts from the initial state."""
    specs = mixed_specs(40, 2, max_side=1536, max_gens=400)
    with golhip.Batch(use_graph=use_graph, persistent=True, round_generations=12) as b:
        for x, y, gens, p, seed in specs:
            b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
        note("persistent batch of", len(specs), "graph", use_graph)
        b.run()
        info = b.info()
        assert info.grids == len(specs)
        for i, (x, y, gens, p, seed) in enumerate(specs):
            ref = ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
            assert np.array_equal(b.fetch(i), ref), (x, y, gens)
        note("persistent batch re-run")
        b.run()
        for i in (0, 5, 17):
            x, y, gens, p, seed = specs[i]
            assert np.array_equal(b.fetch(i), ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, gens))


def test_persistent_split_and_narrow_grids():
    """Split grids with many dependent rounds next to narrow grids advanced in
    one item each, and 0-generation grids: persistent equals per-round launches
    equals the oracle."""
    specs = [(2048, 2048, 150, 0.5, 21), (1088, 1600, 97, 0.3, 22), (64, 2048, 150, 0.4, 23),
             (2048, 64, 64, 0.5, 24), (1984, 1216, 33, 0.45, 25), (4160, 1000, 20, 0.4, 26),
             (512, 512, 700, 0.35, 27), (192, 63, 1000, 0.3, 28), (448, 255, 0, 0.3, 29), (256, 1, 5, 0.5, 30)]
    results = {}
    for persistent in (False, True):
        with golhip.Batch(round_generations=8, persistent=persistent) as b:
            for x, y, gens, p, seed in specs:
                b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
            note("split and narrow, persistent", persistent)
            b.run()
            results[persistent] = [b.fetch(i) for i in range(len(specs))]
    for i, (x, y, gens, p, seed) in enumerate(specs):
        ref = ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
        assert np.array_equal(results[True][i], ref), (x, y, gens)
        assert np.array_equal(results[False][i], results[True][i]), (x, y, gens)


def test_persistent_wait_timeout_is_reported():
    """GOL_BATCH_WAIT_PROBE makes every item that continues a grid wait, with a
    short spin limit, for a count no item reaches: the run stops, every kernel
    ends, gol_batch_run reports GOL_ERR_RUNTIME, and a new batch runs normally."""
    specs = [(2048, 1024, 40, 0.5, 41), (1024, 1024, 30, 0.4, 42), (128, 128, 50, 0.3, 43)]
    with golhip.Batch(round_generations=8, persistent=True, flags=golhip.BATCH_WAIT_PROBE) as b:
        for x, y, gens, p, seed in specs:
            b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
        note("persistent wait probe")
        with pytest.raises(golhip.GolError) as ei:
            b.run()
        assert ei.value.code == golhip.GOL_ERR_RUNTIME
    with golhip.Batch(round_generations=8, persistent=True) as b:
        for x, y, gens, p, seed in specs:
            b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, gens)
        note("persistent batch after the probe")
        b.run()
        for i, (x, y, gens, p, seed) in enumerate(specs):
            assert np.array_equal(b.fetch(i), ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, gens))


def test_persistent_chained_items_of_one_tile_grids():
    """A grid one tile spans is one queue item of up to 2^20 generations; longer
    runs become chained items, each waiting on the device for the one before.
    Narrow and one-tile grids past that length equal the per-round launches
    and the oracle."""
    gens = (1 << 20) + 777
    specs = [(64, 64, gens, 0.3, 51), (128, 40, gens, 0.4, 52), (2048, 64, 3000, 0.5, 53)]
    results = {}
    for persistent in (False, True):
        with golhip.Batch(persistent=persistent) as b:
            for x, y, g, p, seed in specs:
                b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, g)
            note("chained items, persistent", persistent)
            b.run()
            results[persistent] = [b.fetch(i) for i in range(len(specs))]
    for i, (x, y, g, p, seed) in enumerate(specs):
        assert np.array_equal(results[True][i], results[False][i]), (x, y, g)
    x, y, g, p, seed = specs[0]
    assert np.array_equal(results[True][0], ol.bitpar_run(golhip.fill_bernoulli(x, y, seed, p), x, y, g))


def test_persistent_long_waits_of_wide_one_tile_grids():
    """Wide one-tile grids (2048 x 256: not narrow, one workgroup) past 2^20
    generations: each is two chained items of about 1.6 s each on one CU, so a
    chained item can wait through another workgroup's whole item; the wait is
    bounded by device time, not by a poll count. Equal to the per-round
    launches."""
    gens = [(1 << 20) + 1000 + 77 * k for k in range(3)]
    specs = [(2048, 256, g, 0.35, 61 + k) for k, g in enumerate(gens)] + \
        [(128, 128, 5000 + 100 * k, 0.4, 70 + k) for k in range(4)]
    results = {}
    for persistent in (False, True):
        with golhip.Batch(persistent=persistent) as b:
            for x, y, g, p, seed in specs:
                b.add(golhip.fill_bernoulli(x, y, seed, p), x, y, g)
            note("wide one-tile grids past 2^20 generations, persistent", persistent)
            b.run()
            results[persistent] = [b.fetch(i) for i in range(len(specs))]
    for i, (x, y, g, p, seed) in enumerate(specs):
        assert np.array_equal(results[True][i], results[False][i]), (x, y, g)


def test_persistent_batch_of_zero_generation_grids():
    """No items at all: nothing is launched and every grid keeps its initial state."""
    with golhip.Batch(persistent=True) as b:
        words = [golhip.fill_bernoulli(256, 128, 60 + i, 0.4) for i in range(3)]
        for w in words:
            b.add(w, 256, 128, 0)
        note("persistent batch of 0-generation grids")
        b.run()
        for i, w in enumerate(words):
            assert np.array_equal(b.fetch(i), w)


@pytest.mark.parametrize("persistent", [False, True])
def test_select_plan_keeps_results(persistent):
    """GOL_BATCH_SELECT_PLAN times the batch under the measured choices and
    under the built-in model and keeps the faster plan: both plans are timed,
    and every grid is bit-exact whichever plan is kept; a reset restores the
    initial states, which fetch then returns."""
    specs = mixed_specs(24, 9, max_side=2048, max_gens=200)
    try:
        golhip.batch_autotune([(2048, 2048), (1024, 1024)], round_generations=16, generations=32,
                              cells_per_measure=1 << 24)
        with golhip.Batch(round_generations=16, persistent=persistent, flags=golhip.BATCH_SELECT_PLAN) as b:
            words = []
            for x, y, gens, p, seed in specs:
                words.append(golhip.fill_bernoulli(x, y, seed, p))
                b.add(words[-1], x, y, gens)
            note("select plan, persistent", persistent)
            b.run()
            info = b.info()
            assert info.plan_tuned_ms > 0 and info.plan_builtin_ms > 0
            assert info.plan_builtin == (1 if info.plan_builtin_ms <= info.plan_tuned_ms else 0)
            b.run()
            for i, (x, y, gens, p, seed) in enumerate(specs):
                assert np.array_equal(b.fetch(i), ol.bitpar_run(words[i], x, y, gens)), (x, y, gens)
            b.reset()
            for i in (0, 7):
                assert np.array_equal(b.fetch(i), words[i])
    finally:
        golhip.batch_tuning_reset()
