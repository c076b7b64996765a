"""GPU parity of change tracking (GOL_KERNEL_ACTIVE): bit-exact against the
oracle and against the dense kernel, on mostly-stable, fully active and
oscillating inputs, with generation counts that end on partial blocks."""

import json
import os

import numpy as np
import pytest

import golhip
import oracle_lib as ol

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))


def note(*a):
    print("[case]", *a, flush=True)


def run(words, x, y, gens, **kw):
    with golhip.Grid(x, y, kernel=golhip.KERNEL_ACTIVE, **kw) as g:
        g.upload(words)
        g.step(gens)
        return g.download(), g.info()


@pytest.mark.parametrize("steps", [4, 12, 16])
def test_patches_against_naive(steps):
    x, y = 4096, 1024
    w = golhip.fill_patches(x, y, 3, 12, 64, 0.35)
    cells = ol.unpack(w, x, y)
    for gens in (steps * 10, steps * 10 + 5, 301):
        note("active patches", x, y, "steps", steps, "gens", gens)
        out, info = run(w, x, y, gens, temporal_steps=steps)
        ref = ol.naive_run(cells, gens)
        assert np.array_equal(ol.unpack(out, x, y), ref)
        assert info.tracked_blocks == gens // steps


# Tiles are 32 * lane_dwords cells x (64 - 2T) rows: partial last tile
# columns and rows, odd dword counts (rows-32), grids shorter than a
# wavefront, T = 1 and 21, every tile width, and both halo forms: one halo
# dword shared by both sides up to T = 16, a dword on each side from T = 17.
@pytest.mark.parametrize("x,y,steps,word_bits,dw", [
    (32, 1, 12, 32, 8), (96, 5, 3, 32, 8), (2112, 70, 12, 64, 8), (288, 131, 21, 32, 8), (4128, 300, 1, 32, 8),
    (640, 64, 20, 64, 8), (256, 22, 21, 64, 8), (1344, 257, 7, 64, 8), (2080, 100, 16, 32, 8), (2080, 100, 17, 32, 8),
    (32, 3, 12, 32, 4), (160, 41, 12, 32, 4), (2112, 70, 12, 64, 4), (4096, 300, 16, 64, 4), (1344, 257, 21, 64, 4),
    (1088, 90, 16, 64, 4), (1088, 90, 17, 64, 4),
    (64, 1, 12, 32, 2), (96, 5, 3, 32, 2), (2112, 70, 16, 64, 2), (160, 41, 17, 32, 2), (1344, 257, 21, 64, 2),
    (4128, 300, 16, 32, 2),
])
def test_ragged_sizes_against_naive(x, y, steps, word_bits, dw):
    c = ol.bernoulli_cells(x, y, x * 7 + y, 0.4)
    gens = 3 * steps + 2
    note("active ragged", x, y, "steps", steps, "word_bits", word_bits, "lane_dwords", dw, "gens", gens)
    out, info = run(ol.pack(c, word_bits), x, y, gens, word_bits=word_bits, temporal_steps=steps, lane_dwords=dw)
    assert np.array_equal(ol.unpack(out, x, y, word_bits), ol.naive_run(c, gens))
    assert info.lane_dwords == dw
    assert info.tiles == -(-x // (32 * dw)) * -(-y // (64 - 2 * steps))
    assert info.tracked_blocks == gens // steps
    # a wavefront per tile (these grids are far below the resident limit), and at
    # least one workgroup per sub-list of the tile list (8), also for a one-tile grid
    assert info.workgroups == max(-(-info.tiles // 4), 8)


def test_tile_limits():
    with pytest.raises(golhip.GolError):
        golhip.Grid(2048, 2048, kernel=golhip.KERNEL_ACTIVE, temporal_steps=22)
    with pytest.raises(golhip.GolError):
        golhip.Grid(2048, 2048, kernel=golhip.KERNEL_ACTIVE, lane_dwords=1)


def test_skips_tiles_on_a_stable_grid():
    x = y = 8192
    w = golhip.fill_patches(x, y, 5, 4, 64, 0.35)
    note("active skip", x, y)
    out, info = run(w, x, y, 1200, temporal_steps=12)
    tiles = info.tiles
    blocks = info.tracked_blocks
    assert blocks == 100
    # after the first block almost all of the grid is quiet
    assert info.active_tiles_total < 0.25 * tiles * blocks
    ref = ol.bitpar_run(w, x, y, 1200)
    assert np.array_equal(out, ref)


def test_fully_active_random():
    x, y = 2048, 512
    c = ol.random_cells(x, y, 9)
    note("active random", x, y)
    out, _ = run(ol.pack(c), x, y, 123)
    assert np.array_equal(ol.unpack(out, x, y), ol.naive_run(c, 123))


def test_oscillators_stay_correct_between_blocks():
    # blinkers every 4 cells (the reference's always-changing loader,
    # data_loader.hpp:117-134): period 2, unchanged at every even block boundary
    x, y = 4096, 512
    c = np.zeros((y, x), np.uint8)
    for yy in range(1, y - 2, 4):
        for xx in range(0, x - 3, 4):
            c[yy, xx:xx + 3] = 1
    w = ol.pack(c)
    for gens in (12 * 7, 12 * 7 + 1, 12 * 7 + 3):
        note("active blinkers", gens)
        out, info = run(w, x, y, gens, temporal_steps=12)
        assert np.array_equal(ol.unpack(out, x, y), ol.naive_run(c, gens))
    # odd block length: the blinkers change at every block boundary
    out, info = run(w, x, y, 9 * 11, temporal_steps=9)
    assert np.array_equal(ol.unpack(out, x, y), ol.naive_run(c, 99))
    assert info.active_tiles_total == info.tiles * info.tracked_blocks


PULSAR = [
    "..OOO...OOO..",
    ".............",
    "O....O.O....O",
    "O....O.O....O",
    "O....O.O....O",
    "..OOO...OOO..",
    ".............",
    "..OOO...OOO..",
    "O....O.O....O",
    "O....O.O....O",
    "O....O.O....O",
    ".............",
    "..OOO...OOO..",
]


def test_pulsars_quiet_at_12_listed_at_16():
    # period-3 oscillators return to their state at boundaries of 12-generation
    # blocks but not of 16-generation ones (the default): there they keep their
    # tiles listed; the result is exact either way
    x, y = 2048, 512
    c = np.zeros((y, x), np.uint8)
    p = np.array([[ch == "O" for ch in row] for row in PULSAR], np.uint8)
    n = 0
    for yy in range(40, y - 40, 96):
        for xx in range(40, x - 40, 160):
            c[yy:yy + 13, xx:xx + 13] = p
            n += 1
    gens = 16 * 12
    want = ol.naive_run(c, gens)
    note("active pulsars", n)
    out12, info12 = run(ol.pack(c), x, y, gens, temporal_steps=12)
    out16, info16 = run(ol.pack(c), x, y, gens)
    assert info16.temporal_steps == 16
    assert np.array_equal(ol.unpack(out12, x, y), want)
    assert np.array_equal(ol.unpack(out16, x, y), want)
    # 12: every tile in the first block, then nothing changes at a boundary
    assert info12.active_tiles_total == info12.tiles
    # 16: each pulsar keeps at least its own tile listed in every block
    assert info16.active_tiles_total >= info16.tiles + n * (info16.tracked_blocks - 1)


GLIDER_SE = np.array([[0, 1, 0], [0, 0, 1], [1, 1, 1]], np.uint8)


def listed_tiles_model(cells, gens, steps, dw):
    """Tiles a tracked run advances, by the listing rule stepped on the CPU: the
    first block advances every tile; a later block advances a tile that changed
    over the block before, and each neighbour of a changed tile that one of its
    changed cells lies within `steps` cells of (Chebyshev distance). The blocks'
    states come from the oracle's bit-parallel engine."""
    y, x = cells.shape
    tw, tr = 32 * dw, 64 - 2 * steps
    ty, tx = -(-y // tr), -(-x // tw)
    words = ol.pack(cells, 64) if x % 64 == 0 else None
    prev = cells
    total = tx * ty
    for b in range(gens // steps - 1):
        cur = (ol.unpack(ol.bitpar_run(ol.pack(prev, 64), x, y, steps), x, y, 64) if words is not None
               else ol.naive_run(prev, steps))
        d = np.zeros((ty * tr, tx * tw), bool)
        d[:y, :x] = prev != cur
        d = d.reshape(ty, tr, tx, tw)
        r = steps
        anyc = d.any((1, 3))
        part = {(dy, dx): d[:, rows, :, cols].any((1, 3))
                for dy, rows in ((-1, slice(0, r)), (0, slice(None)), (1, slice(tr - r, tr)))
                for dx, cols in ((-1, slice(0, r)), (0, slice(None)), (1, slice(tw - r, tw)))}
        listed = anyc.copy()
        for (dy, dx), m in part.items():
            # a change in region (dy, dx) of tile (i, j) lists tile (i + dy, j + dx)
            src = m[max(0, -dy):ty - max(0, dy), max(0, -dx):tx - max(0, dx)]
            listed[max(0, dy):ty + min(0, dy), max(0, dx):tx + min(0, dx)] |= src
        total += int(listed.sum())
        prev = cur
    return total


# The listing rule is exact and tight: gliders crossing tile corners in all
# four directions, blinkers and random patches near tile edges, at block
# depths with both halo forms and every tile width; the tiles the GPU advanced
# equal the CPU model's count, and the result equals the naive oracle's.
@pytest.mark.parametrize("x,y,steps,dw", [
    (2112, 300, 16, 2), (2048, 257, 9, 2), (2048, 300, 21, 2), (2176, 200, 12, 4), (2048, 129, 5, 8),
])
def test_listing_follows_reach(x, y, steps, dw):
    rng = np.random.default_rng(x + y + steps)
    c = np.zeros((y, x), np.uint8)
    tw, tr = 32 * dw, 64 - 2 * steps
    for k in range(24):
        gy, gx = int(rng.integers(0, y - 8)), int(rng.integers(0, x - 8))
        g = GLIDER_SE
        if k & 1:
            g = g[:, ::-1]
        if k & 2:
            g = g[::-1, :]
        c[gy:gy + 3, gx:gx + 3] |= g
    for k in range(16):  # blinkers hugging tile edges (period 2)
        ty, tx = int(rng.integers(0, -(-y // tr))), int(rng.integers(0, -(-x // tw)))
        yy = min(ty * tr + int(rng.choice([0, 1, steps - 1, steps, tr - steps - 1, tr - 2])), y - 1)
        xx = min(tx * tw + int(rng.choice([0, steps - 2, steps - 1, steps, tw - steps - 1, tw - 3])), x - 3)
        c[yy, xx:xx + 3] = 1
    for k in range(6):  # small random soups
        gy, gx = int(rng.integers(0, y - 12)), int(rng.integers(0, x - 12))
        c[gy:gy + 12, gx:gx + 12] |= (rng.random((12, 12)) < 0.4).astype(np.uint8)
    gens = 14 * steps
    note("active listing", x, y, "steps", steps, "lane_dwords", dw, "gens", gens)
    out, info = run(ol.pack(c), x, y, gens, temporal_steps=steps, lane_dwords=dw)
    assert np.array_equal(ol.unpack(out, x, y), ol.naive_run(c, gens))
    assert info.tracked_blocks == 14
    assert info.active_tiles_total == listed_tiles_model(c, gens, steps, dw)


def test_split_calls_and_reupload():
    x, y = 2048, 600
    w = golhip.fill_patches(x, y, 11, 6, 64, 0.35)
    ref = ol.bitpar_run(w, x, y, 250)
    with golhip.Grid(x, y, kernel=golhip.KERNEL_ACTIVE, temporal_steps=12) as g:
        g.upload(w)
        for k in (7, 24, 100, 1, 118):
            g.step(k)
        assert np.array_equal(g.download(), ref)
        g.upload(w)
        g.step(250)
        assert np.array_equal(g.download(), ref)


def test_width_adapts_on_a_busy_grid():
    # a dense soup lists far more than 7 % of the 64-cell tiles: the step
    # moves to 128-cell tiles once the host has read a chunk's sample; exact
    # throughout, also across calls and after a new upload (which starts on
    # 64-cell tiles again)
    x, y = 4096, 2048
    w = golhip.fill_bernoulli(x, y, 23, 0.35)
    ref1 = ol.bitpar_run(w, x, y, 2500)
    ref2 = ol.bitpar_run(ref1, x, y, 3000)
    with golhip.Grid(x, y, kernel=golhip.KERNEL_ACTIVE) as g:
        g.upload(w)
        note("active width switching", x, y)
        g.step(2500)
        assert np.array_equal(g.download(), ref1)
        first = g.info()
        assert first.tile_width_switches >= 1
        g.step(3000)
        assert np.array_equal(g.download(), ref2)
        g.upload(w)
        assert g.info().lane_dwords == 2
        g.step(2500)
        assert np.array_equal(g.download(), ref1)
    # an explicit width never changes
    with golhip.Grid(x, y, kernel=golhip.KERNEL_ACTIVE, lane_dwords=2) as g:
        g.upload(w)
        g.step(2500)
        assert np.array_equal(g.download(), ref1)
        assert g.info().tile_width_switches == 0


def test_width_returns_when_the_soup_settles():
    # a 1024^2 soup at 0.35 lists over 7 % of the 64-cell tiles for its first
    # few thousand generations and under 5 % of the 128-cell tiles by about
    # generation 8000 (the listing rule stepped on the CPU: 0.49 of 64-cell
    # tiles at 2048, 0.035 of 128-cell tiles at 8192, 0.004 from 12288 on):
    # one step moves to 128-cell tiles and back, exact; in pieces too
    x, y, gens = 1024, 1024, 12000
    w = golhip.fill_bernoulli(x, y, 23, 0.35)
    ref = ol.bitpar_run(w, x, y, gens)
    with golhip.Grid(x, y, kernel=golhip.KERNEL_ACTIVE) as g:
        g.upload(w)
        note("active width back", x, y, gens)
        g.step(gens)
        assert np.array_equal(g.download(), ref)
        info = g.info()
        assert info.tile_width_switches >= 2 and info.lane_dwords == 2
        g.upload(w)
        for _ in range(4):
            g.step(gens // 4)
        assert np.array_equal(g.download(), ref)
        assert g.info().lane_dwords == 2


def reference_workload_input(f):
    """The inputs of the reference's change-tracking workloads
    (cluster-experiments/generate-final-measurements.py:21, 27-30)."""
    x, y = f["x"], f["y"]
    if f["generator"] == "zeros":
        return np.zeros(x * y // 64, np.uint64)
    if f["generator"] == "always-changing":
        # AlwaysChangingSpaceLoader (data_loader.hpp:117-134): a horizontal
        # blinker at every 4th column and row
        c = np.zeros((y, x), np.uint8)
        for k in range(3):
            c[0::4, k::4] = 1
        return ol.pack(c, 64)
    return golhip.fill_lexicon(x, y, f["params"]["expression"])


# The reference's published change-tracking measurements (no work, full work,
# 33 % and 66 % busy), each 16384^2 x 10000, bit-exact against the oracle.
@pytest.mark.parametrize("name", ["zeros-16384", "always-changing-16384", "spacefiller-33-work-16384",
                                  "spacefiller-66-work-16384"])
def test_reference_change_tracking_workloads(name):
    f = [r for r in FIX["bitpar_runs"] if r["name"] == name and r["generations"] == 10000][0]
    w = reference_workload_input(f)
    assert ol.sha256(w) == f["sha256_input"]
    note("active reference workload", name)
    out, info = run(w, f["x"], f["y"], 10000)
    assert ol.sha256(out) == f["sha256"]
    # the busy ones move to 128-cell tiles while their fronts collide
    if "work" in name:
        assert info.tile_width_switches >= 1


def test_full_size_patches_20000():
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "patches-16384" and r["generations"] == 20000][0]
    p = f["params"]
    w = golhip.fill_patches(16384, 16384, 42, p["n_patches"], p["patch"], p["p"])
    note("active 16384^2 patches 20000")
    out, info = run(w, 16384, 16384, 20000)
    assert ol.sha256(out) == f["sha256"]
    # sparse throughout: it stays on 64-cell tiles
    assert info.tile_width_switches == 0
    print("active fraction", info.active_tiles_total / (info.tiles * info.tracked_blocks))
