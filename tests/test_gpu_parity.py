"""GPU parity: the HIP engine, called through the C ABI, against the oracle.

Bit-exact everywhere (integer bit logic, no tolerance). Small cases compare
with the naive oracle on the same seeded input; full benchmark sizes compare
with fixtures written by the oracle's bit-parallel engine
(tests/golden/make_fixtures.py) and with size-independent properties.
"""

import json
import os
import sys

import numpy as np
import pytest

import golhip
import oracle_lib as ol

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))

KERNELS = [
    dict(kernel=golhip.KERNEL_PLAIN),
    dict(kernel=golhip.KERNEL_TEMPORAL),
    dict(kernel=golhip.KERNEL_TEMPORAL, temporal_steps=1),
    dict(kernel=golhip.KERNEL_TEMPORAL, temporal_steps=5),
]
SHAPES = [(1, 8, 8), (1, 16, 8), (1, 24, 8), (1, 32, 8), (1, 32, 16), (2, 16, 8), (4, 8, 8), (4, 8, 16), (8, 4, 16),
          (8, 5, 16), (8, 8, 8), (8, 10, 8)]


def note(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def run_gpu(words, x, y, gens, word_bits=64, **kw):
    with golhip.Grid(x, y, word_bits=word_bits, **kw) as g:
        g.upload(words)
        g.step(gens)
        return g.download()


def cells_for(f):
    if f["generator"] == "random":
        return ol.random_cells(f["x"], f["y"], f["seed"])
    return ol.bernoulli_cells(f["x"], f["y"], f["seed"], f["p"])


@pytest.mark.parametrize("kw", KERNELS, ids=lambda k: "-".join(f"{a}{b}" for a, b in k.items()))
def test_naive_fixtures(kw):
    for f in FIX["naive_runs"]:
        note("fixture", f["x"], f["y"], f["generations"], kw)
        words = ol.pack(cells_for(f), f["word_bits"])
        out = run_gpu(words, f["x"], f["y"], f["generations"], f["word_bits"], **kw)
        assert ol.sha256(out) == f["sha256"], (f, kw)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"dw{s[0]}-br{s[1]}-nw{s[2]}")
@pytest.mark.parametrize("steps", [1, 3, 8, 16])
def test_shapes_against_naive(shape, steps):
    dw, br, nw = shape
    for x, y, gens in [(64, 64, 37), (192, 150, 40), (2048 + 64, 33, 21), (32 * 64 * dw + 128, 70, 19),
                       (128, 1, 9), (64, 2, 9), (4096, 257, 17)]:
        if steps > 32 * dw and x // 32 > 64 * dw:
            continue
        if nw * br - 2 * steps < 1 and y > nw * br:
            continue
        note("shape", shape, "steps", steps, "grid", x, y, gens)
        c = ol.bernoulli_cells(x, y, x + 7 * y + gens, 0.45)
        ref = ol.naive_run(c, gens)
        out = run_gpu(ol.pack(c), x, y, gens, kernel=golhip.KERNEL_TEMPORAL, temporal_steps=steps,
                      lane_dwords=dw, band_rows=br, waves=nw)
        got = ol.unpack(out, x, y)
        if not np.array_equal(got, ref):
            d = np.argwhere(got != ref)
            pytest.fail(f"{len(d)} cells differ, first {d[:5].tolist()}")


@pytest.mark.parametrize("dw", [1, 2, 4, 8])
def test_plain_lane_widths(dw):
    for x, y, gens in [(64, 64, 30), (320, 97, 25), (64 * 64 * dw + 64, 40, 12), (32, 5, 7)]:
        note("plain dw", dw, x, y, gens)
        wb = 64 if x % 64 == 0 else 32
        c = ol.bernoulli_cells(x, y, 3 * x + y, 0.5)
        ref = ol.naive_run(c, gens)
        out = run_gpu(ol.pack(c, wb), x, y, gens, word_bits=wb, kernel=golhip.KERNEL_PLAIN, lane_dwords=dw)
        assert np.array_equal(ol.unpack(out, x, y, wb), ref)


def test_every_generation_matches():
    x, y = 256, 96
    c = ol.random_cells(x, y, 5)
    with golhip.Grid(x, y, kernel=golhip.KERNEL_TEMPORAL, temporal_steps=4) as g:
        g.upload(ol.pack(c))
        for i in range(80):
            note("generation", i)
            g.step(1)
            c = ol.naive_run(c, 1)
            assert np.array_equal(ol.unpack(g.download(), x, y), c), i


def test_split_steps_equal_one_step():
    x = y = 1024
    w = golhip.fill_random(x, y, 11)
    full = run_gpu(w, x, y, 123)
    with golhip.Grid(x, y) as g:
        g.upload(w)
        for k in (1, 2, 3, 17, 100):
            g.step(k)
        assert np.array_equal(g.download(), full)


def test_zero_generations_identity():
    w = golhip.fill_random(512, 64, 3)
    for kw in KERNELS:
        assert np.array_equal(run_gpu(w, 512, 64, 0, **kw), w)


@pytest.mark.parametrize("gens", [1, 100, 1000, 10000])
def test_full_size_random_16384(gens):
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "random-16384" and r["generations"] == gens][0]
    w = golhip.fill_random(16384, 16384, 42)
    assert ol.sha256(w) == f["sha256_input"]
    note("16384^2 random gens", gens)
    out = run_gpu(w, 16384, 16384, gens)
    assert ol.sha256(out) == f["sha256"]


def test_full_size_plain_equals_temporal():
    w = golhip.fill_random(16384, 16384, 42)
    note("16384^2 plain vs temporal, 100 gens")
    a = run_gpu(w, 16384, 16384, 100, kernel=golhip.KERNEL_PLAIN)
    b = run_gpu(w, 16384, 16384, 100, kernel=golhip.KERNEL_TEMPORAL, temporal_steps=7)
    assert np.array_equal(a, b)
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "random-16384" and r["generations"] == 100][0]
    assert ol.sha256(a) == f["sha256"]


@pytest.mark.parametrize("gens", [100, 20000])
def test_full_size_patches(gens):
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "patches-16384" and r["generations"] == gens][0]
    p = f["params"]
    w = golhip.fill_patches(16384, 16384, 42, p["n_patches"], p["patch"], p["p"])
    assert ol.sha256(w) == f["sha256_input"]
    note("16384^2 patches gens", gens)
    out = run_gpu(w, 16384, 16384, gens)
    assert ol.sha256(out) == f["sha256"]


def test_rectangular_4096x8192():
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "random-4096x8192"][0]
    w = golhip.fill_bernoulli(4096, 8192, 7, 0.4)
    note("4096x8192 bernoulli 500")
    out = run_gpu(w, 4096, 8192, 500)
    assert ol.sha256(out) == f["sha256"]


def test_wide_grid_x_tiles():
    # wider than one wavefront row segment: tiles overlap by a one-lane halo
    x, y, gens = 16384 + 4096, 200, 50
    c = ol.bernoulli_cells(x, y, 1, 0.5)
    w = ol.pack(c)
    ref = ol.unpack(ol.bitpar_run(w, x, y, gens), x, y)
    for steps in (8, 25):
        note("wide", x, y, steps)
        out = run_gpu(w, x, y, gens, kernel=golhip.KERNEL_TEMPORAL, temporal_steps=steps)
        assert np.array_equal(ol.unpack(out, x, y), ref)


@pytest.mark.parametrize("x,y,gens,steps", [
    (16384, 2000, 37, 8),     # full-width slabs, ragged height, partial last block
    (16384, 700, 64, 4),
    (8192, 1500, 45, 6),      # narrower than the lanes: columns masked
    (12288, 333, 20, 2),
    (16384, 96, 9, 8),        # fewer generations than one block
])
def test_persistent_slabs_against_oracle(x, y, gens, steps):
    note("persistent", x, y, gens, steps)
    w = ol.pack(ol.bernoulli_cells(x, y, x + y + gens, 0.45))
    ref = ol.bitpar_run(w, x, y, gens)
    with golhip.Grid(x, y, kernel=golhip.KERNEL_TEMPORAL, temporal_steps=steps, lane_dwords=8, band_rows=10,
                     waves=8) as g:
        persistent = g.info().persistent
        g.upload(w)
        g.step(gens)
        out = g.download()
    assert np.array_equal(out, ref)
    with golhip.Grid(x, y, kernel=golhip.KERNEL_TEMPORAL, temporal_steps=steps, lane_dwords=8, band_rows=10,
                     waves=8, flags=golhip.FLAG_NO_PERSISTENT) as g:
        assert g.info().persistent == 0
        g.upload(w)
        g.step(gens)
        assert np.array_equal(g.download(), ref)
    assert persistent in (0, 1)


@pytest.mark.parametrize("band_rows,waves", [(10, 8), (9, 8), (8, 8), (5, 16), (4, 16)])
@pytest.mark.parametrize("steps", [2, 4, 5, 8])
def test_persistent_slab_shapes_against_oracle(band_rows, waves, steps):
    """Every compiled slab shape at several block depths, on a full-width grid
    with a ragged last slab and a partial last block. Where the bands are
    taller than the depth (the 80- and 72-row slabs) every band boundary lies
    among the own rows, so the halo import follows the next block's first
    barrier; elsewhere halo rows sit on band edges and the import comes
    first."""
    x, y = 16384, 1000
    gens = 4 * steps + 3
    note("persistent shape", band_rows, waves, steps, gens)
    w = ol.pack(ol.bernoulli_cells(x, y, band_rows * 100 + steps, 0.4))
    ref = ol.bitpar_run(w, x, y, gens)
    with golhip.Grid(x, y, kernel=golhip.KERNEL_TEMPORAL, temporal_steps=steps, lane_dwords=8,
                     band_rows=band_rows, waves=waves) as g:
        assert g.info().persistent == 1
        g.upload(w)
        g.step(gens)
        assert np.array_equal(g.download(), ref)


@pytest.mark.parametrize("x,y,word_bits,band_rows,steps", [
    (8224, 900, 32, 9, 4),    # 257 dwords: an odd row width in 32-bit words
    (4096, 1300, 64, 9, 3),
    (12352, 777, 64, 10, 8),
    (16320, 2000, 64, 9, 1),  # one generation per block: every block exchanges
])
def test_persistent_slabs_masked_columns(x, y, word_bits, band_rows, steps):
    """The slab launch on rows narrower than a wavefront's 512 dwords (the
    column-masked variant) for the 72- and 80-row slabs, against the oracle."""
    gens = 6 * steps + 1
    note("persistent masked", x, y, word_bits, band_rows, steps, gens)
    cells = ol.bernoulli_cells(x, y, x + steps, 0.45)
    ref = ol.naive_run(cells, gens)
    with golhip.Grid(x, y, word_bits=word_bits, kernel=golhip.KERNEL_TEMPORAL, temporal_steps=steps,
                     lane_dwords=8, band_rows=band_rows, waves=8) as g:
        assert g.info().persistent == 1
        g.upload(ol.pack(cells, word_bits))
        g.step(gens)
        assert np.array_equal(ol.unpack(g.download(), x, y, word_bits), ref)


def test_persistent_repeated_launches_match_fixture():
    """Ten persistent launches of 10000 generations on one handle, the input
    coming alternately from a device buffer and from the host (the benchmark's
    sequence): every result equals the fixture. A halo row published before
    its write-back reached memory would make one of them differ."""
    w = golhip.fill_random(16384, 16384, 42)
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "random-16384" and r["generations"] == 10000][0]
    with golhip.Grid(16384, 16384, kernel=golhip.KERNEL_PLAIN) as src, golhip.Grid(16384, 16384) as g:
        assert g.info().persistent == 1
        src.upload(w)
        src.sync()
        for i in range(10):
            note("persistent repeat", i)
            if i % 2:
                g.upload(w)
            else:
                g.upload_device(src.device_state())
            g.step(10000)
            assert ol.sha256(g.download()) == f["sha256"], f"launch {i}"


def test_persistent_wait_timeout_is_reported():
    """The failure path of the persistent launch, forced from the host: every
    slab waits, with a short spin limit, for a block count no slab publishes,
    so each wait ends by its time-out and every slab leaves; the error reaches
    the caller at the next synchronising call. A normal handle afterwards
    matches the fixture (the device is unharmed)."""
    w = golhip.fill_random(16384, 16384, 42)
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "random-16384" and r["generations"] == 1000][0]
    with golhip.Grid(16384, 16384, flags=golhip.FLAG_SLAB_WAIT_PROBE) as g:
        assert g.info().persistent == 1
        g.upload(w)
        note("slab wait probe")
        g.step(1000)
        with pytest.raises(golhip.GolError) as e:
            g.sync()
        assert e.value.code == golhip.GOL_ERR_RUNTIME and "waited too long" in str(e.value)
        g.sync()  # reported once
    with golhip.Grid(16384, 16384) as g:
        g.upload(w)
        g.step(1000)
        assert ol.sha256(g.download()) == f["sha256"]


def test_persistent_used_for_16384_and_split_calls():
    w = golhip.fill_random(16384, 16384, 42)
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "random-16384" and r["generations"] == 1000][0]
    with golhip.Grid(16384, 16384) as g:
        info = g.info()
        note("persistent default", info.persistent, info.lane_dwords, info.band_rows, info.waves, info.temporal_steps)
        assert info.persistent == 1
        g.upload(w)
        for k in (1, 7, 8, 100, 300, 584):   # 1000 generations in uneven calls
            g.step(k)
        assert ol.sha256(g.download()) == f["sha256"]


# the reference's published workload: one spacefiller in the middle of a
# 16384^2 grid (cluster-experiments/generate-final-measurements.py:33)
@pytest.mark.parametrize("gens,kw", [
    (1000, dict(kernel=golhip.KERNEL_PLAIN)),
    (20000, dict(kernel=golhip.KERNEL_TEMPORAL)),
    (1000, dict(kernel=golhip.KERNEL_TEMPORAL, flags=golhip.FLAG_NO_PERSISTENT)),
    (20000, dict(kernel=golhip.KERNEL_ACTIVE)),
    (20000, dict(kernel=golhip.KERNEL_ACTIVE, lane_dwords=8)),
    (20000, dict(kernel=golhip.KERNEL_ACTIVE, lane_dwords=2)),
])
def test_spacefiller_fixture(gens, kw):
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "spacefiller-16384" and r["generations"] == gens][0]
    w = golhip.fill_lexicon(16384, 16384, f["params"]["expression"])
    assert ol.sha256(w) == f["sha256_input"]
    note("spacefiller", gens, kw)
    with golhip.Grid(16384, 16384, **kw) as g:
        g.upload(w)
        g.step(gens)
        assert ol.sha256(g.download()) == f["sha256"]
