"""GPU parity of the column and tile bit layouts at the C ABI (gol_config.layout):
words in the reference's BitColumns / BitTile layouts (bit_modes.hpp:10-125)
are converted to rows on the device at upload and back at download. Results are
compared, bit-exact, with the oracle's packing of the naive loop's result."""

import ctypes
import json
import os

import numpy as np
import pytest

import golhip
import oracle_lib as ol

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))
LAYOUTS = [(golhip.LAYOUT_COLUMNS, 16), (golhip.LAYOUT_COLUMNS, 32), (golhip.LAYOUT_COLUMNS, 64),
           (golhip.LAYOUT_TILES, 16), (golhip.LAYOUT_TILES, 32), (golhip.LAYOUT_TILES, 64)]


def note(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("layout,w", LAYOUTS)
@pytest.mark.parametrize("x,y", [(2112, 192), (2080, 64), (64, 128)])
def test_round_trip_without_generations(layout, w, x, y):
    c = ol.bernoulli_cells(x, y, x + y + w, 0.5)
    words = ol.pack_layout(c, layout, w)
    note("layout round trip", layout, w, x, y)
    with golhip.Grid(x, y, word_bits=w, layout=layout) as g:
        assert g.info().layout == layout
        g.upload(words)
        g.step(0)
        assert np.array_equal(g.download(), words)


@pytest.mark.parametrize("layout,w", LAYOUTS)
@pytest.mark.parametrize("kernel", [golhip.KERNEL_TEMPORAL, golhip.KERNEL_ACTIVE, golhip.KERNEL_PLAIN])
def test_layouts_against_oracle(layout, w, kernel):
    x, y, gens = 2112, 192, 29
    c = ol.bernoulli_cells(x, y, 3 * w + layout, 0.4)
    want = ol.pack_layout(ol.naive_run(c, gens), layout, w)
    note("layout", layout, w, "kernel", kernel)
    with golhip.Grid(x, y, word_bits=w, kernel=kernel, layout=layout) as g:
        g.upload(ol.pack_layout(c, layout, w))
        g.step(gens)
        assert np.array_equal(g.download(), want)


def device_copy(host: np.ndarray):
    """Device memory holding `host`, allocated with the HIP runtime the engine
    library already loaded (torch brings a runtime of its own)."""
    hip = ctypes.CDLL("libamdhip64.so")
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(host.nbytes)) == 0
    assert hip.hipMemcpy(ptr, ctypes.c_void_p(host.ctypes.data), ctypes.c_size_t(host.nbytes), 1) == 0  # H2D
    return hip, ptr


@pytest.mark.parametrize("layout,w", [(golhip.LAYOUT_COLUMNS, 32), (golhip.LAYOUT_TILES, 64)])
def test_upload_from_device_memory(layout, w):
    x, y, gens = 1024, 128, 40
    c = ol.bernoulli_cells(x, y, 77, 0.45)
    words = ol.pack_layout(c, layout, w)
    hip, ptr = device_copy(words)
    note("upload_device", layout, w)
    try:
        with golhip.Grid(x, y, word_bits=w, layout=layout, device=0) as g:
            g.upload_device(ptr.value)
            g.step(gens)
            assert np.array_equal(g.download(), ol.pack_layout(ol.naive_run(c, gens), layout, w))
    finally:
        hip.hipFree(ptr)


def test_full_size_tiles_against_fixture():
    # the 16384^2 seed-42 grid in the 8x8-tile layout, 1000 generations: the
    # result, converted back to rows, equals the row-layout fixture
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "random-16384" and r["generations"] == 1000][0]
    rows = golhip.fill_random(16384, 16384, 42)
    cells = ol.unpack(rows, 16384, 16384)
    tiles = ol.pack_layout(cells, golhip.LAYOUT_TILES, 64)
    del cells
    note("tiles-64 16384^2 x 1000")
    with golhip.Grid(16384, 16384, word_bits=64, layout=golhip.LAYOUT_TILES) as g:
        g.upload(tiles)
        g.step(1000)
        out = g.download()
    assert ol.sha256(ol.pack(ol.unpack_layout(out, 16384, 16384, golhip.LAYOUT_TILES, 64))) == f["sha256"]


# ---- kernels native to the layouts (GOL_FLAG_NATIVE_LAYOUT): no conversion ----
NATIVE = [(golhip.LAYOUT_TILES, 64), (golhip.LAYOUT_COLUMNS, 32), (golhip.LAYOUT_COLUMNS, 64)]


@pytest.mark.parametrize("layout,w", NATIVE)
@pytest.mark.parametrize("kernel,steps", [(golhip.KERNEL_PLAIN, 0), (golhip.KERNEL_TEMPORAL, 1),
                                          (golhip.KERNEL_TEMPORAL, 3), (golhip.KERNEL_TEMPORAL, 8),
                                          (golhip.KERNEL_TEMPORAL, 9), (golhip.KERNEL_TEMPORAL, 16)])
@pytest.mark.parametrize("x,y", [(2112, 1088), (96, 64), (544, 4160)])
def test_native_layouts_against_oracle(layout, w, kernel, steps, x, y):
    # (2112, 1088): several workgroups across (tile columns / grid columns with
    # halo) and, for tiles, down; (544, 4160): several down for both (4160 rows
    # = 130 column dwords, 520 tile rows); (96, 64): one workgroup, masked lanes
    gens = 37
    c = ol.bernoulli_cells(x, y, x * 3 + y + w + steps, 0.4)
    want = ol.pack_layout(ol.naive_run(c, gens), layout, w)
    note("native layout", layout, w, "kernel", kernel, "steps", steps, x, y)
    with golhip.Grid(x, y, word_bits=w, kernel=kernel, layout=layout, temporal_steps=steps,
                     flags=golhip.FLAG_NATIVE_LAYOUT) as g:
        info = g.info()
        assert info.native_layout == 1 and info.temporal_steps == (1 if kernel == golhip.KERNEL_PLAIN else steps)
        g.upload(ol.pack_layout(c, layout, w))
        g.step(gens)
        assert np.array_equal(g.download(), want)
        assert g.info().launches == -(-gens // info.temporal_steps)


@pytest.mark.parametrize("layout,w", NATIVE)
def test_native_layouts_every_generation(layout, w):
    # split calls of 1..5 generations, each result against the naive loop
    x, y = 640, 512
    c = ol.bernoulli_cells(x, y, 91 + w, 0.45)
    note("native layout every generation", layout, w)
    with golhip.Grid(x, y, word_bits=w, kernel=golhip.KERNEL_TEMPORAL, layout=layout,
                     flags=golhip.FLAG_NATIVE_LAYOUT) as g:
        g.upload(ol.pack_layout(c, layout, w))
        cur = c
        for n in (1, 2, 3, 4, 5, 1, 0, 2):
            g.step(n)
            cur = ol.naive_run(cur, n)
            assert np.array_equal(g.download(), ol.pack_layout(cur, layout, w)), n


@pytest.mark.parametrize("layout,w", NATIVE)
def test_native_layouts_match_converted(layout, w):
    # the native kernels and the conversion to rows give the same words
    x, y, gens = 4096, 2048, 150
    words = ol.pack_layout(ol.bernoulli_cells(x, y, 5, 0.3), layout, w)
    outs = []
    for flags in (0, golhip.FLAG_NATIVE_LAYOUT):
        note("native vs converted", layout, w, flags)
        with golhip.Grid(x, y, word_bits=w, layout=layout, flags=flags) as g:
            g.upload(words)
            g.step(gens)
            outs.append(g.download())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("layout,w", NATIVE)
@pytest.mark.parametrize("kernel", [golhip.KERNEL_TEMPORAL, golhip.KERNEL_PLAIN])
def test_native_full_size_against_fixture(layout, w, kernel):
    # the 16384^2 seed-42 grid in the layout, 1000 generations, converted back
    # to rows on the host, equals the row-layout fixture
    f = [r for r in FIX["bitpar_runs"] if r["name"] == "random-16384" and r["generations"] == 1000][0]
    cells = ol.unpack(golhip.fill_random(16384, 16384, 42), 16384, 16384)
    words = ol.pack_layout(cells, layout, w)
    del cells
    note("native", layout, w, "kernel", kernel, "16384^2 x 1000")
    with golhip.Grid(16384, 16384, word_bits=w, layout=layout, kernel=kernel,
                     flags=golhip.FLAG_NATIVE_LAYOUT) as g:
        g.upload(words)
        g.step(1000)
        out = g.download()
    assert ol.sha256(ol.pack(ol.unpack_layout(out, 16384, 16384, layout, w))) == f["sha256"]
