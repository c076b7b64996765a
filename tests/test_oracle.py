"""Pins the CPU oracle (oracle/gol_oracle.cpp) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement fails it:
  * AN5D alive-count KATs (AN5D/GOL_AN5D.c:92-97) — an independent record of
    the rule and the neighbour count (a dropped neighbour, a wrong birth or
    survival count changes the counts);
  * the B3/S23 table over all 512 3x3 neighbourhoods, centred and at every
    grid edge and corner (pins the dead boundary: cells outside never count);
  * still lifes, oscillators and glider translation (closed forms);
  * the eight symmetries of the square (a transposed operand or a wrong
    index fails these);
  * the bit-parallel engine against the naive loop on ragged sizes;
  * the loaders: pure-Python mt19937 / mt19937_64 written from the C++
    standard's parameters (pinned by the standard's 10000th values) against
    the oracle's loader cells, the Bernoulli threshold in exact rationals,
    and the seed-42 fixture;
  * the batched / served workload fixtures: their smallest grids re-derived
    with the naive loop.
"""

import json
import os

import numpy as np
import pytest

import oracle_lib as ol

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))

# The rule as a table: next state indexed by [alive][live neighbours].
B3S23 = np.array([[0, 0, 0, 1, 0, 0, 0, 0, 0],
                  [0, 0, 1, 1, 0, 0, 0, 0, 0]], dtype=np.uint8)


def test_mt19937_standard_value():
    # [rand.predef]: the 10000th invocation of a default-constructed mt19937 is 4123659995
    assert ol.lib().oracle_mt19937_nth(5489, 10000) == FIX["mt19937_10000"] == 4123659995


class MersenneTwister:
    """The Mersenne twister family of [rand.eng.mers], written from the C++
    standard's parameters (not from any library's code): mt19937 and
    mt19937_64 of [rand.predef]. Independent of the oracle, which calls
    std::mt19937 / std::mt19937_64."""

    def __init__(self, w, n, m, r, a, u, d, s, b, t, c, l, f, seed):
        self.w, self.n, self.m, self.r, self.a = w, n, m, r, a
        self.u, self.d, self.s, self.b, self.t, self.c, self.l = u, d, s, b, t, c, l
        self.mask = (1 << w) - 1
        x = [seed & self.mask]
        for i in range(1, n):
            x.append((f * (x[-1] ^ (x[-1] >> (w - 2))) + i) & self.mask)
        self.x, self.i = x, n

    def __call__(self):
        if self.i == self.n:
            upper = (self.mask << self.r) & self.mask
            lower = (1 << self.r) - 1
            x = self.x
            for k in range(self.n):
                y = (x[k] & upper) | (x[(k + 1) % self.n] & lower)
                x[k] = x[(k + self.m) % self.n] ^ (y >> 1) ^ (self.a if y & 1 else 0)
            self.i = 0
        z = self.x[self.i]
        self.i += 1
        z ^= (z >> self.u) & self.d
        z ^= (z << self.s) & self.b
        z ^= (z << self.t) & self.c
        z ^= z >> self.l
        return z & self.mask


def mt19937(seed=5489):
    return MersenneTwister(32, 624, 397, 31, 0x9908B0DF, 11, 0xFFFFFFFF, 7, 0x9D2C5680, 15, 0xEFC60000, 18,
                           1812433253, seed)


def mt19937_64(seed=5489):
    return MersenneTwister(64, 312, 156, 31, 0xB5026F5AA96619E9, 29, 0x5555555555555555, 17, 0x71D67FFFEDA60000,
                           37, 0xFFF7EEE000000000, 43, 6364136223846793005, seed)


def nth(engine, n):
    for _ in range(n - 1):
        engine()
    return engine()


def test_python_engines_match_the_standard():
    # [rand.predef]: the 10000th outputs of default-constructed mt19937 / mt19937_64
    assert nth(mt19937(), 10000) == 4123659995
    assert nth(mt19937_64(), 10000) == 9981545732273789042


def test_mt19937_64_standard_value():
    # the engine of oracle_load_bernoulli
    assert ol.lib().oracle_mt19937_64_nth(5489, 10000) == 9981545732273789042
    assert ol.lib().oracle_mt19937_64_nth(99, 1000) == nth(mt19937_64(99), 1000)


@pytest.mark.parametrize("seed,p", [(99, 0.35), (7, 0.5), (123456789, 0.1), (2 ** 63 + 5, 0.9)])
def test_bernoulli_loader_against_python_restatement(seed, p):
    # cell i = u_i < p, u_i = (x_i >> 11) / 2^53 for the i-th mt19937_64 output
    # x_i. Computed in exact rationals: (x >> 11) < p * 2^53, so a 32-bit
    # engine, a shift of 12, or <= in place of < fails on some cell.
    from fractions import Fraction

    n = 4096
    got = ol.bernoulli_cells(64, n // 64, seed, p).reshape(-1)
    eng = mt19937_64(seed)
    limit = Fraction(p) * 2 ** 53
    want = np.array([1 if (eng() >> 11) < limit else 0 for _ in range(n)], np.uint8)
    assert np.array_equal(got, want)


def test_bernoulli_loader_threshold_is_strict():
    # p equal to a cell's own u: the cell is dead (u < p fails), and alive for
    # the next double above p
    seed = 31
    x = mt19937_64(seed)()
    u = (x >> 11) / 2.0 ** 53  # exact: 53-bit integer over a power of two
    assert ol.bernoulli_cells(64, 1, seed, u)[0, 0] == 0
    assert ol.bernoulli_cells(64, 1, seed, float(np.nextafter(u, 2.0)))[0, 0] == 1


def test_random_loader_against_python_restatement():
    # RandomOnesZerosDataLoader (data_loader.hpp:57-75) draws
    # uniform_int_distribution<int>(0, 1) from mt19937(seed). libstdc++ maps a
    # 32-bit draw x onto {0, 1} as (x * 2) >> 32 (Lemire's multiply-shift,
    # with no rejection possible for a range of 2): the top bit of x. The seed-42
    # fixture holds the oracle's first 256 cells; here the whole first 64 K.
    n = 65536
    got = ol.random_cells(256, n // 256, 42).reshape(-1)
    eng = mt19937(42)
    want = np.array([(eng() * 2) >> 32 for _ in range(n)], np.uint8)
    assert np.array_equal(got, want)


def test_loader_fixture():
    f = FIX["loader"]
    c = ol.random_cells(f["x"], f["y"], f["seed"])
    assert "".join(str(v) for v in c.reshape(-1)[:256]) == f["first_256"]
    assert ol.sha256(c) == f["sha256_cells"]
    assert int(c.sum()) == f["popcount"]
    assert set(np.unique(c)) <= {0, 1}


@pytest.mark.parametrize("n", [256, 512])
def test_an5d_kat(n):
    g = ol.naive_run(ol.glibc_rand_cells(n, 1985), 1024, periodic=True)
    assert int(g.sum()) == FIX["an5d_kats"][str(n)]


@pytest.mark.slow
def test_an5d_kat_1024():
    g = ol.naive_run(ol.glibc_rand_cells(1024, 1985), 1024, periodic=True)
    assert int(g.sum()) == FIX["an5d_kats"]["1024"]


def test_all_neighbourhoods_centred():
    # each 3x3 pattern in the middle of a 5x5 grid: the centre follows the table
    for pat in range(512):
        g = np.zeros((5, 5), np.uint8)
        bits = [(pat >> k) & 1 for k in range(9)]
        g[1:4, 1:4] = np.array(bits, np.uint8).reshape(3, 3)
        out = ol.naive_run(g, 1)
        alive = bits[4]
        n = sum(bits) - alive
        assert out[2, 2] == B3S23[alive, n], pat


def test_all_neighbourhoods_on_a_3x3_grid():
    # every cell of a 3x3 grid is on an edge: only in-grid cells count
    for pat in range(512):
        g = np.array([(pat >> k) & 1 for k in range(9)], np.uint8).reshape(3, 3)
        out = ol.naive_run(g, 1)
        for y in range(3):
            for x in range(3):
                n = int(g[max(0, y - 1):y + 2, max(0, x - 1):x + 2].sum()) - int(g[y, x])
                assert out[y, x] == B3S23[g[y, x], n], (pat, y, x)


def test_still_lifes_and_oscillators():
    g = np.zeros((12, 12), np.uint8)
    g[1:3, 1:3] = 1                                   # block
    g[6, 7:9] = 1; g[7, 6] = 1; g[7, 9] = 1; g[8, 7:9] = 1  # beehive
    assert np.array_equal(ol.naive_run(g, 1), g)
    assert np.array_equal(ol.naive_run(g, 17), g)
    b = np.zeros((5, 5), np.uint8)
    b[2, 1:4] = 1                                     # blinker (patterns.cpp:1593)
    b1 = ol.naive_run(b, 1)
    assert np.array_equal(b1, b.T)
    assert np.array_equal(ol.naive_run(b, 2), b)
    # a blinker touching the grid edge still oscillates (its outside cells stay dead)
    e = np.zeros((3, 3), np.uint8)
    e[0, :] = 1
    assert np.array_equal(ol.naive_run(e, 1), np.array([[0, 1, 0], [0, 1, 0], [0, 0, 0]], np.uint8))


def test_glider_translates():
    # the lexicon's glider (patterns.cpp:4237: XXX / X.. / .X.) moves (-1,-1) every 4 generations
    g = np.zeros((20, 20), np.uint8)
    g[10, 10:13] = 1
    g[11, 10] = 1
    g[12, 11] = 1
    out = ol.naive_run(g, 4)
    assert np.array_equal(out, np.roll(np.roll(g, -1, 0), -1, 1))
    out = ol.naive_run(g, 32)
    assert np.array_equal(out, np.roll(np.roll(g, -8, 0), -8, 1))


def test_glider_dies_into_block_at_corner():
    # dead boundary: a glider flying into the corner ends as a still life, never wraps
    g = np.zeros((8, 8), np.uint8)
    g[4, 4:7] = 1
    g[5, 4] = 1
    g[6, 5] = 1
    out = ol.naive_run(g, 200)
    assert np.array_equal(ol.naive_run(out, 1), out)
    assert out[6:, :].sum() == 0 and out[:, 6:].sum() == 0


@pytest.mark.parametrize("shape", [(37, 53), (64, 96), (1, 40), (40, 1), (2, 2)])
def test_symmetries(shape):
    rng = np.random.default_rng(sum(shape))
    g = (rng.random(shape) < 0.4).astype(np.uint8)
    ref = ol.naive_run(g, 9)
    for f in (np.transpose, np.fliplr, np.flipud, lambda a: np.rot90(a, 1), lambda a: np.rot90(a, 2),
              lambda a: np.rot90(a, 3), lambda a: np.rot90(a.T, 1)):
        assert np.array_equal(ol.naive_run(np.ascontiguousarray(f(g)), 9), f(ref))


def test_zero_generations_is_identity():
    g = ol.random_cells(64, 8, 3)
    assert np.array_equal(ol.naive_run(g, 0), g)
    w = ol.pack(g)
    assert np.array_equal(ol.bitpar_run(w, 64, 8, 0), w)


@pytest.mark.parametrize("x,y,gens,p", [(64, 64, 40, 0.5), (128, 37, 25, 0.3), (192, 200, 60, 0.45),
                                        (64, 1, 5, 0.6), (512, 96, 33, 0.5)])
def test_bitpar_matches_naive(x, y, gens, p):
    g = ol.bernoulli_cells(x, y, x * 1000 + y, p)
    ref = ol.naive_run(g, gens)
    got = ol.unpack(ol.bitpar_run(ol.pack(g), x, y, gens), x, y)
    assert np.array_equal(got, ref)


def test_bitpar_matches_naive_every_generation():
    g = ol.random_cells(128, 64, 77)
    w = ol.pack(g)
    for _ in range(60):
        g = ol.naive_run(g, 1)
        w = ol.bitpar_run(w, 128, 64, 1)
        assert np.array_equal(ol.unpack(w, 128, 64), g)


def test_pack_places_bits():
    g = np.zeros((3, 128), np.uint8)
    g[1, 0] = 1
    g[2, 127] = 1
    g[0, 65] = 1
    w = ol.pack(g, 64)
    assert w.tolist() == [0, 2, 1, 0, 0, 1 << 63]
    w32 = ol.pack(g, 32)
    assert w32[4] == 1 and w32[11] == 1 << 31 and w32[2] == 2
    # rows-64 and rows-32 are the same bytes on a little-endian host
    assert w.tobytes() == w32.tobytes()
    assert np.array_equal(ol.unpack(w, 128, 3, 64), g)
    assert np.array_equal(ol.unpack(w32, 128, 3, 32), g)


def expected_word_bit(layout, w, x_size, x, y):
    """The reference's closed-form masks (bit_modes.hpp): FullyPackedRows bit x % W;
    BitColumns bit y % W (X_BITS 1, Y_BITS W); BitTile MSB >> (y % Yb * Xb + x % Xb)
    with Xb x Yb = 4x4 / 8x4 / 8x8; word (wx, wy) at wy * (X / X_BITS) + wx."""
    if layout == ol.LAYOUT_ROWS:
        xb, yb, bit = w, 1, x % w
    elif layout == ol.LAYOUT_COLUMNS:
        xb, yb, bit = 1, w, y % w
    else:
        xb, yb = (4, 4) if w == 16 else (8, 4) if w == 32 else (8, 8)
        bit = w - 1 - ((y % yb) * xb + x % xb)
    return (y // yb) * (x_size // xb) + x // xb, bit


@pytest.mark.parametrize("layout", [ol.LAYOUT_ROWS, ol.LAYOUT_COLUMNS, ol.LAYOUT_TILES])
@pytest.mark.parametrize("w", [16, 32, 64])
def test_pack_layout_places_each_cell_by_the_closed_form_mask(layout, w):
    # the oracle packs with the reference's loop (first mask, move to the next
    # mask per cell in y_bit, x_bit order); the closed form is computed apart
    x_size, y_size = 128, 128
    rng = np.random.default_rng(layout * 100 + w)
    for _ in range(40):
        x, y = int(rng.integers(0, x_size)), int(rng.integers(0, y_size))
        c = np.zeros((y_size, x_size), np.uint8)
        c[y, x] = 1
        words = ol.pack_layout(c, layout, w)
        nz = np.flatnonzero(words)
        idx, bit = expected_word_bit(layout, w, x_size, x, y)
        assert list(nz) == [idx], (layout, w, x, y)
        assert int(words[idx]) == 1 << bit, (layout, w, x, y)
        assert np.array_equal(ol.unpack_layout(words, x_size, y_size, layout, w), c)


@pytest.mark.parametrize("layout,w,x,y", [(ol.LAYOUT_COLUMNS, 16, 96, 48), (ol.LAYOUT_COLUMNS, 64, 40, 128),
                                          (ol.LAYOUT_TILES, 16, 36, 20), (ol.LAYOUT_TILES, 32, 64, 12),
                                          (ol.LAYOUT_TILES, 64, 24, 16), (ol.LAYOUT_ROWS, 32, 96, 7)])
def test_pack_layout_round_trips(layout, w, x, y):
    c = ol.bernoulli_cells(x, y, 5, 0.5)
    assert np.array_equal(ol.unpack_layout(ol.pack_layout(c, layout, w), x, y, layout, w), c)
    if layout == ol.LAYOUT_ROWS:
        assert np.array_equal(ol.pack_layout(c, layout, w), ol.pack(c, w))


def test_pack_layout_rejects_sizes_off_the_word_grid():
    with pytest.raises(ValueError):
        ol.pack_layout(np.zeros((20, 64), np.uint8), ol.LAYOUT_COLUMNS, 32)
    with pytest.raises(ValueError):
        ol.pack_layout(np.zeros((8, 12), np.uint8), ol.LAYOUT_TILES, 64)


def test_pack_rejects_ragged_width():
    with pytest.raises(ValueError):
        ol.pack(np.zeros((2, 96), np.uint8), 64)


def test_naive_fixtures_small():
    # re-run the cheap fixture cases (the 2048^2 x 1000 case is covered by a slow test)
    for f in FIX["naive_runs"]:
        if f["x"] * f["y"] * f["generations"] > 5e7:
            continue
        g = (ol.random_cells(f["x"], f["y"], f["seed"]) if f["generator"] == "random"
             else ol.bernoulli_cells(f["x"], f["y"], f["seed"], f["p"]))
        out = ol.naive_run(g, f["generations"])
        assert ol.sha256(ol.pack(out, f["word_bits"])) == f["sha256"], f
        assert int(out.sum()) == f["popcount"]


def test_workload_fixtures_against_the_naive_loop():
    # The batched / served workload fixtures come from the bit-parallel engine;
    # re-derive the smallest grids of each with the naive loop, from the
    # oracle's loader, and check every grid's parameters against the workload
    # generators.
    import workloads

    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "workload_fixtures.json")))
    for name, specs in (("batched_uneven", workloads.batched_uneven(fx["batched_uneven"]["n"])),
                        ("served_stream", workloads.served_stream(fx["served_stream"]["n"], 1.0))):
        grids = fx[name]["grids"]
        assert [(s.side, s.generations) for s in specs] == [tuple(g[:2]) for g in grids[:len(specs)]]
        for i in sorted(range(len(specs)), key=lambda i: specs[i].cell_updates)[:4]:
            s = specs[i]
            cells = ol.bernoulli_cells(s.side, s.side, s.seed, s.density)
            assert ol.sha256(ol.pack(cells, 64))[:16] == grids[i][2]
            out = ol.naive_run(cells, s.generations)
            assert ol.sha256(ol.pack(out, 64))[:16] == grids[i][3], (name, i)
            assert int(out.sum()) == grids[i][4]
