"""gol_step_device_noalloc (no device allocation, scratch in a static device
array) against gol_step on a handle (the persistent slab form), through the C
ABI: two calls in a row, the second on a smaller grid, then the oracle and the
argument limits."""

import hashlib
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


def via_handle(words, x, y, gens):
    with golhip.Grid(x, y, kernel=golhip.KERNEL_TEMPORAL, device=0) as g:
        g.upload(words)
        g.step(gens)
        return g.download()


def via_noalloc(words, x, y, gens):
    # caller-owned device buffers: two plain handles' states
    with golhip.Grid(x, y, kernel=golhip.KERNEL_PLAIN, device=0) as src, \
            golhip.Grid(x, y, kernel=golhip.KERNEL_PLAIN, device=0) as dst:
        src.upload(words)
        dst.upload(np.full_like(words, 0x5A5A5A5A5A5A5A5A))
        golhip.step_device_noalloc(src.device_state(), dst.device_state(), x, y, gens)
        assert np.array_equal(src.download(), words), "the input buffer changed"
        return dst.download()


def test_noalloc_matches_gol_step_two_calls_second_smaller():
    # 16384^2 x 1000 (random, seed 42): 256 slabs at 8 generations per block,
    # then a 4096 x 2048 grid with a partial last block (301 = 37 x 8 + 5)
    cases = [(16384, 16384, 1000, golhip.fill_random(16384, 16384, 42)),
             (4096, 2048, 301, golhip.fill_bernoulli(4096, 2048, 7, 0.4))]
    for x, y, gens, words in cases:
        note("noalloc", x, y, gens)
        got = via_noalloc(words, x, y, gens)
        assert np.array_equal(got, via_handle(words, x, y, gens)), (x, y, gens)
        if (x, y) == (16384, 16384):
            want = next(r for r in FIX["bitpar_runs"] if r["name"] == "random-16384" and r["generations"] == gens)
            assert hashlib.sha256(got.tobytes()).hexdigest() == want["sha256"]
        else:
            assert np.array_equal(got, ol.bitpar_run(words, x, y, gens))


def test_noalloc_zero_generations_and_ragged_rows():
    # 0 generations copies; a 1088-cell row (34 dwords, columns masked) and a
    # height that is no multiple of the slab
    x, y = 1088, 1000
    words = golhip.fill_bernoulli(x, y, 3, 0.35)
    note("noalloc 0 generations", x, y)
    assert np.array_equal(via_noalloc(words, x, y, 0), words)
    for gens in (1, 8, 77):
        note("noalloc", x, y, gens)
        assert np.array_equal(via_noalloc(words, x, y, gens), ol.bitpar_run(words, x, y, gens)), gens


def test_noalloc_limits():
    with golhip.Grid(32768, 64, kernel=golhip.KERNEL_PLAIN, device=0) as a, \
            golhip.Grid(32768, 64, kernel=golhip.KERNEL_PLAIN, device=0) as b:
        note("noalloc limits")
        with pytest.raises(golhip.GolError, match="at most 16384 cells"):
            golhip.step_device_noalloc(a.device_state(), b.device_state(), 32768, 64, 10)
        # a grid one tile spans is not a slab launch
        with pytest.raises(golhip.GolError, match="not one the persistent slab launch takes"):
            golhip.step_device_noalloc(a.device_state(), b.device_state(), 512, 64, 10)
        # more generations than 65536 blocks in one call
        with pytest.raises(golhip.GolError, match="generations per call"):
            golhip.step_device_noalloc(a.device_state(), b.device_state(), 2048, 256, 65536 * 8 + 1)
