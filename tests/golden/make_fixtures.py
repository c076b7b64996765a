"""Writes tests/golden/fixtures.json: known answers for the parity tests.

Sources of each entry:
  * "an5d_kats"     — data from the reference's AN5D/GOL_AN5D.c:92-97 (alive
                      counts after 1024 generations of a glibc srand(1985)
                      rand()%2 grid). They reproduce with a wrap-around
                      boundary, so they pin the rule and the neighbour count of
                      the oracle's naive loop run in its periodic mode.
  * "mt19937_10000" — the C++ standard's required 10000th output of a
                      default-seeded std::mt19937 ([rand.predef]).
  * "loader"        — SHA-256 of the seed-42 cells from the oracle's
                      restatement of RandomOnesZerosDataLoader
                      (data_loader.hpp:57-75) as built by this image's g++.
  * "naive_runs"    — final grids of the oracle's naive loop
                      (gol_cpu_naive.hpp:26-117) on seeded inputs.
  * "bitpar_runs"   — final grids at full benchmark sizes from the oracle's
                      bit-parallel engine, which the CPU tests pin against the
                      naive loop.

Run from the repository root:  python tests/golden/make_fixtures.py
(oracle/liboracle.so and stencil-optimalizations_amd/libgolhip.so must be
built; libgolhip.so is used only for its host-side input generators).
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "stencil-optimalizations_amd"))

import oracle_lib as ol  # noqa: E402
import golhip  # noqa: E402

AN5D_KATS = {"256": 3281, "512": 11072, "1024": 45224, "2048": 182485, "4096": 724393, "8192": 2896683}

# (x, y, generator, seed, p, word_bits, generations list)
NAIVE_CASES = [
    (2048, 2048, "random", 42, 0.5, 64, [1000]),
    (256, 256, "random", 42, 0.5, 64, [0, 1, 2, 3, 10, 100, 1000]),
    (64, 64, "random", 1, 0.5, 64, [1, 50, 300]),
    (128, 320, "random", 3, 0.5, 64, [1, 7, 64, 257]),
    (320, 128, "bernoulli", 11, 0.3, 64, [1, 33, 500]),
    (96, 160, "bernoulli", 5, 0.45, 32, [1, 17, 200]),
    (32, 32, "random", 9, 0.5, 32, [1, 5, 100]),
    (1088, 64, "bernoulli", 13, 0.2, 64, [1, 40, 150]),
    (64, 1536, "bernoulli", 17, 0.5, 64, [1, 40, 150]),
]

# spacefillers on a 3 x 3 and a 6 x 5 lattice (generate-final-measurements.py:29-30)
SPACEFILLERS_33 = "; ".join(f"spacefiller[{x}, {y}]" for x in (4096, 8192, 12288) for y in (4096, 8192, 12288))
SPACEFILLERS_66 = "; ".join(f"spacefiller[{x}, {y}]" for x in (2340, 4680, 7020, 9360, 11700, 14040)
                            for y in (2730, 5460, 8190, 10920, 13650))


def always_changing_words(x, y):
    """AlwaysChangingSpaceLoader (data_loader.hpp:117-134): a horizontal blinker
    (three cells of one row) at every 4th column and row, rows-64 words."""
    c = np.zeros((y, x), np.uint8)
    for k in range(3):
        c[0::4, k::4] = 1
    return ol.pack(c, 64)


BITPAR_CASES = [
    # name, x, y, generator, seed, params, generations list
    ("random-16384", 16384, 16384, "random", 42, {}, [1, 100, 1000, 10000]),
    ("patches-16384", 16384, 16384, "patches", 42, {"n_patches": 256, "patch": 64, "p": 0.35}, [100, 20000]),
    ("random-4096x8192", 4096, 8192, "bernoulli", 7, {"p": 0.4}, [500]),
    # the reference's published workload (cluster-experiments/generate-final-measurements.py:33)
    ("spacefiller-16384", 16384, 16384, "lexicon", 0, {"expression": "spacefiller[8192, 8192]"}, [1000, 20000]),
    # the reference's change-tracking workloads at 16384^2 x 10000
    # (cluster-experiments/generate-final-measurements.py:21, 27-30): no work, full work (blinkers every
    # 4 cells), and 9 / 30 spacefillers
    ("zeros-16384", 16384, 16384, "zeros", 0, {}, [10000]),
    ("always-changing-16384", 16384, 16384, "always-changing", 0, {}, [10000]),
    ("spacefiller-33-work-16384", 16384, 16384, "lexicon", 0, {"expression": SPACEFILLERS_33}, [10000]),
    ("spacefiller-66-work-16384", 16384, 16384, "lexicon", 0, {"expression": SPACEFILLERS_66}, [10000]),
]


def input_cells(x, y, gen, seed, p):
    if gen == "random":
        return ol.random_cells(x, y, seed)
    return ol.bernoulli_cells(x, y, seed, p)


def bitpar_case(name, x, y, gen, seed, params, gens_list):
    if gen == "random":
        w = golhip.fill_random(x, y, seed)
    elif gen == "bernoulli":
        w = golhip.fill_bernoulli(x, y, seed, params["p"])
    elif gen == "lexicon":
        w = golhip.fill_lexicon(x, y, params["expression"])
    elif gen == "zeros":
        w = np.zeros(x * y // 64, np.uint64)
    elif gen == "always-changing":
        w = always_changing_words(x, y)
    else:
        w = golhip.fill_patches(x, y, seed, params["n_patches"], params["patch"], params["p"])
    w0 = w
    done = 0
    runs = []
    for gens in sorted(gens_list):
        t = time.time()
        w = ol.bitpar_run(w, x, y, gens - done)
        done = gens
        runs.append({
            "name": name, "x": x, "y": y, "generator": gen, "seed": seed, "params": params,
            "generations": gens, "sha256_input": ol.sha256(w0), "sha256": ol.sha256(w),
            "popcount": int(np.unpackbits(w.view(np.uint8)).sum()),
        })
        print(f"bitpar {name} gens={gens}: {time.time() - t:.1f}s", flush=True)
    return runs


def main() -> None:
    path = os.path.join(HERE, "fixtures.json")
    if len(sys.argv) > 2 and sys.argv[1] == "--only":
        # recompute the named full-size cases and keep every other fixture
        out = json.load(open(path))
        names = set(sys.argv[2:])
        out["bitpar_runs"] = [r for r in out["bitpar_runs"] if r["name"] not in names]
        for case in BITPAR_CASES:
            if case[0] in names:
                out["bitpar_runs"] += bitpar_case(*case)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
            f.write("\n")
        return
    out = {
        "an5d_kats": AN5D_KATS,
        "mt19937_10000": 4123659995,
        "loader": {},
        "naive_runs": [],
        "bitpar_runs": [],
    }
    c = ol.random_cells(2048, 2048, 42)
    out["loader"] = {
        "x": 2048, "y": 2048, "seed": 42,
        "sha256_cells": ol.sha256(c),
        "first_256": "".join(str(v) for v in c.reshape(-1)[:256]),
        "popcount": int(c.sum()),
    }
    for x, y, gen, seed, p, wb, gens_list in NAIVE_CASES:
        cells = input_cells(x, y, gen, seed, p)
        cur, done = cells, 0
        for gens in sorted(gens_list):
            t = time.time()
            cur = ol.naive_run(cur, gens - done)
            done = gens
            words = ol.pack(cur, wb)
            out["naive_runs"].append({
                "x": x, "y": y, "generator": gen, "seed": seed, "p": p, "word_bits": wb,
                "generations": gens, "sha256": ol.sha256(words), "popcount": int(cur.sum()),
            })
            print(f"naive {x}x{y} {gen} seed={seed} gens={gens}: {time.time() - t:.1f}s", flush=True)
    for case in BITPAR_CASES:
        out["bitpar_runs"] += bitpar_case(*case)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
