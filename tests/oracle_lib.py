"""ctypes access to the CPU oracle (oracle/liboracle.so) — test infrastructure.

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker / baseline. The library is built by
``make -C oracle`` (__graft_entry__.build() runs it).
"""

from __future__ import annotations

import ctypes
import hashlib
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run make -C oracle")
        L = ctypes.CDLL(ORACLE_SO)
        vp, u64 = ctypes.c_void_p, ctypes.c_uint64
        L.oracle_load_random.argtypes = [vp, u64, u64, u64]
        L.oracle_mt19937_nth.argtypes = [u64, u64]
        L.oracle_mt19937_nth.restype = ctypes.c_uint32
        L.oracle_mt19937_64_nth.argtypes = [u64, u64]
        L.oracle_mt19937_64_nth.restype = ctypes.c_uint64
        L.oracle_load_bernoulli.argtypes = [vp, u64, u64, ctypes.c_double]
        L.oracle_load_glibc_rand.argtypes = [vp, u64, ctypes.c_uint]
        L.oracle_naive_run.argtypes = [vp, u64, u64, u64]
        L.oracle_naive_run_boundary.argtypes = [vp, u64, u64, u64, ctypes.c_int]
        L.oracle_pack_rows.argtypes = [vp, u64, u64, ctypes.c_uint32, vp]
        L.oracle_unpack_rows.argtypes = [vp, u64, u64, ctypes.c_uint32, vp]
        L.oracle_pack_layout.argtypes = [vp, u64, u64, ctypes.c_uint32, ctypes.c_uint32, vp]
        L.oracle_unpack_layout.argtypes = [vp, u64, u64, ctypes.c_uint32, ctypes.c_uint32, vp]
        L.oracle_bitpar_run.argtypes = [vp, u64, u64, u64, ctypes.c_int]
        _lib = L
    return _lib


def random_cells(x: int, y: int, seed: int = 42) -> np.ndarray:
    """RandomOnesZerosDataLoader (data_loader.hpp:57-75): (Y, X) uint8."""
    c = np.empty(x * y, np.uint8)
    lib().oracle_load_random(c.ctypes.data, x, y, seed)
    return c.reshape(y, x)


def bernoulli_cells(x: int, y: int, seed: int, p: float) -> np.ndarray:
    c = np.empty(x * y, np.uint8)
    lib().oracle_load_bernoulli(c.ctypes.data, x * y, seed, p)
    return c.reshape(y, x)


def glibc_rand_cells(n: int, seed: int = 1985) -> np.ndarray:
    c = np.empty(n * n, np.uint8)
    lib().oracle_load_glibc_rand(c.ctypes.data, n * n, seed)
    return c.reshape(n, n)


def naive_run(cells: np.ndarray, generations: int, periodic: bool = False) -> np.ndarray:
    c = np.ascontiguousarray(cells, dtype=np.uint8).copy()
    y, x = c.shape
    lib().oracle_naive_run_boundary(c.ctypes.data, x, y, generations, 1 if periodic else 0)
    return c


def pack(cells: np.ndarray, word_bits: int = 64) -> np.ndarray:
    y, x = cells.shape
    dt = np.uint64 if word_bits == 64 else np.uint32
    w = np.empty((x // word_bits) * y, dt)
    rc = lib().oracle_pack_rows(np.ascontiguousarray(cells, dtype=np.uint8).ctypes.data, x, y, word_bits,
                                w.ctypes.data)
    if rc:
        raise ValueError("X must be a multiple of word_bits")
    return w


def unpack(words: np.ndarray, x: int, y: int, word_bits: int = 64) -> np.ndarray:
    c = np.empty(x * y, np.uint8)
    rc = lib().oracle_unpack_rows(np.ascontiguousarray(words).ctypes.data, x, y, word_bits, c.ctypes.data)
    if rc:
        raise ValueError("X must be a multiple of word_bits")
    return c.reshape(y, x)


LAYOUT_ROWS, LAYOUT_COLUMNS, LAYOUT_TILES = 0, 1, 2
_WORD_DTYPE = {16: np.uint16, 32: np.uint32, 64: np.uint64}


def pack_layout(cells: np.ndarray, layout: int, word_bits: int) -> np.ndarray:
    """The reference's GeneralBitGrid packing under the rows / columns / tiles
    policy (general_bit_grid.hpp:217-246, bit_modes.hpp:10-163)."""
    y, x = cells.shape
    w = np.empty(x * y // word_bits, _WORD_DTYPE[word_bits])
    rc = lib().oracle_pack_layout(np.ascontiguousarray(cells, dtype=np.uint8).ctypes.data, x, y, layout, word_bits,
                                  w.ctypes.data)
    if rc:
        raise ValueError("grid size does not fit the layout's words")
    return w


def unpack_layout(words: np.ndarray, x: int, y: int, layout: int, word_bits: int) -> np.ndarray:
    c = np.empty(x * y, np.uint8)
    rc = lib().oracle_unpack_layout(np.ascontiguousarray(words).ctypes.data, x, y, layout, word_bits, c.ctypes.data)
    if rc:
        raise ValueError("grid size does not fit the layout's words")
    return c.reshape(y, x)


def bitpar_run(words64: np.ndarray, x: int, y: int, generations: int, threads: int = 0) -> np.ndarray:
    w = np.ascontiguousarray(words64, dtype=np.uint64).copy()
    lib().oracle_bitpar_run(w.ctypes.data, x, y, generations, threads)
    return w


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
