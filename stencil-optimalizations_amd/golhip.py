"""ctypes binding of the golhip C ABI (include/golhip.h).

This is the Python-side stub a reference maintainer would write to drive the
engine from a script; it mirrors the C calls one to one and raises
``GolError`` carrying the C status code. It loads ``libgolhip.so`` from this
directory and fails loudly when the library has not been built: there is no
CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgolhip.so")

GOL_OK = 0
GOL_ERR_INVALID_ARGUMENT = 1
GOL_ERR_RUNTIME = 2
GOL_ERR_HIP = 3
GOL_ERR_OUT_OF_MEMORY = 4
GOL_ERR_NOT_FOUND = 5
GOL_ERR_BUSY = 6

KERNEL_AUTO = 0
KERNEL_PLAIN = 1
KERNEL_TEMPORAL = 2
KERNEL_ACTIVE = 3

# every symbol include/golhip.h declares
EXPORTED_SYMBOLS = (
    "gol_create", "gol_upload", "gol_upload_device", "gol_step", "gol_step_timed", "gol_download", "gol_sync",
    "gol_device_state", "gol_get_stream", "gol_get_info", "gol_destroy", "gol_step_device_noalloc", "gol_last_error",
    "gol_version", "gol_fill_random", "gol_fill_bernoulli", "gol_fill_patches", "gol_fill_lexicon",
    "gol_batch_create", "gol_batch_add", "gol_batch_run", "gol_batch_reset", "gol_batch_fetch",
    "gol_batch_get_info", "gol_batch_destroy",
    "gol_server_create", "gol_server_submit", "gol_server_poll", "gol_server_wait",
    "gol_server_get_stats", "gol_server_destroy",
    "gol_batch_autotune", "gol_batch_tuning_reset",
)

JOB_QUEUED, JOB_RESIDENT, JOB_DRAINING, JOB_DONE, JOB_FAILED = range(5)

# tile shapes compiled into the library: (dwords per lane, rows per wavefront, wavefronts)
TILE_SHAPES = ((8, 8, 8), (8, 10, 8), (8, 5, 16), (8, 4, 16), (4, 8, 8), (4, 8, 16), (2, 16, 8), (1, 8, 8),
               (1, 16, 8), (1, 24, 8), (1, 32, 8), (1, 32, 16))


class GolError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"golhip error {code}: {message}")
        self.code = code


class GolConfig(ctypes.Structure):
    _fields_ = [
        ("x_cells", ctypes.c_uint64),
        ("y_cells", ctypes.c_uint64),
        ("word_bits", ctypes.c_uint32),
        ("kernel", ctypes.c_uint32),
        ("temporal_steps", ctypes.c_uint32),
        ("lane_dwords", ctypes.c_uint32),
        ("band_rows", ctypes.c_uint32),
        ("waves", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("layout", ctypes.c_uint32),
    ]


FLAG_NO_PERSISTENT = 1
FLAG_SLAB_WAIT_PROBE = 2
FLAG_NATIVE_LAYOUT = 4
# gol_batch_config.flags
BATCH_PERSISTENT = 1
BATCH_WAIT_PROBE = 2
BATCH_SELECT_PLAN = 4

LAYOUT_ROWS, LAYOUT_COLUMNS, LAYOUT_TILES = 0, 1, 2


class GolInfo(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_uint32),
        ("lane_dwords", ctypes.c_uint32),
        ("band_rows", ctypes.c_uint32),
        ("waves", ctypes.c_uint32),
        ("temporal_steps", ctypes.c_uint32),
        ("tiles", ctypes.c_uint32),
        ("persistent", ctypes.c_uint32),
        ("layout", ctypes.c_uint32),
        ("launches", ctypes.c_uint64),
        ("generations", ctypes.c_uint64),
        ("active_tiles_total", ctypes.c_uint64),
        ("tracked_blocks", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("native_layout", ctypes.c_uint32),
        ("workgroups", ctypes.c_uint32),
        ("tile_width_switches", ctypes.c_uint32),
    ]


class GolBatchConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("round_generations", ctypes.c_uint32),
        ("use_graph", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
    ]


class GolBatchInfo(ctypes.Structure):
    _fields_ = [
        ("grids", ctypes.c_uint32),
        ("rounds", ctypes.c_uint32),
        ("items", ctypes.c_uint64),
        ("launches", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("cell_updates", ctypes.c_uint64),
        ("plan_builtin", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("plan_tuned_ms", ctypes.c_double),
        ("plan_builtin_ms", ctypes.c_double),
    ]


class GolTuneConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("round_generations", ctypes.c_uint32),
        ("generations", ctypes.c_uint64),
        ("cells_per_measure", ctypes.c_uint64),
    ]


class GolTuneEntry(ctypes.Structure):
    _fields_ = [
        ("x_cells", ctypes.c_uint64), ("y_cells", ctypes.c_uint64),
        ("lane_dwords", ctypes.c_uint32), ("band_rows", ctypes.c_uint32), ("waves", ctypes.c_uint32),
        ("generations_per_round", ctypes.c_uint32),
        ("candidates", ctypes.c_uint32), ("grids_measured", ctypes.c_uint32),
        ("tuned_cell_updates_per_s", ctypes.c_double),
        ("builtin_cell_updates_per_s", ctypes.c_double),
    ]


class GolServerConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("round_generations", ctypes.c_uint32),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("max_jobs", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class GolJobStats(ctypes.Structure):
    _fields_ = [
        ("queue_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
        ("rounds", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class GolServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64), ("admitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
        ("rounds", ctypes.c_uint64), ("items", ctypes.c_uint64),
        ("cell_updates", ctypes.c_uint64),
        ("resident_bytes", ctypes.c_uint64), ("resident_bytes_peak", ctypes.c_uint64),
        ("resident_jobs", ctypes.c_uint32), ("queued_jobs", ctypes.c_uint32),
        ("pinned_region_bytes", ctypes.c_uint64), ("pinned_region_used", ctypes.c_uint64),
        ("pinned_region_peak", ctypes.c_uint64), ("pinned_own_jobs", ctypes.c_uint64),
    ]


_lib: Optional[ctypes.CDLL] = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libgolhip.so (raises if it is missing: build it first)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.gol_create.argtypes = [ctypes.POINTER(GolConfig), ctypes.POINTER(vp)]
    lib.gol_upload.argtypes = [vp, vp, ctypes.c_size_t]
    lib.gol_upload_device.argtypes = [vp, vp, ctypes.c_size_t]
    lib.gol_step.argtypes = [vp, ctypes.c_uint64]
    lib.gol_step_timed.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    lib.gol_download.argtypes = [vp, vp, ctypes.c_size_t]
    lib.gol_sync.argtypes = [vp]
    lib.gol_device_state.argtypes = [vp, ctypes.POINTER(vp)]
    lib.gol_get_stream.argtypes = [vp, ctypes.POINTER(vp)]
    lib.gol_get_info.argtypes = [vp, ctypes.POINTER(GolInfo)]
    lib.gol_destroy.argtypes = [vp]
    lib.gol_destroy.restype = None
    lib.gol_step_device_noalloc.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, vp]
    lib.gol_step_device_noalloc.restype = ctypes.c_int
    lib.gol_last_error.restype = ctypes.c_char_p
    lib.gol_version.restype = ctypes.c_char_p
    u64 = ctypes.c_uint64
    lib.gol_fill_random.argtypes = [vp, u64, u64, ctypes.c_uint32, u64]
    lib.gol_fill_bernoulli.argtypes = [vp, u64, u64, ctypes.c_uint32, u64, ctypes.c_double]
    lib.gol_fill_lexicon.argtypes = [vp, u64, u64, ctypes.c_uint32, ctypes.c_char_p]
    lib.gol_fill_patches.argtypes = [vp, u64, u64, ctypes.c_uint32, u64, ctypes.c_uint32,
                                     ctypes.c_uint32, ctypes.c_double]
    lib.gol_batch_create.argtypes = [ctypes.POINTER(GolBatchConfig), ctypes.POINTER(vp)]
    lib.gol_batch_add.argtypes = [vp, u64, u64, ctypes.c_uint32, u64, vp, ctypes.POINTER(ctypes.c_uint32)]
    lib.gol_batch_run.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    lib.gol_batch_reset.argtypes = [vp]
    lib.gol_batch_fetch.argtypes = [vp, ctypes.c_uint32, vp, ctypes.c_size_t]
    lib.gol_batch_get_info.argtypes = [vp, ctypes.POINTER(GolBatchInfo)]
    lib.gol_batch_destroy.argtypes = [vp]
    lib.gol_batch_destroy.restype = None
    lib.gol_batch_autotune.argtypes = [ctypes.POINTER(GolTuneConfig), ctypes.POINTER(GolTuneEntry), ctypes.c_uint32]
    lib.gol_batch_autotune.restype = ctypes.c_int
    lib.gol_batch_tuning_reset.argtypes = []
    lib.gol_batch_tuning_reset.restype = ctypes.c_int
    lib.gol_server_create.argtypes = [ctypes.POINTER(GolServerConfig), ctypes.POINTER(vp)]
    lib.gol_server_submit.argtypes = [vp, u64, u64, ctypes.c_uint32, u64, vp, ctypes.POINTER(u64)]
    lib.gol_server_poll.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int)]
    lib.gol_server_wait.argtypes = [vp, u64, vp, ctypes.c_size_t, ctypes.POINTER(GolJobStats)]
    lib.gol_server_get_stats.argtypes = [vp, ctypes.POINTER(GolServerStats)]
    lib.gol_server_destroy.argtypes = [vp]
    lib.gol_server_destroy.restype = None
    for name in ("gol_server_create", "gol_server_submit", "gol_server_poll", "gol_server_wait",
                 "gol_server_get_stats"):
        getattr(lib, name).restype = ctypes.c_int
    for name in ("gol_batch_create", "gol_batch_add", "gol_batch_run", "gol_batch_reset",
                 "gol_batch_fetch", "gol_batch_get_info"):
        getattr(lib, name).restype = ctypes.c_int
    for name in ("gol_create", "gol_upload", "gol_upload_device", "gol_step", "gol_step_timed", "gol_download",
                 "gol_sync", "gol_device_state", "gol_get_stream", "gol_get_info",
                 "gol_fill_random", "gol_fill_bernoulli", "gol_fill_patches", "gol_fill_lexicon"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != GOL_OK:
        msg = load_library().gol_last_error().decode(errors="replace")
        raise GolError(rc, msg)


class Grid:
    """One grid's device state: gol_create / upload / step / download / destroy."""

    def __init__(self, x_cells: int, y_cells: int, word_bits: int = 64, kernel: int = KERNEL_AUTO,
                 temporal_steps: int = 0, lane_dwords: int = 0, band_rows: int = 0, waves: int = 0,
                 device: int = -1, flags: int = 0, layout: int = LAYOUT_ROWS):
        self.lib = load_library()
        cfg = GolConfig(x_cells, y_cells, word_bits, kernel, temporal_steps, lane_dwords, band_rows,
                        waves, device, flags, layout)
        h = ctypes.c_void_p()
        _check(self.lib.gol_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.x_cells, self.y_cells, self.word_bits = x_cells, y_cells, word_bits
        self.dtype = {16: np.uint16, 32: np.uint32, 64: np.uint64}[word_bits]
        self.n_words = x_cells * y_cells // word_bits

    def upload(self, words: np.ndarray) -> None:
        words = np.ascontiguousarray(words, dtype=self.dtype).reshape(-1)
        _check(self.lib.gol_upload(self.handle, words.ctypes.data, words.size))

    def upload_device(self, device_words: int) -> None:
        _check(self.lib.gol_upload_device(self.handle, ctypes.c_void_p(device_words), self.n_words))

    def device_state(self) -> int:
        p = ctypes.c_void_p()
        _check(self.lib.gol_device_state(self.handle, ctypes.byref(p)))
        return p.value or 0

    def step(self, generations: int) -> None:
        _check(self.lib.gol_step(self.handle, generations))

    def step_timed(self, generations: int) -> float:
        ms = ctypes.c_double()
        _check(self.lib.gol_step_timed(self.handle, generations, ctypes.byref(ms)))
        return ms.value

    def download(self) -> np.ndarray:
        out = np.empty(self.n_words, dtype=self.dtype)
        _check(self.lib.gol_download(self.handle, out.ctypes.data, out.size))
        return out

    def sync(self) -> None:
        _check(self.lib.gol_sync(self.handle))

    def info(self) -> GolInfo:
        info = GolInfo()
        _check(self.lib.gol_get_info(self.handle, ctypes.byref(info)))
        return info

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(self.lib.gol_get_stream(self.handle, ctypes.byref(s)))
        return s.value or 0

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.gol_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many grids advanced together (gol_batch_*)."""

    def __init__(self, device: int = -1, round_generations: int = 0, use_graph: bool = True,
                 persistent: bool = False, flags: int = 0):
        self.lib = load_library()
        flags |= BATCH_PERSISTENT if persistent else 0
        cfg = GolBatchConfig(device, round_generations, 1 if use_graph else 0, flags)
        h = ctypes.c_void_p()
        _check(self.lib.gol_batch_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.shapes = []

    def add(self, words: np.ndarray, x_cells: int, y_cells: int, generations: int, word_bits: int = 64) -> int:
        words = np.ascontiguousarray(words, dtype=np.uint64 if word_bits == 64 else np.uint32).reshape(-1)
        gid = ctypes.c_uint32()
        _check(self.lib.gol_batch_add(self.handle, x_cells, y_cells, word_bits, generations,
                                      words.ctypes.data, ctypes.byref(gid)))
        self.shapes.append((x_cells, y_cells, word_bits))
        return gid.value

    def run(self) -> float:
        ms = ctypes.c_double()
        _check(self.lib.gol_batch_run(self.handle, ctypes.byref(ms)))
        return ms.value

    def reset(self) -> None:
        """Restore every grid's initial state (gol_batch_reset)."""
        _check(self.lib.gol_batch_reset(self.handle))

    def fetch(self, gid: int) -> np.ndarray:
        x, y, wb = self.shapes[gid]
        out = np.empty((x // wb) * y, dtype=np.uint64 if wb == 64 else np.uint32)
        _check(self.lib.gol_batch_fetch(self.handle, gid, out.ctypes.data, out.size))
        return out

    def info(self) -> GolBatchInfo:
        info = GolBatchInfo()
        _check(self.lib.gol_batch_get_info(self.handle, ctypes.byref(info)))
        return info

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.gol_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_autotune(sizes, device: int = -1, round_generations: int = 0, generations: int = 0,
                   cells_per_measure: int = 0) -> list:
    """gol_batch_autotune over representative (x_cells, y_cells) sizes; returns one
    dict per size with the chosen shape and the measured rates."""
    lib = load_library()
    sizes = list(sizes)
    entries = (GolTuneEntry * max(1, len(sizes)))()
    for e, (x, y) in zip(entries, sizes):
        e.x_cells, e.y_cells = x, y
    cfg = GolTuneConfig(device, round_generations, generations, cells_per_measure)
    _check(lib.gol_batch_autotune(ctypes.byref(cfg), entries, len(sizes)))
    return [dict(x_cells=e.x_cells, y_cells=e.y_cells, lane_dwords=e.lane_dwords, band_rows=e.band_rows,
                 waves=e.waves, generations_per_round=e.generations_per_round, candidates=e.candidates,
                 grids_measured=e.grids_measured, tuned_cell_updates_per_s=e.tuned_cell_updates_per_s,
                 builtin_cell_updates_per_s=e.builtin_cell_updates_per_s) for e in entries[:len(sizes)]]


def batch_tuning_reset() -> None:
    """gol_batch_tuning_reset: back to the built-in table and cost model."""
    _check(load_library().gol_batch_tuning_reset())


class Server:
    """Continuous-batching job server (gol_server_*). submit/poll/wait are thread-safe."""

    def __init__(self, device: int = -1, round_generations: int = 0, hbm_budget_bytes: int = 0,
                 max_jobs: int = 0):
        self.lib = load_library()
        cfg = GolServerConfig(device, round_generations, hbm_budget_bytes, max_jobs, 0)
        h = ctypes.c_void_p()
        _check(self.lib.gol_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._shapes = {}

    def submit(self, words: np.ndarray, x_cells: int, y_cells: int, generations: int,
               word_bits: int = 64) -> int:
        words = np.ascontiguousarray(words, dtype=np.uint64 if word_bits == 64 else np.uint32).reshape(-1)
        jid = ctypes.c_uint64()
        _check(self.lib.gol_server_submit(self.handle, x_cells, y_cells, word_bits, generations,
                                          words.ctypes.data, ctypes.byref(jid)))
        self._shapes[jid.value] = (x_cells, y_cells, word_bits)
        return jid.value

    def poll(self, jid: int) -> int:
        st = ctypes.c_int()
        _check(self.lib.gol_server_poll(self.handle, jid, ctypes.byref(st)))
        return st.value

    def wait(self, jid: int):
        """Returns (final words, GolJobStats)."""
        x, y, wb = self._shapes.pop(jid)
        out = np.empty((x // wb) * y, dtype=np.uint64 if wb == 64 else np.uint32)
        st = GolJobStats()
        _check(self.lib.gol_server_wait(self.handle, jid, out.ctypes.data, out.size, ctypes.byref(st)))
        return out, st

    def stats(self) -> GolServerStats:
        st = GolServerStats()
        _check(self.lib.gol_server_get_stats(self.handle, ctypes.byref(st)))
        return st

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.gol_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_device_noalloc(device_in: int, device_out: int, x_cells: int, y_cells: int, generations: int,
                        stream: int = 0) -> None:
    """gol_step_device_noalloc: advance the rows-layout grid at device address
    device_in by `generations` into device_out with no device allocation (the
    persistent slab launch, its scratch in a static device array); waits for
    the stream."""
    _check(load_library().gol_step_device_noalloc(device_in, device_out, x_cells, y_cells, generations, stream))


def _words(x_cells: int, y_cells: int, word_bits: int) -> np.ndarray:
    return np.zeros((x_cells // word_bits) * y_cells, dtype=np.uint64 if word_bits == 64 else np.uint32)


def fill_random(x_cells: int, y_cells: int, seed: int = 42, word_bits: int = 64) -> np.ndarray:
    """RandomOnesZerosDataLoader (data_loader.hpp:57-75), packed into row words."""
    w = _words(x_cells, y_cells, word_bits)
    _check(load_library().gol_fill_random(w.ctypes.data, x_cells, y_cells, word_bits, seed))
    return w


def fill_bernoulli(x_cells: int, y_cells: int, seed: int, p: float, word_bits: int = 64) -> np.ndarray:
    w = _words(x_cells, y_cells, word_bits)
    _check(load_library().gol_fill_bernoulli(w.ctypes.data, x_cells, y_cells, word_bits, seed, p))
    return w


def fill_patches(x_cells: int, y_cells: int, seed: int, n_patches: int, patch: int, p: float,
                 word_bits: int = 64) -> np.ndarray:
    w = _words(x_cells, y_cells, word_bits)
    _check(load_library().gol_fill_patches(w.ctypes.data, x_cells, y_cells, word_bits, seed,
                                           n_patches, patch, p))
    return w


def fill_lexicon(x_cells: int, y_cells: int, expression: str, word_bits: int = 64) -> np.ndarray:
    """gol_fill_lexicon: "name[x, y]; ..." Life Lexicon patterns on an empty grid."""
    w = _words(x_cells, y_cells, word_bits)
    _check(load_library().gol_fill_lexicon(w.ctypes.data, x_cells, y_cells, word_bits, expression.encode()))
    return w


def pack_rows(cells: np.ndarray, word_bits: int = 64) -> np.ndarray:
    """Pack a (Y, X) 0/1 array into row words, bit k of word wx = cell word_bits*wx + k
    (src/algorithms/_shared/bitwise/general_bit_grid.hpp:217-246, FullyPackedRows)."""
    y, x = cells.shape
    if x % word_bits:
        raise ValueError(f"Grid dimensions X must be a multiple of {word_bits}")
    bits = np.packbits((cells != 0).astype(np.uint8), axis=1, bitorder="little")
    return bits.view(np.uint64 if word_bits == 64 else np.uint32).reshape(y, x // word_bits).copy()


def unpack_rows(words: np.ndarray, x_cells: int, y_cells: int) -> np.ndarray:
    """Inverse of pack_rows: (Y, X) uint8 array of 0/1 cells."""
    w = np.ascontiguousarray(words).reshape(y_cells, -1)
    return np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")[:, :x_cells]
