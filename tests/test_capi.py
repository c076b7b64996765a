"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, validates configurations before touching a device, and its
host-side input generators equal the oracle's loaders."""

import ctypes
import os
import re

import numpy as np
import pytest

import golhip
import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "golhip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gol_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = golhip.load_library()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(golhip.EXPORTED_SYMBOLS) == names


def test_version_string():
    assert b"gfx950" in golhip.load_library().gol_version()


@pytest.mark.parametrize("kw,code", [
    (dict(x_cells=100, y_cells=64, word_bits=64), golhip.GOL_ERR_INVALID_ARGUMENT),
    (dict(x_cells=64, y_cells=64, word_bits=16), golhip.GOL_ERR_INVALID_ARGUMENT),
    (dict(x_cells=0, y_cells=64, word_bits=64), golhip.GOL_ERR_INVALID_ARGUMENT),
    (dict(x_cells=64, y_cells=64, word_bits=64, kernel=9), golhip.GOL_ERR_INVALID_ARGUMENT),
    (dict(x_cells=64, y_cells=64, word_bits=64, lane_dwords=3), golhip.GOL_ERR_INVALID_ARGUMENT),
])
def test_config_errors_before_device(kw, code):
    # invalid configurations fail in validation, before any HIP call
    with pytest.raises(golhip.GolError) as e:
        golhip.Grid(**kw)
    assert e.value.code == code


@pytest.mark.parametrize("size,rounds", [((100, 64), 0), ((0, 64), 0), ((2048, 2048), 33)])
def test_autotune_errors_before_device(size, rounds):
    # invalid sizes and round depths fail in validation, before any HIP call
    with pytest.raises(golhip.GolError) as e:
        golhip.batch_autotune([size], round_generations=rounds)
    assert e.value.code == golhip.GOL_ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("x,y,w,layout,multiple", [
    (128, 40, 64, golhip.LAYOUT_COLUMNS, "Y must be a multiple of 64"),
    (128, 12, 64, golhip.LAYOUT_TILES, "Y must be a multiple of 8"),
    (128, 6, 32, golhip.LAYOUT_TILES, "Y must be a multiple of 4"),
    (48, 64, 16, golhip.LAYOUT_COLUMNS, "X must be a multiple of 32"),
    (64, 64, 16, golhip.LAYOUT_ROWS, "word_bits must be 32 or 64"),
    (64, 64, 8, golhip.LAYOUT_TILES, "word_bits must be 16, 32 or 64"),
])
def test_layout_size_errors_before_device(x, y, w, layout, multiple):
    with pytest.raises(golhip.GolError) as e:
        golhip.Grid(x, y, word_bits=w, layout=layout)
    assert e.value.code == golhip.GOL_ERR_INVALID_ARGUMENT
    assert multiple in str(e.value)


@pytest.mark.parametrize("kw", [
    dict(x_cells=64, y_cells=64, word_bits=64),                                     # rows
    dict(x_cells=64, y_cells=64, word_bits=32, layout=golhip.LAYOUT_TILES),         # 8x4 tiles
    dict(x_cells=64, y_cells=64, word_bits=16, layout=golhip.LAYOUT_COLUMNS),       # 16-bit columns
    dict(x_cells=64, y_cells=64, word_bits=64, layout=golhip.LAYOUT_TILES, kernel=golhip.KERNEL_ACTIVE),
])
def test_native_layout_flag_errors_before_device(kw):
    with pytest.raises(golhip.GolError) as e:
        golhip.Grid(flags=golhip.FLAG_NATIVE_LAYOUT, **kw)
    assert e.value.code == golhip.GOL_ERR_INVALID_ARGUMENT
    assert "GOL_FLAG_NATIVE_LAYOUT" in str(e.value)


def test_error_message_names_the_multiple():
    with pytest.raises(golhip.GolError) as e:
        golhip.Grid(96, 8, word_bits=64)
    assert "multiple of 64" in str(e.value)


@pytest.mark.parametrize("x,y,seed", [(64, 3, 42), (2048, 16, 42), (96, 5, 7)])
def test_fill_random_matches_loader(x, y, seed):
    wb = 64 if x % 64 == 0 else 32
    got = golhip.fill_random(x, y, seed, wb)
    assert np.array_equal(got, ol.pack(ol.random_cells(x, y, seed), wb))


def test_fill_random_matches_loader_fixture():
    f = ol.random_cells(2048, 2048, 42)
    assert ol.sha256(ol.unpack(golhip.fill_random(2048, 2048, 42), 2048, 2048)) == ol.sha256(f)


@pytest.mark.parametrize("p", [0.1, 0.35, 0.5])
def test_fill_bernoulli_matches_oracle(p):
    got = golhip.fill_bernoulli(128, 40, 99, p)
    assert np.array_equal(got, ol.pack(ol.bernoulli_cells(128, 40, 99, p)))


def test_fill_patches_places_patches():
    w = golhip.fill_patches(512, 256, 5, 3, 64, 0.35)
    cells = ol.unpack(w, 512, 256)
    assert 0 < cells.sum() <= 3 * 64 * 64
    # deterministic for a seed, different across seeds
    assert np.array_equal(golhip.fill_patches(512, 256, 5, 3, 64, 0.35), w)
    assert not np.array_equal(golhip.fill_patches(512, 256, 6, 3, 64, 0.35), w)
    # density inside a lone patch is close to p
    one = ol.unpack(golhip.fill_patches(512, 256, 8, 1, 64, 0.35), 512, 256)
    assert 0.25 < one.sum() / (64 * 64) < 0.45


def test_fill_lexicon_spacefiller_matches_the_reference_table():
    # the 27 x 27 spacefiller of the reference's pattern table
    # (src/infrastructure/gol-lexicon/patterns.cpp:10325-10352): 187 live cells;
    # the hash is of np.packbits over that matrix, row by row
    c = ol.unpack(golhip.fill_lexicon(64, 64, "spacefiller[10, 20]"), 64, 64)
    assert c.sum() == 187
    ys, xs = np.nonzero(c)
    assert (ys.min(), ys.max(), xs.min(), xs.max()) == (20, 46, 10, 36)
    assert ol.sha256(np.packbits(c[20:47, 10:37])) == \
        "85ae4a83ebbfa90bfd769602877f4972560597184999e53e472f134ef100de8f"


def test_fill_lexicon_places_rectangles_and_clips():
    c = ol.unpack(golhip.fill_lexicon(64, 64, "glider[60, 62]; block[0, 0]", word_bits=32), 64, 64, 32)
    assert c.sum() == 8  # the block and the two rows of the glider inside the grid
    assert c[0:2, 0:2].all() and c[62, 60:63].all() and c[63, 60] == 1
    # a later pattern's dead cells overwrite earlier live ones
    c = ol.unpack(golhip.fill_lexicon(64, 64, "block[0, 0]; beehive[0, 0]"), 64, 64)
    assert c[0, 0] == 0 and c[0, 1] == 1 and c[1, 0] == 1
    assert c.sum() == 6
    with pytest.raises(golhip.GolError):
        golhip.fill_lexicon(64, 64, "glider[1, 2")


def test_python_pack_matches_oracle():
    c = ol.bernoulli_cells(192, 7, 3, 0.5)
    assert np.array_equal(golhip.pack_rows(c, 64).reshape(-1), ol.pack(c, 64))
    assert np.array_equal(golhip.pack_rows(c, 32).reshape(-1), ol.pack(c, 32))
    assert np.array_equal(golhip.unpack_rows(ol.pack(c, 64), 192, 7), c)


@pytest.mark.parametrize("args,message", [
    ((0, 1 << 20, 16384, 64, 10), "null device buffer"),
    ((1 << 20, 2 << 20, 100, 64, 10), "multiple of 32"),
    ((1 << 20, 2 << 20, 32768, 64, 10), "at most 16384 cells"),
    ((1 << 20, (1 << 20) + 4096, 16384, 64, 10), "overlap"),
])
def test_noalloc_errors_before_device(args, message):
    # argument errors of gol_step_device_noalloc come before any HIP call (the
    # addresses are never touched)
    with pytest.raises(golhip.GolError) as e:
        golhip.step_device_noalloc(*args)
    assert e.value.code == golhip.GOL_ERR_INVALID_ARGUMENT
    assert message in str(e.value)


# The ctypes mirrors of the header's structs: every field at the offset the C
# compiler gives it, every struct the same size (a field added to one side
# only, or in another order, fails here).
ABI_STRUCTS = [
    ("gol_config", golhip.GolConfig), ("gol_info", golhip.GolInfo),
    ("gol_batch_config", golhip.GolBatchConfig), ("gol_batch_info", golhip.GolBatchInfo),
    ("gol_tune_config", golhip.GolTuneConfig), ("gol_tune_entry", golhip.GolTuneEntry),
    ("gol_server_config", golhip.GolServerConfig), ("gol_job_stats", golhip.GolJobStats),
    ("gol_server_stats", golhip.GolServerStats),
]


def test_ctypes_structs_match_the_header(tmp_path):
    import subprocess

    lines = ["#include <stdio.h>", "#include \"golhip.h\"", "int main(void) {"]
    for cname, cls in ABI_STRUCTS:
        lines.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        for field in cls._fields_:
            lines.append(f'    printf("{cname}.{field[0]} %zu\\n", offsetof({cname}, {field[0]}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)],
                   check=True)
    c_side = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                                  text=True).stdout.splitlines())
    for cname, cls in ABI_STRUCTS:
        assert int(c_side[cname]) == ctypes.sizeof(cls), cname
        for field in cls._fields_:
            assert int(c_side[f"{cname}.{field[0]}"]) == getattr(cls, field[0]).offset, f"{cname}.{field[0]}"
    # and the mirrors name the header's fields, in its order
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for cname, cls in ABI_STRUCTS:
        body = re.search(r"typedef struct " + cname + r" \{(.*?)\} " + cname + ";", text, flags=re.S).group(1)
        names = [re.sub(r"\[.*?\]", "", part).split()[-1]
                 for decl in body.split(";") if decl.strip() for part in decl.split(",")]
        assert names == [f[0] for f in cls._fields_], cname
