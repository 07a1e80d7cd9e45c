"""Command-line front end (src/PGURE-SVT.cpp), TIFF I/O and the hot-pixel
filter (src/hotpixel.hpp): the parts that need no GPU."""

import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pgure-svt_amd", "pgure-svt")


def _lib():
    from pguresvt import _pguresvt as P

    L = P.load()
    L.pgsvt_tiff_info.argtypes = [C.c_char_p] + [C.POINTER(C.c_uint32)] * 4
    L.pgsvt_tiff_read_u16.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.c_void_p]
    L.pgsvt_tiff_write_u16.argtypes = [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
    return L, P


def tiff_write(path, cube):
    L, P = _lib()
    cube = np.asfortranarray(cube.astype(np.uint16))
    H, W, n = cube.shape
    P.check(L.pgsvt_tiff_write_u16(path.encode(), cube.ctypes.data, W, H, n))


def tiff_read(path, first=0, count=None):
    L, P = _lib()
    w, h, b, n = (C.c_uint32() for _ in range(4))
    P.check(L.pgsvt_tiff_info(path.encode(), C.byref(w), C.byref(h), C.byref(b), C.byref(n)))
    count = n.value - first if count is None else count
    out = np.zeros((h.value, w.value, count), dtype=np.uint16, order="F")
    P.check(L.pgsvt_tiff_read_u16(path.encode(), first, count, out.ctypes.data))
    return out, b.value


def parse_tiff_pages(path):
    """Independent baseline-TIFF reader: (bits, pages as (H, W) arrays)."""
    data = open(path, "rb").read()
    e = "<" if data[:2] == b"II" else ">"
    assert struct.unpack(e + "H", data[2:4])[0] == 42
    ifd = struct.unpack(e + "I", data[4:8])[0]
    pages, bits = [], None
    while ifd:
        n = struct.unpack(e + "H", data[ifd:ifd + 2])[0]
        tags = {}
        for i in range(n):
            tag, typ, cnt, val = struct.unpack(e + "HHII", data[ifd + 2 + 12 * i:ifd + 14 + 12 * i])
            if typ == 3 and cnt == 1:
                val = struct.unpack(e + "H", data[ifd + 10 + 12 * i:ifd + 12 + 12 * i])[0]
            tags[tag] = val
        W, H, bits = tags[256], tags[257], tags[258]
        dt = np.dtype(e + ("u2" if bits == 16 else "u1"))
        pages.append(np.frombuffer(data, dtype=dt, count=W * H, offset=tags[273]).reshape(H, W))
        ifd = struct.unpack(e + "I", data[ifd + 2 + 12 * n:ifd + 6 + 12 * n])[0]
    return bits, pages


def test_tiff_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    cube = rng.integers(0, 65536, size=(12, 20, 5), dtype=np.uint16)  # H=12 rows, W=20 columns
    f = str(tmp_path / "seq.tif")
    tiff_write(f, cube)
    bits, pages = parse_tiff_pages(f)
    assert bits == 16 and len(pages) == 5
    for k in range(5):  # page k, row r, column c = cube[r, c, k]
        assert np.array_equal(pages[k], cube[:, :, k])
    back, b = tiff_read(f)
    assert b == 16 and np.array_equal(back, cube)
    sub, _ = tiff_read(f, 2, 2)
    assert np.array_equal(sub, cube[:, :, 2:4])


def _write_baseline(path, pages, bits, endian, override=None):
    """Hand-built uncompressed TIFF with one strip per page; override maps a
    tag to the (type, count, value) written in its place."""
    e = endian
    H, W = pages[0].shape
    dt = np.dtype(e + ("u2" if bits == 16 else "u1"))
    out = bytearray(b"II" if e == "<" else b"MM") + struct.pack(e + "HI", 42, 0)
    link = 4
    for p in pages:
        off = len(out)
        out += p.astype(dt).tobytes()
        if len(out) % 2:
            out += b"\0"
        ifd = len(out)
        out[link:link + 4] = struct.pack(e + "I", ifd)
        ents = [(256, 4, 1, W), (257, 4, 1, H), (258, 3, 1, bits), (259, 3, 1, 1), (262, 3, 1, 1),
                (273, 4, 1, off), (277, 3, 1, 1), (278, 4, 1, H), (279, 4, 1, W * H * bits // 8)]
        ents = [(t, *override[t]) if override and t in override else (t, ty, c, v) for t, ty, c, v in ents]
        out += struct.pack(e + "H", len(ents))
        for tag, typ, cnt, val in ents:
            v = struct.pack(e + "HH", val, 0) if typ == 3 else struct.pack(e + "I", val)
            out += struct.pack(e + "HHI", tag, typ, cnt) + v
        link = len(out)
        out += struct.pack(e + "I", 0)
    open(path, "wb").write(bytes(out))


@pytest.mark.parametrize("bits,endian", [(8, "<"), (16, ">"), (8, ">")])
def test_tiff_reads_baseline_files(tmp_path, bits, endian):
    rng = np.random.default_rng(bits)
    pages = [rng.integers(0, 2 ** bits, size=(9, 9)) for _ in range(3)]
    f = str(tmp_path / "in.tif")
    _write_baseline(f, pages, bits, endian)
    cube, b = tiff_read(f)
    assert b == bits
    for k in range(3):
        assert np.array_equal(cube[:, :, k], pages[k])


@pytest.mark.parametrize("override,msg", [
    ({278: (4, 1, 0)}, b"RowsPerStrip is 0"),                   # row / rows_per_strip would divide by 0
    ({256: (4, 0, 9)}, b"empty field"),                         # a width with no value
    ({273: (4, 0x3fffffff, 8)}, b"field larger than the file"),  # count far beyond the file: no allocation
    ({279: (4, 40, 1 << 30)}, b"field outside the file"),       # values stored past the end
])
def test_tiff_rejects_malformed_fields(tmp_path, override, msg):
    # malformed IFD entries are errors (PGSVT_E_ARG with a message), never a
    # crash, an out-of-range read or a huge allocation
    L, P = _lib()
    f = str(tmp_path / "bad.tif")
    _write_baseline(f, [np.zeros((9, 9))], 16, "<", override)
    out = np.zeros(81, np.uint16)
    w, h, b, n = (C.c_uint32() for _ in range(4))
    rc = L.pgsvt_tiff_info(f.encode(), C.byref(w), C.byref(h), C.byref(b), C.byref(n))
    if rc == 0:
        rc = L.pgsvt_tiff_read_u16(f.encode(), 0, 1, out.ctypes.data)
    assert rc == -1 and msg in L.pgsvt_last_error()


def test_tiff_rejects_bad_files(tmp_path):
    L, P = _lib()
    f = tmp_path / "bad.tif"
    f.write_bytes(b"not a tiff at all")
    w, h, b, n = (C.c_uint32() for _ in range(4))
    assert L.pgsvt_tiff_info(str(f).encode(), C.byref(w), C.byref(h), C.byref(b), C.byref(n)) == -1
    assert b"byte-order" in L.pgsvt_last_error()


# ----------------------------------------------------------- hot pixels --

def test_hotpixel_integer_median_and_border():
    # column medians 10 and 13: Armadillo's even-count median of integers is
    # upper + (lower - upper) / 2 in int arithmetic = 13 + (-3)/2 = 12, and a
    # hot border pixel takes the frame median (src/hotpixel.hpp:54-57)
    F = np.zeros((4, 2, 1), dtype=np.uint16)
    F[:, 0, 0] = 10
    F[:, 1, 0] = 13
    F[0, 0, 0] = 1000
    out = O.hotpixel(F, 3.0)
    expect = F.copy()
    expect[0, 0, 0] = 12
    assert np.array_equal(out, expect)


def test_hotpixel_in_place_order():
    # A 3 x 3 hot block (H) on a flat background (b), threshold 0.  Pixels
    # are replaced in column-major order, in place, each from the mean of
    # the 4th and 5th of its 8 sorted neighbours as they are at that moment
    # (src/hotpixel.hpp:31-53).  Worked by hand, (row, col):
    #   (5,5): 5b 3H -> b            (6,5): b b b [b] H H H H -> (b+H)/2 = 550
    #   (7,5): ... 550 H H -> b      (5,6): b b b b 550 H H H -> (b+550)/2 = 325
    #   (6,6): b b 325 550 H H H H -> (550+H)/2 = 775
    #   (7,6): b b b b 550 775 H H -> 325
    #   (5,7): -> b   (6,7): b b b b 325 325 775 H -> 212 (truncated)   (7,7): -> b
    # Replacing from the original values instead would give H at (6,6).
    b, H = 100, 1000
    F = np.full((12, 12, 1), b, dtype=np.uint16)
    F[5:8, 5:8, 0] = H
    F[0, 0, 0] = b + 1  # a border outlier takes the frame median b
    out = O.hotpixel(F, 0.0)
    expect = np.full((12, 12), b)
    expect[5:8, 5:8] = [[b, 325, b], [550, 775, 212], [b, 325, b]]
    assert np.array_equal(out[:, :, 0], expect)


def test_hotpixel_threshold_keeps_inliers():
    rng = np.random.default_rng(3)
    F = rng.integers(900, 1100, size=(16, 16, 2)).astype(np.uint16)
    out = O.hotpixel(F, 50.0)  # nothing is 50 MADs out
    assert np.array_equal(out, F)
    F[8, 8, 1] = 60000
    out = O.hotpixel(F, 50.0)
    win = np.sort(np.delete(F[7:10, 7:10, 1].ravel(), 4).astype(float))
    assert out[8, 8, 1] == int(0.5 * (win[3] + win[4]))
    assert np.array_equal(np.delete(out.ravel(), np.ravel_multi_index((8, 8, 1), F.shape)),
                          np.delete(F.ravel(), np.ravel_multi_index((8, 8, 1), F.shape)))


# ------------------------------------------------------------------ CLI --

def _run_cli(tmp_path, text):
    pf = tmp_path / "params.txt"
    pf.write_text(text)
    return subprocess.run([CLI, str(pf)], capture_output=True, text=True, cwd=tmp_path, timeout=120)


def test_cli_usage_and_required_parameters(tmp_path):
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Usage" in r.stdout
    r = _run_cli(tmp_path, "filename : x.tif\nstart_frame : 1\n")
    assert r.returncode != 0
    assert "You must specify 'filename', 'start_frame' and 'end_frame'" in r.stderr


def test_cli_lambda_required_without_pgure(tmp_path):
    r = _run_cli(tmp_path, "filename : x.tif\nstart_frame : 1\nend_frame : 5\noptimize_pgure : false\n")
    assert r.returncode != 0 and "no lambda specified" in r.stderr


def test_cli_file_checks(tmp_path):
    r = _run_cli(tmp_path, "filename : missing.tiff   # a comment\nstart_frame : 1\nend_frame : 5\n")
    assert r.returncode != 0 and "File missing.tif not found" in r.stderr  # stem + ".tif"
    tiff_write(str(tmp_path / "rect.tif"), np.zeros((8, 16, 4), dtype=np.uint16))
    r = _run_cli(tmp_path, "filename : rect.tif\nstart_frame : 1\nend_frame : 4\n")
    assert r.returncode != 0 and "not square, got 16x8" in r.stderr
    tiff_write(str(tmp_path / "short.tif"), np.zeros((8, 8, 3), dtype=np.uint16))
    r = _run_cli(tmp_path, "filename : short.tif\nstart_frame : 1\nend_frame : 5\n")
    assert r.returncode != 0 and "Sequence only has 3 frames, expected 5" in r.stderr


def test_cli_parameter_syntax(tmp_path):
    # values are the tokens after ':' concatenated (src/utils.hpp:55-90):
    # "short . tif" names short.tif; a '#' token ends the value
    tiff_write(str(tmp_path / "short.tif"), np.zeros((8, 8, 2), dtype=np.uint16))
    r = _run_cli(tmp_path, "# header\n\nfilename : short . tif # trailing\nstart_frame : 1\nend_frame : 3 #\n")
    assert r.returncode != 0 and "Sequence only has 2 frames, expected 3" in r.stderr
