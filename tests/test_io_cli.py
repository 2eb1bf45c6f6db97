"""Image files and command-line tools, host side (no GPU).

The file formats are checked against encoders and decoders written here in
Python from the format specifications (PNG: zlib stream, CRC-32, the five
row filters; TIFF: baseline tags; Middlebury .flo), so that a mistake in
the C++ codec (a filter predictor, a byte order, a plane order) fails one of
them.  The tools' argument checks and messages are those of
src/main_vnlb.cpp:89-130, lib/tvl1flow/main.c:75-84, 160-196 and
scripts/tvl1flow-seq.sh:13-21; those paths end before any GPU work."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from conftest import bm4dof_amd as A


# ---------------------------------------------------------------- encoders

def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def _png_bytes(samples, depth):
    """samples: (H, W, C) unsigned ints; every row uses filter y % 5."""
    H, W, Cc = samples.shape
    ctype = {1: 0, 2: 4, 3: 2, 4: 6}[Cc]
    bps = depth // 8
    bpp = bps * Cc
    dt = ">u2" if depth == 16 else "u1"
    raw = bytearray()
    prev = bytes(W * bpp)
    for y in range(H):
        row = samples[y].astype(dt).tobytes()
        ft = y % 5
        out = bytearray(len(row))
        for i in range(len(row)):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]
            out[i] = (row[i] - pred) & 255
        raw += bytes([ft]) + out
        prev = row

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)
    ihdr = struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(bytes(raw)))
            + chunk(b"IEND", b""))


def _png_decode(data):
    """Minimal independent PNG decoder (8-bit, no interlace) with CRC checks."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    o, idat, hdr = 8, b"", None
    while o < len(data):
        n, = struct.unpack(">I", data[o:o + 4])
        t, d = data[o + 4:o + 8], data[o + 8:o + 8 + n]
        crc, = struct.unpack(">I", data[o + 8 + n:o + 12 + n])
        assert crc == zlib.crc32(t + d) & 0xffffffff
        if t == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", d)
        elif t == b"IDAT":
            idat += d
        o += 12 + n
    W, H, depth, ctype = hdr[:4]
    assert depth == 8
    Cc = {0: 1, 4: 2, 2: 3, 6: 4}[ctype]
    raw = zlib.decompress(idat)
    stride = W * Cc
    prev = bytearray(stride)
    rows = []
    for y in range(H):
        ft, s = raw[y * (stride + 1)], raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)]
        cur = bytearray(stride)
        for i in range(stride):
            a = cur[i - Cc] if i >= Cc else 0
            b = prev[i]
            c = prev[i - Cc] if i >= Cc else 0
            cur[i] = (s[i] + [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]) & 255
        rows.append(bytes(cur))
        prev = cur
    return np.frombuffer(b"".join(rows), np.uint8).reshape(H, W, Cc)


def _tiff_bytes(planes, bits, fmt, planar, big):
    """Baseline uncompressed TIFF, one strip per plane (planar) or one strip."""
    Cc, H, W = planes.shape
    e = ">" if big else "<"
    dt = {(8, 1): "u1", (16, 1): "u2", (32, 3): "f4"}[(bits, fmt)]
    if planar == 2:
        strips = [planes[c].astype(e + dt).tobytes() for c in range(Cc)]
    else:
        strips = [np.moveaxis(planes, 0, -1).astype(e + dt).tobytes()]
    head = (b"MM" if big else b"II") + struct.pack(e + "HI", 42, 8)
    ntag = 11
    ifd_len = 2 + 12 * ntag + 4
    extra_off = 8 + ifd_len
    offs_pos = extra_off
    cnts_pos = offs_pos + 4 * len(strips)
    bits_pos = cnts_pos + 4 * len(strips)
    data_pos = bits_pos + 2 * Cc
    offs, p = [], data_pos
    for s in strips:
        offs.append(p)
        p += len(s)

    def ent(tag, typ, cnt, val):
        if typ == 3 and cnt == 1:
            return struct.pack(e + "HHIHH", tag, typ, cnt, val, 0)
        return struct.pack(e + "HHII", tag, typ, cnt, val)
    so = ent(273, 4, len(strips), offs_pos) if len(strips) > 1 else ent(273, 4, 1, offs[0])
    sc = ent(279, 4, len(strips), cnts_pos) if len(strips) > 1 else ent(279, 4, 1, len(strips[0]))
    bps = ent(258, 3, Cc, bits_pos) if Cc > 2 else (ent(258, 3, 1, bits) if Cc == 1 else None)
    if Cc == 2:
        bps = struct.pack(e + "HHIHH", 258, 3, 2, bits, bits)
    tags = [ent(256, 4, 1, W), ent(257, 4, 1, H), bps, ent(259, 3, 1, 1), ent(262, 3, 1, 1 if Cc < 3 else 2),
            so, ent(277, 3, 1, Cc), ent(278, 4, 1, H), sc, ent(284, 3, 1, planar), ent(339, 3, 1, fmt)]
    ifd = struct.pack(e + "H", ntag) + b"".join(tags) + struct.pack(e + "I", 0)
    extra = (b"".join(struct.pack(e + "I", o) for o in offs) + b"".join(struct.pack(e + "I", len(s)) for s in strips)
             + b"".join(struct.pack(e + "H", bits) for _ in range(Cc)))
    return head + ifd + extra + b"".join(strips)


# ------------------------------------------------------------------ formats

def test_io_library_exports_header():
    import re
    text = open(os.path.join(os.path.dirname(A.LIB_PATH), "..", "include", "bm4dof_io.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = sorted(set(re.findall(r"\b(bm4dof_io_\w+)\s*\(", text)))
    assert declared == sorted(A.IO_EXPORTS)
    L = A.io_lib()
    assert all(hasattr(L, n) for n in declared)


@pytest.mark.parametrize("channels,depth", [(1, 8), (3, 8), (4, 8), (2, 8), (1, 16), (3, 16)])
def test_png_read_all_filters(tmp_path, channels, depth):
    rng = np.random.default_rng(channels * 100 + depth)
    hi = 255 if depth == 8 else 65535
    s = rng.integers(0, hi + 1, size=(11, 13, channels))
    s[3] = s[2]   # repeated rows and smooth runs exercise the predictors
    s[:, 5:9] = s[:, 5:6]
    p = tmp_path / "a.png"
    p.write_bytes(_png_bytes(s, depth))
    img = A.read_image(p)
    assert img.shape == (channels, 11, 13)
    assert np.array_equal(img, np.moveaxis(s, -1, 0).astype(np.float32))


@pytest.mark.parametrize("channels", [1, 3, 4])
def test_png_write_quantises_like_the_reference(tmp_path, channels):
    """(uint8)clamp(v + 0.5, 0, 255) per sample, and a valid PNG stream."""
    rng = np.random.default_rng(channels)
    v = (rng.random((channels, 9, 10)) * 300 - 25).astype(np.float32)
    v[0, 0, :4] = [-0.6, 254.49, 254.5, 0.5]
    p = tmp_path / "q.png"
    A.write_image(p, v)
    want = np.clip(v + 0.5, 0, 255).astype(np.uint8)
    got = _png_decode(p.read_bytes())
    assert np.array_equal(np.moveaxis(got, -1, 0), want)
    assert np.array_equal(A.read_image(p), want.astype(np.float32))


@pytest.mark.parametrize("bits,fmt,planar,big,channels", [
    (8, 1, 1, False, 1), (16, 1, 1, True, 3), (32, 3, 2, False, 3), (32, 3, 1, True, 2), (16, 1, 2, False, 4)])
def test_tiff_read(tmp_path, bits, fmt, planar, big, channels):
    rng = np.random.default_rng(bits + channels)
    if fmt == 3:
        v = rng.standard_normal((channels, 7, 9)).astype(np.float32) * 100
    else:
        v = rng.integers(0, 2 ** bits, size=(channels, 7, 9)).astype(np.float32)
    p = tmp_path / "a.tif"
    p.write_bytes(_tiff_bytes(v, bits, fmt, planar, big))
    assert np.array_equal(A.read_image(p), v)


def test_tiff_float_round_trip_is_exact(tmp_path):
    v = np.random.default_rng(5).standard_normal((3, 17, 23)).astype(np.float32) * 1e3
    v[0, 0, 0] = np.float32(1e-40)   # subnormal
    p = tmp_path / "r.tiff"
    A.write_image(p, v)
    assert np.array_equal(A.read_image(p), v)
    b = p.read_bytes()
    assert b[:4] == b"II*\x00"


def test_flo_layout(tmp_path):
    """Tag 202021.25, int32 width and height, then (u, v) interleaved."""
    uv = np.random.default_rng(2).standard_normal((2, 5, 6)).astype(np.float32)
    p = tmp_path / "f.flo"
    A.write_image(p, uv)
    b = p.read_bytes()
    tag, w, h = struct.unpack("<fii", b[:12])
    assert (tag, w, h) == (202021.25, 6, 5) and len(b) == 12 + 8 * 30
    inter = np.frombuffer(b[12:], np.float32).reshape(5, 6, 2)
    assert np.array_equal(np.moveaxis(inter, -1, 0), uv)
    assert np.array_equal(A.read_image(p), uv)


def test_gray_read_truncates_integer_samples(tmp_path):
    """3 and 4 channels become .299 R + .587 G + .114 B in double, stored in
    the file's sample type (iio.c:935-1003)."""
    rng = np.random.default_rng(9)
    s = rng.integers(0, 256, size=(6, 7, 3))
    s[0, 0] = 255                       # 0.299+0.587+0.114 in double is just below 1
    p = tmp_path / "c.png"
    p.write_bytes(_png_bytes(s, 8))
    d = s.astype(np.float64)
    want = (0.299 * d[..., 0] + 0.587 * d[..., 1] + 0.114 * d[..., 2]).astype(np.uint8)
    assert np.array_equal(A.read_gray(p), want.astype(np.float32))
    s4 = np.concatenate([s, np.full((6, 7, 1), 255)], -1)
    p.write_bytes(_png_bytes(s4, 8))
    assert np.array_equal(A.read_gray(p), want.astype(np.float32))
    # float samples round to float
    v = rng.random((3, 6, 7)).astype(np.float32) * 255
    q = tmp_path / "c.tif"
    A.write_image(q, v)
    d = v.astype(np.float64)
    assert np.array_equal(A.read_gray(q), (0.299 * d[0] + 0.587 * d[1] + 0.114 * d[2]).astype(np.float32))
    # one channel passes through; two channels are an error
    A.write_image(q, v[:1])
    assert np.array_equal(A.read_gray(q), v[0])
    A.write_image(q, v[:2])
    with pytest.raises(OSError, match="non-scalarizable"):
        A.read_gray(q)


def test_png_header_checks(tmp_path):
    """A short IHDR chunk at the end of the file and sizes of 2^32 samples
    or more are errors, not reads past the buffer or huge allocations."""
    import struct
    import zlib

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d))
    sig = b"\x89PNG\r\n\x1a\n"
    short = sig + chunk(b"IDAT", b"x" * 16) + chunk(b"IHDR", b"\x00\x00\x00\x10\x00")
    p = tmp_path / "short.png"
    p.write_bytes(short)
    with pytest.raises(OSError, match="IHDR"):
        A.read_image(p)
    big = sig + chunk(b"IHDR", struct.pack(">IIBBBBB", 70000, 70000, 8, 6, 0, 0, 0)) + \
        chunk(b"IDAT", zlib.compress(b"\x00" * 16)) + chunk(b"IEND", b"")
    p = tmp_path / "big.png"
    p.write_bytes(big)
    with pytest.raises(OSError, match="too large"):
        A.read_image(p)


def test_read_errors(tmp_path):
    with pytest.raises(OSError):
        A.read_image(tmp_path / "missing.png")
    p = tmp_path / "x.bin"
    p.write_bytes(b"not an image")
    with pytest.raises(OSError, match="unknown format"):
        A.read_image(p)


# -------------------------------------------------------------------- tools

def _tool(name, *args, cwd=None):
    exe = A.TOOLS[name]
    assert os.path.exists(exe), f"{exe} is not built"
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=60)
    return r.returncode, r.stdout, r.stderr


@pytest.mark.parametrize("args,msg", [
    ((), "no input sequence."),
    (("-i", "x_%03d.png", "-px1", "0", "-px2", "4"), "if px1 = 0 and px2 > 0, a basic sequence path must be given."),
    (("-i", "x", "-px1", "-3"), "px1, px2, pt1 and pt2 cannot be negative."),
    (("-i", "x", "-pt2", "-1"), "px1, px2, pt1 and pt2 cannot be negative."),
    (("-i", "x", "-np2", "-2"), "np1 and np2 cannot be negative."),
    (("-i", "x", "-wt1", "-1"), "wx1, wx2, wt1 and wt2 cannot be negative."),
])
def test_bm4d_of_argument_checks(args, msg):
    rc, out, err = _tool("bm4d-of", *args)
    assert rc == 1
    assert err.splitlines()[0].endswith(": " + msg)
    assert "--help' for more information." in err


def test_bm4d_of_flow_and_patch_checks():
    rc, _, err = _tool("bm4d-of", "-i", "x", "-fof", "f_%03d.flo")
    assert rc == 1 and err.startswith("Only one oflow path provided.")
    rc, _, err = _tool("bm4d-of", "-i", "x", "-px1", "0", "-px2", "0", "-b", "y")
    assert rc == 1 and err.startswith("Given patch sizes for both steps are 0. Nothing to do.")


def test_bm4d_of_predictive_search_flags():
    """-vbm3d-search: the flags and printed lines of a VBM3D_SEARCH build
    (main_vnlb.cpp:71-78, 165-172; nldct.cpp:233-239); without it the
    predictive flags change nothing and print nothing."""
    args = ("-i", "x", "-print-prms", "1", "-sigma", "20", "-wxf1", "3", "-npf2", "6", "-dsub1", "2.5")
    rc, out, _ = _tool("bm4d-of", *args, "-vbm3d-search", "1")
    assert rc == 0
    b1, b2 = out.split("\x1b[37;01mParameters for step ")[1:]
    assert "Pred. spatial search window = 3x3" in b1 and "Pred. distance bias (dsub)  = 2.5" in b1
    assert "Similar patches per frame   = 6" in b2 and "Pred. distance bias (dsub)  = 3\n" in b2
    rc, out, _ = _tool("bm4d-of", *args)
    assert rc == 0 and "Pred." not in out


def test_bm4d_of_print_parameters():
    """-print-prms: the defaults for a 0-frame video of the given channels,
    then the overrides, printed in the reference's format."""
    rc, out, _ = _tool("bm4d-of", "-i", "x", "-print-prms", "1", "-sigma", "20", "-px1", "4", "-pt1", "4",
                       "-px2", "4", "-pt2", "4", "-np1", "32", "-np2", "32", "-wx1", "11", "-wx2", "11",
                       "-t2", "12.5", "-b1", "3", "-order-inv", "1")
    assert rc == 0
    blocks = out.split("\x1b[37;01mParameters for step ")[1:]
    assert len(blocks) == 2 and blocks[0].startswith("1:") and blocks[1].startswith("2:")
    assert "Patch size                  = 4x4x4" in blocks[0]
    assert "Number of patches           = 32" in blocks[1]
    assert "Spatial search window       = 11x11" in blocks[0]
    assert "Distance threshold (tau)    = 12.5" in blocks[1]   # -t2 applies when printing
    assert "Beta                        = 3\n" in blocks[0]
    assert "Order invariance            = yes" in blocks[1]
    p1 = A.cli_params(1, 20, (0, 1, 0, 0), px=4, pt=4, wx=11, np_=32, beta=3.0, order_inv=True)
    assert f"Offset                      = {p1.offSet}\n" in blocks[0]
    assert f"Temporal search window      = [-{p1.sizeSearchTimeRangeBwd},{p1.sizeSearchTimeRangeBwd}]" in blocks[0]
    assert "Transform                   = dct" in blocks[0]


def test_bm4d_of_help_lists_options():
    rc, _, err = _tool("bm4d-of", "--help")
    assert rc == 1 and "no input sequence." in err      # help does not stop the program
    for name in ("-i", "-deno", "-bsic", "-fof", "-bof", "-sigma", "-px1", "-np2", "-agg-win2"):
        assert f"    {name:<8s} = " in err


def test_bm4d_of_missing_input_frame(tmp_path):
    rc, _, err = _tool("bm4d-of", "-i", str(tmp_path / "none_%03d.png"), "-f", "1", "-l", "2")
    assert rc == 1 and "Video<T>::loadVideo: loading of " in err and "none_001.png failed" in err


def test_tvl1flow_usage_and_size_mismatch(tmp_path):
    rc, _, err = _tool("tvl1flow", "a.png")
    assert rc == 1 and err.startswith("Usage: ") and "I0 I1 [out nproc tau lambda theta nscales fscale" in err
    A.write_image(tmp_path / "a.png", np.zeros((1, 20, 30)))
    A.write_image(tmp_path / "b.png", np.zeros((1, 21, 30)))
    rc, _, err = _tool("tvl1flow", tmp_path / "a.png", tmp_path / "b.png", tmp_path / "o.flo")
    assert rc == 1 and "ERROR: input images size mismatch 30x20 != 30x21" in err
    rc, _, err = _tool("tvl1flow", tmp_path / "a.png", tmp_path / "none.png")
    assert rc == 1 and f'ERROR: could not read image from file "{tmp_path / "none.png"}"' in err


def test_tvl1flow_seq_checks(tmp_path):
    rc, _, err = _tool("tvl1flow-seq", "x")
    assert rc == 1 and err.startswith("Usage: ")
    A.write_image(tmp_path / "f_000.png", np.zeros((1, 20, 30)))
    rc, out, _ = _tool("tvl1flow-seq", tmp_path / "f_%03d.png", 0, 1, tmp_path / "out")
    assert rc == 1 and out.strip() == f"ERROR: {tmp_path / 'f_001.png'} not found"
    assert (tmp_path / "out").is_dir()       # the folder is made first, as mkdir -p does
