"""Pins of the input-stage oracle (oracle/debayer.py, oracle/ser.py) and the SER reader
of the C ABI on host images (no GPU).

debayer: the oracle against a pure-Python brute force of the definition -- each
missing colour is the mean of the same-coloured sites of the 3x3 neighbourhood (mirrored
without repeating the edge) at the smallest Manhattan distance, summed in raster order
(the operand order of debayer.cc:828-1160), with the integer rounding (n/2 + sum) / n --
plus closed forms.  SER: the header layout of the published format, round trips, the
reference's byte-order field and swap."""
import os
import struct

import numpy as np
import pytest

from oracle import debayer as D
from oracle import ser as SER

PATTERNS = [8, 9, 10, 11, 16, 17, 18, 19]


def brute_nn2(src, colorid):
    h, w = src.shape
    integer = np.issubdtype(src.dtype, np.integer)
    cm = D.color_map(colorid, h, w)
    refl = lambda i, n: -i if i < 0 else (2 * n - 2 - i if i >= n else i)  # noqa: E731
    out = np.zeros((h, w, 3), src.dtype)
    for y in range(h):
        for x in range(w):
            for c in range(3):
                if cm[y, x] == c:
                    out[y, x, c] = src[y, x]
                    continue
                best, terms = None, []
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        yy, xx = refl(y + dy, h), refl(x + dx, w)
                        if (dy or dx) and cm[yy, xx] == c:
                            d = abs(dy) + abs(dx)
                            if best is None or d < best:
                                best, terms = d, [src[yy, xx]]
                            elif d == best:
                                terms.append(src[yy, xx])
                n = len(terms)
                if integer:
                    out[y, x, c] = (n // 2 + sum(int(t) for t in terms)) // n
                else:
                    s = np.float32(0)
                    for t in terms:
                        s = np.float32(s + np.float32(t))
                    out[y, x, c] = np.float32(s / np.float32(n))
    return out


@pytest.mark.parametrize("colorid", PATTERNS)
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("shape", [(2, 2), (4, 6), (6, 4), (8, 10)])
def test_nn2_matches_definition(colorid, dtype, shape):
    rng = np.random.default_rng(hash((colorid, shape)) % 1000)
    if dtype == np.float32:
        src = rng.random(shape).astype(np.float32)
    else:
        src = rng.integers(0, np.iinfo(dtype).max + 1, shape).astype(dtype)
    np.testing.assert_array_equal(D.nn2(src, colorid), brute_nn2(src, colorid))


@pytest.mark.parametrize("colorid", [8, 9, 10, 11])
def test_constant_frames_stay_constant(colorid):
    for src in (np.full((6, 8), 1234, np.uint16), np.full((6, 8), 0.375, np.float32)):
        for f in (D.nn2, D.sp):
            assert np.all(f(src, colorid) == src.flat[0])
        assert np.all(D.avgbp(src) == src.flat[0])


def test_sp_avgbp_planes_matrix_closed_forms():
    src = np.array([[1, 2], [3, 4]], np.uint16)  # RGGB: R=1, G=2,3, B=4
    np.testing.assert_array_equal(D.sp(src, 8)[0, 0], [4, (1 + 2 + 3) // 2, 1])
    np.testing.assert_array_equal(D.sp(src, 11)[0, 0], [1, (1 + 2 + 3) // 2, 4])  # BGGR
    assert D.avgbp(src)[0, 0] == (2 + 1 + 2 + 3 + 4) // 4
    assert [D.plane(src, k)[0, 0] for k in range(4)] == [1, 2, 3, 4]
    m = D.matrix(src, 9)  # GRBG: G R / B G
    np.testing.assert_array_equal(m[0, 0], [0, 1, 0])
    np.testing.assert_array_equal(m[0, 1], [0, 0, 2])
    np.testing.assert_array_equal(m[1, 0], [3, 0, 0])
    f = src.astype(np.float32)
    assert D.avgbp(f)[0, 0] == np.float32(2.5)


def test_read_input_frame_conversion_rules():
    rng = np.random.default_rng(2)
    raw = rng.integers(0, 4096, (6, 8)).astype(np.uint16)
    # no calibration: debayer in 16 bits (integer rounding), then x / 2^bpp
    got = D.read_input_frame(raw, 8, 12)
    np.testing.assert_array_equal(got, D.nn2(raw, 8).astype(np.float32) / np.float32(4096))
    # dark and flat: x / 2^bpp - dark, / flat, debayered in float
    dark = rng.random((6, 8)).astype(np.float32) * 0.01
    flat = (0.9 + 0.2 * rng.random((6, 8))).astype(np.float32)
    want = D.nn2(((raw.astype(np.float32) / np.float32(4096) - dark) / flat).astype(np.float32), 8)
    np.testing.assert_array_equal(D.read_input_frame(raw, 8, 12, dark=dark, flat=flat), want)
    # mono frames are only converted
    np.testing.assert_array_equal(D.read_input_frame(raw, 0, 12), raw.astype(np.float32) / np.float32(4096))
    with pytest.raises(NotImplementedError):
        D.debayer(raw.astype(np.uint8), 8, D.VNG)


@pytest.mark.parametrize("colorid", [8, 9, 10, 11])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_opencv_bilinear_is_nn2_inside_with_copied_borders(colorid, dtype):
    """OpenCV's bilinear demosaic computes every interior pixel with the same integer averages
    as the reference's own nn2 (pinned on its own): the two agree on rows and columns 1 .. n-2;
    the border columns and rows are copies of their inner neighbours."""
    rng = np.random.default_rng(colorid)
    raw = rng.integers(0, np.iinfo(dtype).max + 1, (12, 18)).astype(dtype)
    got = D.debayer(raw, colorid, D.NN)
    assert got.dtype == dtype and got.shape == (12, 18, 3)
    np.testing.assert_array_equal(got[1:-1, 1:-1], D.nn2(raw, colorid)[1:-1, 1:-1])
    np.testing.assert_array_equal(got[1:-1, 0], got[1:-1, 1])
    np.testing.assert_array_equal(got[1:-1, -1], got[1:-1, -2])
    np.testing.assert_array_equal(got[0], got[1])
    np.testing.assert_array_equal(got[-1], got[-2])
    # each pixel keeps its own sample in its own channel
    cm = D.color_map(colorid, 12, 18)
    inner = got[1:-1, 1:-1]
    np.testing.assert_array_equal(np.take_along_axis(inner, cm[1:-1, 1:-1, None], 2)[..., 0], raw[1:-1, 1:-1])


@pytest.mark.parametrize("colorid", [8, 11])
def test_opencv_ea_follows_the_flatter_direction(colorid):
    """Edge-aware green at red / blue sites: on horizontal stripes (rows constant, so |L - R| = 0)
    it is the left/right average, i.e. the site's own row value; on vertical stripes (columns
    constant) it is the up/down average, the site's own column value.  Bilinear mixes in the
    other direction instead.  The other channels are the bilinear ones."""
    h, w = 10, 14
    rows = np.tile((np.arange(h) * 37 % 251)[:, None], (1, w)).astype(np.uint8)
    cols = np.tile((np.arange(w) * 53 % 251)[None, :], (h, 1)).astype(np.uint8)
    cm = D.color_map(colorid, h, w)[1:-1, 1:-1]
    for img, own in ((rows, rows), (cols, cols)):
        ea = D.debayer(img, colorid, D.EA)
        nn = D.debayer(img, colorid, D.NN)
        g = ea[1:-1, 1:-1, 1]
        np.testing.assert_array_equal(g[cm != 1], own[1:-1, 1:-1][cm != 1])
        assert not np.array_equal(nn[1:-1, 1:-1, 1][cm != 1], own[1:-1, 1:-1][cm != 1])
        for k in (0, 2):
            np.testing.assert_array_equal(ea[..., k], nn[..., k])


def test_opencv_demosaic_constants_and_small_images():
    for algo in (D.NN, D.EA):
        for c in (0, 1, 200, 255):
            np.testing.assert_array_equal(D.debayer(np.full((8, 10), c, np.uint8), 8, algo), c)
        np.testing.assert_array_equal(D.debayer(np.full((2, 10), 9, np.uint16), 9, algo), 0)
        np.testing.assert_array_equal(D.debayer(np.full((8, 2), 9, np.uint16), 9, algo), 0)


def test_ser_header_layout_follows_the_format():
    # offsets of SER_Doc_V3b: FileID 0, LuID 14, ColorID 18, LittleEndian 22, Width 26, Height 30,
    # PixelDepthPerPlane 34, FrameCount 38, Observer 42, Instrument 82, Telescope 122,
    # DateTime 162, DateTime_UTC 170; 178 bytes
    b = SER.HEADER.pack(b"LUCAM-RECORDER", 1, 2, 3, 4, 5, 6, 7, b"o" * 40, b"i" * 40, b"t" * 40, 8, 9)
    assert len(b) == 178
    assert struct.unpack_from("<iiiiiii", b, 14) == (1, 2, 3, 4, 5, 6, 7)
    assert b[42:43] == b"o" and b[82:83] == b"i" and b[122:123] == b"t"
    assert struct.unpack_from("<QQ", b, 162) == (8, 9)


@pytest.mark.parametrize("bpp,cid,cn", [(8, 0, 1), (12, 8, 1), (16, 0, 1), (-32, 0, 1), (-64, 0, 1), (24, 0, 1),
                                        (8, 100, 3), (16, 101, 3)])
def test_ser_round_trip(tmp_path, bpp, cid, cn):
    rng = np.random.default_rng(abs(bpp) + cid)
    dt = SER.dtype_of(bpp)
    shape = (5, 7) + ((3,) if cn == 3 else ())
    if dt.kind == "f":
        frames = [rng.random(shape).astype(dt) for _ in range(3)]
    else:
        frames = [rng.integers(0, min(np.iinfo(dt).max, 1 << 20) + 1, shape).astype(dt) for _ in range(3)]
    path = str(tmp_path / "a.ser")
    SER.write_ser(path, frames, cid, bpp, timestamps=[10, 20, 30])
    hdr, got, ts = SER.read_ser(path)
    assert hdr["is_little_endian"] == 1 and hdr["frames_count"] == 3 and hdr["color_id"] == cid
    for a, b in zip(frames, got):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ts, [10, 20, 30])
    SER.write_ser(path, frames, cid, bpp)
    assert SER.read_ser(path)[2] is None


def test_ser_byte_order_field(tmp_path):
    """A non-zero stored field means big-endian frames: 16-bit pixels are swapped back;
    32-bit pixels go through the reference's bswap_16 and keep their low two bytes."""
    f16 = np.array([[0x0102, 0xA0B0]], np.uint16)
    path = str(tmp_path / "b.ser")
    SER.write_ser(path, [f16], 0, 16, stored_endian_field=1, byteswap_frames=True, timestamps=[0x0102])
    hdr, got, ts = SER.read_ser(path)
    assert hdr["is_little_endian"] == 0
    np.testing.assert_array_equal(got[0], f16)
    assert ts[0] == 0x0201  # bswap_16 of the low bytes of the stored value
    f32 = np.array([[0x11223344]], np.int32)
    SER.write_ser(path, [f32], 0, 32, stored_endian_field=1)
    assert SER.read_ser(path)[1][0][0, 0] == 0x4433


def test_ser_reader_abi_on_host_images(tmp_path):
    """c_ser_reader through the C ABI into host arrays (no GPU use): header, timestamps,
    sequential reads, seek, the end of the sequence, non-SER files."""
    import serstacker_amd as S

    rng = np.random.default_rng(9)
    frames = [rng.integers(0, 4096, (6, 10)).astype(np.uint16) for _ in range(4)]
    path = str(tmp_path / "c.ser")
    SER.write_ser(path, frames, 8, 12, timestamps=[5, 6, 7, 8], observer=b"me")
    r = S.c_ser_reader(path)
    assert r.is_open() and r.num_frames() == 4 and r.color_id() == 8 and r.bits_per_plane() == 12
    h = r.header()
    assert h.observer == b"me" and h.is_little_endian == 1 and (h.image_width, h.image_height) == (10, 6)
    np.testing.assert_array_equal(r.timestamps(), [5, 6, 7, 8])
    for i in range(4):
        assert r.curpos() == i
        np.testing.assert_array_equal(r.read(), frames[i])
    with pytest.raises(RuntimeError):
        r.read()
    assert r.seek(2)
    # a host image whose rows are longer than the frame's
    import ctypes as C

    from serstacker_amd import _lib
    big = np.zeros((6, 16), np.uint16)
    v = _lib.Image(big.ctypes.data, 10, 6, big.strides[0], _lib.SSG_16UC1, 0)
    _lib.check(_lib.load().ssg_ser_read(r._h, C.byref(v)), "read")
    np.testing.assert_array_equal(big[:, :10], frames[2])
    assert not big[:, 10:].any()
    assert not r.seek(7)
    # big-endian 16-bit file
    SER.write_ser(path, frames, 0, 16, stored_endian_field=1, byteswap_frames=True)
    r2 = S.c_ser_reader(path)
    np.testing.assert_array_equal(r2.read(), frames[0])
    r2.close()
    bad = str(tmp_path / "d.ser")
    open(bad, "wb").write(b"x" * 200)
    assert not S.c_ser_reader().open(bad)
