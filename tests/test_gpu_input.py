"""GPU input stage against the oracle (oracle/debayer.py, oracle/ser.py), through the C ABI:
debayer() of core/io/debayer.cc (NN2, SP, MATRIX, PLANE_k, AVGBP, DISABLE; OpenCV's bilinear
and edge-aware demosaicing behind NN / EA on 8/16-bit data; the fallbacks of NN/EA/VNG), the raw-frame conversion of read_input_frame
(c_image_stacking_pipeline_base.cc:124-279) and the SER reader into device memory.
Integer and float arithmetic follow the reference exactly: compared bit for bit."""
import numpy as np
import pytest

from oracle import debayer as D
from oracle import ser as SER

pytestmark = pytest.mark.gpu

PATTERNS = [8, 9, 10, 11, 16, 19]
ALGOS = [D.NN2, D.SP, D.MATRIX, D.PLANE_0, D.PLANE_3, D.AVGBP, D.DISABLE, D.DEFAULT]


@pytest.fixture(scope="module")
def S():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def to_gpu(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        return torch.from_numpy(a.view(np.int16)).cuda()
    return torch.from_numpy(a).cuda()


def from_gpu(t, dtype):
    a = t.cpu().numpy()
    return a.view(np.uint16) if dtype == np.uint16 else a


def empty_gpu(shape, dtype):
    import torch

    tdt = {np.uint8: torch.uint8, np.uint16: torch.int16, np.float32: torch.float32}[dtype]
    return torch.empty(shape, dtype=tdt, device="cuda")


def _src(dtype, shape, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.float32:
        return rng.random(shape).astype(np.float32)
    return rng.integers(0, np.iinfo(dtype).max + 1, shape).astype(dtype)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("colorid", PATTERNS)
def test_debayer_bitexact(S, dtype, algo, colorid):
    src = _src(dtype, (34, 46), colorid * 7 + algo)
    want = D.debayer(src, colorid, algo)
    out = empty_gpu(want.shape, dtype)
    S.debayer(to_gpu(src), colorid, algo, out=out)
    np.testing.assert_array_equal(from_gpu(out, dtype), want)
    # host source and destination
    np.testing.assert_array_equal(S.debayer(src, colorid, algo), want)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("algo", [D.NN, D.EA])
@pytest.mark.parametrize("colorid", PATTERNS)
@pytest.mark.parametrize("shape", [(34, 46), (4, 6), (2, 8), (8, 2)])
def test_debayer_opencv_bilinear_and_edge_aware(S, dtype, algo, colorid, shape):
    """DEBAYER_NN and DEBAYER_EA on 8/16-bit data are cv::demosaicing's bilinear and edge-aware
    modes (copied borders; 0 for 2-row / 2-column images): bit-exact, device and host."""
    src = _src(dtype, shape, colorid * 11 + algo + shape[0])
    want = D.debayer(src, colorid, algo)
    out = empty_gpu(want.shape, dtype)
    S.debayer(to_gpu(src), colorid, algo, out=out)
    np.testing.assert_array_equal(from_gpu(out, dtype), want)
    np.testing.assert_array_equal(S.debayer(src, colorid, algo), want)


@pytest.mark.parametrize("algo", [D.NN, D.EA])
def test_read_input_frames_opencv_demosaic(S, algo):
    """12-bit raw frames debayered by OpenCV's modes in 16 bits, then converted to f32 (no
    calibration); with a dark frame the data is float and both modes fall back to nn2."""
    import torch

    n, h, w, bpp = 3, 30, 44, 12
    rng = np.random.default_rng(algo)
    raw = rng.integers(0, 4096, (n, h, w)).astype(np.uint16)
    dark = (rng.random((h, w)) * 0.02).astype(np.float32)
    for dk in (None, dark):
        want = [D.read_input_frame(raw[i], 9, bpp, algo, dk) for i in range(n)]
        slab = to_gpu(raw)
        outs = torch.empty((n,) + want[0].shape, dtype=torch.float32, device="cuda")
        S.read_input_frames([slab[i] for i in range(n)], 9, bpp, algo, dark=dk, outs=[outs[i] for i in range(n)])
        for i in range(n):
            np.testing.assert_array_equal(outs[i].cpu().numpy(), want[i])


def test_debayer_opencv_methods(S):
    """VNG on 8-bit data is OpenCV's VNG demosaicing in the reference: rejected; NN / EA / VNG on
    f32 data and VNG on 16-bit data fall back to NN2, as in debayer() (debayer.cc:1243-1262)."""
    u8, u16, f32 = _src(np.uint8, (8, 8), 1), _src(np.uint16, (8, 8), 2), _src(np.float32, (8, 8), 3)
    with pytest.raises(RuntimeError):
        S.debayer(u8, 8, D.VNG)
    np.testing.assert_array_equal(S.debayer(u16, 8, D.VNG), D.nn2(u16, 8))
    for algo in (D.NN, D.EA, D.VNG):
        np.testing.assert_array_equal(S.debayer(f32, 8, algo), D.nn2(f32, 8))
    with pytest.raises(RuntimeError):
        S.debayer(_src(np.uint16, (7, 8), 4), 8, D.NN2)  # uneven size


@pytest.mark.parametrize("calib", ["none", "dark", "flat", "both"])
@pytest.mark.parametrize("algo", [D.NN2, D.SP, D.AVGBP, D.DISABLE])
def test_read_input_frames_batch(S, calib, algo):
    """Five 12-bit RGGB frames in one device slab (one launch) and as host frames."""
    import torch

    n, h, w, bpp = 5, 30, 44, 12
    rng = np.random.default_rng(algo + 10)
    raw = rng.integers(0, 4096, (n, h, w)).astype(np.uint16)
    dark = (rng.random((h, w)) * 0.02).astype(np.float32) if calib in ("dark", "both") else None
    flat = (0.8 + 0.4 * rng.random((h, w))).astype(np.float32) if calib in ("flat", "both") else None
    want = [D.read_input_frame(raw[i], 8, bpp, algo, dark, flat) for i in range(n)]
    slab = to_gpu(raw)
    outs = torch.empty((n,) + want[0].shape, dtype=torch.float32, device="cuda")
    S.read_input_frames([slab[i] for i in range(n)], 8, bpp, algo, dark=dark, flat=flat,
                        outs=[outs[i] for i in range(n)])
    for i in range(n):
        np.testing.assert_array_equal(outs[i].cpu().numpy(), want[i])
    hosts = [np.zeros(want[0].shape, np.float32) for _ in range(n)]
    S.read_input_frames([raw[i] for i in range(n)], 8, bpp, algo, dark=dark, flat=flat, outs=hosts)
    for i in range(n):
        np.testing.assert_array_equal(hosts[i], want[i])


def test_read_input_frames_mono_rgb_f32(S):
    rng = np.random.default_rng(3)
    mono = rng.integers(0, 256, (2, 10, 12)).astype(np.uint8)
    out = [np.zeros((10, 12), np.float32) for _ in range(2)]
    S.read_input_frames(list(mono), 0, 8, D.NN2, outs=out)
    for i in range(2):
        np.testing.assert_array_equal(out[i], mono[i].astype(np.float32) / np.float32(256))
    rgb = rng.integers(0, 65536, (10, 12, 3)).astype(np.uint16)
    o3 = [np.zeros((10, 12, 3), np.float32)]
    S.read_input_frames([rgb], 100, 16, D.NN2, outs=o3)
    np.testing.assert_array_equal(o3[0], rgb.astype(np.float32) / np.float32(65536))
    f = rng.random((10, 12)).astype(np.float32)
    o = [np.zeros((10, 12, 3), np.float32)]
    S.read_input_frames([f], 9, 16, D.NN2, outs=o)  # f32 raw: no scaling, float debayer
    np.testing.assert_array_equal(o[0], D.nn2(f, 9))


@pytest.mark.parametrize("bpp,cid,endian", [(12, 8, 0), (16, 0, 1), (8, 11, 0), (-32, 0, 0), (8, 100, 0)])
def test_ser_device_reads(S, tmp_path, bpp, cid, endian):
    """read() and read_batch() into device tensors equal the oracle reader; the frames then
    go through the input conversion on the device."""
    import torch

    dt = SER.dtype_of(bpp)
    cn = 3 if cid == 100 else 1
    shape = (24, 32) + ((3,) if cn == 3 else ())
    rng = np.random.default_rng(abs(bpp) + cid)
    if dt.kind == "f":
        frames = [rng.random(shape).astype(dt) for _ in range(6)]
    else:
        frames = [rng.integers(0, min(1 << bpp, np.iinfo(dt).max + 1), shape).astype(dt) for _ in range(6)]
    path = str(tmp_path / "s.ser")
    SER.write_ser(path, frames, cid, bpp, timestamps=list(range(6)), stored_endian_field=endian,
                  byteswap_frames=bool(endian))
    _, want, _ = SER.read_ser(path)
    r = S.c_ser_reader(path)
    one = empty_gpu(shape, dt.type)
    r.read(one)
    np.testing.assert_array_equal(from_gpu(one, dt.type), want[0])
    slab = empty_gpu((5,) + shape, dt.type)
    r.read_batch([slab[i] for i in range(5)])
    assert r.curpos() == 6
    for i in range(5):
        np.testing.assert_array_equal(from_gpu(slab[i], dt.type), want[i + 1])
    if cn == 1 and dt.kind != "f":
        conv = torch.empty((5, 24, 32, 3) if cid != 0 else (5, 24, 32), dtype=torch.float32, device="cuda")
        S.read_input_frames([slab[i] for i in range(5)], cid, abs(bpp), D.NN2, outs=[conv[i] for i in range(5)])
        for i in range(5):
            np.testing.assert_array_equal(conv[i].cpu().numpy(), D.read_input_frame(want[i + 1], cid, abs(bpp)))
