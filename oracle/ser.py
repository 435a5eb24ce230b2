"""SER files for the oracle and the tests (test infrastructure only).

``write_ser`` builds files by the published format ("SER format description
version 3", Heiko Wilkens and Grischa Hahn, 2014, the document
core/io/c_ser_file.h:1-22 follows): a 178-byte little-endian header, the
frames, then an optional array of frames_count 64-bit timestamps.
``read_ser`` restates c_ser_reader (core/io/c_ser_file.cc:246-450), including
its handling of the byte-order field (a stored 0 means little-endian,
c_ser_file.cc:278-279) and of byte swapping (every element size goes through
bswap_16, c_ser_file.cc:25-62, applied to 2- and 4-byte pixels only).
"""
from __future__ import annotations

import struct

import numpy as np

HEADER = struct.Struct("<14s7i40s40s40sQQ")
assert HEADER.size == 178

MONO, RGGB, GRBG, GBRG, BGGR, CYYM, YCMY, YMCY, MYYC, RGB, BGR = 0, 8, 9, 10, 11, 16, 17, 18, 19, 100, 101


def dtype_of(bits_per_plane: int):
    if 0 < bits_per_plane <= 8:
        return np.dtype(np.uint8)
    if 8 < bits_per_plane <= 16:
        return np.dtype(np.uint16)
    if 16 < bits_per_plane <= 32:
        return np.dtype(np.int32)
    if bits_per_plane == -32:
        return np.dtype(np.float32)
    if bits_per_plane == -64:
        return np.dtype(np.float64)
    raise ValueError(bits_per_plane)


def channels_of(color_id: int) -> int:
    return 3 if color_id in (RGB, BGR) else 1


def write_ser(path, frames, color_id=MONO, bits_per_plane=16, timestamps=None, stored_endian_field=0,
              observer=b"observer", instrument=b"instrument", telescope=b"telescope", date_time=0,
              date_time_utc=0, luid=0, byteswap_frames=False):
    """Frames are (H, W) or (H, W, 3) arrays of the depth of bits_per_plane, written in
    their memory byte order (byteswap_frames writes each element's bytes reversed)."""
    frames = [np.asarray(f) for f in frames]
    h, w = frames[0].shape[:2]
    hdr = HEADER.pack(b"LUCAM-RECORDER", luid, color_id, stored_endian_field, w, h, bits_per_plane, len(frames),
                      observer.ljust(40, b"\0"), instrument.ljust(40, b"\0"), telescope.ljust(40, b"\0"),
                      date_time, date_time_utc)
    dt = dtype_of(bits_per_plane)
    with open(path, "wb") as fh:
        fh.write(hdr)
        for f in frames:
            a = np.ascontiguousarray(f, dtype=dt)
            fh.write((a.byteswap() if byteswap_frames else a).tobytes())
        if timestamps is not None:
            fh.write(np.asarray(timestamps, dtype="<u8").tobytes())


def _swap16_all(a: np.ndarray) -> np.ndarray:
    """swap_endianess as written in the reference: bswap_16 of the low 16 bits, stored back
    into the element (the upper bytes become 0)."""
    u = a.view({2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize])
    lo = (u & 0xFFFF).astype(np.uint16)
    sw = ((lo >> 8) | (lo << 8)).astype(np.uint16)
    return sw.astype(u.dtype).view(a.dtype)


def read_ser(path):
    """Returns (header dict, list of frames, timestamps or None) as c_ser_reader reads them."""
    data = open(path, "rb").read()
    (fid, luid, color_id, le_field, w, h, bpp, n, obs, ins, tel, dt, dtu) = HEADER.unpack_from(data, 0)
    if fid != b"LUCAM-RECORDER":
        raise ValueError("not a SER file")
    little = le_field == 0
    hdr = dict(luid=luid, color_id=color_id, is_little_endian=int(little), image_width=w, image_height=h,
               bits_per_plane=bpp, frames_count=n, observer=obs, instrument=ins, telescope=tel, date_time=dt,
               date_time_utc=dtu)
    dtp = dtype_of(bpp)
    cn = channels_of(color_id)
    fb = w * h * cn * dtp.itemsize
    frames = []
    for i in range(n):
        a = np.frombuffer(data, dtype=dtp, count=w * h * cn, offset=178 + i * fb).copy()
        if not little and cn * dtp.itemsize in (2, 4):
            a = _swap16_all(a)
        frames.append(a.reshape((h, w, cn)) if cn == 3 else a.reshape((h, w)))
    ts = None
    off = 178 + n * fb
    if n > 0 and len(data) >= off + 8 * n:
        ts = np.frombuffer(data, dtype="<u8", count=n, offset=off).copy()
        if not little:
            ts = _swap16_all(ts)
    return hdr, frames, ts
