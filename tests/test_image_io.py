"""Image files (no GPU): Radiance RGBE reader against hand-made files with
known values, writer round trips, PFM exact round trip, and the OpenEXR
writer's structure parsed independently of the library."""
import struct

import numpy as np
import pytest

from colvillea_amd import imageio


def rgbe_value(c, e):
    # the reference loader's convention (imageutil.cpp:98-139): c * 2^(e - 136)
    return 0.0 if e == 0 else c * 2.0 ** (e - 136)


def test_hdr_flat_scanlines_known_values(tmp_path):
    quads = [(128, 64, 32, 129), (255, 0, 1, 136), (0, 0, 0, 0), (10, 20, 30, 120), (1, 2, 3, 200), (7, 7, 7, 128)]
    body = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y 2 +X 3\n" + bytes(b for q in quads for b in q)
    p = tmp_path / "flat.hdr"
    p.write_bytes(body)
    img = imageio.load_hdr(p)
    assert img.shape == (2, 3, 4)
    want = np.array([[rgbe_value(c, q[3]) for c in q[:3]] + [1.0] for q in quads], np.float32).reshape(2, 3, 4)
    assert np.array_equal(img, want)


def test_hdr_run_length_scanline(tmp_path):
    w = 10
    # one scanline: R = 5 x 100 then 5 literals, G run of 10, B literals, E run
    r = bytes([128 + 5, 100, 5, 1, 2, 3, 4, 5])
    g = bytes([128 + 10, 50])
    b = bytes([10]) + bytes(range(10))
    e = bytes([128 + 10, 130])
    body = b"#?RGBE\n\n-Y 1 +X 10\n" + bytes([2, 2, 0, w]) + r + g + b + e
    p = tmp_path / "rle.hdr"
    p.write_bytes(body)
    img = imageio.load_hdr(p)[0]
    R = [100] * 5 + [1, 2, 3, 4, 5]
    for x in range(w):
        assert img[x, 0] == rgbe_value(R[x], 130)
        assert img[x, 1] == rgbe_value(50, 130)
        assert img[x, 2] == rgbe_value(x, 130)


def test_hdr_round_trip_precision(tmp_path):
    rng = np.random.default_rng(0)
    img = np.exp(rng.normal(0, 3, (37, 53, 3))).astype(np.float32)
    img[5, :20] = 0.0          # black run
    img[7, :] = [1.5, 2.5, 3.5]  # constant row (runs in every channel)
    p = tmp_path / "rt.hdr"
    imageio.save_hdr(p, img)
    back = imageio.load_hdr(p)[..., :3]
    m = img.max(axis=2, keepdims=True)
    # RGBE keeps 8 mantissa bits of the largest channel: error <= 2^-7 of it
    assert np.all(np.abs(back - img) <= m * 2.0 ** -7 + 1e-30)
    assert np.all(back[5, :20] == 0)


def test_pfm_exact_round_trip_and_orientation(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.normal(size=(5, 7, 3)).astype(np.float32)
    p = tmp_path / "a.pfm"
    imageio.save_pfm(p, img)
    raw = p.read_bytes()
    head = b"PF\n7 5\n-1.0\n"
    assert raw.startswith(head)
    first = np.frombuffer(raw[len(head):len(head) + 12], "<f4")
    assert np.array_equal(first, img[-1, 0])  # bottom row first
    back = imageio.load_pfm(p)
    assert np.array_equal(back[..., :3], img) and np.all(back[..., 3] == 1)


def test_exr_structure_and_pixels(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.normal(size=(4, 6, 4)).astype(np.float32)
    p = tmp_path / "a.exr"
    imageio.save_exr(p, img)
    b = p.read_bytes()
    assert struct.unpack("<II", b[:8]) == (20000630, 2)
    # attributes: name\0 type\0 size value
    pos, attrs = 8, {}
    while b[pos] != 0:
        n = b.index(0, pos)
        name = b[pos:n].decode()
        t = b.index(0, n + 1)
        typ = b[n + 1:t].decode()
        size = struct.unpack("<I", b[t + 1:t + 5])[0]
        attrs[name] = (typ, b[t + 5:t + 5 + size])
        pos = t + 5 + size
    pos += 1
    assert attrs["compression"] == ("compression", b"\x00")
    assert struct.unpack("<4i", attrs["dataWindow"][1]) == (0, 0, 5, 3)
    chl = attrs["channels"][1]
    names = [c.decode() for c in chl.split(b"\x00") if c in (b"A", b"B", b"G", b"R")]
    assert names == ["A", "B", "G", "R"]
    offsets = struct.unpack("<4Q", b[pos:pos + 32])
    for y, off in enumerate(offsets):
        ly, size = struct.unpack("<ii", b[off:off + 8])
        assert ly == y and size == 6 * 16
        data = np.frombuffer(b[off + 8:off + 8 + size], "<f4").reshape(4, 6)  # A, B, G, R planes
        assert np.array_equal(data[3], img[y, :, 0]) and np.array_equal(data[0], img[y, :, 3])


def test_bad_files_raise(tmp_path):
    p = tmp_path / "x.hdr"
    p.write_bytes(b"P6\n1 1\n255\n\x00\x00\x00")
    with pytest.raises(OSError, match="Radiance"):
        imageio.load_hdr(p)
    p.write_bytes(b"#?RADIANCE\n\n+Y 1 +X 1\n\x00\x00\x00\x00")
    with pytest.raises(OSError, match="orientation"):
        imageio.load_hdr(p)
    with pytest.raises(OSError):
        imageio.load_pfm(tmp_path / "missing.pfm")
