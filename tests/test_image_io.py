"""Image files written by mrh_image_write, read back by independent parsers
that follow the published formats: OpenEXR 2.0 (magic, version flags, the
attribute list with its required attributes, the scanline offset table,
per-line blocks of channels sorted by name) and PFM ("PF", width height,
negative scale = little endian, rows bottom to top)."""
import struct

import numpy as np
import pytest

from meturay_amd import native as N


def read_exr(path):
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630
    assert version & 0xFF == 2 and not (version & 0x200)   # version 2, scanline (not tiled)
    pos = 8
    attrs = {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        pos = end + 1
        end = data.index(b"\0", pos)
        typ = data[pos:end].decode()
        pos = end + 1
        (size,) = struct.unpack_from("<i", data, pos)
        pos += 4
        attrs[name] = (typ, data[pos:pos + size])
        pos += size
    pos += 1
    for req in ["channels", "compression", "dataWindow", "displayWindow", "lineOrder", "pixelAspectRatio",
                "screenWindowCenter", "screenWindowWidth"]:
        assert req in attrs, req
    assert attrs["compression"] == ("compression", b"\0")
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chans = []
    raw = attrs["channels"][1]
    p = 0
    while raw[p] != 0:
        end = raw.index(b"\0", p)
        name = raw[p:end].decode()
        ptype, = struct.unpack_from("<i", raw, end + 1)
        xs, ys = struct.unpack_from("<ii", raw, end + 1 + 8)
        assert (xs, ys) == (1, 1)
        chans.append((name, ptype))
        p = end + 1 + 16
    assert [c[0] for c in chans] == sorted(c[0] for c in chans)
    offsets = struct.unpack_from(f"<{h}Q", data, pos)
    img = {c[0]: np.zeros((h, w), np.float32) for c in chans}
    for off in offsets:
        y, size = struct.unpack_from("<ii", data, off)
        q = off + 8
        for name, ptype in chans:
            assert ptype == 2                      # FLOAT
            img[name][y - y0] = np.frombuffer(data, "<f4", w, q)
            q += 4 * w
        assert q - off - 8 == size
    return np.stack([img["R"], img["G"], img["B"]], axis=2)   # top row first


def read_pfm(path):
    data = open(path, "rb").read()
    lines = data.split(b"\n", 3)
    assert lines[0] == b"PF"
    w, h = (int(v) for v in lines[1].split())
    scale = float(lines[2])
    dt = "<f4" if scale < 0 else ">f4"
    return np.frombuffer(lines[3], dt, w * h * 3).reshape(h, w, 3)   # bottom row first


@pytest.fixture
def image():
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 10, (7, 13, 3)).astype(np.float32)
    img[0, 0] = [0.0, np.inf, -1.5]
    return img


def test_exr_round_trip(tmp_path, image):
    p = str(tmp_path / "a.exr")
    N.write_image(p, image)
    back = read_exr(p)
    assert np.array_equal(back[::-1], image)               # the file's first line is the top row


def test_pfm_round_trip(tmp_path, image):
    p = str(tmp_path / "a.pfm")
    N.write_image(p, image)
    assert np.array_equal(read_pfm(p), image)


def test_image_errors(tmp_path, image):
    with pytest.raises(N.MrhError):
        N.write_image(str(tmp_path / "missing" / "a.exr"), image)
    with pytest.raises(KeyError):
        N.write_image(str(tmp_path / "a.png"), image)
    assert N.lib().mrh_image_write(str(tmp_path / "b.exr").encode(), 5, 2, 2, image.ctypes.data) != 0
