"""Film::write_image's file output (film.rs:15-33, 316-354) on the CPU: an rgb
image written as .pfm reads back bit for bit; as .ppm and .png it reads back
as write_img's 8-bit encoding, which is pinned against the formula evaluated
here (float32 power as Rust's f32::powf calls the C library's powf, and a
float64 evaluation within one step), with its clamps, NaN and rounding cases.
The PNG is decoded with Python's zlib and its chunk CRCs checked."""
import struct
import zlib

import numpy as np
import pytest

import pbrt_rust_amd as px

F32 = np.float32


def image(seed=3, h=37, w=53):
    rng = np.random.default_rng(seed)
    rgb = rng.uniform(0, 1.2, (h, w, 3)).astype(F32)
    special = np.array([[0.0, 1.0, 2.0], [np.nan, np.inf, -0.0], [1e-30, 0.5, 0.999]], F32)
    flat = rgb.reshape(-1, 3)
    k = min(len(flat), 3)
    flat[:k] = special[:k]
    if h > 1 and w >= 8:
        rgb[1, :8, 0] = np.float32(1.0) - np.arange(8, dtype=F32) * np.float32(2 ** -24)
    return rgb


def want8(rgb):
    """write_img's to_byte: (255 * p.powf(1/2.2) + 0.5).clamp(0, 255) as u8"""
    with np.errstate(invalid="ignore"):
        v = F32(255.0) * np.power(rgb.astype(F32), F32(1.0) / F32(2.2)) + F32(0.5)
    v = np.where(np.isnan(v), F32(0), np.clip(v, F32(0), F32(255)))
    return v.astype(np.uint8)


def test_encoding_matches_write_img():
    rgb = image()
    got = px.encode8(rgb)
    assert np.array_equal(got, want8(rgb))
    # the float64 evaluation of the same formula differs by at most one step
    f64 = np.clip(255.0 * np.power(np.nan_to_num(rgb.astype(np.float64), nan=0.0, posinf=1e9), 1 / 2.2) + 0.5,
                  0, 255).astype(np.int64)
    assert np.max(np.abs(f64 - got.astype(np.int64))) <= 1
    assert got[0, 0].tolist() == [0, 255, 255] and got[0, 1].tolist() == [0, 255, 0]


def read_pfm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = map(int, parts[1].split())
    assert float(parts[2]) < 0  # little endian
    a = np.frombuffer(parts[3], "<f4").reshape(h, w, 3)
    return a[::-1]  # stored bottom to top


def test_pfm_round_trip(tmp_path):
    rgb = image()
    f = str(tmp_path / "out.pfm")
    px.write_image(f, rgb)
    back = read_pfm(f)
    assert np.array_equal(back.view(np.uint32), rgb.view(np.uint32))


def test_ppm_round_trip(tmp_path):
    rgb = image(5)
    f = str(tmp_path / "out.ppm")
    px.write_image(f, rgb)
    data = open(f, "rb").read()
    head = b"P6\n%d %d\n255\n" % (rgb.shape[1], rgb.shape[0])
    assert data.startswith(head)
    back = np.frombuffer(data[len(head):], np.uint8).reshape(rgb.shape)
    assert np.array_equal(back, want8(rgb))


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, color, comp, filt, inter = ihdr
    assert (depth, color, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(raw[:, 0] == 0)  # filter type 0 on every row
    return raw[:, 1:].reshape(h, w, 3)


@pytest.mark.parametrize("shape", [(37, 53), (1, 1), (300, 250)])
def test_png_round_trip(tmp_path, shape):
    # 300 x 250 spans several stored deflate blocks (65535 bytes each)
    rgb = image(7, *shape)
    f = str(tmp_path / "out.png")
    px.write_image(f, rgb)
    assert np.array_equal(read_png(f), want8(rgb))


def test_bad_names_and_state(tmp_path):
    rgb = image()
    with pytest.raises(px.PbrtxError):
        px.write_image(str(tmp_path / "out.exr"), rgb)
    with pytest.raises(px.PbrtxError):
        px.write_image(str(tmp_path / "missing_dir" / "out.png"), rgb)
