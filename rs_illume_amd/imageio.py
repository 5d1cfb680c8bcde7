"""Image files of rendered images: PFM (float) and the 8-bit PNG and BMP.

``write_pfm`` writes what the reference's write_image_pfm
(core/imageio.rs:264-310) writes: the header ``PF\\n<w> <h>\\n-1\\n`` (three
channels; a negative scale marks little-endian floats), then the float32
pixels row by row from the bottom row of the image to the top, left to right.
``read_pfm`` reads three-channel ``PF`` and one-channel ``Pf`` files of either
byte order and applies |scale|, as pbrt-v3's ReadImagePFM does (the
reference's read_image_pfm, imageio.rs:119-240, swaps the two channel counts
and its header-word reader never advances; the netpbm PFM definition is
followed here instead).

``write_image`` picks the format by the file's suffix, as the reference's
write_image (imageio.rs:30-67) does: ``.pfm`` floats as above; ``.png`` and
``.bmp`` 8-bit RGB, each value v becoming
``clamp(255 * gamma_correct(v) + 0.5, 0, 255)`` truncated to a byte
(gamma_correct is the sRGB curve of core/pbrt.rs:176-183; NaN becomes 0 as
Rust's float-to-u8 cast makes it). The reference writes those two through the
image crate; its TGA and EXR writers are empty (imageio.rs:113-117, 312-324),
so those suffixes are refused here.
"""
import struct
import zlib

import numpy as np


def _host(img):
    if hasattr(img, "detach"):
        img = img.detach().cpu().numpy()
    return np.asarray(img)


def pfm_bytes(img):
    """PFM file contents of an (H, W, 3) image whose first row is the top of
    the picture (the film layout), as little-endian float32."""
    a = _host(img)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("write_pfm: need an (H, W, 3) image, got shape %r" % (a.shape,))
    h, w = a.shape[:2]
    head = b"PF\n%d %d\n-1\n" % (w, h)
    return head + np.ascontiguousarray(a[::-1], "<f4").tobytes()


def write_pfm(path, img):
    with open(path, "wb") as f:
        f.write(pfm_bytes(img))


def _word(data, pos):
    while pos < len(data) and data[pos:pos + 1].isspace():
        pos += 1
    start = pos
    while pos < len(data) and not data[pos:pos + 1].isspace():
        pos += 1
    if start == pos:
        raise ValueError("PFM: truncated header")
    return data[start:pos], pos


def parse_pfm(data):
    """(H, W, C) float32 image (top row first) of PFM file contents."""
    tag, pos = _word(data, 0)
    if tag == b"PF":
        ch = 3
    elif tag == b"Pf":
        ch = 1
    else:
        raise ValueError("PFM: unknown header %r" % tag[:8])
    try:
        w, pos = _word(data, pos)
        h, pos = _word(data, pos)
        s, pos = _word(data, pos)
        w, h, s = int(w), int(h), float(s)
    except ValueError:
        raise ValueError("PFM: bad width, height or scale")
    if w <= 0 or h <= 0 or s == 0:
        raise ValueError("PFM: bad width, height or scale")
    pos += 1  # the single whitespace character after the scale
    n = w * h * ch
    if len(data) - pos < 4 * n:
        raise ValueError("PFM: %d floats expected, file holds %d" % (n, (len(data) - pos) // 4))
    px = np.frombuffer(data, "<f4" if s < 0 else ">f4", n, pos).astype(np.float32)
    if abs(s) != 1.0:
        px = px * np.float32(abs(s))
    return px.reshape(h, w, ch)[::-1].copy()


def read_pfm(path):
    with open(path, "rb") as f:
        return parse_pfm(f.read())


def gamma_correct(v):
    """sRGB transfer curve in float32 (core/pbrt.rs:176-183)."""
    v = np.asarray(v, np.float32)
    with np.errstate(invalid="ignore"):
        hi = np.float32(1.055) * np.power(v, np.float32(1.0 / 2.4)) - np.float32(0.055)
    return np.where(v < np.float32(0.0031308), np.float32(12.92) * v, hi).astype(np.float32)


def to_bytes(img):
    """8-bit values of a float image: clamp(255 * gamma_correct(v) + 0.5, 0,
    255) truncated toward zero; NaN -> 0 (imageio.rs:42-55)."""
    x = np.float32(255.0) * gamma_correct(_host(img)) + np.float32(0.5)
    x = np.where(x < 0, np.float32(0), np.where(x > 255, np.float32(255), x))  # num::clamp keeps NaN
    x = np.where(np.isnan(x), np.float32(0), x)
    return np.trunc(x).astype(np.uint8)


def _rgb8(img):
    a = _host(img)
    if a.ndim != 3 or a.shape[2] != 3 or a.shape[0] == 0 or a.shape[1] == 0:
        raise ValueError("need a non-empty (H, W, 3) image, got shape %r" % (a.shape,))
    return to_bytes(a)


def png_bytes(img):
    """PNG file contents (8-bit RGB, rows top to bottom, no filtering) of an
    (H, W, 3) float image whose first row is the top of the picture."""
    b = _rgb8(img)
    h, w = b.shape[:2]

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    raw = b"".join(b"\x00" + b[y].tobytes() for y in range(h))
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def bmp_bytes(img):
    """BMP file contents (24-bit BGR, rows bottom to top, each padded to 4
    bytes) of an (H, W, 3) float image whose first row is the top."""
    b = _rgb8(img)
    h, w = b.shape[:2]
    stride = (3 * w + 3) & ~3
    rows = np.zeros((h, stride), np.uint8)
    rows[:, :3 * w] = b[::-1, :, ::-1].reshape(h, 3 * w)
    data = rows.tobytes()
    info = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, len(data), 2835, 2835, 0, 0)
    head = struct.pack("<2sIHHI", b"BM", 14 + len(info) + len(data), 0, 0, 14 + len(info))
    return head + info + data


def write_image(path, img):
    """Writes img ((H, W, 3), top row first) in the format of path's suffix:
    .pfm, .png or .bmp."""
    ext = path.rsplit(".", 1)[-1].lower() if "." in path else ""
    if ext == "pfm":
        data = pfm_bytes(img)
    elif ext == "png":
        data = png_bytes(img)
    elif ext == "bmp":
        data = bmp_bytes(img)
    else:
        raise ValueError("write_image: unsupported image suffix %r (pfm, png or bmp)" % ext)
    with open(path, "wb") as f:
        f.write(data)
