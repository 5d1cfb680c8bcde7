"""PFM image files (the on-disk format of rendered images).

``write_pfm`` writes what the reference's write_image_pfm
(core/imageio.rs:264-310) writes: the header ``PF\\n<w> <h>\\n-1\\n`` (three
channels; a negative scale marks little-endian floats), then the float32
pixels row by row from the bottom row of the image to the top, left to right.
``read_pfm`` reads three-channel ``PF`` and one-channel ``Pf`` files of either
byte order and applies |scale|, as pbrt-v3's ReadImagePFM does (the
reference's read_image_pfm, imageio.rs:119-240, swaps the two channel counts
and its header-word reader never advances; the netpbm PFM definition is
followed here instead).
"""
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
