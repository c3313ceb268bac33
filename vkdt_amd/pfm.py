"""Reader for the PFM files o-pfm writes (header padded to 16 bytes,
little-endian RGB float32, rows top to bottom)."""
import numpy as np


def read_pfm(path):
    """returns (H x W x 3 float32 array, byte offset of the pixel data)"""
    with open(path, "rb") as f:
        data = f.read()
    if not data.startswith(b"PF\n"):
        raise ValueError("not a colour PFM")
    nl1 = data.index(b"\n", 3)
    w, h = (int(v) for v in data[3:nl1].split())
    nl2 = data.index(b"\n", nl1 + 1)
    if not data[nl1 + 1:nl2].startswith(b"-"):   # negative scale: little endian
        raise ValueError("big-endian PFM not supported")
    off = nl2 + 1
    img = np.frombuffer(data, dtype="<f4", count=w * h * 3, offset=off).reshape(h, w, 3)
    return img, off
