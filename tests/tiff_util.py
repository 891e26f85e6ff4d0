"""Uncompressed single-channel FP32 TIFF read/write for test datasets."""
import struct

import numpy as np


def write_f32(path, img):
    img = np.ascontiguousarray(img, dtype="<f4")
    h, w = img.shape
    tags = [(256, 4, w), (257, 4, h), (258, 3, 32), (259, 3, 1), (262, 3, 1), (273, 4, 0),
            (277, 3, 1), (278, 4, h), (279, 4, img.nbytes), (284, 3, 1), (339, 3, 3)]
    data_off = 8 + 2 + 12 * len(tags) + 4
    out = bytearray(b"II" + struct.pack("<HI", 42, 8) + struct.pack("<H", len(tags)))
    for tag, typ, val in tags:
        if tag == 273:
            val = data_off
        if typ == 3:
            out += struct.pack("<HHIHH", tag, typ, 1, val, 0)
        else:
            out += struct.pack("<HHII", tag, typ, 1, val)
    out += struct.pack("<I", 0)
    with open(path, "wb") as f:
        f.write(bytes(out) + img.tobytes())


def read_f32(path):
    b = open(path, "rb").read()
    assert b[:2] == b"II"
    ifd = struct.unpack_from("<I", b, 4)[0]
    n = struct.unpack_from("<H", b, ifd)[0]
    vals = {}
    for k in range(n):
        tag, typ, cnt = struct.unpack_from("<HHI", b, ifd + 2 + 12 * k)
        v = struct.unpack_from("<H" if typ == 3 else "<I", b, ifd + 2 + 12 * k + 8)[0]
        vals[tag] = v
    w, h, off = vals[256], vals[257], vals[273]
    return np.frombuffer(b, dtype="<f4", count=w * h, offset=off).reshape(h, w).copy()
