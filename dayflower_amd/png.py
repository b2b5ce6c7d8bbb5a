"""PNG output of a rendered image (the reference saves through ImageIO:
Image.save, M/image/Image.java:450-466, called by FileRendererObserver,
M/renderer/observer/FileRendererObserver.java:313).

An 8-bit RGBA (colour type 6) or RGB (colour type 2) PNG: one IDAT chunk of
zlib-compressed scanlines, each with filter type 0 (none)."""
import struct
import zlib

import numpy as np


def _chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def encode_png(pixels):
    """PNG bytes of an (H, W, 4) or (H, W, 3) uint8 array, row 0 on top."""
    a = np.ascontiguousarray(pixels)
    if a.dtype != np.uint8 or a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError("pixels must be an (H, W, 3) or (H, W, 4) uint8 array")
    h, w, ch = a.shape
    if h < 1 or w < 1:
        raise ValueError("empty image")
    raw = np.empty((h, 1 + w * ch), np.uint8)
    raw[:, 0] = 0
    raw[:, 1:] = a.reshape(h, w * ch)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6 if ch == 4 else 2, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(raw.tobytes(), 6)) +
            _chunk(b"IEND", b""))


def write_png(path, pixels):
    """Writes `pixels` (see encode_png) to `path`, creating parent folders as
    Image.save does."""
    import os
    parent = os.path.dirname(os.path.abspath(path))
    os.makedirs(parent, exist_ok=True)
    with open(path, "wb") as f:
        f.write(encode_png(pixels))


def float_to_bytes(rgb):
    """Encoded colour floats -> bytes: clamp to [0, 1], scale by 255 and round
    half up (the byte step of AbstractImageKernel.imageEnd)."""
    v = np.clip(np.asarray(rgb, np.float32), 0.0, 1.0) * np.float32(255.0) + np.float32(0.5)
    return v.astype(np.int32).astype(np.uint8)
