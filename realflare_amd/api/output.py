"""Writing rendered images (Engine.write_image / write_array,
realflare/api/engine.py:239-291).

The reference converts from the render space (ACES - ACEScg) with OCIO and
writes BGR with OpenCV.  Neither library is available here: a colorspace
other than the render space is reported and the array is written unconverted
(what the reference does when its OCIO processor cannot be built,
utils/ocio.py:91-105).  `.exr` files are written by `write_exr` (OpenEXR
scanline file, no compression, 32-bit float B, G, R channels: the three
channels cv2.imwrite stores for a float32 BGR array); other extensions go
through PIL with OpenCV's float -> 8-bit saturate-cast (no scaling).
"""
from __future__ import annotations

import logging
import os
import struct

import numpy as np

from .data import RealflareError

logger = logging.getLogger(__name__)

RENDER_SPACE = 'ACES - ACEScg'


def _attr(name: str, typ: str, data: bytes) -> bytes:
    return name.encode() + b'\0' + typ.encode() + b'\0' + struct.pack('<i', len(data)) + data


def write_exr(path: str, channels: dict[str, np.ndarray]) -> None:
    """Single-part scanline OpenEXR, NO_COMPRESSION, FLOAT channels.
    `channels` maps names to equal-shape (height, width) float arrays."""
    names = sorted(channels)
    h, w = channels[names[0]].shape
    chlist = b''.join(n.encode() + b'\0' + struct.pack('<iB3xii', 2, 0, 1, 1) for n in names)
    chlist += b'\0'
    box = struct.pack('<iiii', 0, 0, w - 1, h - 1)
    header = b''.join([
        _attr('channels', 'chlist', chlist),
        _attr('compression', 'compression', b'\0'),
        _attr('dataWindow', 'box2i', box),
        _attr('displayWindow', 'box2i', box),
        _attr('lineOrder', 'lineOrder', b'\0'),
        _attr('pixelAspectRatio', 'float', struct.pack('<f', 1.0)),
        _attr('screenWindowCenter', 'v2f', struct.pack('<ff', 0.0, 0.0)),
        _attr('screenWindowWidth', 'float', struct.pack('<f', 1.0)),
    ]) + b'\0'
    start = 8 + len(header) + 8 * h
    row_bytes = 4 * w * len(names)
    offsets = struct.pack(f'<{h}Q', *[start + y * (8 + row_bytes) for y in range(h)])
    planes = [np.ascontiguousarray(channels[n], dtype='<f4') for n in names]
    with open(path, 'wb') as f:
        f.write(struct.pack('<ii', 20000630, 2))
        f.write(header)
        f.write(offsets)
        for y in range(h):
            f.write(struct.pack('<ii', y, row_bytes))
            for p in planes:
                f.write(p[y].tobytes())


def read_exr(path: str) -> dict[str, np.ndarray]:
    """Reader for the files write_exr produces (uncompressed FLOAT scanlines)."""
    with open(path, 'rb') as f:
        data = f.read()
    magic, version = struct.unpack_from('<ii', data, 0)
    if magic != 20000630:
        raise ValueError('not an OpenEXR file')
    pos = 8
    attrs = {}
    while data[pos] != 0:
        end = data.index(b'\0', pos)
        name = data[pos:end].decode()
        pos = end + 1
        end = data.index(b'\0', pos)
        pos = end + 1
        (size,) = struct.unpack_from('<i', data, pos)
        pos += 4
        attrs[name] = data[pos:pos + size]
        pos += size
    pos += 1
    names, cl = [], attrs['channels']
    i = 0
    while cl[i] != 0:
        end = cl.index(b'\0', i)
        names.append(cl[i:end].decode())
        i = end + 1 + 16
    if attrs['compression'] != b'\0':
        raise ValueError('compressed EXR files are not supported')
    x0, y0, x1, y1 = struct.unpack('<iiii', attrs['dataWindow'])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f'<{h}Q', data, pos)
    out = {n: np.empty((h, w), np.float32) for n in names}
    for off in offsets:
        y, size = struct.unpack_from('<ii', data, off)
        row = np.frombuffer(data, '<f4', count=size // 4, offset=off + 8).reshape(len(names), w)
        for k, n in enumerate(names):
            out[n][y - y0] = row[k]
    return out


def write_array(array: np.ndarray, filename: str, colorspace: str) -> None:
    """engine.py:270-291."""
    array = np.array(array, np.float32, copy=True)
    if colorspace and colorspace != RENDER_SPACE:
        logger.warning('No OCIO: %s written in %s, not converted to %s', filename, RENDER_SPACE,
                       colorspace)
    try:
        if not os.path.isdir(os.path.dirname(filename)):
            os.makedirs(os.path.dirname(filename))
        ext = os.path.splitext(filename)[1].lower()
        if ext == '.exr':
            write_exr(filename, {'R': array[..., 0], 'G': array[..., 1], 'B': array[..., 2]})
        else:
            from PIL import Image as PILImage

            rgb = np.clip(np.rint(array[..., :3]), 0, 255).astype(np.uint8)
            PILImage.fromarray(rgb).save(filename)
        logger.info('image written: %s', filename)
    except (OSError, ValueError, KeyError) as e:
        logger.debug(e)
        raise RealflareError(f'Error writing file: {filename}') from None
