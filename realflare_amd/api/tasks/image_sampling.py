"""Light samples of an image-based light (ImageSamplingTask,
realflare/api/tasks/preprocessing.py:106-185).

The image is read, converted to float (8-bit images scaled to [0, 1], other
depths kept, as the reference does), resized to the sample resolution with
bilinear interpolation, and every pixel whose mean intensity is not above the
(1 - samples / pixels) percentile is cleared; the result is RGBA with alpha 0.
The reference reads and resizes with OpenCV (cv2.imread, cv2.resize
INTER_LINEAR); here PIL reads PNG / JPEG / TIFF and the resize restates
OpenCV's INTER_LINEAR arithmetic for float images in numpy.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

from .. import storage
from ..data import Project, RealflareError


@lru_cache(1)
def load_file(path: str) -> np.ndarray:
    """RGB float32 array of an image file (preprocessing.py:110-126)."""
    try:
        from PIL import Image as PILImage

        with PILImage.open(path) as im:
            if im.mode in ('I;16', 'I;16B', 'I;16L', 'I', 'F'):
                a = np.asarray(im)
                a = np.repeat(a[..., None], 3, axis=2)
            else:
                a = np.asarray(im.convert('RGB'))
    except (OSError, ValueError) as e:
        raise RealflareError(f'Invalid Image path for flare light: {path}') from e
    if a.dtype == np.uint8:
        a = np.divide(a, 255)
    return np.float32(a)


def resize_linear(a: np.ndarray, width: int, height: int) -> np.ndarray:
    """Bilinear resize with pixel-centre alignment and edge clamping, the
    arithmetic of OpenCV's INTER_LINEAR for float images."""
    sh, sw = a.shape[:2]

    def axis(n_dst, n_src):
        scale = n_src / n_dst
        f = ((np.arange(n_dst) + 0.5) * scale - 0.5).astype(np.float32)
        i0 = np.floor(f).astype(np.int64)
        t = (f - i0).astype(np.float32)
        low = i0 < 0
        t[low], i0[low] = 0, 0
        high = i0 >= n_src - 1
        t[high], i0[high] = 0, n_src - 1
        i1 = np.minimum(i0 + 1, n_src - 1)
        return i0, i1, t

    y0, y1, ty = axis(height, sh)
    x0, x1, tx = axis(width, sw)
    tx = tx[None, :, None]
    ty = ty[:, None, None]
    top = a[y0][:, x0] * (1 - tx) + a[y0][:, x1] * tx
    bottom = a[y1][:, x0] * (1 - tx) + a[y1][:, x1] * tx
    return np.float32(top * (1 - ty) + bottom * ty)


@lru_cache(1)
def sample_data(path: str, width: int, height: int, samples: int) -> np.ndarray:
    """preprocessing.py:128-160."""
    array = resize_linear(load_file(path), width, height)
    intensity = np.mean(array, axis=2, keepdims=True).ravel()
    try:
        threshold = 1 - (samples / len(intensity))
    except ZeroDivisionError:
        threshold = 0
    threshold = np.clip(threshold * 100, 0, 100)
    percentile = np.percentile(intensity, threshold)
    mask = intensity > percentile
    masked = np.reshape(array.copy(), (-1, 3))
    masked[~mask] = np.zeros((3,), np.float32)
    masked = masked.reshape(array.shape)
    return np.float32(np.dstack((masked, np.zeros(masked.shape[:-1]))))


def sample_resolution(project: Project) -> tuple[int, int]:
    """preprocessing.py:169-179: even width from image_sample_resolution,
    height from the render aspect ratio, made even."""
    width = max(project.flare.light.image_sample_resolution, 1)
    if width % 2 != 0:
        width += 1
    resolution = project.render.resolution
    height = width * resolution.height() / resolution.width()
    if height % 2 != 0:
        height += 1
    return width, int(height)


class ImageSamplingTask:
    def run(self, project: Project) -> np.ndarray:
        path = storage.decode_path(project.flare.light.image_file)
        width, height = sample_resolution(project)
        return sample_data(path, width, height, project.flare.light.image_samples)
