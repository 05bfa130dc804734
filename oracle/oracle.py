"""The oracle: CPU restatement of realflare's render path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, as the checker (or the timed CPU baseline);
the render path in ``realflare_amd`` never does.

* The OpenCL kernels are restated in C (``rf_oracle_impl.h``, built in float
  and double into ``_build/librforacle.so`` by the Makefile) and wrapped here
  with ctypes.
* The host-side numpy transforms are restated here in numpy, in complex128
  like the reference: the starburst power spectrum
  (``realflare/api/tasks/starburst.py:45-81``) and the ghost texture's
  fractional Fourier transform (``realflare/api/tasks/ghost.py:13-36``,
  ``realflare/utils/frft.py:8-62``).
* The lens diagram kernels (``realflare/api/tasks/cl/diagram.cl``) are
  restated in float32 numpy, vectorized over pixels.

Pinning: ``tests/test_oracle.py``, ``tests/test_oracle_pins.py`` and
``tests/test_host.py`` check these functions against the reference's test
vectors, closed forms, invariants, brute force or golden fixtures; the header
of each test names which.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from functools import lru_cache

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, '_build', 'librforacle.so')

LAMBDA_MIN, LAMBDA_MAX = 390, 730
LAMBDA_MID = (LAMBDA_MIN + LAMBDA_MAX) / 2


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(['make', '-s', '-C', HERE], check=True)
    return LIB_PATH


@lru_cache(1)
def lib() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    return ctypes.CDLL(LIB_PATH)


def threads() -> int:
    return int(lib().rfo_threads())


def _real(precision: str):
    return np.float32 if precision == 'f32' else np.float64


def ray_dtype(precision: str = 'f32') -> np.dtype:
    r = _real(precision)
    return np.dtype(
        [('pos', r, 3), ('dir', r, 3), ('pos_apt', r, 2), ('rrel', r), ('reflectance', r)]
    )


def inter_dtype(precision: str = 'f32') -> np.dtype:
    r = _real(precision)
    fields = [('pos', r, 3), ('normal', r, 3), ('incident', r), ('hit', np.int32)]
    dt = np.dtype(fields, align=True)
    return dt


def vertex_dtype(precision: str = 'f32') -> np.dtype:
    r = _real(precision)
    return np.dtype(
        [('pos', r, 2), ('uv', r, 2), ('rrel', r), ('reflectance', r), ('intensity', r)]
    )


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _fn(name: str, precision: str):
    return getattr(lib(), f'{name}_{precision}')


def trace(
    lens: np.ndarray,
    paths,
    wavelengths,
    aperture_index: int,
    coating_min_ior: float,
    grid_count: int,
    grid_length: float,
    direction,
    precision: str = 'f32',
    intersections: bool = False,
):
    """``raytrace`` over the whole (P, W, R) NDRange (raytracing.cl:218-359)."""
    lens = np.ascontiguousarray(lens)
    assert lens.dtype.itemsize == 64
    paths = np.ascontiguousarray(np.asarray(paths, np.int32).reshape(-1, 2))
    wavelengths = np.ascontiguousarray(np.asarray(wavelengths, np.int32))
    P, W, R = len(paths), len(wavelengths), grid_count * grid_count
    out = np.zeros((P, W, R), ray_dtype(precision))
    direction = np.ascontiguousarray(np.asarray(direction, np.float64).reshape(4))
    inters = None
    inters_count = 3 * len(lens) - 1
    if intersections:
        inters = np.zeros((P, W, R, inters_count), inter_dtype(precision))
    f = _fn('rfo_trace', precision)
    f.restype = ctypes.c_int
    f(
        _p(lens), ctypes.c_int(len(lens)), _p(paths), ctypes.c_int(P), _p(wavelengths),
        ctypes.c_int(W), ctypes.c_int(aperture_index), ctypes.c_double(coating_min_ior),
        ctypes.c_int(grid_count), ctypes.c_double(grid_length), _p(direction), _p(out),
        _p(inters) if inters is not None else ctypes.c_void_p(0), ctypes.c_int(inters_count),
    )
    if intersections:
        return out, inters
    return out


def prim_shader(rays: np.ndarray, grid_count: int, area_orig: float, min_area: float,
                precision: str = 'f32'):
    """``prim_shader`` (rasterizing.cl:233-320): bounds[P][Q][4], intensities[P][Q][W]."""
    P, W, R = rays.shape
    Q = (grid_count - 1) ** 2
    rays = np.ascontiguousarray(rays)
    bounds = np.zeros((P, Q, 4), _real(precision))
    intens = np.zeros((P, Q, W), _real(precision))
    _fn('rfo_prim_shader', precision)(
        _p(rays), ctypes.c_int(P), ctypes.c_int(W), ctypes.c_int(grid_count),
        ctypes.c_double(area_orig), ctypes.c_double(min_area), _p(bounds), _p(intens),
    )
    return bounds, intens


def vertex_shader(rays, intensities, grid_count, screen_transform, resolution,
                  precision: str = 'f32'):
    """``vertex_shader`` (rasterizing.cl:322-386): vertexes[P][R][W]."""
    P, W, R = rays.shape
    out = np.zeros((P, R, W), vertex_dtype(precision))
    _fn('rfo_vertex_shader', precision)(
        _p(np.ascontiguousarray(rays)), _p(np.ascontiguousarray(intensities)), ctypes.c_int(P),
        ctypes.c_int(W), ctypes.c_int(grid_count), ctypes.c_double(screen_transform),
        ctypes.c_int(resolution[0]), ctypes.c_int(resolution[1]), _p(out),
    )
    return out


def rasterize(vertexes, bounds, wavelength_sub_count, grid_count, anti_aliasing, intensity,
              ghost_scale, ghost, spectrum, screen_transform, bin_size, resolution,
              rows=None, precision: str = 'f32', binned: bool = True):
    """``binner`` + ``rasterizer`` (rasterizing.cl:388-663) into a cleared image.

    ``intensity`` is the value the host passes to the kernel (light intensity
    times 1e3, rasterizing.py:415).  ``rows`` = (y0, y1) restricts the work to
    a band of unflipped rows; pixels outside stay 0.
    """
    P, R, W = vertexes.shape
    width, height = resolution
    ghost = np.ascontiguousarray(ghost, np.float32)
    spectrum = np.ascontiguousarray(spectrum, np.float32)
    out = np.zeros((height, width, 4), _real(precision))
    y0, y1 = rows if rows is not None else (0, height)
    common = (
        _p(np.ascontiguousarray(vertexes)), _p(np.ascontiguousarray(bounds)), ctypes.c_int(P),
        ctypes.c_int(W), ctypes.c_int(wavelength_sub_count), ctypes.c_int(grid_count),
        ctypes.c_int(anti_aliasing), ctypes.c_double(intensity), ctypes.c_double(ghost_scale),
        _p(ghost), ctypes.c_int(ghost.shape[1]), ctypes.c_int(ghost.shape[0]), _p(spectrum),
        ctypes.c_int(spectrum.shape[0]), ctypes.c_double(screen_transform),
        ctypes.c_int(bin_size), ctypes.c_int(width), ctypes.c_int(height),
    )
    if binned:
        _fn('rfo_rasterize', precision)(*common, ctypes.c_int(y0), ctypes.c_int(y1), _p(out))
    else:
        _fn('rfo_rasterize_unbinned', precision)(*common, _p(out))
    return out


def aperture_shape(resolution, size, blades, rotation, roundness, softness,
                   precision: str = 'f32'):
    """``aperture_shape`` (aperture.cl:5-47); rotation in radians, softness as
    passed to the kernel (the project value / 10, aperture.py:96)."""
    w, h = resolution
    out = np.zeros((h, w), _real(precision))
    _fn('rfo_aperture_shape', precision)(
        ctypes.c_int(w), ctypes.c_int(h), ctypes.c_double(size[0]), ctypes.c_double(size[1]),
        ctypes.c_int(blades), ctypes.c_double(rotation), ctypes.c_double(roundness),
        ctypes.c_double(softness), _p(out),
    )
    return out


def aperture_grating(img, size, strength, density, length, width, softness,
                     precision: str = 'f32'):
    h, w = img.shape
    img = np.ascontiguousarray(img, _real(precision)).copy()
    _fn('rfo_aperture_grating', precision)(
        ctypes.c_int(w), ctypes.c_int(h), ctypes.c_double(size[0]), ctypes.c_double(size[1]),
        ctypes.c_double(strength), ctypes.c_double(density), ctypes.c_double(length),
        ctypes.c_double(width), ctypes.c_double(softness), _p(img),
    )
    return img


def aperture_scratches(img, size, strength, density, length, rotation, rotation_variation,
                       width, parallax, softness, precision: str = 'f32'):
    h, w = img.shape
    img = np.ascontiguousarray(img, _real(precision)).copy()
    _fn('rfo_aperture_scratches', precision)(
        ctypes.c_int(w), ctypes.c_int(h), ctypes.c_double(size[0]), ctypes.c_double(size[1]),
        ctypes.c_double(strength), ctypes.c_double(density), ctypes.c_double(length),
        ctypes.c_double(rotation), ctypes.c_double(rotation_variation), ctypes.c_double(width),
        ctypes.c_double(parallax[0]), ctypes.c_double(parallax[1]), ctypes.c_double(softness),
        _p(img),
    )
    return img


def aperture_dust(img, size, strength, density, radius, parallax, softness,
                  precision: str = 'f32'):
    h, w = img.shape
    img = np.ascontiguousarray(img, _real(precision)).copy()
    _fn('rfo_aperture_dust', precision)(
        ctypes.c_int(w), ctypes.c_int(h), ctypes.c_double(size[0]), ctypes.c_double(size[1]),
        ctypes.c_double(strength), ctypes.c_double(density), ctypes.c_double(radius),
        ctypes.c_double(parallax[0]), ctypes.c_double(parallax[1]), ctypes.c_double(softness),
        _p(img),
    )
    return img


def starburst(resolution, spectrum_img, light, samples, blur, rotation, rotation_weight,
              vignetting, intensity, offset, size, precision: str = 'f32'):
    """``starburst`` kernel (starburst.cl:1-87) with the host's argument
    preparation already applied (starburst.py:96-106)."""
    w, h = resolution
    spectrum_img = np.ascontiguousarray(spectrum_img, np.float32)
    light = np.ascontiguousarray(light, np.float32)
    out = np.zeros((h, w, 4), _real(precision))
    _fn('rfo_starburst', precision)(
        ctypes.c_int(w), ctypes.c_int(h), _p(spectrum_img), ctypes.c_int(spectrum_img.shape[1]),
        ctypes.c_int(spectrum_img.shape[0]), _p(light), ctypes.c_int(light.shape[0]),
        ctypes.c_int(samples), ctypes.c_double(blur), ctypes.c_double(rotation),
        ctypes.c_double(rotation_weight), ctypes.c_double(vignetting[0]),
        ctypes.c_double(vignetting[1]), ctypes.c_double(intensity), ctypes.c_double(offset[0]),
        ctypes.c_double(offset[1]), ctypes.c_double(size[0]), ctypes.c_double(size[1]), _p(out),
    )
    return out


# ---------------------------------------------------------------------------
# host-side numpy restatements
# ---------------------------------------------------------------------------

def cie_table() -> np.ndarray:
    path = os.path.join(os.path.dirname(HERE), 'realflare_amd', 'resources', 'ciexyz_390_790.csv')
    return np.loadtxt(path, delimiter=',', comments='#')


def light_spectrum() -> np.ndarray:
    """CIE XYZ rows with 390 <= lambda < 730 as RGBA (rasterizing.py:169-179)."""
    t = cie_table()
    sel = t[(t[:, 0] >= LAMBDA_MIN) & (t[:, 0] < LAMBDA_MAX)]
    out = np.zeros((len(sel), 4), np.float32)
    out[:, :3] = sel[:, 1:4]
    return out


def power_spectrum(aperture: np.ndarray, distance: float) -> np.ndarray:
    """starburst.py:45-81: |fftshift(fft2(A * exp(i pi/(lambda d) (x^2+y^2))))|^2."""
    wavelength = LAMBDA_MID * 1e-6
    distance = max(1e-9, distance * 1e3)
    h, w = aperture.shape[:2]
    x = np.linspace(-1, 1, w)
    y = np.linspace(-1, 1, h)
    xv, yv = np.meshgrid(x, y)
    exp = np.exp(1j * np.pi / (wavelength * distance) * (xv**2 + yv**2))
    fft = np.fft.fftshift(np.fft.fft2(aperture * exp))
    return np.float32(fft.real**2 + fft.imag**2)


def _chirp(shape) -> np.ndarray:
    """frft.py:8-17: i*pi*sum_k fftshift(n_k)^2 / N_k over an index grid."""
    grids = np.mgrid[tuple(slice(-s // 2, s // 2) for s in shape)]
    if len(shape) == 1:
        grids = [grids[0]] if grids.ndim == 2 else [grids]
    total = 0
    for i, g in enumerate(grids):
        total = total + np.fft.fftshift(g) ** 2 / g.shape[i]
    return 1.0j * np.pi * total


def _normalize(array: np.ndarray, alpha: float):
    """frft.py:20-40: reduce alpha into [0.5, 1.5) with ortho (i)FFTs / flips."""
    rel = alpha % 4
    if rel < 0.5:
        return np.fft.ifftn(array, norm='ortho'), rel + 1
    if rel < 1.5:
        return array, rel
    if rel < 2.5:
        return np.fft.fftn(array, norm='ortho'), rel - 1
    if rel < 3.5:
        return np.flip(array), rel - 2
    return np.fft.ifftn(array, norm='ortho'), rel - 3


def frft(array: np.ndarray, alpha: float) -> np.ndarray:
    """frft.py:43-62: chirp-convolution discrete fractional Fourier transform."""
    array, alpha = _normalize(np.asarray(array), alpha)
    phi = alpha * np.pi / 2.0
    cot = 1.0 / np.tan(phi)
    sq = np.sqrt(1.0 + cot**2)
    scale = np.sqrt(1.0 - 1.0j * cot) / np.sqrt(np.prod(array.shape))
    arg = _chirp(array.shape)
    chirp1 = np.exp(arg * (cot - sq))
    chirp2 = np.exp(arg * sq)
    f1 = np.fft.fftn(chirp2, norm='ortho')
    f2 = np.fft.fftn(chirp1 * array, norm='ortho')
    return scale * chirp1 * np.fft.ifftn(f1 * f2, norm='ortho')


def ghost(aperture: np.ndarray, fstop: float, resolution) -> np.ndarray:
    """ghost.py:13-36: |fftshift(frft(fftshift(A), alpha))| * sqrt(w*h)."""
    alpha = 0.15 * (LAMBDA_MID / 400) * (fstop / 18)
    spectrum = np.fft.fftshift(aperture)
    spectrum = frft(spectrum, alpha)
    spectrum = np.fft.fftshift(spectrum)
    array = np.float32(np.abs(spectrum))
    array *= np.sqrt(resolution[0] * resolution[1])
    return array


def _trunc_eq(a, b):
    """(int) a == (int) b for values inside the int range (C truncation)."""
    return np.trunc(a) == np.trunc(b)


def diagram_lenses(radius, height, center, aperture_index: int, scale: float, width: int,
                   height_px: int) -> np.ndarray:
    """diagram.cl:21-68 (`lenses`), in float32 like the kernel, vectorized over
    pixels: axis row red, flat elements a column (stop and last element green,
    others blue, and the element loop stops there), spherical elements the
    ring trunc(|r|) == trunc(distance) on one half, within the element height.
    ``pow(v, 2)`` is written v * v (the correctly rounded square)."""
    f32 = np.float32
    s = f32(scale)
    x = np.broadcast_to(np.arange(width).astype(f32)[None, :], (height_px, width))
    y = np.broadcast_to((height_px // 2 - np.arange(height_px)).astype(f32)[:, None],
                        (height_px, width))
    rgba = np.zeros((height_px, width, 4), f32)
    rgba[..., 0] = np.where(_trunc_eq(y, f32(0)), f32(1), f32(0))
    done = np.zeros((height_px, width), bool)
    n = len(radius)
    for i in range(n):
        r = f32(radius[i]) * s
        h = f32(height[i]) * s
        c = f32(center[i]) * s
        flat = ~done & (r == 0) & (np.abs(y) < h) & _trunc_eq(x, c)
        rgba[flat, 1 if (i == aperture_index or i == n - 1) else 2] = 1
        dx = x - c
        r_circle = np.sqrt(dx * dx + y * y)
        half = (x < c) if r > 0 else (x > c)
        ring = ~done & ~flat & half & _trunc_eq(np.abs(r), r_circle) & (np.abs(y) < h)
        rgba[ring, 2] = 1
        done |= flat
    return rgba


def diagram_is_line(px, py, x0, y0, x1, y1, width: float = 1.0):
    """diagram.cl:1-19 in float32: strictly between the end caps and within
    width / 2 of the line."""
    f32 = np.float32
    with np.errstate(divide='ignore', invalid='ignore', over='ignore'):
        x0, y0, x1, y1 = f32(x0), f32(y0), f32(x1), f32(y1)
        vx, vy = x1 - x0, y1 - y0
        ax, ay = px - x0, py - y0
        bx, by = px - x1, py - y1
        between = ((vx * ax + vy * ay) > 0) & ((vx * bx + vy * by) < 0)
        d = np.abs(vx * ay - ax * vy) / np.sqrt(vx * vx + vy * vy)
        return between & (d <= f32(width) / f32(2))


def diagram_intersections(image: np.ndarray, pos, normal, hit, records_per_ray: int,
                          ray_count: int, scale: float) -> np.ndarray:
    """diagram.cl:71-138 (`intersections`) over a contiguous slice of records
    [ray][record] (fields pos (N, 3), normal (N, 3), hit (N,)): for each ray in
    order, for each record whose successor is a hit, a normal line from the
    point (-z, -y) * scale overwrites the pixel with (0.25, 0, 0, 0) and, unless
    the successor lies at x == 0, a ray line to the successor adds 0.25 to
    R, G, B.  The successor of a ray's last record is the next ray's first;
    past the slice there is none (the reference reads beyond its buffer)."""
    f32 = np.float32
    s = f32(scale)
    h, w = image.shape[:2]
    out = np.array(image, f32, copy=True)
    px = np.broadcast_to(np.arange(w).astype(f32)[None, :], (h, w))
    py = np.broadcast_to((h // 2 - np.arange(h)).astype(f32)[:, None], (h, w))
    pos = np.asarray(pos, f32)
    normal = np.asarray(normal, f32)
    total = ray_count * records_per_ray
    for ray in range(ray_count):
        for k in range(records_per_ray):
            i1 = ray * records_per_ray + k
            if i1 + 1 >= total or not hit[i1 + 1]:
                break
            p0x, p0y = (pos[i1, 2] * s) * f32(-1), (pos[i1, 1] * s) * f32(-1)
            nx, ny = normal[i1, 2] * s, normal[i1, 1] * s
            pnx, pny = p0x - nx * f32(1), p0y - ny * f32(1)
            m = diagram_is_line(px, py, p0x, p0y, pnx, pny)
            out[m] = (0.25, 0, 0, 0)
            p1x, p1y = (pos[i1 + 1, 2] * s) * f32(-1), (pos[i1 + 1, 1] * s) * f32(-1)
            if p1x == 0:
                break
            m = diagram_is_line(px, py, p0x, p0y, p1x, p1y)
            out[m, :3] += f32(0.25)
    return out


def to_reference_rays(rays: np.ndarray) -> np.ndarray:
    """Repack oracle rays into the reference's 48-byte ``Ray`` records."""
    from realflare_amd.api.dtypes import ray_dtype as ref_ray

    out = np.zeros(rays.shape, ref_ray)
    for name in ('pos', 'dir', 'pos_apt', 'rrel', 'reflectance'):
        out[name] = rays[name]
    return out


def _scalar(name, precision, restype, argtypes):
    f = _fn(name, precision)
    f.restype = restype
    f.argtypes = argtypes
    return f


def fresnel(theta0, n1, n2, precision='f64'):
    """raytracing.cl:1-16."""
    d = ctypes.c_double
    return _scalar('rfo_fresnel', precision, d, [d, d, d])(theta0, n1, n2)


def fresnel_ar(theta0, lam, d1, n0, n1, n2, precision='f64'):
    """raytracing.cl:18-59."""
    d = ctypes.c_double
    return _scalar('rfo_fresnel_ar', precision, d, [d] * 6)(theta0, lam, d1, n0, n1, n2)


def dispersion(lam, coefficients, precision='f64'):
    """raytracing.cl:107-120."""
    c = np.zeros(8, np.float32)
    c[:6] = coefficients
    f = _scalar('rfo_dispersion', precision, ctypes.c_double, [ctypes.c_double, ctypes.c_void_p])
    return f(lam, c.ctypes.data)


def intersect_quad(px, py, xs, ys, precision='f32') -> bool:
    """rasterizing.cl:55-93."""
    x = np.asarray(xs, np.int32)
    y = np.asarray(ys, np.int32)
    f = _scalar('rfo_intersect_quad', precision, ctypes.c_int,
                [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p])
    return bool(f(px, py, x.ctypes.data, y.ctypes.data))


def barycentric_quad(px, py, xs, ys, precision='f32') -> np.ndarray:
    """rasterizing.cl:95-148."""
    x = np.asarray(xs, np.int32)
    y = np.asarray(ys, np.int32)
    w = np.zeros(4, np.float64)
    f = _scalar('rfo_barycentric_quad', precision, None,
                [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p])
    f(px, py, x.ctypes.data, y.ctypes.data, w.ctypes.data)
    return w


def sample_bilinear(tex, s, t, edge=False, precision='f32') -> float:
    """OpenCL 1.2 bilinear filter (normalized coordinates), one channel."""
    tex = np.ascontiguousarray(tex, np.float32)
    f = _scalar('rfo_sample_bilinear', precision, ctypes.c_double,
                [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                 ctypes.c_int])
    return f(tex.ctypes.data, tex.shape[1], tex.shape[0], s, t, int(edge))


def trace_params(radius, ior, height, center, coating, coefficients, disperse, paths, wavelengths,
                 aperture_index, coating_min_ior, grid_count, grid_length, direction,
                 precision='f64'):
    """Trace from element parameters given in double ([L] arrays, coefficients
    [L][6]); used for finite differences of the trace."""
    a = [np.ascontiguousarray(np.asarray(v, np.float64)) for v in
         (radius, ior, height, center, coating)]
    coef = np.ascontiguousarray(np.asarray(coefficients, np.float64).reshape(-1, 6))
    L = len(a[0])
    paths = np.ascontiguousarray(np.asarray(paths, np.int32).reshape(-1, 2))
    wavelengths = np.ascontiguousarray(np.asarray(wavelengths, np.int32))
    P, W, R = len(paths), len(wavelengths), grid_count * grid_count
    out = np.zeros((P, W, R), ray_dtype(precision))
    direction = np.ascontiguousarray(np.asarray(direction, np.float64).reshape(4))
    f = _fn('rfo_trace_params', precision)
    f.restype = ctypes.c_int
    rc = f(*[_p(x) for x in a], _p(coef), ctypes.c_int(L), ctypes.c_int(int(disperse)), _p(paths),
           ctypes.c_int(P), _p(wavelengths), ctypes.c_int(W), ctypes.c_int(aperture_index),
           ctypes.c_double(coating_min_ior), ctypes.c_int(grid_count),
           ctypes.c_double(grid_length), _p(direction), _p(out))
    assert rc == 0
    return out
