"""Parity of the HIP rasterizer (rf_prim_shader, rf_vertex_shader,
rf_rasterize, rf_flare) with the oracle's restatement of rasterizing.cl.

* Same rays in: the GPU trace's planes are fed to both rasterizers (a 48^2
  grid at 512^2, and a 4^2 grid at 1024^2 whose quads each cover thousands of
  tiles).  Bounds,
  intensities and vertices must be bit-identical (same IEEE operations, no
  contraction); the image must match within 1e-6 relative L1 (the oracle
  and the kernel evaluate the same expressions; differences can only come
  from correctly rounded operations the compilers order differently).
* End to end: oracle trace + oracle raster against GPU trace + GPU raster,
  within 1e-3 relative L1 (the reference's image bar); vertex positions
  that differ in the last bit can move a pixel edge.
* Variants: 2x / 4x / 8x anti-aliasing with sub-wavelengths, on bins of
  whole 8x8 tiles (the lane-mask coverage instantiations) and on bins of 36
  and 20 pixels (the generic one), one wavelength (the rasterizer's
  neighbouring-ray quirk), and the pair-capacity overflow/retry path; the
  bin size leaves the image bit-identical at every anti-aliasing level.
"""
import ctypes

import numpy as np
import pytest
import torch

from helpers import trace_inputs
from raster_helpers import (gpu_flare, gpu_trace_planes, make_ghost, oracle_image, params_for,
                            planes_to_oracle_rays, rel_l1)

from realflare_amd.api.color import light_spectrum

pytestmark = pytest.mark.gpu


def _nan_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


@pytest.mark.parametrize('g,res', [(48, 512), (4, 1024)])
def test_raster_stages_same_rays(g, res, oracle_lib):
    from realflare_amd import native

    inp = trace_inputs('small', grid_count=g)
    print(f'case raster/small same-rays g={g} {res}^2', flush=True)
    planes = gpu_trace_planes(inp)
    ghost = make_ghost(oracle_lib)
    prm = params_for(inp, (res, res), ghost.shape)
    rays = planes_to_oracle_rays(planes.cpu().numpy(), oracle_lib)
    ob, oi, ov, oimg = oracle_image(oracle_lib, rays, inp, prm, ghost)

    P, W, R = rays.shape
    Q = (inp['grid_count'] - 1) ** 2
    bounds = torch.empty((P, Q, 4), dtype=torch.float32, device='cuda')
    intens = torch.empty((P, Q, W), dtype=torch.float32, device='cuda')
    vtx = torch.empty((P, R, W, 8), dtype=torch.float32, device='cuda')
    native.call('rf_prim_shader', native.ptr(planes), ctypes.byref(prm), native.ptr(bounds),
                native.ptr(intens), native.stream())
    native.call('rf_vertex_shader', native.ptr(planes), native.ptr(intens), ctypes.byref(prm),
                native.ptr(vtx), native.stream())
    gb, gi, gv = bounds.cpu().numpy(), intens.cpu().numpy(), vtx.cpu().numpy()
    if g == 4:
        # some quad this coarse covers the centres of more than 1024 tiles of
        # 8x8 pixels, the size above which the binner expands a quad step
        # with a whole workgroup
        pos = ov['pos'].reshape(P, g, g, W, 2)
        c = [pos[:, :-1, :-1], pos[:, :-1, 1:], pos[:, 1:, 1:], pos[:, 1:, :-1]]
        valid = ~np.isnan(ob[..., 0]).reshape(P, g - 1, g - 1)
        t = (np.arange(res // 8) * 8 + 4).astype(np.float64)
        tx, ty = np.meshgrid(t, t)

        def side(a, b):
            return ((b[..., 0, None] - a[..., 0, None]) * (ty.ravel() - a[..., 1, None]) -
                    (b[..., 1, None] - a[..., 1, None]) * (tx.ravel() - a[..., 0, None]))

        def inside(a, b, d):
            s1, s2, s3 = side(a, b), side(b, d), side(d, a)
            return ((s1 >= 0) & (s2 >= 0) & (s3 >= 0)) | ((s1 <= 0) & (s2 <= 0) & (s3 <= 0))

        cov = (inside(c[0][..., 0, :], c[1][..., 0, :], c[2][..., 0, :]) |
               inside(c[2][..., 0, :], c[3][..., 0, :], c[0][..., 0, :])).sum(-1)
        assert np.max(np.where(valid, cov, 0)) > 1024
    assert _nan_equal(gb, ob)
    assert np.array_equal(gi, oi)
    for k, name in ((0, 'pos'), (2, 'uv')):
        assert np.array_equal(gv[..., k], ov[name][..., 0]), name
        assert np.array_equal(gv[..., k + 1], ov[name][..., 1]), name
    assert np.array_equal(gv[..., 4], ov['rrel'])
    assert _nan_equal(gv[..., 5], ov['reflectance'])
    assert np.array_equal(gv[..., 6], ov['intensity'])

    img, _ = gpu_flare(planes, ghost, prm)
    err = rel_l1(img, oimg)
    exact = float(np.mean(img == oimg))
    print(f'image rel L1 {err:.3e}, identical values {exact:.6f}, sum {np.abs(oimg).sum():.4e}',
          flush=True)
    assert np.abs(oimg).sum() > 0
    assert err <= 1e-6


@pytest.mark.parametrize('name,g,res', [('small', 64, 512), ('full', 24, 384)])
def test_flare_end_to_end(name, g, res, oracle_lib):
    inp = trace_inputs(name, grid_count=g)
    print(f'case raster/{name} end-to-end g={g} {res}^2', flush=True)
    ghost = make_ghost(oracle_lib)
    prm = params_for(inp, (res, res), ghost.shape)
    ref_rays = oracle_lib.trace(inp['lens'], inp['paths'], inp['wavelengths'],
                                inp['aperture_index'], inp['coating_min_ior'], inp['grid_count'],
                                inp['grid_length'], inp['direction'])
    *_, oimg = oracle_image(oracle_lib, ref_rays, inp, prm, ghost)
    img, _ = gpu_flare(gpu_trace_planes(inp), ghost, prm)
    err = rel_l1(img, oimg)
    print(f'{name}: image rel L1 {err:.3e}', flush=True)
    assert np.abs(oimg).sum() > 0
    assert err <= 1e-3


def test_image_error_against_fp64(oracle_lib):
    """The small configuration end to end (64x64 rays, 15 ghosts, 3
    wavelengths, 512^2): the GPU image and the FP32 oracle image, each
    against the FP64 oracle image (FP64 trace and FP64 rasterizer).  Both
    FP32 renders carry the same kind of error (sensor positions rounded to
    FP32 move coverage edges), so the GPU's error is bounded by the FP32
    oracle's own plus a margin, and by the reference's image bar."""
    inp = trace_inputs('small')
    print(f'case raster/small vs FP64 g={inp["grid_count"]} 512^2', flush=True)
    ghost = make_ghost(oracle_lib)
    prm = params_for(inp, (512, 512), ghost.shape)
    args = (inp['lens'], inp['paths'], inp['wavelengths'], inp['aperture_index'],
            inp['coating_min_ior'], inp['grid_count'], inp['grid_length'], inp['direction'])
    g = inp['grid_count']
    rays64 = oracle_lib.trace(*args, precision='f64')
    b64, i64 = oracle_lib.prim_shader(rays64, g, prm.area_orig, prm.min_area, precision='f64')
    v64 = oracle_lib.vertex_shader(rays64, i64, g, prm.screen_transform, (prm.width, prm.height),
                                   precision='f64')
    img64 = oracle_lib.rasterize(v64, b64, prm.wavelength_sub_count, g, prm.anti_aliasing,
                                 prm.intensity, prm.ghost_scale, ghost, light_spectrum(),
                                 prm.screen_transform, prm.bin_size, (prm.width, prm.height),
                                 precision='f64')
    *_, img32 = oracle_image(oracle_lib, oracle_lib.trace(*args), inp, prm, ghost)
    img, _ = gpu_flare(gpu_trace_planes(inp), ghost, prm)
    e_gpu, e_o32 = rel_l1(img, img64), rel_l1(img32, img64)
    print(f'image rel L1 vs FP64: gpu {e_gpu:.3e}, fp32 oracle {e_o32:.3e}; '
          f'gpu vs fp32 oracle {rel_l1(img, img32):.3e}', flush=True)
    assert np.abs(img64).sum() > 0
    assert e_gpu <= 1e-3
    assert e_gpu <= 2 * e_o32 + 1e-6


@pytest.mark.parametrize('aa,wsub,bin_size,wl', [(4, 3, 64, None), (2, 1, 48, None),
                                                  (1, 1, 64, [550]), (8, 2, 40, [450, 600]),
                                                  (4, 2, 36, None), (1, 1, 20, None)])
def test_raster_variants_same_rays(aa, wsub, bin_size, wl, oracle_lib):
    inp = trace_inputs('small', grid_count=24)
    if wl is not None:
        inp['wavelengths'] = np.array(wl, np.int32)
    print(f'case raster/variant aa={aa} wsub={wsub} bin={bin_size} wl={wl}', flush=True)
    planes = gpu_trace_planes(inp)
    ghost = make_ghost(oracle_lib, res=64)
    prm = params_for(inp, (256, 192), ghost.shape, bin_size=bin_size, aa=aa, wsub=wsub)
    rays = planes_to_oracle_rays(planes.cpu().numpy(), oracle_lib)
    *_, oimg = oracle_image(oracle_lib, rays, inp, prm, ghost)
    img, _ = gpu_flare(planes, ghost, prm)
    err = rel_l1(img, oimg)
    print(f'variant rel L1 {err:.3e}', flush=True)
    assert np.abs(oimg).sum() > 0
    assert err <= 1e-6


@pytest.mark.parametrize('aa', [1, 2, 4, 8])
def test_raster_bin_size_does_not_change_the_image(aa):
    """A pixel adds the fragments of the quads of its bin in binning order and
    a quad that covers none of its samples adds nothing, so the bin size does
    not change the image.  Bins of whole 8x8 tiles (64, 8) run the
    lane-mask coverage instantiation, bins of 36 the generic one: the images
    are identical."""
    inp = trace_inputs('small', grid_count=24)
    print(f'case raster/bin sizes aa={aa}', flush=True)
    planes = gpu_trace_planes(inp)
    ghost = np.random.default_rng(2).random((48, 64)).astype(np.float32)
    imgs = []
    for bin_size in (64, 36, 8):
        prm = params_for(inp, (264, 200), ghost.shape, bin_size=bin_size, aa=aa, wsub=2)
        imgs.append(gpu_flare(planes, ghost, prm)[0])
    assert np.abs(imgs[0]).sum() > 0
    for img in imgs[1:]:
        np.testing.assert_array_equal(img, imgs[0])


def test_raster_capacity_overflow_retry(oracle_lib):
    from realflare_amd.api.tasks.rasterizing import RasterWorkspace, flare
    from realflare_amd.api.color import light_spectrum

    inp = trace_inputs('small', grid_count=24)
    planes = gpu_trace_planes(inp)
    ghost = make_ghost(oracle_lib, res=64)
    prm = params_for(inp, (256, 256), ghost.shape)
    ref, ws_big = gpu_flare(planes, ghost, prm, capacity=1 << 22)
    required, overflow = ws_big.status()
    assert not overflow and required > 1000
    ws = RasterWorkspace(planes.device)
    ws.capacity = 1000
    img = flare(planes, torch.from_numpy(ghost).cuda(), torch.from_numpy(light_spectrum()).cuda(),
                prm, ws)
    assert ws.capacity >= required
    assert np.array_equal(img.cpu().numpy(), ref)
