"""Parity of the HIP aperture, FFT, power spectrum, ghost FrFT, starburst
and composition with the oracle, and of the Engine with an oracle-composed
FLARE_STARBURST frame.

Bars: aperture images with soft edges within 1e-5 absolute, hard edges may
flip pixels lying on the edge (<= 0.1% of pixels); scratch and dust masks
within 1e-2 relative L1, because their centres come from
fract(sin(x) * 43758.5453) at arguments up to ~3e5, where one ulp of sin
moves a centre by about half a pixel: the kernels round that sine
correctly, the oracle takes the C library's sinf, which is off by one ulp
for about 1.3% of such arguments; FFT within 1e-5 of the
norm; power spectrum and ghost texture (complex64 FFTs against numpy's
complex128) within 1e-4 relative L1; starburst within 1e-5 relative L1 with
the reference defaults (no blur/rotation jitter) and 1e-3 with jitter, whose
hash noise meets the same sine differences; a whole FLARE_STARBURST frame
within 1e-3 relative L1.
"""
import ctypes

import numpy as np
import pytest
import torch

from raster_helpers import rel_l1

pytestmark = pytest.mark.gpu


def _gpu_aperture(aperture, res, sp=(0, 0), dp=(0, 0)):
    from realflare_amd import native
    from realflare_amd.api.tasks.aperture import aperture_params

    prm = aperture_params(aperture, res, sp, dp)
    out = torch.empty((res[1], res[0]), dtype=torch.float32, device='cuda')
    native.call('rf_aperture', ctypes.byref(prm), native.ptr(out), native.stream())
    return out.cpu().numpy(), prm


def _oracle_aperture(oracle, prm):
    img = oracle.aperture_shape((prm.width, prm.height), tuple(prm.size), prm.blades,
                                prm.rotation, prm.roundness, prm.softness)
    if prm.grating_strength > 0:
        img = oracle.aperture_grating(img, tuple(prm.size), prm.grating_strength,
                                      prm.grating_density, prm.grating_length, prm.grating_width,
                                      prm.grating_softness)
    if prm.scratches_strength > 0:
        img = oracle.aperture_scratches(img, tuple(prm.size), prm.scratches_strength,
                                        prm.scratches_density, prm.scratches_length,
                                        prm.scratches_rotation,
                                        prm.scratches_rotation_variation, prm.scratches_width,
                                        tuple(prm.scratches_parallax), prm.scratches_softness)
    if prm.dust_strength > 0:
        img = oracle.aperture_dust(img, tuple(prm.size), prm.dust_strength, prm.dust_density,
                                   prm.dust_radius, tuple(prm.dust_parallax), prm.dust_softness)
    return img


@pytest.mark.parametrize('kind', ['soft', 'hard', 'grating', 'noise_masks'])
def test_aperture_matches_oracle(kind, oracle_lib):
    from realflare_amd.api.data import Aperture

    a = Aperture()
    if kind == 'soft':
        a.shape.blades, a.shape.softness, a.shape.rotation, a.shape.roundness = 6, 0.5, 15, 0.2
    elif kind == 'hard':
        a.shape.blades = 9
    elif kind == 'grating':
        a.shape.blades, a.shape.softness = 7, 0.3
        a.grating.strength, a.grating.softness = 0.6, 0.4
    else:
        a.shape.blades, a.shape.softness = 7, 0.3
        a.scratches.strength, a.scratches.rotation_variation, a.scratches.softness = 0.8, 0.5, 0.3
        a.dust.strength, a.dust.softness = 0.7, 0.5
    print(f'case aperture/{kind}', flush=True)
    img, prm = _gpu_aperture(a, (192, 160), (0.1, 0.05), (0.2, -0.1))
    ref = _oracle_aperture(oracle_lib, prm)
    diff = np.abs(img - ref)
    if kind == 'hard':
        assert np.mean(diff > 1e-5) <= 1e-3
    elif kind == 'noise_masks':
        # scratch/dust centres come from fract(sin(x) * 43758.5453) at arguments
        # up to ~3e5: one ulp of sin moves a centre by ~1e-3 of the aperture
        err = rel_l1(img, ref)
        print(f'noise masks rel L1 {err:.2e}, max {diff.max():.3f}', flush=True)
        assert err <= 1e-2
    else:
        assert diff.max() <= 1e-5, diff.max()


def test_fft2_matches_numpy():
    from realflare_amd import native

    rng = np.random.default_rng(1)
    # powers of two, and sides of any length (Bluestein), odd, prime, mixed,
    # and the largest of each kind
    for h, w in [(128, 256), (64, 64), (1024, 32), (1000, 600), (37, 12), (5, 1), (97, 128),
                 (2048, 3), (6, 4096), (1, 1), (1, 7), (2, 1), (1, 2)]:
        x = (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)
        for inverse in (0, 1):
            t = torch.from_numpy(x.view(np.float32).copy()).cuda()
            size = native.lib().rf_fft_workspace_size
            size.restype = ctypes.c_size_t
            ws = torch.empty((size(w, h),), dtype=torch.uint8, device='cuda')
            scale = 1.0 / (w * h) if inverse else 1.0
            native.call('rf_fft2', native.ptr(t), w, h, inverse, ctypes.c_float(scale),
                        native.ptr(ws), ctypes.c_size_t(ws.numel()), native.stream())
            got = t.cpu().numpy().view(np.complex64)
            ref = np.fft.ifft2(x.astype(np.complex128)) if inverse else np.fft.fft2(
                x.astype(np.complex128))
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            print(f'fft {h}x{w} inverse={inverse} rel err {err:.2e}', flush=True)
            assert err <= 1e-5


def test_fft_size_limits():
    from realflare_amd import native
    from realflare_amd.api.data import RealflareError

    x = torch.zeros((2 * 2049,), dtype=torch.float32, device='cuda')
    ws = torch.empty((1 << 20,), dtype=torch.uint8, device='cuda')
    for w, h in ((2049, 1), (1, 8192)):
        with pytest.raises(RealflareError, match='power of two up to 4096'):
            native.call('rf_fft2', native.ptr(x), w, h, 0, ctypes.c_float(1.0), native.ptr(ws),
                        ctypes.c_size_t(ws.numel()), native.stream())


@pytest.mark.parametrize('size', [(1000, 1000), (300, 200), (301, 199)])
def test_power_spectrum_and_ghost_any_size(size, oracle_lib):
    """Sizes a project may hold that are not powers of two (a 1000x1000
    starburst or ghost): the same bars as the power-of-two sizes."""
    from realflare_amd import native
    from realflare_amd.api.data import Aperture

    w, h = size
    a = Aperture()
    a.shape.blades, a.shape.softness = 7, 0.1
    ap, _ = _gpu_aperture(a, (w, h))
    ap_t = torch.from_numpy(ap).cuda()
    fn = native.lib().rf_power_spectrum_workspace_size
    fn.restype = ctypes.c_size_t
    ws = torch.empty((fn(w, h),), dtype=torch.uint8, device='cuda')
    out = torch.empty((h, w), dtype=torch.float32, device='cuda')
    native.call('rf_power_spectrum', native.ptr(ap_t), w, h, ctypes.c_float(1.0), native.ptr(ws),
                ctypes.c_size_t(ws.numel()), native.ptr(out), native.stream())
    ref = oracle_lib.power_spectrum(ap.astype(np.float64), 1.0)
    err = rel_l1(out.cpu().numpy(), ref)
    print(f'power spectrum {w}x{h} rel L1 {err:.2e}', flush=True)
    assert err <= 1e-4
    fn = native.lib().rf_ghost_workspace_size
    fn.restype = ctypes.c_size_t
    ws = torch.empty((fn(w, h),), dtype=torch.uint8, device='cuda')
    for fstop in (2.0, 16.0, 120.0):
        native.call('rf_ghost', native.ptr(ap_t), w, h, ctypes.c_float(fstop), native.ptr(ws),
                    ctypes.c_size_t(ws.numel()), native.ptr(out), native.stream())
        ref = oracle_lib.ghost(ap, fstop, (w, h))
        err = rel_l1(out.cpu().numpy(), ref)
        print(f'ghost {w}x{h} fstop={fstop} rel L1 {err:.2e}', flush=True)
        assert err <= 1e-4


def test_power_spectrum_and_ghost_match_oracle(oracle_lib):
    from realflare_amd import native
    from realflare_amd.api.data import Aperture

    a = Aperture()
    a.shape.blades, a.shape.softness = 9, 0.1
    ap, _ = _gpu_aperture(a, (512, 512))
    ap_t = torch.from_numpy(ap).cuda()
    size = native.lib().rf_power_spectrum_workspace_size
    size.restype = ctypes.c_size_t
    ws = torch.empty((size(512, 512),), dtype=torch.uint8, device='cuda')
    out = torch.empty((512, 512), dtype=torch.float32, device='cuda')
    for distance in (1.0, 0.05):
        native.call('rf_power_spectrum', native.ptr(ap_t), 512, 512, ctypes.c_float(distance),
                    native.ptr(ws), ctypes.c_size_t(ws.numel()), native.ptr(out), native.stream())
        ref = oracle_lib.power_spectrum(ap.astype(np.float64), distance)
        err = rel_l1(out.cpu().numpy(), ref)
        print(f'power spectrum d={distance} rel L1 {err:.2e}', flush=True)
        assert err <= 1e-4

    g, _ = _gpu_aperture(a, (256, 256))
    g_t = torch.from_numpy(g).cuda()
    size = native.lib().rf_ghost_workspace_size
    size.restype = ctypes.c_size_t
    ws = torch.empty((size(256, 256),), dtype=torch.uint8, device='cuda')
    out = torch.empty((256, 256), dtype=torch.float32, device='cuda')
    for fstop in (0.0, 2.0, 16.0, 50.0, 120.0, 200.0, 300.0):
        native.call('rf_ghost', native.ptr(g_t), 256, 256, ctypes.c_float(fstop), native.ptr(ws),
                    ctypes.c_size_t(ws.numel()), native.ptr(out), native.stream())
        ref = oracle_lib.ghost(g, fstop, (256, 256))
        err = rel_l1(out.cpu().numpy(), ref)
        print(f'ghost fstop={fstop} rel L1 {err:.2e}', flush=True)
        assert err <= 1e-4


@pytest.mark.parametrize('jitter', [False, True, 'blur'])
def test_starburst_matches_oracle(jitter, oracle_lib):
    from realflare_amd import native
    from realflare_amd.api.color import light_spectrum
    from realflare_amd.api.data import Flare
    from realflare_amd.api.tasks.starburst import starburst_params

    n = 256
    ap = oracle_lib.aperture_shape((n, n), (0.75, 0.75), 6, 0.0, 0.0, 0.02)
    spec = oracle_lib.power_spectrum(ap.astype(np.float64), 1.0)
    cfg = Flare.Starburst()
    cfg.vignetting_enabled = True
    if jitter == 'blur':  # the reference benchmark project's setting
        cfg.blur, cfg.rotation, cfg.rotation_weight = 1.0, 0.0, 1.0
    elif jitter:
        cfg.blur, cfg.rotation, cfg.rotation_weight = 1.0, 10.0, 2.0
    prm = starburst_params(cfg, (200, 150), 24, (0.3, 0.2), (1.0, 1.0))
    spec_t = torch.from_numpy(spec).cuda()
    light = torch.from_numpy(light_spectrum()).cuda()
    out = torch.empty((150, 200, 4), dtype=torch.float32, device='cuda')
    native.call('rf_starburst', native.ptr(spec_t), n, n, native.ptr(light), light.shape[0],
                ctypes.byref(prm), native.ptr(out), native.stream())
    ref = oracle_lib.starburst((200, 150), spec, light_spectrum(), prm.samples, prm.blur,
                               prm.rotation, prm.rotation_weight, tuple(prm.vignetting),
                               prm.intensity, tuple(prm.offset), tuple(prm.size))
    err = rel_l1(out.cpu().numpy(), ref)
    print(f'starburst jitter={jitter} rel L1 {err:.2e}', flush=True)
    assert np.abs(ref).sum() > 0
    assert err <= (1e-3 if jitter else 1e-5)


@pytest.mark.parametrize('blur, rotation', [(0.0, 0.0), (2.0, 0.0), (0.0, 3.0), (1.0, 3.0)])
def test_starburst_long_sample_loop_matches_oracle(blur, rotation, oracle_lib):
    """More samples than the kernel's per-sample table holds (1024), so the
    table is refilled mid-loop; a non-square power spectrum and a light far
    off centre with a small size, so that many lookups fall outside the
    spectrum (border 0) on both axes; each of the four jitter / rotation
    instantiations."""
    from realflare_amd import native
    from realflare_amd.api.color import light_spectrum
    from realflare_amd.api.data import Flare
    from realflare_amd.api.tasks.starburst import starburst_params

    sw, sh = 120, 80
    ap = oracle_lib.aperture_shape((sw, sh), (0.75, 0.75), 5, 0.0, 0.0, 0.05)
    spec = oracle_lib.power_spectrum(ap.astype(np.float64), 1.0)
    cfg = Flare.Starburst()
    cfg.blur, cfg.rotation, cfg.rotation_weight = blur, rotation, 1.0
    res = (37, 29)
    prm = starburst_params(cfg, res, 2100, (0.8, -0.7), (0.5, 0.6))
    spec_t = torch.from_numpy(spec).cuda()
    light = torch.from_numpy(light_spectrum()).cuda()
    out = torch.empty((res[1], res[0], 4), dtype=torch.float32, device='cuda')
    print(f'case starburst {res} x {prm.samples} samples blur={blur} rotation={rotation}',
          flush=True)
    native.call('rf_starburst', native.ptr(spec_t), sw, sh, native.ptr(light), light.shape[0],
                ctypes.byref(prm), native.ptr(out), native.stream())
    ref = oracle_lib.starburst(res, spec, light_spectrum(), prm.samples, prm.blur, prm.rotation,
                               prm.rotation_weight, tuple(prm.vignetting), prm.intensity,
                               tuple(prm.offset), tuple(prm.size))
    got = out.cpu().numpy()
    err = rel_l1(got, ref)
    print(f'rel L1 {err:.2e}', flush=True)
    assert np.abs(ref).sum() > 0 and np.count_nonzero(ref[..., 1]) < ref[..., 1].size
    assert err <= (1e-3 if blur or rotation else 1e-5)


@pytest.mark.parametrize('blur', [0.0, 1.0])
def test_starburst_skipped_rotation_is_exact(blur):
    """Rotation 0 skips the rotation's noise and trigonometry; a rotation of
    1e-30 evaluates them and moves no sample coordinate by a representable
    amount, so the two images are identical."""
    from realflare_amd import native
    from realflare_amd.api.color import light_spectrum
    from realflare_amd.api.data import Flare
    from realflare_amd.api.tasks.starburst import starburst_params

    n = 128
    rng = np.random.default_rng(5)
    spec_t = torch.from_numpy(rng.random((n, n)).astype(np.float32)).cuda()
    light = torch.from_numpy(light_spectrum()).cuda()
    outs = []
    for rotation in (0.0, 1e-30):
        cfg = Flare.Starburst()
        cfg.blur, cfg.rotation, cfg.rotation_weight = blur, rotation, 1.0
        prm = starburst_params(cfg, (96, 80), 40, (0.1, -0.2), (1.0, 1.0))
        out = torch.empty((80, 96, 4), dtype=torch.float32, device='cuda')
        native.call('rf_starburst', native.ptr(spec_t), n, n, native.ptr(light), light.shape[0],
                    ctypes.byref(prm), native.ptr(out), native.stream())
        outs.append(out.cpu().numpy())
    assert np.abs(outs[0]).sum() > 0
    np.testing.assert_array_equal(outs[0], outs[1])


def test_engine_flare_starburst_matches_oracle(oracle_lib):
    from helpers import trace_inputs
    from raster_helpers import oracle_image, params_for
    from realflare_amd.api.color import light_spectrum
    from realflare_amd.api.data import RenderElement, Size
    from realflare_amd.api.engine import Engine

    inp = trace_inputs('small', grid_count=32)
    wl = inp['workload']
    p = wl.project
    p.render.grid_count = 32
    p.render.resolution = Size(256, 256)
    p.render.starburst.resolution = Size(128, 128)
    p.render.starburst.samples = 16
    p.render.ghost.resolution = Size(64, 64)
    p.render.wavelengths = tuple(wl.wavelengths)
    p.render.ghost_paths = 'all'
    p.flare.lens.lens_model_path = wl.lens_model
    p.flare.ghost.fstop = 4.0
    p.flare.starburst.vignetting_enabled = True
    engine = Engine()
    engine.set_elements([RenderElement.FLARE_STARBURST])
    got = {}
    engine.image_rendered.append(lambda img, el: got.__setitem__(el, img))
    assert engine.render(p)
    img = got[RenderElement.FLARE_STARBURST].array

    # oracle composition of the same frame
    ga = oracle_lib.aperture_shape((64, 64), (0.75, 0.75), 6, 0.0, 0.0, 0.0)
    ghost = oracle_lib.ghost(ga, 4.0, (64, 64))
    prm = params_for(inp, (256, 256), ghost.shape, fstop=p.flare.lens.fstop)
    rays = oracle_lib.trace(inp['lens'], inp['paths'], inp['wavelengths'], inp['aperture_index'],
                            inp['coating_min_ior'], 32, inp['grid_length'], inp['direction'])
    *_, flare = oracle_image(oracle_lib, rays, inp, prm, ghost)
    sa = oracle_lib.aperture_shape((128, 128), (0.75, 0.75), 6, 0.0, 0.0, 0.0)
    spec = oracle_lib.power_spectrum(sa.astype(np.float64), 1.0)
    from realflare_amd.api.tasks.starburst import starburst_params
    sp = starburst_params(p.flare.starburst, (256, 256), 16, (0.3, 0.2), (1.0, 1.0))
    star = oracle_lib.starburst((256, 256), spec, light_spectrum(), 16, sp.blur, sp.rotation,
                                sp.rotation_weight, tuple(sp.vignetting), sp.intensity,
                                tuple(sp.offset), tuple(sp.size))
    ref = flare + (flare + star)
    err = rel_l1(img, ref)
    print(f'engine FLARE_STARBURST rel L1 {err:.2e}', flush=True)
    assert np.abs(flare).sum() > 0 and np.abs(star).sum() > 0
    assert err <= 1e-3
