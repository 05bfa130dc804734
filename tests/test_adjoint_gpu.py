"""The hand-written HIP adjoint of the trace (rf_trace_adjoint behind
realflare_amd.adjoint.GhostTrace).

* Its forward pass is bit-identical to rf_trace_planes_coated.
* Its gradient of the ghost loss (squared error of sensor positions and
  scaled reflectances against targets from a seeded 1% perturbation of
  radii, thicknesses, Sellmeier B and coating wavelengths) matches the
  autograd gradient of the FP64 PyTorch reference trace.
* On the differentiable configuration (12-surface lens, all 66 ghosts,
  64x64 rays, 8 wavelengths) it matches the FP64 central-difference
  gradient (h = 1e-6 relative) of the C oracle's double trace, checked along
  200 random relative parameter directions.
* Checkpoint spacing does not change the result.
* Edge cases: no ray with a target gives a zero loss and an exactly zero
  gradient; the smallest trace (2 x 2 rays, one wavelength) matches FP64
  central differences.
Bars are stated relative to the gradient's scale |g| |v| per direction,
since an FP32 trace differentiates the FP32 function.
"""
import numpy as np
import pytest
import torch

from adjoint_helpers import directions, oracle_outputs, setup
from realflare_amd.adjoint import GhostTrace, LensParams, TraceSpec, ghost_loss, perturbed

pytestmark = pytest.mark.gpu


def _spec(inp, static, K=4):
    return TraceSpec(static, inp['paths'], inp['wavelengths'], inp['coating_min_ior'],
                     inp['grid_count'], inp['grid_length'], inp['direction'], checkpoint_every=K)


def _gpu_grad(spec, params, target, mask, scale):
    tp = LensParams(*(t.clone().requires_grad_(True) for t in params.tensors()))
    x, y, r = GhostTrace.apply(*tp.tensors(), spec)
    tx, ty, tr = (torch.from_numpy(np.where(mask, t, np.nan).astype(np.float32)).cuda()
                  for t in target)
    loss = ghost_loss(x, y, r, (tx, ty, tr), scale)
    loss.backward()
    return float(loss), torch.cat([t.grad.reshape(-1) for t in tp.tensors()]).numpy()


def _target(oracle_lib, params, static, inp):
    target = oracle_outputs(oracle_lib, perturbed(params, static, 0.01, seed=11), static, inp)
    base = oracle_outputs(oracle_lib, params, static, inp)
    mask = np.all([np.isfinite(a) for a in (*base, *target)], axis=0)
    scale = float(np.sqrt(np.mean(target[2][mask] ** 2)))
    return target, mask, scale


def test_adjoint_forward_is_the_trace(oracle_lib):
    inp, params, static = setup('small', grid_count=17)
    spec = _spec(inp, static)
    planes = spec.forward(params)
    fwd = torch.empty((3,) + spec.shape, dtype=torch.float32, device='cuda')
    z = torch.zeros(spec.shape, dtype=torch.float32, device='cuda')
    spec.adjoint(params, z, z, z, forward_out=fwd)
    f = fwd.cpu().numpy()
    pl = planes.cpu().numpy()
    for a, b in ((f[0], pl[0]), (f[1], pl[1]), (f[2], pl[5])):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def test_split_forward_backward_equals_fused():
    inp, params, static = setup('full', grid_count=12)
    for K in (1, 4, 6):
        spec = _spec(inp, static, K)
        rng = np.random.default_rng(K)
        gx, gy, gr = (torch.from_numpy(rng.standard_normal(spec.shape).astype(np.float32)).cuda()
                      for _ in range(3))
        fused_fwd = torch.empty((3,) + spec.shape, dtype=torch.float32, device='cuda')
        g_fused = spec.adjoint(params, gx, gy, gr, forward_out=fused_fwd)
        fwd, tape = spec.forward_checkpointed(params)
        g_split = spec.backward(tape, gx, gy, gr)
        a, b = fwd.cpu().numpy(), fused_fwd.cpu().numpy()
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])
        np.testing.assert_array_equal(g_split, g_fused)
        assert np.isfinite(g_split).all() and np.abs(g_split).max() > 0


def test_adjoint_matches_torch_fp64_reference(oracle_lib):
    import torch_trace

    inp, params, static = setup('small', grid_count=9)
    target, mask, scale = _target(oracle_lib, params, static, inp)
    spec = _spec(inp, static)
    _, g_gpu = _gpu_grad(spec, params, target, mask, scale)

    tp = LensParams(*(t.clone().requires_grad_(True) for t in params.tensors()))
    x, y, r = torch_trace.trace(*tp.tensors(), static, [tuple(p) for p in inp['paths']],
                                list(inp['wavelengths']), inp['coating_min_ior'], 9,
                                inp['grid_length'], inp['direction'])
    m = torch.from_numpy(mask)
    tx, ty, tr = (torch.from_numpy(t) for t in target)
    dx, dy, dr = (x - tx)[m], (y - ty)[m], ((r - tr) / scale)[m]
    ((dx * dx + dy * dy + dr * dr).sum() / m.sum()).backward()
    g_ref = torch.cat([t.grad.reshape(-1) for t in tp.tensors()]).numpy()
    err = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    print(f'adjoint vs torch fp64 autograd: rel err {err:.2e}', flush=True)
    assert err <= 1e-3


def test_adjoint_matches_fp64_finite_differences(oracle_lib):
    """The differentiable configuration: 12-surface lens, all 66 ghosts at
    64x64 rays, 8 wavelengths.  FP64 central differences (h = 1e-6 relative)
    of the oracle's loss along every parameter with a nonzero magnitude (the
    stop radius and the air gaps' Sellmeier terms have none) give the FD
    gradient; 200 random relative directions are checked as projections of
    it (a directional central difference equals the projection to O(h^2))."""
    inp, params, static = setup('full', grid_count=64)
    print('case adjoint/full g=64: FP64 FD of the oracle loss', flush=True)
    target, mask, scale = _target(oracle_lib, params, static, inp)
    flat = params.flat().detach().double().numpy()
    h = 1e-6
    coords = np.flatnonzero(flat != 0)

    def sq_err(p):
        x, y, r = oracle_outputs(oracle_lib, LensParams.unflat(torch.tensor(p), params), static,
                                 inp)
        dx, dy, dr = x - target[0], y - target[1], (r - target[2]) / scale
        return dx * dx + dy * dy + dr * dr

    # Per-ray central differences; a ray that starts or stops reaching the
    # sensor within +-h sits on a visibility edge where the loss is not
    # differentiable, so it leaves the mask for every coordinate.
    diffs = np.zeros((len(coords),) + mask.shape, np.float32)
    edge = np.zeros(mask.shape, bool)
    for k, i in enumerate(coords):
        step = np.zeros_like(flat)
        step[i] = h * abs(flat[i])
        d = sq_err(flat + step) - sq_err(flat - step)
        edge |= mask & ~np.isfinite(d)
        diffs[k] = np.where(np.isfinite(d), d, 0.0)
    mask = mask & ~edge
    n = max(int(mask.sum()), 1)
    fd = np.zeros_like(flat)
    for k, i in enumerate(coords):
        fd[i] = diffs[k][mask].sum(dtype=np.float64) / n / (2 * h * abs(flat[i]))
    del diffs

    loss, g = _gpu_grad(_spec(inp, static), params, target, mask, scale)
    vs = directions(params, 200, seed=7)
    rel = np.array([abs(float(np.dot(g - fd, v))) / (np.linalg.norm(g) * np.linalg.norm(v))
                    for v in vs])
    sc = np.where(flat != 0, np.abs(flat), 0.0)
    coord = np.abs(g - fd)[coords] * sc[coords] / np.linalg.norm(g * sc)
    print(f'adjoint vs FP64 FD gradient ({len(coords)} coordinates, '
          f'{int(edge.sum())} visibility-edge rays of {int(mask.sum())} masked): loss {loss:.4e}, '
          f'200 directions rel err median {np.median(rel):.2e} p90 {np.percentile(rel, 90):.2e} '
          f'max {rel.max():.2e}; per coordinate max {coord.max():.2e}', flush=True)
    assert edge.sum() <= 1e-3 * mask.sum()
    assert np.median(rel) <= 1e-4
    assert rel.max() <= 1e-3


def test_checkpoint_spacing_is_invariant(oracle_lib):
    inp, params, static = setup('small', grid_count=9)
    target, mask, scale = _target(oracle_lib, params, static, inp)
    grads = [_gpu_grad(_spec(inp, static, K), params, target, mask, scale)[1] for K in (1, 3, 4, 8)]
    for g in grads[1:]:
        np.testing.assert_array_equal(g, grads[0])


def test_adjoint_without_targets_is_zero():
    """No ray has a target (all NaN): the loss is 0 and every parameter's
    gradient is exactly 0, with no NaN from the rays the loss leaves out."""
    inp, params, static = setup('small', grid_count=9)
    shape = (len(inp['paths']), len(inp['wavelengths']), 81)
    nan = np.full(shape, np.nan)
    loss, g = _gpu_grad(_spec(inp, static), params, (nan, nan, nan), np.ones(shape, bool), 1.0)
    assert loss == 0.0
    assert np.array_equal(g, np.zeros_like(g))


@pytest.mark.parametrize('K', [1, 4])
def test_adjoint_minimum_grid_single_wavelength(K, oracle_lib):
    """The smallest trace (2 x 2 rays per ghost, one wavelength): the gradient
    against FP64 central differences of the oracle's loss along 40 relative
    directions (rays on a visibility edge within +-h leave the mask)."""
    inp, params, static = setup('small', grid_count=2)
    inp['wavelengths'] = np.array([560], np.int32)
    inp['grid_length'] = 3.0
    print(f'case adjoint/grid 2 x 2, one wavelength, K={K}', flush=True)
    target, mask, scale = _target(oracle_lib, params, static, inp)
    assert mask.sum() > 0
    flat = params.flat().detach().double().numpy()
    h = 1e-6
    vs = directions(params, 40, seed=5)

    def sq_err(p):
        x, y, r = oracle_outputs(oracle_lib, LensParams.unflat(torch.tensor(p), params), static,
                                 inp)
        dx, dy, dr = x - target[0], y - target[1], (r - target[2]) / scale
        return dx * dx + dy * dy + dr * dr

    diffs = []
    edge = np.zeros(mask.shape, bool)
    for v in vs:
        d = sq_err(flat + h * v) - sq_err(flat - h * v)
        edge |= mask & ~np.isfinite(d)
        diffs.append(d)
    mask = mask & ~edge
    n = max(int(mask.sum()), 1)
    loss, g = _gpu_grad(_spec(inp, static, K), params, target, mask, scale)
    rel = []
    for v, d in zip(vs, diffs):
        fd = float(np.where(mask, d, 0.0).sum() / n / (2 * h))
        rel.append(abs(float(np.dot(g, v)) - fd) / (np.linalg.norm(g) * np.linalg.norm(v)))
    rel = np.array(rel)
    print(f'grid 2 x 2: {int(mask.sum())} rays, loss {loss:.3e}, rel err median '
          f'{np.median(rel):.2e} max {rel.max():.2e}', flush=True)
    assert np.median(rel) <= 1e-4
    assert rel.max() <= 1e-3
