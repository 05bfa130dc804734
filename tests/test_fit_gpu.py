"""Batched fit step (rf_trace_fit, realflare_amd.fit) on several lens instances.

* Forward: for every (instance, light, ghost) segment the fit kernel's
  outputs equal rf_trace_planes_coated's for that lens alone (bit-exact).
* Gradient (squared and Huber loss): per instance it equals the single-lens adjoint
  (rf_trace_adjoint, itself checked against FP64 finite differences and an
  FP64 PyTorch trace) fed with the same loss gradient and summed over the
  instance's light positions, within 1e-5 relative (only the summation order
  differs); it does not depend on the checkpoint spacing (bit-exact).
* A 100-step fit from the base lenses towards traces of 1%-perturbed lenses
  lowers the best loss by 10x or more; at every step the FP32 forward of a sample
  segment is compared with the FP64 CPU trace (the oracle) of the same
  parameters, and the error is reported per step and bounded.
"""
import numpy as np
import pytest
import torch

import oracle
from adjoint_helpers import element_arrays
from realflare_amd.adjoint import GhostTrace, TraceSpec, perturbed
from realflare_amd.api.tasks.raytracing import light_direction
from realflare_amd.fit import FitBatch, fit, make_instances
from realflare_amd.synthetic import lens_grid_length, synthetic_lens

pytestmark = pytest.mark.gpu

RES = (512, 512)
SENSOR = (36.0, 24.0)
WL = [450, 550, 650]


def _instances(grid=16, positions=((0.3, 0.2), (-0.2, 0.1))):
    models = [synthetic_lens(6, 2024, fstop=4.0), synthetic_lens(8, 108, fstop=2.8)]
    inst = make_instances(models, positions, grid_range=(grid, grid), sensor=SENSOR, seed=1)
    return inst


def _segment_rays(fb, j):
    s = fb.seg[j]
    n = fb.W * int(s['grid_count']) ** 2
    return fb.fwd[:, s['ray_offset']:s['ray_offset'] + n].cpu().numpy()


def test_fit_forward_is_the_trace():
    inst = _instances(grid=12)
    fb = FitBatch(inst, WL, RES, SENSOR, checkpoint_every=4, rays_per_lane=2, workgroups=64)
    fb.set_targets([i.params for i in inst])
    for j, s in enumerate(fb.seg):
        it = inst[int(s['lens'])]
        pos = it.positions[int(s['frame']) % len(it.positions)]
        d = light_direction(pos, RES, SENSOR, it.model.focal_length)
        spec = TraceSpec(it.static, [it.paths[int(s['path']) - sum(len(x.paths) for x in inst[:int(s['lens'])])]],
                         WL, 1.38, int(s['grid_count']), lens_grid_length(it.model), d)
        planes = spec.forward(it.params).cpu().numpy()
        got = _segment_rays(fb, j)
        ref = np.stack([planes[0, 0].reshape(-1), planes[1, 0].reshape(-1),
                        planes[5, 0].reshape(-1)])
        # a ray that stops keeps its last position in the planes; the fit
        # kernel reports it as NaN
        alive = ~np.isnan(ref[2])
        np.testing.assert_array_equal(np.isnan(got[2]), ~alive)
        assert np.isnan(got[:2, ~alive]).all()
        np.testing.assert_array_equal(got[:, alive], ref[:, alive])


def _huber(r, delta):
    if delta <= 0:
        return r * r
    return torch.where(r.abs() > delta, 2 * delta * r.abs() - delta * delta, r * r)


def _reference_grads(fb, inst, params, targets, delta=0.0):
    """Single-lens adjoint per (instance, light), summed over lights."""
    tx, ty, tr = (t.cpu().numpy() for t in targets)
    out = []
    for k, it in enumerate(inst):
        acc = None
        p0 = sum(len(x.paths) for x in inst[:k])
        for f, pos in enumerate(it.positions):
            d = light_direction(pos, RES, SENSOR, it.model.focal_length)
            g = int(it.grids[0])
            spec = TraceSpec(it.static, it.paths, WL, 1.38, g, lens_grid_length(it.model), d)
            segs = [j for j, s in enumerate(fb.seg) if s['lens'] == k and
                    (s['frame'] - sum(len(x.positions) for x in inst[:k])) == f]
            sl = [slice(int(fb.seg[j]['ray_offset']), int(fb.seg[j]['ray_offset']) + fb.W * g * g)
                  for j in segs]
            t = [np.stack([a[s_].reshape(fb.W, g * g) for s_ in sl]) for a in (tx, ty, tr)]
            tp = [x.clone().requires_grad_(True) for x in params[k].tensors()]
            x, y, r = GhostTrace.apply(*tp, spec)
            xt, yt, rt = (torch.from_numpy(a).cuda() for a in t)
            m = ~(torch.isnan(x) | torch.isnan(xt) | torch.isnan(rt))
            z = torch.zeros_like(x)
            dx = torch.where(m, x - xt, z)
            dy = torch.where(m, y - yt, z)
            dr = torch.where(m, (r - rt) / fb.refl_scale, z)
            (fb.loss_scale * (_huber(dx, delta) + _huber(dy, delta) + _huber(dr, delta)).sum()).backward()
            gk = torch.cat([q.grad.reshape(-1) for q in tp]).numpy()
            acc = gk if acc is None else acc + gk
        out.append(acc)
    return out


def test_fit_gradient_matches_single_lens_adjoint():
    inst = _instances(grid=16)
    fb = FitBatch(inst, WL, RES, SENSOR, checkpoint_every=4, rays_per_lane=2, workgroups=96)
    target = [perturbed(i.params, i.static, 0.01, seed=5 + k) for k, i in enumerate(inst)]
    fb.set_targets(target)
    params = [i.params for i in inst]
    _, grads = fb.step(params)
    ref = _reference_grads(fb, inst, params, fb.targets)
    for g, r in zip(grads, ref):
        g = torch.cat([t.reshape(-1) for t in g]).numpy()
        err = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f'fit vs single-lens adjoint: rel err {err:.2e}', flush=True)
        assert err <= 1e-5
    # checkpoint spacing and launch shape change nothing
    for K, rpl, wgs in ((1, 2, 33), (3, 4, 128)):
        fb2 = FitBatch(inst, WL, RES, SENSOR, checkpoint_every=K, rays_per_lane=rpl,
                       workgroups=wgs)
        fb2.targets, fb2.refl_scale, fb2.loss_scale = fb.targets, fb.refl_scale, fb.loss_scale
        _, g2 = fb2.step(params)
        if rpl == fb.rpl:
            for a, b in zip(grads, g2):
                for ta, tb in zip(a, b):
                    np.testing.assert_array_equal(ta.numpy(), tb.numpy())
        else:
            for a, b in zip(grads, g2):
                ga = torch.cat([t.reshape(-1) for t in a]).numpy()
                gb = torch.cat([t.reshape(-1) for t in b]).numpy()
                assert np.linalg.norm(ga - gb) <= 1e-5 * np.linalg.norm(ga)


def test_fit_huber_gradient_matches_single_lens_adjoint():
    # the Huber form of the loss (residuals beyond delta pull linearly): the
    # fit kernel's gradient equals the single-lens adjoint fed with the
    # gradient of the same loss, and delta = 0 is the squared loss bit for bit
    inst = _instances(grid=16)
    fb = FitBatch(inst, WL, RES, SENSOR, checkpoint_every=4, rays_per_lane=2, workgroups=96)
    target = [perturbed(i.params, i.static, 0.01, seed=5 + k) for k, i in enumerate(inst)]
    fb.set_targets(target)
    params = [i.params for i in inst]
    _, g_sq = fb.step(params)
    fb.huber_delta = fb.residual_quantile(params, 0.5)
    assert fb.huber_delta > 0
    losses, grads = fb.step(params)
    ref = _reference_grads(fb, inst, params, fb.targets, delta=fb.huber_delta)
    for g, r, q in zip(grads, ref, g_sq):
        g = torch.cat([t.reshape(-1) for t in g]).numpy()
        q = torch.cat([t.reshape(-1) for t in q]).numpy()
        err = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f'huber delta {fb.huber_delta:.3e}: rel err {err:.2e}, '
              f'differs from squared by {np.linalg.norm(g - q) / np.linalg.norm(q):.2e}', flush=True)
        assert err <= 1e-5
        assert np.linalg.norm(g - q) > 1e-3 * np.linalg.norm(q)
    fb.huber_delta = 0.0
    _, g0 = fb.step(params)
    for a, b in zip(g_sq, g0):
        for ta, tb in zip(a, b):
            np.testing.assert_array_equal(ta.numpy(), tb.numpy())


def test_fit_100_steps_with_fp64_error_per_step():
    inst = _instances(grid=16)
    fb = FitBatch(inst, WL, RES, SENSOR, checkpoint_every=4, rays_per_lane=2)
    target = [perturbed(i.params, i.static, 0.01, seed=11 + k) for k, i in enumerate(inst)]
    fb.set_targets(target)
    it = inst[0]
    g = int(it.grids[0])
    d = light_direction(it.positions[0], RES, SENSOR, it.model.focal_length)

    def check(step, params):
        # segment 0: instance 0, first light, first ghost
        radius, ior, height, center, coating, coef = element_arrays(params[0], it.static)
        rays = oracle.trace_params(radius, ior, height, center, coating, coef, it.static.disperse,
                                   [it.paths[0]], WL, it.static.aperture_index, 1.38, g,
                                   lens_grid_length(it.model), d)
        ref = np.stack([rays['pos'][0, :, :, 0].reshape(-1), rays['pos'][0, :, :, 1].reshape(-1)])
        got = _segment_rays(fb, 0)[:2]
        ok = ~np.isnan(got[0]) & ~np.isnan(rays['reflectance'][0].reshape(-1))
        scale = np.abs(ref[:, ok]).max()
        err = np.abs(got[:, ok] - ref[:, ok]) / scale
        return {'fp64_pos_err_median': float(np.median(err)),
                'fp64_pos_err_p99': float(np.percentile(err, 99)),
                'fp64_pos_err_max': float(err.max())}

    _, hist = fit(fb, [i.params for i in inst], 100, lr=1e-4, on_step=check)
    for h in hist[::10] + [hist[-1]]:
        print(f"step {h['step']:3d} loss {h['loss']:.4e} best {h['best_loss']:.4e} fp64 pos err median "
              f"{h['fp64_pos_err_median']:.2e} p99 {h['fp64_pos_err_p99']:.2e} max "
              f"{h['fp64_pos_err_max']:.2e} gpu {h['gpu_ms']:.2f} ms", flush=True)
    assert hist[-1]['best_loss'] <= 0.1 * hist[0]['loss']
    assert max(h['fp64_pos_err_median'] for h in hist) <= 1e-5
    assert max(h['fp64_pos_err_p99'] for h in hist) <= 1e-3


def test_fit_counts_lost_rays_in_the_launch():
    """The launch's per-instance count of rays that have a target but that the
    trace loses equals the count over its forward outputs (NaN reflectance or
    a non-finite position where the target's reflectance is defined), and the
    reported loss adds lost_penalty per such ray.  The lenses are evaluated
    30 % away from the ones the targets come from, so that rays cross
    visibility edges."""
    from realflare_amd.adjoint import perturbed

    inst = _instances(grid=16)
    fb = FitBatch(inst, WL, RES, SENSOR, checkpoint_every=4, rays_per_lane=2, workgroups=64)
    fb.set_targets([i.params for i in inst], image_window=False)
    losses, _ = fb.step([perturbed(i.params, i.static, 0.3, seed=3 + k) for k, i in enumerate(inst)])
    torch.cuda.synchronize()
    counted = fb.grad[:, (3 + fb.W) * fb.LF + 1].cpu().numpy()
    expect = fb.lost_rays()
    print(f'lost rays: launch {counted.tolist()}, forward outputs {expect}', flush=True)
    assert sum(expect) > 0
    np.testing.assert_array_equal(counted, np.array(expect, np.float32))
    base = fb.grad[:, (3 + fb.W) * fb.LF].double().cpu().numpy()
    np.testing.assert_allclose(losses, base + fb.loss_scale * fb.lost_penalty * np.array(expect),
                               rtol=1e-12)
