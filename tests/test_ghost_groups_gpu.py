"""Ghost groups: the tracer and its adjoint give each workgroup four
consecutive ghosts, which share the straight path's events up to each
ghost's first reflection (raytracing.cl:288: the reflection test is false
below b1).  The straight state is advanced in path order and restarted from
init_ray when a ghost's b1 lies behind it, so the results must not depend on
the order of the path list, on repeated paths, or on paths that never
reflect.

* Trace: a shuffled path list with repeated paths and a path without
  reflection (-1, -1) gives, for every path, the same bits as tracing that
  path alone (a group of one, which shares nothing), and the FP32 oracle's
  positions.
* Adjoint: the shuffled list's gradient equals the ordered list's up to the
  order of the sums (the shared chain's reverse takes the ghosts' adjoints in
  decreasing b1), and its forward outputs are the same bits; with a path
  without reflection added, the gradient matches the FP64 PyTorch reference
  trace's autograd gradient."""
import numpy as np
import pytest
import torch

from adjoint_helpers import oracle_outputs, setup
from helpers import trace_inputs
from realflare_amd.adjoint import GhostTrace, LensParams, TraceSpec, ghost_loss, perturbed

pytestmark = pytest.mark.gpu


def _planes(inp, paths):
    from realflare_amd import native
    from realflare_amd.api.tasks.raytracing import trace_params

    dev = torch.device('cuda')
    lens = torch.from_numpy(inp['lens'].view(np.uint8).copy()).to(dev)
    pt = torch.from_numpy(np.ascontiguousarray(paths, np.int32)).to(dev)
    wl = torch.from_numpy(inp['wavelengths']).to(dev)
    prm = trace_params(len(inp['lens']), len(paths), len(inp['wavelengths']),
                       inp['aperture_index'], inp['coating_min_ior'], inp['grid_count'],
                       inp['grid_length'], inp['direction'])
    P, W, R = len(paths), len(inp['wavelengths']), inp['grid_count'] ** 2
    planes = torch.empty((6, P, W, R), dtype=torch.float32, device=dev)
    native.call('rf_trace_planes', native.ptr(lens), native.ptr(pt), native.ptr(wl),
                __import__('ctypes').byref(prm), native.ptr(planes), native.stream())
    return planes.cpu().numpy()


def _mixed_paths(paths, seed):
    rng = np.random.default_rng(seed)
    mixed = np.concatenate([paths, paths[[3, 3, 0]], np.array([[-1, -1]], np.int32)])
    return mixed[rng.permutation(len(mixed))].astype(np.int32)


def test_trace_path_order_repeats_and_no_bounce(oracle_lib):
    inp = trace_inputs('full', grid_count=24)
    paths = _mixed_paths(inp['paths'], seed=3)
    print(f'case trace: {len(paths)} shuffled paths with repeats and (-1, -1)', flush=True)
    grouped = _planes(inp, paths)
    for i, p in enumerate(paths):
        alone = _planes(inp, p[None])
        np.testing.assert_array_equal(grouped[:, i].view(np.uint32), alone[:, 0].view(np.uint32))
    ref = oracle_lib.trace(inp['lens'], paths, inp['wavelengths'], inp['aperture_index'],
                           inp['coating_min_ior'], inp['grid_count'], inp['grid_length'],
                           inp['direction'])
    ok = ~np.isnan(ref['reflectance']) & ~np.isnan(grouped[5])
    assert ok.mean() > 0.3
    np.testing.assert_array_equal(grouped[0][ok], ref['pos'][..., 0][ok])
    np.testing.assert_array_equal(grouped[1][ok], ref['pos'][..., 1][ok])
    np.testing.assert_array_equal(np.isnan(grouped[5]), np.isnan(ref['reflectance']))


def _spec(inp, static, paths, K=2):
    return TraceSpec(static, paths, inp['wavelengths'], inp['coating_min_ior'],
                     inp['grid_count'], inp['grid_length'], inp['direction'], checkpoint_every=K)


def test_adjoint_path_order():
    inp, params, static = setup('full', grid_count=12)
    paths = inp['paths']
    perm = np.random.default_rng(5).permutation(len(paths))
    print(f'case adjoint: {len(paths)} paths in order and shuffled', flush=True)
    a, b = _spec(inp, static, paths), _spec(inp, static, paths[perm])
    rng = np.random.default_rng(9)
    g3 = [rng.standard_normal(a.shape).astype(np.float32) for _ in range(3)]
    fa = torch.empty((3,) + a.shape, dtype=torch.float32, device='cuda')
    fb = torch.empty((3,) + b.shape, dtype=torch.float32, device='cuda')
    ga = a.adjoint(params, *(torch.from_numpy(g).cuda() for g in g3), forward_out=fa)
    gb = b.adjoint(params, *(torch.from_numpy(g[perm].copy()).cuda() for g in g3), forward_out=fb)
    x, y = fa.cpu().numpy()[:, perm], fb.cpu().numpy()
    assert np.array_equal(np.isnan(x), np.isnan(y))
    assert np.array_equal(x[~np.isnan(x)], y[~np.isnan(y)])
    err = np.linalg.norm(ga - gb) / np.linalg.norm(ga)
    print(f'gradient, ordered vs shuffled paths: rel {err:.2e}', flush=True)
    assert np.abs(ga).max() > 0 and err <= 1e-5


def test_adjoint_no_bounce_path_matches_torch_fp64(oracle_lib):
    import torch_trace

    inp, params, static = setup('small', grid_count=9)
    paths = _mixed_paths(inp['paths'], seed=8)
    target = oracle_outputs(oracle_lib, perturbed(params, static, 0.01, seed=11), static, inp,
                            paths=paths)
    base = oracle_outputs(oracle_lib, params, static, inp, paths=paths)
    mask = np.all([np.isfinite(v) for v in (*base, *target)], axis=0)
    scale = float(np.sqrt(np.mean(target[2][mask] ** 2)))
    spec = _spec(inp, static, paths, K=4)
    tp = LensParams(*(t.clone().requires_grad_(True) for t in params.tensors()))
    x, y, r = GhostTrace.apply(*tp.tensors(), spec)
    tt = [torch.from_numpy(np.where(mask, t, np.nan).astype(np.float32)).cuda() for t in target]
    ghost_loss(x, y, r, tt, scale).backward()
    g_gpu = torch.cat([t.grad.reshape(-1) for t in tp.tensors()]).numpy()

    tq = LensParams(*(t.clone().requires_grad_(True) for t in params.tensors()))
    x, y, r = torch_trace.trace(*tq.tensors(), static, [tuple(p) for p in paths],
                                list(inp['wavelengths']), inp['coating_min_ior'], 9,
                                inp['grid_length'], inp['direction'])
    m = torch.from_numpy(mask)
    tx, ty, tr = (torch.from_numpy(t) for t in target)
    dx, dy, dr = (x - tx)[m], (y - ty)[m], ((r - tr) / scale)[m]
    ((dx * dx + dy * dy + dr * dr).sum() / m.sum()).backward()
    g_ref = torch.cat([t.grad.reshape(-1) for t in tq.tensors()]).numpy()
    err = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    print(f'adjoint with shuffled, repeated and (-1, -1) paths vs torch fp64: rel {err:.2e}',
          flush=True)
    assert err <= 1e-3
