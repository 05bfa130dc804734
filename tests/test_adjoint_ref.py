"""Pins the FP64 PyTorch reference trace used to check the HIP adjoint:
its values equal the C oracle's double-precision trace, and its autograd
gradient equals FP64 central finite differences of that trace
(h = 1e-6 relative) along random parameter directions."""
import numpy as np
import pytest
import torch

from adjoint_helpers import directions, np_loss, oracle_outputs, setup
from realflare_amd.adjoint import LensParams, perturbed


def _torch_outputs(torch_trace, params, static, inp, paths, wavelengths, g):
    return torch_trace.trace(*params.tensors(), static, paths, wavelengths,
                             inp['coating_min_ior'], g, inp['grid_length'], inp['direction'])


def test_torch_reference_matches_oracle_f64(oracle_lib):
    import torch_trace

    inp, params, static = setup('small', grid_count=7)
    paths = [tuple(p) for p in inp['paths'][:6]] + [(-1, -1)]
    wl = list(inp['wavelengths'])
    x, y, r = _torch_outputs(torch_trace, params, static, inp, paths, wl, 7)
    ox, oy, orr = oracle_outputs(oracle_lib, params, static, inp, paths, wl, 7)
    assert np.array_equal(np.isnan(x.numpy()), np.isnan(ox))
    v = ~np.isnan(ox)
    assert v.mean() > 0.5
    np.testing.assert_allclose(x.numpy()[v], ox[v], rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(y.numpy()[v], oy[v], rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(r.numpy()[v], orr[v], rtol=1e-10, atol=0)


def test_torch_reference_gradient_matches_fd(oracle_lib):
    import torch_trace

    inp, params, static = setup('small', grid_count=5)
    paths = [tuple(p) for p in inp['paths'][:5]]
    wl = list(inp['wavelengths'][:2])
    target_params = perturbed(params, static, 0.01, seed=11)
    target = oracle_outputs(oracle_lib, target_params, static, inp, paths, wl, 5)
    base = oracle_outputs(oracle_lib, params, static, inp, paths, wl, 5)
    mask = np.all([np.isfinite(a) for a in (*base, *target)], axis=0)
    scale = float(np.sqrt(np.mean(target[2][mask] ** 2)))

    tp = LensParams(*(t.clone().requires_grad_(True) for t in params.tensors()))
    x, y, r = _torch_outputs(torch_trace, tp, static, inp, paths, wl, 5)
    m = torch.from_numpy(mask)
    tx, ty, tr = (torch.from_numpy(t) for t in target)
    dx, dy, dr = (x - tx)[m], (y - ty)[m], ((r - tr) / scale)[m]
    loss = (dx * dx + dy * dy + dr * dr).sum() / m.sum()
    loss.backward()
    grad = torch.cat([t.grad.reshape(-1) for t in tp.tensors()]).numpy()

    flat = params.flat().detach().double().numpy()
    for v in directions(params, 6, seed=5):
        h = 1e-6
        lp = np_loss(oracle_outputs(oracle_lib, LensParams.unflat(torch.tensor(flat + h * v), params),
                                    static, inp, paths, wl, 5), target, mask, scale)
        lm = np_loss(oracle_outputs(oracle_lib, LensParams.unflat(torch.tensor(flat - h * v), params),
                                    static, inp, paths, wl, 5), target, mask, scale)
        fd = (lp - lm) / (2 * h)
        ad = float(np.dot(grad, v))
        assert ad == pytest.approx(fd, rel=1e-5, abs=1e-9), (ad, fd)
