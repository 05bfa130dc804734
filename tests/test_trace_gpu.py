"""Parity of the HIP ghost tracer (rf_trace / rf_trace_planes /
rf_trace_intersections) with the oracle's restatement of raytracing.cl,
on the standard small and full workloads.

Bar: rays valid in one and not the other are < 0.1% of all rays; on rays
valid in both, sensor position, aperture position, rrel and reflectance
agree within 1e-5 relative for 99.9% of rays (positions with a 1e-2 mm
floor on the denominator).  The FP32 error of both against the FP64
oracle trace is printed.
"""
import numpy as np
import pytest
import torch

from helpers import rel_err, trace_inputs

pytestmark = pytest.mark.gpu


def _gpu_inputs(inp):
    from realflare_amd.api.tasks.raytracing import trace_params

    dev = torch.device('cuda')
    lens = torch.from_numpy(inp['lens'].view(np.uint8).copy()).to(dev)
    paths = torch.from_numpy(inp['paths']).to(dev)
    wl = torch.from_numpy(inp['wavelengths']).to(dev)
    prm = trace_params(len(inp['lens']), len(inp['paths']), len(inp['wavelengths']),
                       inp['aperture_index'], inp['coating_min_ior'], inp['grid_count'],
                       inp['grid_length'], inp['direction'])
    return dict(lens=lens, paths=paths, wavelengths=wl, prm=prm)


def _oracle(oracle_lib, inp, precision='f32', intersections=False):
    return oracle_lib.trace(inp['lens'], inp['paths'], inp['wavelengths'], inp['aperture_index'],
                            inp['coating_min_ior'], inp['grid_count'], inp['grid_length'],
                            inp['direction'], precision=precision, intersections=intersections)


def _compare(gpu, ref, label):
    gv = ~np.isnan(gpu['reflectance'])
    rv = ~np.isnan(ref['reflectance'])
    flips = np.mean(gv != rv)
    both = gv & rv
    stats = {'valid': float(rv.mean()), 'flips': float(flips)}
    errs = {
        'pos': rel_err(gpu['pos'][both][:, :2], ref['pos'][both][:, :2], 1e-2),
        'pos_apt': rel_err(gpu['pos_apt'][both], ref['pos_apt'][both], 1e-3),
        'rrel': rel_err(gpu['rrel'][both], ref['rrel'][both], 1e-3),
        'reflectance': rel_err(gpu['reflectance'][both], ref['reflectance'][both], 1e-30),
    }
    for k, e in errs.items():
        stats[k] = (float(np.percentile(e, 99.9)), float(e.max()))
    print(label, stats, flush=True)
    return flips, errs


@pytest.mark.parametrize('name', ['small', 'full'])
def test_trace_matches_oracle(name, oracle_lib):
    from realflare_amd.api.tasks.raytracing import trace_records

    inp = trace_inputs(name)
    print(f'case trace/{name} g={inp["grid_count"]}', flush=True)
    gpu = trace_records(**_gpu_inputs(inp))
    ref = _oracle(oracle_lib, inp)
    assert gpu.shape == ref.shape
    flips, errs = _compare(gpu, ref, f'{name} gpu-vs-oracle32')
    assert flips < 1e-3
    for k, e in errs.items():
        assert np.percentile(e, 99.9) <= 1e-5, k
    ref64 = _oracle(oracle_lib, inp, 'f64')
    _compare(gpu, ref64, f'{name} gpu-vs-oracle64')
    _compare(ref, ref64, f'{name} oracle32-vs-oracle64')


def test_trace_planes_equal_records(oracle_lib):
    from realflare_amd import native
    from realflare_amd.api.tasks.raytracing import trace_records
    import ctypes

    inp = trace_inputs('small', grid_count=33)
    gi = _gpu_inputs(inp)
    rec = trace_records(**gi)
    P, W, R = rec.shape
    planes = torch.empty((6, P, W, R), dtype=torch.float32, device='cuda')
    native.call('rf_trace_planes', native.ptr(gi['lens']), native.ptr(gi['paths']),
                native.ptr(gi['wavelengths']), ctypes.byref(gi['prm']), native.ptr(planes),
                native.stream())
    pl = planes.cpu().numpy()
    np.testing.assert_array_equal(pl[0], rec['pos'][..., 0])
    np.testing.assert_array_equal(pl[1], rec['pos'][..., 1])
    np.testing.assert_array_equal(pl[2], rec['pos_apt'][..., 0])
    np.testing.assert_array_equal(pl[3], rec['pos_apt'][..., 1])
    np.testing.assert_array_equal(pl[4], rec['rrel'])
    np.testing.assert_array_equal(pl[5], rec['reflectance'])


def test_trace_intersections(oracle_lib):
    from realflare_amd.api.tasks.raytracing import trace_records

    inp = trace_inputs('small', grid_count=9)
    gi = _gpu_inputs(inp)
    rays, inters = trace_records(**gi, intersections=True)
    ref, ref_inters = _oracle(oracle_lib, inp, intersections=True)
    rays_only = trace_records(**gi)
    np.testing.assert_array_equal(rays.view(np.uint8), rays_only.view(np.uint8))
    np.testing.assert_array_equal(inters['hit'], ref_inters['hit'])
    hit = ref_inters['hit'] == 1
    e = rel_err(inters['pos'][hit], ref_inters['pos'][hit], 1e-2)
    assert np.percentile(e, 99.9) <= 1e-5
    e = np.abs(inters['incident'][hit] - ref_inters['incident'][hit])
    assert np.percentile(e, 99.9) <= 1e-5
