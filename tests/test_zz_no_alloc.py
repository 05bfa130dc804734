"""rf_flare_static (no workspace argument, no allocation) against rf_flare
with a caller workspace, through the C ABI: two calls in a row, the second on
a smaller frame, each image bit-identical to rf_flare's and its status equal
to what rf_raster_status reports for rf_flare."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import trace_inputs
from raster_helpers import gpu_trace_planes, params_for
from realflare_amd import native
from realflare_amd.api.color import light_spectrum
from realflare_amd.api.tasks.rasterizing import RasterWorkspace, flare

pytestmark = pytest.mark.gpu


def _ghost(n=128):
    y, x = np.mgrid[0:n, 0:n] / (n - 1) * 2 - 1
    return np.clip(1.2 - np.hypot(x, y), 0, 1).astype(np.float32)


def test_flare_static_matches_flare_twice():
    dev = torch.device('cuda')
    spec = torch.from_numpy(light_spectrum()).to(dev)
    ghost = _ghost()
    ghost_t = torch.from_numpy(ghost).to(dev)
    status = torch.zeros(2, dtype=torch.int64, device=dev)
    for g, res in ((48, 512), (24, 256)):  # the second frame is the smaller
        print(f'case no-alloc rf_flare_static small g={g} {res}^2', flush=True)
        inp = trace_inputs('small', grid_count=g)
        planes = gpu_trace_planes(inp)
        prm = params_for(inp, (res, res), ghost.shape)
        ws = RasterWorkspace(dev)
        ref = flare(planes, ghost_t, spec, prm, ws)
        required, overflow = ws.status()
        assert not overflow and required > 0
        out = torch.full((res, res, 4), np.nan, dtype=torch.float32, device=dev)
        status.fill_(-1)
        native.call('rf_flare_static', native.ptr(planes), native.ptr(ghost_t), native.ptr(spec),
                    ctypes.byref(prm), native.ptr(out), native.ptr(status), native.stream())
        torch.cuda.synchronize()
        a, b = out.cpu().numpy(), ref.cpu().numpy()
        assert np.abs(b).sum() > 0
        np.testing.assert_array_equal(a, b)
        st = status.cpu().numpy()
        assert int(st[0]) == required and int(st[1]) == 0
