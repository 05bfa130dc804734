"""The batched configuration at its full size, against the oracle: the
12-surface lens (66 ghosts), 64 light positions on a seeded spiral, per-ghost
grids 32^2..256^2 from the footprint pre-trace, 16 wavelengths over 400-700
nm, 2048x2048 FLARE images, the whole batch captured as a HIP graph.

* Graph replays (with eager work between them) equal the eager batch.
* Sampled segments of the persistent trace launch meet the trace bar against
  the FP32 and FP64 oracle traces of the same ghost, grid and light
  (raytracing.cl:218-359).
* Rows of four frames equal the oracle's rasterization of the same traced
  planes, ghost by ghost, summed (rasterizing.cl:233-663): within 1e-5
  relative L1 per row (the batch sums every ghost's fragments before the AP1
  conversion; the sum of per-ghost images differs only by rounding).
"""
import numpy as np
import pytest
import torch

from parity import oracle_ghost_rows, planes_fields, record_fields, trace_parity
from realflare_amd.api.dtypes import lens_element_dtype
from realflare_amd.batch import BatchRenderer, Tuning, footprint_grids, spiral_positions
from realflare_amd.synthetic import workload

pytestmark = pytest.mark.gpu

FRAMES = 64
WAVELENGTHS = [int(round(400 + 300 * i / 15)) for i in range(16)]


@pytest.fixture(scope='module')
def batch():
    wl = workload('full')
    positions = spiral_positions(FRAMES, seed=7)
    grids = footprint_grids(wl.project, wl.lens_model, wl.paths, positions)
    print(f'case batched config: {FRAMES} frames x {len(wl.paths)} ghosts x '
          f'{len(WAVELENGTHS)} wavelengths, grids {grids.min()}..{grids.max()}', flush=True)
    br = BatchRenderer(wl.project, wl.lens_model, wl.paths, WAVELENGTHS, positions, grids,
                       Tuning(rays_per_lane=2, trace_workgroups=4096, frames_per_pass=8,
                              super_block=8))
    br.fit_capacity()
    yield wl, br
    del br
    torch.cuda.empty_cache()


def test_batched_config_graph_and_trace(batch, oracle_lib):
    wl, br = batch
    assert br.grids.min() == 32 and br.grids.max() == 256
    eager = br.render().clone()
    graph = br.capture()
    for _ in range(2):
        br.images.fill_(-1.0)
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(br.images, eager)
        br.render()
        torch.cuda.synchronize()
        assert torch.equal(br.images, eager)
    del eager

    # sampled segments: spread over frames and ghosts, with the largest and
    # smallest grids among them
    P = br.P
    flat = br.grids.reshape(-1)
    picks = {int(np.argmax(flat)), int(np.argmin(flat))}
    picks |= {f * P + (7 * f) % P for f in (3, 20, 41, 60)}
    lens = br.tr.lens_t.cpu().numpy().view(lens_element_dtype)
    gpu, r32, r64 = [], [], []
    for j in sorted(picks):
        s = br.seg[j]
        g = int(s['grid_count'])
        n = br.W * g * g
        ro = int(s['ray_offset'])
        gpu.append(planes_fields(br.planes[:, ro:ro + n].cpu().numpy()))
        args = (lens, [br.tr.paths[int(s['path'])]], WAVELENGTHS, br.tr.aperture_index,
                wl.project.flare.lens.coating_min_ior, g, float(s['grid_length']),
                np.array([*s['direction'], 1.0]))
        r32.append(record_fields(oracle_lib.trace(*args)))
        r64.append(record_fields(oracle_lib.trace(*args, precision='f64')))
    cat = [{k: np.concatenate([d[k] for d in ds]) for k in ds[0]} for ds in (gpu, r32, r64)]
    trace_parity(*cat, label=f'batched config, segments {sorted(picks)}')


@pytest.mark.parametrize('frame', [0, 21, 42, 63])
def test_batched_config_rows_match_oracle(batch, frame, oracle_lib):
    wl, br = batch
    img = br.images[frame].cpu().numpy()
    H, W = img.shape[:2]
    P = br.P
    ghost = br.ghost_tex.cpu().numpy()
    spectrum = br.spectrum.cpu().numpy()
    s0 = br.seg[frame * P]
    lo = int(s0['ray_offset'])
    hi = int(br.seg[(frame + 1) * P]['ray_offset']) if frame + 1 < br.F else br.N
    planes = br.planes[:, lo:hi].cpu().numpy()
    # rows through the brightest band and two fixed ones
    lum = np.abs(img).sum(axis=(1, 2))
    rows = sorted({int(np.argmax(lum)), H // 3, (2 * H) // 3})
    ref = np.zeros((len(rows), W, 4))
    for p in range(P):
        s = br.seg[frame * P + p]
        g = int(s['grid_count'])
        ro = int(s['ray_offset']) - lo
        got = oracle_ghost_rows(oracle_lib, planes[:, ro:ro + br.W * g * g], g, br.prm, ghost,
                                spectrum, float(s['grid_length']), float(br.prm.min_area), rows)
        for k in range(len(rows)):
            ref[k] += got[k]
    total = 0.0
    for k, out_row in enumerate(rows):
        a, b = img[out_row].astype(np.float64), ref[k]
        den = np.abs(b).sum()
        err = np.abs(a - b).sum() / max(den, 1e-30)
        print(f'frame {frame} row {out_row}: rel L1 {err:.2e} (sum {den:.3e})', flush=True)
        assert err <= 1e-5
        total += den
    assert total > 0
