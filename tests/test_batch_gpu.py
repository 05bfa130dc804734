"""Batched form (realflare_amd.batch) against the plain path.

* rf_trace_batch, for every rays-per-lane choice, writes for every (frame,
  ghost) segment exactly the planes rf_trace_planes writes for that ghost
  alone with the same grid and light direction (bit-exact: same per-ray code).
* The batched FLARE images equal the plain per-ghost form (one trace and one
  rasterization per ghost and frame, added into the frame) within 1e-5
  relative L1: the two sum the same fragments; the plain form converts each
  ghost to AP1 separately and adds images, the batch converts per pixel once.
* One frame of the batch equals the oracle's rasterization of the same
  traced planes, ghost by ghost, summed (1e-5 relative L1).
* Frames per pass and the XCD super-block size do not change the images
  (bit-exact); a captured HIP graph replays to the eager result, with eager
  work between replays.
"""
import ctypes

import numpy as np
import pytest
import torch

from realflare_amd import native
from realflare_amd.api.tasks.raytracing import trace_params
from realflare_amd.batch import BatchRenderer, PlainPerGhost, Tuning, footprint_grids, spiral_positions
from realflare_amd.synthetic import workload

pytestmark = pytest.mark.gpu


def _setup(frames=3, seed=5, gmin=6, gmax=40):
    wl = workload('small')
    rng = np.random.default_rng(seed)
    grids = rng.integers(gmin, gmax + 1, size=(frames, len(wl.paths)))
    return wl, spiral_positions(frames, seed=seed), grids


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-30))


@pytest.mark.parametrize('rpl', [1, 2, 4])
def test_batch_trace_equals_plain_trace(rpl):
    wl, pos, grids = _setup()
    br = BatchRenderer(wl.project, wl.lens_model, wl.paths, wl.wavelengths, pos, grids,
                       Tuning(rays_per_lane=rpl, trace_workgroups=64, frames_per_pass=3))
    br.trace()
    planes = br.planes.cpu().numpy()
    W = br.W
    for s in br.seg:
        g = int(s['grid_count'])
        tp = trace_params(br.tr.lens_count, 1, W, br.tr.aperture_index,
                          wl.project.flare.lens.coating_min_ior, g, wl.project.render.grid_length,
                          (*s['direction'], 1.0))
        out = torch.empty((6, W * g * g), dtype=torch.float32, device='cuda')
        native.call('rf_trace_planes', native.ptr(br.tr.lens_t),
                    native.ptr_at(br.tr.paths_t, 8 * int(s['path'])), native.ptr(br.tr.wl_t),
                    ctypes.byref(tp), native.ptr(out), native.stream())
        ref = out.cpu().numpy()
        got = planes[:, s['ray_offset']:s['ray_offset'] + W * g * g]
        np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
        np.testing.assert_array_equal(np.nan_to_num(got), np.nan_to_num(ref))


def test_batch_images_equal_plain_per_ghost():
    wl, pos, grids = _setup()
    br = BatchRenderer(wl.project, wl.lens_model, wl.paths, wl.wavelengths, pos, grids,
                       Tuning(frames_per_pass=2))
    br.fit_capacity()
    batch = br.render().cpu().numpy()
    plain = PlainPerGhost(br)
    plain.fit_capacity(range(br.F))
    for f in range(br.F):
        plain.render_frame(f)
    ref = plain.images.cpu().numpy()
    assert np.abs(ref).sum() > 0
    for f in range(br.F):
        err = _rel_l1(batch[f], ref[f])
        print(f'frame {f}: rel L1 {err:.2e}', flush=True)
        assert err <= 1e-5


def test_batch_frame_matches_oracle(oracle_lib):
    from parity import oracle_ghost_rows

    wl, pos, grids = _setup()
    br = BatchRenderer(wl.project, wl.lens_model, wl.paths, wl.wavelengths, pos, grids,
                       Tuning(frames_per_pass=2))
    br.fit_capacity()
    imgs = br.render().cpu().numpy()
    ghost = br.ghost_tex.cpu().numpy()
    spectrum = br.spectrum.cpu().numpy()
    planes = br.planes.cpu().numpy()
    for f in range(br.F):
        ref = np.zeros(imgs[f].shape)
        for p in range(br.P):
            s = br.seg[f * br.P + p]
            g = int(s['grid_count'])
            ro = int(s['ray_offset'])
            ref += oracle_ghost_rows(oracle_lib, planes[:, ro:ro + br.W * g * g], g, br.prm, ghost,
                                     spectrum, float(s['grid_length']), float(br.prm.min_area),
                                     None)
        err = _rel_l1(imgs[f], ref)
        print(f'frame {f} vs oracle: rel L1 {err:.2e}', flush=True)
        assert np.abs(ref).sum() > 0
        assert err <= 1e-5


def test_batch_passes_and_super_block_do_not_change_images():
    wl, pos, grids = _setup(frames=4, seed=9)
    outs = []
    for fpp, sb in ((4, 4), (1, 2), (3, 8), (2, 1)):
        br = BatchRenderer(wl.project, wl.lens_model, wl.paths, wl.wavelengths, pos, grids,
                           Tuning(frames_per_pass=fpp, super_block=sb))
        br.fit_capacity()
        outs.append(br.render().cpu().numpy())
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_batch_graph_replays_match_eager():
    wl, pos, grids = _setup(frames=2, seed=3)
    br = BatchRenderer(wl.project, wl.lens_model, wl.paths, wl.wavelengths, pos, grids)
    br.fit_capacity()
    eager = br.render().clone()
    graph = br.capture()
    for _ in range(3):
        br.images.fill_(-1.0)
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(br.images, eager)
        br.render()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(br.images, eager)


def test_footprint_grids_range():
    wl, pos, _ = _setup(frames=4)
    g = footprint_grids(wl.project, wl.lens_model, wl.paths, pos, g_min=8, g_max=64)
    assert g.shape == (4, len(wl.paths))
    assert g.min() >= 8 and g.max() <= 64
    assert len(np.unique(g)) > 1  # ghosts differ in size


def test_autotune_choice_is_recorded_and_keeps_the_images(tmp_path):
    from realflare_amd import autotune

    wl, pos, grids = _setup(frames=2, seed=4)
    br = BatchRenderer(wl.project, wl.lens_model, wl.paths, wl.wavelengths, pos, grids)
    br.fit_capacity()
    ref = br.render().cpu().numpy()
    path = str(tmp_path / 'autotune.json')
    best = autotune.tune(br, reps=1, path=path, key='small')
    assert best.rays_per_lane in autotune.RAYS_PER_LANE
    assert best.super_block in autotune.SUPER_BLOCKS
    assert best.frames_per_pass in autotune.FRAMES_PER_PASS
    rec = br.autotune_record
    assert len(rec['raster_ms_per_frame']) == len(autotune.SUPER_BLOCKS) * 2  # F = 2: fpp 1, 2
    assert all(v > 0 for v in rec['trace_ms'].values())
    assert autotune.load('small', path) == best
    np.testing.assert_array_equal(br.render().cpu().numpy(), ref)
