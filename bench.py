"""Benchmark: rays traced per second through a full 2048x2048 FLARE_STARBURST
frame on one MI355X.

Workload (the reference's drop-in full-size configuration): a seeded
synthetic 12-surface double-Gauss-like lens (f/2 stop mid-stack), all 66
two-reflection ghosts, 128x128 rays per ghost, 8 wavelengths over 400-700
nm, 2048x2048 RGBA fp32 output, 9-blade starburst aperture at 1024x1024
with 8 spectral samples; the remaining render settings are the reference
defaults.  A step is one whole frame: ghost aperture + FrFT, trace,
prim/vertex shaders, binner, rasterizer, starburst aperture, power
spectrum, starburst, composition.  All inputs are resident on the device.

value = rays per frame / seconds per frame (rays = ghosts x wavelengths x
grid^2).  Per-stage times use the reference's stage names.  The roofline
object describes the dominant kernel, timed with HIP events on the stream
it runs on.  cpu_baseline times the oracle (the C restatement of the
reference kernels, OpenMP over all host cores) on a bounded sample of the
same frame and extrapolates to a frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

# Algorithmic work per unit (frozen table; see the per-event counts in the
# design notes): FLOPs per traced ray of the full workload, and bytes the
# rasterizer must move at minimum per frame.
TRACE_FLOP_PER_EVENT_REFRACT = 89
TRACE_FLOP_PER_EVENT_REFLECT = 155
TRACE_FLOP_INIT = 76
TRACE_FLOP_STOP = 11
TRACE_FLOP_SENSOR = 14
TRACE_FLOP_SELLMEIER = 15


def trace_flops(pipe) -> float:
    """Algorithmic FLOPs of one trace launch: per ray, init + every surface
    event of its ghost path (counted from raytracing.cl, each +,-,*,/,
    sqrt and transcendental = 1), Sellmeier per glass side."""
    L = pipe.lens_count
    stop = pipe.lens_model.aperture_index
    ior = [e.refractive_index for e in pipe.lens_model.lens_elements] + [1.0]
    total = 0.0
    for b1, b2 in pipe.paths:
        seq = list(range(0, b1 + 1)) + list(range(b1 - 1, b2 - 1, -1)) + list(range(b2 + 1, L))
        f = TRACE_FLOP_INIT
        for k, el in enumerate(seq):
            if el == stop:
                f += TRACE_FLOP_STOP
            elif el == L - 1:
                f += TRACE_FLOP_SENSOR
            else:
                reflect = (el == b1 and k == b1) or (el == b2 and k == 2 * b1 - b2)
                f += TRACE_FLOP_PER_EVENT_REFLECT if reflect else TRACE_FLOP_PER_EVENT_REFRACT
                glass = (ior[el] > 1) + (el > 0 and ior[el - 1] > 1)
                f += TRACE_FLOP_SELLMEIER * glass
        total += f
    return total * pipe.W * pipe.R


def raster_bytes(pipe) -> float:
    """Minimum bytes the rasterizer moves per frame: each vertex's 28 useful
    bytes and each quad's bounds (16 bytes) read once, the RGBA fp32 image
    written once."""
    w, h = pipe.resolution
    return pipe.P * pipe.R * pipe.W * 28.0 + pipe.P * pipe.Q * 16.0 + w * h * 16.0


def measured_traffic(kernel: str = 'k_raster') -> float | None:
    """HBM bytes per dispatch of `kernel` on the full workload, as the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary records them
    (counter values scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, 'profiles', 'hbm_traffic_full_frame.txt')
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f'{kernel}_traffic_bytes:'):
                    return float(line.split(':', 1)[1])
    except OSError:
        return None
    return None


def cpu_baseline(pipe, inp_name: str = 'full') -> dict:
    """Oracle (C restatement of the reference kernels, OpenMP) on a bounded
    sample of the frame, extrapolated to one frame."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    import numpy as np

    import oracle
    from helpers import trace_inputs
    from raster_helpers import params_for

    from realflare_amd.api.color import light_spectrum

    oracle.build()
    inp = trace_inputs(inp_name)
    g = inp['grid_count']
    t0 = time.perf_counter()
    rays = oracle.trace(inp['lens'], inp['paths'], inp['wavelengths'], inp['aperture_index'],
                        inp['coating_min_ior'], g, inp['grid_length'], inp['direction'])
    t_trace = time.perf_counter() - t0
    ghost = pipe.ghost_tex.cpu().numpy()
    prm = params_for(inp, pipe.resolution, ghost.shape)
    t0 = time.perf_counter()
    bounds, intens = oracle.prim_shader(rays, g, prm.area_orig, prm.min_area)
    vtx = oracle.vertex_shader(rays, intens, g, prm.screen_transform, pipe.resolution)
    t_pv = time.perf_counter() - t0
    w, h = pipe.resolution
    rows = [int((i + 0.5) * h / 24) for i in range(24)]
    t0 = time.perf_counter()
    for y in rows:
        oracle.rasterize(vtx, bounds, prm.wavelength_sub_count, g, prm.anti_aliasing,
                         prm.intensity, prm.ghost_scale, ghost, light_spectrum(),
                         prm.screen_transform, prm.bin_size, pipe.resolution, rows=(y, y + 1))
    t_rows = time.perf_counter() - t0
    t_raster = t_rows / len(rows) * h
    # starburst: power spectrum (numpy, as the reference) + kernel on 4 rows
    ap = pipe.star_ap.cpu().numpy()
    t0 = time.perf_counter()
    spec = oracle.power_spectrum(ap, pipe.project.flare.starburst.distance)
    t_fft = time.perf_counter() - t0
    sp = pipe.star_prm
    t0 = time.perf_counter()
    oracle.starburst((w, 4), spec, light_spectrum(), sp.samples, sp.blur, sp.rotation,
                     sp.rotation_weight, tuple(sp.vignetting), sp.intensity, tuple(sp.offset),
                     tuple(sp.size))
    t_star = (time.perf_counter() - t0) / 4 * h
    frame_s = t_trace + t_pv + t_raster + t_fft + t_star
    return {
        'value': pipe.rays / frame_s,
        'unit': 'rays/s',
        'cores': oracle.threads(),
        'kind': 'port',
        'sample': (f'oracle trace of all {pipe.rays} rays ({t_trace:.2f} s), prim+vertex shaders '
                   f'({t_pv:.2f} s), rasterizer on {len(rows)} full-width rows spread over the '
                   f'frame ({t_rows:.2f} s, x{h / len(rows):.0f} to a frame), numpy power '
                   f'spectrum ({t_fft:.2f} s), starburst on 4 rows (x{h / 4:.0f}); '
                   f'est. {frame_s:.1f} s per frame'),
        'frame_s_estimate': frame_s,
        'stages_s': {'RaytracingTask.run': t_trace, 'prim+vertex_shader': t_pv,
                     'rasterizer': t_raster, 'fourier_spectrum': t_fft, 'starburst': t_star},
    }


def adjoint_timing(wl, steps: int) -> dict:
    """Differentiable configuration: the same 12-surface lens, 66 ghosts,
    64x64 rays, 8 wavelengths, upstream gradients of ones.  Device time per
    batch of: the plain forward trace (rf_trace_planes_coated), the forward
    pass that also writes the checkpoint tape, the backward pass from that
    tape, and forward + backward fused in one launch."""
    import ctypes

    import torch

    from realflare_amd import native
    from realflare_amd.adjoint import TraceSpec, lens_params
    from realflare_amd.api.tasks.raytracing import light_direction

    p = wl.project
    sensor = (p.flare.lens.sensor_size.width(), p.flare.lens.sensor_size.height())
    res = (p.render.resolution.width(), p.render.resolution.height())
    params, static = lens_params(wl.lens_model, sensor, p.flare.lens.glasses_path,
                                 p.flare.lens.abbe_nr_adjustment, p.flare.lens.coating)
    direction = light_direction((p.flare.light.position.x(), p.flare.light.position.y()), res,
                                sensor, wl.lens_model.focal_length)
    out = {}
    for K in (1, 2, 3, 4):
        spec = TraceSpec(static, wl.paths, wl.wavelengths, p.flare.lens.coating_min_ior, 64,
                         p.render.grid_length, direction, checkpoint_every=K)
        ones = torch.ones(spec.shape, dtype=torch.float32, device='cuda')
        lens, coat = spec.upload(params)
        P, W, R = spec.shape
        planes = torch.empty((6, P, W, R), dtype=torch.float32, device='cuda')
        fwd_out = torch.empty((3, P, W, R), dtype=torch.float32, device='cuda')
        tape = torch.empty((spec.ws_bytes,), dtype=torch.uint8, device='cuda')
        gout = torch.empty((3 * (static.n_model + 1) * (1 + W),), dtype=torch.float32, device='cuda')
        common = (native.ptr(lens), native.ptr(coat), native.ptr(spec.paths_t),
                  native.ptr(spec.wl_t), ctypes.byref(spec.prm))
        grads = (native.ptr(ones), native.ptr(ones), native.ptr(ones))
        ws = (native.ptr(tape), ctypes.c_size_t(spec.ws_bytes))
        fns = {
            'forward': lambda: native.call('rf_trace_planes_coated', *common, native.ptr(planes),
                                           native.stream()),
            'forward_checkpointed': lambda: native.call(
                'rf_trace_forward_checkpointed', *common, spec.K, *ws, native.ptr(fwd_out),
                native.stream()),
            'backward': lambda: native.call('rf_trace_backward', *common, spec.K, *grads, *ws,
                                            native.ptr(gout), native.stream()),
            'fused': lambda: native.call('rf_trace_adjoint', *common, spec.K, *grads, *ws,
                                         native.ptr(gout), None, native.stream()),
        }
        times = {}
        for name in ('forward', 'forward_checkpointed', 'backward', 'fused'):
            fn = fns[name]
            if name == 'backward':
                fns['forward_checkpointed']()
            fn()
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(steps):
                fn()
            b.record()
            torch.cuda.synchronize()
            times[name + '_ms'] = a.elapsed_time(b) / steps
        times['backward_over_forward'] = times['backward_ms'] / times['forward_ms']
        times['fused_over_forward'] = times['fused_ms'] / times['forward_ms']
        times['gradient_over_forward'] = ((times['forward_checkpointed_ms'] + times['backward_ms'])
                                          / times['forward_ms'])
        out[f'K{K}'] = times
    out['rays'] = 66 * 8 * 64 * 64
    # GhostTrace's default checkpoint spacing (TraceSpec checkpoint_every=2)
    d = out['K2']
    out['default'] = {'K': 2, 'fused_over_forward': d['fused_over_forward'],
                      'gradient_over_forward': d['gradient_over_forward']}
    return out


def batched_timing(steps: int, warmup: int, n_frames: int = 64, tune_frames: int = 8,
                   plain_frames: int = 4) -> dict:
    """Batched configuration: the 12-surface lens, `n_frames` light positions
    on a seeded spiral, per-ghost grids 32^2..256^2 from a 3x3 footprint
    pre-trace, 16 wavelengths over 400-700 nm, 2048x2048 FLARE image per frame.
    The batch (one persistent trace launch, prim/vertex passes, binning and
    rasterization in passes of frames) is captured into one HIP graph and
    replayed; launch choices come from the autotuner run on `tune_frames`
    frames of the batch.  The plain per-ghost form (one trace and one
    rasterization per ghost and frame) is timed on `plain_frames` frames."""
    import numpy as np
    import torch

    from realflare_amd import autotune
    from realflare_amd.batch import (BatchRenderer, PlainPerGhost, Tuning, footprint_grids,
                                     spiral_positions)
    from realflare_amd.synthetic import workload

    wl = workload('full')
    wavelengths = [int(round(400 + 300 * i / 15)) for i in range(16)]
    positions = spiral_positions(n_frames, seed=7)
    grids = footprint_grids(wl.project, wl.lens_model, wl.paths, positions)
    idx = np.unique(np.linspace(0, n_frames - 1, tune_frames).round().astype(int))
    sample = BatchRenderer(wl.project, wl.lens_model, wl.paths, wavelengths,
                           [positions[i] for i in idx], grids[idx], Tuning())
    tuning = autotune.tune(sample, reps=2, key='batched-12s-66g-16w-2048px')
    record = sample.autotune_record
    del sample
    torch.cuda.empty_cache()
    print(f'# autotuned: {tuning}', file=sys.stderr, flush=True)

    br = BatchRenderer(wl.project, wl.lens_model, wl.paths, wavelengths, positions, grids, tuning)
    pairs = br.fit_capacity()
    graph = br.capture()
    for _ in range(warmup):
        graph.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        graph.replay()
    b.record()
    torch.cuda.synchronize()
    ms_batch = a.elapsed_time(b) / steps
    # stage split of one eager batch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    br.trace()
    ev[1].record()
    br.shade()
    ev[2].record()
    for d in br.passes:
        br.raster_pass(d)
    ev[3].record()
    torch.cuda.synchronize()
    split = {'trace_ms': ev[0].elapsed_time(ev[1]), 'prim_vertex_ms': ev[1].elapsed_time(ev[2]),
             'bin_raster_ms': ev[2].elapsed_time(ev[3])}
    batch_images = br.images

    plain = PlainPerGhost(br)
    frames = [int(f) for f in np.linspace(0, n_frames - 1, plain_frames).round()]
    plain.fit_capacity(frames)
    plain.render_frame(frames[0])
    torch.cuda.synchronize()
    a.record()
    for f in frames:
        plain.render_frame(f)
    b.record()
    torch.cuda.synchronize()
    ms_plain = a.elapsed_time(b) / len(frames)
    errs = []
    for f in frames:
        d = (plain.images[f] - batch_images[f]).abs().sum()
        errs.append(float(d / batch_images[f].abs().sum().clamp(min=1e-30)))
    ms_frame = ms_batch / n_frames
    return {
        'frames': n_frames, 'wavelengths': len(wavelengths),
        'grid_min': int(grids.min()), 'grid_median': float(np.median(grids)),
        'grid_max': int(grids.max()), 'rays_per_batch': br.rays,
        'rays_per_frame': br.rays / n_frames, 'max_pass_pairs': pairs,
        'ms_per_batch': ms_batch, 'ms_per_frame': ms_frame, 'frames_per_s': 1000.0 / ms_frame,
        'rays_per_s': br.rays / (ms_batch * 1e-3), 'stage_split_eager': split,
        'hipgraph': True, 'tuning': record,
        'plain_per_ghost_ms_per_frame': ms_plain, 'plain_frames_timed': frames,
        'speedup_vs_plain_per_ghost': ms_plain / ms_frame,
        'batched_vs_plain_rel_l1_max': max(errs),
    }


def fit_timing(fit_steps: int = 100, lr: float = 3e-3) -> dict:
    """Batched + differentiable configuration: 16 synthetic lenses of 8..16
    surfaces (all 28..120 refracting-pair ghosts each), 32 light positions
    each on a seeded spiral, a seeded grid of 64^2..128^2 rays per ghost, 8
    wavelengths, 2048^2 image; one rf_trace_fit launch per gradient step
    (forward, loss against traces of 1%-perturbed lenses, adjoint).  A
    `fit_steps`-step Adam fit of all 16 lenses runs on that batch (squared
    error of the rays whose target lands in the image; step size 3e-3 of each
    parameter's magnitude, the best of 1e-4, 1e-3, 3e-3 measured); every step compares
    one segment's FP32 forward with the FP64 oracle trace of the same
    parameters.  Reported: device ms per step, the summed and per-lens losses
    at the start and at the best points, and the FP64 error per step."""
    import numpy as np
    import torch

    from realflare_amd.adjoint import perturbed
    from realflare_amd.api.tasks.raytracing import light_direction
    from realflare_amd.batch import spiral_positions
    from realflare_amd.fit import FitBatch, fit, make_instances
    from realflare_amd.synthetic import lens_grid_length, synthetic_lens

    res, sensor = (2048, 2048), (36.0, 24.0)
    wavelengths = [int(round(400 + 300 * i / 7)) for i in range(8)]
    models = [synthetic_lens(8 + i % 9, 300 + 17 * i, fstop=2.8) for i in range(16)]
    positions = spiral_positions(32, seed=11)
    inst = make_instances(models, positions, grid_range=(64, 128), sensor=sensor, seed=3)
    target = [perturbed(i.params, i.static, 0.01, seed=40 + k) for k, i in enumerate(inst)]
    fb = FitBatch(inst, wavelengths, res, sensor, checkpoint_every=4, rays_per_lane=4)
    fb.set_targets(target)
    start = [i.params for i in inst]

    # FP64 check of segment 0 (lens 0, its first light and ghost) per step
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), 'oracle'))
    import oracle

    it, s0 = inst[0], fb.seg[0]
    g0 = int(s0['grid_count'])
    n0 = fb.W * g0 * g0
    d0 = light_direction(it.positions[0], res, sensor, it.model.focal_length)

    def check(step, params):
        p = params[0]
        n = it.static.n_model
        radii = p.radii.detach().double().numpy()
        dists = p.distances.detach().double().numpy()
        b = p.sellmeier_b.detach().double().numpy()
        radius = np.append(radii, 0.0)
        center = np.concatenate([[0.0], np.cumsum(dists)]) + radius
        coating = np.append(p.coating.detach().double().numpy(), it.static.sensor_coating)
        coef = np.zeros((n + 1, 6))
        for i in range(n):
            if np.any(b[i] != 0) or it.static.has_glass[i]:
                coef[i, 0::2] = b[i]
                coef[i, 1::2] = it.static.sellmeier_c[i]
        r = oracle.trace_params(radius, it.static.ior, it.static.height, center, coating, coef,
                                it.static.disperse, [it.paths[0]], wavelengths,
                                it.static.aperture_index, 1.38, g0, lens_grid_length(it.model), d0)
        ref = np.stack([r['pos'][0, :, :, 0].reshape(-1), r['pos'][0, :, :, 1].reshape(-1)])
        got = fb.fwd[:2, :n0].cpu().numpy()
        ok = ~np.isnan(got[0]) & ~np.isnan(r['reflectance'][0].reshape(-1))
        err = np.abs(got[:, ok] - ref[:, ok]) / np.abs(ref[:, ok]).max()
        if step % 10 == 0:
            print(f'# fit step {step}', file=sys.stderr, flush=True)
        return {'fp64_pos_err_median': float(np.median(err)),
                'fp64_pos_err_p99': float(np.percentile(err, 99))}

    best_params, hist = fit(fb, start, fit_steps, lr=lr, on_step=check)
    first = hist[0]['instance_loss']
    best = hist[-1]['instance_best']
    ratio = [b_ / f_ if f_ > 0 else 0.0 for f_, b_ in zip(first, best)]
    out = {
        'lens_instances': len(inst), 'surfaces': [len(m.lens_elements) - 1 for m in models],
        'ghosts': sum(len(i.paths) for i in inst), 'light_positions': len(positions),
        'wavelengths': len(wavelengths), 'rays_per_step': fb.rays,
        'ms_per_step': float(np.median([h['gpu_ms'] for h in hist])),
        'fit': {'steps': fit_steps, 'adam_lr': lr,
                'loss': 'squared error, rays whose target lands in the image',
                'loss_first': sum(first), 'loss_best': sum(best),
                'loss_best_over_first': sum(best) / sum(first),
                'best_loss_at': {str(k): hist[k]['best_loss'] for k in (0, 10, 25, 50, 99)
                                 if k < len(hist)},
                'instance_loss_first': first, 'instance_loss_best': best,
                'instances_reduced_10x': sum(r_ <= 0.1 for r_ in ratio),
                'instance_best_over_first_median': float(sorted(ratio)[len(ratio) // 2]),
                'fp64_pos_err_median_max': max(h['fp64_pos_err_median'] for h in hist),
                'fp64_pos_err_p99_max': max(h['fp64_pos_err_p99'] for h in hist)},
    }
    out['rays_per_s_forward_plus_adjoint'] = fb.rays / (out['ms_per_step'] * 1e-3)
    del fb
    torch.cuda.empty_cache()
    return out


def reference_benchmark(frames: int = 3) -> dict:
    """The reference's own benchmark (benchmark/benchmark.py with
    benchmark/nikon_ai_50_135mm): the Nikon AI 50-135mm project (1920x1080,
    anti-aliasing 4, 5 wavelengths x 8 sub-steps, 129^2 rays, cull 0.5, ghost
    512^2, starburst 512^2 x 1000 samples), FLARE_STARBURST rendered by
    Engine.render for the animation's light positions (0.6, 0.3), (0.3, 0.6),
    the last held for the frames after it.  The stage times are parsed from
    the timer log lines as that script parses them (device-synchronized
    stages); the first frame (workspace sizing) is reported apart from the
    mean of the others."""
    import logging
    import re
    import statistics

    from realflare_amd.api.data import Point, Project, RenderElement, cast
    from realflare_amd.api.engine import Engine

    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, 'tests', 'golden', 'nikon_benchmark_project.json')) as f:
        project = cast(Project, json.load(f))
    project.output.write = False
    positions = [(0.6, 0.3), (0.3, 0.6)]
    lines = []

    class Grab(logging.Handler):
        def emit(self, record):
            lines.append(record.getMessage())

    log = logging.getLogger('realflare_amd.utils.timing')
    handler = Grab(level=logging.INFO)
    old_level = log.level
    log.setLevel(logging.INFO)
    log.addHandler(handler)
    pattern = re.compile(r'([\w.]+):\s*(\d+(?:\.\d+)?)ms$')
    per_frame = []
    try:
        engine = Engine()
        engine.set_elements([RenderElement.FLARE_STARBURST])
        for i in range(frames):
            project.flare.light.position = Point(*positions[min(i, len(positions) - 1)])
            lines.clear()
            assert engine.render(project)
            times: dict = {}
            for line in lines:
                m = pattern.search(line.strip())
                if m:
                    times.setdefault(m.group(1), []).append(float(m.group(2)))
            per_frame.append({k: sum(v) for k, v in times.items()})
    finally:
        log.removeHandler(handler)
        log.setLevel(old_level)
    warm = per_frame[1:] or per_frame
    keys = list(per_frame[0])
    return {
        'project': 'benchmark/nikon_ai_50_135mm (1920x1080, AA 4, 5 x 8 wavelength steps, '
                   '129^2 rays, cull 0.5, starburst 512^2 x 1000 samples)',
        'frames': frames,
        'stage_ms_mean_after_first': {k: statistics.mean(f.get(k, 0.0) for f in warm) for k in keys},
        'stage_ms_first_frame': per_frame[0],
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='full', choices=['full', 'small'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-graph', action='store_true',
                    help='time frames launched one by one instead of graph replays')
    ap.add_argument('--no-batched', action='store_true', help='skip the batched configuration')
    ap.add_argument('--batch-frames', type=int, default=64)
    ap.add_argument('--no-fit', action='store_true', help='skip the batched fit configuration')
    ap.add_argument('--no-reference-benchmark', action='store_true',
                    help="skip the reference's own benchmark project")
    ap.add_argument('--reference-benchmark-only', action='store_true',
                    help="run only the reference's own benchmark project and print it")
    ap.add_argument('--copy-calibration', action='store_true',
                    help='also run a device-to-device copy of 256 MiB (for counter scaling)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py measures one GPU; --gpus {args.gpus} is not supported (the flare path '
              'runs on one GPU; frames or ghosts are not sharded across GPUs)', file=sys.stderr)
        return 2

    import torch

    from realflare_amd.pipeline import FramePipeline
    from realflare_amd.synthetic import workload

    assert torch.cuda.is_available(), 'bench.py needs a HIP device'
    if args.reference_benchmark_only:
        print(json.dumps({'reference_benchmark': reference_benchmark()}), flush=True)
        return 0
    wl = workload(args.workload)
    pipe = FramePipeline(wl.project, wl.lens_model, wl.paths, wl.wavelengths)
    print(f'# workload {args.workload}: {pipe.P} ghosts x {pipe.W} wavelengths x {pipe.g}^2 rays '
          f'= {pipe.rays} rays, {pipe.resolution[0]}x{pipe.resolution[1]}', file=sys.stderr,
          flush=True)
    pairs = pipe.fit_capacity()
    print(f'# binning pairs per frame: {pairs}', file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        pipe.render()
    torch.cuda.synchronize()

    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # frames launched one by one (about 40 launches each), for reference
    torch.cuda.synchronize()
    start.record()
    for _ in range(args.steps):
        pipe.render()
    stop.record()
    torch.cuda.synchronize()
    eager_ms = start.elapsed_time(stop) / args.steps

    # timed region: exactly `steps` frames, each one replay of the frame
    # captured as a HIP graph (the same launches, without their host-side gaps)
    graph = None
    if not args.no_graph:
        graph = pipe.capture()
        for _ in range(max(1, args.warmup)):
            graph.replay()
    frame = graph.replay if graph is not None else pipe.render
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        frame()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop) / args.steps
    _, overflow = pipe.raster_status()
    assert not overflow

    # per-stage times (separate pass, events around every stage)
    events: dict = {}
    for _ in range(args.steps):
        pipe.render_timed(events)
    torch.cuda.synchronize()
    stages = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in events.items()}
    print('# stage ms: ' + ', '.join(f'{k} {v:.3f}' for k, v in stages.items()), file=sys.stderr,
          flush=True)

    dominant = max(stages, key=stages.get)
    if dominant == 'raytrace':
        flops = trace_flops(pipe)
        achieved = flops / (stages['raytrace'] * 1e-3) / 1e12
        roofline = {'kernel': 'k_trace (rf_trace_planes)', 'bound': 'valu',
                    'achieved': achieved, 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                    'frac': achieved / FP32_PEAK_TFLOPS, 'traffic': None}
    else:
        nbytes = raster_bytes(pipe)
        achieved = nbytes / (stages[dominant] * 1e-3) / 1e9
        traffic = measured_traffic() if (dominant == 'rasterizer' and args.workload == 'full') else None
        roofline = {'kernel': 'k_raster (rf_raster_draw)' if dominant == 'rasterizer' else dominant,
                    'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                    'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic,
                    'algorithmic_bytes': nbytes,
                    'traffic_source': 'profiles/hbm_traffic_full_frame.txt' if traffic else None}
    trace_tflops = trace_flops(pipe) / (stages['raytrace'] * 1e-3) / 1e12

    line = {
        'metric': 'rays traced/s & ms per 2048^2 flare frame, batched vs plain; adjoint ms/frame',
        'value': pipe.rays / (ms * 1e-3),
        'unit': 'rays/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms,
        'frames_per_s': 1000.0 / ms,
        'frame_launch': 'hip graph replay' if graph is not None else 'eager',
        'eager_ms_per_step': eager_ms,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded 12-surface lens, reference default render settings)',
        'config': {'workload': f'{args.workload}: FLARE_STARBURST frame, {pipe.P} ghosts x '
                               f'{pipe.W} wavelengths x {pipe.g}^2 rays, '
                               f'{pipe.resolution[0]}x{pipe.resolution[1]}, starburst '
                               f'{pipe.sres[0]}^2 x {pipe.star_prm.samples} samples',
                   'rays_per_frame': pipe.rays, 'binning_pairs': pairs},
        'wall_s': wall,
        'stages_ms': stages,
        'trace_rays_per_s': pipe.rays / (stages['raytrace'] * 1e-3),
        'trace_tflops': trace_tflops,
        'roofline': roofline,
        'trace_roofline': {'kernel': 'k_trace (rf_trace_planes)', 'bound': 'valu',
                           'achieved': trace_tflops, 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                           'frac': trace_tflops / FP32_PEAK_TFLOPS},
    }
    if args.workload == 'full':
        line['adjoint'] = adjoint_timing(wl, args.steps)
        print(f"# adjoint: {line['adjoint']}", file=sys.stderr, flush=True)
    if args.workload == 'full' and not args.no_batched:
        line['batched'] = batched_timing(max(1, min(args.steps, 3)), 1, args.batch_frames)
        print(f"# batched: {line['batched']}", file=sys.stderr, flush=True)
    if args.workload == 'full' and not args.no_fit:
        line['batched_fit'] = fit_timing()
        print(f"# batched fit: {line['batched_fit']}", file=sys.stderr, flush=True)
    if args.workload == 'full' and not args.no_reference_benchmark:
        line['reference_benchmark'] = reference_benchmark()
        print(f"# reference benchmark: {line['reference_benchmark']}", file=sys.stderr, flush=True)
    if args.copy_calibration:
        src = torch.empty((64 << 20,), dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        line['copy_calibration_bytes'] = 2 * src.numel() * 4
    if not args.no_cpu_baseline:
        print('# cpu baseline (oracle on a bounded sample)...', file=sys.stderr, flush=True)
        line['cpu_baseline'] = cpu_baseline(pipe, args.workload)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
