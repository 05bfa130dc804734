"""Autotuner of the batched form: picks rays per lane and persistent
workgroups for rf_trace_batch, and the rasterizer's XCD super-block size
(the bin edge in pixels, 8 x super_block) and frames per raster pass, by
timing candidates on the batch itself (or a sample of its frames), and
writes the choice to a JSON file keyed by device and workload.
"""
from __future__ import annotations

import dataclasses
import json
import os
import time

import torch

from .batch import BatchRenderer, Tuning

RAYS_PER_LANE = (1, 2, 4)
WORKGROUPS_PER_CU = (4, 8, 16)
SUPER_BLOCKS = (1, 2, 4, 8)  # bin edges of 8, 16, 32 and 64 px
FRAMES_PER_PASS = (1, 2, 4, 8)


def default_path() -> str:
    base = os.environ.get('XDG_CACHE_HOME') or os.path.join(os.path.expanduser('~'), '.cache')
    return os.path.join(base, 'realflare_amd', 'autotune.json')


def _time(fn, reps: int = 3) -> float:
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune(br: BatchRenderer, reps: int = 3, path: str | None = None, key: str | None = None) -> Tuning:
    """Time the candidates on `br` (its frames and grids), leave `br` set to
    the fastest, write {key: choice and timings} into `path`."""
    cus = torch.cuda.get_device_properties(br.device).multi_processor_count
    base = br.tuning
    trace_ms = {}
    for rpl in RAYS_PER_LANE:
        for wpc in WORKGROUPS_PER_CU:
            t = dataclasses.replace(base, rays_per_lane=rpl, trace_workgroups=wpc * cus)
            br.set_tuning(t)
            trace_ms[(rpl, t.trace_workgroups)] = _time(br.trace, reps)
    (rpl, wgs) = min(trace_ms, key=trace_ms.get)
    base = dataclasses.replace(base, rays_per_lane=rpl, trace_workgroups=wgs)
    br.set_tuning(base)
    br.trace()
    br.shade()
    raster_ms = {}
    for sb in SUPER_BLOCKS:
        for fpp in FRAMES_PER_PASS:
            if fpp > br.F:
                continue
            t = dataclasses.replace(base, super_block=sb, frames_per_pass=fpp)
            br.set_tuning(t)
            br.fit_capacity()

            def passes():
                for d in br.passes:
                    br.raster_pass(d)

            raster_ms[(sb, fpp)] = _time(passes, reps) / br.F
    (sb, fpp) = min(raster_ms, key=raster_ms.get)
    best = dataclasses.replace(base, super_block=sb, frames_per_pass=fpp)
    br.set_tuning(best)
    br.fit_capacity()
    record = {
        'device': torch.cuda.get_device_name(br.device),
        'tuning': dataclasses.asdict(best),
        'bin_pixels': best.bin_pixels(),
        'trace_ms': {f'rpl{k[0]}_wg{k[1]}': v for k, v in trace_ms.items()},
        'raster_ms_per_frame': {f'sb{k[0]}_fpp{k[1]}': v for k, v in raster_ms.items()},
        'time': time.strftime('%Y-%m-%dT%H:%M:%S'),
    }
    path = path or default_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = {}
        if os.path.exists(path):
            with open(path) as f:
                data = json.load(f)
        data[key or 'default'] = record
        with open(path, 'w') as f:
            json.dump(data, f, indent=1)
    except OSError:
        pass
    br.autotune_record = record
    return best


def load(key: str, path: str | None = None) -> Tuning | None:
    """A previously written choice for `key`, if any."""
    try:
        with open(path or default_path()) as f:
            rec = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    return Tuning(**rec['tuning']) if rec else None
