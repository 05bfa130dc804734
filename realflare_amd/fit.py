"""Batched lens fit: several lens instances, each seen from several light
positions, all ghosts with their own grids and all wavelengths, traced forward
and differentiated in one persistent launch per gradient step
(`rf_trace_fit`), then an Adam step on every instance's parameters.

Per step the kernel evaluates, for every ray that reaches the sensor in both
the current trace and the target,
    loss_scale * ((x - x_t)^2 + (y - y_t)^2 + ((r - r_t) / refl_scale)^2)
and back-propagates it through the ray in place, so no per-ray gradient array
is materialized.  Targets are traces of perturbed parameters made by the same
kernel.  Gradients are chained to (radii, thicknesses, Sellmeier B, coating
wavelength) per instance exactly as `realflare_amd.adjoint.GhostTrace` does.
The reference has no gradients; its trace semantics are those of
raytracing.cl:218-359.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np
import torch

from . import native
from .adjoint import LensParams, LensStatic, chain_grads, lens_params, pack
from .api.data import LensModel
from .api.dtypes import lens_element_dtype
from .api.lens import all_pair_paths
from .api.tasks.raytracing import light_direction
from .batch import SEGMENT_DTYPE, BatchDesc
from .synthetic import lens_grid_length

LANES = 64


@dataclasses.dataclass
class FitInstance:
    model: LensModel
    params: LensParams
    static: LensStatic
    paths: list
    grids: np.ndarray       # per ghost
    positions: list         # light positions (normalized screen coordinates)


def _size(name):
    f = getattr(native.lib(), name)
    f.restype = ctypes.c_size_t
    return f


class FitBatch:
    """Device layout of a batched fit: lens records of every instance
    (concatenated), ghost paths (concatenated), and segments ordered by
    instance, then light position, then ghost."""

    def __init__(self, instances: list[FitInstance], wavelengths, resolution, sensor,
                 coating_min_ior: float = 1.38, checkpoint_every: int = 4,
                 rays_per_lane: int = 4, workgroups: int | None = None, device=None) -> None:
        self.device = dev = device or torch.device('cuda', torch.cuda.current_device())
        self.inst = instances
        self.wavelengths = [int(w) for w in wavelengths]
        self.W = len(self.wavelengths)
        self.K = int(checkpoint_every)
        self.rpl = int(rays_per_lane)
        self.cmin = float(coating_min_ior)
        self.I = len(instances)
        self.LF = max(i.static.n_model + 1 for i in instances)
        cus = torch.cuda.get_device_properties(dev).multi_processor_count
        self.workgroups = int(workgroups or 8 * cus)

        table, paths, segs = [], [], []
        first = path_off = ro = qo = frame = 0
        seg_off = [0]
        for k, it in enumerate(instances):
            L = it.static.n_model + 1
            table.append((first, L, it.static.aperture_index))
            paths += [tuple(p) for p in it.paths]
            gl = lens_grid_length(it.model)
            for pos in it.positions:
                d = light_direction(pos, resolution, sensor, it.model.focal_length)
                for p, g in enumerate(it.grids):
                    g = int(g)
                    segs.append((ro, qo, frame, path_off + p, k, g, gl, d[:3]))
                    ro += self.W * g * g
                    qo += (g - 1) * (g - 1)
                frame += 1
            seg_off.append(len(segs))
            first += L
            path_off += len(it.paths)
        seg = np.zeros(len(segs), SEGMENT_DTYPE)
        for j, (r_, q_, f_, p_, l_, g_, gl_, d_) in enumerate(segs):
            seg[j]['ray_offset'], seg[j]['quad_offset'], seg[j]['frame'] = r_, q_, f_
            seg[j]['path'], seg[j]['lens'], seg[j]['grid_count'] = p_, l_, g_
            seg[j]['grid_length'] = np.float32(gl_)
            seg[j]['direction'] = np.asarray(d_, np.float32)
        self.seg = seg
        self.rays = int(ro)
        self.n_elements = first
        self.seg_t = torch.from_numpy(seg.view(np.uint8).copy()).to(dev)
        per = LANES * self.rpl
        g = seg['grid_count'].astype(np.int64)
        n = self.W * ((g * g + per - 1) // per)
        offs = np.concatenate([[0], np.cumsum(n)]).astype(np.uint64)
        self.chunk_count = int(offs[-1])
        self.chunks_t = torch.from_numpy(offs.view(np.int64)).to(dev)
        self.seg_off_t = torch.tensor(np.asarray(seg_off, np.int64), device=dev)
        self.table_t = torch.tensor(table, dtype=torch.int32, device=dev)
        self.paths_t = torch.tensor(paths, dtype=torch.int32, device=dev).reshape(-1, 2)
        self.wl_t = torch.tensor(self.wavelengths, dtype=torch.int32, device=dev)
        self.desc = BatchDesc(self.seg_t.data_ptr(), len(seg), 0, frame, 0, 0, self.rays, 0,
                              int(qo), self.rays)
        self.lens_t = torch.empty((first * 64,), dtype=torch.uint8, device=dev)
        self.coat_t = torch.empty((first,), dtype=torch.float32, device=dev)
        ws = _size('rf_trace_fit_workspace_size')(self.LF, self.K, self.workgroups,
                                                  ctypes.c_uint64(self.chunk_count))
        if ws == 0:
            raise ValueError('rf_trace_fit_workspace_size rejected the batch')
        self.ws = torch.empty((ws,), dtype=torch.uint8, device=dev)
        self.nout = (3 + self.W) * self.LF + 2
        self.grad = torch.empty((self.I, self.nout), dtype=torch.float32, device=dev)
        self.fwd = torch.empty((3, self.rays), dtype=torch.float32, device=dev)
        nan = torch.full((self.rays,), float('nan'), dtype=torch.float32, device=dev)
        self.targets = (nan, nan, nan)
        self.refl_scale = 1.0
        self.loss_scale = 1.0
        self.huber_delta = 0.0  # 0: squared error; > 0: Huber beyond this residual
        # half extents in mm of the sensor region the image shows: pixel =
        # pos * st + res / 2 with st = width / (sensor diagonal / 2)
        # (rasterizing.py:158-167, rasterizing.cl:368)
        st = resolution[0] / (math.hypot(sensor[0], sensor[1]) / 2)
        self.window = (resolution[0] / 2 / st, resolution[1] / 2 / st)
        # a ray that reaches the sensor in the target but not in the current
        # trace adds this (times loss_scale) to the loss, with no gradient:
        # the squared image diagonal, as far as a ray in the image can be
        # from its target.  Without it, parameters that lose every ray would
        # score a loss of 0.
        self.lost_penalty = 4.0 * (self.window[0] ** 2 + self.window[1] ** 2)
        first_ray = [int(seg[seg_off[k]]['ray_offset']) if seg_off[k] < len(seg) else self.rays
                     for k in range(self.I)]
        self.instance_rays = [(first_ray[k], first_ray[k + 1] if k + 1 < self.I else self.rays)
                              for k in range(self.I)]

    def upload(self, params: list[LensParams]) -> None:
        recs = np.zeros(self.n_elements, lens_element_dtype)  # keeps the 64-byte records
        coats = np.zeros(self.n_elements, np.float32)
        first = 0
        for it, p in zip(self.inst, params):
            arr, coat = pack(p, it.static)
            recs[first:first + len(arr)] = arr
            coats[first:first + len(arr)] = coat
            first += len(arr)
        self.lens_t.copy_(torch.from_numpy(recs.view(np.uint8).copy()))
        self.coat_t.copy_(torch.from_numpy(coats))

    def launch(self) -> None:
        tx, ty, tr = self.targets
        p = native.ptr
        native.call('rf_trace_fit', p(self.lens_t), p(self.coat_t), p(self.table_t), p(self.paths_t),
                    p(self.wl_t), self.W, ctypes.c_float(self.cmin), ctypes.byref(self.desc),
                    p(self.chunks_t), ctypes.c_uint64(self.chunk_count), p(self.seg_off_t), self.I,
                    self.LF, self.rpl, self.K, self.workgroups, p(tx), p(ty), p(tr),
                    ctypes.c_float(self.refl_scale), ctypes.c_float(self.loss_scale),
                    ctypes.c_float(self.huber_delta), p(self.ws),
                    ctypes.c_size_t(self.ws.numel()), p(self.grad), p(self.fwd), native.stream())

    def set_targets(self, params: list[LensParams], image_window: bool = True) -> None:
        """Targets = the forward trace of `params` (made by the same kernel
        with no target, so it runs no reverse pass).  With `image_window`,
        only rays whose target lands inside the rendered image are fitted
        (the others get no target): ghost rays that reach the sensor plane at
        grazing angles land up to ~1e6 mm away, never show in the flare, and
        would otherwise dominate the squared error."""
        nan = torch.full((self.rays,), float('nan'), dtype=torch.float32, device=self.device)
        self.targets = (nan, nan, nan)
        self.upload(params)
        self.launch()
        t = self.fwd.clone()
        if image_window:
            xw, yw = self.window
            for i in range(0, self.rays, 1 << 26):
                x, y, r = t[0, i:i + (1 << 26)], t[1, i:i + (1 << 26)], t[2, i:i + (1 << 26)]
                r.masked_fill_((x.abs() > xw) | (y.abs() > yw), float('nan'))
        self.targets = (t[0], t[1], t[2])
        # RMS of the target reflectances and their count, in chunks (the
        # batch can hold billions of rays)
        total, count = 0.0, 0
        r = t[2]
        for i in range(0, r.numel(), 1 << 26):
            c = r[i:i + (1 << 26)]
            ok = ~torch.isnan(c)
            total += float(torch.where(ok, c * c, torch.zeros_like(c)).double().sum())
            count += int(ok.sum())
        self.refl_scale = math.sqrt(total / count) if count else 1.0
        self.loss_scale = 1.0 / max(count, 1)

    def residual_quantile(self, params: list[LensParams], q: float) -> float:
        """Quantile q of the position residuals |x - x_target|, |y - y_target|
        over the rays that reach the sensor in both (a strided sample of at most
        2^24 values per component)."""
        self.upload(params)
        self.launch()
        vals = []
        stride = max(1, self.rays // (1 << 24))
        for c in range(2):
            r = (self.fwd[c, ::stride] - self.targets[c][::stride]).abs()
            vals.append(r[~torch.isnan(r)])
        v = torch.cat(vals)
        return float(torch.quantile(v.float(), q)) if v.numel() else 0.0

    def lost_rays(self) -> list[int]:
        """Per instance: rays with a target that the last launch's trace lost
        (did not reach the sensor, or reached it at a non-finite position)."""
        tr = self.targets[2]
        out = []
        for lo, hi in self.instance_rays:
            n = 0
            for i in range(lo, hi, 1 << 26):
                j = min(hi, i + (1 << 26))
                f = self.fwd[:, i:j]
                gone = torch.isnan(f[2]) | ~torch.isfinite(f[0]) | ~torch.isfinite(f[1])
                n += int((~torch.isnan(tr[i:j]) & gone).sum())
            out.append(n)
        return out

    def step(self, params: list[LensParams]):
        """Loss per instance (with the lost-ray penalty) and parameter
        gradients per instance."""
        self.upload(params)
        self.launch()
        g = self.grad.double().cpu().numpy()
        # the launch counts each instance's lost rays (row[(3 + W) LF + 1])
        losses, grads = [], []
        LF, W = self.LF, self.W
        for k, (it, p) in enumerate(zip(self.inst, params)):
            L = it.static.n_model + 1
            row = g[k]
            compact = np.concatenate([row[b * LF:b * LF + L] for b in range(3 + W)])
            grads.append(chain_grads(it.static, self.wavelengths, p, compact))
            lost = row[(3 + W) * LF + 1]
            losses.append(row[(3 + W) * LF] + self.loss_scale * self.lost_penalty * lost)
        return np.array(losses), grads


class Adam:
    """Adam on parameters measured relative to their initial magnitude, with
    one step size per lens instance."""

    def __init__(self, params: list[LensParams], lr: float = 1e-4, betas=(0.9, 0.999),
                 eps: float = 1e-12) -> None:
        self.lr = [lr] * len(params)
        self.b1, self.b2, self.eps = betas[0], betas[1], eps
        self.scale = [self._scale(p) for p in params]
        self.m = [np.zeros_like(s) for s in self.scale]
        self.v = [np.zeros_like(s) for s in self.scale]
        self.t = [0] * len(params)

    @staticmethod
    def _scale(p: LensParams) -> np.ndarray:
        f = p.flat().detach().double().numpy()
        return np.where(f != 0, np.abs(f), 1.0)

    def restart(self, k: int, lr: float) -> None:
        """Forget instance k's moments and continue with step size lr."""
        self.lr[k] = lr
        self.m[k][:] = 0
        self.v[k][:] = 0
        self.t[k] = 0

    def update(self, params: list[LensParams], grads) -> list[LensParams]:
        """One step per instance; an instance whose gradient is not finite
        keeps its parameters and moments."""
        out = []
        for k, (p, g) in enumerate(zip(params, grads)):
            if not all(bool(torch.isfinite(x).all()) for x in g):
                out.append(p)
                continue
            self.t[k] += 1
            t = self.t[k]
            gf = torch.cat([x.reshape(-1) for x in g]).double().numpy() * self.scale[k]
            self.m[k] = self.b1 * self.m[k] + (1 - self.b1) * gf
            self.v[k] = self.b2 * self.v[k] + (1 - self.b2) * gf * gf
            mh = self.m[k] / (1 - self.b1 ** t)
            vh = self.v[k] / (1 - self.b2 ** t)
            step = self.lr[k] * mh / (np.sqrt(vh) + self.eps) * self.scale[k]
            flat = p.flat().detach().double().numpy() - step
            out.append(LensParams.unflat(torch.tensor(flat), p))
        return out


def perturb(params: LensParams, static: LensStatic, rel: float, seed: int) -> LensParams:
    """Seeded relative perturbation (the stop radius stays 0)."""
    from .adjoint import perturbed

    return perturbed(params, static, rel, seed)


def make_instances(models: list[LensModel], positions, grid_range=(64, 128), sensor=(36.0, 24.0),
                   glasses_path='$GLASS/schott', abbe_nr_adjustment=0.0, seed: int = 0):
    """Instances with all refracting-surface ghost pairs and a seeded grid per
    ghost in `grid_range`."""
    rng = np.random.default_rng(seed)
    out = []
    for m in models:
        n = len(m.lens_elements)
        params, static = lens_params(m, sensor, glasses_path, abbe_nr_adjustment, [550] * (n + 1))
        paths = all_pair_paths(m)
        grids = rng.integers(grid_range[0], grid_range[1] + 1, size=len(paths))
        out.append(FitInstance(m, params, static, paths, grids, list(positions)))
    return out


def fit(batch: FitBatch, start: list[LensParams], steps: int, lr: float = 1e-4,
        on_step=None, patience: float = 2.0, robust: float = 0.0
        ) -> tuple[list[LensParams], list[dict]]:
    """`steps` Adam steps from `start`.  The instances are independent, so
    each keeps its own best point and step size.  The loss is not convex
    (ghost rays appear and vanish at aperture edges), so a step that raises an
    instance's loss is kept unless the loss exceeds `patience` times that
    instance's best; then the instance returns to its best point with half the
    step size.  `on_step(step, params)` runs after each gradient
    evaluation.  With `robust` = q > 0 the loss is Huber's beyond the q-quantile
    of the starting position residuals (rays whose ghost path changes between
    the trace and the target then pull linearly).  Returns the best parameters
    of every instance.  History:
    summed and per-instance losses of every evaluated point and the best
    losses so far, GPU time."""
    if robust > 0:
        batch.huber_delta = batch.residual_quantile(start, robust)
    opt = Adam(start, lr)
    params = list(start)
    n = len(start)
    best_p = list(start)
    best_l = [float('inf')] * n
    hist = []
    for s in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        losses, grads = batch.step(params)
        b.record()
        torch.cuda.synchronize()
        per = [float(x) for x in losses.reshape(-1)]
        rec = {'step': s, 'loss': sum(per), 'gpu_ms': a.elapsed_time(b)}
        if on_step is not None:
            rec.update(on_step(s, params) or {})
        restart = []
        for k in range(n):
            finite = all(bool(torch.isfinite(t).all()) for t in grads[k])
            if finite and per[k] < best_l[k]:
                best_l[k], best_p[k] = per[k], params[k]
            elif not finite or not math.isfinite(per[k]) or per[k] > patience * best_l[k]:
                restart.append(k)
        rec['best_loss'] = sum(best_l)
        rec['instance_loss'] = per
        rec['instance_best'] = list(best_l)
        hist.append(rec)
        params = opt.update(params, grads)
        for k in restart:
            params[k] = best_p[k]
            opt.restart(k, opt.lr[k] * 0.5)
    return best_p, hist
