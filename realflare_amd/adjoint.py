"""Differentiable ghost trace: lens parameters -> sensor positions and
intensities, with the hand-written HIP adjoint as the backward pass.

Parameters (float64 master copies, one entry per model element):
  radii       surface radius of curvature (mm); the stop's 0 stays fixed
  distances   thickness / air gap after each element (mm)
  sellmeier_b B1..B3 of each element's catalog glass (C stays fixed)
  coating     anti-reflection coating design wavelength (nm), which sets
              the coating thickness d = lambda_c / (4 n_c) (raytracing.cl:316-335)

They are packed exactly as ``realflare/api/lens.py:40-86`` packs a model
(centre = sum of previous distances + radius, sensor appended).
``GhostTrace.apply`` runs ``rf_trace_forward_checkpointed``, whose outputs are
bit-identical to ``rf_trace_planes_coated``'s, and returns (x, y, reflectance)
per ray with NaN for rays that do not reach the sensor; the checkpoints it
stores are the tape.  The backward pass runs ``rf_trace_backward`` on that tape
and chains its per-element gradients (radius, centre, coating, effective index per wavelength) to the
parameters above: centres through the prefix sums of distances, effective
indices through the Sellmeier equation.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from . import native
from .api import lens as api_lens
from .api import glass
from .api import storage
from .api.data import LensModel, RealflareError
from .api.dtypes import lens_element_dtype
from .api.tasks.raytracing import trace_params


@dataclasses.dataclass
class LensParams:
    radii: torch.Tensor
    distances: torch.Tensor
    sellmeier_b: torch.Tensor
    coating: torch.Tensor

    def tensors(self):
        return self.radii, self.distances, self.sellmeier_b, self.coating

    def flat(self) -> torch.Tensor:
        return torch.cat([t.reshape(-1) for t in self.tensors()])

    @staticmethod
    def unflat(v: torch.Tensor, like: 'LensParams') -> 'LensParams':
        out, i = [], 0
        for t in like.tensors():
            n = t.numel()
            out.append(v[i:i + n].reshape(t.shape))
            i += n
        return LensParams(*out)


@dataclasses.dataclass
class LensStatic:
    """Everything of the packed lens that is not differentiated."""
    ior: np.ndarray           # [L]
    height: np.ndarray        # [L]
    sellmeier_c: np.ndarray   # [L, 3]
    has_glass: np.ndarray     # [L] bool: Sellmeier B is a parameter
    disperse: bool
    aperture_index: int
    focal_length: float
    n_model: int              # model elements (L - 1)
    sensor_coating: float = 0.0


def lens_params(lens_model: LensModel, sensor_size, glasses_path: str, abbe_nr_adjustment: float,
                coating) -> tuple[LensParams, LensStatic]:
    """Initial parameters and static data, as lens.elements packs them."""
    elements = list(lens_model.lens_elements)
    n = len(elements)
    glasses = tuple(glass.glasses_from_path(storage.decode_path(glasses_path)))
    radii = np.array([e.radius for e in elements], np.float64)
    dists = np.array([e.distance for e in elements], np.float64)
    b = np.zeros((n + 1, 3))
    c = np.zeros((n + 1, 3))
    has_glass = np.zeros(n + 1, bool)
    for i, e in enumerate(elements):
        g = glass.closest_glass(glasses, e.refractive_index, e.abbe_nr, abbe_nr_adjustment)
        if g:
            b[i] = g.coefficients[0::2]
            c[i] = g.coefficients[1::2]
            has_glass[i] = e.refractive_index > 1
    coat = np.zeros(n + 1)
    for i in range(min(len(coating), n + 1)):
        coat[i] = coating[i]
    ior = np.array([e.refractive_index for e in elements] + [1.0])
    height = np.array([e.height for e in elements] + [np.linalg.norm(sensor_size) / 2])
    params = LensParams(torch.tensor(radii), torch.tensor(dists), torch.tensor(b[:n]),
                        torch.tensor(coat[:n]))
    static = LensStatic(ior, height, c, has_glass, bool(glasses), lens_model.aperture_index,
                        float(lens_model.focal_length), n, float(coat[n]))
    return params, static


def pack(params: LensParams, static: LensStatic) -> tuple[np.ndarray, np.ndarray]:
    """Packed LensElement records (float32) and coating wavelengths (float32)."""
    n = static.n_model
    radii = params.radii.detach().double().cpu().numpy()
    dists = params.distances.detach().double().cpu().numpy()
    b = params.sellmeier_b.detach().double().cpu().numpy()
    coat = params.coating.detach().double().cpu().numpy()
    arr = np.zeros(n + 1, lens_element_dtype)
    offset = 0.0
    for i in range(n + 1):
        r = radii[i] if i < n else 0.0
        arr[i]['radius'] = r
        arr[i]['ior'] = static.ior[i]
        arr[i]['height'] = static.height[i]
        arr[i]['center'] = offset + r
        arr[i]['coating'] = int(coat[i]) if i < n else int(static.sensor_coating)
        if i < n:
            cf = np.zeros(8)
            cf[0:6:2] = b[i]
            cf[1:6:2] = static.sellmeier_c[i]
            if np.any(cf[:6] != 0) or static.has_glass[i]:
                arr[i]['coefficients'] = cf
            offset += dists[i]
    if not static.disperse:
        arr[0]['coefficients'][0] = np.nan
    coating = np.zeros(n + 1, np.float32)
    coating[:n] = coat
    coating[n] = static.sensor_coating
    return arr, coating


def dn_db(b: np.ndarray, c: np.ndarray, wavelength: float) -> tuple[float, np.ndarray]:
    """Sellmeier n(lambda) and dn/dB (lambda in nm; l = float32(lambda * 1e-3))."""
    l = float(np.float32(wavelength * 1e-3))
    l2 = l * l
    terms = l2 / (l2 - c)
    n = np.sqrt(1 + np.sum(b * terms))
    return n, terms / (2 * n)


class TraceSpec:
    """Non-differentiable inputs of one traced batch, resident on the device."""

    def __init__(self, static: LensStatic, paths, wavelengths, coating_min_ior: float,
                 grid_count: int, grid_length: float, direction, checkpoint_every: int = 2,
                 device=None) -> None:
        self.static = static
        self.paths = [tuple(p) for p in paths]
        L = static.n_model + 1
        for b1, b2 in self.paths:
            if b1 < 0 and b2 < 0:
                continue
            if not (b1 > b2 >= 0 and b1 < L - 1 and static.aperture_index not in (b1, b2)):
                raise RealflareError(f'adjoint: unsupported bounce pair {(b1, b2)}')
        self.wavelengths = [int(w) for w in wavelengths]
        self.device = device or torch.device('cuda', torch.cuda.current_device())
        self.paths_t = torch.tensor(self.paths, dtype=torch.int32, device=self.device).reshape(-1, 2)
        self.wl_t = torch.tensor(self.wavelengths, dtype=torch.int32, device=self.device)
        self.prm = trace_params(L, len(self.paths), len(self.wavelengths), static.aperture_index,
                                coating_min_ior, grid_count, grid_length, direction)
        self.K = int(checkpoint_every)
        size = native.lib().rf_trace_adjoint_workspace_size
        size.restype = ctypes.c_size_t
        self.ws_bytes = size(ctypes.byref(self.prm), self.K)
        self.workspace = None
        self.shape = (len(self.paths), len(self.wavelengths), grid_count * grid_count)

    def upload(self, params: LensParams):
        arr, coating = pack(params, self.static)
        lens = torch.from_numpy(arr.view(np.uint8).copy()).to(self.device)
        coat = torch.from_numpy(coating).to(self.device)
        return lens, coat

    def forward(self, params: LensParams) -> torch.Tensor:
        lens, coat = self.upload(params)
        P, W, R = self.shape
        planes = torch.empty((6, P, W, R), dtype=torch.float32, device=self.device)
        native.call('rf_trace_planes_coated', native.ptr(lens), native.ptr(coat),
                    native.ptr(self.paths_t), native.ptr(self.wl_t), ctypes.byref(self.prm),
                    native.ptr(planes), native.stream())
        return planes

    def _grad_out(self) -> torch.Tensor:
        L = self.static.n_model + 1
        return torch.empty((3 * L + len(self.wavelengths) * L,), dtype=torch.float32,
                           device=self.device)

    def _common(self, lens, coat):
        return (native.ptr(lens), native.ptr(coat), native.ptr(self.paths_t),
                native.ptr(self.wl_t), ctypes.byref(self.prm), self.K)

    def adjoint(self, params: LensParams, gx, gy, gr, forward_out=None) -> np.ndarray:
        """Forward and reverse pass in one launch (rf_trace_adjoint)."""
        lens, coat = self.upload(params)
        if self.workspace is None:
            self.workspace = torch.empty((self.ws_bytes,), dtype=torch.uint8, device=self.device)
        out = self._grad_out()
        native.call('rf_trace_adjoint', *self._common(lens, coat), native.ptr(gx), native.ptr(gy),
                    native.ptr(gr), native.ptr(self.workspace), ctypes.c_size_t(self.ws_bytes),
                    native.ptr(out),
                    native.ptr(forward_out) if forward_out is not None else None,
                    native.stream())
        return out.double().cpu().numpy()

    def forward_checkpointed(self, params: LensParams):
        """(x, y, reflectance) per ray and the tape (a workspace holding the
        checkpoints) that ``backward`` consumes."""
        lens, coat = self.upload(params)
        tape = torch.empty((self.ws_bytes,), dtype=torch.uint8, device=self.device)
        fwd = torch.empty((3,) + self.shape, dtype=torch.float32, device=self.device)
        native.call('rf_trace_forward_checkpointed', *self._common(lens, coat), native.ptr(tape),
                    ctypes.c_size_t(self.ws_bytes), native.ptr(fwd), native.stream())
        return fwd, (lens, coat, tape)

    def backward(self, tape, gx, gy, gr) -> np.ndarray:
        """Per-element gradients from a tape of ``forward_checkpointed``."""
        lens, coat, ws = tape
        out = self._grad_out()
        native.call('rf_trace_backward', *self._common(lens, coat), native.ptr(gx),
                    native.ptr(gy), native.ptr(gr), native.ptr(ws), ctypes.c_size_t(self.ws_bytes),
                    native.ptr(out), native.stream())
        return out.double().cpu().numpy()

    def chain(self, params: LensParams, g: np.ndarray) -> tuple[torch.Tensor, ...]:
        """Per-element gradients (radius, centre, coating, n_eff[w]) -> parameters."""
        return chain_grads(self.static, self.wavelengths, params, g)


def chain_grads(st: LensStatic, wavelengths, params: LensParams,
                g: np.ndarray) -> tuple[torch.Tensor, ...]:
    """Per-element gradients laid out [radius L][centre L][coating L][n_eff W x L]
    -> gradients of (radii, distances, Sellmeier B, coating)."""
    n, L, W = st.n_model, st.n_model + 1, len(wavelengths)
    g_rad, g_cen, g_coat = g[:L], g[L:2 * L], g[2 * L:3 * L]
    g_n = g[3 * L:3 * L + W * L].reshape(W, L)
    radii = g_rad[:n] + g_cen[:n]
    # centre_i = sum_{k<i} d_k + r_i  ->  dL/dd_k = sum_{i>k} dL/dcentre_i
    suffix = np.cumsum(g_cen[::-1])[::-1].copy()
    dists = suffix[1:L].copy()
    b = params.sellmeier_b.detach().double().cpu().numpy()
    gb = np.zeros((n, 3))
    if st.disperse:
        for i in range(n):
            if not st.has_glass[i]:
                continue
            for w, lam in enumerate(wavelengths):
                _, d = dn_db(b[i], st.sellmeier_c[i], lam)
                gb[i] += g_n[w, i] * d
    radii = radii.copy()
    radii[st.aperture_index] = 0.0
    return (torch.tensor(radii), torch.tensor(dists), torch.tensor(gb),
            torch.tensor(g_coat[:n].copy()))


class GhostTrace(torch.autograd.Function):
    """(radii, distances, sellmeier_b, coating) -> (x, y, reflectance) per ray."""

    @staticmethod
    def forward(ctx, radii, distances, sellmeier_b, coating, spec: TraceSpec):
        params = LensParams(radii, distances, sellmeier_b, coating)
        fwd, tape = spec.forward_checkpointed(params)
        ctx.spec = spec
        ctx.tape = tape
        ctx.params = LensParams(*(t.detach().clone() for t in params.tensors()))
        return fwd[0], fwd[1], fwd[2]

    @staticmethod
    def backward(ctx, gx, gy, gr):
        spec: TraceSpec = ctx.spec
        dev = spec.device

        def clean(t):
            if t is None:
                return torch.zeros(spec.shape, dtype=torch.float32, device=dev)
            return torch.nan_to_num(t.float(), nan=0.0, posinf=0.0, neginf=0.0).contiguous()

        g = spec.backward(ctx.tape, clean(gx), clean(gy), clean(gr))
        ctx.tape = None
        grads = spec.chain(ctx.params, g)
        return (*[gr_.to(p.dtype) for gr_, p in zip(grads, ctx.params.tensors())], None)


def ghost_loss(x, y, r, target, refl_scale: float):
    """Mean squared error of sensor positions plus scaled reflectances over
    rays that reach the sensor in both the trace and the target.  A ray that
    was stopped keeps its last position and carries a NaN reflectance, so
    all three outputs of both sides decide the mask."""
    xt, yt, rt = target
    mask = ~(torch.isnan(x) | torch.isnan(y) | torch.isnan(r) |
             torch.isnan(xt) | torch.isnan(yt) | torch.isnan(rt))
    n = mask.sum().clamp(min=1)
    z = torch.zeros_like(x)
    dx = torch.where(mask, x - xt, z)
    dy = torch.where(mask, y - yt, z)
    dr = torch.where(mask, (r - rt) / refl_scale, z)
    return (dx * dx + dy * dy + dr * dr).sum() / n


def perturbed(params: LensParams, static: LensStatic, rel: float, seed: int) -> LensParams:
    """Seeded relative perturbation of radii, distances, Sellmeier B and coating."""
    rng = np.random.default_rng(seed)
    out = []
    for t in params.tensors():
        v = t.detach().double().cpu().numpy()
        out.append(torch.tensor(v * (1 + rel * rng.standard_normal(v.shape))))
    p = LensParams(*out)
    p.radii[static.aperture_index] = 0.0
    return p
