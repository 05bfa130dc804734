"""Shared setup for the adjoint tests: lens parameters of the standard
workloads and an FP64 loss evaluated with the C oracle."""
from __future__ import annotations

import numpy as np
import torch

from helpers import trace_inputs
from realflare_amd.adjoint import lens_params


def setup(name='small', grid_count=9):
    inp = trace_inputs(name, grid_count=grid_count)
    wl = inp['workload']
    p = wl.project
    params, static = lens_params(wl.lens_model, inp['sensor'], p.flare.lens.glasses_path,
                                 p.flare.lens.abbe_nr_adjustment, p.flare.lens.coating)
    return inp, params, static


def element_arrays(params, static):
    """Per-element double arrays (radius, ior, height, centre, coating, coef[L][6])."""
    n = static.n_model
    radii = params.radii.detach().double().numpy()
    dists = params.distances.detach().double().numpy()
    b = params.sellmeier_b.detach().double().numpy()
    coat = params.coating.detach().double().numpy()
    radius = np.append(radii, 0.0)
    center = np.concatenate([[0.0], np.cumsum(dists)]) + radius
    coating = np.append(coat, static.sensor_coating)
    coef = np.zeros((n + 1, 6))
    for i in range(n):
        if np.any(b[i] != 0) or static.has_glass[i]:
            coef[i, 0::2] = b[i]
            coef[i, 1::2] = static.sellmeier_c[i]
    return radius, static.ior, static.height, center, coating, coef


def oracle_outputs(oracle, params, static, inp, paths=None, wavelengths=None, grid_count=None):
    radius, ior, height, center, coating, coef = element_arrays(params, static)
    rays = oracle.trace_params(radius, ior, height, center, coating, coef, static.disperse,
                               inp['paths'] if paths is None else paths,
                               inp['wavelengths'] if wavelengths is None else wavelengths,
                               static.aperture_index, inp['coating_min_ior'],
                               grid_count or inp['grid_count'], inp['grid_length'],
                               inp['direction'])
    return rays['pos'][..., 0], rays['pos'][..., 1], rays['reflectance']


def np_loss(out, target, mask, refl_scale):
    x, y, r = out
    xt, yt, rt = target
    dx, dy, dr = x - xt, y - yt, (r - rt) / refl_scale
    return float(np.sum((dx * dx + dy * dy + dr * dr)[mask]) / max(mask.sum(), 1))


def directions(params, n, seed):
    """Random parameter directions scaled to each parameter's magnitude
    (relative perturbations); the stop radius never moves."""
    rng = np.random.default_rng(seed)
    flat = params.flat().detach().double().numpy()
    scale = np.where(flat != 0, np.abs(flat), 0.0)
    return [rng.standard_normal(flat.shape) * scale for _ in range(n)]
