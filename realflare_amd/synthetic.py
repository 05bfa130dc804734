"""Seeded synthetic lenses and the standard render workloads.

A synthetic lens has ``n_surfaces`` spherical refracting surfaces plus a flat
stop element.  Radii are drawn with magnitude U[20, 200] mm and a random sign,
thicknesses U[1, 15] mm, glasses by (n_d, V_d) with n_d U[1.45, 1.95] and
V_d U[20, 80] (mapped to the closest catalog glass by the lens packer, as the
reference does).  Media follow a fixed glass/air pattern so the stop sits in
air.  Candidates are drawn from ``seed, seed+1, ...`` until one has a
positive paraxial focal length, a back focal distance inside the sensor
range, clear apertures smaller than the sphere radii, and a grid that hits
the first surface; the accepted sub-seed is recorded in the model's notes.

Workloads:

* ``small``: 6 surfaces, stop at element 3, all 15 two-reflection ghosts,
  64x64 rays per ghost, wavelengths 450/550/650 nm, light at (0.3, 0.2),
  512x512 output, 6-blade starburst aperture at 256x256.
* ``full``: 12 surfaces (double-Gauss-like media, f/2 stop mid-stack), 66
  ghosts, 128x128 rays, 8 wavelengths round(400 + 300 i / 7) nm, 2048x2048
  output, 9-blade starburst aperture at 1024x1024 with 8 spectral samples.

Unspecified render settings keep the reference defaults
(``realflare/api/data.py:209-241``): bin size 64, anti-aliasing 1, one
sub-wavelength, min_area 0.01, coating_min_ior 1.38, sensor 36x24.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .api.data import Flare, LensModel, Point, Project, Render, Size

# glass (True) / air (False) medium after each refracting surface
_MEDIA = {
    6: [True, True, False, True, True, False],
    12: [True, False, True, True, True, False, True, True, True, False, True, False],
}
_STOP = {6: 3, 12: 6}


def _paraxial(radii, thicknesses, iors, y0, u0):
    """y-nu paraxial trace; returns (y, u) after the last surface and the ray
    heights at every surface."""
    y, u, n = y0, u0, 1.0
    heights = []
    for i, (r, n2) in enumerate(zip(radii, iors)):
        heights.append(y)
        c = 0.0 if r == 0 else 1.0 / r
        u = (n * u - y * (n2 - n) * c) / n2
        n = n2
        if i < len(thicknesses):
            y = y + u * thicknesses[i]
    return y, u, heights


def media_pattern(n_surfaces: int) -> tuple[list, int]:
    """Glass/air pattern after each refracting surface and the index of the
    surface the stop sits in front of.  6 and 12 surfaces use the fixed
    patterns above; other counts alternate singlets (glass, air) and cemented
    doublets (glass, glass, air), end in air, and put the stop in the air gap
    nearest the middle."""
    if n_surfaces in _MEDIA:
        return _MEDIA[n_surfaces], _STOP[n_surfaces]
    if n_surfaces < 3:
        raise ValueError('need at least 3 surfaces')
    media, k = [], 0
    while len(media) < n_surfaces:
        media += [True, False] if k % 2 == 0 else [True, True, False]
        k += 1
    media = media[:n_surfaces]
    media[-1] = False
    if not media[-2]:
        media[-2] = True
    gaps = [i for i in range(1, n_surfaces) if not media[i - 1]]
    stop_at = min(gaps, key=lambda i: abs(i - n_surfaces / 2))
    return media, stop_at


def synthetic_lens(n_surfaces: int, seed: int, fstop: float = 2.0, max_tries: int = 10000,
                   grid_margin: float = 1.15) -> LensModel:
    media, stop_at = media_pattern(n_surfaces)
    for k in range(max_tries):
        rng = np.random.default_rng(seed + k)
        radii = rng.uniform(20, 200, n_surfaces) * rng.choice([-1.0, 1.0], n_surfaces)
        thick = rng.uniform(1, 15, n_surfaces)
        nd = rng.uniform(1.45, 1.95, n_surfaces)
        vd = rng.uniform(20, 80, n_surfaces)
        stop_gap = rng.uniform(1, 15)

        # element list with the stop inserted before surface `stop_at`
        e_r, e_t, e_n, e_v = [], [], [], []
        for i in range(n_surfaces):
            if i == stop_at:
                e_r.append(0.0)
                e_t.append(stop_gap)
                e_n.append(1.0)
                e_v.append(0.0)
            e_r.append(float(radii[i]))
            e_t.append(float(thick[i]))
            e_n.append(float(nd[i]) if media[i] else 1.0)
            e_v.append(float(vd[i]) if media[i] else 0.0)

        y_end, u_end, _ = _paraxial(e_r, e_t[:-1], e_n, 1.0, 0.0)
        if u_end >= 0:
            continue
        efl = -1.0 / u_end
        bfl = -y_end / u_end
        if not (30 <= efl <= 150 and 5 <= bfl <= 80):
            continue
        e_t[-1] = bfl

        # marginal ray at f/fstop and a chief-ish ray for half the sensor
        # diagonal, to size clear apertures
        ep = efl / (2 * fstop)
        _, _, ym = _paraxial(e_r, e_t[:-1], e_n, ep, 0.0)
        _, _, yc = _paraxial(e_r, e_t[:-1], e_n, 0.0, -21.6 / efl)
        stop_h = abs(ym[stop_at])
        if stop_h <= 0.5:
            continue
        heights = []
        ok = True
        for i, r in enumerate(e_r):
            if i == stop_at:
                heights.append(stop_h)
                continue
            h = 1.2 * (abs(ym[i]) + 0.5 * abs(yc[i])) + 1.0
            if h >= 0.9 * abs(r):
                ok = False
                break
            heights.append(h)
        if not ok:
            continue
        grid_length = 2 * ep * grid_margin
        if grid_length / 2 * math.sqrt(2) >= 0.95 * abs(e_r[0]):
            continue

        elements = [
            LensModel.LensElement(
                radius=round(r, 4), distance=round(t, 4), refractive_index=round(n, 4),
                abbe_nr=round(v, 3), height=round(h, 4)
            )
            for r, t, n, v, h in zip(e_r, e_t, e_n, e_v, heights)
        ]
        return LensModel(
            name=f'synthetic-{n_surfaces}s-seed{seed}',
            notes=f'sub_seed={seed + k} efl={efl:.3f} bfl={bfl:.3f} grid_length={grid_length:.4f}',
            focal_length=int(round(efl)),
            fstop=fstop,
            aperture_index=stop_at,
            lens_elements=elements,
        )
    raise RuntimeError('no synthetic lens candidate accepted')


def lens_grid_length(model: LensModel) -> float:
    for token in model.notes.split():
        if token.startswith('grid_length='):
            return float(token.split('=')[1])
    return 50.0


@dataclasses.dataclass
class Workload:
    name: str
    project: Project
    lens_model: LensModel
    paths: list
    wavelengths: list


def _project(model: LensModel, resolution, grid_count, wavelengths, light, starburst_res,
             blades, starburst_samples, n_elements) -> Project:
    p = Project()
    p.render = Render(
        resolution=Size(*resolution),
        grid_count=grid_count,
        grid_length=lens_grid_length(model),
        wavelength_count=len(wavelengths),
    )
    p.render.starburst.resolution = Size(starburst_res, starburst_res)
    p.render.starburst.samples = starburst_samples
    p.flare = Flare()
    p.flare.light.position = Point(*light)
    p.flare.lens.glasses_path = '$GLASS/schott'
    p.flare.lens.coating = tuple([550] * (n_elements + 1))
    p.flare.lens.fstop = model.fstop
    p.flare.starburst_aperture.shape.blades = blades
    p.flare.ghost_aperture.shape.blades = blades
    return p


def workload(name: str, seed: int = 2024) -> Workload:
    from .api.lens import all_pair_paths

    if name == 'small':
        model = synthetic_lens(6, seed, fstop=4.0)
        wl = [450, 550, 650]
        proj = _project(model, (512, 512), 64, wl, (0.3, 0.2), 256, 6, 100,
                        len(model.lens_elements))
    elif name == 'full':
        model = synthetic_lens(12, seed, fstop=2.0)
        wl = [int(round(400 + 300 * i / 7)) for i in range(8)]
        proj = _project(model, (2048, 2048), 128, wl, (0.3, 0.2), 1024, 9, 8,
                        len(model.lens_elements))
    else:
        raise ValueError(f'unknown workload {name!r}')
    return Workload(name, proj, model, all_pair_paths(model), wl)
