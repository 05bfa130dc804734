"""The binner leaves out quad steps whose fragments are all exactly 0
(realflare_amd/csrc/raster.hip, k_setup): steps with a corner whose ray died
(NaN reflectance, so max(weighted reflectance, 0) is 0 at every pixel,
rasterizing.cl:498-503), and, with one sample per pixel, strictly convex steps
whose four corners map past the same edge of the ghost texture by more than
their spread plus one texel (the lookup then reads only its zero border,
rasterizing.cl:484-485).

These checks run the oracle (the reference's rasterizer restated) on the small
workload's traced rays, with the light at (0.3, 0.2), where no ray dies, and
far off axis at (0.9, 0.9), where 13 % of the rays die (the prim shader culls
every quad those touch); and, for a quad with one dead corner, which the prim
shader keeps, with 3 % of the (ghost, wavelength, ray) records of the first
case marked dead (NaN reflectance) before the prim and vertex shaders.  The
quads whose every step the rule removes are rasterized alone and give an image
of exact zeros; the other quads alone give
the full frame's image bit for bit (a pixel adds the same non-zero terms in the
same order, and an exact 0 added to a sum leaves it unchanged).  A rule that
dropped a contributing quad fails the second check; a plausible slip in the
rule (a wrong texel mapping, a missing margin, the wrong border) fails the
first."""
import numpy as np
import pytest

from helpers import trace_inputs
from realflare_amd.api.color import light_spectrum
from realflare_amd.api.tasks.rasterizing import area_orig, screen_transform
from realflare_amd.api.tasks.raytracing import light_direction


def _ghost(oracle, res=128, blades=6, fstop=4.0):
    ap = oracle.aperture_shape((res, res), (0.75, 0.75), blades, 0.0, 0.0, 0.01)
    return oracle.ghost(ap.astype(np.float32), fstop, (res, res)).astype(np.float32)


def _strictly_convex(x, y):
    """Integer corner positions [..., 4] -> all four turns of one sign."""
    c = []
    for i in range(4):
        j, k = (i + 1) % 4, (i + 2) % 4
        c.append((x[..., j] - x[..., i]) * (y[..., k] - y[..., j])
                 - (y[..., j] - y[..., i]) * (x[..., k] - x[..., j]))
    c = np.stack(c, -1)
    return np.all(c > 0, -1) | np.all(c < 0, -1)


def _past_border(t, size):
    lo, hi = t.min(-1), t.max(-1)
    spread = hi - lo
    return (lo >= size + 1 + spread) | (hi <= -3 - spread)


def _zero_steps(vtx, g, aa, ghost_shape, fstop):
    """[P, Q, S] mask of the quad steps the binner leaves out (one
    sub-wavelength step, so a step's corners are the vertices of wavelength
    `step` and their blend with the next wavelength is by 0)."""
    P, R, W = vtx.shape
    S = max(W - 1, 1)
    q = np.arange((g - 1) ** 2)
    c0 = (q // (g - 1)) * g + q % (g - 1)
    corners = [c0, c0 + 1, c0 + 1 + g, c0 + g]
    gs = np.float32(1 - fstop / 32)
    gh, gw = ghost_shape
    ku, kv = np.float32(0.5) / gs * np.float32(gw), np.float32(0.5) / gs * np.float32(gh)
    cu, cv = np.float32(0.5 * gw - 0.5), np.float32(0.5 * gh - 0.5)
    out = np.zeros((P, len(q), S), bool)
    for s in range(S):
        a = vtx[:, :, s]
        b = vtx[:, :, s + 1] if s + 1 < W else np.zeros_like(a)
        with np.errstate(invalid='ignore'):
            refl = a['reflectance'] + (b['reflectance'] - a['reflectance']) * np.float32(0)
        st = lambda f: np.stack([f[:, c] for c in corners], -1)
        x = np.trunc(st(a['pos'][..., 0]) * aa).astype(np.int64)
        y = np.trunc(st(a['pos'][..., 1]) * aa).astype(np.int64)
        tu = st(a['uv'][..., 0]) * ku + cu
        tv = st(a['uv'][..., 1]) * kv + cv
        ok = np.all(np.abs(tu) <= 1e6, -1) & np.all(np.abs(tv) <= 1e6, -1)
        dead = np.any(np.isnan(st(refl)), -1)
        small = ((x.max(-1) - x.min(-1)) < 4096) & ((y.max(-1) - y.min(-1)) < 4096)
        border = (aa == 1) & small & (_past_border(tu, gw) | _past_border(tv, gh)) \
            & _strictly_convex(x, y)
        out[:, :, s] = ok & (dead | border)
    return out


@pytest.mark.parametrize('aa, light, kill', [(1, (0.3, 0.2), 0.0), (1, (0.9, 0.9), 0.0),
                                             (1, (0.3, 0.2), 0.03), (4, (0.3, 0.2), 0.03)])
def test_left_out_quads_contribute_exact_zeros(oracle_lib, aa, light, kill):
    oracle = oracle_lib
    inp = trace_inputs('small')
    fstop = inp['workload'].project.flare.lens.fstop
    g, res = inp['grid_count'], inp['resolution']
    direction = light_direction(light, res, inp['sensor'], inp['workload'].lens_model.focal_length)
    rays = oracle.trace(inp['lens'], inp['paths'], inp['wavelengths'], inp['aperture_index'],
                        inp['coating_min_ior'], g, inp['grid_length'],
                        np.asarray(direction, np.float64))
    if kill:
        dead = np.random.default_rng(7).random(rays.shape) < kill
        rays['reflectance'][dead] = np.nan
    ao = area_orig(g, inp['grid_length'])
    bounds, intens = oracle.prim_shader(rays, g, ao, 0.01 * ao)
    st = screen_transform(res, inp['sensor'])
    vtx = oracle.vertex_shader(rays, intens, g, st, res)
    ghost = _ghost(oracle, fstop=fstop)
    zero = _zero_steps(vtx, g, aa, ghost.shape, fstop)
    drawn = ~np.isnan(bounds[..., 0])
    whole = zero.all(-1) & drawn
    assert whole.sum() > 100, int(whole.sum())  # the rule removes something here

    def render(b):
        return oracle.rasterize(vtx, b, 1, g, aa, 1e3, np.float32(1 - fstop / 32), ghost,
                                light_spectrum(), st, 64, res)

    only_zero = bounds.copy()
    only_zero[~whole] = np.nan
    img_zero = render(only_zero)
    assert not np.any(img_zero), 'quads the binner leaves out lit a pixel'

    rest = bounds.copy()
    rest[whole] = np.nan
    full = render(bounds)
    assert np.abs(full).sum() > 0
    np.testing.assert_array_equal(render(rest).view(np.uint32), full.view(np.uint32))
