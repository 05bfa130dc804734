"""Pins the oracle (CPU restatement of the reference kernels) without a GPU.

Each check is chosen so a plausible restatement mistake fails it:
  * Fresnel (raytracing.cl:1-16): normal incidence ((n1-n2)/(n1+n2))^2 and
    Brewster's angle (rp = 0).
  * AR coating (raytracing.cl:18-59): its own normal-incidence limit
    R = r01^2 + b^2 + 2 r01 b cos(phi), b = t01^2 r12 (derived by hand from
    the formula; a swapped index or sign changes it).
  * Sellmeier (raytracing.cl:107-120): catalog n_d at 587.56 nm.
  * Trace (raytracing.cl:218-359): unit directions, hit points on the
    sphere surfaces, Snell's law at every refraction, on-axis ray stays on
    axis, paraxial focus, float and double restatements agree.
  * Coverage (rasterizing.cl:55-93): top-left rule tiles shared edges
    exactly once; collinear quads cover nothing.
  * Mean value coordinates (rasterizing.cl:95-148): vertices and centre.
  * Binning (rasterizing.cl:388-469): binned raster == unbinned brute force.
  * Bilinear filter: exact texel centres, midpoints, border / edge clamp.
  * Power spectrum: rectangle aperture at far field -> Dirichlet kernel.
  * FrFT (frft.py): golden vectors from the reference's own frft.py, and
    the closed forms alpha=1 (ortho FFT), alpha=2 (flip).
  * Aperture: n-gon area and rotational symmetry.
"""
import os

import numpy as np
import pytest

from helpers import trace_inputs

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def test_fresnel_closed_forms(oracle_lib):
    for n1, n2 in [(1.0, 1.5), (1.7, 1.0), (1.5, 1.9)]:
        r = oracle_lib.fresnel(1e-9, n1, n2)
        assert r == pytest.approx(((n1 - n2) / (n1 + n2)) ** 2, rel=1e-9)
    n1, n2 = 1.0, 1.6
    tb = np.arctan(n2 / n1)
    t1 = np.arcsin(np.sin(tb) * n1 / n2)
    rs = (n1 * np.cos(tb) - n2 * np.cos(t1)) / (n1 * np.cos(tb) + n2 * np.cos(t1))
    assert oracle_lib.fresnel(tb, n1, n2) == pytest.approx(rs * rs / 2, rel=1e-9)


@pytest.mark.parametrize('n0,n2,lam,lc', [(1.0, 1.5, 550, 550), (1.0, 1.8, 450, 550),
                                          (1.6, 1.0, 650, 500)])
def test_fresnel_ar_normal_incidence_limit(oracle_lib, n0, n2, lam, lc):
    nc = max(np.sqrt(n0 * n2), 1.38)
    d1 = lc / (4 * nc)
    r01 = (nc - n0) / (nc + n0)
    t01 = 2 * n0 / (n0 + nc)
    b = t01 ** 2 * (n2 - nc) / (n2 + nc)
    phi = 4 * np.pi / lam * d1 * nc
    expect = r01 ** 2 + b ** 2 + 2 * r01 * b * np.cos(phi)
    assert oracle_lib.fresnel_ar(1e-9, lam, d1, n0, nc, n2) == pytest.approx(expect, rel=1e-7)


def test_dispersion_catalog(oracle_lib):
    from realflare_amd.api import glass, storage

    for g in glass.glasses_from_path(storage.decode_path('$GLASS/schott'))[:40]:
        n = oracle_lib.dispersion(587.56, np.float32(g.coefficients))
        assert n == pytest.approx(glass.sellmeier(np.float32(g.coefficients).astype(float), 587.56),
                                  rel=1e-12)
        assert abs(n - g.n) < 2e-3


def test_trace_invariants(oracle_lib):
    inp = trace_inputs('small', grid_count=9)
    rays, inters = oracle_lib.trace(inp['lens'], inp['paths'], inp['wavelengths'],
                                    inp['aperture_index'], inp['coating_min_ior'], 9,
                                    inp['grid_length'], inp['direction'], precision='f64',
                                    intersections=True)
    valid = ~np.isnan(rays['reflectance'])
    assert valid.mean() > 0.5
    d = rays['dir'][valid]
    np.testing.assert_allclose(np.linalg.norm(d, axis=-1), 1, atol=1e-12)
    lensel = inp['lens']
    hit = inters['hit'] == 1
    # every recorded hit on a sphere lies on it
    P, W, R, K = inters.shape
    for p in range(P):
        b1, b2 = inp['paths'][p]
        seq = list(range(0, b1 + 1)) + list(range(b1 - 1, b2 - 1, -1)) + \
            list(range(b2 + 1, len(lensel)))
        for k, el in enumerate(seq[:K]):
            m = hit[p, :, :, k]
            if not m.any():
                continue
            pos = inters['pos'][p, :, :, k][m]
            rad, cen = float(lensel[el]['radius']), float(lensel[el]['center'])
            if rad == 0:
                np.testing.assert_allclose(pos[:, 2], -cen, atol=1e-9)
            else:
                dist = np.linalg.norm(pos - np.array([0, 0, -cen]), axis=-1)
                np.testing.assert_allclose(dist, abs(rad), rtol=1e-9)
    assert np.all(rays['reflectance'][valid] > 0) and np.all(rays['reflectance'][valid] < 1)


def test_trace_snell_on_axis_and_paraxial(oracle_lib):
    from realflare_amd.api import lens as api_lens

    inp = trace_inputs('small', grid_count=9)
    # straight-through path, light on axis: centre ray stays on axis,
    # near-axis rays converge to the paraxial focus on the sensor
    direction = np.array([0, 0, 50, 1.0])
    rays = oracle_lib.trace(inp['lens'], [(-1, -1)], [550], inp['aperture_index'], 1.38, 9,
                            inp['grid_length'] * 0.05, direction, precision='f64')
    r = rays[0, 0]
    centre = r[40]
    assert abs(centre['pos'][0]) < 1e-12 and abs(centre['pos'][1]) < 1e-12
    h0 = inp['grid_length'] * 0.05 / 2
    assert np.max(np.abs(r['pos'][:, :2])) < 0.05 * h0
    # Snell at every refraction of a tilted ray (f64 with intersections)
    direction = np.array([3.0, 2.0, 50, 1.0])
    rays, inters = oracle_lib.trace(inp['lens'], [(-1, -1)], [550], inp['aperture_index'], 1.38,
                                    5, inp['grid_length'] * 0.5, direction, precision='f64',
                                    intersections=True)
    el = inp['lens']
    coef = el['coefficients'][:, :6].astype(np.float64)
    nvals = [oracle_lib.dispersion(550, el['coefficients'][i][:6]) if el['ior'][i] > 1 else 1.0
             for i in range(len(el))]
    for ray_id in range(25):
        if np.isnan(rays['reflectance'][0, 0, ray_id]):
            continue
        rec = inters[0, 0, ray_id]
        for k in range(len(el) - 1):
            if rec['hit'][k] != 1 or rec['hit'][k + 1] != 1 or k == inp['aperture_index']:
                continue
            d_in = rec['pos'][k] - (rec['pos'][k - 1] if k > 0 else rec['pos'][k] - [0, 0, 1])
            if k == 0:
                continue
            d_out = rec['pos'][k + 1] - rec['pos'][k]
            nrm = rec['normal'][k]
            d_in = d_in / np.linalg.norm(d_in)
            d_out = d_out / np.linalg.norm(d_out)
            s_in = np.linalg.norm(np.cross(d_in, nrm))
            s_out = np.linalg.norm(np.cross(d_out, nrm))
            n1 = nvals[k - 1] if k > 0 else 1.0
            assert n1 * s_in == pytest.approx(nvals[k] * s_out, rel=1e-9, abs=1e-12)


def test_trace_float_vs_double(oracle_lib):
    inp = trace_inputs('small', grid_count=33)
    args = (inp['lens'], inp['paths'], inp['wavelengths'], inp['aperture_index'],
            inp['coating_min_ior'], 33, inp['grid_length'], inp['direction'])
    a = oracle_lib.trace(*args, precision='f32')
    b = oracle_lib.trace(*args, precision='f64')
    both = ~np.isnan(a['reflectance']) & ~np.isnan(b['reflectance'])
    assert both.mean() > 0.99
    e = np.abs(a['pos'][both] - b['pos'][both]) / np.maximum(np.abs(b['pos'][both]), 1e-2)
    assert np.median(e) < 1e-5


def test_top_left_rule_tiles_shared_edges(oracle_lib):
    # a 3x3 grid of unit-spaced (in sample units: 7) squares, slightly skewed
    g = 4
    xs = np.array([[x * 7 + y for x in range(g)] for y in range(g)])
    ys = np.array([[y * 7 - x for x in range(g)] for y in range(g)])
    counts = np.zeros((40, 40), int)
    for qy in range(g - 1):
        for qx in range(g - 1):
            vx = [xs[qy, qx], xs[qy, qx + 1], xs[qy + 1, qx + 1], xs[qy + 1, qx]]
            vy = [ys[qy, qx], ys[qy, qx + 1], ys[qy + 1, qx + 1], ys[qy + 1, qx]]
            for py in range(-5, 30):
                for px in range(-5, 30):
                    counts[py + 5, px + 5] += oracle_lib.intersect_quad(px, py, vx, vy)
    assert counts.max() == 1
    # interior of the union is covered: total = polygon area (Pick: area of 3x3 cells of 7x7
    # with the skew preserving area: 9 * (7*7 + 1*1))
    assert counts.sum() == 9 * 50
    assert not oracle_lib.intersect_quad(1, 1, [0, 2, 4, 6], [0, 2, 4, 6])


def test_mean_value_coordinates(oracle_lib):
    vx, vy = [0, 10, 10, 0], [0, 0, 10, 10]
    np.testing.assert_allclose(oracle_lib.barycentric_quad(5, 5, vx, vy), [0.25] * 4)
    np.testing.assert_allclose(oracle_lib.barycentric_quad(10, 0, vx, vy), [0, 1, 0, 0])
    w = oracle_lib.barycentric_quad(3, 7, vx, vy)
    assert w.sum() == pytest.approx(1, abs=1e-6)
    # linear precision: sum w_i v_i reproduces the point
    assert np.dot(w, vx) == pytest.approx(3, abs=1e-5)
    assert np.dot(w, vy) == pytest.approx(7, abs=1e-5)


def test_binning_changes_nothing(oracle_lib):
    from raster_helpers import make_ghost, oracle_image, params_for

    inp = trace_inputs('small', grid_count=12)
    ghost = make_ghost(oracle_lib, res=32)
    rays = oracle_lib.trace(inp['lens'], inp['paths'], inp['wavelengths'], inp['aperture_index'],
                            inp['coating_min_ior'], 12, inp['grid_length'], inp['direction'])
    for bin_size in (16, 64, 20):
        prm = params_for(inp, (96, 80), ghost.shape, bin_size=bin_size, aa=2, wsub=2)
        *_, binned = oracle_image(oracle_lib, rays, inp, prm, ghost)
        *_, brute = oracle_image(oracle_lib, rays, inp, prm, ghost, binned=False)
        assert np.abs(binned).sum() > 0
        np.testing.assert_array_equal(binned, brute)


def test_bilinear_filter(oracle_lib):
    tex = np.arange(12, dtype=np.float32).reshape(3, 4)
    # texel centres
    assert oracle_lib.sample_bilinear(tex, 1.5 / 4, 2.5 / 3) == tex[2, 1]
    # midpoint between 4 texels
    assert oracle_lib.sample_bilinear(tex, 2 / 4, 1 / 3) == pytest.approx(tex[0:2, 1:3].mean())
    # border colour 0 vs clamp to edge at the image corner
    assert oracle_lib.sample_bilinear(tex, 0, 0) == pytest.approx(tex[0, 0] / 4)
    assert oracle_lib.sample_bilinear(tex, 0, 0, edge=True) == tex[0, 0]


def test_power_spectrum_far_field_rectangle(oracle_lib):
    n, k = 64, 9
    ap = np.zeros((n, n))
    ap[n // 2 - k // 2:n // 2 + k // 2 + 1, n // 2 - k // 2:n // 2 + k // 2 + 1] = 1
    ps = oracle_lib.power_spectrum(ap, 1e9)  # distance -> far field, chirp ~ 1
    f = np.fft.fftshift(np.fft.fftfreq(n))
    with np.errstate(invalid='ignore', divide='ignore'):
        dk = np.where(f == 0, k, np.sin(np.pi * k * f) / np.sin(np.pi * f))
    expect = (np.outer(dk, dk)) ** 2
    np.testing.assert_allclose(ps, expect, rtol=1e-4, atol=1e-3)


def test_frft_reference_golden_vectors(oracle_lib):
    z = np.load(os.path.join(GOLDEN, 'frft_vectors.npz'))
    for si in range(int(z['nshapes'])):
        x = z[f'x{si}']
        for ai, a in enumerate(z['alphas']):
            np.testing.assert_allclose(oracle_lib.frft(x, a), z[f'y{si}_{ai}'], rtol=1e-10,
                                       atol=1e-10)


def test_frft_closed_forms(oracle_lib):
    # frft.py's chirp decomposition at alpha = 1 is the orthonormal DFT with an
    # extra 1/sqrt(N) (its `scale` and both ortho FFTs); alpha = 2 first
    # applies an ortho FFT (frft.py:32-34) and then the alpha = 1 transform.
    rng = np.random.default_rng(3)
    x = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    N = x.size
    f1 = np.fft.fftn(x, norm='ortho') / np.sqrt(N)
    np.testing.assert_allclose(oracle_lib.frft(x, 1.0), f1, atol=1e-12)
    f2 = np.fft.fftn(np.fft.fftn(x, norm='ortho'), norm='ortho') / np.sqrt(N)
    np.testing.assert_allclose(oracle_lib.frft(x, 2.0), f2, atol=1e-12)
    np.testing.assert_allclose(oracle_lib.frft(x, 5.0), f1, atol=1e-12)  # period 4


def test_aperture_polygon_area_and_symmetry(oracle_lib):
    n = 512
    for blades in (5, 6, 9):
        img = oracle_lib.aperture_shape((n, n), (0.75, 0.75), blades, 0.0, 0.0, 0.0,
                                        precision='f64')
        # regular n-gon with apothem 0.75 in ndc; pixel area (2/(n-1))^2
        area = blades * np.tan(np.pi / blades) * 0.75 ** 2
        measured = img.sum() * (2 / (n - 1)) ** 2
        assert measured == pytest.approx(area, rel=0.01)
        assert img[n // 2, n // 2] == 1 and img[0, 0] == 0
    img = oracle_lib.aperture_shape((n, n), (0.75, 0.75), 6, 0.0, 0.0, 0.1, precision='f64')
    np.testing.assert_allclose(img, img[::-1, ::-1], atol=1e-9)  # point symmetric (even blades)
