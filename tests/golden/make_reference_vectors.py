"""Writes reference_vectors.npz: vectors of two reference functions that no
closed form pins completely.

* ``fresnel_ar`` away from normal incidence: the reference's own numpy
  statement of the coated-Fresnel formula (realflare tests/lens.py:14-45,
  the formula of raytracing.cl:18-59), evaluated in float64 over a grid of
  angles, wavelengths, coating thicknesses and indices.
* The CIE colour-matching table (realflare/utils/ciexyz.py:4-406, the CIE
  2006 10-degree functions), all rows.

Run in a container that holds the realflare sources:
    python tests/golden/make_reference_vectors.py /root/reference
Each module file is loaded by path; both need only numpy (tests/lens.py also
imports logging).  Nothing else of the reference is imported.
"""
import importlib.util
import os
import sys

import numpy as np


def _load(path: str, name: str):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main(src: str) -> None:
    lens_tests = _load(os.path.join(src, 'tests', 'lens.py'), 'reference_lens_tests')
    ciexyz = _load(os.path.join(src, 'realflare', 'utils', 'ciexyz.py'), 'reference_ciexyz')

    rng = np.random.default_rng(11)
    n = 400
    theta0 = rng.uniform(1e-6, 1.2, n)          # up to ~69 degrees
    wavelength = rng.uniform(390.0, 730.0, n)   # nm
    lambda_c = rng.uniform(400.0, 700.0, n)     # coating design wavelength, nm
    n0 = np.where(rng.uniform(size=n) < 0.5, 1.0, rng.uniform(1.45, 1.95, n))
    n2 = np.where(n0 > 1.0, 1.0, rng.uniform(1.45, 1.95, n))
    nc = np.maximum(np.sqrt(n0 * n2), 1.38)
    d1 = lambda_c / (4.0 * nc)
    # keep only angles below total reflection in every layer
    ok = (np.sin(theta0) * n0 / nc < 0.999) & (np.sin(theta0) * n0 / n2 < 0.999)
    theta0, wavelength, d1, n0, nc, n2 = (a[ok] for a in (theta0, wavelength, d1, n0, nc, n2))
    r = np.array([lens_tests.fresnel_ar(t, w, d, a, b, c)
                  for t, w, d, a, b, c in zip(theta0, wavelength, d1, n0, nc, n2)])

    table = np.array(ciexyz.CIEXYZ, np.float64)
    out = dict(fresnel_ar_in=np.stack([theta0, wavelength, d1, n0, nc, n2], axis=1),
               fresnel_ar_out=r, ciexyz=table)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                     'reference_vectors.npz'), **out)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
