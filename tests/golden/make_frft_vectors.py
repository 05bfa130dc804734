"""Writes frft_vectors.npz: input/output vectors of the reference's fractional
Fourier transform (realflare/utils/frft.py), the one step of the render path
that no closed form pins for general alpha.

Run in a container that holds the realflare sources:
    python tests/golden/make_frft_vectors.py /root/reference
The module file is loaded by path (it needs only numpy); nothing else of the
reference is imported.
"""
import importlib.util
import os
import sys

import numpy as np


def main(src: str) -> None:
    path = os.path.join(src, 'realflare', 'utils', 'frft.py')
    spec = importlib.util.spec_from_file_location('reference_frft', path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rng = np.random.default_rng(7)
    out = {}
    shapes = [(8, 8), (16, 8), (32, 32)]
    alphas = [0.0, 0.0117, 0.35, 0.7, 1.0, 1.3, 1.9, 2.6, 3.2, 3.8, 4.5]
    for si, shape in enumerate(shapes):
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        out[f'x{si}'] = x
        for ai, a in enumerate(alphas):
            out[f'y{si}_{ai}'] = mod.frft(x.copy(), a)
    out['alphas'] = np.array(alphas)
    out['nshapes'] = np.array(len(shapes))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'frft_vectors.npz'),
                        **out)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
