"""Extract the data tables the render path needs from a realflare checkout
into ``realflare_amd/resources``.

A one-off step that made those data files, kept beside the other scripts
that write data from the reference: it reads data files (never code) from a
realflare source tree and writes compact files into the package, so the
package carries no dependency on that tree.  tests/test_oracle_pins.py
checks the CIE table against the reference's own rows.

* ``glass/schott/catalog.json``: name, n_d, V_d and the six "formula 2"
  Sellmeier coefficients of every SCHOTT glass in
  ``realflare/resources/glass/schott/*.yml`` (refractiveindex.info data,
  CC0).  Parsing follows ``realflare/api/glass.py:32-70``: the first
  ``formula 2`` record, coefficients split on single spaces with the
  leading constant dropped, ``SPECS.nd``/``Nd`` and ``vd``/``Vd``.
* ``ciexyz_390_790.csv``: the CIE 2006 10-degree XYZ colour matching
  functions at 1 nm (cvrl.org), read as numbers from the table text in
  ``realflare/utils/ciexyz.py``.
* ``model/<maker>/<name>.json``: the lens prescriptions shipped in
  ``realflare/resources/model``.

Usage: ``python tests/golden/extract_resources.py /root/reference``
"""
from __future__ import annotations

import json
import os
import re
import shutil
import sys

import yaml


HERE = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                    'realflare_amd', 'resources')


def extract_glasses(src: str) -> list[dict]:
    glass_dir = os.path.join(src, 'realflare', 'resources', 'glass', 'schott')
    out = []
    for file_name in sorted(os.listdir(glass_dir)):
        if not file_name.endswith(('.yml', '.yaml')):
            continue
        with open(os.path.join(glass_dir, file_name)) as f:
            data = yaml.safe_load(f)
        coefficients = []
        for item in data.get('DATA', []):
            if item.get('type') == 'formula 2':
                coefficients = [float(c) for c in item['coefficients'].split(' ')[1:]]
                break
        specs = data.get('SPECS', {})
        nd = specs.get('nd') or specs.get('Nd')
        vd = specs.get('vd') or specs.get('Vd')
        if not coefficients or not nd or not vd:
            continue
        out.append(
            {
                'name': os.path.splitext(file_name)[0],
                'n': nd,
                'v': vd,
                'coefficients': coefficients,
            }
        )
    return out


def extract_ciexyz(src: str) -> list[list[float]]:
    path = os.path.join(src, 'realflare', 'utils', 'ciexyz.py')
    rows = []
    number = r'([-+0-9.eE]+)'
    pattern = re.compile(
        r'^\s*\[\s*' + r'\s*,\s*'.join([number] * 4) + r'\s*\]\s*,?\s*$'
    )
    with open(path) as f:
        for line in f:
            m = pattern.match(line)
            if m:
                rows.append([float(g) for g in m.groups()])
    return rows


def main(src: str) -> None:
    glasses = extract_glasses(src)
    with open(os.path.join(HERE, 'glass', 'schott', 'catalog.json'), 'w') as f:
        json.dump({'manufacturer': 'schott', 'glasses': glasses}, f, indent=0)

    rows = extract_ciexyz(src)
    with open(os.path.join(HERE, 'ciexyz_390_790.csv'), 'w') as f:
        f.write('# CIE 2006 10-deg XYZ CMFs (cvrl.org), 1 nm: wavelength,x,y,z\n')
        for w, x, y, z in rows:
            f.write(f'{int(w)},{x:.6E},{y:.6E},{z:.6E}\n')

    model_src = os.path.join(src, 'realflare', 'resources', 'model')
    for maker in sorted(os.listdir(model_src)):
        maker_dir = os.path.join(model_src, maker)
        if not os.path.isdir(maker_dir):
            continue
        os.makedirs(os.path.join(HERE, 'model', maker), exist_ok=True)
        for name in sorted(os.listdir(maker_dir)):
            if name.endswith('.json'):
                shutil.copyfile(
                    os.path.join(maker_dir, name), os.path.join(HERE, 'model', maker, name)
                )
    print(f'{len(glasses)} glasses, {len(rows)} CIE rows')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
