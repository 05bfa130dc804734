"""CLI animation and output writing (cli/app.py:22-102, engine.py:239-291),
host side: per-frame animation values (the reference's CLI animation
fixture), frame-number substitution in output paths, and the EXR writer
read back (header fields and every float bit preserved)."""
import json
import os

import numpy as np

from realflare_amd.api.data import Flare, Point, Project, Render, Size, basic, cast
from realflare_amd.api.output import read_exr, write_array, write_exr
from realflare_amd.api.storage import parse_output_path
from realflare_amd.animation import compile_tracks, frame_project, frames
from realflare_amd.cli import RenderJob, argument_parser

HERE = os.path.dirname(os.path.abspath(__file__))


def test_animation_fixture_sets_light_positions_per_frame():
    """The reference's CLI fixture (three positions): frames 0-2 take them,
    frame 3 keeps the last; each frame is a copy, the base is untouched."""
    with open(os.path.join(HERE, 'golden', 'cli_animation.json')) as f:
        anim = json.load(f)
    p = Project()
    base = (p.flare.light.position.x(), p.flare.light.position.y())
    tracks = compile_tracks(p, anim)
    assert [t.path for t in tracks] == [('flare', 'light', 'position')]
    seen = [frame_project(p, tracks, i).flare.light.position for i in range(4)]
    assert all(isinstance(v, Point) for v in seen)
    assert [(v.x(), v.y()) for v in seen] == [(0.5, 0.5), (1.0, 1.0), (1.5, 1.5), (1.5, 1.5)]
    assert (p.flare.light.position.x(), p.flare.light.position.y()) == base


def test_animation_tracks_types_and_unknown_keys():
    p = Project()
    anim = {'render': {'grid_count': [16, 32], 'resolution': [[64, 32]],
                       'no_such_group': {'x': [1]}},
            'flare': {'light': {'intensity': []}}, 'missing': {'a': [1]}}
    tracks = compile_tracks(p, anim)
    assert sorted(t.path for t in tracks) == [('render', 'grid_count'), ('render', 'resolution')]
    f0 = frame_project(p, tracks, 0)
    f5 = frame_project(p, tracks, 5)
    assert f0.render.grid_count == 16 and f5.render.grid_count == 32
    assert isinstance(f5.render.resolution, Size)
    assert (f5.render.resolution.width(), f5.render.resolution.height()) == (64, 32)
    assert f5.flare.light.intensity == p.flare.light.intensity
    numbers = [(n, q.output.frame) for n, q in frames(p, tracks, 3, 5)]
    assert numbers == [(3, 3), (4, 4), (5, 5)]


def test_output_path_frames(tmp_path):
    assert parse_output_path(str(tmp_path / 'out.$F4.exr'), 7).endswith('out.0007.exr')
    assert parse_output_path(str(tmp_path / 'out.%04d.exr'), 12).endswith('out.0012.exr')


def test_exr_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    chans = {c: rng.standard_normal((5, 7)).astype(np.float32) for c in 'BGR'}
    chans['R'][0, 0] = np.inf
    path = str(tmp_path / 'a.exr')
    write_exr(path, chans)
    back = read_exr(path)
    assert sorted(back) == ['B', 'G', 'R']
    for c in 'BGR':
        np.testing.assert_array_equal(back[c].view(np.uint32), chans[c].view(np.uint32))
    with open(path, 'rb') as f:
        head = f.read(8)
    assert head == (20000630).to_bytes(4, 'little') + (2).to_bytes(4, 'little')


def test_write_array_exr_keeps_rgb(tmp_path):
    a = np.random.default_rng(2).random((6, 4, 4)).astype(np.float32)
    path = str(tmp_path / 'sub' / 'img.0001.exr')
    write_array(a, path, 'ACES - ACEScg')
    back = read_exr(path)
    for k, c in enumerate('RGB'):
        np.testing.assert_array_equal(back[c], a[..., k])


def test_argument_parser_defaults():
    args = argument_parser().parse_args(['--project', 'p.json'])
    assert args.frame_start == 1 and args.frame_end == 1
    assert args.colorspace == 'ACES - ACEScg'


def test_project_json_round_trip_keeps_nested_types():
    # Flare.Ghost (fstop) and Render.Ghost (resolution) share a class name
    p = Project()
    p.flare.ghost.fstop = 5.6
    p.render.ghost.resolution = Size(96, 48)
    q = cast(Project, json.loads(json.dumps(basic(p))))
    assert type(q.flare.ghost) is Flare.Ghost and q.flare.ghost.fstop == 5.6
    assert type(q.render.ghost) is Render.Ghost
    assert q.render.ghost.resolution == Size(96, 48)
    assert basic(q) == basic(p)


def test_exr_layout_follows_the_file_format(tmp_path):
    # OpenEXR file layout (single-part scanline, version 2, no flags): magic,
    # version, attributes "name\0type\0<i32 size><value>" ending in "\0", the
    # offset table (one u64 per scanline block), then per block <i32 y>
    # <i32 size> and the channels' rows in alphabetical channel order.
    import struct
    chans = {c: np.full((3, 5), v, np.float32) for c, v in (('R', 1.0), ('G', 2.0), ('B', 3.0))}
    path = str(tmp_path / 'layout.exr')
    write_exr(path, chans)
    with open(path, 'rb') as f:
        data = f.read()
    assert struct.unpack_from('<ii', data, 0) == (20000630, 2)
    pos, attrs = 8, {}
    while data[pos] != 0:
        name_end = data.index(b'\0', pos)
        type_end = data.index(b'\0', name_end + 1)
        size = struct.unpack_from('<i', data, type_end + 1)[0]
        attrs[data[pos:name_end].decode()] = (data[name_end + 1:type_end].decode(),
                                              data[type_end + 5:type_end + 5 + size])
        pos = type_end + 5 + size
    pos += 1
    required = {'channels': 'chlist', 'compression': 'compression', 'dataWindow': 'box2i',
                'displayWindow': 'box2i', 'lineOrder': 'lineOrder', 'pixelAspectRatio': 'float',
                'screenWindowCenter': 'v2f', 'screenWindowWidth': 'float'}
    assert {k: v[0] for k, v in attrs.items()} == required
    assert attrs['compression'][1] == b'\0' and attrs['lineOrder'][1] == b'\0'
    assert struct.unpack('<iiii', attrs['dataWindow'][1]) == (0, 0, 4, 2)
    ch = attrs['channels'][1]
    names, q = [], 0
    while ch[q] != 0:
        e = ch.index(b'\0', q)
        names.append(ch[q:e].decode())
        ptype, plinear, xs, ys = struct.unpack_from('<iB3xii', ch, e + 1)
        assert (ptype, plinear, xs, ys) == (2, 0, 1, 1)  # FLOAT, 1x1 sampling
        q = e + 1 + 16
    assert names == ['B', 'G', 'R']
    offsets = struct.unpack_from('<3Q', data, pos)
    for y, off in enumerate(offsets):
        assert struct.unpack_from('<ii', data, off) == (y, 3 * 5 * 4)
        row = np.frombuffer(data, '<f4', 15, off + 8)
        np.testing.assert_array_equal(row, [3.0] * 5 + [2.0] * 5 + [1.0] * 5)
    assert offsets[-1] + 8 + 60 == len(data)


def test_render_job_from_files_applies_overrides(tmp_path):
    p = Project()
    (tmp_path / 'p.json').write_text(json.dumps(basic(p)))
    (tmp_path / 'a.json').write_text(json.dumps({'flare': {'light': {'position': [[0.1, 0.2]]}}}))
    job = RenderJob.from_files(str(tmp_path / 'p.json'), str(tmp_path / 'a.json'),
                               str(tmp_path / 'o.$F4.exr'), 'ACES - ACEScg', 'STARBURST', 4, 6)
    assert job.project.output.write and job.project.output.path.endswith('o.$F4.exr')
    assert job.project.output.element.name == 'STARBURST'
    got = [(n, q.flare.light.position.x()) for n, q in job.frames()]
    assert got == [(4, 0.1), (5, 0.1), (6, 0.1)]
    import pytest
    from realflare_amd.api.data import RealflareError
    with pytest.raises(RealflareError, match='project path not valid'):
        RenderJob.from_files(str(tmp_path / 'none.json'))
    (tmp_path / 'bad.json').write_text('{')
    with pytest.raises(RealflareError, match='animation is not valid'):
        RenderJob.from_files(str(tmp_path / 'p.json'), str(tmp_path / 'bad.json'))
