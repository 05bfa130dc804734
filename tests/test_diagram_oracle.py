"""Pins the oracle's lens diagram (diagram.cl restated in float32 numpy)
without a GPU.

* is_line (diagram.cl:1-19) against exact rational geometry: the pixel lies
  strictly between the end caps and 4 cross^2 <= |v|^2; on a 1/8 grid every
  product is exact in float32, so the two may differ only at near ties.
* lenses (diagram.cl:21-68) against closed forms on integer geometry: the
  axis row; a flat element is the column trunc(c) over |y| < h, green for the
  stop and the last element, blue otherwise, and stops the element loop; a
  spherical element is the lattice ring r^2 <= (x-c)^2 + y^2 < (r+1)^2 on the
  side the sign of r picks (a flat element off its column is the ring of
  radius 0: the points within 1 on its right).
* intersections (diagram.cl:71-138): axis-aligned ray and normal segments
  draw the pixel sets geometry fixes; a later normal overwrites, crossing
  rays add; the walk stops at the first record without a hit and after a
  successor at x == 0, and a ray's last record pairs with the next ray's
  first.
"""
from fractions import Fraction

import numpy as np

import oracle


def test_is_line_matches_exact_geometry():
    rng = np.random.default_rng(7)
    px, py = np.meshgrid(np.arange(-25, 26, dtype=np.float32), np.arange(-25, 26, dtype=np.float32))
    near_ties = 0
    for _ in range(60):
        x0, y0, x1, y1 = (rng.integers(-160, 161, size=4) / 8.0).tolist()
        got = oracle.diagram_is_line(px, py, x0, y0, x1, y1)
        vx, vy = Fraction(x1) - Fraction(x0), Fraction(y1) - Fraction(y0)
        len2 = vx * vx + vy * vy
        for j in range(px.shape[0]):
            for i in range(px.shape[1]):
                X, Y = Fraction(int(px[j, i])), Fraction(int(py[j, i]))
                ax, ay = X - Fraction(x0), Y - Fraction(y0)
                bx, by = X - Fraction(x1), Y - Fraction(y1)
                between = vx * ax + vy * ay > 0 and vx * bx + vy * by < 0
                cross = vx * ay - ax * vy
                if len2 == 0:
                    exact = False
                else:
                    exact = between and 4 * cross * cross <= len2
                if bool(got[j, i]) != exact:
                    # only a near tie of the distance test may round either
                    # way; an exact tie (d == 1/2) is computed exactly and counts
                    gap = abs(4 * cross * cross - len2)
                    assert between and 0 < gap <= Fraction(1, 10 ** 6) * len2
                    near_ties += 1
    assert near_ties <= 5
    # exact ties: a line at y = 1/2 lies 1/2 from the rows y = 0 and y = 1
    got = oracle.diagram_is_line(px, py, 0.0, 0.5, 10.0, 0.5)
    exp = (px > 0) & (px < 10) & ((py == 0) | (py == 1))
    np.testing.assert_array_equal(got, exp)


def test_lenses_closed_forms():
    # elements: sphere r=+30 at c=200, flat stop at c=260, sphere r=-40 at
    # c=330, flat at c=380 (not stop), flat sensor at c=450; scale 1
    radius = np.array([30, 0, -40, 0, 0], np.float32)
    height = np.array([25, 18, 35, 12, 60], np.float32)
    center = np.array([200, 260.6, 330, 380, 450.25], np.float32)
    W, H = 480, 161
    img = oracle.diagram_lenses(radius, height, center, 1, 1.0, W, H)
    ys = H // 2 - np.arange(H)
    xs = np.arange(W)
    X, Y = np.meshgrid(xs, ys)
    red = (Y == 0)
    np.testing.assert_array_equal(img[..., 0] == 1, red)
    green = ((X == 260) & (np.abs(Y) < 18)) | ((X == 450) & (np.abs(Y) < 60))
    np.testing.assert_array_equal(img[..., 1] == 1, green)
    d2 = (X - 200) ** 2 + Y ** 2
    ring1 = (X < 200) & (d2 >= 30 ** 2) & (d2 < 31 ** 2) & (np.abs(Y) < 25)
    d2 = (X - 330) ** 2 + Y ** 2
    ring2 = (X > 330) & (d2 >= 40 ** 2) & (d2 < 41 ** 2) & (np.abs(Y) < 35)
    col = (X == 380) & (np.abs(Y) < 12)
    # a flat element away from its column still takes the ring branch with
    # trunc(0) == trunc(distance): the points within 1 on its right
    dots = np.zeros_like(col)
    for c, h in ((260.6, 18), (380, 12), (450.25, 60)):
        dots |= (X > c) & ((X - c) ** 2 + Y ** 2 < 1) & (np.abs(Y) < h)
    assert dots.sum() == 2  # (261, 0) and (451, 0)
    blue = (ring1 | ring2 | col | dots) & ~green
    np.testing.assert_array_equal(img[..., 2] == 1, blue)
    assert not img[..., 3].any()
    # a flat element stops the loop: a ring through its column is not drawn there
    radius2 = np.array([0, 5], np.float32)
    img2 = oracle.diagram_lenses(radius2, np.array([10, 10], np.float32),
                                 np.array([20, 25], np.float32), 0, 1.0, 40, 21)
    assert img2[10, 20, 1] == 1 and img2[10, 20, 2] == 0   # (20, 0): stop, ring skipped
    assert img2[10, 20 - 0, 2] == 0
    assert img2[10, 18, 2] == 0  # d = 7, not on the ring
    assert img2[10 - 3, 21, 2] == 1  # (21, 3): d^2 = 16 + 9 = 25, ring of radius 5


def _records(points, normals, hits):
    n = len(points)
    pos = np.zeros((n, 3), np.float32)
    nor = np.zeros((n, 3), np.float32)
    for k, ((x, y), (nx, ny)) in enumerate(zip(points, normals)):
        # drawn point (x, y) = (-z, -y) * scale; the normal line runs from
        # it to (x, y) - (normal.z, normal.y) * scale, given here as the offset
        pos[k] = (0, -y, -x)
        nor[k] = (0, -ny, -nx)
    return pos, nor, np.asarray(hits, np.int32)


def test_intersections_segments_and_order():
    W, H = 64, 33
    base = np.zeros((H, W, 4), np.float32)
    ys = H // 2 - np.arange(H)
    X, Y = np.meshgrid(np.arange(W), ys)
    # one ray: (10, 0) -> (40, 0) -> (40, 8); normals of length 0 (no line)
    pos, nor, hit = _records([(10, 0), (40, 0), (40, 8), (0, 0)], [(0, 0)] * 4, [1, 1, 1, 0])
    img = oracle.diagram_intersections(base, pos, nor, hit, 4, 1, 1.0)
    seg1 = (Y == 0) & (X > 10) & (X < 40)
    seg2 = (X == 40) & (Y > 0) & (Y < 8)
    exp = np.zeros((H, W, 4), np.float32)
    exp[seg1 | seg2, :3] = 0.25
    np.testing.assert_array_equal(img, exp)
    # crossing rays add
    pos, nor, hit = _records([(5, 4), (30, 4), (20, -6), (20, 10), (12, 4), (13, 4)],
                             [(0, 0)] * 6, [1] * 6)
    # rays of 3 records: (5,4)->(30,4)->(20,-6); then (20,10)->(12,4)->(13,4);
    # the last record of ray 0 pairs with ray 1's first: (20,-6)->(20,10)
    img = oracle.diagram_intersections(base, pos, nor, hit, 3, 2, 1.0)
    assert img[H // 2 - 4, 20, 0] == np.float32(0.5)   # (20, 4): two ray lines
    assert img[H // 2 - 0, 20, 1] == np.float32(0.25)  # (20, 0): the pairing line
    # a later normal overwrites: B's normal runs back over the line A->B
    pos, nor, hit = _records([(5, 0), (30, 0), (31, 0)], [(0, 0), (-20, 0), (0, 0)], [1, 1, 1])
    img = oracle.diagram_intersections(base, pos, nor, hit, 3, 1, 1.0)
    row = img[H // 2]
    for x in range(6, 11):
        np.testing.assert_array_equal(row[x], [0.25, 0.25, 0.25, 0])
    for x in range(11, 30):
        np.testing.assert_array_equal(row[x], [0.25, 0, 0, 0])
    assert row[5].sum() == 0 and row[30].sum() == 0
    # walk stops: no hit on the successor, or successor at x == 0
    pos, nor, hit = _records([(10, 0), (40, 0), (50, 0)], [(0, 0)] * 3, [1, 1, 0])
    img = oracle.diagram_intersections(base, pos, nor, hit, 3, 1, 1.0)
    assert img[H // 2, 45].sum() == 0 and img[H // 2, 20, 0] == 0.25
    pos, nor, hit = _records([(10, 0), (0, 3), (50, 0)], [(0, 2), (0, 0), (0, 0)], [1, 1, 1])
    img = oracle.diagram_intersections(base, pos, nor, hit, 3, 1, 1.0)
    np.testing.assert_array_equal(img[H // 2 - 1, 10], [0.25, 0, 0, 0])  # normal drawn
    assert img[..., 1].sum() == 0  # no ray line at all
