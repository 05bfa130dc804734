"""FP64 PyTorch restatement of the ghost trace, differentiable by autograd
(test infrastructure: the reference against which the hand-written HIP
adjoint is checked).

Follows realflare/api/tasks/cl/raytracing.cl:123-359 like the C oracle
(the same event sequence, Fresnel / AR-coating formulas, Sellmeier
dispersion, n1 chosen by the sign of dir.z), vectorized over the rays of one
(ghost, wavelength); rays leave the computation when they miss a surface or
totally reflect, and come back as NaN.  Pinned by tests/test_adjoint_ref.py
against the C oracle's double-precision trace (values) and FP64 central
finite differences of that trace (gradients).
"""
from __future__ import annotations

import math

import torch

PI = math.pi


def _dot(a, b):
    return (a * b).sum(-1, keepdim=True)


def _dispersion(lam, b, c):
    l = lam * 1e-3
    l2 = l * l
    return torch.sqrt(1 + (b * l2 / (l2 - c)).sum())


def _fresnel(theta0, n1, n2):
    theta1 = torch.asin(torch.sin(theta0) * n1 / n2)
    cosi, cost = torch.cos(theta0), torch.cos(theta1)
    rs = (n1 * cosi - n2 * cost) / (n1 * cosi + n2 * cost)
    rp = (n1 * cost - n2 * cosi) / (n1 * cost + n2 * cosi)
    return (rs * rs + rp * rp) / 2


def _fresnel_ar(theta0, lam, d1, n0, n1, n2):
    s0 = torch.sin(theta0)
    theta1 = torch.asin(s0 * n0 / n1)
    theta2 = torch.asin(s0 * n0 / n2)
    rs01 = -torch.sin(theta0 - theta1) / torch.sin(theta0 + theta1)
    rp01 = torch.tan(theta0 - theta1) / torch.tan(theta0 + theta1)
    ts01 = 2 * torch.sin(theta1) * torch.cos(theta0) / torch.sin(theta0 + theta1)
    tp01 = ts01 * torch.cos(theta0 - theta1)
    rs12 = -torch.sin(theta1 - theta2) / torch.sin(theta1 + theta2)
    rp12 = torch.tan(theta1 - theta2) / torch.tan(theta1 + theta2)
    ris = ts01 * ts01 * rs12
    rip = tp01 * tp01 * rp12
    dy = d1 * n1
    dx = torch.tan(theta1) * dy
    delay = torch.sqrt(dx * dx + dy * dy)
    ph = 4 * PI / lam * (delay - dx * s0)
    out_s2 = rs01 * rs01 + ris * ris + 2 * rs01 * ris * torch.cos(ph)
    out_p2 = rp01 * rp01 + rip * rip + 2 * rp01 * rip * torch.cos(ph)
    return (out_s2 + out_p2) / 2


def elements(radii, distances, sellmeier_b, coating, static):
    """Per-element tensors (radius, centre, coating, b) of the packed lens."""
    n = static.n_model
    zero = radii.new_zeros(1)
    radius = torch.cat([radii, zero])
    offsets = torch.cat([zero, torch.cumsum(distances, 0)])
    center = offsets + radius
    coat = torch.cat([coating, coating.new_tensor([static.sensor_coating])])
    b = torch.cat([sellmeier_b, sellmeier_b.new_zeros(1, 3)])
    return radius, center, coat, b


def trace(radii, distances, sellmeier_b, coating, static, paths, wavelengths, coating_min_ior,
          grid_count, grid_length, direction):
    """Returns x, y, reflectance as float64 tensors [P][W][R] (NaN: dead)."""
    radius, center, coat, b = elements(radii, distances, sellmeier_b, coating, static)
    L = static.n_model + 1
    stop = static.aperture_index
    g = grid_count
    R = g * g
    f32 = lambda v: float(torch.tensor(v, dtype=torch.float32))  # noqa: E731
    ids = torch.arange(R)
    ys = torch.trunc(ids.double() / float(g)).long()
    xs = ids - ys * g
    gl = f32(grid_length)
    gx = gl * (xs.double() / (g - 1) - 0.5)
    gy = gl * (0.5 - ys.double() / (g - 1))
    coating_min_ior = f32(coating_min_ior)
    d = torch.tensor([-f32(direction[0]), -f32(direction[1]), -f32(direction[2])],
                     dtype=torch.float64)
    d = d / torch.sqrt((d * d).sum())
    outs = []
    for (b1, b2) in paths:
        per_w = []
        for lam in wavelengths:
            n_eff = []
            for i in range(L):
                if static.disperse and static.ior[i] > 1:
                    n_eff.append(_dispersion(lam, b[i], torch.tensor(static.sellmeier_c[i]
                                                                    if i < static.n_model else
                                                                    [0.0, 0.0, 0.0])))
                else:
                    n_eff.append(radius.new_tensor(float(static.ior[i])))
            # init_ray: vertical probe onto element 0, back off along d
            p0 = torch.stack([gx, gy, torch.ones(R, dtype=torch.float64)], -1)
            down = torch.tensor([0.0, 0.0, -1.0], dtype=torch.float64).expand(R, 3)
            hit0, pos = _intersect(p0, down, radius[0], center[0])
            pos = pos - d
            dirv = d.expand(R, 3).clone()
            refl = torch.ones(R, dtype=torch.float64)
            idx = torch.arange(R)
            if b1 >= 0:
                seq = list(range(0, b1 + 1)) + list(range(b1 - 1, b2 - 1, -1)) + \
                    list(range(b2 + 1, L))
                refl_at = {b1, 2 * b1 - b2}
            else:
                seq, refl_at = list(range(L)), set()
            for k, el in enumerate(seq):
                reflect = k in refl_at
                hit, p, nrm, inc = _intersect(pos, dirv, radius[el], center[el], normals=True)
                keep = hit
                pos, dirv, refl, idx, p, nrm, inc = (t[keep] for t in (pos, dirv, refl, idx, p,
                                                                       nrm, inc))
                pos = p
                if el == stop or el == L - 1:
                    continue
                down_travel = dirv[:, 2] < 0
                nlo = n_eff[el - 1] if el - 1 >= 0 else radius.new_tensor(1.0)
                nhi = n_eff[el + 1] if el + 1 < L else radius.new_tensor(1.0)
                n1 = torch.where(down_travel, nlo, nhi)
                n2 = n_eff[el]
                if reflect:
                    dirv = dirv + nrm * (2.0 * _dot(nrm, -dirv))
                    theta = inc + 1e-9
                    if float(coat[el].detach()) > 0:
                        nc = torch.clamp(torch.sqrt(n1 * n2), min=coating_min_ior)
                        d1 = coat[el] / (4.0 * nc)
                        r = _fresnel_ar(theta, float(lam), d1, n1, nc, n2)
                    else:
                        r = _fresnel(theta, n1, n2)
                    refl = torch.where(r > 0, refl * r, refl)
                else:
                    o = (n1 / n2).unsqueeze(-1)
                    cost = _dot(-dirv, nrm)
                    sint2 = o * o * (1 - cost * cost)
                    ok = (sint2 < 1).squeeze(-1)
                    t = o * dirv + (o * cost - torch.sqrt(torch.abs(1 - sint2))) * nrm
                    pos, t, refl, idx = pos[ok], t[ok], refl[ok], idx[ok]
                    dirv = t
            x = torch.full((R,), float('nan'), dtype=torch.float64)
            y = x.clone()
            rr = x.clone()
            x = x.index_put((idx,), pos[:, 0])
            y = y.index_put((idx,), pos[:, 1])
            rr = rr.index_put((idx,), refl)
            per_w.append(torch.stack([x, y, rr]))
        outs.append(torch.stack(per_w))
    out = torch.stack(outs)  # [P, W, 3, R]
    return out[:, :, 0], out[:, :, 1], out[:, :, 2]


def _intersect(pos, dirv, radius, center, normals=False):
    """raytracing.cl:123-175 for a batch of rays (returns hit mask, p[, n, incident])."""
    R = pos.shape[0]
    if float(radius.detach()) == 0:
        dz = -center - pos[:, 2]
        p = pos + dirv * (dz / dirv[:, 2]).unsqueeze(-1)
        n = torch.zeros_like(pos)
        n[:, 2] = torch.where(dirv[:, 2] < 0, 1.0, -1.0)
        hit = dirv[:, 2] != 0
        inc = torch.zeros(R, dtype=pos.dtype)
        return (hit, p, n, inc) if normals else (hit, p)
    r = torch.abs(radius)
    C = torch.stack([radius.new_zeros(()), radius.new_zeros(()), -center])
    U = C - pos
    U1 = _dot(U, dirv) * dirv
    V = U - U1
    vv = (V * V).sum(-1)
    # |V| with a zero (not NaN) derivative where a ray passes through the centre
    dd = torch.where(vv > 0, torch.sqrt(torch.where(vv > 0, vv, torch.ones_like(vv))),
                     torch.zeros_like(vv))
    hit = (dd <= r) & (dirv[:, 2] != 0)
    sgn = torch.where(radius * dirv[:, 2] > 0, -1.0, 1.0).unsqueeze(-1).to(pos.dtype)
    mm = r * r - dd * dd
    m = torch.where(mm > 0, torch.sqrt(torch.where(mm > 0, mm, torch.ones_like(mm))),
                    torch.zeros_like(mm)).unsqueeze(-1)
    p = pos + U1 - m * dirv * sgn
    w = p - C
    n = w / torch.sqrt((w * w).sum(-1, keepdim=True)) * sgn
    if not normals:
        return hit, p
    c = _dot(-dirv, n).squeeze(-1)
    inner = (c > -1.0) & (c < 1.0)
    inc = torch.where(inner, torch.acos(torch.where(inner, c, torch.zeros_like(c))),
                      torch.acos(torch.clamp(c, -1.0, 1.0)).detach())
    return hit, p, n, inc
