"""Plain PyTorch float64 restatement of the differentiable NVE observable.

TEST INFRASTRUCTURE ONLY: the reference gradient for the hand-written adjoint
(autograd through every step of this program), checked itself against
finite differences of oracle/cpu_md.c (orc_adj_histogram_f64).  Never
imported by the product.

Model (same definitions as the C checker):
  LJ_12_6_sigma_epsilon (LJ_12_6_sigma_epsilon.py:25-36) with gamdpy's
  shifted-force cutoff (apply_shifted_force_cutoff.py:29-37, s -= s(rc)) at a fixed cut:
    s(r) = 24 eps (2 (sig/r)^12 - (sig/r)^6) / r^2 - s(rc),   F_i = sum_j s(r_ij) (r_i - r_j)
  leap-frog NVE (NVE.py:59-105): v += F/m dt; r += v dt; PBC wrap (orthorhombic.py)
  H[c][k] = sum over frames n % rdf_every == 0 (n = 0..nsteps) and unordered
  pairs of class c of max(0, exp(-(d - r_k)^2 / 2w^2) - exp(-12.5)) for
  |d - r_k| < 5w, r_k = (k + 1/2) rmax / bins.
"""
import math

import torch


def type_classes(nt):
    return [(a, b) for a in range(nt) for b in range(a, nt)]


def _mic(d, L):
    return torch.where(2 * d > L, d - L, torch.where(2 * d < -L, d + L, d))


def _pairs(r, L):
    d = r[:, None, :] - r[None, :, :]
    return _mic(d, L)


def forces(r, types, sig, eps, cut, L):
    """F (N,3) for (nt,nt) tensors sig, eps, cut."""
    N = r.shape[0]
    d = _pairs(r, L)
    d2 = (d * d).sum(-1)
    eye = torch.eye(N, dtype=torch.bool, device=r.device)
    S = sig[types][:, types]
    E = eps[types][:, types]
    C = cut[types][:, types]
    inside = (d2 < C * C) & ~eye
    d2s = torch.where(inside, d2, torch.ones_like(d2))
    x = S * S / d2s
    x3 = x * x * x
    x6 = x3 * x3
    xc6 = (S / C) ** 6
    sc = 24 * E * (2 * xc6 * xc6 - xc6) / (C * C)
    s = 24 * E * (2 * x6 - x3) / d2s - sc
    s = torch.where(inside, s, torch.zeros_like(s))
    return (s[:, :, None] * d).sum(1)


def frame_histogram(r, types, nt, L, bins, rmax, w):
    N = r.shape[0]
    d = _pairs(r, L)
    d2 = (d * d).sum(-1)
    iu = torch.triu_indices(N, N, offset=1, device=r.device)
    dd2 = d2[iu[0], iu[1]]
    ti, tj = types[iu[0]], types[iu[1]]
    rc = rmax + 5 * w
    keep = dd2 < rc * rc
    dist = torch.sqrt(dd2[keep])
    a = torch.minimum(ti[keep], tj[keep])
    b = torch.maximum(ti[keep], tj[keep])
    cls = a * nt - a * (a - 1) // 2 + (b - a)
    dr = rmax / bins
    rk = (torch.arange(bins, dtype=r.dtype, device=r.device) + 0.5) * dr
    x = dist[:, None] - rk[None, :]
    kv = torch.exp(-0.5 * x * x / (w * w)) - math.exp(-12.5)
    kv = torch.where((x.abs() < 5 * w) & (kv > 0), kv, torch.zeros_like(kv))
    ncls = nt * (nt + 1) // 2
    H = torch.zeros(ncls, bins, dtype=r.dtype, device=r.device)
    return H.index_add(0, cls, kv)


def trajectory_histogram(r0, v0, mass, types, sig, eps, cut, L, dt, nsteps, rdf_every, bins, rmax, w):
    """Differentiable in r0, sig, eps (float64 tensors)."""
    nt = sig.shape[0]
    r, v = r0, v0
    H = None
    m = mass[:, None]
    for n in range(nsteps + 1):
        if n % rdf_every == 0:
            h = frame_histogram(r, types, nt, L, bins, rmax, w)
            H = h if H is None else H + h
        if n == nsteps:
            break
        F = forces(r, types, sig, eps, cut, L)
        v = v + F / m * dt
        x = r + v * dt
        r = torch.where(2 * x > L, x - L, torch.where(2 * x < -L, x + L, x))
    return H
