"""CPU checks of the two float64 restatements used to check the adjoint:
the C trajectory histogram (oracle/cpu_md.c) and the PyTorch autograd
program (oracle/adjoint_ref.py).

* both compute the same H (independent code);
* the shifted-force pair force equals the oracle's SF wrapper of
  LJ_12_6_sigma_epsilon (md_oracle.pair_eval, apply_shifted_force_cutoff.py:29-37);
* autograd gradients of a loss of H equal central finite differences of the
  C program, for sigma, epsilon and a direction in the initial positions.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_ref, c_oracle
from oracle import md_oracle as orc

KA_SIG = np.array([[1.00, 0.80], [0.80, 0.88]])
KA_EPS = np.array([[1.00, 1.50], [1.50, 0.50]])


def _system(N=216, rho=1.2, T=1.0, seed=3):
    """Jittered simple-cubic start; the box (L = 5.65) exceeds twice the largest cut (2.5)."""
    rng = np.random.default_rng(seed)
    n = int(round(N ** (1 / 3)))
    L = (N / rho) ** (1 / 3)
    g = (np.arange(n) + 0.5) / n * L - L / 2
    r = np.stack(np.meshgrid(g, g, g, indexing='ij'), -1).reshape(-1, 3)[:N]
    r = r + rng.uniform(-0.05, 0.05, r.shape)
    v = rng.normal(0, np.sqrt(T), r.shape)
    v -= v.mean(0)
    types = np.zeros(N, np.int32)
    types[rng.permutation(N)[:N // 5]] = 1
    return r, v, types, np.array([L, L, L])


SETTINGS = dict(dt=0.004, nsteps=12, rdf_every=4, bins=40, rmax=1.6, w=0.05)


def _params(sig, eps, cut):
    return np.stack([sig, eps, cut], -1)


def _torch_H(r, v, types, box, sig, eps, cut):
    t = lambda a: torch.tensor(a, dtype=torch.float64)
    s = SETTINGS
    return adjoint_ref.trajectory_histogram(t(r), t(v), torch.ones(len(r), dtype=torch.float64),
                                            torch.tensor(types, dtype=torch.long), t(sig), t(eps), t(cut),
                                            t(box), s['dt'], s['nsteps'], s['rdf_every'], s['bins'],
                                            s['rmax'], s['w'])


def _c_H(r, v, types, box, sig, eps, cut):
    s = SETTINGS
    H, _ = c_oracle.adj_histogram_f64(r, v, np.ones(len(r)), types, box, _params(sig, eps, cut), s['dt'],
                                      s['nsteps'], s['rdf_every'], s['bins'], s['rmax'], s['w'])
    return H


def test_torch_and_c_histograms_agree():
    r, v, types, box = _system()
    cut = 2.5 * KA_SIG
    Ht = _torch_H(r, v, types, box, KA_SIG, KA_EPS, cut).numpy()
    Hc = _c_H(r, v, types, box, KA_SIG, KA_EPS, cut)
    assert Ht.sum() > 100
    np.testing.assert_allclose(Ht, Hc, rtol=1e-9, atol=1e-9)


def test_shifted_force_matches_the_oracle_wrapper():
    """s(r) - s(rc) of apply_shifted_force_cutoff.py:29-37, via md_oracle.pair_eval (float32)."""
    r = torch.tensor([[0.0, 0.0, 0.0], [1.13, 0.0, 0.0]], dtype=torch.float64)
    for (a, b) in [(0, 0), (0, 1), (1, 1)]:
        types = torch.tensor([a, b])
        F = adjoint_ref.forces(r, types, torch.tensor(KA_SIG), torch.tensor(KA_EPS), torch.tensor(2.5 * KA_SIG),
                               torch.tensor([10.0, 10.0, 10.0], dtype=torch.float64))
        _, s, _ = orc.pair_eval('LJ_12_6_sigma_epsilon', 'SF', np.float32(1.13), np.float32(KA_SIG[a, b]),
                                np.float32(KA_EPS[a, b]), np.float32(2.5 * KA_SIG[a, b]))
        assert abs(float(F[0, 0]) - (-1.13 * float(s))) < 2e-6 * max(1.0, abs(1.13 * float(s)))


def _loss_np(H, target):
    return float(np.sum((H - target) ** 2)) / H.size


def test_autograd_gradient_equals_finite_differences():
    r, v, types, box = _system()
    cut = 2.5 * KA_SIG
    target = _c_H(r, v, types, box, KA_SIG * np.array([[1, 1], [1, 1.02]]), KA_EPS * np.array([[1, 0.95], [0.95, 1]]),
                  cut)
    sig = torch.tensor(KA_SIG, requires_grad=True)
    eps = torch.tensor(KA_EPS, requires_grad=True)
    r0 = torch.tensor(r, requires_grad=True)
    s = SETTINGS
    H = adjoint_ref.trajectory_histogram(r0, torch.tensor(v), torch.ones(len(r), dtype=torch.float64),
                                         torch.tensor(types, dtype=torch.long), sig, eps, torch.tensor(cut),
                                         torch.tensor(box), s['dt'], s['nsteps'], s['rdf_every'], s['bins'],
                                         s['rmax'], s['w'])
    loss = ((H - torch.tensor(target)) ** 2).sum() / H.numel()
    loss.backward()
    # symmetric parameters: d/d theta_ab = grad[a,b] + grad[b,a] (a != b)
    for which, base, grad in (("sig", KA_SIG, sig.grad.numpy()), ("eps", KA_EPS, eps.grad.numpy())):
        for a, b in [(0, 0), (0, 1), (1, 1)]:
            h = 1e-6 * base[a, b]
            def L_at(x):
                p = base.copy()
                p[a, b] = p[b, a] = x
                args = (p, KA_EPS) if which == "sig" else (KA_SIG, p)
                return _loss_np(_c_H(r, v, types, box, args[0], args[1], cut), target)
            fd = (L_at(base[a, b] + h) - L_at(base[a, b] - h)) / (2 * h)
            ad = grad[a, b] + (grad[b, a] if a != b else 0.0)
            assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-8) + 1e-9, (which, a, b, ad, fd)
    rng = np.random.default_rng(9)
    u = rng.normal(size=r.shape)
    h = 1e-6
    Lp = _loss_np(_c_H(r + h * u, v, types, box, KA_SIG, KA_EPS, cut), target)
    Lm = _loss_np(_c_H(r - h * u, v, types, box, KA_SIG, KA_EPS, cut), target)
    fd = (Lp - Lm) / (2 * h)
    ad = float(np.sum(r0.grad.numpy() * u))
    assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-9, (ad, fd)
