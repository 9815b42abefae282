"""Differentiable NVE trajectories: gradients of a structural observable with
respect to the LJ parameters and the initial positions.

The forward pass is a float32 NVE trajectory (leap-frog as integrators/NVE.py)
of an LJ_12_6_sigma_epsilon system with the shifted-force cutoff at a fixed
cut; the backward pass is the hand-written adjoint in libgamdpy_amd.so
(force Jacobian-vector products, parameter gradients, recomputation from
checkpoints every ``ckpt_every`` steps).  The observable is the
trajectory-summed, Gaussian-smoothed pair histogram per type class; ``rdf``
turns it into partial g(r).

    traj = DifferentiableNVE(conf, dt=0.004, nsteps=200, cut=cut_matrix)
    g = traj.rdf(r0, sigma, epsilon)           # torch tensors, autograd-enabled
    loss = ((g - g_target) ** 2).mean()
    loss.backward()                            # sigma.grad, epsilon.grad, r0.grad

PyTorch is used only for this autograd wrapper; the arithmetic runs in the
HIP kernels.
"""
import ctypes
import math

import numpy as np

from . import _native


GmdAdjDesc = _native.GmdAdjDesc
_c_double_p = _native.c_double_p


def type_classes(ntypes):
    """Unordered type pairs a <= b in the library's class order."""
    return [(a, b) for a in range(ntypes) for b in range(a, ntypes)]


class AdjointEngine:
    """Owns one gmd_adj: a trajectory of fixed length, types and masses."""

    def __init__(self, N, ntypes, box, dt, nsteps, ptype, mass, ckpt_every=20, rdf_every=10, bins=200,
                 rmax=4.0, width=0.02):
        self.lib = _native.load()
        d = GmdAdjDesc()
        d.N, d.ntypes = int(N), int(ntypes)
        d.box = (ctypes.c_float * 3)(*[float(x) for x in box])
        d.dt, d.nsteps, d.ckpt_every = float(dt), int(nsteps), int(ckpt_every)
        d.rdf_every, d.rdf_bins, d.rdf_rmax, d.rdf_width = int(rdf_every), int(bins), float(rmax), float(width)
        self.desc = d
        self.N, self.ntypes, self.bins = d.N, d.ntypes, d.rdf_bins
        self.ncls = ntypes * (ntypes + 1) // 2
        self.ptype = np.ascontiguousarray(ptype, dtype=np.int32)
        self.mass = np.ascontiguousarray(mass, dtype=np.float32)
        h = ctypes.c_void_p()
        _native.check(self.lib.gmd_adj_create(ctypes.byref(d), _native.iptr(self.ptype), _native.fptr(self.mass),
                                              ctypes.byref(h)), "gmd_adj_create")
        self.h = h
        self.nframes = len(range(0, int(nsteps) + 1, int(rdf_every)))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.gmd_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, params, r0, v0):
        """params (nt, nt, 3) float32 (sigma, epsilon, cut) -> H (ncls, bins) float64."""
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        r = np.ascontiguousarray(r0, dtype=np.float32)
        v = np.ascontiguousarray(v0, dtype=np.float32)
        H = np.zeros((self.ncls, self.bins), np.float64)
        _native.check(self.lib.gmd_adj_forward(self.h, _native.fptr(p), _native.fptr(r), _native.fptr(v),
                                               H.ctypes.data_as(_c_double_p), None, None), "gmd_adj_forward")
        return H

    def backward(self, dLdH):
        """-> (grad (ncls, 2) float64 [d/dsigma, d/depsilon], grad_r0 (N,3), grad_v0 (N,3))."""
        g = np.ascontiguousarray(dLdH, dtype=np.float64)
        gp = np.zeros((self.ncls, 2), np.float64)
        gr = np.zeros((self.N, 3), np.float32)
        gv = np.zeros((self.N, 3), np.float32)
        _native.check(self.lib.gmd_adj_backward(self.h, g.ctypes.data_as(_c_double_p), gp.ctypes.data_as(_c_double_p),
                                                _native.fptr(gr), _native.fptr(gv)), "gmd_adj_backward")
        return gp, gr, gv

    def last_ms(self):
        f, b = ctypes.c_float(), ctypes.c_float()
        _native.check(self.lib.gmd_adj_last_ms(self.h, ctypes.byref(f), ctypes.byref(b)), "gmd_adj_last_ms")
        return float(f.value), float(b.value)


def _function():
    import torch

    class TrajectoryHistogram(torch.autograd.Function):
        """H(r0, sigma, epsilon): trajectory-summed smoothed pair histogram."""

        @staticmethod
        def forward(ctx, r0, sigma, epsilon, engine, cut, v0):
            nt = engine.ntypes
            params = np.zeros((nt, nt, 3), np.float32)
            params[:, :, 0] = sigma.detach().cpu().numpy()
            params[:, :, 1] = epsilon.detach().cpu().numpy()
            params[:, :, 2] = cut
            H = engine.forward(params, r0.detach().cpu().numpy(), v0)
            ctx.engine = engine
            ctx.meta = (r0.dtype, r0.device, sigma.dtype, sigma.device)
            return torch.as_tensor(H, dtype=torch.float64, device=r0.device)

        @staticmethod
        def backward(ctx, gH):
            engine = ctx.engine
            rdt, rdev, sdt, sdev = ctx.meta
            gp, gr, _ = engine.backward(gH.detach().cpu().numpy().astype(np.float64))
            nt = engine.ntypes
            gs = np.zeros((nt, nt))
            ge = np.zeros((nt, nt))
            for c, (a, b) in enumerate(type_classes(nt)):
                # a symmetric matrix entry: split the class gradient over (a,b) and (b,a)
                f = 1.0 if a == b else 0.5
                gs[a, b] += f * gp[c, 0]
                ge[a, b] += f * gp[c, 1]
                if a != b:
                    gs[b, a] += f * gp[c, 0]
                    ge[b, a] += f * gp[c, 1]
            return (torch.as_tensor(gr, dtype=rdt, device=rdev), torch.as_tensor(gs, dtype=sdt, device=sdev),
                    torch.as_tensor(ge, dtype=sdt, device=sdev), None, None, None)

    return TrajectoryHistogram


_FN = None


def trajectory_histogram(r0, sigma, epsilon, engine, cut, v0):
    """Autograd entry point: H = sum over frames of the smoothed pair histogram."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(r0, sigma, epsilon, engine, cut, v0)


class DifferentiableNVE:
    """An NVE trajectory of ``configuration`` (its types, masses, box and
    velocities) whose partial g(r) is differentiable in (r0, sigma, epsilon).

    sigma and epsilon are symmetric (ntypes, ntypes) torch tensors; the cut
    matrix is fixed.  ``rdf`` returns g[c, k] for the type classes
    ``type_classes(ntypes)`` at bin centres ``self.r``.
    """

    def __init__(self, configuration, dt, nsteps, cut, ckpt_every=20, rdf_every=10, bins=200, rmax=4.0,
                 width=0.02):
        conf = configuration
        self.N = conf.N
        self.ntypes = int(np.max(conf.ptype)) + 1
        self.box = np.asarray(conf.simbox.lengths, np.float64)
        self.cut = np.asarray(cut, np.float32) * np.ones((self.ntypes, self.ntypes), np.float32)
        self.v0 = np.ascontiguousarray(conf['v'], dtype=np.float32)
        self.engine = AdjointEngine(conf.N, self.ntypes, self.box, dt, nsteps, conf.ptype, conf['m'],
                                    ckpt_every=ckpt_every, rdf_every=rdf_every, bins=bins, rmax=rmax, width=width)
        self.bins, self.rmax, self.width = bins, rmax, width
        self.dr = rmax / bins
        self.r = (np.arange(bins) + 0.5) * self.dr
        counts = np.bincount(conf.ptype, minlength=self.ntypes).astype(np.float64)
        V = float(np.prod(self.box))
        # expected smoothed count per frame of an ideal gas for class (a,b):
        #   pairs N_a N_b (a != b) or N_a (N_a - 1)/2 (a == b), times the shell
        #   density 4 pi r^2 / V, times the kernel integral sqrt(2 pi) w
        norm = []
        for a, b in type_classes(self.ntypes):
            npairs = counts[a] * counts[b] if a != b else counts[a] * (counts[a] - 1) / 2
            norm.append(npairs * 4 * math.pi * self.r ** 2 / V * math.sqrt(2 * math.pi) * width)
        self.norm = np.array(norm) * self.engine.nframes

    def histogram(self, r0, sigma, epsilon):
        return trajectory_histogram(r0, sigma, epsilon, self.engine, self.cut, self.v0)

    def rdf(self, r0, sigma, epsilon):
        import torch
        H = self.histogram(r0, sigma, epsilon)
        return H / torch.as_tensor(self.norm, dtype=H.dtype, device=H.device)

    def last_ms(self):
        return self.engine.last_ms()
