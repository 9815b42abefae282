"""Pair-potential functions and cutoff wrappers.

Each function returns ``(u, s, umm)`` = (u(r), -u'(r)/r, u''(r)) like the
reference (interactions/potential_functions/).  On the device the kernels do
not call Python: every supported function carries a descriptor
(``_gmd_base``, ``_gmd_cutoff``) that selects the matching HIP template
instantiation.  A callable without a descriptor is rejected by PairPotential.
"""
import numpy as np

F32 = np.float32


def _f32(x):
    return np.asarray(x, dtype=F32)


def _out(*vals):
    vals = [np.asarray(v, F32) for v in vals]
    if all(v.ndim == 0 for v in vals):
        return tuple(F32(v) for v in vals)
    return tuple(vals)


def LJ_12_6(dist, params):
    """u(r) = A12 r^-12 + A6 r^-6 (LJ_12_6.py:6-42).  params: A12, A6[, cut]."""
    a12, a6 = _f32(params[0]), _f32(params[1])
    inv = F32(1.0) / _f32(dist)
    p2 = inv * inv
    p6 = p2 * p2 * p2
    p12 = p6 * p6
    p8 = p6 * p2
    p14 = p12 * p2
    u = a12 * p12 + a6 * p6
    s = (F32(12.0) * a12) * p14 + (F32(6.0) * a6) * p8
    umm = (F32(156.0) * a12) * p14 + (F32(42.0) * a6) * p8
    return _out(u, s, umm)


LJ_12_6._gmd_base = "LJ_12_6"
LJ_12_6._gmd_cutoff = None


def LJ_12_6_sigma_epsilon(dist, params):
    """u(r) = 4 eps ((r/sigma)^-12 - (r/sigma)^-6) (LJ_12_6_sigma_epsilon.py:6-36).
    params: sigma, epsilon[, cut]."""
    sigma, eps = _f32(params[0]), _f32(params[1])
    od = F32(1.0) / _f32(dist)
    x = sigma * od
    x2 = x * x
    x6 = x2 * x2 * x2
    x12 = x6 * x6
    od2 = od * od
    u = (F32(4.0) * eps) * (x12 - x6)
    s = ((F32(24.0) * eps) * (F32(2.0) * x12 - x6)) * od2
    umm = ((F32(24.0) * eps) * (F32(26.0) * x12 - F32(7.0) * x6)) * od2
    return _out(u, s, umm)


LJ_12_6_sigma_epsilon._gmd_base = "LJ_12_6_sigma_epsilon"
LJ_12_6_sigma_epsilon._gmd_cutoff = None


def _base_of(pair_potential):
    base = getattr(pair_potential, "_gmd_base", None)
    if base is None or getattr(pair_potential, "_gmd_cutoff", None) is not None:
        raise ValueError(f"{getattr(pair_potential, '__name__', pair_potential)!r} is not a base pair function "
                         "supported by the MI355X kernels (LJ_12_6, LJ_12_6_sigma_epsilon)")
    return base


def apply_shifted_potential_cutoff(pair_potential):
    """u(r) - u(rc) (apply_shifted_potential_cutoff.py:6-49)."""
    base = _base_of(pair_potential)

    def potential(dist, params):
        cut = params[-1]
        u, s, umm = pair_potential(dist, params)
        u_cut, s_cut, umm_cut = pair_potential(cut, params)
        return _out(u - u_cut, s, umm)

    potential._gmd_base = base
    potential._gmd_cutoff = "SP"
    potential.__name__ = f"shifted_potential({base})"
    return potential


def apply_shifted_force_cutoff(pair_potential):
    """u(r) - u(rc) + s(rc) rc (r - rc), s(r) - s(rc) (apply_shifted_force_cutoff.py:6-38)."""
    base = _base_of(pair_potential)

    def potential(dist, params):
        cut = _f32(params[-1])
        u, s, umm = pair_potential(dist, params)
        u_cut, s_cut, umm_cut = pair_potential(cut, params)
        u = u - (u_cut - (s_cut * cut) * (_f32(dist) - cut))
        return _out(u, s - s_cut, umm)

    potential._gmd_base = base
    potential._gmd_cutoff = "SF"
    potential.__name__ = f"shifted_force({base})"
    return potential


def pair_function_descriptor(func):
    """(base, cutoff) of a supported pair function, else ValueError."""
    base = getattr(func, "_gmd_base", None)
    if base is None:
        raise ValueError(f"pair function {getattr(func, '__name__', func)!r} has no HIP implementation; supported: "
                         "LJ_12_6, LJ_12_6_sigma_epsilon, optionally wrapped by apply_shifted_potential_cutoff "
                         "or apply_shifted_force_cutoff")
    return base, getattr(func, "_gmd_cutoff", None)
