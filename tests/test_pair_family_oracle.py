"""CPU checks of the pair-function family (make_IPL_n, make_LJ_m_n,
harmonic_repulsion, hertzian): the oracle restatements in oracle/md_oracle.py
and the package's host functions.

Pinned by the reference's own known answers (test_potential_functions.py:16-70),
by closed forms (s = -u'/r from central differences of u in float64, u'' from
differences of u'), by LJ_m_n(12, 6) == LJ_12_6, and by the two independent
float32 codes agreeing.
"""
import numpy as np
import pytest

import gamdpy_amd as gp
from oracle import md_oracle as orc

F32 = np.float32


def test_ipl_reference_known_answers():
    """test_potential_functions.py:16-20."""
    for n, r, a in [(12, 1, 1), (12, 2 ** (1 / 6), 3), (6, 2 ** (1 / 6), 1), (6, 2, 4), (1, 2, 3)]:
        expected = (a * r ** (-n), n * a * r ** (-n - 2), a * n * (n + 1) * r ** (-n - 2))
        assert np.all(np.isclose(orc.ipl_n(F32(r), F32(a), 0.0, (n, 0.0)), expected))
        assert np.all(np.isclose(gp.make_IPL_n(n)(r, (a,)), expected))


def test_harmonic_repulsion_reference_known_answers():
    """test_potential_functions.py:44-57."""
    for f in (lambda r, p: orc.harmonic_repulsion(r, p[0], p[1]), gp.harmonic_repulsion):
        u, s, umm = f(F32(0.5), (2.0, 1.0))
        assert np.isclose(u, 0.25) and np.isclose(s, 2.0) and np.isclose(umm, 2.0)
        eps, sig, r = 1.43, 1.37, 0.98
        u, s, umm = f(F32(r), (eps, sig))
        assert np.isclose(u, np.float32(0.5 * eps * (1.0 - r / sig) ** 2))
        assert np.isclose(s, eps * (1.0 - r / sig) / sig / r)
        assert np.isclose(umm, eps / sig ** 2)


def test_hertzian_reference_known_answers():
    """test_potential_functions.py:59-70."""
    for f in (lambda r, p: orc.hertzian(r, p[0], p[2], (p[1], 0.0)), gp.hertzian):
        u, s, umm = f(F32(0.5), (1.0, 2.0, 1.0))
        assert np.isclose(u, 0.25) and np.isclose(s, 2.0) and np.isclose(umm, 2.0)
        eps, alpha, sig, r = 1.43, 3.1, 1.24, 0.98
        u, s, umm = f(F32(r), (eps, alpha, sig))
        assert np.isclose(u, eps * (1.0 - r / sig) ** alpha)
        assert np.isclose(s, alpha * eps * (1.0 - r / sig) ** (alpha - 1) / sig / r)
        assert np.isclose(umm, eps * alpha * (alpha - 1) * (1.0 - r / sig) ** (alpha - 2) / sig / sig)


def test_lj_m_n_12_6_is_lj_12_6():
    """The reference's add_potential_functions check (test_potential_functions.py:22-26) in Mie form."""
    for r, a12, a6 in [(1.0, 1.0, -1.0), (2 ** (1 / 6), 3.0, -3.0), (2.0, 4.0, 4.0), (1.37, 4.0, -4.0)]:
        got = orc.lj_m_n(F32(r), a12, a6, (12, 6))
        ref = orc.lj_12_6(F32(r), a12, a6)
        assert np.allclose(got, ref, rtol=2e-6, atol=1e-5)
        assert np.allclose(gp.make_LJ_m_n(12, 6)(r, (a12, a6)), ref, rtol=2e-6, atol=1e-5)


CASES = [("IPL_n", (1.7, 0.0), (12, 0.0)), ("IPL_n", (0.9, 0.0), (7.5, 0.0)), ("LJ_m_n", (4.0, -4.0), (12, 6)),
         ("LJ_m_n", (2.5, -3.1), (9.0, 4.5)), ("harmonic_repulsion", (1.6, 1.2), None),
         ("hertzian", (1.3, 1.15), (2.5, 0.0))]


@pytest.mark.parametrize("base,p,exps", CASES)
def test_closed_forms_by_finite_differences(base, p, exps):
    """s = -u'/r and u'' from central differences of the float64 evaluation of the oracle's formula."""
    f = orc.BASE_FUNCTIONS[base]
    e = exps if exps is not None else (0.0, 0.0)
    for r in (0.8, 0.95, 1.1):
        u, s, umm = (float(x) for x in f(F32(r), p[0], p[1], e))
        h = 1e-3
        up, um = (float(f(F32(r + h), p[0], p[1], e)[0]), float(f(F32(r - h), p[0], p[1], e)[0]))
        du = (up - um) / (2 * h)
        assert abs(-du / r - s) <= 2e-3 * max(1.0, abs(s)), (r, s, -du / r)
        sp, sm = (float(f(F32(r + h), p[0], p[1], e)[1]), float(f(F32(r - h), p[0], p[1], e)[1]))
        # u'' = d(-r s)/dr... = -s - r ds/dr
        d2u = -s - r * (sp - sm) / (2 * h)
        assert abs(d2u - umm) <= 5e-3 * max(1.0, abs(umm)), (r, umm, d2u)


@pytest.mark.parametrize("base,p,exps", CASES)
def test_host_functions_match_the_oracle(base, p, exps):
    """Two float32 codes: the package's host functions vs the oracle."""
    r = np.linspace(0.75, 1.1, 40).astype(F32)   # inside every cut (hertzian: sigma 1.15)
    if base == "IPL_n":
        pkg = gp.make_IPL_n(exps[0])(r, (p[0],))
    elif base == "LJ_m_n":
        pkg = gp.make_LJ_m_n(*exps)(r, p)
    elif base == "harmonic_repulsion":
        pkg = gp.harmonic_repulsion(r, p)
    else:
        pkg = gp.hertzian(r, (p[0], exps[0], p[1]))
    ref = orc.BASE_FUNCTIONS[base](r, p[0], p[1], exps if exps is not None else (0.0, 0.0))
    for a, b in zip(pkg, ref):
        # pointwise near the zero crossing of a Mie term the cancellation
        # leaves an absolute, not relative, difference
        np.testing.assert_allclose(a, b, rtol=3e-6, atol=3e-6 * float(np.max(np.abs(b))))


def test_param_table_maps_each_function():
    full = lambda x: np.full((2, 2), x)
    pp = gp.PairPotential(gp.hertzian, params=[[[1.3, 1.4], [1.4, 1.5]], full(2.5), [[1.0, 1.1], [1.1, 1.2]],
                                               full(1.2)], max_num_nbs=64)
    table, max_cut = pp.param_table()
    assert pp.exps == (2.5, 0.0)
    np.testing.assert_allclose(table[..., 0], [[1.3, 1.4], [1.4, 1.5]])
    np.testing.assert_allclose(table[..., 1], [[1.0, 1.1], [1.1, 1.2]])
    np.testing.assert_allclose(table[..., 2], 1.2)
    pp = gp.PairPotential(gp.harmonic_repulsion, params=[1.5, 1.1], max_num_nbs=64)
    table, max_cut = pp.param_table()
    assert table[0, 0, 1] == np.float32(1.1) and table[0, 0, 2] == np.float32(1.1) and max_cut == np.float32(1.1)
    pp = gp.PairPotential(gp.apply_shifted_force_cutoff(gp.make_LJ_m_n(9, 3)), params=[2.0, -3.0, 2.5],
                          max_num_nbs=64)
    table, _ = pp.param_table()
    assert pp.exps == (9.0, 3.0) and tuple(table[0, 0]) == (2.0, -3.0, 2.5)
    pp = gp.PairPotential(gp.make_IPL_n(12), params=[1.7, 2.5], max_num_nbs=64)
    table, _ = pp.param_table()
    assert pp.exps == (12.0, 0.0) and tuple(table[0, 0]) == (np.float32(1.7), 0.0, 2.5)


def test_unsupported_forms_raise():
    with pytest.raises(ValueError):
        gp.PairPotential(gp.hertzian, params=[np.ones((2, 2)), [[2.5, 2.0], [2.0, 2.5]], np.ones((2, 2)),
                                              np.ones((2, 2))], max_num_nbs=64).param_table()
    with pytest.raises(ValueError):
        gp.PairPotential(gp.make_IPL_n(6, first_parameter=1), params=[1.0, 1.0, 2.5], max_num_nbs=64)
