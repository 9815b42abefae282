"""Exclusions (PairPotential(..., exclusions=), nblist.py:108-124 and the
same test in nblist_linked_lists.py) on the GPU: neighbour sets of both list
kinds equal the oracle's with the same exclusions, and forces, energies and
virials match the float32 oracle over those lists (relative 2e-5)."""
import numpy as np
import pytest

import gamdpy_amd as gp
from oracle import md_oracle as orc
from helpers import lj_fcc, plan, rows_from_array
from test_exclusions_oracle import random_exclusions

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nblist", ["N squared", "linked lists"])
def test_exclusions_lists_and_forces(nblist):
    conf = lj_fcc([8, 8, 8], jitter=0.15)
    r, box, N = conf['r'], conf.simbox.lengths, conf.N
    full = orc.nblist_nsquared(r, box, 2.5, 0.5)
    ex = random_exclusions(full, N, seed=11)
    pp = gp.PairPotential(gp.apply_shifted_potential_cutoff(gp.LJ_12_6_sigma_epsilon), params=[1.0, 1.0, 2.5],
                          max_num_nbs=200, exclusions=ex)
    ev = gp.Evaluator(conf, pp, compute_plan=plan(tp=4, nblist=nblist))
    ev.evaluate()
    rows = orc.nblist_nsquared(r, box, 2.5, 0.5, exclusions=ex)
    assert sum(len(a) for a in full) - sum(len(a) for a in rows) == int(ex[:, -1].sum())
    got = rows_from_array(ev.ctx.nblist())
    for a, b in zip(got, rows):
        np.testing.assert_array_equal(a, b)
    table, _ = pp.param_table()
    f, U, W, _ = orc.pair_forces(r, conf.ptype, box, table, rows)
    rel = np.max(np.abs(conf['f'] - f)) / np.max(np.abs(f))
    print(nblist, rel, int(ex[:, -1].sum()))
    assert rel < 2e-5
    np.testing.assert_allclose(conf['U'], U, rtol=2e-5, atol=2e-5 * (1 + np.abs(U).max()))
    np.testing.assert_allclose(conf['W'], W, rtol=2e-5, atol=2e-5 * (1 + np.abs(W).max()))
