"""Host-side logic of the batched replicas and the adjoint wrapper on the CPU
(no device call): per-replica plans, block assignment, autotune trial set,
and the adjoint's class bookkeeping and normalisation."""
import math

import numpy as np
import pytest

import gamdpy_amd as gp
from gamdpy_amd import batch as gb
from gamdpy_amd.adjoint import type_classes

SIG = [[1.00, 0.80], [0.80, 0.88]]
EPS = [[1.00, 1.50], [1.50, 0.50]]


def _replica(N, rho, T=1.0):
    conf = gp.Configuration(D=3)
    conf.make_positions(N=N, rho=rho)
    conf['m'] = 1.0
    pt = np.zeros(conf.N, np.int32)
    pt[: conf.N // 5] = 1
    conf.ptype = pt
    func = gp.apply_shifted_potential_cutoff(gp.LJ_12_6_sigma_epsilon)
    pp = gp.PairPotential(func, params=[SIG, EPS, np.array(SIG) * 2.5], max_num_nbs=256)
    return gp.Replica(conf, pp, gp.NVT(temperature=T, tau=0.2, dt=0.004))


def test_replica_plan_uses_linked_lists_where_the_box_admits_them():
    small = _replica(1000, 1.3)      # L = 9.2: 6 cells of (2.5 + 0.5)/2 -> N squared
    large = _replica(4000, 1.1)      # L = 15.4: 10 cells -> linked lists
    assert gb.replica_plan(small, 4, 0.5)['nblist'] == 'N squared'
    assert gb.replica_plan(large, 4, 0.5)['nblist'] == 'linked lists'
    p = gb.replica_plan(large, 2, 0.3)
    assert p['tp'] == 2 and p['pb'] == 128 and p['skin'] == 0.3 and not p['gridsync']


def test_assign_blocks_is_proportional_to_work_and_capped():
    reps = [_replica(n, 1.2) for n in (1000, 2000, 4000, 8000)]
    blocks = gb.assign_blocks(reps, 4, 0.5, total=60)
    assert np.all(blocks >= 1)
    # proportional to N (same density, same list length) within rounding
    np.testing.assert_array_equal(blocks, [4, 8, 16, 32])
    # at most one particle slot per thread: ceil(N tp / 256)
    caps = [math.ceil(r.configuration.N * 4 / 256) for r in reps]
    np.testing.assert_array_equal(gb.assign_blocks(reps, 4, 0.5, total=10000), caps)
    # at least one block each, however small the budget
    assert np.all(gb.assign_blocks(reps, 4, 0.5, total=2) >= 1)


def test_replica_work_counts_list_length():
    r = _replica(2000, 1.2)
    n_list = r.list_length(0.5)
    assert abs(n_list - 1.2 * 4 * math.pi / 3 * 3.0 ** 3) < 1e-6 * n_list
    assert r.work(0.5) == pytest.approx(2000 * (n_list + 20.0))


def test_batch_rejects_mixed_timesteps_before_any_device_call():
    a = _replica(1000, 1.2)
    b = _replica(1000, 1.2)
    b.integrator = gp.NVT(temperature=1.0, tau=0.2, dt=0.005)
    with pytest.raises(ValueError):
        gp.BatchedSimulation([a, b], num_timeblocks=1, steps_per_timeblock=10)
    with pytest.raises(ValueError):
        gp.BatchedSimulation([], num_timeblocks=1, steps_per_timeblock=10)


def test_adjoint_type_classes():
    assert type_classes(1) == [(0, 0)]
    assert type_classes(2) == [(0, 0), (0, 1), (1, 1)]
    assert type_classes(3) == [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
