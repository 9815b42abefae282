"""Batched replicas on the GPU.

A replica advanced inside a batch runs the same per-particle arithmetic as the
same system advanced alone (forces and integration of a particle do not
depend on which block computes them; lists are in a fixed order), so NVE
trajectories without cross-block sums are bitwise equal to single-system
runs with the same lanes per particle and list kind.  Sums over particles
(thermostat, saver rows) are grouped by block and agree to rounding.
"""
import numpy as np
import pytest

import gamdpy_amd as gp
from helpers import KA_SIG, KA_EPS

pytestmark = pytest.mark.gpu


def _ka(cells, rho, T, seed):
    conf = gp.Configuration(D=3)
    conf.make_lattice(gp.unit_cells.FCC, cells=cells, rho=rho)
    conf['m'] = 1.0
    rng = np.random.default_rng(seed)
    pt = np.zeros(conf.N, np.int32)
    pt[rng.permutation(conf.N)[:int(round(0.2 * conf.N))]] = 1
    conf.ptype = pt
    conf.randomize_velocities(temperature=T, seed=seed)
    return conf


def _pp():
    func = gp.apply_shifted_potential_cutoff(gp.LJ_12_6_sigma_epsilon)
    return gp.PairPotential(func, params=[KA_SIG, KA_EPS, np.array(KA_SIG) * 2.5], max_num_nbs=300)


# (cells, rho, T): N = 4 * prod(cells); the last two use linked lists
SYSTEMS = [((5, 5, 5), 1.2, 0.8), ((6, 5, 7), 1.1, 1.5), ((4, 4, 4), 1.3, 0.45), ((12, 12, 12), 1.2, 1.0),
           ((10, 12, 14), 1.15, 2.0)]


def _replicas(integ_factory):
    reps = []
    for k, (cells, rho, T) in enumerate(SYSTEMS):
        conf = _ka(list(cells), rho, T, seed=k)
        reps.append(gp.Replica(conf, _pp(), integ_factory(T)))
    return reps


@pytest.mark.parametrize("driver", ["persistent", "graph"])
def test_batched_nve_equals_single_runs(driver):
    tp, skin, steps = 4, 0.5, 300
    reps = _replicas(lambda T: gp.NVE(dt=0.004))
    singles = []
    for r in reps:
        conf = gp.Configuration(D=3)
        conf['r'] = r.configuration['r'].copy()
        conf['v'] = r.configuration['v'].copy()
        conf['m'] = 1.0
        conf.ptype = r.configuration.ptype.copy()
        conf.simbox = gp.Orthorhombic(3, r.configuration.simbox.lengths.copy())
        singles.append(conf)
    batch = gp.BatchedSimulation(reps, num_timeblocks=2, steps_per_timeblock=steps // 2, tp=tp, skin=skin,
                                 driver=driver)
    for _ in batch.run_timeblocks():
        pass
    for r, conf in zip(reps, singles):
        plan = gp.batch.replica_plan(r, tp, skin)
        plan = dict(plan, graph=True)
        pp = _pp()
        sim = gp.Simulation(conf, pp, gp.NVE(dt=0.004), [], num_timeblocks=2, steps_per_timeblock=steps // 2,
                            compute_plan=plan, storage='memory', steps_in_kernel_test=0)
        # the batch builds its lists at step 0 of its first run; so does this
        # run (a rebuild step sums forces in list-build order, so the rebuild
        # steps must coincide for bitwise equality)
        built_at_setup = int(pp.nblist.d_nbflag.copy_to_host()[2])
        sim.ctx.force_rebuild()
        for _ in sim.run_timeblocks():
            pass
        assert np.array_equal(conf['r'], r.configuration['r']), plan['nblist']
        assert np.array_equal(conf['v'], r.configuration['v'])
        assert np.array_equal(conf.r_im, r.configuration.r_im)
        assert pp.nblist.d_nbflag.copy_to_host()[2] - built_at_setup == r.ctx.nbflag()[2]


def test_batched_nvt_state_points():
    """Each replica thermostats to its own temperature (NVT.py:75-124 per replica)."""
    reps = _replicas(lambda T: gp.NVT(temperature=T, tau=0.2, dt=0.004))
    for r in reps:
        r.runtime_actions.append(gp.ScalarSaver(16))
    batch = gp.BatchedSimulation(reps, num_timeblocks=4, steps_per_timeblock=1024, tp=4, skin=0.5,
                                 driver='persistent')
    for _ in batch.run_timeblocks():
        pass
    for (cells, rho, T), r, out in zip(SYSTEMS, reps, batch.outputs):
        (K,) = gp.extract_scalars(out, ['K'], first_block=2)
        N = r.configuration.N
        Tk = np.mean(2 * K / (3 * (N - 1)))
        print(cells, rho, T, Tk)
        assert abs(Tk - T) < 0.05 * T + 3.0 / np.sqrt(N)


def test_batch_layout_and_errors():
    reps = _replicas(lambda T: gp.NVE(dt=0.004))
    b = gp.BatchedSimulation(reps, num_timeblocks=1, steps_per_timeblock=10, tp=4, driver='persistent',
                             blocks=[1, 2, 3, 4, 5])
    n, grid = b.batch.size()
    assert n == 5 and grid % 8 == 0 and grid >= 15
    with pytest.raises(gp._native.NativeError):
        b.replicas[0].ctx.run(0.0, 1)            # a replica runs only with its batch
    with pytest.raises(ValueError):
        gp.BatchedSimulation(reps, num_timeblocks=1, steps_per_timeblock=10, blocks=[1, 2])
    with pytest.raises(gp._native.NativeError):
        gp.BatchedSimulation(reps, num_timeblocks=1, steps_per_timeblock=10, blocks=[100000] * 5)


def test_autotune_runs_and_keeps_state():
    reps = _replicas(lambda T: gp.NVE(dt=0.004))
    r0 = [r.configuration['r'].copy() for r in reps]
    b = gp.BatchedSimulation(reps, num_timeblocks=1, steps_per_timeblock=50, tp=4)
    best = b.autotune(steps=50, tps=(2, 4), fractions=(1.0, 2.0), verbose=True)
    assert best['driver'] in ('persistent', 'graph')
    for r, x in zip(reps, r0):
        assert np.array_equal(r.configuration['r'], x)
    for _ in b.run_timeblocks():
        pass
