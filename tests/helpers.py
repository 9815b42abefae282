"""Shared builders for the tests."""
import os

import numpy as np

import gamdpy_amd as gp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

KA_SIG = [[1.00, 0.80], [0.80, 0.88]]
KA_EPS = [[1.00, 1.50], [1.50, 0.50]]


def lj_fcc(cells, rho=0.8442, T=1.44, seed=0, jitter=0.0, flags=None):
    conf = gp.Configuration(D=3, compute_flags=flags)
    conf.make_lattice(gp.unit_cells.FCC, cells=cells, rho=rho)
    conf['m'] = 1.0
    if jitter:
        rng = np.random.default_rng(seed + 100)
        conf['r'] = conf['r'] + rng.uniform(-jitter, jitter, conf['r'].shape).astype(np.float32)
    conf.randomize_velocities(temperature=T, seed=seed)
    return conf


def ka_fcc(cells, rho=1.2, T=0.45, seed=0, jitter=0.0, flags=None):
    conf = lj_fcc(cells, rho=rho, T=T, seed=seed, jitter=jitter, flags=flags)
    rng = np.random.default_rng(seed)
    N = conf.N
    nB = int(round(0.2 * N))
    pt = np.zeros(N, np.int32)
    pt[rng.permutation(N)[:nB]] = 1
    conf.ptype = pt
    return conf


def ka_pair_potential(max_nbs=300, cutoff="SP"):
    base = gp.LJ_12_6_sigma_epsilon
    func = gp.apply_shifted_potential_cutoff(base) if cutoff == "SP" else gp.apply_shifted_force_cutoff(base)
    cut = np.array(KA_SIG) * 2.5
    return gp.PairPotential(func, params=[KA_SIG, KA_EPS, cut], max_num_nbs=max_nbs)


def lj_pair_potential(max_nbs=300, cutoff="SP", cut=2.5):
    base = gp.LJ_12_6_sigma_epsilon
    func = {"SP": gp.apply_shifted_potential_cutoff, "SF": gp.apply_shifted_force_cutoff,
            None: (lambda f: f)}[cutoff](base)
    return gp.PairPotential(func, params=[1.0, 1.0, cut], max_num_nbs=max_nbs)


def rumd3_conf(k=0):
    d = np.load(os.path.join(GOLDEN, f"ka_rumd3_restart000{k}.npz"))
    conf = gp.Configuration(D=3)
    conf['r'] = d['r']
    conf.simbox = gp.Orthorhombic(3, d['box'])
    conf['m'] = 1.0
    conf['v'] = d['v']
    conf.ptype = d['type'].astype(np.int32)
    return conf, d


def plan(tp=4, gridsync=False, nblist='N squared', skin=0.5, graph=True, pb=None):
    return {'pb': pb or max(1, 256 // tp), 'tp': tp, 'skin': skin, 'UtilizeNIII': False, 'gridsync': gridsync,
            'nblist': nblist, 'graph': graph}


def rows_from_array(arr):
    """Reference-layout list (N, max+1) -> list of sorted rows."""
    return [np.sort(row[:min(row[-1], arr.shape[1] - 1)]) for row in arr]
