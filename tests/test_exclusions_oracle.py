"""The oracle's exclusion rule (nblist.py:108-124: j is left out of row i when
it appears among the first exclusions[i, -1] entries of exclusions[i]) pinned
by a set difference against the list without exclusions."""
import numpy as np

from oracle import md_oracle as orc
from helpers import lj_fcc


def random_exclusions(rows, N, seed, most=4):
    """Up to `most` of each particle's list neighbours, count in the last column."""
    rng = np.random.default_rng(seed)
    ex = np.zeros((N, most + 1), np.int32)
    for i in range(N):
        k = int(rng.integers(0, most + 1))
        pick = rng.choice(rows[i], size=min(k, len(rows[i])), replace=False) if len(rows[i]) else []
        ex[i, :len(pick)] = pick
        ex[i, -1] = len(pick)
    return ex


def test_oracle_exclusions_are_a_set_difference():
    conf = lj_fcc([4, 4, 4], jitter=0.1)
    r, box, N = conf['r'], conf.simbox.lengths, conf.N
    rows = orc.nblist_nsquared(r, box, 2.5, 0.5)
    ex = random_exclusions(rows, N, seed=3)
    # stale entries past the count must be ignored: fill them with a real neighbour
    for i in range(N):
        if ex[i, -1] < ex.shape[1] - 1 and len(rows[i]):
            ex[i, ex[i, -1]] = rows[i][0] if rows[i][0] not in ex[i, :ex[i, -1]] else ex[i, ex[i, -1]]
    got = orc.nblist_nsquared(r, box, 2.5, 0.5, exclusions=ex)
    removed = 0
    for i in range(N):
        want = sorted(set(rows[i].tolist()) - set(ex[i, :ex[i, -1]].tolist()))
        assert got[i].tolist() == want
        removed += len(rows[i]) - len(want)
    assert removed == int(ex[:, -1].sum())
