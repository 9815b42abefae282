"""Radial distribution function (calculators/calculator_radial_distribution.py:12-184).

``update()`` histograms all ordered pairs i != j with |r_ij| < min(L)/2 on the
GPU (LDS-privatised integer histograms, each unordered pair visited once);
``read()`` normalises on the host exactly as the reference does.
"""
import math

import numpy as np

from . import _native


class CalculatorRadialDistribution:
    """g(r) of a configuration, accumulated over calls to update().

    Parameters
    ----------
    configuration : Configuration
    bins : number of bins over [0, min(L)/2)
    compute_plan : accepted for interface compatibility (unused)
    ptype : optional type array replacing configuration.ptype
    """

    def __init__(self, configuration, bins, compute_plan=None, ptype=None) -> None:
        if configuration.D != 3:
            raise NotImplementedError("CalculatorRadialDistribution: the MI355X kernel is three-dimensional")
        self.configuration = configuration
        self.ptype_user = None if ptype is None else np.asarray(ptype, np.int32)
        self.bins = bins
        self.count = 0
        self.compute_plan = compute_plan
        self._rdf_list = []
        self._pending = False
        nptypes = int(np.max(ptype if ptype is not None else configuration.ptype)) + 1
        self.gr_bins = np.zeros((nptypes, nptypes, self.bins), dtype=np.float64)
        self._h = _native.RdfHandle(configuration.N, nptypes, bins)

    def _ptype(self):
        return self.ptype_user if self.ptype_user is not None else self.configuration.ptype

    @property
    def rdf_list(self):
        """Counts of every update so far (an overlapped update is waited for here)."""
        self._collect()
        return self._rdf_list

    def _collect(self):
        if self._pending:
            self._pending = False
            self._rdf_list.append(self._h.collect().astype(np.float64))

    def update(self, overlap=False):
        """Add the current configuration's pair counts (calculator_radial_distribution.py:111-119).

        overlap=True (configuration on the device): snapshot the positions and
        histogram them on the calculator's own stream while the simulation's
        next steps run; the counts are the same and are collected at the next
        update or when rdf_list / read() is used."""
        self.count += 1
        conf = self.configuration
        ctx = getattr(conf, "_ctx", None)
        self._collect()
        if ctx is not None:
            if overlap:
                self._h.launch_ctx(ctx, self.ptype_user)
                self._pending = True
                return
            counts = self._h.compute_ctx(ctx, self.ptype_user)
        else:
            counts = self._h.compute_host(conf['r'], self._ptype(), conf.simbox.lengths)
        self._rdf_list.append(counts.astype(np.float64))

    def read(self):
        """Normalised total and partial g(r) (calculator_radial_distribution.py:121-159)."""
        bins = self.rdf_list[0].shape[2]
        min_box_dim = np.min(self.configuration.simbox.lengths)
        bin_width = (min_box_dim / 2) / bins
        rdf_ptype = np.array(self.rdf_list)
        rho = self.configuration.N / np.prod(self.configuration.simbox.lengths)
        D = self.configuration.D
        for i in range(bins):
            r_outer = (i + 1) * bin_width
            r_inner = i * bin_width
            if D % 2 == 0:
                n = D // 2
                unit_hypersphere_volume = math.pi ** n / math.factorial(n)
            else:
                n = (D - 1) // 2
                unit_hypersphere_volume = (2 ** D * math.pi ** n * math.factorial(n)) / math.factorial(D)
            shell_volume = unit_hypersphere_volume * (r_outer ** D - r_inner ** D)
            expected_num = rho * shell_volume
            rdf_ptype[:, :, :, i] /= (expected_num * self.configuration.N)
        rdf = rdf_ptype.sum(axis=1).sum(axis=1)
        ptype = np.array(self._ptype())
        for j in range(rdf_ptype.shape[1]):
            n_j = np.sum(ptype == j) / len(ptype)
            rdf_ptype[:, j, :, :] /= n_j
        for k in range(rdf_ptype.shape[2]):
            n_k = np.sum(ptype == k) / len(ptype)
            rdf_ptype[:, :, k, :] /= n_k
        distances = np.arange(0, bins) * bin_width
        return {'distances': distances, 'rdf': rdf, 'rdf_ptype': rdf_ptype, "ptype": ptype}

    def save_average(self, output_filename="rdf.dat", save_ptype=False) -> None:
        rdf_dict = self.read()
        rdf_total = np.mean(rdf_dict['rdf'], axis=0)
        rdf_ij = []
        header_ij = " "
        for i in range(rdf_dict["rdf_ptype"].shape[1]):
            for j in range(rdf_dict["rdf_ptype"].shape[2]):
                rdf_ij.append(np.mean(rdf_dict['rdf_ptype'][:, i, j, :], axis=0))
                header_ij += f"g[{i}-{j}](r) "
        np.savetxt(output_filename, np.array([rdf_dict['distances'], rdf_total, *rdf_ij]).T,
                   header=f"r g(r) {header_ij} ptype")
        np.savetxt(f"ptype_{output_filename}", np.array(rdf_dict["ptype"]).T, header="ptype")
