"""Convert the rumd3 Kob-Andersen snapshots shipped with the reference into
small .npz fixtures (data only: positions, images, velocities, forces,
per-particle potential energy and virial, box).

Source data: examples/Data/NVT_N4000_T2.0_rho1.2_KABLJ_rumd3/TrajectoryFiles/
restart000{0,1}.xyz.gz of the gamdpy checkout (written by rumd3, an
independent C++/CUDA MD code; KA-LJ N=4000, rho=1.2, T=2.0, shifted-potential
cutoff 2.5 sigma_ij).  Also copies the rumd3 mean energies and the 1000-bin
partial g(r) table of the same directory.

    python tests/golden/make_rumd3_fixtures.py /root/reference
"""
import gzip
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SUB = "examples/Data/NVT_N4000_T2.0_rho1.2_KABLJ_rumd3"


def read_xyz_gz(path):
    with gzip.open(path, "rt") as fh:
        n = int(fh.readline())
        header = fh.readline()
        meta = dict(tok.split("=", 1) for tok in header.split() if "=" in tok)
        cols = meta["columns"].split(",")
        data = np.loadtxt(fh, max_rows=n)
    box = np.array([float(x) for x in meta["sim_box"].split(",")[1:4]])
    out = {"box": box.astype(np.float64)}
    col = {c: i for i, c in enumerate(cols)}
    out["type"] = data[:, col["type"]].astype(np.int32)
    out["r"] = data[:, [col["x"], col["y"], col["z"]]].astype(np.float32)
    out["im"] = data[:, [col["imx"], col["imy"], col["imz"]]].astype(np.int32)
    out["v"] = data[:, [col["vx"], col["vy"], col["vz"]]].astype(np.float32)
    out["f"] = data[:, [col["fx"], col["fy"], col["fz"]]].astype(np.float32)
    out["pe"] = data[:, col["pe"]].astype(np.float32)
    out["vir"] = data[:, col["vir"]].astype(np.float32)
    return out


def main(ref_root):
    src = os.path.join(ref_root, SUB)
    for k in (0, 1):
        d = read_xyz_gz(os.path.join(src, "TrajectoryFiles", f"restart000{k}.xyz.gz"))
        np.savez_compressed(os.path.join(HERE, f"ka_rumd3_restart000{k}.npz"), **d)
    en = np.loadtxt(os.path.join(src, "energies_mean.dat"))
    rdf = np.loadtxt(os.path.join(src, "KABLJ_rdf_R1.200_T0.800_rumd3.dat"))
    np.savez_compressed(os.path.join(HERE, "ka_rumd3_means_rdf.npz"),
                        energies_mean=en, energies_columns=np.array("ke pe p sxy sxz syz T Etot W".split()),
                        rdf_T0800=rdf.astype(np.float32))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
