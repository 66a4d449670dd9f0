"""Writes tests/golden/ta_json.npz: the tantalum DFT training set of the upstream fitting example
(fitting/snap/Ta/JSON, 12 folders, 363 configurations, 4,224 atoms) as mdp_amd.dataset reads it with the
options of fitting/snap/Ta/fitsnap.jl:4-52 (species Ta, transposelattice = true, all data, not shuffled)
and the per-folder weights of energies, forces and stresses of fitsnap.jl:18-29.  Data only: positions
(rotated as the upstream reader rotates them), cells, energies, forces and stresses.

    python tests/golden/make_ta_dataset.py [json_root] [outdir]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mdp_amd import dataset  # noqa: E402

FOLDERS = ["Displaced_A15", "Displaced_BCC", "Displaced_FCC", "Elastic_BCC", "Elastic_FCC", "GSF_110", "GSF_112",
           "Liquid", "Surface", "Volume_A15", "Volume_BCC", "Volume_FCC"]
WEIGHTS = [[100, 1, 1.00E-08], [100, 1, 1.00E-08], [100, 1, 1.00E-08], [1.00E-08, 1.00E-08, 0.0001],
           [1.00E-09, 1.00E-09, 1.00E-09], [100, 1, 1.00E-08], [100, 1, 1.00E-08], [4.67E+02, 1, 1.00E-08],
           [100, 1, 1.00E-08], [1.00E+00, 1.00E-09, 1.00E-09], [1.00E+00, 1.00E-09, 1.00E-09],
           [1.00E+00, 1.00E-09, 1.00E-09]]


def main(src, outdir):
    groups = [dataset.read_folder(os.path.join(src, f), ("Ta",), "json", True, w) for f, w in zip(FOLDERS, WEIGHTS)]
    path = os.path.join(outdir, "ta_json.npz")
    dataset.save_npz(path, groups)
    print(f"wrote {path}: {sum(g.nconfigs for g in groups)} configurations, "
          f"{sum(int(g.natom.sum()) for g in groups)} atoms")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/fitting/snap/Ta/JSON",
         sys.argv[2] if len(sys.argv) > 2 else HERE)
