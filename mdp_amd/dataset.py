"""Training data of the linear SNAP fit, read as the upstream Julia preprocessing reads it.

Mirrors src/Julia/Preprocessing/readJSON.jl:8-199 (one JSON file: a "Dataset" -> "Data" list of frames
with NumAtoms, Lattice, Stress, Energy, Positions, Forces, AtomTypes), readconfigfile.jl:99-131 (every
file of a folder with the wanted extension, in sorted name order, frames concatenated),
readconfigfile.jl:8-60 (the optional transpose of the lattice, then every lattice rotated into the form
a = (ax, 0, 0), b = (bx, by, 0), c = (cx, cy, cz) by rotatelattice.jl:1-21 and the positions by
R = [a b c] [A B C]^-1, transformcoords.jl:1-11), and readconfigdata/adddata (datastruct.jl:17-60,
readconfigfile.jl:133-176: percentage, weights of energies, forces and stresses per folder).

As upstream, forces and stresses are taken as stored (they are not rotated with the positions), and
stresses enter the fit as (xx, yy, zz, yz, xz, xy), the entries [1, 5, 9, 6, 3, 2] of the stored 3x3
matrix flattened row by row (emldescriptors.jl:252-257).
"""
import json
import os
from dataclasses import dataclass, field

import numpy as np

STRESS_ORDER = [0, 4, 8, 5, 2, 1]   # 0-based [1, 5, 9, 6, 3, 2]


@dataclass
class Frames:
    """Configurations of one data folder: per frame natom, cell (3, 3) rows a, b, c, energy, stress (6,);
    per atom positions, forces and 1-based types, frames concatenated."""
    natom: np.ndarray
    cell: np.ndarray
    x: np.ndarray
    f: np.ndarray
    e: np.ndarray
    stress: np.ndarray
    t: np.ndarray
    weight: np.ndarray = field(default_factory=lambda: np.ones(3))
    name: str = ""

    @property
    def nconfigs(self):
        return len(self.natom)

    def config(self, c):
        a0 = int(self.natom[:c].sum())
        sl = slice(a0, a0 + int(self.natom[c]))
        return self.x[sl], self.cell[c], self.t[sl]


def _json_frames(path, species):
    """readJSON.jl: the frames of one file (lattice and stress flattened as Julia's hcat of the rows)."""
    text = open(path).read()
    data = json.loads(text[text.index("{"):])["Dataset"]["Data"]
    out = []
    for d in data:
        n = int(d["NumAtoms"])
        t = np.ones(n, np.int32)
        if "AtomTypes" in d:
            at = np.asarray(d["AtomTypes"]).ravel()
            for k, sp in enumerate(species):
                t[at == sp] = k + 1
        out.append(dict(natom=n, lattice=np.asarray(d["Lattice"], np.float64).reshape(9),
                        stress=np.asarray(d.get("Stress", np.zeros((3, 3))), np.float64).reshape(9),
                        e=float(np.asarray(d["Energy"], np.float64).ravel()[0]),
                        x=np.asarray(d["Positions"], np.float64).reshape(n, 3),
                        f=np.asarray(d["Forces"], np.float64).reshape(n, 3), t=t))
    return out


def rotatelattice(A, B, C):
    """rotatelattice.jl:1-21: the same lattice with a along x and b in the xy plane."""
    An, Bn, Cn = np.linalg.norm(A), np.linalg.norm(B), np.linalg.norm(C)
    Ah = A / An
    ax = An
    bx = np.dot(B, Ah)
    by = np.sqrt(Bn ** 2 - bx ** 2)
    cx = np.dot(C, Ah)
    cy = (np.dot(B, C) - bx * cx) / by
    cz = np.sqrt(Cn ** 2 - cx ** 2 - cy ** 2)
    return np.array([ax, 0.0, 0.0]), np.array([bx, by, 0.0]), np.array([cx, cy, cz])


def read_folder(path, species=("Ta",), extension="json", transposelattice=False, weight=(1.0, 1.0, 1.0),
                percentage=100.0):
    """readconfigdata of one DataStruct (randomize = false): a Frames object."""
    files = sorted(f for f in os.listdir(path) if f.rsplit(".", 1)[-1].lower() == extension.lower())
    frames = []
    for fn in files:
        frames.extend(_json_frames(os.path.join(path, fn), list(species)))
    nkeep = int(round(percentage * len(frames) / 100.0))
    frames = frames[:nkeep]
    cells, xs = [], []
    for fr in frames:
        lat = fr["lattice"]
        # Julia keeps the rows of the stored matrix as columns (hcat); "transposelattice" swaps them back
        M = lat.reshape(3, 3).T if not transposelattice else lat.reshape(3, 3)
        A, B, C = M[:, 0], M[:, 1], M[:, 2]
        a, b, c = rotatelattice(A, B, C)
        R = np.column_stack([a, b, c]) @ np.linalg.inv(np.column_stack([A, B, C]))
        xs.append((R @ fr["x"].T).T)
        cells.append(np.stack([a, b, c]))
    return Frames(natom=np.array([fr["natom"] for fr in frames], np.int64), cell=np.array(cells),
                  x=np.concatenate(xs), f=np.concatenate([fr["f"] for fr in frames]),
                  e=np.array([fr["e"] for fr in frames]),
                  stress=np.array([fr["stress"][STRESS_ORDER] for fr in frames]),
                  t=np.concatenate([fr["t"] for fr in frames]), weight=np.asarray(weight, np.float64),
                  name=os.path.basename(os.path.normpath(path)))


def save_npz(path, groups):
    """All groups in one .npz (arrays only)."""
    out = {"names": np.array([g.name for g in groups]), "weights": np.stack([g.weight for g in groups])}
    for k, g in enumerate(groups):
        for key in ("natom", "cell", "x", "f", "e", "stress", "t"):
            out[f"{k}_{key}"] = getattr(g, key)
    np.savez_compressed(path, **out)


def load_npz(path):
    d = np.load(path, allow_pickle=False)
    groups = []
    for k, name in enumerate(d["names"]):
        groups.append(Frames(**{key: d[f"{k}_{key}"] for key in ("natom", "cell", "x", "f", "e", "stress", "t")},
                             weight=d["weights"][k], name=str(name)))
    return groups
