"""Writes the golden SNAP fixtures of tests/golden/*.npz from the upstream SNAP C library.

Run in a container that holds the upstream tree: ``make -C oracle ref`` compiles
src/Julia/Potential/cpuSnap.cpp (unmodified, -O2, no fast-math) into oracle/_ref/libcpusnap_ref.so;
this script then drives that library through the same call chain as the Julia driver
(snapcpp.jl:323-426 for energy/force/virial, :539-665 for descriptors) on seeded inputs and stores
inputs and outputs.  Each fixture holds its own pair arrays so that it does not depend on any
neighbour builder.

At twojmax >= 7 the upstream force routine cpuSnapComputeFi writes past its stack arrays
(cpuSnap.cpp:1286-1306; the process aborts with "stack smashing detected"), so for the twojmax=8
fixture the reference forces and virials are taken from the upstream descriptor route instead:
F = sum_k beta_k bd_k and W = sum_k beta_k bv_k (cpuComputeDbidrj + cpuSnapTallyBispectrumDeriv/
Virial), which is free of that defect.  The fixture records this in ``force_source``.

    python tests/golden/make_golden.py [outdir]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mdp_amd import _abi, lattice, snapcpp  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "libcpusnap_ref.so")

# name: (kind, ncell/natoms, a, sigma, twojmax, ntypes, flags dict, seed)
CASES = {
    "j4_bcc3": dict(kind="bcc", n=3, a=3.316, sigma=0.1, J=4, ntypes=1, seed=11),
    "j6_bcc3": dict(kind="bcc", n=3, a=3.316, sigma=0.1, J=6, ntypes=1, seed=12),
    "j8_bcc2": dict(kind="bcc", n=2, a=3.316, sigma=0.1, J=8, ntypes=1, seed=13),
    "j6_liquid40": dict(kind="liquid", n=40, a=0.0554, sigma=0.0, J=6, ntypes=1, seed=14),
    "j4_fcc2_chem": dict(kind="fcc", n=2, a=4.15, sigma=0.1, J=4, ntypes=2, seed=15, chemflag=1),
    "j4_bcc2_flags": dict(kind="bcc", n=2, a=3.316, sigma=0.1, J=4, ntypes=1, seed=16,
                          bnormflag=1, wselfallflag=1, quadraticflag=1),
}

RCUTFAC, RFAC0, RMIN0 = 4.67637, 0.99363, 0.0


def build_case(lib, c):
    if c["kind"] == "liquid":
        x, box = lattice.random_packing(c["n"], c["a"], 2.2, seed=c["seed"])
        x = x + np.random.default_rng(c["seed"]).normal(0, 0.0, x.shape)
    else:
        x, box = lattice.crystal(c["kind"], c["n"], c["a"], c["sigma"], c["seed"])
    n = len(x)
    ntypes = c["ntypes"]
    rng = np.random.default_rng(c["seed"] + 100)
    t = (rng.integers(1, ntypes + 1, size=n) if ntypes > 1 else np.ones(n)).astype(np.int32)
    param = np.array([ntypes, c["J"], RCUTFAC, RFAC0, RMIN0, 1, 1, c.get("quadraticflag", 0),
                      c.get("chemflag", 0), c.get("bnormflag", 0), c.get("wselfallflag", 0)], float)
    elemradius = np.array([0.5, 0.45][:ntypes])
    elemweight = np.array([1.0, 0.8][:ntypes])
    s = snapcpp.initsna(lib, param, elemradius, elemweight)
    coeff = np.concatenate([lattice.snap_coefficients(s.ncoeff, seed=c["seed"] + 1000 * k)
                            for k in range(ntypes)])
    s = snapcpp.initsna(lib, param, elemradius, elemweight, coeff)
    B = np.diag(box)
    p = snapcpp.pairs_from_box(x, t, B[0], B[1], B[2], [1, 1, 1], s)
    return x, box, t, param, elemradius, elemweight, coeff, s, p


def main(outdir=HERE):
    lib = _abi.bind(REF)
    for name, c in CASES.items():
        x, box, t, param, er, ew, coeff, s, p = build_case(lib, c)
        N, iu = p["n"], s.idxu_max
        nel = s.nelements
        # U after compute_ui + wself/mirror, Y, E (upstream chain, snapcpp.jl:386-414)
        Ur, Ui = np.zeros(N * iu * nel), np.zeros(N * iu * nel)
        lib.cpuSnapComputeUi(Ur, Ui, s.rootpqarray, p["rij"], s.wjelem, s.radelem, s.rmin0, s.rfac0,
                             s.rcutfac, s.map, p["ai"], p["ti"], p["tj"], s.twojmax, iu, N, p["ijnum"],
                             s.switchflag, s.chemflag)
        lib.cpuAddWself2Ui(Ur, Ui, s.wself, s.idxu_block, p["atomtype"], s.map, s.wselfallflag,
                           s.chemflag, iu, nel, s.twojmax, N)
        e = np.zeros(N)
        lib.cpuSnapComputeEi(e, Ur, Ui, s.cglist, s.bzero, s.coeffelem, p["ilist"], s.map, p["atomtype"],
                             s.idxb, s.idxcg_block, s.idxu_block, s.twojmax, s.idxb_max, iu, nel,
                             s.ncoeffall, s.bnormflag, s.bzeroflag, s.wselfallflag, s.chemflag, N)
        Yr, Yi = np.zeros(N * iu * nel), np.zeros(N * iu * nel)
        lib.cpuSnapComputeYi(Yr, Yi, Ur, Ui, s.cglist, s.coeffelem, s.map, p["atomtype"], s.idxz,
                             s.idxb_block, s.idxu_block, s.idxcg_block, s.twojmax, s.idxb_max, iu,
                             s.idxz_max, nel, s.ncoeffall, s.bnormflag, s.chemflag, N)
        bi, bd, bv, extra = snapcpp.snapdescriptors_pairs(lib, s, p)
        if s.twojmax <= 6:
            _, f, v = snapcpp.snappotential_pairs(lib, s, p)
            force_source = "cpuSnapComputeFi"
        else:
            # descriptor route, per-element coefficient blocks (single element here)
            beta = coeff[1:s.ncoeffall]
            f = np.einsum("k,kna->na", beta, bd)
            vtot = np.einsum("k,ka->a", beta, bv)[[0, 1, 2, 5, 4, 3]]
            v = np.full((N, 6), np.nan)
            v[0] = vtot   # only the total virial is defined by this route (row 0 holds it)
            force_source = "descriptor_route"
        out = os.path.join(outdir, f"snap_{name}.npz")
        np.savez_compressed(
            out, x=x, box=box, types=t, param=param, elemradius=er, elemweight=ew, coeff=coeff,
            rij=p["rij"], ai=p["ai"], aj=p["aj"], ti=p["ti"], tj=p["tj"],
            rootpqarray=s.rootpqarray, cglist=s.cglist, idxz=s.idxz, idxb=s.idxb,
            idxu_block=s.idxu_block, idxcg_block=s.idxcg_block, idxz_block=s.idxz_block,
            idxb_block=s.idxb_block, idx_max=s.idx_max,
            Ur=Ur, Ui=Ui, eatom=e, Yr=Yr, Yi=Yi, fatom=f, vatom=v, force_source=force_source,
            bi=bi, bd=bd, bv=bv, blist=extra["blist"])
        print(f"{out}: N={N} pairs={p['ijnum']} J={s.twojmax} force={force_source} "
              f"{os.path.getsize(out) // 1024} KiB")


if __name__ == "__main__":
    main(*(sys.argv[1:2]))
