"""Algorithmic FP64 operation counts of the SNAP force path (an FMA counts 2, a multiply or add 1),
enumerated from the loop bounds of the index spaces -- the denominators of the roofline figures.

* ``ui``   per pair : half-U row-chain recursion (ma = 0: 8, interior: 18, ma = j: 8 per row update)
                      plus the sfac-weighted accumulation (4 per emitted element).
* ``dedr`` per pair : U recursion as above plus three derivative recursions (48 / 102 / 48 per row
                      update) plus the contraction with Y (32 per contracted element: 2 for dsfac*U and
                      10 per Cartesian component).
* ``dedr_gen`` per pair: dE/dr through the SU(2) generators: the U recursion as in ``ui`` plus, per
                      contracted element, 4 for Re(conj(Y) U), 5 for the diagonal generator and 10 for each
                      existing neighbour ma +- 1 (two generators), plus 90 per pair for the projection of
                      dq/dr onto the three fields and the combination.
* ``yi``   per atom : the upstream compute_yi work (cpuSnap.cpp:891-989): sum over the idxz entries of
                      10*na*nb (complex product and CG weight) + 4*nb (row weight) + 4 (beta * Z).
* ``yi_md`` per atom: the same restricted to the Y entries the force needs (middle row of even layers
                      only up to ma = mb) plus 4 per entry for the energy partial U.Y.
* ``yi_md_sym`` per atom: what compute_yi executes when the (j1, j1) couplings take each unordered pair
                      of U rows once (the swap symmetry of their two factors): for those couplings only the
                      row pairs mb1 < mb2 (whole) and mb1 = mb2 (the terms m1 <= m2) are computed.
* ``ei``   per atom : the upstream fused energy (cpuSnap.cpp:544-719): Z over the half domain of each
                      bispectrum triple (10*na*nb + 4*nb) + 4 for the U.Z product + 2 for beta*B.
* ``zb``   per atom : every Z entry (cpuSnap.cpp:1899-1969: 10*na*nb + 4*nb per idxz entry) plus B of
                      the bispectrum triples (4 per half-domain U.Z product).
* ``dbdr`` per pair : the U and dU/dr recursions of ``dedr`` plus, per Cartesian component, 4 per
                      (bispectrum term, half-domain element) of the three-term contraction of
                      cpuSnap.cpp:2101-2240 and 2 per term for its weight.
"""
from functools import lru_cache


def _cg_n(m, j1, j2, j):
    lo = max(0, (2 * m - j - j2 + j1) // 2) if (2 * m - j - j2 + j1) >= 0 else 0
    return min(j1, (2 * m - j + j2 + j1) // 2) - lo + 1


def _triples(J):
    for j1 in range(J + 1):
        for j2 in range(j1 + 1):
            for j in range(j1 - j2, min(J, j1 + j2) + 1, 2):
                yield j1, j2, j


def _row_updates(J, derivs):
    first, mid, last = (8, 18, 8) if not derivs else (8 + 48, 18 + 102, 8 + 48)
    tot = 0
    for mb in range(J // 2 + 1):
        for j in range(max(1, 2 * mb), J + 1):
            tot += first + mid * (j - 1) + last
    return tot


@lru_cache(maxsize=None)
def snap_flops(J):
    emitted = sum(j + 1 for j in range(J + 1) for mb in range(j // 2 + 1))
    contracted = sum((mb + 1 if j == 2 * mb else j + 1) for j in range(J + 1) for mb in range(j // 2 + 1))
    ui = _row_updates(J, False) + 4 * emitted
    dedr = _row_updates(J, True) + 32 * contracted
    dedr_gen = _row_updates(J, False) + 90
    for j in range(1, J + 1):
        for mb in range(j // 2 + 1):
            for ma in range((mb if j == 2 * mb else j) + 1):
                dedr_gen += 4 + 5 + 10 * (ma < j) + 10 * (ma > 0)
    yi = yi_md = yi_md_sym = ei = zb = 0
    terms = elems = 0
    half = lambda l: sum((mb + 1 if l == 2 * mb else l + 1) for mb in range(l // 2 + 1))
    for j1, j2, j in _triples(J):
        if j >= j1:
            terms += 3
            elems += half(j) + half(j1) + half(j2)
            zb += 4 * half(j) + 2
        for mb in range(j // 2 + 1):
            nb_ = _cg_n(mb, j1, j2, j)
            for ma in range(j + 1):
                zb += 10 * _cg_n(ma, j1, j2, j) * nb_ + 4 * nb_
    for j1, j2, j in _triples(J):
        for mb in range(j // 2 + 1):
            nb = _cg_n(mb, j1, j2, j)
            for ma in range(j + 1):
                na = _cg_n(ma, j1, j2, j)
                w = 10 * na * nb + 4 * nb
                yi += w + 4
                if 2 * mb < j or ma <= mb:
                    yi_md += w + 4 + 4
                    if j1 == j2:   # rows mb1 < mb2 once, mb1 = mb2 with the terms m1 <= m2
                        c = mb + (j1 + j2 - j) // 2
                        lo = max(0, c - j2)
                        hi = min(j1, c)
                        n_lt = sum(1 for mb1 in range(lo, hi + 1) if mb1 < c - mb1)
                        n_eq = sum(1 for mb1 in range(lo, hi + 1) if mb1 == c - mb1)
                        yi_md_sym += 10 * na * n_lt + 4 * n_lt + n_eq * (10 * ((na + 1) // 2) + 4) + 4 + 4
                    else:
                        yi_md_sym += w + 4 + 4
                    if j >= j1:
                        ei += w + 4
        if j >= j1:
            ei += 2
    dbdr = _row_updates(J, True) + 3 * (4 * elems + 2 * terms)
    return dict(ui=ui, dedr=dedr, dedr_gen=dedr_gen, yi=yi, yi_md=yi_md, yi_md_sym=yi_md_sym, ei=ei, zb=zb,
                dbdr=dbdr)


def dedr_key(variant):
    """Operation count that matches an MD dE/dr kernel variant (8 to 13 use the SU(2) generators)."""
    return "dedr_gen" if 8 <= variant <= 13 else "dedr"


def step_flops_per_atom(J, pairs_per_atom, dedr_variant=9):
    """Plain MD step (Ui + Yi with energy + dE/dr) per atom."""
    f = snap_flops(J)
    return pairs_per_atom * (f["ui"] + f[dedr_key(dedr_variant)]) + f["yi_md"]
