// Launchers of the SNAP device kernels.  All pointers are device pointers; arrays use the layouts
// of src/Julia/Potential/cpuSnap.cpp (atom-fastest U/Y/Z/B blocks, AoS r_ij, 1-based types).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace mdp {

// Per-parameter-set constants every kernel needs (device pointers to the staged tables).
struct SnapDev {
    int J = 0, idxu_max = 0, idxz_max = 0, idxb_max = 0, idxcg_max = 0;
    double rcutfac = 0, rfac0 = 0, rmin0 = 0;
    int switchflag = 1;
    const double* rootpq = nullptr;   // [(J+2)^2]
    const double* cg = nullptr;       // [idxcg_max]
    const double* radelem = nullptr;  // [ntypes+1], 1-based
    const double* wjelem = nullptr;   // [ntypes+1], 1-based
    const int* map = nullptr;         // [ntypes+1] type -> element
    const int* idxz = nullptr;        // [10*idxz_max]
    const int* idxb = nullptr;        // [3*idxb_max]
    const int* idxu_block = nullptr;  // [J+1]
    const int* idxz_block = nullptr;  // [(J+1)^3]
    const int* idxb_block = nullptr;  // [(J+1)^3]
    const int* idxcg_block = nullptr; // [(J+1)^3]
};

bool templated_j(int J);   // J has the compile-time specialised kernels

// ---- force path ---------------------------------------------------------------------------------

enum UiMode { UI_ACCUMULATE = 0, UI_FULL_WSELF = 1, UI_HALF_WSELF = 2 };
// Utot of every atom from its pairs (pair_off CSR over atoms, pairs optionally permuted by perm).
//   UI_ACCUMULATE : U[half rows] += sum_p sfac*U_p (cpuSnapComputeUi semantics, j=0 included)
//   UI_FULL_WSELF : U = wself*I + sum_p sfac*U_p on all (j+1)^2 elements (mirror applied)
//   UI_HALF_WSELF : as UI_FULL_WSELF but only rows 2mb <= j, in the compact half layout
//                   (element (j,mb,ma) at h_block(j) + (j+1)*mb + ma)
// elem_sel >= 0 keeps only pairs whose neighbour element map[tj] equals elem_sel.
void launch_ui(const SnapDev& s, UiMode mode, double* Ur, double* Ui, const double* rij, const int* pair_off,
               const int* perm, const int* ti, const int* tj, int natoms, double wself, int elem_sel,
               hipStream_t st);
// compute_ui on the half layout with wself (UI_HALF_WSELF) from per-pair geometry written by the MD
// pair list: pgeo[5p..5p+4] = (a_r, a_i, b_r, b_i, sfac * w_j) of pair p, pairs of atom i in
// [pair_off[i], pair_off[i+1]).
void launch_ui_geo(const SnapDev& s, double* Ur, double* Ui, const double* pgeo, const int* pair_off, int natoms,
                   double wself, hipStream_t st);
// pgeo[5q..5q+4] = (a_r, a_i, b_r, b_i, sfac w_j) of every pair q < *np_dev (cap: the launch's pair capacity)
void launch_pair_geo(const SnapDev& s, const double* rij, const int* ti, const int* tj, const int* np_dev, int cap,
                     double* pgeo, hipStream_t st);
// The same result with the per-atom sums formed in LDS, one lane per pair (at most 96 pairs per atom);
// persistent over `groups` workgroups of 512 lanes (0: 512).
void launch_ui_geo_lds(const SnapDev& s, double* Ur, double* Ui, const double* pgeo, const int* pair_off,
                       int natoms, double wself, int groups, hipStream_t st);

// Y[half rows] of every atom (Y zeroed by the caller where untouched).  beta_t[triple] holds the
// effective weight of each coupling triple (multiplicity, (j1+1)/(j+1) and 1/(j+1) factors folded
// in).  full_mid: also the ma > mb entries of the middle row of even layers (upstream output);
// epart != nullptr: per (row, atom) partial of sum' U.Y for the energy identity.
void launch_yi(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
               double* Yi, double* epart, int natoms, bool full_mid, hipStream_t st);

// Y with one weight per (coupling triple, atom): beta_t[triple*N + atom]; full layout, every idxz entry.
// Returns false for a twojmax without a compiled kernel.
// Per-atom weights of Y from per-atom bispectrum weights: beta_t[t*N + a] = fac[t] * g[jjb[t]*N + a].
void launch_beta_peratom(const double* g, const int* jjb, const double* fac, int ntrip, int N, double* beta_t,
                         hipStream_t st);
bool launch_yi_peratom(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_t, double* Yr,
                       double* Yi, int natoms, hipStream_t st);

// MD path: Y (half layout, entries the force uses) + energy partials from a half-layout U.  The half U
// of 64 atoms is staged in LDS; the workgroup's waves take output rows from a shared counter in the
// order row_order (device array of n_half_rows(J) row ids, most expensive first).
void launch_yi_half(const SnapDev& s, const double* Uhr, const double* Uhi, const double* beta_t, double* Yhr,
                    double* Yhi, double* epart, const int* row_order, int natoms, hipStream_t st, int waves = 8);
// waves-per-workgroup variants available for twojmax J (ids = waves); returns their number
int yi_half_variants(int J, int* ids);
// row ids of the half domain sorted by decreasing compute_yi work
void yi_row_order(int J, int* out);
// The learned surrogate's twojmax 2 / 4 descriptor path on the half layout (rows 2mb <= j, [e][N]): Y
// with per-atom weights beta_pa[triple * N + atom] (no energy partials; order = yi_row_order(J)), and
// every B (bispectrum triples j >= j1, blist[jjb * N + atom], no bzero subtraction).  Both stage the
// half U of 64 atoms in LDS.  Return false for other twojmax.
bool launch_yi_half_pa(const SnapDev& s, const double* Ur, const double* Ui, const double* beta_pa, double* Yr,
                       double* Yi, const int* order, int N, hipStream_t st);
bool launch_b_half(const SnapDev& s, const double* Ur, const double* Ui, double* blist, int N, int bnorm,
                   hipStream_t st);

// B of the idxb triples: blist[ii + N*jjb] = 2 sum' U*(j).Z(j1,j2,j)  (before bzero; Z/(j+1) if bnorm).
// U1/U2/U3 element blocks given by their offsets (elements) into Ur/Ui.
void launch_bi(const SnapDev& s, const double* Ur, const double* Ui, long off1, long off2, long off3,
               double* blist, int natoms, int bnorm, hipStream_t st);

// zlist (all idxz entries) for element pair (off1, off2) into z + N*jjz layout, divided by (j+1) if bnorm.
void launch_zi(const SnapDev& s, const double* Ur, const double* Ui, long off1, long off2, double* zr, double* zi,
               int natoms, int bnorm, hipStream_t st);

// dE/dr_ij per pair: 2 sum' d(sfac U_ij)/dr . Y(atom of the pair); pairs p in [0, *npairs) (device count)
// or [0, npairs_host) when npairs_dev == nullptr.  yatom[p] = row of Y used (aii), ybase[p] optional
// element offset (elements) for chemflag.  The full layout always runs the SU(2)-generator kernel; the
// half layout runs the given variant (dedr_variants).
void launch_dedr(const SnapDev& s, const double* Yr, const double* Yi, const double* rij, const int* yatom,
                 const int* ti, const int* tj, const long* ybase, double* dedr, int natoms, int npairs_host,
                 const int* npairs_dev, hipStream_t st, bool half_layout = false, int variant = 0);
// dE/dr with one centre atom per wave (pairs sorted by centre atom, CSR offsets pair_off): Y (half
// layout) is first copied atom-major into yt (2 * n_half_elems(J) * natoms doubles) so that each
// wave reads its atom's Y through the scalar cache.
void launch_dedr_atom(const SnapDev& s, const double* Yr, const double* Yi, double* yt, const double* rij,
                      const int* ti, const int* tj, const int* pair_off, double* dedr, int natoms, hipStream_t st);
// Variants of the MD path's dE/dr (0: all three components per lane; 2, 3, 4: one component per lane
// with at least that many waves per SIMD; 5: launch_dedr_atom; 6, 7: variant 2 with 2 or 4 waves per
// workgroup on consecutive pair tiles; 8, 9: lane = pair, derivatives from U through the SU(2)
// generators, 1 or 4 waves per workgroup; 10, 11: the same with Y rows loaded one row ahead; 12, 13:
// 10, 11 with each XCD walking one contiguous eighth of the pairs).
// Writes at most 16 ids, returns their number.
int dedr_variants(int* ids);

// ---- generic helpers ----------------------------------------------------------------------------

// exclusive prefix sum of n ints: out[0..n] (out[n] = total); tmp needs scan_tmp_ints(n) ints
size_t scan_tmp_ints(int n);
void exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t st);

// Utot += sfac w_j U over each listed atom's pairs, U pair-major, Utot atom-major (cpuSnap.cpp:358-395);
// sfw: scratch of one double per pair (the pairs' sfac w_j, formed once per pair)
void launch_add_utot_rij(double* utr, double* uti, const double* ur, const double* ui, const double* rij,
                         const double* wj, const double* rad, double rmin0, double rcutfac, const int* ilist,
                         const int* type, const int* pairnum, const int* pairnumsum, const int* map, const int* tj,
                         int idxu_max, int ne, int inum, int sw, int chem, double* sfw, hipStream_t st);

// cell-list neighbour search of the upstream core (cpuNeighborList3D.cpp:334-443, cpuNeighborList2D.cpp:298-308)
void launch_cell_list3d(int* clist, const double* x, const double* eta1, const double* eta2, const double* eta3,
                        const double* s2rmap, const int nc[3], int inum, int natom, hipStream_t st);
// c2alist: atoms by cell (stable), c2anumsum[ncell+1]: first entry of each cell, c2anum: sorted cell ids;
// a stable radix sort (cellsort.hip) whose temporary storage comes from temp_alloc(bytes)
void launch_cell2atom_sorted(int* c2alist, int* c2anumsum, int* c2anum, const int* clist, int natom, int ncell,
                             void* (*temp_alloc)(size_t), hipStream_t st);
void launch_full_neigh_cell(int* nl, int* nn, const double* x, double rc2, const int* alist, const int* clist,
                            const int* c2alist, const int* c2anumsum, const int nc[3], int inum, int jnum,
                            hipStream_t st);
void launch_full_neigh_all(int* nl, int* nn, const double* x, double rc2, int anum, int inum, int jnum,
                           hipStream_t st);

}  // namespace mdp

namespace mdp {

// ---- descriptor path, tallies and pair arrays (generic kernels) ------------------------------

// cpuComputeSij / cpuComputeUij: per-pair U and d(sfac U)/dr; dS[k] are the x/y/z derivative arrays
// (each [idxu_max*ijnum] at ij + ijnum*jju).
void launch_uij(const SnapDev& s, double* Sr, double* Si, double* const dSr[3], double* const dSi[3],
                const double* rij, const int* ti, const int* tj, int ijnum, hipStream_t st);

// Stot = wself on the diagonal of the own-element block (element read through type[ai[ii]]).
void launch_zero_utot(const SnapDev& s, double* Sr, double* Si, double wself, const int* type, const int* ai,
                      int wselfall, int chemflag, int nelem, int inum, hipStream_t st);

// Stot[ai + N*jju + N*idxu_max*elem(tj)] += S[ij + ijnum*jju]
void launch_add_utot(const SnapDev& s, double* Stotr, double* Stoti, const double* Sr, const double* Si,
                     const int* ai, const int* tj, int chemflag, int inum, int ijnum, hipStream_t st);

// blist = 2 sum' Stot*.Z from a precomputed zlist, all element triples; then -bzero.
void launch_bi2(const SnapDev& s, double* blist, const double* zr, const double* zi, const double* Sr,
                const double* Si, const double* bzero, const int* ilist, const int* type, int bzero_flag,
                int wselfall, int chemflag, int nelem, int inum, hipStream_t st);

// subtract bzero[j] from blist of the triples upstream does (own-element triple, or all)
void launch_bzero(const SnapDev& s, double* blist, const double* bzero, const int* ilist, const int* type,
                  int wselfall, int chemflag, int nelem, int inum, int use_ilist, hipStream_t st);

// eatom[ilist[ii]] += beta0 + sum beta_k B_k from a blist holding all element triples (bzero applied)
void launch_ei_from_b(const SnapDev& s, double* eatom, const double* blist, const double* coeffelem,
                      const int* ilist, const int* type, int chemflag, int nelem, int ncoeffall, int inum,
                      hipStream_t st);

// dblist of cpuComputeDbidrj (zeroed by the caller)
// dB/dr of the per-pair route (cpuSnap.cpp:1050-1266): Z atom-major, dU and dB pair-major (AoS in k);
// writes every dB entry of the ijnum pairs (no prior clear needed).
void launch_dbidrj_atom_major(const SnapDev& s, double* db, const double* zr, const double* zi, const double* dur,
                              const double* dui, const int* ai, const int* tj, int ne, int bnorm, int chem, int ijnum,
                              hipStream_t st);
void launch_dbidrj(const SnapDev& s, double* dblist, const double* zr, const double* zi, const double* dur,
                   const double* dui, const int* ai, const int* tj, int bnorm, int chemflag, int nelem, int inum,
                   int ijnum, hipStream_t st);

void launch_tally_bi(double* bi, const double* b, const int* ilist, const int* type, int inum, int ncoeff,
                     int nperdim, int quadratic, hipStream_t st);
void launch_tally_bd(double* db, const double* b, const double* dbdr, const int* aii, const int* ai,
                     const int* aj, const int* ti, int inum, int ijnum, int ncoeff, int nperdim, int quadratic,
                     hipStream_t st);
void launch_tally_bv(double* bv, const double* b, const double* dbdr, const double* rij, const int* aii,
                     const int* ti, int inum, int ijnum, int ncoeff, int nperdim, int quadratic, hipStream_t st);

// fatom[3*ai] += dE, fatom[3*aj] -= dE, vatom[6*ai + (xx,yy,zz,xy,xz,yz)] += -r (x) dE  (atomic)
void launch_tally_force_atomic(double* fatom, double* vatom, const double* dedr, const double* rij, const int* ai,
                               const int* aj, int ijnum, hipStream_t st);
// same forces; virial in the SoA layout vatom[k * anum + atom], -r(x)dE/2 to both atoms of the pair
void launch_tally_force_atomic_soa(double* fatom, double* vatom, const double* dedr, const double* rij, const int* ai,
                                   const int* aj, int ijnum, int anum, hipStream_t st);

// cpuNeighPairList: pairnum[ii] = count of neighbours j with d2 < rcutsq(type pair), d2 > 1e-20
void launch_pair_count(int* pairnum, const double* x, const double* rcutsq, const int* ilist, const int* neighlist,
                       const int* neighnum, const int* atomtype, const int* alist, int inum, int dim, int ntypes,
                       hipStream_t st);
void launch_pair_fill(int* pairlist, const int* pairoff, const double* x, const double* rcutsq, const int* ilist,
                      const int* neighlist, const int* neighnum, const int* atomtype, const int* alist, int inum,
                      int dim, int ntypes, hipStream_t st);
// cpuNeighPairs
void launch_pair_arrays(double* xij, const double* x, int* aii, int* ai, int* aj, int* ti, int* tj,
                        const int* pairlist, const int* pairnumsum, const int* ilist, const int* atomtype,
                        const int* alist, int inum, int dim, hipStream_t st);

// the fixed-stride forms of the upstream C++ core: neighbours of atom i at neighlist[l + jnum*i],
// kept pairs of the ii-th atom at pairlist[l + jnum*ii] with pairnum[ii] of them
void launch_pair_list_strided(int* pairnum, int* pairlist, const double* x, const double* rcutsq, const int* ilist,
                              const int* neighlist, const int* neighnum, const int* atomtype, const int* alist,
                              int inum, int jnum, int dim, int ntypes, hipStream_t st);
void launch_pair_list_strided_cut(int* pairnum, int* pairlist, const double* x, double cut2, const int* ilist,
                                  const int* neighlist, const int* neighnum, int inum, int jnum, int dim,
                                  hipStream_t st);
void launch_pair_arrays_strided(double* xij, const double* x, int* aii, int* ai, int* aj, int* ti, int* tj,
                                const int* pairnum, const int* pairlist, const int* pairnumsum, const int* ilist,
                                const int* atomtype, const int* alist, int inum, int jnum, int dim, hipStream_t st);
// Stot += S summed over each atom's pair range [pairnumsum[ii], pairnumsum[ii] + pairnum[ii])
void launch_add_utot_pairs(double* Stotr, double* Stoti, const double* Sr, const double* Si, const int* pairnum,
                           const int* pairnumsum, const int* map, const int* tj, int idxu_max, int inum, int ijnum,
                           int chemflag, hipStream_t st);

// energy / force / virial tallies of the upstream C++ core: eatom[ii] += beta . B (AoS or SoA B),
// fatom += f_ij on ai and -= on aj (atomics), SoA half/half virial of -r (x) f_ij (atomics)
void launch_tally_energy_full(double* eatom, const double* b, const double* coeff, const int* ilist, const int* map,
                              const int* type, int inum, int ncoeff, int ncoeffall, int quadratic, int soa,
                              hipStream_t st);
void launch_tally_force_full(double* fatom, const double* fij, const int* ai, const int* aj, int ijnum,
                             hipStream_t st);
void launch_tally_virial_full(double* vatom, const double* fij, const double* rij, const int* ai, const int* aj,
                              int inum, int ijnum, hipStream_t st);

// AddWself2Ui on device (wself on the diagonal, mirror of rows 2mb < j)
void launch_add_wself(const SnapDev& s, double* Ur, double* Ui, double wself, const int* type, int wselfall,
                      int chemflag, int nelem, int inum, hipStream_t st);

// generic compute_yi accumulate for element blocks: Y[e3] += sum beta_t(e1,e2,e3,ielem) Z(e1,e2) over
// the atoms whose element (map[type[ii]]) equals isel (all atoms if isel < 0); per_atom: beta_t[triple*N + atom]
void launch_yi_acc(const SnapDev& s, const double* Ur, const double* Ui, long off1, long off2, long off3,
                   const double* beta_t, double* Yr, double* Yi, const int* atom_elem, int isel, int N,
                   hipStream_t st, bool per_atom = false);

}  // namespace mdp
