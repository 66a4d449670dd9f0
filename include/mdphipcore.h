/* The hip* functions that the backend-3 branches of the upstream C++ core dispatch
 * (src/C++/Common/mdpcore.h, `#ifdef USE_HIP`) call for the SNAP force path and for the descriptor
 * chain of the fitting driver (snappotential.cpp:762-910), as templates over the
 * upstream dstype.  The double instantiations forward to the part-5 entry points of libmdpsnap
 * (include/mdpsnap.h); other types are not provided.  Include this header where mdpcore.h expects
 * the HIP backend declarations and link libmdpsnap. */
#ifndef MDP_HIPCORE_H
#define MDP_HIPCORE_H

#include "mdpsnap.h"

template <typename T>
void hipSnapComputeUi(T *Stotr, T *Stoti, T *rootpqarray, T *rij, T *wjelem, T *radelem, T rmin0, T rfac0, T rcutfac,
                      int *idxu_block, int *map, int *ai, int *ti, int *tj, int twojmax, int idxu_max, int inum,
                      int ijnum, int switchflag, int chemflag);
template <>
inline void hipSnapComputeUi<double>(double *Stotr, double *Stoti, double *rootpqarray, double *rij, double *wjelem,
                                     double *radelem, double rmin0, double rfac0, double rcutfac, int *idxu_block,
                                     int *map, int *ai, int *ti, int *tj, int twojmax, int idxu_max, int inum,
                                     int ijnum, int switchflag, int chemflag) {
    mdp_hipSnapComputeUi(Stotr, Stoti, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac, idxu_block, map, ai,
                         ti, tj, twojmax, idxu_max, inum, ijnum, switchflag, chemflag);
}

template <typename T>
void hipAddWself2Ui(T *ulisttot_r, T *ulisttot_i, T wself, int *idxu_block, int *type, int *map, int *ai,
                    int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum);
template <>
inline void hipAddWself2Ui<double>(double *ulisttot_r, double *ulisttot_i, double wself, int *idxu_block, int *type,
                                   int *map, int *ai, int wselfall_flag, int chemflag, int idxu_max, int nelements,
                                   int twojmax, int inum) {
    mdp_hipAddWself2Ui(ulisttot_r, ulisttot_i, wself, idxu_block, type, map, ai, wselfall_flag, chemflag, idxu_max,
                       nelements, twojmax, inum);
}

template <typename T>
void hipSnapComputeEi(T *eatom, T *ulisttot_r, T *ulisttot_i, T *cglist, T *bzero, T *coeffelem, int *ilist, int *map,
                      int *type, int *idxb, int *idxcg_block, int *idxu_block, int twojmax, int idxb_max, int idxu_max,
                      int nelements, int ncoeffall, int bnorm_flag, int bzero_flag, int wselfall_flag,
                      int quadraticflag, int inum);
template <>
inline void hipSnapComputeEi<double>(double *eatom, double *ulisttot_r, double *ulisttot_i, double *cglist,
                                     double *bzero, double *coeffelem, int *ilist, int *map, int *type, int *idxb,
                                     int *idxcg_block, int *idxu_block, int twojmax, int idxb_max, int idxu_max,
                                     int nelements, int ncoeffall, int bnorm_flag, int bzero_flag, int wselfall_flag,
                                     int quadraticflag, int inum) {
    mdp_hipSnapComputeEi(eatom, ulisttot_r, ulisttot_i, cglist, bzero, coeffelem, ilist, map, type, idxb, idxcg_block,
                         idxu_block, twojmax, idxb_max, idxu_max, nelements, ncoeffall, bnorm_flag, bzero_flag,
                         wselfall_flag, quadraticflag, inum);
}

template <typename T>
void hipSnapComputeYi(T *ylist_r, T *ylist_i, T *ulisttot_r, T *ulisttot_i, T *cglist, T *betalist, int *map,
                      int *type, int *idxz, int *idxb_block, int *idxu_block, int *idxcg_block, int twojmax,
                      int idxb_max, int idxu_max, int idxz_max, int nelements, int ncoeffall, int bnorm_flag, int inum);
template <>
inline void hipSnapComputeYi<double>(double *ylist_r, double *ylist_i, double *ulisttot_r, double *ulisttot_i,
                                     double *cglist, double *betalist, int *map, int *type, int *idxz, int *idxb_block,
                                     int *idxu_block, int *idxcg_block, int twojmax, int idxb_max, int idxu_max,
                                     int idxz_max, int nelements, int ncoeffall, int bnorm_flag, int inum) {
    mdp_hipSnapComputeYi(ylist_r, ylist_i, ulisttot_r, ulisttot_i, cglist, betalist, map, type, idxz, idxb_block,
                         idxu_block, idxcg_block, twojmax, idxb_max, idxu_max, idxz_max, nelements, ncoeffall,
                         bnorm_flag, inum);
}

template <typename T>
void hipSnapComputeYi(T *ylist_r, T *ylist_i, T *ulisttot_r, T *ulisttot_i, T *cglist, T *betalist, int *idxz,
                      int *idxb_block, int *idxu_block, int *idxcg_block, int twojmax, int idxb_max, int idxu_max,
                      int idxz_max, int nelements, int bnorm_flag, int inum);
template <>
inline void hipSnapComputeYi<double>(double *ylist_r, double *ylist_i, double *ulisttot_r, double *ulisttot_i,
                                     double *cglist, double *betalist, int *idxz, int *idxb_block, int *idxu_block,
                                     int *idxcg_block, int twojmax, int idxb_max, int idxu_max, int idxz_max,
                                     int nelements, int bnorm_flag, int inum) {
    mdp_hipSnapComputeYiPerAtom(ylist_r, ylist_i, ulisttot_r, ulisttot_i, cglist, betalist, idxz, idxb_block,
                                idxu_block, idxcg_block, twojmax, idxb_max, idxu_max, idxz_max, nelements, bnorm_flag,
                                inum);
}

template <typename T>
void hipSnapComputeFi(T *fatom, T *vatom, T *ylist_r, T *ylist_i, T *rootpqarray, T *rij, T *wjelem, T *radelem,
                      T rmin0, T rfac0, T rcutfac, int *idxu_block, int *map, int *aii, int *ai, int *aj, int *ti,
                      int *tj, int twojmax, int idxu_max, int inum, int anum, int ijnum, int switchflag, int chemflag);
template <>
inline void hipSnapComputeFi<double>(double *fatom, double *vatom, double *ylist_r, double *ylist_i,
                                     double *rootpqarray, double *rij, double *wjelem, double *radelem, double rmin0,
                                     double rfac0, double rcutfac, int *idxu_block, int *map, int *aii, int *ai,
                                     int *aj, int *ti, int *tj, int twojmax, int idxu_max, int inum, int anum,
                                     int ijnum, int switchflag, int chemflag) {
    mdp_hipSnapComputeFi(fatom, vatom, ylist_r, ylist_i, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac,
                         idxu_block, map, aii, ai, aj, ti, tj, twojmax, idxu_max, inum, anum, ijnum, switchflag,
                         chemflag);
}

template <typename T>
void hipNeighPairList(int *pairnum, int *pairlist, T *x, T *rcutsq, int *ilist, int *neighlist, int *neighnum,
                      int *atomtype, int *alist, int inum, int jnum, int dim, int ntypes);
template <>
inline void hipNeighPairList<double>(int *pairnum, int *pairlist, double *x, double *rcutsq, int *ilist, int *neighlist,
                                     int *neighnum, int *atomtype, int *alist, int inum, int jnum, int dim,
                                     int ntypes) {
    mdp_hipNeighPairList(pairnum, pairlist, x, rcutsq, ilist, neighlist, neighnum, atomtype, alist, inum, jnum, dim,
                         ntypes);
}

template <typename T>
void hipNeighPairList(int *pairnum, int *pairlist, T *x, T rcutsq, int *ilist, int *neighlist, int *neighnum, int inum,
                      int jnum, int dim);
template <>
inline void hipNeighPairList<double>(int *pairnum, int *pairlist, double *x, double rcutsq, int *ilist, int *neighlist,
                                     int *neighnum, int inum, int jnum, int dim) {
    mdp_hipNeighPairListCut(pairnum, pairlist, x, rcutsq, ilist, neighlist, neighnum, inum, jnum, dim);
}

inline void hipCumsum(int *d_out, int *d_in, int *d_sums, int *d_incr, int length) {
    mdp_hipCumsum(d_out, d_in, d_sums, d_incr, length);
}

template <typename T>
void hipNeighPairs(T *xij, T *x, int *aii, int *ai, int *aj, int *ti, int *tj, int *pairnum, int *pairlist,
                   int *pairnumsum, int *ilist, int *atomtype, int *alist, int inum, int jnum, int dim);
template <>
inline void hipNeighPairs<double>(double *xij, double *x, int *aii, int *ai, int *aj, int *ti, int *tj, int *pairnum,
                                  int *pairlist, int *pairnumsum, int *ilist, int *atomtype, int *alist, int inum,
                                  int jnum, int dim) {
    mdp_hipNeighPairs(xij, x, aii, ai, aj, ti, tj, pairnum, pairlist, pairnumsum, ilist, atomtype, alist, inum, jnum,
                      dim);
}

template <typename T>
void hipComputeSij(T *Sr, T *Si, T *Srx, T *Six, T *Sry, T *Siy, T *Srz, T *Siz, T *rootpqarray, T *rij, T *wjelem,
                   T *radelem, T rmin0, T rfac0, T rcutfac, int *idxu_block, int *ti, int *tj, int twojmax,
                   int idxu_max, int ijnum, int switchflag);
template <>
inline void hipComputeSij<double>(double *Sr, double *Si, double *Srx, double *Six, double *Sry, double *Siy,
                                  double *Srz, double *Siz, double *rootpqarray, double *rij, double *wjelem,
                                  double *radelem, double rmin0, double rfac0, double rcutfac, int *idxu_block,
                                  int *ti, int *tj, int twojmax, int idxu_max, int ijnum, int switchflag) {
    mdp_hipComputeSij(Sr, Si, Srx, Six, Sry, Siy, Srz, Siz, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac,
                      idxu_block, ti, tj, twojmax, idxu_max, ijnum, switchflag);
}

template <typename T>
void hipZeroUarraytot2(T *ulisttot_r, T *ulisttot_i, T wself, int *idxu_block, int *type, int *map, int *ai,
                       int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum);
template <>
inline void hipZeroUarraytot2<double>(double *ulisttot_r, double *ulisttot_i, double wself, int *idxu_block,
                                      int *type, int *map, int *ai, int wselfall_flag, int chemflag, int idxu_max,
                                      int nelements, int twojmax, int inum) {
    mdp_hipZeroUarraytot2(ulisttot_r, ulisttot_i, wself, idxu_block, type, map, ai, wselfall_flag, chemflag, idxu_max,
                          nelements, twojmax, inum);
}

template <typename T>
void hipAddUarraytot(T *ulisttot_r, T *ulisttot_i, T *ulist_r, T *ulist_i, int *map, int *ai, int *tj, int idxu_max,
                     int inum, int ijnum, int chemflag);
template <>
inline void hipAddUarraytot<double>(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i, int *map,
                                    int *ai, int *tj, int idxu_max, int inum, int ijnum, int chemflag) {
    mdp_hipAddUarraytot(ulisttot_r, ulisttot_i, ulist_r, ulist_i, map, ai, tj, idxu_max, inum, ijnum, chemflag);
}

template <typename T>
void hipAddUarraytot(T *ulisttot_r, T *ulisttot_i, T *ulist_r, T *ulist_i, int *pairnum, int *pairnumsum, int *map,
                     int *tj, int idxu_max, int nelements, int inum, int ijnum, int chemflag);
template <>
inline void hipAddUarraytot<double>(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i,
                                    int *pairnum, int *pairnumsum, int *map, int *tj, int idxu_max, int nelements,
                                    int inum, int ijnum, int chemflag) {
    mdp_hipAddUarraytotPairs(ulisttot_r, ulisttot_i, ulist_r, ulist_i, pairnum, pairnumsum, map, tj, idxu_max,
                             nelements, inum, ijnum, chemflag);
}

template <typename T>
void hipComputeZi2(T *zlist_r, T *zlist_i, T *ulisttot_r, T *ulisttot_i, T *cglist, int *idxz, int *idxu_block,
                   int *idxcg_block, int twojmax, int idxu_max, int idxz_max, int nelements, int bnorm_flag, int inum);
template <>
inline void hipComputeZi2<double>(double *zlist_r, double *zlist_i, double *ulisttot_r, double *ulisttot_i,
                                  double *cglist, int *idxz, int *idxu_block, int *idxcg_block, int twojmax,
                                  int idxu_max, int idxz_max, int nelements, int bnorm_flag, int inum) {
    mdp_hipComputeZi2(zlist_r, zlist_i, ulisttot_r, ulisttot_i, cglist, idxz, idxu_block, idxcg_block, twojmax,
                      idxu_max, idxz_max, nelements, bnorm_flag, inum);
}

template <typename T>
void hipComputeBi2(T *blist, T *zlist_r, T *zlist_i, T *ulisttot_r, T *ulisttot_i, T *bzero, int *ilist, int *type,
                   int *map, int *idxb, int *idxu_block, int *idxz_block, int twojmax, int idxb_max, int idxu_max,
                   int idxz_max, int nelements, int bzero_flag, int wselfall_flag, int chemflag, int inum);
template <>
inline void hipComputeBi2<double>(double *blist, double *zlist_r, double *zlist_i, double *ulisttot_r,
                                  double *ulisttot_i, double *bzero, int *ilist, int *type, int *map, int *idxb,
                                  int *idxu_block, int *idxz_block, int twojmax, int idxb_max, int idxu_max,
                                  int idxz_max, int nelements, int bzero_flag, int wselfall_flag, int chemflag,
                                  int inum) {
    mdp_hipComputeBi2(blist, zlist_r, zlist_i, ulisttot_r, ulisttot_i, bzero, ilist, type, map, idxb, idxu_block,
                      idxz_block, twojmax, idxb_max, idxu_max, idxz_max, nelements, bzero_flag, wselfall_flag,
                      chemflag, inum);
}

template <typename T>
void hipComputeDbidrj(T *dblist, T *zlist_r, T *zlist_i, T *dulist_r, T *dulist_i, int *idxb, int *idxu_block,
                      int *idxz_block, int *map, int *ai, int *tj, int twojmax, int idxb_max, int idxu_max,
                      int idxz_max, int nelements, int bnorm_flag, int chemflag, int inum, int ijnum);
template <>
inline void hipComputeDbidrj<double>(double *dblist, double *zlist_r, double *zlist_i, double *dulist_r,
                                     double *dulist_i, int *idxb, int *idxu_block, int *idxz_block, int *map, int *ai,
                                     int *tj, int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements,
                                     int bnorm_flag, int chemflag, int inum, int ijnum) {
    mdp_hipComputeDbidrj(dblist, zlist_r, zlist_i, dulist_r, dulist_i, idxb, idxu_block, idxz_block, map, ai, tj,
                         twojmax, idxb_max, idxu_max, idxz_max, nelements, bnorm_flag, chemflag, inum, ijnum);
}

template <typename T>
void hipComputeDbidrj(T *dblist, T *zlist_r, T *zlist_i, T *dulist_r, T *dulist_i, int *idxb, int *idxu_block,
                      int *idxz_block, int *map, int *ai, int *aj, int *ti, int *tj, int twojmax, int idxb_max,
                      int idxu_max, int idxz_max, int nelements, int bnorm_flag, int chemflag, int ijnum);
template <>
inline void hipComputeDbidrj<double>(double *dblist, double *zlist_r, double *zlist_i, double *dulist_r,
                                     double *dulist_i, int *idxb, int *idxu_block, int *idxz_block, int *map, int *ai,
                                     int *aj, int *ti, int *tj, int twojmax, int idxb_max, int idxu_max, int idxz_max,
                                     int nelements, int bnorm_flag, int chemflag, int ijnum) {
    mdp_hipComputeDbidrjAtomMajor(dblist, zlist_r, zlist_i, dulist_r, dulist_i, idxb, idxu_block, idxz_block, map, ai,
                                  aj, ti, tj, twojmax, idxb_max, idxu_max, idxz_max, nelements, bnorm_flag, chemflag,
                                  ijnum);
}

template <typename T>
void hipAddUarraytot(T *ulisttot_r, T *ulisttot_i, T *ulist_r, T *ulist_i, T *rij, T *wjelem, T *radelem, T rmin0,
                     T rcutfac, int *idxu_block, int *ilist, int *type, int *pairnum, int *pairnumsum, int *map,
                     int *tj, int twojmax, int idxu_max, int nelements, int inum, int switch_flag, int chemflag);
template <>
inline void hipAddUarraytot<double>(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i,
                                    double *rij, double *wjelem, double *radelem, double rmin0, double rcutfac,
                                    int *idxu_block, int *ilist, int *type, int *pairnum, int *pairnumsum, int *map,
                                    int *tj, int twojmax, int idxu_max, int nelements, int inum, int switch_flag,
                                    int chemflag) {
    mdp_hipAddUarraytotRij(ulisttot_r, ulisttot_i, ulist_r, ulist_i, rij, wjelem, radelem, rmin0, rcutfac, idxu_block,
                           ilist, type, pairnum, pairnumsum, map, tj, twojmax, idxu_max, nelements, inum, switch_flag,
                           chemflag);
}

template <typename T>
void hipCellList3D(int *clist, T *x, T *eta1, T *eta2, T *eta3, T *s2rmap, int *nc, int inum, int natom, int dim);
template <>
inline void hipCellList3D<double>(int *clist, double *x, double *eta1, double *eta2, double *eta3, double *s2rmap,
                                  int *nc, int inum, int natom, int dim) {
    mdp_hipCellList3D(clist, x, eta1, eta2, eta3, s2rmap, nc, inum, natom, dim);
}

inline void hipCell2AtomList(int *c2alist, int *c2anumsum, int *c2anum, int *clist, int *dtemp, int natom, int ncell) {
    mdp_hipCell2AtomList(c2alist, c2anumsum, c2anum, clist, dtemp, natom, ncell);
}

template <typename T>
void hipFullNeighborList3D(int *neighlist, int *neighnum, T *x, T *rcutsq, int anum, int inum, int jnum, int dim);
template <>
inline void hipFullNeighborList3D<double>(int *neighlist, int *neighnum, double *x, double *rcutsq, int anum, int inum,
                                          int jnum, int dim) {
    mdp_hipFullNeighborList3D(neighlist, neighnum, x, rcutsq, anum, inum, jnum, dim);
}

/* the cell-list form; the dispatch's backend-3 branch calls it hipFullNeighborList23 */
template <typename T>
void hipFullNeighborList3D(int *neighlist, int *neighnum, T *x, T *rcutsq, int *alist, int *clist, int *c2alist,
                           int *c2anumsum, int *nc, int inum, int jnum, int dim);
template <>
inline void hipFullNeighborList3D<double>(int *neighlist, int *neighnum, double *x, double *rcutsq, int *alist,
                                          int *clist, int *c2alist, int *c2anumsum, int *nc, int inum, int jnum,
                                          int dim) {
    mdp_hipFullNeighborList3DCell(neighlist, neighnum, x, rcutsq, alist, clist, c2alist, c2anumsum, nc, inum, jnum,
                                  dim);
}
template <typename T>
inline void hipFullNeighborList23(int *neighlist, int *neighnum, T *x, T *rcutsq, int *alist, int *clist,
                                  int *c2alist, int *c2anumsum, int *nc, int inum, int jnum, int dim) {
    hipFullNeighborList3D<T>(neighlist, neighnum, x, rcutsq, alist, clist, c2alist, c2anumsum, nc, inum, jnum, dim);
}

template <typename T>
void hipSnapTallyBispectrum(T *bi, T *bispectrum, int *ilist, int *type, int inum, int ncoeff, int nperdim, int ntype,
                            int quadraticflag);
template <>
inline void hipSnapTallyBispectrum<double>(double *bi, double *bispectrum, int *ilist, int *type, int inum, int ncoeff,
                                           int nperdim, int ntype, int quadraticflag) {
    mdp_hipSnapTallyBispectrum(bi, bispectrum, ilist, type, inum, ncoeff, nperdim, ntype, quadraticflag);
}

template <typename T>
void hipSnapTallyBispectrumDeriv(T *db, T *bispectrum, T *dbdr, int *aii, int *ai, int *aj, int *ti, int inum,
                                 int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag);
template <>
inline void hipSnapTallyBispectrumDeriv<double>(double *db, double *bispectrum, double *dbdr, int *aii, int *ai,
                                                int *aj, int *ti, int inum, int ijnum, int ncoeff, int nperdim,
                                                int ntype, int quadraticflag) {
    mdp_hipSnapTallyBispectrumDeriv(db, bispectrum, dbdr, aii, ai, aj, ti, inum, ijnum, ncoeff, nperdim, ntype,
                                    quadraticflag);
}

template <typename T>
void hipSnapTallyBispectrumVirial(T *bv, T *bispectrum, T *dbdr, T *rij, int *aii, int *ai, int *aj, int *ti,
                                  int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag);
template <>
inline void hipSnapTallyBispectrumVirial<double>(double *bv, double *bispectrum, double *dbdr, double *rij, int *aii,
                                                 int *ai, int *aj, int *ti, int inum, int ijnum, int ncoeff,
                                                 int nperdim, int ntype, int quadraticflag) {
    mdp_hipSnapTallyBispectrumVirial(bv, bispectrum, dbdr, rij, aii, ai, aj, ti, inum, ijnum, ncoeff, nperdim, ntype,
                                     quadraticflag);
}

template <typename T>
void hipSnapTallyEnergyFull(T *eatom, T *bispectrum, T *coeffelem, int *ilist, int *map, int *type, int inum,
                            int ncoeff, int ncoeffall, int quadraticflag);
template <>
inline void hipSnapTallyEnergyFull<double>(double *eatom, double *bispectrum, double *coeffelem, int *ilist, int *map,
                                           int *type, int inum, int ncoeff, int ncoeffall, int quadraticflag) {
    mdp_hipSnapTallyEnergyFull(eatom, bispectrum, coeffelem, ilist, map, type, inum, ncoeff, ncoeffall, quadraticflag);
}

template <typename T>
void hipSnapTallyEnergyFull2(T *eatom, T *bispectrum, T *coeffelem, int *ilist, int *map, int *type, int inum,
                             int ncoeff, int ncoeffall, int quadraticflag);
template <>
inline void hipSnapTallyEnergyFull2<double>(double *eatom, double *bispectrum, double *coeffelem, int *ilist, int *map,
                                            int *type, int inum, int ncoeff, int ncoeffall, int quadraticflag) {
    mdp_hipSnapTallyEnergyFull2(eatom, bispectrum, coeffelem, ilist, map, type, inum, ncoeff, ncoeffall,
                                quadraticflag);
}

template <typename T>
void hipSnapTallyForceFull(T *fatom, T *fij, int *ai, int *aj, int *alist, int ijnum);
template <>
inline void hipSnapTallyForceFull<double>(double *fatom, double *fij, int *ai, int *aj, int *alist, int ijnum) {
    mdp_hipSnapTallyForceFull(fatom, fij, ai, aj, alist, ijnum);
}

template <typename T>
void hipSnapTallyVirialFull(T *vatom, T *fij, T *rij, int *ai, int *aj, int inum, int ijnum);
template <>
inline void hipSnapTallyVirialFull<double>(double *vatom, double *fij, double *rij, int *ai, int *aj, int inum,
                                           int ijnum) {
    mdp_hipSnapTallyVirialFull(vatom, fij, rij, ai, aj, inum, ijnum);
}

/* reference pair potential (part 6): the HIP counterparts of the generated pair gradient and the
 * full-list pair tallies of the upstream empirical-potential route */
template <typename T>
void hipPairaGradient(T *u, T *u_xij, T *xij, T *qi, T *qj, int *ti, int *tj, int *ai, int *aj, T *mu, T *eta,
                      int *kappa, int dim, int ncq, int nmu, int neta, int nkappa, int ng, int potnum);
template <>
inline void hipPairaGradient<double>(double *u, double *u_xij, double *xij, double *qi, double *qj, int *ti, int *tj,
                                     int *ai, int *aj, double *mu, double *eta, int *kappa, int dim, int ncq, int nmu,
                                     int neta, int nkappa, int ng, int potnum) {
    mdp_hipPairaGradient(u, u_xij, xij, qi, qj, ti, tj, ai, aj, mu, eta, kappa, dim, ncq, nmu, neta, nkappa, ng,
                         potnum);
}

template <typename T>
void hipFullNeighPairDecomposition(T *e, T *f, T *eij, T *fij, int *ai, int ijnum, int dim);
template <>
inline void hipFullNeighPairDecomposition<double>(double *e, double *f, double *eij, double *fij, int *ai, int ijnum,
                                                  int dim) {
    mdp_hipFullNeighPairDecomposition(e, f, eij, fij, ai, ijnum, dim);
}

template <typename T>
void hipVirialPairTally(T *v, T *fij, T *rij, T factor, int *ai, int dim, int inum, int ijnum);
template <>
inline void hipVirialPairTally<double>(double *v, double *fij, double *rij, double factor, int *ai, int dim, int inum,
                                       int ijnum) {
    mdp_hipVirialPairTally(v, fij, rij, factor, ai, dim, inum, ijnum);
}

#endif
