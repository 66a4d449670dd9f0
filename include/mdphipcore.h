/* The hip* functions that the backend-3 branches of the upstream C++ core dispatch
 * (src/C++/Common/mdpcore.h, `#ifdef USE_HIP`) call for the SNAP force path, as templates over the
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

#endif
