// Checker only (tests): the upstream C++ core -- src/C++/Core/cpuCore.cpp, which gathers the CPU
// kernels of the upstream C++ route into one translation unit -- compiled unmodified from its sources
// where they lie, with C entry points for the double instantiations of the SNAP force path that the
// backend-3 entry points of libmdpsnap (include/mdpsnap.h, part 5) replace.  Built by the Makefile into
// _ref/ only when the upstream tree is present; never part of the product path.
#include "cpuCore.cpp"

extern "C" {

void r2_cpuSnapComputeUi(double *Stotr, double *Stoti, double *rootpqarray, double *rij, double *wjelem,
                         double *radelem, double rmin0, double rfac0, double rcutfac, int *idxu_block, int *map,
                         int *ai, int *ti, int *tj, int twojmax, int idxu_max, int inum, int ijnum, int switchflag,
                         int chemflag) {
    cpuSnapComputeUi(Stotr, Stoti, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac, idxu_block, map, ai, ti,
                     tj, twojmax, idxu_max, inum, ijnum, switchflag, chemflag);
}

void r2_cpuAddWself2Ui(double *ulisttot_r, double *ulisttot_i, double wself, int *idxu_block, int *type, int *map,
                       int *ai, int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum) {
    cpuAddWself2Ui(ulisttot_r, ulisttot_i, wself, idxu_block, type, map, ai, wselfall_flag, chemflag, idxu_max,
                   nelements, twojmax, inum);
}

void r2_cpuSnapComputeEi(double *eatom, double *ulisttot_r, double *ulisttot_i, double *cglist, double *bzero,
                         double *coeffelem, int *ilist, int *map, int *type, int *idxb, int *idxcg_block,
                         int *idxu_block, int twojmax, int idxb_max, int idxu_max, int nelements, int ncoeffall,
                         int bnorm_flag, int bzero_flag, int wselfall_flag, int quadraticflag, int inum) {
    cpuSnapComputeEi(eatom, ulisttot_r, ulisttot_i, cglist, bzero, coeffelem, ilist, map, type, idxb, idxcg_block,
                     idxu_block, twojmax, idxb_max, idxu_max, nelements, ncoeffall, bnorm_flag, bzero_flag,
                     wselfall_flag, quadraticflag, inum);
}

void r2_cpuSnapComputeYi(double *ylist_r, double *ylist_i, double *ulisttot_r, double *ulisttot_i, double *cglist,
                         double *betalist, int *map, int *type, int *idxz, int *idxb_block, int *idxu_block,
                         int *idxcg_block, int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements,
                         int ncoeffall, int bnorm_flag, int inum) {
    cpuSnapComputeYi(ylist_r, ylist_i, ulisttot_r, ulisttot_i, cglist, betalist, map, type, idxz, idxb_block,
                     idxu_block, idxcg_block, twojmax, idxb_max, idxu_max, idxz_max, nelements, ncoeffall, bnorm_flag,
                     inum);
}

void r2_cpuSnapComputeFi(double *fatom, double *vatom, double *ylist_r, double *ylist_i, double *rootpqarray,
                         double *rij, double *wjelem, double *radelem, double rmin0, double rfac0, double rcutfac,
                         int *idxu_block, int *map, int *aii, int *ai, int *aj, int *ti, int *tj, int twojmax,
                         int idxu_max, int inum, int anum, int ijnum, int switchflag, int chemflag) {
    cpuSnapComputeFi(fatom, vatom, ylist_r, ylist_i, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac,
                     idxu_block, map, aii, ai, aj, ti, tj, twojmax, idxu_max, inum, anum, ijnum, switchflag,
                     chemflag);
}

}  // extern "C"
