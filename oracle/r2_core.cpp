// Checker only (tests): the upstream C++ core -- src/C++/Core/cpuCore.cpp, which gathers the CPU
// kernels of the upstream C++ route into one translation unit -- compiled unmodified from its sources
// where they lie, with C entry points for the double instantiations of the SNAP force path and of the
// fitting driver's descriptor chain that the backend-3 entry points of libmdpsnap (include/mdpsnap.h,
// part 5) replace, and of the Nose-Hoover NVT integrator that the MD engine's thermostat follows.  Built by the Makefile into
// _ref/ only when the upstream tree is present; never part of the product path.
#include <vector>

#include "cpuCore.cpp"
// the generated ZBL pair gradient of the Ta application (src/C++/Potentials, generated from
// simulating/snap/ZBLpotential.jl with cut_inner 4, cut_outer 4.8, z 73)
#include "cpuPairaGradient.cpp"

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

void r2_cpuSnapComputeYiPerAtom(double *ylist_r, double *ylist_i, double *ulisttot_r, double *ulisttot_i,
                                double *cglist, double *betalist, int *idxz, int *idxb_block, int *idxu_block,
                                int *idxcg_block, int twojmax, int idxb_max, int idxu_max, int idxz_max,
                                int nelements, int bnorm_flag, int inum) {
    cpuSnapComputeYi(ylist_r, ylist_i, ulisttot_r, ulisttot_i, cglist, betalist, idxz, idxb_block, idxu_block,
                     idxcg_block, twojmax, idxb_max, idxu_max, idxz_max, nelements, bnorm_flag, inum);
}

void r2_cpuSnapComputeFi(double *fatom, double *vatom, double *ylist_r, double *ylist_i, double *rootpqarray,
                         double *rij, double *wjelem, double *radelem, double rmin0, double rfac0, double rcutfac,
                         int *idxu_block, int *map, int *aii, int *ai, int *aj, int *ti, int *tj, int twojmax,
                         int idxu_max, int inum, int anum, int ijnum, int switchflag, int chemflag) {
    cpuSnapComputeFi(fatom, vatom, ylist_r, ylist_i, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac,
                     idxu_block, map, aii, ai, aj, ti, tj, twojmax, idxu_max, inum, anum, ijnum, switchflag,
                     chemflag);
}

// descriptor chain of ComputeSnap (src/C++/Descriptors/snappotential.cpp:762-910)
void r2_cpuNeighPairList(int *pairnum, int *pairlist, double *x, double *rcutsq, int *ilist, int *neighlist,
                         int *neighnum, int *atomtype, int *alist, int inum, int jnum, int dim, int ntypes) {
    cpuNeighPairList(pairnum, pairlist, x, rcutsq, ilist, neighlist, neighnum, atomtype, alist, inum, jnum, dim,
                     ntypes);
}

void r2_cpuNeighPairListCut(int *pairnum, int *pairlist, double *x, double rcutsq, int *ilist, int *neighlist,
                            int *neighnum, int inum, int jnum, int dim) {
    cpuNeighPairList(pairnum, pairlist, x, rcutsq, ilist, neighlist, neighnum, inum, jnum, dim);
}

void r2_cpuCumsum(int *d_out, int *d_in, int *d_sums, int *d_incr, int length) {
    (void)d_sums; (void)d_incr;
    cpuCumsum(d_out, d_in, length);
}

void r2_cpuNeighPairs(double *xij, double *x, int *aii, int *ai, int *aj, int *ti, int *tj, int *pairnum,
                      int *pairlist, int *pairnumsum, int *ilist, int *atomtype, int *alist, int inum, int jnum,
                      int dim) {
    cpuNeighPairs(xij, x, aii, ai, aj, ti, tj, pairnum, pairlist, pairnumsum, ilist, atomtype, alist, inum, jnum, dim);
}

void r2_cpuComputeSij(double *Sr, double *Si, double *Srx, double *Six, double *Sry, double *Siy, double *Srz,
                      double *Siz, double *rootpqarray, double *rij, double *wjelem, double *radelem, double rmin0,
                      double rfac0, double rcutfac, int *idxu_block, int *ti, int *tj, int twojmax, int idxu_max,
                      int ijnum, int switchflag) {
    cpuComputeSij(Sr, Si, Srx, Six, Sry, Siy, Srz, Siz, rootpqarray, rij, wjelem, radelem, rmin0, rfac0, rcutfac,
                  idxu_block, ti, tj, twojmax, idxu_max, ijnum, switchflag);
}

void r2_cpuZeroUarraytot2(double *ulisttot_r, double *ulisttot_i, double wself, int *idxu_block, int *type, int *map,
                          int *ai, int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax,
                          int inum) {
    cpuZeroUarraytot2(ulisttot_r, ulisttot_i, wself, idxu_block, type, map, ai, wselfall_flag, chemflag, idxu_max,
                      nelements, twojmax, inum);
}

void r2_cpuAddUarraytot(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i, int *map, int *ai,
                        int *tj, int idxu_max, int inum, int ijnum, int chemflag) {
    cpuAddUarraytot(ulisttot_r, ulisttot_i, ulist_r, ulist_i, map, ai, tj, idxu_max, inum, ijnum, chemflag);
}

void r2_cpuAddUarraytotPairs(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i, int *pairnum,
                             int *pairnumsum, int *map, int *tj, int idxu_max, int nelements, int inum, int ijnum,
                             int chemflag) {
    cpuAddUarraytot(ulisttot_r, ulisttot_i, ulist_r, ulist_i, pairnum, pairnumsum, map, tj, idxu_max, nelements, inum,
                    ijnum, chemflag);
}

void r2_cpuComputeZi2(double *zlist_r, double *zlist_i, double *ulisttot_r, double *ulisttot_i, double *cglist,
                      int *idxz, int *idxu_block, int *idxcg_block, int twojmax, int idxu_max, int idxz_max,
                      int nelements, int bnorm_flag, int inum) {
    cpuComputeZi2(zlist_r, zlist_i, ulisttot_r, ulisttot_i, cglist, idxz, idxu_block, idxcg_block, twojmax, idxu_max,
                  idxz_max, nelements, bnorm_flag, inum);
}

void r2_cpuComputeBi2(double *blist, double *zlist_r, double *zlist_i, double *ulisttot_r, double *ulisttot_i,
                      double *bzero, int *ilist, int *type, int *map, int *idxb, int *idxu_block, int *idxz_block,
                      int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements, int bzero_flag,
                      int wselfall_flag, int chemflag, int inum) {
    cpuComputeBi2(blist, zlist_r, zlist_i, ulisttot_r, ulisttot_i, bzero, ilist, type, map, idxb, idxu_block,
                  idxz_block, twojmax, idxb_max, idxu_max, idxz_max, nelements, bzero_flag, wselfall_flag, chemflag,
                  inum);
}

void r2_cpuComputeDbidrj(double *dblist, double *zlist_r, double *zlist_i, double *dulist_r, double *dulist_i,
                         int *idxb, int *idxu_block, int *idxz_block, int *map, int *ai, int *tj, int twojmax,
                         int idxb_max, int idxu_max, int idxz_max, int nelements, int bnorm_flag, int chemflag,
                         int inum, int ijnum) {
    cpuComputeDbidrj(dblist, zlist_r, zlist_i, dulist_r, dulist_i, idxb, idxu_block, idxz_block, map, ai, tj, twojmax,
                     idxb_max, idxu_max, idxz_max, nelements, bnorm_flag, chemflag, inum, ijnum);
}

// the per-pair route's form (cpuSnap.cpp:1050-1266)
void r2_cpuComputeDbidrjAtomMajor(double *dblist, double *zlist_r, double *zlist_i, double *dulist_r,
                                  double *dulist_i, int *idxb, int *idxu_block, int *idxz_block, int *map, int *ai,
                                  int *aj, int *ti, int *tj, int twojmax, int idxb_max, int idxu_max, int idxz_max,
                                  int nelements, int bnorm_flag, int chemflag, int ijnum) {
    cpuComputeDbidrj(dblist, zlist_r, zlist_i, dulist_r, dulist_i, idxb, idxu_block, idxz_block, map, ai, aj, ti, tj,
                     twojmax, idxb_max, idxu_max, idxz_max, nelements, bnorm_flag, chemflag, ijnum);
}

// the per-pair route's rij form (cpuSnap.cpp:358-395)
void r2_cpuAddUarraytotRij(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i, double *rij,
                           double *wjelem, double *radelem, double rmin0, double rcutfac, int *idxu_block, int *ilist,
                           int *type, int *pairnum, int *pairnumsum, int *map, int *tj, int twojmax, int idxu_max,
                           int nelements, int inum, int switch_flag, int chemflag) {
    cpuAddUarraytot(ulisttot_r, ulisttot_i, ulist_r, ulist_i, rij, wjelem, radelem, rmin0, rcutfac, idxu_block, ilist,
                    type, pairnum, pairnumsum, map, tj, twojmax, idxu_max, nelements, inum, switch_flag, chemflag);
}

// cell-list neighbour search (cpuNeighborList3D.cpp:334-441, cpuNeighborList2D.cpp:298-308)
void r2_cpuCellList3D(int *clist, double *x, double *eta1, double *eta2, double *eta3, double *s2rmap, int *nc,
                      int inum, int natom, int dim) {
    cpuCellList3D(clist, x, eta1, eta2, eta3, s2rmap, nc, inum, natom, dim);
}
void r2_cpuCell2AtomList(int *c2alist, int *c2anumsum, int *c2anum, int *clist, int natom, int ncell) {
    cpuCell2AtomList(c2alist, c2anumsum, c2anum, clist, natom, ncell);
}
void r2_cpuFullNeighborList3DCell(int *neighlist, int *neighnum, double *x, double *rcutsq, int *alist, int *clist,
                                  int *c2alist, int *c2anumsum, int *nc, int inum, int jnum, int dim) {
    cpuFullNeighborList3D(neighlist, neighnum, x, rcutsq, alist, clist, c2alist, c2anumsum, nc, inum, jnum, dim);
}
void r2_cpuFullNeighborList3D(int *neighlist, int *neighnum, double *x, double *rcutsq, int anum, int inum, int jnum,
                              int dim) {
    cpuFullNeighborList3D(neighlist, neighnum, x, rcutsq, anum, inum, jnum, dim);
}

void r2_cpuSnapTallyBispectrum(double *bi, double *bispectrum, int *ilist, int *type, int inum, int ncoeff,
                               int nperdim, int ntype, int quadraticflag) {
    cpuSnapTallyBispectrum(bi, bispectrum, ilist, type, inum, ncoeff, nperdim, ntype, quadraticflag);
}

void r2_cpuSnapTallyBispectrumDeriv(double *db, double *bispectrum, double *dbdr, int *aii, int *ai, int *aj, int *ti,
                                    int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag) {
    cpuSnapTallyBispectrumDeriv(db, bispectrum, dbdr, aii, ai, aj, ti, inum, ijnum, ncoeff, nperdim, ntype,
                                quadraticflag);
}

void r2_cpuSnapTallyBispectrumVirial(double *bv, double *bispectrum, double *dbdr, double *rij, int *aii, int *ai,
                                     int *aj, int *ti, int inum, int ijnum, int ncoeff, int nperdim, int ntype,
                                     int quadraticflag) {
    cpuSnapTallyBispectrumVirial(bv, bispectrum, dbdr, rij, aii, ai, aj, ti, inum, ijnum, ncoeff, nperdim, ntype,
                                 quadraticflag);
}

// Nose-Hoover NVT integrator halves (src/C++/Core/cpuVelocityIntegration.cpp:602-660)
void r2_cpuNVTInitialIntegrate(double *x, double *v, double *f, double *mass, double *dtarray, double *tarray,
                               double *eta_mass, double *eta, double *eta_dot, double *eta_dotdot, int *type,
                               int *ilist, int eta_mass_flag, int biasflag, int mtchain, int nc_tchain, int dim,
                               int inum) {
    cpuNVTInitialIntegrate(x, v, f, mass, dtarray, tarray, eta_mass, eta, eta_dot, eta_dotdot, type, ilist,
                           eta_mass_flag, biasflag, mtchain, nc_tchain, dim, inum);
}

void r2_cpuNVTFinalIntegrate(double *x, double *v, double *f, double *mass, double *dtarray, double *tarray,
                             double *eta_mass, double *eta, double *eta_dot, double *eta_dotdot, int *type,
                             int *ilist, int eta_mass_flag, int biasflag, int mtchain, int nc_tchain, int dim,
                             int inum) {
    cpuNVTFinalIntegrate(x, v, f, mass, dtarray, tarray, eta_mass, eta, eta_dot, eta_dotdot, type, ilist,
                         eta_mass_flag, biasflag, mtchain, nc_tchain, dim, inum);
}

// the upstream velocity draws themselves, all loop modes (cpuVelocityIntegration.cpp:54-135)
void r2_cpuVelocityCreate(double *x, double *v, double *mass, double *second, int *seed, int *save, int *map,
                          int *type, int seed0, int sum_flag, int dist_flag, int loop_flag, int dim, int mpiRank,
                          int nlocal, int natoms) {
    cpuVelocityCreate(x, v, mass, second, seed, save, map, type, seed0, sum_flag, dist_flag, loop_flag, dim, mpiRank,
                      nlocal, natoms);
}

// initial velocities: the call sequence of readdata.cpp:609-640 on the upstream functions
void r2_velocity_create(double *x, double *v, double *mass, int *type, int natoms, double t_desired, int seed0,
                        int dist_flag, int loop_flag, double mvv2e, double boltz) {
    std::vector<int> ilist(natoms), seed(1, 0), save(1, 0);
    std::vector<double> second(1, 0.0), amass(natoms);
    for (int i = 0; i < natoms; ++i) ilist[i] = i;
    cpuVelocityCreate(x, v, mass, second.data(), seed.data(), save.data(), ilist.data(), type, seed0, 0, dist_flag,
                      loop_flag, 3, 0, natoms, natoms);
    const double masstotal = cpuComputeMass(amass.data(), mass, type, ilist.data(), natoms);
    double vcm[3];
    cpuComputeVCM(vcm, v, mass, masstotal, ilist.data(), type, 3, natoms);
    cpuVelocityZeroMomentum(v, vcm, 3, natoms);
    const double tdof = (natoms - 1) * 3;
    const double tfactor = mvv2e / (tdof * boltz);
    const double temp = cpuComputeTempScalar(v, mass, tfactor, type, ilist.data(), 3, natoms);
    cpuArrayMultiplyScalar(v, sqrt(t_desired / temp), natoms * 3);
}

// energy / force / virial tallies of the split force path (src/C++/Core/cpuSnap.cpp:1746-1831)
void r2_cpuSnapTallyEnergyFull(double *eatom, double *bispectrum, double *coeffelem, int *ilist, int *map, int *type,
                               int inum, int ncoeff, int ncoeffall, int quadraticflag) {
    cpuSnapTallyEnergyFull(eatom, bispectrum, coeffelem, ilist, map, type, inum, ncoeff, ncoeffall, quadraticflag);
}
void r2_cpuSnapTallyEnergyFull2(double *eatom, double *bispectrum, double *coeffelem, int *ilist, int *map, int *type,
                                int inum, int ncoeff, int ncoeffall, int quadraticflag) {
    cpuSnapTallyEnergyFull2(eatom, bispectrum, coeffelem, ilist, map, type, inum, ncoeff, ncoeffall, quadraticflag);
}
void r2_cpuSnapTallyForceFull(double *fatom, double *fij, int *ai, int *aj, int *alist, int ijnum) {
    cpuSnapTallyForceFull(fatom, fij, ai, aj, alist, ijnum);
}
void r2_cpuSnapTallyVirialFull(double *vatom, double *fij, double *rij, int *ai, int *aj, int inum, int ijnum) {
    cpuSnapTallyVirialFull(vatom, fij, rij, ai, aj, inum, ijnum);
}

// the empirical pair route of the upstream app (src/C++/Potentials/cpuPotentials.cpp:1587-1649): the
// generated pair gradient (potnum 1), the full-list energy/force decomposition and the pair virial
void r2_cpuPairaGradient(double *u, double *u_xij, double *xij, int ng) {
    cpuPairaGradient(u, u_xij, xij, (double *)nullptr, (double *)nullptr, (int *)nullptr, (int *)nullptr,
                     (int *)nullptr, (int *)nullptr, (double *)nullptr, (double *)nullptr, (int *)nullptr, 3, 0, 0, 0,
                     0, ng, 1);
}
void r2_cpuFullNeighPairDecomposition(double *e, double *f, double *eij, double *fij, int *ai, int ijnum, int dim) {
    cpuFullNeighPairDecomposition(e, f, eij, fij, ai, ijnum, dim);
}
void r2_cpuVirialPairTally(double *v, double *fij, double *rij, double factor, int *ai, int dim, int inum,
                           int ijnum) {
    cpuVirialPairTally(v, fij, rij, factor, ai, dim, inum, ijnum);
}

// lattice atoms of the application (readdata.cpp:490-551, 647-676): cpuLattice, the orthogonal box of the
// region in lattice spacings (cpuSetGlobalBox, cpuSetLocalOrthBox, cpuBoundingSubbox), the lattice-space
// bounding box (cpuLatticeBoundingBox), the index range of latticestruct::latticebounds
// (src/C++/Common/mdpcommon.h:393-413), cpuAtomLattice and cpuAtomAdd; metal units, one process.
// Returns the number of atoms written (at most cap), or -(needed) when cap is too small.
int r2_app_lattice_atoms(double *x, int *type, int cap, int style, double scale, int spaceflag, const int *orient9,
                         const double *origin3, const double *spacing3, const double *a9, const double *basis_custom,
                         const int *btype, int nbasis_custom, const double *reglo3, const double *reghi3,
                         double *spacing_out3) {
    double basis[8 * 3] = {0}, primitive[9], priminv[9], rotaterow[9], rotatecol[9], origin[3], spacing[3];
    double a1[3], a2[3], a3[3];
    int ox[3], oy[3], oz[3];
    for (int d = 0; d < 3; ++d) {
        origin[d] = origin3[d];
        spacing[d] = spacing3[d];
        a1[d] = a9[d];
        a2[d] = a9[3 + d];
        a3[d] = a9[6 + d];
        ox[d] = orient9[d];
        oy[d] = orient9[3 + d];
        oz[d] = orient9[6 + d];
    }
    for (int k = 0; k < 3 * nbasis_custom && k < 24; ++k) basis[k] = basis_custom[k];
    int nbasis = cpuLattice(basis, primitive, rotaterow, priminv, rotatecol, origin, spacing, a1, a2, a3, scale, ox, oy,
                            oz, style, 2, spaceflag, 3);
    if (style == 9) nbasis = nbasis_custom;
    for (int d = 0; d < 3; ++d) spacing_out3[d] = spacing[d];
    double boxlo[3], boxhi[3], tilt[3] = {0, 0, 0}, h[6], hinv[6], blo[3], bhi[3], sublo[3], subhi[3];
    double lo_l[3] = {0, 0, 0}, hi_l[3] = {1, 1, 1}, bsublo[3], bsubhi[3], lmin[3], lmax[3];
    for (int d = 0; d < 3; ++d) {
        boxhi[d] = spacing[d] * reghi3[d];
        boxlo[d] = spacing[d] * reglo3[d];
    }
    cpuSetGlobalBox(h, hinv, blo, bhi, boxhi, boxlo, tilt, 0);
    cpuSetLocalOrthBox(subhi, sublo, boxhi, boxlo, hi_l, lo_l, 3);
    cpuBoundingSubbox(bsublo, bsubhi, sublo, subhi, lo_l, hi_l, boxlo, h, 0);
    cpuLatticeBoundingBox(lmin, lmax, bsublo, bsubhi, primitive, rotaterow, priminv, rotatecol, origin, spacing, scale);
    int ilo = static_cast<int>(lmin[0]) - 1, jlo = static_cast<int>(lmin[1]) - 1, klo = static_cast<int>(lmin[2]) - 1;
    int ihi = static_cast<int>(lmax[0]) + 1, jhi = static_cast<int>(lmax[1]) + 1, khi = static_cast<int>(lmax[2]) + 1;
    if (lmin[0] < 0.0) ilo--;
    if (lmin[1] < 0.0) jlo--;
    if (lmin[2] < 0.0) klo--;
    const int natom = nbasis * (ihi - ilo + 1) * (jhi - jlo + 1) * (khi - klo + 1);
    std::vector<double> y(3 * static_cast<size_t>(natom));
    std::vector<int> ilist(natom), bt(btype, btype + nbasis);
    cpuAtomLattice(y.data(), ilist.data(), basis, primitive, rotaterow, origin, spacing, scale, bt.data(), nbasis, 0,
                   ilo, ihi, jlo, jhi, klo, khi, 3);
    std::vector<double> xs(3 * static_cast<size_t>(natom));
    std::vector<int> ts(natom);
    const int n = cpuAtomAdd(xs.data(), ts.data(), y.data(), hinv, boxlo, lo_l, hi_l, ilist.data(), 3, 0, natom);
    if (n > cap) return -n;
    for (int q = 0; q < 3 * n; ++q) x[q] = xs[q];
    for (int q = 0; q < n; ++q) type[q] = ts[q];
    return n;
}

}  // extern "C"
