/*
 * mdpsnap.h -- C ABI of libmdpsnap, the MI355X (gfx950) implementation of the SNAP path of MDP.
 *
 * Part 1 is the drop-in: the 23 entry points of the upstream SNAP C library
 * (src/Julia/Potential/cpuSnap.h:4-88) with identical names, signatures and host-pointer semantics,
 * so that the upstream Julia driver (src/Julia/Potential/snapcpp.jl) binds this library by changing
 * only its library path.  Each call stages its host arrays to the GPU, runs the HIP kernels and
 * copies the results back before returning.  Outputs the upstream function accumulates into (+=)
 * are accumulated here too; outputs it overwrites are overwritten.  Atom indices are 0-based,
 * element types 1-based, exactly as upstream.  twojmax must be 2, 4, 6 or 8 (every upstream example
 * uses 4, 6 or 8); other values set an error and leave the outputs untouched.
 *
 * Part 2 is this library's own device-resident MD engine (no upstream counterpart in this ABI; it
 * plays the role of the upstream C++ velocity-Verlet driver, src/C++/Integration/Integration.cpp,
 * around the same SNAP kernels): neighbour lists, SNAP forces and NVE integration stay on the GPU.
 *
 * Errors: the drop-in functions return void like upstream; a failure leaves a message readable
 * with mdp_snap_last_error() (empty string when the last call succeeded).  Part 2 returns 0 on
 * success, nonzero on failure (message likewise).
 */
#ifndef MDPSNAP_H
#define MDPSNAP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------- */
/* Part 1: drop-in for src/Julia/Potential/cpuSnap.h                                               */
/* ---------------------------------------------------------------------------------------------- */

/* cpuSnap.cpp:34-45 */
int idxcgcount(int twojmax);
/* cpuSnap.cpp:60-69 */
int idxucount(int twojmax);
/* cpuSnap.cpp:23-32 */
int idxbcount(int twojmax);
/* cpuSnap.cpp:47-58 */
int idxzcount(int twojmax);

/* cpuSnap.cpp:248-256: index lists, sqrt(p/q) table and Clebsch-Gordan table */
void cpuInitSna(double *rootpqarray, double *cglist, double *factorial, int *idx_max, int *idxz,
      int *idxz_block, int *idxb, int *idxb_block, int *idxu_block, int *idxcg_block, int twojmax);

/* cpuSnap.cpp:258-286: per-atom pair filter d^2 < rcutsq[jtype + itype*(ntypes+1)], d^2 > 1e-20 */
void cpuNeighPairList(int *pairnum, int *pairlist, double *x, double *rcutsq, int *ilist, int *neighlist,
        int *neighnum, int *atomtype, int *alist, int inum, int dim, int ntypes);

/* cpuSnap.cpp:289-310: pair arrays r_ij, aii, ai, aj (owner), ti, tj */
void cpuNeighPairs(double *xij, double *x, int *aii, int *ai, int *aj, int *ti, int *tj,
        int *pairlist, int *pairnumsum, int *ilist, int *atomtype, int *alist, int inum, int dim);

/* cpuSnap.cpp:312-490: Utot += sum over pairs of sfac*wj*U (half rows) */
void cpuSnapComputeUi(double *Utotr, double *Utoti, double *rootpqarray, double *rij, double *wjelem,
        double *radelem, double rmin0, double rfac0, double rcutfac, int *map, int *aii, int *ti, int *tj,
        int twojmax, int idxu_max, int inum, int ijnum, int switchflag, int chemflag);

/* cpuSnap.cpp:492-542: + wself on the diagonal, mirror of rows 2mb < j */
void cpuAddWself2Ui(double *Utotr, double *Utoti, double wself, int *idxu_block, int *type, int *map,
        int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum);

/* cpuSnap.cpp:544-719: eatom[ilist[ii]] += beta0 + sum_k beta_k B_k */
void cpuSnapComputeEi(double *eatom, double *Utotr, double *Utoti, double *cglist, double *bzero,
        double *coeffelem, int *ilist, int *map, int *type, int *idxb, int *idxcg_block, int *idxu_block,
        int twojmax, int idxb_max, int idxu_max, int nelements, int ncoeffall, int bnorm_flag,
        int bzero_flag, int wselfall_flag, int chemflag, int inum);

/* cpuSnap.cpp:721-889: blist = B (overwrite) */
void cpuSnapComputeBi(double *blist, double *Utotr, double *Utoti, double *cglist, double *bzero,
        int *idxb, int *idxcg_block, int *idxu_block, int *idxz_block, int twojmax, int idxb_max,
        int idxu_max, int idxz_max, int nelements, int bnorm_flag, int bzero_flag, int wselfall_flag,
        int inum);

/* cpuSnap.cpp:891-989: Y (zeroed, then sum of beta-weighted Z) */
void cpuSnapComputeYi(double *ylist_r, double *ylist_i, double *Utotr, double *Utoti, double *cglist,
        double *beta, int *map, int *type, int *idxz, int *idxb_block, int *idxu_block, int *idxcg_block,
        int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements, int ncoeffall,
        int bnorm_flag, int chemflag, int inum);

/* cpuSnap.cpp:991-1410: fatom += dE/dr on ai, -= on aj; vatom[6*ai+..] += -r (x) dE/dr (AoS) */
void cpuSnapComputeFi(double *fatom, double *vatom, double *ylist_r, double *ylist_i, double *rootpqarray,
        double *rij, double *wjelem, double *radelem, double rmin0, double rfac0, double rcutfac, int *map,
        int *aii, int *ai, int *aj, int *ti, int *tj, int twojmax, int idxu_max, int inum, int anum,
        int ijnum, int switchflag, int chemflag);

/* cpuSnap.cpp:1412-1618: per-pair U and d(sfac U)/dr, separate x/y/z arrays */
void cpuComputeSij(double *Sr, double *Si, double *Srx, double *Six, double *Sry, double *Siy, double *Srz,
        double *Siz, double *rootpqarray, double *rij, double *wjelem, double *radelem, double rmin0,
        double rfac0, double rcutfac, int *idxu_block, int *ti, int *tj, int twojmax, int idxu_max,
        int ijnum, int switch_flag);

/* cpuSnap.cpp:1620-1833: per-pair U and d(sfac U)/dr, derivative blocks packed [k][jju][ij] */
void cpuComputeUij(double *Sr, double *Si, double *dSr, double *dSi, double *rootpqarray, double *rij,
        double *wjelem, double *radelem, double rmin0, double rfac0, double rcutfac, int *idxu_block,
        int *ti, int *tj, int twojmax, int idxu_max, int ijnum, int switch_flag);

/* cpuSnap.cpp:1835-1863: Stot = wself on the own-element diagonal */
void cpuZeroUarraytot2(double *Stotr, double *Stoti, double wself, int *idxu_block, int *type, int *map,
        int *ai, int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum);

/* cpuSnap.cpp:1865-1897: Stot[ai] += S */
void cpuAddUarraytot(double *Stotr, double *Stoti, double *Sr, double *Si, int *map, int *ai, int *tj,
        int idxu_max, int inum, int ijnum, int chemflag);

/* cpuSnap.cpp:1899-1969: zlist (overwrite) */
void cpuComputeZi2(double *zlist_r, double *zlist_i, double *Stotr, double *Stoti, double *cglist,
        int *idxz, int *idxu_block, int *idxcg_block, int twojmax, int idxu_max, int idxz_max,
        int nelements, int bnorm_flag, int inum);

/* cpuSnap.cpp:1971-2072: blist = 2 sum' Stot*.Z - bzero (overwrite) */
void cpuComputeBi2(double *blist, double *zlist_r, double *zlist_i, double *Stotr, double *Stoti,
        double *bzero, int *ilist, int *type, int *map, int *idxb, int *idxu_block, int *idxz_block,
        int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements, int bzero_flag,
        int wselfall_flag, int chemflag, int inum);

/* cpuSnap.cpp:2074-2263: dB/dr per pair (zeroed, then filled) */
void cpuComputeDbidrj(double *dblist, double *zlist_r, double *zlist_i, double *dulist_r, double *dulist_i,
        int *idxb, int *idxu_block, int *idxz_block, int *map, int *ai, int *tj, int twojmax,
        int idxb_max, int idxu_max, int idxz_max, int nelements, int bnorm_flag, int chemflag,
        int inum, int ijnum);

/* cpuSnap.cpp:2265-2291 */
void cpuSnapTallyBispectrum(double *bi, double *bispectrum, int *ilist, int *type, int inum, int ncoeff,
        int nperdim, int ntype, int quadraticflag);

/* cpuSnap.cpp:2293-2365 */
void cpuSnapTallyBispectrumDeriv(double *db, double *bispectrum, double *dbdr, int *aii, int *ai, int *aj,
        int *ti, int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag);

/* cpuSnap.cpp:2367-2470 */
void cpuSnapTallyBispectrumVirial(double *bv, double *bispectrum, double *dbdr, double *rij, int *aii,
        int *ti, int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag);

/* Message of the last failed call on this thread ("" after a successful call). */
const char *mdp_snap_last_error(void);

/* ---------------------------------------------------------------------------------------------- */
/* Part 2: device-resident MD engine                                                               */
/* ---------------------------------------------------------------------------------------------- */

typedef struct mdp_md mdp_md;   /* opaque */

typedef struct {
    int twojmax;                /* 2, 4, 6 or 8 */
    int ntypes;                 /* element types 1..ntypes (chemflag = 0: one element block) */
    double rcutfac, rfac0, rmin0;
    int switchflag, bzeroflag, bnormflag, wselfallflag;
    const double *elemradius;   /* [ntypes] */
    const double *elemweight;   /* [ntypes] */
    const double *coeff;        /* [ntypes * (ncoeff+1)]: beta0, beta_1..beta_ncoeff per type */
    const double *mass;         /* [ntypes], amu */
    double dt;                  /* time step, ps (metal units) */
    double skin;                /* neighbour skin, Angstrom */
    int rebuild_every;          /* neighbour rebuild period in steps (>= 1) */
} mdp_md_params;

/* Positions x[3N] (Angstrom, AoS), velocities v[3N] (Angstrom/ps), types[N] (1-based), orthogonal
 * periodic box box[3].  Allocates all device state on the current device. */
int mdp_md_create(mdp_md **md, const mdp_md_params *p, int natoms, const double *box, const double *x,
                  const double *v, const int *types);
int mdp_md_destroy(mdp_md *md);

/* Rebuild the neighbour list and evaluate energies, forces and virials at the current positions. */
int mdp_md_compute(mdp_md *md);

/* Advance nsteps velocity-Verlet NVE steps (forces must be current: call mdp_md_compute first). */
int mdp_md_run(mdp_md *md, int nsteps);

/* Copy state back to the host; any pointer may be NULL.  vatom is 6 per atom (xx,yy,zz,xy,xz,yz),
 * the whole pair virial -r_ij (x) dE/dr_ij on atom i as in cpuSnap.cpp:1384-1396. */
int mdp_md_get(mdp_md *md, double *x, double *v, double *f, double *eatom, double *vatom);

/* Host copy of the neighbour list of the last rebuild: for atom i its entries are
 * [off[i], off[i+1]) of (j, image shift sx, sy, sz).  With nbj == NULL only *nentries is set. */
int mdp_md_neighbors(mdp_md *md, int *off, int *nbj, int *nbshift, long *nentries);

/* Statistics: [0] pairs in the last force call, [1] neighbour entries, [2] rebuilds,
 * [3] rebuilds where an atom had moved more than skin/2 (list possibly incomplete),
 * [4] steps run, [5] neighbour-list overflows detected (run aborted). */
int mdp_md_stats(mdp_md *md, long *out6);

/* Per-kernel timing over the steps run since it was enabled: with on = 1 the engine records a
 * hipEvent pair around each kernel class on its stream.  names are fixed: see mdp_md_timer_name. */
int mdp_md_set_timing(mdp_md *md, int on);
int mdp_md_timers(mdp_md *md, double *ms_total, long *launches, int n);
const char *mdp_md_timer_name(int k);
int mdp_md_num_timers(void);

/* Launch variants chosen by the engine's autotuner after the first force evaluation, as a JSON object
 * with the measured ms per launch of every variant (set MDP_AUTOTUNE=0 to keep the defaults). */
int mdp_md_tuning(mdp_md *md, char *buf, int len);

/* Total potential and kinetic energy (eV) of the current state. */
int mdp_md_energy(mdp_md *md, double *epot, double *ekin);

#ifdef __cplusplus
}
#endif
#endif
