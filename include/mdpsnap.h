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
 * around the same SNAP kernels): neighbour lists, SNAP forces and NVE or Nose-Hoover NVT integration
 * stay on the GPU.
 *
 * Part 3 computes fitting descriptors (B and dB/dr) of many configurations in one batch.
 * Part 4 holds the kernels of the learned surrogate (per-atom MLP over bispectrum descriptors).
 * Part 5 is the backend-3 (HIP) branch of the upstream C++ core dispatch (src/C++/Common/mdpcore.h) for
 * the SNAP force path and the descriptor path of the upstream fitting driver: device arrays in and out;
 * include/mdphipcore.h wraps it as the hip* templates.
 * Part 6 is the reference pair potential (ZBL) the SNAP application adds to SNAP, as backend-3 entry
 * points of the upstream empirical-potential route.
 * Part 7 reads the upstream application's binary inputs (app.bin, config.bin, lattice.bin, ...) and sets
 * up the MD engine from them.
 *
 * Errors: the drop-in functions return void like upstream; a failure leaves a message readable
 * with mdp_snap_last_error() (empty string when the last call succeeded).  Part 2 returns 0 on
 * success, nonzero on failure (message likewise).
 */
#ifndef MDPSNAP_H
#define MDPSNAP_H

#include <stddef.h>

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
 * hipEvent pair around each kernel class on its stream, with on = 2 around compute_ui, compute_yi and
 * compute_dedr only (0 switches it off).  names are fixed: see mdp_md_timer_name. */
int mdp_md_set_timing(mdp_md *md, int on);
int mdp_md_timers(mdp_md *md, double *ms_total, long *launches, int n);
const char *mdp_md_timer_name(int k);
int mdp_md_num_timers(void);

/* Launch variants chosen by the engine's autotuner after the first force evaluation, as a JSON object
 * with the measured ms per launch of every variant (set MDP_AUTOTUNE=0 to keep the defaults). */
int mdp_md_tuning(mdp_md *md, char *buf, int len);

/* Mean number of pairs per force evaluation since timing was last switched (mdp_md_set_timing). */
int mdp_md_mean_pairs(mdp_md *md, double *mean);

/* Total potential and kinetic energy (eV) of the current state. */
int mdp_md_energy(mdp_md *md, double *epot, double *ekin);
/* Thermo output of the upstream app (ThermoOutput, Calculation.cpp:436-465) for the current state:
 * out4 = potential energy per atom (eV), temperature (K, tdof = 3 (N - 1)), kinetic energy per atom
 * (eV, 0.5 tdof k T / N) and pressure (bar: (tdof k T + W_xx + W_yy + W_zz) / (3 V) nktv2p). */
int mdp_md_thermo(mdp_md *md, double *out4);

/* Nose-Hoover chain NVT instead of NVE for the following mdp_md_run steps (cpuNVTInitialIntegrate /
 * cpuNVTFinalIntegrate / cpuNoseHooverThermostat, src/C++/Core/cpuVelocityIntegration.cpp:505-648):
 * a thermostat half-step before the first velocity half-step and after the second, the chain masses
 * recomputed from the target every half-step (eta_mass_flag = 1).  The target ramps linearly,
 * t_start + (step - beginstep) / (endstep - beginstep) (t_stop - t_start), where step counts from
 * currentstep and is advanced at the start of every step.  t_period is the damping time (ps;
 * t_freq = 1 / t_period, readdata.cpp:432); tdof <= 0 selects 3 (N - 1) (Calculation.cpp:450);
 * mtchain 1..9 (upstream default 3), nc_tchain >= 1 (default 1).  The chain starts at rest. */
int mdp_md_set_nvt(mdp_md *md, double t_start, double t_stop, double t_period, double drag, double tdof, int mtchain,
                   int nc_tchain, long beginstep, long endstep, long currentstep);
/* back to NVE */
int mdp_md_clear_nvt(mdp_md *md);
/* New positions x[3N] and, when v is not NULL, velocities v[3N] from the host (a restart, or a state
 * prepared elsewhere).  Non-finite values are rejected and leave the engine unchanged; the next compute
 * or step rebuilds the neighbour lists and recomputes the forces.  (During a run, a position that turns
 * non-finite or runs away is reported by the next neighbour rebuild as an error.) */
int mdp_md_set_state(mdp_md *md, const double *x, const double *v);
/* Adds the ZBL reference pair potential of the SNAP Ta application (simulating/snap/ZBLpotential.jl with
 * mu = [cut_inner, cut_outer, z], snapTA06A.jl:62-66: 4, 4.8, 73) to SNAP from the next force evaluation
 * on, with the upstream full-neighbour-list conventions (src/C++/Potentials/cpuPotentials.cpp:1587-1649):
 * e_i += sum_j u(r_ij)/2, F_i += sum_j du/dxij, per-atom virial -1/2 sum_j xij (x) du/dxij, pairs with
 * r_ij^2 <= cut_outer^2.  The neighbour list is then built at max(SNAP cutoff, cut_outer) + skin.
 * Needs 0 < cut_inner < cut_outer and z > 0. */
int mdp_md_set_zbl(mdp_md *md, double cut_inner, double cut_outer, double z);
/* Initial velocities as the upstream app creates them (readdata.cpp:609-640; host code): per-atom
 * uniform (dist_flag 0) or normal (1) draws of cpuVelocityCreate in loop mode 2 ("geom", the mode of the
 * upstream examples; the other modes are rejected), scaled by 1/sqrt(mass), then zero total momentum and
 * rescaling to t_desired with tdof = 3 (natoms - 1).  mass[t] for type t (mass[0] unused); mvv2e and
 * boltz as the application's units (metal: 1.0364269e-4, 8.617343e-5).  v[3*natoms] is overwritten. */
int mdp_velocity_create(const double *x, const int *type, const double *mass, int natoms, double t_desired,
                        int seed0, int dist_flag, int loop_flag, double mvv2e, double boltz, double *v);
/* cpuVelocityCreate (src/C++/Core/cpuVelocityIntegration.cpp:54-135; host code) with the generator
 * state held by the caller: seed[0] (Park-Miller state), save[0] and second[0] (a waiting normal
 * deviate), read at entry and written back.  loop_flag 0 ("all"): every atom of the system draws in
 * global-ID order, ID k = 1..natoms, and map[k] (natoms + 1 entries) names its local index, which takes
 * the draw when 0 <= map[k] < nlocal; 1 ("local"): seed[0] += mpi_rank, 100 discarded draws, then the
 * nlocal atoms in order; 2 ("geom", dim = 3): each atom's stream re-seeded from seed0 and its
 * coordinates.  Three uniform - 0.5 (dist_flag 0) or normal (1) draws per atom, the first dim of them
 * added as v += draw / sqrt(mass[type[i]]); v[dim*nlocal] is zeroed first unless sum_flag. */
int mdp_velocity_create_loop(const double *x, double *v, const double *mass, double *second, int *seed, int *save,
                             const int *map, const int *type, int seed0, int sum_flag, int dist_flag, int loop_flag,
                             int dim, int mpi_rank, int nlocal, int natoms);
/* out[43]: eta_mass[10], eta[10], eta_dot[10], eta_dotdot[10], the temperature after the last
 * thermostat half-step, its target, and the step counter. */
int mdp_md_nvt_state(mdp_md *md, double *out);

/* Learned surrogate (no upstream counterpart): from the next force evaluation on, forces and
 * energies come from the per-atom MLP of part 4 (params: its FP32 parameters; aff: mu[16] |
 * 1/sigma[16] | e_scale | e_shift over descriptors B - bzero_j of twojmax 2 or 4, nin = their count),
 * through the descriptor chain rule.  Every audit_every steps (0 = never) plain SNAP is evaluated on
 * the same state; if the RMS force difference exceeds audit_tol (eV/A) that step uses the plain
 * forces.  mdp_md_clear_surrogate returns to plain SNAP. */
int mdp_md_set_surrogate(mdp_md *md, const float *params, const double *aff, int nin, int twojmax,
                         int audit_every, double audit_tol);
int mdp_md_clear_surrogate(mdp_md *md);
/* [0] audits, [1] fallbacks to plain forces, [2] last, [3] max, [4] mean RMS force difference (eV/A) */
int mdp_md_audits(mdp_md *md, double *out5);

/* ---------------------------------------------------------------------------------------------- */
/* Part 3: batched fitting descriptors                                                             */
/* ---------------------------------------------------------------------------------------------- */

/* Per-atom bispectrum B and per-pair dB/dr of many independent periodic configurations at once.
 * Upstream computes them one configuration at a time (cpuComputeUij -> cpuZeroUarraytot2 ->
 * cpuAddUarraytot -> cpuComputeZi2 -> cpuComputeBi2 -> cpuComputeDbidrj, cpuSnap.cpp:1620-2263,
 * driven by snapcpp.jl:539-665 per configuration).  Here the pair build, Utot, Z + B and dB/dr of
 * the whole batch are one captured hipGraph; dB/dr is a persistent kernel over all atoms.
 * Single chemical element block (chemflag = 0); wself = 1. */
typedef struct mdp_fit mdp_fit;   /* opaque */

typedef struct {
    int twojmax;                /* 2, 4, 6 or 8 */
    int ntypes;                 /* atom types 1..ntypes */
    double rcutfac, rfac0, rmin0;
    int switchflag, bzeroflag, bnormflag;
    const double *elemradius;   /* [ntypes] */
    const double *elemweight;   /* [ntypes] */
} mdp_fit_params;

/* nconfig configurations; configuration c has natoms[c] atoms (concatenated in x[3*sum] AoS and
 * types[sum], 1-based) in the orthogonal periodic box boxes[3c..3c+2]; coordinates must be finite
 * and within [-L, 2L) of their box.  Pairs of atom i are all
 * periodic images x_j + s*L (j of the same configuration, the atom itself at s = 0 excluded) with
 * |x_j + s*L - x_i|^2 <= ((r_ti + r_tj) rcutfac)^2, ordered by j then (sz, sy, sx). */
int mdp_fit_create(mdp_fit **h, const mdp_fit_params *p, int nconfig, const int *natoms, const double *boxes,
                   const double *x, const int *types);
/* The same for periodic cells that need not be orthogonal: cells[9c..9c+8] holds the cell vectors a, b, c
 * of configuration c in the form a = (Lx, 0, 0), b = (bx, Ly, 0), c = (cx, cy, Lz) (the form
 * src/Julia/Preprocessing/rotatelattice.jl:1-21 rotates every lattice into), tilts smaller than the box
 * lengths.  Pairs of atom i are all images x_j + s1 a + s2 b + s3 c within the cutoff, ordered by j then
 * (s3, s2, s1); orthogonal cells give exactly the pairs of mdp_fit_create.  Positions must be finite and
 * within [-2L, 3L) per component when a cell is tilted. */
int mdp_fit_create_cells(mdp_fit **h, const mdp_fit_params *p, int nconfig, const int *natoms, const double *cells,
                         const double *x, const int *types);
int mdp_fit_destroy(mdp_fit *h);
/* Replace all positions (same configurations and atom counts). */
int mdp_fit_set_positions(mdp_fit *h, const double *x);
/* Pairs, Utot, Z, B and dB/dr of every configuration (the first call also autotunes the dB/dr
 * launch and captures the graph); synchronous. */
int mdp_fit_compute(mdp_fit *h);
/* Replay the captured graph nreps times without waiting (timing loops); mdp_fit_sync waits. */
int mdp_fit_replay(mdp_fit *h, int nreps);
int mdp_fit_sync(mdp_fit *h);
/* atom_off[nconfig+1], pair_off[nconfig+1]: first atom / pair of each configuration (totals last). */
int mdp_fit_layout(mdp_fit *h, long *atom_off, long *pair_off);
/* Outputs of the last compute, configuration after configuration, each in the upstream layout of
 * that configuration alone: blist block of configuration c at atom_off[c]*idxb_max, element
 * [ii + N_c*jjb] (cpuComputeBi2); dblist block at pair_off[c]*3*idxb_max, element
 * [ij + P_c*k + 3*P_c*jjb] (cpuComputeDbidrj).  Either pointer may be NULL. */
int mdp_fit_get(mdp_fit *h, double *blist, double *dblist);
/* Pair arrays of the last compute: ai/aj local 0-based atom indices within each configuration (aj is
 * the owner of the image), rij[3P] AoS.  Any pointer may be NULL. */
int mdp_fit_pairs(mdp_fit *h, int *ai, int *aj, double *rij);
/* Autotuner choices and timings (JSON). */
int mdp_fit_tuning(mdp_fit *h, char *buf, int len);
/* ms of one uncaptured run of each stage: [0] pairs, [1] Utot, [2] Z+B, [3] dB/dr. */
int mdp_fit_profile(mdp_fit *h, double *ms4);
/* Tallies of the last compute, as the upstream descriptor driver returns them for every configuration
 * (snapcpp.jl:645-664 with cpuSnapTallyBispectrum, ...Deriv, ...Virial, cpuSnap.cpp:2265-2470), with
 * Kt = ntypes*idxb_max columns, column (t-1)*idxb_max + k for the atoms (pairs) of type t (owner type):
 * bi[c*Kt + col]; bd block of configuration c at 3*atom_off[c]*Kt, element [comp + 3*ii + 3*N_c*col]
 * (sign as upstream: + on the pair's owner, - on its neighbour, so forces are bd . beta); bv[c*6*Kt +
 * 6*col + m], m = xx, yy, zz, xy, xz, yz of -r (x) dB/dr.  Any pointer may be NULL. */
int mdp_fit_descriptors(mdp_fit *h, double *bi, double *bd, double *bv);
/* Weighted normal equations of the linear fit (linearfit.jl:168-200) from the last compute:
 * L[g*Kt*Kt + k*Kt + l] = sum_r w_r^2 a_r[k] a_r[l], hv[g*Kt + k] = sum_r w_r^2 a_r[k] b_r for the groups
 * g = 0 energy (rows bi_c, targets e[c], both divided by N_c when normalizeenergy), 1 force (the 3 N_c
 * rows of bd_c, targets f[3*atom + comp] over the batch), 2 stress (the virial rows in the upstream order
 * xx, yy, zz, yz, xz, xy (snapcpp.jl:660-662), targets s[6c + m]).  Weights per configuration (we, wf,
 * ws; NULL = 1).  A NULL target array leaves its group zero.  Sums run in a fixed order. */
int mdp_fit_normal_equations(mdp_fit *h, const double *e, const double *f, const double *s, const double *we,
                             const double *wf, const double *ws, int normalizeenergy, double *L, double *hv);
/* lsqsolve (linearfit.jl:29-75) on the normal equations: ce = Le\he, cf = Lf\hf, cef = (Le+Lf)\(he+hf),
 * cefs = (Le+Lf+Ls)\(he+hf+hs), each when its inputs are not all zero, the diagonal scaled by 1 + 1e-15
 * (and a(0,0) of the force system raised to 1e-10 when smaller), LU with partial pivoting.  solved[4]
 * flags which were computed; NULL outputs are skipped.  Host only. */
int mdp_lsqsolve(int K, const double *Le, const double *Lf, const double *Ls, const double *he, const double *hf,
                 const double *hs, double *ce, double *cf, double *cef, double *cefs, int *solved);

/* ---------------------------------------------------------------------------------------------- */
/* Part 4: learned surrogate kernels (no upstream counterpart)                                     */
/* ---------------------------------------------------------------------------------------------- */

/* Per-atom MLP x -> 64 -> 64 -> 1 (SiLU) over nin <= 16 standardized bispectrum components, BF16
 * weights and activations on the matrix cores with FP32 accumulation.  Parameters are an FP32 array
 * W1[16][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3; mdp_mlp_pack turns them into the BF16
 * operand tiles the kernels read.  aff = mu[16] | 1/sigma[16] | e_scale | e_shift (doubles):
 * E_atom = e_shift + e_scale * f((B - mu)/sigma).  Descriptor-shaped arrays are component-major
 * (x[k*n + atom]).  All pointers are device pointers owned by the caller; calls are asynchronous on
 * `stream` (a hipStream_t, NULL = the library stream) and allocate nothing. */
int mdp_mlp_sizes(int *nparam, int *npack, int *nin_max, int *nhidden);
/* Device buffers for the part-4 kernels (copies run on the library stream after its pending kernels
 * and return when done; mdp_dev_zero is asynchronous on it). */
int mdp_dev_malloc(void **p, size_t bytes);
int mdp_dev_free(void *p);
int mdp_dev_h2d(void *dst, const void *src, size_t bytes);
int mdp_dev_d2h(void *dst, const void *src, size_t bytes);
int mdp_dev_zero(void *p, size_t bytes);
int mdp_dev_sync(void);
int mdp_mlp_pack(const float *params, unsigned short *packed, void *stream);
/* e[n] = E_atom; g[k*n + a] = dE_atom/dB_k (g may be NULL) */
int mdp_mlp_eval(const unsigned short *packed, const float *params, const double *aff, int nin, const double *x,
                 int n, double *e, double *g, void *stream);
/* grad[nparam] = d/dparams sum_a (alpha_a E_a + u_a . g_a); partials: nwaves * nparam floats */
int mdp_mlp_train(const unsigned short *packed, const float *params, const double *aff, int nin, const double *x,
                  const double *u, const double *alpha, int n, float *partials, int nwaves, float *grad,
                  void *stream);
/* Adam step `step` (>= 1) with decoupled weight decay on the weight matrices */
int mdp_mlp_adam(float *params, const float *grad, float *m, float *v, float lr, float beta1, float beta2,
                 float eps, int step, float weight_decay, void *stream);

/* Descriptor chain rule (pairs sorted by centre atom; dbdr[K][3][P] = dB_k(i)/dr_ij, r_ij = x_j - x_i):
 * s[c][p] = sum_k g[k][pi[p]] dbdr[k][c][p]; F_a = sum over own pairs of s - sum over incoming pairs
 * (in_idx lists, per atom, the pairs whose neighbour is that atom); R = F - fref. */
int mdp_sur_pair_s(const double *g, const double *dbdr, const int *pi, int K, int n, long P, double *s,
                   void *stream);
int mdp_sur_atom_f(const double *s, const int *own_off, const int *in_off, const int *in_idx, long P, int n,
                   const double *fref, double *F, double *R, void *stream);
/* U[k][a] = coef sum_{pairs p of a} (R_a - R_pj) . dbdr[k][:][p]  (adjoint of coef/2 sum |R|^2), K <= 16 */
int mdp_sur_u(const double *R, const double *dbdr, const int *own_off, const int *pj, int K, int n, long P,
              double coef, double *U, void *stream);
/* Ec[c] = sum of e over configuration c (atoms cfg_off[c]..cfg_off[c+1]-1); alpha_a (may be NULL) =
 * coef 2 (Ec - eref_c) / N_c^2 */
int mdp_sur_energy(const double *e, const int *cfg_off, const double *eref, int C, double coef, double *Ec,
                   double *alpha, void *stream);
/* partials[b] (b < nblocks): block sums of v^2 */
int mdp_sumsq(const double *v, long n, double *partials, int nblocks, void *stream);

/* ---------------------------------------------------------------------------------------------- */
/* Part 5: backend-3 entry points of src/C++/Common/mdpcore.h (SNAP force and descriptor paths)   */
/* ---------------------------------------------------------------------------------------------- */
/* Same argument lists as the dispatch's `backend == 3` branches (without `backend`), dstype = double,
 * every pointer a device pointer.  Semantics of the upstream C++ core (src/C++/Core/cpuSnap4.cpp):
 * Ui and AddWself2Ui, Ei and Fi accumulate into their outputs, Yi zeroes and writes; the element of
 * atom i is map[type[i]] in Ei and Yi; Fi writes forces AoS and the virial SoA vatom[k*anum + atom]
 * with -r(x)dE/2 added to both atoms of each pair.  Small index and table arrays are read back to the
 * host once per call; the large arrays are used in place.  Errors: mdp_snap_last_error(). */
/* replaces hipSnapComputeUi in SnapComputeUi (mdpcore.h:3093-3116) */
void mdp_hipSnapComputeUi(double *Stotr, double *Stoti, double *rootpqarray, double *rij, double *wjelem,
                          double *radelem, double rmin0, double rfac0, double rcutfac, int *idxu_block, int *map,
                          int *ai, int *ti, int *tj, int twojmax, int idxu_max, int inum, int ijnum, int switchflag,
                          int chemflag);
/* replaces hipAddWself2Ui in AddWself2Ui (mdpcore.h:3118-3140) */
void mdp_hipAddWself2Ui(double *ulisttot_r, double *ulisttot_i, double wself, int *idxu_block, int *type, int *map,
                        int *ai, int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum);
/* replaces hipSnapComputeEi in SnapComputeEi (mdpcore.h:3167-3194) */
void mdp_hipSnapComputeEi(double *eatom, double *ulisttot_r, double *ulisttot_i, double *cglist, double *bzero,
                          double *coeffelem, int *ilist, int *map, int *type, int *idxb, int *idxcg_block,
                          int *idxu_block, int twojmax, int idxb_max, int idxu_max, int nelements, int ncoeffall,
                          int bnorm_flag, int bzero_flag, int wselfall_flag, int quadraticflag, int inum);
/* replaces hipSnapComputeYi in the coefficient form of SnapComputeYi (mdpcore.h:3221-3244) */
void mdp_hipSnapComputeYi(double *ylist_r, double *ylist_i, double *ulisttot_r, double *ulisttot_i, double *cglist,
                          double *betalist, int *map, int *type, int *idxz, int *idxb_block, int *idxu_block,
                          int *idxcg_block, int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements,
                          int ncoeffall, int bnorm_flag, int inum);
/* replaces hipSnapComputeYi in the per-atom-weight form of SnapComputeYi (mdpcore.h:3196-3219;
   src/C++/Core/cpuSnap4.cpp:601-690): betalist[(elemblock * idxb_max + jjb) * inum + atom] */
void mdp_hipSnapComputeYiPerAtom(double *ylist_r, double *ylist_i, double *ulisttot_r, double *ulisttot_i,
                                 double *cglist, double *betalist, int *idxz, int *idxb_block, int *idxu_block,
                                 int *idxcg_block, int twojmax, int idxb_max, int idxu_max, int idxz_max,
                                 int nelements, int bnorm_flag, int inum);
/* replaces hipSnapComputeFi in SnapComputeFi (mdpcore.h:3246-3269) */
void mdp_hipSnapComputeFi(double *fatom, double *vatom, double *ylist_r, double *ylist_i, double *rootpqarray,
                          double *rij, double *wjelem, double *radelem, double rmin0, double rfac0, double rcutfac,
                          int *idxu_block, int *map, int *aii, int *ai, int *aj, int *ti, int *tj, int twojmax,
                          int idxu_max, int inum, int anum, int ijnum, int switchflag, int chemflag);

/* The descriptor chain of the upstream fitting driver ComputeSnap (snappotential.cpp:762-910):
 * NeighborPairList -> Cumsum -> NeighborPairs -> ComputeSij -> ZeroUarraytot2 -> AddUarraytot ->
 * ComputeZi2 -> ComputeBi2 -> ComputeDbidrj -> SnapTallyBispectrum{,Deriv,Virial}, same conventions
 * (src/C++/Core/cpuSnap.cpp:1883-1935, src/C++/Core/cpuSnap2.cpp).  Neighbour and pair lists use the
 * fixed stride jnum: neighbours of atom i at neighlist[l + jnum*i] (l < neighnum[i]), kept pairs of
 * the ii-th listed atom at pairlist[l + jnum*ii] (l < pairnum[ii]).  ComputeSij, ZeroUarraytot2,
 * ComputeZi2, ComputeBi2, ComputeDbidrj and the three tallies overwrite their outputs; AddUarraytot
 * accumulates. */
/* replaces hipNeighPairList in the typed-cutoff NeighborPairList (mdpcore.h:3830-3851): pair kept when
 * 1e-20 < |x_j - x_i|^2 < rcutsq[atomtype[alist[j]] + atomtype[i]*(ntypes+1)] */
void mdp_hipNeighPairList(int *pairnum, int *pairlist, double *x, double *rcutsq, int *ilist, int *neighlist,
                          int *neighnum, int *atomtype, int *alist, int inum, int jnum, int dim, int ntypes);
/* replaces hipNeighPairList in the one-cutoff NeighborPairList (mdpcore.h:3807-3828): pair kept when
 * |x_j - x_i|^2 <= rcutsq */
void mdp_hipNeighPairListCut(int *pairnum, int *pairlist, double *x, double rcutsq, int *ilist, int *neighlist,
                             int *neighnum, int inum, int jnum, int dim);
/* replaces hipCumsum in Cumsum (mdpcore.h:1114-1131): d_out[0] = 0, d_out[j] = sum of d_in[0..j-1] for
 * j < length; d_sums and d_incr (caller scratch upstream) are not used */
void mdp_hipCumsum(int *d_out, int *d_in, int *d_sums, int *d_incr, int length);
/* replaces hipNeighPairs in NeighborPairs (mdpcore.h:3853-3875): per pair k = pairnumsum[ii] + l,
 * xij = x_j - x_i, aii = ii, ai = i, aj = alist[j], ti = atomtype[i], tj = atomtype[alist[j]] */
void mdp_hipNeighPairs(double *xij, double *x, int *aii, int *ai, int *aj, int *ti, int *tj, int *pairnum,
                       int *pairlist, int *pairnumsum, int *ilist, int *atomtype, int *alist, int inum, int jnum,
                       int dim);
/* replaces hipComputeSij in ComputeSij (mdpcore.h:2837-2861) */
void mdp_hipComputeSij(double *Sr, double *Si, double *Srx, double *Six, double *Sry, double *Siy, double *Srz,
                       double *Siz, double *rootpqarray, double *rij, double *wjelem, double *radelem, double rmin0,
                       double rfac0, double rcutfac, int *idxu_block, int *ti, int *tj, int twojmax, int idxu_max,
                       int ijnum, int switchflag);
/* replaces hipZeroUarraytot2 in ZeroUarraytot2 (mdpcore.h:2887-2909) */
void mdp_hipZeroUarraytot2(double *ulisttot_r, double *ulisttot_i, double wself, int *idxu_block, int *type, int *map,
                           int *ai, int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax,
                           int inum);
/* replaces hipAddUarraytot in the (map, ai, tj) form of AddUarraytot (mdpcore.h:2964-2986) */
void mdp_hipAddUarraytot(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i, int *map, int *ai,
                         int *tj, int idxu_max, int inum, int ijnum, int chemflag);
/* replaces hipAddUarraytot in the (pairnum, pairnumsum) form of AddUarraytot (mdpcore.h:2940-2962):
 * the pairs of the ii-th atom are pairnumsum[ii] .. pairnumsum[ii] + pairnum[ii] - 1 */
void mdp_hipAddUarraytotPairs(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i, int *pairnum,
                              int *pairnumsum, int *map, int *tj, int idxu_max, int nelements, int inum, int ijnum,
                              int chemflag);
/* replaces hipComputeZi2 in ComputeZi2 (mdpcore.h:3011-3033) */
void mdp_hipComputeZi2(double *zlist_r, double *zlist_i, double *ulisttot_r, double *ulisttot_i, double *cglist,
                       int *idxz, int *idxu_block, int *idxcg_block, int twojmax, int idxu_max, int idxz_max,
                       int nelements, int bnorm_flag, int inum);
/* replaces hipComputeBi2 in ComputeBi2 (mdpcore.h:3064-3090) */
void mdp_hipComputeBi2(double *blist, double *zlist_r, double *zlist_i, double *ulisttot_r, double *ulisttot_i,
                       double *bzero, int *ilist, int *type, int *map, int *idxb, int *idxu_block, int *idxz_block,
                       int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements, int bzero_flag,
                       int wselfall_flag, int chemflag, int inum);
/* replaces hipComputeDbidrj in the (inum, ijnum) form of ComputeDbidrj (mdpcore.h:3395-3420) */
void mdp_hipComputeDbidrj(double *dblist, double *zlist_r, double *zlist_i, double *dulist_r, double *dulist_i,
                          int *idxb, int *idxu_block, int *idxz_block, int *map, int *ai, int *tj, int twojmax,
                          int idxb_max, int idxu_max, int idxz_max, int nelements, int bnorm_flag, int chemflag,
                          int inum, int ijnum);
/* replaces hipAddUarraytot in the rij form of AddUarraytot (mdpcore.h:2911-2938; src/C++/Core/cpuSnap.cpp:358-395):
   U pair-major, Utot atom-major, switching and element weights from rij */
void mdp_hipAddUarraytotRij(double *ulisttot_r, double *ulisttot_i, double *ulist_r, double *ulist_i, double *rij,
                            double *wjelem, double *radelem, double rmin0, double rcutfac, int *idxu_block, int *ilist,
                            int *type, int *pairnum, int *pairnumsum, int *map, int *tj, int twojmax, int idxu_max,
                            int nelements, int inum, int switch_flag, int chemflag);
/* replaces hipCellList3D in CellList3D (mdpcore.h:1332-1349; src/C++/Core/cpuNeighborList3D.cpp:334-356) */
void mdp_hipCellList3D(int *clist, double *x, double *eta1, double *eta2, double *eta3, double *s2rmap, int *nc,
                       int inum, int natom, int dim);
/* replaces hipCell2AtomList in Cell2AtomList (mdpcore.h:1351-1368; src/C++/Core/cpuNeighborList2D.cpp:298-308);
   dtemp is not used */
void mdp_hipCell2AtomList(int *c2alist, int *c2anumsum, int *c2anum, int *clist, int *dtemp, int natom, int ncell);
/* No upstream counterpart: mdp_hipCell2AtomList for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments and outputs;
   it calls no hipMalloc/hipFree: the sort's temporary storage is a fixed 64 MiB device array of the
   library (one per GPU, held until the process ends), enough for natom up to about 4 million; larger
   natom, ncell <= 0 or natom < 0 are rejected with an error before any work.  Cell ids outside
   [0, ncell) set an error after the call (the outputs are then unspecified).  Calls that overlap on one
   GPU share that array and race; serialize them.  dtemp is not used. */
void mdp_hipCell2AtomListNoAlloc(int *c2alist, int *c2anumsum, int *c2anum, int *clist, int *dtemp, int natom,
                                 int ncell);
/* replaces hipFullNeighborList23 in the cell-list form of FullNeighborList3D (mdpcore.h:1432-1453;
   src/C++/Core/cpuNeighborList3D.cpp:361-412) */
void mdp_hipFullNeighborList3DCell(int *neighlist, int *neighnum, double *x, double *rcutsq, int *alist, int *clist,
                                   int *c2alist, int *c2anumsum, int *nc, int inum, int jnum, int dim);
/* replaces hipFullNeighborList3D in the all-pairs form of FullNeighborList3D (mdpcore.h:1390-1408;
   src/C++/Core/cpuNeighborList3D.cpp:417-441) */
void mdp_hipFullNeighborList3D(int *neighlist, int *neighnum, double *x, double *rcutsq, int anum, int inum, int jnum,
                               int dim);
/* replaces hipComputeDbidrj in the (ai, aj, ti, tj) form of ComputeDbidrj (mdpcore.h:3366-3393;
   src/C++/Core/cpuSnap.cpp:1050-1266): Z atom-major, dU and dB pair-major; aj and ti are not read */
void mdp_hipComputeDbidrjAtomMajor(double *dblist, double *zlist_r, double *zlist_i, double *dulist_r,
                                   double *dulist_i, int *idxb, int *idxu_block, int *idxz_block, int *map, int *ai,
                                   int *aj, int *ti, int *tj, int twojmax, int idxb_max, int idxu_max, int idxz_max,
                                   int nelements, int bnorm_flag, int chemflag, int ijnum);
/* replaces hipSnapTallyBispectrum in SnapTallyBispectrum (mdpcore.h:3750-3767) */
void mdp_hipSnapTallyBispectrum(double *bi, double *bispectrum, int *ilist, int *type, int inum, int ncoeff,
                                int nperdim, int ntype, int quadraticflag);
/* replaces hipSnapTallyBispectrumDeriv in SnapTallyBispectrumDeriv (mdpcore.h:3769-3786) */
void mdp_hipSnapTallyBispectrumDeriv(double *db, double *bispectrum, double *dbdr, int *aii, int *ai, int *aj, int *ti,
                                     int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag);
/* replaces hipSnapTallyBispectrumVirial in SnapTallyBispectrumVirial (mdpcore.h:3788-3805); ai and aj
 * are read but unused upstream too */
void mdp_hipSnapTallyBispectrumVirial(double *bv, double *bispectrum, double *dbdr, double *rij, int *aii, int *ai,
                                      int *aj, int *ti, int inum, int ijnum, int ncoeff, int nperdim, int ntype,
                                      int quadraticflag);

/* Energy, force and virial tallies of the force path's split form (src/C++/Core/cpuSnap.cpp:1746-1831,
 * cpuSnap2.cpp:1265-1292): eatom[ii] += beta0 + beta . B (+ quadratic terms) of element
 * map[type[ilist[ii]]], fatom[ai] += f, fatom[aj] -= f, and the SoA virial vatom[c*inum + atom] +=
 * -r (x) f / 2 for both atoms.  The force and virial sums use atomics, so their order (last bits) can
 * vary from run to run. */
/* replaces hipSnapTallyEnergyFull in SnapTallyEnergyFull (mdpcore.h:3630-3651), B at [ii*ncoeff + k] */
void mdp_hipSnapTallyEnergyFull(double *eatom, double *bispectrum, double *coeffelem, int *ilist, int *map, int *type,
                                int inum, int ncoeff, int ncoeffall, int quadraticflag);
/* replaces hipSnapTallyEnergyFull2 in SnapTallyEnergyFull2 (mdpcore.h:3653-3674), B at [ii + inum*k] */
void mdp_hipSnapTallyEnergyFull2(double *eatom, double *bispectrum, double *coeffelem, int *ilist, int *map,
                                 int *type, int inum, int ncoeff, int ncoeffall, int quadraticflag);
/* replaces hipSnapTallyForceFull in SnapTallyForceFull (mdpcore.h:3676-3691) */
void mdp_hipSnapTallyForceFull(double *fatom, double *fij, int *ai, int *aj, int *alist, int ijnum);
/* replaces hipSnapTallyVirialFull in SnapTallyVirialFull (mdpcore.h:3712-3730) */
void mdp_hipSnapTallyVirialFull(double *vatom, double *fij, double *rij, int *ai, int *aj, int inum, int ijnum);

/* ---------------------------------------------------------------------------------------------- */
/* Part 6: reference pair potential (ZBL) of the SNAP application, backend-3 entry points          */
/* ---------------------------------------------------------------------------------------------- */
/* The empirical pair route of the upstream application (src/C++/Potentials/cpuPotentials.cpp:
 * 1587-1649 for a full neighbour list, then the force sign flip of :1837), device arrays in and out.
 * replaces gpuPairaGradient (src/C++/Common/mdpgpucore.h:373; the generated
 * src/C++/Potentials/cpuPairaGradient.cpp:1-52 on the CPU): u[p] and u_xij[3p..3p+2] = du/dxij of each
 * pair vector xij = x_j - x_i.  potnum must be 1, the ZBL potential of simulating/snap/ZBLpotential.jl,
 * with mu = [cut_inner, cut_outer, z] (snapTA06A.jl:62-66); dim must be 3; qi, qj, ti, tj, ai, aj, eta
 * and kappa are not read (the potential does not use them). */
void mdp_hipPairaGradient(double *u, double *u_xij, double *xij, double *qi, double *qj, int *ti, int *tj, int *ai,
                          int *aj, double *mu, double *eta, int *kappa, int dim, int ncq, int nmu, int neta,
                          int nkappa, int ng, int potnum);
/* replaces cpuFullNeighPairDecomposition (src/C++/Core/cpuEnergyForceTally.cpp:216-227) on the device:
 * e[ai] += eij/2, f[ai*dim + d] -= fij[p*dim + d] (accumulates) */
void mdp_hipFullNeighPairDecomposition(double *e, double *f, double *eij, double *fij, int *ai, int ijnum, int dim);
/* replaces cpuVirialPairTally (src/C++/Core/cpuVirialTally.cpp:34-57) on the device:
 * v[k*inum + ai] += factor * (rij (x) fij)_k, k = xx, yy, zz, xy, xz, yz (accumulates) */
void mdp_hipVirialPairTally(double *v, double *fij, double *rij, double factor, int *ai, int dim, int inum,
                            int ijnum);

/* ---------------------------------------------------------------------------------------------- */
/* Part 7: the upstream application's binary inputs                                                */
/* ---------------------------------------------------------------------------------------------- */
/* The files the Julia preprocessing writes for the C++ application (mdp(app), Preprocessing.jl:60-321:
 * <prefix>app.bin, and when present config.bin, lattice.bin, region.bin, domain.bin), read as the upstream
 * reader does (src/C++/Configuration/readdata.cpp:11-290).  Host code; no GPU is touched until
 * mdp_md_create_app.  Errors: return value (nonzero / -1) and mdp_snap_last_error(). */
typedef struct mdp_app mdp_app;   /* opaque */
int mdp_app_read(mdp_app **h, const char *prefix);
int mdp_app_destroy(mdp_app *h);
/* Values of one field, by the upstream name: app.bin's ndims, flag, bcs, pbc, boxoffset, atomnumbers,
 * atommasses, atomcharges, simparam, solparam, eta, kappa, muml, mu1a..mu4b, pot1a..pot4b, rcutsqml,
 * rcutsq2a..rcutsq4b, rcutsqmax, atom1b..atom4b, traininglist, validatelist, nveparam, nvtparam,
 * snaparam, snapelemradius, snapelemweight, snapcoeff, createvelocity, setatomtypefraction; config.bin's
 * cfg_ndims, natoms, a, b, c, e, t, x, q, v, f, we, wf, ws, stress, cfg_pbc, move, group.  Copies at most
 * cap values to out (may be NULL) and returns the field's length, -1 for an unknown name. */
long mdp_app_field(const mdp_app *h, const char *name, double *out, long cap);
/* flags4 = which of config.bin, lattice.bin, region.bin, domain.bin were present */
int mdp_app_has(const mdp_app *h, int *flags4);
/* The atoms the application starts from: the first configuration of config.bin, or else the lattice
 * points inside the region (readdata.cpp:490-551, 647-676; cpuLattice, cpuAtomLattice, cpuAtomAdd), the
 * same positions and 1-based types bit for bit, relative to the box origin; box3 = the orthogonal box
 * lengths.  Returns the atom count (x, type may be NULL to query it; cap = their capacity in atoms), or
 * -1 (tilted cells, units other than metal, setatomtypefraction). */
long mdp_app_atoms(const mdp_app *h, double *x, int *type, double *box3, long cap);
/* An MD engine (part 2) for the application: SNAP from snaparam/snapelemradius/snapelemweight/snapcoeff,
 * masses, dt (simparam[1]), neighbour skin (solparam[1]), rebuilds every neighevery steps, the ZBL pair
 * potential when pot2a = [1] (mu2a), NVE or Nose-Hoover NVT (nvtparam), and the initial velocities of
 * config.bin and createvelocity.  Run it with mdp_md_run for the application's ntimesteps (ndims[3]). */
int mdp_md_create_app(mdp_md **md, const mdp_app *h);

#ifdef __cplusplus
}
#endif
#endif
