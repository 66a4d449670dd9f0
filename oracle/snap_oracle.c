/*
 * CPU oracle for the SNAP path -- TEST INFRASTRUCTURE ONLY (never linked into the product).
 *
 * Each function restates, from the mathematics of the SNAP bispectrum (Thompson et al., J. Comput.
 * Phys. 285 (2015) 316; Wigner-U recursion and Clebsch-Gordan products as in Varshalovich,
 * Moskalev & Khersonskii, "Quantum Theory of Angular Momentum"), the upstream C entry point cited
 * above it (src/Julia/Potential/cpuSnap.cpp), including its array layouts, index conventions and
 * accumulate-vs-overwrite behaviour.  Floating-point expressions are evaluated in the same order as
 * the cited code (this file is compiled with -ffp-contract=off) so that, at twojmax <= 6, the outputs
 * can be compared with the upstream build bit for bit.
 *
 * Deliberate difference: the force recursion mirrors the middle row of each even layer for
 * ma <= j/2 only.  The upstream loop at cpuSnap.cpp:1286-1306 runs ma = 0..j, which reads and
 * writes outside its 11-element stack arrays for twojmax >= 7; its only in-bounds effect (restoring
 * the sign of the imaginary part of the derivative's centre element) is kept here.
 */
#include <math.h>
#include <string.h>
#include "snap_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define JMAX_ORACLE 16   /* recursion scratch bound; twojmax <= 14 is accepted */

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------------------------ */
/* a1: index spaces and tables                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* Coupled triples (j1 >= j2, j = |j1-j2| .. min(J, j1+j2) step 2) in canonical order.
 * cpuSnap.cpp:23-69 */
#define FOR_TRIPLES(J, j1, j2, j)                      \
    for (int j1 = 0; j1 <= (J); j1++)                  \
        for (int j2 = 0; j2 <= j1; j2++)               \
            for (int j = j1 - j2; j <= imin((J), j1 + j2); j += 2)

int idxbcount(int J) { int n = 0; FOR_TRIPLES(J, j1, j2, j) if (j >= j1) n++; return n; }
int idxcgcount(int J) { int n = 0; FOR_TRIPLES(J, j1, j2, j) n += (j1 + 1) * (j2 + 1); return n; }
int idxzcount(int J) { int n = 0; FOR_TRIPLES(J, j1, j2, j) n += (j / 2 + 1) * (j + 1); return n; }
int idxucount(int J) { int n = 0; for (int j = 0; j <= J; j++) n += (j + 1) * (j + 1); return n; }

/* Range of the first-factor index m1 (and second m2 = its partner) contributing to output m of the
 * coupling (j1, j2) -> j: m1 in [m1min, m1min+n-1], m2 runs down from m2max.  cpuSnap.cpp:155-163 */
static void cg_range(int m, int j1, int j2, int j, int *m1min, int *m2max, int *n)
{
    int lo = imax(0, (2 * m - j - j2 + j1) / 2);
    *m1min = lo;
    *m2max = (2 * m - j - (2 * lo - j1) + j2) / 2;
    *n = imin(j1, (2 * m - j + j2 + j1) / 2) - lo + 1;
}

/* cpuSnap.cpp:71-172 (cpuBuildIndexList) */
static void build_index_lists(int *idx_max, int *idxz, int *idxz_block, int *idxb, int *idxb_block,
                              int *idxu_block, int *idxcg_block, int J)
{
    const int jd = J + 1;
    int ncg = 0;
    FOR_TRIPLES(J, j1, j2, j) {
        idxcg_block[j + j2 * jd + j1 * jd * jd] = ncg;
        ncg += (j1 + 1) * (j2 + 1);
    }
    idx_max[0] = ncg;

    int nu = 0;
    for (int j = 0; j <= J; j++) { idxu_block[j] = nu; nu += (j + 1) * (j + 1); }
    idx_max[1] = nu;

    int nb = 0;
    FOR_TRIPLES(J, j1, j2, j) {
        if (j < j1) continue;
        idxb[3 * nb + 0] = j1; idxb[3 * nb + 1] = j2; idxb[3 * nb + 2] = j;
        idxb_block[j + j2 * jd + j1 * jd * jd] = nb;
        nb++;
    }
    idx_max[2] = nb;

    int nz = 0;
    FOR_TRIPLES(J, j1, j2, j) {
        idxz_block[j + j2 * jd + j1 * jd * jd] = nz;
        for (int mb = 0; 2 * mb <= j; mb++)
            for (int ma = 0; ma <= j; ma++) {
                int *e = idxz + 10 * nz;
                e[0] = j1; e[1] = j2; e[2] = j;
                cg_range(ma, j1, j2, j, &e[3], &e[4], &e[5]);
                cg_range(mb, j1, j2, j, &e[6], &e[7], &e[8]);
                e[9] = idxu_block[j] + (j + 1) * mb + ma;
                nz++;
            }
    }
    idx_max[3] = nz;
}

/* Racah's closed form of the Clebsch-Gordan coefficient <j1 m1, j2 m2 | j m> in the doubled-index
 * convention (aa2 = 2*m1 - j1, ...).  fac[] is the caller's factorial table.  cpuSnap.cpp:182-246 */
static double triangle_coeff(const double *fac, int j1, int j2, int j)
{
    return sqrt(fac[(j1 + j2 - j) / 2] * fac[(j1 - j2 + j) / 2] * fac[(-j1 + j2 + j) / 2]
                / fac[(j1 + j2 + j) / 2 + 1]);
}

static double clebsch_gordan(const double *fac, int j1, int j2, int j, int aa2, int bb2)
{
    int m = (aa2 + bb2 + j) / 2;
    if (m < 0 || m > j) return 0.0;
    int zlo = imax(0, imax(-(j - j2 + aa2) / 2, -(j - j1 - bb2) / 2));
    int zhi = imin((j1 + j2 - j) / 2, imin((j1 - aa2) / 2, (j2 + bb2) / 2));
    double s = 0.0;
    for (int z = zlo; z <= zhi; z++) {
        double sgn = (z % 2) ? -1.0 : 1.0;
        s += sgn / (fac[z] * fac[(j1 + j2 - j) / 2 - z] * fac[(j1 - aa2) / 2 - z] * fac[(j2 + bb2) / 2 - z]
                    * fac[(j - j2 + aa2) / 2 + z] * fac[(j - j1 - bb2) / 2 + z]);
    }
    int cc2 = 2 * m - j;
    double norm = sqrt(fac[(j1 + aa2) / 2] * fac[(j1 - aa2) / 2] * fac[(j2 + bb2) / 2] * fac[(j2 - bb2) / 2]
                       * fac[(j + cc2) / 2] * fac[(j - cc2) / 2] * (j + 1));
    return s * triangle_coeff(fac, j1, j2, j) * norm;
}

/* cpuSnap.cpp:248-256 (cpuInitSna = index lists + sqrt(p/q) table + CG table) */
void cpuInitSna(double *rootpqarray, double *cglist, double *factorial, int *idx_max, int *idxz,
                int *idxz_block, int *idxb, int *idxb_block, int *idxu_block, int *idxcg_block, int twojmax)
{
    const int J = twojmax, jd = J + 1;
    build_index_lists(idx_max, idxz, idxz_block, idxb, idxb_block, idxu_block, idxcg_block, J);
    for (int p = 1; p <= J; p++)
        for (int q = 1; q <= J; q++)
            rootpqarray[p * jd + q] = sqrt((double)p / q);
    int k = 0;
    FOR_TRIPLES(J, j1, j2, j)
        for (int m1 = 0; m1 <= j1; m1++)
            for (int m2 = 0; m2 <= j2; m2++)
                cglist[k++] = clebsch_gordan(factorial, j1, j2, j, 2 * m1 - j1, 2 * m2 - j2);
}

/* ------------------------------------------------------------------------------------------ */
/* a2: pair filter and pair arrays                                                             */
/* ------------------------------------------------------------------------------------------ */

/* cpuSnap.cpp:258-286.  neighnum is a CSR offset array indexed by atom; the per-type-pair cutoff
 * table is rcutsq[jtype + itype*(ntypes+1)] (1-based types); coincident points are dropped. */
void cpuNeighPairList(int *pairnum, int *pairlist, double *x, double *rcutsq, int *ilist, int *neighlist,
                      int *neighnum, int *atomtype, int *alist, int inum, int dim, int ntypes)
{
    int total = 0;
    for (int ii = 0; ii < inum; ii++) {
        const int i = ilist[ii], it = atomtype[i];
        const int beg = neighnum[i], cnt = neighnum[i + 1] - beg;
        int kept = 0;
        for (int l = 0; l < cnt; l++) {
            const int j = neighlist[beg + l];
            const int jt = atomtype[alist[j]];
            double dx = x[j * dim] - x[i * dim];
            double dy = x[j * dim + 1] - x[i * dim + 1];
            double dz = x[j * dim + 2] - x[i * dim + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < rcutsq[jt + it * (ntypes + 1)] && d2 > 1e-20)
                pairlist[total + kept++] = j;
        }
        pairnum[ii] = kept;
        total += kept;
    }
}

/* cpuSnap.cpp:289-310.  aj holds the owner of the (possibly ghost) neighbour. */
void cpuNeighPairs(double *xij, double *x, int *aii, int *ai, int *aj, int *ti, int *tj,
                   int *pairlist, int *pairnumsum, int *ilist, int *atomtype, int *alist, int inum, int dim)
{
    for (int ii = 0; ii < inum; ii++) {
        const int i = ilist[ii];
        for (int k = pairnumsum[ii]; k < pairnumsum[ii + 1]; k++) {
            const int j = pairlist[k];
            aii[k] = ii; ai[k] = i; aj[k] = alist[j];
            ti[k] = atomtype[i]; tj[k] = atomtype[alist[j]];
            for (int d = 0; d < dim; d++) xij[k * dim + d] = x[j * dim + d] - x[i * dim + d];
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Per-pair geometry shared by the U/dU recursions                                             */
/* ------------------------------------------------------------------------------------------ */

/* Switching function value and radial derivative on (rmin0, rc]; 1 and 0 outside, as upstream
 * (cpuSnap.cpp:327-342 and :1010-1031). */
static void switching(double r, double rc, double rmin0, int switchflag, double *sfac, double *dsfac)
{
    *sfac = 1.0; *dsfac = 0.0;
    if (switchflag == 1 && r > rmin0 && r <= rc) {
        double w = M_PI / (rc - rmin0);
        *sfac = 0.5 * (cos((r - rmin0) * w) + 1.0);
        *dsfac = -0.5 * sin((r - rmin0) * w) * w;
    }
}

/* Cayley-Klein parameters a = (z0 - i z)/r0, b = (y - i x)/r0 of the rotation that maps the
 * 4-sphere point (x, y, z, z0) to the pole, and their Cartesian derivatives.  r0inv = 1/|(r, z0)|. */
typedef struct {
    double ar, ai, br, bi;
    double dar[3], dai[3], dbr[3], dbi[3];
    double u[3];           /* unit vector */
    double sfac, dsfac;
} ck_t;

/* Geometry as evaluated in the fused force path, cpuSnap.cpp:996-1067 (theta0 grouping
 * ((r-rmin0)*rfac0*pi)/(rc-rmin0)). */
static void ck_force(ck_t *g, double x, double y, double z, double rc, double rmin0, double rfac0,
                     int switchflag, double wj)
{
    double rsq = x * x + y * y + z * z;
    double r = sqrt(rsq);
    double rinv = 1.0 / r;
    g->u[0] = x * rinv; g->u[1] = y * rinv; g->u[2] = z * rinv;
    double z0 = r / tan((r - rmin0) * rfac0 * M_PI / (rc - rmin0));
    double dz0dr = z0 / r - (r * rfac0 * M_PI / (rc - rmin0)) * (rsq + z0 * z0) / rsq;
    switching(r, rc, rmin0, switchflag, &g->sfac, &g->dsfac);
    g->sfac *= wj; g->dsfac *= wj;
    double r0inv = 1.0 / sqrt(r * r + z0 * z0);
    g->ar = r0inv * z0; g->ai = -r0inv * z; g->br = r0inv * y; g->bi = -r0inv * x;
    double dr0 = -pow(r0inv, 3.0) * (r + z0 * dz0dr);
    for (int k = 0; k < 3; k++) {
        double w = dr0 * g->u[k];
        double dz = dz0dr * g->u[k];
        g->dar[k] = dz * r0inv + z0 * w;
        g->dai[k] = -z * w;
        g->dbr[k] = y * w;
        g->dbi[k] = -x * w;
    }
    g->dai[2] += -r0inv;
    g->dbi[0] += -r0inv;
    g->dbr[1] += r0inv;
}

/* Geometry as evaluated in the per-pair descriptor path, cpuSnap.cpp:1417-1493 (theta0 grouping
 * (r-rmin0)*(rfac0*pi/(rc-rmin0))). */
static void ck_desc(ck_t *g, double x, double y, double z, double rc, double rmin0, double rfac0,
                    int switchflag, double wj)
{
    double rsq = x * x + y * y + z * z;
    double r = sqrt(rsq);
    double rinv = 1.0 / r;
    g->u[0] = x * rinv; g->u[1] = y * rinv; g->u[2] = z * rinv;
    double rscale0 = rfac0 * M_PI / (rc - rmin0);
    double z0 = r / tan((r - rmin0) * rscale0);
    double dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;
    switching(r, rc, rmin0, switchflag, &g->sfac, &g->dsfac);
    g->sfac *= wj; g->dsfac *= wj;
    double r0inv = 1.0 / sqrt(r * r + z0 * z0);
    g->ar = r0inv * z0; g->ai = -r0inv * z; g->br = r0inv * y; g->bi = -r0inv * x;
    double dr0 = -pow(r0inv, 3.0) * (r + z0 * dz0dr);
    double drv[3], dzv[3];
    for (int k = 0; k < 3; k++) { drv[k] = dr0 * g->u[k]; dzv[k] = dz0dr * g->u[k]; }
    for (int k = 0; k < 3; k++) { g->dar[k] = dzv[k] * r0inv + z0 * drv[k]; g->dai[k] = -z * drv[k]; }
    g->dai[2] += -r0inv;
    for (int k = 0; k < 3; k++) { g->dbr[k] = y * drv[k]; g->dbi[k] = -x * drv[k]; }
    g->dbi[0] += -r0inv;
    g->dbr[1] += r0inv;
}

/* ------------------------------------------------------------------------------------------ */
/* Row-chain half-U recursion (fused path).                                                    */
/*                                                                                            */
/* Row mb of layer j (elements ma = 0..j) follows from row mb of layer j-1:                    */
/*   U_j(ma,mb) = sqrt((j-ma)/(j-mb)) conj(a) U_{j-1}(ma,mb) - sqrt(ma/(j-mb)) conj(b) U_{j-1}(ma-1,mb) */
/* Row mb starts at layer 2mb from the mirror of row mb-1 at layer 2mb-1,                      */
/*   U_j(j-ma, j-mb) = (-1)^(ma-mb) conj U_j(ma,mb),                                           */
/* and the middle row of an even layer is completed by the same mirror.  Only rows 2mb <= j    */
/* are produced.  An optional derivative carries d/dx,y,z along (nd = 0 or 3).                 */
/* ------------------------------------------------------------------------------------------ */

typedef struct { double r, i; } cplx;

/* new = rp1*conj(a)*w - rp2*conj(b)*u, evaluated as cpuSnap.cpp:1123-1130 */
static inline void rec_mid(cplx *out, cplx *dout, const ck_t *g, double rp1, double rp2,
                           cplx w, const cplx *dw, cplx u, const cplx *du, int nd)
{
    out->r = rp1 * (g->ar * w.r + g->ai * w.i) - rp2 * (g->br * u.r + g->bi * u.i);
    out->i = rp1 * (g->ar * w.i - g->ai * w.r) - rp2 * (g->br * u.i - g->bi * u.r);
    for (int k = 0; k < nd; k++) {
        dout[k].r = rp1 * (g->dar[k] * w.r + g->dai[k] * w.i + g->ar * dw[k].r + g->ai * dw[k].i)
                  - rp2 * (g->dbr[k] * u.r + g->dbi[k] * u.i + g->br * du[k].r + g->bi * du[k].i);
        dout[k].i = rp1 * (g->dar[k] * w.i - g->dai[k] * w.r + g->ar * dw[k].i - g->ai * dw[k].r)
                  - rp2 * (g->dbr[k] * u.i - g->dbi[k] * u.r + g->br * du[k].i - g->bi * du[k].r);
    }
}
static inline void rec_first(cplx *out, cplx *dout, const ck_t *g, double rp1, cplx w, const cplx *dw, int nd)
{
    out->r = rp1 * (g->ar * w.r + g->ai * w.i);
    out->i = rp1 * (g->ar * w.i - g->ai * w.r);
    for (int k = 0; k < nd; k++) {
        dout[k].r = rp1 * (g->dar[k] * w.r + g->dai[k] * w.i + g->ar * dw[k].r + g->ai * dw[k].i);
        dout[k].i = rp1 * (g->dar[k] * w.i - g->dai[k] * w.r + g->ar * dw[k].i - g->ai * dw[k].r);
    }
}
static inline void rec_last(cplx *out, cplx *dout, const ck_t *g, double rp2, cplx u, const cplx *du, int nd)
{
    out->r = -rp2 * (g->br * u.r + g->bi * u.i);
    out->i = -rp2 * (g->br * u.i - g->bi * u.r);
    for (int k = 0; k < nd; k++) {
        dout[k].r = -rp2 * (g->dbr[k] * u.r + g->dbi[k] * u.i + g->br * du[k].r + g->bi * du[k].i);
        dout[k].i = -rp2 * (g->dbr[k] * u.i - g->dbi[k] * u.r + g->br * du[k].i - g->bi * du[k].r);
    }
}

/* Callback receiving one finished row (j, mb) of the half U (and derivative) of a pair. */
typedef void (*row_fn)(void *ctx, int j, int mb, const cplx *P, const cplx (*dP)[3]);

static void halfu_rowchain(const ck_t *g, const double *rootpq, int J, int nd, row_fn emit, void *ctx)
{
    const int jd = J + 1;
    cplx P[JMAX_ORACLE + 1], Q[JMAX_ORACLE + 1];
    cplx dP[JMAX_ORACLE + 1][3], dQ[JMAX_ORACLE + 1][3];
    P[0].r = 1.0; P[0].i = 0.0;
    for (int k = 0; k < 3; k++) { dP[0][k].r = 0.0; dP[0][k].i = 0.0; }
    emit(ctx, 0, 0, P, (const cplx (*)[3])dP);

    for (int mb = 0; 2 * mb <= J; mb++) {
        int jstart = (mb == 0) ? 1 : 2 * mb;
        if (mb > 0)
            for (int ma = 0; ma < 2 * mb; ma++) { P[ma] = Q[ma]; for (int k = 0; k < nd; k++) dP[ma][k] = dQ[ma][k]; }
        for (int j = jstart; j <= J; j++) {
            /* in-place update along the row; u/du keep the previous-layer value of P[ma-1] */
            cplx u = P[0], du[3];
            for (int k = 0; k < nd; k++) du[k] = dP[0][k];
            rec_first(&P[0], dP[0], g, rootpq[j * jd + (j - mb)], u, du, nd);
            for (int ma = 1; ma < j; ma++) {
                cplx w = P[ma], dw[3];
                for (int k = 0; k < nd; k++) dw[k] = dP[ma][k];
                rec_mid(&P[ma], dP[ma], g, rootpq[(j - ma) * jd + (j - mb)], rootpq[ma * jd + (j - mb)],
                        w, dw, u, du, nd);
                u = w;
                for (int k = 0; k < nd; k++) du[k] = dw[k];
            }
            rec_last(&P[j], dP[j], g, rootpq[j * jd + (j - mb)], u, du, nd);

            if (j == 2 * mb) {
                /* middle row: right half from the left half by the mirror symmetry.  The centre
                 * element's imaginary part is negated for U but not for dU (net upstream effect). */
                int s = 1;
                for (int ma = 0; ma <= j / 2; ma++) {
                    const int src = j / 2 - ma, dst = j / 2 + ma;
                    if (s == 1) { P[dst].r = P[src].r; P[dst].i = -P[src].i; }
                    else        { P[dst].r = -P[src].r; P[dst].i = P[src].i; }
                    if (ma > 0)
                        for (int k = 0; k < nd; k++) {
                            if (s == 1) { dP[dst][k].r = dP[src][k].r; dP[dst][k].i = -dP[src][k].i; }
                            else        { dP[dst][k].r = -dP[src][k].r; dP[dst][k].i = dP[src][k].i; }
                        }
                    s = -s;
                }
            }
            if (j == 2 * mb + 1) {
                /* seed of row mb+1: mirror of this row at layer 2mb+1 */
                int s = (((j + 1) / 2) % 2 == 0) ? -1 : 1;
                for (int ma = 0; ma <= j; ma++) {
                    if (s == 1) { Q[j - ma].r = P[ma].r; Q[j - ma].i = -P[ma].i; }
                    else        { Q[j - ma].r = -P[ma].r; Q[j - ma].i = P[ma].i; }
                    for (int k = 0; k < nd; k++) {
                        if (s == 1) { dQ[j - ma][k].r = dP[ma][k].r; dQ[j - ma][k].i = -dP[ma][k].i; }
                        else        { dQ[j - ma][k].r = -dP[ma][k].r; dQ[j - ma][k].i = dP[ma][k].i; }
                    }
                    s = -s;
                }
            }
            emit(ctx, j, mb, P, (const cplx (*)[3])dP);
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a3: compute_ui                                                                              */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    double *Ur, *Ui;
    long base, stride;   /* element (jju) of this atom lives at base + stride*jju */
    double sfac;
} ui_ctx;

static void ui_emit(void *vc, int j, int mb, const cplx *P, const cplx (*dP)[3])
{
    (void)dP;
    ui_ctx *c = (ui_ctx *)vc;
    if (j == 0) { c->Ur[c->base] += c->sfac; return; }
    int k = idxucount(j - 1) + (j + 1) * mb;
    for (int ma = 0; ma <= j; ma++, k++) {
        c->Ur[c->base + c->stride * k] += c->sfac * P[ma].r;
        c->Ui[c->base + c->stride * k] += c->sfac * P[ma].i;
    }
}

/* cpuSnap.cpp:312-490.  Accumulates sfac*wj*U (half rows) of every pair into Utot of atom aii. */
void cpuSnapComputeUi(double *Utotr, double *Utoti, double *rootpqarray, double *rij, double *wjelem,
                      double *radelem, double rmin0, double rfac0, double rcutfac, int *map, int *aii, int *ti,
                      int *tj, int twojmax, int idxu_max, int inum, int ijnum, int switchflag, int chemflag)
{
    for (int ij = 0; ij < ijnum; ij++) {
        double x = rij[3 * ij], y = rij[3 * ij + 1], z = rij[3 * ij + 2];
        double r = sqrt(x * x + y * y + z * z);
        double rc = (radelem[ti[ij]] + radelem[tj[ij]]) * rcutfac;
        double z0 = r / tan((r - rmin0) * rfac0 * M_PI / (rc - rmin0));
        double sfac, dsfac;
        switching(r, rc, rmin0, switchflag, &sfac, &dsfac);
        sfac *= wjelem[tj[ij]];
        ck_t g;
        double r0inv = 1.0 / sqrt(r * r + z0 * z0);
        g.ar = r0inv * z0; g.ai = -r0inv * z; g.br = r0inv * y; g.bi = -r0inv * x;
        ui_ctx c;
        c.Ur = Utotr; c.Ui = Utoti; c.stride = inum; c.sfac = sfac;
        c.base = aii[ij] + (chemflag ? (long)inum * idxu_max * map[tj[ij]] : 0);
        halfu_rowchain(&g, rootpqarray, twojmax, 0, ui_emit, &c);
    }
}

/* cpuSnap.cpp:492-542.  Adds wself on the diagonal of every layer and completes the rows
 * 2mb > j from the computed half by U(j-mb, j-ma) = (-1)^(ma+mb) conj U(mb, ma). */
void cpuAddWself2Ui(double *Utotr, double *Utoti, double wself, int *idxu_block, int *type, int *map,
                    int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax, int inum)
{
    for (int jelem = 0; jelem < nelements; jelem++)
        for (int j = 0; j <= twojmax; j++)
            for (int ii = 0; ii < inum; ii++) {
                const int ielem = chemflag ? map[type[ii]] : 0;
                const long base = ii + (long)inum * idxu_max * jelem;
                if (jelem == ielem || wselfall_flag)
                    for (int m = 0; m <= j; m++)
                        Utotr[base + (long)inum * (idxu_block[j] + (j + 1) * m + m)] += wself;
                int lo = idxu_block[j], hi = idxu_block[j] + (j + 1) * (j + 1) - 1;
                for (int mb = 0; 2 * mb < j; mb++)
                    for (int ma = 0; ma <= j; ma++, lo++, hi--) {
                        const long s = base + (long)inum * lo, d = base + (long)inum * hi;
                        if (((ma + mb) & 1) == 0) { Utotr[d] = Utotr[s]; Utoti[d] = -Utoti[s]; }
                        else                      { Utotr[d] = -Utotr[s]; Utoti[d] = Utoti[s]; }
                    }
            }
}

/* ------------------------------------------------------------------------------------------ */
/* a4: Z, B and energy                                                                         */
/* ------------------------------------------------------------------------------------------ */

/* One Z(j1,j2,j)(ma,mb) element of one atom, double CG sum in the unfactored order
 * ztmp += (cg_b*cg_a)*(u1*u2) (cpuSnap.cpp:587-616).  u1/u2 point at element 0 of the atom's
 * U array for elem1/elem2, s is the atom stride. */
static void z_unfactored(const double *cg, const int *idxu_block, int j1, int j2, int j, int ma, int mb,
                         const double *u1r, const double *u1i, const double *u2r, const double *u2i, long s,
                         int bnorm, double *zr, double *zi)
{
    int ma1min, ma2max, na, mb1min, mb2max, nb;
    cg_range(ma, j1, j2, j, &ma1min, &ma2max, &na);
    cg_range(mb, j1, j2, j, &mb1min, &mb2max, &nb);
    int jju1 = idxu_block[j1] + (j1 + 1) * mb1min;
    int jju2 = idxu_block[j2] + (j2 + 1) * mb2max;
    int icgb = mb1min * (j2 + 1) + mb2max;
    double tr = 0.0, ti = 0.0;
    for (int ib = 0; ib < nb; ib++) {
        int ma1 = ma1min, ma2 = ma2max, icga = ma1min * (j2 + 1) + ma2max;
        for (int ia = 0; ia < na; ia++) {
            const double ar = u1r[s * (jju1 + ma1)], ai = u1i[s * (jju1 + ma1)];
            const double br = u2r[s * (jju2 + ma2)], bi = u2i[s * (jju2 + ma2)];
            tr += cg[icgb] * cg[icga] * (ar * br - ai * bi);
            ti += cg[icgb] * cg[icga] * (ar * bi + ai * br);
            ma1++; ma2--; icga += j2;
        }
        jju1 += j1 + 1; jju2 -= j2 + 1; icgb += j2;
    }
    if (bnorm) { ti /= j + 1; tr /= j + 1; }
    *zr = tr; *zi = ti;
}

/* Half-domain contraction 2 * sum' U*(j) . Z(j1,j2,j) with the middle column of an even layer
 * weighted 1/2 (cpuSnap.cpp:583-697). */
static double bsum_unfactored(const double *cg, const int *idxu_block, int j1, int j2, int j,
                              const double *u1r, const double *u1i, const double *u2r, const double *u2i,
                              const double *u3r, const double *u3i, long s, int bnorm)
{
    int jju = idxu_block[j];
    double sum = 0.0, zr, zi;
    for (int mb = 0; 2 * mb < j; mb++)
        for (int ma = 0; ma <= j; ma++, jju++) {
            z_unfactored(cg, idxu_block, j1, j2, j, ma, mb, u1r, u1i, u2r, u2i, s, bnorm, &zr, &zi);
            sum += u3r[s * jju] * zr + u3i[s * jju] * zi;
        }
    if (j % 2 == 0) {
        const int mb = j / 2;
        for (int ma = 0; ma < mb; ma++, jju++) {
            z_unfactored(cg, idxu_block, j1, j2, j, ma, mb, u1r, u1i, u2r, u2i, s, bnorm, &zr, &zi);
            sum += u3r[s * jju] * zr + u3i[s * jju] * zi;
        }
        z_unfactored(cg, idxu_block, j1, j2, j, mb, mb, u1r, u1i, u2r, u2i, s, bnorm, &zr, &zi);
        sum += 0.5 * (u3r[s * jju] * zr + u3i[s * jju] * zi);
    }
    return sum * 2.0;
}

/* cpuSnap.cpp:544-719.  eatom[ilist[ii]] += beta0 + sum_k beta_k B_k; accumulation order per atom
 * is (element triple, jjb). */
void cpuSnapComputeEi(double *eatom, double *Utotr, double *Utoti, double *cglist, double *bzero,
                      double *coeffelem, int *ilist, int *map, int *type, int *idxb, int *idxcg_block,
                      int *idxu_block, int twojmax, int idxb_max, int idxu_max, int nelements, int ncoeffall,
                      int bnorm_flag, int bzero_flag, int wselfall_flag, int chemflag, int inum)
{
    const int jd = twojmax + 1, ne = nelements;
    const long nu = (long)idxu_max * inum;
    for (int ii = 0; ii < inum; ii++) {
        const int ielem = chemflag ? map[type[ii]] : 0;
        for (int e1 = 0; e1 < ne; e1++) for (int e2 = 0; e2 < ne; e2++) for (int e3 = 0; e3 < ne; e3++)
            for (int jjb = 0; jjb < idxb_max; jjb++) {
                const int jelem = (e1 * ne + e2) * ne + e3;
                const int j1 = idxb[3 * jjb], j2 = idxb[3 * jjb + 1], j = idxb[3 * jjb + 2];
                const double *cg = cglist + idxcg_block[j + j2 * jd + j1 * jd * jd];
                double b = bsum_unfactored(cg, idxu_block, j1, j2, j,
                                           Utotr + ii + nu * e1, Utoti + ii + nu * e1,
                                           Utotr + ii + nu * e2, Utoti + ii + nu * e2,
                                           Utotr + ii + nu * e3, Utoti + ii + nu * e3, inum, bnorm_flag);
                if (bzero_flag && (wselfall_flag || (e1 == e2 && e1 == e3))) b -= bzero[j];
                const int icoeff = jjb + idxb_max * jelem;
                if (icoeff == 0) eatom[ilist[ii]] += coeffelem[ielem * ncoeffall];
                eatom[ilist[ii]] += coeffelem[1 + icoeff + ielem * ncoeffall] * b;
            }
    }
}

/* cpuSnap.cpp:721-889.  blist[ii + inum*jjb + inum*idxb_max*jelem] = B (overwrite). */
void cpuSnapComputeBi(double *blist, double *Utotr, double *Utoti, double *cglist, double *bzero,
                      int *idxb, int *idxcg_block, int *idxu_block, int *idxz_block, int twojmax, int idxb_max,
                      int idxu_max, int idxz_max, int nelements, int bnorm_flag, int bzero_flag,
                      int wselfall_flag, int inum)
{
    (void)idxz_block; (void)idxz_max;
    const int jd = twojmax + 1, ne = nelements;
    const long nu = (long)idxu_max * inum;
    for (int e1 = 0; e1 < ne; e1++) for (int e2 = 0; e2 < ne; e2++) for (int e3 = 0; e3 < ne; e3++)
        for (int jjb = 0; jjb < idxb_max; jjb++)
            for (int ii = 0; ii < inum; ii++) {
                const int jelem = (e1 * ne + e2) * ne + e3;
                const int j1 = idxb[3 * jjb], j2 = idxb[3 * jjb + 1], j = idxb[3 * jjb + 2];
                const double *cg = cglist + idxcg_block[j + j2 * jd + j1 * jd * jd];
                double b = bsum_unfactored(cg, idxu_block, j1, j2, j,
                                           Utotr + ii + nu * e1, Utoti + ii + nu * e1,
                                           Utotr + ii + nu * e2, Utoti + ii + nu * e2,
                                           Utotr + ii + nu * e3, Utoti + ii + nu * e3, inum, bnorm_flag);
                if (bzero_flag && (wselfall_flag || (e1 == e2 && e1 == e3))) b -= bzero[j];
                blist[ii + (long)inum * jjb + (long)inum * idxb_max * jelem] = b;
            }
}

/* ------------------------------------------------------------------------------------------ */
/* a5: compute_yi                                                                              */
/* ------------------------------------------------------------------------------------------ */

/* Weight with which Z(j1,j2,j) enters Y(j): the coefficient of the bispectrum component that
 * contains this coupling, times its multiplicity among the (j1,j2,j) orderings of the same
 * component, times (j1+1)/(j+1) when the stored order is not j >= j1 (cpuSnap.cpp:959-981). */
static double y_weight(const double *beta, const int *idxb_block, int jd, int j1, int j2, int j,
                       int e1, int e2, int e3, int ne, int idxb_max, int ielem, int ncoeffall, int bnorm)
{
    double w;
    if (j >= j1) {
        int jjb = idxb_block[j + j2 * jd + j1 * jd * jd];
        int it = ((e1 * ne + e2) * ne + e3) * idxb_max + jjb + 1 + ielem * ncoeffall;
        if (j1 == j) w = (j2 == j) ? 3 * beta[it] : 2 * beta[it];
        else w = beta[it];
    } else if (j >= j2) {
        int jjb = idxb_block[j1 + j2 * jd + j * jd * jd];
        int it = ((e3 * ne + e2) * ne + e1) * idxb_max + jjb + 1 + ielem * ncoeffall;
        w = (j2 == j) ? 2 * beta[it] : beta[it];
    } else {
        int jjb = idxb_block[j1 + j * jd + j2 * jd * jd];
        int it = ((e2 * ne + e3) * ne + e1) * idxb_max + jjb + 1 + ielem * ncoeffall;
        w = beta[it];
    }
    if (!bnorm && j1 > j) w *= (j1 + 1) / (j + 1.0);
    return w;
}

/* cpuSnap.cpp:891-989.  Y (zeroed here) = sum over jjz of weight * Z(jjz), scattered to jju. */
void cpuSnapComputeYi(double *ylist_r, double *ylist_i, double *Utotr, double *Utoti, double *cglist,
                      double *beta, int *map, int *type, int *idxz, int *idxb_block, int *idxu_block,
                      int *idxcg_block, int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements,
                      int ncoeffall, int bnorm_flag, int chemflag, int inum)
{
    const int jd = twojmax + 1, ne = nelements;
    const long nu = (long)idxu_max * inum;
    memset(ylist_r, 0, sizeof(double) * nu * ne);
    memset(ylist_i, 0, sizeof(double) * nu * ne);
    for (int ii = 0; ii < inum; ii++) {
        const int ielem = chemflag ? map[type[ii]] : 0;
        for (int jjz = 0; jjz < idxz_max; jjz++) {
            const int *e = idxz + 10 * jjz;
            const int j1 = e[0], j2 = e[1], j = e[2], jju = e[9];
            const double *cg = cglist + idxcg_block[j + j2 * jd + j1 * jd * jd];
            /* recover (ma, mb) of this entry from jju */
            const int off = jju - idxu_block[j], mb = off / (j + 1), ma = off % (j + 1);
            for (int e1 = 0; e1 < ne; e1++) for (int e2 = 0; e2 < ne; e2++) {
                double zr, zi;
                z_unfactored(cg, idxu_block, j1, j2, j, ma, mb,
                             Utotr + ii + nu * e1, Utoti + ii + nu * e1,
                             Utotr + ii + nu * e2, Utoti + ii + nu * e2, inum, bnorm_flag, &zr, &zi);
                for (int e3 = 0; e3 < ne; e3++) {
                    double w = y_weight(beta, idxb_block, jd, j1, j2, j, e1, e2, e3, ne, idxb_max, ielem,
                                        ncoeffall, bnorm_flag);
                    ylist_r[ii + (long)inum * jju + nu * e3] += w * zr;
                    ylist_i[ii + (long)inum * jju + nu * e3] += w * zi;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a6: compute_duidrj / compute_deidrj fused with the force and virial tally                   */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    const double *Yr, *Yi;
    long base, stride;
    const ck_t *g;
    double de[3];
} fi_ctx;

static void fi_emit(void *vc, int j, int mb, const cplx *P, const cplx (*dP)[3])
{
    fi_ctx *c = (fi_ctx *)vc;
    const ck_t *g = c->g;
    if (j == 0) {
        double y0 = 0.5 * c->Yr[c->base];
        for (int k = 0; k < 3; k++) c->de[k] = (g->dsfac * g->u[k]) * y0;
        return;
    }
    int kk = idxucount(j - 1) + (j + 1) * mb;
    int last = (j == 2 * mb) ? mb : j;
    for (int ma = 0; ma <= last; ma++, kk++) {
        double yr = c->Yr[c->base + c->stride * kk], yi = c->Yi[c->base + c->stride * kk];
        if (j == 2 * mb && ma == mb) { yr = 0.5 * yr; yi = 0.5 * yi; }
        for (int k = 0; k < 3; k++)
            c->de[k] += (g->dsfac * P[ma].r * g->u[k] + g->sfac * dP[ma][k].r) * yr
                      + (g->dsfac * P[ma].i * g->u[k] + g->sfac * dP[ma][k].i) * yi;
    }
}

/* cpuSnap.cpp:991-1410.  dE/dr_ij = 2 * sum' (d(sfac U)/dr . Y_i); fatom[ai] += dE, fatom[aj] -= dE;
 * vatom[6*ai + (xx,yy,zz,xy,xz,yz)] += -(r_a dE_b) -- the whole pair virial on atom i (AoS). */
void cpuSnapComputeFi(double *fatom, double *vatom, double *ylist_r, double *ylist_i, double *rootpqarray,
                      double *rij, double *wjelem, double *radelem, double rmin0, double rfac0, double rcutfac,
                      int *map, int *aii, int *ai, int *aj, int *ti, int *tj, int twojmax, int idxu_max, int inum,
                      int anum, int ijnum, int switchflag, int chemflag)
{
    (void)anum;
    for (int ij = 0; ij < ijnum; ij++) {
        double x = rij[3 * ij], y = rij[3 * ij + 1], z = rij[3 * ij + 2];
        double rc = (radelem[ti[ij]] + radelem[tj[ij]]) * rcutfac;
        ck_t g;
        ck_force(&g, x, y, z, rc, rmin0, rfac0, switchflag, wjelem[tj[ij]]);
        fi_ctx c;
        c.Yr = ylist_r; c.Yi = ylist_i; c.stride = inum; c.g = &g;
        c.base = aii[ij] + (chemflag ? (long)inum * idxu_max * map[tj[ij]] : 0);
        halfu_rowchain(&g, rootpqarray, twojmax, 3, fi_emit, &c);
        double dx = 2.0 * c.de[0], dy = 2.0 * c.de[1], dz = 2.0 * c.de[2];
        const int a = ai[ij], b = aj[ij];
        fatom[3 * a] += dx; fatom[3 * a + 1] += dy; fatom[3 * a + 2] += dz;
        fatom[3 * b] -= dx; fatom[3 * b + 1] -= dy; fatom[3 * b + 2] -= dz;
        vatom[6 * a + 0] += -1.0 * x * dx;
        vatom[6 * a + 1] += -1.0 * y * dy;
        vatom[6 * a + 2] += -1.0 * z * dz;
        vatom[6 * a + 3] += -1.0 * x * dy;
        vatom[6 * a + 4] += -1.0 * x * dz;
        vatom[6 * a + 5] += -1.0 * y * dz;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a3/a7: per-pair U and dU/dr, layer-by-layer recursion (descriptor path)                     */
/* ------------------------------------------------------------------------------------------ */

/* One pair of cpuComputeSij / cpuComputeUij (cpuSnap.cpp:1412-1618, 1620-1833): full layers are
 * built from the previous layer for rows 2mb <= j, the remaining rows follow by the mirror
 * symmetry (derivatives mirrored only for the row that the next layer's recursion reads), then the
 * derivative rows 2mb <= j become d(sfac U) and U is scaled by sfac.  Element jju of this pair
 * lives at index ij + ijnum*jju of each array. */
static void uij_pair(double *Sr, double *Si, double *dSr[3], double *dSi[3], const ck_t *g,
                     const double *rootpq, const int *idxu_block, int J, int idxu_max, long ij, long np)
{
    const int jd = J + 1;
#define AT(arr, jju) arr[ij + np * (long)(jju)]
    AT(Sr, 0) = 1.0; AT(Si, 0) = 0.0;
    for (int k = 0; k < 3; k++) { AT(dSr[k], 0) = 0.0; AT(dSi[k], 0) = 0.0; }
    for (int j = 1; j <= J; j++) {
        int jju = idxu_block[j], jjup = idxu_block[j - 1];
        for (int mb = 0; 2 * mb <= j; mb++) {
            AT(Sr, jju) = 0.0; AT(Si, jju) = 0.0;
            for (int k = 0; k < 3; k++) { AT(dSr[k], jju) = 0.0; AT(dSi[k], jju) = 0.0; }
            for (int ma = 0; ma < j; ma++) {
                double rp = rootpq[(j - ma) * jd + (j - mb)];
                double ur = AT(Sr, jjup), ui = AT(Si, jjup);
                double dur[3], dui[3];
                for (int k = 0; k < 3; k++) { dur[k] = AT(dSr[k], jjup); dui[k] = AT(dSi[k], jjup); }
                AT(Sr, jju) += rp * (g->ar * ur + g->ai * ui);
                AT(Si, jju) += rp * (g->ar * ui - g->ai * ur);
                for (int k = 0; k < 3; k++) {
                    AT(dSr[k], jju) += rp * (g->dar[k] * ur + g->dai[k] * ui + g->ar * dur[k] + g->ai * dui[k]);
                    AT(dSi[k], jju) += rp * (g->dar[k] * ui - g->dai[k] * ur + g->ar * dui[k] - g->ai * dur[k]);
                }
                rp = rootpq[(ma + 1) * jd + (j - mb)];
                AT(Sr, jju + 1) = -rp * (g->br * ur + g->bi * ui);
                AT(Si, jju + 1) = -rp * (g->br * ui - g->bi * ur);
                for (int k = 0; k < 3; k++) {
                    AT(dSr[k], jju + 1) = -rp * (g->dbr[k] * ur + g->dbi[k] * ui + g->br * dur[k] + g->bi * dui[k]);
                    AT(dSi[k], jju + 1) = -rp * (g->dbr[k] * ui - g->dbi[k] * ur + g->br * dui[k] - g->bi * dur[k]);
                }
                jju++; jjup++;
            }
            jju++;
        }
        /* mirror: element (mb, ma) -> (j-mb, j-ma), walking forward and backward from both ends */
        int lo = idxu_block[j], hi = lo + (j + 1) * (j + 1) - 1;
        for (int mb = 0; 2 * mb <= j; mb++) {
            const int with_d = (j % 2 == 1 && mb == j / 2);
            for (int ma = 0; ma <= j; ma++, lo++, hi--) {
                const int even = ((ma + mb) & 1) == 0;
                if (even) { AT(Sr, hi) = AT(Sr, lo); AT(Si, hi) = -AT(Si, lo); }
                else      { AT(Sr, hi) = -AT(Sr, lo); AT(Si, hi) = AT(Si, lo); }
                if (with_d)
                    for (int k = 0; k < 3; k++) {
                        if (even) { AT(dSr[k], hi) = AT(dSr[k], lo); AT(dSi[k], hi) = -AT(dSi[k], lo); }
                        else      { AT(dSr[k], hi) = -AT(dSr[k], lo); AT(dSi[k], hi) = AT(dSi[k], lo); }
                    }
            }
        }
    }
    for (int j = 0; j <= J; j++) {
        int jju = idxu_block[j];
        for (int mb = 0; 2 * mb <= j; mb++)
            for (int ma = 0; ma <= j; ma++, jju++)
                for (int k = 0; k < 3; k++) {
                    AT(dSr[k], jju) = g->dsfac * AT(Sr, jju) * g->u[k] + g->sfac * AT(dSr[k], jju);
                    AT(dSi[k], jju) = g->dsfac * AT(Si, jju) * g->u[k] + g->sfac * AT(dSi[k], jju);
                }
    }
    for (int k = 0; k < idxu_max; k++) { AT(Sr, k) = g->sfac * AT(Sr, k); AT(Si, k) = g->sfac * AT(Si, k); }
#undef AT
}

/* cpuSnap.cpp:1412-1618: separate x/y/z derivative arrays, each [idxu_max*ijnum]. */
void cpuComputeSij(double *Sr, double *Si, double *Srx, double *Six, double *Sry, double *Siy, double *Srz,
                   double *Siz, double *rootpqarray, double *rij, double *wjelem, double *radelem, double rmin0,
                   double rfac0, double rcutfac, int *idxu_block, int *ti, int *tj, int twojmax, int idxu_max,
                   int ijnum, int switch_flag)
{
    double *dr[3] = {Srx, Sry, Srz}, *di[3] = {Six, Siy, Siz};
    for (int ij = 0; ij < ijnum; ij++) {
        ck_t g;
        double rc = (radelem[ti[ij]] + radelem[tj[ij]]) * rcutfac;
        ck_desc(&g, rij[3 * ij], rij[3 * ij + 1], rij[3 * ij + 2], rc, rmin0, rfac0, switch_flag, wjelem[tj[ij]]);
        uij_pair(Sr, Si, dr, di, &g, rootpqarray, idxu_block, twojmax, idxu_max, ij, ijnum);
    }
}

/* cpuSnap.cpp:1620-1833: derivative arrays packed as dS[ij + ijnum*jju + ijnum*idxu_max*k]. */
void cpuComputeUij(double *Sr, double *Si, double *dSr, double *dSi, double *rootpqarray, double *rij,
                   double *wjelem, double *radelem, double rmin0, double rfac0, double rcutfac, int *idxu_block,
                   int *ti, int *tj, int twojmax, int idxu_max, int ijnum, int switch_flag)
{
    const long n = (long)idxu_max * ijnum;
    double *dr[3] = {dSr, dSr + n, dSr + 2 * n}, *di[3] = {dSi, dSi + n, dSi + 2 * n};
    for (int ij = 0; ij < ijnum; ij++) {
        ck_t g;
        double rc = (radelem[ti[ij]] + radelem[tj[ij]]) * rcutfac;
        ck_desc(&g, rij[3 * ij], rij[3 * ij + 1], rij[3 * ij + 2], rc, rmin0, rfac0, switch_flag, wjelem[tj[ij]]);
        uij_pair(Sr, Si, dr, di, &g, rootpqarray, idxu_block, twojmax, idxu_max, ij, ijnum);
    }
}

/* cpuSnap.cpp:1835-1863.  Stot = wself on the diagonal of the own-element block, 0 elsewhere.
 * (The element of atom ii is read through ai[ii], as upstream.) */
void cpuZeroUarraytot2(double *Stotr, double *Stoti, double wself, int *idxu_block, int *type, int *map,
                       int *ai, int wselfall_flag, int chemflag, int idxu_max, int nelements, int twojmax,
                       int inum)
{
    for (int jelem = 0; jelem < nelements; jelem++)
        for (int j = 0; j <= twojmax; j++)
            for (int ii = 0; ii < inum; ii++) {
                const int ielem = chemflag ? map[type[ai[ii]]] : 0;
                int jju = idxu_block[j];
                for (int mb = 0; mb <= j; mb++)
                    for (int ma = 0; ma <= j; ma++, jju++) {
                        const long n = ii + (long)inum * jju + (long)inum * idxu_max * jelem;
                        Stotr[n] = (ma == mb && (jelem == ielem || wselfall_flag)) ? wself : 0.0;
                        Stoti[n] = 0.0;
                    }
            }
}

/* cpuSnap.cpp:1865-1897.  Stot[ai] += S, per element (jju) in ascending pair order. */
void cpuAddUarraytot(double *Stotr, double *Stoti, double *Sr, double *Si, int *map, int *ai, int *tj,
                     int idxu_max, int inum, int ijnum, int chemflag)
{
    const long n1 = (long)inum * idxu_max;
    for (int jju = 0; jju < idxu_max; jju++)
        for (int ij = 0; ij < ijnum; ij++) {
            const long src = ij + (long)ijnum * jju;
            const long dst = ai[ij] + (long)inum * jju + (chemflag ? n1 * map[tj[ij]] : 0);
            Stotr[dst] += Sr[src];
            Stoti[dst] += Si[src];
        }
}

/* cpuSnap.cpp:1899-1969.  Z in the factored order q += cg_b * (sum_ia cg_a u1 u2). */
void cpuComputeZi2(double *zlist_r, double *zlist_i, double *Stotr, double *Stoti, double *cglist,
                   int *idxz, int *idxu_block, int *idxcg_block, int twojmax, int idxu_max, int idxz_max,
                   int nelements, int bnorm_flag, int inum)
{
    const int jd = twojmax + 1, ne = nelements;
    const long nu = (long)inum * idxu_max;
    for (int e1 = 0; e1 < ne; e1++) for (int e2 = 0; e2 < ne; e2++)
        for (int jjz = 0; jjz < idxz_max; jjz++)
            for (int ii = 0; ii < inum; ii++) {
                const int *e = idxz + 10 * jjz;
                const int j1 = e[0], j2 = e[1], j = e[2];
                const int ma1min = e[3], ma2max = e[4], na = e[5], mb1min = e[6], mb2max = e[7], nb = e[8];
                const double *cg = cglist + idxcg_block[j + j2 * jd + j1 * jd * jd];
                int jju1 = idxu_block[j1] + (j1 + 1) * mb1min;
                int jju2 = idxu_block[j2] + (j2 + 1) * mb2max;
                int icgb = mb1min * (j2 + 1) + mb2max;
                double qr = 0.0, qi = 0.0;
                for (int ib = 0; ib < nb; ib++) {
                    const double *u1r = Stotr + ii + (long)inum * jju1 + nu * e1;
                    const double *u1i = Stoti + ii + (long)inum * jju1 + nu * e1;
                    const double *u2r = Stotr + ii + (long)inum * jju2 + nu * e2;
                    const double *u2i = Stoti + ii + (long)inum * jju2 + nu * e2;
                    double sr = 0.0, si = 0.0;
                    int ma1 = ma1min, ma2 = ma2max, icga = ma1min * (j2 + 1) + ma2max;
                    for (int ia = 0; ia < na; ia++) {
                        sr += cg[icga] * (u1r[(long)inum * ma1] * u2r[(long)inum * ma2] - u1i[(long)inum * ma1] * u2i[(long)inum * ma2]);
                        si += cg[icga] * (u1r[(long)inum * ma1] * u2i[(long)inum * ma2] + u1i[(long)inum * ma1] * u2r[(long)inum * ma2]);
                        ma1++; ma2--; icga += j2;
                    }
                    qr += cg[icgb] * sr;
                    qi += cg[icgb] * si;
                    jju1 += j1 + 1; jju2 -= j2 + 1; icgb += j2;
                }
                if (bnorm_flag) { qr /= (j + 1); qi /= (j + 1); }
                const long idx = ii + (long)inum * jjz + (long)inum * idxz_max * (e1 * ne + e2);
                zlist_r[idx] = qr; zlist_i[idx] = qi;
            }
}

/* cpuSnap.cpp:1971-2072.  B = 2 sum' Stot* . Z (overwrite), then -bzero on the own-element triple
 * (or on every triple with wselfall). */
void cpuComputeBi2(double *blist, double *zlist_r, double *zlist_i, double *Stotr, double *Stoti,
                   double *bzero, int *ilist, int *type, int *map, int *idxb, int *idxu_block, int *idxz_block,
                   int twojmax, int idxb_max, int idxu_max, int idxz_max, int nelements, int bzero_flag,
                   int wselfall_flag, int chemflag, int inum)
{
    const int jd = twojmax + 1, ne = nelements;
    const long nz = (long)idxz_max * inum, nu = (long)idxu_max * inum, nb = (long)idxb_max * inum;
    for (int e1 = 0; e1 < ne; e1++) for (int e2 = 0; e2 < ne; e2++) for (int e3 = 0; e3 < ne; e3++)
        for (int jjb = 0; jjb < idxb_max; jjb++)
            for (int ii = 0; ii < inum; ii++) {
                const int j1 = idxb[3 * jjb], j2 = idxb[3 * jjb + 1], j = idxb[3 * jjb + 2];
                const long zoff = ii + nz * (e2 + ne * e1), uoff = ii + nu * e3;
                int jjz = idxz_block[j + j2 * jd + j1 * jd * jd], jju = idxu_block[j];
                double s = 0.0;
                for (int mb = 0; 2 * mb < j; mb++)
                    for (int ma = 0; ma <= j; ma++, jjz++, jju++)
                        s += Stotr[uoff + (long)inum * jju] * zlist_r[zoff + (long)inum * jjz]
                           + Stoti[uoff + (long)inum * jju] * zlist_i[zoff + (long)inum * jjz];
                if (j % 2 == 0) {
                    const int mb = j / 2;
                    for (int ma = 0; ma < mb; ma++, jjz++, jju++)
                        s += Stotr[uoff + (long)inum * jju] * zlist_r[zoff + (long)inum * jjz]
                           + Stoti[uoff + (long)inum * jju] * zlist_i[zoff + (long)inum * jjz];
                    s += 0.5 * (Stotr[uoff + (long)inum * jju] * zlist_r[zoff + (long)inum * jjz]
                              + Stoti[uoff + (long)inum * jju] * zlist_i[zoff + (long)inum * jjz]);
                }
                blist[ii + (long)inum * jjb + nb * ((e1 * ne + e2) * ne + e3)] = 2.0 * s;
            }
    if (!bzero_flag) return;
    if (!wselfall_flag) {
        for (int jjb = 0; jjb < idxb_max; jjb++)
            for (int ii = 0; ii < inum; ii++) {
                const int ie = chemflag ? map[type[ilist[ii]]] : 0;
                blist[ii + (long)inum * jjb + nb * ((ie * ne + ie) * ne + ie)] -= bzero[idxb[3 * jjb + 2]];
            }
    } else {
        for (long t = 0; t < (long)ne * ne * ne; t++)
            for (int jjb = 0; jjb < idxb_max; jjb++)
                for (int ii = 0; ii < inum; ii++)
                    blist[ii + (long)inum * jjb + nb * t] -= bzero[idxb[3 * jjb + 2]];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a7: dB/dr and the descriptor tallies                                                        */
/* ------------------------------------------------------------------------------------------ */

/* 2 * sum' dU(jl)* . Z over the half of layer jl (middle column 1/2), for the three Cartesian
 * components.  zoff/jjz address the Z block, n2 the pair, nu the dU component stride. */
static void dz_contract(double out[3], const double *zr, const double *zi, const double *dur, const double *dui,
                        long zi0, long zs, int jjz, int jju, int jl, long n2, long ns, long nu)
{
    out[0] = out[1] = out[2] = 0.0;
    for (int mb = 0; 2 * mb < jl; mb++)
        for (int ma = 0; ma <= jl; ma++, jjz++, jju++) {
            const double a = zr[zi0 + zs * jjz], b = zi[zi0 + zs * jjz];
            for (int k = 0; k < 3; k++)
                out[k] += dur[n2 + ns * jju + nu * k] * a + dui[n2 + ns * jju + nu * k] * b;
        }
    if (jl % 2 == 0) {
        const int mb = jl / 2;
        for (int ma = 0; ma < mb; ma++, jjz++, jju++) {
            const double a = zr[zi0 + zs * jjz], b = zi[zi0 + zs * jjz];
            for (int k = 0; k < 3; k++)
                out[k] += dur[n2 + ns * jju + nu * k] * a + dui[n2 + ns * jju + nu * k] * b;
        }
        const double a = zr[zi0 + zs * jjz], b = zi[zi0 + zs * jjz];
        for (int k = 0; k < 3; k++)
            out[k] += (dur[n2 + ns * jju + nu * k] * a + dui[n2 + ns * jju + nu * k] * b) * 0.5;
    }
}

/* cpuSnap.cpp:2074-2263.  dB_k/dr_ij as the sum of the three symmetric Z contractions of dU. The
 * result lands in dblist[ij + ijnum*k + 3*ijnum*jjb + 3*ijnum*idxb_max*itriple] for the triple
 * (elem1, elem2, elem3(j)) of the first contraction (all three add there, as upstream). */
void cpuComputeDbidrj(double *dblist, double *zlist_r, double *zlist_i, double *dulist_r, double *dulist_i,
                      int *idxb, int *idxu_block, int *idxz_block, int *map, int *ai, int *tj, int twojmax,
                      int idxb_max, int idxu_max, int idxz_max, int nelements, int bnorm_flag, int chemflag,
                      int inum, int ijnum)
{
    const int jd = twojmax + 1, ne = nelements;
    const long nz = (long)idxz_max * inum, nb = (long)idxb_max * ijnum, nu = (long)idxu_max * ijnum;
    memset(dblist, 0, sizeof(double) * nb * 3 * ne * ne * ne);
    for (int jjb = 0; jjb < idxb_max; jjb++)
        for (int ij = 0; ij < ijnum; ij++) {
            const int e3 = chemflag ? map[tj[ij]] : 0;
            const int i = ai[ij];
            const int j1 = idxb[3 * jjb], j2 = idxb[3 * jjb + 1], j = idxb[3 * jjb + 2];
            for (int e1 = 0; e1 < ne; e1++) for (int e2 = 0; e2 < ne; e2++) {
                const long zi0 = i + nz * (e1 * ne + e2);
                double *dbdr = dblist + nb * 3 * ((e1 * ne + e2) * ne + e3);
                double s[3];
                const long o = ij + (long)ijnum * 3 * jjb;
                dz_contract(s, zlist_r, zlist_i, dulist_r, dulist_i, zi0, inum,
                            idxz_block[j + j2 * jd + j1 * jd * jd], idxu_block[j], j, ij, ijnum, nu);
                for (int k = 0; k < 3; k++) dbdr[o + (long)ijnum * k] += 2.0 * s[k];

                const double j1fac = (j + 1) / (j1 + 1.0);
                dz_contract(s, zlist_r, zlist_i, dulist_r, dulist_i, zi0, inum,
                            idxz_block[j1 + j2 * jd + j * jd * jd], idxu_block[j1], j1, ij, ijnum, nu);
                for (int k = 0; k < 3; k++)
                    dbdr[o + (long)ijnum * k] += bnorm_flag ? 2.0 * s[k] : 2.0 * s[k] * j1fac;

                const double j2fac = (j + 1) / (j2 + 1.0);
                dz_contract(s, zlist_r, zlist_i, dulist_r, dulist_i, zi0, inum,
                            idxz_block[j2 + j1 * jd + j * jd * jd], idxu_block[j2], j2, ij, ijnum, nu);
                for (int k = 0; k < 3; k++)
                    dbdr[o + (long)ijnum * k] += bnorm_flag ? 2.0 * s[k] : 2.0 * s[k] * j2fac;
            }
        }
}

/* Start of row icoeff of the packed upper triangle of quadratic terms, after the ncoeff linear ones. */
static int quad_row(int ncoeff, int icoeff)
{
    return ncoeff + ncoeff * (ncoeff + 1) / 2 - (ncoeff - icoeff) * (ncoeff - icoeff + 1) / 2;
}

/* cpuSnap.cpp:2265-2291.  bi[icoeff + nperdim*(type-1)] = sum over atoms (zeroed here).  The
 * quadratic branch keeps the upstream placement: the first cross term shares the slot of the
 * square term. */
void cpuSnapTallyBispectrum(double *bi, double *bispectrum, int *ilist, int *type, int inum, int ncoeff,
                            int nperdim, int ntype, int quadraticflag)
{
    memset(bi, 0, sizeof(double) * nperdim * ntype);
    for (int ic = 0; ic < ncoeff; ic++)
        for (int ii = 0; ii < inum; ii++) {
            const int n = nperdim * (type[ilist[ii]] - 1);
            const double b = bispectrum[ii + (long)inum * ic];
            bi[ic + n] += b;
            if (quadraticflag == 1) {
                int k = n + quad_row(ncoeff, ic);
                bi[k] += 0.5 * b * b;
                for (int jc = ic + 1; jc < ncoeff; jc++) bi[k++] += b * bispectrum[ii + (long)inum * jc];
            }
        }
}

/* cpuSnap.cpp:2293-2365.  db[3*atom + inum*3*(icoeff + nperdim*(type-1))]: +dB on i, -dB on j. */
void cpuSnapTallyBispectrumDeriv(double *db, double *bispectrum, double *dbdr, int *aii, int *ai, int *aj,
                                 int *ti, int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag)
{
    memset(db, 0, sizeof(double) * (long)inum * 3 * nperdim * ntype);
    for (int ic = 0; ic < ncoeff; ic++)
        for (int ij = 0; ij < ijnum; ij++) {
            const int ii = aii[ij], i = ai[ij], j = aj[ij];
            const int n = nperdim * (ti[ij] - 1);
            const long nij = (long)ijnum * 3 * ic;
            const double bx = dbdr[ij + nij], by = dbdr[ij + ijnum + nij], bz = dbdr[ij + 2L * ijnum + nij];
            long nii = (long)inum * 3 * (ic + n);
            db[3 * i + nii] += bx; db[1 + 3 * i + nii] += by; db[2 + 3 * i + nii] += bz;
            db[3 * j + nii] -= bx; db[1 + 3 * j + nii] -= by; db[2 + 3 * j + nii] -= bz;
            if (quadraticflag) {
                const double b = bispectrum[ii + (long)inum * ic];
                int k = quad_row(ncoeff, ic);
                double tx = b * bx, ty = b * by, tz = b * bz;
                nii = (long)inum * 3 * (k + n);
                db[3 * i + nii] += tx; db[1 + 3 * i + nii] += ty; db[2 + 3 * i + nii] += tz;
                db[3 * j + nii] -= tx; db[1 + 3 * j + nii] -= ty; db[2 + 3 * j + nii] -= tz;
                for (int jc = ic + 1; jc < ncoeff; jc++) {
                    const long nj = (long)ijnum * 3 * jc;
                    const double cx = dbdr[ij + nj], cy = dbdr[ij + ijnum + nj], cz = dbdr[ij + 2L * ijnum + nj];
                    const double c = bispectrum[ii + (long)inum * jc];
                    tx = b * cx + bx * c; ty = b * cy + by * c; tz = b * cz + bz * c;
                    k += 1;
                    nii = (long)inum * 3 * (k + n);
                    db[3 * i + nii] += tx; db[1 + 3 * i + nii] += ty; db[2 + 3 * i + nii] += tz;
                    db[3 * j + nii] -= tx; db[1 + 3 * j + nii] -= ty; db[2 + 3 * j + nii] -= tz;
                }
            }
        }
}

static void add_virial(double *v, double dx, double dy, double dz, double bx, double by, double bz)
{
    v[0] += 1.0 * dx * bx; v[1] += 1.0 * dy * by; v[2] += 1.0 * dz * bz;
    v[3] += 1.0 * dx * by; v[4] += 1.0 * dx * bz; v[5] += 1.0 * dy * bz;
}

/* cpuSnap.cpp:2367-2470.  bv[6*(icoeff + nperdim*(type-1))] += -r_ij (x) dB (xx,yy,zz,xy,xz,yz). */
void cpuSnapTallyBispectrumVirial(double *bv, double *bispectrum, double *dbdr, double *rij, int *aii,
                                  int *ti, int inum, int ijnum, int ncoeff, int nperdim, int ntype, int quadraticflag)
{
    memset(bv, 0, sizeof(double) * 6 * nperdim * ntype);
    for (int ic = 0; ic < ncoeff; ic++)
        for (int ij = 0; ij < ijnum; ij++) {
            const int ii = aii[ij];
            const int n = nperdim * (ti[ij] - 1);
            const long nij = (long)ijnum * 3 * ic;
            const double dx = -rij[3 * ij], dy = -rij[3 * ij + 1], dz = -rij[3 * ij + 2];
            const double bx = dbdr[ij + nij], by = dbdr[ij + ijnum + nij], bz = dbdr[ij + 2L * ijnum + nij];
            add_virial(bv + 6 * (ic + n), dx, dy, dz, bx, by, bz);
            if (quadraticflag) {
                const double b = bispectrum[ii + (long)inum * ic];
                int k = quad_row(ncoeff, ic);
                add_virial(bv + 6 * (k + n), dx, dy, dz, b * bx, b * by, b * bz);
                for (int jc = ic + 1; jc < ncoeff; jc++) {
                    const long nj = (long)ijnum * 3 * jc;
                    const double c = bispectrum[ii + (long)inum * jc];
                    k += 1;
                    add_virial(bv + 6 * (k + n), dx, dy, dz,
                               b * dbdr[ij + nj] + bx * c, b * dbdr[ij + ijnum + nj] + by * c,
                               b * dbdr[ij + 2L * ijnum + nj] + bz * c);
                }
            }
        }
}

/* ------------------------------------------------------------------------------------------ */
/* Brute-force periodic neighbour list (oracle-only)                                           */
/* ------------------------------------------------------------------------------------------ */

long oracle_neighbor_bruteforce(const double *x, int n, const double *box, double rc,
                                int *offsets, int *nbr_j, int *nbr_shift, long maxnb)
{
    int reach[3];
    for (int d = 0; d < 3; d++) reach[d] = (int)ceil(rc / box[d]) + 1;
    const double rc2 = rc * rc;
    long cnt = 0;
    for (int i = 0; i < n; i++) {
        offsets[i] = (int)(cnt < maxnb ? cnt : maxnb);
        for (int j = 0; j < n; j++)
            for (int sz = -reach[2]; sz <= reach[2]; sz++)
                for (int sy = -reach[1]; sy <= reach[1]; sy++)
                    for (int sx = -reach[0]; sx <= reach[0]; sx++) {
                        if (j == i && sx == 0 && sy == 0 && sz == 0) continue;
                        double gx = x[3 * j] + sx * box[0], gy = x[3 * j + 1] + sy * box[1], gz = x[3 * j + 2] + sz * box[2];
                        double dx = gx - x[3 * i], dy = gy - x[3 * i + 1], dz = gz - x[3 * i + 2];
                        double d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 <= rc2) {
                            if (cnt < maxnb) {
                                nbr_j[cnt] = j;
                                nbr_shift[3 * cnt] = sx; nbr_shift[3 * cnt + 1] = sy; nbr_shift[3 * cnt + 2] = sz;
                            }
                            cnt++;
                        }
                    }
    }
    offsets[n] = (int)(cnt < maxnb ? cnt : maxnb);
    return cnt <= maxnb ? cnt : -cnt;
}

/* ------------------------------------------------------------------------------------------------
 * ZBL reference pair potential (the empirical pair term used beside SNAP in the Ta application).
 *
 * Restated from simulating/snap/ZBLpotential.jl:13-59 (ezbl) and :61-108 (Paira) with the
 * parameters mu = [cut_inner, cut_outer, z] (the mu-driven form of ZBLpotential.jl:68-71;
 * snapTA06A.jl:62-66 passes [4, 4.8, 73]).  u(r) = E(r) + S(r) with
 *   E(r) = zze/r * sum_k c_k exp(-d_k a r),   a = (z^0.23 + z^0.23)/0.4685,  zze = z z 14.399645,
 *   S(r) = sw5 + A(t) t^3 (sw3 + sw4 t),      t = r - cut_inner,  A = 1/2 + tanh(100 t)/2,
 * the switching constants sw3..sw5 from E, E', E'' at cut_outer (ZBLpotential.jl:79-89).  The
 * generated code the upstream application compiles (src/C++/Potentials/cpuPairaGradient.cpp:1-43)
 * differentiates the same expression symbolically, tanh step included; here
 *   du/dr = E'(r) + 50 (1 - tanh^2(100 t)) t^3 (sw3 + sw4 t) + A(t) (3 sw3 t^2 + 4 sw4 t^3),
 * and u_xij = du/dr * xij / r, with xij = x_j - x_i.
 * ---------------------------------------------------------------------------------------------- */
static void zbl_e(double r, double ainv, double zze, double *e, double *ep, double *epp)
{
    static const double c[4] = {0.02817, 0.28022, 0.50986, 0.18175};
    static const double d[4] = {0.20162, 0.40290, 0.94229, 3.19980};
    double s = 0.0, sp = 0.0, spp = 0.0;
    for (int k = 0; k < 4; k++) {
        const double da = d[k] * ainv, ek = exp(-da * r);
        s += c[k] * ek;
        sp -= c[k] * da * ek;
        spp += c[k] * ek * da * da;
    }
    const double rinv = 1.0 / r;
    *e = zze * s * rinv;
    *ep = zze * (sp - s * rinv) * rinv;
    *epp = zze * (spp - 2.0 * sp * rinv + 2.0 * s * rinv * rinv) * rinv;
}

void oracle_PairaGradient(double *u, double *u_xij, const double *xij, const double *mu, int ng)
{
    const double rin = mu[0], rout = mu[1], z = mu[2];
    const double ainv = (pow(z, 0.23) + pow(z, 0.23)) / 0.46850;
    const double zze = z * z * 14.399645;
    double fc, fcp, fcpp;
    zbl_e(rout, ainv, zze, &fc, &fcp, &fcpp);
    const double tc = rout - rin;
    const double sw3 = (-3.0 * fcp + tc * fcpp) / (tc * tc) / 3.0;
    const double sw4 = (2.0 * fcp - tc * fcpp) / (tc * tc * tc) / 4.0;
    const double sw5 = -fc + (tc / 2.0) * fcp - (tc * tc / 12.0) * fcpp;
    for (int p = 0; p < ng; p++) {
        const double x = xij[3 * p], y = xij[3 * p + 1], zz = xij[3 * p + 2];
        const double r = sqrt(x * x + y * y + zz * zz);
        double e, ep, epp;
        zbl_e(r, ainv, zze, &e, &ep, &epp);
        const double t = r - rin, th = tanh(100.0 * t), a = 0.5 + 0.5 * th;
        const double poly = t * t * t * (sw3 + sw4 * t);
        u[p] = e + sw5 + a * poly;
        const double dudr = ep + 50.0 * (1.0 - th * th) * poly + a * (3.0 * sw3 * t * t + 4.0 * sw4 * t * t * t);
        u_xij[3 * p] = dudr * x / r;
        u_xij[3 * p + 1] = dudr * y / r;
        u_xij[3 * p + 2] = dudr * zz / r;
    }
}

/* cpuFullNeighPairDecomposition (src/C++/Core/cpuEnergyForceTally.cpp:216-227): e[ai] += eij / 2,
 * f[ai] -= fij, over a full pair list (each pair seen from both atoms). */
void oracle_FullNeighPairDecomposition(double *e, double *f, const double *eij, const double *fij, const int *ai,
                                       int ijnum, int dim)
{
    for (int p = 0; p < ijnum; p++) {
        const int i = ai[p];
        e[i] += 0.5 * eij[p];
        for (int d = 0; d < dim; d++) f[dim * i + d] += -fij[dim * p + d];
    }
}

/* cpuVirialPairTally (src/C++/Core/cpuVirialTally.cpp:34-57): v[k*inum + ai] += factor * (rij (x) fij)_k
 * for k = xx, yy, zz, xy, xz, yz (SoA). */
void oracle_VirialPairTally(double *v, const double *fij, const double *rij, double factor, const int *ai, int dim,
                            int inum, int ijnum)
{
    for (int p = 0; p < ijnum; p++) {
        const int i = ai[p];
        const double dx = rij[dim * p], dy = rij[dim * p + 1], dz = dim == 3 ? rij[dim * p + 2] : 0.0;
        const double fx = fij[dim * p], fy = fij[dim * p + 1], fz = dim == 3 ? fij[dim * p + 2] : 0.0;
        v[0 * inum + i] += factor * dx * fx;
        v[1 * inum + i] += factor * dy * fy;
        v[2 * inum + i] += factor * dz * fz;
        v[3 * inum + i] += factor * dx * fy;
        v[4 * inum + i] += factor * dx * fz;
        v[5 * inum + i] += factor * dy * fz;
    }
}
