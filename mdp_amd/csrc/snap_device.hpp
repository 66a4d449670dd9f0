// Device building blocks of the SNAP kernels: compile-time loops, per-pair Cayley-Klein geometry,
// and the row-chain Wigner-U recursion with optional Cartesian derivatives.
//
// The recursion (src/Julia/Potential/cpuSnap.cpp:366-488 for U, :1091-1368 with derivatives) builds
// row mb of layer j from row mb of layer j-1,
//   U_j(ma,mb) = r(j-ma,j-mb) conj(a) U_{j-1}(ma,mb) - r(ma,j-mb) conj(b) U_{j-1}(ma-1,mb),
// with r(p,q) = sqrt(p/q); row mb starts at layer 2mb from the mirror of row mb-1 at layer 2mb-1,
// U_j(j-ma, j-mb) = (-1)^(ma-mb) conj U_j(ma,mb).  Only rows 2mb <= j are produced.  Here every
// loop bound and array index is a compile-time constant (J is a template parameter), so the whole
// recursion state lives in registers.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace mdp {

template <int V>
using ic = std::integral_constant<int, V>;

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(ic<I>{});
        static_for<I + 1, N>(static_cast<F&&>(f));
    }
}

// ---- compile-time index spaces --------------------------------------------------------------

__host__ __device__ constexpr int cx_min(int a, int b) { return a < b ? a : b; }
__host__ __device__ constexpr int cx_max(int a, int b) { return a > b ? a : b; }

// first element of layer j in the full U array
__host__ __device__ constexpr int u_block(int j) {
    int n = 0;
    for (int k = 0; k < j; ++k) n += (k + 1) * (k + 1);
    return n;
}

// first element of layer j in the compact half layout (rows 2mb <= j only), and its size
__host__ __device__ constexpr int h_block(int j) {
    int n = 0;
    for (int k = 0; k < j; ++k) n += (k / 2 + 1) * (k + 1);
    return n;
}
__host__ __device__ constexpr int n_half_elems(int J) { return h_block(J + 1); }

// coupling triple (j1 >= j2 -> j) exists
__host__ __device__ constexpr bool coupled(int j1, int j2, int j) {
    return j2 <= j1 && j >= j1 - j2 && j <= j1 + j2 && ((j1 + j2 + j) % 2 == 0);
}

// offset of the CG block of (j1, j2, j) in the CG table, in canonical triple order
__host__ __device__ constexpr int cg_block(int J, int j1, int j2, int j) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) {
                if (a == j1 && b == j2 && c == j) return n;
                n += (a + 1) * (b + 1);
            }
    return -1;
}

// ordinal of the triple (j1, j2, j) in canonical order
__host__ __device__ constexpr int triple_id(int J, int j1, int j2, int j) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) {
                if (a == j1 && b == j2 && c == j) return n;
                ++n;
            }
    return -1;
}

// ordinal of (j1, j2, j) among the bispectrum triples (j >= j1) in canonical order, and their number
__host__ __device__ constexpr int bisp_index(int J, int j1, int j2, int j) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) {
                if (c < a) continue;
                if (a == j1 && b == j2 && c == j) return n;
                ++n;
            }
    return -1;
}
__host__ __device__ constexpr int n_bisp(int J) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) n += (c >= a);
    return n;
}

__host__ __device__ constexpr int n_triples(int J) {
    int n = 0;
    for (int a = 0; a <= J; ++a)
        for (int b = 0; b <= a; ++b)
            for (int c = a - b; c <= cx_min(J, a + b); c += 2) ++n;
    return n;
}

// first m1 / partner m2 / count of the CG sum producing output m of (j1, j2) -> j
__host__ __device__ constexpr int cg_lo(int m, int j1, int j2, int j) { return cx_max(0, (2 * m - j - j2 + j1) / 2); }
__host__ __device__ constexpr int cg_hi2(int m, int j1, int j2, int j) {
    return (2 * m - j - (2 * cg_lo(m, j1, j2, j) - j1) + j2) / 2;
}
__host__ __device__ constexpr int cg_n(int m, int j1, int j2, int j) {
    return cx_min(j1, (2 * m - j + j2 + j1) / 2) - cg_lo(m, j1, j2, j) + 1;
}

// half-U rows (j, mb) with 2mb <= j, enumerated j-major: row id -> (j, mb)
__host__ __device__ constexpr int n_half_rows(int J) {
    int n = 0;
    for (int j = 0; j <= J; ++j) n += j / 2 + 1;
    return n;
}
__host__ __device__ constexpr int half_row_j(int r) {
    int j = 0;
    while (r >= j / 2 + 1) { r -= j / 2 + 1; ++j; }
    return j;
}
__host__ __device__ constexpr int half_row_mb(int r) {
    int j = 0;
    while (r >= j / 2 + 1) { r -= j / 2 + 1; ++j; }
    return r;
}

// ---- compile-time Clebsch-Gordan coefficients ------------------------------------------------------

// sqrt of a non-negative argument at compile time (Newton from above; within an ulp of sqrt)
__host__ __device__ constexpr double cx_sqrt(double x) {
    if (x <= 0.0) return 0.0;
    double r = x > 1.0 ? x : 1.0;
    for (int k = 0; k < 200; ++k) {
        const double nr = 0.5 * (r + x / r);
        if (nr >= r) break;
        r = nr;
    }
    return r;
}
__host__ __device__ constexpr double cx_fact(int n) {
    double f = 1.0;
    for (int k = 2; k <= n; ++k) f *= k;
    return f;
}
// <j1 m1 j2 m2 | j m> with doubled j and m1, m2 counted from -j1/2, -j2/2 (the layout of the CG table:
// entry m1 * (j2 + 1) + m2 of block cg_block(j1, j2, j)), by Racah's formula (Varshalovich 8.2.1), the
// same expression the host tables evaluate
__host__ __device__ constexpr double cg_const(int j1, int j2, int j, int m1, int m2) {
    const int a2 = 2 * m1 - j1, b2 = 2 * m2 - j2, m = (a2 + b2 + j) / 2;
    if (m < 0 || m > j) return 0.0;
    const int zmin = cx_max(0, cx_max(-(j - j2 + a2) / 2, -(j - j1 - b2) / 2));
    const int zmax = cx_min((j1 + j2 - j) / 2, cx_min((j1 - a2) / 2, (j2 + b2) / 2));
    double s = 0.0;
    for (int z = zmin; z <= zmax; ++z) {
        const double sg = (z & 1) ? -1.0 : 1.0;
        s += sg / (cx_fact(z) * cx_fact((j1 + j2 - j) / 2 - z) * cx_fact((j1 - a2) / 2 - z) *
                   cx_fact((j2 + b2) / 2 - z) * cx_fact((j - j2 + a2) / 2 + z) * cx_fact((j - j1 - b2) / 2 + z));
    }
    const int c2 = 2 * m - j;
    const double tri = cx_sqrt(cx_fact((j1 + j2 - j) / 2) * cx_fact((j1 - j2 + j) / 2) * cx_fact((-j1 + j2 + j) / 2) /
                               cx_fact((j1 + j2 + j) / 2 + 1));
    const double nrm = cx_sqrt(cx_fact((j1 + a2) / 2) * cx_fact((j1 - a2) / 2) * cx_fact((j2 + b2) / 2) *
                               cx_fact((j2 - b2) / 2) * cx_fact((j + c2) / 2) * cx_fact((j - c2) / 2) * (j + 1));
    return s * tri * nrm;
}

// ---- per-pair geometry ------------------------------------------------------------------------

struct PairGeom {
    double ar, ai, br, bi;                       // Cayley-Klein parameters
    double dar[3], dai[3], dbr[3], dbi[3];       // d/dx, d/dy, d/dz
    double u[3];                                 // unit vector of r_ij
    double sfac, dsfac;                          // switching * w_j and its radial derivative
};

// r_ij -> (a, b): theta0 = (r - rmin0) rfac0 pi / (rc - rmin0), z0 = r / tan(theta0),
// a = (z0 - i z)/r0, b = (y - i x)/r0, r0 = |(r, z0)|   (cpuSnap.cpp:996-1067)
template <bool DERIV>
__device__ __forceinline__ void pair_geometry(double x, double y, double z, double rc, double rmin0, double rfac0,
                                              int switchflag, double wj, PairGeom& g) {
    const double rsq = x * x + y * y + z * z;
    const double r = sqrt(rsq);
    const double rscale = rfac0 * M_PI / (rc - rmin0);
    const double theta0 = (r - rmin0) * rscale;
    const double z0 = r / tan(theta0);
    const double r0inv = 1.0 / sqrt(rsq + z0 * z0);
    g.ar = r0inv * z0;
    g.ai = -r0inv * z;
    g.br = r0inv * y;
    g.bi = -r0inv * x;
    double sfac = (switchflag == 0) ? 1.0 : 0.0, dsfac = 0.0;
    if (switchflag == 1) {
        if (r <= rmin0 || r > rc) {
            sfac = 1.0;
        } else {
            const double w = M_PI / (rc - rmin0);
            double s, c;
            sincos((r - rmin0) * w, &s, &c);
            sfac = 0.5 * (c + 1.0);
            dsfac = -0.5 * s * w;
        }
    }
    g.sfac = sfac * wj;
    g.dsfac = dsfac * wj;
    if constexpr (DERIV) {
        const double rinv = 1.0 / r;
        g.u[0] = x * rinv;
        g.u[1] = y * rinv;
        g.u[2] = z * rinv;
        const double dz0dr = z0 * rinv - (r * rscale) * (rsq + z0 * z0) / rsq;
        const double dr0 = -(r0inv * r0inv * r0inv) * (r + z0 * dz0dr);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double w = dr0 * g.u[k];
            g.dar[k] = dz0dr * g.u[k] * r0inv + z0 * w;
            g.dai[k] = -z * w;
            g.dbr[k] = y * w;
            g.dbi[k] = -x * w;
        }
        g.dai[2] -= r0inv;
        g.dbi[0] -= r0inv;
        g.dbr[1] += r0inv;
    }
}

// ---- row-chain recursion ----------------------------------------------------------------------

// Runs the recursion for rows 0..RHI.  Rows below RLO are advanced only as far as the seed of the
// next row (layer 2mb+1); rows RLO..RHI run to layer J and every finished (j, mb) row is passed to
// emit(ic<j>, ic<mb>, pr, pi, dpr, dpi); derivative slot k carries Cartesian component k + KOFF
// (ND = 0, 1 or 3).
// The middle row of an even layer is completed by the mirror for ma > j/2; the centre element keeps
// the upstream sign convention (imaginary part of U negated, of dU not; cpuSnap.cpp:1260-1306).
template <int J, int ND, int RLO, int RHI, int KOFF = 0, class Emit>
__device__ __forceinline__ void row_chain(const PairGeom& g, const double* __restrict__ rootpq, Emit&& emit) {
    constexpr int JD = J + 1;
    double pr[J + 1], pi[J + 1], qr[J + 1], qi[J + 1];
    double dpr[ND > 0 ? ND : 1][J + 1], dpi[ND > 0 ? ND : 1][J + 1];
    double dqr[ND > 0 ? ND : 1][J + 1], dqi[ND > 0 ? ND : 1][J + 1];
    pr[0] = 1.0;
    pi[0] = 0.0;
#pragma unroll
    for (int k = 0; k < ND; ++k) { dpr[k][0] = 0.0; dpi[k][0] = 0.0; }
    if constexpr (RLO == 0) emit(ic<0>{}, ic<0>{}, pr, pi, dpr, dpi);

    static_for<0, RHI + 1>([&](auto MBc) {
        constexpr int MB = decltype(MBc)::value;
        constexpr int JSTART = (MB == 0) ? 1 : 2 * MB;
        constexpr int JEND = (MB < RLO) ? cx_min(J, 2 * MB + 1) : J;
        if constexpr (MB > 0) {
#pragma unroll
            for (int ma = 0; ma < 2 * MB; ++ma) {
                pr[ma] = qr[ma];
                pi[ma] = qi[ma];
#pragma unroll
                for (int k = 0; k < ND; ++k) { dpr[k][ma] = dqr[k][ma]; dpi[k][ma] = dqi[k][ma]; }
            }
        }
        static_for<JSTART, JEND + 1>([&](auto Jc) {
            constexpr int JJ = decltype(Jc)::value;
            // ma = 0: only the conj(a) term
            double ur = pr[0], ui = pi[0];
            double dur[ND > 0 ? ND : 1], dui[ND > 0 ? ND : 1];
#pragma unroll
            for (int k = 0; k < ND; ++k) { dur[k] = dpr[k][0]; dui[k] = dpi[k][0]; }
            {
                const double rp = rootpq[JJ * JD + (JJ - MB)];
                pr[0] = rp * (g.ar * ur + g.ai * ui);
                pi[0] = rp * (g.ar * ui - g.ai * ur);
#pragma unroll
                for (int k = 0; k < ND; ++k) {
                    dpr[k][0] = rp * (g.dar[k + KOFF] * ur + g.dai[k + KOFF] * ui + g.ar * dur[k] + g.ai * dui[k]);
                    dpi[k][0] = rp * (g.dar[k + KOFF] * ui - g.dai[k + KOFF] * ur + g.ar * dui[k] - g.ai * dur[k]);
                }
            }
#pragma unroll
            for (int ma = 1; ma < JJ; ++ma) {
                const double wr = pr[ma], wi = pi[ma];
                double dwr[ND > 0 ? ND : 1], dwi[ND > 0 ? ND : 1];
#pragma unroll
                for (int k = 0; k < ND; ++k) { dwr[k] = dpr[k][ma]; dwi[k] = dpi[k][ma]; }
                const double r1 = rootpq[(JJ - ma) * JD + (JJ - MB)];
                const double r2 = rootpq[ma * JD + (JJ - MB)];
                pr[ma] = r1 * (g.ar * wr + g.ai * wi) - r2 * (g.br * ur + g.bi * ui);
                pi[ma] = r1 * (g.ar * wi - g.ai * wr) - r2 * (g.br * ui - g.bi * ur);
#pragma unroll
                for (int k = 0; k < ND; ++k) {
                    dpr[k][ma] = r1 * (g.dar[k + KOFF] * wr + g.dai[k + KOFF] * wi + g.ar * dwr[k] + g.ai * dwi[k]) -
                                 r2 * (g.dbr[k + KOFF] * ur + g.dbi[k + KOFF] * ui + g.br * dur[k] + g.bi * dui[k]);
                    dpi[k][ma] = r1 * (g.dar[k + KOFF] * wi - g.dai[k + KOFF] * wr + g.ar * dwi[k] - g.ai * dwr[k]) -
                                 r2 * (g.dbr[k + KOFF] * ui - g.dbi[k + KOFF] * ur + g.br * dui[k] - g.bi * dur[k]);
                }
                ur = wr;
                ui = wi;
#pragma unroll
                for (int k = 0; k < ND; ++k) { dur[k] = dwr[k]; dui[k] = dwi[k]; }
            }
            {
                const double r2 = rootpq[JJ * JD + (JJ - MB)];
                pr[JJ] = -r2 * (g.br * ur + g.bi * ui);
                pi[JJ] = -r2 * (g.br * ui - g.bi * ur);
#pragma unroll
                for (int k = 0; k < ND; ++k) {
                    dpr[k][JJ] = -r2 * (g.dbr[k + KOFF] * ur + g.dbi[k + KOFF] * ui + g.br * dur[k] + g.bi * dui[k]);
                    dpi[k][JJ] = -r2 * (g.dbr[k + KOFF] * ui - g.dbi[k + KOFF] * ur + g.br * dui[k] - g.bi * dur[k]);
                }
            }
            if constexpr (JJ == 2 * MB) {
#pragma unroll
                for (int ma = 0; ma <= JJ / 2; ++ma) {
                    const int s = JJ / 2 - ma, d = JJ / 2 + ma;
                    const bool even = (ma % 2) == 0;
                    pr[d] = even ? pr[s] : -pr[s];
                    pi[d] = even ? -pi[s] : pi[s];
                    if (ma > 0) {
#pragma unroll
                        for (int k = 0; k < ND; ++k) {
                            dpr[k][d] = even ? dpr[k][s] : -dpr[k][s];
                            dpi[k][d] = even ? -dpi[k][s] : dpi[k][s];
                        }
                    }
                }
            }
            if constexpr (JJ == 2 * MB + 1 && MB < RHI) {
                constexpr int S0 = (((JJ + 1) / 2) % 2 == 0) ? -1 : 1;
#pragma unroll
                for (int ma = 0; ma <= JJ; ++ma) {
                    const bool pos = ((ma % 2) == 0) == (S0 == 1);
                    qr[JJ - ma] = pos ? pr[ma] : -pr[ma];
                    qi[JJ - ma] = pos ? -pi[ma] : pi[ma];
#pragma unroll
                    for (int k = 0; k < ND; ++k) {
                        dqr[k][JJ - ma] = pos ? dpr[k][ma] : -dpr[k][ma];
                        dqi[k][JJ - ma] = pos ? -dpi[k][ma] : dpi[k][ma];
                    }
                }
            }
            if constexpr (MB >= RLO) emit(ic<JJ>{}, ic<MB>{}, pr, pi, dpr, dpi);
        });
    });
}

// Number of half-U elements emitted for rows RLO..RHI (plus the j=0 element when RLO == 0).
__host__ __device__ constexpr int rows_elems(int J, int rlo, int rhi) {
    int n = 0;
    for (int mb = rlo; mb <= rhi; ++mb)
        for (int j = 2 * mb; j <= J; ++j) n += j + 1;
    return n;
}
// accumulator slot of (j, mb, ma) inside the group rows RLO..RHI (row-major by mb, then j)
__host__ __device__ constexpr int rows_slot(int J, int rlo, int j, int mb, int ma) {
    int n = 0;
    for (int b = rlo; b < mb; ++b)
        for (int jj = 2 * b; jj <= J; ++jj) n += jj + 1;
    for (int jj = 2 * mb; jj < j; ++jj) n += jj + 1;
    return n + ma;
}

}  // namespace mdp
