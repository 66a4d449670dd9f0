// SNAP index spaces and Clebsch-Gordan table.  See snap_tables.hpp.
#include "snap_tables.hpp"

#include <algorithm>
#include <cmath>

namespace mdp {

namespace {

// Visit the coupled triples (j1 >= j2, j = j1-j2 .. min(J, j1+j2) step 2) in canonical order
// (cpuSnap.cpp:23-69).
template <class F>
void each_triple(int J, F&& f) {
    for (int j1 = 0; j1 <= J; ++j1)
        for (int j2 = 0; j2 <= j1; ++j2)
            for (int j = j1 - j2; j <= std::min(J, j1 + j2); j += 2) f(j1, j2, j);
}

// m1 range coupling to output m (cpuSnap.cpp:158-163): first index, partner start, count.
void coupling_range(int m, int j1, int j2, int j, int& lo, int& hi2, int& n) {
    lo = std::max(0, (2 * m - j - j2 + j1) / 2);
    hi2 = (2 * m - j - (2 * lo - j1) + j2) / 2;
    n = std::min(j1, (2 * m - j + j2 + j1) / 2) - lo + 1;
}

double fact_exact(int n) {
    double f = 1.0;
    for (int k = 2; k <= n; ++k) f *= k;
    return f;
}

}  // namespace

int count_u(int J) { int n = 0; for (int j = 0; j <= J; ++j) n += (j + 1) * (j + 1); return n; }
int count_z(int J) { int n = 0; each_triple(J, [&](int, int, int j) { n += (j / 2 + 1) * (j + 1); }); return n; }
int count_b(int J) { int n = 0; each_triple(J, [&](int j1, int, int j) { n += (j >= j1); }); return n; }
int count_cg(int J) { int n = 0; each_triple(J, [&](int j1, int j2, int) { n += (j1 + 1) * (j2 + 1); }); return n; }

int SnapIndex::half_count() const {
    int n = 0;
    for (int j = 0; j <= J; ++j) n += (j / 2 + 1) * (j + 1);
    return n;
}

void SnapIndex::build(int twojmax, const double* factorial) {
    J = twojmax;
    const int jd = J + 1;
    idxu_max = count_u(J);
    idxz_max = count_z(J);
    idxb_max = count_b(J);
    idxcg_max = count_cg(J);
    idxu_block.assign(jd, 0);
    idxz.assign(10 * idxz_max, 0);
    idxz_block.assign(jd * jd * jd, 0);
    idxb.assign(3 * idxb_max, 0);
    idxb_block.assign(jd * jd * jd, 0);
    idxcg_block.assign(jd * jd * jd, 0);
    cglist.assign(idxcg_max, 0.0);
    rootpq.assign((J + 2) * (J + 2), 0.0);

    for (int j = 0, acc = 0; j <= J; ++j) { idxu_block[j] = acc; acc += (j + 1) * (j + 1); }
    int ncg = 0, nb = 0, nz = 0;
    each_triple(J, [&](int j1, int j2, int j) {
        const int key = j + j2 * jd + j1 * jd * jd;
        idxcg_block[key] = ncg;
        ncg += (j1 + 1) * (j2 + 1);
        if (j >= j1) {
            idxb[3 * nb] = j1; idxb[3 * nb + 1] = j2; idxb[3 * nb + 2] = j;
            idxb_block[key] = nb++;
        }
        idxz_block[key] = nz;
        for (int mb = 0; 2 * mb <= j; ++mb)
            for (int ma = 0; ma <= j; ++ma) {
                int* e = &idxz[10 * nz++];
                e[0] = j1; e[1] = j2; e[2] = j;
                coupling_range(ma, j1, j2, j, e[3], e[4], e[5]);
                coupling_range(mb, j1, j2, j, e[6], e[7], e[8]);
                e[9] = idxu_block[j] + (j + 1) * mb + ma;
            }
    });

    for (int p = 1; p <= J; ++p)
        for (int q = 1; q <= J; ++q) rootpq[p * jd + q] = std::sqrt(static_cast<double>(p) / q);

    // Clebsch-Gordan coefficients <j1 m1 j2 m2 | j m> (doubled indices) by Racah's formula
    // (Varshalovich 8.2.1), with the caller's factorial table when given.
    std::vector<double> fac(3 * J + 4);
    for (size_t k = 0; k < fac.size(); ++k) fac[k] = factorial ? factorial[k] : fact_exact(static_cast<int>(k));
    int k = 0;
    each_triple(J, [&](int j1, int j2, int j) {
        for (int m1 = 0; m1 <= j1; ++m1)
            for (int m2 = 0; m2 <= j2; ++m2) {
                const int a2 = 2 * m1 - j1, b2 = 2 * m2 - j2, m = (a2 + b2 + j) / 2;
                double cg = 0.0;
                if (m >= 0 && m <= j) {
                    const int zmin = std::max(0, std::max(-(j - j2 + a2) / 2, -(j - j1 - b2) / 2));
                    const int zmax = std::min((j1 + j2 - j) / 2, std::min((j1 - a2) / 2, (j2 + b2) / 2));
                    double s = 0.0;
                    for (int z = zmin; z <= zmax; ++z) {
                        const double sg = (z & 1) ? -1.0 : 1.0;
                        s += sg / (fac[z] * fac[(j1 + j2 - j) / 2 - z] * fac[(j1 - a2) / 2 - z] *
                                   fac[(j2 + b2) / 2 - z] * fac[(j - j2 + a2) / 2 + z] * fac[(j - j1 - b2) / 2 + z]);
                    }
                    const int c2 = 2 * m - j;
                    const double tri = std::sqrt(fac[(j1 + j2 - j) / 2] * fac[(j1 - j2 + j) / 2] *
                                                 fac[(-j1 + j2 + j) / 2] / fac[(j1 + j2 + j) / 2 + 1]);
                    const double nrm = std::sqrt(fac[(j1 + a2) / 2] * fac[(j1 - a2) / 2] * fac[(j2 + b2) / 2] *
                                                 fac[(j2 - b2) / 2] * fac[(j + c2) / 2] * fac[(j - c2) / 2] * (j + 1));
                    cg = s * tri * nrm;
                }
                cglist[k++] = cg;
            }
    });
}

}  // namespace mdp
