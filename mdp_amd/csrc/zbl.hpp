// ZBL reference pair potential of the SNAP Ta application (simulating/snap/ZBLpotential.jl:13-108;
// parameters mu = [cut_inner, cut_outer, z], snapTA06A.jl:62-66 passes [4, 4.8, 73]).
//
//   u(r) = E(r) + sw5 + A(t) t^3 (sw3 + sw4 t),   t = r - cut_inner,   A(t) = 1/2 + tanh(100 t)/2
//   E(r) = zze/r * sum_k c_k exp(-d_k a r),       a = (z^0.23 + z^0.23)/0.4685,   zze = z^2 * 14.399645
//
// with sw3, sw4, sw5 from E, E', E'' at cut_outer (ZBLpotential.jl:79-89), so that u and u' vanish
// there.  The upstream application differentiates this expression symbolically (tanh step included)
// into src/C++/Potentials/cpuPairaGradient.cpp; the same derivative in closed form is
//   du/dr = E'(r) + 50 (1 - tanh^2(100 t)) t^3 (sw3 + sw4 t) + A(t) (3 sw3 t^2 + 4 sw4 t^3),
// and the gradient with respect to the pair vector xij = x_j - x_i is du/dr * xij / r.
#pragma once

#include <cmath>

struct ZblParams {
    double rin = 0, rout2 = 0, zze = 0, sw3 = 0, sw4 = 0, sw5 = 0;
    double c[4] = {0, 0, 0, 0}, da[4] = {0, 0, 0, 0};   // c_k and d_k * a
};

// ZBL screening sum and its first two radial derivatives (ZBLpotential.jl:13-59)
inline void zbl_host_e(const ZblParams& p, double r, double& e, double& ep, double& epp) {
    double s = 0.0, sp = 0.0, spp = 0.0;
    for (int k = 0; k < 4; ++k) {
        const double ek = std::exp(-p.da[k] * r);
        s += p.c[k] * ek;
        sp -= p.c[k] * p.da[k] * ek;
        spp += p.c[k] * ek * p.da[k] * p.da[k];
    }
    const double rinv = 1.0 / r;
    e = p.zze * s * rinv;
    ep = p.zze * (sp - s * rinv) * rinv;
    epp = p.zze * (spp - 2.0 * sp * rinv + 2.0 * s * rinv * rinv) * rinv;
}

// constants of one (cut_inner, cut_outer, z) parameter set; cut_outer must exceed cut_inner
inline ZblParams zbl_params(double cut_inner, double cut_outer, double z) {
    ZblParams p;
    static const double c[4] = {0.02817, 0.28022, 0.50986, 0.18175};
    static const double d[4] = {0.20162, 0.40290, 0.94229, 3.19980};
    const double ainv = (std::pow(z, 0.23) + std::pow(z, 0.23)) / 0.46850;
    for (int k = 0; k < 4; ++k) {
        p.c[k] = c[k];
        p.da[k] = d[k] * ainv;
    }
    p.zze = z * z * 14.399645;
    p.rin = cut_inner;
    p.rout2 = cut_outer * cut_outer;
    double fc, fcp, fcpp;
    zbl_host_e(p, cut_outer, fc, fcp, fcpp);
    const double tc = cut_outer - cut_inner;
    p.sw3 = (-3.0 * fcp + tc * fcpp) / (tc * tc) / 3.0;
    p.sw4 = (2.0 * fcp - tc * fcpp) / (tc * tc * tc) / 4.0;
    p.sw5 = -fc + (tc / 2.0) * fcp - (tc * tc / 12.0) * fcpp;
    return p;
}

#ifdef __HIPCC__
// u(r) and du/dr of one pair at distance r
__device__ __forceinline__ void zbl_eval(const ZblParams& p, double r, double& u, double& dudr) {
    double s = 0.0, sp = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double ek = exp(-p.da[k] * r);
        s += p.c[k] * ek;
        sp -= p.c[k] * p.da[k] * ek;
    }
    const double rinv = 1.0 / r;
    const double t = r - p.rin, th = tanh(100.0 * t), a = 0.5 + 0.5 * th;
    const double t2 = t * t, t3 = t2 * t;
    const double poly = t3 * (p.sw3 + p.sw4 * t);
    u = p.zze * s * rinv + p.sw5 + a * poly;
    dudr = p.zze * (sp - s * rinv) * rinv + 50.0 * (1.0 - th * th) * poly + a * (3.0 * p.sw3 * t2 + 4.0 * p.sw4 * t3);
}
#endif
