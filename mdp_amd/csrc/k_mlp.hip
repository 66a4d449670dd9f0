// Per-atom MLP of the learned SNAP surrogate on gfx950: x (<= 16 standardized bispectrum components)
// -> 64 -> 64 -> 1 with SiLU, BF16 weights and activations on the matrix cores (v_mfma_f32_16x16x16_bf16),
// FP32 accumulation.  No upstream counterpart (the reference's potentials are linear/quadratic SNAP).
//
// Every layer is computed transposed, D^T = W^T X^T with the atoms as the N dimension: in the MFMA
// register layout a lane (r = lane % 16, q = lane / 16) then holds output rows 4q..4q+3 of atom r,
// which is exactly the B-operand layout (K rows 4q..4q+3, column r) the next layer needs, so layers
// chain in registers with no data movement.  Weights are pre-swizzled once per update into the
// A-operand layout of each 16 x 16 tile (k_mlp_pack), 40 tiles of 64 lanes x 4 BF16:
//   tiles  0..3  A1[m]      = W1^T tile (hidden 16m.., inputs)      lane: W1[4q+i][16m+r]
//   tiles  4..19 A2[m'][m]  = W2^T tile (h2 16m'.., h1 16m..)       lane: W2[16m+4q+i][16m'+r]
//   tiles 20..35 B2[m][m']  = W2   tile (h1 16m.., h2 16m'..)       lane: W2[16m+r][16m'+4q+i]
//   tiles 36..39 B1[m]      = W1   tile (inputs, hidden 16m..)      lane: W1[r][16m+4q+i]
// Parameters (FP32 master copy): W1[16][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.
// Descriptors and outputs are component-major: x[k*n + atom], g[k*n + atom].
#include <hip/hip_runtime.h>

#include "mlp_launch.hpp"

namespace mdp {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int OFF_W1 = 0, OFF_B1 = 1024, OFF_W2 = 1088, OFF_B2 = 5184, OFF_W3 = 5248, OFF_B3 = 5312;
static_assert(OFF_B3 + 1 == MLP_NPARAM, "parameter layout");

__device__ __forceinline__ unsigned short f2bf(float f) {   // round to nearest even
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ s4 pack4(float a, float b, float c, float d) {
    s4 v;
    v.x = static_cast<short>(f2bf(a));
    v.y = static_cast<short>(f2bf(b));
    v.z = static_cast<short>(f2bf(c));
    v.w = static_cast<short>(f2bf(d));
    return v;
}
__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

}  // namespace

// ---- packing and optimiser -------------------------------------------------------------------------

__global__ void k_mlp_pack(const float* __restrict__ p, unsigned short* __restrict__ q) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= MLP_NPACK) return;
    const int t = idx / 256, lane = (idx / 4) % 64, i = idx % 4;
    const int r = lane % 16, qq = lane / 16;
    float v;
    if (t < 4) {
        const int m = t;
        v = p[OFF_W1 + (4 * qq + i) * 64 + 16 * m + r];
    } else if (t < 20) {
        const int mp = (t - 4) / 4, m = (t - 4) % 4;
        v = p[OFF_W2 + (16 * m + 4 * qq + i) * 64 + 16 * mp + r];
    } else if (t < 36) {
        const int m = (t - 20) / 4, mp = (t - 20) % 4;
        v = p[OFF_W2 + (16 * m + r) * 64 + 16 * mp + 4 * qq + i];
    } else {
        const int m = t - 36;
        v = p[OFF_W1 + r * 64 + 16 * m + 4 * qq + i];
    }
    q[idx] = f2bf(v);
}

__global__ void k_mlp_reduce(const float* __restrict__ part, int nparts, float* __restrict__ grad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= MLP_NPARAM) return;
    float s = 0.0f;
    for (int w = 0; w < nparts; ++w) s += part[static_cast<size_t>(w) * MLP_NPARAM + j];
    grad[j] = s;
}

// Adam with bias correction (Kingma & Ba 2015, Algorithm 1) and decoupled weight decay on the weights.
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                           float wd) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= MLP_NPARAM) return;
    const float gj = g[j];
    const float mj = b1 * m[j] + (1.0f - b1) * gj;
    const float vj = b2 * v[j] + (1.0f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    const bool weight = (j < OFF_B1) || (j >= OFF_W2 && j < OFF_B2) || (j >= OFF_W3 && j < OFF_B3);
    float pj = p[j];
    if (weight) pj -= lr * wd * pj;
    p[j] = pj - lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

// ---- evaluation: energy and its gradient with respect to the descriptors -----------------------

struct Weights {
    s4 A1[4], A2[16], B2[16], B1[4];
    float b1[16], b2[16], w3[16], b3;
};

template <bool NEED_B1>
__device__ __forceinline__ void load_weights(const unsigned short* __restrict__ q, const float* __restrict__ p, int lane,
                                             Weights& w) {
    const s4* __restrict__ qt = reinterpret_cast<const s4*>(q);
    const int qq = lane / 16;
#pragma unroll
    for (int m = 0; m < 4; ++m) w.A1[m] = qt[m * 64 + lane];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        w.A2[t] = qt[(4 + t) * 64 + lane];
        w.B2[t] = qt[(20 + t) * 64 + lane];
    }
    if constexpr (NEED_B1) {
#pragma unroll
        for (int m = 0; m < 4; ++m) w.B1[m] = qt[(36 + m) * 64 + lane];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int h = 16 * m + 4 * qq + i;
            w.b1[4 * m + i] = p[OFF_B1 + h];
            w.b2[4 * m + i] = p[OFF_B2 + h];
            w.w3[4 * m + i] = p[OFF_W3 + h];
        }
    w.b3 = p[OFF_B3];
}

// aff: mu[16] | 1/sigma[16] (0 for unused inputs) | e_scale | e_shift:
//   x^ = (B - mu) / sigma,  E_atom = e_shift + e_scale * f(x^),  g = dE_atom/dB = e_scale/sigma * df/dx^.
// One wave per 16-atom tile, grid-stride over tiles.
__global__ __launch_bounds__(256) void k_mlp_eval(const unsigned short* __restrict__ q, const float* __restrict__ p,
                                                  const double* __restrict__ aff, int nin,
                                                  const double* __restrict__ X, int n, double* __restrict__ e,
                                                  double* __restrict__ g) {
    const int lane = threadIdx.x & 63, r = lane & 15, qq = lane >> 4;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    Weights w;
    load_weights<true>(q, p, lane, w);
    double mu[4], sc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * qq + i;
        mu[i] = k < nin ? aff[k] : 0.0;
        sc[i] = k < nin ? aff[16 + k] : 0.0;
    }
    const double escale = aff[32], eshift = aff[33];
    const int ntiles = (n + 15) / 16;
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int t = wave; t < ntiles; t += nwaves) {
        const int a = t * 16 + r;
        const bool live = a < n;
        float xv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = 4 * qq + i;
            xv[i] = (live && k < nin) ? static_cast<float>((X[static_cast<size_t>(k) * n + a] - mu[i]) * sc[i]) : 0.0f;
        }
        const s4 bx = pack4(xv[0], xv[1], xv[2], xv[3]);
        float z1[16], s1[16];
        s4 h1b[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const f4 d = mfma(w.A1[m], bx, zero);
            float h[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float z = d[i] + w.b1[4 * m + i];
                const float s = sigm(z);
                z1[4 * m + i] = z;
                s1[4 * m + i] = s;
                h[i] = z * s;
            }
            h1b[m] = pack4(h[0], h[1], h[2], h[3]);
        }
        f4 d2[4];
#pragma unroll
        for (int mp = 0; mp < 4; ++mp) {
            d2[mp] = zero;
#pragma unroll
            for (int m = 0; m < 4; ++m) d2[mp] = mfma(w.A2[4 * mp + m], h1b[m], d2[mp]);
        }
        float part = 0.0f;
        float dz2[16];
#pragma unroll
        for (int mp = 0; mp < 4; ++mp)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float z = d2[mp][i] + w.b2[4 * mp + i];
                const float s = sigm(z);
                part += w.w3[4 * mp + i] * (z * s);
                dz2[4 * mp + i] = w.w3[4 * mp + i] * s * (1.0f + z * (1.0f - s));
            }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        if (qq == 0 && live) e[a] = eshift + escale * static_cast<double>(part + w.b3);
        if (g) {
            s4 d2b[4];
#pragma unroll
            for (int mp = 0; mp < 4; ++mp)
                d2b[mp] = pack4(dz2[4 * mp], dz2[4 * mp + 1], dz2[4 * mp + 2], dz2[4 * mp + 3]);
            s4 d1b[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                f4 acc = zero;
#pragma unroll
                for (int mp = 0; mp < 4; ++mp) acc = mfma(w.B2[4 * m + mp], d2b[mp], acc);
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float z = z1[4 * m + i], s = s1[4 * m + i];
                    v[i] = acc[i] * s * (1.0f + z * (1.0f - s));
                }
                d1b[m] = pack4(v[0], v[1], v[2], v[3]);
            }
            f4 gg = zero;
#pragma unroll
            for (int m = 0; m < 4; ++m) gg = mfma(w.B1[m], d1b[m], gg);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 4 * qq + i;
                if (live && k < nin) g[static_cast<size_t>(k) * n + a] = escale * sc[i] * static_cast<double>(gg[i]);
            }
        }
    }
}

// ---- training: gradient of  sum_a [ alpha_a E_a + u_a . g_a ]  with respect to the parameters ------
//
// alpha_a = dL/dE_a (energy loss), u_a = dL/dg_a (force loss, in descriptor units).  The u term is the
// derivative of the directional derivative u . grad_x f, obtained by one tangent-mode pass along
// u^ = e_scale u / sigma followed by one reverse pass through both the primal and the tangent graph:
//   z1 = W1^T x + b1, h1 = s(z1);  z1' = W1^T u^, h1' = s'(z1) z1'
//   z2 = W2^T h1 + b2, h2 = s(z2); z2' = W2^T h1', h2' = s'(z2) z2';  f = w3.h2 + b3
//   adjoints: D2 = w3 s'(z2) (of z2'), P2 = a' w3 s'(z2) + w3 s''(z2) z2' (of z2),
//             D1 = s'(z1) W2 D2, P1 = s'(z1) W2 P2 + s''(z1) z1' (W2 D2)
//   dW1 = x P1^T + u^ D1^T, db1 = P1, dW2 = h1 P2^T + h1' D2^T, db2 = P2, dw3 = a' h2 + h2', db3 = a'
// with a' = e_scale alpha.  The sums over atoms of dW1/dW2 run on the matrix cores after an LDS
// transpose of the per-tile activations (atoms become the K dimension).  Each wave writes its partial
// gradient to part[wave][MLP_NPARAM]; k_mlp_reduce sums them in a fixed order.
__global__ __launch_bounds__(64) void k_mlp_train(const unsigned short* __restrict__ q, const float* __restrict__ p,
                                                  const double* __restrict__ aff, int nin,
                                                  const double* __restrict__ X, const double* __restrict__ U,
                                                  const double* __restrict__ alpha, int n, float* __restrict__ part) {
    const int lane = threadIdx.x, r = lane & 15, qq = lane >> 4;
    const int wave = blockIdx.x, nwaves = gridDim.x;
    __shared__ unsigned short sH1[64 * 16], sH1t[64 * 16], sP2[64 * 16], sD2[64 * 16], sP1[64 * 16], sD1[64 * 16];
    __shared__ unsigned short sX[16 * 16], sU[16 * 16];
    Weights w;
    load_weights<false>(q, p, lane, w);
    double mu[4], sc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * qq + i;
        mu[i] = k < nin ? aff[k] : 0.0;
        sc[i] = k < nin ? aff[16 + k] : 0.0;
    }
    const float escale = static_cast<float>(aff[32]);
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f4 gW2[16], gW1[4];
#pragma unroll
    for (int t = 0; t < 16; ++t) gW2[t] = zero;
#pragma unroll
    for (int t = 0; t < 4; ++t) gW1[t] = zero;
    float gb1[16], gb2[16], gw3[16], gb3 = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) { gb1[k] = 0.0f; gb2[k] = 0.0f; gw3[k] = 0.0f; }

    const int ntiles = (n + 15) / 16;
    for (int t = wave; t < ntiles; t += nwaves) {
        const int a = t * 16 + r;
        const bool live = a < n;
        float xv[4], uv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = 4 * qq + i;
            const bool ok = live && k < nin;
            xv[i] = ok ? static_cast<float>((X[static_cast<size_t>(k) * n + a] - mu[i]) * sc[i]) : 0.0f;
            uv[i] = ok ? static_cast<float>(U[static_cast<size_t>(k) * n + a] * sc[i]) * escale : 0.0f;
        }
        const float ap = live ? escale * static_cast<float>(alpha[a]) : 0.0f;
        const s4 bx = pack4(xv[0], xv[1], xv[2], xv[3]);
        const s4 bu = pack4(uv[0], uv[1], uv[2], uv[3]);
        // layer 1, primal and tangent
        float z1[16], s1[16], t1[16];
        s4 h1b[4], h1tb[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const f4 d = mfma(w.A1[m], bx, zero);
            const f4 dt = mfma(w.A1[m], bu, zero);
            float h[4], ht[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float z = d[i] + w.b1[4 * m + i];
                const float s = sigm(z);
                z1[4 * m + i] = z;
                s1[4 * m + i] = s;
                t1[4 * m + i] = dt[i];
                h[i] = z * s;
                ht[i] = s * (1.0f + z * (1.0f - s)) * dt[i];
            }
            h1b[m] = pack4(h[0], h[1], h[2], h[3]);
            h1tb[m] = pack4(ht[0], ht[1], ht[2], ht[3]);
        }
        // layer 2, primal and tangent; adjoints of z2 and z2'
        float p2[16], dd2[16];
#pragma unroll
        for (int mp = 0; mp < 4; ++mp) {
            f4 d = zero, dt = zero;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                d = mfma(w.A2[4 * mp + m], h1b[m], d);
                dt = mfma(w.A2[4 * mp + m], h1tb[m], dt);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 4 * mp + i;
                const float z = d[i] + w.b2[k];
                const float s = sigm(z);
                const float s1p = s * (1.0f + z * (1.0f - s));
                const float s2p = s * (1.0f - s) * (2.0f + z * (1.0f - 2.0f * s));
                dd2[k] = w.w3[k] * s1p;
                p2[k] = ap * w.w3[k] * s1p + w.w3[k] * s2p * dt[i];
                gw3[k] += ap * (z * s) + s1p * dt[i];
                gb2[k] += p2[k];
            }
        }
        gb3 += (qq == 0) ? ap : 0.0f;
        s4 p2b[4], d2b[4];
#pragma unroll
        for (int mp = 0; mp < 4; ++mp) {
            p2b[mp] = pack4(p2[4 * mp], p2[4 * mp + 1], p2[4 * mp + 2], p2[4 * mp + 3]);
            d2b[mp] = pack4(dd2[4 * mp], dd2[4 * mp + 1], dd2[4 * mp + 2], dd2[4 * mp + 3]);
        }
        // adjoints of z1 and z1'
        float p1[16], dd1[16];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f4 ah = zero, aht = zero;
#pragma unroll
            for (int mp = 0; mp < 4; ++mp) {
                ah = mfma(w.B2[4 * m + mp], p2b[mp], ah);
                aht = mfma(w.B2[4 * m + mp], d2b[mp], aht);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 4 * m + i;
                const float z = z1[k], s = s1[k];
                const float s1p = s * (1.0f + z * (1.0f - s));
                const float s2p = s * (1.0f - s) * (2.0f + z * (1.0f - 2.0f * s));
                dd1[k] = s1p * aht[i];
                p1[k] = s1p * ah[i] + s2p * t1[k] * aht[i];
                gb1[k] += p1[k];
            }
        }
        // stage the tile transposed ([unit][atom]) so that atoms become the K dimension
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int h = 16 * m + 4 * qq + i, k = 4 * m + i;
                sH1[h * 16 + r] = static_cast<unsigned short>(h1b[m][i]);
                sH1t[h * 16 + r] = static_cast<unsigned short>(h1tb[m][i]);
                sP2[h * 16 + r] = static_cast<unsigned short>(p2b[m][i]);
                sD2[h * 16 + r] = static_cast<unsigned short>(d2b[m][i]);
                sP1[h * 16 + r] = f2bf(p1[k]);
                sD1[h * 16 + r] = f2bf(dd1[k]);
            }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            sX[(4 * qq + i) * 16 + r] = static_cast<unsigned short>(bx[i]);
            sU[(4 * qq + i) * 16 + r] = static_cast<unsigned short>(bu[i]);
        }
        __syncthreads();
        // dW2[h1][h2] += sum_a H1[a][h1] P2[a][h2] + H1'[a][h1] D2[a][h2]
        const s4* vH1 = reinterpret_cast<const s4*>(sH1);
        const s4* vH1t = reinterpret_cast<const s4*>(sH1t);
        const s4* vP2 = reinterpret_cast<const s4*>(sP2);
        const s4* vD2 = reinterpret_cast<const s4*>(sD2);
        const s4* vP1 = reinterpret_cast<const s4*>(sP1);
        const s4* vD1 = reinterpret_cast<const s4*>(sD1);
        const s4* vX = reinterpret_cast<const s4*>(sX);
        const s4* vU = reinterpret_cast<const s4*>(sU);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const s4 ah = vH1[(16 * m + r) * 4 + qq], aht = vH1t[(16 * m + r) * 4 + qq];
#pragma unroll
            for (int mp = 0; mp < 4; ++mp) {
                gW2[4 * m + mp] = mfma(ah, vP2[(16 * mp + r) * 4 + qq], gW2[4 * m + mp]);
                gW2[4 * m + mp] = mfma(aht, vD2[(16 * mp + r) * 4 + qq], gW2[4 * m + mp]);
            }
        }
        // dW1[k][h] += sum_a x^[a][k] P1[a][h] + u^[a][k] D1[a][h]
        const s4 ax = vX[r * 4 + qq], au = vU[r * 4 + qq];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            gW1[m] = mfma(ax, vP1[(16 * m + r) * 4 + qq], gW1[m]);
            gW1[m] = mfma(au, vD1[(16 * m + r) * 4 + qq], gW1[m]);
        }
    }
    // per-unit sums over the 16 atoms columns (lanes with equal q), then one partial per wave
#pragma unroll
    for (int k = 0; k < 16; ++k)
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            gb1[k] += __shfl_xor(gb1[k], o);
            gb2[k] += __shfl_xor(gb2[k], o);
            gw3[k] += __shfl_xor(gw3[k], o);
        }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) gb3 += __shfl_xor(gb3, o);
    float* __restrict__ out = part + static_cast<size_t>(wave) * MLP_NPARAM;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            // C layout: lane holds row 4q+i, column r of each tile
            out[OFF_W1 + (4 * qq + i) * 64 + 16 * m + r] = gW1[m][i];
#pragma unroll
            for (int mp = 0; mp < 4; ++mp)
                out[OFF_W2 + (16 * m + 4 * qq + i) * 64 + 16 * mp + r] = gW2[4 * m + mp][i];
            if (r == 0) {
                const int h = 16 * m + 4 * qq + i;
                out[OFF_B1 + h] = gb1[4 * m + i];
                out[OFF_B2 + h] = gb2[4 * m + i];
                out[OFF_W3 + h] = gw3[4 * m + i];
            }
        }
    if (lane == 0) out[OFF_B3] = gb3;
}

// ---- launchers ------------------------------------------------------------------------------------

void launch_mlp_pack(const float* p, unsigned short* q, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_NPACK + 255) / 256), dim3(256), 0, st, p, q);
}

void launch_mlp_eval(const unsigned short* q, const float* p, const double* aff, int nin, const double* X, int n,
                     double* e, double* g, hipStream_t st) {
    if (n <= 0) return;
    const int tiles = (n + 15) / 16;
    const int blocks = std::min((tiles + 3) / 4, 2048);
    hipLaunchKernelGGL(k_mlp_eval, dim3(blocks), dim3(256), 0, st, q, p, aff, nin, X, n, e, g);
}

void launch_mlp_train(const unsigned short* q, const float* p, const double* aff, int nin, const double* X,
                      const double* U, const double* alpha, int n, float* part, int nwaves, float* grad,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_train, dim3(nwaves), dim3(64), 0, st, q, p, aff, nin, X, U, alpha, n, part);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, part, nwaves, grad);
}

void launch_mlp_adam(float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps, int t,
                     float wd, hipStream_t st) {
    const float c1 = 1.0f - powf(b1, static_cast<float>(t));
    const float c2 = 1.0f - powf(b2, static_cast<float>(t));
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, p, g, m, v, lr, b1, b2, eps,
                       c1, c2, wd);
}

}  // namespace mdp
