// mfc_mlp.hip -- training and inference kernels of the learned WENO5 weights
// (model in mfc_mlp.hpp) and the extraction of coarse/fine training pairs
// from solver fields, behind the C ABI of include/mfc_rhs.h.
#include "mfc_device.hpp"
#include "mfc_mlp.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

extern "C" {
#include "../../include/mfc_rhs.h"
}

using namespace mfc;

namespace {

constexpr int kTrainBlocksMax = 512;
constexpr int kRed = kMlpParams + 1; // gradient entries + the loss

thread_local std::string g_mlp_err;

int mfail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_mlp_err = buf;
    return code;
}

#define MHIP(expr)                                                                      \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) return mfail(-3, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct MlpState {
    bool init = false;
    float *params = nullptr, *m1 = nullptr, *m2 = nullptr;
    float *partials = nullptr, *grad = nullptr;
    MlpFrags *frags = nullptr;
    long step = 0;
    hipStream_t stream = nullptr;
} M;

// ---------------------------------------------------------------------------
// kernels

__global__ void k_mlp_pack(const float *P, MlpFrags *F) {
    const int l = threadIdx.x; // 64 lanes
    const int j = l & 15, k0 = 4 * (l >> 4);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        bf16x4 a, c, d;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            a[e] = (short)f2bf(P[kOffW1 + (k0 + e) * 32 + 16 * n + j]);
            c[e] = (short)f2bf(P[kOffW3 + (16 * n + k0 + e) * 16 + j]);
            d[e] = (short)f2bf(P[kOffW3 + (16 * n + j) * 16 + (k0 + e)]);
        }
        F->w1[n][l] = a;
        F->w3[n][l] = c;
        F->w3t[n][l] = d;
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            bf16x4 w, t;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                w[e] = (short)f2bf(P[kOffW2 + (16 * cc + k0 + e) * 32 + 16 * n + j]);
                t[e] = (short)f2bf(P[kOffW2 + (16 * n + j) * 32 + 16 * cc + k0 + e]);
            }
            F->w2[cc][n][l] = w;
            F->w2t[cc][n][l] = t;
        }
    }
    if (l < 32) {
        F->b1[l] = P[kOffB1 + l];
        F->b2[l] = P[kOffB2 + l];
    }
    if (l < 16) F->b3[l] = P[kOffB3 + l];
    // transposed inference: W3^T rows shifted to 4q..4q+2, bias fragments
    // (element e of lane l is unit 16n + 4*(l/16) + e)
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            bf16x4 w;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int o = j - 4 * q; // output row of this lane's A row
                w[e] = (short)((o >= 0 && o < kMlpOut) ? f2bf(P[kOffW3 + (16 * c + k0 + e) * 16 + o]) : 0);
            }
            F->w3q[q][c][l] = w;
        }
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        f32x4 b1, b2;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            b1[e] = P[kOffB1 + 16 * n + k0 + e];
            b2[e] = P[kOffB2 + 16 * n + k0 + e];
        }
        F->c1[n][l] = b1;
        F->c2[n][l] = b2;
    }
}

struct TrainLds {
    unsigned short x[16][16];
    MlpActLds act;
    float z[16][4];
    unsigned short dz[16][16];
    unsigned short da2[16][32];
    unsigned short da1[16][32];
};

__device__ __forceinline__ float lane16_sum(float v) {
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One fused forward + backward pass over n samples (structure of arrays,
// 12 planes of n floats: x[5], p[3], w_plain[3], t).  The loss is the mean
// of (sum_k w_k p_k - t)^2.  Each wave takes tiles of 16 samples and keeps
// the weight gradients in MFMA accumulators (FP32); block partials (gradient
// and loss) go to partials[block][kRed].
__global__ __launch_bounds__(256) void k_mlp_train(const float *S, long n, const MlpFrags *F, float *partials) {
    __shared__ TrainLds W[4];
    __shared__ float red[4][kRed];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = lane & 15, i0 = 4 * (lane >> 4);
    TrainLds &L = W[w];
    MlpRegs R;
    mlp_load_regs(F, lane, R);
    const float inv_n = 1.0f / (float)n;

    f32x4 acc1[2], acc2[2][2], acc3[2];
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        acc1[a] = zero;
        acc3[a] = zero;
        acc2[a][0] = zero;
        acc2[a][1] = zero;
    }
    float db1[2] = {0.f, 0.f}, db2[2] = {0.f, 0.f}, db3[3] = {0.f, 0.f, 0.f}, lsum = 0.f;

    const long ntile = (n + 15) / 16;
    for (long t = (long)blockIdx.x * 4 + w; t < ntile; t += (long)gridDim.x * 4) {
        const long s0 = 16 * t;
        // features, bf16, zero padded to 16 columns
        {
            const int s = lane & 15;
            const long idx = s0 + s;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = i0 + e;
                const float v = (idx < n && k < kMlpIn) ? S[(long)k * n + idx] : 0.f;
                L.x[s][k] = f2bf(v);
            }
        }
        MlpFwdState st;
        mlp_forward_tile(R, L.x, L.act, L.z, lane, st);
        // loss and d loss / d z, one lane per sample
        if (lane < 16) {
            const long idx = s0 + lane;
            float dzv[3] = {0.f, 0.f, 0.f};
            if (idx < n) {
                float z[3], wp[3], p[3], wgt[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    z[k] = L.z[lane][k];
                    p[k] = S[(long)(5 + k) * n + idx];
                    wp[k] = S[(long)(8 + k) * n + idx];
                }
                const float tt = S[11L * n + idx];
                const float g = mlp_weights(z, wp, wgt);
                const float y = wgt[0] * p[0] + wgt[1] * p[1] + wgt[2] * p[2];
                const float r = y - tt;
                lsum += r * r;
                const float dy = 2.f * r * inv_n;
                const float sdw = wgt[0] * dy * p[0] + wgt[1] * dy * p[1] + wgt[2] * dy * p[2];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    dzv[k] = g * wgt[k] * (dy * p[k] - sdw);
                    db3[k] += dzv[k];
                }
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) L.dz[lane][k] = k < 3 ? f2bf(dzv[k]) : (unsigned short)0;
        }
        wave_lds_sync();
        // dW3 += H2^T dZ ; dH2 = dZ W3^T
        const bf16x4 bz = load_b<16>(L.dz, lane, 0);
        const bf16x4 az = load_a<16>(L.dz, lane, 0);
#pragma unroll
        for (int m = 0; m < 2; ++m) acc3[m] = mfma16(load_at<32>(L.act.h2, lane, m), bz, acc3[m]);
#pragma unroll
        for (int nn = 0; nn < 2; ++nn) {
            f32x4 d = mfma16(az, R.w3t[nn], zero);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                d[e] = st.a2[nn][e] > 0.f ? d[e] : 0.f;
                db2[nn] += d[e];
            }
            store_d_bf16<32>(L.da2, lane, nn, d);
        }
        wave_lds_sync();
        // dW2 += H1^T dA2 ; dH1 = dA2 W2^T
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int nn = 0; nn < 2; ++nn)
                acc2[m][nn] = mfma16(load_at<32>(L.act.h1, lane, m), load_b<32>(L.da2, lane, nn), acc2[m][nn]);
        const bf16x4 ad0 = load_a<32>(L.da2, lane, 0), ad1 = load_a<32>(L.da2, lane, 1);
#pragma unroll
        for (int nn = 0; nn < 2; ++nn) {
            f32x4 d = mfma16(ad0, R.w2t[0][nn], zero);
            d = mfma16(ad1, R.w2t[1][nn], d);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                d[e] = st.a1[nn][e] > 0.f ? d[e] : 0.f;
                db1[nn] += d[e];
            }
            store_d_bf16<32>(L.da1, lane, nn, d);
        }
        wave_lds_sync();
        // dW1 += X^T dA1
        const bf16x4 axt = load_at<16>(L.x, lane, 0);
#pragma unroll
        for (int nn = 0; nn < 2; ++nn) acc1[nn] = mfma16(axt, load_b<32>(L.da1, lane, nn), acc1[nn]);
        wave_lds_sync();
    }

    // per-wave partials into LDS (every entry written exactly once)
    float *rw = red[w];
#pragma unroll
    for (int nn = 0; nn < 2; ++nn)
#pragma unroll
        for (int e = 0; e < 4; ++e) rw[kOffW1 + (i0 + e) * 32 + 16 * nn + j] = acc1[nn][e];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int nn = 0; nn < 2; ++nn)
#pragma unroll
            for (int e = 0; e < 4; ++e) rw[kOffW2 + (16 * m + i0 + e) * 32 + 16 * nn + j] = acc2[m][nn][e];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e) rw[kOffW3 + (16 * m + i0 + e) * 16 + j] = acc3[m][e];
#pragma unroll
    for (int nn = 0; nn < 2; ++nn) {
        float v1 = db1[nn], v2 = db2[nn];
        v1 += __shfl_xor(v1, 16, 64);
        v1 += __shfl_xor(v1, 32, 64);
        v2 += __shfl_xor(v2, 16, 64);
        v2 += __shfl_xor(v2, 32, 64);
        if (lane < 16) {
            rw[kOffB1 + 16 * nn + lane] = v1;
            rw[kOffB2 + 16 * nn + lane] = v2;
        }
    }
    float b3s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) b3s[k] = lane16_sum(lane < 16 ? db3[k] : 0.f);
    const float ls = lane16_sum(lane < 16 ? lsum : 0.f);
    if (lane < 16) rw[kOffB3 + lane] = lane < 3 ? b3s[lane < 3 ? lane : 0] : 0.f;
    if (lane == 0) rw[kMlpParams] = ls;
    __syncthreads();
    for (int p = threadIdx.x; p < kRed; p += 256)
        partials[(long)blockIdx.x * kRed + p] = red[0][p] + red[1][p] + red[2][p] + red[3][p];
}

// Sums block partials in block order: grad[p] = sum_b partials[b][p];
// grad[kMlpParams] is the loss (mean of squared errors).
__global__ void k_mlp_reduce(const float *partials, int nblk, long n, float *grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kRed) return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += partials[(long)b * kRed + p];
    grad[p] = p == kMlpParams ? s / (float)n : s;
}

__global__ void k_mlp_adam(const float *grad, float *P, float *m1, float *m2, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    const float a = b1 * m1[p] + (1.f - b1) * g;
    const float v = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = a;
    m2[p] = v;
    P[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// Inference: out has 4 planes of n floats (w0, w1, w2, learned face value in
// normalised units).
__global__ __launch_bounds__(256) void k_mlp_infer(const float *S, long n, const MlpFrags *F, float *out) {
    __shared__ unsigned short X[4][16][16];
    __shared__ MlpActLds A[4];
    __shared__ float Z[4][16][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = 4 * (lane >> 4);
    MlpRegs R;
    mlp_load_regs(F, lane, R);
    const long ntile = (n + 15) / 16;
    for (long t = (long)blockIdx.x * 4 + w; t < ntile; t += (long)gridDim.x * 4) {
        const long s0 = 16 * t;
        const int s = lane & 15;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = i0 + e;
            X[w][s][k] = f2bf((s0 + s < n && k < kMlpIn) ? S[(long)k * n + s0 + s] : 0.f);
        }
        MlpFwdState st;
        mlp_forward_tile(R, X[w], A[w], Z[w], lane, st);
        if (lane < 16 && s0 + lane < n) {
            const long idx = s0 + lane;
            float z[3], wp[3], p[3], wgt[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                z[k] = Z[w][lane][k];
                p[k] = S[(long)(5 + k) * n + idx];
                wp[k] = S[(long)(8 + k) * n + idx];
            }
            mlp_weights(z, wp, wgt);
#pragma unroll
            for (int k = 0; k < 3; ++k) out[(long)k * n + idx] = wgt[k];
            out[3L * n + idx] = wgt[0] * p[0] + wgt[1] * p[1] + wgt[2] * p[2];
        }
        wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------
// coarse/fine training pairs

struct Coef27 {
    double c[kNCoef];
};

// Right-face candidates p_k and plain weights w_k of a 5-cell stencil (FP64,
// one set of WENO5 coefficients; m_weno.fpp:476-531 for the right side).
__device__ __forceinline__ void weno_r_parts(const double *v, const Coef27 &cf, double eps, double *p, double *w) {
    const double d1 = v[4] - v[3], d0 = v[3] - v[2], dm1 = v[2] - v[1], dm2 = v[1] - v[0];
    p[0] = v[2] + cf.c[kCbR + 0] * d1 + cf.c[kCbR + 1] * d0;
    p[1] = v[2] + cf.c[kCbR + 2] * d0 + cf.c[kCbR + 3] * dm1;
    p[2] = v[2] + cf.c[kCbR + 4] * dm1 + cf.c[kCbR + 5] * dm2;
    const double b0 = cf.c[kBeta + 0] * d1 * d1 + cf.c[kBeta + 1] * d1 * d0 + cf.c[kBeta + 2] * d0 * d0 + eps;
    const double b1 = cf.c[kBeta + 3] * d0 * d0 + cf.c[kBeta + 4] * d0 * dm1 + cf.c[kBeta + 5] * dm1 * dm1 + eps;
    const double b2 = cf.c[kBeta + 6] * dm1 * dm1 + cf.c[kBeta + 7] * dm1 * dm2 + cf.c[kBeta + 8] * dm2 * dm2 + eps;
    const double a0 = cf.c[kDR + 0] / (b0 * b0), a1 = cf.c[kDR + 1] / (b1 * b1), a2 = cf.c[kDR + 2] / (b2 * b2);
    const double s = a0 + a1 + a2;
    w[0] = a0 / s;
    w[1] = a1 / s;
    w[2] = a2 / s;
}

__device__ __forceinline__ double weno_r_value(const double *v, const Coef27 &cf, double eps) {
    double p[3], w[3];
    weno_r_parts(v, cf, eps, p, w);
    return w[0] * p[0] + w[1] * p[1] + w[2] * p[2];
}

// One thread per (coarse interior cell, side).  coarse: one variable of the
// coarse ghosted field ((m+1) x (n+1) x (p+1) interior, b ghosts in every
// direction present, BC-filled); fine: the same variable on the 2x finer
// grid.  The face value target is the mean over the 1, 2 or 4 fine faces
// that tile the coarse face of the fine WENO5 face values from both sides.
// Output: 13 planes (x[5], p[3], w_plain[3], t, valid) of
// 2 (m+1)(n+1)(p+1) samples; sample 2c is the right face of coarse cell c,
// 2c+1 its left face (stencil reversed).
__global__ void k_ml_samples(int dir, int m, int n, int p, int b, const double *coarse, const double *fine,
                             Coef27 cf, double eps, float *out) {
    const long ncell = (long)(m + 1) * (n + 1) * (p + 1);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * ncell) return;
    const long c = t >> 1;
    const int side = (int)(t & 1);
    const int i = (int)(c % (m + 1)), jj = (int)((c / (m + 1)) % (n + 1)), kk = (int)(c / ((long)(m + 1) * (n + 1)));
    const int bj = n > 0 ? b : 0, bk = p > 0 ? b : 0;
    const long cnx = m + 2 * b + 1, cny = n + 2 * bj + 1;
    const long fnx = 2 * m + 1 + 2 * b + 1, fny = n > 0 ? 2 * n + 1 + 2 * b + 1 : 1;
    auto C = [&](int x, int y, int z) {
        return coarse[(long)(x + b) + cnx * ((long)(y + bj) + cny * (long)(z + bk))];
    };
    auto Fv = [&](int x, int y, int z) {
        return fine[(long)(x + b) + fnx * ((long)(y + bj) + fny * (long)(z + bk))];
    };
    // coarse stencil along dir, oriented so the face is between v[2] and v[3]
    double v[5];
    for (int k = 0; k < 5; ++k) {
        const int o = side == 0 ? k - 2 : 2 - k;
        v[k] = dir == 0 ? C(i + o, jj, kk) : dir == 1 ? C(i, jj + o, kk) : C(i, jj, kk + o);
    }
    // fine face: right face of coarse cell i is between fine cells 2i+1 and
    // 2i+2; left face between 2i and 2i-1 (oriented the same way).  Fine
    // cell at index `x` along dir and transverse offset r (bit 0: the first
    // transverse direction present, bit 1: the second).
    const int along = dir == 0 ? i : dir == 1 ? jj : kk;
    const int fa = side == 0 ? 2 * along + 1 : 2 * along; // fine cell on the v[2] side
    const int sgn = side == 0 ? 1 : -1;
    const int nrow = 1 << ((n > 0) + (p > 0));              // fine faces on the coarse face
    auto F = [&](int x, int r) {
        const int t1 = r & 1, t2 = r >> 1;
        if (dir == 0) return Fv(x, n > 0 ? 2 * jj + t1 : 0, p > 0 ? 2 * kk + t2 : 0);
        if (dir == 1) return Fv(2 * i + t1, x, p > 0 ? 2 * kk + t2 : 0);
        return Fv(2 * i + t1, 2 * jj + t2, x);
    };
    double tsum = 0.0;
    for (int r = 0; r < nrow; ++r) {
        double a[5], bb[5];
        for (int k = 0; k < 5; ++k) {
            a[k] = F(fa + sgn * (k - 2), r);
            bb[k] = F(fa + sgn * (1 + 2 - k), r);
        }
        tsum += 0.5 * (weno_r_value(a, cf, eps) + weno_r_value(bb, cf, eps));
    }
    const double tface = tsum / nrow;
    double pc[3], w[3];
    weno_r_parts(v, cf, eps, pc, w);
    double s = 0.0, vmax = 0.0;
    for (int k = 0; k < 4; ++k) s = fmax(s, fabs(v[k + 1] - v[k]));
    for (int k = 0; k < 5; ++k) vmax = fmax(vmax, fabs(v[k]));
    const bool valid = s > 1e-6 * vmax && s > 1e-300; // variation above rounding noise
    const long ns = 2 * ncell;
    const double inv = valid ? 1.0 / s : 0.0;
    for (int k = 0; k < 5; ++k) out[(long)k * ns + t] = (float)((v[k] - v[2]) * inv);
    for (int k = 0; k < 3; ++k) {
        out[(long)(5 + k) * ns + t] = (float)((pc[k] - v[2]) * inv);
        out[(long)(8 + k) * ns + t] = (float)w[k];
    }
    out[11L * ns + t] = (float)((tface - v[2]) * inv);
    out[12L * ns + t] = valid ? 1.f : 0.f;
}

// Coarse cell averages of one variable: each coarse interior cell is the
// mean of the 2 (1D), 2x2 (2D) or 2x2x2 (3D) fine cells it covers.
__global__ void k_ml_restrict(int m, int n, int p, int b, const double *fine, double *coarse) {
    const long ncell = (long)(m + 1) * (n + 1) * (p + 1);
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int i = (int)(c % (m + 1)), jj = (int)((c / (m + 1)) % (n + 1)), kk = (int)(c / ((long)(m + 1) * (n + 1)));
    const int bj = n > 0 ? b : 0, bk = p > 0 ? b : 0;
    const long cnx = m + 2 * b + 1, cny = n + 2 * bj + 1;
    const long fnx = 2 * m + 1 + 2 * b + 1, fny = n > 0 ? 2 * n + 1 + 2 * b + 1 : 1;
    auto Fv = [&](int x, int y, int z) {
        return fine[(long)(x + b) + fnx * ((long)(y + bj) + fny * (long)(z + bk))];
    };
    double v;
    if (p > 0) {
        const int x = 2 * i, y = 2 * jj, z = 2 * kk;
        v = 0.125 * (((Fv(x, y, z) + Fv(x + 1, y, z)) + (Fv(x, y + 1, z) + Fv(x + 1, y + 1, z))) +
                     ((Fv(x, y, z + 1) + Fv(x + 1, y, z + 1)) + (Fv(x, y + 1, z + 1) + Fv(x + 1, y + 1, z + 1))));
    } else if (n > 0) {
        v = 0.25 * ((Fv(2 * i, 2 * jj, 0) + Fv(2 * i + 1, 2 * jj, 0)) +
                    (Fv(2 * i, 2 * jj + 1, 0) + Fv(2 * i + 1, 2 * jj + 1, 0)));
    } else {
        v = 0.5 * (Fv(2 * i, 0, 0) + Fv(2 * i + 1, 0, 0));
    }
    coarse[(long)(i + b) + cnx * ((long)(jj + bj) + cny * (long)(kk + bk))] = v;
}

int mlp_require() {
    if (!M.init) return mfail(-4, "mfc_mlp_init has not been called");
    return 0;
}

int mlp_pack() {
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, M.stream, M.params, M.frags);
    MHIP(hipGetLastError());
    return 0;
}

int mlp_grad(const float *samples, long n, float *loss) {
    if (n <= 0) return mfail(-1, "no samples");
    const long ntile = (n + 15) / 16;
    const int nblk = (int)std::min<long>(kTrainBlocksMax, (ntile + 3) / 4);
    hipLaunchKernelGGL(k_mlp_train, dim3(nblk), dim3(256), 0, M.stream, samples, n, M.frags, M.partials);
    MHIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kRed + 255) / 256), dim3(256), 0, M.stream, M.partials, nblk, n, M.grad);
    MHIP(hipGetLastError());
    if (loss) {
        MHIP(hipMemcpyAsync(loss, M.grad + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, M.stream));
        MHIP(hipStreamSynchronize(M.stream));
    }
    return 0;
}

} // namespace

namespace mfc {
// The packed weights the learned sweep reads (nullptr before mfc_mlp_init).
const MlpFrags *mlp_frags() { return M.init ? M.frags : nullptr; }
} // namespace mfc

extern "C" {

const char *mfc_mlp_last_error(void) { return g_mlp_err.c_str(); }

int mfc_mlp_init(const float *params) {
    if (!params) return mfail(-1, "NULL parameters");
    if (!M.init) {
        MHIP(hipStreamCreateWithFlags(&M.stream, hipStreamNonBlocking));
        MHIP(hipMalloc(&M.params, sizeof(float) * kMlpParams));
        MHIP(hipMalloc(&M.m1, sizeof(float) * kMlpParams));
        MHIP(hipMalloc(&M.m2, sizeof(float) * kMlpParams));
        MHIP(hipMalloc(&M.grad, sizeof(float) * kRed));
        MHIP(hipMalloc(&M.partials, sizeof(float) * kRed * kTrainBlocksMax));
        MHIP(hipMalloc(&M.frags, sizeof(MlpFrags)));
        M.init = true;
    }
    MHIP(hipMemcpyAsync(M.params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, M.stream));
    MHIP(hipMemsetAsync(M.m1, 0, sizeof(float) * kMlpParams, M.stream));
    MHIP(hipMemsetAsync(M.m2, 0, sizeof(float) * kMlpParams, M.stream));
    M.step = 0;
    if (int rc = mlp_pack()) return rc;
    MHIP(hipStreamSynchronize(M.stream));
    return 0;
}

int mfc_mlp_get(float *params) {
    if (int rc = mlp_require()) return rc;
    MHIP(hipMemcpyAsync(params, M.params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, M.stream));
    MHIP(hipStreamSynchronize(M.stream));
    return 0;
}

int mfc_mlp_gradient(const float *samples, long n, float *grad, float *loss) {
    if (int rc = mlp_require()) return rc;
    if (int rc = mlp_grad(samples, n, loss)) return rc;
    if (grad) {
        MHIP(hipMemcpyAsync(grad, M.grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, M.stream));
        MHIP(hipStreamSynchronize(M.stream));
    }
    return 0;
}

int mfc_mlp_train_step(const float *samples, long n, float lr, float *loss) {
    if (int rc = mlp_require()) return rc;
    if (!(lr > 0.f)) return mfail(-1, "learning rate must be positive");
    if (int rc = mlp_grad(samples, n, loss)) return rc;
    M.step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)M.step), c2 = 1.f - powf(b2, (float)M.step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, M.stream, M.grad, M.params, M.m1,
                       M.m2, lr, b1, b2, 1e-8f, c1, c2);
    MHIP(hipGetLastError());
    if (int rc = mlp_pack()) return rc;
    MHIP(hipStreamSynchronize(M.stream));
    return 0;
}

int mfc_mlp_infer(const float *samples, long n, float *out) {
    if (int rc = mlp_require()) return rc;
    if (n <= 0) return 0;
    const long ntile = (n + 15) / 16;
    const int nblk = (int)std::min<long>(4096, (ntile + 3) / 4);
    hipLaunchKernelGGL(k_mlp_infer, dim3(nblk), dim3(256), 0, M.stream, samples, n, M.frags, out);
    MHIP(hipGetLastError());
    MHIP(hipStreamSynchronize(M.stream));
    return 0;
}

int mfc_ml_restrict3(int m, int n, int p, int b, const double *fine, double *coarse) {
    if (m < 0 || n < 0 || p < 0 || (p > 0 && n == 0) || b < 0 || !fine || !coarse) return mfail(-1, "bad arguments");
    hipStream_t st = M.init ? M.stream : nullptr;
    const long ncell = (long)(m + 1) * (n + 1) * (p + 1);
    hipLaunchKernelGGL(k_ml_restrict, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st, m, n, p, b, fine,
                       coarse);
    MHIP(hipGetLastError());
    MHIP(hipStreamSynchronize(st));
    return 0;
}

int mfc_ml_restrict(int m, int n, int b, const double *fine, double *coarse) {
    return mfc_ml_restrict3(m, n, 0, b, fine, coarse);
}

int mfc_ml_samples3(int dir, int m, int n, int p, int b, const double *coarse, const double *fine,
                    const double *coef27, double eps, float *out) {
    if (dir < 0 || dir > 2 || (dir == 1 && n == 0) || (dir == 2 && p == 0) || (p > 0 && n == 0) || b < 4 ||
        !coarse || !fine || !coef27 || !out)
        return mfail(-1, "bad arguments");
    Coef27 cf;
    for (int q = 0; q < kNCoef; ++q) cf.c[q] = coef27[q];
    hipStream_t st = M.init ? M.stream : nullptr;
    const long ns = 2L * (m + 1) * (n + 1) * (p + 1);
    hipLaunchKernelGGL(k_ml_samples, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, dir, m, n, p, b, coarse,
                       fine, cf, eps, out);
    MHIP(hipGetLastError());
    MHIP(hipStreamSynchronize(st));
    return 0;
}

int mfc_ml_samples(int dir, int m, int n, int b, const double *coarse, const double *fine, const double *coef27,
                   double eps, float *out) {
    if (dir > 1) return mfail(-1, "bad arguments");
    return mfc_ml_samples3(dir, m, n, 0, b, coarse, fine, coef27, eps, out);
}

} // extern "C"
