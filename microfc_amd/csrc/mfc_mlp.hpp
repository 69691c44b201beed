// mfc_mlp.hpp -- the learned WENO5 weights: a 5 -> 32 -> 32 -> 3 MLP evaluated
// and trained with BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulate) on
// tiles of 16 faces per wave.
//
// Model.  For a face, the 5-cell stencil v[0..4] (cells j-2..j+2, oriented so
// the face is j+1/2) is normalised as x_k = (v_k - v_2)/s with s the largest
// neighbour difference.  The network maps x to logits z (3); the weights are
// the classical WENO5 weights modulated by them,
//     w_k = w_plain_k exp(g z_k) / sum_m w_plain_m exp(g z_m),
// with g in [0, 1] a gate on how far the plain weights are from ideal
// (mlp_gate), so a zero last layer, or smooth resolved flow, reproduces
// plain WENO5, and the plain weights are the per-face fallback.  The face value is v_2 + s * sum_k w_k p_k with p_k the
// normalised candidate polynomials.  Training minimises the mean of
// (sum_k w_k p_k - t)^2 with t the normalised fine-grid face value.
//
// MFMA lane layout (16x16x16, wave64, checked on gfx950): A[i][k] is element
// k%4 of lane i + 16*(k/4); B[k][j] is element k%4 of lane j + 16*(k/4);
// D[i][j] is element i%4 of lane j + 16*(i/4).
#pragma once

#include <hip/hip_runtime.h>

namespace mfc {

constexpr int kMlpIn = 5, kMlpH = 32, kMlpOut = 3;
// Parameter block (FP32 master copy; padded entries stay zero):
// W1[16][32] (rows >= 5 zero), b1[32], W2[32][32], b2[32], W3[32][16]
// (columns >= 3 zero), b3[16].
constexpr int kOffW1 = 0, kOffB1 = 512, kOffW2 = 544, kOffB2 = 1568, kOffW3 = 1600, kOffB3 = 2112;
constexpr int kMlpParams = 2128;
// Training sample: x[5], p[3], w_plain[3], t  (structure of arrays, 12 planes)
constexpr int kSampleFields = 12;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 with round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// BF16 copies of the weights used as MFMA operands, laid out so each lane
// loads its B fragment with one 8-byte read: frag[tile][lane] = 4 bf16.
// W1: 2 N-tiles (K = 16).  W2: 2 K-chunks x 2 N-tiles.  W3: 2 K-chunks x 1
// N-tile.  Transposed copies for the backward pass: W3T (K = 16 outputs,
// 2 N-tiles of hidden units) and W2T (2 K-chunks x 2 N-tiles).
struct MlpFrags {
    bf16x4 w1[2][64];
    bf16x4 w2[2][2][64];
    bf16x4 w3[2][64];
    bf16x4 w3t[2][64];
    bf16x4 w2t[2][2][64];
    float b1[32], b2[32], b3[16];
    // Transposed inference (activations as units x samples, see
    // mlp_infer_lane): W3^T with its three output rows moved to rows
    // 4q..4q+2 for lane group q, and the biases as accumulator fragments.
    bf16x4 w3q[4][2][64];
    f32x4 c1[2][64], c2[2][64];
};

// The fragments one lane needs, held in registers for a whole kernel.
struct MlpRegs {
    bf16x4 w1[2], w2[2][2], w3[2], w3t[2], w2t[2][2];
    float b1[2], b2[2], b3; // b1[16n + j], b2[16n + j], b3[j] with j = lane % 16
};

__device__ __forceinline__ void mlp_load_regs(const MlpFrags *F, int lane, MlpRegs &R) {
    const int j = lane & 15;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        R.w1[n] = F->w1[n][lane];
        R.w3[n] = F->w3[n][lane];
        R.w3t[n] = F->w3t[n][lane];
        R.b1[n] = F->b1[16 * n + j];
        R.b2[n] = F->b2[16 * n + j];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            R.w2[c][n] = F->w2[c][n][lane];
            R.w2t[c][n] = F->w2t[c][n][lane];
        }
    }
    R.b3 = F->b3[j];
}

// Orders the LDS traffic of one wave: stores issued before are visible to
// loads issued after, by any lane of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A fragment of rows [16] x cols [16*c, 16*c+16) of a row-major bf16 tile.
template <int COLS>
__device__ __forceinline__ bf16x4 load_a(const unsigned short (*t)[COLS], int lane, int c) {
    const int i = lane & 15, k0 = 16 * c + 4 * (lane >> 4);
    bf16x4 a;
#pragma unroll
    for (int e = 0; e < 4; ++e) a[e] = (short)t[i][k0 + e];
    return a;
}
// A fragment of the TRANSPOSE of a row-major tile t[16 samples][COLS]:
// A[i][k] = t[k][16*m + i] (rows = columns of t, K = samples).
template <int COLS>
__device__ __forceinline__ bf16x4 load_at(const unsigned short (*t)[COLS], int lane, int m) {
    const int i = lane & 15, k0 = 4 * (lane >> 4);
    bf16x4 a;
#pragma unroll
    for (int e = 0; e < 4; ++e) a[e] = (short)t[k0 + e][16 * m + i];
    return a;
}
// B fragment of rows [16 samples] x cols [16*n, 16*n+16) of a row-major tile
// (B[k][j] = t[k][16n + j], K = samples).
template <int COLS>
__device__ __forceinline__ bf16x4 load_b(const unsigned short (*t)[COLS], int lane, int n) {
    const int j = lane & 15, k0 = 4 * (lane >> 4);
    bf16x4 b;
#pragma unroll
    for (int e = 0; e < 4; ++e) b[e] = (short)t[k0 + e][16 * n + j];
    return b;
}
// Stores a D fragment (16 x 16, columns 16*n..) into a row-major bf16 tile.
template <int COLS>
__device__ __forceinline__ void store_d_bf16(unsigned short (*t)[COLS], int lane, int n, f32x4 d) {
    const int j = lane & 15, i0 = 4 * (lane >> 4);
#pragma unroll
    for (int e = 0; e < 4; ++e) t[i0 + e][16 * n + j] = f2bf(d[e]);
}

// Forward pass of one tile of 16 samples: x (bf16, [16][16], features in
// columns 0..4) -> h1, h2 (bf16 in LDS, FP32 in st, D layout) -> logits
// written to z[16][4] (columns 0..2).
struct MlpFwdState {
    f32x4 a1[2], a2[2];
};

struct MlpActLds {
    unsigned short h1[16][32];
    unsigned short h2[16][32];
};

__device__ __forceinline__ void mlp_forward_tile(const MlpRegs &R, const unsigned short (*x)[16], MlpActLds &A,
                                                 float (*z)[4], int lane, MlpFwdState &st) {
    const int j = lane & 15;
    wave_lds_sync();
    // layer 1: [16 x 16] x [16 x 32]
    const bf16x4 ax = load_a<16>(x, lane, 0);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = mfma16(ax, R.w1[n], d);
#pragma unroll
        for (int e = 0; e < 4; ++e) d[e] = fmaxf(d[e] + R.b1[n], 0.f);
        st.a1[n] = d;
        store_d_bf16<32>(A.h1, lane, n, d);
    }
    wave_lds_sync();
    // layer 2: [16 x 32] x [32 x 32]
    const bf16x4 a20 = load_a<32>(A.h1, lane, 0), a21 = load_a<32>(A.h1, lane, 1);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = mfma16(a20, R.w2[0][n], d);
        d = mfma16(a21, R.w2[1][n], d);
#pragma unroll
        for (int e = 0; e < 4; ++e) d[e] = fmaxf(d[e] + R.b2[n], 0.f);
        st.a2[n] = d;
        store_d_bf16<32>(A.h2, lane, n, d);
    }
    wave_lds_sync();
    // layer 3: [16 x 32] x [32 x 16]; only columns 0..2 are outputs
    const bf16x4 a30 = load_a<32>(A.h2, lane, 0), a31 = load_a<32>(A.h2, lane, 1);
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
    d = mfma16(a30, R.w3[0], d);
    d = mfma16(a31, R.w3[1], d);
    const int i0 = 4 * (lane >> 4);
    if (j < 4) {
#pragma unroll
        for (int e = 0; e < 4; ++e) z[i0 + e][j] = d[e] + R.b3;
    }
    wave_lds_sync();
}

// ---------------------------------------------------------------------------
// Transposed inference for the sweeps.  With activations stored as units x
// samples (H^T = W^T X^T), the D fragment of one layer -- unit i of sample j
// in element i%4 of lane j + 16*(i/4) -- is, after ReLU and BF16 rounding,
// exactly the B fragment (row k = unit) of the next layer, so the three
// layers chain in registers.  The weight A fragments are the ones of the
// sample-major layout (the A and B lane layouts coincide).  For the output
// layer, lane group q = lane/16 uses W3^T with its three rows moved to rows
// 4q..4q+2: the logits of the sample owned by lane 16q + j then land in
// elements 0..2 of that lane's accumulator, and the other groups' passes add
// exact zeros there, so one accumulator per side collects every lane's own
// logits over the four groups.
struct MlpTRegs {
    bf16x4 w1[2], w2[2][2], w3q[4][2];
    f32x4 c1[2], c2[2], c3;
};

__device__ __forceinline__ void mlp_load_tregs(const MlpFrags *F, int lane, MlpTRegs &R) {
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        R.w1[n] = F->w1[n][lane];
        R.c1[n] = F->c1[n][lane];
        R.c2[n] = F->c2[n][lane];
#pragma unroll
        for (int c = 0; c < 2; ++c) R.w2[c][n] = F->w2[c][n][lane];
#pragma unroll
        for (int q = 0; q < 4; ++q) R.w3q[q][n] = F->w3q[q][n][lane];
    }
    R.c3 = f32x4{F->b3[0], F->b3[1], F->b3[2], 0.f};
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// ReLU and BF16 (round to nearest even) of a D fragment: v_cvt_pk_bf16_f32,
// then the ReLU as a signed 16-bit max (a negative BF16 is a negative
// int16), which commutes with the rounding.
__device__ __forceinline__ bf16x4 relu_bf16(f32x4 d) {
    const s16x2 lo = __builtin_elementwise_max(
        __builtin_bit_cast(s16x2, __builtin_convertvector((f32x2){d[0], d[1]}, bf16x2_t)), (s16x2){0, 0});
    const s16x2 hi = __builtin_elementwise_max(
        __builtin_bit_cast(s16x2, __builtin_convertvector((f32x2){d[2], d[3]}, bf16x2_t)), (s16x2){0, 0});
    return bf16x4{lo[0], lo[1], hi[0], hi[1]};
}

// Logits of the two samples (sides) of every lane of a wave: xs[side][lane]
// holds the BF16 features (5, padded to 8) of that lane's sample; returns
// this lane's logits in z[side][0..2].  32 MFMAs per side.
__device__ __forceinline__ void mlp_infer_lane(const MlpTRegs &R, const unsigned short (*xs)[64][8], int lane,
                                               f32x4 *z) {
    const int j = lane & 15, g = lane >> 4;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        f32x4 acc = R.c3;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            bf16x4 b1 = {0, 0, 0, 0};
            if (g < 2) b1 = *reinterpret_cast<const bf16x4 *>(&xs[side][16 * q + j][4 * g]);
            const bf16x4 h0 = relu_bf16(mfma16(R.w1[0], b1, R.c1[0]));
            const bf16x4 h1 = relu_bf16(mfma16(R.w1[1], b1, R.c1[1]));
            const bf16x4 g0 = relu_bf16(mfma16(R.w2[1][0], h1, mfma16(R.w2[0][0], h0, R.c2[0])));
            const bf16x4 g1 = relu_bf16(mfma16(R.w2[1][1], h1, mfma16(R.w2[0][1], h0, R.c2[1])));
            acc = mfma16(R.w3q[q][1], g1, mfma16(R.w3q[q][0], g0, acc));
        }
        z[side] = acc;
    }
}

// Gate of the logits: with dev = sum_k |w_plain_k - d_k| and d the ideal
// weights of a uniform grid for the oriented face (0.3, 0.6, 0.1),
// g = min(1, max(0, dev - kGateOff) / kGateDev).  Where the plain weights are
// near ideal (smooth, resolved flow, in which WENO5 is already close to its
// optimal linear scheme) g is zero and the face keeps its plain weights
// exactly; near discontinuities and under-resolved features the logits act
// in full.  A wave whose faces are all closed skips the network.
constexpr float kGateOff = 0.005f, kGateDev = 0.03f;
__device__ __forceinline__ float mlp_gate(const float *wp) {
    const float dev = fabsf(wp[0] - 0.3f) + fabsf(wp[1] - 0.6f) + fabsf(wp[2] - 0.1f);
    return fminf(1.f, fmaxf(0.f, dev - kGateOff) / kGateDev);
}

// Learned weights of one face from its logits and plain weights:
// w_k = w_plain_k exp(g z_k) / sum_m w_plain_m exp(g z_m), or w = w_plain
// when g = 0.  Returns g.
__device__ __forceinline__ float mlp_weights(const float *z, const float *wp, float *w) {
    const float g = mlp_gate(wp);
    if (g == 0.f) {
        w[0] = wp[0];
        w[1] = wp[1];
        w[2] = wp[2];
        return 0.f;
    }
    const float z0 = g * z[0], z1 = g * z[1], z2 = g * z[2];
    const float zm = fmaxf(z0, fmaxf(z1, z2));
    const float u0 = wp[0] * __expf(z0 - zm), u1 = wp[1] * __expf(z1 - zm), u2 = wp[2] * __expf(z2 - zm);
    const float inv = 1.f / (u0 + u1 + u2);
    w[0] = u0 * inv;
    w[1] = u1 * inv;
    w[2] = u2 * inv;
    return g;
}

} // namespace mfc
