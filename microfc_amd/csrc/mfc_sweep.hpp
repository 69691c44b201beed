// mfc_sweep.hpp -- the directional sweep kernels of the RHS.
//
// One sweep per direction fuses: conservative -> primitive conversion of the
// stencil cells, WENO reconstruction of both face states, the HLLC flux and
// contact velocity, the flux divergence and the alpha*div(u) source, and --
// in the last direction of a time step -- the TVD-RK stage update.  Face
// states and fluxes never touch HBM.  Per cell and direction the HBM traffic
// is the stencil read of the S conservative variables (halo reads mostly hit
// L2) plus the rhs read-modify-write, or the RK base/output in the fused last
// sweep.
//
//  * k_sweep_x (x, the contiguous direction): one wave per tile of one grid
//    row.  The 64+2*polyn+1 stencil cells are converted to primitive and
//    staged in LDS; lane l computes face x0-1/2+l; fluxes go through LDS so
//    lane l can difference faces l and l+1 for cell x0+l (63 or 62 cells per
//    wave, face- or cell-parallel reconstruction).
//  * k_sweep_t (y or z, strided, face-parallel; FP64 with per-cell
//    coefficients, i.e. non-uniform grids): one 256-thread block per
//    tile of 16 x-columns by 16 faces along the sweep (15 cells).  Stencil
//    rows of 16 consecutive x are staged in LDS, thread (x, f) solves face f
//    of column x, and the fluxes go back through LDS for the divergence.
//  * k_sweep_tc (y or z, cell-parallel): tiles of C x-columns by K threads
//    along the sweep (K = 32 for the FP64 y sweep, else 16; C = 32 for the
//    FP32 z sweep, else 16); K lanes of a
//    wave hold one column, thread (x, c) reconstructs both faces of cell c,
//    the neighbouring face value moves one lane within the column, and the
//    fluxes go through LDS to the x-fastest layout (K - 2 cells per column).
#pragma once

#include "mfc_device.hpp"
#include "mfc_mlp.hpp"

#include <type_traits>

namespace mfc {

struct Fields {
    double *v[kMaxSys];
};

// Stage kinds of the TVD-RK update fused into the last sweep
// (m_time_steppers.fpp:129-362).  base is q_cons_ts(1), in is the stage input.
enum RkKind {
    kRkNone = 0,   // write rhs_vf
    kRkEuler = 1,  // out = base + dt*rhs
    kRk2Final = 2, // out = (base + in + dt*rhs)/2
    kRk3Stage2 = 3,// out = (3*base + in + dt*rhs)/4
    kRk3Stage3 = 4 // out = (base + 2*in + 2*dt*rhs)/3
};

struct SweepArgs {
    Fields qin, qbase, qout, rhs;
    const double *coef;   // 27 planes of ncell entries for this direction
    const double *inv_ds; // 1/ds(i), i = 0..s
    Eos eos;
    double eps, dt;
    long nxg, nyg;        // ghosted pitches in elements
    long ncell;           // s + 2b + 1
    int m, n, p, b, bj, bk;
    int bc_beg, bc_end;   // of this direction
    int rk;               // RkKind
    int accumulate;       // 0: first direction assigns rhs, 1: adds
    int base_is_in;       // the RK base is the stage input (first stages)
    // Contracted FP64 form, 2D/3D: an RK stage out = cb*base + ci*in +
    // cdt*L spread over the sweeps (rk = kRkNone in each).  The first sweep
    // stores cdt*L_d into rhs, the middle one adds cdt*L_d, the last writes
    // rhs + cdt*L_d + ci*in to qout, and the sweeps before the last add
    // cb*base, each for a range of the variables, so the last sweep does
    // not read the base.
    int fold;             // 0: off, 1: first sweep, 2: middle, 3: last
    int base_lo, base_hi; // this sweep adds cb*base of variables [base_lo, base_hi)
    double cb, ci, cdt;
    int chunk, nchunk;    // cells per tile along the sweep and tile count (k_sweep_t)
    int ntile;            // x tiles (k_sweep_x) or x column groups (k_sweep_t)
    long nblocks;
    const MlpFrags *mlp;  // learned WENO5 weights (MFC_PREC_FP32_LEARNED)
    // Uniform grids: when every cell the sweep reconstructs has the same 27
    // coefficients (bit for bit), they travel here as kernel arguments (read
    // into scalar registers) instead of being loaded per cell.
    int uniform;
    double ucf[kNCoef];
    float ucf32[kNCoef];
    // k_sweep_xy: the y direction's cell widths, boundary codes and uniform
    // coefficients (the fields above are x's)
    const double *inv_ds2;
    int bc2_beg, bc2_end;
    double ucf2[kNCoef];
    float ucf2_32[kNCoef];
};

// Threads per column of the cell-parallel y/z tile by form.  Measured on
// MI355X: in the plain FP64 form 32 (512-thread blocks, 30 cells per column)
// makes the y sweep 2-3 % faster than 16 (14 cells) through less halo
// staging; in the contracted FP64 form 16 is faster (512^3 y sweep 6.32
// against 6.83 ms, 2048^2 0.204 against 0.219 ms: with less arithmetic per
// cell, four 256-thread blocks per CU hide the finish's rhs traffic better
// than two of 512).  The z sweep, whose stencil rows are planes apart,
// gained from 32 at 256^3 but lost 6 % at 512^3 (36 planes per block) while
// it came last; first, storing rhs in the contracted form, 32 and 16 are
// level at 512^3 (5.52 against 5.53 ms).  In FP32, where the 3D sweeps near
// the HBM rate, 32 is up to 11 % slower.
constexpr int tc_rows_of(bool f64, bool fast_f64, int dir) { return f64 && !fast_f64 && dir == 1 ? 32 : 16; }
// x columns per cell-parallel tile: 16 in every form.  (32 for the FP32
// z sweeps, whose stencil rows are planes apart, had been 7.5 % faster at
// 512^3 while z came last and carried the RK stage; first, in the reversed
// order, 16 is faster: FP32 bulk z sweep 5.55 against 6.18 ms per launch,
// learned weights 65.0 against 66.1 ms per step.)
constexpr int tc_cols_of(bool /*learned*/, int /*dir*/) { return 16; }

// Cells per tile of the fused x and y sweep (k_sweep_xy): 30 along x (32
// x-phase lanes per row), 14 along y (16 y-phase lanes per column).  It is
// built for one or two fluids: from three on, its staged tile (8 B x 720
// cells x S) and the per-lane face states outgrow the LDS and registers of
// two blocks per CU (scratch spills, up to 145 KB of LDS at ten fluids).
constexpr int kXyCellsX = 30, kXyCellsY = 14, kXyMaxFluids = 2;
constexpr int kXyThreads = (kXyCellsX + 2) * (kXyCellsY + 2);

// Kernels and their helpers belong to the unit's arithmetic form (see
// mfc_device.hpp); the argument types above are shared by every form.
inline namespace MFC_FORM {

__device__ __forceinline__ long ghost_off(const SweepArgs &a, int i, int j, int k) {
    return (long)(i + a.b) + a.nxg * ((long)(j + a.bj) + a.nyg * (long)(k + a.bk));
}
__device__ __forceinline__ long interior_off(const SweepArgs &a, int i, int j, int k) {
    return (long)i + (long)(a.m + 1) * ((long)j + (long)(a.n + 1) * (long)k);
}

// Blocks are dealt to the 8 XCDs round-robin by id.  Renumber so that each
// XCD gets a contiguous range of logical blocks: neighbouring tiles, which
// share stencil halos, then run on the same XCD and share its L2.
__device__ __forceinline__ long xcd_logical_block(long bid, long nblocks) {
    constexpr long kXcd = 8;
    const long full = (nblocks / kXcd) * kXcd;
    if (bid >= full) return bid;
    const long per = full / kXcd;
    return (bid % kXcd) * per + bid / kXcd;
}

template <typename R>
__device__ __forceinline__ double to_d(R x) { return (double)x; }

// An element offset split into a part that is the same on every lane of a
// block and a 32-bit per-lane part.  el() adds the uniform part to the
// array's address in scalar registers and the lane part as a 32-bit byte
// offset, so the load or store takes the scalar-base global address form
// with no 64-bit address arithmetic per lane.  The lane part must stay
// below 2^29 elements (the host checks the largest block span).  Kernels
// whose offsets have no uniform part pass a plain element offset.
struct Off {
    long base;    // elements, uniform
    unsigned rel; // elements, per lane
};
template <typename T>
__device__ __forceinline__ T &el(T *p, Off o) {
    unsigned b = o.rel * (unsigned)sizeof(T);
    // opaque to the optimiser, so the 32-bit offset is not widened before
    // the address is formed (a widened one takes a 64-bit VALU add)
    asm("" : "+v"(b));
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(p + o.base) + b);
}
template <typename T>
__device__ __forceinline__ T &el(T *p, long o) {
    return p[o];
}

// Operands of finish_cell that come from HBM: the rhs accumulated by the
// earlier sweeps, the RK base and the stage input (the volume fractions of
// the stage input feed the alpha*div(u) source in every sweep).
template <int S>
struct FinishIn {
    double rhs[S], base[S], in[S];
};

// Which finish modes a sweep kernel can meet.  The contracted FP64, the
// FP32-bulk and the learned forms of a 2D/3D case never pass an RK stage to
// a sweep: they spread the stage over the sweeps (SweepArgs::fold),
// their x sweep, always the last, meets folds 0 and 3, their y/z sweeps
// folds 0, 1 and 2.  The other forms and cases meet rk and never fold.
// Fixing the set at compile time keeps the modes a kernel cannot meet out
// of its code (and its registers).
enum FinishModes { kModesAll = 0, kModesFoldLast = 1, kModesFoldFirst = 2 };
template <int D, typename R, bool ML>
constexpr bool folded_form() {
    return (kFastF64 && sizeof(R) == 8 && D > 1) || (sizeof(R) == 4 && D > 1);
}
template <int D, bool LAST, typename R, bool ML>
constexpr int finish_modes() {
    return !folded_form<D, R, ML>() ? kModesAll : LAST ? kModesFoldLast : kModesFoldFirst;
}

template <int NF, int D, int M, typename O>
__device__ __forceinline__ void finish_load(const SweepArgs &a, O go, O ro, FinishIn<Sys<NF, D>::S> &f) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    if (a.accumulate) {
#pragma unroll
        for (int v = 0; v < S; ++v) f.rhs[v] = el(a.rhs.v[v], ro);
    }
    const bool fold_in = M == kModesFoldLast && a.fold == 3 && a.ci != 0.0;
    if ((a.rk == kRkNone || a.rk == kRkEuler) && !fold_in) {
#pragma unroll
        for (int i = 0; i < NF; ++i) f.in[Y::ADV + i] = el(a.qin.v[Y::ADV + i], go);
    } else {
#pragma unroll
        for (int v = 0; v < S; ++v) f.in[v] = el(a.qin.v[v], go);
    }
    if (M == kModesAll && a.rk != kRkNone) {
#pragma unroll
        for (int v = 0; v < S; ++v) f.base[v] = el(a.qbase.v[v], go);
    } else if (M == kModesFoldFirst) {
#pragma unroll
        for (int v = 0; v < S; ++v)
            if (v >= a.base_lo && v < a.base_hi) f.base[v] = el(a.qbase.v[v], go);
    }
}

// The same with the stage input of the cell taken from the copy a sweep
// staged in LDS (element v at cin[v * stride]); the RK base too when it is
// the stage input.
template <int NF, int D, int M, typename O>
__device__ __forceinline__ void finish_load_staged(const SweepArgs &a, O go, O ro, const double *cin, int stride,
                                                   FinishIn<Sys<NF, D>::S> &f) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    if (a.accumulate) {
#pragma unroll
        for (int v = 0; v < S; ++v) f.rhs[v] = el(a.rhs.v[v], ro);
    }
#pragma unroll
    for (int v = 0; v < S; ++v) f.in[v] = cin[v * stride];
    if (M == kModesAll && a.rk != kRkNone) {
        if (a.base_is_in) {
#pragma unroll
            for (int v = 0; v < S; ++v) f.base[v] = f.in[v];
        } else {
#pragma unroll
            for (int v = 0; v < S; ++v) f.base[v] = el(a.qbase.v[v], go);
        }
    } else if (M == kModesFoldFirst) {
        // (a spread stage never adds a base that is its input: the last
        // sweep adds (cb + ci)*in instead)
#pragma unroll
        for (int v = 0; v < S; ++v)
            if (v >= a.base_lo && v < a.base_hi) f.base[v] = el(a.qbase.v[v], go);
    }
}

// The same for a contracted 2D/3D y/z sweep (folds 0, 1, 2), which keeps
// only the stage input's volume fractions in LDS (element i at
// cadv[i * stride]): the finish needs no other stage-input value, and the
// RK base it adds (stages 2 and 3) comes from HBM.  Keeping all S
// variables took 36.9 KB of LDS per 256-thread block in 3D, four blocks
// per CU; the volume fractions alone leave 25.3 KB, five.
template <int NF, int D>
__device__ __forceinline__ void finish_load_adv(const SweepArgs &a, Off go, Off ro, const double *cadv, int stride,
                                                FinishIn<Sys<NF, D>::S> &f) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    if (a.accumulate) {
#pragma unroll
        for (int v = 0; v < S; ++v) f.rhs[v] = el(a.rhs.v[v], ro);
    }
#pragma unroll
    for (int i = 0; i < NF; ++i) f.in[Y::ADV + i] = cadv[i * stride];
#pragma unroll
    for (int v = 0; v < S; ++v)
        if (v >= a.base_lo && v < a.base_hi) f.base[v] = el(a.qbase.v[v], go);
}

// x/3 of the third RK3 stage: the IEEE quotient in the plain form, a
// multiply by the rounded 1/3 in the contracted FP64 form.
__device__ __forceinline__ double third(double x) {
    if constexpr (kFastF64)
        return x * (1.0 / 3.0);
    else
        return x / 3.0;
}

// TVD-RK stage combination of the base, the stage input and rhs
// (m_time_steppers.fpp:129-362).
template <int S>
__device__ __forceinline__ void rk_combine(const SweepArgs &a, const FinishIn<S> &f, const double *r, double *o) {
    switch (a.rk) {
    case kRkEuler:
#pragma unroll
        for (int v = 0; v < S; ++v) o[v] = f.base[v] + a.dt * r[v];
        break;
    case kRk2Final:
#pragma unroll
        for (int v = 0; v < S; ++v) o[v] = (f.base[v] + f.in[v] + a.dt * r[v]) / 2.0;
        break;
    case kRk3Stage2:
#pragma unroll
        for (int v = 0; v < S; ++v) o[v] = (3.0 * f.base[v] + f.in[v] + a.dt * r[v]) / 4.0;
        break;
    default:
#pragma unroll
        for (int v = 0; v < S; ++v) o[v] = third(f.base[v] + 2.0 * f.in[v] + 2.0 * a.dt * r[v]);
        break;
    }
}

// Finishes one cell: flux divergence, alpha*div(u) source, then either the
// rhs store or the fused RK stage (m_rhs.fpp:565-652, m_time_steppers.fpp).
template <int NF, int D, typename R, int M, typename O>
__device__ __forceinline__ void finish_cell(const SweepArgs &a, O go, O ro, double inv,
                                            const R *Fm, const R *Fp, R usm, R usp,
                                            const FinishIn<Sys<NF, D>::S> &f) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    double r[S];
    const double dus = to_d(usp) - to_d(usm);
    if constexpr (M != kModesAll) {
        if (a.fold) {
            // sc*dF: this direction's part of cdt*L
            const double sc = a.cdt * inv;
#pragma unroll
            for (int v = 0; v < S; ++v) {
                const double df = to_d(Fm[v]) - to_d(Fp[v]);
                if constexpr (M == kModesFoldFirst) {
                    // sc*dF enters through the last fused multiply-add: in
                    // a momentum component that is small against the
                    // pressure, dF is a cancellation whose separate
                    // rounding doubled the error against the oracle
                    double acc = a.fold == 1 ? 0.0 : f.rhs[v];
                    if (v >= a.base_lo && v < a.base_hi) acc = fma(a.cb, f.base[v], acc);
                    r[v] = fma(sc, df, acc);
                } else
                    r[v] = f.rhs[v] + sc * df;
            }
#pragma unroll
            for (int i = 0; i < NF; ++i) r[Y::ADV + i] = r[Y::ADV + i] + sc * f.in[Y::ADV + i] * dus;
            if constexpr (M == kModesFoldFirst) {
#pragma unroll
                for (int v = 0; v < S; ++v) el(a.rhs.v[v], ro) = r[v];
                return;
            }
            if (a.ci != 0.0) {
#pragma unroll
                for (int v = 0; v < S; ++v) r[v] = r[v] + a.ci * f.in[v];
            }
            // (Euler: the stage input is not part of the output, and
            // finish_load leaves its non-advected entries unset)
            if constexpr (!std::is_same<R, double>::value) {
                // FP32 bulk: volume fractions clipped to [0, 1] as below
#pragma unroll
                for (int i = 0; i < NF; ++i) r[Y::ADV + i] = min(max(r[Y::ADV + i], 0.0), 1.0);
            }
#pragma unroll
            for (int v = 0; v < S; ++v) el(a.qout.v[v], go) = r[v];
            return;
        }
    }
#pragma unroll
    for (int v = 0; v < S; ++v) {
        const double df = to_d(Fm[v]) - to_d(Fp[v]);
        r[v] = a.accumulate ? f.rhs[v] + inv * df : inv * df;
    }
#pragma unroll
    for (int i = 0; i < NF; ++i) r[Y::ADV + i] = r[Y::ADV + i] + inv * f.in[Y::ADV + i] * dus;
    if constexpr (!std::is_same<R, double>::value) {
        // FP32 bulk: the RK stage accumulates in FP64 and clips the volume
        // fractions to [0, 1] (positivity; the reference has no clip).
        if (a.rk != kRkNone) {
            double o[S];
            rk_combine<S>(a, f, r, o);
#pragma unroll
            for (int i = 0; i < NF; ++i) o[Y::ADV + i] = min(max(o[Y::ADV + i], 0.0), 1.0);
#pragma unroll
            for (int v = 0; v < S; ++v) el(a.qout.v[v], go) = o[v];
            return;
        }
    }
    if constexpr (M != kModesAll) {
        // no RK stage reaches these sweeps
#pragma unroll
        for (int v = 0; v < S; ++v) el(a.rhs.v[v], ro) = r[v];
        return;
    }
    switch (a.rk) {
    case kRkNone:
#pragma unroll
        for (int v = 0; v < S; ++v) el(a.rhs.v[v], ro) = r[v];
        break;
    case kRkEuler:
#pragma unroll
        for (int v = 0; v < S; ++v) el(a.qout.v[v], go) = f.base[v] + a.dt * r[v];
        break;
    case kRk2Final:
#pragma unroll
        for (int v = 0; v < S; ++v) el(a.qout.v[v], go) = (f.base[v] + f.in[v] + a.dt * r[v]) / 2.0;
        break;
    case kRk3Stage2:
#pragma unroll
        for (int v = 0; v < S; ++v) el(a.qout.v[v], go) = (3.0 * f.base[v] + f.in[v] + a.dt * r[v]) / 4.0;
        break;
    default:
#pragma unroll
        for (int v = 0; v < S; ++v) el(a.qout.v[v], go) = third(f.base[v] + 2.0 * f.in[v] + 2.0 * a.dt * r[v]);
        break;
    }
}

template <int NF, int D>
__device__ __forceinline__ void load_prim(const SweepArgs &a, long go, double *p) {
    constexpr int S = Sys<NF, D>::S;
    double c[S];
#pragma unroll
    for (int v = 0; v < S; ++v) c[v] = a.qin.v[v][go];
    cons_to_prim<NF, D>(c, p, a.eos);
}
template <int NF, int D>
__device__ __forceinline__ void load_cons_prim(const SweepArgs &a, long go, double *c, double *p) {
    constexpr int S = Sys<NF, D>::S;
#pragma unroll
    for (int v = 0; v < S; ++v) c[v] = a.qin.v[v][go];
    cons_to_prim<NF, D>(c, p, a.eos);
}

template <typename R, int WO, bool UNI>
__device__ __forceinline__ void load_coef(const SweepArgs &a, int cell, R *cf) {
    if constexpr (WO > 1 && UNI) {
#pragma unroll
        for (int q = 0; q < kNCoef; ++q) cf[q] = sizeof(R) == 8 ? (R)a.ucf[q] : (R)a.ucf32[q];
    } else if constexpr (WO > 1) {
        const double *t = a.coef + (long)(cell + a.b);
#pragma unroll
        for (int q = 0; q < kNCoef; ++q) cf[q] = (R)t[(long)q * a.ncell];
    }
}

// ---------------------------------------------------------------------------
// Learned WENO5 weights inside a sweep (FP32 bulk).  Per variable each lane
// owns two reconstructions: the right face of its left cell and the left
// face of its right cell.  Both are oriented so the face lies between
// stencil entries 2 and 3 (the left-face stencil is reversed, its candidates
// and weights listed in the reversed order), their BF16 features go to LDS,
// the wave evaluates the MLP for its 128 samples with the transposed,
// register-chained MFMA passes of mlp_infer_lane, and each lane combines its
// own logits with its candidates.  Per face the plain weights are the
// fallback when the learned value is not finite or breaks positivity
// (partial densities and volume fractions >= 0, volume fractions <= 1,
// pressure > 0 where the plain value is).

// Candidates and plain weights of one side of a WENO5 cell (FP32 bulk; the
// same arithmetic as weno_cell), in the oriented order.
template <bool RIGHT>
__device__ __forceinline__ void weno5_parts(const float *v, const float *cf, float eps, float *p, float *w) {
    const float vj = v[2];
    const float d1 = v[4] - v[3], d0 = v[3] - v[2], dm1 = v[2] - v[1], dm2 = v[1] - v[0];
    const float b0 = cf[kBeta + 0] * d1 * d1 + cf[kBeta + 1] * d1 * d0 + cf[kBeta + 2] * d0 * d0 + eps;
    const float b1 = cf[kBeta + 3] * d0 * d0 + cf[kBeta + 4] * d0 * dm1 + cf[kBeta + 5] * dm1 * dm1 + eps;
    const float b2 = cf[kBeta + 6] * dm1 * dm1 + cf[kBeta + 7] * dm1 * dm2 + cf[kBeta + 8] * dm2 * dm2 + eps;
    const int C = RIGHT ? kCbR : kCbL, Dd = RIGHT ? kDR : kDL;
    float pp[3], a[3];
    pp[0] = vj + cf[C + 0] * d1 + cf[C + 1] * d0;
    pp[1] = vj + cf[C + 2] * d0 + cf[C + 3] * dm1;
    pp[2] = vj + cf[C + 4] * dm1 + cf[C + 5] * dm2;
    a[0] = bdiv(cf[Dd + 0], b0 * b0);
    a[1] = bdiv(cf[Dd + 1], b1 * b1);
    a[2] = bdiv(cf[Dd + 2], b2 * b2);
    const float s = a[0] + a[1] + a[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // the left face's candidate k is the oriented candidate 2-k
        const int o = RIGHT ? k : 2 - k;
        p[o] = pp[k];
        w[o] = bdiv(a[k], s);
    }
}

// Features of an oriented stencil into one LDS row: 5 BF16 values and 3
// zero pads.  Returns the scale s.
__device__ __forceinline__ float ml_features(const float *v, unsigned short *row) {
    float sc = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) sc = fmaxf(sc, fabsf(v[k + 1] - v[k]));
    const float inv = sc > 0.f ? 1.f / sc : 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) row[k] = k < kMlpIn ? f2bf((v[k] - v[2]) * inv) : (unsigned short)0;
    return sc;
}

// Whether the network may act on a stencil: its variation must stand above
// rounding noise (largest neighbour difference > 1e-6 of the largest
// magnitude, the rule that admits training pairs) and the gate must be open.
__device__ __forceinline__ bool ml_open(const float *v, const float *wp) {
    float sc = 0.f, vmax = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) sc = fmaxf(sc, fabsf(v[k + 1] - v[k]));
#pragma unroll
    for (int k = 0; k < 5; ++k) vmax = fmaxf(vmax, fabsf(v[k]));
    return sc > 1e-6f * vmax && mlp_gate(wp) > 0.f;
}

template <int NF, int D>
__device__ __forceinline__ float ml_value(int var, bool open, const float *z, const float *p, const float *wp) {
    using Y = Sys<NF, D>;
    if (!open) return wp[0] * p[0] + wp[1] * p[1] + wp[2] * p[2];
    float w[3];
    mlp_weights(z, wp, w);
    const float learned = w[0] * p[0] + w[1] * p[1] + w[2] * p[2];
    const float plain = wp[0] * p[0] + wp[1] * p[1] + wp[2] * p[2];
    bool ok = isfinite(learned);
    if (var < NF || var >= Y::ADV) ok = ok && !(learned < 0.f && plain >= 0.f);
    if (var >= Y::ADV) ok = ok && !(learned > 1.f && plain <= 1.f);
    if (var == Y::E) ok = ok && !(learned <= 0.f && plain > 0.f);
    return ok ? learned : plain;
}

// Learned face values of one variable for the wave's 64 lanes: the value at
// the upper face of stencil stR (cells j-2..j+2, coefficients cfR) and at
// the lower face of stencil stL (cfL); the cell-parallel sweeps pass one
// cell's stencil as both.  xs is the wave's feature rows.
template <int NF, int D>
__device__ __forceinline__ void ml_face_states(int var, const float *stR, const float *stL, const float *cfR,
                                               const float *cfL, float eps, const MlpTRegs &T,
                                               unsigned short (*xs)[64][8], int lane, float &qLv, float &qRv) {
    float pR[3], wR[3], pL[3], wL[3], vo[5];
    weno5_parts<true>(stR, cfR, eps, pR, wR);
    weno5_parts<false>(stL, cfL, eps, pL, wL);
    // the network runs for the whole wave (its MFMA passes need every lane)
    // unless it may act on neither side of any of the 64 faces
#pragma unroll
    for (int k = 0; k < 5; ++k) vo[k] = stL[4 - k];
    const bool openR = ml_open(stR, wR), openL = ml_open(vo, wL);
    if (__any(openR || openL)) {
        ml_features(stR, xs[0][lane]);
        ml_features(vo, xs[1][lane]);
        wave_lds_sync();
        f32x4 z[2];
        mlp_infer_lane(T, xs, lane, z);
        wave_lds_sync(); // the next variable rewrites xs
        const float zR[3] = {z[0][0], z[0][1], z[0][2]}, zL[3] = {z[1][0], z[1][1], z[1][2]};
        qLv = ml_value<NF, D>(var, openR, zR, pR, wR);
        qRv = ml_value<NF, D>(var, openL, zL, pL, wL);
    } else {
        qLv = wR[0] * pR[0] + wR[1] * pR[1] + wR[2] * pR[2];
        qRv = wL[0] * pL[0] + wL[1] * pL[1] + wL[2] * pL[2];
    }
}

// ---------------------------------------------------------------------------
// x sweep: one wave per tile of a row.  Face-parallel (FP64 on non-uniform
// grids): lane L reconstructs the two sides of face x0-1+L + 1/2 from the
// stencils of its two cells; 63 cells per tile are finished.  Cell-parallel
// (FP32 bulk, the learned form, and FP64 with the uniform-grid
// coefficients in scalar registers): lane L reconstructs both faces of cell
// x0-1+L (the smoothness indicators and the reciprocals of their squares
// serve both), the left-face values move one lane down through LDS, and
// lane L solves face x0-1+L + 1/2; 62 cells per tile.  Measured on MI355X:
// cell-parallel is 6 % faster in FP32 and with uniform FP64 coefficients,
// and no faster for FP64 with per-cell coefficient loads, whose registers
// then cost a wave of occupancy.
template <bool ML, typename R, bool UNI>
constexpr bool x_cell_parallel() { return ML || sizeof(R) == 4 || UNI; }
template <bool ML, typename R, bool UNI>
constexpr int x_tile_cells() { return x_cell_parallel<ML, R, UNI>() ? 62 : 63; }

template <int NF, int D, int WO, typename R, bool ML, bool UNI>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(ML ? 2 : 4, 8))) void k_sweep_x(const SweepArgs a) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    constexpr int FM = finish_modes<D, true, R, ML>();
    constexpr int PN = (WO - 1) / 2;
    constexpr int NC = 64 + 2 * PN + 1;
    constexpr int TW = x_tile_cells<ML, R, UNI>();
    // Cell-parallel plain forms exchange face states and fluxes between
    // lanes by shuffles: LDS holds only the staged rows (widened to 128 so
    // the two face states of a variable fit over its row), 8 KB per wave
    // for FP64 instead of 9 KB of staging and flux rows.
    constexpr bool kShfl = x_cell_parallel<ML, R, UNI>() && !ML;
    constexpr int SPW = kShfl && NC < 128 ? 128 : NC;
    __shared__ R sp[S][SPW];
    __shared__ R sf[kShfl ? 1 : S + 1][64];

    const long bid = xcd_logical_block(blockIdx.x, a.nblocks);
    const int tile = (int)(bid % a.ntile);
    const long row = bid / a.ntile;
    const int j = (int)(row % (a.n + 1));
    const int k = (int)(row / (a.n + 1));
    const int lane = threadIdx.x;
    const int x0 = tile * TW;
    // conservative values of this lane's finish cell x0+lane, taken from the
    // staging loads (forms that load before converting) so the finish does
    // not read the stage input again
    constexpr bool kStagedFinish = !(sizeof(R) == 8 && !UNI);
    double cin[S];

    // 1. stage primitive stencil cells x0-1-PN .. x0+63+PN.  FP32 forms and
    // uniform-grid FP64 issue every load before the first conversion waits
    // on one; FP64 with per-cell coefficient loads, with no registers to
    // spare at 4 waves per SIMD, converts as loads arrive (measured: hoisting
    // is 17 % faster in FP32, 2-3 % faster in uniform FP64, 3 % slower in
    // FP64 with per-cell coefficients).
    if constexpr (sizeof(R) == 8 && !UNI) {
        for (int c = lane; c < NC; c += 64) {
            const int x = x0 - 1 - PN + c;
            double pr[S];
            if (x <= a.m + a.b) {
                load_prim<NF, D>(a, ghost_off(a, x, j, k), pr);
            } else {
#pragma unroll
                for (int v = 0; v < S; ++v) pr[v] = 0.0;
            }
#pragma unroll
            for (int v = 0; v < S; ++v) sp[v][c] = (R)pr[v];
        }
    } else {
        constexpr int NIT = (NC + 63) / 64;
        double cv[NIT][S];
        // staging entry c is cell x0-1-PN+c (entry 1+PN: cell x0)
        const long sb = ghost_off(a, x0 - 1 - PN, j, k);
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int c = lane + 64 * it, x = x0 - 1 - PN + c;
            const bool ok = c < NC && x <= a.m + a.b;
            // entries past the ghost layers take cell x0's values (their
            // faces are never finished; a select per value would cost VALU)
            const Off go{sb, (unsigned)(ok ? c : 1 + PN)};
#pragma unroll
            for (int v = 0; v < S; ++v) cv[it][v] = el(a.qin.v[v], go);
        }
        // cell x0+lane is staging entry lane+1+PN: loaded by lane
        // (lane+1+PN) mod 64 in pass 0, or in pass 1 for the last 1+PN lanes
        static_assert(NIT == 2, "two staging passes");
#pragma unroll
        for (int v = 0; v < S; ++v) {
            const double send = lane < 1 + PN ? cv[1][v] : cv[0][v];
            cin[v] = __shfl(send, (lane + 1 + PN) & 63, 64);
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int c = lane + 64 * it;
            if (c < NC) {
                double pr[S];
                cons_to_prim<NF, D>(cv[it], pr, a.eos);
#pragma unroll
                for (int v = 0; v < S; ++v) sp[v][c] = (R)pr[v];
            }
        }
    }
    __syncthreads();

    // 2. face jf + 1/2, jf = x0-1+lane: qL left of it, qR right of it
    const int jf = x0 - 1 + lane;
    R qL[S], qR[S];
    if constexpr (ML) {
        // cell-parallel like the plain forms: lane = cell jf, whose upper-face
        // value is face jf+1/2's left state and whose lower-face value is
        // face jf-1/2's right state, read back one lane over.  Every lane
        // takes part in the wave's MFMA passes, valid cell or not; one
        // variable at a time, results through LDS over the variable's own
        // staged row and the flux rows (the block is one wave, and each value
        // is written after the reads it depends on)
        __shared__ unsigned short xs[2][64][8];
        MlpTRegs T;
        mlp_load_tregs(a.mlp, lane, T);
        R cf[kNCoef];
        load_coef<R, WO, UNI>(a, min(jf, a.m + 1), cf);
        const R eps = (R)a.eps;
#pragma unroll 1
        for (int v = 0; v < S; ++v) {
            R st[5], vr, vl;
#pragma unroll
            for (int w = 0; w < 5; ++w) st[w] = sp[v][lane + w];
            ml_face_states<NF, D>(v, st, st, cf, cf, eps, T, xs, lane, vr, vl);
            sp[v][lane] = vl;
            sf[v][lane] = vr;
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < S; ++v) {
            qL[v] = sf[v][lane];
            qR[v] = sp[v][min(lane + 1, 63)];
        }
    } else if constexpr (x_cell_parallel<ML, R, UNI>()) {
        // both reconstructions of cell jf (coefficients of a cell past the
        // last face are clamped; those lanes' values are not used), one
        // variable at a time to bound the live registers.  The wave has read
        // the variable's stencil row before any lane writes it, so the face
        // states of face jf+1/2 go over that row: this lane's upper-face
        // value and, by shuffle, the next lane's lower-face value; each lane
        // reads back only its own slots.
        R cf[kNCoef];
        load_coef<R, WO, UNI>(a, min(jf, a.m + 1), cf);
        const R eps = (R)a.eps;
#pragma unroll 1
        for (int v = 0; v < S; ++v) {
            R st[WO], vl, vr;
#pragma unroll
            for (int w = 0; w < WO; ++w) st[w] = sp[v][lane + w];
            weno_cell<WO, R, true, true>(st, cf, eps, vl, vr);
            const R vn = __shfl_down(vl, 1, 64);
            sp[v][lane] = vr;
            sp[v][64 + lane] = vn;
        }
#pragma unroll
        for (int v = 0; v < S; ++v) {
            qL[v] = sp[v][lane];
            qR[v] = sp[v][64 + lane];
        }
    } else if (jf <= a.m) {
        R cfR[kNCoef], cfL[kNCoef];
        load_coef<R, WO, UNI>(a, jf, cfR);
        load_coef<R, WO, UNI>(a, jf + 1, cfL);
        const R eps = (R)a.eps;
#pragma unroll
        for (int v = 0; v < S; ++v) {
            R st[WO], dummy;
#pragma unroll
            for (int w = 0; w < WO; ++w) st[w] = sp[v][lane + w];
            weno_cell<WO, R, false, true>(st, cfR, eps, dummy, qL[v]);
#pragma unroll
            for (int w = 0; w < WO; ++w) st[w] = sp[v][lane + 1 + w];
            weno_cell<WO, R, true, false>(st, cfL, eps, qR[v], dummy);
        }
    }
    if constexpr (kShfl) {
        // every lane solves its face (lanes past the last face on values
        // nobody uses), then the upper flux of each cell comes from the
        // next lane
        R f[S], us;
        if (jf == -1 && a.bc_beg == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qL[v] = qR[v];
        }
        if (jf == a.m && a.bc_end == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qR[v] = qL[v];
        }
        hllc_face<NF, D, 0, R>(qL, qR, a.eos, f, us);
        R fp[S];
#pragma unroll
        for (int v = 0; v < S; ++v) fp[v] = __shfl_down(f[v], 1, 64);
        const R usp = __shfl_down(us, 1, 64);
        const int i = x0 + lane;
        if (lane < TW && i <= a.m) {
            const Off go{ghost_off(a, x0, j, k), (unsigned)lane}, ro{interior_off(a, x0, j, k), (unsigned)lane};
            FinishIn<S> fin;
            finish_load_staged<NF, D, FM>(a, go, ro, cin, 1, fin);
            finish_cell<NF, D, R, FM>(a, go, ro, a.inv_ds[i], f, fp, us, usp, fin);
        }
        return;
    }
    if (lane < TW + 1 && jf <= a.m) {
        R f[S], us;
        if (jf == -1 && a.bc_beg == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qL[v] = qR[v];
        }
        if (jf == a.m && a.bc_end == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qR[v] = qL[v];
        }
        hllc_face<NF, D, 0, R>(qL, qR, a.eos, f, us);
#pragma unroll
        for (int v = 0; v < S; ++v) sf[v][lane] = f[v];
        sf[S][lane] = us;
    }
    __syncthreads();

    // 3. cell x0+lane
    const int i = x0 + lane;
    if (lane < TW && i <= a.m) {
        R Fm[S], Fp[S];
#pragma unroll
        for (int v = 0; v < S; ++v) {
            Fm[v] = sf[v][lane];
            Fp[v] = sf[v][lane + 1];
        }
        const Off go{ghost_off(a, x0, j, k), (unsigned)lane}, ro{interior_off(a, x0, j, k), (unsigned)lane};
        FinishIn<S> fin;
        if constexpr (kStagedFinish)
            finish_load_staged<NF, D, FM>(a, go, ro, cin, 1, fin);
        else
            finish_load<NF, D, FM>(a, go, ro, fin);
        finish_cell<NF, D, R, FM>(a, go, ro, a.inv_ds[i], Fm, Fp, sf[S][lane], sf[S][lane + 1], fin);
    }
}

// ---------------------------------------------------------------------------
// y / z sweep (DIR = 1 or 2): a TX x TS tile of faces per block.
// Thread (tx, ts) computes face ts of column x0+tx; the stencil rows are
// staged in LDS (rows of TX consecutive x, so every global load and store
// is a contiguous 8*TX-byte segment) and the fluxes reuse the same LDS.
constexpr int kTX = 16;
// Faces per column of a tile (one thread each).  Measured on MI355X for the
// FP64 2D y sweep: 16 (256-thread blocks, four per CU) 0.33 ms per launch,
// 32 (512-thread blocks, two per CU, 19 % less halo re-fetch) 0.38 ms.
template <bool ML>
constexpr int t_faces() { return 16; }
template <bool ML>
constexpr int t_threads() { return kTX * t_faces<ML>(); }
// The y/z sweep is cell-parallel (k_sweep_tc) in the forms whose x sweep is.
template <bool ML, typename R, bool UNI>
constexpr bool t_cell_parallel() { return x_cell_parallel<ML, R, UNI>(); }
// Threads per column of the cell-parallel tile (a power of two <= 64; the
// column's lanes are consecutive within a wave): tc_rows_of (plain and
// contracted FP64 differ, so the host picks by form too).
template <typename R>
constexpr int tc_rows(int dir) { return tc_rows_of(sizeof(R) == 8, kFastF64, dir); }
template <typename R, bool ML>
constexpr int tc_cols(int dir) { return tc_cols_of(ML, dir); }
template <bool ML, typename R, bool UNI>
constexpr int t_tile_cells(int dir) {
    return t_cell_parallel<ML, R, UNI>() ? tc_rows<R>(dir) - 2 : t_faces<ML>() - 1;
}

template <int NF, int D, int DIR, int WO, typename R, bool ML, bool UNI>
__global__ __launch_bounds__(t_threads<ML>()) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_sweep_t(
    const SweepArgs a) {
    static_assert(!ML, "the learned form runs the cell-parallel sweep");
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    constexpr int FM = finish_modes<D, false, R, ML>();
    constexpr int kTS = t_faces<ML>();
    constexpr int NT = t_threads<ML>();
    constexpr int PN = (WO - 1) / 2;
    constexpr int NR = kTS + 2 * PN + 1; // staged rows
    constexpr int PRIM_WORDS = S * NR * kTX;
    constexpr int FLUX_WORDS = (S + 1) * kTS * kTX;
    constexpr int LDS_BYTES = sizeof(R) * (PRIM_WORDS > FLUX_WORDS ? PRIM_WORDS : FLUX_WORDS);
    // The conservative stage input of the tile's own cells stays in LDS for
    // the finish (rhs accumulation, RK stage) instead of being read from HBM
    // a second time, while the block still fits four to a CU.
    constexpr bool KEEP = LDS_BYTES + 8 * S * (kTS - 1) * kTX <= 40 * 1024 * (NT / 256);
    __shared__ R lds[PRIM_WORDS > FLUX_WORDS ? PRIM_WORDS : FLUX_WORDS];
    __shared__ double lcons[KEEP ? S : 1][kTS - 1][kTX];
    R(*sp)[NR][kTX] = reinterpret_cast<R(*)[NR][kTX]>(lds);
    R(*sf)[kTS][kTX] = reinterpret_cast<R(*)[kTS][kTX]>(lds);

    // x tiles are the fastest block index: neighbouring tiles share the
    // 128-byte lines their 16-wide row segments straddle, so they run back to
    // back on one XCD (ordering by sweep tile instead doubled the fetched
    // bytes on MI355X).
    const long bid = xcd_logical_block(blockIdx.x, a.nblocks);
    const int xt = (int)(bid % a.ntile);
    const long rest = bid / a.ntile;
    const int chunk = (int)(rest % a.nchunk);
    const int t = (int)(rest / a.nchunk);
    const int s = (DIR == 1) ? a.n : a.p;
    const int x0 = xt * kTX;
    const int y0 = chunk * (kTS - 1); // first cell of this tile along the sweep
    const int tid = threadIdx.x;
    const int tx = tid % kTX, ts = tid / kTX;

    auto gcell = [&](int i, int c) -> long {
        return (DIR == 1) ? ghost_off(a, i, c, t) : ghost_off(a, i, t, c);
    };
    auto rcell = [&](int i, int c) -> long {
        return (DIR == 1) ? interior_off(a, i, c, t) : interior_off(a, i, t, c);
    };

    // 1. stage primitive rows y0-1-PN .. y0+kTS-1+PN of columns x0..x0+kTX-1
    // (as in the x sweep: every load issued before the first conversion,
    // except FP64 with per-cell coefficient loads)
    if constexpr (sizeof(R) == 8 && !UNI) {
        for (int e = tid; e < NR * kTX; e += NT) {
            const int ex = e % kTX, er = e / kTX;
            const int c = y0 - 1 - PN + er;
            const int i = min(x0 + ex, a.m);
            double pr[S], cv[S];
            if (c <= s + a.b) {
                load_cons_prim<NF, D>(a, gcell(i, c), cv, pr);
            } else {
#pragma unroll
                for (int v = 0; v < S; ++v) pr[v] = cv[v] = 0.0;
            }
#pragma unroll
            for (int v = 0; v < S; ++v) sp[v][er][ex] = (R)pr[v];
            if constexpr (KEEP) {
                if (er > PN && er < PN + kTS) {
#pragma unroll
                    for (int v = 0; v < S; ++v) lcons[v][er - PN - 1][ex] = cv[v];
                }
            }
        }
    } else {
        constexpr int NIT = (NR * kTX + NT - 1) / NT;
        double cv[NIT][S];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int e = tid + NT * it;
            const int ex = e % kTX, er = e / kTX;
            const int c = y0 - 1 - PN + er;
            const bool ok = e < NR * kTX && c <= s + a.b;
            // rows past the ghost layers take row y0's values (their faces
            // are never finished)
            const long go = gcell(min(x0 + ex, a.m), ok ? c : y0);
#pragma unroll
            for (int v = 0; v < S; ++v) cv[it][v] = a.qin.v[v][go];
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int e = tid + NT * it;
            if (e < NR * kTX) {
                const int ex = e % kTX, er = e / kTX;
                double pr[S];
                cons_to_prim<NF, D>(cv[it], pr, a.eos);
#pragma unroll
                for (int v = 0; v < S; ++v) sp[v][er][ex] = (R)pr[v];
                if constexpr (KEEP) {
                    if (er > PN && er < PN + kTS) {
#pragma unroll
                        for (int v = 0; v < S; ++v) lcons[v][er - PN - 1][ex] = cv[it][v];
                    }
                }
            }
        }
    }
    __syncthreads();

    // 2. face jf+1/2 with jf = y0-1+ts
    const int jf = y0 - 1 + ts;
    R f[S], us = R(0);
    const bool face_ok = jf <= s;
    R qL[S], qR[S];
    if (face_ok) {
        R cfR[kNCoef], cfL[kNCoef];
        load_coef<R, WO, UNI>(a, jf, cfR);
        load_coef<R, WO, UNI>(a, jf + 1, cfL);
        const R eps = (R)a.eps;
#pragma unroll
        for (int v = 0; v < S; ++v) {
            R st[WO], dummy;
#pragma unroll
            for (int w = 0; w < WO; ++w) st[w] = sp[v][ts + w][tx];
            weno_cell<WO, R, false, true>(st, cfR, eps, dummy, qL[v]);
#pragma unroll
            for (int w = 0; w < WO; ++w) st[w] = sp[v][ts + 1 + w][tx];
            weno_cell<WO, R, true, false>(st, cfL, eps, qR[v], dummy);
        }
        if (jf == -1 && a.bc_beg == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qL[v] = qR[v];
        }
        if (jf == s && a.bc_end == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qR[v] = qL[v];
        }
        hllc_face<NF, D, DIR, R>(qL, qR, a.eos, f, us);
    }
    __syncthreads(); // the stencil is dead; the fluxes take its place
    if (face_ok) {
#pragma unroll
        for (int v = 0; v < S; ++v) sf[v][ts][tx] = f[v];
        sf[S][ts][tx] = us;
    }
    __syncthreads();

    // 3. cell y0+ts (ts < kTS-1)
    const int i = x0 + tx, cc = y0 + ts;
    if (ts < kTS - 1 && i <= a.m && cc <= s) {
        R Fm[S], Fp[S];
#pragma unroll
        for (int v = 0; v < S; ++v) {
            Fm[v] = sf[v][ts][tx];
            Fp[v] = sf[v][ts + 1][tx];
        }
        FinishIn<S> fin;
        if constexpr (KEEP)
            finish_load_staged<NF, D, FM>(a, gcell(i, cc), rcell(i, cc), &lcons[0][ts][tx], (kTS - 1) * kTX, fin);
        else
            finish_load<NF, D, FM>(a, gcell(i, cc), rcell(i, cc), fin);
        finish_cell<NF, D, R, FM>(a, gcell(i, cc), rcell(i, cc), a.inv_ds[cc], Fm, Fp, sf[S][ts][tx],
                                  sf[S][ts + 1][tx], fin);
    }
}

// ---------------------------------------------------------------------------
// y / z sweep, cell-parallel (the forms with a cell-parallel x sweep), K =
// tc_rows<R>(DIR) threads per column, C = tc_cols<R, ML>(DIR) columns, C * K
// threads per block.  Lane l of wave w works on column
// x0 + (64/K)w + l/K; its position ts = l%K among the column's K lanes is
// cell y0-1+ts, of which it reconstructs both faces (one set of smoothness
// indicators and reciprocals for the two).  The upper-face value is the
// left state of face y0-1+ts + 1/2, and the right state is the lower-face
// value of cell y0+ts, held by the next lane of the column, from which it
// moves by a shuffle (no LDS, no barrier).  K-1 faces per column, cells
// y0 .. y0+K-3 finished.  Staged columns are stored with an odd pitch so
// the K lanes of a column read K consecutive values.
template <int W, typename R>
__device__ __forceinline__ R row_next(R x) {
    return __shfl_down(x, 1u, W);
}

template <int NF, int D, int DIR, int WO, typename R, bool ML, bool UNI>
__global__ __launch_bounds__((tc_cols<R, ML>(DIR) * tc_rows<R>(DIR)))
    __attribute__((amdgpu_waves_per_eu(ML ? 2 : 4, 8))) void k_sweep_tc(const SweepArgs a) {
    static_assert(t_cell_parallel<ML, R, UNI>(), "cell-parallel forms only");
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    constexpr int FM = finish_modes<D, false, R, ML>();
    constexpr int XC = tc_cols<R, ML>(DIR), kTS = tc_rows<R>(DIR), NT = XC * kTS;
    constexpr int TC = t_tile_cells<ML, R, UNI>(DIR);
    constexpr int PN = (WO - 1) / 2;
    constexpr int NR = kTS + 2 * PN; // staged cells y0-1-PN .. y0+kTS-2+PN
    constexpr int P = NR | 1;        // odd column pitches
    constexpr int KP = TC | 1;
    constexpr int DP = XC + 1;       // flux rows
    constexpr int SP_WORDS = S * XC * P, FL_WORDS = (S + 1) * kTS * DP;
    constexpr int WORDS = SP_WORDS > FL_WORDS ? SP_WORDS : FL_WORDS;
    // the stage input of the finished cells stays in LDS for the finish
    // while four 256-thread blocks (or the same threads in fewer blocks)
    // still fit a CU (not in the learned form, whose MLP staging takes
    // that LDS)
    // (the contracted 2D/3D sweeps, and the 3D y sweep of the forms that
    // sweep x, y, z, which is never the last, keep only the volume
    // fractions)
    constexpr bool ADVK = !ML && (FM == kModesFoldFirst || (FM == kModesAll && D == 3 && DIR == 1));
    constexpr int KS = ADVK ? NF : S;
    constexpr bool KEEP = !ML && sizeof(R) * WORDS + 8 * KS * XC * KP <= 40 * 1024 * (NT / 256);
    __shared__ R lds[WORDS];
    __shared__ double lcons[KEEP ? KS : 1][XC][KEEP ? KP : 1];
    R(*sp)[XC][P] = reinterpret_cast<R(*)[XC][P]>(lds);
    R(*fl)[kTS][DP] = reinterpret_cast<R(*)[kTS][DP]>(lds);

    const long bid = xcd_logical_block(blockIdx.x, a.nblocks);
    const int xt = (int)(bid % a.ntile);
    const long rest = bid / a.ntile;
    const int chunk = (int)(rest % a.nchunk);
    const int t = (int)(rest / a.nchunk);
    const int s = (DIR == 1) ? a.n : a.p;
    const int x0 = xt * XC;
    const int y0 = chunk * TC;
    const int tid = threadIdx.x;
    const int tx = (64 / kTS) * (tid >> 6) + ((tid & 63) / kTS), ts = tid & (kTS - 1);

    auto gcell = [&](int i, int c) -> long {
        return (DIR == 1) ? ghost_off(a, i, c, t) : ghost_off(a, i, t, c);
    };
    auto rcell = [&](int i, int c) -> long {
        return (DIR == 1) ? interior_off(a, i, c, t) : interior_off(a, i, t, c);
    };

    // Offsets within the block's cells: staged cell (x0 + ex, y0-1-PN + er)
    // at sb + ex + er*gp, interior cell (x0 + ex, y0 + er) at rb + ex + er*rp
    // (uniform bases, 32-bit lane parts)
    const long sb = gcell(x0, y0 - 1 - PN), rb = rcell(x0, y0);
    const unsigned gp = (unsigned)((DIR == 1) ? a.nxg : a.nxg * a.nyg);
    const unsigned rp = (unsigned)((DIR == 1) ? a.m + 1 : (a.m + 1) * (a.n + 1));

    // 1. stage primitive rows y0-1-PN .. y0+kTS-2+PN of columns x0..x0+XC-1
    // (loads along x, every load issued before the first conversion)
    {
        constexpr int NIT = (NR * XC + NT - 1) / NT;
        double cv[NIT][S];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int e = tid + NT * it;
            const int ex = e % XC, er = e / XC;
            const int c = y0 - 1 - PN + er;
            const bool ok = e < NR * XC && c <= s + a.b;
            // rows past the ghost layers take row y0's values (their faces
            // are never finished; a select per value would cost VALU)
            const Off go{sb, (unsigned)(min(x0 + ex, a.m) - x0) + (unsigned)(ok ? er : 1 + PN) * gp};
#pragma unroll
            for (int v = 0; v < S; ++v) cv[it][v] = el(a.qin.v[v], go);
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int e = tid + NT * it;
            if (e < NR * XC) {
                const int ex = e % XC, er = e / XC;
                double pr[S];
                cons_to_prim<NF, D>(cv[it], pr, a.eos);
#pragma unroll
                for (int v = 0; v < S; ++v) sp[v][ex][er] = (R)pr[v];
                if constexpr (KEEP) {
                    if (er > PN && er <= PN + TC) {
#pragma unroll
                        for (int v = 0; v < KS; ++v) lcons[v][ex][er - PN - 1] = cv[it][ADVK ? Y::ADV + v : v];
                    }
                }
            }
        }
    }
    __syncthreads();

    // 2. cell jf = y0-1+ts: both faces; face jf + 1/2
    const int jf = y0 - 1 + ts;
    R qL[S], qR[S];
    if constexpr (ML) {
        // learned: one variable at a time (the wave's MFMA passes take
        // every lane), each thread's two face values through its own LDS
        // slots; the lower-face value then moves one lane down the column
        __shared__ unsigned short xs[NT / 64][2][64][8];
        __shared__ R mres[2][S][NT];
        const int lane = tid & 63, wv = tid >> 6;
        MlpTRegs T;
        mlp_load_tregs(a.mlp, lane, T);
        R cf[kNCoef];
        load_coef<R, WO, UNI>(a, min(jf, s + 1), cf);
        const R eps = (R)a.eps;
#pragma unroll 1
        for (int v = 0; v < S; ++v) {
            R st[5], vr, vl;
#pragma unroll
            for (int w = 0; w < 5; ++w) st[w] = sp[v][tx][ts + w];
            ml_face_states<NF, D>(v, st, st, cf, cf, eps, T, xs[wv], lane, vr, vl);
            mres[0][v][tid] = vr;
            mres[1][v][tid] = vl;
        }
#pragma unroll
        for (int v = 0; v < S; ++v) {
            qL[v] = mres[0][v][tid];
            qR[v] = row_next<kTS>(mres[1][v][tid]);
        }
    } else {
        R cf[kNCoef];
        load_coef<R, WO, UNI>(a, min(jf, s + 1), cf);
        const R eps = (R)a.eps;
#pragma unroll
        for (int v = 0; v < S; ++v) {
            R st[WO], vl;
#pragma unroll
            for (int w = 0; w < WO; ++w) st[w] = sp[v][tx][ts + w];
            weno_cell<WO, R, true, true>(st, cf, eps, vl, qL[v]);
            qR[v] = row_next<kTS>(vl);
        }
    }
    // every lane solves its face, valid or not (lane 15 of a row and lanes
    // past the last face; their fluxes are never used): under a branch the
    // compiler would sink the upper-face reconstruction into it and keep
    // every variable's intermediates live across the loop
    R f[S], us;
    if (jf == -1 && a.bc_beg == -4) {
#pragma unroll
        for (int v = 0; v < S; ++v) qL[v] = qR[v];
    }
    if (jf == s && a.bc_end == -4) {
#pragma unroll
        for (int v = 0; v < S; ++v) qR[v] = qL[v];
    }
    hllc_face<NF, D, DIR, R>(qL, qR, a.eos, f, us);
    // 3. the fluxes go through the dead stencil columns to the x-fastest
    // thread layout (C consecutive x per C threads), which finishes cell
    // y0+fs of column x0+fx: the finish's HBM reads and writes then cover
    // rows of C doubles instead of 64/K-column pieces (measured with K = 16,
    // C = 16: 9 % faster z sweep)
    __syncthreads();
#pragma unroll
    for (int v = 0; v < S; ++v) fl[v][ts][tx] = f[v];
    fl[S][ts][tx] = us;
    __syncthreads();
    const int fx = tid % XC, fs = tid / XC;
    const int i = x0 + fx, cc = y0 + fs;
    if (fs < TC && i <= a.m && cc <= s) {
        R Fm[S], Fp[S];
#pragma unroll
        for (int v = 0; v < S; ++v) {
            Fm[v] = fl[v][fs][fx];
            Fp[v] = fl[v][fs + 1][fx];
        }
        const Off go{sb, (unsigned)fx + (unsigned)(fs + 1 + PN) * gp}, ro{rb, (unsigned)fx + (unsigned)fs * rp};
        FinishIn<S> fin;
        if constexpr (KEEP && ADVK)
            finish_load_adv<NF, D>(a, go, ro, &lcons[0][fx][fs], XC * KP, fin);
        else if constexpr (KEEP)
            finish_load_staged<NF, D, FM>(a, go, ro, &lcons[0][fx][fs], XC * KP, fin);
        else
            finish_load<NF, D, FM>(a, go, ro, fin);
        finish_cell<NF, D, R, FM>(a, go, ro, a.inv_ds[cc], Fm, Fp, fl[S][fs][fx], fl[S][fs + 1][fx], fin);
    }
}

// ---------------------------------------------------------------------------
// Fused x and y sweep of the plain and contracted FP64 and the FP32-bulk
// forms (WENO5 on grids uniform in x and y; the FP32-bulk form reconstructs
// and solves in FP32 and finishes in FP64 with its volume-fraction clip; the
// plain form finishes as its separate x and y sweeps would): a 512-thread
// block finishes a tile of XT = 30 x YT = 14 cells of one z plane, so
// neither the rhs between the two directions nor a second pass over the
// stencil input goes through HBM.
//  1. The 36 x 20 cells x0-3..x0+32, y0-3..y0+16 are staged (cons->prim) in
//     LDS as columns of odd pitch; the tile's central 32 x 14 cells first,
//     one per thread of threads 0..447, which keep their conservative
//     values for the finish, then the halo ring.
//  2. y faces: 32 columns of 16 lanes (cell y0-1+lane of column x0+col,
//     both faces; the right state of face y+1/2 from the next lane), HLLC
//     of face y0-1+lane+1/2 by every lane.
//  3. x faces: 14 rows of 32 lanes (cell x0-1+lane of row y0+row, both
//     faces, the right state from the next lane), HLLC of face
//     x0-1+lane+1/2; the lower flux of a cell from the previous lane.
//  4. The y fluxes go to LDS over the dead staged cells and
//  5. every x-phase lane with a tile cell finishes it: both directions'
//     flux differences and alpha*div(u) sources, then the rhs store or the
//     spread RK stage (rhs from the z sweep in 3D, the rest of the base,
//     ci*in from the kept conservative values).
template <int NCX, int NRY, int PN>
__device__ __forceinline__ void xy_halo_cell(int h, int &cx, int &ry) {
    // the halo ring of the NCX x NRY staging around the central cells
    // (columns PN..NCX-PN-1 of rows PN+1..NRY-PN-2): the PN+1 rows at
    // either end whole, then the PN columns at either end of the rows
    // between
    constexpr int ROWS = PN + 1;
    if (h < 2 * ROWS * NCX) {
        const int band = h / NCX;
        cx = h % NCX;
        ry = band < ROWS ? band : band + NRY - 2 * ROWS;
    } else {
        const int hh = h - 2 * ROWS * NCX, q = hh % (2 * PN);
        ry = ROWS + hh / (2 * PN);
        cx = q < PN ? q : q + NCX - 2 * PN;
    }
}

template <int NF, int D, typename R>
__global__ __launch_bounds__(kXyThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_sweep_xy(const SweepArgs a) {
    static_assert(sizeof(R) == 8 || sizeof(R) == 4, "FP64 or FP32 reconstruction");
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    constexpr int PN = 2;
    constexpr int XT = kXyCellsX, XL = XT + 2; // finished cells per row, x-phase lanes per row
    constexpr int YT = kXyCellsY, YK = YT + 2; // finished rows, y-phase lanes per column
    constexpr int YC = XL;              // y-phase columns (XT of them finished)
    constexpr int NCX = XL + 2 * PN;    // staged columns x0-1-PN .. x0+XL-2+PN
    constexpr int NRY = YK + 2 * PN;    // staged rows y0-1-PN .. y0+YK-2+PN
    constexpr int P = NRY | 1;          // odd column pitch
    constexpr int NT = kXyThreads;
    constexpr int NCEN = XL * YT;       // central cells, the x phase's own
    constexpr int NHALO = NCX * NRY - NCEN;
    static_assert(XL == 32 && NCEN <= NT && NHALO <= 2 * NT - NCEN && YC * YK == NT, "tile shape");
    // y-flux rows padded to YC + 1: the 16 lanes of a column write one
    // column of the rows (with YC, all to one bank)
    constexpr int FP = YC + 1;
    constexpr int SP_WORDS = S * NCX * P, FY_WORDS = (S + 1) * (YT + 1) * FP;
    static_assert(FY_WORDS <= SP_WORDS, "y fluxes fit over the staged cells");
    __shared__ R lds[SP_WORDS];
    R(*sp)[NCX][P] = reinterpret_cast<R(*)[NCX][P]>(lds);
    R(*fyl)[YT + 1][FP] = reinterpret_cast<R(*)[YT + 1][FP]>(lds);

    const long bid = xcd_logical_block(blockIdx.x, a.nblocks);
    const int xt = (int)(bid % a.ntile);
    const long rest = bid / a.ntile;
    const int yt = (int)(rest % a.nchunk);
    const int k = (int)(rest / a.nchunk);
    const int x0 = xt * XT, y0 = yt * YT;
    const int tid = threadIdx.x;
    const unsigned gp = (unsigned)a.nxg;

    // 1. staging; cells past the ghost layers take the clamped cell's values
    // (their faces are never finished)
    double cin[S];
    {
        const long sb = ghost_off(a, x0 - 1 - PN, y0 - 1 - PN, k);
        auto soff = [&](int cx, int ry) -> Off {
            const int x = min(x0 - 1 - PN + cx, a.m + a.b), y = min(y0 - 1 - PN + ry, a.n + a.b);
            return Off{sb, (unsigned)(x - (x0 - 1 - PN)) + (unsigned)(y - (y0 - 1 - PN)) * gp};
        };
        int cx[2], ry[2];
        if (tid < NCEN) {
            cx[0] = tid % XL + PN;
            ry[0] = tid / XL + 1 + PN;
        } else {
            xy_halo_cell<NCX, NRY, PN>(tid - NCEN, cx[0], ry[0]);
        }
        const int h1 = tid + (NT - NCEN);
        const bool two = h1 < NHALO;
        if (two)
            xy_halo_cell<NCX, NRY, PN>(h1, cx[1], ry[1]);
        else
            cx[1] = cx[0], ry[1] = ry[0];
        double cv[2][S];
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const Off o = soff(cx[it], ry[it]);
#pragma unroll
            for (int v = 0; v < S; ++v) cv[it][v] = el(a.qin.v[v], o);
        }
#pragma unroll
        for (int v = 0; v < S; ++v) cin[v] = cv[0][v];
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            if (it == 0 || two) {
                double pr[S];
                cons_to_prim<NF, D>(cv[it], pr, a.eos);
#pragma unroll
                for (int v = 0; v < S; ++v) sp[v][cx[it]][ry[it]] = (R)pr[v];
            }
        }
    }
    __syncthreads();

    const R eps = (R)a.eps;
    // 2. y faces: column yc (cells x0+yc), lane yr: cell y0-1+yr
    R fy[S], usy;
    {
        const int yc = tid / YK, yr = tid % YK, cx = yc + 1 + PN;
        R cf[kNCoef], qL[S], qR[S];
#pragma unroll
        for (int q = 0; q < kNCoef; ++q) cf[q] = sizeof(R) == 8 ? (R)a.ucf2[q] : (R)a.ucf2_32[q];
#pragma unroll
        for (int v = 0; v < S; ++v) {
            R st[5], vl;
#pragma unroll
            for (int w = 0; w < 5; ++w) st[w] = sp[v][cx][yr + w];
            weno_cell<5, R, true, true>(st, cf, eps, vl, qL[v]);
            qR[v] = __shfl_down(vl, 1u, YK);
        }
        const int jf = y0 - 1 + yr;
        if (jf == -1 && a.bc2_beg == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qL[v] = qR[v];
        }
        if (jf == a.n && a.bc2_end == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qR[v] = qL[v];
        }
        hllc_face<NF, D, 1, R>(qL, qR, a.eos, fy, usy);
    }

    // 3. x faces: row xr (cells y0+xr), lane c: cell x0-1+c (wave 7 has none)
    const int xr = tid / XL, c = tid % XL;
    R fx[S], usx = R(0), fxm[S], usxm = R(0);
    if (tid < NCEN) {
        const int ry = xr + 1 + PN;
        R cf[kNCoef], qL[S], qR[S];
#pragma unroll
        for (int q = 0; q < kNCoef; ++q) cf[q] = sizeof(R) == 8 ? (R)a.ucf[q] : (R)a.ucf32[q];
#pragma unroll
        for (int v = 0; v < S; ++v) {
            R st[5], vl;
#pragma unroll
            for (int w = 0; w < 5; ++w) st[w] = sp[v][c + w][ry];
            weno_cell<5, R, true, true>(st, cf, eps, vl, qL[v]);
            qR[v] = __shfl_down(vl, 1u, XL);
        }
        const int jf = x0 - 1 + c;
        if (jf == -1 && a.bc_beg == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qL[v] = qR[v];
        }
        if (jf == a.m && a.bc_end == -4) {
#pragma unroll
            for (int v = 0; v < S; ++v) qR[v] = qL[v];
        }
        hllc_face<NF, D, 0, R>(qL, qR, a.eos, fx, usx);
#pragma unroll
        for (int v = 0; v < S; ++v) fxm[v] = __shfl_up(fx[v], 1u, XL);
        usxm = __shfl_up(usx, 1u, XL);
    }
    __syncthreads();

    // 4. y fluxes of faces y0-1/2 .. y0+YT-1/2 over the staged cells
    {
        const int yc = tid / YK, yr = tid % YK;
        if (yr <= YT) {
#pragma unroll
            for (int v = 0; v < S; ++v) fyl[v][yr][yc] = fy[v];
            fyl[S][yr][yc] = usy;
        }
    }
    __syncthreads();

    // 5. finish cell (x0-1+c, y0+xr), c = 1..XT
    const int i = x0 - 1 + c, j = y0 + xr;
    if (tid < NCEN && c >= 1 && c <= XT && i <= a.m && j <= a.n) {
        const int yc = c - 1;
        const Off go{ghost_off(a, x0, y0, k), (unsigned)yc + (unsigned)xr * gp};
        const Off ro{interior_off(a, x0, y0, k), (unsigned)yc + (unsigned)xr * (unsigned)(a.m + 1)};
        // flux differences in FP64 (the FP32-bulk form converts its fluxes)
        double dfy[S], dfx[S];
#pragma unroll
        for (int v = 0; v < S; ++v) {
            dfy[v] = to_d(fyl[v][xr][yc]) - to_d(fyl[v][xr + 1][yc]);
            dfx[v] = to_d(fxm[v]) - to_d(fx[v]);
        }
        const double dusy = to_d(fyl[S][xr + 1][yc]) - to_d(fyl[S][xr][yc]), dusx = to_d(usx) - to_d(usxm);
        const double ix = a.inv_ds[i], iy = a.inv_ds2[j];
        double r[S];
        if constexpr (!kFastF64 && sizeof(R) == 8) {
            // plain FP64: the x sweep's finish (the first direction: it
            // assigns) and then the y sweep's, operation for operation, so
            // the result is bitwise theirs; then the rhs store or the RK
            // stage (2D, where y is the last direction)
#pragma unroll
            for (int v = 0; v < S; ++v) r[v] = ix * dfx[v];
#pragma unroll
            for (int q = 0; q < NF; ++q) r[Y::ADV + q] = r[Y::ADV + q] + ix * cin[Y::ADV + q] * dusx;
#pragma unroll
            for (int v = 0; v < S; ++v) r[v] = r[v] + iy * dfy[v];
#pragma unroll
            for (int q = 0; q < NF; ++q) r[Y::ADV + q] = r[Y::ADV + q] + iy * cin[Y::ADV + q] * dusy;
            if (a.rk == kRkNone) {
#pragma unroll
                for (int v = 0; v < S; ++v) el(a.rhs.v[v], ro) = r[v];
            } else {
                FinishIn<S> f;
#pragma unroll
                for (int v = 0; v < S; ++v) {
                    f.in[v] = cin[v];
                    f.base[v] = a.base_is_in ? cin[v] : el(a.qbase.v[v], go);
                }
                double o[S];
                rk_combine<S>(a, f, r, o);
#pragma unroll
                for (int v = 0; v < S; ++v) el(a.qout.v[v], go) = o[v];
            }
            return;
        }
        if (a.fold) {
            // spread RK stage: rhs (z's part in 3D) + the rest of cb*base +
            // cdt*(L_x + L_y) + ci*in
            const double sx = a.cdt * ix, sy = a.cdt * iy;
#pragma unroll
            for (int v = 0; v < S; ++v) {
                double acc = a.accumulate ? el(a.rhs.v[v], ro) : 0.0;
                if (v >= a.base_lo && v < a.base_hi) acc = fma(a.cb, el(a.qbase.v[v], go), acc);
                acc = fma(sy, dfy[v], acc);
                r[v] = fma(sx, dfx[v], acc);
            }
#pragma unroll
            for (int q = 0; q < NF; ++q)
                r[Y::ADV + q] = r[Y::ADV + q] + (sx * dusx + sy * dusy) * cin[Y::ADV + q];
            if (a.ci != 0.0) {
#pragma unroll
                for (int v = 0; v < S; ++v) r[v] = r[v] + a.ci * cin[v];
            }
            if constexpr (sizeof(R) == 4) {
                // FP32 bulk: volume fractions clipped to [0, 1], as finish_cell
#pragma unroll
                for (int q = 0; q < NF; ++q) r[Y::ADV + q] = min(max(r[Y::ADV + q], 0.0), 1.0);
            }
#pragma unroll
            for (int v = 0; v < S; ++v) el(a.qout.v[v], go) = r[v];
        } else {
#pragma unroll
            for (int v = 0; v < S; ++v) {
                const double acc = a.accumulate ? el(a.rhs.v[v], ro) : 0.0;
                r[v] = acc + iy * dfy[v] + ix * dfx[v];
            }
#pragma unroll
            for (int q = 0; q < NF; ++q)
                r[Y::ADV + q] = r[Y::ADV + q] + (ix * dusx + iy * dusy) * cin[Y::ADV + q];
#pragma unroll
            for (int v = 0; v < S; ++v) el(a.rhs.v[v], ro) = r[v];
        }
    }
}

} // inline namespace MFC_FORM
} // namespace mfc
