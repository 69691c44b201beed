// mfc_launch.hpp -- dispatch of the sweep kernels over (num_dims, weno_order,
// precision, uniform-grid coefficients) for one number of fluids.  Each mfc_sweep_nf<N>.hip instantiates
// launch_sweep<N> so the kernel variants compile in parallel.
#pragma once

#include "mfc_sweep.hpp"

namespace mfc {

struct SweepLaunch {
    int dir;       // 0, 1, 2; 3: the fused x and y sweep (k_sweep_xy)
    int num_dims;  // 1..3
    int weno_order;
    int precision; // MFC_PREC_*
};

// Arithmetic forms, one set of translation units each: the plain FP64 form
// (built without FMA contraction), the FP32-bulk and learned forms, and the
// contracted FP64 form (the last two built with contraction).
enum SweepForm { kFormF64 = 0, kFormF32 = 1, kFormF64Fast = 2 };
template <int NF, int FORM>
int launch_sweep(const SweepLaunch &L, const SweepArgs &a, hipStream_t st);

#define MFC_SWEEP_HELPERS(NF)                                                                           \
    template <int D, int WO, typename R, bool ML, bool UNI>                                             \
    static int launch_nf##NF##_dwrm(int dir, const SweepArgs &a, hipStream_t st) {                      \
        const dim3 grid((unsigned)a.nblocks);                                                           \
        if (dir == 3) {                                                                                 \
            if constexpr (NF <= kXyMaxFluids && D >= 2 && WO == 5 && !ML && UNI) hipLaunchKernelGGL((k_sweep_xy<NF, D, R>), grid, dim3(kXyThreads), 0, st, a); \
            else return -1;                                                                             \
        } else if (dir == 0) {                                                                          \
            hipLaunchKernelGGL((k_sweep_x<NF, D, WO, R, ML, UNI>), grid, dim3(64), 0, st, a);                \
        } else if (dir == 1) {                                                                          \
            if constexpr (D >= 2 && t_cell_parallel<ML, R, UNI>()) hipLaunchKernelGGL((k_sweep_tc<NF, D, 1, WO, R, ML, UNI>), grid, dim3(tc_cols<R, ML>(1) * tc_rows<R>(1)), 0, st, a); \
            else if constexpr (D >= 2) hipLaunchKernelGGL((k_sweep_t<NF, D, 1, WO, R, ML, UNI>), grid, dim3(t_threads<ML>()), 0, st, a); \
            else return -1;                                                                             \
        } else {                                                                                        \
            if constexpr (D >= 3 && t_cell_parallel<ML, R, UNI>()) hipLaunchKernelGGL((k_sweep_tc<NF, D, 2, WO, R, ML, UNI>), grid, dim3(tc_cols<R, ML>(2) * tc_rows<R>(2)), 0, st, a); \
            else if constexpr (D >= 3) hipLaunchKernelGGL((k_sweep_t<NF, D, 2, WO, R, ML, UNI>), grid, dim3(t_threads<ML>()), 0, st, a); \
            else return -1;                                                                             \
        }                                                                                               \
        return hipGetLastError() == hipSuccess ? 0 : -2;                                                \
    }                                                                                                   \
    template <int D, int WO, typename R>                                                                \
    static int launch_nf##NF##_dwr(int dir, const SweepArgs &a, hipStream_t st) {                       \
        if constexpr (WO == 5)                                                                          \
            if (a.uniform) return launch_nf##NF##_dwrm<D, 5, R, false, true>(dir, a, st);               \
        return launch_nf##NF##_dwrm<D, WO, R, false, false>(dir, a, st);                                \
    }                                                                                                   \
    template <int D, typename R>                                                                        \
    static int launch_nf##NF##_dr(const SweepLaunch &L, const SweepArgs &a, hipStream_t st) {           \
        switch (L.weno_order) {                                                                         \
        case 1: return launch_nf##NF##_dwr<D, 1, R>(L.dir, a, st);                                      \
        case 3: return launch_nf##NF##_dwr<D, 3, R>(L.dir, a, st);                                      \
        case 5: return launch_nf##NF##_dwr<D, 5, R>(L.dir, a, st);                                      \
        default: return -1;                                                                             \
        }                                                                                               \
    }                                                                                                   \
    template <typename R>                                                                               \
    static int launch_nf##NF##_r(const SweepLaunch &L, const SweepArgs &a, hipStream_t st) {            \
        switch (L.num_dims) {                                                                           \
        case 1: return launch_nf##NF##_dr<1, R>(L, a, st);                                              \
        case 2: return launch_nf##NF##_dr<2, R>(L, a, st);                                              \
        case 3: return launch_nf##NF##_dr<3, R>(L, a, st);                                              \
        default: return -1;                                                                             \
        }                                                                                               \
    }

#define MFC_SWEEP_LAUNCH_F64(NF)                                                                        \
    MFC_SWEEP_HELPERS(NF)                                                                               \
    template <>                                                                                         \
    int launch_sweep<NF, kFormF64>(const SweepLaunch &L, const SweepArgs &a, hipStream_t st) {          \
        return L.precision == 0 ? launch_nf##NF##_r<double>(L, a, st) : -1;                             \
    }

#define MFC_SWEEP_LAUNCH_F64FAST(NF)                                                                    \
    MFC_SWEEP_HELPERS(NF)                                                                               \
    template <>                                                                                         \
    int launch_sweep<NF, kFormF64Fast>(const SweepLaunch &L, const SweepArgs &a, hipStream_t st) {      \
        return L.precision == 3 ? launch_nf##NF##_r<double>(L, a, st) : -1;                             \
    }

#define MFC_SWEEP_LAUNCH_F32(NF)                                                                        \
    MFC_SWEEP_HELPERS(NF)                                                                               \
    template <>                                                                                         \
    int launch_sweep<NF, kFormF32>(const SweepLaunch &L, const SweepArgs &a, hipStream_t st) {          \
        if (L.precision == 2) {                                                                         \
            if (L.weno_order != 5) return -1;                                                           \
            switch (L.num_dims) {                                                                       \
            case 1: return a.uniform ? launch_nf##NF##_dwrm<1, 5, float, true, true>(L.dir, a, st) : launch_nf##NF##_dwrm<1, 5, float, true, false>(L.dir, a, st);                       \
            case 2: return a.uniform ? launch_nf##NF##_dwrm<2, 5, float, true, true>(L.dir, a, st) : launch_nf##NF##_dwrm<2, 5, float, true, false>(L.dir, a, st);                       \
            case 3: return a.uniform ? launch_nf##NF##_dwrm<3, 5, float, true, true>(L.dir, a, st) : launch_nf##NF##_dwrm<3, 5, float, true, false>(L.dir, a, st);                       \
            default: return -1;                                                                         \
            }                                                                                           \
        }                                                                                               \
        return L.precision == 1 ? launch_nf##NF##_r<float>(L, a, st) : -1;                              \
    }

} // namespace mfc
