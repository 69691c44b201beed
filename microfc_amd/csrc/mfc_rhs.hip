// mfc_rhs.hip -- C ABI of the MI355X RHS library (declared in include/mfc_rhs.h)
// plus the small kernels around the sweeps: ghost-cell fill, cons->prim,
// the CFL max-reduction and the pointwise RK update.
#include "mfc_launch.hpp"
#include "mfc_viscous.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

extern "C" {
#include "../../include/mfc_rhs.h"
}

namespace mfc {
const MlpFrags *mlp_frags();
#define MFC_DECLARE_SWEEP(NF, FORM) \
    template <> int launch_sweep<NF, FORM>(const SweepLaunch &, const SweepArgs &, hipStream_t);
#define MFC_DECLARE_SWEEPS(FORM) \
    MFC_DECLARE_SWEEP(1, FORM) MFC_DECLARE_SWEEP(2, FORM) MFC_DECLARE_SWEEP(3, FORM) MFC_DECLARE_SWEEP(4, FORM)
#define MFC_DECLARE_SWEEPS_WIDE(FORM)                                                                 \
    MFC_DECLARE_SWEEPS(FORM) MFC_DECLARE_SWEEP(5, FORM) MFC_DECLARE_SWEEP(6, FORM) MFC_DECLARE_SWEEP(7, FORM) \
    MFC_DECLARE_SWEEP(8, FORM) MFC_DECLARE_SWEEP(9, FORM) MFC_DECLARE_SWEEP(10, FORM)
MFC_DECLARE_SWEEPS_WIDE(kFormF64)
MFC_DECLARE_SWEEPS(kFormF32)
MFC_DECLARE_SWEEPS_WIDE(kFormF64Fast)
#undef MFC_DECLARE_SWEEPS_WIDE
#undef MFC_DECLARE_SWEEPS
#undef MFC_DECLARE_SWEEP
} // namespace mfc

using namespace mfc;

constexpr int kReduceBlocks = 8192; // 32 blocks of 256 threads per CU, two cells per thread at 2048^2
constexpr int kMaxDevices = 64;

// Block partials of the allocation-free CFL reduction (mfc_cfl_dt_noalloc):
// one copy of this array per GPU, 64 KiB, held until the process ends.  In
// host code its name is not a device address; kernels get the address that
// hipGetSymbolAddress returned for the current device (g_scratch_dev).
static __device__ __align__(256) double scratch[kReduceBlocks];

namespace {

// Device address of `scratch` on each GPU, looked up once per device by
// mfc_rhs_init with that device current.
double *g_scratch_dev[kMaxDevices] = {};

constexpr int kNumFluidsBuilt = 10;   // the FP64 forms (every num_fluids the reference allows)
constexpr int kNumFluidsBuiltF32 = 4; // the FP32-bulk and learned forms
// Variables of the RK base that the z sweep of a contracted 3D RK stage
// adds (the y sweep adds the rest).  The step moves the same bytes for any
// split (512^3: 50.6-50.7 ms per step with 0, 2, 3 or 4); half on each
// balances the two sweeps (launches of 5.5 and 5.8 ms instead of 5.0 and
// 6.3 with the whole base on y).
constexpr int kBaseOnZ = 4;
// A grid direction counts as uniform in the contracted and FP32 forms when
// every cell's WENO coefficients are within this relative distance of the
// first reconstructed cell's.  Equal widths on paper differ by the rounding
// of the cell boundaries: 768 cells on [0, 1] give coefficients 2.6e-13
// apart, which moves the RHS by about as much relative to its scale, well
// inside the contracted form's 1e-12.  The plain FP64 form keeps the
// per-cell coefficients unless they are equal bit for bit.
constexpr double kUniformRelTol = 5e-13;
// The fused x and y sweep, and the RK-base variables the z sweep adds before
// it (512^3: 47.7, 46.0, 44.9, 44.8, 45.9 ms per step with 0, 2, 4, 6, 8 of
// the 8; the fused sweep is bound by VALU issue, the z sweep less so; FP32
// bulk the same way: 42.4, 41.0, 40.8, 41.3 ms with 2, 4, 6, 8).
constexpr bool kFuseXY = true;
constexpr int kBaseOnZFused = 6;

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return fail(-3, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct Dir {
    int s = 0;
    double *coef = nullptr;   // device, 27 x ncell
    double *inv_ds = nullptr; // device, s+1
    double *ds_dev = nullptr; // device, ds(0..s)
    double *cc_dev = nullptr; // device, ghosted cell centres s_cc(-b..s+b)
    double *inv_dsg = nullptr; // device, 1/ds(i) for i = -b..s+b (weno_Re_flux gradients)
    std::vector<double> coef_h, ds_h;
    bool ready = false;
    // every reconstructed cell has the same coefficients: bit for bit in the
    // plain FP64 form, to kUniformRelTol in the others
    bool uniform = false;
    double ucf[kNCoef] = {};
    bool uniform_wide = false; // the same over every cell the coefficients are set for (-b+polyn..s+b-polyn)
};

struct State {
    bool init = false;
    mfc_params prm{};
    int D = 1, S = 0, b = 0, polyn = 0;
    long nxg = 0, nyg = 0, nzg = 0, ng = 0, nint = 0;
    Dir dir[3];
    hipStream_t own = nullptr, stream = nullptr;
    double *partials = nullptr; // 2 * kReduceBlocks doubles
    double *scratch = nullptr;  // this device's `scratch` array (kReduceBlocks doubles)
    double *scratch_state = nullptr; // S*ng doubles, RK3 stage buffer
    bool visc = false;               // any fluid_pp%Re > 0
    ViscParams vp{};
    double *visc_fs = nullptr;       // (S+2+D) x faces of one direction: HLLC flux, contact velocity, viscous flux
    long visc_nface = 0;
    bool wrf = false;                // viscous with weno_Re_flux: WENO-reconstructed velocity gradients
    double *visc_grad = nullptr;     // D*D planes of ng: cell-centred velocity gradients (weno_Re_flux)
    double *visc_fgrad = nullptr;    // D*D planes of visc_nface: their face averages (weno_Re_flux)
} G;

// ---------------------------------------------------------------------------
// Optional per-kernel timing with HIP events on the library's stream.
enum KernelId { kKBcFill = 0, kKConsToPrim, kKSweepX, kKSweepY, kKSweepZ, kKIcfl, kKRk1, kKViscous, kKSweepXY, kKCount };
const char *const kKernelNames[kKCount] = {"bc_fill", "cons_to_prim", "sweep_x", "sweep_y", "sweep_z", "icfl",
                                           "rk1_update", "viscous", "sweep_xy"};

struct Timing {
    bool on = false;
    struct Rec {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[kKCount] = {};
    long count[kKCount] = {};
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
} T;

struct TimedScope {
    int id;
    hipEvent_t a = nullptr;
    explicit TimedScope(int k) : id(k) { // k < 0: not timed
        if (k >= 0 && T.on && (a = T.get())) (void)hipEventRecord(a, G.stream);
    }
    ~TimedScope() {
        if (!a) return;
        hipEvent_t b = T.get();
        if (!b) return;
        (void)hipEventRecord(b, G.stream);
        T.pending.push_back({id, a, b});
    }
};

int timing_flush() {
    if (T.pending.empty()) return 0;
    HIPCHK(hipStreamSynchronize(G.stream));
    for (auto &r : T.pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        T.total_ms[r.id] += ms;
        T.count[r.id] += 1;
        T.pool.push_back(r.a);
        T.pool.push_back(r.b);
    }
    T.pending.clear();
    return 0;
}

// ---------------------------------------------------------------------------
// input checks: src/simulation/m_start_up.fpp:117-290 (the parts that apply)
int check_params(const mfc_params &p) {
    if (p.m <= 0) return fail(-1, "Unsupported value of m");
    if (p.n < 0) return fail(-1, "Unsupported value of n");
    if (p.p < 0) return fail(-1, "Unsupported value of p");
    if (p.p > 0 && p.n == 0) return fail(-1, "p > 0 requires n > 0");
    if (p.num_fluids < 1 || p.num_fluids > MFC_NUM_FLUIDS_MAX) return fail(-1, "Unsupported value of num_fluids");
    if (p.num_fluids > kNumFluidsBuilt)
        return fail(-1, "num_fluids = %d: kernels are built for 1..%d fluids", p.num_fluids, kNumFluidsBuilt);
    if (p.weno_order != 1 && p.weno_order != 3 && p.weno_order != 5) return fail(-1, "Unsupported value of weno_order");
    if (p.m + 1 < 5 * p.weno_order) return fail(-1, "Unsupported combination of values of m and weno_order");
    if (p.n > 0 && p.n + 1 < 5 * p.weno_order) return fail(-1, "Unsupported combination of values of n and weno_order");
    if (p.p > 0 && p.p + 1 < 5 * p.weno_order) return fail(-1, "Unsupported combination of values of p and weno_order");
    if (!(p.weno_eps > 0.0) || p.weno_eps > 1e-6) return fail(-1, "Unsupported value of weno_eps");
    if (p.time_stepper < 1 || p.time_stepper > 3) return fail(-1, "Unsupported value of time_stepper");
    const int bcs[3][2] = {{p.bc_x_beg, p.bc_x_end}, {p.bc_y_beg, p.bc_y_end}, {p.bc_z_beg, p.bc_z_end}};
    const char *nm[3] = {"bc_x", "bc_y", "bc_z"};
    const int D = 1 + (p.n > 0) + (p.p > 0);
    for (int d = 0; d < D; ++d) {
        for (int e = 0; e < 2; ++e)
            if (bcs[d][e] < -12 || bcs[d][e] > -1)
                return fail(-1, "Unsupported value of %s%%%s", nm[d], e ? "end" : "beg");
        if ((bcs[d][0] == -1) != (bcs[d][1] == -1))
            return fail(-1, "Unsupported combination of values of %s%%beg and %s%%end", nm[d], nm[d]);
    }
    for (int i = 0; i < p.num_fluids; ++i) {
        if (!(p.gammas[i] > 0.0)) return fail(-1, "Unsupported value of fluid_pp(%d)%%gamma", i + 1);
        if (p.pi_infs[i] < 0.0) return fail(-1, "Unsupported value of fluid_pp(%d)%%pi_inf", i + 1);
    }
    if (p.precision != MFC_PREC_FP64 && p.precision != MFC_PREC_FP32_BULK && p.precision != MFC_PREC_FP32_LEARNED &&
        p.precision != MFC_PREC_FP64_FAST)
        return fail(-1, "Unsupported precision");
    const bool f64 = p.precision == MFC_PREC_FP64 || p.precision == MFC_PREC_FP64_FAST;
    if (!f64 && p.num_fluids > kNumFluidsBuiltF32)
        return fail(-1, "num_fluids = %d: the FP32 forms are built for 1..%d fluids", p.num_fluids,
                    kNumFluidsBuiltF32);
    if (p.precision == MFC_PREC_FP32_LEARNED && p.weno_order != 5)
        return fail(-1, "learned WENO weights need weno_order = 5");
    bool visc = false;
    for (int i = 0; i < p.num_fluids; ++i)
        for (int r = 0; r < 2; ++r) visc = visc || p.Re[i][r] > 0.0;
    if (visc && D > 2) return fail(-1, "viscous terms are 1D and 2D only, as in the reference");
    if (visc && p.precision != MFC_PREC_FP64) return fail(-1, "viscous terms run in the plain FP64 form only");
    if (visc && p.num_fluids > kViscFluidsMax)
        return fail(-1, "viscous kernels are built for 1..%d fluids", kViscFluidsMax);
    // The FP64 weight divisions (ndiv) equal IEEE division while the squared
    // smoothness indicators stay within [1e-200, 1e100], which weno_eps >=
    // 1e-100 guarantees from below.
    if (f64 && p.weno_order > 1 && !(p.weno_eps >= 1e-100))
        return fail(-1, "weno_eps below 1e-100 is not supported in FP64 (the weight division is exact for "
                        "weno_eps >= 1e-100)");
    // The contracted form scales the WENO5 weights by the product of the
    // three squared smoothness indicators and normalises both faces of a
    // cell with one reciprocal of the product of their weight sums (>=
    // weno_eps^8 / 100): normal for weno_eps >= 1e-25, finite while the
    // stencil differences stay below about 1e18.
    if (p.precision == MFC_PREC_FP64_FAST && p.weno_order > 1 && !(p.weno_eps >= 1e-25))
        return fail(-1, "weno_eps below 1e-25 is not supported in the contracted FP64 form (the scaled weights "
                        "and the product of a cell's two weight sums must stay normal)");
    // In FP32, beta^2 >= weno_eps^2 must stay a normal float and d/beta^2 finite.
    if (!f64 && p.weno_order > 1 && !(p.weno_eps >= 1e-18))
        return fail(-1, "weno_eps below 1e-18 is not supported in the FP32 forms (its square underflows)");
    return 0;
}

// ---------------------------------------------------------------------------
// WENO coefficient tables on the host: src/simulation/m_weno.fpp:168-363.
// Same plane layout as mfc_device.hpp; x**2d0 is evaluated as x*x.
inline double sq(double x) { return x * x; }

void weno_tables(int s, int b, int order, const double *cb0, std::vector<double> &out) {
    const long nc = s + 2 * b + 1;
    const double *cb = cb0 + 1 + b; // cb[i] = s_cb(i)
    const int polyn = (order - 1) / 2;
    out.assign(27 * nc, 0.0);
    if (order == 1) return;
    auto C = [&](int plane, int j) -> double & { return out[(long)plane * nc + j + b]; };
    for (int i = -b - 1 + polyn; i <= s + b - 1 - polyn; ++i) {
        const int j = i + 1;
        const double S0 = cb[i], S1 = cb[i + 1], S2 = cb[i + 2], Sm1 = cb[i - 1];
        if (order == 3) {
            C(kCbR + 0, j) = (S0 - S1) / (S0 - S2);
            C(kCbR + 2, j) = (S0 - S1) / (Sm1 - S1);
            C(kCbL + 0, j) = -C(kCbR + 0, j);
            C(kCbL + 2, j) = -C(kCbR + 2, j);
            C(kDR + 0, j) = (Sm1 - S1) / (Sm1 - S2);
            C(kDL + 0, j) = (Sm1 - S0) / (Sm1 - S2);
            C(kDR + 1, j) = 1.0 - C(kDR + 0, j);
            C(kDL + 1, j) = 1.0 - C(kDL + 0, j);
            C(kBeta + 0, j) = 4.0 * sq(S0 - S1) / sq(S0 - S2);
            C(kBeta + 3, j) = 4.0 * sq(S0 - S1) / sq(Sm1 - S1);
            continue;
        }
        const double S3 = cb[i + 3], Sm2 = cb[i - 2];
        const double h4 = 4.0 * sq(S0 - S1);
        C(kCbR + 0, j) = ((S0 - S1) * (S1 - S2)) / ((S0 - S3) * (S3 - S1));
        C(kCbR + 2, j) = ((Sm1 - S1) * (S1 - S0)) / ((Sm1 - S2) * (S2 - S0));
        C(kCbR + 3, j) = ((S0 - S1) * (S1 - S2)) / ((Sm1 - S1) * (Sm1 - S2));
        C(kCbR + 5, j) = ((S0 - S1) * (S1 - Sm1)) / ((Sm2 - S0) * (Sm2 - S1));
        C(kCbL + 0, j) = ((S1 - S0) * (S0 - S2)) / ((S0 - S3) * (S3 - S1));
        C(kCbL + 2, j) = ((S0 - Sm1) * (S0 - S1)) / ((Sm1 - S2) * (S0 - S2));
        C(kCbL + 3, j) = ((S1 - S0) * (S0 - S2)) / ((Sm1 - S1) * (Sm1 - S2));
        C(kCbL + 5, j) = ((Sm1 - S0) * (S0 - S1)) / ((Sm2 - S0) * (Sm2 - S1));
        C(kCbR + 1, j) = ((S0 - S2) + (S1 - S3)) / ((S0 - S2) * (S0 - S3)) * ((S0 - S1));
        C(kCbR + 4, j) = ((Sm2 - S1) + (Sm1 - S1)) / ((Sm1 - S1) * (S1 - Sm2)) * ((S1 - S0));
        C(kCbL + 1, j) = ((S0 - S2) + (S0 - S3)) / ((S0 - S2) * (S0 - S3)) * ((S1 - S0));
        C(kCbL + 4, j) = ((Sm2 - S0) + (Sm1 - S1)) / ((Sm2 - S1) * (S1 - Sm1)) * ((S0 - S1));
        C(kDR + 0, j) = ((Sm2 - S1) * (S1 - Sm1)) / ((Sm2 - S3) * (S3 - Sm1));
        C(kDR + 2, j) = ((S1 - S2) * (S1 - S3)) / ((Sm2 - S2) * (Sm2 - S3));
        C(kDL + 0, j) = ((Sm2 - S0) * (S0 - Sm1)) / ((Sm2 - S3) * (S3 - Sm1));
        C(kDL + 2, j) = ((S0 - S2) * (S0 - S3)) / ((Sm2 - S2) * (Sm2 - S3));
        C(kDR + 1, j) = 1.0 - C(kDR + 0, j) - C(kDR + 2, j);
        C(kDL + 1, j) = 1.0 - C(kDL + 0, j) - C(kDL + 2, j);
        const double a10 = S1 - S0, a21 = S2 - S1, a20 = S2 - S0, a31 = S3 - S1, a0m = S0 - Sm1;
        C(kBeta + 0, j) = h4 * (10.0 * sq(a10) + a10 * a21 + sq(a21)) / (sq(S0 - S3) * sq(S1 - S3));
        C(kBeta + 1, j) = h4 * (19.0 * sq(a10) - a10 * a31 + 2.0 * a20 * (a20 + a31)) /
                          ((S0 - S2) * sq(S0 - S3) * a31);
        C(kBeta + 2, j) = h4 * (10.0 * sq(a10) + a10 * (a20 + a31) + sq(a20 + a31)) / (sq(S0 - S2) * sq(S0 - S3));
        C(kBeta + 3, j) = h4 * (10.0 * sq(a10) + sq(a0m) + a0m * a10) / (sq(Sm1 - S2) * sq(S0 - S2));
        C(kBeta + 4, j) = h4 * ((S0 - S1) * (a0m + 20.0 * a10) + (2.0 * a0m + a10) * a20) /
                          ((S1 - Sm1) * sq(Sm1 - S2) * a20);
        C(kBeta + 5, j) = h4 * (10.0 * sq(a10) + a10 * a21 + sq(a21)) / (sq(Sm1 - S1) * sq(Sm1 - S2));
        const double c2m = (S0 - Sm2) + (S0 - Sm1);
        C(kBeta + 6, j) = h4 * (12.0 * sq(a10) + sq(c2m) + 3.0 * c2m * a10) / (sq(Sm2 - S1) * sq(Sm1 - S1));
        C(kBeta + 7, j) = h4 * (19.0 * sq(a10) + ((S0 - Sm2) * (S0 - S1)) + 2.0 * (S1 - Sm1) * ((S0 - Sm2) + (S1 - Sm1))) /
                          ((Sm2 - S0) * sq(Sm2 - S1) * (S1 - Sm1));
        C(kBeta + 8, j) = h4 * (10.0 * sq(a10) + sq(a0m) + a0m * a10) / (sq(Sm2 - S0) * sq(Sm2 - S1));
    }
}

// ---------------------------------------------------------------------------
// small kernels

struct Box {
    int m, n, p, b, bj, bk;
    long nxg, nyg;
};

__device__ __forceinline__ long box_off(const Box &x, int i, int j, int k) {
    return (long)(i + x.b) + x.nxg * ((long)(j + x.bj) + x.nyg * (long)(k + x.bk));
}

// Source index of a ghost coordinate for one direction's boundary condition
// (m_rhs.fpp:692-905): <= -3 extrapolation, -2 reflection, -1 periodic.
__device__ __forceinline__ int bc_source(int bcb, int bce, int s, int i, bool &flip) {
    if (i < 0) {
        if (bcb <= -3) return 0;
        if (bcb == -2) { flip = !flip; return -i - 1; }
        return s + 1 + i;
    }
    if (i > s) {
        if (bce <= -3) return s;
        if (bce == -2) { flip = !flip; return 2 * s + 1 - i; }
        return i - (s + 1);
    }
    return i;
}

// Ghost fill.  The reference fills x ghosts for interior rows, then y ghosts
// over all x, then z ghosts over all x and y, so a corner holds the interior
// cell reached by applying each direction's map in turn, with the momentum
// component normal to each reflecting boundary negated.  Each thread here
// fills one ghost cell with that composed map.  blockIdx.y selects the slab:
// 0: x ghosts (all j, k), 1: y ghosts (interior i, all k), 2: z ghosts
// (interior i, j); together they cover every ghost cell once.
__global__ void k_bc_fill(Fields q, int S, int momxb, Box x, int bxb, int bxe, int byb, int bye, int bzb,
                          int bze, long n0, long n1, long n2) {
    const int slab = blockIdx.y;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    const int gx = 2 * x.b, gy = 2 * x.bj;
    const int NY = x.n + 1 + gy, NX = x.m + 1;
    if (slab == 0) {
        if (t >= n0) return;
        const int gi = (int)(t % gx);
        const long r = t / gx;
        const int jj = (int)(r % NY);
        const int kk = (int)(r / NY);
        i = gi < x.b ? gi - x.b : x.m + 1 + (gi - x.b);
        j = jj - x.bj;
        k = kk - x.bk;
    } else if (slab == 1) {
        if (t >= n1) return;
        const int ii = (int)(t % NX);
        const long r = t / NX;
        const int gj = (int)(r % gy);
        const int kk = (int)(r / gy);
        i = ii;
        j = gj < x.bj ? gj - x.bj : x.n + 1 + (gj - x.bj);
        k = kk - x.bk;
    } else {
        if (t >= n2) return;
        const int ii = (int)(t % NX);
        const long r = t / NX;
        const int jj = (int)(r % (x.n + 1));
        const int gk = (int)(r / (x.n + 1));
        i = ii;
        j = jj;
        k = gk < x.bk ? gk - x.bk : x.p + 1 + (gk - x.bk);
    }
    bool fx = false, fy = false, fz = false;
    const int si = bc_source(bxb, bxe, x.m, i, fx);
    const int sj = x.bj ? bc_source(byb, bye, x.n, j, fy) : j;
    const int sk = x.bk ? bc_source(bzb, bze, x.p, k, fz) : k;
    const long dst = box_off(x, i, j, k), src = box_off(x, si, sj, sk);
    for (int v = 0; v < S; ++v) {
        double val = q.v[v][src];
        if ((fx && v == momxb) || (fy && v == momxb + 1) || (fz && v == momxb + 2)) val = -val;
        q.v[v][dst] = val;
    }
}

template <int NF, int D>
__global__ void k_cons_to_prim(Fields qc, Fields qp, Eos eos, long ng) {
    constexpr int S = Sys<NF, D>::S;
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= ng) return;
    double c[S], p[S];
#pragma unroll
    for (int v = 0; v < S; ++v) c[v] = qc.v[v][o];
    cons_to_prim<NF, D>(c, p, eos);
#pragma unroll
    for (int v = 0; v < S; ++v) qp.v[v][o] = p[v];
}

// NaN-propagating maximum: once a NaN is seen it stays.
__device__ __forceinline__ double max_nan(double a, double b) {
    return (a != a || b != b) ? (a != a ? a : b) : (a > b ? a : b);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max_nan(v, __shfl_xor(v, off, 64));
    return v;
}

// ICFL of one cell (m_data_output.fpp:197-244) from its S values: primitive
// ones, or conservative ones converted first (FROM_CONS; the values are
// identical because the reference's q_prim_vf is that same conversion).
template <int NF, int D, bool FROM_CONS>
__device__ __forceinline__ double icfl_cell(const double *cv, const Eos &eos, const double *dx, const double *dy,
                                            const double *dz, int i, int j, int k, double dt) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    double pr[S];
    if constexpr (FROM_CONS) {
        cons_to_prim<NF, D>(cv, pr, eos);
    } else {
#pragma unroll
        for (int v = 0; v < S; ++v) pr[v] = cv[v];
    }
    double rho = 0.0, gam = 0.0, pinf = 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        rho = rho + pr[f];
        gam = gam + pr[Y::ADV + f] * eos.gammas[f];
        pinf = pinf + pr[Y::ADV + f] * eos.pi_infs[f];
    }
    const double pres = pr[Y::E];
    const double c = sqrt(((gam + 1.0) * pres + pinf) / (gam * rho));
    if constexpr (D == 1) {
        return (dt / dx[i]) * (fabs(pr[Y::MOM]) + c);
    } else if constexpr (D == 2) {
        return dt / min_ref(dx[i] / (fabs(pr[Y::MOM]) + c), dy[j] / (fabs(pr[Y::MOM + 1]) + c));
    } else {
        return dt / min_ref(min_ref(dx[i] / (fabs(pr[Y::MOM]) + c), dy[j] / (fabs(pr[Y::MOM + 1]) + c)),
                            dz[k] / (fabs(pr[Y::MOM + 2]) + c));
    }
}

// Maximum ICFL over the interior and a block max-reduction by wavefront
// shuffles; block results go to partials[blockIdx.x].  A block takes chunks
// of 512 consecutive interior cells (x fastest), two per thread with both
// cells' loads issued before either is converted.
template <int NF, int D, bool FROM_CONS>
__global__ __launch_bounds__(256) void k_icfl(Fields q, Eos eos, Box x, const double *dx, const double *dy,
                                              const double *dz, double dt, long nint, double *partials) {
    constexpr int S = Sys<NF, D>::S;
    double mx = -INFINITY;
    for (long c0 = (long)blockIdx.x * 512; c0 < nint; c0 += (long)gridDim.x * 512) {
        double cv[2][S];
        int ci[2], cj[2], ck[2];
        bool ok[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const long e = c0 + threadIdx.x + 256 * h;
            ok[h] = e < nint;
            const long ee = ok[h] ? e : 0;
            const long r = ee / (x.m + 1);
            ci[h] = (int)(ee - r * (x.m + 1));
            cj[h] = (int)(r % (x.n + 1));
            ck[h] = (int)(r / (x.n + 1));
            const long o = box_off(x, ci[h], cj[h], ck[h]);
#pragma unroll
            for (int v = 0; v < S; ++v) cv[h][v] = q.v[v][o];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (ok[h]) mx = max_nan(mx, icfl_cell<NF, D, FROM_CONS>(cv[h], eos, dx, dy, dz, ci[h], cj[h], ck[h], dt));
    }
    __shared__ double red[4];
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = max_nan(v, red[w]);
        partials[blockIdx.x] = v;
    }
}

// Stability criteria of one cell (m_data_output.fpp:214-263) from its
// primitive values, for viscous cases: the mixture Reynolds numbers of
// s_convert_species_to_mixture_variables_acc (m_variables_conversion.fpp:
// 210-223), VCFL = maxval(dt/Re)/min(dx, dy)**2 and
// Rc = min(dx(|u|+c), dy(|v|+c))/maxval(1/Re) (1D: dx alone).
template <int NF, int D>
__device__ __forceinline__ void visc_criteria(const double *pr, const Eos &eos, const ViscParams &vp, double dxi,
                                              double dyj, double dt, double &vcfl, double &rc) {
    using Y = Sys<NF, D>;
    double rho = 0.0, gam = 0.0, pinf = 0.0, Re[2];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        rho = rho + pr[f];
        gam = gam + pr[Y::ADV + f] * eos.gammas[f];
        pinf = pinf + pr[Y::ADV + f] * eos.pi_infs[f];
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        double v = -1e6; // dflt_real
        if (vp.size[r] > 0) v = 0.0;
        for (int q = 0; q < vp.size[r]; ++q) v = pr[Y::ADV + vp.idx[r][q]] / vp.res[r][q] + v;
        Re[r] = 1.0 / max_ref(v, 1e-16);
    }
    const double c = sqrt(((gam + 1.0) * pr[Y::E] + pinf) / (gam * rho));
    const double dtre = max_ref(dt / Re[0], dt / Re[1]);
    const double inv = max_ref(1.0 / Re[0], 1.0 / Re[1]);
    if constexpr (D == 1) {
        vcfl = dtre / (dxi * dxi);
        rc = dxi * (fabs(pr[Y::MOM]) + c) / inv;
    } else {
        const double h = min_ref(dxi, dyj);
        vcfl = dtre / (h * h);
        rc = min_ref(dxi * (fabs(pr[Y::MOM]) + c), dyj * (fabs(pr[Y::MOM + 1]) + c)) / inv;
    }
}

// VCFL maximum and Rc minimum over the interior (as the negated maximum of
// -Rc, so both reduce with max_nan); block results to partials[blockIdx.x]
// and partials[gridDim.x + blockIdx.x].
template <int NF, int D>
__global__ __launch_bounds__(256) void k_visc_criteria(Fields q, Eos eos, ViscParams vp, Box x, const double *dx,
                                                       const double *dy, double dt, long nint, double *partials) {
    constexpr int S = Sys<NF, D>::S;
    double mv = -INFINITY, mr = -INFINITY;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < nint; e += (long)gridDim.x * 256) {
        const long r = e / (x.m + 1);
        const int i = (int)(e - r * (x.m + 1));
        const int j = (int)(r % (x.n + 1));
        const long o = box_off(x, i, j, 0);
        double pr[S];
#pragma unroll
        for (int v = 0; v < S; ++v) pr[v] = q.v[v][o];
        double vcfl, rc;
        visc_criteria<NF, D>(pr, eos, vp, dx[i], D > 1 ? dy[j] : 0.0, dt, vcfl, rc);
        mv = max_nan(mv, vcfl);
        mr = max_nan(mr, -rc);
    }
    __shared__ double red[2][4];
    mv = wave_max(mv);
    mr = wave_max(mr);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = mv;
        red[1][threadIdx.x >> 6] = mr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0][0], b = red[1][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            a = max_nan(a, red[0][w]);
            b = max_nan(b, red[1][w]);
        }
        partials[blockIdx.x] = a;
        partials[gridDim.x + blockIdx.x] = b;
    }
}

// Final pass over the block partials: one block of 1024 threads (eight
// partials per thread at 8192 blocks).
__global__ __launch_bounds__(1024) void k_reduce_max(double *partials, int n) {
    double mx = -INFINITY;
    for (int t = threadIdx.x; t < n; t += blockDim.x) mx = max_nan(mx, partials[t]);
    __shared__ double red[16];
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v = max_nan(v, red[w]);
        partials[0] = v;
    }
}

// Pointwise RK1 update q = q + dt*rhs on the interior
// (m_time_steppers.fpp:165-175); in place, so it cannot be fused into a
// stencil sweep that reads q.
__global__ void k_rk1_update(Fields q, Fields rhs, int S, Box x, double dt, long nint) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nint) return;
    const int i = (int)(t % (x.m + 1));
    const long r = t / (x.m + 1);
    const int j = (int)(r % (x.n + 1));
    const int k = (int)(r / (x.n + 1));
    const long o = box_off(x, i, j, k);
    for (int v = 0; v < S; ++v) q.v[v][o] = q.v[v][o] + dt * rhs.v[v][t];
}

// ---------------------------------------------------------------------------
// host helpers

Box box() {
    Box x;
    x.m = G.prm.m;
    x.n = G.prm.n;
    x.p = G.prm.p;
    x.b = G.b;
    x.bj = G.prm.n > 0 ? G.b : 0;
    x.bk = G.prm.p > 0 ? G.b : 0;
    x.nxg = G.nxg;
    x.nyg = G.nyg;
    return x;
}

Eos eos() {
    Eos e;
    for (int i = 0; i < kNumFluidsMax; ++i) {
        e.gammas[i] = G.prm.gammas[i];
        e.pi_infs[i] = G.prm.pi_infs[i];
    }
    return e;
}

int fields(double *const *ptrs, Fields &f, const char *what) {
    if (!ptrs) return fail(-1, "%s: NULL field array", what);
    for (int v = 0; v < kMaxSys; ++v) f.v[v] = nullptr;
    for (int v = 0; v < G.S; ++v) {
        if (!ptrs[v]) return fail(-1, "%s: NULL pointer for variable %d", what, v + 1);
        f.v[v] = ptrs[v];
    }
    return 0;
}

int require_ready() {
    if (!G.init) return fail(-4, "mfc_rhs_init has not been called");
    for (int d = 0; d < G.D; ++d)
        if (!G.dir[d].ready) return fail(-4, "mfc_rhs_set_grid has not been called for direction %d", d);
    return 0;
}

template <typename F>
int dispatch_nf_d(F &&f) {
    const int nf = G.prm.num_fluids, D = G.D;
#define MFC_CASE(NF, DD) \
    if (nf == NF && D == DD) return f(std::integral_constant<int, NF>{}, std::integral_constant<int, DD>{});
    MFC_CASE(1, 1) MFC_CASE(1, 2) MFC_CASE(1, 3)
    MFC_CASE(2, 1) MFC_CASE(2, 2) MFC_CASE(2, 3)
    MFC_CASE(3, 1) MFC_CASE(3, 2) MFC_CASE(3, 3)
    MFC_CASE(4, 1) MFC_CASE(4, 2) MFC_CASE(4, 3)
    MFC_CASE(5, 1) MFC_CASE(5, 2) MFC_CASE(5, 3)
    MFC_CASE(6, 1) MFC_CASE(6, 2) MFC_CASE(6, 3)
    MFC_CASE(7, 1) MFC_CASE(7, 2) MFC_CASE(7, 3)
    MFC_CASE(8, 1) MFC_CASE(8, 2) MFC_CASE(8, 3)
    MFC_CASE(9, 1) MFC_CASE(9, 2) MFC_CASE(9, 3)
    MFC_CASE(10, 1) MFC_CASE(10, 2) MFC_CASE(10, 3)
#undef MFC_CASE
    return fail(-1, "no kernels for num_fluids=%d, num_dims=%d", nf, D);
}

int bc_fill(const Fields &q) {
    const Box x = box();
    const long n0 = (long)(2 * x.b) * (x.n + 1 + 2 * x.bj) * (x.p + 1 + 2 * x.bk);
    const long n1 = (long)(x.m + 1) * (2 * x.bj) * (x.p + 1 + 2 * x.bk);
    const long n2 = (long)(x.m + 1) * (x.n + 1) * (2 * x.bk);
    const long nmax = std::max(n0, std::max(n1, n2));
    const dim3 grid((unsigned)((nmax + 255) / 256), 1 + (G.D >= 2) + (G.D >= 3)), block(256);
    TimedScope ts(kKBcFill);
    hipLaunchKernelGGL(k_bc_fill, grid, block, 0, G.stream, q, G.S, G.prm.num_fluids, x, G.prm.bc_x_beg,
                       G.prm.bc_x_end, G.prm.bc_y_beg, G.prm.bc_y_end, G.prm.bc_z_beg, G.prm.bc_z_end, n0, n1, n2);
    HIPCHK(hipGetLastError());
    return 0;
}

int cons_to_prim_launch(const Fields &qc, const Fields &qp) {
    const Eos e = eos();
    TimedScope ts(kKConsToPrim);
    return dispatch_nf_d([&](auto NF, auto D) -> int {
        hipLaunchKernelGGL((k_cons_to_prim<decltype(NF)::value, decltype(D)::value>),
                           dim3((unsigned)((G.ng + 255) / 256)), dim3(256), 0, G.stream, qc, qp, e, G.ng);
        HIPCHK(hipGetLastError());
        return 0;
    });
}

SweepArgs sweep_args(int d) {
    SweepArgs a;
    memset(&a, 0, sizeof a);
    a.coef = G.dir[d].coef;
    a.inv_ds = G.dir[d].inv_ds;
    a.eos = eos();
    a.eps = G.prm.weno_eps;
    a.nxg = G.nxg;
    a.nyg = G.nyg;
    a.ncell = G.dir[d].s + 2 * G.b + 1;
    a.m = G.prm.m;
    a.n = G.prm.n;
    a.p = G.prm.p;
    a.b = G.b;
    a.bj = G.prm.n > 0 ? G.b : 0;
    a.bk = G.prm.p > 0 ? G.b : 0;
    const int bcs[3][2] = {{G.prm.bc_x_beg, G.prm.bc_x_end}, {G.prm.bc_y_beg, G.prm.bc_y_end},
                           {G.prm.bc_z_beg, G.prm.bc_z_end}};
    a.bc_beg = bcs[d][0];
    a.bc_end = bcs[d][1];
    a.uniform = G.dir[d].uniform ? 1 : 0;
    for (int q = 0; q < kNCoef; ++q) {
        a.ucf[q] = G.dir[d].ucf[q];
        a.ucf32[q] = (float)G.dir[d].ucf[q];
    }
    if (d == 0) {
        // the launcher picks the uniform-coefficient kernels for WENO5 on uniform grids
        const bool uni = G.prm.weno_order == 5 && G.dir[0].uniform;
        const int tw = G.prm.precision == MFC_PREC_FP32_LEARNED ? x_tile_cells<true, float, true>()
                       : G.prm.precision == MFC_PREC_FP32_BULK
                           ? (uni ? x_tile_cells<false, float, true>() : x_tile_cells<false, float, false>())
                           : (uni ? x_tile_cells<false, double, true>() : x_tile_cells<false, double, false>());
        a.ntile = (G.prm.m + 1 + tw - 1) / tw;
        a.nblocks = (long)a.ntile * (G.prm.n + 1) * (G.prm.p + 1);
    } else {
        const int s = G.dir[d].s;
        const bool uni = G.prm.weno_order == 5 && G.dir[d].uniform;
        const bool cp = G.prm.precision == MFC_PREC_FP32_LEARNED ? t_cell_parallel<true, float, true>()
                        : G.prm.precision == MFC_PREC_FP32_BULK
                            ? (uni ? t_cell_parallel<false, float, true>() : t_cell_parallel<false, float, false>())
                            : (uni ? t_cell_parallel<false, double, true>() : t_cell_parallel<false, double, false>());
        const bool f64 = G.prm.precision == MFC_PREC_FP64 || G.prm.precision == MFC_PREC_FP64_FAST;
        const int xc = !cp ? kTX : tc_cols_of(G.prm.precision == MFC_PREC_FP32_LEARNED, d);
        a.ntile = (G.prm.m + 1 + xc - 1) / xc;
        a.chunk = G.prm.precision == MFC_PREC_FP32_LEARNED ? t_tile_cells<true, float, true>(d)
                  : G.prm.precision == MFC_PREC_FP32_BULK
                      ? (uni ? t_tile_cells<false, float, true>(d) : t_tile_cells<false, float, false>(d))
                  : !uni ? t_tile_cells<false, double, false>(d)
                         : tc_rows_of(true, G.prm.precision == MFC_PREC_FP64_FAST, d) - 2;
        a.nchunk = (s + 1 + a.chunk - 1) / a.chunk;
        const int other = d == 1 ? G.prm.p + 1 : G.prm.n + 1;
        a.nblocks = (long)a.ntile * a.nchunk * other;
    }
    return a;
}

// Arguments of the fused x and y sweep: the x direction's, the y
// direction's cell widths, boundary codes and coefficients, and its tiles.
SweepArgs xy_args() {
    SweepArgs a = sweep_args(0);
    a.inv_ds2 = G.dir[1].inv_ds;
    a.bc2_beg = G.prm.bc_y_beg;
    a.bc2_end = G.prm.bc_y_end;
    for (int q = 0; q < kNCoef; ++q) {
        a.ucf2[q] = G.dir[1].ucf[q];
        a.ucf2_32[q] = (float)G.dir[1].ucf[q];
    }
    a.ntile = (G.prm.m + 1 + kXyCellsX - 1) / kXyCellsX;
    a.nchunk = (G.prm.n + 1 + kXyCellsY - 1) / kXyCellsY;
    a.nblocks = (long)a.ntile * a.nchunk * (G.prm.p + 1);
    return a;
}

// d = 0, 1, 2: one direction; d = 3: the fused x and y sweep.
int sweep(int d, SweepArgs &a) {
    SweepLaunch L{d, G.D, G.prm.weno_order, G.prm.precision};
    if (G.prm.precision == MFC_PREC_FP32_LEARNED) {
        a.mlp = mlp_frags();
        if (!a.mlp) return fail(-4, "learned weights requested but mfc_mlp_init has not been called");
    }
    TimedScope ts(d == 3 ? kKSweepXY : kKSweepX + d);
    int rc;
    const int form = G.prm.precision == MFC_PREC_FP64        ? kFormF64
                     : G.prm.precision == MFC_PREC_FP64_FAST ? kFormF64Fast
                                                             : kFormF32;
    auto go = [&](auto nf) -> int {
        constexpr int NF = decltype(nf)::value;
        if (form == kFormF64) return launch_sweep<NF, kFormF64>(L, a, G.stream);
        if (form == kFormF64Fast) return launch_sweep<NF, kFormF64Fast>(L, a, G.stream);
        if constexpr (NF <= kNumFluidsBuiltF32) return launch_sweep<NF, kFormF32>(L, a, G.stream);
        return -1;
    };
    switch (G.prm.num_fluids) {
    case 1: rc = go(std::integral_constant<int, 1>{}); break;
    case 2: rc = go(std::integral_constant<int, 2>{}); break;
    case 3: rc = go(std::integral_constant<int, 3>{}); break;
    case 4: rc = go(std::integral_constant<int, 4>{}); break;
    case 5: rc = go(std::integral_constant<int, 5>{}); break;
    case 6: rc = go(std::integral_constant<int, 6>{}); break;
    case 7: rc = go(std::integral_constant<int, 7>{}); break;
    case 8: rc = go(std::integral_constant<int, 8>{}); break;
    case 9: rc = go(std::integral_constant<int, 9>{}); break;
    case 10: rc = go(std::integral_constant<int, 10>{}); break;
    default: rc = -1;
    }
    if (rc != 0) return fail(-3, "sweep launch failed (dir %d): %s", d, hipGetErrorString(hipGetLastError()));
    return 0;
}

// Direction d of a viscous RHS evaluation (inviscid and viscous fluxes of
// every face, then the finish: m_rhs.fpp:565-652), with sa the direction's
// finish arguments (rhs assignment or accumulation, or the RK stage).
int viscous(int d, const Fields &qin, const Fields &rhs, const SweepArgs &sa) {
    ViscArgs v;
    memset(&v, 0, sizeof v);
    v.q = qin;
    v.rhs = rhs;
    v.fs = G.visc_fs;
    v.nface = d == 0 ? (long)(G.prm.m + 2) * (G.prm.n + 1) : (long)(G.prm.m + 1) * (G.prm.n + 2);
    v.coef = G.dir[d].coef;
    v.ncell = G.dir[d].s + 2 * G.b + 1;
    v.uniform = G.dir[d].uniform ? 1 : 0;
    for (int q = 0; q < kNCoef; ++q) v.ucf[q] = G.dir[d].ucf[q];
    v.ccx = G.dir[0].cc_dev;
    v.ccy = G.D > 1 ? G.dir[1].cc_dev : nullptr;
    v.inv_ds = G.dir[d].inv_ds;
    v.eos = eos();
    v.eps = G.prm.weno_eps;
    v.nxg = G.nxg;
    v.m = G.prm.m;
    v.n = G.prm.n;
    v.b = G.b;
    v.bj = G.prm.n > 0 ? G.b : 0;
    const int bcs[2][2] = {{G.prm.bc_x_beg, G.prm.bc_x_end}, {G.prm.bc_y_beg, G.prm.bc_y_end}};
    v.bc_beg = bcs[d][0];
    v.bc_end = bcs[d][1];
    v.vp = G.vp;
    v.grad = G.visc_grad;
    v.fgrad = G.visc_fgrad;
    v.ng = G.ng;
    v.wrf = G.wrf ? 1 : 0;
    TimedScope ts(kKViscous);
    if (launch_viscous(G.prm.num_fluids, G.D, d, G.prm.weno_order, v, sa, G.stream))
        return fail(-3, "viscous launch failed (dir %d): %s", d, hipGetErrorString(hipGetLastError()));
    return 0;
}

// All directional sweeps of one RHS evaluation on the ghost-filled field qin.
// The last sweep either writes rhs (rk = kRkNone) or applies the RK stage.
// Viscous runs use the viscous kernels for each direction instead, with the
// same finish semantics.
int sweeps(const Fields &qin, const Fields &rhs, int rk, const Fields *qbase, const Fields *qout, double dt) {
    // The plain form adds the directions in the reference's order x, y, z
    // (m_rhs.fpp:469-656; it fixes the rounding).  The contracted FP64,
    // FP32-bulk and learned forms run them in the reverse order: the sum is
    // the same up to rounding, and the RK stage's output is then written by
    // the x sweep (or the fused x and y sweep), whose loads are contiguous
    // rows (learned form, 512^3: 68.2 -> 64.5 ms per step, 2048^2 1.47 ->
    // 1.35 ms).
    const bool reverse = (G.prm.precision == MFC_PREC_FP64_FAST || G.prm.precision == MFC_PREC_FP32_BULK ||
                          G.prm.precision == MFC_PREC_FP32_LEARNED) &&
                         !G.visc;
    if (G.wrf) {
        // cell-centred velocity gradients of the stage input (s_get_viscous,
        // m_rhs.fpp:456-464), before the directions
        for (int e = 0; e < G.D; ++e) {
            ViscGradArgs g;
            memset(&g, 0, sizeof g);
            g.q = qin;
            g.grad = G.visc_grad;
            g.coef = G.dir[e].coef;
            g.ncell = G.dir[e].s + 2 * G.b + 1;
            g.inv_dsg = G.dir[e].inv_dsg;
            g.eps = G.prm.weno_eps;
            g.nxg = G.nxg;
            g.ng = G.ng;
            g.m = G.prm.m;
            g.n = G.prm.n;
            g.b = G.b;
            g.bj = G.prm.n > 0 ? G.b : 0;
            g.s = G.dir[e].s;
            g.uniform = G.dir[e].uniform_wide ? 1 : 0;
            for (int q = 0; q < kNCoef; ++q) g.ucf[q] = G.dir[e].ucf[q];
            TimedScope ts(kKViscous);
            if (launch_visc_grad(G.prm.num_fluids, G.D, e, G.prm.weno_order, g, G.stream))
                return fail(-3, "viscous gradient launch failed (dir %d): %s", e,
                            hipGetErrorString(hipGetLastError()));
        }
    }
    // The contracted form spreads an RK stage over its sweeps (out = cb*base
    // + ci*in + cdt*L; the sweeps before the last read the base, the last
    // the stage input's own cells, which it stages anyway), so the x sweep
    // does not carry a fourth array.
    const bool fold = reverse && rk != kRkNone && G.D > 1;
    double cb = 1.0, ci = 0.0, cdt = dt;
    if (rk == kRk2Final) {
        cb = 0.5, ci = 0.5, cdt = 0.5 * dt;
    } else if (rk == kRk3Stage2) {
        cb = 0.75, ci = 0.25, cdt = 0.25 * dt;
    } else if (rk == kRk3Stage3) {
        cb = 1.0 / 3.0, ci = 2.0 / 3.0, cdt = 2.0 * dt / 3.0;
    }
    if (fold && qbase->v[0] == qin.v[0]) {
        // base = in (first stages): (cb + ci)*in, added by the last sweep
        // from its staged cells, and no sweep reads the base
        ci += cb;
        cb = 0.0;
    }
    // The contracted and FP32-bulk forms on grids uniform in x and y, WENO5,
    // one or two fluids, fuse the y and x sweeps (k_sweep_xy): in 3D after
    // the z sweep, in 2D alone.
    const bool fuse = reverse && kFuseXY && G.D >= 2 &&
                      (G.prm.precision == MFC_PREC_FP64_FAST || G.prm.precision == MFC_PREC_FP32_BULK) &&
                      G.prm.weno_order == 5 && G.prm.num_fluids <= kXyMaxFluids && G.dir[0].uniform &&
                      G.dir[1].uniform;
    if (fuse) {
        const int split = G.D == 3 ? kBaseOnZFused : 0;
        auto stage = [&](SweepArgs &a, int fold_kind, int lo, int hi) {
            a.qin = qin;
            a.rhs = rhs;
            a.dt = dt;
            a.rk = kRkNone;
            if (fold) {
                a.qbase = *qbase;
                a.qout = *qout;
                a.base_is_in = qbase->v[0] == qin.v[0];
                a.fold = fold_kind;
                if (cb != 0.0) a.base_lo = lo, a.base_hi = hi;
                a.cb = cb;
                a.ci = ci;
                a.cdt = cdt;
            }
        };
        if (G.D == 3) {
            SweepArgs a = sweep_args(2);
            stage(a, 1, 0, split);
            a.accumulate = 0;
            if (int rc = sweep(2, a)) return rc;
        }
        SweepArgs a = xy_args();
        stage(a, 3, split, G.S);
        a.accumulate = G.D == 3;
        return sweep(3, a);
    }
    // The plain form fuses its first two directions the same way (bitwise
    // the separate x and y sweeps): in 3D the z sweep then adds its part and
    // applies the RK stage, in 2D the fused sweep applies it.
    const bool fuse_plain = kFuseXY && !G.visc && G.D >= 2 && G.prm.precision == MFC_PREC_FP64 &&
                            G.prm.weno_order == 5 && G.prm.num_fluids <= kXyMaxFluids && G.dir[0].uniform &&
                            G.dir[1].uniform;
    if (fuse_plain) {
        auto finish = [&](SweepArgs &a, bool last) {
            a.qin = qin;
            a.rhs = rhs;
            a.dt = dt;
            a.rk = last ? rk : kRkNone;
            if (a.rk != kRkNone) {
                a.qbase = *qbase;
                a.qout = *qout;
                a.base_is_in = qbase->v[0] == qin.v[0];
            }
        };
        SweepArgs a = xy_args();
        finish(a, G.D == 2);
        a.accumulate = 0;
        if (int rc = sweep(3, a)) return rc;
        if (G.D == 3) {
            SweepArgs z = sweep_args(2);
            finish(z, true);
            z.accumulate = 1;
            return sweep(2, z);
        }
        return 0;
    }
    for (int k = 0; k < G.D; ++k) {
        const int d = reverse ? G.D - 1 - k : k;
        SweepArgs a = sweep_args(d);
        a.qin = qin;
        a.rhs = rhs;
        a.accumulate = k > 0;
        a.dt = dt;
        const bool last = k == G.D - 1;
        a.rk = last && !fold ? rk : kRkNone;
        if (a.rk != kRkNone || fold) {
            a.qbase = *qbase;
            a.qout = *qout;
            a.base_is_in = qbase->v[0] == qin.v[0];
        }
        if (fold) {
            a.fold = k == 0 ? 1 : last ? 3 : 2;
            // cb*base: in 3D variables [0, kBaseOnZ) on the z sweep, the rest
            // on y; in 2D all on y (the first sweep)
            const int split = G.D == 3 ? kBaseOnZ : 0;
            if (cb != 0.0 && k == 0 && G.D == 3) a.base_lo = 0, a.base_hi = split;
            if (cb != 0.0 && k == G.D - 2) a.base_lo = split, a.base_hi = G.S;
            a.cb = cb;
            a.ci = ci;
            a.cdt = cdt;
        }
        if (int rc = G.visc ? viscous(d, qin, rhs, a) : sweep(d, a)) return rc;
    }
    return 0;
}

} // namespace

// ===========================================================================
// C ABI

extern "C" {

const char *mfc_last_error(void) { return g_err.c_str(); }

int mfc_rhs_init(const mfc_params *prm) {
    if (!prm) return fail(-1, "NULL params");
    if (int rc = check_params(*prm)) return rc;
    if (G.init) mfc_rhs_finalize();
    G.prm = *prm;
    G.D = 1 + (prm->n > 0) + (prm->p > 0);
    G.S = 2 * prm->num_fluids + G.D + 1;
    G.polyn = (prm->weno_order - 1) / 2;
    // buff_size (m_global_parameters.fpp:355-359): viscous cases keep
    // 2*polyn + 2 ghost layers, inviscid ones polyn + 2
    bool visc = false;
    for (int i = 0; i < prm->num_fluids; ++i)
        for (int r = 0; r < 2; ++r) visc = visc || prm->Re[i][r] > 0.0;
    G.b = visc ? 2 * G.polyn + 2 : G.polyn + 2;
    G.nxg = prm->m + 2 * G.b + 1;
    G.nyg = prm->n > 0 ? prm->n + 2 * G.b + 1 : 1;
    G.nzg = prm->p > 0 ? prm->p + 2 * G.b + 1 : 1;
    G.ng = G.nxg * G.nyg * G.nzg;
    // The sweeps address a block's cells as a uniform base plus a 32-bit
    // byte offset per lane (Off in mfc_sweep.hpp); a z-sweep block spans at
    // most 40 ghosted planes
    if (G.nzg > 1 && 40L * G.nxg * G.nyg * (long)sizeof(double) >= (1L << 32))
        return fail(-1, "a z plane of %ld x %ld ghosted cells exceeds the sweeps' 32-bit block offsets", (long)G.nxg,
                    (long)G.nyg);
    G.nint = (long)(prm->m + 1) * (prm->n + 1) * (prm->p + 1);
    G.dir[0].s = prm->m;
    G.dir[1].s = prm->n;
    G.dir[2].s = prm->p;
    HIPCHK(hipStreamCreateWithFlags(&G.own, hipStreamNonBlocking));
    G.stream = G.own;
    HIPCHK(hipMalloc(&G.partials, sizeof(double) * 2 * kReduceBlocks));
    {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        if (dev < 0 || dev >= kMaxDevices) return fail(-3, "device ordinal %d outside 0..%d", dev, kMaxDevices - 1);
        if (!g_scratch_dev[dev]) {
            void *ptr = nullptr;
            HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            g_scratch_dev[dev] = static_cast<double *>(ptr);
        }
        G.scratch = g_scratch_dev[dev];
    }
    // Re_size, Re_idx, Res (m_global_parameters.fpp:314-338, m_viscous.fpp:36-42)
    for (int r = 0; r < 2; ++r)
        for (int i = 0; i < prm->num_fluids; ++i)
            if (prm->Re[i][r] > 0.0) {
                G.vp.idx[r][G.vp.size[r]] = i;
                G.vp.res[r][G.vp.size[r]] = prm->Re[i][r];
                G.vp.size[r]++;
            }
    G.visc = G.vp.size[0] > 0 || G.vp.size[1] > 0;
    if (G.visc) {
        G.visc_nface = std::max((long)(prm->m + 2) * (prm->n + 1), (long)(prm->m + 1) * (prm->n + 2));
        HIPCHK(hipMalloc(&G.visc_fs, sizeof(double) * (G.S + 2 + G.D) * G.visc_nface));
        G.wrf = prm->weno_Re_flux != 0;
        if (G.wrf) {
            HIPCHK(hipMalloc(&G.visc_grad, sizeof(double) * G.D * G.D * G.ng));
            HIPCHK(hipMalloc(&G.visc_fgrad, sizeof(double) * G.D * G.D * G.visc_nface));
        }
    }
    G.init = true;
    return 0;
}

int mfc_rhs_set_grid(int dir, const double *s_cb, const double *ds) {
    if (!G.init) return fail(-4, "mfc_rhs_init has not been called");
    if (dir < 0 || dir >= G.D) return fail(-1, "direction %d outside 0..%d", dir, G.D - 1);
    if (!s_cb || !ds) return fail(-1, "NULL grid array");
    Dir &D = G.dir[dir];
    const int s = D.s, b = G.b;
    for (int i = 0; i <= s; ++i)
        if (!(ds[i + b] > 0.0)) return fail(-1, "non-positive cell width at index %d of direction %d", i, dir);
    weno_tables(s, b, G.prm.weno_order, s_cb, D.coef_h);
    // Uniform coefficients over the cells the sweeps reconstruct (-1..s+1)?
    // Bit for bit in the plain FP64 form; within kUniformRelTol of cell -1's
    // in the others, whose uniform-grid kernels then use cell -1's.
    {
        const long nc = s + 2 * b + 1;
        const double tol = G.prm.precision == MFC_PREC_FP64 ? 0.0 : kUniformRelTol;
        D.uniform = true;
        for (int q = 0; q < kNCoef && D.uniform; ++q) {
            const double v0 = D.coef_h[(long)q * nc + (-1 + b)];
            D.ucf[q] = v0;
            for (int j = 0; j <= s + 1 && D.uniform; ++j) {
                const double v = D.coef_h[(long)q * nc + (j + b)];
                D.uniform = std::memcmp(&v0, &v, sizeof(double)) == 0 || std::fabs(v - v0) <= tol * std::fabs(v0);
            }
        }
    }
    // ... and over every cell with coefficients (the weno_Re_flux gradient
    // pass reconstructs the ghost cells too)
    {
        const long nc = s + 2 * b + 1;
        const int pn = (G.prm.weno_order - 1) / 2;
        D.uniform_wide = D.uniform;
        for (int q = 0; q < kNCoef && D.uniform_wide; ++q)
            for (int j = -b + pn; j <= s + b - pn && D.uniform_wide; ++j)
                D.uniform_wide = std::memcmp(&D.ucf[q], &D.coef_h[(long)q * nc + (j + b)], sizeof(double)) == 0;
    }
    std::vector<double> inv(s + 1), invg(s + 2 * b + 1);
    for (int i = 0; i <= s; ++i) inv[i] = 1.0 / ds[i + b];
    for (int i = -b; i <= s + b; ++i) invg[i + b] = 1.0 / ds[i + b]; // 1d0/dx(j), m_viscous.fpp:386-393
    // cell centres with ghosts (m_start_up.fpp:356-359, :551, :582)
    std::vector<double> cc(s + 2 * b + 1);
    for (int i = 0; i <= s; ++i) cc[i + b] = s_cb[i - 1 + 1 + b] + ds[i + b] / 2.0;
    for (int i = 1; i <= b; ++i) cc[-i + b] = cc[1 - i + b] - (ds[1 - i + b] + ds[-i + b]) / 2.0;
    for (int i = 1; i <= b; ++i) cc[s + i + b] = cc[s + (i - 1) + b] + (ds[s + (i - 1) + b] + ds[s + i + b]) / 2.0;
    D.ds_h.assign(ds, ds + s + 2 * b + 1);
    if (D.coef) HIPCHK(hipFree(D.coef));
    if (D.inv_ds) HIPCHK(hipFree(D.inv_ds));
    if (D.ds_dev) HIPCHK(hipFree(D.ds_dev));
    if (D.cc_dev) HIPCHK(hipFree(D.cc_dev));
    if (D.inv_dsg) HIPCHK(hipFree(D.inv_dsg));
    D.inv_dsg = nullptr;
    D.coef = nullptr;
    D.inv_ds = nullptr;
    D.ds_dev = nullptr;
    D.cc_dev = nullptr;
    HIPCHK(hipMalloc(&D.coef, sizeof(double) * D.coef_h.size()));
    HIPCHK(hipMalloc(&D.inv_ds, sizeof(double) * inv.size()));
    HIPCHK(hipMalloc(&D.ds_dev, sizeof(double) * (s + 1)));
    HIPCHK(hipMalloc(&D.cc_dev, sizeof(double) * cc.size()));
    HIPCHK(hipMemcpy(D.cc_dev, cc.data(), sizeof(double) * cc.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.ds_dev, ds + b, sizeof(double) * (s + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.coef, D.coef_h.data(), sizeof(double) * D.coef_h.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.inv_ds, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice));
    if (G.wrf) {
        HIPCHK(hipMalloc(&D.inv_dsg, sizeof(double) * invg.size()));
        HIPCHK(hipMemcpy(D.inv_dsg, invg.data(), sizeof(double) * invg.size(), hipMemcpyHostToDevice));
    }
    D.ready = true;
    return 0;
}

int mfc_compute_rhs(double *const *q_cons_vf, double *const *q_prim_vf, double *const *rhs_vf, int t_step) {
    if (int rc = require_ready()) return rc;
    Fields qc, qp, rhs;
    if (int rc = fields(q_cons_vf, qc, "q_cons_vf")) return rc;
    if (int rc = bc_fill(qc)) return rc;
    if (t_step == G.prm.t_step_stop) return 1;
    if (int rc = fields(rhs_vf, rhs, "rhs_vf")) return rc;
    if (q_prim_vf) {
        if (int rc = fields(q_prim_vf, qp, "q_prim_vf")) return rc;
        if (int rc = cons_to_prim_launch(qc, qp)) return rc;
    }
    return sweeps(qc, rhs, kRkNone, nullptr, nullptr, 0.0);
}

int mfc_tvd_rk_step(int t_step, double dt, double *const *q_cons_ts1, double *const *q_cons_ts2,
                    double *const *q_prim_vf, double *const *rhs_vf) {
    if (int rc = require_ready()) return rc;
    Fields q1, q2, qp, rhs;
    if (int rc = fields(q_cons_ts1, q1, "q_cons_ts(1)")) return rc;
    if (int rc = fields(rhs_vf, rhs, "rhs_vf")) return rc;
    if (G.prm.time_stepper > 1)
        if (int rc = fields(q_cons_ts2, q2, "q_cons_ts(2)")) return rc;
    if (int rc = bc_fill(q1)) return rc;
    if (t_step == G.prm.t_step_stop) return 1;
    if (q_prim_vf) {
        if (int rc = fields(q_prim_vf, qp, "q_prim_vf")) return rc;
        if (int rc = cons_to_prim_launch(q1, qp)) return rc;
    }
    if (G.prm.time_stepper == 1) {
        if (int rc = sweeps(q1, rhs, kRkNone, nullptr, nullptr, dt)) return rc;
        const Box x = box();
        TimedScope ts(kKRk1);
        hipLaunchKernelGGL(k_rk1_update, dim3((unsigned)((G.nint + 255) / 256)), dim3(256), 0, G.stream, q1, rhs,
                           G.S, x, dt, G.nint);
        HIPCHK(hipGetLastError());
        return 0;
    }
    if (G.prm.time_stepper == 2) {
        if (int rc = sweeps(q1, rhs, kRkEuler, &q1, &q2, dt)) return rc;
        if (int rc = bc_fill(q2)) return rc;
        return sweeps(q2, rhs, kRk2Final, &q1, &q1, dt);
    }
    if (!G.scratch_state) HIPCHK(hipMalloc(&G.scratch_state, sizeof(double) * G.S * G.ng));
    Fields q3;
    for (int v = 0; v < kMaxSys; ++v) q3.v[v] = v < G.S ? G.scratch_state + (long)v * G.ng : nullptr;
    // stage 1: q3 = q1 + dt*L(q1); stage 2: q2 = (3 q1 + q3 + dt*L(q3))/4;
    // stage 3: q1 = (q1 + 2 q2 + 2 dt*L(q2))/3.  Each stage reads its input's
    // stencil and writes a different array, or only its own cell of q1.
    if (int rc = sweeps(q1, rhs, kRkEuler, &q1, &q3, dt)) return rc;
    if (int rc = bc_fill(q3)) return rc;
    if (int rc = sweeps(q3, rhs, kRk3Stage2, &q1, &q2, dt)) return rc;
    if (int rc = bc_fill(q2)) return rc;
    return sweeps(q2, rhs, kRk3Stage3, &q1, &q1, dt);
}

int mfc_cons_to_prim(double *const *q_cons_vf, double *const *q_prim_vf) {
    if (!G.init) return fail(-4, "mfc_rhs_init has not been called");
    Fields qc, qp;
    if (int rc = fields(q_cons_vf, qc, "q_cons_vf")) return rc;
    if (int rc = fields(q_prim_vf, qp, "q_prim_vf")) return rc;
    return cons_to_prim_launch(qc, qp);
}

// The ICFL max-reduction into `partials` (capacity doubles), a grid-stride
// pass of at most kReduceBlocks blocks, so every grid size fits.  `timed`
// records HIP events around the pass (event creation may allocate, so the
// allocation-free entry point passes false).
static int icfl_reduce(double *const *q, bool from_cons, double dt, double *out, double *partials, int capacity,
                       bool timed) {
    if (int rc = require_ready()) return rc;
    if (!out) return fail(-1, "NULL output");
    Fields f;
    if (int rc = fields(q, f, from_cons ? "q_cons_vf" : "q_prim_vf")) return rc;
    const Box x = box();
    const Eos e = eos();
    double *dxs[3] = {G.dir[0].ds_dev, G.dir[1].ds_dev, G.dir[2].ds_dev};
    const int nb = (int)std::min<long>(kReduceBlocks, std::max<long>(1, (G.nint + 511) / 512));
    if (!partials || nb > capacity)
        return fail(-1, "CFL reduction needs %d partials, the buffer holds %d", nb, capacity);
    int rc;
    {
    TimedScope ts(timed ? kKIcfl : -1); // the per-cell pass (the single-block final reduction is not included)
    rc = dispatch_nf_d([&](auto NF, auto D) -> int {
        if (from_cons)
            hipLaunchKernelGGL((k_icfl<decltype(NF)::value, decltype(D)::value, true>), dim3(nb), dim3(256), 0,
                               G.stream, f, e, x, dxs[0], dxs[1], dxs[2], dt, G.nint, partials);
        else
            hipLaunchKernelGGL((k_icfl<decltype(NF)::value, decltype(D)::value, false>), dim3(nb), dim3(256), 0,
                               G.stream, f, e, x, dxs[0], dxs[1], dxs[2], dt, G.nint, partials);
        HIPCHK(hipGetLastError());
        return 0;
    });
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, G.stream, partials, nb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, partials, sizeof(double), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    return 0;
}

int mfc_icfl_max(double *const *q_prim_vf, double dt, double *icfl_max) {
    return icfl_reduce(q_prim_vf, false, dt, icfl_max, G.partials, kReduceBlocks, true);
}

int mfc_stability_criteria(double *const *q_prim_vf, double dt, double *out) {
    if (!out) return fail(-1, "NULL output");
    out[1] = 0.0;
    out[2] = 0.0;
    if (int rc = icfl_reduce(q_prim_vf, false, dt, &out[0], G.partials, kReduceBlocks, true)) return rc;
    if (!G.visc) return 0;
    Fields f;
    if (int rc = fields(q_prim_vf, f, "q_prim_vf")) return rc;
    const Box x = box();
    const Eos e = eos();
    const int nb = (int)std::min<long>(kReduceBlocks, std::max<long>(1, (G.nint + 255) / 256));
    const double *dy = G.D > 1 ? G.dir[1].ds_dev : nullptr;
    int rc = 0;
    if (G.D == 1 || G.D == 2) {
        rc = dispatch_nf_d([&](auto NF, auto D) -> int {
            if constexpr (decltype(D)::value <= 2 && decltype(NF)::value <= kViscFluidsMax) {
                hipLaunchKernelGGL((k_visc_criteria<decltype(NF)::value, decltype(D)::value>), dim3(nb), dim3(256),
                                   0, G.stream, f, e, G.vp, x, G.dir[0].ds_dev, dy, dt, G.nint, G.partials);
                HIPCHK(hipGetLastError());
                return 0;
            } else {
                return fail(-1, "viscous criteria are 1D and 2D only");
            }
        });
    }
    if (rc) return rc;
    double h[2];
    for (int k = 0; k < 2; ++k) {
        hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, G.stream, G.partials + (long)k * nb, nb);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&h[k], G.partials + (long)k * nb, sizeof(double), hipMemcpyDeviceToHost, G.stream));
    }
    HIPCHK(hipStreamSynchronize(G.stream));
    out[1] = h[0];
    out[2] = -h[1];
    return 0;
}

static int cfl_dt(double *const *q_cons_vf, double cfl, double *dt_out, double *partials, bool timed) {
    if (!(cfl > 0.0)) return fail(-1, "cfl must be positive");
    if (!dt_out) return fail(-1, "NULL output");
    double mx = 0.0;
    if (int rc = icfl_reduce(q_cons_vf, true, 1.0, &mx, partials, kReduceBlocks, timed)) return rc;
    if (!(mx > 0.0) || mx != mx) return fail(-5, "CFL reduction gave %g (NaN or non-positive wave speed)", mx);
    *dt_out = cfl / mx;
    return 0;
}

int mfc_cfl_dt(double *const *q_cons_vf, double cfl, double *dt_out) {
    return cfl_dt(q_cons_vf, cfl, dt_out, G.partials, true);
}

int mfc_cfl_dt_noalloc(double *const *q_cons_vf, double cfl, double *dt_out) {
    if (!G.init) return fail(-4, "mfc_rhs_init has not been called");
    return cfl_dt(q_cons_vf, cfl, dt_out, G.scratch, false);
}

int mfc_sizes(int *sys_size, int *buff_size, long *ghost_len, long *interior_len) {
    if (!G.init) return fail(-4, "mfc_rhs_init has not been called");
    if (sys_size) *sys_size = G.S;
    if (buff_size) *buff_size = G.b;
    if (ghost_len) *ghost_len = G.ng;
    if (interior_len) *interior_len = G.nint;
    return 0;
}

int mfc_weno_coefficients(int dir, double *out) {
    if (!G.init) return fail(-4, "mfc_rhs_init has not been called");
    if (dir < 0 || dir >= G.D || !G.dir[dir].ready) return fail(-1, "no grid for direction %d", dir);
    memcpy(out, G.dir[dir].coef_h.data(), sizeof(double) * G.dir[dir].coef_h.size());
    return 0;
}

int mfc_device_alloc(size_t bytes, void **ptr) {
    if (!ptr) return fail(-1, "NULL pointer");
    HIPCHK(hipMalloc(ptr, bytes));
    return 0;
}

int mfc_device_free(void *ptr) {
    HIPCHK(hipFree(ptr));
    return 0;
}

int mfc_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    hipStream_t st = G.init ? G.stream : nullptr;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int mfc_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    hipStream_t st = G.init ? G.stream : nullptr;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int mfc_sync(void) {
    if (!G.init) return 0;
    HIPCHK(hipStreamSynchronize(G.stream));
    return 0;
}

void *mfc_stream(void) { return G.init ? (void *)G.stream : nullptr; }

int mfc_timing_enable(int on) {
    if (!G.init) return fail(-4, "mfc_rhs_init has not been called");
    if (int rc = timing_flush()) return rc;
    T.on = on != 0;
    return 0;
}

int mfc_timing_reset(void) {
    if (G.init)
        if (int rc = timing_flush()) return rc;
    for (int k = 0; k < kKCount; ++k) {
        T.total_ms[k] = 0.0;
        T.count[k] = 0;
    }
    return 0;
}

int mfc_timing_read(int kernel, const char **name, double *total_ms, long *count) {
    if (kernel < 0 || kernel >= kKCount) return fail(-1, "kernel id %d outside 0..%d", kernel, kKCount - 1);
    if (G.init)
        if (int rc = timing_flush()) return rc;
    if (name) *name = kKernelNames[kernel];
    if (total_ms) *total_ms = T.total_ms[kernel];
    if (count) *count = T.count[kernel];
    return 0;
}

int mfc_set_stream(void *stream) {
    if (!G.init) return fail(-4, "mfc_rhs_init has not been called");
    G.stream = stream ? (hipStream_t)stream : G.own;
    return 0;
}

int mfc_rhs_finalize(void) {
    if (!G.init) return 0;
    // best-effort teardown: errors here leave nothing to recover
    (void)hipStreamSynchronize(G.stream);
    timing_flush();
    for (auto e : T.pool) (void)hipEventDestroy(e);
    T = Timing();
    for (int d = 0; d < 3; ++d) {
        if (G.dir[d].coef) (void)hipFree(G.dir[d].coef);
        if (G.dir[d].inv_ds) (void)hipFree(G.dir[d].inv_ds);
        if (G.dir[d].ds_dev) (void)hipFree(G.dir[d].ds_dev);
        if (G.dir[d].cc_dev) (void)hipFree(G.dir[d].cc_dev);
        if (G.dir[d].inv_dsg) (void)hipFree(G.dir[d].inv_dsg);
        G.dir[d] = Dir();
    }
    if (G.partials) (void)hipFree(G.partials);
    if (G.scratch_state) (void)hipFree(G.scratch_state);
    if (G.visc_fs) (void)hipFree(G.visc_fs);
    if (G.visc_grad) (void)hipFree(G.visc_grad);
    if (G.visc_fgrad) (void)hipFree(G.visc_fgrad);
    if (G.own) (void)hipStreamDestroy(G.own);
    G = State();
    return 0;
}

} // extern "C"
