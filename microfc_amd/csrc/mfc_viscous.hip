// mfc_viscous.hip -- the RHS and RK stage of viscous runs: inviscid and
// viscous fluxes, with velocity gradients by finite differences or
// (weno_Re_flux) by the divergence theorem on WENO face values.
//
// The viscous flux needs, per face, the HLLC star velocity and the
// reconstructed volume fractions (m_riemann_solvers.fpp:169-231, :313-323),
// which the inviscid sweeps keep in registers and LDS only.  Viscous runs
// therefore use these kernels in place of the sweeps.  The face kernels
// reconstruct with the sweeps' own weno_cell and hllc_face (hence the same
// bits), each cell converted to primitive once into an LDS tile and
// reconstructed once (both faces), one variable at a time.  The
// finite-difference gradients come from the primitive velocities (of the
// tile, and for the x faces' neighbours in y of the conservative field);
// with weno_Re_flux, gradient planes and their face averages are computed
// by two passes before the face kernels.  Per face they store the HLLC flux,
// the contact velocity and the viscous source flux
// (m_riemann_solvers.fpp:683-902); a finish kernel per direction applies the
// flux divergence, the alpha*div(u) source and the viscous divergence in the
// reference's order (m_rhs.fpp:565-652), and in the last direction the RK
// stage.
#include "mfc_viscous.hpp"

namespace mfc {

namespace {

__device__ __forceinline__ long goff(const ViscArgs &a, int i, int j) {
    return (long)(i + a.b) + a.nxg * (long)(j + a.bj);
}

// Velocity component d of cell (i, j) as cons->prim computes it
// (m_variables_conversion.fpp:349-357): rho = max(sum alpha_i rho_i, sgm_eps).
template <int NF, int D>
__device__ __forceinline__ double cell_vel(const ViscArgs &a, int i, int j, int d) {
    const long o = goff(a, i, j);
    double rho = 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) rho = rho + a.q.v[f][o];
    rho = max_ref(rho, 1e-16);
    return a.q.v[Sys<NF, D>::MOM + d][o] / rho;
}

// Harmonic face Reynolds numbers (m_riemann_solvers.fpp:169-231).
template <int NF, int D>
__device__ __forceinline__ void re_avg(const ViscArgs &a, const double *ql, const double *qr, double *ra) {
    using Y = Sys<NF, D>;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        ra[r] = 0.0;
        if (a.vp.size[r] == 0) continue;
        double reL = 0.0, reR = 0.0;
        for (int q = 0; q < a.vp.size[r]; ++q) {
            reL = ql[Y::ADV + a.vp.idx[r][q]] / a.vp.res[r][q] + reL;
            reR = qr[Y::ADV + a.vp.idx[r][q]] / a.vp.res[r][q] + reR;
        }
        reL = 1.0 / max_ref(reL, 1e-16);
        reR = 1.0 / max_ref(reR, 1e-16);
        ra[r] = 2.0 / (1.0 / reL + 1.0 / reR);
    }
}

// WENO coefficients of cell c of the direction: kernel arguments (scalar
// registers) on uniform grids, as in the sweeps, else per-cell loads; the
// same values either way.
template <bool UNI>
__device__ __forceinline__ void cell_coef(const ViscArgs &a, int c, double *cf) {
#pragma unroll
    for (int q = 0; q < kNCoef; ++q) cf[q] = UNI ? a.ucf[q] : a.coef[(long)q * a.ncell + (c + a.b)];
}

// Viscous source flux of one face (m_riemann_solvers.fpp:683-902) from its
// two face states, star velocity and the face averages of the velocity
// gradients: ax[c] = dvel_avg_dx(c), ay[c] = dvel_avg_dy(c), each
// 5e-1*(solver-left + solver-right value) as the reference forms it (the
// solver's left gradients are the caller's dqR_prim_d*_n, m_rhs.fpp:545-556).
template <int NF, int D, int DIR>
__device__ __forceinline__ void face_flux(const ViscArgs &a, const double *ql, const double *qr, const double *ax,
                                          const double *ay, double *flux, double &ustar, double *fs) {
    double vsrc[D];
    hllc_face<NF, D, DIR, double, true>(ql, qr, a.eos, flux, ustar, vsrc);
    double ra[2];
    re_avg<NF, D>(a, ql, qr, ra);
#pragma unroll
    for (int v = 0; v <= D; ++v) fs[v] = 0.0;
    if constexpr (DIR == 0) {
        if (a.vp.size[0] > 0) {
            const double t11 = (4.0 / 3.0) * ax[0] / ra[0];
            fs[0] = fs[0] - t11;
            fs[D] = fs[D] - vsrc[0] * t11;
        }
        if (a.vp.size[1] > 0) {
            const double t11 = ax[0] / ra[1];
            fs[0] = fs[0] - t11;
            fs[D] = fs[D] - vsrc[0] * t11;
        }
        if constexpr (D == 2) {
            if (a.vp.size[0] > 0) {
                double tau[2];
                tau[0] = -((2.0 / 3.0) * ay[1] / ra[0]);
                tau[1] = (ay[0] + ax[1]) / ra[0];
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    fs[d] = fs[d] - tau[d];
                    fs[D] = fs[D] - vsrc[d] * tau[d];
                }
            }
            if (a.vp.size[1] > 0) {
                const double t11 = ay[1] / ra[1];
                fs[0] = fs[0] - t11;
                fs[D] = fs[D] - vsrc[0] * t11;
            }
        }
    } else {
        if (a.vp.size[0] > 0) {
            double tau[2];
            tau[0] = (ay[0] + ax[1]) / ra[0];
            tau[1] = (4.0 * ay[1] - 2.0 * ax[0]) / (3.0 * ra[0]);
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                fs[d] = fs[d] - tau[d];
                fs[D] = fs[D] - vsrc[d] * tau[d];
            }
        }
        if (a.vp.size[1] > 0) {
            const double t22 = (ax[0] + ay[1]) / ra[1];
            fs[1] = fs[1] - t22;
            fs[D] = fs[D] - vsrc[1] * t22;
        }
    }
}

// Face averages of finite-difference gradients: both sides of a face hold
// the same difference (dqR_prim_d*_n of the left cell and dqL_prim_d*_n of
// the right one, m_viscous.fpp:231-336), so each average is 5e-1*(g + g).
// gn: differences across the face, gt: the transverse four-point averages.
template <int D, int DIR>
__device__ __forceinline__ void fd_averages(const double *gn, const double *gt, double *ax, double *ay) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const double n2 = 5e-1 * (gn[c] + gn[c]), t2 = 5e-1 * (gt[c] + gt[c]);
        ax[c] = DIR == 0 ? n2 : t2;
        ay[c] = DIR == 0 ? t2 : n2;
    }
}

// The face averages precomputed by k_visc_face_grad (weno_Re_flux).
template <int D>
__device__ __forceinline__ void weno_averages(const ViscArgs &a, long t, double *ax, double *ay) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        ax[c] = c < D ? a.fgrad[(long)c * a.nface + t] : 0.0;
        ay[c] = c < D && D == 2 ? a.fgrad[(long)(D + c) * a.nface + t] : 0.0;
    }
}

// The face record: the HLLC flux (S), the contact velocity and the viscous
// source flux (D+1), plane p of face t at fs[p * nface + t].
template <int NF, int D>
__device__ __forceinline__ void store_face(const ViscArgs &a, long t, const double *flux, double ustar,
                                           const double *fs) {
    constexpr int S = Sys<NF, D>::S;
#pragma unroll
    for (int v = 0; v < S; ++v) a.fs[(long)v * a.nface + t] = flux[v];
    a.fs[(long)S * a.nface + t] = ustar;
#pragma unroll
    for (int v = 0; v <= D; ++v) a.fs[(long)(S + 1 + v) * a.nface + t] = fs[v];
}

// Gradient planes of direction E (weno_Re_flux), one thread per ghosted
// cell: each velocity component reconstructed along E with WENO over cells
// [-b+polyn, s+b-polyn] (s_reconstruct_cell_boundary_values_visc,
// m_viscous.fpp:400-469), then d u_c / d x_E = 1/ds * (vR - vL)
// (s_apply_scalar_divergence_theorem, :372-398) over [-b+1, s+b-1].  Cells
// where the reference reads face values it never wrote, or never writes the
// gradient, get NaN: their reconstructions reach no face flux the RHS uses.
template <int NF, int D, int E, int WO, bool UNI>
__global__ __launch_bounds__(256) void k_visc_grad(ViscGradArgs g) {
    using Y = Sys<NF, D>;
    constexpr int P = (WO - 1) / 2;
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= g.ng) return;
    const int i = (int)(o % g.nxg) - g.b;
    const int j = (int)(o / g.nxg) - g.bj;
    const int je = E == 0 ? i : j;
    const bool set = je >= -g.b + P && je <= g.s + g.b - P && je >= -g.b + 1 && je <= g.s + g.b - 1;
    const long step = E == 0 ? 1 : g.nxg;
    double cf[kNCoef];
    if (set) {
#pragma unroll
        for (int q = 0; q < kNCoef; ++q) cf[q] = UNI ? g.ucf[q] : g.coef[(long)q * g.ncell + (je + g.b)];
    }
    double st[D][WO];
    if (set) {
#pragma unroll
        for (int r = 0; r < WO; ++r) {
            const long oc = o + (r - P) * step;
            double rho = 0.0;
#pragma unroll
            for (int f = 0; f < NF; ++f) rho = rho + g.q.v[f][oc];
            rho = max_ref(rho, 1e-16);
#pragma unroll
            for (int c = 0; c < D; ++c) st[c][r] = g.q.v[Y::MOM + c][oc] / rho;
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
        double val = __builtin_nan("");
        if (set) {
            double vl, vr;
            weno_cell<WO, double, true, true>(st[c], cf, g.eps, vl, vr);
            val = g.inv_dsg[je + g.b] * (vr - vl);
        }
        g.grad[(long)(E * D + c) * g.ng + o] = val;
    }
}

// Face averages of the gradients for the faces of direction DIR
// (weno_Re_flux), one thread per face: each gradient plane reconstructed
// along DIR (s_reconstruct_cell_boundary_values_visc_deriv,
// m_viscous.fpp:471-591) at the upper face of the left cell (dqR_prim_d*_n)
// and the lower face of the right cell (dqL_prim_d*_n), the bc = -4 copies
// of s_populate_riemann_states_variables_buffers (m_riemann_solvers.fpp:
// 488-608, with the caller's L/R swap), then 5e-1*(R + L) as
// s_compute_viscous_source_flux averages them (:683-902).
template <int NF, int D, int DIR, int WO, bool UNI>
__global__ __launch_bounds__(256) void k_visc_face_grad(ViscArgs a) {
    constexpr int P = (WO - 1) / 2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nface) return;
    int x, y, f, s;
    if constexpr (DIR == 0) {
        f = (int)(t % (a.m + 2)) - 1;
        y = (int)(t / (a.m + 2));
        x = f;
        s = a.m;
    } else {
        x = (int)(t % (a.m + 1));
        f = (int)(t / (a.m + 1)) - 1;
        y = f;
        s = a.n;
    }
    const long step = DIR == 0 ? 1 : a.nxg;
    const long o0 = (long)(x + a.b) + a.nxg * (long)(y + a.bj); // cell f of the face's pair
    double cf0[kNCoef], cf1[kNCoef];
    cell_coef<UNI>(a, f, cf0);
    cell_coef<UNI>(a, f + 1, cf1);
#pragma unroll
    for (int pl = 0; pl < D * D; ++pl) {
        const double *G = a.grad + (long)pl * a.ng;
        double st[WO], dummy, R, L;
#pragma unroll
        for (int r = 0; r < WO; ++r) st[r] = G[o0 + (r - P) * step];
        weno_cell<WO, double, false, true>(st, cf0, a.eps, dummy, R);
#pragma unroll
        for (int r = 0; r < WO; ++r) st[r] = G[o0 + (r + 1 - P) * step];
        weno_cell<WO, double, true, false>(st, cf1, a.eps, L, dummy);
        if (f == -1 && a.bc_beg == -4) R = L;
        if (f == s && a.bc_end == -4) L = R;
        a.fgrad[(long)pl * a.nface + t] = 5e-1 * (R + L);
    }
}

// x faces: one wave per 63 faces f0 .. f0+62 of row y (f0 = -1 + 63*tile).
// Lane l converts stencil cells into LDS once, then reconstructs both faces
// of cell c = f0 + l (the smoothness indicators shared, as in the
// cell-parallel sweeps); face f = c takes its right state, the left-face
// value of cell c+1, from the next lane.  No register cap: the kernel takes
// 230 VGPRs (2 waves per SIMD) for two fluids in 2D; capped at 4 or 3 waves
// it spills and is slower.
template <int NF, int D, int WO, bool UNI, bool WRF>
__global__ __launch_bounds__(64) void k_visc_face_x(ViscArgs a, int ntile) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    constexpr int P = (WO - 1) / 2;
    constexpr int NS = 65 + 2 * P; // staged cells f0-P .. f0+64+P
    __shared__ double sp[S][NS];
    const int lane = threadIdx.x;
    const int tile = (int)(blockIdx.x % ntile);
    const int y = (int)(blockIdx.x / ntile);
    const int f0 = -1 + 63 * tile;
    const int hi = a.m + a.b; // last ghost cell
    for (int k = lane; k < NS; k += 64) {
        const int c = min(f0 - P + k, hi);
        const long o = goff(a, c, y);
        double cv[S], pv[S];
#pragma unroll
        for (int v = 0; v < S; ++v) cv[v] = a.q.v[v][o];
        cons_to_prim<NF, D>(cv, pv, a.eos);
#pragma unroll
        for (int v = 0; v < S; ++v) sp[v][k] = pv[v];
    }
    __syncthreads();
    const int c = min(f0 + lane, hi - P); // this lane's cell (clamped past the last face)
    double cf[kNCoef];
    cell_coef<UNI>(a, c, cf);
    // one variable at a time (bounded registers); the face states go to LDS
    // rows only this lane reads back: its upper-face value, and the next
    // lane's lower-face value
    __shared__ double sq[2][S][64];
#pragma unroll 1
    for (int v = 0; v < S; ++v) {
        double st[WO], vl, vr;
#pragma unroll
        for (int r = 0; r < WO; ++r) st[r] = sp[v][lane + r];
        weno_cell<WO, double, true, true>(st, cf, a.eps, vl, vr);
        sq[0][v][lane] = vr;
        sq[1][v][lane] = __shfl_down(vl, 1, 64);
    }
    double ql[S], qr[S];
#pragma unroll
    for (int v = 0; v < S; ++v) {
        ql[v] = sq[0][v][lane];
        qr[v] = sq[1][v][lane];
    }
    // velocities of rows y-1 and y+1 of this lane's cell, then of the next
    // lane's (2D transverse gradients)
    double gt[2] = {0.0, 0.0};
    if constexpr (D == 2 && !WRF) {
        double um[2], up[2], um1[2], up1[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            um[d] = cell_vel<NF, D>(a, c, y - 1, d);
            up[d] = cell_vel<NF, D>(a, c, y + 1, d);
        }
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            um1[d] = __shfl_down(um[d], 1, 64);
            up1[d] = __shfl_down(up[d], 1, 64);
        }
        const double dym = a.ccy[y + a.b] - a.ccy[y - 1 + a.b], dyp = a.ccy[y + 1 + a.b] - a.ccy[y + a.b];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double u0 = sp[Y::MOM + d][lane + P], u1 = sp[Y::MOM + d][lane + P + 1];
            // four-point average of the one-sided y differences of cells
            // x+1 and x (dqR_prim_dy_n(1) at x, m_viscous.fpp:318-336)
            const double sum = (u1 - um1[d]) / dym + (up1[d] - u1) / dyp + (u0 - um[d]) / dym + (up[d] - u0) / dyp;
            gt[d] = 25e-2 * sum;
        }
    }
    const int f = f0 + lane;
    if (lane == 63 || f > a.m) return;
    if (f == -1 && a.bc_beg == -4) {
#pragma unroll
        for (int v = 0; v < S; ++v) ql[v] = qr[v];
    }
    if (f == a.m && a.bc_end == -4) {
#pragma unroll
        for (int v = 0; v < S; ++v) qr[v] = ql[v];
    }
    const long t = (long)y * (a.m + 2) + (f + 1);
    double ax[2], ay[2];
    if constexpr (WRF) {
        weno_averages<D>(a, t, ax, ay);
    } else {
        // one-sided x differences across the face (dqR_prim_dx_n(1) at x)
        double gxn[2] = {0.0, 0.0};
        const double dxf = a.ccx[f + 1 + a.b] - a.ccx[f + a.b];
#pragma unroll
        for (int d = 0; d < D; ++d) gxn[d] = (sp[Y::MOM + d][lane + P + 1] - sp[Y::MOM + d][lane + P]) / dxf;
        fd_averages<D, 0>(gxn, gt, ax, ay);
    }
    double flux[S], ustar, fs[D + 1];
    face_flux<NF, D, 0>(a, ql, qr, ax, ay, flux, ustar, fs);
    store_face<NF, D>(a, t, flux, ustar, fs);
}

// y faces (2D), cell-parallel on an LDS tile: a 256-thread block covers
// kYC x-columns by kYK cells along y, the K lanes of a column consecutive
// in the wave.  The primitive stencil rows of the tile (with one x-halo
// column on each side, for the transverse velocity differences) are
// converted once into LDS; thread (column, c) reconstructs both faces of
// cell f0+c one variable at a time (the WENO of the sweeps), its upper-face
// value and the next lane's lower-face value go to LDS as the states of
// face f0+c+1/2, and after a barrier the first K-1 lanes of each column
// solve their faces (HLLC, face Reynolds numbers, gradients, viscous flux)
// and store the face records.  Each cell is converted and reconstructed
// once, with the per-variable reconstruction bounded to one variable's
// registers (a thread marching along y with the stencil window in registers
// held 256 VGPRs and 72 AGPRs: one wave per SIMD).
constexpr int kYC = 16; // x-columns per tile
constexpr int kYK = 16; // cells per column (faces: kYK - 1)

template <int NF, int WO, bool UNI, bool WRF>
__global__ __launch_bounds__(kYC *kYK) void k_visc_face_yc(ViscArgs a, int nxb) {
    constexpr int D = 2;
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    constexpr int P = (WO - 1) / 2;
    constexpr int R = kYK + 2 * P; // stencil rows f0-P .. f0+K-1+P
    constexpr int XC = kYC + 2;     // columns x0-1 .. x0+C
    // stencil rows of each column contiguous (odd pitch R + 1): the lanes of
    // a column read consecutive entries
    __shared__ double sp[S][XC][R + 1];
    // face f0+c+1/2 of column col: left state (upper face of cell f0+c) and
    // right state (lower face of cell f0+c+1), c fastest (the lanes of a
    // column are consecutive: conflict-free LDS access)
    __shared__ double sL[S][kYC][kYK];
    __shared__ double sR[S][kYC][kYK];
    const int tid = threadIdx.x;
    const int col = tid / kYK, c = tid % kYK;
    const int x0 = (int)(blockIdx.x % nxb) * kYC;
    const int f0 = -1 + (kYK - 1) * (int)(blockIdx.x / nxb);
    const int hx = a.m + a.b, hy = a.n + a.b; // last ghost cells
    for (int e = tid; e < R * XC; e += kYC * kYK) {
        const int r = e / XC, xc = e % XC;
        const long o = goff(a, min(x0 - 1 + xc, hx), min(f0 - P + r, hy));
        double cv[S], pv[S];
#pragma unroll
        for (int v = 0; v < S; ++v) cv[v] = a.q.v[v][o];
        cons_to_prim<NF, D>(cv, pv, a.eos);
#pragma unroll
        for (int v = 0; v < S; ++v) sp[v][xc][r] = pv[v];
    }
    __syncthreads();
    const int x = x0 + col;
    const int cy = f0 + c; // this lane's cell
    double cf[kNCoef];
    cell_coef<UNI>(a, min(cy, hy - P), cf);
#pragma unroll 1
    for (int v = 0; v < S; ++v) {
        double st[WO], vl, vr;
#pragma unroll
        for (int r = 0; r < WO; ++r) st[r] = sp[v][col + 1][c + r];
        weno_cell<WO, double, true, true>(st, cf, a.eps, vl, vr);
        sL[v][col][c] = vr;
        // lane c+1 of the same column holds cell f0+c+1 (the last lane's value
        // is not used)
        sR[v][col][c] = __shfl_down(vl, 1, 64);
    }
    __syncthreads();
    const int f = cy;
    if (c == kYK - 1 || x > a.m || f > a.n) return;
    double ql[S], qr[S];
#pragma unroll
    for (int v = 0; v < S; ++v) {
        ql[v] = sL[v][col][c];
        qr[v] = sR[v][col][c];
    }
    if (f == -1 && a.bc_beg == -4) {
#pragma unroll
        for (int v = 0; v < S; ++v) ql[v] = qr[v];
    }
    if (f == a.n && a.bc_end == -4) {
#pragma unroll
        for (int v = 0; v < S; ++v) qr[v] = ql[v];
    }
    const long t = (long)(f + 1) * (a.m + 1) + x;
    double ax[2], ay[2];
    if constexpr (WRF) {
        weno_averages<D>(a, t, ax, ay);
    } else {
        // velocities of cells (x-1 .. x+1, f) and (x-1 .. x+1, f+1)
        const int r0 = c + P, r1 = c + P + 1;
        const double dyf = a.ccy[f + 1 + a.b] - a.ccy[f + a.b];
        const double dxm = a.ccx[x + a.b] - a.ccx[x - 1 + a.b], dxp = a.ccx[x + 1 + a.b] - a.ccx[x + a.b];
        double gxn[2], gt[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double uo = sp[Y::MOM + d][col + 1][r0], uw = sp[Y::MOM + d][col][r0], ue = sp[Y::MOM + d][col + 2][r0];
            const double no = sp[Y::MOM + d][col + 1][r1], nw = sp[Y::MOM + d][col][r1], ne = sp[Y::MOM + d][col + 2][r1];
            gxn[d] = (no - uo) / dyf;
            // four-point average of the one-sided x differences of rows f+1
            // and f (dqR_prim_dx_n(2) at f, m_viscous.fpp:285-315)
            const double sum = (no - nw) / dxm + (ne - no) / dxp + (uo - uw) / dxm + (ue - uo) / dxp;
            gt[d] = 25e-2 * sum;
        }
        fd_averages<D, 1>(gxn, gt, ax, ay);
    }
    double flux[S], ustar, fs[D + 1];
    face_flux<NF, D, 1>(a, ql, qr, ax, ay, flux, ustar, fs);
    store_face<NF, D>(a, t, flux, ustar, fs);
}

// Finish of one direction for every interior cell, in the reference's
// order (m_rhs.fpp:565-652): the flux divergence (assigned in the first
// direction, added after), the alpha*div(u) source, the viscous divergence
// of momentum and energy; then the rhs store, or the RK stage in the last
// direction (the sweeps' finish_cell and rk_combine arithmetic).
template <int NF, int D, int DIR>
__global__ __launch_bounds__(256) void k_visc_finish(ViscArgs a, SweepArgs s, long nint) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nint) return;
    const int i = (int)(t % (a.m + 1));
    const int j = (int)(t / (a.m + 1));
    long fl, fr;
    double inv;
    if constexpr (DIR == 0) {
        fl = (long)j * (a.m + 2) + i; // face i-1/2 is stored at (i-1)+1
        fr = fl + 1;
        inv = a.inv_ds[i];
    } else {
        fl = (long)j * (a.m + 1) + i;
        fr = fl + (a.m + 1);
        inv = a.inv_ds[j];
    }
    const long go = ghost_off(s, i, j, 0);
    auto F = [&](int p, long f) { return a.fs[(long)p * a.nface + f]; };
    double r[S];
#pragma unroll
    for (int v = 0; v < S; ++v) {
        const double df = F(v, fl) - F(v, fr);
        r[v] = s.accumulate ? s.rhs.v[v][t] + inv * df : inv * df;
    }
    const double dus = F(S, fr) - F(S, fl);
#pragma unroll
    for (int k = 0; k < NF; ++k) r[Y::ADV + k] = r[Y::ADV + k] + inv * s.qin.v[Y::ADV + k][go] * dus;
#pragma unroll
    for (int v = 0; v <= D; ++v) r[Y::MOM + v] = r[Y::MOM + v] + inv * (F(S + 1 + v, fl) - F(S + 1 + v, fr));
    if (s.rk == kRkNone) {
#pragma unroll
        for (int v = 0; v < S; ++v) s.rhs.v[v][t] = r[v];
        return;
    }
    FinishIn<S> fin;
    double o[S];
#pragma unroll
    for (int v = 0; v < S; ++v) {
        fin.base[v] = s.qbase.v[v][go];
        fin.in[v] = s.qin.v[v][go];
    }
    rk_combine<S>(s, fin, r, o);
#pragma unroll
    for (int v = 0; v < S; ++v) s.qout.v[v][go] = o[v];
}

template <int NF, int D, int DIR, int WO, bool UNI, bool WRF>
void launch_face_u(const ViscArgs &a, hipStream_t st) {
    if constexpr (DIR == 0) {
        const int ntile = (a.m + 2 + 62) / 63;
        hipLaunchKernelGGL((k_visc_face_x<NF, D, WO, UNI, WRF>), dim3((unsigned)((long)ntile * (a.n + 1))), dim3(64),
                           0, st, a, ntile);
    } else {
        const int nxb = (a.m + 1 + kYC - 1) / kYC;
        const int ntile = (a.n + 2 + kYK - 2) / (kYK - 1); // faces -1 .. n in tiles of kYK - 1
        hipLaunchKernelGGL((k_visc_face_yc<NF, WO, UNI, WRF>), dim3((unsigned)((long)nxb * ntile)),
                           dim3(kYC * kYK), 0, st, a, nxb);
    }
}

template <int NF, int D, int DIR, int WO>
void launch_face(const ViscArgs &a, hipStream_t st) {
    if (a.wrf) {
        const dim3 grid((unsigned)((a.nface + 255) / 256));
        if (a.uniform) {
            hipLaunchKernelGGL((k_visc_face_grad<NF, D, DIR, WO, true>), grid, dim3(256), 0, st, a);
            launch_face_u<NF, D, DIR, WO, true, true>(a, st);
        } else {
            hipLaunchKernelGGL((k_visc_face_grad<NF, D, DIR, WO, false>), grid, dim3(256), 0, st, a);
            launch_face_u<NF, D, DIR, WO, false, true>(a, st);
        }
    } else {
        if (a.uniform) launch_face_u<NF, D, DIR, WO, true, false>(a, st);
        else launch_face_u<NF, D, DIR, WO, false, false>(a, st);
    }
}

template <int NF, int D, int DIR>
int launch_dir(int weno_order, const ViscArgs &a, const SweepArgs &sa, hipStream_t st) {
    switch (weno_order) {
    case 1: launch_face<NF, D, DIR, 1>(a, st); break;
    case 3: launch_face<NF, D, DIR, 3>(a, st); break;
    case 5: launch_face<NF, D, DIR, 5>(a, st); break;
    default: return -1;
    }
    if (hipGetLastError() != hipSuccess) return -1;
    const long nint = (long)(a.m + 1) * (a.n + 1);
    hipLaunchKernelGGL((k_visc_finish<NF, D, DIR>), dim3((unsigned)((nint + 255) / 256)), dim3(256), 0, st, a, sa,
                       nint);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NF>
int launch_nf(int D, int dir, int weno_order, const ViscArgs &a, const SweepArgs &sa, hipStream_t st) {
    if (D == 1 && dir == 0) return launch_dir<NF, 1, 0>(weno_order, a, sa, st);
    if (D == 2 && dir == 0) return launch_dir<NF, 2, 0>(weno_order, a, sa, st);
    if (D == 2 && dir == 1) return launch_dir<NF, 2, 1>(weno_order, a, sa, st);
    return -1;
}

template <int NF, int D, int E>
int launch_grad_e(int weno_order, const ViscGradArgs &g, hipStream_t st) {
    const dim3 grid((unsigned)((g.ng + 255) / 256));
    const bool u = g.uniform != 0;
    switch (weno_order) {
    case 1: hipLaunchKernelGGL((k_visc_grad<NF, D, E, 1, false>), grid, dim3(256), 0, st, g); break;
    case 3:
        if (u) hipLaunchKernelGGL((k_visc_grad<NF, D, E, 3, true>), grid, dim3(256), 0, st, g);
        else hipLaunchKernelGGL((k_visc_grad<NF, D, E, 3, false>), grid, dim3(256), 0, st, g);
        break;
    case 5:
        if (u) hipLaunchKernelGGL((k_visc_grad<NF, D, E, 5, true>), grid, dim3(256), 0, st, g);
        else hipLaunchKernelGGL((k_visc_grad<NF, D, E, 5, false>), grid, dim3(256), 0, st, g);
        break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NF>
int launch_grad_nf(int D, int e, int weno_order, const ViscGradArgs &g, hipStream_t st) {
    if (D == 1 && e == 0) return launch_grad_e<NF, 1, 0>(weno_order, g, st);
    if (D == 2 && e == 0) return launch_grad_e<NF, 2, 0>(weno_order, g, st);
    if (D == 2 && e == 1) return launch_grad_e<NF, 2, 1>(weno_order, g, st);
    return -1;
}

} // namespace

int launch_visc_grad(int nf, int D, int e, int weno_order, const ViscGradArgs &g, hipStream_t st) {
    switch (nf) {
    case 1: return launch_grad_nf<1>(D, e, weno_order, g, st);
    case 2: return launch_grad_nf<2>(D, e, weno_order, g, st);
    case 3: return launch_grad_nf<3>(D, e, weno_order, g, st);
    case 4: return launch_grad_nf<4>(D, e, weno_order, g, st);
    default: return -1;
    }
}

int launch_viscous(int nf, int D, int dir, int weno_order, const ViscArgs &a, const SweepArgs &sa,
                   hipStream_t st) {
    switch (nf) {
    case 1: return launch_nf<1>(D, dir, weno_order, a, sa, st);
    case 2: return launch_nf<2>(D, dir, weno_order, a, sa, st);
    case 3: return launch_nf<3>(D, dir, weno_order, a, sa, st);
    case 4: return launch_nf<4>(D, dir, weno_order, a, sa, st);
    default: return -1;
    }
}

} // namespace mfc
