// mfc_viscous.hip -- viscous terms of the RHS with finite-difference velocity
// gradients, and the pointwise TVD-RK stage the viscous path applies after
// them.
//
// The inviscid sweeps keep face states in registers and LDS only, so the
// viscous flux, which needs the HLLC star velocity and the reconstructed
// volume fractions of every face (m_riemann_solvers.fpp:169-231, :313-323),
// recomputes them here: one thread per face reconstructs the two face states
// of its stencil (the same weno_cell and hllc_face as the sweeps, hence the
// same bits), takes the velocity gradients from the conservative field, and
// writes the momentum and energy source flux of the face
// (m_riemann_solvers.fpp:683-902).  A second kernel adds its divergence to
// rhs after the direction's inviscid terms, in the reference's order
// (m_rhs.fpp:591-603, :639-651).  No viscous benchmark case exists, so these
// kernels favour plainness over speed.
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

// One-sided x difference across the face between cells x and x+1
// (dqR_prim_dx_n(1) at x, dqL_prim_dx_n(1) at x+1, m_viscous.fpp:234-255).
template <int NF, int D>
__device__ __forceinline__ double dx1(const ViscArgs &a, int x, int y, int d) {
    return (cell_vel<NF, D>(a, x + 1, y, d) - cell_vel<NF, D>(a, x, y, d)) / (a.ccx[x + 1 + a.b] - a.ccx[x + a.b]);
}
// One-sided y difference across the face between cells y and y+1
// (dqR_prim_dy_n(2) at y, dqL_prim_dy_n(2) at y+1, :259-283).
template <int NF, int D>
__device__ __forceinline__ double dy2(const ViscArgs &a, int x, int y, int d) {
    return (cell_vel<NF, D>(a, x, y + 1, d) - cell_vel<NF, D>(a, x, y, d)) / (a.ccy[y + 1 + a.b] - a.ccy[y + a.b]);
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

// Viscous source flux of one face.  DIR 0: faces f = -1..m of row y, left
// cell (f, y); DIR 1: faces f = -1..n of column x, left cell (x, f).
template <int NF, int D, int DIR, int WO>
__global__ __launch_bounds__(256) void k_visc_face(ViscArgs a) {
    using Y = Sys<NF, D>;
    constexpr int S = Y::S;
    constexpr int P = (WO - 1) / 2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nface) return;
    int f, x, y, s;
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
    // face states: right face of cell f and left face of cell f+1
    double pr[2 * P + 2][S];
#pragma unroll
    for (int r = 0; r < 2 * P + 2; ++r) {
        const int c = f - P + r;
        const long o = DIR == 0 ? goff(a, c, y) : goff(a, x, c);
        double cv[S];
#pragma unroll
        for (int v = 0; v < S; ++v) cv[v] = a.q.v[v][o];
        cons_to_prim<NF, D>(cv, pr[r], a.eos);
    }
    double cf0[kNCoef], cf1[kNCoef];
#pragma unroll
    for (int q = 0; q < kNCoef; ++q) {
        cf0[q] = a.coef[(long)q * a.ncell + (f + a.b)];
        cf1[q] = a.coef[(long)q * a.ncell + (f + 1 + a.b)];
    }
    double ql[S], qr[S];
#pragma unroll
    for (int v = 0; v < S; ++v) {
        double st0[WO], st1[WO], dummy;
#pragma unroll
        for (int r = 0; r < WO; ++r) {
            st0[r] = pr[r][v];
            st1[r] = pr[r + 1][v];
        }
        weno_cell<WO, double, false, true>(st0, cf0, a.eps, dummy, ql[v]);
        weno_cell<WO, double, true, false>(st1, cf1, a.eps, qr[v], dummy);
    }
    // Riemann-state extrapolation (bc = -4, m_riemann_solvers.fpp:480-550)
    if (f == -1 && a.bc_beg == -4) {
#pragma unroll
        for (int v = 0; v < S; ++v) ql[v] = qr[v];
    }
    if (f == s && a.bc_end == -4) {
#pragma unroll
        for (int v = 0; v < S; ++v) qr[v] = ql[v];
    }
    double flux[S], ustar, vsrc[D];
    hllc_face<NF, D, DIR, double, true>(ql, qr, a.eos, flux, ustar, vsrc);
    double ra[2];
    re_avg<NF, D>(a, ql, qr, ra);

    // The reference passes dqR_prim_d*_n as the solver's left gradients and
    // dqL_prim_d*_n as its right ones (m_rhs.fpp:545-556); both sides of a
    // face evaluate the same difference, so each average below is of two
    // equal terms, kept as written (m_riemann_solvers.fpp:719-720 etc.).
    double fs[D + 1];
#pragma unroll
    for (int v = 0; v <= D; ++v) fs[v] = 0.0;
    if constexpr (DIR == 0) {
        const double gxu = dx1<NF, D>(a, x, y, 0);
        if (a.vp.size[0] > 0) {
            const double dvx = 5e-1 * (gxu + gxu);
            const double t11 = (4.0 / 3.0) * dvx / ra[0];
            fs[0] = fs[0] - t11;
            fs[D] = fs[D] - vsrc[0] * t11;
        }
        if (a.vp.size[1] > 0) {
            const double dvx = 5e-1 * (gxu + gxu);
            const double t11 = dvx / ra[1];
            fs[0] = fs[0] - t11;
            fs[D] = fs[D] - vsrc[0] * t11;
        }
        if constexpr (D == 2) {
            // y derivatives at the x face: four-point averages of the
            // one-sided y differences of cells x and x+1
            // (dqR_prim_dy_n(1) at x, dqL_prim_dy_n(1) at x+1, m_viscous.fpp:318-336)
            double gy[2];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const double sum = dy2<NF, D>(a, x + 1, y - 1, d) + dy2<NF, D>(a, x + 1, y, d) +
                                   dy2<NF, D>(a, x, y - 1, d) + dy2<NF, D>(a, x, y, d);
                gy[d] = 25e-2 * sum;
            }
            if (a.vp.size[0] > 0) {
                double dvy[2], tau[2];
#pragma unroll
                for (int d = 0; d < 2; ++d) dvy[d] = 5e-1 * (gy[d] + gy[d]);
                const double gxv = dx1<NF, D>(a, x, y, 1);
                const double dvx2 = 5e-1 * (gxv + gxv);
                tau[0] = -((2.0 / 3.0) * dvy[1] / ra[0]);
                tau[1] = (dvy[0] + dvx2) / ra[0];
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    fs[d] = fs[d] - tau[d];
                    fs[D] = fs[D] - vsrc[d] * tau[d];
                }
            }
            if (a.vp.size[1] > 0) {
                const double dvy2 = 5e-1 * (gy[1] + gy[1]);
                const double t11 = dvy2 / ra[1];
                fs[0] = fs[0] - t11;
                fs[D] = fs[D] - vsrc[0] * t11;
            }
        }
    } else {
        // x derivatives at the y face: four-point averages of the one-sided
        // x differences of rows y and y+1 (dqR_prim_dx_n(2) at y,
        // dqL_prim_dx_n(2) at y+1, m_viscous.fpp:285-315)
        double gx[2], gyv[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double sum = dx1<NF, D>(a, x - 1, y + 1, d) + dx1<NF, D>(a, x, y + 1, d) +
                               dx1<NF, D>(a, x - 1, y, d) + dx1<NF, D>(a, x, y, d);
            gx[d] = 25e-2 * sum;
            gyv[d] = dy2<NF, D>(a, x, y, d);
        }
        if (a.vp.size[0] > 0) {
            double dvx[2], dvy[2], tau[2];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                dvx[d] = 5e-1 * (gx[d] + gx[d]);
                dvy[d] = 5e-1 * (gyv[d] + gyv[d]);
            }
            tau[0] = (dvy[0] + dvx[1]) / ra[0];
            tau[1] = (4.0 * dvy[1] - 2.0 * dvx[0]) / (3.0 * ra[0]);
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                fs[d] = fs[d] - tau[d];
                fs[D] = fs[D] - vsrc[d] * tau[d];
            }
        }
        if (a.vp.size[1] > 0) {
            const double dvx1 = 5e-1 * (gx[0] + gx[0]);
            const double dvy2 = 5e-1 * (gyv[1] + gyv[1]);
            const double t22 = (dvx1 + dvy2) / ra[1];
            fs[1] = fs[1] - t22;
            fs[D] = fs[D] - vsrc[1] * t22;
        }
    }
#pragma unroll
    for (int v = 0; v <= D; ++v) a.fs[(long)v * a.nface + t] = fs[v];
}

// rhs(mom..E) += (1/ds)(fs(c-1) - fs(c)) over the interior
// (m_rhs.fpp:591-603 for x, :639-651 for y).
template <int NF, int D, int DIR>
__global__ __launch_bounds__(256) void k_visc_add(ViscArgs a, long nint) {
    using Y = Sys<NF, D>;
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
#pragma unroll
    for (int v = 0; v <= D; ++v) {
        double *r = a.rhs.v[Y::MOM + v];
        r[t] = r[t] + inv * (a.fs[(long)v * a.nface + fl] - a.fs[(long)v * a.nface + fr]);
    }
}

template <int NF, int D>
__global__ __launch_bounds__(256) void k_rk_stage(SweepArgs a, long nint) {
    constexpr int S = Sys<NF, D>::S;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nint) return;
    const int i = (int)(t % (a.m + 1));
    const int j = (int)(t / (a.m + 1));
    const long go = ghost_off(a, i, j, 0);
    FinishIn<S> f;
    double r[S], o[S];
#pragma unroll
    for (int v = 0; v < S; ++v) {
        f.base[v] = a.qbase.v[v][go];
        f.in[v] = a.qin.v[v][go];
        r[v] = a.rhs.v[v][t];
    }
    rk_combine<S>(a, f, r, o);
#pragma unroll
    for (int v = 0; v < S; ++v) a.qout.v[v][go] = o[v];
}

template <int NF, int D, int DIR>
int launch_dir(int weno_order, const ViscArgs &a, hipStream_t st) {
    const dim3 blk(256), gf((unsigned)((a.nface + 255) / 256));
    switch (weno_order) {
    case 1: hipLaunchKernelGGL((k_visc_face<NF, D, DIR, 1>), gf, blk, 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_visc_face<NF, D, DIR, 3>), gf, blk, 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_visc_face<NF, D, DIR, 5>), gf, blk, 0, st, a); break;
    default: return -1;
    }
    if (hipGetLastError() != hipSuccess) return -1;
    const long nint = (long)(a.m + 1) * (a.n + 1);
    hipLaunchKernelGGL((k_visc_add<NF, D, DIR>), dim3((unsigned)((nint + 255) / 256)), blk, 0, st, a, nint);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NF>
int launch_nf(int D, int dir, int weno_order, const ViscArgs &a, hipStream_t st) {
    if (D == 1 && dir == 0) return launch_dir<NF, 1, 0>(weno_order, a, st);
    if (D == 2 && dir == 0) return launch_dir<NF, 2, 0>(weno_order, a, st);
    if (D == 2 && dir == 1) return launch_dir<NF, 2, 1>(weno_order, a, st);
    return -1;
}

template <int NF>
int rk_nf(int D, const SweepArgs &a, long nint, hipStream_t st) {
    const dim3 blk(256), g((unsigned)((nint + 255) / 256));
    if (D == 1) hipLaunchKernelGGL((k_rk_stage<NF, 1>), g, blk, 0, st, a, nint);
    else if (D == 2) hipLaunchKernelGGL((k_rk_stage<NF, 2>), g, blk, 0, st, a, nint);
    else return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

} // namespace

int launch_viscous(int nf, int D, int dir, int weno_order, const ViscArgs &a, hipStream_t st) {
    switch (nf) {
    case 1: return launch_nf<1>(D, dir, weno_order, a, st);
    case 2: return launch_nf<2>(D, dir, weno_order, a, st);
    case 3: return launch_nf<3>(D, dir, weno_order, a, st);
    case 4: return launch_nf<4>(D, dir, weno_order, a, st);
    default: return -1;
    }
}

int launch_rk_stage(int nf, int D, const SweepArgs &a, long nint, hipStream_t st) {
    switch (nf) {
    case 1: return rk_nf<1>(D, a, nint, st);
    case 2: return rk_nf<2>(D, a, nint, st);
    case 3: return rk_nf<3>(D, a, nint, st);
    case 4: return rk_nf<4>(D, a, nint, st);
    default: return -1;
    }
}

} // namespace mfc
