// mfc_viscous.hpp -- viscous terms with finite-difference velocity gradients
// (weno_Re_flux = F): src/simulation/m_viscous.fpp:159-370,
// m_riemann_solvers.fpp:169-231 (face Reynolds numbers) and :683-902 (the
// viscous source flux), m_rhs.fpp:591-603 and :639-651 (its divergence).
// 1D and 2D, FP64, as in the reference.
#pragma once

#include "mfc_sweep.hpp"

namespace mfc {

constexpr int kViscFluidsMax = 4; // fluids the kernels are built for

// Re_size, Re_idx, Res (m_global_parameters.fpp:314-338, m_viscous.fpp:36-42).
struct ViscParams {
    int size[2];
    int idx[2][kViscFluidsMax]; // 0-based fluid index
    double res[2][kViscFluidsMax];
};

struct ViscArgs {
    Fields q;             // ghost-filled conservative stage input
    Fields rhs;           // interior rhs the divergence is added to
    double *fs;           // face fluxes, (D+1) planes of nface doubles
    long nface;
    const double *coef;   // WENO coefficients of the sweep direction, 27 planes of ncell
    long ncell;
    int uniform;          // every reconstructed cell has the coefficients ucf (bit for bit)
    double ucf[kNCoef];
    const double *ccx, *ccy; // ghosted cell centres x_cc(-b:m+b), y_cc(-b:n+b)
    const double *inv_ds; // 1/ds(i), i = 0..s, of the sweep direction
    Eos eos;
    double eps;
    long nxg;
    int m, n, b, bj;
    int bc_beg, bc_end;   // of the sweep direction
    ViscParams vp;
};

// Face fluxes (k_visc_face) then their divergence (k_visc_add) for direction
// dir, after that direction's inviscid sweep has written or added to rhs.
int launch_viscous(int nf, int D, int dir, int weno_order, const ViscArgs &a, hipStream_t stream);

// Pointwise TVD-RK stage a.rk from a.qbase, a.qin and a.rhs into a.qout over
// the interior (the stage update the inviscid path fuses into its last sweep).
int launch_rk_stage(int nf, int D, const SweepArgs &a, long nint, hipStream_t stream);

} // namespace mfc
