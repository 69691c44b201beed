// mfc_viscous.hpp -- viscous terms: velocity gradients by finite
// differences (weno_Re_flux = F, src/simulation/m_viscous.fpp:216-355) or by
// the divergence theorem on WENO face values, reconstructed to the faces
// with WENO (weno_Re_flux = T, m_viscous.fpp:200-214, :372-591,
// m_rhs.fpp:514-528); m_riemann_solvers.fpp:169-231 (face Reynolds numbers)
// and :683-902 (the viscous source flux), m_rhs.fpp:591-603 and :639-651
// (its divergence).  1D and 2D, FP64, as in the reference.
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
    double *fs;           // face records, S+2+D planes of nface doubles (HLLC flux, contact velocity, viscous flux)
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
    // weno_Re_flux: cell-centred velocity gradients, 2*D planes of ng doubles
    // (plane e*D + c: d u_c / d x_e over the ghosted grid), and their face
    // averages in the sweep direction, 2*D planes of nface doubles (same
    // order, face t as in the face record)
    const double *grad;
    double *fgrad;
    long ng;
    int wrf;
};

// Velocity gradients along direction e (weno_Re_flux): the WENO face values
// of each velocity component along e and the divergence theorem.
struct ViscGradArgs {
    Fields q;               // ghost-filled conservative stage input
    double *grad;           // the gradient planes (ViscArgs::grad)
    const double *coef;     // WENO coefficients of direction e, 27 planes of ncell
    long ncell;
    const double *inv_dsg;  // 1/ds(i) of direction e over the ghosted range, i = -b..s+b at i+b
    double eps;
    long nxg, ng;
    int m, n, b, bj, s;     // s: last interior index along e
    int uniform;            // every cell the pass reconstructs has the coefficients ucf (bit for bit)
    double ucf[kNCoef];
};

// Direction dir of a viscous RHS evaluation: the face records, then the
// finish, which assigns (sa.accumulate = 0) or adds to sa.rhs, or applies
// the RK stage sa.rk from sa.qbase and sa.qin into sa.qout.
int launch_viscous(int nf, int D, int dir, int weno_order, const ViscArgs &a, const SweepArgs &sa,
                   hipStream_t stream);

// The gradient planes of direction e (weno_Re_flux), before the directions
// of a viscous RHS evaluation.
int launch_visc_grad(int nf, int D, int e, int weno_order, const ViscGradArgs &g, hipStream_t stream);

} // namespace mfc
