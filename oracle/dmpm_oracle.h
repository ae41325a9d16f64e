/*
 * CPU FP64 oracle for the differentiable MLS-MPM step (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it.  It restates,
 * in plain C and double precision, the reference algorithm of DiffMPMLib3D:
 *
 *   Interpolation.h:5-30            cubic B-spline N(x), N'(x)
 *   ForwardSimulation.cpp:6-157     P_op_1, P2G, G_op, G2P, P_op_2
 *   ForwardSimulation.cpp:262-290   CalculatePointCloudVolumes
 *   Elasticity.cpp:74-178,231-235   fixed-corotated energy, PK1, H:dF, Lame
 *   BackPropagation.cpp:6-125       Back_Timestep (exact adjoint by default,
 *                                   reference defects (a)-(d) behind flags)
 *   CompGraph.cpp:110-244           EndLayerMassLoss
 *   qrsvd/ImplicitQRSVD.h:687-875   SVD/polar conventions (sorted by
 *                                   magnitude, sign on the smallest sigma,
 *                                   U and V rotations, R = U V^T)
 *
 * orc_real is double.  The Makefile also builds liboracle_dmpm_f32.so from
 * the same source with ORC_REAL=float and single-precision literals, the
 * reference's own precision (DiffMPMLib3D/pch.h:24-25): it is only timed, as
 * the FP32 CPU baseline, and never used as a checker.
 *
 * Layouts (host, row-major AoS): x,v are [n][3]; F,C,dFc are [n][9] with
 * element (i,j) at 3*i+j; grids are flat with index i*dy*dz + j*dz + k
 * (Grid.h:22-31) and node position min + (i,j,k)*dx (Grid.cpp:22).
 */
#ifndef DMPM_ORACLE_H
#define DMPM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL orc_real;

/* Reference-compat switches (BackPropagation.cpp / PointCloud.cpp quirks). */
#define ORC_COMPAT_DUP_FBAR     1  /* (a) (I+dtC')^T dLdF' added twice, :20 and :121 */
#define ORC_COMPAT_NO_1MS       2  /* (b) dLdC_next lacks (1-s), :24 */
#define ORC_COMPAT_NO_DRAG_X    4  /* (c) x-gradient uses m v, not m(1-dt drag)v, :109 */
#define ORC_COMPAT_DLDM_NO_FEXT 8  /* (d) dLdm uses v instead of v - dt f_ext, :57 */
#define ORC_COMPAT_ALL (ORC_COMPAT_DUP_FBAR | ORC_COMPAT_NO_1MS | ORC_COMPAT_NO_DRAG_X | ORC_COMPAT_DLDM_NO_FEXT)

typedef struct orc_params {
    int n;              /* particles */
    int dim[3];         /* grid nodes per axis */
    orc_real dx;
    orc_real min_point[3];
    orc_real dt, drag, smoothing;
    orc_real f_ext[3];
    int compat;         /* ORC_COMPAT_* bit set, 0 = exact adjoint */
} orc_params;

void orc_set_threads(int nthreads);

/* Interpolation.h */
orc_real orc_cubic_bspline(orc_real x);
orc_real orc_cubic_bspline_slope(orc_real x);

/* 3x3 SVD with the JIXIE conventions; returns sweeps used. */
int orc_svd3(const orc_real A[9], orc_real U[9], orc_real sigma[3], orc_real V[9]);
void orc_polar3(const orc_real A[9], orc_real R[9], orc_real S[9]);

/* Elasticity.cpp */
orc_real orc_fc_energy(const orc_real F[9], orc_real lam, orc_real mu);
void orc_pk1(const orc_real F[9], orc_real lam, orc_real mu, orc_real P[9]);
/* exact = 1: derivative of the clamped PK1; exact = 0: the reference's
 * formula structure (J clamp ignored in d(J F^-T)). Identical when det F > 1e-6. */
void orc_hess_mult(const orc_real F[9], orc_real lam, orc_real mu, const orc_real dF[9], orc_real out[9], int exact);
void orc_lame(orc_real E, orc_real nu, orc_real* lam, orc_real* mu);

/* Batched forms for tests. */
void orc_pk1_batch(int n, const orc_real* F, const orc_real* lam, const orc_real* mu, orc_real* P);
void orc_hess_mult_batch(int n, const orc_real* F, const orc_real* lam, const orc_real* mu,
                         const orc_real* dF, orc_real* out, int exact);

/* ForwardSimulation.cpp:11-51 with grid reset; grid_p is [G][3] momentum. */
void orc_p2g(const orc_params* prm, const orc_real* x, const orc_real* v, const orc_real* C,
             const orc_real* F, const orc_real* dFc, const orc_real* m, const orc_real* vol,
             const orc_real* lam, const orc_real* mu, orc_real dt, orc_real drag,
             orc_real* grid_m, orc_real* grid_p);

/* One ForwardTimeStep (smoothing overload, ForwardSimulation.cpp:136-157).
 * grid_m / grid_v receive the updated grid (needed by the adjoint). */
void orc_forward_step(const orc_params* prm,
                      const orc_real* x, const orc_real* v, const orc_real* C, const orc_real* F,
                      const orc_real* dFc, const orc_real* m, const orc_real* vol,
                      const orc_real* lam, const orc_real* mu,
                      orc_real* grid_m, orc_real* grid_v,
                      orc_real* x1, orc_real* v1, orc_real* C1, orc_real* F1);

/* MultiThreadMPM::ForwardTimeStep (MultiThreadForwardSimulation.cpp:7-225):
 * the proxy-grid forward with `sections` particle runs, no smoothing.
 * Returns 0, or -1 for sections < 1, -2 when the proxy grids cannot be
 * allocated. */
int orc_forward_step_proxy(const orc_params* prm, int sections,
                           const orc_real* x, const orc_real* v, const orc_real* C, const orc_real* F,
                           const orc_real* dFc, const orc_real* m, const orc_real* vol,
                           const orc_real* lam, const orc_real* mu,
                           orc_real* grid_m, orc_real* grid_v,
                           orc_real* x1, orc_real* v1, orc_real* C1, orc_real* F1);

/* ForwardSimulation.cpp:262-290 (vol = m / sum_i w m_i, no dx^3). */
void orc_point_volumes(const orc_params* prm, const orc_real* x, const orc_real* m, orc_real* vol);

/* P2G mass only (target-grid construction, MPMPolyscope.cpp:158-160). */
void orc_p2g_mass(const orc_params* prm, const orc_real* x, const orc_real* m, orc_real* grid_m);

/* CompGraph.cpp:110-244. Writes grid_m, grid_dLdm and dLdx; returns the loss. */
orc_real orc_end_mass_loss(const orc_params* prm, const orc_real* x, const orc_real* m,
                         const orc_real* target_m, orc_real* grid_m, orc_real* grid_dLdm,
                         orc_real* dLdx);

/* One Back_Timestep (BackPropagation.cpp:6-125).
 * Inputs: layer n state and constants, the updated grid of step n,
 * layer n+1's C1 and gradients (gx1, gv1, gF1, gC1).
 * Outputs: layer n gradients (gx, gv, gF, gC) and dL/d(dFc_n) in gdFc.
 * grid_vbar ([G][3]) and grid_mbar ([G]) are work arrays (overwritten). */
void orc_backward_step(const orc_params* prm,
                       const orc_real* x, const orc_real* v, const orc_real* C, const orc_real* F,
                       const orc_real* dFc, const orc_real* m, const orc_real* vol,
                       const orc_real* lam, const orc_real* mu,
                       const orc_real* grid_m, const orc_real* grid_v,
                       const orc_real* C1, const orc_real* gx1, const orc_real* gv1,
                       const orc_real* gF1, const orc_real* gC1,
                       orc_real* grid_vbar, orc_real* grid_mbar,
                       orc_real* gx, orc_real* gv, orc_real* gF, orc_real* gC, orc_real* gdFc);

#ifdef __cplusplus
}
#endif
#endif
