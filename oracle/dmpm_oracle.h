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
 * Layouts (host, row-major AoS): x,v are [n][3]; F,C,dFc are [n][9] with
 * element (i,j) at 3*i+j; grids are flat with index i*dy*dz + j*dz + k
 * (Grid.h:22-31) and node position min + (i,j,k)*dx (Grid.cpp:22).
 */
#ifndef DMPM_ORACLE_H
#define DMPM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Reference-compat switches (BackPropagation.cpp / PointCloud.cpp quirks). */
#define ORC_COMPAT_DUP_FBAR     1  /* (a) (I+dtC')^T dLdF' added twice, :20 and :121 */
#define ORC_COMPAT_NO_1MS       2  /* (b) dLdC_next lacks (1-s), :24 */
#define ORC_COMPAT_NO_DRAG_X    4  /* (c) x-gradient uses m v, not m(1-dt drag)v, :109 */
#define ORC_COMPAT_DLDM_NO_FEXT 8  /* (d) dLdm uses v instead of v - dt f_ext, :57 */
#define ORC_COMPAT_ALL (ORC_COMPAT_DUP_FBAR | ORC_COMPAT_NO_1MS | ORC_COMPAT_NO_DRAG_X | ORC_COMPAT_DLDM_NO_FEXT)

typedef struct orc_params {
    int n;              /* particles */
    int dim[3];         /* grid nodes per axis */
    double dx;
    double min_point[3];
    double dt, drag, smoothing;
    double f_ext[3];
    int compat;         /* ORC_COMPAT_* bit set, 0 = exact adjoint */
} orc_params;

void orc_set_threads(int nthreads);

/* Interpolation.h */
double orc_cubic_bspline(double x);
double orc_cubic_bspline_slope(double x);

/* 3x3 SVD with the JIXIE conventions; returns sweeps used. */
int orc_svd3(const double A[9], double U[9], double sigma[3], double V[9]);
void orc_polar3(const double A[9], double R[9], double S[9]);

/* Elasticity.cpp */
double orc_fc_energy(const double F[9], double lam, double mu);
void orc_pk1(const double F[9], double lam, double mu, double P[9]);
/* exact = 1: derivative of the clamped PK1; exact = 0: the reference's
 * formula structure (J clamp ignored in d(J F^-T)). Identical when det F > 1e-6. */
void orc_hess_mult(const double F[9], double lam, double mu, const double dF[9], double out[9], int exact);
void orc_lame(double E, double nu, double* lam, double* mu);

/* Batched forms for tests. */
void orc_pk1_batch(int n, const double* F, const double* lam, const double* mu, double* P);
void orc_hess_mult_batch(int n, const double* F, const double* lam, const double* mu,
                         const double* dF, double* out, int exact);

/* ForwardSimulation.cpp:11-51 with grid reset; grid_p is [G][3] momentum. */
void orc_p2g(const orc_params* prm, const double* x, const double* v, const double* C,
             const double* F, const double* dFc, const double* m, const double* vol,
             const double* lam, const double* mu, double dt, double drag,
             double* grid_m, double* grid_p);

/* One ForwardTimeStep (smoothing overload, ForwardSimulation.cpp:136-157).
 * grid_m / grid_v receive the updated grid (needed by the adjoint). */
void orc_forward_step(const orc_params* prm,
                      const double* x, const double* v, const double* C, const double* F,
                      const double* dFc, const double* m, const double* vol,
                      const double* lam, const double* mu,
                      double* grid_m, double* grid_v,
                      double* x1, double* v1, double* C1, double* F1);

/* MultiThreadMPM::ForwardTimeStep (MultiThreadForwardSimulation.cpp:7-225):
 * the proxy-grid forward with `sections` particle runs, no smoothing.
 * Returns 0, or -1 for sections < 1, -2 when the proxy grids cannot be
 * allocated. */
int orc_forward_step_proxy(const orc_params* prm, int sections,
                           const double* x, const double* v, const double* C, const double* F,
                           const double* dFc, const double* m, const double* vol,
                           const double* lam, const double* mu,
                           double* grid_m, double* grid_v,
                           double* x1, double* v1, double* C1, double* F1);

/* ForwardSimulation.cpp:262-290 (vol = m / sum_i w m_i, no dx^3). */
void orc_point_volumes(const orc_params* prm, const double* x, const double* m, double* vol);

/* P2G mass only (target-grid construction, MPMPolyscope.cpp:158-160). */
void orc_p2g_mass(const orc_params* prm, const double* x, const double* m, double* grid_m);

/* CompGraph.cpp:110-244. Writes grid_m, grid_dLdm and dLdx; returns the loss. */
double orc_end_mass_loss(const orc_params* prm, const double* x, const double* m,
                         const double* target_m, double* grid_m, double* grid_dLdm,
                         double* dLdx);

/* One Back_Timestep (BackPropagation.cpp:6-125).
 * Inputs: layer n state and constants, the updated grid of step n,
 * layer n+1's C1 and gradients (gx1, gv1, gF1, gC1).
 * Outputs: layer n gradients (gx, gv, gF, gC) and dL/d(dFc_n) in gdFc.
 * grid_vbar ([G][3]) and grid_mbar ([G]) are work arrays (overwritten). */
void orc_backward_step(const orc_params* prm,
                       const double* x, const double* v, const double* C, const double* F,
                       const double* dFc, const double* m, const double* vol,
                       const double* lam, const double* mu,
                       const double* grid_m, const double* grid_v,
                       const double* C1, const double* gx1, const double* gv1,
                       const double* gF1, const double* gC1,
                       double* grid_vbar, double* grid_mbar,
                       double* gx, double* gv, double* gF, double* gC, double* gdFc);

#ifdef __cplusplus
}
#endif
#endif
