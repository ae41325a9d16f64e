/*
 * CPU FP64 oracle for the differentiable MLS-MPM step.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker that tests/, the smoke entry
 * point and bench.py's CPU-baseline leg compare the HIP path against.  The
 * product never links, loads or calls it.
 *
 * Each function names the reference lines it restates (paths relative to the
 * reference repository root).  The arithmetic is double precision throughout;
 * the reference itself is float (DiffMPMLib3D/pch.h:24-25).
 *
 * Scatter phases (P2G and the transposed G2P) are parallel over particles
 * with OpenMP atomics, like ForwardSimulation.cpp:45-46; every other phase is
 * a per-particle or per-node gather.  With orc_set_threads(1) the oracle is
 * deterministic.  orc_forward_step_proxy restates the proxy-grid forward of
 * MultiThreadForwardSimulation.cpp instead: one grid per particle section,
 * reduced in section order, deterministic for a fixed section count.
 */
#include "dmpm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* std::numeric_limits<float>::epsilon(), the node-mass threshold of
 * ForwardSimulation.cpp:55 and BackPropagation.cpp:55. */
#define ORC_MASS_EPS 1.1920928955078125e-07
/* CompGraph.cpp:115,122-123 */
#define ORC_LOSS_EPS 1e-4
#define ORC_MIN_MASS 1e-3
#define ORC_PENALTY 10.0
/* Elasticity.cpp:95 */
#define ORC_J_MIN 1e-6

void orc_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* ------------------------------------------------------------------ */
/* Interpolation.h:5-30                                                 */
/* ------------------------------------------------------------------ */
double orc_cubic_bspline(double x)
{
    double a = fabs(x);
    if (a < 1.0) return 0.5 * a * a * a - a * a + 2.0 / 3.0;
    if (a < 2.0) return (2.0 - a) * (2.0 - a) * (2.0 - a) / 6.0;
    return 0.0;
}

double orc_cubic_bspline_slope(double x)
{
    double a = fabs(x);
    if (a < 1.0) return 1.5 * x * a - 2.0 * x;
    if (a < 2.0) return -x * a / 2.0 + 2.0 * x - 2.0 * x / a;
    return 0.0;
}

/* ------------------------------------------------------------------ */
/* 3x3 helpers (row-major, (i,j) at 3*i+j)                              */
/* ------------------------------------------------------------------ */
static void m3_mul(const double* A, const double* B, double* out)
{
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
    memcpy(out, r, sizeof r);
}
/* A^T B */
static void m3_tmul(const double* A, const double* B, double* out)
{
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
    memcpy(out, r, sizeof r);
}
/* A B^T */
static void m3_mult(const double* A, const double* B, double* out)
{
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
    memcpy(out, r, sizeof r);
}
static void m3_transpose(const double* A, double* out)
{
    double r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = A[3 * j + i];
    memcpy(out, r, sizeof r);
}
static double m3_det(const double* A)
{
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}
/* inverse via the adjugate */
static void m3_inverse(const double* A, double* out)
{
    double d = m3_det(A);
    double r[9];
    r[0] = (A[4] * A[8] - A[5] * A[7]) / d;
    r[1] = (A[2] * A[7] - A[1] * A[8]) / d;
    r[2] = (A[1] * A[5] - A[2] * A[4]) / d;
    r[3] = (A[5] * A[6] - A[3] * A[8]) / d;
    r[4] = (A[0] * A[8] - A[2] * A[6]) / d;
    r[5] = (A[2] * A[3] - A[0] * A[5]) / d;
    r[6] = (A[3] * A[7] - A[4] * A[6]) / d;
    r[7] = (A[1] * A[6] - A[0] * A[7]) / d;
    r[8] = (A[0] * A[4] - A[1] * A[3]) / d;
    memcpy(out, r, sizeof r);
}
static double m3_inner(const double* A, const double* B)
{
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += A[i] * B[i];
    return s;
}

/* ------------------------------------------------------------------ */
/* SVD / polar: one-sided (Hestenes) Jacobi, then the conventions of    */
/* qrsvd/ImplicitQRSVD.h:680-686 and :853-875.                          */
/* ------------------------------------------------------------------ */
static void swap_col(double* M, int a, int b)
{
    for (int k = 0; k < 3; ++k) {
        double t = M[3 * k + a];
        M[3 * k + a] = M[3 * k + b];
        M[3 * k + b] = t;
    }
}

int orc_svd3(const double A[9], double U[9], double s[3], double V[9])
{
    double B[9];
    memcpy(B, A, sizeof B);
    for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    int sweep;
    for (sweep = 0; sweep < 64; ++sweep) {
        int rotated = 0;
        for (int r = 0; r < 3; ++r) {
            int p = PQ[r][0], q = PQ[r][1];
            double al = 0, be = 0, ga = 0;
            for (int k = 0; k < 3; ++k) {
                al += B[3 * k + p] * B[3 * k + p];
                be += B[3 * k + q] * B[3 * k + q];
                ga += B[3 * k + p] * B[3 * k + q];
            }
            if (ga == 0.0 || fabs(ga) <= DBL_EPSILON * sqrt(al * be)) continue;
            rotated = 1;
            double zeta = (be - al) / (2.0 * ga);
            double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
            for (int k = 0; k < 3; ++k) {
                double bp = B[3 * k + p], bq = B[3 * k + q];
                B[3 * k + p] = c * bp - sn * bq;
                B[3 * k + q] = sn * bp + c * bq;
                double vp = V[3 * k + p], vq = V[3 * k + q];
                V[3 * k + p] = c * vp - sn * vq;
                V[3 * k + q] = sn * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    for (int j = 0; j < 3; ++j)
        s[j] = sqrt(B[j] * B[j] + B[3 + j] * B[3 + j] + B[6 + j] * B[6 + j]);
    /* sort by decreasing singular value */
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2 - a; ++b)
            if (s[b] < s[b + 1]) {
                double t = s[b];
                s[b] = s[b + 1];
                s[b + 1] = t;
                swap_col(B, b, b + 1);
                swap_col(V, b, b + 1);
            }
    double smax = s[0] > 0 ? s[0] : 1.0;
    double tiny = 64.0 * DBL_EPSILON * smax;
    /* left vectors; complete a basis where singular values vanish */
    if (s[0] <= tiny) {
        for (int i = 0; i < 9; ++i) U[i] = (i % 4 == 0) ? 1.0 : 0.0;
    } else {
        for (int k = 0; k < 3; ++k) U[3 * k] = B[3 * k] / s[0];
        if (s[1] <= tiny) {
            /* any unit vector orthogonal to u0 */
            double u0[3] = {U[0], U[3], U[6]};
            int m = (fabs(u0[0]) < fabs(u0[1])) ? ((fabs(u0[0]) < fabs(u0[2])) ? 0 : 2)
                                                : ((fabs(u0[1]) < fabs(u0[2])) ? 1 : 2);
            double e[3] = {0, 0, 0};
            e[m] = 1.0;
            double d = e[0] * u0[0] + e[1] * u0[1] + e[2] * u0[2];
            double w[3] = {e[0] - d * u0[0], e[1] - d * u0[1], e[2] - d * u0[2]};
            double nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            for (int k = 0; k < 3; ++k) U[3 * k + 1] = w[k] / nw;
        } else {
            for (int k = 0; k < 3; ++k) U[3 * k + 1] = B[3 * k + 1] / s[1];
        }
        if (s[2] <= tiny) {
            U[2] = U[3] * U[7] - U[6] * U[4];
            U[5] = U[6] * U[1] - U[0] * U[7];
            U[8] = U[0] * U[4] - U[3] * U[1];
        } else {
            for (int k = 0; k < 3; ++k) U[3 * k + 2] = B[3 * k + 2] / s[2];
        }
    }
    /* V a rotation: flipping column 2 of both U and V keeps A = U S V^T */
    if (m3_det(V) < 0.0) {
        for (int k = 0; k < 3; ++k) {
            V[3 * k + 2] = -V[3 * k + 2];
            U[3 * k + 2] = -U[3 * k + 2];
        }
    }
    /* U a rotation: the sign moves onto the smallest singular value */
    if (m3_det(U) < 0.0) {
        for (int k = 0; k < 3; ++k) U[3 * k + 2] = -U[3 * k + 2];
        s[2] = -s[2];
    }
    return sweep;
}

void orc_polar3(const double A[9], double R[9], double S[9])
{
    double U[9], s[3], V[9];
    orc_svd3(A, U, s, V);
    m3_mult(U, V, R); /* U V^T */
    double VS[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) VS[3 * i + j] = V[3 * i + j] * s[j];
    m3_mult(VS, V, S); /* V S V^T */
}

/* ------------------------------------------------------------------ */
/* Elasticity.cpp                                                        */
/* ------------------------------------------------------------------ */
/* Elasticity.cpp:74-86 */
double orc_fc_energy(const double F[9], double lam, double mu)
{
    double U[9], s[3], V[9];
    orc_svd3(F, U, s, V);
    double sum = (s[0] - 1) * (s[0] - 1) + (s[1] - 1) * (s[1] - 1) + (s[2] - 1) * (s[2] - 1);
    double J = m3_det(F);
    if (J < ORC_J_MIN) J = ORC_J_MIN;
    return mu * sum + 0.5 * lam * (J - 1) * (J - 1);
}

/* Elasticity.cpp:88-98: P = 2 mu (F - R) + lam (J-1) J F^-T, J clamped. */
void orc_pk1(const double F[9], double lam, double mu, double P[9])
{
    double R[9], S[9], Fi[9];
    orc_polar3(F, R, S);
    double J = m3_det(F);
    if (J < ORC_J_MIN) J = ORC_J_MIN;
    m3_inverse(F, Fi);
    double k = lam * (J - 1.0) * J;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            P[3 * i + j] = 2.0 * mu * (F[3 * i + j] - R[3 * i + j]) + k * Fi[3 * j + i];
}

/* Elasticity.cpp:100-178.  The reference computes d(J F^-T)[dF] by float
 * central differences (:7-72,119); this is its closed form
 * J[(F^-T:dF) F^-T - F^-T dF^T F^-T].  dR solves the same 3x3 system
 * (:144-173), written in the singular basis: with M = U^T dF V,
 * dR = U w V^T, w_ij = (M_ij - M_ji)/(s_i + s_j); a vanishing s_i + s_j
 * takes the pseudo-inverse (w_ij = 0). */
void orc_hess_mult(const double F[9], double lam, double mu, const double dF[9], double out[9],
                   int exact)
{
    double U[9], s[3], V[9];
    orc_svd3(F, U, s, V);
    double Ju = m3_det(F);
    double J = Ju < ORC_J_MIN ? ORC_J_MIN : Ju;
    double Fi[9], Fit[9];
    m3_inverse(F, Fi);
    m3_transpose(Fi, Fit);
    for (int i = 0; i < 9; ++i) out[i] = 2.0 * mu * dF[i];
    /* Fit dF^T Fit */
    double t1[9], FdF[9];
    m3_mult(Fit, dF, t1); /* Fit dF^T */
    m3_mul(t1, Fit, FdF);
    if (exact && Ju <= ORC_J_MIN) {
        double k = lam * (J - 1.0) * J;
        for (int i = 0; i < 9; ++i) out[i] -= k * FdF[i];
    } else {
        double inner = J * m3_inner(Fit, dF);
        for (int i = 0; i < 9; ++i) out[i] += lam * J * Fit[i] * inner;
        double fd = m3_inner(Fit, dF);
        double k = lam * (J - 1.0) * Ju;
        for (int i = 0; i < 9; ++i) out[i] += k * (fd * Fit[i] - FdF[i]);
    }
    double M[9], tmp[9];
    m3_tmul(U, dF, tmp);
    m3_mul(tmp, V, M);
    double w[9] = {0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            if (i == j) continue;
            double den = s[i] + s[j];
            double scale = fabs(s[0]) > 1.0 ? fabs(s[0]) : 1.0;
            w[3 * i + j] = (fabs(den) > 1e-12 * scale) ? (M[3 * i + j] - M[3 * j + i]) / den : 0.0;
        }
    double dR[9];
    m3_mul(U, w, tmp);
    m3_mult(tmp, V, dR);
    for (int i = 0; i < 9; ++i) out[i] -= 2.0 * mu * dR[i];
}

/* Elasticity.cpp:231-235 */
void orc_lame(double E, double nu, double* lam, double* mu)
{
    *lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    *mu = E / (2.0 + 2.0 * nu);
}

void orc_pk1_batch(int n, const double* F, const double* lam, const double* mu, double* P)
{
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p) orc_pk1(F + 9 * p, lam[p], mu[p], P + 9 * p);
}

void orc_hess_mult_batch(int n, const double* F, const double* lam, const double* mu,
                         const double* dF, double* out, int exact)
{
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p)
        orc_hess_mult(F + 9 * p, lam[p], mu[p], dF + 9 * p, out + 9 * p, exact);
}

/* ------------------------------------------------------------------ */
/* Stencil (ForwardSimulation.cpp:15-37, Grid.cpp:40-81)                */
/* ------------------------------------------------------------------ */
typedef struct {
    int base[3];
    double N[3][4];  /* N(u) per axis and offset */
    double dN[3][4]; /* N'(u) */
    double d[3][4];  /* x_i - x_p per axis and offset */
} orc_stencil;

static void stencil(const orc_params* prm, const double* xp, orc_stencil* st, int with_slope)
{
    for (int a = 0; a < 3; ++a) {
        double rel = xp[a] - prm->min_point[a];
        /* clamped to +-2^20 cells (inf and NaN included) before the int
         * conversion: every node of such a particle is outside the grid, and
         * base + 3 cannot overflow */
        st->base[a] = (int)fmin(fmax(floor(rel / prm->dx), -1048576.0), 1048576.0) - 1;
        for (int o = 0; o < 4; ++o) {
            double xi = prm->min_point[a] + (double)(st->base[a] + o) * prm->dx;
            double d = xi - xp[a];
            double u = d / prm->dx;
            st->d[a][o] = d;
            st->N[a][o] = orc_cubic_bspline(u);
            st->dN[a][o] = with_slope ? orc_cubic_bspline_slope(u) : 0.0;
        }
    }
}

static long node_index(const orc_params* prm, const orc_stencil* st, int i, int j, int k)
{
    int a = st->base[0] + i, b = st->base[1] + j, c = st->base[2] + k;
    if (a < 0 || a >= prm->dim[0] || b < 0 || b >= prm->dim[1] || c < 0 || c >= prm->dim[2]) return -1;
    return ((long)a * prm->dim[1] + b) * prm->dim[2] + c;
}

static long grid_size(const orc_params* prm)
{
    return (long)prm->dim[0] * prm->dim[1] * prm->dim[2];
}

/* ------------------------------------------------------------------ */
/* P2G: ForwardSimulation.cpp:11-51 (after G_Reset, Grid.cpp:349-361)   */
/* ------------------------------------------------------------------ */
/* One particle's scatter (SingleParticle_to_grid, ForwardSimulation.cpp:
 * 11-51), into a shared grid with OpenMP atomics (shared = 1) or into a grid
 * only this thread writes (shared = 0: the proxy grids of
 * MultiThreadForwardSimulation.cpp:171-181). */
static void p2g_particle(const orc_params* prm, int p, const double* x, const double* v, const double* C,
                         const double* F, const double* dFc, const double* m, const double* vol,
                         const double* lam, const double* mu, double dt, double drag,
                         double* grid_m, double* grid_p, int shared)
{
    double Dinv = 3.0 / (prm->dx * prm->dx);
    double Fc[9], P[9], PF[9], A[9];
    for (int i = 0; i < 9; ++i) Fc[i] = F[9 * p + i] + dFc[9 * p + i];
    orc_pk1(Fc, lam[p], mu[p], P);
    m3_mult(P, Fc, PF); /* P Fc^T */
    for (int i = 0; i < 9; ++i) A[i] = -Dinv * dt * vol[p] * PF[i] + m[p] * C[9 * p + i];
    double mv[3];
    for (int a = 0; a < 3; ++a) mv[a] = m[p] * v[3 * p + a] * (1.0 - dt * drag);
    orc_stencil st;
    stencil(prm, x + 3 * p, &st, 0);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                long id = node_index(prm, &st, i, j, k);
                if (id < 0) continue;
                double w = st.N[0][i] * st.N[1][j] * st.N[2][k];
                double d[3] = {st.d[0][i], st.d[1][j], st.d[2][k]};
                double dm = w * m[p];
                if (shared) {
#pragma omp atomic
                    grid_m[id] += dm;
                } else {
                    grid_m[id] += dm;
                }
                for (int a = 0; a < 3; ++a) {
                    double dp = w * (mv[a] + A[3 * a] * d[0] + A[3 * a + 1] * d[1] + A[3 * a + 2] * d[2]);
                    if (shared) {
#pragma omp atomic
                        grid_p[3 * id + a] += dp;
                    } else {
                        grid_p[3 * id + a] += dp;
                    }
                }
            }
}

void orc_p2g(const orc_params* prm, const double* x, const double* v, const double* C,
             const double* F, const double* dFc, const double* m, const double* vol,
             const double* lam, const double* mu, double dt, double drag,
             double* grid_m, double* grid_p)
{
    long G = grid_size(prm);
    memset(grid_m, 0, sizeof(double) * G);
    memset(grid_p, 0, sizeof(double) * 3 * G);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p)
        p2g_particle(prm, p, x, v, C, F, dFc, m, vol, lam, mu, dt, drag, grid_m, grid_p, 1);
}

void orc_p2g_mass(const orc_params* prm, const double* x, const double* m, double* grid_m)
{
    long G = grid_size(prm);
    memset(grid_m, 0, sizeof(double) * G);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 0);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    double dm = st.N[0][i] * st.N[1][j] * st.N[2][k] * m[p];
#pragma omp atomic
                    grid_m[id] += dm;
                }
    }
}

/* ------------------------------------------------------------------ */
/* ForwardTimeStep: ForwardSimulation.cpp:136-157                       */
/* ------------------------------------------------------------------ */
/* G_op on one node: ForwardSimulation.cpp:53-58,184-200 (v stays 0 from
 * the reset); grid_v holds the momentum on entry. */
static void grid_op_node(const orc_params* prm, long i, const double* grid_m, double* grid_v)
{
    if (fabs(grid_m[i]) > ORC_MASS_EPS) {
        for (int a = 0; a < 3; ++a) grid_v[3 * i + a] = grid_v[3 * i + a] / grid_m[i] + prm->dt * prm->f_ext[a];
    } else {
        for (int a = 0; a < 3; ++a) grid_v[3 * i + a] = 0.0;
    }
}

/* G2P + P_op_2: ForwardSimulation.cpp:60-106,108-124 (s = 0 is the
 * overload without smoothing, :108-112) */
static void g2p_update(const orc_params* prm, double s, const double* x, const double* F, const double* dFc,
                       const double* grid_v, double* x1, double* v1, double* C1, double* F1)
{
    double Dinv = 3.0 / (prm->dx * prm->dx);
    double dt = prm->dt;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 0);
        double vp[3] = {0, 0, 0}, Cp[9] = {0};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    double w = st.N[0][i] * st.N[1][j] * st.N[2][k];
                    double d[3] = {st.d[0][i], st.d[1][j], st.d[2][k]};
                    for (int a = 0; a < 3; ++a) {
                        double vi = grid_v[3 * id + a];
                        vp[a] += w * vi;
                        for (int b = 0; b < 3; ++b) Cp[3 * a + b] += Dinv * w * vi * d[b];
                    }
                }
        double Fc[9], IC[9], Fn[9];
        for (int i = 0; i < 9; ++i) {
            Fc[i] = F[9 * p + i] + dFc[9 * p + i];
            IC[i] = ((i % 4 == 0) ? 1.0 : 0.0) + dt * Cp[i];
        }
        m3_mul(IC, Fc, Fn);
        for (int i = 0; i < 9; ++i) {
            F1[9 * p + i] = Fn[i] * (1.0 - s) + F[9 * p + i] * s;
            C1[9 * p + i] = Cp[i];
        }
        for (int a = 0; a < 3; ++a) {
            v1[3 * p + a] = vp[a];
            x1[3 * p + a] = x[3 * p + a] + dt * vp[a];
        }
    }
}

void orc_forward_step(const orc_params* prm,
                      const double* x, const double* v, const double* C, const double* F,
                      const double* dFc, const double* m, const double* vol,
                      const double* lam, const double* mu,
                      double* grid_m, double* grid_v,
                      double* x1, double* v1, double* C1, double* F1)
{
    long G = grid_size(prm);
    /* P_op_1 + G_Reset + P2G; grid_v holds the momentum until G_op */
    orc_p2g(prm, x, v, C, F, dFc, m, vol, lam, mu, prm->dt, prm->drag, grid_m, grid_v);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < G; ++i) grid_op_node(prm, i, grid_m, grid_v);
    g2p_update(prm, prm->smoothing, x, F, dFc, grid_v, x1, v1, C1, F1);
}

/* ------------------------------------------------------------------ */
/* MultiThreadMPM::ForwardTimeStep: MultiThreadForwardSimulation.cpp:7-225 */
/* ------------------------------------------------------------------ */
/* The particles are cut into `sections` contiguous runs, the first taking
 * the remainder (:21-33); each run scatters into a proxy grid of its own
 * with plain additions (:171-181), then every node sums the proxy grids in
 * section order and applies G_op (:183-207); G2P and P_op_2 without
 * smoothing follow (:210-225).  With a fixed section count the result does
 * not depend on the thread count or schedule. */
int orc_forward_step_proxy(const orc_params* prm, int sections,
                           const double* x, const double* v, const double* C, const double* F,
                           const double* dFc, const double* m, const double* vol,
                           const double* lam, const double* mu,
                           double* grid_m, double* grid_v,
                           double* x1, double* v1, double* C1, double* F1)
{
    if (sections < 1) return -1;
    long G = grid_size(prm);
    double* proxy = (double*)calloc((size_t)sections * 4 * (size_t)G, sizeof(double));
    if (!proxy) return -2;
    int per = prm->n / sections, first = per + prm->n % sections;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < sections; ++t) {
        int begin = t == 0 ? 0 : first + (t - 1) * per, count = t == 0 ? first : per;
        double* pm = proxy + (size_t)t * 4 * G;
        for (int p = begin; p < begin + count; ++p)
            p2g_particle(prm, p, x, v, C, F, dFc, m, vol, lam, mu, prm->dt, prm->drag, pm, pm + G, 0);
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < G; ++i) {
        double gm = 0.0, gp[3] = {0.0, 0.0, 0.0};
        for (int t = 0; t < sections; ++t) {
            const double* pm = proxy + (size_t)t * 4 * G;
            gm += pm[i];
            for (int a = 0; a < 3; ++a) gp[a] += pm[G + 3 * i + a];
        }
        grid_m[i] = gm;
        for (int a = 0; a < 3; ++a) grid_v[3 * i + a] = gp[a];
        grid_op_node(prm, i, grid_m, grid_v);
    }
    free(proxy);
    g2p_update(prm, 0.0, x, F, dFc, grid_v, x1, v1, C1, F1);
    return 0;
}


/* ForwardSimulation.cpp:262-290: vol = m / sum_i w_ip m_i (no dx^3 factor) */
void orc_point_volumes(const orc_params* prm, const double* x, const double* m, double* vol)
{
    long G = grid_size(prm);
    double* gm = (double*)malloc(sizeof(double) * (size_t)G);
    orc_p2g_mass(prm, x, m, gm);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 0);
        double mg = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    mg += st.N[0][i] * st.N[1][j] * st.N[2][k] * gm[id];
                }
        vol[p] = (fabs(mg) > ORC_MASS_EPS) ? m[p] / mg : 0.0;
    }
    free(gm);
}

/* ------------------------------------------------------------------ */
/* EndLayerMassLoss: CompGraph.cpp:110-244                              */
/* ------------------------------------------------------------------ */
double orc_end_mass_loss(const orc_params* prm, const double* x, const double* m,
                         const double* target_m, double* grid_m, double* grid_dLdm,
                         double* dLdx)
{
    long G = grid_size(prm);
    orc_p2g_mass(prm, x, m, grid_m);
    double loss = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : loss)
    for (long i = 0; i < G; ++i) {
        double cm = grid_m[i], tm = target_m[i];
        double c = log(cm + 1.0 + ORC_LOSS_EPS) - log(tm + 1.0 + ORC_LOSS_EPS);
        double l = 0.5 * c * c;
        double g = c / (cm + 1.0 + ORC_LOSS_EPS);
        if (cm < ORC_MIN_MASS) {
            double diff = ORC_MIN_MASS - cm;
            l += ORC_PENALTY * diff * diff;
            g += -2.0 * ORC_PENALTY * diff;
        }
        loss += l;
        grid_dLdm[i] = g;
    }
    double inv_dx = 1.0 / prm->dx;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 1);
        double gx[3] = {0, 0, 0};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    double gw[3] = {-inv_dx * st.dN[0][i] * st.N[1][j] * st.N[2][k],
                                    -inv_dx * st.N[0][i] * st.dN[1][j] * st.N[2][k],
                                    -inv_dx * st.N[0][i] * st.N[1][j] * st.dN[2][k]};
                    for (int a = 0; a < 3; ++a) gx[a] += m[p] * grid_dLdm[id] * gw[a];
                }
        for (int a = 0; a < 3; ++a) dLdx[3 * p + a] = gx[a];
    }
    return loss;
}

/* ------------------------------------------------------------------ */
/* Back_Timestep: BackPropagation.cpp:6-125                             */
/*                                                                      */
/* Exact adjoint of orc_forward_step.  Bars are dL/d(.), primes layer   */
/* n+1, Fc = F + dFc, D = 3/dx^2, k = D dt vol:                         */
/*   vt = v' + dt x'                         (BackPropagation.cpp:23)   */
/*   Ct = C' + (1-s) dt F' Fc^T              (:24; (b) drops 1-s)       */
/*   vbar_i = sum_p w (vt + D Ct d)          (:27-44)                   */
/*   pbar_i = vbar_i/m_i, mbar_i = -(v_i - dt f).vbar_i/m_i  (:46-59)   */
/*   Gb = sum_i w pbar_i d^T,  Cbar = m Gb,  Pbar = -k Gb Fc            */
/*   Fcbar = (1-s)(I+dtC')^T F' - k Gb^T P + H(Fc):Pbar                 */
/*   xbar  = x' + sum_i [m mbar_i gw + gw (b + G d).pbar_i - w G^T pbar_i*/
/*           + gw (v_i.vt) + D((Ct : v_i d^T) gw - w Ct^T v_i)]         */
/*   vbar  = sum_i w m (1 - dt drag) pbar_i                             */
/* with G = -k P Fc^T + m C and b = m (1 - dt drag) v.                  */
/* Outputs: gdFc = Fcbar, gF = Fcbar + s F'.  In compat mode gF = gdFc  */
/* = the reference's dLdF, including defect (a).                        */
/* ------------------------------------------------------------------ */
void orc_backward_step(const orc_params* prm,
                       const double* x, const double* v, const double* C, const double* F,
                       const double* dFc, const double* m, const double* vol,
                       const double* lam, const double* mu,
                       const double* grid_m, const double* grid_v,
                       const double* C1, const double* gx1, const double* gv1,
                       const double* gF1, const double* gC1,
                       double* grid_vbar, double* grid_mbar,
                       double* gx, double* gv, double* gF, double* gC, double* gdFc)
{
    long G = grid_size(prm);
    const int cp = prm->compat;
    const double dt = prm->dt, s = prm->smoothing, drag = prm->drag;
    const double Dinv = 3.0 / (prm->dx * prm->dx);
    const double inv_dx = 1.0 / prm->dx;
    const double cs = (cp & ORC_COMPAT_NO_1MS) ? 1.0 : (1.0 - s);

    /* G2P transposed: scatter vbar_i */
    memset(grid_vbar, 0, sizeof(double) * 3 * G);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        double Fc[9], vt[3], Ct[9], t[9];
        for (int i = 0; i < 9; ++i) Fc[i] = F[9 * p + i] + dFc[9 * p + i];
        for (int a = 0; a < 3; ++a) vt[a] = gv1[3 * p + a] + dt * gx1[3 * p + a];
        m3_mult(gF1 + 9 * p, Fc, t); /* F' Fc^T */
        for (int i = 0; i < 9; ++i) Ct[i] = gC1[9 * p + i] + cs * dt * t[i];
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 0);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    double w = st.N[0][i] * st.N[1][j] * st.N[2][k];
                    double d[3] = {st.d[0][i], st.d[1][j], st.d[2][k]};
                    for (int a = 0; a < 3; ++a) {
                        double val = w * (vt[a] + Dinv * (Ct[3 * a] * d[0] + Ct[3 * a + 1] * d[1] + Ct[3 * a + 2] * d[2]));
#pragma omp atomic
                        grid_vbar[3 * id + a] += val;
                    }
                }
    }

    /* G_op transposed; pbar overwrites vbar */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < G; ++i) {
        double mi = grid_m[i];
        if (fabs(mi) > ORC_MASS_EPS) {
            double dotv = 0.0;
            for (int a = 0; a < 3; ++a) {
                double vi = grid_v[3 * i + a];
                if (!(cp & ORC_COMPAT_DLDM_NO_FEXT)) vi -= dt * prm->f_ext[a];
                dotv += vi * grid_vbar[3 * i + a];
            }
            grid_mbar[i] = -dotv / mi;
            for (int a = 0; a < 3; ++a) grid_vbar[3 * i + a] /= mi;
        } else {
            grid_mbar[i] = 0.0;
            for (int a = 0; a < 3; ++a) grid_vbar[3 * i + a] = 0.0;
        }
    }

    /* P2G transposed (gather) + P_op_1 transposed */
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        const double mp = m[p];
        double Fc[9], P[9], PF[9], Gm[9], t[9];
        for (int i = 0; i < 9; ++i) Fc[i] = F[9 * p + i] + dFc[9 * p + i];
        orc_pk1(Fc, lam[p], mu[p], P);
        const double kk = Dinv * dt * vol[p];
        m3_mult(P, Fc, PF);
        for (int i = 0; i < 9; ++i) Gm[i] = -kk * PF[i] + mp * C[9 * p + i];
        double b[3];
        const double bscale = (cp & ORC_COMPAT_NO_DRAG_X) ? 1.0 : (1.0 - dt * drag);
        for (int a = 0; a < 3; ++a) b[a] = mp * bscale * v[3 * p + a];
        double vt[3], Ct[9];
        for (int a = 0; a < 3; ++a) vt[a] = gv1[3 * p + a] + dt * gx1[3 * p + a];
        m3_mult(gF1 + 9 * p, Fc, t);
        for (int i = 0; i < 9; ++i) Ct[i] = gC1[9 * p + i] + cs * dt * t[i];

        double Gb[9] = {0}, xb[3] = {0, 0, 0}, vb[3] = {0, 0, 0};
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 1);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    double w = st.N[0][i] * st.N[1][j] * st.N[2][k];
                    double gw[3] = {-inv_dx * st.dN[0][i] * st.N[1][j] * st.N[2][k],
                                    -inv_dx * st.N[0][i] * st.dN[1][j] * st.N[2][k],
                                    -inv_dx * st.N[0][i] * st.N[1][j] * st.dN[2][k]};
                    double d[3] = {st.d[0][i], st.d[1][j], st.d[2][k]};
                    const double* pb = grid_vbar + 3 * id;
                    const double* vi = grid_v + 3 * id;
                    double mb = grid_mbar[id];
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c) Gb[3 * a + c] += w * pb[a] * d[c];
                    /* (b + G d) . pbar */
                    double bgd = 0.0, vdot = 0.0, cvd = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        double gd = Gm[3 * a] * d[0] + Gm[3 * a + 1] * d[1] + Gm[3 * a + 2] * d[2];
                        bgd += (b[a] + gd) * pb[a];
                        vdot += vi[a] * vt[a];
                        cvd += vi[a] * (Ct[3 * a] * d[0] + Ct[3 * a + 1] * d[1] + Ct[3 * a + 2] * d[2]);
                    }
                    for (int a = 0; a < 3; ++a) {
                        double GTp = Gm[a] * pb[0] + Gm[3 + a] * pb[1] + Gm[6 + a] * pb[2];
                        double CTv = Ct[a] * vi[0] + Ct[3 + a] * vi[1] + Ct[6 + a] * vi[2];
                        xb[a] += mp * mb * gw[a] + gw[a] * bgd - w * GTp + gw[a] * vdot +
                                 Dinv * (cvd * gw[a] - w * CTv);
                        vb[a] += w * mp * (1.0 - dt * drag) * pb[a];
                    }
                }
        /* Pbar = -k Gb Fc ; stress term of Fcbar = -k Gb^T P */
        double Pb[9], GTP[9], H[9], IC[9], base[9];
        m3_mul(Gb, Fc, t);
        for (int i = 0; i < 9; ++i) Pb[i] = -kk * t[i];
        m3_tmul(Gb, P, GTP);
        orc_hess_mult(Fc, lam[p], mu[p], Pb, H, cp ? 0 : 1);
        for (int i = 0; i < 9; ++i) IC[i] = ((i % 4 == 0) ? 1.0 : 0.0) + dt * C1[9 * p + i];
        m3_tmul(IC, gF1 + 9 * p, base); /* (I + dt C')^T F' */
        for (int i = 0; i < 9; ++i) {
            double fcb;
            if (cp & ORC_COMPAT_DUP_FBAR) {
                /* BackPropagation.cpp:20-21 then :121-122 */
                fcb = (1.0 - s) * base[i] + s * gF1[9 * p + i] - kk * GTP[i] + base[i] + H[i];
                gF[9 * p + i] = fcb;
                gdFc[9 * p + i] = fcb;
            } else {
                fcb = (1.0 - s) * base[i] - kk * GTP[i] + H[i];
                gdFc[9 * p + i] = fcb;
                gF[9 * p + i] = fcb + s * gF1[9 * p + i];
            }
            gC[9 * p + i] = mp * Gb[i];
        }
        for (int a = 0; a < 3; ++a) {
            gx[3 * p + a] = gx1[3 * p + a] + xb[a];
            gv[3 * p + a] = vb[a];
        }
    }
}
