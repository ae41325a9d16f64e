/*
 * CPU FP64 oracle for the differentiable MLS-MPM step.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker that tests/, the smoke entry
 * point and bench.py's CPU-baseline leg compare the HIP path against.  The
 * product never links, loads or calls it.
 *
 * Each function names the reference lines it restates (paths relative to the
 * reference repository root).  The arithmetic is double precision throughout
 * (orc_real; the float build of the same source is only a timing baseline);
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
#include <tgmath.h> /* fabs, sqrt, floor, ... of the argument's type */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EPSILON (sizeof(orc_real) == sizeof(float) ? FLT_EPSILON : DBL_EPSILON)

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
orc_real orc_cubic_bspline(orc_real x)
{
    orc_real a = fabs(x);
    if (a < 1.0) return 0.5 * a * a * a - a * a + 2.0 / 3.0;
    if (a < 2.0) return (2.0 - a) * (2.0 - a) * (2.0 - a) / 6.0;
    return 0.0;
}

orc_real orc_cubic_bspline_slope(orc_real x)
{
    orc_real a = fabs(x);
    if (a < 1.0) return 1.5 * x * a - 2.0 * x;
    if (a < 2.0) return -x * a / 2.0 + 2.0 * x - 2.0 * x / a;
    return 0.0;
}

/* ------------------------------------------------------------------ */
/* 3x3 helpers (row-major, (i,j) at 3*i+j)                              */
/* ------------------------------------------------------------------ */
static void m3_mul(const orc_real* A, const orc_real* B, orc_real* out)
{
    orc_real r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
    memcpy(out, r, sizeof r);
}
/* A^T B */
static void m3_tmul(const orc_real* A, const orc_real* B, orc_real* out)
{
    orc_real r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
    memcpy(out, r, sizeof r);
}
/* A B^T */
static void m3_mult(const orc_real* A, const orc_real* B, orc_real* out)
{
    orc_real r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[3 * i + j] = A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
    memcpy(out, r, sizeof r);
}
static void m3_transpose(const orc_real* A, orc_real* out)
{
    orc_real r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = A[3 * j + i];
    memcpy(out, r, sizeof r);
}
static orc_real m3_det(const orc_real* A)
{
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}
/* inverse via the adjugate */
static void m3_inverse(const orc_real* A, orc_real* out)
{
    orc_real d = m3_det(A);
    orc_real r[9];
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
static orc_real m3_inner(const orc_real* A, const orc_real* B)
{
    orc_real s = 0.0;
    for (int i = 0; i < 9; ++i) s += A[i] * B[i];
    return s;
}

/* ------------------------------------------------------------------ */
/* SVD / polar: one-sided (Hestenes) Jacobi, then the conventions of    */
/* qrsvd/ImplicitQRSVD.h:680-686 and :853-875.                          */
/* ------------------------------------------------------------------ */
static void swap_col(orc_real* M, int a, int b)
{
    for (int k = 0; k < 3; ++k) {
        orc_real t = M[3 * k + a];
        M[3 * k + a] = M[3 * k + b];
        M[3 * k + b] = t;
    }
}

int orc_svd3(const orc_real A[9], orc_real U[9], orc_real s[3], orc_real V[9])
{
    orc_real B[9];
    memcpy(B, A, sizeof B);
    for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    int sweep;
    for (sweep = 0; sweep < 64; ++sweep) {
        int rotated = 0;
        for (int r = 0; r < 3; ++r) {
            int p = PQ[r][0], q = PQ[r][1];
            orc_real al = 0, be = 0, ga = 0;
            for (int k = 0; k < 3; ++k) {
                al += B[3 * k + p] * B[3 * k + p];
                be += B[3 * k + q] * B[3 * k + q];
                ga += B[3 * k + p] * B[3 * k + q];
            }
            if (ga == 0.0 || fabs(ga) <= ORC_EPSILON * sqrt(al * be)) continue;
            rotated = 1;
            orc_real zeta = (be - al) / (2.0 * ga);
            orc_real t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            orc_real c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
            for (int k = 0; k < 3; ++k) {
                orc_real bp = B[3 * k + p], bq = B[3 * k + q];
                B[3 * k + p] = c * bp - sn * bq;
                B[3 * k + q] = sn * bp + c * bq;
                orc_real vp = V[3 * k + p], vq = V[3 * k + q];
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
                orc_real t = s[b];
                s[b] = s[b + 1];
                s[b + 1] = t;
                swap_col(B, b, b + 1);
                swap_col(V, b, b + 1);
            }
    orc_real smax = s[0] > 0 ? s[0] : 1.0;
    orc_real tiny = 64.0 * ORC_EPSILON * smax;
    /* left vectors; complete a basis where singular values vanish */
    if (s[0] <= tiny) {
        for (int i = 0; i < 9; ++i) U[i] = (i % 4 == 0) ? 1.0 : 0.0;
    } else {
        for (int k = 0; k < 3; ++k) U[3 * k] = B[3 * k] / s[0];
        if (s[1] <= tiny) {
            /* any unit vector orthogonal to u0 */
            orc_real u0[3] = {U[0], U[3], U[6]};
            int m = (fabs(u0[0]) < fabs(u0[1])) ? ((fabs(u0[0]) < fabs(u0[2])) ? 0 : 2)
                                                : ((fabs(u0[1]) < fabs(u0[2])) ? 1 : 2);
            orc_real e[3] = {0, 0, 0};
            e[m] = 1.0;
            orc_real d = e[0] * u0[0] + e[1] * u0[1] + e[2] * u0[2];
            orc_real w[3] = {e[0] - d * u0[0], e[1] - d * u0[1], e[2] - d * u0[2]};
            orc_real nw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
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

void orc_polar3(const orc_real A[9], orc_real R[9], orc_real S[9])
{
    orc_real U[9], s[3], V[9];
    orc_svd3(A, U, s, V);
    m3_mult(U, V, R); /* U V^T */
    orc_real VS[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) VS[3 * i + j] = V[3 * i + j] * s[j];
    m3_mult(VS, V, S); /* V S V^T */
}

/* ------------------------------------------------------------------ */
/* Elasticity.cpp                                                        */
/* ------------------------------------------------------------------ */
/* Elasticity.cpp:74-86 */
orc_real orc_fc_energy(const orc_real F[9], orc_real lam, orc_real mu)
{
    orc_real U[9], s[3], V[9];
    orc_svd3(F, U, s, V);
    orc_real sum = (s[0] - 1) * (s[0] - 1) + (s[1] - 1) * (s[1] - 1) + (s[2] - 1) * (s[2] - 1);
    orc_real J = m3_det(F);
    if (J < ORC_J_MIN) J = ORC_J_MIN;
    return mu * sum + 0.5 * lam * (J - 1) * (J - 1);
}

/* Elasticity.cpp:88-98: P = 2 mu (F - R) + lam (J-1) J F^-T, J clamped. */
void orc_pk1(const orc_real F[9], orc_real lam, orc_real mu, orc_real P[9])
{
    orc_real R[9], S[9], Fi[9];
    orc_polar3(F, R, S);
    orc_real J = m3_det(F);
    if (J < ORC_J_MIN) J = ORC_J_MIN;
    m3_inverse(F, Fi);
    orc_real k = lam * (J - 1.0) * J;
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
void orc_hess_mult(const orc_real F[9], orc_real lam, orc_real mu, const orc_real dF[9], orc_real out[9],
                   int exact)
{
    orc_real U[9], s[3], V[9];
    orc_svd3(F, U, s, V);
    orc_real Ju = m3_det(F);
    orc_real J = Ju < ORC_J_MIN ? ORC_J_MIN : Ju;
    orc_real Fi[9], Fit[9];
    m3_inverse(F, Fi);
    m3_transpose(Fi, Fit);
    for (int i = 0; i < 9; ++i) out[i] = 2.0 * mu * dF[i];
    /* Fit dF^T Fit */
    orc_real t1[9], FdF[9];
    m3_mult(Fit, dF, t1); /* Fit dF^T */
    m3_mul(t1, Fit, FdF);
    if (exact && Ju <= ORC_J_MIN) {
        orc_real k = lam * (J - 1.0) * J;
        for (int i = 0; i < 9; ++i) out[i] -= k * FdF[i];
    } else {
        orc_real inner = J * m3_inner(Fit, dF);
        for (int i = 0; i < 9; ++i) out[i] += lam * J * Fit[i] * inner;
        orc_real fd = m3_inner(Fit, dF);
        orc_real k = lam * (J - 1.0) * Ju;
        for (int i = 0; i < 9; ++i) out[i] += k * (fd * Fit[i] - FdF[i]);
    }
    orc_real M[9], tmp[9];
    m3_tmul(U, dF, tmp);
    m3_mul(tmp, V, M);
    orc_real w[9] = {0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            if (i == j) continue;
            orc_real den = s[i] + s[j];
            orc_real scale = fabs(s[0]) > 1.0 ? fabs(s[0]) : 1.0;
            w[3 * i + j] = (fabs(den) > 1e-12 * scale) ? (M[3 * i + j] - M[3 * j + i]) / den : 0.0;
        }
    orc_real dR[9];
    m3_mul(U, w, tmp);
    m3_mult(tmp, V, dR);
    for (int i = 0; i < 9; ++i) out[i] -= 2.0 * mu * dR[i];
}

/* Elasticity.cpp:231-235 */
void orc_lame(orc_real E, orc_real nu, orc_real* lam, orc_real* mu)
{
    *lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    *mu = E / (2.0 + 2.0 * nu);
}

void orc_pk1_batch(int n, const orc_real* F, const orc_real* lam, const orc_real* mu, orc_real* P)
{
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; ++p) orc_pk1(F + 9 * p, lam[p], mu[p], P + 9 * p);
}

void orc_hess_mult_batch(int n, const orc_real* F, const orc_real* lam, const orc_real* mu,
                         const orc_real* dF, orc_real* out, int exact)
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
    orc_real N[3][4];  /* N(u) per axis and offset */
    orc_real dN[3][4]; /* N'(u) */
    orc_real d[3][4];  /* x_i - x_p per axis and offset */
} orc_stencil;

static void stencil(const orc_params* prm, const orc_real* xp, orc_stencil* st, int with_slope)
{
    for (int a = 0; a < 3; ++a) {
        orc_real rel = xp[a] - prm->min_point[a];
        /* clamped to +-2^20 cells (inf and NaN included) before the int
         * conversion: every node of such a particle is outside the grid, and
         * base + 3 cannot overflow */
        st->base[a] = (int)fmin(fmax(floor(rel / prm->dx), -1048576.0), 1048576.0) - 1;
        for (int o = 0; o < 4; ++o) {
            orc_real xi = prm->min_point[a] + (orc_real)(st->base[a] + o) * prm->dx;
            orc_real d = xi - xp[a];
            orc_real u = d / prm->dx;
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
static void p2g_particle(const orc_params* prm, int p, const orc_real* x, const orc_real* v, const orc_real* C,
                         const orc_real* F, const orc_real* dFc, const orc_real* m, const orc_real* vol,
                         const orc_real* lam, const orc_real* mu, orc_real dt, orc_real drag,
                         orc_real* grid_m, orc_real* grid_p, int shared)
{
    orc_real Dinv = 3.0 / (prm->dx * prm->dx);
    orc_real Fc[9], P[9], PF[9], A[9];
    for (int i = 0; i < 9; ++i) Fc[i] = F[9 * p + i] + dFc[9 * p + i];
    orc_pk1(Fc, lam[p], mu[p], P);
    m3_mult(P, Fc, PF); /* P Fc^T */
    for (int i = 0; i < 9; ++i) A[i] = -Dinv * dt * vol[p] * PF[i] + m[p] * C[9 * p + i];
    orc_real mv[3];
    for (int a = 0; a < 3; ++a) mv[a] = m[p] * v[3 * p + a] * (1.0 - dt * drag);
    orc_stencil st;
    stencil(prm, x + 3 * p, &st, 0);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                long id = node_index(prm, &st, i, j, k);
                if (id < 0) continue;
                orc_real w = st.N[0][i] * st.N[1][j] * st.N[2][k];
                orc_real d[3] = {st.d[0][i], st.d[1][j], st.d[2][k]};
                orc_real dm = w * m[p];
                if (shared) {
#pragma omp atomic
                    grid_m[id] += dm;
                } else {
                    grid_m[id] += dm;
                }
                for (int a = 0; a < 3; ++a) {
                    orc_real dp = w * (mv[a] + A[3 * a] * d[0] + A[3 * a + 1] * d[1] + A[3 * a + 2] * d[2]);
                    if (shared) {
#pragma omp atomic
                        grid_p[3 * id + a] += dp;
                    } else {
                        grid_p[3 * id + a] += dp;
                    }
                }
            }
}

void orc_p2g(const orc_params* prm, const orc_real* x, const orc_real* v, const orc_real* C,
             const orc_real* F, const orc_real* dFc, const orc_real* m, const orc_real* vol,
             const orc_real* lam, const orc_real* mu, orc_real dt, orc_real drag,
             orc_real* grid_m, orc_real* grid_p)
{
    long G = grid_size(prm);
    memset(grid_m, 0, sizeof(orc_real) * G);
    memset(grid_p, 0, sizeof(orc_real) * 3 * G);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p)
        p2g_particle(prm, p, x, v, C, F, dFc, m, vol, lam, mu, dt, drag, grid_m, grid_p, 1);
}

void orc_p2g_mass(const orc_params* prm, const orc_real* x, const orc_real* m, orc_real* grid_m)
{
    long G = grid_size(prm);
    memset(grid_m, 0, sizeof(orc_real) * G);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 0);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    orc_real dm = st.N[0][i] * st.N[1][j] * st.N[2][k] * m[p];
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
static void grid_op_node(const orc_params* prm, long i, const orc_real* grid_m, orc_real* grid_v)
{
    if (fabs(grid_m[i]) > ORC_MASS_EPS) {
        for (int a = 0; a < 3; ++a) grid_v[3 * i + a] = grid_v[3 * i + a] / grid_m[i] + prm->dt * prm->f_ext[a];
    } else {
        for (int a = 0; a < 3; ++a) grid_v[3 * i + a] = 0.0;
    }
}

/* G2P + P_op_2: ForwardSimulation.cpp:60-106,108-124 (s = 0 is the
 * overload without smoothing, :108-112) */
static void g2p_update(const orc_params* prm, orc_real s, const orc_real* x, const orc_real* F, const orc_real* dFc,
                       const orc_real* grid_v, orc_real* x1, orc_real* v1, orc_real* C1, orc_real* F1)
{
    orc_real Dinv = 3.0 / (prm->dx * prm->dx);
    orc_real dt = prm->dt;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 0);
        orc_real vp[3] = {0, 0, 0}, Cp[9] = {0};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    orc_real w = st.N[0][i] * st.N[1][j] * st.N[2][k];
                    orc_real d[3] = {st.d[0][i], st.d[1][j], st.d[2][k]};
                    for (int a = 0; a < 3; ++a) {
                        orc_real vi = grid_v[3 * id + a];
                        vp[a] += w * vi;
                        for (int b = 0; b < 3; ++b) Cp[3 * a + b] += Dinv * w * vi * d[b];
                    }
                }
        orc_real Fc[9], IC[9], Fn[9];
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
                      const orc_real* x, const orc_real* v, const orc_real* C, const orc_real* F,
                      const orc_real* dFc, const orc_real* m, const orc_real* vol,
                      const orc_real* lam, const orc_real* mu,
                      orc_real* grid_m, orc_real* grid_v,
                      orc_real* x1, orc_real* v1, orc_real* C1, orc_real* F1)
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
                           const orc_real* x, const orc_real* v, const orc_real* C, const orc_real* F,
                           const orc_real* dFc, const orc_real* m, const orc_real* vol,
                           const orc_real* lam, const orc_real* mu,
                           orc_real* grid_m, orc_real* grid_v,
                           orc_real* x1, orc_real* v1, orc_real* C1, orc_real* F1)
{
    if (sections < 1) return -1;
    long G = grid_size(prm);
    orc_real* proxy = (orc_real*)calloc((size_t)sections * 4 * (size_t)G, sizeof(orc_real));
    if (!proxy) return -2;
    int per = prm->n / sections, first = per + prm->n % sections;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < sections; ++t) {
        int begin = t == 0 ? 0 : first + (t - 1) * per, count = t == 0 ? first : per;
        orc_real* pm = proxy + (size_t)t * 4 * G;
        for (int p = begin; p < begin + count; ++p)
            p2g_particle(prm, p, x, v, C, F, dFc, m, vol, lam, mu, prm->dt, prm->drag, pm, pm + G, 0);
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < G; ++i) {
        orc_real gm = 0.0, gp[3] = {0.0, 0.0, 0.0};
        for (int t = 0; t < sections; ++t) {
            const orc_real* pm = proxy + (size_t)t * 4 * G;
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
void orc_point_volumes(const orc_params* prm, const orc_real* x, const orc_real* m, orc_real* vol)
{
    long G = grid_size(prm);
    orc_real* gm = (orc_real*)malloc(sizeof(orc_real) * (size_t)G);
    orc_p2g_mass(prm, x, m, gm);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 0);
        orc_real mg = 0.0;
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
orc_real orc_end_mass_loss(const orc_params* prm, const orc_real* x, const orc_real* m,
                         const orc_real* target_m, orc_real* grid_m, orc_real* grid_dLdm,
                         orc_real* dLdx)
{
    long G = grid_size(prm);
    orc_p2g_mass(prm, x, m, grid_m);
    orc_real loss = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : loss)
    for (long i = 0; i < G; ++i) {
        orc_real cm = grid_m[i], tm = target_m[i];
        orc_real c = log(cm + 1.0 + ORC_LOSS_EPS) - log(tm + 1.0 + ORC_LOSS_EPS);
        orc_real l = 0.5 * c * c;
        orc_real g = c / (cm + 1.0 + ORC_LOSS_EPS);
        if (cm < ORC_MIN_MASS) {
            orc_real diff = ORC_MIN_MASS - cm;
            l += ORC_PENALTY * diff * diff;
            g += -2.0 * ORC_PENALTY * diff;
        }
        loss += l;
        grid_dLdm[i] = g;
    }
    orc_real inv_dx = 1.0 / prm->dx;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 1);
        orc_real gx[3] = {0, 0, 0};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    orc_real gw[3] = {-inv_dx * st.dN[0][i] * st.N[1][j] * st.N[2][k],
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
                       const orc_real* x, const orc_real* v, const orc_real* C, const orc_real* F,
                       const orc_real* dFc, const orc_real* m, const orc_real* vol,
                       const orc_real* lam, const orc_real* mu,
                       const orc_real* grid_m, const orc_real* grid_v,
                       const orc_real* C1, const orc_real* gx1, const orc_real* gv1,
                       const orc_real* gF1, const orc_real* gC1,
                       orc_real* grid_vbar, orc_real* grid_mbar,
                       orc_real* gx, orc_real* gv, orc_real* gF, orc_real* gC, orc_real* gdFc)
{
    long G = grid_size(prm);
    const int cp = prm->compat;
    const orc_real dt = prm->dt, s = prm->smoothing, drag = prm->drag;
    const orc_real Dinv = 3.0 / (prm->dx * prm->dx);
    const orc_real inv_dx = 1.0 / prm->dx;
    const orc_real cs = (cp & ORC_COMPAT_NO_1MS) ? 1.0 : (1.0 - s);

    /* G2P transposed: scatter vbar_i */
    memset(grid_vbar, 0, sizeof(orc_real) * 3 * G);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < prm->n; ++p) {
        orc_real Fc[9], vt[3], Ct[9], t[9];
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
                    orc_real w = st.N[0][i] * st.N[1][j] * st.N[2][k];
                    orc_real d[3] = {st.d[0][i], st.d[1][j], st.d[2][k]};
                    for (int a = 0; a < 3; ++a) {
                        orc_real val = w * (vt[a] + Dinv * (Ct[3 * a] * d[0] + Ct[3 * a + 1] * d[1] + Ct[3 * a + 2] * d[2]));
#pragma omp atomic
                        grid_vbar[3 * id + a] += val;
                    }
                }
    }

    /* G_op transposed; pbar overwrites vbar */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < G; ++i) {
        orc_real mi = grid_m[i];
        if (fabs(mi) > ORC_MASS_EPS) {
            orc_real dotv = 0.0;
            for (int a = 0; a < 3; ++a) {
                orc_real vi = grid_v[3 * i + a];
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
        const orc_real mp = m[p];
        orc_real Fc[9], P[9], PF[9], Gm[9], t[9];
        for (int i = 0; i < 9; ++i) Fc[i] = F[9 * p + i] + dFc[9 * p + i];
        orc_pk1(Fc, lam[p], mu[p], P);
        const orc_real kk = Dinv * dt * vol[p];
        m3_mult(P, Fc, PF);
        for (int i = 0; i < 9; ++i) Gm[i] = -kk * PF[i] + mp * C[9 * p + i];
        orc_real b[3];
        const orc_real bscale = (cp & ORC_COMPAT_NO_DRAG_X) ? 1.0 : (1.0 - dt * drag);
        for (int a = 0; a < 3; ++a) b[a] = mp * bscale * v[3 * p + a];
        orc_real vt[3], Ct[9];
        for (int a = 0; a < 3; ++a) vt[a] = gv1[3 * p + a] + dt * gx1[3 * p + a];
        m3_mult(gF1 + 9 * p, Fc, t);
        for (int i = 0; i < 9; ++i) Ct[i] = gC1[9 * p + i] + cs * dt * t[i];

        orc_real Gb[9] = {0}, xb[3] = {0, 0, 0}, vb[3] = {0, 0, 0};
        orc_stencil st;
        stencil(prm, x + 3 * p, &st, 1);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    long id = node_index(prm, &st, i, j, k);
                    if (id < 0) continue;
                    orc_real w = st.N[0][i] * st.N[1][j] * st.N[2][k];
                    orc_real gw[3] = {-inv_dx * st.dN[0][i] * st.N[1][j] * st.N[2][k],
                                    -inv_dx * st.N[0][i] * st.dN[1][j] * st.N[2][k],
                                    -inv_dx * st.N[0][i] * st.N[1][j] * st.dN[2][k]};
                    orc_real d[3] = {st.d[0][i], st.d[1][j], st.d[2][k]};
                    const orc_real* pb = grid_vbar + 3 * id;
                    const orc_real* vi = grid_v + 3 * id;
                    orc_real mb = grid_mbar[id];
                    for (int a = 0; a < 3; ++a)
                        for (int c = 0; c < 3; ++c) Gb[3 * a + c] += w * pb[a] * d[c];
                    /* (b + G d) . pbar */
                    orc_real bgd = 0.0, vdot = 0.0, cvd = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        orc_real gd = Gm[3 * a] * d[0] + Gm[3 * a + 1] * d[1] + Gm[3 * a + 2] * d[2];
                        bgd += (b[a] + gd) * pb[a];
                        vdot += vi[a] * vt[a];
                        cvd += vi[a] * (Ct[3 * a] * d[0] + Ct[3 * a + 1] * d[1] + Ct[3 * a + 2] * d[2]);
                    }
                    for (int a = 0; a < 3; ++a) {
                        orc_real GTp = Gm[a] * pb[0] + Gm[3 + a] * pb[1] + Gm[6 + a] * pb[2];
                        orc_real CTv = Ct[a] * vi[0] + Ct[3 + a] * vi[1] + Ct[6 + a] * vi[2];
                        xb[a] += mp * mb * gw[a] + gw[a] * bgd - w * GTp + gw[a] * vdot +
                                 Dinv * (cvd * gw[a] - w * CTv);
                        vb[a] += w * mp * (1.0 - dt * drag) * pb[a];
                    }
                }
        /* Pbar = -k Gb Fc ; stress term of Fcbar = -k Gb^T P */
        orc_real Pb[9], GTP[9], H[9], IC[9], base[9];
        m3_mul(Gb, Fc, t);
        for (int i = 0; i < 9; ++i) Pb[i] = -kk * t[i];
        m3_tmul(Gb, P, GTP);
        orc_hess_mult(Fc, lam[p], mu[p], Pb, H, cp ? 0 : 1);
        for (int i = 0; i < 9; ++i) IC[i] = ((i % 4 == 0) ? 1.0 : 0.0) + dt * C1[9 * p + i];
        m3_tmul(IC, gF1 + 9 * p, base); /* (I + dt C')^T F' */
        for (int i = 0; i < 9; ++i) {
            orc_real fcb;
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
