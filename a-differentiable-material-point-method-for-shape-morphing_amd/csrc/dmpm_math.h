// Device math for the differentiable MLS-MPM step on gfx950.
//
// 3x3 algebra in registers, the cubic B-spline stencil written per offset
// (no branches), an FP64 3x3 SVD/polar, the fixed-corotated PK1 and its
// directional derivative H(F):dF.  Matrices are row-major m[3*i+j].
//
// Reference semantics followed (paths relative to the reference root):
//   Interpolation.h:5-30               N(x), N'(x)
//   ForwardSimulation.cpp:15-20        base = floor((x-min)/dx) - 1
//   Elasticity.cpp:88-98               P = 2mu(F-R) + lam(J-1)J F^-T, J>=1e-6
//   Elasticity.cpp:100-178             H:dF with dR from R^T dF - dF^T R = WS + SW
//   qrsvd/ImplicitQRSVD.h:680-686,853-875  U,V rotations, |s| sorted
//                                      descending, sign on the smallest
#pragma once

#include <hip/hip_runtime.h>

namespace dmpm {

constexpr double kMassEps = 1.1920928955078125e-07;  // FLT_EPSILON (ForwardSimulation.cpp:55)
constexpr double kJMin = 1e-6;                        // Elasticity.cpp:95

template <typename T>
struct M3 {
    T a[9];
};

template <typename T>
__device__ __forceinline__ M3<T> m3_zero()
{
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.a[i] = T(0);
    return r;
}

template <typename T>
__device__ __forceinline__ M3<T> m3_eye()
{
    M3<T> r = m3_zero<T>();
    r.a[0] = r.a[4] = r.a[8] = T(1);
    return r;
}

// A B
template <typename T>
__device__ __forceinline__ M3<T> mul(const M3<T>& A, const M3<T>& B)
{
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.a[3 * i + j] = A.a[3 * i] * B.a[j] + A.a[3 * i + 1] * B.a[3 + j] + A.a[3 * i + 2] * B.a[6 + j];
    return r;
}

// A^T B
template <typename T>
__device__ __forceinline__ M3<T> tmul(const M3<T>& A, const M3<T>& B)
{
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.a[3 * i + j] = A.a[i] * B.a[j] + A.a[3 + i] * B.a[3 + j] + A.a[6 + i] * B.a[6 + j];
    return r;
}

// A B^T
template <typename T>
__device__ __forceinline__ M3<T> mult(const M3<T>& A, const M3<T>& B)
{
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.a[3 * i + j] = A.a[3 * i] * B.a[3 * j] + A.a[3 * i + 1] * B.a[3 * j + 1] + A.a[3 * i + 2] * B.a[3 * j + 2];
    return r;
}

template <typename T>
__device__ __forceinline__ M3<T> transpose(const M3<T>& A)
{
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[3 * i + j] = A.a[3 * j + i];
    return r;
}

template <typename T>
__device__ __forceinline__ T det(const M3<T>& A)
{
    const T* a = A.a;
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// cofactor matrix: cof(A) = det(A) A^-T
template <typename T>
__device__ __forceinline__ M3<T> cofactor(const M3<T>& A)
{
    const T* a = A.a;
    M3<T> r;
    r.a[0] = a[4] * a[8] - a[5] * a[7];
    r.a[1] = a[5] * a[6] - a[3] * a[8];
    r.a[2] = a[3] * a[7] - a[4] * a[6];
    r.a[3] = a[2] * a[7] - a[1] * a[8];
    r.a[4] = a[0] * a[8] - a[2] * a[6];
    r.a[5] = a[1] * a[6] - a[0] * a[7];
    r.a[6] = a[1] * a[5] - a[2] * a[4];
    r.a[7] = a[2] * a[3] - a[0] * a[5];
    r.a[8] = a[0] * a[4] - a[1] * a[3];
    return r;
}

template <typename T>
__device__ __forceinline__ T inner(const M3<T>& A, const M3<T>& B)
{
    T s = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) s += A.a[i] * B.a[i];
    return s;
}

template <typename To, typename From>
__device__ __forceinline__ M3<To> cast(const M3<From>& A)
{
    M3<To> r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.a[i] = To(A.a[i]);
    return r;
}

// ---------------------------------------------------------------------------
// Cubic B-spline stencil of one axis.  With u = (x - min)/dx, base =
// floor(u) - 1 and f = u - floor(u) in [0,1), node offset o in 0..3 sits at
// signed distance (o - 1 - f) dx from the particle, so
//   N: (1-f)^3/6, f^3/2 - f^2 + 2/3, (1-f)^3/2 - (1-f)^2 + 2/3, f^3/6
//   N': (1-f)^2/2, 2f - 3f^2/2, 3(1-f)^2/2 - 2(1-f), -f^2/2
// which is N and N' of Interpolation.h evaluated at those distances.
// ---------------------------------------------------------------------------
template <typename T>
struct Axis {
    int base;
    T f;
    T w[4];
    T dw[4];  // N'(u)
};

// Stencil bases are clamped to +-2^20 cells before the integer conversion: a
// particle that far out (or at inf or NaN, which the clamp maps to -2^20) has
// every node outside any grid either way, and base + 3 then cannot overflow.
// Unclamped, a saturated conversion (INT_MAX) plus the node offset is signed
// overflow, and the compiler may then test the offset node against the grid
// bounds in a form that lets the wrapped index through.
constexpr double kBaseClamp = 1048576.0;

template <typename T>
__device__ __forceinline__ int stencil_base(T fl)
{
    return int(fmin(fmax(fl, T(-kBaseClamp)), T(kBaseClamp))) - 1;
}

template <typename T>
__device__ __forceinline__ void axis_weights(T u, Axis<T>& ax, bool slopes)
{
    T fl = floor(u);
    ax.base = stencil_base(fl);
    T f = u - fl;
    ax.f = f;
    T g = T(1) - f;
    T f2 = f * f, g2 = g * g;
    ax.w[0] = g2 * g * T(1.0 / 6.0);
    ax.w[1] = T(0.5) * f2 * f - f2 + T(2.0 / 3.0);
    ax.w[2] = T(0.5) * g2 * g - g2 + T(2.0 / 3.0);
    ax.w[3] = f2 * f * T(1.0 / 6.0);
    if (slopes) {
        ax.dw[0] = T(0.5) * g2;
        ax.dw[1] = T(2) * f - T(1.5) * f2;
        ax.dw[2] = T(1.5) * g2 - T(2) * g;
        ax.dw[3] = T(-0.5) * f2;
    }
}

// ---------------------------------------------------------------------------
// FP64 3x3 SVD: cyclic Jacobi eigen-decomposition of A^T A for V, sort by
// eigenvalue, then Givens QR of A V for U and the signed singular values
// (the scheme of McAdams et al., "Computing the Singular Value
// Decomposition of 3x3 matrices with minimal branching and elementary
// floating point operations", 2011, run to convergence in double).
// U and V are rotations; the QR leaves s0, s1 >= 0 and puts det's sign on s2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void jacobi_rot(double S[3][3], double V[3][3], int p, int q)
{
    double apq = S[p][q];
    if (apq == 0.0) return;
    double theta = (S[q][q] - S[p][p]) / (2.0 * apq);
    double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
    double c = rsqrt(t * t + 1.0);
    double s = t * c;
    // S' = J^T S J with J = [[c, s], [-s, c]] in the (p,q) plane
    double spp = S[p][p], sqq = S[q][q];
    S[p][p] = spp - t * apq;
    S[q][q] = sqq + t * apq;
    S[p][q] = S[q][p] = 0.0;
    int r = 3 - p - q;
    double srp = S[r][p], srq = S[r][q];
    S[r][p] = S[p][r] = c * srp - s * srq;
    S[r][q] = S[q][r] = s * srp + c * srq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double vp = V[k][p], vq = V[k][q];
        V[k][p] = c * vp - s * vq;
        V[k][q] = s * vp + c * vq;
    }
}

__device__ __forceinline__ void givens_qr_step(double B[3][3], double U[3][3], int a, int b, int col)
{
    // zero B[b][col] against B[a][col] with a row rotation; U accumulates Q
    double x = B[a][col], y = B[b][col];
    double rho = sqrt(x * x + y * y);
    double c = 1.0, s = 0.0;
    if (rho > 0.0) {
        c = x / rho;
        s = y / rho;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double ba = B[a][k], bb = B[b][k];
        B[a][k] = c * ba + s * bb;
        B[b][k] = -s * ba + c * bb;
        double ua = U[k][a], ub = U[k][b];
        U[k][a] = c * ua + s * ub;
        U[k][b] = -s * ua + c * ub;
    }
}

__device__ __forceinline__ void svd3(const M3<double>& A, M3<double>& Uo, double sig[3], M3<double>& Vo)
{
    double S[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            S[i][j] = A.a[i] * A.a[j] + A.a[3 + i] * A.a[3 + j] + A.a[6 + i] * A.a[6 + j];
            V[i][j] = (i == j) ? 1.0 : 0.0;
        }
    double scale = S[0][0] + S[1][1] + S[2][2];
    for (int sweep = 0; sweep < 12; ++sweep) {
        double off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
        if (off <= 1e-34 * scale * scale) break;
        jacobi_rot(S, V, 0, 1);
        jacobi_rot(S, V, 0, 2);
        jacobi_rot(S, V, 1, 2);
    }
    double ev[3] = {S[0][0], S[1][1], S[2][2]};
    // sort eigenvalues descending, swapping V columns (with a sign flip so V stays a rotation)
#pragma unroll
    for (int pass = 0; pass < 2; ++pass)
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (ev[i] < ev[i + 1]) {
                double t = ev[i];
                ev[i] = ev[i + 1];
                ev[i + 1] = t;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double a = V[k][i];
                    V[k][i] = -V[k][i + 1];
                    V[k][i + 1] = a;
                }
            }
    // guard det(V) = +1 against rounding of the initial rotation products
    double dv = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) - V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
                V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
    if (dv < 0.0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) V[k][2] = -V[k][2];
    }
    // B = A V
    double B[3][3], U[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            B[i][j] = A.a[3 * i] * V[0][j] + A.a[3 * i + 1] * V[1][j] + A.a[3 * i + 2] * V[2][j];
            U[i][j] = (i == j) ? 1.0 : 0.0;
        }
    givens_qr_step(B, U, 0, 1, 0);
    givens_qr_step(B, U, 0, 2, 0);
    givens_qr_step(B, U, 1, 2, 1);
    sig[0] = B[0][0];
    sig[1] = B[1][1];
    sig[2] = B[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Uo.a[3 * i + j] = U[i][j];
            Vo.a[3 * i + j] = V[i][j];
        }
}

// ---------------------------------------------------------------------------
// Polar rotation by Newton's iteration (FP64).  For det F > 0 the rotation
// R = U V^T of the SVD above is the orthogonal polar factor of F, which the
// scaled Newton iteration X <- (g X + X^-T / g) / 2 (Higham, "Computing the
// polar decomposition -- with applications", 1986) reaches in 3-5 steps for
// the near-rotations of an elastic solid, at a fraction of the Jacobi SVD's
// cost.  Scaling g = sqrt(|X^-1|_F / |X|_F) is used in the first two steps;
// a step that moves X by at most 1e-8 ends the iteration: a step leaves X
// within O(|X - R|^2) of R and moves it by |X - R| to first order, so the
// error after it is ~1e-16, FP64 rounding.  Returns false (the
// caller then takes the SVD) unless det F > 1e-6 |F|_F^3 and the iteration
// converged in 16 steps.  detF = det F is an input.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool polar_newton(const M3<double>& F, double detF, M3<double>& R)
{
    double nf = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) nf += F.a[i] * F.a[i];
    if (!(detF > 1e-6 * nf * sqrt(nf))) return false;
    M3<double> X = F, Y;
    for (int it = 0; it < 16; ++it) {
        {
            const M3<double> cf = cofactor(X);
            const double inv = 1.0 / (X.a[0] * cf.a[0] + X.a[1] * cf.a[1] + X.a[2] * cf.a[2]);
#pragma unroll
            for (int i = 0; i < 9; ++i) Y.a[i] = cf.a[i] * inv;
        }
        // X <- (g X + X^-T / g) / 2 keeps the polar factor for any g > 0
        // (both terms are R times a symmetric positive matrix), so g comes
        // from the FP32 hardware reciprocal and square root: it only sets the
        // speed of convergence.  1/g itself is FP64-accurate (hardware
        // reciprocal, two Newton steps): h hi = 1/4 makes the step's error
        // second order, where an FP32 1/g would leave X off R by a factor
        // 1 + O(1e-7) for the next step to remove.  nx = ny (the identity, the
        // rest state) keeps g = 1 exactly, so that X stays fixed.
        double h = 0.5, hi = 0.5;
        if (it < 2) {
            double nx = 0.0, ny = 0.0;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                nx += X.a[i] * X.a[i];
                ny += Y.a[i] * Y.a[i];
            }
            if (nx != ny) {
                const double g = double(__builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(float(ny) * __builtin_amdgcn_rcpf(float(nx)))));
                double r = __builtin_amdgcn_rcp(g);
                r = fma(r, fma(-g, r, 1.0), r);
                r = fma(r, fma(-g, r, 1.0), r);
                h = 0.5 * g;
                hi = 0.5 * r;
            }
        }
        double step = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const double xn = h * X.a[i] + hi * Y.a[i];
            const double d = xn - X.a[i];
            step += d * d;
            X.a[i] = xn;
        }
        if (step <= 1e-16) {
            R = X;
            return true;
        }
    }
    return false;
}

// ---------------------------------------------------------------------------
// Fixed-corotated elasticity in FP64.
// ---------------------------------------------------------------------------
struct FCE {
    M3<double> U, V, R;
    M3<double> S;   // R^T F symmetrised (polar path)
    double s[3];
    double Ju;      // det F (unclamped)
    double J;       // max(det F, 1e-6)
    M3<double> Fit; // F^-T (unclamped)
    bool svd;       // U, s, V valid (else the polar path: R, S)
};

__device__ __forceinline__ void fce_prepare(const M3<double>& F, FCE& e)
{
    e.Ju = det(F);
    e.J = e.Ju < kJMin ? kJMin : e.Ju;
    M3<double> cf = cofactor(F);
    double inv = 1.0 / e.Ju;
#pragma unroll
    for (int i = 0; i < 9; ++i) e.Fit.a[i] = cf.a[i] * inv;
    e.svd = !polar_newton(F, e.Ju, e.R);
    if (e.svd) {
        svd3(F, e.U, e.s, e.V);
        e.R = mult(e.U, e.V);
    } else {
        const M3<double> RF = tmul(e.R, F);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) e.S.a[3 * i + j] = 0.5 * (RF.a[3 * i + j] + RF.a[3 * j + i]);
    }
}

// P(F) F^T for the PK1 below: P = 2 mu (F - R) + lam (J - 1) J F^-T with the
// clamped J and the unclamped F^-T, and F^-T F^T = I, so
// P F^T = 2 mu (F - R) F^T + lam (J - 1) J I: only the rotation is needed
// (the forward P2G's affine term).
__device__ __forceinline__ M3<double> pk1_times_ft(const M3<double>& F, double lam, double mu)
{
    const double Ju = det(F);
    const double J = Ju < kJMin ? kJMin : Ju;
    M3<double> R;
    if (!polar_newton(F, Ju, R)) {
        M3<double> U, V;
        double sv[3];
        svd3(F, U, sv, V);
        R = mult(U, V);
    }
    M3<double> D;
#pragma unroll
    for (int i = 0; i < 9; ++i) D.a[i] = F.a[i] - R.a[i];
    M3<double> PF = mult(D, F);
    const double k = lam * (J - 1.0) * J;
#pragma unroll
    for (int i = 0; i < 9; ++i) PF.a[i] *= 2.0 * mu;
    PF.a[0] += k;
    PF.a[4] += k;
    PF.a[8] += k;
    return PF;
}

// Elasticity.cpp:74-86: psi = mu sum_i (s_i - 1)^2 + lam/2 (J - 1)^2 with the
// clamped J and the signed singular values (the sign of det F on the last).
// On the polar path R^T F = S = V diag(s) V^T, so sum_i (s_i - 1)^2 = |S - I|_F^2
// without an eigen-decomposition (and without cancellation near the rest state).
__device__ __forceinline__ double fce_energy(const FCE& e, double lam, double mu)
{
    double sum = 0.0;
    if (e.svd) {
#pragma unroll
        for (int a = 0; a < 3; ++a) sum += (e.s[a] - 1.0) * (e.s[a] - 1.0);
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double d = e.S.a[3 * i + j] - (i == j ? 1.0 : 0.0);
                sum += d * d;
            }
    }
    return mu * sum + 0.5 * lam * (e.J - 1.0) * (e.J - 1.0);
}

// Elasticity.cpp:88-98
__device__ __forceinline__ M3<double> fce_pk1(const M3<double>& F, const FCE& e, double lam, double mu)
{
    M3<double> P;
    double k = lam * (e.J - 1.0) * e.J;
#pragma unroll
    for (int i = 0; i < 9; ++i) P.a[i] = 2.0 * mu * (F.a[i] - e.R.a[i]) + k * e.Fit.a[i];
    return P;
}

// Elasticity.cpp:100-178 in closed form: d(J F^-T)[dF] = J[(F^-T:dF)F^-T - F^-T dF^T F^-T]
// and dR = U w V^T with w_ij = (M_ij - M_ji)/(s_i + s_j), M = U^T dF V.
// exact: where the J clamp is active, differentiate the clamped PK1.
__device__ __forceinline__ M3<double> fce_hess_mult(const FCE& e, double lam, double mu, const M3<double>& dF, bool exact)
{
    M3<double> out;
#pragma unroll
    for (int i = 0; i < 9; ++i) out.a[i] = 2.0 * mu * dF.a[i];
    M3<double> FdF = mul(mult(e.Fit, dF), e.Fit);  // F^-T dF^T F^-T
    if (exact && e.Ju <= kJMin) {
        double k = lam * (e.J - 1.0) * e.J;
#pragma unroll
        for (int i = 0; i < 9; ++i) out.a[i] -= k * FdF.a[i];
    } else {
        double fd = inner(e.Fit, dF);
        double k1 = lam * e.J * e.J * fd;
        double k2 = lam * (e.J - 1.0) * e.Ju;
#pragma unroll
        for (int i = 0; i < 9; ++i) out.a[i] += k1 * e.Fit.a[i] + k2 * (fd * e.Fit.a[i] - FdF.a[i]);
    }
    M3<double> dR;
    if (e.svd) {
        M3<double> M = mul(tmul(e.U, dF), e.V);
        double scale = fabs(e.s[0]) > 1.0 ? fabs(e.s[0]) : 1.0;
        M3<double> w = m3_zero<double>();
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (i == j) continue;
                double den = e.s[i] + e.s[j];
                w.a[3 * i + j] = (fabs(den) > 1e-12 * scale) ? (M.a[3 * i + j] - M.a[3 * j + i]) / den : 0.0;
            }
        dR = mult(mul(e.U, w), e.V);
    } else {
        // F = R S: R^T dF - dF^T R = W S + S W with W = R^T dR skew, i.e.
        // (tr S I - S) w = axial(R^T dF - dF^T R) for the axial vector w of W
        const M3<double> K = tmul(e.R, dF);
        const double b0 = K.a[7] - K.a[5], b1 = K.a[2] - K.a[6], b2 = K.a[3] - K.a[1];
        const double tr = e.S.a[0] + e.S.a[4] + e.S.a[8];
        M3<double> A;
#pragma unroll
        for (int i = 0; i < 9; ++i) A.a[i] = -e.S.a[i];
        A.a[0] += tr;
        A.a[4] += tr;
        A.a[8] += tr;
        const M3<double> ca = cofactor(A);  // A symmetric: A^-1 = cof(A) / det A
        const double inv = 1.0 / (A.a[0] * ca.a[0] + A.a[1] * ca.a[1] + A.a[2] * ca.a[2]);
        const double w0 = (ca.a[0] * b0 + ca.a[1] * b1 + ca.a[2] * b2) * inv;
        const double w1 = (ca.a[3] * b0 + ca.a[4] * b1 + ca.a[5] * b2) * inv;
        const double w2 = (ca.a[6] * b0 + ca.a[7] * b1 + ca.a[8] * b2) * inv;
        M3<double> W;
        W.a[0] = 0.0;
        W.a[1] = -w2;
        W.a[2] = w1;
        W.a[3] = w2;
        W.a[4] = 0.0;
        W.a[5] = -w0;
        W.a[6] = -w1;
        W.a[7] = w0;
        W.a[8] = 0.0;
        dR = mul(e.R, W);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) out.a[i] -= 2.0 * mu * dR.a[i];
    return out;
}

}  // namespace dmpm
