// The phases of one ForwardTimeStep as separate FP64 kernels over host-layout
// (AoS) arrays: the reference's SingleThreadMPM phase API
// (ForwardSimulation.h:23-45, paths relative to the reference root), for
// callers that run the phases one by one on their own particle and grid
// containers.  The fused episode kernels (dmpm_kernels.hip, dmpm_blocks.hip)
// are the fast path; these follow the reference's phase boundaries instead:
// P2G reads the particle's stored stress P (P_op_1's output) and accumulates
// into the grid it is given, G_op leaves the velocity of nodes at or below
// the mass threshold as it was, G2P and P_op_2 are separate passes.
//
// Layouts: particles x, v [n][3]; C, F, dFc, P [n][3][3] row-major; grid
// masses [G], momenta and velocities [G][3], node (i, j, k) at
// i dy dz + j dz + k with position min_point + (i, j, k) dx.
#include "dmpm_kernels.h"
#include "dmpm_math.h"

namespace dmpm {

namespace {

constexpr int kPhaseBlock = 256;

inline unsigned phase_blocks(long long n) { return (unsigned)((n + kPhaseBlock - 1) / kPhaseBlock); }

__device__ __forceinline__ M3<double> ld9(const double* a, int p)
{
    M3<double> M;
#pragma unroll
    for (int i = 0; i < 9; ++i) M.a[i] = a[9 * (long long)p + i];
    return M;
}

__device__ __forceinline__ void stencil_of(const StepParams& g, const double x[3], Axis<double> ax[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) axis_weights<double>((x[a] - g.min_point[a]) * g.inv_dx, ax[a], false);
}

// SingleParticle_to_grid (ForwardSimulation.cpp:11-51):
//   m_i += w m,  p_i += w [m (1 - dt drag) v + (-(3/dx^2) dt vol P (F + dFc)^T + m C) d]
__global__ __launch_bounds__(kPhaseBlock) void k_phase_p2g(StepParams g, const double* x, const double* v, const double* C,
                                                          const double* F, const double* dFc, const double* P, const double* m,
                                                          const double* vol, double* gm, double* gp)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n) return;
    const double mp = m[p];
    const M3<double> Fc = [&] {
        M3<double> a = ld9(F, p), b = ld9(dFc, p);
#pragma unroll
        for (int i = 0; i < 9; ++i) a.a[i] += b.a[i];
        return a;
    }();
    const M3<double> PF = mult(ld9(P, p), Fc);  // P Fc^T
    const M3<double> Cp = ld9(C, p);
    const double k = -3.0 * g.inv_dx * g.inv_dx * g.dt * vol[p];
    M3<double> A;
#pragma unroll
    for (int i = 0; i < 9; ++i) A.a[i] = k * PF.a[i] + mp * Cp.a[i];
    double mv[3], xp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        xp[a] = x[3 * (long long)p + a];
        mv[a] = mp * v[3 * (long long)p + a] * (1.0 - g.dt * g.drag);
    }
    Axis<double> ax[3];
    stencil_of(g, xp, ax);
    for (int i = 0; i < 4; ++i) {
        const int gi = ax[0].base + i;
        if (gi < 0 || gi >= g.dim[0]) continue;
        const double d0 = (double(i - 1) - ax[0].f) * g.dx;
        for (int j = 0; j < 4; ++j) {
            const int gj = ax[1].base + j;
            if (gj < 0 || gj >= g.dim[1]) continue;
            const double d1 = (double(j - 1) - ax[1].f) * g.dx;
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = ax[2].base + kk;
                if (gk < 0 || gk >= g.dim[2]) continue;
                const double d2 = (double(kk - 1) - ax[2].f) * g.dx;
                const double w = ax[0].w[i] * ax[1].w[j] * ax[2].w[kk];
                const long long q = ((long long)gi * g.dim[1] + gj) * g.dim[2] + gk;
                unsafeAtomicAdd(gm + q, w * mp);
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    unsafeAtomicAdd(gp + 3 * q + a, w * (mv[a] + A.a[3 * a] * d0 + A.a[3 * a + 1] * d1 + A.a[3 * a + 2] * d2));
            }
        }
    }
}

// SingleNode_op (ForwardSimulation.cpp:53-58) over every node: v = p/m + dt f
// where |m| > FLT_EPSILON, other nodes keep v.  f is f_ext, or with
// gravity_mag != 0 the pull towards gravity_point, |f| = gravity_mag (0 at
// the point itself; ForwardSimulation.cpp:202-224).
__global__ __launch_bounds__(kPhaseBlock) void k_phase_grid_op(StepParams g, const double* gm, const double* gp, double* gv,
                                                              int point_mode, double mag)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.G) return;
    const double mq = gm[q];
    if (!(fabs(mq) > kMassEps)) return;
    double f[3] = {g.f_ext[0], g.f_ext[1], g.f_ext[2]};
    if (point_mode) {
        const long long i = q / ((long long)g.dim[1] * g.dim[2]);
        const long long j = (q / g.dim[2]) % g.dim[1];
        const long long k = q % g.dim[2];
        const double xn[3] = {g.min_point[0] + double(i) * g.dx, g.min_point[1] + double(j) * g.dx, g.min_point[2] + double(k) * g.dx};
        double r[3], nn = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            r[a] = g.f_ext[a] - xn[a];  // f_ext carries the gravity point here
            nn += r[a] * r[a];
        }
        const double s = nn > 0.0 ? mag / sqrt(nn) : 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) f[a] = r[a] * s;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) gv[3 * q + a] = gp[3 * q + a] / mq + g.dt * f[a];
}

// Grid_to_SingleParticle (ForwardSimulation.cpp:60-106): v' = sum w v_i,
// C' = (3/dx^2) sum w v_i d^T
__global__ __launch_bounds__(kPhaseBlock) void k_phase_g2p(StepParams g, const double* x, const double* gv, double* vn, double* Cn)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n) return;
    double xp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) xp[a] = x[3 * (long long)p + a];
    Axis<double> ax[3];
    stencil_of(g, xp, ax);
    const double D = 3.0 * g.inv_dx * g.inv_dx;
    double v[3] = {0.0, 0.0, 0.0};
    M3<double> Cp = m3_zero<double>();
    for (int i = 0; i < 4; ++i) {
        const int gi = ax[0].base + i;
        if (gi < 0 || gi >= g.dim[0]) continue;
        const double d0 = (double(i - 1) - ax[0].f) * g.dx;
        for (int j = 0; j < 4; ++j) {
            const int gj = ax[1].base + j;
            if (gj < 0 || gj >= g.dim[1]) continue;
            const double d1 = (double(j - 1) - ax[1].f) * g.dx;
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = ax[2].base + kk;
                if (gk < 0 || gk >= g.dim[2]) continue;
                const double d2 = (double(kk - 1) - ax[2].f) * g.dx;
                const double w = ax[0].w[i] * ax[1].w[j] * ax[2].w[kk];
                const long long q = ((long long)gi * g.dim[1] + gj) * g.dim[2] + gk;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double wv = w * gv[3 * q + a];
                    v[a] += wv;
                    Cp.a[3 * a] += D * wv * d0;
                    Cp.a[3 * a + 1] += D * wv * d1;
                    Cp.a[3 * a + 2] += D * wv * d2;
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) vn[3 * (long long)p + a] = v[a];
#pragma unroll
    for (int i = 0; i < 9; ++i) Cn[9 * (long long)p + i] = Cp.a[i];
}

// SingleParticle_op_2 (ForwardSimulation.cpp:108-124): F' = (I + dt C')(F + dFc),
// blended F' (1 - s) + F s, x' = x + dt v'
__global__ __launch_bounds__(kPhaseBlock) void k_phase_p_op2(int n, const double* x, const double* F, const double* dFc,
                                                            const double* vn, const double* Cn, double s, double dt, double* xn,
                                                            double* Fn)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    M3<double> IC = ld9(Cn, p), Fc = ld9(F, p);
    const M3<double> F0 = Fc, dF = ld9(dFc, p);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        IC.a[i] *= dt;
        Fc.a[i] += dF.a[i];
    }
    IC.a[0] += 1.0;
    IC.a[4] += 1.0;
    IC.a[8] += 1.0;
    const M3<double> Fp = mul(IC, Fc);
#pragma unroll
    for (int i = 0; i < 9; ++i) Fn[9 * (long long)p + i] = Fp.a[i] * (1.0 - s) + F0.a[i] * s;
#pragma unroll
    for (int a = 0; a < 3; ++a) xn[3 * (long long)p + a] = x[3 * (long long)p + a] + dt * vn[3 * (long long)p + a];
}

// G2G_Mass (ForwardSimulation.cpp:329-367): every node of grid 2 takes the
// mass sum w m_i of the grid-1 nodes around its position
__global__ __launch_bounds__(kPhaseBlock) void k_phase_g2g_mass(StepParams g1, StepParams g2, const double* m1, double* m2)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g2.G) return;
    const long long i = q / ((long long)g2.dim[1] * g2.dim[2]);
    const long long j = (q / g2.dim[2]) % g2.dim[1];
    const long long k = q % g2.dim[2];
    const double xp[3] = {g2.min_point[0] + double(i) * g2.dx, g2.min_point[1] + double(j) * g2.dx,
                          g2.min_point[2] + double(k) * g2.dx};
    Axis<double> ax[3];
    stencil_of(g1, xp, ax);
    double acc = 0.0;
    for (int a = 0; a < 4; ++a) {
        const int gi = ax[0].base + a;
        if (gi < 0 || gi >= g1.dim[0]) continue;
        for (int b = 0; b < 4; ++b) {
            const int gj = ax[1].base + b;
            if (gj < 0 || gj >= g1.dim[1]) continue;
            for (int c = 0; c < 4; ++c) {
                const int gk = ax[2].base + c;
                if (gk < 0 || gk >= g1.dim[2]) continue;
                acc += ax[0].w[a] * ax[1].w[b] * ax[2].w[c] * m1[((long long)gi * g1.dim[1] + gj) * g1.dim[2] + gk];
            }
        }
    }
    m2[q] = acc;
}

}  // namespace

void launch_phase_p2g(const StepParams& g, const double* x, const double* v, const double* C, const double* F, const double* dFc,
                      const double* P, const double* m, const double* vol, double* gm, double* gp, hipStream_t s)
{
    if (g.n <= 0) return;
    hipLaunchKernelGGL(k_phase_p2g, dim3(phase_blocks(g.n)), dim3(kPhaseBlock), 0, s, g, x, v, C, F, dFc, P, m, vol, gm, gp);
}

void launch_phase_grid_op(const StepParams& g, const double* gm, const double* gp, double* gv, int point_mode, double mag,
                          hipStream_t s)
{
    if (g.G <= 0) return;
    hipLaunchKernelGGL(k_phase_grid_op, dim3(phase_blocks(g.G)), dim3(kPhaseBlock), 0, s, g, gm, gp, gv, point_mode, mag);
}

void launch_phase_g2p(const StepParams& g, const double* x, const double* gv, double* vn, double* Cn, hipStream_t s)
{
    if (g.n <= 0) return;
    hipLaunchKernelGGL(k_phase_g2p, dim3(phase_blocks(g.n)), dim3(kPhaseBlock), 0, s, g, x, gv, vn, Cn);
}

void launch_phase_p_op2(int n, const double* x, const double* F, const double* dFc, const double* vn, const double* Cn, double sm,
                        double dt, double* xn, double* Fn, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_phase_p_op2, dim3(phase_blocks(n)), dim3(kPhaseBlock), 0, s, n, x, F, dFc, vn, Cn, sm, dt, xn, Fn);
}

void launch_phase_g2g_mass(const StepParams& g1, const StepParams& g2, const double* m1, double* m2, hipStream_t s)
{
    if (g2.G <= 0) return;
    hipLaunchKernelGGL(k_phase_g2g_mass, dim3(phase_blocks(g2.G)), dim3(kPhaseBlock), 0, s, g1, g2, m1, m2);
}

}  // namespace dmpm
