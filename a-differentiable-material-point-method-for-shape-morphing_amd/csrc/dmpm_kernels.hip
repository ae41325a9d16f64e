// HIP kernels of the MLS-MPM step and its adjoint for gfx950.
//
// One thread per particle (or per node); particle state is structure-of-
// arrays so every per-particle load and store is a coalesced 64-lane access.
// Scatters use hardware FP atomics into the dense grid.  Templated on the
// state precision: Real = double is the plain FP64 path, Real = float the
// FP32 path; both evaluate the polar/SVD, PK1 and H:dF in FP64 and reduce
// the loss in FP64.
//
// Phase order and formulas (paths relative to the reference root):
//   k_p2g        P_op_1 + P2G           ForwardSimulation.cpp:6-51,159-182
//   k_grid_update G_op                  ForwardSimulation.cpp:53-58,184-200
//   k_g2p        G2P + P_op_2           ForwardSimulation.cpp:60-124,226-259
//   k_loss_*     EndLayerMassLoss       CompGraph.cpp:110-244
//   k_g2p_adj    P_op_2^T + G2P^T       BackPropagation.cpp:14-44
//   k_grid_adj   G_op^T                 BackPropagation.cpp:46-59
//   k_p2g_adj    P2G^T + P_op_1^T       BackPropagation.cpp:61-124
#include "dmpm_kernels.h"
#include "dmpm_math.h"

namespace dmpm {

constexpr int kBlock = 256;

__host__ __device__ inline unsigned nblocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename Real>
__device__ __forceinline__ void atomic_add(Real* p, Real v)
{
    unsafeAtomicAdd(p, v);
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_fill(Real* p, long long count, Real value)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long step = (long long)gridDim.x * blockDim.x;
    for (; i < count; i += step) p[i] = value;
}

// 16-byte stores where p is 16-byte aligned (every buffer of this library is)
__global__ __launch_bounds__(kBlock) void k_fill4(float4* p, long long count4, float value)
{
    const float4 v = make_float4(value, value, value, value);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count4; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

template <typename Real>
void launch_fill(Real* p, long long count, Real value, hipStream_t s)
{
    if (count <= 0) return;
    if (sizeof(Real) == 4 && ((uintptr_t)p & 15) == 0 && (count & 3) == 0) {
        const long long c4 = count / 4;
        unsigned nb = nblocks(c4, kBlock);
        if (nb > 8192) nb = 8192;
        hipLaunchKernelGGL(k_fill4, dim3(nb), dim3(kBlock), 0, s, reinterpret_cast<float4*>(p), c4, float(value));
        return;
    }
    unsigned nb = nblocks(count, kBlock);
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_fill<Real>, dim3(nb), dim3(kBlock), 0, s, p, count, value);
}

// Stencil of one particle: per-axis weights plus the node index base.
template <typename Real>
struct Stencil {
    Axis<Real> ax[3];
};

template <typename Real>
__device__ __forceinline__ void make_stencil(const StepParams& sp, const Real x[3], Stencil<Real>& st, bool slopes)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) axis_weights<Real>((x[a] - Real(sp.min_point[a])) * Real(sp.inv_dx), st.ax[a], slopes);
}

template <typename Real>
__device__ __forceinline__ void load_vec3(const Real* base, int field, int S, int p, Real out[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = base[(long long)(field + a) * S + p];
}

template <typename Real>
__device__ __forceinline__ M3<Real> load_mat(const Real* base, int field, int S, int p)
{
    M3<Real> M;
#pragma unroll
    for (int i = 0; i < 9; ++i) M.a[i] = base[(long long)(field + i) * S + p];
    return M;
}

template <typename Real>
__device__ __forceinline__ void store_mat(Real* base, int field, int S, int p, const M3<Real>& M)
{
#pragma unroll
    for (int i = 0; i < 9; ++i) base[(long long)(field + i) * S + p] = M.a[i];
}

// ---------------------------------------------------------------- P2G ------
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_p2g(StepParams sp, StepIO<Real> io)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sp.n) return;
    const int S = sp.stride;
    Real x[3], v[3];
    load_vec3(io.cur, FX, S, p, x);
    load_vec3(io.cur, FV, S, p, v);
    M3<Real> C = load_mat(io.cur, FC, S, p);
    M3<Real> F = load_mat(io.cur, FF, S, p);
    M3<Real> dF = load_mat(io.ctrl, 0, S, p);
    const Real m = io.cst[KM * S + p];
    const Real vol = io.cst[KVOL * S + p];
    const double lam = io.cst[KLAM * S + p], mu = io.cst[KMU * S + p];

    M3<double> Fc;
#pragma unroll
    for (int i = 0; i < 9; ++i) Fc.a[i] = double(F.a[i]) + double(dF.a[i]);
    FCE e;
    fce_prepare(Fc, e);
    M3<double> P = fce_pk1(Fc, e, lam, mu);
    M3<double> PF = mult(P, Fc);
    const double k = -3.0 * sp.inv_dx * sp.inv_dx * sp.dt * double(vol);
    M3<Real> A;
#pragma unroll
    for (int i = 0; i < 9; ++i) A.a[i] = Real(k * PF.a[i]) + m * C.a[i];
    const Real sdrag = Real(1) - Real(sp.dt) * Real(sp.drag);
    Real mv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) mv[a] = m * v[a] * sdrag;

    Stencil<Real> st;
    make_stencil(sp, x, st, false);
    const Real dx = Real(sp.dx);
    Real* gm = io.grid;
    Real* gp = io.grid + sp.G;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st.ax[0].base + i;
        if (gi < 0 || gi >= sp.dim[0]) continue;
        const Real d0 = (Real(i - 1) - st.ax[0].f) * dx;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st.ax[1].base + j;
            if (gj < 0 || gj >= sp.dim[1]) continue;
            const Real d1 = (Real(j - 1) - st.ax[1].f) * dx;
            const Real wij = st.ax[0].w[i] * st.ax[1].w[j];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st.ax[2].base + kk;
                if (gk < 0 || gk >= sp.dim[2]) continue;
                const Real d2 = (Real(kk - 1) - st.ax[2].f) * dx;
                const Real w = wij * st.ax[2].w[kk];
                const long long node = ((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk;
                atomic_add(gm + node, w * m);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const Real val = w * (mv[a] + A.a[3 * a] * d0 + A.a[3 * a + 1] * d1 + A.a[3 * a + 2] * d2);
                    atomic_add(gp + a * sp.G + node, val);
                }
            }
        }
    }
}

template <typename Real>
void launch_p2g(const StepParams& sp, const StepIO<Real>& io, hipStream_t s)
{
    if (sp.n <= 0) return;
    hipLaunchKernelGGL(k_p2g<Real>, dim3(nblocks(sp.n, kBlock)), dim3(kBlock), 0, s, sp, io);
}

// ------------------------------------------------------------ grid update --
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_grid_update(StepParams sp, Real* grid)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sp.G) return;
    const Real m = grid[i];
    Real* g = grid + sp.G;
    if (fabs(double(m)) > kMassEps) {
        const Real inv = Real(1) / m;
#pragma unroll
        for (int a = 0; a < 3; ++a) g[a * sp.G + i] = g[a * sp.G + i] * inv + Real(sp.dt) * Real(sp.f_ext[a]);
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) g[a * sp.G + i] = Real(0);
    }
}

template <typename Real>
void launch_grid_update(const StepParams& sp, Real* grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_grid_update<Real>, dim3(nblocks(sp.G, kBlock)), dim3(kBlock), 0, s, sp, grid);
}

// ---------------------------------------------------------------- G2P ------
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_g2p(StepParams sp, StepIO<Real> io)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sp.n) return;
    const int S = sp.stride;
    Real x[3];
    load_vec3(io.cur, FX, S, p, x);
    Stencil<Real> st;
    make_stencil(sp, x, st, false);
    const Real* gv = io.grid + sp.G;
    const Real Dd = Real(3.0 * sp.inv_dx);  // D^-1 d = (3/dx)(o - 1 - f)
    Real vp[3] = {0, 0, 0};
    M3<Real> Cp = m3_zero<Real>();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st.ax[0].base + i;
        if (gi < 0 || gi >= sp.dim[0]) continue;
        const Real e0 = (Real(i - 1) - st.ax[0].f) * Dd;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st.ax[1].base + j;
            if (gj < 0 || gj >= sp.dim[1]) continue;
            const Real e1 = (Real(j - 1) - st.ax[1].f) * Dd;
            const Real wij = st.ax[0].w[i] * st.ax[1].w[j];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st.ax[2].base + kk;
                if (gk < 0 || gk >= sp.dim[2]) continue;
                const Real e2 = (Real(kk - 1) - st.ax[2].f) * Dd;
                const Real w = wij * st.ax[2].w[kk];
                const long long node = ((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const Real wv = w * gv[a * sp.G + node];
                    vp[a] += wv;
                    Cp.a[3 * a] += wv * e0;
                    Cp.a[3 * a + 1] += wv * e1;
                    Cp.a[3 * a + 2] += wv * e2;
                }
            }
        }
    }
    M3<Real> F = load_mat(io.cur, FF, S, p);
    M3<Real> dF = load_mat(io.ctrl, 0, S, p);
    M3<Real> Fc, IC;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Fc.a[i] = F.a[i] + dF.a[i];
        IC.a[i] = Real(sp.dt) * Cp.a[i];
    }
    IC.a[0] += Real(1);
    IC.a[4] += Real(1);
    IC.a[8] += Real(1);
    M3<Real> Fn = mul(IC, Fc);
    const Real s = Real(sp.smoothing);
#pragma unroll
    for (int i = 0; i < 9; ++i) Fn.a[i] = Fn.a[i] * (Real(1) - s) + F.a[i] * s;
    Real* nx = io.next;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        nx[(long long)(FX + a) * S + p] = x[a] + Real(sp.dt) * vp[a];
        nx[(long long)(FV + a) * S + p] = vp[a];
    }
    store_mat(nx, FC, S, p, Cp);
    store_mat(nx, FF, S, p, Fn);
}

template <typename Real>
void launch_g2p(const StepParams& sp, const StepIO<Real>& io, hipStream_t s)
{
    if (sp.n <= 0) return;
    hipLaunchKernelGGL(k_g2p<Real>, dim3(nblocks(sp.n, kBlock)), dim3(kBlock), 0, s, sp, io);
}

// ------------------------------------------------------- P2G mass / vol ----
// one point's mass m at x into the node masses gm (HBM atomics)
template <typename Real>
__device__ __forceinline__ void p2g_mass_point(const StepParams& sp, const Real x[3], Real m, Real* gm)
{
    Stencil<Real> st;
    make_stencil(sp, x, st, false);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st.ax[0].base + i;
        if (gi < 0 || gi >= sp.dim[0]) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st.ax[1].base + j;
            if (gj < 0 || gj >= sp.dim[1]) continue;
            const Real wij = st.ax[0].w[i] * st.ax[1].w[j] * m;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st.ax[2].base + kk;
                if (gk < 0 || gk >= sp.dim[2]) continue;
                atomic_add(gm + ((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk, wij * st.ax[2].w[kk]);
            }
        }
    }
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_p2g_mass(StepParams sp, const Real* xs, const Real* ms, int S, Real* gm)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sp.n) return;
    Real x[3];
    load_vec3(xs, 0, S, p, x);
    p2g_mass_point<Real>(sp, x, ms[p], gm);
}

template <typename Real>
void launch_p2g_mass(const StepParams& sp, const Real* x, const Real* m, int stride, Real* gm, hipStream_t s)
{
    if (sp.n <= 0) return;
    hipLaunchKernelGGL(k_p2g_mass<Real>, dim3(nblocks(sp.n, kBlock)), dim3(kBlock), 0, s, sp, x, m, stride, gm);
}

// P2G mass of host-layout points: AoS double positions x[3p + a] (shifted by
// -org[a] and rounded to Real, as the host staging of the SoA kernel does) and
// double masses m[p]
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_p2g_mass_aos(StepParams sp, const double* xa, const double* ma, double o0, double o1,
                                                         double o2, Real* gm)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sp.n) return;
    const Real x[3] = {Real(xa[3 * (long long)p] - o0), Real(xa[3 * (long long)p + 1] - o1), Real(xa[3 * (long long)p + 2] - o2)};
    p2g_mass_point<Real>(sp, x, Real(ma[p]), gm);
}

template <typename Real>
void launch_p2g_mass_aos(const StepParams& sp, const double* x, const double* m, const double org[3], Real* gm, hipStream_t s)
{
    if (sp.n <= 0) return;
    hipLaunchKernelGGL(k_p2g_mass_aos<Real>, dim3(nblocks(sp.n, kBlock)), dim3(kBlock), 0, s, sp, x, m, org[0], org[1], org[2], gm);
}

// ForwardSimulation.cpp:262-290: vol = m / sum_i w m_i
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_volumes(StepParams sp, const Real* xs, Real* cst, const Real* gm)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sp.n) return;
    const int S = sp.stride;
    Real x[3];
    load_vec3(xs, FX, S, p, x);
    Stencil<Real> st;
    make_stencil(sp, x, st, false);
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st.ax[0].base + i;
        if (gi < 0 || gi >= sp.dim[0]) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st.ax[1].base + j;
            if (gj < 0 || gj >= sp.dim[1]) continue;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st.ax[2].base + kk;
                if (gk < 0 || gk >= sp.dim[2]) continue;
                const Real w = st.ax[0].w[i] * st.ax[1].w[j] * st.ax[2].w[kk];
                acc += double(w) * double(gm[((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk]);
            }
        }
    }
    const Real m = cst[KM * S + p];
    cst[KVOL * S + p] = (fabs(acc) > kMassEps) ? Real(double(m) / acc) : Real(0);
}

template <typename Real>
void launch_volumes(const StepParams& sp, const Real* x, Real* cst, const Real* gm, hipStream_t s)
{
    if (sp.n <= 0) return;
    hipLaunchKernelGGL(k_volumes<Real>, dim3(nblocks(sp.n, kBlock)), dim3(kBlock), 0, s, sp, x, cst, gm);
}

// ------------------------------------------------------------------ loss ---
// CompGraph.cpp:143-187 in FP64: c = ln(m+1+eps) - ln(t+1+eps),
// l = c^2/2 + [m < 1e-3] 10 (1e-3 - m)^2, dL/dm = c/(m+1+eps) - [m < 1e-3] 20 (1e-3 - m).
// dL/dm is kept in FP64: at small node masses every node sits under the
// soft minimum, dL/dm ~ -0.02 + O(m), and the particle gather below relies on
// sum_i grad w = 0 to cancel the constant part exactly.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_loss(StepParams sp, const Real* gm, const Real* target, double* dLdm, double* part)
{
    __shared__ double red[kBlock / 64];
    double local = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < sp.G; i += (long long)gridDim.x * blockDim.x) {
        const double cm = gm[i], tm = target[i];
        const double c = log(cm + 1.0 + 1e-4) - log(tm + 1.0 + 1e-4);
        double l = 0.5 * c * c;
        double g = c / (cm + 1.0 + 1e-4);
        if (cm < 1e-3) {
            const double diff = 1e-3 - cm;
            l += 10.0 * diff * diff;
            g -= 20.0 * diff;
        }
        local += l;
        dLdm[i] = g;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// Row sums of partials in a fixed order (one workgroup per row): the loss is
// the same bit for bit whenever the node masses are.
__global__ __launch_bounds__(kBlock) void k_sum_rows(const double* part, int len, double* out)
{
    __shared__ double red[kBlock / 64];
    const double* row = part + (long long)blockIdx.x * len;
    double local = 0.0;
    for (int i = threadIdx.x; i < len; i += blockDim.x) local += row[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) s += red[w];
        out[blockIdx.x] = s;
    }
}

void launch_sum_rows(const double* part, int rows, int len, double* out, hipStream_t s)
{
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_sum_rows, dim3(rows), dim3(kBlock), 0, s, part, len, out);
}

template <typename Real>
void launch_loss(const StepParams& sp, const Real* gm, const Real* target, double* dLdm, double* part, double* loss, hipStream_t s)
{
    unsigned nb = nblocks(sp.G, kBlock);
    if (nb > kLossBlocks) nb = kLossBlocks;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_loss<Real>, dim3(nb), dim3(kBlock), 0, s, sp, gm, target, dLdm, part);
    launch_sum_rows(part, 1, (int)nb, loss, s);
}

// CompGraph.cpp:192-241: dLdx = sum_i m dLdm_i grad w; dLdv = dLdF = dLdC = 0.
// Weights and the sum in FP64 for both precisions (a gradient reduction).
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_loss_particles(StepParams sp, const Real* layer, const Real* cst, const double* dLdm, Real* g)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sp.n) return;
    const int S = sp.stride;
    double x[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) x[a] = double(layer[(long long)(FX + a) * S + p]);
    Stencil<double> st;
    make_stencil<double>(sp, x, st, true);
    double gx[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st.ax[0].base + i;
        if (gi < 0 || gi >= sp.dim[0]) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st.ax[1].base + j;
            if (gj < 0 || gj >= sp.dim[1]) continue;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st.ax[2].base + kk;
                if (gk < 0 || gk >= sp.dim[2]) continue;
                const double dl = dLdm[((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk];
                gx[0] += dl * st.ax[0].dw[i] * st.ax[1].w[j] * st.ax[2].w[kk];
                gx[1] += dl * st.ax[0].w[i] * st.ax[1].dw[j] * st.ax[2].w[kk];
                gx[2] += dl * st.ax[0].w[i] * st.ax[1].w[j] * st.ax[2].dw[kk];
            }
        }
    }
    const double sc = -double(cst[KM * S + p]) * sp.inv_dx;
#pragma unroll
    for (int a = 0; a < 3; ++a) g[(long long)(FX + a) * S + p] = Real(sc * gx[a]);
#pragma unroll
    for (int f = FV; f < kLayerFields; ++f) g[(long long)f * S + p] = Real(0);
}

template <typename Real>
void launch_loss_particles(const StepParams& sp, const Real* layer, const Real* cst, const double* dLdm, Real* g, hipStream_t s)
{
    if (sp.n <= 0) return;
    hipLaunchKernelGGL(k_loss_particles<Real>, dim3(nblocks(sp.n, kBlock)), dim3(kBlock), 0, s, sp, layer, cst, dLdm, g);
}

// --------------------------------------------------------------- adjoint ---
// P_op_2^T and the G2P scatter: vt = v' + dt x', Ct = C' + (1-s) dt F' Fc^T,
// vbar_i += w (vt + D Ct d).  vt and Ct are kept for the gather kernel.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_g2p_adj(StepParams sp, AdjIO<Real> io)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sp.n) return;
    const int S = sp.stride;
    Real x[3], gx1[3], gv1[3];
    load_vec3(io.cur, FX, S, p, x);
    load_vec3(io.g1, FX, S, p, gx1);
    load_vec3(io.g1, FV, S, p, gv1);
    M3<Real> F = load_mat(io.cur, FF, S, p);
    M3<Real> dF = load_mat(io.ctrl, 0, S, p);
    M3<Real> gF1 = load_mat(io.g1, FF, S, p);
    M3<Real> gC1 = load_mat(io.g1, FC, S, p);
    M3<Real> Fc;
#pragma unroll
    for (int i = 0; i < 9; ++i) Fc.a[i] = F.a[i] + dF.a[i];
    const Real cs = (sp.compat & 2) ? Real(1) : Real(1) - Real(sp.smoothing);
    M3<Real> FF_ = mult(gF1, Fc);
    M3<Real> Ct;
    Real vt[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) Ct.a[i] = gC1.a[i] + cs * Real(sp.dt) * FF_.a[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) vt[a] = gv1[a] + Real(sp.dt) * gx1[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) io.tmp[(long long)a * S + p] = vt[a];
    store_mat(io.tmp, 3, S, p, Ct);

    Stencil<Real> st;
    make_stencil(sp, x, st, false);
    const Real Dd = Real(3.0 * sp.inv_dx);
    Real* vb = io.agrid;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st.ax[0].base + i;
        if (gi < 0 || gi >= sp.dim[0]) continue;
        const Real e0 = (Real(i - 1) - st.ax[0].f) * Dd;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st.ax[1].base + j;
            if (gj < 0 || gj >= sp.dim[1]) continue;
            const Real e1 = (Real(j - 1) - st.ax[1].f) * Dd;
            const Real wij = st.ax[0].w[i] * st.ax[1].w[j];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st.ax[2].base + kk;
                if (gk < 0 || gk >= sp.dim[2]) continue;
                const Real e2 = (Real(kk - 1) - st.ax[2].f) * Dd;
                const Real w = wij * st.ax[2].w[kk];
                const long long node = ((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    atomic_add(vb + a * sp.G + node, w * (vt[a] + Ct.a[3 * a] * e0 + Ct.a[3 * a + 1] * e1 + Ct.a[3 * a + 2] * e2));
            }
        }
    }
}

template <typename Real>
void launch_g2p_adj(const StepParams& sp, const AdjIO<Real>& io, hipStream_t s)
{
    if (sp.n <= 0) return;
    hipLaunchKernelGGL(k_g2p_adj<Real>, dim3(nblocks(sp.n, kBlock)), dim3(kBlock), 0, s, sp, io);
}

// G_op^T: pbar = vbar/m, mbar = -((v - dt f) . vbar)/m  (compat: v . vbar)
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_grid_adj(StepParams sp, AdjIO<Real> io)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sp.G) return;
    const Real m = io.grid[i];
    Real* vb = io.agrid;
    Real* mb = io.agrid + 3 * sp.G;
    if (fabs(double(m)) > kMassEps) {
        const Real inv = Real(1) / m;
        Real dot = Real(0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            Real vi = io.grid[(1 + a) * sp.G + i];
            if (!(sp.compat & 8)) vi -= Real(sp.dt) * Real(sp.f_ext[a]);
            const Real b = vb[a * sp.G + i];
            dot += vi * b;
            vb[a * sp.G + i] = b * inv;
        }
        mb[i] = -dot * inv;
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) vb[a * sp.G + i] = Real(0);
        mb[i] = Real(0);
    }
}

template <typename Real>
void launch_grid_adj(const StepParams& sp, const AdjIO<Real>& io, hipStream_t s)
{
    hipLaunchKernelGGL(k_grid_adj<Real>, dim3(nblocks(sp.G, kBlock)), dim3(kBlock), 0, s, sp, io);
}

// P2G^T gather + P_op_1^T.  With G = -k P Fc^T + m C, b = m(1 - dt drag) v:
//   Gb = sum w pbar d^T, Cbar = m Gb, Pbar = -k Gb Fc,
//   Fcbar = (1-s)(I + dt C')^T F' - k Gb^T P + H(Fc):Pbar,
//   xbar = x' + sum [m mbar gw + gw (b + G d).pbar - w G^T pbar + gw (v_i.vt)
//                    + D((Ct : v_i d^T) gw - w Ct^T v_i)],
//   vbar = sum w m (1 - dt drag) pbar.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_p2g_adj(StepParams sp, AdjIO<Real> io)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= sp.n) return;
    const int S = sp.stride;
    const int cp = sp.compat;
    Real x[3], v[3];
    load_vec3(io.cur, FX, S, p, x);
    load_vec3(io.cur, FV, S, p, v);
    M3<Real> C = load_mat(io.cur, FC, S, p);
    M3<Real> F = load_mat(io.cur, FF, S, p);
    M3<Real> dF = load_mat(io.ctrl, 0, S, p);
    const Real m = io.cst[KM * S + p];
    const double vol = io.cst[KVOL * S + p];
    const double lam = io.cst[KLAM * S + p], mu = io.cst[KMU * S + p];

    M3<double> Fc;
#pragma unroll
    for (int i = 0; i < 9; ++i) Fc.a[i] = double(F.a[i]) + double(dF.a[i]);
    FCE e;
    fce_prepare(Fc, e);
    M3<double> P = fce_pk1(Fc, e, lam, mu);
    const double kk_ = 3.0 * sp.inv_dx * sp.inv_dx * sp.dt * vol;
    M3<double> PF = mult(P, Fc);
    M3<Real> Gm;
#pragma unroll
    for (int i = 0; i < 9; ++i) Gm.a[i] = Real(-kk_ * PF.a[i]) + m * C.a[i];
    const Real sdrag = Real(1) - Real(sp.dt) * Real(sp.drag);
    const Real bsc = (cp & 4) ? Real(1) : sdrag;
    Real b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) b[a] = m * bsc * v[a];
    Real vt[3];
    load_vec3(io.tmp, 0, S, p, vt);
    M3<Real> Ct = load_mat(io.tmp, 3, S, p);

    Stencil<Real> st;
    make_stencil(sp, x, st, true);
    const Real dx = Real(sp.dx);
    const Real Dinv = Real(3.0 * sp.inv_dx * sp.inv_dx);
    const Real mgw = -Real(sp.inv_dx);
    const Real* gv = io.grid + sp.G;
    const Real* pb = io.agrid;
    const Real* mb = io.agrid + 3 * sp.G;
    M3<Real> Gb = m3_zero<Real>();
    Real xb[3] = {0, 0, 0}, pacc[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st.ax[0].base + i;
        if (gi < 0 || gi >= sp.dim[0]) continue;
        const Real d0 = (Real(i - 1) - st.ax[0].f) * dx;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st.ax[1].base + j;
            if (gj < 0 || gj >= sp.dim[1]) continue;
            const Real d1 = (Real(j - 1) - st.ax[1].f) * dx;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st.ax[2].base + kk;
                if (gk < 0 || gk >= sp.dim[2]) continue;
                const Real d2 = (Real(kk - 1) - st.ax[2].f) * dx;
                const Real w0 = st.ax[0].w[i], w1 = st.ax[1].w[j], w2 = st.ax[2].w[kk];
                const Real w = w0 * w1 * w2;
                const Real gw[3] = {mgw * st.ax[0].dw[i] * w1 * w2, mgw * w0 * st.ax[1].dw[j] * w2, mgw * w0 * w1 * st.ax[2].dw[kk]};
                const long long node = ((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk;
                Real pbi[3], vi[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    pbi[a] = pb[a * sp.G + node];
                    vi[a] = gv[a * sp.G + node];
                }
                const Real mbi = mb[node];
                const Real d[3] = {d0, d1, d2};
                Real bgd = Real(0), vdot = Real(0), cvd = Real(0);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const Real gd = Gm.a[3 * a] * d0 + Gm.a[3 * a + 1] * d1 + Gm.a[3 * a + 2] * d2;
                    bgd += (b[a] + gd) * pbi[a];
                    vdot += vi[a] * vt[a];
                    cvd += vi[a] * (Ct.a[3 * a] * d0 + Ct.a[3 * a + 1] * d1 + Ct.a[3 * a + 2] * d2);
                    pacc[a] += w * pbi[a];
#pragma unroll
                    for (int c = 0; c < 3; ++c) Gb.a[3 * a + c] += w * pbi[a] * d[c];
                }
                const Real sgw = m * mbi + bgd + vdot + Dinv * cvd;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const Real GTp = Gm.a[a] * pbi[0] + Gm.a[3 + a] * pbi[1] + Gm.a[6 + a] * pbi[2];
                    const Real CTv = Ct.a[a] * vi[0] + Ct.a[3 + a] * vi[1] + Ct.a[6 + a] * vi[2];
                    xb[a] += gw[a] * sgw - w * (GTp + Dinv * CTv);
                }
            }
        }
    }
    // Pbar = -k Gb Fc (FP64), stress term -k Gb^T P, H(Fc):Pbar
    M3<double> Gbd = cast<double>(Gb);
    M3<double> Pb = mul(Gbd, Fc);
#pragma unroll
    for (int i = 0; i < 9; ++i) Pb.a[i] *= -kk_;
    M3<double> H = fce_hess_mult(e, lam, mu, Pb, cp == 0);
    M3<double> GTP = tmul(Gbd, P);
    M3<Real> C1 = load_mat(io.next, FC, S, p);
    M3<Real> gF1 = load_mat(io.g1, FF, S, p);
    M3<Real> IC;
#pragma unroll
    for (int i = 0; i < 9; ++i) IC.a[i] = Real(sp.dt) * C1.a[i];
    IC.a[0] += Real(1);
    IC.a[4] += Real(1);
    IC.a[8] += Real(1);
    M3<Real> base = tmul(IC, gF1);
    const Real s = Real(sp.smoothing);
    M3<Real> gF, gdFc;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const Real rest = Real(-kk_ * GTP.a[i] + H.a[i]);
        if (cp & 1) {
            const Real f = (Real(1) - s) * base.a[i] + s * gF1.a[i] + base.a[i] + rest;
            gF.a[i] = f;
            gdFc.a[i] = f;
        } else {
            gdFc.a[i] = (Real(1) - s) * base.a[i] + rest;
            gF.a[i] = gdFc.a[i] + s * gF1.a[i];
        }
    }
    Real gx1[3];
    load_vec3(io.g1, FX, S, p, gx1);
    const Real vsc = m * sdrag;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        io.g0[(long long)(FX + a) * S + p] = gx1[a] + xb[a];
        io.g0[(long long)(FV + a) * S + p] = vsc * pacc[a];
    }
    M3<Real> gC;
#pragma unroll
    for (int i = 0; i < 9; ++i) gC.a[i] = m * Gb.a[i];
    store_mat(io.g0, FC, S, p, gC);
    store_mat(io.g0, FF, S, p, gF);
    if (io.gctrl) store_mat(io.gctrl, 0, S, p, gdFc);
}

template <typename Real>
void launch_p2g_adj(const StepParams& sp, const AdjIO<Real>& io, hipStream_t s)
{
    if (sp.n <= 0) return;
    hipLaunchKernelGGL(k_p2g_adj<Real>, dim3(nblocks(sp.n, kBlock)), dim3(kBlock), 0, s, sp, io);
}

// ------------------------------------------------- batched constitutive ----
__global__ __launch_bounds__(kBlock) void k_eval_pk1(int n, const double* F, const double* lam, const double* mu, double* P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    M3<double> A;
#pragma unroll
    for (int i = 0; i < 9; ++i) A.a[i] = F[9 * p + i];
    FCE e;
    fce_prepare(A, e);
    M3<double> R = fce_pk1(A, e, lam[p], mu[p]);
#pragma unroll
    for (int i = 0; i < 9; ++i) P[9 * p + i] = R.a[i];
}

__global__ __launch_bounds__(kBlock) void k_eval_hess(int n, const double* F, const double* lam, const double* mu,
                                                      const double* dF, double* out, int exact)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    M3<double> A, D;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        A.a[i] = F[9 * p + i];
        D.a[i] = dF[9 * p + i];
    }
    FCE e;
    fce_prepare(A, e);
    M3<double> R = fce_hess_mult(e, lam[p], mu[p], D, exact != 0);
#pragma unroll
    for (int i = 0; i < 9; ++i) out[9 * p + i] = R.a[i];
}

__global__ __launch_bounds__(kBlock) void k_eval_svd(int n, const double* F, double* U, double* sig, double* V)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    M3<double> A, Uo, Vo;
#pragma unroll
    for (int i = 0; i < 9; ++i) A.a[i] = F[9 * p + i];
    double s[3];
    svd3(A, Uo, s, Vo);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        U[9 * p + i] = Uo.a[i];
        V[9 * p + i] = Vo.a[i];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) sig[3 * p + a] = s[a];
}

void launch_eval_pk1(int n, const double* F, const double* lam, const double* mu, double* P, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eval_pk1, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, s, n, F, lam, mu, P);
}

void launch_eval_hess(int n, const double* F, const double* lam, const double* mu, const double* dF, double* out,
                      int exact, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eval_hess, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, s, n, F, lam, mu, dF, out, exact);
}

void launch_eval_svd(int n, const double* F, double* U, double* sig, double* V, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eval_svd, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, s, n, F, U, sig, V);
}

// Elasticity.cpp:74-86 for a batch
__global__ __launch_bounds__(kBlock) void k_eval_energy(int n, const double* F, const double* lam, const double* mu, double* psi)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    M3<double> A;
#pragma unroll
    for (int i = 0; i < 9; ++i) A.a[i] = F[9 * p + i];
    FCE e;
    fce_prepare(A, e);
    psi[p] = fce_energy(e, lam[p], mu[p]);
}

void launch_eval_energy(int n, const double* F, const double* lam, const double* mu, double* psi, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eval_energy, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, s, n, F, lam, mu, psi);
}

// PointCloud::GetPointDeterminants / GetPointElasticEnergies (PointCloud.cpp:441-467)
// over the S slots of one layer: Fe = F + dFc in FP64 from the stored values,
// out[p] = det Fe, out[S + p] = psi(Fe) (Elasticity.cpp:74-86).  ctrl may be
// null (a layer without a control: dFc = 0).
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_point_diag(long long S, const Real* layer, const Real* ctrl, const Real* cst, double* out)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= S) return;
    M3<double> A;
#pragma unroll
    for (int i = 0; i < 9; ++i)
        A.a[i] = double(layer[(long long)(FF + i) * S + p]) + (ctrl ? double(ctrl[(long long)i * S + p]) : 0.0);
    FCE e;
    fce_prepare(A, e);
    out[p] = e.Ju;
    out[S + p] = fce_energy(e, double(cst[(long long)KLAM * S + p]), double(cst[(long long)KMU * S + p]));
}

template <typename Real>
void launch_point_diag(long long S, const Real* layer, const Real* ctrl, const Real* cst, double* out, hipStream_t s)
{
    if (S <= 0) return;
    hipLaunchKernelGGL(k_point_diag<Real>, dim3(nblocks(S, kBlock)), dim3(kBlock), 0, s, S, layer, ctrl, cst, out);
}

// ---------------------------------------------------- instantiations -------
// ------------------------------------------------------------ optimizer ----
// PointCloud::Descend_* (PointCloud.cpp:83-179) per control entry, in the
// path's precision (the reference's float on the FP32 path):
// grad = dL/d(dFc) / gradient_norm, then
//   Adam:    m = b1 m + (1-b1) grad, v = b2 v + (1-b2) grad^2,
//            dFc -= alpha (m / bc1) / (sqrt(v / bc2) + eps)
//   AMSGrad: as Adam with vmax = max(vmax, v / bc2) in the denominator
//   Lion:    m = b1 m + (1-b1) sign(grad), dFc -= alpha m
//   GD:      dFc -= alpha grad
// with bc1 = 1 - b1^t, bc2 = 1 - b2^t.  State m, v, vmax (MaterialPoint
// momentum, vector, vector_max) is per particle and control layer.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_descend(long long count, DescendParams dp, Real* dFc, const Real* g, Real* mom, Real* vec,
                                                    Real* vmax)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Real b1 = Real(dp.beta1), b2 = Real(dp.beta2), alpha = Real(dp.alpha);
    const Real grad = g[i] / Real(dp.gradient_norm);
    Real d = dFc[i];
    if (dp.rule == 2) {  // Lion
        const Real sg = grad > Real(0) ? Real(1) : (grad < Real(0) ? Real(-1) : Real(0));
        const Real m = b1 * mom[i] + (Real(1) - b1) * sg;
        mom[i] = m;
        d -= alpha * m;
    } else if (dp.rule == 3) {  // gradient descent
        d -= alpha * grad;
    } else {  // Adam, AMSGrad
        const Real m = b1 * mom[i] + (Real(1) - b1) * grad;
        const Real v = b2 * vec[i] + (Real(1) - b2) * grad * grad;
        mom[i] = m;
        vec[i] = v;
        const Real mh = m / Real(dp.bc1), vh = v / Real(dp.bc2);
        Real den = vh;
        if (dp.rule == 1) {
            den = vmax[i] > vh ? vmax[i] : vh;
            vmax[i] = den;
        }
        d -= alpha * (mh / (sqrt(den) + Real(dp.epsilon)));
    }
    dFc[i] = d;
}

template <typename Real>
void launch_descend(long long count, const DescendParams& dp, Real* dFc, const Real* g, Real* mom, Real* vec, Real* vmax,
                    hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_descend<Real>, dim3(nblocks(count, kBlock)), dim3(kBlock), 0, s, count, dp, dFc, g, mom, vec, vmax);
}

// sum of squares in FP64 (PointCloud::Compute_dLdF_Norm, PointCloud.cpp:24-35):
// workgroup partials in part, their fixed-order sum in *out
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_sumsq(const Real* a, long long count, double* part)
{
    __shared__ double red[kBlock / 64];
    double local = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
        const double v = double(a[i]);
        local += v * v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

template <typename Real>
void launch_sumsq(const Real* a, long long count, double* part, double* out, hipStream_t s)
{
    unsigned nb = count > 0 ? nblocks(count, kBlock) : 1;
    if (nb > kLossBlocks) nb = kLossBlocks;
    hipLaunchKernelGGL(k_sumsq<Real>, dim3(nb), dim3(kBlock), 0, s, a, count, part);
    launch_sum_rows(part, 1, (int)nb, out, s);
}

#define DMPM_INSTANTIATE(R)                                                                                        \
    template void launch_descend<R>(long long, const DescendParams&, R*, const R*, R*, R*, R*, hipStream_t);      \
    template void launch_sumsq<R>(const R*, long long, double*, double*, hipStream_t);                             \
    template void launch_point_diag<R>(long long, const R*, const R*, const R*, double*, hipStream_t);             \
    template void launch_fill<R>(R*, long long, R, hipStream_t);                                                   \
    template void launch_p2g<R>(const StepParams&, const StepIO<R>&, hipStream_t);                                 \
    template void launch_grid_update<R>(const StepParams&, R*, hipStream_t);                                       \
    template void launch_g2p<R>(const StepParams&, const StepIO<R>&, hipStream_t);                                 \
    template void launch_p2g_mass<R>(const StepParams&, const R*, const R*, int, R*, hipStream_t);                 \
    template void launch_p2g_mass_aos<R>(const StepParams&, const double*, const double*, const double*, R*, hipStream_t); \
    template void launch_volumes<R>(const StepParams&, const R*, R*, const R*, hipStream_t);                       \
    template void launch_loss<R>(const StepParams&, const R*, const R*, double*, double*, double*, hipStream_t);        \
    template void launch_loss_particles<R>(const StepParams&, const R*, const R*, const double*, R*, hipStream_t);      \
    template void launch_g2p_adj<R>(const StepParams&, const AdjIO<R>&, hipStream_t);                              \
    template void launch_grid_adj<R>(const StepParams&, const AdjIO<R>&, hipStream_t);                             \
    template void launch_p2g_adj<R>(const StepParams&, const AdjIO<R>&, hipStream_t);

DMPM_INSTANTIATE(double)
DMPM_INSTANTIATE(float)

}  // namespace dmpm
