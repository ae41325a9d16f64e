// FP32 path: particle-grid transfers staged per grid block in LDS (gfx950).
//
// Particles are stored sorted by the 4^3-cell block of their cell; one
// workgroup processes the particle range of one block (CSR ranges built at
// sort time, at most 512 particles).  The workgroup first reduces the
// bounding box of its particles' stencil bases, then:
//   scatters (P2G, transposed G2P): accumulate every particle's 64-node
//     contribution in an LDS window, then add each non-zero window node to
//     HBM with one atomic;
//   gathers (G2P, transposed P2G): copy the window's grid channels from HBM
//     into LDS once, then every particle reads its 64 nodes from LDS.
// Fast path: when the bases span at most 5 cells per axis (particles within
// one cell of their block) the window is a fixed 8^3 box at the smallest
// base.  Nodes outside the grid are kept as zero-padded LDS slots (a gather
// reads 0, a scatter's contribution is never flushed), so the 64-node loops
// carry no bounds checks and every LDS offset is a compile-time constant
// relative to the particle's base slot.  Otherwise a general window (the
// clamped box, up to `cap` nodes, with per-node bounds checks) is used, and
// past that, direct HBM access.
//
// The FP32 grid of a step keeps the P2G sums (m, p): G2P and the transposed
// P2G apply the grid update (and its transpose) to each node as they load it,
// so no node-parallel kernel runs between the particle kernels.
//
// Scatter accumulation is fixed point in LDS integer words: LDS float atomics
// are serialised lane by lane on gfx950 (ds_add_f32 ~0.3 lane-ops/clk/CU
// against ~9 for ds_add_u64 and ~12 for ds_add_u32).  Each channel c of a
// workgroup of nP particles gets a power-of-two scale s_c with
// s_c * max_p B_pc <= 2^E, where B_pc bounds |val| for every contribution
// w * val of particle p to channel c; each contribution is rounded to an
// integer multiple of 1/s_c.
//   64-bit words (P2G): E = 30, sums in int64 (a contribution is a rounded
//   int32, sign-extended).
//   32-bit words (transposed G2P, I32): E is the largest exponent with
//   2^E nP (2/3)^3 + nP/2 < 2^31 (a node receives at most nP contributions,
//   each of weight <= (2/3)^3), so every node sum of every channel fits an
//   int32 and the wrapping 32-bit adds of the rounded two's complement
//   contributions are exact, at a resolution 2^(E-30) coarser than the
//   64-bit words (E = 23 for 512 particles); no sign extension per add.
// Window sums are thus independent of the order of the additions; the flush
// converts them back to float for one HBM atomic per node and channel.
//
// Formulas are those of the per-particle kernels (ForwardSimulation.cpp:6-124,
// BackPropagation.cpp:14-124 in the reference) with the same FP64 polar/SVD,
// PK1 and H:dF.
#include "dmpm_kernels.h"
#include "dmpm_math.h"

#include <climits>
#include <type_traits>

namespace dmpm {

namespace {

constexpr int kGat = 256;    // gather kernels
constexpr int kScat = 512;   // the loss mass scatter: one particle per thread of a block range
constexpr int kMaxWaves = 8;
constexpr int kFW = 8;       // fixed window edge (nodes)
constexpr int kFW3 = kFW * kFW * kFW;
// LDS layout of the fixed window: node (i, j, k) at slot i*kSI + j*kSJ + k.
// With kSJ = 12 and kSI = 112 the 64 cells of a 4^3 block (one per lane of a
// wave at a given stencil offset) fall on 32 distinct banks mod 32 and 16
// distinct 8-byte bank pairs: the fewest conflicts possible for 64 lanes,
// half (floats) or a quarter (64-bit) of those of the dense 8x8x8 layout.
constexpr int kSJ = 12;
constexpr int kSI = 112;
constexpr int kFWS = kFW * kSI;  // slots per channel

__device__ __forceinline__ int fixed_slot_of(int idx)  // idx: x-major node of the 8^3 window
{
    return (idx >> 6) * kSI + ((idx >> 3) & 7) * kSJ + (idx & 7);
}

// Slot of particle slot p in the home order of the constants, controls and
// control gradients (bt.home: null when the particles are in that order).
__device__ __forceinline__ int home_of(const Batch& bt, int p) { return bt.home ? bt.home[p] : p; }

struct Win {
    int lo[3];   // general window: clamped box
    int w[3];
    int vol;
    bool lds;
    int flo[3];  // fixed window origin (smallest base, unclamped)
    bool fixed;
};

__device__ __forceinline__ void stencil_f(const StepParams& sp, const float x[3], Axis<float> st[3], bool slopes)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) axis_weights<float>((x[a] - float(sp.min_point[a])) * float(sp.inv_dx), st[a], slopes);
}

__device__ __forceinline__ int base_of(const StepParams& sp, float x, int a)
{
    return stencil_base(floorf((x - float(sp.min_point[a])) * float(sp.inv_dx)));
}

// extend a thread's base box v (min x3, max x3) by the stencil base of position x
__device__ __forceinline__ void box_add(const StepParams& sp, const float x[3], int v[6])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int b = base_of(sp, x[a], a);
        v[a] = min(v[a], b);
        v[3 + a] = max(v[3 + a], b);
    }
}

// wave-wide min (first 3) / max (last 3) of the base boxes, in every lane
__device__ __forceinline__ void wave_box(int v[6])
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], __shfl_xor(v[a], off, 64));
            v[3 + a] = max(v[3 + a], __shfl_xor(v[3 + a], off, 64));
        }
    }
}

// the window from the waves' boxes red[c][wave] (one thread)
__device__ Win box_window(const StepParams& sp, const int (*red)[kMaxWaves], int nw, int cap)
{
    Win W;
    long long vol = 1;
    bool empty = false;
    W.fixed = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int lo = INT_MAX, hi = INT_MIN;
        for (int w = 0; w < nw; ++w) {
            lo = min(lo, red[a][w]);
            hi = max(hi, red[3 + a][w]);
        }
        W.flo[a] = lo;
        if ((long long)hi - (long long)lo > kFW - 4) W.fixed = false;
        long long l = lo < 0 ? 0 : lo;
        long long h = (long long)hi + 3;
        if (h > sp.dim[a] - 1) h = sp.dim[a] - 1;
        if (h < l) empty = true;
        W.lo[a] = (int)(l > INT_MAX ? INT_MAX : l);
        W.w[a] = empty ? 0 : (int)(h - l + 1);
        vol *= empty ? 0 : (h - l + 1);
    }
    W.vol = empty ? 0 : (int)(vol > INT_MAX ? INT_MAX : vol);
    W.lds = !empty && vol <= cap;
    return W;
}

// Window of a workgroup from every thread's base box (all threads).
__device__ Win window_of(const StepParams& sp, int v[6], int cap)
{
    __shared__ int red[6][kMaxWaves];
    __shared__ Win win;
    wave_box(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 6; ++c) red[c][wid] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) win = box_window(sp, red, (int)(blockDim.x >> 6), cap);
    __syncthreads();
    return win;
}

// Bounding box of the stencil bases of particles [r0, r1).
__device__ Win range_window(const StepParams& sp, const float* lay, int S, int r0, int r1, int cap)
{
    int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    for (int p = r0 + (int)threadIdx.x; p < r1; p += blockDim.x) {
        float x[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) x[a] = lay[(long long)(FX + a) * S + p];
        box_add(sp, x, v);
    }
    return window_of(sp, v, cap);
}

__device__ __forceinline__ long long win_node(const StepParams& sp, const Win& W, int idx)
{
    const int lk = idx % W.w[2];
    const int t = idx / W.w[2];
    const int lj = t % W.w[1];
    const int li = t / W.w[1];
    return ((long long)(W.lo[0] + li) * sp.dim[1] + (W.lo[1] + lj)) * sp.dim[2] + (W.lo[2] + lk);
}

__device__ __forceinline__ int win_local(const Win& W, int gi, int gj, int gk)
{
    return ((gi - W.lo[0]) * W.w[1] + (gj - W.lo[1])) * W.w[2] + (gk - W.lo[2]);
}

// fixed window: node of slot idx, or -1 outside the grid
__device__ __forceinline__ long long fixed_node(const StepParams& sp, const Win& W, int idx)
{
    const int g0 = W.flo[0] + (idx >> 6), g1 = W.flo[1] + ((idx >> 3) & 7), g2 = W.flo[2] + (idx & 7);
    if (g0 < 0 || g0 >= sp.dim[0] || g1 < 0 || g1 >= sp.dim[1] || g2 < 0 || g2 >= sp.dim[2]) return -1;
    return ((long long)g0 * sp.dim[1] + g1) * sp.dim[2] + g2;
}

__device__ __forceinline__ int fixed_slot(const Win& W, const Axis<float> st[3])
{
    return (st[0].base - W.flo[0]) * kSI + (st[1].base - W.flo[1]) * kSJ + (st[2].base - W.flo[2]);
}

// load NC channels of the window into LDS; ld(node, out) gives the NC
// channel values of a grid node
template <int NC, typename Ld>
__device__ __forceinline__ void win_load(const StepParams& sp, const Win& W, float* lds, Ld&& ld)
{
    if (W.fixed) {
        for (int idx = threadIdx.x; idx < kFW3; idx += blockDim.x) {
            const long long node = fixed_node(sp, W, idx);
            float v[NC];
            if (node >= 0) ld(node, v);
            else
#pragma unroll
                for (int c = 0; c < NC; ++c) v[c] = 0.f;
#pragma unroll
            for (int c = 0; c < NC; ++c) lds[fixed_slot_of(idx) * NC + c] = v[c];
        }
    } else if (W.lds) {
        for (int idx = threadIdx.x; idx < W.vol; idx += blockDim.x) {
            float v[NC];
            ld(win_node(sp, W, idx), v);
#pragma unroll
            for (int c = 0; c < NC; ++c) lds[idx * NC + c] = v[c];
        }
    }
}

// G_op (ForwardSimulation.cpp:53-58) of one node from its P2G sums m, p:
// v = p/m + dt f where |m| > FLT_EPSILON, else 0.  FP32 grids keep m and p;
// the kernels that read node velocities apply this as they load a node.
__device__ __forceinline__ void node_velocity(float m, const float p[3], const float fdt[3], float v[3])
{
    const bool ok = fabsf(m) > float(kMassEps);
    const float inv = 1.f / m;
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = ok ? fmaf(p[a], inv, fdt[a]) : 0.f;
}

template <int NC>
struct Scales {
    float s[NC];
    float inv[NC];
    bool ok;
};

// wave-wide max of NC bounds, in every lane (a NaN wins)
template <int NC>
__device__ __forceinline__ void wave_max(float v[NC])
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const float o = __shfl_xor(v[c], off, 64);
            v[c] = (v[c] == v[c] && !(o > v[c])) ? v[c] : o;  // max that keeps NaN
        }
}

// power-of-two scales from the waves' bounds red[c][wave] (one thread), with
// s_c * max <= 2^E: E = 30 for 64-bit words, for 32-bit words (I32) the
// largest exponent with 2^E nP (2/3)^3 + nP/2 < 2^31 (capped at 30)
template <int NC, bool I32>
__device__ Scales<NC> bound_scales(const float (*red)[kMaxWaves], int nw, int nP)
{
    Scales<NC> S;
    S.ok = true;
    int E = 30;
    if (I32) {
        // (2/3)^3 rounded up, so that float weights cannot exceed it
        E = (int)floor(log2((2147483647.0 - 0.5 * nP) / (0.29630 * (nP > 1 ? nP : 1))));
        if (E > 30) E = 30;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        float t = 0.f;
        for (int w = 0; w < nw; ++w) {
            const float r = red[c][w];
            if (!(r < 3.0e38f)) S.ok = false;  // inf or NaN somewhere in the block
            else t = fmaxf(t, r);
        }
        int e = E;
        if (t > 0.f) {
            int ex;
            frexpf(t, &ex);  // t < 2^ex
            e = E - ex;
        }
        e = e < -120 ? -120 : (e > 120 ? 120 : e);
        S.s[c] = ldexpf(1.f, e);
        S.inv[c] = ldexpf(1.f, -e);
    }
    return S;
}

// block-wide max of NC per-thread contribution bounds -> scales (all threads)
template <int NC, bool I32>
__device__ Scales<NC> block_scales(const float bound[NC], int nP)
{
    __shared__ float red[NC][kMaxWaves];
    __shared__ Scales<NC> sc;
    float v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = bound[c];
    wave_max<NC>(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NC; ++c) red[c][wid] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) sc = bound_scales<NC, I32>(red, (int)(blockDim.x >> 6), nP);
    __syncthreads();
    return sc;
}

// The window and the scales of a scatter kernel in one round of barriers.
template <int NC, bool I32>
__device__ void block_setup(const StepParams& sp, int box[6], const float bound[NC], int nP, int cap, Win& W, Scales<NC>& S)
{
    __shared__ int redi[6][kMaxWaves];
    __shared__ float redf[NC][kMaxWaves];
    __shared__ Win win;
    __shared__ Scales<NC> sc;
    float v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = bound[c];
    wave_box(box);
    wave_max<NC>(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 6; ++c) redi[c][wid] = box[c];
#pragma unroll
        for (int c = 0; c < NC; ++c) redf[c][wid] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        win = box_window(sp, redi, nw, cap);
        sc = bound_scales<NC, I32>(redf, nw, nP);
    }
    __syncthreads();
    W = win;
    S = sc;
}

// |y| < 2^30 rounded to an int32: floor(y + 1/2), one v_cvt_rpi_i32_f32
// (rintf and a conversion take two; the ties of a half go up instead of to even)
__device__ __forceinline__ int round_i32(float y)
{
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(y));
    return r;
}

// |y| < 2^30: rounded to an int32, sign-extended to the int64 accumulator
// (NONNEG: y >= 0, so the high word is 0)
template <bool NONNEG = false>
__device__ __forceinline__ unsigned long long fix32(float y)
{
    if (NONNEG) return (unsigned long long)(unsigned)round_i32(y);
    return (unsigned long long)(long long)round_i32(y);
}

// the integer sum of channel c at window slot `slot` (stride: the channel stride in words), and as a float
template <bool I32>
__device__ __forceinline__ long long lds_word(const unsigned long long* lds, int c, int stride, int slot)
{
    if (I32) return reinterpret_cast<const int*>(lds)[c * stride + slot];
    return (long long)lds[c * stride + slot];
}

template <bool I32>
__device__ __forceinline__ float word_value(long long q, float inv)
{
    if (I32) return (float)(int)q * inv;
    return (float)((double)q * (double)inv);
}

// MASS: channel 0 is the node mass; a node whose window mass is zero adds
// nothing (its other channels hold only contributions of weights below the
// fixed-point resolution), so a node left at mass 0 keeps every channel at 0.
// flags (optional): tile flags of the target grid, set for every node
// flushed (nofs: the instance's first flat node).
// I32: 32-bit words (lds holds NC int32 channels at the same strides).
template <int NC, bool MASS, bool I32>
__device__ __forceinline__ void flush_fixed(const StepParams& sp, const Win& W, const unsigned long long* lds,
                                            const Scales<NC>& S, float* base, long long cstride, unsigned* flags = nullptr,
                                            long long nofs = 0)
{
    const int n = W.fixed ? kFW3 : W.vol;
    const int stride = W.fixed ? kFWS : W.vol;
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        long long node;
        int slot;
        if (W.fixed) {
            node = fixed_node(sp, W, idx);
            if (node < 0) continue;
            slot = fixed_slot_of(idx);
        } else {
            node = win_node(sp, W, idx);
            slot = idx;
        }
        float val[NC];
        bool any = true;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const long long q = lds_word<I32>(lds, c, stride, slot);
            if (MASS && c == 0) any = q != 0;
            val[c] = word_value<I32>(q, S.inv[c]);
        }
        if (!any) continue;
        if (flags) flags[(nofs + node) >> kTileShift] = 1u;
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (val[c] != 0.f) unsafeAtomicAdd(base + c * cstride + node, val[c]);
    }
}

template <typename T>
__device__ __forceinline__ void ld3(const T* base, int field, int S, int p, T out[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = base[(long long)(field + a) * S + p];
}

template <typename T>
__device__ __forceinline__ M3<T> ldm(const T* base, int field, int S, int p)
{
    M3<T> M;
#pragma unroll
    for (int i = 0; i < 9; ++i) M.a[i] = base[(long long)(field + i) * S + p];
    return M;
}

// the control dFc of slot h, zero when ctrl is null (the layer's F field holds Fc)
__device__ __forceinline__ M3<float> ld_ctrl(const float* ctrl, int S, int h)
{
    if (ctrl) return ldm(ctrl, 0, S, h);
    M3<float> z;
#pragma unroll
    for (int i = 0; i < 9; ++i) z.a[i] = 0.f;
    return z;
}

template <typename T>
__device__ __forceinline__ void stm(T* base, int field, int S, int p, const M3<T>& M)
{
#pragma unroll
    for (int i = 0; i < 9; ++i) base[(long long)(field + i) * S + p] = M.a[i];
}

// Scatter of NC channels over a particle's 64 nodes.  val(i, j, k, out)
// fills the NC contributions (already in scaled units for the LDS modes).
// I32: 32-bit words (NC int32 channels at the strides of the 64-bit layout);
// otherwise channel 0 may be declared non-negative (NONNEG0: the mass).
// MODE 0: fixed window, 1: general LDS window, 2: HBM float atomics (the tile
// flags of the nodes written are set when flags is given).
template <int NC, int MODE, bool I32, bool NONNEG0 = false, typename ValFn>
__device__ __forceinline__ void scatter64(const StepParams& sp, const Win& W, const Axis<float> st[3],
                                          unsigned long long* lds, float* hbm, ValFn&& val, unsigned* flags = nullptr,
                                          long long nofs = 0)
{
    const int slot = MODE == 0 ? fixed_slot(W, st) : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st[0].base + i;
        if (MODE != 0 && (gi < 0 || gi >= sp.dim[0])) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st[1].base + j;
            if (MODE != 0 && (gj < 0 || gj >= sp.dim[1])) continue;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st[2].base + kk;
                if (MODE != 0 && (gk < 0 || gk >= sp.dim[2])) continue;
                float v[NC];
                val(i, j, kk, v);
                if (MODE != 2) {
                    const int o = MODE == 0 ? slot + i * kSI + j * kSJ + kk : win_local(W, gi, gj, gk);
                    const int stride = MODE == 0 ? kFWS : W.vol;
                    if (I32) {
                        unsigned* l32 = reinterpret_cast<unsigned*>(lds);
#pragma unroll
                        for (int c = 0; c < NC; ++c) atomicAdd(&l32[c * stride + o], (unsigned)round_i32(v[c]));
                    } else {
#pragma unroll
                        for (int c = 0; c < NC; ++c) atomicAdd(&lds[c * stride + o], c == 0 && NONNEG0 ? fix32<true>(v[c]) : fix32(v[c]));
                    }
                } else {
                    const long long node = ((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk;
                    if (flags) flags[(nofs + node) >> kTileShift] = 1u;
#pragma unroll
                    for (int c = 0; c < NC; ++c) unsafeAtomicAdd(hbm + c * sp.G + node, v[c]);
                }
            }
        }
    }
}

// Gather over a particle's 64 nodes: fn(mode, i, j, k, ch) with ch[c] the
// node's channel c, then row(mode, i, j) after the nodes of each (i, j) row of
// 4 (the kernels sum a row with the k-weights first and apply the (i, j)
// factors once per row); mode is MODE as a std::integral_constant.  MODE 0:
// fixed window (every node visited), 1: general LDS window, 2: HBM (nodes
// outside the grid skipped: a row may miss its first nodes, or be skipped
// with its row call when its (i, j) is outside).
template <int NC, int MODE, typename Ld, typename Fn, typename RowFn>
__device__ __forceinline__ void gather64(const StepParams& sp, const Win& W, const Axis<float> st[3], const float* lds, Ld&& ld,
                                         Fn&& fn, RowFn&& row)
{
    const int slot = MODE == 0 ? fixed_slot(W, st) : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = st[0].base + i;
        if (MODE != 0 && (gi < 0 || gi >= sp.dim[0])) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = st[1].base + j;
            if (MODE != 0 && (gj < 0 || gj >= sp.dim[1])) continue;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int gk = st[2].base + kk;
                if (MODE != 0 && (gk < 0 || gk >= sp.dim[2])) continue;
                float ch[NC];
                if (MODE == 0) {
                    const int o = slot + i * kSI + j * kSJ + kk;
#pragma unroll
                    for (int c = 0; c < NC; ++c) ch[c] = lds[o * NC + c];
                } else if (MODE == 1) {
                    const int o = win_local(W, gi, gj, gk);
#pragma unroll
                    for (int c = 0; c < NC; ++c) ch[c] = lds[o * NC + c];
                } else {
                    ld(((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk, ch);
                }
                fn(std::integral_constant<int, MODE>{}, i, j, kk, ch);
            }
            row(std::integral_constant<int, MODE>{}, i, j);
        }
    }
}

// node offsets (o - 1 - f) * h of one axis
__device__ __forceinline__ void offsets4(const Axis<float>& ax, float h, float out[4])
{
#pragma unroll
    for (int o = 0; o < 4; ++o) out[o] = (float(o - 1) - ax.f) * h;
}

}  // namespace

// ------------------------------------------------------------ tile clears --
// instance owning flat node q (nofs ascending, B+1 entries)
__device__ __forceinline__ int instance_of(const Batch& b, long long q)
{
    int lo = 0, hi = b.B - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.nofs[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Zero the 4 channels of every node of the flagged tiles and clear the flags,
// as wave `wave` of `nwaves`.  Each wave reads the flags of 64 consecutive
// tiles (one per lane, so every flag is read and reset by one lane), then
// clears the flagged ones: lanes 0-31 channels 0 and 2, lanes 32-63 channels
// 1 and 3 of the tile's 32 nodes.
__device__ __forceinline__ void clear_tiles(const Batch& bt, float* grid, unsigned* flags, long long ntiles, long long wave,
                                            long long nwaves)
{
    const int lane = threadIdx.x & 63;
    const int c0 = lane >> 5;
    const long long Gtot = bt.Gtot;
    for (long long t0 = wave * 64; t0 < ntiles; t0 += nwaves * 64) {
        const long long t = t0 + lane;
        const unsigned f = t < ntiles ? flags[t] : 0u;
        unsigned long long mask = __ballot(f != 0u);
        if (f != 0u) flags[t] = 0u;
        while (mask) {
            const int b = __ffsll((long long)mask) - 1;
            mask &= mask - 1;
            const long long q = ((t0 + b) << kTileShift) + (lane & 31);
            if (q >= Gtot) continue;
            long long n0 = 0, Gi = Gtot;
            if (bt.B > 1) {
                const int i = instance_of(bt, q);
                n0 = bt.nofs[i];
                Gi = bt.nofs[i + 1] - n0;
            }
            float* g = grid + 4 * n0 + (q - n0);
            g[c0 * Gi] = 0.f;
            g[(c0 + 2) * Gi] = 0.f;
        }
    }
}

// a kernel's workgroups from `first` on clear tc's grid (k_g2p_blk and the
// elastic kernels carry the clear of a grid the next kernels scatter into)
__device__ __forceinline__ void appended_clear(const Batch& bt, const TileClear& tc, unsigned first)
{
    const long long nb = (long long)gridDim.x - first;
    clear_tiles(bt, tc.grid, tc.flags, num_tiles(bt.Gtot), ((long long)(blockIdx.x - first) * blockDim.x + threadIdx.x) >> 6,
                (nb * blockDim.x) >> 6);
}

// The ranges of a launch.  Without a queue (bt.queue null) workgroup b takes
// range b and the workgroups past the last range clear tc.  With one the
// launch is persistent: its workgroups (as many as fit on the chip at once)
// take the next range from the counter queue[0] until none is left, then
// share the clear of tc; the last workgroup to finish (counter queue[1])
// resets both counters for the next launch.
// The two forms are separate instantiations (PERSIST), so that the loop of
// the persistent one leaves the register allocation of the other unchanged.
template <bool PERSIST, typename Body>
__device__ __forceinline__ void for_ranges(const Batch& bt, const TileClear& tc, Body&& body)
{
    if constexpr (!PERSIST) {
        if (blockIdx.x >= (unsigned)bt.nranges) {
            appended_clear(bt, tc, (unsigned)bt.nranges);
            return;
        }
        body(bt.ranges[blockIdx.x]);
        return;
    }
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(bt.queue, 1);
        __syncthreads();
        const int r = next;
        __syncthreads();  // every thread holds r before thread 0 takes the next range
        if (r >= bt.nranges) break;
        body(bt.ranges[r]);
    }
    if (tc.grid)
        clear_tiles(bt, tc.grid, tc.flags, num_tiles(bt.Gtot), ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6,
                    ((long long)gridDim.x * blockDim.x) >> 6);
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(bt.queue + 1, 1) == (int)gridDim.x - 1) {
            atomicExch(bt.queue, 0);
            atomicExch(bt.queue + 1, 0);
        }
    }
}

// ------------------------------------------------------------------ P2G ----
// One particle's P2G inputs: position, m (1 - dt drag) v, m, and the affine
// matrix A = -(3/dx^2) dt vol P(Fc) Fc^T + m C (FP64 polar and PK1), which is
// also cached in the layer for the transposed P2G.
struct P2GPart {
    float x[3], mv[3], m;
    M3<float> A;
};

// q from Fc = F + dF and the constants; rest(x, v, C, m) supplies the fields
// the elastic part does not need after it, so that they need not occupy
// registers across the FP64 polar.  aux (or null): the layer that caches A.
// bound: running per-channel bounds of the contributions.
template <typename Rest>
__device__ __forceinline__ void p2g_part(const StepParams& sp, const M3<float>& F, const M3<float>& dF, double vol, double lam,
                                         double mu, Rest&& rest, float* aux, int p, P2GPart& q, float bound[4])
{
    const float dx = float(sp.dx);
    const float sdrag = 1.f - float(sp.dt) * float(sp.drag);
    const double kst = -3.0 * sp.inv_dx * sp.inv_dx * sp.dt;
    M3<double> Fc;
#pragma unroll
    for (int i = 0; i < 9; ++i) Fc.a[i] = double(F.a[i]) + double(dF.a[i]);
    const M3<double> PF = pk1_times_ft(Fc, lam, mu);
    const double k = kst * vol;
    float kPF[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) kPF[i] = float(k * PF.a[i]);
    float v[3];
    M3<float> C;
    rest(q.x, v, C, q.m);
#pragma unroll
    for (int i = 0; i < 9; ++i) q.A.a[i] = kPF[i] + q.m * C.a[i];
    if (aux) stm(aux, FA, sp.stride, p, q.A);
#pragma unroll
    for (int a = 0; a < 3; ++a) q.mv[a] = q.m * v[a] * sdrag;
    // every contribution: |w (mv + A d)| <= |mv| + 2 dx sum_b |A_ab|  (w <= 1, |d| <= 2 dx)
    bound[0] = fmaxf(bound[0], fabsf(q.m));
#pragma unroll
    for (int a = 0; a < 3; ++a)
        bound[1 + a] = fmaxf(bound[1 + a], fabsf(q.mv[a]) + 2.f * dx * (fabsf(q.A.a[3 * a]) + fabsf(q.A.a[3 * a + 1]) +
                                                                         fabsf(q.A.a[3 * a + 2])));
}

// h: the particle's home slot (constants, control)
__device__ __forceinline__ void p2g_prepare(const StepParams& sp, const StepIO<float>& io, int p, int h, P2GPart& q, float bound[4])
{
    const int S = sp.stride;
    const M3<float> F = ldm(io.cur, FF, S, p);
    const M3<float> dF = ldm(io.ctrl, 0, S, h);
    const double vol = io.cst[KVOL * S + h];
    const double lam = io.cst[KLAM * S + h], mu = io.cst[KMU * S + h];
    auto rest = [&](float x[3], float v[3], M3<float>& C, float& m) {
        ld3(io.cur, FX, S, p, x);
        ld3(io.cur, FV, S, p, v);
        C = ldm(io.cur, FC, S, p);
        m = io.cst[KM * S + h];
    };
    p2g_part(sp, F, dF, vol, lam, mu, rest, io.aux, p, q, bound);
}

__device__ __forceinline__ void p2g_scale(P2GPart& q, const Scales<4>& SC)
{
    q.m *= SC.s[0];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        q.mv[a] *= SC.s[1 + a];
#pragma unroll
        for (int b = 0; b < 3; ++b) q.A.a[3 * a + b] *= SC.s[1 + a];
    }
}

// Two particles per thread: particles t and t + kP2G of the range.  With the
// block-sorted order (rank in cell, then cell) these are particles of one
// cell, so they usually share their stencil base; a wave whose lanes all do
// (or have no second particle) adds both with one LDS atomic per node and
// channel, which halves the atomics that bound this kernel.  Other waves add
// the two particles one after the other.
constexpr int kP2G = 256;

// The scatter of a range's prepared particles q0 (valid v0), q1 (v1) into
// io.grid (already offset to the instance), from their base box and bounds:
// window and scales, LDS accumulation, flush.  Starts with a barrier, so the
// caller may have used the LDS before.  ZEROED: the caller zeroed the first
// 4 kFWS words of lds64 before that barrier (windows then hold at most kFWS
// nodes: bt.cap <= kFWS).
template <bool ZEROED = false>
__device__ __forceinline__ void p2g_scatter(const Batch& bt, const StepParams& sp, const StepIO<float>& io, long long nofs, int nP,
                                            P2GPart& q0, P2GPart& q1, bool v0, bool v1, const float bound[4], int box[6],
                                            unsigned long long* lds64)
{
    const float dx = float(sp.dx);
    Win W;
    Scales<4> SC;
    block_setup<4, false>(sp, box, bound, nP, bt.cap, W, SC);
    const int mode = (W.fixed && SC.ok) ? 0 : ((W.lds && SC.ok) ? 1 : 2);
    if (threadIdx.x == 0 && bt.wstats) atomicAdd(bt.wstats + mode, 1ull);
    if (!ZEROED) {
        if (mode == 0)
            for (int i = threadIdx.x; i < 4 * kFWS; i += blockDim.x) lds64[i] = 0ull;
        else if (mode == 1)
            for (int i = threadIdx.x; i < 4 * W.vol; i += blockDim.x) lds64[i] = 0ull;
        __syncthreads();
    }
    if (mode != 2) {  // scaled units
        p2g_scale(q0, SC);
        p2g_scale(q1, SC);
    }
    Axis<float> s0[3], s1[3];
    stencil_f(sp, q0.x, s0, false);
    stencil_f(sp, q1.x, s1, false);
    const bool same = v1 && s0[0].base == s1[0].base && s0[1].base == s1[1].base && s0[2].base == s1[2].base;
    if (mode == 0 && __all(!v1 || same)) {
        if (v0) {
            float a0[4], a1[4], a2[4], b0[4], b1[4], b2[4];
            offsets4(s0[0], dx, a0);
            offsets4(s0[1], dx, a1);
            offsets4(s0[2], dx, a2);
            offsets4(s1[0], dx, b0);
            offsets4(s1[1], dx, b1);
            offsets4(s1[2], dx, b2);
            // without a second particle q1 is all zero: its terms vanish
            auto val = [&](int i, int j, int kk, float* out) {
                const float wa = s0[0].w[i] * s0[1].w[j] * s0[2].w[kk];
                const float wb = s1[0].w[i] * s1[1].w[j] * s1[2].w[kk];
                out[0] = wa * q0.m + wb * q1.m;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    out[1 + a] = wa * (q0.mv[a] + q0.A.a[3 * a] * a0[i] + q0.A.a[3 * a + 1] * a1[j] + q0.A.a[3 * a + 2] * a2[kk]) +
                                 wb * (q1.mv[a] + q1.A.a[3 * a] * b0[i] + q1.A.a[3 * a + 1] * b1[j] + q1.A.a[3 * a + 2] * b2[kk]);
            };
            scatter64<4, 0, false, true>(sp, W, s0, lds64, io.grid, val);
        }
    } else {
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            if (!(h ? v1 : v0)) continue;
            const P2GPart& q = h ? q1 : q0;
            Axis<float> st[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) st[c] = h ? s1[c] : s0[c];
            float d0s[4], d1s[4], d2s[4];
            offsets4(st[0], dx, d0s);
            offsets4(st[1], dx, d1s);
            offsets4(st[2], dx, d2s);
            auto val = [&](int i, int j, int kk, float* out) {
                const float w = st[0].w[i] * st[1].w[j] * st[2].w[kk];
                out[0] = w * q.m;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    out[1 + a] = w * (q.mv[a] + q.A.a[3 * a] * d0s[i] + q.A.a[3 * a + 1] * d1s[j] + q.A.a[3 * a + 2] * d2s[kk]);
            };
            if (mode == 0) scatter64<4, 0, false, true>(sp, W, st, lds64, io.grid, val);
            else if (mode == 1) scatter64<4, 1, false, true>(sp, W, st, lds64, io.grid, val);
            else scatter64<4, 2, false>(sp, W, st, lds64, io.grid, val, io.gflags, nofs);
        }
    }
    if (mode != 2) {
        __syncthreads();
        flush_fixed<4, true, false>(sp, W, lds64, SC, io.grid, sp.G, io.gflags, nofs);
    }
}

// tc: a grid to clear in appended workgroups (the next step's), or none
__device__ __forceinline__ void p2g_range(const Batch& bt, StepIO<float> io, const int4 rg, unsigned long long* lds64)
{
    const StepParams sp = bt.sps[rg.z];
    const long long nofs = bt.nofs[rg.z];
    io.grid += 4 * nofs;
    const int p0 = rg.x + (int)threadIdx.x, p1 = p0 + kP2G;
    const bool v0 = p0 < rg.y, v1 = p1 < rg.y;
    P2GPart q0{}, q1{};
    float bound[4] = {0.f, 0.f, 0.f, 0.f};
    int box[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    if (v0) {
        p2g_prepare(sp, io, p0, home_of(bt, p0), q0, bound);
        box_add(sp, q0.x, box);
    }
    if (v1) {
        p2g_prepare(sp, io, p1, home_of(bt, p1), q1, bound);
        box_add(sp, q1.x, box);
    }
    p2g_scatter(bt, sp, io, nofs, rg.y - rg.x, q0, q1, v0, v1, bound, box, lds64);
}

template <bool PERSIST>
__global__ __launch_bounds__(kP2G, 4) void k_p2g_blk(Batch bt, StepIO<float> io, TileClear tc)
{
    extern __shared__ unsigned long long lds64[];
    for_ranges<PERSIST>(bt, tc, [&](const int4 rg) { p2g_range(bt, io, rg, lds64); });
}

// --------------------------------------------------- P2G of mass only ----
// Node masses of a layer (the EndLayerMassLoss grid, CompGraph.cpp:110-141),
// with the window scatter of k_p2g_blk.
__global__ __launch_bounds__(kScat) void k_p2g_mass_blk(Batch bt, const float* lay, const float* cst, float* gm)
{
    extern __shared__ unsigned long long lds64[];
    const int4 rg = bt.ranges[blockIdx.x];
    const StepParams sp = bt.sps[rg.z];
    gm += bt.nofs[rg.z];
    const int S = sp.stride;
    const int p = rg.x + (int)threadIdx.x;
    const bool valid = p < rg.y;
    const Win W = range_window(sp, lay, S, rg.x, rg.y, bt.cap);
    float x[3] = {0.f, 0.f, 0.f}, m = 0.f;
    float bound[1] = {0.f};
    if (valid) {
        ld3(lay, FX, S, p, x);
        m = cst[KM * S + home_of(bt, p)];
        bound[0] = fabsf(m);
    }
    const Scales<1> SC = block_scales<1, false>(bound, rg.y - rg.x);
    const int mode = (W.fixed && SC.ok) ? 0 : ((W.lds && SC.ok) ? 1 : 2);
    if (mode == 0)
        for (int i = threadIdx.x; i < kFWS; i += blockDim.x) lds64[i] = 0ull;
    else if (mode == 1)
        for (int i = threadIdx.x; i < W.vol; i += blockDim.x) lds64[i] = 0ull;
    __syncthreads();
    if (valid) {
        Axis<float> st[3];
        stencil_f(sp, x, st, false);
        if (mode != 2) m *= SC.s[0];
        auto val = [&](int i, int j, int kk, float* out) { out[0] = st[0].w[i] * st[1].w[j] * st[2].w[kk] * m; };
        if (mode == 0) scatter64<1, 0, false, true>(sp, W, st, lds64, gm, val);
        else if (mode == 1) scatter64<1, 1, false, true>(sp, W, st, lds64, gm, val);
        else scatter64<1, 2, false>(sp, W, st, lds64, gm, val);
    }
    if (mode != 2) {
        __syncthreads();
        flush_fixed<1, true, false>(sp, W, lds64, SC, gm, sp.G);
    }
}

// ------------------------------------------------------------------ G2P ----
// node velocities of an FP32 grid (m, p) as G2P loads them
struct NodeVel {
    const float* gr;
    long long G;
    float fdt[3];
    __device__ NodeVel(const StepParams& sp, const float* grid) : gr(grid), G(sp.G)
    {
        const float dt = float(sp.dt);
#pragma unroll
        for (int a = 0; a < 3; ++a) fdt[a] = dt * float(sp.f_ext[a]);
    }
    __device__ __forceinline__ void operator()(long long node, float* v) const
    {
        const float p[3] = {gr[G + node], gr[2 * G + node], gr[3 * G + node]};
        node_velocity(gr[node], p, fdt, v);
    }
};

// G2P and P_op_2 (ForwardSimulation.cpp:60-124) of a particle at x: its next
// state from the window W of node velocities in lds; fdf(F, dF) gives its
// deformation gradient and control after the gather (a load, or values the
// caller loaded earlier)
template <typename FdF>
__device__ __forceinline__ void g2p_particle(const StepParams& sp, const Win& W, const float* lds, const NodeVel& ld,
                                             const float x[3], FdF&& fdf, float x1[3], float v1[3], M3<float>& C1, M3<float>& F1)
{
    const float dt = float(sp.dt), s = float(sp.smoothing);
    const float Dd = float(3.0 * sp.inv_dx);
    Axis<float> st[3];
    stencil_f(sp, x, st, false);
    float e0s[4], e1s[4], e2s[4];
    offsets4(st[0], Dd, e0s);
    offsets4(st[1], Dd, e1s);
    offsets4(st[2], Dd, e2s);
    float vp[3] = {0.f, 0.f, 0.f};
    M3<float> Cp = m3_zero<float>();
    // row sums R = sum_k w_k v, Re = sum_k w_k e_k v: the fixed window visits
    // every node, so a row's first node starts them; the other modes may skip
    // it, so their sums start at 0 and restart after every row
    float we2[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) we2[k] = st[2].w[k] * e2s[k];
    float R[3] = {0.f, 0.f, 0.f}, Re[3] = {0.f, 0.f, 0.f};
    auto acc = [&](auto mode, int, int, int kk, const float* vi) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (decltype(mode)::value == 0 && kk == 0) {
                R[a] = st[2].w[0] * vi[a];
                Re[a] = we2[0] * vi[a];
            } else {
                R[a] += st[2].w[kk] * vi[a];
                Re[a] += we2[kk] * vi[a];
            }
        }
    };
    auto row = [&](auto mode, int i, int j) {
        const float w01 = st[0].w[i] * st[1].w[j];
        const float c0 = w01 * e0s[i], c1 = w01 * e1s[j];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vp[a] += w01 * R[a];
            Cp.a[3 * a] += c0 * R[a];
            Cp.a[3 * a + 1] += c1 * R[a];
            Cp.a[3 * a + 2] += w01 * Re[a];
            if (decltype(mode)::value != 0) {
                R[a] = 0.f;
                Re[a] = 0.f;
            }
        }
    };
    if (W.fixed) gather64<3, 0>(sp, W, st, lds, ld, acc, row);
    else if (W.lds) gather64<3, 1>(sp, W, st, lds, ld, acc, row);
    else gather64<3, 2>(sp, W, st, lds, ld, acc, row);
    M3<float> F, dF;
    fdf(F, dF);
    M3<float> Fc, IC;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Fc.a[i] = F.a[i] + dF.a[i];
        IC.a[i] = dt * Cp.a[i];
    }
    IC.a[0] += 1.f;
    IC.a[4] += 1.f;
    IC.a[8] += 1.f;
    F1 = mul(IC, Fc);
#pragma unroll
    for (int i = 0; i < 9; ++i) F1.a[i] = F1.a[i] * (1.f - s) + F.a[i] * s;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        x1[a] = x[a] + dt * vp[a];
        v1[a] = vp[a];
    }
    C1 = Cp;
}

// full false: a layer that only the next step's G2P reads (x and F)
__device__ __forceinline__ void store_state(const StepParams& sp, float* lay, int p, const float x[3], const float v[3],
                                            const M3<float>& C, const M3<float>& F, bool full = true)
{
    const int S = sp.stride;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lay[(long long)(FX + a) * S + p] = x[a];
        if (full) lay[(long long)(FV + a) * S + p] = v[a];
    }
    if (full) stm(lay, FC, S, p, C);
    stm(lay, FF, S, p, F);
}

// the window of the node velocities that a range's G2P reads, loaded into lds;
// xs: the thread's (at most two: ranges hold <= 512) particle positions
__device__ __forceinline__ Win g2p_window(const Batch& bt, const StepParams& sp, const StepIO<float>& io, int4 rg, float xs[2][3],
                                          float* lds, const NodeVel& ld)
{
    int box[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = rg.x + (int)threadIdx.x + h * kGat;
        if (p < rg.y) {
            ld3(io.cur, FX, sp.stride, p, xs[h]);
            box_add(sp, xs[h], box);
        }
    }
    const Win W = window_of(sp, box, bt.cap);
    win_load<3>(sp, W, lds, ld);
    __syncthreads();
    return W;
}

__device__ __forceinline__ void g2p_range(const Batch& bt, StepIO<float> io, const int4 rg, float* lds)
{
    const StepParams sp = bt.sps[rg.z];
    io.grid += 4 * bt.nofs[rg.z];
    const NodeVel ld(sp, io.grid);
    float xs[2][3];
    const Win W = g2p_window(bt, sp, io, rg, xs, lds, ld);
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        const int p = rg.x + (int)threadIdx.x + h * kGat;
        if (p >= rg.y) break;
        float x[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) x[a] = h ? xs[1][a] : xs[0][a];
        float x1[3], v1[3];
        M3<float> C1, F1;
        const int hp = home_of(bt, p);
        auto fdf = [&](M3<float>& F, M3<float>& dF) {
            F = ldm(io.cur, FF, sp.stride, p);
            dF = ld_ctrl(io.ctrl, sp.stride, hp);
        };
        g2p_particle(sp, W, lds, ld, x, fdf, x1, v1, C1, F1);
        store_state(sp, io.next, p, x1, v1, C1, F1);
    }
}

template <bool PERSIST>
__global__ __launch_bounds__(kGat) void k_g2p_blk(Batch bt, StepIO<float> io, TileClear tc)
{
    extern __shared__ float lds[];
    for_ranges<PERSIST>(bt, tc, [&](const int4 rg) { g2p_range(bt, io, rg, lds); });
}

// ------------------------------------------------------- fused step ------
// G2P of step i - 1 (gio: layer i - 1 -> layer i, grid of step i - 1) and P2G
// of step i (pio: dFc_i, grid of step i, A cached in layer i when pio.aux is
// set) in one launch, one workgroup per range: each thread's new particle
// state goes from the gather straight into the scatter's inputs, and the
// VALU-bound scatter of one workgroup overlaps the memory-bound gathers of
// others.  The grid of step i must be clear; tc clears step i + 1's.  The
// results equal those of k_g2p_blk then k_p2g_blk: the state written is the
// state scattered, and the scatter sums are independent of their order.
// The LDS holds the scatter's window (4 kFWS words, zeroed as the gather's
// window loads) and after it the gather's (3 kFWS floats): windows hold at
// most kFWS nodes (bt.cap <= kFWS).
static_assert(kP2G == kGat, "the fused step maps particles to threads as both kernels do");
__device__ __forceinline__ void step_range(const Batch& bt, StepIO<float> gio, StepIO<float> pio, const int4 rg,
                                           unsigned long long* lds64)
{
    const StepParams sp = bt.sps[rg.z];
    const long long nofs = bt.nofs[rg.z];
    gio.grid += 4 * nofs;
    pio.grid += 4 * nofs;
    const int S = sp.stride;
    const NodeVel ld(sp, gio.grid);
    float* glds = reinterpret_cast<float*>(lds64 + 4 * kFWS);
    for (int i = threadIdx.x; i < 4 * kFWS; i += blockDim.x) lds64[i] = 0ull;
    float xs[2][3];
    const Win W = g2p_window(bt, sp, gio, rg, xs, glds, ld);
    const int p0 = rg.x + (int)threadIdx.x, p1 = p0 + kP2G;
    const bool v0 = p0 < rg.y, v1 = p1 < rg.y;
    P2GPart q0{}, q1{};
    float bound[4] = {0.f, 0.f, 0.f, 0.f};
    int box[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    auto one = [&](int p, const float x[3], P2GPart& q) {
        float xn[3], vn[3];
        M3<float> Cn, Fn;
        const int h = home_of(bt, p);
        auto fdf = [&](M3<float>& F, M3<float>& dF_) {
            F = ldm(gio.cur, FF, S, p);
            dF_ = ld_ctrl(gio.ctrl, S, h);
        };
        g2p_particle(sp, W, glds, ld, x, fdf, xn, vn, Cn, Fn);
        const M3<float> dF = ldm(pio.ctrl, 0, S, h);
        // v and C of a lean layer go straight into this step's P2G below; an Fc
        // layer's F field takes Fn + dF, the sum its readers form in FP32
        M3<float> Fs;
#pragma unroll
        for (int i = 0; i < 9; ++i) Fs.a[i] = gio.fc ? Fn.a[i] + dF.a[i] : Fn.a[i];
        store_state(sp, gio.next, p, xn, vn, Cn, Fs, !gio.lean);
        const double vol = pio.cst[KVOL * S + h];
        const double lam = pio.cst[KLAM * S + h], mu = pio.cst[KMU * S + h];
        auto rest = [&](float x_[3], float v_[3], M3<float>& C_, float& m_) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                x_[a] = xn[a];
                v_[a] = vn[a];
            }
            C_ = Cn;
            m_ = pio.cst[KM * S + h];
        };
        p2g_part(sp, Fn, dF, vol, lam, mu, rest, pio.aux, p, q, bound);
        box_add(sp, q.x, box);
    };
    if (v0) one(p0, xs[0], q0);
    if (v1) one(p1, xs[1], q1);
    // the scatter's window was zeroed before g2p_window's barriers
    p2g_scatter<true>(bt, sp, pio, nofs, rg.y - rg.x, q0, q1, v0, v1, bound, box, lds64);
}

template <bool PERSIST>
__global__ __launch_bounds__(kP2G, 4) void k_step_blk(Batch bt, StepIO<float> gio, StepIO<float> pio, TileClear tc)
{
    extern __shared__ unsigned long long lds64[];
    for_ranges<PERSIST>(bt, tc, [&](const int4 rg) { step_range(bt, gio, pio, rg, lds64); });
}

// -------------------------------------------------------- transposed G2P ---
// P_op_2^T + G2P^T: vt = v' + dt x', Ct = C' + (1-s) dt F' Fc^T (kept in tmp),
// vbar_i += w (vt + D Ct d), into the adjoint grid (3 channels, 32-bit
// words; a node of zero forward mass may receive vbar: the transposed P2G
// gives it pbar = mbar = 0 whatever vbar is).  Two particles per thread,
// paired as in P2G.
struct G2PAdjPart {
    float x[3], vt[3];
    M3<float> Ct;
};

__device__ __forceinline__ void g2p_adj_prepare(const StepParams& sp, const AdjIO<float>& io, int p, int h, G2PAdjPart& q, float bound[3])
{
    const int S = sp.stride;
    const float Dd = float(3.0 * sp.inv_dx);
    const float dt = float(sp.dt);
    const float cs = (sp.compat & 2) ? 1.f : 1.f - float(sp.smoothing);
    float gx1[3], gv1[3];
    ld3(io.cur, FX, S, p, q.x);
    ld3(io.g1, FX, S, p, gx1);
    ld3(io.g1, FV, S, p, gv1);
    const M3<float> F = ldm(io.cur, FF, S, p);
    // without a control the load repeats F's (a cache hit) and its value is
    // dropped: no branch between this particle's loads
    const bool hc = io.ctrl != nullptr;
    const M3<float> dF = ldm(hc ? io.ctrl : io.cur, hc ? 0 : FF, S, hc ? h : p);
    const M3<float> gF1 = ldm(io.g1, FF, S, p);
    const M3<float> gC1 = ldm(io.g1, FC, S, p);
    M3<float> Fc;
#pragma unroll
    for (int i = 0; i < 9; ++i) Fc.a[i] = F.a[i] + (hc ? dF.a[i] : 0.f);
    const M3<float> FF_ = mult(gF1, Fc);
#pragma unroll
    for (int i = 0; i < 9; ++i) q.Ct.a[i] = gC1.a[i] + cs * dt * FF_.a[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) q.vt[a] = gv1[a] + dt * gx1[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) io.tmp[(long long)a * S + p] = q.vt[a];
    stm(io.tmp, 3, S, p, q.Ct);
    // every contribution: |w (vt + Ct e)| <= |vt| + (6/dx) sum_b |Ct_ab|  (w <= 1, |e| = |D d| <= 6/dx)
#pragma unroll
    for (int a = 0; a < 3; ++a)
        bound[a] = fmaxf(bound[a], fabsf(q.vt[a]) + 2.f * Dd * (fabsf(q.Ct.a[3 * a]) + fabsf(q.Ct.a[3 * a + 1]) +
                                                                fabsf(q.Ct.a[3 * a + 2])));
}

__device__ __forceinline__ void g2p_adj_scale(G2PAdjPart& q, const Scales<3>& SC)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        q.vt[a] *= SC.s[a];
#pragma unroll
        for (int b = 0; b < 3; ++b) q.Ct.a[3 * a + b] *= SC.s[a];
    }
}

constexpr int kG2PAdj = 256;

__device__ __forceinline__ void g2p_adj_range(const Batch& bt, AdjIO<float> io, const int4 rg, unsigned long long* lds64)
{
    const StepParams sp = bt.sps[rg.z];
    const long long nofs = bt.nofs[rg.z];
    io.agrid += 4 * nofs;
    const int p0 = rg.x + (int)threadIdx.x, p1 = p0 + kG2PAdj;
    const bool v0 = p0 < rg.y, v1 = p1 < rg.y;
    const float Dd = float(3.0 * sp.inv_dx);
    G2PAdjPart q0{}, q1{};
    float bound[3] = {0.f, 0.f, 0.f};
    int box[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    if (v0) {
        g2p_adj_prepare(sp, io, p0, home_of(bt, p0), q0, bound);
        box_add(sp, q0.x, box);
    }
    if (v1) {
        g2p_adj_prepare(sp, io, p1, home_of(bt, p1), q1, bound);
        box_add(sp, q1.x, box);
    }
    Win W;
    Scales<3> SC;
    block_setup<3, true>(sp, box, bound, rg.y - rg.x, bt.cap, W, SC);
    const int mode = (W.fixed && SC.ok) ? 0 : ((W.lds && SC.ok) ? 1 : 2);
    if (mode == 0)
        for (int i = threadIdx.x; i < 3 * kFWS; i += blockDim.x) reinterpret_cast<unsigned*>(lds64)[i] = 0u;
    else if (mode == 1)
        for (int i = threadIdx.x; i < 3 * W.vol; i += blockDim.x) reinterpret_cast<unsigned*>(lds64)[i] = 0u;
    __syncthreads();
    if (mode != 2) {
        g2p_adj_scale(q0, SC);
        g2p_adj_scale(q1, SC);
    }
    Axis<float> s0[3], s1[3];
    stencil_f(sp, q0.x, s0, false);
    stencil_f(sp, q1.x, s1, false);
    const bool same = v1 && s0[0].base == s1[0].base && s0[1].base == s1[1].base && s0[2].base == s1[2].base;
    if (mode == 0 && __all(!v1 || same)) {
        if (v0) {
            float a0[4], a1[4], a2[4], b0[4], b1[4], b2[4];
            offsets4(s0[0], Dd, a0);
            offsets4(s0[1], Dd, a1);
            offsets4(s0[2], Dd, a2);
            offsets4(s1[0], Dd, b0);
            offsets4(s1[1], Dd, b1);
            offsets4(s1[2], Dd, b2);
            // without a second particle q1 is all zero: its terms vanish
            auto val = [&](int i, int j, int kk, float* out) {
                const float wa = s0[0].w[i] * s0[1].w[j] * s0[2].w[kk];
                const float wb = s1[0].w[i] * s1[1].w[j] * s1[2].w[kk];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    out[a] = wa * (q0.vt[a] + q0.Ct.a[3 * a] * a0[i] + q0.Ct.a[3 * a + 1] * a1[j] + q0.Ct.a[3 * a + 2] * a2[kk]) +
                             wb * (q1.vt[a] + q1.Ct.a[3 * a] * b0[i] + q1.Ct.a[3 * a + 1] * b1[j] + q1.Ct.a[3 * a + 2] * b2[kk]);
            };
            scatter64<3, 0, true>(sp, W, s0, lds64, io.agrid, val);
        }
    } else {
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            if (!(h ? v1 : v0)) continue;
            const G2PAdjPart& q = h ? q1 : q0;
            Axis<float> st[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) st[c] = h ? s1[c] : s0[c];
            float e0s[4], e1s[4], e2s[4];
            offsets4(st[0], Dd, e0s);
            offsets4(st[1], Dd, e1s);
            offsets4(st[2], Dd, e2s);
            auto val = [&](int i, int j, int kk, float* out) {
                const float w = st[0].w[i] * st[1].w[j] * st[2].w[kk];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    out[a] = w * (q.vt[a] + q.Ct.a[3 * a] * e0s[i] + q.Ct.a[3 * a + 1] * e1s[j] + q.Ct.a[3 * a + 2] * e2s[kk]);
            };
            if (mode == 0) scatter64<3, 0, true>(sp, W, st, lds64, io.agrid, val);
            else if (mode == 1) scatter64<3, 1, true>(sp, W, st, lds64, io.agrid, val);
            else scatter64<3, 2, true>(sp, W, st, lds64, io.agrid, val, io.aflags, nofs);
        }
    }
    if (mode != 2) {
        __syncthreads();
        flush_fixed<3, false, true>(sp, W, lds64, SC, io.agrid, sp.G, io.aflags, nofs);
    }
}

template <bool PERSIST>
__global__ __launch_bounds__(kG2PAdj, 4) void k_g2p_adj_blk(Batch bt, AdjIO<float> io)
{
    extern __shared__ unsigned long long lds64[];
    for_ranges<PERSIST>(bt, TileClear{}, [&](const int4 rg) { g2p_adj_range(bt, io, rg, lds64); });
}

// Elastic half of P2G^T + P_op_1^T for particle p (FP64 polar, PK1, H:dF):
//   Pbar = -k Gb Fc,  Fcbar = (1-s)(I + dt C')^T F' - k Gb^T P + H(Fc):Pbar
// with k = 3 dt vol / dx^2.  rest = -k Gb^T P + H(Fc):Pbar is formed in FP64
// and rounded; this is the FP32 remainder and the stores.
__device__ __forceinline__ void elastic_adj_store(const StepParams& sp, const AdjIO<float>& io, int p, int h, const M3<float>& rest)
{
    const int S = sp.stride;
    const int cp = sp.compat;
    const float dt = float(sp.dt);
    const M3<float> C1 = ldm(io.next, FC, S, p);
    const M3<float> gF1 = ldm(io.g1, FF, S, p);
    M3<float> IC;
#pragma unroll
    for (int i = 0; i < 9; ++i) IC.a[i] = dt * C1.a[i];
    IC.a[0] += 1.f;
    IC.a[4] += 1.f;
    IC.a[8] += 1.f;
    const M3<float> base = tmul(IC, gF1);
    const float s = float(sp.smoothing);
    M3<float> gF, gdFc;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (cp & 1) {
            const float f = (1.f - s) * base.a[i] + s * gF1.a[i] + base.a[i] + rest.a[i];
            gF.a[i] = f;
            gdFc.a[i] = f;
        } else {
            gdFc.a[i] = (1.f - s) * base.a[i] + rest.a[i];
            gF.a[i] = gdFc.a[i] + s * gF1.a[i];
        }
    }
    stm(io.g0, FF, S, p, gF);
    if (io.gctrl) stm(io.gctrl, 0, S, h, gdFc);
}

// symmetric 3x3, entries (00, 11, 22, 01, 02, 12)
struct Sym3 {
    double a[6];
    __device__ __forceinline__ double operator()(int i, int j) const { return i == j ? a[i] : a[i + j + 2]; }
};

__device__ __forceinline__ Sym3 sym_inverse(const Sym3& s)
{
    Sym3 c;
    c.a[0] = s.a[1] * s.a[2] - s.a[5] * s.a[5];
    c.a[1] = s.a[0] * s.a[2] - s.a[4] * s.a[4];
    c.a[2] = s.a[0] * s.a[1] - s.a[3] * s.a[3];
    c.a[3] = s.a[4] * s.a[5] - s.a[3] * s.a[2];
    c.a[4] = s.a[3] * s.a[5] - s.a[4] * s.a[1];
    c.a[5] = s.a[3] * s.a[4] - s.a[0] * s.a[5];
    const double inv = 1.0 / (s.a[0] * c.a[0] + s.a[3] * c.a[3] + s.a[4] * c.a[4]);
#pragma unroll
    for (int i = 0; i < 6; ++i) c.a[i] *= inv;
    return c;
}

// rest of the polar path (Fc = R S, R from the Newton iteration), in the
// frame of the rotation: F^-T = R S^-1, P = R Sig with
// Sig = 2 mu (S - I) + lam (J - 1) J S^-1, and with Gh = R^T Gb R
//   Gb^T P = R Gh^T Sig,   Pbar = R Pb' with Pb' = -k Gh S,
//   F^-T:Pbar = tr(S^-1 Pb') = -k tr Gb,   F^-T Pbar^T F^-T = -k R Gh^T S^-1,
//   dR = R [w]x with (tr S I - S) w = axial(Pb' - Pb'^T),
// so rest = R (Gh^T T + 2 mu (Pb' - [w]x) + c1 S^-1) with the symmetric
// T = c2 S^-1 - k Sig: the closed form of fce_hess_mult (Elasticity.cpp:100-178)
// in terms of S, S^-1 and Gh (c1, c2 below; the J-clamped branch of the exact
// adjoint keeps only the F^-T variation).  False when Fc needs the SVD
// (det Fc <= 1e-6 |Fc|^3 or no Newton convergence); the caller then takes
// elastic_adj_svd.
__device__ __forceinline__ bool elastic_rest_polar(const M3<double>& Fc, const M3<float>& Gb, double k, double lam, double mu,
                                                   bool exact, M3<float>& rest)
{
    const double Ju = det(Fc);
    M3<double> R;
    if (!polar_newton(Fc, Ju, R)) return false;
    const double J = Ju < kJMin ? kJMin : Ju;
    Sym3 S;
    {
        const M3<double> RF = tmul(R, Fc);
        S.a[0] = RF.a[0];
        S.a[1] = RF.a[4];
        S.a[2] = RF.a[8];
        S.a[3] = 0.5 * (RF.a[1] + RF.a[3]);
        S.a[4] = 0.5 * (RF.a[2] + RF.a[6]);
        S.a[5] = 0.5 * (RF.a[5] + RF.a[7]);
    }
    const double fd = -k * (double(Gb.a[0]) + double(Gb.a[4]) + double(Gb.a[8]));
    const M3<double> Gh = tmul(R, mul(cast<double>(Gb), R));
    M3<double> N;  // Pb' first
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) N.a[3 * i + j] = -k * (Gh.a[3 * i] * S(0, j) + Gh.a[3 * i + 1] * S(1, j) + Gh.a[3 * i + 2] * S(2, j));
    double w0, w1, w2;
    {
        const double b0 = N.a[7] - N.a[5], b1 = N.a[2] - N.a[6], b2 = N.a[3] - N.a[1];
        Sym3 A;
        A.a[0] = S.a[1] + S.a[2];
        A.a[1] = S.a[0] + S.a[2];
        A.a[2] = S.a[0] + S.a[1];
        A.a[3] = -S.a[3];
        A.a[4] = -S.a[4];
        A.a[5] = -S.a[5];
        const Sym3 Ai = sym_inverse(A);
        w0 = Ai(0, 0) * b0 + Ai(0, 1) * b1 + Ai(0, 2) * b2;
        w1 = Ai(1, 0) * b0 + Ai(1, 1) * b1 + Ai(1, 2) * b2;
        w2 = Ai(2, 0) * b0 + Ai(2, 1) * b1 + Ai(2, 2) * b2;
    }
    double c1, c2;
    if (exact && Ju <= kJMin) {
        c1 = 0.0;
        c2 = k * lam * (J - 1.0) * J;
    } else {
        const double k2 = lam * (J - 1.0) * Ju;
        c1 = (lam * J * J + k2) * fd;
        c2 = k2 * k;
    }
    const double m2 = 2.0 * mu;
    const double kJ = lam * (J - 1.0) * J;
    const Sym3 Si = sym_inverse(S);
    Sym3 T;
#pragma unroll
    for (int d = 0; d < 6; ++d) T.a[d] = c2 * Si.a[d] - k * (m2 * (d < 3 ? S.a[d] - 1.0 : S.a[d]) + kJ * Si.a[d]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            N.a[3 * i + j] = m2 * N.a[3 * i + j] + c1 * Si(i, j) + (Gh.a[i] * T(0, j) + Gh.a[3 + i] * T(1, j) + Gh.a[6 + i] * T(2, j));
    // - 2 mu [w]x
    N.a[1] += m2 * w2;
    N.a[2] -= m2 * w1;
    N.a[3] -= m2 * w2;
    N.a[5] += m2 * w0;
    N.a[6] += m2 * w1;
    N.a[7] -= m2 * w0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            rest.a[3 * i + j] = float(R.a[3 * i] * N.a[j] + R.a[3 * i + 1] * N.a[3 + j] + R.a[3 * i + 2] * N.a[6 + j]);
    return true;
}

__device__ __forceinline__ M3<double> elastic_fc(const StepParams& sp, const AdjIO<float>& io, int p, int h)
{
    const int S = sp.stride;
    const M3<float> F = ldm(io.cur, FF, S, p);
    const M3<float> dF = ld_ctrl(io.ctrl, S, h);
    M3<double> Fc;
#pragma unroll
    for (int i = 0; i < 9; ++i) Fc.a[i] = double(F.a[i]) + double(dF.a[i]);
    return Fc;
}

// The elastic half on the polar path; false (nothing written) when Fc needs the SVD.
__device__ __forceinline__ bool elastic_adj_polar(const StepParams& sp, const AdjIO<float>& io, int p, int h, const M3<float>& Gb)
{
    const int S = sp.stride;
    const M3<double> Fc = elastic_fc(sp, io, p, h);
    const double vol = io.cst[KVOL * S + h];
    const double lam = io.cst[KLAM * S + h], mu = io.cst[KMU * S + h];
    M3<float> rest;
    if (!elastic_rest_polar(Fc, Gb, 3.0 * sp.inv_dx * sp.inv_dx * sp.dt * vol, lam, mu, sp.compat == 0, rest)) return false;
    elastic_adj_store(sp, io, p, h, rest);
    return true;
}

// The elastic half on the SVD path (inverted or nearly flat elements).
__device__ __forceinline__ void elastic_adj_svd(const StepParams& sp, const AdjIO<float>& io, int p, int h, const M3<float>& Gb)
{
    const int S = sp.stride;
    const M3<double> Fc = elastic_fc(sp, io, p, h);
    const double vol = io.cst[KVOL * S + h];
    const double lam = io.cst[KLAM * S + h], mu = io.cst[KMU * S + h];
    FCE e;
    e.svd = true;
    e.Ju = det(Fc);
    e.J = e.Ju < kJMin ? kJMin : e.Ju;
    const M3<double> cf = cofactor(Fc);
    const double inv = 1.0 / e.Ju;
#pragma unroll
    for (int i = 0; i < 9; ++i) e.Fit.a[i] = cf.a[i] * inv;
    svd3(Fc, e.U, e.s, e.V);
    e.R = mult(e.U, e.V);
    const double k = 3.0 * sp.inv_dx * sp.inv_dx * sp.dt * vol;
    const M3<double> Gbd = cast<double>(Gb);
    // Gb^T P first, so that P dies before H:dF
    const M3<double> GTP = tmul(Gbd, fce_pk1(Fc, e, lam, mu));
    M3<double> Pb = mul(Gbd, Fc);
#pragma unroll
    for (int i = 0; i < 9; ++i) Pb.a[i] *= -k;
    const M3<double> H = fce_hess_mult(e, lam, mu, Pb, sp.compat == 0);
    M3<float> rest;
#pragma unroll
    for (int i = 0; i < 9; ++i) rest.a[i] = float(-k * GTP.a[i] + H.a[i]);
    elastic_adj_store(sp, io, p, h, rest);
}

// ------------------------------------------------------- transposed P2G ----
// P2G^T + G_op^T + P_op_1^T, one thread per particle of the range.  The
// window load applies the transposed grid update (BackPropagation.cpp:46-59)
// to each node: pbar = vbar/m, mbar = -((v - dt f).vbar)/m where
// |m| > FLT_EPSILON, else 0, next to the node velocity v.  With the P2G
// affine matrix G = -k P Fc^T + m C (cached in the layer by P2G) and
// b = m(1 - dt drag) v, the gather over the 64 nodes gives
//   Gb   = sum w pbar d^T,  Cbar = m Gb,  vbar = m (1 - dt drag) sum w pbar,
//   xbar = x' + sum [gw (m mbar + (b + G d).pbar + v_i.vt + D (Ct : v_i d^T))
//                    - w (G^T pbar + D Ct^T v_i)],
// then the elastic part (FP64 polar, PK1 and H:dF):
//   Pbar = -k Gb Fc,  Fcbar = (1-s)(I + dt C')^T F' - k Gb^T P + H(Fc):Pbar.
template <bool FUSE>
__device__ __forceinline__ void p2g_adj_range(const Batch& bt, AdjIO<float> io, const int4 rg, float* lds)
{
    const StepParams sp = bt.sps[rg.z];
    io.grid += 4 * bt.nofs[rg.z];
    io.agrid += 4 * bt.nofs[rg.z];
    const int cap = bt.cap;
    const int S = sp.stride;
    const int cp = sp.compat;
    const Win W = range_window(sp, io.cur, S, rg.x, rg.y, cap);
    const float dt = float(sp.dt);
    const float fdt[3] = {dt * float(sp.f_ext[0]), dt * float(sp.f_ext[1]), dt * float(sp.f_ext[2])};
    const bool fsub = !(cp & 8);
    const float* gr = io.grid;
    const float* ag = io.agrid;
    const long long G = sp.G;
    // channels: pbar (3), mbar, v (3)
    auto ld = [&](long long node, float* ch) {
        const float m = gr[node];
        const float p[3] = {gr[G + node], gr[2 * G + node], gr[3 * G + node]};
        const float vb[3] = {ag[node], ag[G + node], ag[2 * G + node]};
        node_velocity(m, p, fdt, ch + 4);
        const bool ok = fabsf(m) > float(kMassEps);
        const float inv = 1.f / m;
        float dot = 0.f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float vi = fsub ? ch[4 + a] - fdt[a] : ch[4 + a];
            dot += vi * vb[a];
            ch[a] = ok ? vb[a] * inv : 0.f;
        }
        ch[3] = ok ? -dot * inv : 0.f;
    };
    win_load<7>(sp, W, lds, ld);
    __syncthreads();
    const float dx = float(sp.dx);
    const float Dinv = float(3.0 * sp.inv_dx * sp.inv_dx);
    const float mgw = -float(sp.inv_dx);
    const float sdrag = 1.f - dt * float(sp.drag);
    const float bsc = (cp & 4) ? 1.f : sdrag;
    const double kst = -3.0 * sp.inv_dx * sp.inv_dx * sp.dt;
    for (int p = rg.x + (int)threadIdx.x; p < rg.y; p += blockDim.x) {
        const float m = io.cst[KM * S + home_of(bt, p)];
        const M3<float> Gm = ldm(io.cur, FA, S, p);
        float x[3], v[3], vt[3];
        ld3(io.cur, FX, S, p, x);
        ld3(io.cur, FV, S, p, v);
        ld3(io.tmp, 0, S, p, vt);
        const M3<float> Ct = ldm(io.tmp, 3, S, p);
        float b[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) b[a] = m * bsc * v[a];
        Axis<float> st[3];
        stencil_f(sp, x, st, true);
        float d0s[4], d1s[4], d2s[4];
        offsets4(st[0], dx, d0s);
        offsets4(st[1], dx, d1s);
        offsets4(st[2], dx, d2s);
        // the sums over nodes are regrouped so that the node loop does the
        // least work: sum_i w_i G^T pbar_i = G^T (sum_i w_i pbar_i) and
        // sum_i w_i D Ct^T v_i = D Ct^T (sum_i w_i v_i) leave the loop, and the
        // weights' (i, j) factors leave the k-rows: a row sums only the
        // k-weighted channels (w2, w2 d2 and dw2 times them), which the row's
        // (i, j) factors then scale
        M3<float> CtD;
#pragma unroll
        for (int i = 0; i < 9; ++i) CtD.a[i] = Dinv * Ct.a[i];
        M3<float> Gb = m3_zero<float>();
        float xg[3] = {0.f, 0.f, 0.f}, pacc[3] = {0.f, 0.f, 0.f}, vacc[3] = {0.f, 0.f, 0.f};
        float wzd[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) wzd[k] = st[2].w[k] * d2s[k];
        // row sums: rs = sum_k w2 pbar, rsd = sum_k w2 d2 pbar, vs = sum_k w2 v,
        // s0 = sum_k w2 sgw, s2 = sum_k dw2 sgw.  The fixed window visits every
        // node, so its first node of a row starts the sums; the other modes
        // may skip it, so their sums start at 0 and restart after every row.
        float rs[3] = {0.f, 0.f, 0.f}, rsd[3] = {0.f, 0.f, 0.f}, vs[3] = {0.f, 0.f, 0.f}, s0 = 0.f, s2 = 0.f;
        auto acc = [&](auto mode, int i, int j, int kk, const float* ch) {
            constexpr bool first = decltype(mode)::value == 0;
            const float d0 = d0s[i], d1 = d1s[j], d2 = d2s[kk];
            float sgw = m * ch[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // q = b + G d, r = vt + D Ct d (the (i, j) parts are shared by the 4 nodes of a k-row)
                const float q = (b[a] + Gm.a[3 * a] * d0 + Gm.a[3 * a + 1] * d1) + Gm.a[3 * a + 2] * d2;
                const float r = (vt[a] + CtD.a[3 * a] * d0 + CtD.a[3 * a + 1] * d1) + CtD.a[3 * a + 2] * d2;
                sgw += ch[a] * q + ch[4 + a] * r;
                if (first && kk == 0) {
                    rs[a] = st[2].w[0] * ch[a];
                    rsd[a] = wzd[0] * ch[a];
                    vs[a] = st[2].w[0] * ch[4 + a];
                } else {
                    rs[a] += st[2].w[kk] * ch[a];
                    rsd[a] += wzd[kk] * ch[a];
                    vs[a] += st[2].w[kk] * ch[4 + a];
                }
            }
            if (first && kk == 0) {
                s0 = st[2].w[0] * sgw;
                s2 = st[2].dw[0] * sgw;
            } else {
                s0 += st[2].w[kk] * sgw;
                s2 += st[2].dw[kk] * sgw;
            }
        };
        auto row = [&](auto mode, int i, int j) {
            const float w01 = st[0].w[i] * st[1].w[j];
            const float a0 = w01 * d0s[i], a1 = w01 * d1s[j];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                pacc[a] += w01 * rs[a];
                Gb.a[3 * a] += a0 * rs[a];
                Gb.a[3 * a + 1] += a1 * rs[a];
                Gb.a[3 * a + 2] += w01 * rsd[a];
                vacc[a] += w01 * vs[a];
            }
            xg[0] += (st[0].dw[i] * st[1].w[j]) * s0;
            xg[1] += (st[0].w[i] * st[1].dw[j]) * s0;
            xg[2] += w01 * s2;
            if (decltype(mode)::value != 0) {
#pragma unroll
                for (int a = 0; a < 3; ++a) rs[a] = rsd[a] = vs[a] = 0.f;
                s0 = s2 = 0.f;
            }
            // keep the scheduler from interleaving the rows (their hoisted
            // row terms and loads would raise the register count)
            __builtin_amdgcn_sched_barrier(0);
        };
        if (W.fixed) gather64<7, 0>(sp, W, st, lds, ld, acc, row);
        else if (W.lds) gather64<7, 1>(sp, W, st, lds, ld, acc, row);
        else gather64<7, 2>(sp, W, st, lds, ld, acc, row);
#pragma unroll
        for (int a = 0; a < 3; ++a) xg[a] *= mgw;
        float xb[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            xb[a] = xg[a] - (Gm.a[a] * pacc[0] + Gm.a[3 + a] * pacc[1] + Gm.a[6 + a] * pacc[2]) -
                    (CtD.a[a] * vacc[0] + CtD.a[3 + a] * vacc[1] + CtD.a[6 + a] * vacc[2]);
        float gx1[3];
        ld3(io.g1, FX, S, p, gx1);
        const float vsc = m * sdrag;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            io.g0[(long long)(FX + a) * S + p] = gx1[a] + xb[a];
            io.g0[(long long)(FV + a) * S + p] = vsc * pacc[a];
        }
        M3<float> gC;
#pragma unroll
        for (int i = 0; i < 9; ++i) gC.a[i] = m * Gb.a[i];
        stm(io.g0, FC, S, p, gC);
        // vt, Ct in tmp are dead for this particle: Gb takes their place when
        // the elastic half runs in a kernel of its own (FUSE false, or the SVD
        // path of a particle deferred to k_elastic_defer)
        if (FUSE) {
            if (!elastic_adj_polar(sp, io, p, home_of(bt, p), Gb)) {
                stm(io.tmp, 0, S, p, Gb);
                io.defer[atomicAdd(io.defer_n, 1)] = make_int2(p, rg.z);
            }
        } else {
            stm(io.tmp, 0, S, p, Gb);
        }
    }
}

// OCC: the waves per SIMD the register allocation must allow (1: no limit)
template <bool FUSE, bool PERSIST, int OCC>
__global__ __launch_bounds__(kGat, OCC) void k_p2g_adj_blk(Batch bt, AdjIO<float> io)
{
    extern __shared__ float lds[];
    for_ranges<PERSIST>(bt, TileClear{}, [&](const int4 rg) { p2g_adj_range<FUSE>(bt, io, rg, lds); });
}

// Elastic half as a kernel of its own (Gb from tmp) over the block ranges, for
// persistent launches (k_elastic_flat otherwise): one thread per particle of
// a range; the SVD path is deferred to k_elastic_defer as in the fused gather.
__device__ __forceinline__ void elastic_range(const Batch& bt, const AdjIO<float>& io, const int4 rg)
{
    const StepParams sp = bt.sps[rg.z];
    const int S = sp.stride;
    for (int p = rg.x + (int)threadIdx.x; p < rg.y; p += blockDim.x)
        if (!elastic_adj_polar(sp, io, p, home_of(bt, p), ldm(io.tmp, 0, S, p))) io.defer[atomicAdd(io.defer_n, 1)] = make_int2(p, rg.z);
}

// tc: the adjoint grid, cleared for the next step by appended workgroups
template <bool PERSIST>
__global__ __launch_bounds__(kGat) void k_elastic_adj(Batch bt, AdjIO<float> io, TileClear tc)
{
    for_ranges<PERSIST>(bt, tc, [&](const int4 rg) { elastic_range(bt, io, rg); });
}

// The elastic half over every particle slot [0, n), one thread per particle:
// every lane of a workgroup holds a particle (a block range leaves lanes
// idle), and a thread issues its particle's loads at once instead of after
// the previous particle's FP64 work.  A particle that needs the SVD takes it
// here, through a call (no deferred list).  poff: the slot prefix of the
// instances (B + 1 entries; instance i holds slots [poff[i], poff[i + 1])).
// Workgroups from nwork on clear tc.
// The SVD path of one particle out of line: k_elastic_flat's register
// allocation stays that of its polar path (the call, rare, spills instead).
__device__ __noinline__ void elastic_svd_call(const StepParams* sps, int inst, AdjIO<float> io, int p, int h)
{
    const StepParams sp = sps[inst];
    elastic_adj_svd(sp, io, p, h, ldm(io.tmp, 0, sp.stride, p));
}

__global__ __launch_bounds__(kGat, 3) void k_elastic_flat(Batch bt, AdjIO<float> io, TileClear tc, unsigned nwork, int n,
                                                       const int* poff)
{
    if (blockIdx.x >= nwork) {
        appended_clear(bt, tc, nwork);
        return;
    }
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= n) return;
    int lo = 0, hi = bt.B - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (poff[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    const StepParams sp = bt.sps[lo];
    const int h = home_of(bt, p);
    if (!elastic_adj_polar(sp, io, p, h, ldm(io.tmp, 0, sp.stride, p))) elastic_svd_call(bt.sps, lo, io, p, h);
}

// The elastic half of the particles deferred to the SVD path (Gb in tmp):
// workgroups [0, nwork) share the list; the last of them to finish resets
// its two counters for the next step (every one has read the entry count
// before it adds to the done count).  Workgroups from nwork on clear tc.
__global__ __launch_bounds__(kGat) void k_elastic_defer(Batch bt, AdjIO<float> io, TileClear tc, unsigned nwork)
{
    if (blockIdx.x >= nwork) {
        appended_clear(bt, tc, nwork);
        return;
    }
    const int n = io.defer_n[0];
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(nwork * blockDim.x)) {
        const int2 d = io.defer[i];
        const StepParams sp = bt.sps[d.y];
        elastic_adj_svd(sp, io, d.x, home_of(bt, d.x), ldm(io.tmp, 0, sp.stride, d.x));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(io.defer_n + 1, 1) == (int)nwork - 1) {
            atomicExch(io.defer_n, 0);
            atomicExch(io.defer_n + 1, 0);
        }
    }
}

// --------------------------------------------------- batched grid kernels --
// Zero the flagged tiles of a grid (clear_tiles) in a launch of its own.
__global__ __launch_bounds__(kGat) void k_clear_tiles(Batch bt, float* grid, unsigned* flags, long long ntiles)
{
    clear_tiles(bt, grid, flags, ntiles, ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6,
                ((long long)gridDim.x * blockDim.x) >> 6);
}

__global__ __launch_bounds__(kGat) void k_set_flags(unsigned* flags, long long ntiles)
{
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += (long long)gridDim.x * blockDim.x)
        flags[t] = 1u;
}

// EndLayerMassLoss node terms (CompGraph.cpp:143-187) in FP64.  blockIdx.y is
// the instance; gm and target hold one node mass per node (instance i at
// nofs[i]); workgroup partial sums go to part[instance * gridDim.x + x].
__global__ __launch_bounds__(kGat) void k_loss_b(Batch bt, const float* gm, const float* target, double* dLdm, double* part)
{
    __shared__ double red[kGat / 64];
    const int inst = blockIdx.y;
    const long long n0 = bt.nofs[inst], n1 = bt.nofs[inst + 1];
    double local = 0.0;
    for (long long q = n0 + (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n1; q += (long long)gridDim.x * blockDim.x) {
        const double cm = gm[q], tm = target[q];
        const double c = log(cm + 1.0 + 1e-4) - log(tm + 1.0 + 1e-4);
        double l = 0.5 * c * c;
        double g = c / (cm + 1.0 + 1e-4);
        if (cm < 1e-3) {
            const double diff = 1e-3 - cm;
            l += 10.0 * diff * diff;
            g -= 20.0 * diff;
        }
        local += l;
        dLdm[q] = g;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < kGat / 64; ++w) sum += red[w];
        part[(long long)inst * gridDim.x + blockIdx.x] = sum;
    }
}

// EndLayerMassLoss particle gather (CompGraph.cpp:192-241) in FP64 over the ranges
__global__ __launch_bounds__(kGat) void k_loss_particles_b(Batch bt, const float* layer, const float* cst, const double* dLdm, float* g)
{
    const int4 rg = bt.ranges[blockIdx.x];
    const StepParams sp = bt.sps[rg.z];
    const int S = sp.stride;
    const double* dl = dLdm + bt.nofs[rg.z];
    for (int p = rg.x + (int)threadIdx.x; p < rg.y; p += blockDim.x) {
        Axis<double> st[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            axis_weights<double>((double(layer[(long long)(FX + a) * S + p]) - sp.min_point[a]) * sp.inv_dx, st[a], true);
        double gx[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int gi = st[0].base + i;
            if (gi < 0 || gi >= sp.dim[0]) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int gj = st[1].base + j;
                if (gj < 0 || gj >= sp.dim[1]) continue;
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int gk = st[2].base + kk;
                    if (gk < 0 || gk >= sp.dim[2]) continue;
                    const double d = dl[((long long)gi * sp.dim[1] + gj) * sp.dim[2] + gk];
                    gx[0] += d * st[0].dw[i] * st[1].w[j] * st[2].w[kk];
                    gx[1] += d * st[0].w[i] * st[1].dw[j] * st[2].w[kk];
                    gx[2] += d * st[0].w[i] * st[1].w[j] * st[2].dw[kk];
                }
            }
        }
        const double sc = -double(cst[KM * S + home_of(bt, p)]) * sp.inv_dx;
#pragma unroll
        for (int a = 0; a < 3; ++a) g[(long long)(FX + a) * S + p] = float(sc * gx[a]);
#pragma unroll
        for (int f = FV; f < kLayerFields; ++f) g[(long long)f * S + p] = 0.f;
    }
}

// ------------------------------------------------------------ launchers ----
// LDS: max(cap, fixed-window slots) nodes of NC 64-bit words (P2G scatters), NC 32-bit words (the transposed
// G2P) or NC floats (gathers)
static size_t win_nodes(int cap) { return (size_t)(cap > kFWS ? cap : kFWS); }


void launch_p2g_mass_blk(const Batch& b, const float* lay, const float* cst, float* gm, hipStream_t s)
{
    if (b.nranges <= 0) return;
    hipLaunchKernelGGL(k_p2g_mass_blk, dim3(b.nranges), dim3(kScat), win_nodes(b.cap) * sizeof(unsigned long long), s, b, lay, cst,
                       gm);
}

// workgroups appended for a tile clear: a wave per 64 tiles, 4 waves per block
static unsigned clear_blocks(const Batch& b, const TileClear& tc)
{
    if (!tc.grid) return 0;
    long long nb = (num_tiles(b.Gtot) + 4 * 64 - 1) / (4 * 64);
    if (nb > 4096) nb = 4096;
    return (unsigned)nb;
}

// The grid of a range kernel and the batch it takes: one workgroup per range
// plus the appended clear workgroups, or, with a queue, a persistent launch of
// as many workgroups as fit on the chip at once (at most one per range) on the
// counters of the kernel's class (cls 0: forward, 1: adjoint; a forward
// recompute on the second stream overlaps only adjoint kernels).
template <typename Kern>
static dim3 range_grid(const Batch& b, Batch& bq, const TileClear& tc, Kern kernel, int block, size_t lds, int cls)
{
    bq = b;
    if (!b.queue) return dim3(b.nranges + clear_blocks(b, tc));
    bq.queue = b.queue + 2 * cls;
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
        cus = 1;
    long long g = (long long)per_cu * cus;
    if (g > b.nranges) g = b.nranges;
    return dim3((unsigned)(g < 1 ? 1 : g));
}

void launch_p2g_blk(const Batch& b, const StepIO<float>& io, const TileClear& tc, hipStream_t s)
{
    if (b.nranges <= 0) {
        if (tc.grid) launch_clear_tiles(b, tc.grid, tc.flags, s);
        return;
    }
    const size_t lds = 4 * win_nodes(b.cap) * sizeof(unsigned long long);
    Batch bq;
    if (b.queue) {
        const dim3 g = range_grid(b, bq, tc, k_p2g_blk<true>, kP2G, lds, 0);
        hipLaunchKernelGGL(k_p2g_blk<true>, g, dim3(kP2G), lds, s, bq, io, tc);
    } else {
        const dim3 g = range_grid(b, bq, tc, k_p2g_blk<false>, kP2G, lds, 0);
        hipLaunchKernelGGL(k_p2g_blk<false>, g, dim3(kP2G), lds, s, bq, io, tc);
    }
}

void launch_step_blk(const Batch& b, const StepIO<float>& gio, const StepIO<float>& pio, const TileClear& tc, hipStream_t s)
{
    if (b.nranges <= 0) {
        if (tc.grid) launch_clear_tiles(b, tc.grid, tc.flags, s);
        return;
    }
    // the scatter's window (4 64-bit words) and the gather's (3 floats) of at most kFWS nodes each
    const size_t lds = kFWS * (4 * sizeof(unsigned long long) + 3 * sizeof(float));
    Batch bc = b;
    if (bc.cap > kFWS) bc.cap = kFWS;
    Batch bq;
    if (b.queue) {
        const dim3 g = range_grid(bc, bq, tc, k_step_blk<true>, kP2G, lds, 0);
        hipLaunchKernelGGL(k_step_blk<true>, g, dim3(kP2G), lds, s, bq, gio, pio, tc);
    } else {
        const dim3 g = range_grid(bc, bq, tc, k_step_blk<false>, kP2G, lds, 0);
        hipLaunchKernelGGL(k_step_blk<false>, g, dim3(kP2G), lds, s, bq, gio, pio, tc);
    }
}

void launch_g2p_blk(const Batch& b, const StepIO<float>& io, const TileClear& tc, hipStream_t s)
{
    if (b.nranges <= 0) {
        if (tc.grid) launch_clear_tiles(b, tc.grid, tc.flags, s);
        return;
    }
    const size_t lds = 3 * win_nodes(b.cap) * sizeof(float);
    Batch bq;
    if (b.queue) {
        const dim3 g = range_grid(b, bq, tc, k_g2p_blk<true>, kGat, lds, 0);
        hipLaunchKernelGGL(k_g2p_blk<true>, g, dim3(kGat), lds, s, bq, io, tc);
    } else {
        const dim3 g = range_grid(b, bq, tc, k_g2p_blk<false>, kGat, lds, 0);
        hipLaunchKernelGGL(k_g2p_blk<false>, g, dim3(kGat), lds, s, bq, io, tc);
    }
}

void launch_g2p_adj_blk(const Batch& b, const AdjIO<float>& io, hipStream_t s)
{
    if (b.nranges <= 0) return;
    const size_t lds = 3 * win_nodes(b.cap) * sizeof(unsigned);
    Batch bq;
    if (b.queue) {
        const dim3 g = range_grid(b, bq, TileClear{}, k_g2p_adj_blk<true>, kG2PAdj, lds, 1);
        hipLaunchKernelGGL(k_g2p_adj_blk<true>, g, dim3(kG2PAdj), lds, s, bq, io);
    } else {
        const dim3 g = range_grid(b, bq, TileClear{}, k_g2p_adj_blk<false>, kG2PAdj, lds, 1);
        hipLaunchKernelGGL(k_g2p_adj_blk<false>, g, dim3(kG2PAdj), lds, s, bq, io);
    }
}

// the deferred-list workgroups of k_elastic_defer
constexpr unsigned kDeferBlocks = 16;

void launch_p2g_adj_blk(const Batch& b, const AdjIO<float>& io, const TileClear& tc, int elastic, int n, const int* poff,
                        hipStream_t s)
{
    if (b.nranges <= 0) {
        if (tc.grid) launch_clear_tiles(b, tc.grid, tc.flags, s);
        return;
    }
    const size_t lds = 7 * win_nodes(b.cap) * sizeof(float);
    Batch bq;
    dim3 g;
    TileClear dtc = tc;  // the clear rides on the elastic kernel when it has one, else on k_elastic_defer
    if (elastic == 0) {
        if (b.queue) {
            g = range_grid(b, bq, TileClear{}, k_p2g_adj_blk<false, true, 3>, kGat, lds, 1);
            hipLaunchKernelGGL((k_p2g_adj_blk<false, true, 3>), g, dim3(kGat), lds, s, bq, io);
            g = range_grid(b, bq, tc, k_elastic_adj<true>, kGat, 0, 1);
            hipLaunchKernelGGL(k_elastic_adj<true>, g, dim3(kGat), 0, s, bq, io, tc);
        } else {
            hipLaunchKernelGGL((k_p2g_adj_blk<false, false, 3>), dim3(b.nranges), dim3(kGat), lds, s, b, io);
            const unsigned nwork = (unsigned)((n + kGat - 1) / kGat);
            hipLaunchKernelGGL(k_elastic_flat, dim3(nwork + clear_blocks(b, tc)), dim3(kGat), 0, s, b, io, tc, nwork, n, poff);
            return;  // k_elastic_flat runs the SVD path itself: nothing is deferred
        }
        dtc = TileClear{};
    } else if (elastic == 1) {
        if (b.queue) {
            g = range_grid(b, bq, TileClear{}, k_p2g_adj_blk<true, true, 1>, kGat, lds, 1);
            hipLaunchKernelGGL((k_p2g_adj_blk<true, true, 1>), g, dim3(kGat), lds, s, bq, io);
        } else {
            hipLaunchKernelGGL((k_p2g_adj_blk<true, false, 1>), dim3(b.nranges), dim3(kGat), lds, s, b, io);
        }
    } else {
        if (b.queue) {
            g = range_grid(b, bq, TileClear{}, k_p2g_adj_blk<true, true, 3>, kGat, lds, 1);
            hipLaunchKernelGGL((k_p2g_adj_blk<true, true, 3>), g, dim3(kGat), lds, s, bq, io);
        } else {
            hipLaunchKernelGGL((k_p2g_adj_blk<true, false, 3>), dim3(b.nranges), dim3(kGat), lds, s, b, io);
        }
    }
    hipLaunchKernelGGL(k_elastic_defer, dim3(kDeferBlocks + clear_blocks(b, dtc)), dim3(kGat), 0, s, b, io, dtc, kDeferBlocks);
}

void launch_clear_tiles(const Batch& b, float* grid, unsigned* flags, hipStream_t s)
{
    const long long nt = num_tiles(b.Gtot);
    if (nt <= 0) return;
    long long nb = (nt + 4 * 64 - 1) / (4 * 64);  // a wave per 64 tiles, 4 waves per block
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_clear_tiles, dim3((unsigned)nb), dim3(kGat), 0, s, b, grid, flags, nt);
}

void launch_set_flags(unsigned* flags, long long ntiles, hipStream_t s)
{
    if (ntiles <= 0) return;
    long long nb = (ntiles + kGat - 1) / kGat;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_set_flags, dim3((unsigned)nb), dim3(kGat), 0, s, flags, ntiles);
}

void launch_loss_b(const Batch& b, const float* gm, const float* target, double* dLdm, double* part, double* acc, long long gmax,
                   hipStream_t s)
{
    if (b.Gtot <= 0 || b.B <= 0) return;
    long long nb = (gmax + kGat - 1) / kGat;
    const long long lim = loss_partials(b.B) / b.B;
    if (nb > lim) nb = lim;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_loss_b, dim3((unsigned)nb, (unsigned)b.B), dim3(kGat), 0, s, b, gm, target, dLdm, part);
    launch_sum_rows(part, b.B, (int)nb, acc, s);
}

void launch_loss_particles_b(const Batch& b, const float* layer, const float* cst, const double* dLdm, float* g, hipStream_t s)
{
    if (b.nranges <= 0) return;
    hipLaunchKernelGGL(k_loss_particles_b, dim3(b.nranges), dim3(kGat), 0, s, b, layer, cst, dLdm, g);
}

}  // namespace dmpm
