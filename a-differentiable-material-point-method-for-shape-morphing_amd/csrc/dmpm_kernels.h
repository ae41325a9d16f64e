// Kernel-side parameter blocks and host launch wrappers of the MLS-MPM step.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace dmpm {

// Particle layer in HBM: structure of arrays, field f of particle p at
// base[f * stride + p].  Fields: x0..2, v0..2, C00..C22, F00..F22.
enum : int { FX = 0, FV = 3, FC = 6, FF = 15, kLayerFields = 24 };
// FP32 layers carry 9 more fields: the P2G affine matrix
// A = m C - (3/dx^2) dt vol P(Fc) Fc^T of the step that starts at the layer,
// written by P2G and read back by the transposed P2G (no second polar solve).
enum : int { FA = 24, kLayerFieldsF32 = 33 };
// Per-particle constants: m, vol, lam, mu.
enum : int { KM = 0, KVOL = 1, KLAM = 2, KMU = 3, kConstFields = 4 };
// Grid of one step: m then v (or momentum before the update) per axis.
enum : int { kGridFields = 4 };

struct StepParams {
    int n;            // particles
    int stride;       // SoA field stride (>= n)
    int dim[3];
    long long G;      // nodes
    double dx, inv_dx;
    double min_point[3];
    double dt, drag, smoothing;
    double f_ext[3];
    int compat;       // reference-compat bits (same as the C ABI)
};

template <typename Real>
struct StepIO {
    const Real* cur;   // layer n state
    Real* next;        // layer n+1 state (forward)
    const Real* ctrl;  // dFc_n (9 fields); FP32 G2P: null when cur's F field holds Fc_n = F_n + dFc_n
    const Real* cst;   // constants
    Real* grid;        // grid m, p -> v (FP64); FP32 keeps m, p (kGridFields * G)
    Real* aux;         // writable alias of cur for derived fields (FP32: A), or null
    unsigned* gflags;  // FP32: tile flags of grid (set where P2G writes), else null
    int lean = 0;      // FP32 fused step: the layer written is read only for x and F (v, C not stored)
    int fc = 0;        // FP32 fused step: the layer written holds Fc = F + dFc (pio.ctrl) in its F field
};

template <typename Real>
struct AdjIO {
    const Real* cur;     // layer n state
    const Real* next;    // layer n+1 state (C' is read)
    const Real* ctrl;    // dFc_n; FP32: null when cur's F field holds Fc_n
    const Real* cst;
    const Real* grid;    // grid of step n: m, v after the update (FP64) or m, p (FP32)
    Real* agrid;         // adjoint grid: vbar -> pbar (3G), mbar (G) (FP64); vbar (FP32)
    const Real* g1;      // gradients of layer n+1 (24 fields, same layout as a layer)
    Real* g0;            // gradients of layer n
    Real* gctrl;         // dL/d(dFc_n) (9 fields), or null: not stored (no reader)
    Real* tmp;           // per particle vt (3) and Ct (9), from the transposed G2P
    const unsigned* gflags;  // FP32: tile flags of grid, else null
    unsigned* aflags;        // FP32: tile flags of agrid (set where the transposed G2P writes)
    // FP32: particles whose elastic half needs the SVD, deferred by the
    // transposed P2G to a kernel of their own ({slot, instance}, S entries),
    // and two counters (entries, workgroups done) that start and end at 0
    int2* defer;
    int* defer_n;
};

// most workgroups of a one-row FP64 reduction (loss, norm)
constexpr int kLossBlocks = 2048;

// control update of one layer (PointCloud::Descend_*): rule 0 Adam, 1 AMSGrad, 2 Lion, 3 gradient descent
struct DescendParams {
    int rule;
    float alpha, gradient_norm, beta1, beta2, epsilon;
    float bc1, bc2;  // 1 - beta1^t, 1 - beta2^t
};

template <typename Real> void launch_fill(Real* p, long long count, Real value, hipStream_t s);
template <typename Real> void launch_descend(long long count, const DescendParams& dp, Real* dFc, const Real* g, Real* mom,
                                             Real* vec, Real* vmax, hipStream_t s);
// sum of squares in FP64 (partials in part, at most kLossBlocks), fixed-order total in *out
template <typename Real> void launch_sumsq(const Real* a, long long count, double* part, double* out, hipStream_t s);
template <typename Real> void launch_p2g(const StepParams& sp, const StepIO<Real>& io, hipStream_t s);
template <typename Real> void launch_grid_update(const StepParams& sp, Real* grid, hipStream_t s);
template <typename Real> void launch_g2p(const StepParams& sp, const StepIO<Real>& io, hipStream_t s);
template <typename Real> void launch_p2g_mass(const StepParams& sp, const Real* x, const Real* m, int stride, Real* gm, hipStream_t s);
// the same from host-layout points: x AoS [n][3] doubles less org, m [n] doubles (n = sp.n)
template <typename Real> void launch_p2g_mass_aos(const StepParams& sp, const double* x, const double* m, const double org[3], Real* gm,
                                                  hipStream_t s);
template <typename Real> void launch_volumes(const StepParams& sp, const Real* x, Real* cst, const Real* gm, hipStream_t s);
// loss node terms: dL/dm in FP64, per-workgroup partial sums (at most
// kLossBlocks) in part, their fixed-order sum in *loss
template <typename Real> void launch_loss(const StepParams& sp, const Real* gm, const Real* target, double* dLdm,
                                          double* part, double* loss, hipStream_t s);
// out[r] = sum of part[r * len .. r * len + len) in a fixed order
void launch_sum_rows(const double* part, int rows, int len, double* out, hipStream_t s);
template <typename Real> void launch_loss_particles(const StepParams& sp, const Real* layer, const Real* cst,
                                                    const double* dLdm, Real* g, hipStream_t s);
template <typename Real> void launch_g2p_adj(const StepParams& sp, const AdjIO<Real>& io, hipStream_t s);
template <typename Real> void launch_grid_adj(const StepParams& sp, const AdjIO<Real>& io, hipStream_t s);
template <typename Real> void launch_p2g_adj(const StepParams& sp, const AdjIO<Real>& io, hipStream_t s);

// FP32 path: many instances (independent morphs) per launch.  Particles of
// all instances share one SoA layer (instance-major, block-sorted within an
// instance); each instance's grid is a block of the concatenated grids.
// Work items are particle ranges of one 4^3-cell block of one instance
// (ranges[i] = {begin, end, instance, 0} in device order, at most 512
// particles); the node window is staged in LDS (cap = general-window
// capacity in nodes; larger windows fall back to direct HBM access).
struct Batch {
    const StepParams* sps;  // per instance: dims, G, dx, min point, dt, ...
    const long long* nofs;  // node prefix, B+1 entries: instance i owns nodes [nofs[i], nofs[i+1])
    const int4* ranges;
    int nranges;
    int B;
    long long Gtot;         // total nodes
    int cap;
    long long gmax;         // nodes of the largest instance
    unsigned long long* wstats;  // P2G workgroups per window mode (fixed, general, HBM), or null
    const int* home;        // slot of each particle slot in the order of the constants, controls and
                            // control gradients (null: the same order)
    int* queue;             // persistent launches: 4 zeroed counters (ranges taken, workgroups done) for
                            // the forward kernels, then for the adjoint kernels; null: a workgroup per range
};
// 4-channel grids (m or mbar, then 3 vector channels) of instance i start at 4 * nofs[i];
// 1-channel arrays (target, dL/dm) at nofs[i].
//
// Tile flags (FP32 grids): flat node q (instance i's node l is q = nofs[i] + l)
// lies in tile q >> kTileShift; a grid's flag array holds one word per tile,
// non-zero where a kernel may have written a node of the tile since the last
// clear.  Nodes of unflagged tiles are 0 in every channel, so the clear and
// the node-parallel kernels visit flagged tiles only.
constexpr int kTileShift = 5;
__host__ __device__ inline long long num_tiles(long long Gtot) { return (Gtot + (1LL << kTileShift) - 1) >> kTileShift; }
void launch_clear_tiles(const Batch& b, float* grid, unsigned* flags, hipStream_t s);
// a tile clear carried by a kernel that does not touch the grid (null: none)
struct TileClear {
    float* grid;
    unsigned* flags;
};
void launch_set_flags(unsigned* flags, long long ntiles, hipStream_t s);
void launch_p2g_blk(const Batch& b, const StepIO<float>& io, const TileClear& tc, hipStream_t s);
// G2P of a step (gio) and P2G of the next (pio) in one launch; tc: the grid after
void launch_step_blk(const Batch& b, const StepIO<float>& gio, const StepIO<float>& pio, const TileClear& tc, hipStream_t s);
void launch_p2g_mass_blk(const Batch& b, const float* lay, const float* cst, float* gm, hipStream_t s);
void launch_g2p_blk(const Batch& b, const StepIO<float>& io, const TileClear& tc, hipStream_t s);
void launch_g2p_adj_blk(const Batch& b, const AdjIO<float>& io, hipStream_t s);
// elastic: where the elastic half of the transposed P2G runs (DMPM_OPT_ELASTIC): 0 in a
// kernel of its own, 1 in the gather kernel, 2 there at 3 waves per SIMD; particles that
// need the SVD run in k_elastic_defer, whose appended workgroups clear tc
// n particle slots, poff the instances' slot prefix (device, B + 1 entries): the elastic half in a kernel of
// its own runs one thread per particle (k_elastic_flat), or per block range in a persistent launch
void launch_p2g_adj_blk(const Batch& b, const AdjIO<float>& io, const TileClear& tc, int elastic, int n, const int* poff,
                        hipStream_t s);
// loss over all instances from node masses gm (instance i at nofs[i]):
// per-instance losses in acc[0..B) (fixed-order sums of workgroup partials
// kept in part, at most loss_partials(B) doubles); gmax = largest instance node count
inline long long loss_partials(int B) { return (long long)B * (4096 / B > 16 ? 4096 / B : 16); }
void launch_loss_b(const Batch& b, const float* gm, const float* target, double* dLdm, double* part, double* acc, long long gmax,
                   hipStream_t s);
void launch_loss_particles_b(const Batch& b, const float* layer, const float* cst, const double* dLdm, float* g, hipStream_t s);

// the reference's ForwardTimeStep phases one by one (dmpm_phases.hip): FP64,
// host-layout AoS arrays; g carries the grid (dim, G, dx, inv_dx, min_point),
// n, dt, drag and f_ext (grid_op's gravity point when point_mode is set)
void launch_phase_p2g(const StepParams& g, const double* x, const double* v, const double* C, const double* F, const double* dFc,
                      const double* P, const double* m, const double* vol, double* gm, double* gp, hipStream_t s);
void launch_phase_grid_op(const StepParams& g, const double* gm, const double* gp, double* gv, int point_mode, double mag,
                          hipStream_t s);
void launch_phase_g2p(const StepParams& g, const double* x, const double* gv, double* vn, double* Cn, hipStream_t s);
void launch_phase_p_op2(int n, const double* x, const double* F, const double* dFc, const double* vn, const double* Cn, double sm,
                        double dt, double* xn, double* Fn, hipStream_t s);
void launch_phase_g2g_mass(const StepParams& g1, const StepParams& g2, const double* m1, double* m2, hipStream_t s);

// Device order of the FP32 particles (dmpm_sort.hip): per instance by T^3-cell
// block (T = 2^tshift, 4 or 8), rank in cell, cell in block.  Buffers of N entries; after
// launch_sort_order, v1[d] is the current slot of the particle that goes to
// slot d and blk[d] its global block (instance i's blocks from boff[i]).
struct SortArgs {
    const StepParams* sps;   // per instance (device)
    const int* poff;         // slot prefix, B+1 entries (device)
    const long long* boff;   // block prefix, B+1 entries (device)
    int B;
    int tshift;              // log2 of the block edge in cells (2 or 3)
    long long nblocks;       // boff[B]
    const float* x;          // layer holding the positions (fields FX.., stride S)
    long long S;
    int N;
    unsigned long long *k0, *k1;
    int *v0, *v1, *head, *blk;
    void* temp;
    size_t temp_bytes;       // sort_temp_bytes(N, nblocks)
};
size_t sort_temp_bytes(int N, long long nblocks);
hipError_t launch_sort_order(const SortArgs& a, hipStream_t s);
// dst[f S + d] = src[f S + perm[d]], f < nf
template <typename T> void launch_gather(const T* src, T* dst, long long S, int N, int nf, const int* perm, hipStream_t s);
// dst[f S + idx[d]] = src[f S + d] for nf fields
template <typename T> void launch_scatter(const T* src, T* dst, long long S, int N, int nf, const int* idx, hipStream_t s);
// inv[a[d]] = d
void launch_invert(const int* a, int* inv, int N, hipStream_t s);
void launch_iota(int* a, int N, hipStream_t s);

// batched constitutive evaluations (AoS 9 doubles per matrix)
void launch_eval_pk1(int n, const double* F, const double* lam, const double* mu, double* P, hipStream_t s);
void launch_eval_hess(int n, const double* F, const double* lam, const double* mu, const double* dF, double* out,
                      int exact, hipStream_t s);
void launch_eval_svd(int n, const double* F, double* U, double* sig, double* V, hipStream_t s);
void launch_eval_energy(int n, const double* F, const double* lam, const double* mu, double* psi, hipStream_t s);
// per-slot det(F + dFc) into out[0..S) and elastic energy into out[S..2S) of one layer
template <typename Real>
void launch_point_diag(long long S, const Real* layer, const Real* ctrl, const Real* cst, double* out, hipStream_t s);

}  // namespace dmpm
