// C ABI, device memory and episode driver of the differentiable MLS-MPM step.
//
// A context holds one morph (or, FP32, a batch of independent morphs, each
// with its own particles and grid) on one GPU: per-particle constants, per-layer
// controls, the forward trajectory (every layer, or checkpoints plus one
// segment tape), two ping-pong gradient layers and the work grids.  The
// passes follow CompGraph (CompGraph.cpp:246-267): forward over layers
// [start, L-1), EndLayerMassLoss on the last layer, backward from L-2 down to
// the control layer.  In checkpoint mode the forward keeps layer 0, every
// ckpt_stride-th layer and the whole last segment; the backward recomputes
// one segment at a time into the tape before stepping back through it.
#include "../../include/dmpm.h"
#include "dmpm_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

// Staging of dmpm_target_from_points_noalloc, which allocates nothing: the
// points' AoS positions (3 n doubles), then their masses (n doubles), for at
// most kNoAllocPoints points (32 MiB, held on every GPU the library runs on
// until the process ends).  Every call on one GPU shares that GPU's copy.
constexpr long long kNoAllocPoints = 1LL << 20;
static __device__ __align__(256) double scratch[4 * kNoAllocPoints];

namespace {

using namespace dmpm;

thread_local std::string g_err;

// Device address of each GPU's copy of scratch (the array's name is a
// host-side stand-in, not a device address): looked up once per device, with
// the device current, by the first dmpm_create on it.
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(DMPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct ProfRec {
    int kernel;
    hipEvent_t a, b;
};

struct GraphEntry {
    long long key;
    hipGraphExec_t exec;
};

}  // namespace

struct dmpm_ctx {
    dmpm_config cfg{};
    bool f32 = false;
    size_t esz = 8;
    int lf = kLayerFields;  // fields per layer slot (FP32: + the cached P2G affine matrix)
    int N = 0, S = 0, L = 0, K = 0;
    long long G = 0;
    hipStream_t stream = nullptr;
    // device buffers
    void* cst = nullptr;      // 4 x S
    void* ctrl = nullptr;     // (L-1) x 9 x S
    void* layers = nullptr;   // layer slots x 24 x S
    void* grids = nullptr;    // grid slots x 4 x G
    void* wgrid = nullptr;    // 4 x G: loss mass + dLdm / checkpoint-mode forward grid (even steps)
    void* wgrid2 = nullptr;   // 4 x G: checkpoint-mode forward grid (steps i % 3 == 1; FP64: odd steps)
    void* wgrid3 = nullptr;   // 4 x G (FP32): checkpoint-mode forward grid (steps i % 3 == 2)
    void* agrid = nullptr;    // 4 x G: vbar/pbar + mbar
    // FP32: tile flags of every grid buffer (grid slots, then wgrid, agrid, wgrid2, wgrid3)
    unsigned* tflags = nullptr;
    long long ntiles = 0;
    size_t grid_slots = 0;
    // FP32 checkpoint tape: the grid of every checkpoint step (its own buffer after the two tapes' grids)
    int ck_grids = 0;
    // gradient layers: the last layer's (the loss seed, kept across backward
    // passes as the reference keeps the last layer's dLdx) in gbuf[2], the
    // others ping-pong between gbuf[0] and gbuf[1] by layer parity
    void* gbuf[3] = {nullptr, nullptr, nullptr};
    void* gctrl = nullptr;    // (L-1) x 9 x S or 9 x S
    void* tmp = nullptr;      // 12 x S
    void* target = nullptr;   // G
    double* dacc = nullptr;   // FP64 accumulators: per-instance losses [0, B), norm at B
    double* dpart = nullptr;  // per-workgroup partial sums of the loss
    std::vector<void*> opt;   // per control layer: optimizer state m, v, vmax (27 x S), allocated on first use
    uint64_t bytes = 0;
    // instances: particles of instance i are API/device slots [poff[i], poff[i+1]),
    // its grid nodes [nofs[i], nofs[i+1]) (4-channel grids at 4 * nofs[i])
    int B = 1;
    std::vector<dmpm_instance> inst;
    std::vector<int> poff;
    std::vector<long long> nofs;
    // device positions are kept relative to each instance's grid centre
    // (origin[3 i + a]): an FP32 coordinate near 0 has a finer spacing than one
    // near the far side of the box, and every step's weights and node offsets
    // inherit that spacing.  Uploads subtract it, downloads add it back, and the
    // kernels see min_point - origin.
    std::vector<double> origin;
    long long gmax = 0;
    std::vector<double> inst_loss;
    // FP32 path: device order of the particles (sorted on the device by 4^3-cell
    // block per instance); pos[p] is the device slot of API particle p (empty =
    // identity), d_api its inverse on the device
    std::vector<int> pos;
    int* d_api = nullptr;           // API index of each device slot (N)
    int* d_poff = nullptr;          // slot prefix per instance (B+1)
    long long* d_boff = nullptr;    // block prefix per instance (B+1)
    std::vector<long long> boff;
    int tile = 4;                   // block edge in cells (4 or 8): the LDS tile of a workgroup's particles
    int* queue = nullptr;           // persistent launches: range counters of the forward and adjoint kernels (4),
                                    // then the two counters of the deferred elastic list (2)
    int2* defer = nullptr;          // FP32: particles deferred to the SVD path of the elastic adjoint (S)
    bool persistent = false;
    long long nblocks4 = 0;         // blocks of every instance at tile 4 (the most of either tile)
    void* sortbuf = nullptr;        // sort keys, values, heads, blocks and rocPRIM scratch (first sort)
    size_t sort_temp = 0;
    // P2G workgroups per window mode (fixed, general, HBM) since the forward
    // pass began (3 per segment, nstat segments), and their sums at the last EndLayerMassLoss
    unsigned long long* wstats = nullptr;
    long long last_wstats[3] = {0, 0, 0};
    int resort_mode = 2;            // 0: at set_particles only; 1: + every chain_episode; 2: chain_episode when stale; 1, 2: forward_pass(0) when stale
    int resort_permille = 50;       // stale: more than this share of P2G workgroups off the fixed window
    long long n_sorts = 0;
    std::vector<int4> host_blocks;  // {begin, end, instance} of every non-empty 4^3-cell block
    // FP32 checkpoint tape: segment s >= 1 may keep its layers (and the
    // gradients of its layers) in a particle order of its own, sorted by the
    // positions of its first layer, when its P2G windows went stale in the last
    // forward pass that reached the loss (seg_resort).  The constants, controls,
    // control gradients and optimiser states stay in the home order (the order
    // of layer 0), which the kernels reach through home[].  A layer on a
    // segment boundary is gathered into the next segment's order (bidx); the
    // backward pass scatters its gradients back.
    struct SegOrder {
        bool own = false;         // an order of its own (else the home order)
        int* home = nullptr;      // N: home slot of each slot
        int* inv = nullptr;       // N: slot of each home slot
        int* bidx = nullptr;      // N: slot in segment s - 1's order of each slot of this segment's (orders differ)
        bool bidx_ok = false;
        int4* ranges = nullptr;   // block ranges of this order
        int nranges = 0;
        std::vector<int4> host_blocks;
        size_t ranges_cap = 0;
    };
    std::vector<SegOrder> sord;
    int seg_resort = 1;
    long long n_seg_sorts = 0;
    int nstat = 1;                       // window statistics kept per segment (FP32 checkpoint) or for the pass
    std::vector<long long> last_swstats; // 3 per segment, snapshot at EndLayerMassLoss
    int4* ranges = nullptr;         // per-workgroup {begin, end, instance}: blocks split to <= range_max
    std::vector<int4> host_ranges;
    int n_ranges = 0;
    int range_max = 512;            // particles per workgroup range (64..512)
    int lds_cap = 1024;             // general-window LDS capacity of the block kernels (nodes)
    // captured passes, replayed while graphs are on (kernel arguments are
    // fixed per context; the cache is dropped when ranges or options change)
    bool use_graphs = false;
    std::vector<GraphEntry> graphs;
    StepParams* d_sps = nullptr;
    long long* d_nofs = nullptr;
    Batch batch{};
    // checkpoint-mode slot layout
    // checkpoint mode: two tapes, segment s kept in tape s & 1 (slots from
    // tape_base[t], grids from grid slot t K); tape_hold[t]: the segment tape t
    // holds, or -1.  The backward pass recomputes segment s - 1 into one tape on
    // stream2 while it steps back through segment s in the other.
    int n_ckpt = 0, seg_last = 0, roll_base = 0, n_slots = 0;
    int tape_base[2] = {0, 0};
    int tape_hold[2] = {-1, -1};
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    bool overlap = true;
    // FP32: G2P of a step and P2G of the next in one launch (k_step_blk)
    bool fuse = true;
    // FP32: order of the block ranges (launch order): 0 by block, 1 longest first
    int range_order = 1;
    // FP32: where the elastic half of the transposed P2G runs (DMPM_OPT_ELASTIC)
    int elastic = 0;
    // validity
    bool have_particles = false;
    bool have_target = false;
    int fwd_valid = -1;      // layers [0, fwd_valid] are consistent with the controls
    bool loss_valid = false; // last layer gradients seeded
    int grads_lo = -1;       // lowest layer holding gradients
    int bwd_control = -1;    // control layer of the last backward pass
    // profiling
    bool prof = false;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    double prof_ms[DMPM_K_COUNT] = {0};
    int64_t prof_n[DMPM_K_COUNT] = {0};
};

namespace {

// Every device allocation of the library: a failed hipMalloc returns
// DMPM_ERR_NOMEM and clears the error hipMalloc also records as the thread's
// last error (the next launch check of any context would report it again).
int dmalloc(void** p, size_t bytes, const char* what)
{
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(DMPM_ERR_NOMEM, std::string(what) + ": hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
    }
    return DMPM_OK;
}

// a context buffer (counted in dmpm_memory_bytes)
int dalloc(dmpm_ctx* c, void** p, size_t bytes)
{
    const int rc = dmalloc(p, bytes, "context buffer");
    if (rc == DMPM_OK) c->bytes += bytes == 0 ? 256 : bytes;
    return rc;
}

// The device address of scratch on `device` (current): looked up on first use.
int scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(DMPM_ERR_INVALID, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return DMPM_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return fail(DMPM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_dev[device] = ptr;
    return DMPM_OK;
}

void* scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
}

hipEvent_t pool_event(dmpm_ctx* c)
{
    if (c->pool_used == c->pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->pool.push_back(e);
    }
    return c->pool[c->pool_used++];
}

// Run a launch, bracketed by events when profiling is on.
template <typename Fn>
void timed(dmpm_ctx* c, int kernel, Fn&& fn)
{
    if (!c->prof) {
        fn();
        return;
    }
    hipEvent_t a = pool_event(c), b = pool_event(c);
    if (!a || !b) {
        fn();
        return;
    }
    hipEventRecord(a, c->stream);
    fn();
    hipEventRecord(b, c->stream);
    c->recs.push_back({kernel, a, b});
}

int drain_profile(dmpm_ctx* c)
{
    if (c->recs.empty()) return DMPM_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto& r : c->recs) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
        c->prof_ms[r.kernel] += ms;
        c->prof_n[r.kernel] += 1;
    }
    c->recs.clear();
    c->pool_used = 0;
    return DMPM_OK;
}

// Step parameters of instance i (particle count, grid, time step).
void clear_graphs(dmpm_ctx* c)
{
    for (auto& g : c->graphs) hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

// Run the launches of body() on the context's stream.  With graphs on (and
// profiling off) the first run of a key is captured into a HIP graph and
// every run replays it; body's host-side effects happen only when captured,
// so callers set host state themselves after the call.
template <typename Fn>
int run_graph(dmpm_ctx* c, long long key, Fn&& body)
{
    if (!c->use_graphs || c->prof) {
        body();
        return DMPM_OK;
    }
    for (auto& g : c->graphs)
        if (g.key == key) {
            HIP_TRY(hipGraphLaunch(g.exec, c->stream));
            return DMPM_OK;
        }
    constexpr size_t kMaxGraphs = 64;
    if (c->graphs.size() >= kMaxGraphs) clear_graphs(c);
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    body();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (e != hipSuccess) return fail(DMPM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(DMPM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    c->graphs.push_back({key, exec});
    HIP_TRY(hipGraphLaunch(exec, c->stream));
    return DMPM_OK;
}

StepParams inst_params(const dmpm_ctx* c, int i)
{
    const dmpm_instance& in = c->inst[i];
    StepParams sp{};
    sp.n = in.n_particles;
    sp.stride = c->S;
    for (int a = 0; a < 3; ++a) {
        sp.dim[a] = in.dim[a];
        sp.min_point[a] = in.min_point[a] - c->origin[3 * (size_t)i + a];
        sp.f_ext[a] = c->cfg.f_ext[a];
    }
    sp.G = c->nofs[i + 1] - c->nofs[i];
    sp.dx = in.dx;
    sp.inv_dx = 1.0 / in.dx;
    sp.dt = c->cfg.dt;
    sp.drag = c->cfg.drag;
    sp.smoothing = c->cfg.smoothing;
    sp.compat = c->cfg.compat;
    return sp;
}

// Default checkpoint stride of `steps` steps: the divisor of steps nearest to
// sqrt(steps) within a factor 1.5 (the larger of two equally near), else
// round(sqrt(steps)).  A divisor makes every segment whole, so the last two
// segments, which the forward pass leaves on the two tapes, hold 2K steps:
// 50 steps take K = 5 (40 recomputed, 25 layer slots and 10 grids) instead of
// 7 (42 recomputed, 26 slots and 14 grids).
int default_stride(int steps)
{
    const double r = std::sqrt((double)steps);
    int best = 0;
    double bd = 0.0;
    for (int d = 1; d <= steps; ++d) {
        if (steps % d || d < r / 1.5 || d > r * 1.5) continue;
        const double dd = std::fabs(d - r);
        if (!best || dd <= bd) {
            best = d;
            bd = dd;
        }
    }
    return best ? best : (int)std::lround(r);
}

// Parameters of the single-instance (FP64) path.
StepParams step_params(const dmpm_ctx* c) { return inst_params(c, 0); }

// ------------------------------------------------------------ slot layout --
template <typename Real>
Real* slot(dmpm_ctx* c, int s)
{
    return static_cast<Real*>(c->layers) + (size_t)s * c->lf * c->S;
}

// the tape holding segment seg, or -1
int tape_of(const dmpm_ctx* c, int seg)
{
    for (int t = 0; t < 2; ++t)
        if (c->tape_hold[t] == seg) return t;
    return -1;
}

void drop_tapes(dmpm_ctx* c) { c->tape_hold[0] = c->tape_hold[1] = -1; }

// tapes holding a segment that reaches past layer lo (its controls changed) are stale
void drop_tapes_from(dmpm_ctx* c, int lo)
{
    for (int& h : c->tape_hold)
        if (h >= 0 && h * c->K + c->K > lo) h = -1;
}

// layer j of segment j / K in tape t (a checkpoint when j is a multiple of K)
template <typename Real>
Real* tape_slot(dmpm_ctx* c, int t, int j)
{
    if (j % c->K == 0) return slot<Real>(c, j / c->K);
    return slot<Real>(c, c->tape_base[t] + (j - (j / c->K) * c->K));
}

// grid of step i of segment i / K in tape t
template <typename Real>
Real* tape_grid(dmpm_ctx* c, int t, int i)
{
    return static_cast<Real*>(c->grids) + ((size_t)t * c->K + (size_t)(i - (i / c->K) * c->K)) * kGridFields * c->G;
}

// The grid of checkpoint step i (a multiple of K) when the context keeps one
// (FP32 checkpoint tape: the forward pass writes it, together with the step's
// affine matrix in the checkpoint layer, so a recompute of the segment starts
// with its first G2P), else null.
template <typename Real>
Real* ck_grid(dmpm_ctx* c, int i)
{
    if (!c->ck_grids || i % c->K != 0) return nullptr;
    return static_cast<Real*>(c->grids) + (2 * (size_t)c->K + (size_t)(i / c->K)) * kGridFields * c->G;
}

// grid of step i of segment i / K held in tape t, or its checkpoint grid
template <typename Real>
Real* seg_grid(dmpm_ctx* c, int t, int i)
{
    Real* g = ck_grid<Real>(c, i);
    return g ? g : tape_grid<Real>(c, t, i);
}

template <typename Real>
Real* tape_layer(dmpm_ctx* c, int j)  // layer j of a segment held by a tape
{
    if (c->cfg.tape_mode == DMPM_TAPE_STORE_ALL) return slot<Real>(c, j);
    const int t = tape_of(c, j / c->K);
    return tape_slot<Real>(c, t < 0 ? 0 : t, j);
}

template <typename Real>
Real* grid_of_step(dmpm_ctx* c, int i)
{
    if (c->cfg.tape_mode == DMPM_TAPE_STORE_ALL) return static_cast<Real*>(c->grids) + (size_t)i * kGridFields * c->G;
    if (Real* g = ck_grid<Real>(c, i)) return g;
    const int t = tape_of(c, i / c->K);
    if (t >= 0) return tape_grid<Real>(c, t, i);
    return static_cast<Real*>(c->wgrid);
}

template <typename Real>
Real* ctrl_of(dmpm_ctx* c, int layer)
{
    return static_cast<Real*>(c->ctrl) + (size_t)layer * 9 * c->S;
}

// FP32 checkpoint tape with the fused step and no smoothing: F enters every
// kernel only as Fc = F + dFc, so the passes store Fc in the F field of the
// layers that are neither checkpoints nor the last, and their readers skip
// the control (a null ctrl).  The API's layer reads recompute true F.
bool fc_layers(const dmpm_ctx* c)
{
    return c->f32 && c->cfg.tape_mode == DMPM_TAPE_CHECKPOINT && c->fuse && c->cfg.smoothing == 0.0;
}
bool fc_layer(const dmpm_ctx* c, int j)
{
    return fc_layers(c) && j % c->K != 0 && j < c->L - 1;
}

template <typename Real>
Real* grads_of(dmpm_ctx* c, int layer)
{
    return static_cast<Real*>(c->gbuf[layer == c->L - 1 ? 2 : (layer & 1)]);
}

template <typename Real>
Real* gctrl_of(dmpm_ctx* c, int layer)
{
    return static_cast<Real*>(c->gctrl) + (c->cfg.keep_control_grads ? (size_t)layer * 9 * c->S : 0);
}

// --------------------------------------------------------------- kernels ---
template <typename Real>
void fill(dmpm_ctx* c, Real* p, long long count, Real v)
{
    timed(c, DMPM_K_FILL, [&] { launch_fill<Real>(p, count, v, c->stream); });
}

// Tile flags of an FP32 grid buffer (null for the FP64 path).
unsigned* flags_of(dmpm_ctx* c, const void* g)
{
    if (!c->tflags) return nullptr;
    const size_t gb = (size_t)kGridFields * c->G * sizeof(float);
    size_t k;
    if (g == c->wgrid) k = c->grid_slots;
    else if (g == c->agrid) k = c->grid_slots + 1;
    else if (g == c->wgrid2) k = c->grid_slots + 2;
    else if (g == c->wgrid3) k = c->grid_slots + 3;
    else k = (size_t)(static_cast<const char*>(g) - static_cast<const char*>(c->grids)) / gb;
    return c->tflags + k * (size_t)c->ntiles;
}

// Zero a 4-channel grid buffer before a scatter into it: FP32 clears the
// flagged tiles only (every other node is already 0), FP64 the whole buffer.
void clear_grid(dmpm_ctx* c, float* g)
{
    timed(c, DMPM_K_FILL, [&] { launch_clear_tiles(c->batch, g, flags_of(c, g), c->stream); });
}

void clear_grid(dmpm_ctx* c, double* g) { fill<double>(c, g, (long long)kGridFields * c->G, 0.0); }

// restart the window-mode counts (a kernel, so that graphs may hold it)
void zero_wstats(dmpm_ctx* c)
{
    if (c->wstats) launch_fill<double>(reinterpret_cast<double*>(c->wstats), 3LL * c->nstat, 0.0, c->stream);
}

// wgrid used as scratch (node masses, dL/dm): every tile may be non-zero
void taint_wgrid(dmpm_ctx* c)
{
    if (c->tflags) launch_set_flags(flags_of(c, c->wgrid), c->ntiles, c->stream);
}

// Particle-grid kernels: the plain FP64 path runs one thread per particle
// with direct HBM atomics; the FP32 path runs one workgroup per block range
// with the node window staged in LDS.
inline void p2g_dispatch(dmpm_ctx* c, const StepParams& sp, const StepIO<double>& io) { launch_p2g<double>(sp, io, c->stream); }
inline void p2g_dispatch(dmpm_ctx* c, const StepParams&, const StepIO<float>& io) { launch_p2g_blk(c->batch, io, TileClear{}, c->stream); }
inline void grid_update_dispatch(dmpm_ctx* c, const StepParams& sp, double* g) { launch_grid_update<double>(sp, g, c->stream); }
inline void g2p_dispatch(dmpm_ctx* c, const StepParams& sp, const StepIO<double>& io) { launch_g2p<double>(sp, io, c->stream); }
inline void g2p_dispatch(dmpm_ctx* c, const StepParams&, const StepIO<float>& io, const TileClear& tc)
{
    launch_g2p_blk(c->batch, io, tc, c->stream);
}
inline void g2p_adj_dispatch(dmpm_ctx* c, const StepParams& sp, const AdjIO<double>& io) { launch_g2p_adj<double>(sp, io, c->stream); }
inline void g2p_adj_dispatch(dmpm_ctx* c, const StepParams&, const AdjIO<float>& io) { launch_g2p_adj_blk(c->batch, io, c->stream); }
inline void grid_adj_dispatch(dmpm_ctx* c, const StepParams& sp, const AdjIO<double>& io) { launch_grid_adj<double>(sp, io, c->stream); }
inline void p2g_adj_dispatch(dmpm_ctx* c, const StepParams& sp, const AdjIO<double>& io) { launch_p2g_adj<double>(sp, io, c->stream); }
inline void p2g_adj_dispatch(dmpm_ctx* c, const StepParams&, const AdjIO<float>& io, const TileClear& tc)
{
    launch_p2g_adj_blk(c->batch, io, tc, c->elastic, c->N, c->d_poff, c->stream);
}

// One ForwardTimeStep: cur -> nxt, grid of the step in g.  g_clean: the
// previous step of the same launch sequence cleared g already.  gnext (FP32):
// the grid of the next step of the sequence, another buffer, which G2P's
// appended workgroups clear; returns whether they did.  keep_a (FP32): cur is
// a layer the backward pass reads, so P2G caches its affine matrix there.
// with_g2p false: the grid of the step only (its next layer is held already).
template <typename Real>
bool step_forward(dmpm_ctx* c, const Real* cur, Real* nxt, const Real* ctrl, Real* g, bool g_clean = false, Real* gnext = nullptr,
                  bool keep_a = true, bool with_g2p = true)
{
    StepParams sp = step_params(c);
    StepIO<Real> io{cur, nxt, ctrl, static_cast<const Real*>(c->cst), g, c->f32 && keep_a ? const_cast<Real*>(cur) : nullptr,
                    flags_of(c, g)};
    if (!g_clean) clear_grid(c, g);
    timed(c, DMPM_K_P2G, [&] { p2g_dispatch(c, sp, io); });
    if (!with_g2p) {
        if constexpr (std::is_same<Real, double>::value) timed(c, DMPM_K_GRID_UPDATE, [&] { grid_update_dispatch(c, sp, g); });
        return false;
    }
    if constexpr (std::is_same<Real, double>::value) {
        timed(c, DMPM_K_GRID_UPDATE, [&] { grid_update_dispatch(c, sp, g); });
        timed(c, DMPM_K_G2P, [&] { g2p_dispatch(c, sp, io); });
        return false;
    } else {
        // FP32: G2P applies the grid update as it loads the nodes
        TileClear tc{};
        if (gnext && gnext != g) tc = TileClear{gnext, flags_of(c, gnext)};
        timed(c, DMPM_K_G2P, [&] { g2p_dispatch(c, sp, io, tc); });
        return tc.grid != nullptr;
    }
}

// One Back_Timestep: gradients of layer i+1 -> layer i.  agrid_clean: the
// previous backward step of the same launch sequence cleared the adjoint
// grid; clear_next (FP32): the elastic adjoint's appended workgroups clear it
// for the next one (returns whether they did).
template <typename Real>
bool step_backward(dmpm_ctx* c, int i, const Real* cur, const Real* nxt, const Real* g, bool agrid_clean = false,
                   bool clear_next = false)
{
    StepParams sp = step_params(c);
    AdjIO<Real> io{};
    io.cur = cur;
    io.next = nxt;
    io.ctrl = fc_layer(c, i) ? nullptr : ctrl_of<Real>(c, i);
    io.cst = static_cast<const Real*>(c->cst);
    io.grid = g;
    io.agrid = static_cast<Real*>(c->agrid);
    io.g1 = grads_of<Real>(c, i + 1);
    io.g0 = grads_of<Real>(c, i);
    // without keep_control_grads every layer's control gradient goes to one
    // block and only the last step's (the control layer's, clear_next false)
    // is read: the steps above it skip that store
    io.gctrl = (c->cfg.keep_control_grads || !clear_next) ? gctrl_of<Real>(c, i) : nullptr;
    io.tmp = static_cast<Real*>(c->tmp);
    io.gflags = flags_of(c, g);
    io.aflags = flags_of(c, c->agrid);
    io.defer = c->defer;
    io.defer_n = c->queue ? c->queue + 4 : nullptr;
    if constexpr (std::is_same<Real, double>::value) {
        fill<Real>(c, io.agrid, 3 * c->G, Real(0));  // the vbar channels
        timed(c, DMPM_K_G2P_ADJ, [&] { g2p_adj_dispatch(c, sp, io); });
        timed(c, DMPM_K_GRID_ADJ, [&] { grid_adj_dispatch(c, sp, io); });
        timed(c, DMPM_K_P2G_ADJ, [&] { p2g_adj_dispatch(c, sp, io); });
        return false;
    } else {
        // FP32: the flagged tiles; the transposed P2G applies the transposed
        // grid update as it loads the nodes
        if (!agrid_clean) clear_grid(c, io.agrid);
        timed(c, DMPM_K_G2P_ADJ, [&] { g2p_adj_dispatch(c, sp, io); });
        TileClear tc{};
        if (clear_next) tc = TileClear{io.agrid, io.aflags};
        timed(c, DMPM_K_P2G_ADJ, [&] { p2g_adj_dispatch(c, sp, io, tc); });
        return clear_next;
    }
}

// --------------------------------------------------------- segment orders --
bool seg_own(const dmpm_ctx* c, int s) { return s >= 1 && s < (int)c->sord.size() && c->sord[s].own; }

// the orders of segments s - 1 and s differ: layer sK is gathered between them
bool boundary_permuted(const dmpm_ctx* c, int s) { return seg_own(c, s - 1) || seg_own(c, s); }

// the segment whose order holds layer j as the passes leave it: a checkpoint
// its own segment's, the last layer the last segment's
int layer_seg(const dmpm_ctx* c, int j) { return j >= c->L - 1 ? c->seg_last : j / std::max(c->K, 1); }

// the launch configuration of segment s's steps: its block ranges and home
// map, and its window statistics
Batch batch_of(const dmpm_ctx* c, int s)
{
    Batch b = c->batch;
    b.ranges = c->ranges;
    b.nranges = c->n_ranges;
    b.home = nullptr;
    b.wstats = c->wstats ? c->wstats + 3 * (size_t)std::min(std::max(s, 0), c->nstat - 1) : nullptr;
    if (seg_own(c, s)) {
        const auto& o = c->sord[s];
        b.ranges = o.ranges;
        b.nranges = o.nranges;
        b.home = o.home;
    }
    return b;
}

// c->batch is segment s's for the launches of a scope
struct BatchScope {
    dmpm_ctx* c;
    Batch saved;
    BatchScope(dmpm_ctx* ctx, int s) : c(ctx), saved(ctx->batch) { c->batch = batch_of(ctx, s); }
    ~BatchScope() { c->batch = saved; }
};

// slot in segment s - 1's order of each slot of segment s's order, or null when
// the two orders are the same; computed on the stream when first needed after
// either order changed (prepare_bidx does it before a graph capture)
const int* bidx_of(dmpm_ctx* c, int s)
{
    if (!boundary_permuted(c, s)) return nullptr;
    auto& o = c->sord[s];
    if (!o.bidx_ok) {
        const size_t bytes = (size_t)c->N * sizeof(int);
        if (seg_own(c, s) && seg_own(c, s - 1))
            launch_gather<int>(c->sord[s - 1].inv, o.bidx, (long long)c->S, c->N, 1, o.home, c->stream);
        else if (seg_own(c, s))
            hipMemcpyAsync(o.bidx, o.home, bytes, hipMemcpyDeviceToDevice, c->stream);
        else
            hipMemcpyAsync(o.bidx, c->sord[s - 1].inv, bytes, hipMemcpyDeviceToDevice, c->stream);
        o.bidx_ok = true;
    }
    return o.bidx;
}

void prepare_bidx(dmpm_ctx* c)
{
    for (int s = 1; s < (int)c->sord.size(); ++s) bidx_of(c, s);
}

// segment s's P2G windows in the last forward pass that reached the loss:
// more than resort_permille per mille of its workgroups off the fixed window
bool seg_stale(const dmpm_ctx* c, int s)
{
    if ((int)c->last_swstats.size() < 3 * (s + 1)) return false;
    const long long* w = &c->last_swstats[3 * (size_t)s];
    const long long tot = w[0] + w[1] + w[2];
    return tot > 0 && (w[1] + w[2]) * 1000 > (long long)c->resort_permille * tot;
}

// checkpoint-mode forward grid of a step outside the tape: buffers in
// rotation, so that a step's G2P can clear the next step's grid (FP64: two),
// also while it runs as one launch with the next step's P2G (FP32: three)
template <typename Real>
Real* wgrid_of(dmpm_ctx* c, int i)
{
    if (c->wgrid3) {
        void* const g[3] = {c->wgrid, c->wgrid2, c->wgrid3};
        return static_cast<Real*>(g[i % 3]);
    }
    return static_cast<Real*>((i & 1) ? c->wgrid2 : c->wgrid);
}

// Steps [i0, i1) of one launch sequence.  lay(j): the slot of layer j (read
// for j = i0, written for i0 < j <= i1); grid(i): the grid of step i, three
// consecutive ones distinct buffers; keep(i): the backward pass reads layer i,
// so P2G caches its affine matrix there; last_g2p: the last step's G2P runs
// (else only its grid: the layer after it is held already).  FP32 with
// c->fuse runs the G2P of each step with the P2G of the next as one launch;
// there p2g0 false starts from grid(i0) and the affine matrix of layer i0 as
// an earlier pass left them (no P2G of step i0), and the layers j < lean_below
// that are not checkpoints are written lean (x and F: the next step's G2P is
// their only reader; the same launch's P2G takes v and C from registers).
// fc: the layers of fc_layer hold Fc in their F field (written by the fused
// step, read with a null control); i0 and i1 are checkpoints or the last layer.
template <typename Real, typename LayFn, typename GridFn, typename KeepFn>
void forward_seq(dmpm_ctx* c, int i0, int i1, LayFn&& lay, GridFn&& grid, KeepFn&& keep, bool last_g2p = true,
                 bool p2g0 = true, int lean_below = -1, bool fc = false)
{
    if (i1 <= i0) return;
    if constexpr (std::is_same<Real, float>::value) {
        if (c->fuse) {
            const StepParams sp = step_params(c);
            auto pio_of = [&](int i) {
                float* g = grid(i);
                return StepIO<float>{lay(i), nullptr, ctrl_of<float>(c, i), static_cast<const float*>(c->cst), g,
                                     keep(i) ? lay(i) : nullptr, flags_of(c, g)};
            };
            auto clear_of = [&](int i) { return i < i1 ? TileClear{grid(i), flags_of(c, grid(i))} : TileClear{}; };
            // the P2G of step i counts its window modes for step i's segment
            auto batch_at = [&](int i) {
                Batch b = c->batch;
                if (c->nstat > 1 && c->wstats) b.wstats = c->wstats + 3 * (size_t)std::min(i / c->K, c->nstat - 1);
                return b;
            };
            if (p2g0) {
                clear_grid(c, grid(i0));
                timed(c, DMPM_K_P2G, [&] { launch_p2g_blk(batch_at(i0), pio_of(i0), clear_of(i0 + 1), c->stream); });
            } else if (i0 + 1 < i1) {
                clear_grid(c, grid(i0 + 1));  // the P2G of step i0 would have carried this clear
            }
            auto is_fc = [&](int j) { return fc && fc_layer(c, j); };
            auto ctrl_in = [&](int j) { return is_fc(j) ? nullptr : ctrl_of<float>(c, j); };
            for (int i = i0 + 1; i < i1; ++i) {
                const StepIO<float> gio{lay(i - 1), lay(i), ctrl_in(i - 1), static_cast<const float*>(c->cst), grid(i - 1),
                                        nullptr, nullptr, i < lean_below && i % c->K != 0, is_fc(i)};
                timed(c, DMPM_K_STEP, [&] { launch_step_blk(batch_at(i), gio, pio_of(i), clear_of(i + 1), c->stream); });
            }
            if (last_g2p) {
                const StepIO<float> gio{lay(i1 - 1), lay(i1), ctrl_in(i1 - 1), static_cast<const float*>(c->cst),
                                        grid(i1 - 1), nullptr, nullptr};
                timed(c, DMPM_K_G2P, [&] { g2p_dispatch(c, sp, gio, TileClear{}); });
            }
            return;
        }
    }
    bool clean = false;
    unsigned long long* const ws = c->batch.wstats;
    for (int i = i0; i < i1; ++i) {
        if (c->nstat > 1 && c->wstats) c->batch.wstats = c->wstats + 3 * (size_t)std::min(i / c->K, c->nstat - 1);
        clean = step_forward<Real>(c, lay(i), lay(i + 1), ctrl_of<Real>(c, i), grid(i), clean, i + 1 < i1 ? grid(i + 1) : nullptr,
                                   keep(i), i + 1 < i1 || last_g2p);
    }
    c->batch.wstats = ws;
}

// Bring layer j (<= fwd_valid) into a readable slot and return it.  Checkpoint
// mode recomputes from the checkpoint below j into the rolling slots when j
// is neither a checkpoint nor in a tape segment, or its tape holds Fc for F.
template <typename Real>
Real* materialize(dmpm_ctx* c, int j)
{
    if (c->cfg.tape_mode == DMPM_TAPE_STORE_ALL) return slot<Real>(c, j);
    if (j % c->K == 0) return slot<Real>(c, j / c->K);
    if (tape_of(c, j / c->K) >= 0 && !fc_layer(c, j)) return tape_layer<Real>(c, j);
    const int j0 = (j / c->K) * c->K;
    auto lay = [&](int k) { return k == j0 ? slot<Real>(c, j0 / c->K) : slot<Real>(c, c->roll_base + (k & 1)); };
    BatchScope bs(c, j / c->K);
    forward_seq<Real>(c, j0, j, lay, [&](int i) { return wgrid_of<Real>(c, i); }, [](int) { return false; });
    return lay(j);
}

// Recompute segment seg into its tape, seg & 1 (checkpoint mode), with the
// launches on c->stream.  The caller records which segment the tape holds.
template <typename Real>
void recompute_segment(dmpm_ctx* c, int seg)
{
    const int t = seg & 1;
    const int s0 = seg * c->K, s1 = std::min(s0 + c->K, c->L - 1);
    // a segment that ends on a checkpoint needs only the grid of its last
    // step: the backward pass reads the next layer from the checkpoint, unless
    // the checkpoint holds it in another order; then the last G2P writes it in
    // this segment's order to the tape's boundary slot
    const bool perm_b = s1 < c->L - 1 && boundary_permuted(c, seg + 1);
    Real* bslot = slot<Real>(c, c->tape_base[t] + c->K);
    BatchScope bs(c, seg);
    // with checkpoint grids the forward pass left step s0's grid and affine matrix: no P2G of it
    forward_seq<Real>(c, s0, s1, [&](int j) { return (j == s1 && perm_b) ? bslot : tape_slot<Real>(c, t, j); },
                      [&](int i) { return seg_grid<Real>(c, t, i); }, [](int) { return true; }, s1 % c->K != 0 || perm_b,
                      c->ck_grids == 0, -1, fc_layers(c));
}

// ------------------------------------------------------------ block sort ---
// block prefix of the instances at the current tile (host; the caller uploads it)
void set_block_prefix(dmpm_ctx* c)
{
    const int t = c->tile;
    c->boff.assign(c->B + 1, 0);
    for (int i = 0; i < c->B; ++i) {
        const dmpm_instance& in = c->inst[i];
        c->boff[i + 1] = c->boff[i] + (long long)((in.dim[0] + t - 1) / t) * ((in.dim[1] + t - 1) / t) * ((in.dim[2] + t - 1) / t);
    }
}

// Workgroup ranges of blocks {begin, end, instance}: at most range_max
// particles each, longest first with range_order 1 (workgroups start in range
// order, so the short ones fill the last round of a launch).
std::vector<int4> make_ranges(const dmpm_ctx* c, const std::vector<int4>& blocks)
{
    std::vector<int4> rg;
    for (const int4& b : blocks) {
        const int len = b.y - b.x;
        const int parts = (len + c->range_max - 1) / c->range_max;
        for (int k = 0; k < parts; ++k)
            rg.push_back(make_int4(b.x + (int)((long long)len * k / parts), b.x + (int)((long long)len * (k + 1) / parts), b.z, 0));
    }
    if (c->range_order == 1)
        std::stable_sort(rg.begin(), rg.end(), [](const int4& a, const int4& b) { return a.y - a.x > b.y - b.x; });
    return rg;
}

// ranges of a segment order into its device array (grown when short)
int upload_seg_ranges(dmpm_ctx* c, dmpm_ctx::SegOrder& o)
{
    const std::vector<int4> rg = make_ranges(c, o.host_blocks);
    if (rg.size() > o.ranges_cap) {
        if (o.ranges) hipFree(o.ranges);
        o.ranges = nullptr;
        o.ranges_cap = 0;
        const int rc = dalloc(c, (void**)&o.ranges, rg.size() * sizeof(int4));
        if (rc) return rc;
        o.ranges_cap = rg.size();
    }
    if (!rg.empty())
        HIP_TRY(hipMemcpyAsync(o.ranges, rg.data(), rg.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    o.nranges = (int)rg.size();
    return DMPM_OK;
}

// The block order of the particles whose positions are the x field of lay
// (dmpm_sort.hip): a.v1[d] is the slot of lay whose particle goes to slot d,
// a.blk[d] its global block.
int run_sort(dmpm_ctx* c, const float* lay, SortArgs& a)
{
    const int N = c->N;
    const long long nblocks = c->boff[c->B];
    if (!c->sortbuf) {
        // scratch for the widest keys of either tile
        c->sort_temp = sort_temp_bytes(N, std::max(nblocks, c->nblocks4));
        if (c->sort_temp == 0) return fail(DMPM_ERR_HIP, "rocPRIM sort: temporary storage query failed");
        const int rc = dalloc(c, &c->sortbuf, (size_t)N * (2 * 8 + 4 * 4) + 256 + c->sort_temp);
        if (rc) return rc;
    }
    auto* k0 = static_cast<unsigned long long*>(c->sortbuf);
    a = SortArgs{};
    a.sps = c->d_sps;
    a.poff = c->d_poff;
    a.boff = c->d_boff;
    a.B = c->B;
    a.tshift = c->tile == 8 ? 3 : 2;
    a.nblocks = nblocks;
    a.x = lay;
    a.S = (long long)c->S;
    a.N = N;
    a.k0 = k0;
    a.k1 = k0 + N;
    a.v0 = reinterpret_cast<int*>(a.k1 + N);
    a.v1 = a.v0 + N;
    a.head = a.v1 + N;
    a.blk = a.head + N;
    a.temp = reinterpret_cast<char*>(c->sortbuf) + (((size_t)N * (2 * 8 + 4 * 4) + 255) & ~(size_t)255);
    a.temp_bytes = c->sort_temp;
    HIP_TRY(launch_sort_order(a, c->stream));
    return DMPM_OK;
}

// {begin, end, instance} of every non-empty block from the sorted block ids
std::vector<int4> blocks_from(const dmpm_ctx* c, const std::vector<int>& blk)
{
    std::vector<int4> out;
    const int N = (int)blk.size();
    int in = 0;
    for (int d0 = 0, d = 1; d <= N; ++d) {
        if (d == N || blk[d] != blk[d0]) {
            while (in + 1 < c->B && c->boff[in + 1] <= blk[d0]) ++in;
            out.push_back(make_int4(d0, d, in, 0));
            d0 = d;
        }
    }
    return out;
}

// every segment back on the home order: layers held in segment orders are
// no longer readable, so the caller invalidates the pass
void drop_seg_orders(dmpm_ctx* c)
{
    bool any = false;
    for (auto& o : c->sord) {
        any = any || o.own;
        o.own = false;
        o.bidx_ok = false;
    }
    if (any) clear_graphs(c);
}

// A new order for segment s >= 1 from the positions of its first layer, held
// in lay in segment s - 1's order: by block as the home order is, composed
// with segment s - 1's map to the home order.
int seg_sort(dmpm_ctx* c, int s, const float* lay)
{
    auto& o = c->sord[s];
    const int N = c->N;
    const size_t bytes = (size_t)N * sizeof(int);
    int rc;
    if (!o.home && ((rc = dalloc(c, (void**)&o.home, bytes)) || (rc = dalloc(c, (void**)&o.inv, bytes)))) return rc;
    for (int k = s; k <= std::min(s + 1, c->seg_last); ++k)
        if (!c->sord[k].bidx && (rc = dalloc(c, (void**)&c->sord[k].bidx, bytes))) return rc;
    SortArgs a;
    if ((rc = run_sort(c, lay, a))) return rc;
    HIP_TRY(hipMemcpyAsync(o.bidx, a.v1, bytes, hipMemcpyDeviceToDevice, c->stream));
    if (seg_own(c, s - 1)) launch_gather<int>(c->sord[s - 1].home, o.home, (long long)c->S, N, 1, a.v1, c->stream);
    else HIP_TRY(hipMemcpyAsync(o.home, a.v1, bytes, hipMemcpyDeviceToDevice, c->stream));
    launch_invert(o.home, o.inv, N, c->stream);
    HIP_TRY(hipGetLastError());
    std::vector<int> blk(N);
    HIP_TRY(hipMemcpyAsync(blk.data(), a.blk, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    o.host_blocks = blocks_from(c, blk);
    if ((rc = upload_seg_ranges(c, o))) return rc;
    o.own = true;
    o.bidx_ok = true;
    if (s + 1 <= c->seg_last) c->sord[s + 1].bidx_ok = false;
    for (int k = 0; k < 3; ++k) c->last_swstats[3 * (size_t)s + k] = 0;
    ++c->n_seg_sorts;
    clear_graphs(c);
    return DMPM_OK;
}

// Gradients of layer s K (in segment s's order) into segment s - 1's order,
// in place, through tmp (12 fields at a time).
void grads_to_prev(dmpm_ctx* c, float* g, const int* bidx)
{
    float* t = static_cast<float*>(c->tmp);
    for (int f = 0; f < kLayerFields; f += 12) {
        const int k = std::min(12, kLayerFields - f);
        launch_scatter<float>(g + (size_t)f * c->S, t, (long long)c->S, c->N, k, bidx, c->stream);
        hipMemcpyAsync(g + (size_t)f * c->S, t, (size_t)k * c->S * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    }
}

// FP32 checkpoint forward pass from checkpoint `from`, one launch sequence per
// segment, whose P2G window statistics are kept apart.  Segments whose
// windows went stale in the last pass that reached the loss are ordered
// afresh from their first layer (that pass runs without a graph: the sort
// reads block counts back); each boundary layer is gathered into the next
// segment's order when the two differ.
int forward_segments(dmpm_ctx* c, int from)
{
    const int K = c->K;
    const int s0_last = c->seg_last * K;
    const int tl = c->seg_last & 1;
    const int s_from = from / K;
    // the segment before the last goes to the other tape (its backward steps
    // then need no recompute) when this pass runs it
    const int sp = c->seg_last - 1;
    const bool tape_prev = sp >= s_from;
    const int s0_prev = tape_prev ? sp * K : s0_last;
    std::vector<char> sort_at(c->seg_last + 2, 0);
    bool any_sort = false;
    if (c->seg_resort)
        for (int s = s_from + 1; s <= c->seg_last; ++s)
            if (seg_stale(c, s)) {
                sort_at[s] = 1;
                any_sort = true;
            }
    // checkpoints and the last two segments go to the tapes, other layers to the rolling slots
    auto lay = [&](int j) {
        if (j % K == 0 || j / K == c->seg_last) return tape_slot<float>(c, tl, j);
        if (j >= s0_prev) return tape_slot<float>(c, tl ^ 1, j);
        return slot<float>(c, c->roll_base + (j & 1));
    };
    auto grid_at = [&](int k) {
        if (float* g = ck_grid<float>(c, k)) return g;
        if (k >= s0_last) return tape_grid<float>(c, tl, k);
        if (k >= s0_prev) return tape_grid<float>(c, tl ^ 1, k);
        return wgrid_of<float>(c, k);
    };
    // steps of earlier segments are recomputed for the backward pass: their affine matrices are not
    // read, but those of checkpoint steps with a grid of their own (the recompute starts after them)
    auto keep = [&](int i) { return i >= s0_prev || ck_grid<float>(c, i) != nullptr; };
    int err = DMPM_OK;
    auto pass = [&] {
        zero_wstats(c);
        for (int s = s_from; s <= c->seg_last;) {
            // one launch sequence over segments s .. e, joined by boundaries in one order
            int e = s;
            while (e < c->seg_last && !sort_at[e + 1] && !boundary_permuted(c, e + 1)) ++e;
            const int a = s * K, b = std::min((e + 1) * K, c->L - 1);
            const bool perm_b = b < c->L - 1;  // the boundary after e changes the order
            // a taped segment before the last keeps that boundary layer, in its own
            // order, in its tape's boundary slot (as a recompute leaves it)
            float* scratch = (tape_prev && e == sp) ? slot<float>(c, c->tape_base[tl ^ 1] + K) : slot<float>(c, c->roll_base + (b & 1));
            {
                BatchScope bs(c, s);
                forward_seq<float>(c, a, b, [&](int j) { return (j == b && perm_b) ? scratch : lay(j); }, grid_at, keep, true,
                                   true, s0_prev, fc_layers(c));
            }
            if (perm_b) {
                if (sort_at[e + 1] && (err = seg_sort(c, e + 1, scratch))) return;
                launch_gather<float>(scratch, lay(b), (long long)c->S, c->N, kLayerFields, bidx_of(c, e + 1), c->stream);
            }
            s = e + 1;
        }
    };
    if (any_sort) {
        pass();
        if (err) return err;
        HIP_TRY(hipGetLastError());
        return DMPM_OK;
    }
    prepare_bidx(c);
    const int rc = run_graph(c, from, pass);
    return rc ? rc : err;
}

// the segment before the last, held by the other tape after a checkpoint
// forward pass from `from` (that pass ran it), or -1
int prev_taped(const dmpm_ctx* c, int from)
{
    const int sp = c->seg_last - 1;
    return sp >= 0 && sp >= from / std::max(c->K, 1) ? sp : -1;
}

// -------------------------------------------------------------- passes -----
int resort(dmpm_ctx* c, bool controls_live);
bool order_stale(const dmpm_ctx* c);

template <typename Real>
int forward_pass(dmpm_ctx* c, int start)
{
    if (!c->have_particles) return fail(DMPM_ERR_STATE, "forward pass before dmpm_set_particles");
    if (start < 0 || start >= c->L - 1) return fail(DMPM_ERR_INVALID, "start_layer out of range");
    // the home order: the whole pass's windows, or with segment orders segment 0's
    if (c->f32 && start == 0 && c->resort_mode >= 1 && (c->nstat > 1 ? seg_stale(c, 0) : order_stale(c))) {
        const int rs = resort(c, true);
        if (rs) return rs;
    }
    int from = std::min(start, c->fwd_valid);
    int rc = DMPM_OK;
    if (c->cfg.tape_mode == DMPM_TAPE_STORE_ALL) {
        rc = run_graph(c, from, [&] {
            zero_wstats(c);
            forward_seq<Real>(c, from, c->L - 1, [&](int j) { return slot<Real>(c, j); }, [&](int i) { return grid_of_step<Real>(c, i); },
                              [](int) { return true; });
        });
        if (rc) return rc;
    } else {
        from = (from / c->K) * c->K;  // restart from a checkpoint so the tape holds whole segments
        const int s0_last = c->seg_last * c->K;
        const int tl = c->seg_last & 1;  // the last segment's tape
        const int sp = prev_taped(c, from);  // the segment before it, in the other tape
        const int s0_prev = sp >= 0 ? sp * c->K : s0_last;
        auto pass = [&] {
            zero_wstats(c);
            // checkpoints and the last two segments go to the tapes, other layers to the rolling slots
            auto lay = [&](int j) {
                if (j % c->K == 0 || j / c->K == c->seg_last) return tape_slot<Real>(c, tl, j);
                if (j >= s0_prev) return tape_slot<Real>(c, tl ^ 1, j);
                return slot<Real>(c, c->roll_base + (j & 1));
            };
            auto grid_at = [&](int k) {
                if (Real* g = ck_grid<Real>(c, k)) return g;
                if (k >= s0_last) return tape_grid<Real>(c, tl, k);
                if (k >= s0_prev) return tape_grid<Real>(c, tl ^ 1, k);
                return wgrid_of<Real>(c, k);
            };
            // steps of earlier segments are recomputed for the backward pass:
            // their affine matrices are not read, but those of checkpoint steps
            // with a grid of their own (the recompute starts after them)
            forward_seq<Real>(c, from, c->L - 1, lay, grid_at,
                              [&](int i) { return i >= s0_prev || ck_grid<Real>(c, i) != nullptr; }, true, true, s0_prev,
                              fc_layers(c));
        };
        if (c->f32 && c->nstat > 1) {
            if constexpr (std::is_same<Real, float>::value) rc = forward_segments(c, from);
        } else {
            rc = run_graph(c, from, pass);
        }
        if (rc) return rc;
        c->tape_hold[tl] = c->seg_last;
        c->tape_hold[tl ^ 1] = sp;
    }
    HIP_TRY(hipGetLastError());
    c->fwd_valid = c->L - 1;
    c->loss_valid = false;
    c->grads_lo = -1;
    return DMPM_OK;
}

// Loss node terms and particle seed: per-particle kernels (FP64 path) or the
// batched kernels over the block ranges (FP32 path).
void loss_kernels(dmpm_ctx* c, const double* lay, const double* gm, double* dLdm)
{
    const StepParams sp = step_params(c);
    const double* cst = static_cast<const double*>(c->cst);
    launch_loss<double>(sp, gm, static_cast<const double*>(c->target), dLdm, c->dpart, c->dacc, c->stream);
    launch_loss_particles<double>(sp, lay, cst, dLdm, grads_of<double>(c, c->L - 1), c->stream);
}

// node masses of the end layer: per-particle HBM atomics (FP64) or the
// window scatter over the ranges (FP32)
void end_masses(dmpm_ctx* c, const double* lay, double* gm)
{
    const double* cst = static_cast<const double*>(c->cst);
    for (int i = 0; i < c->B; ++i)
        launch_p2g_mass<double>(inst_params(c, i), lay + c->poff[i], cst + KM * c->S + c->poff[i], c->S, gm + c->nofs[i], c->stream);
}

void end_masses(dmpm_ctx* c, const float* lay, float* gm)
{
    launch_p2g_mass_blk(c->batch, lay, static_cast<const float*>(c->cst), gm, c->stream);
}

void loss_kernels(dmpm_ctx* c, const float* lay, const float* gm, double* dLdm)
{
    launch_loss_b(c->batch, gm, static_cast<const float*>(c->target), dLdm, c->dpart, c->dacc, c->gmax, c->stream);
    launch_loss_particles_b(c->batch, lay, static_cast<const float*>(c->cst), dLdm, grads_of<float>(c, c->L - 1), c->stream);
}

template <typename Real>
int end_loss(dmpm_ctx* c, double* loss)
{
    if (!c->have_target) return fail(DMPM_ERR_STATE, "loss before the target grid is set");
    if (c->fwd_valid < c->L - 1) return fail(DMPM_ERR_STATE, "loss before the forward pass reached the last layer");
    const int last = c->L - 1;
    Real* lay = (c->cfg.tape_mode == DMPM_TAPE_STORE_ALL) ? slot<Real>(c, last) : materialize<Real>(c, last);
    Real* gm = static_cast<Real*>(c->wgrid);
    // node masses in the first G reals; FP64 dL/dm after 2G reals of the
    // 4G-real work grid (8-byte aligned for both precisions)
    double* dLdm = reinterpret_cast<double*>(gm + 2 * c->G);
    fill<Real>(c, gm, c->G, Real(0));
    {
        BatchScope bs(c, layer_seg(c, last));
        timed(c, DMPM_K_LOSS, [&] {
            end_masses(c, lay, gm);
            loss_kernels(c, lay, gm, dLdm);
        });
    }
    taint_wgrid(c);
    HIP_TRY(hipGetLastError());
    c->inst_loss.assign(c->B, 0.0);
    HIP_TRY(hipMemcpyAsync(c->inst_loss.data(), c->dacc, c->B * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    std::vector<unsigned long long> ws(3 * (size_t)c->nstat, 0ull);
    if (c->wstats) HIP_TRY(hipMemcpyAsync(ws.data(), c->wstats, ws.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->last_swstats.assign(ws.begin(), ws.end());
    for (int k = 0; k < 3; ++k) {
        c->last_wstats[k] = 0;
        for (int i = 0; i < c->nstat; ++i) c->last_wstats[k] += (long long)ws[3 * (size_t)i + k];
    }
    double h = 0.0;
    for (double v : c->inst_loss) h += v;
    if (loss) *loss = h;
    c->loss_valid = true;
    c->grads_lo = last;
    // the control gradients of the last backward pass stay held (as the
    // reference's per-point dLdF): a line search reads them after new losses
    return DMPM_OK;
}

template <typename Real>
int backward_pass(dmpm_ctx* c, int control)
{
    if (!c->loss_valid) return fail(DMPM_ERR_STATE, "backward pass before EndLayerMassLoss");
    if (control < 0 || control > c->L - 2) return fail(DMPM_ERR_INVALID, "control_layer out of range");
    // graph key: backward flag, control layer, segments the tapes hold at the start
    const long long key = (1LL << 62) | ((long long)control << 32) | ((long long)(c->tape_hold[0] + 1) << 16) |
                          (long long)(c->tape_hold[1] + 1);
    int rc;
    if (c->cfg.tape_mode == DMPM_TAPE_STORE_ALL) {
        rc = run_graph(c, key, [&] {
            bool clean = false;
            for (int i = c->L - 2; i >= control; --i)
                clean = step_backward<Real>(c, i, slot<Real>(c, i), slot<Real>(c, i + 1), grid_of_step<Real>(c, i), clean, i > control);
        });
        if (rc) return rc;
    } else {
        const int seg_lo = control / c->K;
        const int hold0[2] = {c->tape_hold[0], c->tape_hold[1]};
        const bool ov = c->overlap && !c->prof && c->stream2;
        prepare_bidx(c);
        rc = run_graph(c, key, [&] {
            // the adjoint grid chain runs across the recomputed segments (they do not touch it)
            int hold[2] = {hold0[0], hold0[1]};
            bool pending[2] = {false, false};
            bool clean = false;
            for (int seg = c->seg_last; seg >= seg_lo; --seg) {
                const int t = seg & 1;
                const int s0 = seg * c->K, s1 = std::min(s0 + c->K, c->L - 1);
                if (pending[t]) {
                    hipStreamWaitEvent(c->stream, c->ev_join[t], 0);
                    pending[t] = false;
                } else if (hold[t] != seg) {
                    recompute_segment<Real>(c, seg);
                    hold[t] = seg;
                }
                // recompute the segment below into the other tape on stream2 while
                // this one steps back (the fork waits for the steps that last read
                // that tape, those of segment seg + 1)
                const int sp = seg - 1;
                if (ov && sp >= seg_lo && hold[sp & 1] != sp) {
                    hipEventRecord(c->ev_fork, c->stream);
                    hipStreamWaitEvent(c->stream2, c->ev_fork, 0);
                    std::swap(c->stream, c->stream2);
                    recompute_segment<Real>(c, sp);
                    std::swap(c->stream, c->stream2);
                    hipEventRecord(c->ev_join[sp & 1], c->stream2);
                    hold[sp & 1] = sp;
                    pending[sp & 1] = true;
                }
                // a boundary layer in another order: its gradients into this
                // segment's order, the layer itself from the tape's boundary slot
                const bool perm_b = s1 < c->L - 1 && boundary_permuted(c, seg + 1);
                if constexpr (std::is_same<Real, float>::value)
                    if (perm_b) grads_to_prev(c, grads_of<float>(c, s1), bidx_of(c, seg + 1));
                BatchScope bs(c, seg);
                for (int i = s1 - 1; i >= std::max(s0, control); --i) {
                    const Real* nxt = (i + 1 == s1 && perm_b) ? slot<Real>(c, c->tape_base[t] + c->K) : tape_slot<Real>(c, t, i + 1);
                    clean = step_backward<Real>(c, i, tape_slot<Real>(c, t, i), nxt, seg_grid<Real>(c, t, i), clean, i > control);
                }
            }
        });
        if (rc) return rc;
        // the pass ends holding the control layer's segment and the one above it
        c->tape_hold[seg_lo & 1] = seg_lo;
        if (seg_lo + 1 <= c->seg_last) c->tape_hold[(seg_lo + 1) & 1] = seg_lo + 1;
    }
    HIP_TRY(hipGetLastError());
    c->grads_lo = control;
    c->bwd_control = control;
    return DMPM_OK;
}

// ------------------------------------------------------- host conversion ---
template <typename Real>
int upload_fields(dmpm_ctx* c, Real* dst, int field0, int nfield, const double* src, int n)
{
    // AoS [n][nfield] double -> SoA fields field0.. of stride S
    std::vector<Real> h((size_t)nfield * c->S, Real(0));
    const bool perm = !c->pos.empty();
    for (int p = 0; p < n; ++p) {
        const size_t d = perm ? (size_t)c->pos[p] : (size_t)p;
        for (int f = 0; f < nfield; ++f) h[(size_t)f * c->S + d] = Real(src[(size_t)p * nfield + f]);
    }
    HIP_TRY(hipMemcpyAsync(dst + (size_t)field0 * c->S, h.data(), h.size() * sizeof(Real), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMPM_OK;
}

template <typename Real>
int download_fields(dmpm_ctx* c, const Real* src, int field0, int nfield, double* dst, int n)
{
    std::vector<Real> h((size_t)nfield * c->S);
    HIP_TRY(hipMemcpyAsync(h.data(), src + (size_t)field0 * c->S, h.size() * sizeof(Real), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const bool perm = !c->pos.empty();
    for (int p = 0; p < n; ++p) {
        const size_t d = perm ? (size_t)c->pos[p] : (size_t)p;
        for (int f = 0; f < nfield; ++f) dst[(size_t)p * nfield + f] = double(h[(size_t)f * c->S + d]);
    }
    return DMPM_OK;
}

// Positions (AoS [n][3], API order) to and from a layer's x field: shifted by
// the origin of the particle's instance.
template <typename Real>
int upload_positions(dmpm_ctx* c, Real* layer, const double* x)
{
    std::vector<double> xs((size_t)3 * c->N);
    for (int i = 0; i < c->B; ++i)
        for (int p = c->poff[i]; p < c->poff[i + 1]; ++p)
            for (int a = 0; a < 3; ++a) xs[3 * (size_t)p + a] = x[3 * (size_t)p + a] - c->origin[3 * (size_t)i + a];
    return upload_fields<Real>(c, layer, FX, 3, xs.data(), c->N);
}

template <typename Real>
int download_positions(dmpm_ctx* c, const Real* layer, double* x)
{
    const int rc = download_fields<Real>(c, layer, FX, 3, x, c->N);
    if (rc) return rc;
    for (int i = 0; i < c->B; ++i)
        for (int p = c->poff[i]; p < c->poff[i + 1]; ++p)
            for (int a = 0; a < 3; ++a) x[3 * (size_t)p + a] += c->origin[3 * (size_t)i + a];
    return DMPM_OK;
}

// Workgroup ranges: every block split into equal parts of at most range_max
// particles (the particle kernels hold two particles per thread of a
// 256-thread workgroup, so range_max <= 512).
int build_ranges(dmpm_ctx* c)
{
    std::vector<int4> rg = make_ranges(c, c->host_blocks);
    clear_graphs(c);
    for (auto& o : c->sord)
        if (o.own) {
            const int rc = upload_seg_ranges(c, o);
            if (rc) return rc;
        }
    if (c->ranges) {
        hipFree(c->ranges);
        c->ranges = nullptr;
    }
    c->host_ranges.swap(rg);
    c->n_ranges = (int)c->host_ranges.size();
    if (c->n_ranges > 0) {
        const int rc = dmalloc((void**)&c->ranges, c->host_ranges.size() * sizeof(int4), "block ranges");
        if (rc) {
            c->n_ranges = 0;
            c->host_ranges.clear();
            c->batch.ranges = nullptr;
            c->batch.nranges = 0;
            return rc;
        }
        HIP_TRY(hipMemcpy(c->ranges, c->host_ranges.data(), c->host_ranges.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    c->batch.ranges = c->ranges;
    c->batch.nranges = c->n_ranges;
    return DMPM_OK;
}

// Re-order the FP32 particles on the device (per instance by 4^3-cell block,
// rank in cell, cell in block; see dmpm_sort.hip) and permute every array
// that lives per particle across calls: layer 0, the constants, the API index
// map, the optimiser states, the controls (controls_live) and the held
// control gradients.  The trajectory after layer 0 and the layer gradients
// are invalidated; the block ranges are rebuilt from the new order.
int resort(dmpm_ctx* c, bool controls_live)
{
    const int N = c->N;
    const size_t S = c->S;
    std::vector<int> blk(N), api(N);
    if (N > 0) {
        SortArgs a;
        const int src = run_sort(c, slot<float>(c, 0), a);
        if (src) return src;
        const int* perm = a.v1;
        // every per-particle array through tmp (12 fields at a time) in the new order
        float* t = static_cast<float*>(c->tmp);
        auto permute = [&](float* base, long long nf) -> int {
            for (long long f = 0; f < nf; f += 12) {
                const int k = (int)std::min<long long>(12, nf - f);
                launch_gather<float>(base + f * S, t, (long long)S, N, k, perm, c->stream);
                HIP_TRY(hipMemcpyAsync(base + f * S, t, (size_t)k * S * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            }
            return DMPM_OK;
        };
        int rc;
        if ((rc = permute(slot<float>(c, 0), kLayerFields)) || (rc = permute(static_cast<float*>(c->cst), kConstFields)))
            return rc;
        if (controls_live && (rc = permute(static_cast<float*>(c->ctrl), (long long)(c->L - 1) * 9))) return rc;
        if (c->bwd_control >= 0 &&
            (rc = permute(static_cast<float*>(c->gctrl), (c->cfg.keep_control_grads ? (long long)(c->L - 1) : 1) * 9)))
            return rc;
        for (void* o : c->opt)
            if (o && (rc = permute(static_cast<float*>(o), 27))) return rc;
        launch_gather<int>(c->d_api, reinterpret_cast<int*>(t), (long long)S, N, 1, perm, c->stream);
        HIP_TRY(hipMemcpyAsync(c->d_api, t, (size_t)N * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(blk.data(), a.blk, (size_t)N * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(api.data(), c->d_api, (size_t)N * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    c->pos.assign(N, 0);
    for (int d = 0; d < N; ++d) c->pos[api[d]] = d;
    c->host_blocks = blocks_from(c, blk);
    // the segment orders map to the old home slots
    drop_seg_orders(c);
    const int rc = build_ranges(c);
    if (rc) return rc;
    ++c->n_sorts;
    for (long long& w : c->last_wstats) w = 0;
    std::fill(c->last_swstats.begin(), c->last_swstats.end(), 0LL);
    c->fwd_valid = std::min(c->fwd_valid, 0);
    c->loss_valid = false;
    c->grads_lo = -1;
    drop_tapes(c);
    return DMPM_OK;
}

// stale: in the last forward pass that reached the loss, more than
// resort_permille per mille of the P2G workgroups left the fixed window
bool order_stale(const dmpm_ctx* c)
{
    const long long tot = c->last_wstats[0] + c->last_wstats[1] + c->last_wstats[2];
    return tot > 0 && (c->last_wstats[1] + c->last_wstats[2]) * 1000 > (long long)c->resort_permille * tot;
}

template <typename Real>
int set_particles(dmpm_ctx* c, const double* x, const double* v, const double* F, const double* C, const double* m,
                  const double* vol, const double* lam, const double* mu)
{
    if (!x || !v || !F || !C || !m || !lam || !mu) return fail(DMPM_ERR_INVALID, "null particle array");
    if (c->f32) {
        // uploaded in API order, then ordered on the device
        c->pos.clear();
        launch_iota(c->d_api, c->N, c->stream);
    }
    // a new particle set starts with zero controls (MaterialPoint.h:31)
    fill<Real>(c, static_cast<Real*>(c->ctrl), (long long)(c->L - 1) * 9 * c->S, Real(0));
    Real* l0 = slot<Real>(c, 0);
    int rc;
    if ((rc = upload_positions<Real>(c, l0, x))) return rc;
    if ((rc = upload_fields<Real>(c, l0, FV, 3, v, c->N))) return rc;
    if ((rc = upload_fields<Real>(c, l0, FC, 9, C, c->N))) return rc;
    if ((rc = upload_fields<Real>(c, l0, FF, 9, F, c->N))) return rc;
    Real* cst = static_cast<Real*>(c->cst);
    if ((rc = upload_fields<Real>(c, cst, KM, 1, m, c->N))) return rc;
    if (vol) {
        if ((rc = upload_fields<Real>(c, cst, KVOL, 1, vol, c->N))) return rc;
    } else {
        fill<Real>(c, cst + (size_t)KVOL * c->S, c->S, Real(0));
    }
    if ((rc = upload_fields<Real>(c, cst, KLAM, 1, lam, c->N))) return rc;
    if ((rc = upload_fields<Real>(c, cst, KMU, 1, mu, c->N))) return rc;
    if (c->f32 && (rc = resort(c, false))) return rc;
    c->have_particles = true;
    c->fwd_valid = 0;
    c->loss_valid = false;
    c->grads_lo = -1;
    drop_tapes(c);
    return DMPM_OK;
}

template <typename Real>
int compute_volumes(dmpm_ctx* c)
{
    if (!c->have_particles) return fail(DMPM_ERR_STATE, "volumes before dmpm_set_particles");
    Real* gm = static_cast<Real*>(c->wgrid);
    Real* cst = static_cast<Real*>(c->cst);
    fill<Real>(c, gm, c->G, Real(0));
    for (int i = 0; i < c->B; ++i) {
        const StepParams sp = inst_params(c, i);
        const int p0 = c->poff[i];
        launch_p2g_mass<Real>(sp, slot<Real>(c, 0) + p0, cst + KM * c->S + p0, c->S, gm + c->nofs[i], c->stream);
        launch_volumes<Real>(sp, slot<Real>(c, 0) + p0, cst + p0, gm + c->nofs[i], c->stream);
    }
    taint_wgrid(c);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->fwd_valid = std::min(c->fwd_valid, 0);
    return DMPM_OK;
}

// P2G mass of point clouds into the target grids: counts[i] points for
// instance i, concatenated in instance order.
template <typename Real>
int target_from_points(dmpm_ctx* c, const double* x, const double* m, const int* counts)
{
    long long n = 0;
    for (int i = 0; i < c->B; ++i) {
        if (counts[i] < 0) return fail(DMPM_ERR_INVALID, "bad target point count");
        n += counts[i];
    }
    if (n > INT32_MAX) return fail(DMPM_ERR_INVALID, "too many target points");
    if (n > 0 && (!x || !m)) return fail(DMPM_ERR_INVALID, "bad target point cloud");
    Real* tg = static_cast<Real*>(c->target);
    void* buf = nullptr;
    if (n > 0) {
        // allocated before the target is touched: a failure leaves the old target in place
        const int rc = dmalloc(&buf, (size_t)4 * n * sizeof(Real), "target points");
        if (rc) return rc;
    }
    fill<Real>(c, tg, c->G, Real(0));
    if (n > 0) {
        // SoA [4][n]: x, y, z, m; instance i's points at column offset sum(counts[<i])
        std::vector<Real> h((size_t)4 * n);
        long long p = 0;
        for (int i = 0; i < c->B; ++i)
            for (int k = 0; k < counts[i]; ++k, ++p) {
                for (int a = 0; a < 3; ++a) h[(size_t)a * n + p] = Real(x[3 * (size_t)p + a] - c->origin[3 * (size_t)i + a]);
                h[(size_t)3 * n + p] = Real(m[p]);
            }
        Real* d = static_cast<Real*>(buf);
        hipError_t e2 = hipMemcpyAsync(d, h.data(), h.size() * sizeof(Real), hipMemcpyHostToDevice, c->stream);
        if (e2 == hipSuccess) {
            long long q0 = 0;
            for (int i = 0; i < c->B; ++i) {
                StepParams sp = inst_params(c, i);
                sp.n = counts[i];
                if (counts[i] > 0)
                    launch_p2g_mass<Real>(sp, d + q0, d + (size_t)3 * n + q0, (int)n, tg + c->nofs[i], c->stream);
                q0 += counts[i];
            }
        }
        hipError_t e3 = hipStreamSynchronize(c->stream);
        hipFree(buf);
        if (e2 != hipSuccess || e3 != hipSuccess) {
            c->have_target = false;  // the target was cleared
            return fail(DMPM_ERR_HIP, std::string("target P2G: ") + hipGetErrorString(e2 != hipSuccess ? e2 : e3));
        }
    }
    HIP_TRY(hipGetLastError());
    c->have_target = true;
    c->loss_valid = false;
    return DMPM_OK;
}

// target_from_points for one morph with the points staged in the file-scope
// scratch array (device address from scratch_init): no allocation at all.
template <typename Real>
int target_from_points_noalloc(dmpm_ctx* c, const double* x, const double* m, int n)
{
    if (c->B != 1) return fail(DMPM_ERR_STATE, "batched context: use dmpm_target_from_points_batch");
    if (n < 0) return fail(DMPM_ERR_INVALID, "bad target point count");
    if (n > kNoAllocPoints)
        return fail(DMPM_ERR_INVALID, "dmpm_target_from_points_noalloc takes at most " + std::to_string(kNoAllocPoints) +
                                          " points (got " + std::to_string(n) + "); use dmpm_target_from_points");
    if (n > 0 && (!x || !m)) return fail(DMPM_ERR_INVALID, "bad target point cloud");
    double* sx = static_cast<double*>(scratch_of(c->cfg.device));
    if (!sx) return fail(DMPM_ERR_STATE, "scratch not set up on this device");
    Real* tg = static_cast<Real*>(c->target);
    if (n > 0) {
        double* sm = sx + 3 * (size_t)n;
        HIP_TRY(hipMemcpyAsync(sx, x, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(sm, m, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    fill<Real>(c, tg, c->G, Real(0));
    if (n > 0) {
        StepParams sp = inst_params(c, 0);
        sp.n = n;
        launch_p2g_mass_aos<Real>(sp, sx, sx + 3 * (size_t)n, &c->origin[0], tg, c->stream);
    }
    HIP_TRY(hipGetLastError());
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->have_target = false;
        return fail(DMPM_ERR_HIP, std::string("target P2G: ") + hipGetErrorString(e));
    }
    c->have_target = true;
    c->loss_valid = false;
    return DMPM_OK;
}

template <typename Real>
int grid_download(dmpm_ctx* c, const Real* src, long long count, double* dst)
{
    std::vector<Real> h(count);
    HIP_TRY(hipMemcpyAsync(h.data(), src, count * sizeof(Real), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (long long i = 0; i < count; ++i) dst[i] = double(h[i]);
    return DMPM_OK;
}

template <typename Real>
int grid_upload(dmpm_ctx* c, Real* dst, long long count, const double* src)
{
    std::vector<Real> h(count);
    for (long long i = 0; i < count; ++i) h[i] = Real(src[i]);
    HIP_TRY(hipMemcpyAsync(dst, h.data(), count * sizeof(Real), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMPM_OK;
}

// A layer (or gradient layer) l held in segment s's order, gathered into the
// home order in a rolling slot (scratch between passes) other than l; l itself
// when segment s uses the home order.
template <typename Real>
Real* home_copy(dmpm_ctx* c, Real* l, int s)
{
    if constexpr (std::is_same<Real, float>::value) {
        if (!seg_own(c, s)) return l;
        float* st = slot<float>(c, c->roll_base) == l ? slot<float>(c, c->roll_base + 1) : slot<float>(c, c->roll_base);
        launch_gather<float>(l, st, (long long)c->S, c->N, kLayerFields, c->sord[s].inv, c->stream);
        return st;
    }
    return l;
}

template <typename Real>
int get_layer_state(dmpm_ctx* c, int layer, double* x, double* v, double* F, double* C)
{
    if (layer < 0 || layer >= c->L) return fail(DMPM_ERR_INVALID, "layer out of range");
    if (layer > c->fwd_valid) return fail(DMPM_ERR_STATE, "layer not computed by a forward pass");
    Real* l = (c->cfg.tape_mode == DMPM_TAPE_STORE_ALL) ? slot<Real>(c, layer) : materialize<Real>(c, layer);
    l = home_copy<Real>(c, l, layer_seg(c, layer));
    int rc;
    if (x && (rc = download_positions<Real>(c, l, x))) return rc;
    if (v && (rc = download_fields<Real>(c, l, FV, 3, v, c->N))) return rc;
    if (F && (rc = download_fields<Real>(c, l, FF, 9, F, c->N))) return rc;
    if (C && (rc = download_fields<Real>(c, l, FC, 9, C, c->N))) return rc;
    return DMPM_OK;
}

// PointCloud::GetPointDeterminants / GetPointElasticEnergies (PointCloud.cpp:441-467)
// of one layer, in FP64 on the device; the tmp buffer (12 S values of the
// path's precision, free between steps) holds the 2 S doubles.
template <typename Real>
int get_point_diagnostics(dmpm_ctx* c, int layer, double* J, double* psi)
{
    if (layer < 0 || layer >= c->L) return fail(DMPM_ERR_INVALID, "layer out of range");
    if (layer > c->fwd_valid) return fail(DMPM_ERR_STATE, "layer not computed by a forward pass");
    const Real* l = home_copy<Real>(c, materialize<Real>(c, layer), layer_seg(c, layer));
    const Real* ctrl = layer < c->L - 1 ? ctrl_of<Real>(c, layer) : nullptr;
    double* out = static_cast<double*>(c->tmp);
    launch_point_diag<Real>(c->S, l, ctrl, static_cast<const Real*>(c->cst), out, c->stream);
    HIP_TRY(hipGetLastError());
    int rc;
    if (J && (rc = download_fields<double>(c, out, 0, 1, J, c->N))) return rc;
    if (psi && (rc = download_fields<double>(c, out, 1, 1, psi, c->N))) return rc;
    return DMPM_OK;
}

template <typename Real>
int set_layer_state(dmpm_ctx* c, int layer, const double* x, const double* v, const double* F, const double* C)
{
    if (layer < 0 || layer >= c->L) return fail(DMPM_ERR_INVALID, "layer out of range");
    if (layer > 0 && c->cfg.tape_mode != DMPM_TAPE_STORE_ALL)
        return fail(DMPM_ERR_STATE, "checkpoint tape: only layer 0 can be written");
    if (layer > c->fwd_valid) return fail(DMPM_ERR_STATE, "layer not computed by a forward pass");
    Real* l = slot<Real>(c, layer);
    int rc;
    if (x && (rc = upload_positions<Real>(c, l, x))) return rc;
    if (v && (rc = upload_fields<Real>(c, l, FV, 3, v, c->N))) return rc;
    if (F && (rc = upload_fields<Real>(c, l, FF, 9, F, c->N))) return rc;
    if (C && (rc = upload_fields<Real>(c, l, FC, 9, C, c->N))) return rc;
    c->fwd_valid = layer;
    c->loss_valid = false;
    c->grads_lo = -1;
    if (c->cfg.tape_mode != DMPM_TAPE_STORE_ALL) drop_tapes(c);
    return DMPM_OK;
}

template <typename Real>
int get_layer_grads(dmpm_ctx* c, int layer, double* gx, double* gv, double* gF, double* gC)
{
    const bool seed = layer == c->L - 1 && c->loss_valid;  // the last layer's seed stays held
    if (!seed && (c->grads_lo < 0 || layer < c->grads_lo || layer > std::min(c->grads_lo + 1, c->L - 1)))
        return fail(DMPM_ERR_STATE,
                    "gradients of this layer are not held (only the last layer, the lowest layer of the last pass and the one above)");
    // held in the order of the lowest layer's segment (the last layer's: the last segment's)
    Real* g = home_copy<Real>(c, grads_of<Real>(c, layer), layer == c->L - 1 ? c->seg_last : layer_seg(c, c->grads_lo));
    int rc;
    if (gx && (rc = download_fields<Real>(c, g, FX, 3, gx, c->N))) return rc;
    if (gv && (rc = download_fields<Real>(c, g, FV, 3, gv, c->N))) return rc;
    if (gF && (rc = download_fields<Real>(c, g, FF, 9, gF, c->N))) return rc;
    if (gC && (rc = download_fields<Real>(c, g, FC, 9, gC, c->N))) return rc;
    return DMPM_OK;
}

template <typename Real>
int set_layer_grads(dmpm_ctx* c, int layer, const double* gx, const double* gv, const double* gF, const double* gC)
{
    if (layer < 0 || layer >= c->L) return fail(DMPM_ERR_INVALID, "layer out of range");
    if (layer > c->fwd_valid) return fail(DMPM_ERR_STATE, "layer not computed by a forward pass");
    Real* dst = grads_of<Real>(c, layer);
    // in the order of the layer's segment: through a rolling slot in the home order
    const int sg = layer_seg(c, layer);
    Real* g = dst;
    if constexpr (std::is_same<Real, float>::value)
        if (seg_own(c, sg)) g = slot<float>(c, c->roll_base);
    fill<Real>(c, g, (long long)kLayerFields * c->S, Real(0));
    int rc;
    if (gx && (rc = upload_fields<Real>(c, g, FX, 3, gx, c->N))) return rc;
    if (gv && (rc = upload_fields<Real>(c, g, FV, 3, gv, c->N))) return rc;
    if (gF && (rc = upload_fields<Real>(c, g, FF, 9, gF, c->N))) return rc;
    if (gC && (rc = upload_fields<Real>(c, g, FC, 9, gC, c->N))) return rc;
    if constexpr (std::is_same<Real, float>::value)
        if (g != dst) launch_gather<float>(g, dst, (long long)c->S, c->N, kLayerFields, c->sord[sg].home, c->stream);
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->grads_lo = layer;
    c->bwd_control = -1;
    c->loss_valid = (layer == c->L - 1);
    return DMPM_OK;
}

// The control gradient of `layer` is held: the control layer of the last
// backward pass, or any layer it reached when every layer's gradient is kept.
int check_control_grad(dmpm_ctx* c, int layer)
{
    if (c->bwd_control < 0) return fail(DMPM_ERR_STATE, "no backward pass has run");
    bool ok = c->cfg.keep_control_grads ? (layer >= c->bwd_control && layer <= c->L - 2) : (layer == c->bwd_control);
    if (!ok) return fail(DMPM_ERR_STATE, "control gradient of this layer is not kept");
    return DMPM_OK;
}

template <typename Real>
int get_control_grad(dmpm_ctx* c, int layer, double* out)
{
    const int rc = check_control_grad(c, layer);
    if (rc) return rc;
    return download_fields<Real>(c, gctrl_of<Real>(c, layer), 0, 9, out, c->N);
}

template <typename Real>
int chain_episode(dmpm_ctx* c)
{
    if (c->fwd_valid < c->L - 1) return fail(DMPM_ERR_STATE, "chain before the forward pass reached the last layer");
    Real* last = (c->cfg.tape_mode == DMPM_TAPE_STORE_ALL) ? slot<Real>(c, c->L - 1) : materialize<Real>(c, c->L - 1);
    bool own_any = false;
    if constexpr (std::is_same<Real, float>::value) {
        for (int s = 1; s <= c->seg_last && s < (int)c->sord.size(); ++s) own_any = own_any || seg_own(c, s);
        if (seg_own(c, c->seg_last))  // back to the home order
            launch_gather<float>(last, slot<float>(c, 0), (long long)c->S, c->N, kLayerFields, c->sord[c->seg_last].inv, c->stream);
        else
            HIP_TRY(hipMemcpyAsync(slot<Real>(c, 0), last, (size_t)kLayerFields * c->S * sizeof(Real), hipMemcpyDeviceToDevice, c->stream));
        // the segment orders were those of the episode that ended
        drop_seg_orders(c);
    } else {
        HIP_TRY(hipMemcpyAsync(slot<Real>(c, 0), last, (size_t)kLayerFields * c->S * sizeof(Real), hipMemcpyDeviceToDevice, c->stream));
    }
    fill<Real>(c, static_cast<Real*>(c->ctrl), (long long)(c->L - 1) * 9 * c->S, Real(0));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->fwd_valid = 0;
    c->loss_valid = false;
    c->grads_lo = -1;
    c->bwd_control = -1;
    drop_tapes(c);
    // the particles moved over the episode: order the new layer 0 afresh
    // (mode 1), or when the episode's P2G windows showed a stale order or a
    // segment needed an order of its own (mode 2)
    if (c->f32 && (c->resort_mode == 1 || (c->resort_mode == 2 && (order_stale(c) || own_any)))) return resort(c, false);
    return DMPM_OK;
}

// Grid of one step: node masses (G) and velocities ([G][3]), instances
// concatenated.  The FP32 grid holds the P2G sums m, p: the velocities are
// the grid update's, v = p/m + dt f where |m| > FLT_EPSILON, else 0, in the
// float operations the kernels use (correctly rounded 1/m, then one fma).
template <typename Real>
int get_grid(dmpm_ctx* c, int step, double* m, double* v)
{
    const Real* g = grid_of_step<Real>(c, step);
    std::vector<double> t((size_t)kGridFields * c->G);
    int rc = grid_download<Real>(c, g, kGridFields * c->G, t.data());
    if (rc) return rc;
    for (int i = 0; i < c->B; ++i) {
        const long long n0 = c->nofs[i], Gi = c->nofs[i + 1] - n0;
        const double* gi = t.data() + 4 * n0;
        const StepParams sp = inst_params(c, i);
        float fdt[3];
        for (int a = 0; a < 3; ++a) fdt[a] = float(sp.dt) * float(sp.f_ext[a]);
        for (long long q = 0; q < Gi; ++q) {
            if (m) m[n0 + q] = gi[q];
            if (!v) continue;
            if (c->f32) {
                const float mq = float(gi[q]);
                const bool ok = std::fabs(mq) > 1.1920928955078125e-07f;  // FLT_EPSILON
                const float inv = 1.f / mq;
                for (int a = 0; a < 3; ++a)
                    v[3 * (n0 + q) + a] = ok ? double(std::fmaf(float(gi[(1 + a) * Gi + q]), inv, fdt[a])) : 0.0;
            } else {
                for (int a = 0; a < 3; ++a) v[3 * (n0 + q) + a] = gi[(1 + a) * Gi + q];
            }
        }
    }
    return DMPM_OK;
}

#define DISPATCH(ctx, fn, ...) ((ctx)->f32 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

int check_ctx(dmpm_ctx* c)
{
    if (!c) return fail(DMPM_ERR_INVALID, "null context");
    HIP_TRY(hipSetDevice(c->cfg.device));
    return DMPM_OK;
}

}  // namespace

// ================================================================ C ABI ====
extern "C" {

int dmpm_version(void) { return 100; }

const char* dmpm_last_error(void) { return g_err.c_str(); }

int dmpm_config_init(dmpm_config* cfg)
{
    if (!cfg) return fail(DMPM_ERR_INVALID, "null config");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->precision = DMPM_PREC_F64;
    cfg->num_layers = 2;
    cfg->dx = 1.0;
    cfg->dt = 1.0 / 120.0;
    cfg->drag = 0.5;
    cfg->smoothing = 0.1;
    cfg->tape_mode = DMPM_TAPE_STORE_ALL;
    return DMPM_OK;
}

}  // extern "C"

namespace {

int create_impl(const dmpm_config* cfg, const std::vector<dmpm_instance>& insts, dmpm_ctx** out)
{
    *out = nullptr;
    if (cfg->precision != DMPM_PREC_F64 && cfg->precision != DMPM_PREC_F32) return fail(DMPM_ERR_INVALID, "precision");
    long long ntot = 0;
    for (const dmpm_instance& in : insts) {
        if (in.n_particles < 0) return fail(DMPM_ERR_INVALID, "n_particles < 0");
        if (in.dim[0] <= 0 || in.dim[1] <= 0 || in.dim[2] <= 0) return fail(DMPM_ERR_INVALID, "grid dims must be > 0");
        if (!(in.dx > 0.0)) return fail(DMPM_ERR_INVALID, "dx must be > 0");
        ntot += in.n_particles;
    }
    if (ntot > INT32_MAX - 64) return fail(DMPM_ERR_INVALID, "too many particles");
    if (cfg->num_layers < 2) return fail(DMPM_ERR_INVALID, "num_layers must be >= 2");
    if (cfg->tape_mode != DMPM_TAPE_STORE_ALL && cfg->tape_mode != DMPM_TAPE_CHECKPOINT) return fail(DMPM_ERR_INVALID, "tape_mode");
    if (cfg->compat < 0 || cfg->compat > DMPM_COMPAT_REFERENCE) return fail(DMPM_ERR_INVALID, "compat bits");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (cfg->device < 0 || cfg->device >= ndev) return fail(DMPM_ERR_INVALID, "device ordinal out of range");
    HIP_TRY(hipSetDevice(cfg->device));
    if (const int rs = scratch_init(cfg->device)) return rs;

    dmpm_ctx* c = new dmpm_ctx();
    c->cfg = *cfg;
    c->f32 = cfg->precision == DMPM_PREC_F32;
    c->esz = c->f32 ? 4 : 8;
    c->lf = c->f32 ? kLayerFieldsF32 : kLayerFields;
    c->B = (int)insts.size();
    c->inst = insts;
    c->poff.assign(c->B + 1, 0);
    c->nofs.assign(c->B + 1, 0);
    for (int i = 0; i < c->B; ++i) {
        const dmpm_instance& in = insts[i];
        const long long Gi = (long long)in.dim[0] * in.dim[1] * in.dim[2];
        c->poff[i + 1] = c->poff[i] + in.n_particles;
        c->nofs[i + 1] = c->nofs[i] + Gi;
        c->gmax = std::max(c->gmax, Gi);
        for (int d = 0; d < 3; ++d) c->origin.push_back(in.min_point[d] + 0.5 * in.dim[d] * in.dx);
    }
    c->N = c->poff[c->B];
    c->cfg.n_particles = c->N;
    c->S = ((c->N + 63) / 64) * 64;
    if (c->S == 0) c->S = 64;
    c->L = cfg->num_layers;
    c->G = c->nofs[c->B];
    const int steps = c->L - 1;
    if (cfg->tape_mode == DMPM_TAPE_CHECKPOINT) {
        int K = cfg->ckpt_stride > 0 ? cfg->ckpt_stride : default_stride(steps);
        if (K < 1) K = 1;
        if (K > steps) K = steps;
        c->K = K;
        c->n_ckpt = (c->L - 1) / K + 1;  // layers 0, K, 2K, ... <= L-1
        c->seg_last = (c->L - 2) / K;
        // two tapes of K+1 slots (slot 0 unused: the segment start is a
        // checkpoint; slot K: the segment's end layer when the next segment
        // holds its checkpoint in another particle order)
        c->tape_base[0] = c->n_ckpt;
        c->tape_base[1] = c->n_ckpt + K + 1;
        c->roll_base = c->tape_base[1] + K + 1;
        c->n_slots = c->roll_base + 2;
    } else {
        c->K = steps;
        c->n_slots = c->L;
    }
    // FP32 checkpoint tape: window statistics and particle orders per segment
    c->nstat = (c->f32 && cfg->tape_mode == DMPM_TAPE_CHECKPOINT) ? c->seg_last + 1 : 1;
    c->sord.resize(c->nstat > 1 ? (size_t)c->nstat : 0);
    c->last_swstats.assign(3 * (size_t)c->nstat, 0);
    const size_t e = c->esz;
    const size_t S = c->S;
    const size_t G = c->G;
    int rc = DMPM_OK;
    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete c;
        return fail(DMPM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    if (cfg->tape_mode == DMPM_TAPE_CHECKPOINT) {
        // the recompute stream (default priority) and its fork / join events
        hipError_t e1 = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
        hipError_t e2 = e1 == hipSuccess ? hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) : e1;
        hipError_t e3 = e2 == hipSuccess ? hipEventCreateWithFlags(&c->ev_join[0], hipEventDisableTiming) : e2;
        hipError_t e4 = e3 == hipSuccess ? hipEventCreateWithFlags(&c->ev_join[1], hipEventDisableTiming) : e3;
        if (e4 != hipSuccess) {
            dmpm_destroy(c);
            return fail(DMPM_ERR_HIP, std::string("recompute stream: ") + hipGetErrorString(e4));
        }
    }
    // FP32 checkpoint tape: the two tapes' grids, then one grid per checkpoint step
    c->ck_grids = (c->f32 && cfg->tape_mode == DMPM_TAPE_CHECKPOINT) ? c->seg_last + 1 : 0;
    const size_t grid_slots = cfg->tape_mode == DMPM_TAPE_CHECKPOINT ? 2 * (size_t)c->K + c->ck_grids : (size_t)steps;
    c->grid_slots = grid_slots;
    c->ntiles = c->f32 ? num_tiles((long long)G) : 0;
    const size_t gctrl_layers = cfg->keep_control_grads ? (size_t)steps : 1;
    if ((rc = dalloc(c, &c->cst, kConstFields * S * e)) || (rc = dalloc(c, &c->ctrl, (size_t)steps * 9 * S * e)) ||
        (rc = dalloc(c, &c->layers, (size_t)c->n_slots * c->lf * S * e)) ||
        (rc = dalloc(c, &c->grids, grid_slots * kGridFields * G * e)) || (rc = dalloc(c, &c->wgrid, kGridFields * G * e)) ||
        (rc = dalloc(c, &c->wgrid2, kGridFields * G * e)) || (c->f32 && (rc = dalloc(c, &c->wgrid3, kGridFields * G * e))) ||
        (rc = dalloc(c, &c->agrid, kGridFields * G * e)) || (rc = dalloc(c, &c->gbuf[0], kLayerFields * S * e)) ||
        (rc = dalloc(c, &c->gbuf[1], kLayerFields * S * e)) || (rc = dalloc(c, &c->gbuf[2], kLayerFields * S * e)) ||
        (rc = dalloc(c, &c->gctrl, gctrl_layers * 9 * S * e)) ||
        (rc = dalloc(c, &c->tmp, 12 * S * e)) || (rc = dalloc(c, &c->target, G * e)) ||
        (rc = dalloc(c, (void**)&c->dacc, std::max(64, c->B + 8) * sizeof(double))) ||
        (rc = dalloc(c, (void**)&c->dpart, std::max<long long>(kLossBlocks, loss_partials(c->B)) * sizeof(double))) ||
        (rc = dalloc(c, (void**)&c->d_sps, c->B * sizeof(StepParams))) ||
        (rc = dalloc(c, (void**)&c->d_nofs, (c->B + 1) * sizeof(long long))) ||
        (c->f32 && (rc = dalloc(c, (void**)&c->tflags, (grid_slots + 4) * (size_t)c->ntiles * sizeof(unsigned)))) ||
        (c->f32 && (rc = dalloc(c, (void**)&c->d_api, S * sizeof(int)))) ||
        (c->f32 && (rc = dalloc(c, (void**)&c->d_poff, (c->B + 1) * sizeof(int)))) ||
        (c->f32 && (rc = dalloc(c, (void**)&c->d_boff, (c->B + 1) * sizeof(long long)))) ||
        (c->f32 && (rc = dalloc(c, (void**)&c->wstats, 3 * (size_t)c->nstat * sizeof(unsigned long long)))) ||
        (c->f32 && (rc = dalloc(c, (void**)&c->queue, 6 * sizeof(int)))) ||
        (c->f32 && (rc = dalloc(c, (void**)&c->defer, S * sizeof(int2))))) {
        std::string msg = g_err;
        dmpm_destroy(c);
        return fail(rc, msg);
    }
    // zero controls and layer 0 so a context is well defined before uploads
    if (c->f32) {
        launch_fill<float>(static_cast<float*>(c->gctrl), (long long)gctrl_layers * 9 * S, 0.f, c->stream);
        launch_fill<float>(static_cast<float*>(c->ctrl), (long long)steps * 9 * S, 0.f, c->stream);
        launch_fill<float>(static_cast<float*>(c->layers), (long long)c->lf * S, 0.f, c->stream);
        launch_fill<float>(static_cast<float*>(c->cst), (long long)kConstFields * S, 0.f, c->stream);
        // grids start at 0 with no tile flagged
        launch_fill<float>(static_cast<float*>(c->grids), (long long)(grid_slots * kGridFields * G), 0.f, c->stream);
        launch_fill<float>(static_cast<float*>(c->wgrid), (long long)(kGridFields * G), 0.f, c->stream);
        launch_fill<float>(static_cast<float*>(c->wgrid2), (long long)(kGridFields * G), 0.f, c->stream);
        launch_fill<float>(static_cast<float*>(c->wgrid3), (long long)(kGridFields * G), 0.f, c->stream);
        launch_fill<float>(static_cast<float*>(c->agrid), (long long)(kGridFields * G), 0.f, c->stream);
        launch_fill<float>(reinterpret_cast<float*>(c->tflags), (long long)(grid_slots + 4) * c->ntiles, 0.f, c->stream);
    } else {
        launch_fill<double>(static_cast<double*>(c->gctrl), (long long)gctrl_layers * 9 * S, 0.0, c->stream);
        launch_fill<double>(static_cast<double*>(c->ctrl), (long long)steps * 9 * S, 0.0, c->stream);
        launch_fill<double>(static_cast<double*>(c->layers), (long long)kLayerFields * S, 0.0, c->stream);
        launch_fill<double>(static_cast<double*>(c->cst), (long long)kConstFields * S, 0.0, c->stream);
    }
    set_block_prefix(c);
    c->nblocks4 = c->boff[c->B];
    {
        std::vector<StepParams> sps(c->B);
        for (int i = 0; i < c->B; ++i) sps[i] = inst_params(c, i);
        hipError_t e1 = hipMemcpyAsync(c->d_sps, sps.data(), sps.size() * sizeof(StepParams), hipMemcpyHostToDevice, c->stream);
        hipError_t e2 = hipMemcpyAsync(c->d_nofs, c->nofs.data(), c->nofs.size() * sizeof(long long), hipMemcpyHostToDevice, c->stream);
        hipError_t e3 = hipSuccess, e4 = hipSuccess;
        if (c->f32) {
            e3 = hipMemcpyAsync(c->d_poff, c->poff.data(), c->poff.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
            e4 = hipMemcpyAsync(c->d_boff, c->boff.data(), c->boff.size() * sizeof(long long), hipMemcpyHostToDevice, c->stream);
            launch_fill<double>(reinterpret_cast<double*>(c->wstats), 3LL * c->nstat, 0.0, c->stream);
            launch_fill<float>(reinterpret_cast<float*>(c->queue), 6, 0.f, c->stream);  // the bits of int 0
        }
        he = hipStreamSynchronize(c->stream);
        if (he == hipSuccess) he = e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : (e3 != hipSuccess ? e3 : e4));
    }
    c->batch.sps = c->d_sps;
    c->batch.nofs = c->d_nofs;
    c->batch.ranges = nullptr;
    c->batch.nranges = 0;
    c->batch.B = c->B;
    c->batch.Gtot = c->G;
    c->batch.gmax = c->gmax;
    c->batch.cap = c->lds_cap;
    c->batch.wstats = c->wstats;
    if (he != hipSuccess) {
        dmpm_destroy(c);
        return fail(DMPM_ERR_HIP, std::string("init: ") + hipGetErrorString(he));
    }
    *out = c;
    return DMPM_OK;
}

}  // namespace

extern "C" {

int dmpm_create(const dmpm_config* cfg, dmpm_ctx** out)
{
    if (!cfg || !out) return fail(DMPM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->n_particles < 0) return fail(DMPM_ERR_INVALID, "n_particles < 0");
    dmpm_instance in{};
    in.n_particles = cfg->n_particles;
    for (int a = 0; a < 3; ++a) {
        in.dim[a] = cfg->dim[a];
        in.min_point[a] = cfg->min_point[a];
    }
    in.dx = cfg->dx;
    return create_impl(cfg, {in}, out);
}

int dmpm_create_batch(const dmpm_config* cfg, int32_t n_instances, const dmpm_instance* instances, dmpm_ctx** out)
{
    if (!cfg || !out || !instances) return fail(DMPM_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n_instances < 1) return fail(DMPM_ERR_INVALID, "n_instances must be >= 1");
    if (cfg->precision != DMPM_PREC_F32) return fail(DMPM_ERR_INVALID, "batched contexts run the FP32 path (precision DMPM_PREC_F32)");
    return create_impl(cfg, std::vector<dmpm_instance>(instances, instances + n_instances), out);
}

int dmpm_destroy(dmpm_ctx* c)
{
    if (!c) return DMPM_OK;
    hipSetDevice(c->cfg.device);
    if (c->stream) hipStreamSynchronize(c->stream);
    clear_graphs(c);
    void* bufs[] = {c->cst,  c->ctrl, c->layers, c->grids, c->wgrid, c->wgrid2, c->wgrid3, c->agrid, c->gbuf[0], c->gbuf[1], c->gbuf[2],
                    c->gctrl, c->tmp, c->target, c->dacc,  c->ranges, c->d_sps, c->d_nofs, c->dpart, c->tflags,
                    c->d_api, c->d_poff, c->d_boff, c->sortbuf, c->wstats, c->queue, c->defer};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (void* b : c->opt)
        if (b) hipFree(b);
    for (auto& o : c->sord)
        for (void* b : {(void*)o.home, (void*)o.inv, (void*)o.bidx, (void*)o.ranges})
            if (b) hipFree(b);
    for (hipEvent_t ev : c->pool) hipEventDestroy(ev);
    for (hipEvent_t ev : {c->ev_fork, c->ev_join[0], c->ev_join[1]})
        if (ev) hipEventDestroy(ev);
    if (c->stream2) {
        hipStreamSynchronize(c->stream2);
        hipStreamDestroy(c->stream2);
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return DMPM_OK;
}

int dmpm_set_particles(dmpm_ctx* c, const double* x, const double* v, const double* F, const double* C,
                       const double* m, const double* vol, const double* lam, const double* mu)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, set_particles, c, x, v, F, C, m, vol, lam, mu);
}

int dmpm_compute_volumes(dmpm_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, compute_volumes, c);
}

int dmpm_get_volumes(dmpm_ctx* c, double* vol)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!vol) return fail(DMPM_ERR_INVALID, "null output");
    if (c->f32) return download_fields<float>(c, static_cast<float*>(c->cst), KVOL, 1, vol, c->N);
    return download_fields<double>(c, static_cast<double*>(c->cst), KVOL, 1, vol, c->N);
}

int dmpm_set_control(dmpm_ctx* c, int32_t layer, const double* dFc)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!dFc) return fail(DMPM_ERR_INVALID, "null control");
    if (layer < -1 || layer > c->L - 2) return fail(DMPM_ERR_INVALID, "control layer out of range");
    const int lo = layer < 0 ? 0 : layer, hi = layer < 0 ? c->L - 2 : layer;
    for (int l = lo; l <= hi; ++l) {
        rc = c->f32 ? upload_fields<float>(c, ctrl_of<float>(c, l), 0, 9, dFc, c->N)
                    : upload_fields<double>(c, ctrl_of<double>(c, l), 0, 9, dFc, c->N);
        if (rc) return rc;
    }
    c->fwd_valid = std::min(c->fwd_valid, lo);
    c->loss_valid = false;
    if (c->cfg.tape_mode != DMPM_TAPE_STORE_ALL) drop_tapes_from(c, lo);
    return DMPM_OK;
}

int dmpm_get_control(dmpm_ctx* c, int32_t layer, double* dFc)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!dFc) return fail(DMPM_ERR_INVALID, "null output");
    if (layer < 0 || layer > c->L - 2) return fail(DMPM_ERR_INVALID, "control layer out of range");
    return c->f32 ? download_fields<float>(c, ctrl_of<float>(c, layer), 0, 9, dFc, c->N)
                  : download_fields<double>(c, ctrl_of<double>(c, layer), 0, 9, dFc, c->N);
}

int dmpm_set_target_grid(dmpm_ctx* c, const double* mass)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!mass) return fail(DMPM_ERR_INVALID, "null target");
    rc = c->f32 ? grid_upload<float>(c, static_cast<float*>(c->target), c->G, mass)
                : grid_upload<double>(c, static_cast<double*>(c->target), c->G, mass);
    if (rc) return rc;
    c->have_target = true;
    c->loss_valid = false;
    return DMPM_OK;
}

int dmpm_target_from_points(dmpm_ctx* c, const double* x, const double* m, int32_t n)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->B != 1) return fail(DMPM_ERR_STATE, "batched context: use dmpm_target_from_points_batch");
    const int counts[1] = {n};
    return DISPATCH(c, target_from_points, c, x, m, counts);
}

int dmpm_target_from_points_noalloc(dmpm_ctx* c, const double* x, const double* m, int32_t n)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, target_from_points_noalloc, c, x, m, n);
}

int dmpm_target_from_points_batch(dmpm_ctx* c, const double* x, const double* m, const int32_t* counts)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!counts) return fail(DMPM_ERR_INVALID, "null counts");
    return DISPATCH(c, target_from_points, c, x, m, counts);
}

int dmpm_num_instances(dmpm_ctx* c, int32_t* n)
{
    if (!c || !n) return fail(DMPM_ERR_INVALID, "null argument");
    *n = c->B;
    return DMPM_OK;
}

int dmpm_get_instance_losses(dmpm_ctx* c, double* losses)
{
    if (!c || !losses) return fail(DMPM_ERR_INVALID, "null argument");
    if ((int)c->inst_loss.size() != c->B) return fail(DMPM_ERR_STATE, "no loss has been evaluated");
    for (int i = 0; i < c->B; ++i) losses[i] = c->inst_loss[i];
    return DMPM_OK;
}

int dmpm_get_target_grid(dmpm_ctx* c, double* mass)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!mass) return fail(DMPM_ERR_INVALID, "null output");
    return c->f32 ? grid_download<float>(c, static_cast<float*>(c->target), c->G, mass)
                  : grid_download<double>(c, static_cast<double*>(c->target), c->G, mass);
}

int dmpm_forward_pass(dmpm_ctx* c, int32_t start)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, forward_pass, c, start);
}

int dmpm_end_layer_mass_loss(dmpm_ctx* c, double* loss)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, end_loss, c, loss);
}

int dmpm_backward_pass(dmpm_ctx* c, int32_t control)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, backward_pass, c, control);
}

int dmpm_forward_step(dmpm_ctx* c, int32_t layer)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->cfg.tape_mode != DMPM_TAPE_STORE_ALL) return fail(DMPM_ERR_STATE, "single steps need the store-all tape");
    if (layer < 0 || layer > c->L - 2) return fail(DMPM_ERR_INVALID, "layer out of range");
    if (layer > c->fwd_valid) return fail(DMPM_ERR_STATE, "input layer not computed");
    if (c->f32)
        step_forward<float>(c, slot<float>(c, layer), slot<float>(c, layer + 1), ctrl_of<float>(c, layer), grid_of_step<float>(c, layer));
    else
        step_forward<double>(c, slot<double>(c, layer), slot<double>(c, layer + 1), ctrl_of<double>(c, layer), grid_of_step<double>(c, layer));
    HIP_TRY(hipGetLastError());
    c->fwd_valid = layer + 1;
    c->loss_valid = false;
    return DMPM_OK;
}

int dmpm_backward_step(dmpm_ctx* c, int32_t layer)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->cfg.tape_mode != DMPM_TAPE_STORE_ALL) return fail(DMPM_ERR_STATE, "single steps need the store-all tape");
    if (layer < 0 || layer > c->L - 2) return fail(DMPM_ERR_INVALID, "layer out of range");
    if (c->grads_lo != layer + 1) return fail(DMPM_ERR_STATE, "gradients of layer+1 are not the current ones");
    if (layer + 1 > c->fwd_valid) return fail(DMPM_ERR_STATE, "forward layers missing");
    if (c->f32)
        step_backward<float>(c, layer, slot<float>(c, layer), slot<float>(c, layer + 1), grid_of_step<float>(c, layer));
    else
        step_backward<double>(c, layer, slot<double>(c, layer), slot<double>(c, layer + 1), grid_of_step<double>(c, layer));
    HIP_TRY(hipGetLastError());
    c->grads_lo = layer;
    c->bwd_control = layer;
    return DMPM_OK;
}

int dmpm_get_layer_state(dmpm_ctx* c, int32_t layer, double* x, double* v, double* F, double* C)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, get_layer_state, c, layer, x, v, F, C);
}

int dmpm_get_point_diagnostics(dmpm_ctx* c, int32_t layer, double* J, double* psi)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, get_point_diagnostics, c, layer, J, psi);
}

int dmpm_set_layer_state(dmpm_ctx* c, int32_t layer, const double* x, const double* v, const double* F, const double* C)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, set_layer_state, c, layer, x, v, F, C);
}

int dmpm_get_layer_grads(dmpm_ctx* c, int32_t layer, double* dLdx, double* dLdv, double* dLdF, double* dLdC)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, get_layer_grads, c, layer, dLdx, dLdv, dLdF, dLdC);
}

int dmpm_set_layer_grads(dmpm_ctx* c, int32_t layer, const double* dLdx, const double* dLdv, const double* dLdF,
                         const double* dLdC)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, set_layer_grads, c, layer, dLdx, dLdv, dLdF, dLdC);
}

int dmpm_get_control_grad(dmpm_ctx* c, int32_t layer, double* out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!out) return fail(DMPM_ERR_INVALID, "null output");
    return DISPATCH(c, get_control_grad, c, layer, out);
}

int dmpm_control_grad_norm(dmpm_ctx* c, int32_t layer, double* norm)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!norm) return fail(DMPM_ERR_INVALID, "null output");
    if ((rc = check_control_grad(c, layer))) return rc;
    double* acc = c->dacc + c->B;
    if (c->f32) launch_sumsq<float>(gctrl_of<float>(c, layer), 9LL * c->S, c->dpart, acc, c->stream);
    else launch_sumsq<double>(gctrl_of<double>(c, layer), 9LL * c->S, c->dpart, acc, c->stream);
    HIP_TRY(hipGetLastError());
    double h = 0.0;
    HIP_TRY(hipMemcpyAsync(&h, acc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *norm = std::sqrt(h);
    return DMPM_OK;
}

int dmpm_descend(dmpm_ctx* c, int32_t layer, int32_t rule, float alpha, float gradient_norm, float beta1, float beta2,
                 float epsilon, int32_t timestep)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (rule < DMPM_DESCEND_ADAM || rule > DMPM_DESCEND_GD) return fail(DMPM_ERR_INVALID, "descend rule");
    if ((rc = check_control_grad(c, layer))) return rc;
    if (c->opt.empty()) c->opt.assign(c->L - 1, nullptr);
    const size_t n = (size_t)9 * c->S;
    if (!c->opt[layer]) {
        if ((rc = dalloc(c, &c->opt[layer], 3 * n * c->esz))) return rc;
        if (c->f32) launch_fill<float>(static_cast<float*>(c->opt[layer]), 3 * (long long)n, 0.f, c->stream);
        else launch_fill<double>(static_cast<double*>(c->opt[layer]), 3 * (long long)n, 0.0, c->stream);
    }
    DescendParams dp{};
    dp.rule = rule;
    dp.alpha = alpha;
    dp.gradient_norm = gradient_norm;
    dp.beta1 = beta1;
    dp.beta2 = beta2;
    dp.epsilon = epsilon;
    dp.bc1 = (float)(1.0 - std::pow((double)beta1, (double)timestep));
    dp.bc2 = (float)(1.0 - std::pow((double)beta2, (double)timestep));
    if (c->f32) {
        float* st = static_cast<float*>(c->opt[layer]);
        launch_descend<float>((long long)n, dp, ctrl_of<float>(c, layer), gctrl_of<float>(c, layer), st, st + n, st + 2 * n, c->stream);
    } else {
        double* st = static_cast<double*>(c->opt[layer]);
        launch_descend<double>((long long)n, dp, ctrl_of<double>(c, layer), gctrl_of<double>(c, layer), st, st + n, st + 2 * n, c->stream);
    }
    HIP_TRY(hipGetLastError());
    // the control changed: layers after `layer` are stale (as dmpm_set_control)
    c->fwd_valid = std::min(c->fwd_valid, (int)layer);
    c->loss_valid = false;
    if (c->cfg.tape_mode != DMPM_TAPE_STORE_ALL) drop_tapes_from(c, layer);
    return DMPM_OK;
}

int dmpm_get_grid(dmpm_ctx* c, int32_t step, double* m, double* v)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->cfg.tape_mode != DMPM_TAPE_STORE_ALL) return fail(DMPM_ERR_STATE, "grids are kept by the store-all tape only");
    if (step < 0 || step > c->L - 2 || step + 1 > c->fwd_valid) return fail(DMPM_ERR_STATE, "grid of this step not computed");
    return DISPATCH(c, get_grid, c, step, m, v);
}

int dmpm_chain_episode(dmpm_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return DISPATCH(c, chain_episode, c);
}

// FP32: the block tile (4 or 8 cells per edge); a context holding particles
// re-orders them (resort: the pass and the segment orders are dropped)
int set_tile(dmpm_ctx* c, int t)
{
    if (t == c->tile) return DMPM_OK;
    c->tile = t;
    set_block_prefix(c);
    if (!c->f32) return DMPM_OK;
    HIP_TRY(hipMemcpyAsync(c->d_boff, c->boff.data(), c->boff.size() * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return c->have_particles ? resort(c, true) : DMPM_OK;
}

int dmpm_set_option(dmpm_ctx* c, int32_t option, int64_t value)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    switch (option) {
    case DMPM_OPT_GRAPHS:
        if (value != 0 && value != 1) return fail(DMPM_ERR_INVALID, "DMPM_OPT_GRAPHS is 0 or 1");
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->use_graphs = value != 0;
        clear_graphs(c);
        return DMPM_OK;
    case DMPM_OPT_RANGE_MAX:
        if (value < 64 || value > 512) return fail(DMPM_ERR_INVALID, "DMPM_OPT_RANGE_MAX must be in [64, 512]");
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->range_max = (int)value;
        return c->f32 && !c->host_blocks.empty() ? build_ranges(c) : DMPM_OK;
    case DMPM_OPT_LDS_CAP:
        if (value < 512 || value > 4096) return fail(DMPM_ERR_INVALID, "DMPM_OPT_LDS_CAP must be in [512, 4096]");
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->lds_cap = (int)value;
        c->batch.cap = c->lds_cap;
        clear_graphs(c);
        return DMPM_OK;
    case DMPM_OPT_RESORT:
        if (value < 0 || value > 2) return fail(DMPM_ERR_INVALID, "DMPM_OPT_RESORT is 0, 1 or 2");
        c->resort_mode = (int)value;
        return DMPM_OK;
    case DMPM_OPT_RESORT_PERMILLE:
        if (value < 0 || value > 1000) return fail(DMPM_ERR_INVALID, "DMPM_OPT_RESORT_PERMILLE must be in [0, 1000]");
        c->resort_permille = (int)value;
        return DMPM_OK;
    case DMPM_OPT_OVERLAP:
        if (value != 0 && value != 1) return fail(DMPM_ERR_INVALID, "DMPM_OPT_OVERLAP is 0 or 1");
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->overlap = value != 0;
        clear_graphs(c);
        return DMPM_OK;
    case DMPM_OPT_PERSISTENT:
        if (value != 0 && value != 1) return fail(DMPM_ERR_INVALID, "DMPM_OPT_PERSISTENT is 0 or 1");
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->persistent = value != 0 && c->queue;
        c->batch.queue = c->persistent ? c->queue : nullptr;
        clear_graphs(c);
        return DMPM_OK;
    case DMPM_OPT_TILE:
        if (value != 4 && value != 8) return fail(DMPM_ERR_INVALID, "DMPM_OPT_TILE is 4 or 8");
        HIP_TRY(hipStreamSynchronize(c->stream));
        return set_tile(c, (int)value);
    case DMPM_OPT_SEG_RESORT:
        if (value != 0 && value != 1) return fail(DMPM_ERR_INVALID, "DMPM_OPT_SEG_RESORT is 0 or 1");
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->seg_resort = (int)value;
        if (!value) {
            bool any = false;
            for (int sg = 1; sg < (int)c->sord.size(); ++sg) any = any || seg_own(c, sg);
            if (any) {
                // layers held in segment orders: recomputed from layer 0
                drop_seg_orders(c);
                c->fwd_valid = std::min(c->fwd_valid, 0);
                c->loss_valid = false;
                c->grads_lo = -1;
                drop_tapes(c);
            }
        }
        return DMPM_OK;
    case DMPM_OPT_RANGE_ORDER:
        if (value != 0 && value != 1) return fail(DMPM_ERR_INVALID, "DMPM_OPT_RANGE_ORDER is 0 or 1");
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->range_order = (int)value;
        return c->f32 && !c->host_blocks.empty() ? build_ranges(c) : DMPM_OK;
    case DMPM_OPT_FUSE:
        if (value != 0 && value != 1) return fail(DMPM_ERR_INVALID, "DMPM_OPT_FUSE is 0 or 1");
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->fuse != (value != 0)) {
            const bool fc = fc_layers(c);
            c->fuse = value != 0;
            if (fc != fc_layers(c)) {
                // the tapes' layers hold F or Fc by the old setting: run the pass again
                c->fwd_valid = std::min(c->fwd_valid, 0);
                c->loss_valid = false;
                c->grads_lo = -1;
                drop_tapes(c);
            }
        }
        clear_graphs(c);
        return DMPM_OK;
    case DMPM_OPT_ELASTIC:
        if (value < 0 || value > 2) return fail(DMPM_ERR_INVALID, "DMPM_OPT_ELASTIC is 0, 1 or 2");
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->elastic = (int)value;
        clear_graphs(c);
        return DMPM_OK;
    default:
        return fail(DMPM_ERR_INVALID, "unknown option");
    }
}

int dmpm_get_option(dmpm_ctx* c, int32_t option, int64_t* value)
{
    if (!c || !value) return fail(DMPM_ERR_INVALID, "null argument");
    switch (option) {
    case DMPM_OPT_GRAPHS: *value = c->use_graphs ? 1 : 0; return DMPM_OK;
    case DMPM_OPT_RANGE_MAX: *value = c->range_max; return DMPM_OK;
    case DMPM_OPT_LDS_CAP: *value = c->lds_cap; return DMPM_OK;
    case DMPM_OPT_RESORT: *value = c->resort_mode; return DMPM_OK;
    case DMPM_OPT_RESORT_PERMILLE: *value = c->resort_permille; return DMPM_OK;
    case DMPM_OPT_OVERLAP: *value = c->overlap ? 1 : 0; return DMPM_OK;
    case DMPM_OPT_FUSE: *value = c->fuse ? 1 : 0; return DMPM_OK;
    case DMPM_OPT_RANGE_ORDER: *value = c->range_order; return DMPM_OK;
    case DMPM_OPT_SEG_RESORT: *value = c->seg_resort; return DMPM_OK;
    case DMPM_OPT_TILE: *value = c->tile; return DMPM_OK;
    case DMPM_OPT_PERSISTENT: *value = c->persistent ? 1 : 0; return DMPM_OK;
    case DMPM_OPT_ELASTIC: *value = c->elastic; return DMPM_OK;
    default: return fail(DMPM_ERR_INVALID, "unknown option");
    }
}

int dmpm_autotune(dmpm_ctx* c, double* best_ms)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->f32) return fail(DMPM_ERR_INVALID, "autotune tunes the FP32 block kernels");
    if (!c->have_particles) return fail(DMPM_ERR_STATE, "autotune before dmpm_set_particles");
    const bool prof = c->prof;
    c->prof = false;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    float* l0 = slot<float>(c, 0);
    float* l1 = slot<float>(c, 1);
    float* g = static_cast<float*>(c->grids);
    const int tiles[] = {4, 8};
    const int ranges[] = {128, 256, 512};
    const int caps[] = {512, 1024, 2048};
    double best = 1e30;
    int best_r = c->range_max, best_c = c->lds_cap, best_t = c->tile;
    bool best_p = c->persistent;
    for (int tile : tiles) {
      // the positions of layer 0 are those of the particles' order: re-ordering moves them with it
      if ((rc = set_tile(c, tile))) break;
      for (int pers = 0; pers < 2 && !rc; ++pers) {
      c->batch.queue = pers ? c->queue : nullptr;
      for (int cap : caps) {
        c->lds_cap = cap;
        c->batch.cap = cap;
        for (int r : ranges) {
            c->range_max = r;
            if ((rc = build_ranges(c))) break;
            float ms_min = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {  // first repetition warms up
                launch_fill<float>(grads_of<float>(c, 1), (long long)kLayerFields * c->S, 0.f, c->stream);
                hipEventRecord(a, c->stream);
                step_forward<float>(c, l0, l1, ctrl_of<float>(c, 0), g);
                step_backward<float>(c, 0, l0, l1, g);
                hipEventRecord(b, c->stream);
                if (hipEventSynchronize(b) != hipSuccess) {
                    rc = fail(DMPM_ERR_HIP, "autotune step failed");
                    break;
                }
                float ms = 0.f;
                hipEventElapsedTime(&ms, a, b);
                if (rep > 0) ms_min = std::min(ms_min, ms);
            }
            if (rc) break;
            if (ms_min < best) {
                best = ms_min;
                best_r = r;
                best_c = cap;
                best_t = tile;
                best_p = pers != 0;
            }
        }
        if (rc) break;
      }
      }
      if (rc) break;
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    c->prof = prof;
    c->lds_cap = best_c;
    c->batch.cap = best_c;
    c->range_max = best_r;
    c->persistent = best_p && c->queue;
    c->batch.queue = c->persistent ? c->queue : nullptr;
    int rc2 = rc ? DMPM_OK : set_tile(c, best_t);
    if (!rc2) rc2 = build_ranges(c);
    c->fwd_valid = 0;
    c->loss_valid = false;
    c->grads_lo = -1;
    c->bwd_control = -1;
    drop_tapes(c);
    if (rc) return rc;
    if (rc2) return rc2;
    HIP_TRY(hipGetLastError());
    if (best_ms) *best_ms = best;
    return DMPM_OK;
}

int dmpm_sort_particles(dmpm_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->f32) return DMPM_OK;
    if (!c->have_particles) return fail(DMPM_ERR_STATE, "sort before dmpm_set_particles");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return resort(c, true);
}

int dmpm_get_window_stats(dmpm_ctx* c, int64_t* stats)
{
    if (!c || !stats) return fail(DMPM_ERR_INVALID, "null argument");
    for (int k = 0; k < 3; ++k) stats[k] = c->last_wstats[k];
    stats[3] = c->n_sorts;
    return DMPM_OK;
}

int dmpm_get_segment_window_stats(dmpm_ctx* c, int32_t segment, int64_t* stats)
{
    if (!c || !stats) return fail(DMPM_ERR_INVALID, "null argument");
    if (segment < 0 || segment >= c->nstat) return fail(DMPM_ERR_INVALID, "segment out of range");
    for (int k = 0; k < 3; ++k) stats[k] = c->last_swstats[3 * (size_t)segment + k];
    stats[3] = seg_own(c, segment) ? 1 : 0;
    stats[4] = c->n_seg_sorts;
    return DMPM_OK;
}

int dmpm_synchronize(dmpm_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return DMPM_OK;
}

int dmpm_memory_bytes(dmpm_ctx* c, uint64_t* bytes)
{
    if (!c || !bytes) return fail(DMPM_ERR_INVALID, "null argument");
    *bytes = c->bytes;
    return DMPM_OK;
}

int dmpm_profile_enable(dmpm_ctx* c, int32_t on)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!on && c->prof) {
        rc = drain_profile(c);
        if (rc) return rc;
    }
    c->prof = on != 0;
    return DMPM_OK;
}

int dmpm_profile_reset(dmpm_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    rc = drain_profile(c);
    if (rc) return rc;
    for (int k = 0; k < DMPM_K_COUNT; ++k) {
        c->prof_ms[k] = 0.0;
        c->prof_n[k] = 0;
    }
    return DMPM_OK;
}

int dmpm_profile_read(dmpm_ctx* c, int32_t kernel, double* total_ms, int64_t* launches)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (kernel < 0 || kernel >= DMPM_K_COUNT) return fail(DMPM_ERR_INVALID, "kernel id");
    rc = drain_profile(c);
    if (rc) return rc;
    if (total_ms) *total_ms = c->prof_ms[kernel];
    if (launches) *launches = c->prof_n[kernel];
    return DMPM_OK;
}

}  // extern "C"

// ------------------------------------------------------ batched evaluators --
namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) hipFree(p);
    }
};

using HostIn = std::pair<const double*, size_t>;   // host array, element count
using HostOut = std::pair<double*, size_t>;

// Run launch(ptrs) on device 0 over device copies of host arrays: `ins` are
// copied to the device, `inouts` to the device and back, `outs` back only.
// ptrs lists the device copies in that order.
int dev_call(const std::vector<HostIn>& ins, const std::vector<HostOut>& inouts, const std::vector<HostOut>& outs,
             const std::function<void(std::vector<double*>&)>& launch)
{
    HIP_TRY(hipSetDevice(0));
    std::vector<DevBuf> bufs(ins.size() + inouts.size() + outs.size());
    std::vector<double*> ptrs;
    size_t k = 0;
    auto alloc = [&](size_t count, const char* what) -> int {
        if (const int rc = dmalloc(&bufs[k].p, count * sizeof(double), what)) return rc;
        ptrs.push_back(static_cast<double*>(bufs[k].p));
        ++k;
        return DMPM_OK;
    };
    for (const HostIn& in : ins) {
        if (!in.first && in.second) return fail(DMPM_ERR_INVALID, "null input");
        if (const int rc = alloc(in.second, "input")) return rc;
        if (in.second) HIP_TRY(hipMemcpy(ptrs.back(), in.first, in.second * sizeof(double), hipMemcpyHostToDevice));
    }
    for (const HostOut& io : inouts) {
        if (!io.first && io.second) return fail(DMPM_ERR_INVALID, "null input/output");
        if (const int rc = alloc(io.second, "input/output")) return rc;
        if (io.second) HIP_TRY(hipMemcpy(ptrs.back(), io.first, io.second * sizeof(double), hipMemcpyHostToDevice));
    }
    for (const HostOut& o : outs) {
        if (!o.first && o.second) return fail(DMPM_ERR_INVALID, "null output");
        if (const int rc = alloc(o.second, "output")) return rc;
    }
    launch(ptrs);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    k = ins.size();
    for (const HostOut& io : inouts) {
        if (io.second) HIP_TRY(hipMemcpy(io.first, ptrs[k], io.second * sizeof(double), hipMemcpyDeviceToHost));
        ++k;
    }
    for (const HostOut& o : outs) {
        if (o.second) HIP_TRY(hipMemcpy(o.first, ptrs[k], o.second * sizeof(double), hipMemcpyDeviceToHost));
        ++k;
    }
    return DMPM_OK;
}

// n matrices: `per` doubles of each input and output per matrix
int eval_common(int n, std::initializer_list<std::pair<const double*, size_t>> ins,
                std::initializer_list<std::pair<double*, size_t>> outs,
                const std::function<void(int, std::vector<double*>&)>& launch)
{
    if (n < 0) return fail(DMPM_ERR_INVALID, "n < 0");
    if (n == 0) return DMPM_OK;
    std::vector<HostIn> vi;
    std::vector<HostOut> vo;
    for (auto& in : ins) {
        if (!in.first) return fail(DMPM_ERR_INVALID, "null input");
        vi.push_back({in.first, in.second * n});
    }
    for (auto& o : outs) {
        if (!o.first) return fail(DMPM_ERR_INVALID, "null output");
        vo.push_back({o.first, o.second * n});
    }
    return dev_call(vi, {}, vo, [&](std::vector<double*>& p) { launch(n, p); });
}

// A phase grid (dmpm_grid_desc) as step parameters: dims, node count, spacing, origin.
int phase_grid(const dmpm_grid_desc* g, StepParams& sp)
{
    if (!g) return fail(DMPM_ERR_INVALID, "null grid");
    sp = StepParams{};
    long long G = 1;
    for (int a = 0; a < 3; ++a) {
        if (g->dim[a] <= 0) return fail(DMPM_ERR_INVALID, "grid dims must be > 0");
        G *= g->dim[a];
        sp.dim[a] = g->dim[a];
        sp.min_point[a] = g->min_point[a];
    }
    if (!(g->dx > 0.0)) return fail(DMPM_ERR_INVALID, "dx must be > 0");
    if (G > (1LL << 34)) return fail(DMPM_ERR_INVALID, "grid too large");
    sp.G = G;
    sp.dx = g->dx;
    sp.inv_dx = 1.0 / g->dx;
    return DMPM_OK;
}

}  // namespace

extern "C" {

int dmpm_eval_pk1(int32_t n, const double* F, const double* lam, const double* mu, double* P)
{
    return eval_common(n, {{F, 9}, {lam, 1}, {mu, 1}}, {{P, 9}}, [](int n_, std::vector<double*>& p) {
        launch_eval_pk1(n_, p[0], p[1], p[2], p[3], nullptr);
    });
}

int dmpm_eval_hessian(int32_t n, const double* F, const double* lam, const double* mu, const double* dF, double* out,
                      int32_t exact)
{
    return eval_common(n, {{F, 9}, {lam, 1}, {mu, 1}, {dF, 9}}, {{out, 9}}, [exact](int n_, std::vector<double*>& p) {
        launch_eval_hess(n_, p[0], p[1], p[2], p[3], p[4], exact, nullptr);
    });
}

int dmpm_eval_energy(int32_t n, const double* F, const double* lam, const double* mu, double* psi)
{
    return eval_common(n, {{F, 9}, {lam, 1}, {mu, 1}}, {{psi, 1}}, [](int n_, std::vector<double*>& p) {
        launch_eval_energy(n_, p[0], p[1], p[2], p[3], nullptr);
    });
}

int dmpm_eval_svd(int32_t n, const double* F, double* U, double* sigma, double* V)
{
    return eval_common(n, {{F, 9}}, {{U, 9}, {sigma, 3}, {V, 9}}, [](int n_, std::vector<double*>& p) {
        launch_eval_svd(n_, p[0], p[1], p[2], p[3], nullptr);
    });
}

// ------------------------------------------------------------- phases ------
int dmpm_phase_p2g(const dmpm_grid_desc* g, int32_t n, const double* x, const double* v, const double* C, const double* F,
                   const double* dFc, const double* P, const double* m, const double* vol, double dt, double drag, double* grid_m,
                   double* grid_p)
{
    StepParams sp;
    if (const int rc = phase_grid(g, sp)) return rc;
    if (n < 0) return fail(DMPM_ERR_INVALID, "n < 0");
    if (!grid_m || !grid_p) return fail(DMPM_ERR_INVALID, "null grid arrays");
    if (n == 0) return DMPM_OK;
    sp.n = n;
    sp.dt = dt;
    sp.drag = drag;
    const size_t N = (size_t)n, G = (size_t)sp.G;
    return dev_call({{x, 3 * N}, {v, 3 * N}, {C, 9 * N}, {F, 9 * N}, {dFc, 9 * N}, {P, 9 * N}, {m, N}, {vol, N}},
                    {{grid_m, G}, {grid_p, 3 * G}}, {}, [&](std::vector<double*>& p) {
                        launch_phase_p2g(sp, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], nullptr);
                    });
}

static int phase_grid_op(const dmpm_grid_desc* g, const double* grid_m, const double* grid_p, double dt, const double* f3,
                         int point_mode, double mag, double* grid_v)
{
    StepParams sp;
    if (const int rc = phase_grid(g, sp)) return rc;
    if (!f3 || !grid_m || !grid_p || !grid_v) return fail(DMPM_ERR_INVALID, "null argument");
    sp.dt = dt;
    for (int a = 0; a < 3; ++a) sp.f_ext[a] = f3[a];
    const size_t G = (size_t)sp.G;
    return dev_call({{grid_m, G}, {grid_p, 3 * G}}, {{grid_v, 3 * G}}, {}, [&](std::vector<double*>& p) {
        launch_phase_grid_op(sp, p[0], p[1], p[2], point_mode, mag, nullptr);
    });
}

int dmpm_phase_grid_op(const dmpm_grid_desc* g, const double* grid_m, const double* grid_p, double dt, const double* f_ext,
                       double* grid_v)
{
    return phase_grid_op(g, grid_m, grid_p, dt, f_ext, 0, 0.0, grid_v);
}

int dmpm_phase_grid_op_point(const dmpm_grid_desc* g, const double* grid_m, const double* grid_p, double dt,
                             const double* gravity_point, double gravity_mag, double* grid_v)
{
    return phase_grid_op(g, grid_m, grid_p, dt, gravity_point, 1, gravity_mag, grid_v);
}

int dmpm_phase_g2p(const dmpm_grid_desc* g, int32_t n, const double* x, const double* grid_v, double* v_next, double* C_next)
{
    StepParams sp;
    if (const int rc = phase_grid(g, sp)) return rc;
    if (n < 0) return fail(DMPM_ERR_INVALID, "n < 0");
    if (!grid_v || !v_next || !C_next) return fail(DMPM_ERR_INVALID, "null argument");
    if (n == 0) return DMPM_OK;
    sp.n = n;
    const size_t N = (size_t)n;
    return dev_call({{x, 3 * N}, {grid_v, 3 * (size_t)sp.G}}, {}, {{v_next, 3 * N}, {C_next, 9 * N}},
                    [&](std::vector<double*>& p) { launch_phase_g2p(sp, p[0], p[1], p[2], p[3], nullptr); });
}

int dmpm_phase_p_op2(int32_t n, const double* x, const double* F, const double* dFc, const double* v_next, const double* C_next,
                     double smoothing, double dt, double* x_next, double* F_next)
{
    if (n < 0) return fail(DMPM_ERR_INVALID, "n < 0");
    if (n == 0) return DMPM_OK;
    if (!x_next || !F_next) return fail(DMPM_ERR_INVALID, "null output");
    const size_t N = (size_t)n;
    return dev_call({{x, 3 * N}, {F, 9 * N}, {dFc, 9 * N}, {v_next, 3 * N}, {C_next, 9 * N}}, {}, {{x_next, 3 * N}, {F_next, 9 * N}},
                    [&](std::vector<double*>& p) {
                        launch_phase_p_op2(n, p[0], p[1], p[2], p[3], p[4], smoothing, dt, p[5], p[6], nullptr);
                    });
}

int dmpm_phase_p2g_mass(const dmpm_grid_desc* g, int32_t n, const double* x, const double* m, double* grid_m)
{
    StepParams sp;
    if (const int rc = phase_grid(g, sp)) return rc;
    if (n < 0) return fail(DMPM_ERR_INVALID, "n < 0");
    if (!grid_m) return fail(DMPM_ERR_INVALID, "null grid array");
    if (n == 0) return DMPM_OK;
    sp.n = n;
    const size_t N = (size_t)n;
    const double org[3] = {0.0, 0.0, 0.0};
    return dev_call({{x, 3 * N}, {m, N}}, {{grid_m, (size_t)sp.G}}, {},
                    [&](std::vector<double*>& p) { launch_p2g_mass_aos<double>(sp, p[0], p[1], org, p[2], nullptr); });
}

int dmpm_phase_g2g_mass(const dmpm_grid_desc* g1, const double* grid1_m, const dmpm_grid_desc* g2, double* grid2_m)
{
    StepParams s1, s2;
    if (const int rc = phase_grid(g1, s1)) return rc;
    if (const int rc = phase_grid(g2, s2)) return rc;
    if (!grid2_m) return fail(DMPM_ERR_INVALID, "null output");
    return dev_call({{grid1_m, (size_t)s1.G}}, {}, {{grid2_m, (size_t)s2.G}},
                    [&](std::vector<double*>& p) { launch_phase_g2g_mass(s1, s2, p[0], p[1], nullptr); });
}

}  // extern "C"
