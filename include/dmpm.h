/*
 * dmpm.h -- C ABI of the MI355X-native differentiable MLS-MPM step.
 *
 * The library replaces the hot path of DiffMPMLib3D (paths relative to the
 * reference root): the forward step, the end-layer mass loss and the
 * adjoint step, driven over one episode of layers.  Each entry point names
 * the reference interface it stands in for.  Host arrays are row-major AoS
 * doubles: x, v are [n][3]; F, C, dFc are [n][3][3] with (i,j) at 3*i+j;
 * grids are flat [dim_x*dim_y*dim_z] with index i*dy*dz + j*dz + k
 * (Grid.h:22-31) and node position min_point + (i,j,k)*dx (Grid.cpp:22).
 *
 * Every call returns DMPM_OK (0) or a negative error code; dmpm_last_error()
 * returns the message of the calling thread's last failure.  A context owns
 * one HIP stream; calls on one context are serialised on it, distinct
 * contexts are independent.  Calls that return host data synchronise.
 */
#ifndef DMPM_H
#define DMPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMPM_OK 0
#define DMPM_ERR_INVALID (-1) /* bad argument or configuration */
#define DMPM_ERR_HIP (-2)     /* HIP runtime failure (message has the HIP error) */
#define DMPM_ERR_STATE (-3)   /* call order: e.g. backward before the loss */
#define DMPM_ERR_NOMEM (-4)   /* device allocation failed */

/* precision of particle and grid state */
#define DMPM_PREC_F64 0 /* plain FP64 path: the drop-in */
#define DMPM_PREC_F32 1 /* FP32 state, FP64 polar/SVD, PK1, H:dF and reductions */

/* tape: how the forward trajectory is kept for the adjoint */
#define DMPM_TAPE_STORE_ALL 0  /* every layer and grid (the reference's CompGraph, CompGraph.cpp:7-17) */
#define DMPM_TAPE_CHECKPOINT 1 /* checkpoints every ckpt_stride layers + one recomputed segment */

/* reference-compat switches for the adjoint (BackPropagation.cpp), 0 = exact adjoint */
#define DMPM_COMPAT_DUP_FBAR 1     /* (I+dtC')^T dLdF' added twice, :20 and :121 */
#define DMPM_COMPAT_NO_1MS 2       /* dLdC_next without (1-s), :24 */
#define DMPM_COMPAT_NO_DRAG_X 4    /* x-gradient uses m v, not m(1-dt drag) v, :109 */
#define DMPM_COMPAT_DLDM_NO_FEXT 8 /* dLdm uses v, not v - dt f_ext, :57 */
#define DMPM_COMPAT_REFERENCE 15

/* kernel ids for the profiling counters */
#define DMPM_K_P2G 0
#define DMPM_K_GRID_UPDATE 1
#define DMPM_K_G2P 2
#define DMPM_K_G2P_ADJ 3
#define DMPM_K_GRID_ADJ 4
#define DMPM_K_P2G_ADJ 5
#define DMPM_K_LOSS 6
#define DMPM_K_FILL 7
#define DMPM_K_STEP 8 /* FP32: G2P of a step with P2G of the next in one launch (DMPM_OPT_FUSE) */
#define DMPM_K_COUNT 9

typedef struct dmpm_ctx dmpm_ctx;

typedef struct dmpm_config {
    int32_t precision;          /* DMPM_PREC_* */
    int32_t n_particles;
    int32_t dim[3];             /* grid nodes per axis (Grid::dim_x/y/z) */
    int32_t num_layers;         /* layers per episode, >= 2 (CompGraph::SetUpCompGraph num_layers) */
    double dx;                  /* Grid::dx */
    double min_point[3];        /* Grid::min_point */
    double dt, drag, smoothing; /* CompGraph.h:55-58 */
    double f_ext[3];
    int32_t compat;             /* DMPM_COMPAT_* bits */
    int32_t tape_mode;          /* DMPM_TAPE_* */
    int32_t ckpt_stride;        /* checkpoint mode: layers between checkpoints, 0 = the divisor of L-1 nearest
                                 * sqrt(L-1) within a factor 1.5, else round(sqrt(L-1)) */
    int32_t keep_control_grads; /* 1: keep dL/d(dFc_i) of every layer of the backward pass */
    int32_t device;             /* HIP device ordinal */
    int32_t reserved;
} dmpm_config;

/* Library version (major*10000 + minor*100 + patch). */
int dmpm_version(void);
const char* dmpm_last_error(void);

/* Defaults of CompGraph.h:55-58 (dt = 1/120, drag = 0.5, s = 0.1, f_ext = 0), FP64, store-all. */
int dmpm_config_init(dmpm_config* cfg);

/* Lifetime: replaces CompGraph::CompGraph (CompGraph.cpp:269-278) + SetUpCompGraph (:7-17). */
int dmpm_create(const dmpm_config* cfg, dmpm_ctx** out);
int dmpm_destroy(dmpm_ctx* ctx);

/* One morph of a batched context: its particle count and grid (Grid.h:22-31, Grid.cpp:22). */
typedef struct dmpm_instance {
    int32_t n_particles;
    int32_t dim[3];
    double dx;
    double min_point[3];
} dmpm_instance;

/* Batched context: n_instances independent morphs advanced together, every launch covering all
 * of them (the reference runs one CompGraph per morph, DiffMPMAnimator3D.cpp:421-471).  FP32 only.
 * cfg gives precision, num_layers, dt, drag, smoothing, f_ext, compat, tape and device; its
 * n_particles, dim, dx and min_point are ignored.  Every particle array of the other calls holds
 * the instances' particles concatenated in instance order; grid arrays (target, dmpm_get_grid)
 * hold the instances' grids concatenated.  dmpm_end_layer_mass_loss returns the sum of the
 * instance losses (instances are independent, so each instance's gradients are its own). */
int dmpm_create_batch(const dmpm_config* cfg, int32_t n_instances, const dmpm_instance* instances, dmpm_ctx** out);
int dmpm_num_instances(dmpm_ctx* ctx, int32_t* n);
/* EndLayerMassLoss of every instance from the last dmpm_end_layer_mass_loss. */
int dmpm_get_instance_losses(dmpm_ctx* ctx, double* losses);
/* dmpm_target_from_points per instance: point clouds concatenated in instance order, counts[i] each. */
int dmpm_target_from_points_batch(dmpm_ctx* ctx, const double* x, const double* m, const int32_t* counts);

/* Layer 0 particles: MaterialPoint x, v, F, C, m, vol, lam, mu (MaterialPoint.h:18-28).
 * vol may be NULL (then call dmpm_compute_volumes).  Invalidates every later layer. */
int dmpm_set_particles(dmpm_ctx* ctx, const double* x, const double* v, const double* F, const double* C,
                       const double* m, const double* vol, const double* lam, const double* mu);
/* SingleThreadMPM::CalculatePointCloudVolumes (ForwardSimulation.cpp:262-290) on layer 0. */
int dmpm_compute_volumes(dmpm_ctx* ctx);
int dmpm_get_volumes(dmpm_ctx* ctx, double* vol);

/* Control dFc of one layer (MaterialPoint::dFc, MaterialPoint.h:31); layer -1 sets all layers,
 * as SetUpCompGraph's copy of layer 0 does.  Invalidates the layers after it. */
int dmpm_set_control(dmpm_ctx* ctx, int32_t layer, const double* dFc);
int dmpm_get_control(dmpm_ctx* ctx, int32_t layer, double* dFc);

/* Target grid of EndLayerMassLoss (CompGraph::target_grid). */
int dmpm_set_target_grid(dmpm_ctx* ctx, const double* mass);
/* Target grid as the P2G mass of a point cloud (MPMPolyscope.cpp:156-160). */
int dmpm_target_from_points(dmpm_ctx* ctx, const double* x, const double* m, int32_t n);
int dmpm_get_target_grid(dmpm_ctx* ctx, double* mass);
/* dmpm_target_from_points without any allocation: no reference counterpart; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  The points
 * are staged in a fixed device array of the library (32 MiB per GPU, held until the process
 * ends) instead of a buffer allocated per call, so n is at most 1048576 (DMPM_ERR_INVALID
 * beyond).  Every call on one GPU shares that array: calls that overlap on one GPU (contexts
 * of one device used from several threads) race.  Single-morph contexts only. */
#define DMPM_NOALLOC_MAX_POINTS 1048576
int dmpm_target_from_points_noalloc(dmpm_ctx* ctx, const double* x, const double* m, int32_t n);

/* CompGraph::ComputeForwardPass(start_layer) (CompGraph.cpp:246-256). */
int dmpm_forward_pass(dmpm_ctx* ctx, int32_t start_layer);
/* CompGraph::EndLayerMassLoss (CompGraph.cpp:110-244): loss and the adjoint seed. */
int dmpm_end_layer_mass_loss(dmpm_ctx* ctx, double* loss);
/* CompGraph::ComputeBackwardPass(control_layer) (CompGraph.cpp:258-267). */
int dmpm_backward_pass(dmpm_ctx* ctx, int32_t control_layer);

/* One SingleThreadMPM::ForwardTimeStep layer -> layer+1 (ForwardSimulation.h:18) and one
 * Back_Timestep layer+1 -> layer (BackPropagation.h:7); store-all tape only. */
int dmpm_forward_step(dmpm_ctx* ctx, int32_t layer);
int dmpm_backward_step(dmpm_ctx* ctx, int32_t layer);

/* State of a layer (any pointer may be NULL); checkpoint mode recomputes it when needed. */
int dmpm_get_layer_state(dmpm_ctx* ctx, int32_t layer, double* x, double* v, double* F, double* C);
/* Per-particle diagnostics of a layer (either pointer may be NULL), computed on the GPU in FP64
 * from the layer's stored F and control dFc (dFc = 0 for the last layer, which has no control):
 * J = det(F + dFc) (PointCloud::GetPointDeterminants, PointCloud.cpp:441-452) and
 * psi = FixedCorotatedElasticity(F + dFc, lam, mu) (PointCloud::GetPointElasticEnergies,
 * PointCloud.cpp:455-467). */
int dmpm_get_point_diagnostics(dmpm_ctx* ctx, int32_t layer, double* J, double* psi);
/* Store-all tape: replace a computed layer's state (the C++ host mirror's write-back). */
int dmpm_set_layer_state(dmpm_ctx* ctx, int32_t layer, const double* x, const double* v, const double* F,
                         const double* C);
/* Gradients dLdx, dLdv, dLdF, dLdC (MaterialPoint.h:34-37) of the lowest layer reached by the last
 * backward pass or of the layer above it (the last layer after the loss). */
int dmpm_get_layer_grads(dmpm_ctx* ctx, int32_t layer, double* dLdx, double* dLdv, double* dLdF, double* dLdC);
/* Seed the gradients of a layer (MaterialPoint dLdx, dLdv, dLdF, dLdC; NULL = zero): on the last
 * layer it replaces EndLayerMassLoss's seed (a caller-defined loss) for dmpm_backward_pass; on any
 * layer it is the input of dmpm_backward_step(layer-1) (Back_Timestep reads layer n+1's gradients). */
int dmpm_set_layer_grads(dmpm_ctx* ctx, int32_t layer, const double* dLdx, const double* dLdv, const double* dLdF,
                         const double* dLdC);
/* dL/d(dFc_layer): the control layer of the last backward pass, or any of its layers with
 * keep_control_grads.  Exact mode: dL/d(dFc); compat mode: the reference's dLdF. */
int dmpm_get_control_grad(dmpm_ctx* ctx, int32_t layer, double* out);
/* PointCloud::Compute_dLdF_Norm (PointCloud.cpp:24-35) of a control gradient, FP64 reduction on the device. */
int dmpm_control_grad_norm(dmpm_ctx* ctx, int32_t layer, double* norm);
/* PointCloud::Descend_Adam / Descend_AMSGrad / Descend_Lion / Descend_GradientDescent
 * (PointCloud.cpp:83-179) on the device: dFc of `layer` updated from its control gradient of the
 * last backward pass (dmpm_get_control_grad) divided by gradient_norm, with the per-particle
 * momentum / vector / vector_max of that layer (zero at first use, MaterialPoint.h).  Float
 * arithmetic as the reference; timestep t enters as 1 - beta^t.  Invalidates later layers. */
#define DMPM_DESCEND_ADAM 0
#define DMPM_DESCEND_AMSGRAD 1
#define DMPM_DESCEND_LION 2 /* beta1 is Lion's beta */
#define DMPM_DESCEND_GD 3
int dmpm_descend(dmpm_ctx* ctx, int32_t layer, int32_t rule, float alpha, float gradient_norm, float beta1, float beta2,
                 float epsilon, int32_t timestep);
/* Grid of one step after G_op (node m and v), store-all tape. */
int dmpm_get_grid(dmpm_ctx* ctx, int32_t step, double* m, double* v);

/* Episode chaining of DiffMPMAnimator3D.cpp:424-429: layer 0 <- last layer, controls <- 0. */
int dmpm_chain_episode(dmpm_ctx* ctx);

/* Execution options (no reference counterpart: launch configuration of this library). */
#define DMPM_OPT_GRAPHS 1    /* 1: capture each pass into a HIP graph on first use, then replay it (default 0) */
#define DMPM_OPT_RANGE_MAX 2 /* FP32: most particles per workgroup range, 64..512 (default 512) */
#define DMPM_OPT_LDS_CAP 3   /* FP32: node capacity of the general LDS window, 512..4096 (default 1024) */
#define DMPM_OPT_RESORT 4    /* FP32: when the device re-orders the particles by grid block (dmpm_sort_particles):
                              * 0: at dmpm_set_particles only; 1: also at every dmpm_chain_episode and at
                              * dmpm_forward_pass(0) when the order is stale (DMPM_OPT_RESORT_PERMILLE); 2 (default):
                              * at dmpm_chain_episode and dmpm_forward_pass(0) when the order is stale */
#define DMPM_OPT_RESORT_PERMILLE 5 /* FP32: stale = in the last forward pass that reached the loss, more than this
                                    * many per mille of the P2G workgroups left the fixed 8^3 window (default 50) */
#define DMPM_OPT_OVERLAP 6   /* checkpoint tape: 1 (default) recomputes the next segment down on a second stream
                              * while the backward pass steps through the current one (two tapes); 0: one stream */
#define DMPM_OPT_FUSE 7      /* FP32: 1 (default) runs the G2P of each step and the P2G of the next as one launch
                              * inside a pass; 0: separate launches (the same results) */
#define DMPM_OPT_RANGE_ORDER 8 /* FP32: launch order of the block ranges: 1 (default) longest first, 0 block order */
#define DMPM_OPT_SEG_RESORT 9  /* FP32 checkpoint tape: 1 (default) orders a segment's particles afresh from its first
                                * layer when its P2G windows went stale (DMPM_OPT_RESORT_PERMILLE) in the last forward
                                * pass that reached the loss, and keeps that order; 0: every segment in layer 0's order */
#define DMPM_OPT_TILE 10       /* FP32: block tile, the cells per edge of the blocks the particles are ordered by and
                                * each workgroup's range is drawn from: 4 (default; fixed 8^3-node LDS window) or 8
                                * (ranges span 9 cells: the general LDS window up to DMPM_OPT_LDS_CAP nodes) */
#define DMPM_OPT_PERSISTENT 11 /* FP32: 1 runs the particle kernels as persistent launches (as many workgroups as fit
                                * on the chip, taking every instance's block ranges from a counter); 0 (default): one
                                * workgroup per range */
#define DMPM_OPT_ELASTIC 12    /* FP32: where the elastic half of the transposed P2G (FP64 polar, PK1 and H:dF) runs:
                                * 0 in a kernel of its own after the gather (one thread per particle), 1 in the
                                * gather kernel, 2 there with registers for 3 waves per SIMD; particles that need
                                * the SVD (inverted or nearly flat elements) take it inside mode 0's kernel, and in
                                * a small kernel after the step in modes 1 and 2 and with persistent launches (the
                                * same results) */
int dmpm_set_option(dmpm_ctx* ctx, int32_t option, int64_t value);
int dmpm_get_option(dmpm_ctx* ctx, int32_t option, int64_t* value);
/* FP32: time a forward and a backward step of layer 0 for every candidate block tile (4, 8), launch form
 * (a workgroup per range or persistent), range size (128, 256, 512 particles: one or two per thread) and
 * LDS capacity (512, 1024, 2048 nodes), keep
 * the fastest (best_ms: its time).  Needs particles; the particles are re-ordered for each tile, the
 * forward trajectory (layers after 0) and gradients are invalidated. */
int dmpm_autotune(dmpm_ctx* ctx, double* best_ms);

/* FP32: re-order the particles on the device by the DMPM_OPT_TILE block of their layer-0 cell (then rank in
 * cell, then cell in block), as dmpm_set_particles does; every per-particle array follows (controls,
 * optimiser states, held control gradients), the API order of every other call is unchanged, and the
 * layers after 0 are invalidated.  A no-op on the FP64 path, which keeps the API order.  No reference
 * counterpart: the reference's particles have no device order (DiffMPMAnimator3D.cpp:424-429 chains
 * episodes on the host). */
int dmpm_sort_particles(dmpm_ctx* ctx);
/* FP32: stats[0..2] = P2G workgroups on the fixed window, the general LDS window and direct HBM atomics
 * over the last forward pass that reached dmpm_end_layer_mass_loss; stats[3] = device sorts so far. */
int dmpm_get_window_stats(dmpm_ctx* ctx, int64_t* stats);
/* FP32 checkpoint tape (no reference counterpart): stats[0..2] as dmpm_get_window_stats for the steps of one
 * checkpoint segment (segments 0 .. ceil((num_layers - 1) / stride) - 1; other contexts have segment 0 only);
 * stats[3] = 1 when the segment holds its particles in an order of its own (DMPM_OPT_SEG_RESORT);
 * stats[4] = segment sorts so far. */
int dmpm_get_segment_window_stats(dmpm_ctx* ctx, int32_t segment, int64_t* stats);

int dmpm_synchronize(dmpm_ctx* ctx);
/* Device bytes held by the context. */
int dmpm_memory_bytes(dmpm_ctx* ctx, uint64_t* bytes);
/* HIP events around every launch on the context's stream; per-kernel totals in milliseconds. */
int dmpm_profile_enable(dmpm_ctx* ctx, int32_t on);
int dmpm_profile_reset(dmpm_ctx* ctx);
int dmpm_profile_read(dmpm_ctx* ctx, int32_t kernel, double* total_ms, int64_t* launches);

/* Batched constitutive model on the GPU (device 0, host arrays, n matrices):
 * PK_FixedCorotatedElasticity (Elasticity.cpp:88-98),
 * d2_FCE_psi_dF2_mult_by_dF (Elasticity.cpp:100-178; exact = 1 differentiates the J clamp),
 * and the 3x3 SVD of qrsvd/ImplicitQRSVD.h:687-850 (U, V rotations, sign on sigma[2]). */
int dmpm_eval_pk1(int32_t n, const double* F, const double* lam, const double* mu, double* P);
int dmpm_eval_hessian(int32_t n, const double* F, const double* lam, const double* mu, const double* dF,
                      double* out, int32_t exact);
int dmpm_eval_svd(int32_t n, const double* F, double* U, double* sigma, double* V);
/* FixedCorotatedElasticity (Elasticity.cpp:74-86) for n matrices: psi = mu sum (sigma_i - 1)^2
 * + lam/2 (J - 1)^2, J = max(det F, 1e-6), FP64. */
int dmpm_eval_energy(int32_t n, const double* F, const double* lam, const double* mu, double* psi);

/* ------------------------------------------------------------------------------------------
 * The phases of one ForwardTimeStep as separate calls (SingleThreadMPM, ForwardSimulation.h:23-45),
 * for callers that run them one by one on their own containers: FP64 on device 0, host arrays in
 * the reference's per-particle layout (x, v [n][3]; C, F, dFc, P [n][3][3]) and grids as masses [G]
 * and momenta / velocities [G][3] (node order and positions as above).  The episode calls above
 * fuse these phases and keep the data on the device; these copy their arrays in and out. */
typedef struct dmpm_grid_desc {
    int32_t dim[3];      /* Grid::dim_x, dim_y, dim_z */
    int32_t reserved;
    double dx;           /* Grid::dx */
    double min_point[3]; /* Grid::min_point */
} dmpm_grid_desc;
/* SingleThreadMPM::P2G (ForwardSimulation.cpp:11-51,174-182): adds each particle's w m to grid_m and
 * w [m (1 - dt drag) v + (-(3/dx^2) dt vol P (F + dFc)^T + m C) d] to grid_p (accumulated: G_Reset
 * clears them); P is the particle's stored stress (P_op_1). */
int dmpm_phase_p2g(const dmpm_grid_desc* g, int32_t n, const double* x, const double* v, const double* C, const double* F,
                   const double* dFc, const double* P, const double* m, const double* vol, double dt, double drag, double* grid_m,
                   double* grid_p);
/* SingleThreadMPM::G_op (ForwardSimulation.cpp:53-58,184-200): grid_v = p/m + dt f_ext where
 * |m| > FLT_EPSILON; the other nodes keep grid_v. */
int dmpm_phase_grid_op(const dmpm_grid_desc* g, const double* grid_m, const double* grid_p, double dt, const double* f_ext,
                       double* grid_v);
/* SingleThreadMPM::G_op with a gravity point (ForwardSimulation.cpp:202-224): the force on a node is
 * gravity_mag towards gravity_point (zero at the point). */
int dmpm_phase_grid_op_point(const dmpm_grid_desc* g, const double* grid_m, const double* grid_p, double dt,
                             const double* gravity_point, double gravity_mag, double* grid_v);
/* SingleThreadMPM::G2P (ForwardSimulation.cpp:60-106,226-236): v' = sum w v_i, C' = (3/dx^2) sum w v_i d^T. */
int dmpm_phase_g2p(const dmpm_grid_desc* g, int32_t n, const double* x, const double* grid_v, double* v_next, double* C_next);
/* SingleThreadMPM::P_op_2 (ForwardSimulation.cpp:108-124,238-259): F' = (1 - s)(I + dt C')(F + dFc) + s F,
 * x' = x + dt v' (s = 0: the overload without smoothing). */
int dmpm_phase_p_op2(int32_t n, const double* x, const double* F, const double* dFc, const double* v_next, const double* C_next,
                     double smoothing, double dt, double* x_next, double* F_next);
/* SingleThreadMPM::P2G_Mass (ForwardSimulation.cpp:292-327): adds each point's w m to grid_m. */
int dmpm_phase_p2g_mass(const dmpm_grid_desc* g, int32_t n, const double* x, const double* m, double* grid_m);
/* SingleThreadMPM::G2G_Mass (ForwardSimulation.cpp:329-367): every node of grid 2 gets the sum w m_i of
 * the grid-1 node masses around its position. */
int dmpm_phase_g2g_mass(const dmpm_grid_desc* g1, const double* grid1_m, const dmpm_grid_desc* g2, double* grid2_m);

#ifdef __cplusplus
}
#endif
#endif /* DMPM_H */
