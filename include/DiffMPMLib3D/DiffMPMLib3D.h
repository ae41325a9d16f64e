// DiffMPMLib3D.h -- C++ host mirror of the reference's MLS-MPM API, backed by
// the MI355X C ABI (include/dmpm.h).
//
// Names, signatures and data members follow the reference headers (paths
// relative to the reference root): MaterialPoint.h, PointCloud.h, GridNode.h,
// Grid.h, CompGraph.h, ForwardSimulation.h, BackPropagation.h, Elasticity.h.
// The containers hold host data exactly as the reference does (per-layer
// point clouds and grids, float fields); every hot call (forward step,
// passes, loss, adjoint, constitutive model) runs on the GPU through the C
// ABI and writes its results back into these containers.  Errors raise
// DiffMPMLib3D::DeviceError (the reference has no error channel; a silent
// CPU fallback is not offered).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct dmpm_ctx;

namespace DiffMPMLib3D {

// ---------------------------------------------------------------- types ----
// Vec3 / Mat3: float, as the reference's Eigen::Vector3f / Matrix3f (pch.h:24-25).
struct Vec3 {
    float d[3] = {0.f, 0.f, 0.f};
    Vec3() = default;
    Vec3(float a, float b, float c) : d{a, b, c} {}
    float& operator[](int i) { return d[i]; }
    float operator[](int i) const { return d[i]; }
    float& operator()(int i) { return d[i]; }
    float operator()(int i) const { return d[i]; }
    static Vec3 Zero() { return Vec3(); }
    void setZero() { d[0] = d[1] = d[2] = 0.f; }
    Vec3 operator+(const Vec3& o) const { return {d[0] + o.d[0], d[1] + o.d[1], d[2] + o.d[2]}; }
    Vec3 operator-(const Vec3& o) const { return {d[0] - o.d[0], d[1] - o.d[1], d[2] - o.d[2]}; }
    Vec3 operator*(float s) const { return {d[0] * s, d[1] * s, d[2] * s}; }
    Vec3& operator+=(const Vec3& o)
    {
        for (int i = 0; i < 3; ++i) d[i] += o.d[i];
        return *this;
    }
    float dot(const Vec3& o) const { return d[0] * o.d[0] + d[1] * o.d[1] + d[2] * o.d[2]; }
    float squaredNorm() const { return dot(*this); }
    float norm() const { return std::sqrt(squaredNorm()); }
};
inline Vec3 operator*(float s, const Vec3& v) { return v * s; }

struct Mat3 {
    float d[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // row-major
    float& operator()(int i, int j) { return d[3 * i + j]; }
    float operator()(int i, int j) const { return d[3 * i + j]; }
    static Mat3 Zero() { return Mat3(); }
    static Mat3 Identity()
    {
        Mat3 m;
        m.d[0] = m.d[4] = m.d[8] = 1.f;
        return m;
    }
    void setZero()
    {
        for (float& x : d) x = 0.f;
    }
    Mat3 transpose() const
    {
        Mat3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r(i, j) = (*this)(j, i);
        return r;
    }
    Mat3 operator+(const Mat3& o) const
    {
        Mat3 r;
        for (int i = 0; i < 9; ++i) r.d[i] = d[i] + o.d[i];
        return r;
    }
    Mat3 operator-(const Mat3& o) const
    {
        Mat3 r;
        for (int i = 0; i < 9; ++i) r.d[i] = d[i] - o.d[i];
        return r;
    }
    Mat3 operator*(float s) const
    {
        Mat3 r;
        for (int i = 0; i < 9; ++i) r.d[i] = d[i] * s;
        return r;
    }
    Mat3 operator*(const Mat3& o) const
    {
        Mat3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r(i, j) = (*this)(i, 0) * o(0, j) + (*this)(i, 1) * o(1, j) + (*this)(i, 2) * o(2, j);
        return r;
    }
    Vec3 operator*(const Vec3& v) const
    {
        return {d[0] * v[0] + d[1] * v[1] + d[2] * v[2], d[3] * v[0] + d[4] * v[1] + d[5] * v[2],
                d[6] * v[0] + d[7] * v[1] + d[8] * v[2]};
    }
    Mat3& operator+=(const Mat3& o)
    {
        for (int i = 0; i < 9; ++i) d[i] += o.d[i];
        return *this;
    }
    Mat3& operator-=(const Mat3& o)
    {
        for (int i = 0; i < 9; ++i) d[i] -= o.d[i];
        return *this;
    }
    float squaredNorm() const
    {
        float s = 0.f;
        for (float x : d) s += x * x;
        return s;
    }
    float norm() const { return std::sqrt(squaredNorm()); }
    float determinant() const
    {
        return d[0] * (d[4] * d[8] - d[5] * d[7]) - d[1] * (d[3] * d[8] - d[5] * d[6]) + d[2] * (d[3] * d[7] - d[4] * d[6]);
    }
};
inline Mat3 operator*(float s, const Mat3& m) { return m * s; }

// pch.h:27-30
inline float InnerProduct(const Mat3& A, const Mat3& B)
{
    float s = 0.f;
    for (int i = 0; i < 9; ++i) s += A.d[i] * B.d[i];
    return s;
}

class DeviceError : public std::runtime_error {
   public:
    explicit DeviceError(const std::string& m) : std::runtime_error(m) {}
};

// How the mirror drives the GPU.  Defaults: the plain FP64 path with every
// layer stored and the reference's own adjoint (compat), i.e. the drop-in.
struct DeviceOptions {
    int precision = 0;   // DMPM_PREC_F64 / DMPM_PREC_F32
    int tape = 0;        // DMPM_TAPE_STORE_ALL / DMPM_TAPE_CHECKPOINT
    int compat = 15;     // DMPM_COMPAT_REFERENCE; 0 = exact adjoint
    int device = 0;
};
DeviceOptions& GetDeviceOptions();

// ------------------------------------------------------- MaterialPoint ----
// MaterialPoint.h:8-75
struct MaterialPoint {
    void ResetGradients();
    bool IsEqualToOtherPoint(const MaterialPoint& other_mp) const;

    Vec3 x = Vec3::Zero();
    Vec3 v = Vec3::Zero();
    Mat3 F = Mat3::Identity();
    Mat3 C = Mat3::Zero();
    Mat3 P = Mat3::Zero();

    float m = 0.f;
    float vol = 0.f;

    float lam = 0.f;
    float mu = 0.f;

    Mat3 dFc = Mat3::Zero();

    Vec3 dLdx = Vec3::Zero();
    Vec3 dLdv = Vec3::Zero();
    Mat3 dLdF = Mat3::Zero();
    Mat3 dLdC = Mat3::Zero();
    Mat3 dLdP = Mat3::Zero();

    Vec3 dLdv_next = Vec3::Zero();
    Mat3 dLdC_next = Mat3::Zero();

    Mat3 momentum = Mat3::Zero();
    Mat3 vector = Mat3::Zero();
    Mat3 vector_max = Mat3::Zero();
};

// ----------------------------------------------------------- PointCloud ----
// PointCloud.h:10-66 (optimiser updates of PointCloud.cpp:83-178)
struct PointCloud {
    void ResetGradients();
    float Compute_dLdF_Norm();
    void Descend_Adam(float alpha, float gradient_norm, float beta1, float beta2, float epsilon, int timestep);
    void Descend_AMSGrad(float alpha, float gradient_norm, float beta1, float beta2, float epsilon, int timestep);
    void Descend_GradientDescent(float alpha, float gradient_norm);
    void Descend_Lion(float alpha, float beta, float gradient_norm);
    void Descend_dLdF(float alpha, float gradient_norm);
    std::vector<int> generateMinibatchIndices(int total_size, int minibatch_size);
    float Compute_dLdF_Norm_Stochastic(const std::vector<int>& minibatch_indices);
    void Descend_Adam_Stochastic(float alpha, float gradient_norm, float beta1, float beta2, float epsilon, int timestep,
                                 const std::vector<int>& minibatch_indices);
    inline float SignFloat(float x) { return (x > 0.f) ? 1.f : ((x < 0.f) ? -1.f : 0.f); }
    void RemovePoint(size_t point_index);

    // .mpmbin: the reference's cereal binary archive of the point cloud (PointCloud.cpp:334-362)
    void WriteEntirePointCloudToBinaryFile(std::string file_path);
    bool ReadEntirePointCloudFromBinaryFile(std::string file_path);
    bool IsEqualToOtherPointCloud(const PointCloud& other_pc);

    std::vector<Vec3> GetPointPositions() const;
    std::vector<float> GetPointMasses() const;
    std::vector<float> GetPointVolumes() const;
    std::vector<Mat3> GetPointDefGrads() const;
    // det(F + dFc) and FixedCorotatedElasticity(F + dFc, lam, mu) per point (PointCloud.cpp:441-467)
    std::vector<float> GetPointDeterminants() const;
    std::vector<float> GetPointElasticEnergies() const;

    std::vector<MaterialPoint> points;
};

// ------------------------------------------------------------------ Grid ----
// GridNode.h:5-24, Grid.h:7-57
struct GridNode {
    void ResetGradients();

    Vec3 x;
    Vec3 v;
    Vec3 p;
    float m = 0.f;

    Vec3 dLdx;
    Vec3 dLdv;
    Vec3 dLdp;
    float dLdm = 0.f;
};

struct Grid {
    Grid(int _dim_x, int _dim_y, int dim_z, float _dx, Vec3 _min_point);
    Grid(const Grid& grid) = default;

    // Grid.cpp:40-81: the grid nodes of a point's 4^3 cubic B-spline stencil (base
    // floor((x - min)/dx) - 1, x-major), skipping those outside the grid; their
    // (i, j, k) are appended to indices when given
    std::vector<std::reference_wrapper<GridNode>> QueryPoint_CubicBSpline(Vec3 point,
                                                                          std::vector<std::array<int, 3>>* indices = nullptr);
    std::vector<std::reference_wrapper<const GridNode>> QueryPointConst_CubicBSpline(
        Vec3 point, std::vector<std::array<int, 3>>* indices = nullptr) const;

    std::vector<GridNode> flatten() const;
    void unflatten(const std::vector<GridNode>& flat_nodes);
    std::vector<Vec3> GetNodePositions() const;
    std::vector<float> GetNodeMasses() const;
    std::vector<Vec3> GetNodeVelocities() const;

    void ResetGradients();
    void ResetValues();

    int dim_x;
    int dim_y;
    int dim_z;
    float dx;
    Vec3 min_point;
    std::vector<std::vector<std::vector<GridNode>>> nodes;
};

// ------------------------------------------------------------- CompGraph ----
// CompGraph.h:11-67
struct CompGraphLayer {
    std::shared_ptr<PointCloud> point_cloud = nullptr;
    std::shared_ptr<Grid> grid = nullptr;
};

class CompGraph {
   public:
    CompGraph(std::shared_ptr<PointCloud> initial_point_cloud, std::shared_ptr<Grid> grid, std::shared_ptr<const Grid> _target_grid);
    ~CompGraph();
    CompGraph(const CompGraph&) = delete;
    CompGraph& operator=(const CompGraph&) = delete;

    void OptimizeDefGradControlSequence(int num_steps, float _dt, float _drag, Vec3 _f_ext, int control_stride,
                                        int max_gd_iters, int max_line_search_iters, float initial_alpha, float gd_tol,
                                        float smoothing_factor);

    void SetUpCompGraph(int num_layers);
    float EndLayerMassLoss();
    void ComputeForwardPass(size_t start_layer);
    void ComputeBackwardPass(size_t control_layer);
    void FiniteDifferencesGradientTest(int num_steps, size_t particle_id);

    // simulation parameters (private in the reference, CompGraph.h:55-58)
    void SetSimulationParameters(float _dt, float _drag, Vec3 _f_ext, float _smoothing_factor);

    std::vector<CompGraphLayer> layers;
    std::shared_ptr<const Grid> target_grid;

    std::vector<float> gradientNormHistory;

   private:
    void EnsureDevice();
    void ReleaseDevice();

    Vec3 f_ext = Vec3::Zero();
    float dt = 1.0f / 120.0f;
    float drag = 0.5f;
    float smoothing_factor = 0.1f;
    int timestep = 1;

    dmpm_ctx* ctx = nullptr;
    int ctx_layers = 0;
    bool ctx_fresh = false;
    std::array<double, 6> ctx_params{};
};

// ------------------------------------------------------ ForwardSimulation ----
// ForwardSimulation.h:7-45.  ForwardTimeStep runs the fused step on the GPU;
// the phase functions run each phase as its own GPU call on these containers
// (FP64, device 0), with the reference's boundaries: P2G reads the stored P
// and accumulates into the grid, G_op leaves nodes at or below the mass
// threshold untouched, G2P writes v and C of next, P_op_2 its F and x.  The
// single-particle and single-node forms make one GPU call each.
namespace SingleThreadMPM {
void SingleParticle_op_1(MaterialPoint& mp);
void SingleParticle_to_grid(const MaterialPoint& mp, Grid& grid, float dt, float drag);
void SingleNode_op(GridNode& node, float dt, Vec3 f_ext);
void Grid_to_SingleParticle(MaterialPoint& next_timestep_mp, const MaterialPoint& curr_timestep_mp, const Grid& grid);
void SingleParticle_op_2(MaterialPoint& next_timestep_mp, const MaterialPoint& curr_timestep_mp, float dt);
void SingleParticle_op_2(MaterialPoint& next_timestep_mp, const MaterialPoint& curr_timestep_mp, float smoothing_factor, float dt);

void ForwardTimeStep(PointCloud& next_point_cloud, PointCloud& curr_point_cloud, Grid& grid, float dt, float drag, Vec3 f_ext);
void ForwardTimeStep(PointCloud& next_point_cloud, PointCloud& curr_point_cloud, Grid& grid, float smoothing_factor,
                     float dt, float drag, Vec3 f_ext);

// F <- F (1 - s) + F_curr s (ForwardSimulation.cpp:114-116), host arithmetic on one particle
void smooth_deformation_gradient(MaterialPoint& next_timestep_mp, const MaterialPoint& curr_timestep_mp, float smoothing_factor);

void P_op_1(PointCloud& curr_point_cloud);
void G_Reset(Grid& grid);
void P2G(const PointCloud& curr_point_cloud, Grid& grid, float dt, float drag);
void G_op(Grid& grid, float dt, Vec3 f_ext);
void G_op(Grid& grid, float dt, Vec3 gravity_point, float gravity_mag);
void G2P(PointCloud& next_point_cloud, const PointCloud& curr_point_cloud, Grid& grid);
void P_op_2(PointCloud& next_point_cloud, const PointCloud& curr_point_cloud, float dt);
void P_op_2(PointCloud& next_point_cloud, const PointCloud& curr_point_cloud, float smoothing_factor, float dt);

void CalculatePointCloudVolumes(PointCloud& curr_point_cloud, Grid& grid);
void P2G_Mass(const std::vector<Vec3> points, Grid& grid, float mp_m);
void G2G_Mass(Grid& grid_1, Grid& grid_2);

// not in the reference: grid.m <- P2G mass of a point cloud's own masses, the other node
// fields cleared (the target-grid construction of MPMPolyscope.cpp:156-160 in one call)
void P2G_PointCloudMass(const PointCloud& point_cloud, Grid& grid);
}  // namespace SingleThreadMPM

// ----------------------------------------- MultiThreadForwardSimulation ----
// MultiThreadForwardSimulation.h:6-11: the step the animator replays in place,
// next and curr the same cloud (DiffMPMAnimator3D.cpp:83-111): no smoothing.
// In the reference proxy_grids holds one grid per std::thread for a race-free
// P2G and its size sets the thread count; the GPU scatter needs no copies, so
// the proxy grids are not written here.  An empty vector (a division by zero
// in the reference) raises std::invalid_argument.
namespace MultiThreadMPM {
void ForwardTimeStep(PointCloud& next_point_cloud, PointCloud& curr_point_cloud, Grid& main_grid,
                     std::vector<std::shared_ptr<Grid>>& proxy_grids, float dt, float drag, Vec3 f_ext);
}  // namespace MultiThreadMPM

// ------------------------------------------------------- BackPropagation ----
// BackPropagation.h:7
void Back_Timestep(CompGraphLayer& layer_nplus1, CompGraphLayer& layer_n, float drag, float dt, float smoothing_factor);

// ------------------------------------------------------------ Elasticity ----
// Elasticity.h:9-23
float FixedCorotatedElasticity(const Mat3& F, float lam, float mu);
Mat3 PK_FixedCorotatedElasticity(const Mat3& F, float lam, float mu);
Mat3 d2_FCE_psi_dF2_mult_by_dF(const Mat3& F, float lam, float mu, const Mat3& dF);
void CalculateLameParameters(float young_mod, float poisson, float& lam, float& mu);

}  // namespace DiffMPMLib3D
