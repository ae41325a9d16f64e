// C++ host mirror of the reference's MLS-MPM API over the C ABI.
//
// Containers (MaterialPoint, PointCloud, Grid, CompGraph layers) keep host
// data as the reference does; the step, the passes, the loss, the adjoint and
// the constitutive model run on the GPU through include/dmpm.h.  Reference
// files cited per function are relative to the reference root.
#include "DiffMPMLib3D/DiffMPMLib3D.h"

#include "dmpm.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <stdexcept>

namespace DiffMPMLib3D {

namespace {

void check(int rc, const char* what)
{
    if (rc != DMPM_OK) throw DeviceError(std::string(what) + ": " + dmpm_last_error());
}

// host AoS double staging of one point cloud
struct Staging {
    std::vector<double> x, v, F, C, m, vol, lam, mu, dFc;
    explicit Staging(size_t n) : x(3 * n), v(3 * n), F(9 * n), C(9 * n), m(n), vol(n), lam(n), mu(n), dFc(9 * n) {}
};

Staging stage(const PointCloud& pc)
{
    const size_t n = pc.points.size();
    Staging s(n);
    for (size_t p = 0; p < n; ++p) {
        const MaterialPoint& mp = pc.points[p];
        for (int a = 0; a < 3; ++a) {
            s.x[3 * p + a] = mp.x[a];
            s.v[3 * p + a] = mp.v[a];
        }
        for (int i = 0; i < 9; ++i) {
            s.F[9 * p + i] = mp.F.d[i];
            s.C[9 * p + i] = mp.C.d[i];
            s.dFc[9 * p + i] = mp.dFc.d[i];
        }
        s.m[p] = mp.m;
        s.vol[p] = mp.vol;
        s.lam[p] = mp.lam;
        s.mu[p] = mp.mu;
    }
    return s;
}

std::vector<double> controls_of(const PointCloud& pc)
{
    std::vector<double> c(9 * pc.points.size());
    for (size_t p = 0; p < pc.points.size(); ++p)
        for (int i = 0; i < 9; ++i) c[9 * p + i] = pc.points[p].dFc.d[i];
    return c;
}

void unstage_state(PointCloud& pc, const std::vector<double>& x, const std::vector<double>& v,
                   const std::vector<double>& F, const std::vector<double>& C)
{
    for (size_t p = 0; p < pc.points.size(); ++p) {
        MaterialPoint& mp = pc.points[p];
        for (int a = 0; a < 3; ++a) {
            mp.x[a] = (float)x[3 * p + a];
            mp.v[a] = (float)v[3 * p + a];
        }
        for (int i = 0; i < 9; ++i) {
            mp.F.d[i] = (float)F[9 * p + i];
            mp.C.d[i] = (float)C[9 * p + i];
        }
    }
}

std::vector<double> grid_masses(const Grid& g)
{
    std::vector<double> m((size_t)g.dim_x * g.dim_y * g.dim_z);
    size_t k = 0;
    for (int i = 0; i < g.dim_x; ++i)
        for (int j = 0; j < g.dim_y; ++j)
            for (int l = 0; l < g.dim_z; ++l) m[k++] = g.nodes[i][j][l].m;
    return m;
}

dmpm_config make_config(size_t n, const Grid& g, int layers, float dt, float drag, float s, Vec3 f, bool keep)
{
    const DeviceOptions& o = GetDeviceOptions();
    dmpm_config cfg;
    check(dmpm_config_init(&cfg), "dmpm_config_init");
    cfg.precision = o.precision;
    cfg.tape_mode = o.tape;
    cfg.compat = o.compat;
    cfg.device = o.device;
    cfg.n_particles = (int32_t)n;
    cfg.dim[0] = g.dim_x;
    cfg.dim[1] = g.dim_y;
    cfg.dim[2] = g.dim_z;
    cfg.num_layers = layers;
    cfg.dx = g.dx;
    for (int a = 0; a < 3; ++a) {
        cfg.min_point[a] = g.min_point[a];
        cfg.f_ext[a] = f[a];
    }
    cfg.dt = dt;
    cfg.drag = drag;
    cfg.smoothing = s;
    cfg.keep_control_grads = keep ? 1 : 0;
    return cfg;
}

struct Ctx {
    dmpm_ctx* c = nullptr;
    explicit Ctx(const dmpm_config& cfg) { check(dmpm_create(&cfg, &c), "dmpm_create"); }
    ~Ctx() { dmpm_destroy(c); }
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
};

void upload_particles(dmpm_ctx* c, const PointCloud& pc)
{
    Staging s = stage(pc);
    check(dmpm_set_particles(c, s.x.data(), s.v.data(), s.F.data(), s.C.data(), s.m.data(), s.vol.data(), s.lam.data(),
                             s.mu.data()),
          "dmpm_set_particles");
}

void download_layer(dmpm_ctx* c, int layer, PointCloud& pc)
{
    const size_t n = pc.points.size();
    std::vector<double> x(3 * n), v(3 * n), F(9 * n), C(9 * n);
    check(dmpm_get_layer_state(c, layer, x.data(), v.data(), F.data(), C.data()), "dmpm_get_layer_state");
    unstage_state(pc, x, v, F, C);
}

// grid of one step -> GridNode m, v and p = m (v - dt f) (the momentum before G_op)
void download_grid(dmpm_ctx* c, int step, Grid& g, float dt, Vec3 f)
{
    const size_t G = (size_t)g.dim_x * g.dim_y * g.dim_z;
    std::vector<double> m(G), v(3 * G);
    check(dmpm_get_grid(c, step, m.data(), v.data()), "dmpm_get_grid");
    size_t k = 0;
    for (int i = 0; i < g.dim_x; ++i)
        for (int j = 0; j < g.dim_y; ++j)
            for (int l = 0; l < g.dim_z; ++l, ++k) {
                GridNode& nd = g.nodes[i][j][l];
                nd.m = (float)m[k];
                for (int a = 0; a < 3; ++a) {
                    nd.v[a] = (float)v[3 * k + a];
                    nd.p[a] = (float)(m[k] * (v[3 * k + a] - (m[k] > 1.1920928955078125e-07 ? dt * f[a] : 0.0)));
                }
            }
}

void download_grads(dmpm_ctx* c, int layer, PointCloud& pc)
{
    const size_t n = pc.points.size();
    std::vector<double> gx(3 * n), gv(3 * n), gF(9 * n), gC(9 * n);
    check(dmpm_get_layer_grads(c, layer, gx.data(), gv.data(), gF.data(), gC.data()), "dmpm_get_layer_grads");
    for (size_t p = 0; p < n; ++p) {
        MaterialPoint& mp = pc.points[p];
        for (int a = 0; a < 3; ++a) {
            mp.dLdx[a] = (float)gx[3 * p + a];
            mp.dLdv[a] = (float)gv[3 * p + a];
        }
        for (int i = 0; i < 9; ++i) {
            mp.dLdF.d[i] = (float)gF[9 * p + i];
            mp.dLdC.d[i] = (float)gC[9 * p + i];
        }
    }
}

}  // namespace

DeviceOptions& GetDeviceOptions()
{
    static DeviceOptions opts;
    return opts;
}

// ------------------------------------------------------- MaterialPoint ----
// MaterialPoint.cpp:4-17
void MaterialPoint::ResetGradients()
{
    dLdx.setZero();
    dLdv.setZero();
    dLdF.setZero();
    dLdC.setZero();
    dLdP.setZero();
    dLdv_next.setZero();
    dLdC_next.setZero();
}

bool MaterialPoint::IsEqualToOtherPoint(const MaterialPoint& o) const
{
    auto eq3 = [](const Vec3& a, const Vec3& b) { return std::memcmp(a.d, b.d, sizeof a.d) == 0; };
    auto eq9 = [](const Mat3& a, const Mat3& b) { return std::memcmp(a.d, b.d, sizeof a.d) == 0; };
    return eq3(x, o.x) && eq3(v, o.v) && eq9(F, o.F) && eq9(C, o.C) && eq9(P, o.P) && m == o.m && vol == o.vol &&
           lam == o.lam && mu == o.mu;
}

// ----------------------------------------------------------- PointCloud ----
// PointCloud.cpp:16-22
void PointCloud::ResetGradients()
{
    for (MaterialPoint& mp : points) mp.ResetGradients();
}

// PointCloud.cpp:24-35 (accumulated in double)
float PointCloud::Compute_dLdF_Norm()
{
    double s = 0.0;
    for (const MaterialPoint& mp : points) s += mp.dLdF.squaredNorm();
    return (float)std::sqrt(s);
}

// PointCloud.cpp:83-99
void PointCloud::Descend_Adam(float alpha, float gradient_norm, float beta1, float beta2, float epsilon, int ts)
{
    const float c1 = 1.0f - std::pow(beta1, (float)ts), c2 = 1.0f - std::pow(beta2, (float)ts);
    for (MaterialPoint& mp : points)
        for (int i = 0; i < 9; ++i) {
            const float g = mp.dLdF.d[i] / gradient_norm;
            mp.momentum.d[i] = beta1 * mp.momentum.d[i] + (1.0f - beta1) * g;
            mp.vector.d[i] = beta2 * mp.vector.d[i] + (1.0f - beta2) * g * g;
            const float mh = mp.momentum.d[i] / c1, vh = mp.vector.d[i] / c2;
            mp.dFc.d[i] -= alpha * (mh / (std::sqrt(vh) + epsilon));
        }
}

// PointCloud.cpp:101-132
void PointCloud::Descend_AMSGrad(float alpha, float gradient_norm, float beta1, float beta2, float epsilon, int ts)
{
    const float c1 = 1.f - std::pow(beta1, (float)ts), c2 = 1.f - std::pow(beta2, (float)ts);
    for (MaterialPoint& mp : points)
        for (int i = 0; i < 9; ++i) {
            const float g = mp.dLdF.d[i] / gradient_norm;
            mp.momentum.d[i] = beta1 * mp.momentum.d[i] + (1.f - beta1) * g;
            mp.vector.d[i] = beta2 * mp.vector.d[i] + (1.f - beta2) * g * g;
            const float mh = mp.momentum.d[i] / c1, vh = mp.vector.d[i] / c2;
            mp.vector_max.d[i] = std::max(mp.vector_max.d[i], vh);
            mp.dFc.d[i] -= alpha * (mh / (std::sqrt(mp.vector_max.d[i]) + epsilon));
        }
}

// PointCloud.cpp:134-162
void PointCloud::Descend_Lion(float alpha, float beta, float gradient_norm)
{
    for (MaterialPoint& mp : points)
        for (int i = 0; i < 9; ++i) {
            const float g = mp.dLdF.d[i] / gradient_norm;
            mp.momentum.d[i] = beta * mp.momentum.d[i] + (1.f - beta) * SignFloat(g);
            mp.dFc.d[i] -= alpha * mp.momentum.d[i];
        }
}

// PointCloud.cpp:164-178
void PointCloud::Descend_GradientDescent(float alpha, float gradient_norm)
{
    for (MaterialPoint& mp : points)
        for (int i = 0; i < 9; ++i) mp.dFc.d[i] -= alpha * (mp.dLdF.d[i] / gradient_norm);
}

// PointCloud.cpp:181-204: clipped direction (|g| in [1e-3, 1], sign kept) into
// a heavy-ball momentum (0.9), dFc -= momentum
void PointCloud::Descend_dLdF(float alpha, float gradient_norm)
{
    const float beta = 0.9f, lo = 1e-3f, hi = 1.0f;
    for (MaterialPoint& mp : points)
        for (int i = 0; i < 9; ++i) {
            float g = mp.dLdF.d[i] / gradient_norm;
            if (g > hi) g = hi;
            else if (g < -hi) g = -hi;
            else if (std::fabs(g) < lo) g = g < 0.f ? -lo : lo;
            mp.momentum.d[i] = beta * mp.momentum.d[i] + alpha * g;
            mp.dFc.d[i] -= mp.momentum.d[i];
        }
}

// PointCloud.cpp:37-49: minibatch_size indices drawn without replacement, clock-seeded
std::vector<int> PointCloud::generateMinibatchIndices(int total_size, int minibatch_size)
{
    std::vector<int> idx(total_size);
    for (int i = 0; i < total_size; ++i) idx[i] = i;
    std::mt19937 g((unsigned)std::chrono::system_clock::now().time_since_epoch().count());
    std::shuffle(idx.begin(), idx.end(), g);
    idx.resize(minibatch_size);
    return idx;
}

// PointCloud.cpp:51-63 (accumulated in double)
float PointCloud::Compute_dLdF_Norm_Stochastic(const std::vector<int>& minibatch_indices)
{
    double s = 0.0;
    for (int i : minibatch_indices) s += points[i].dLdF.squaredNorm();
    return (float)std::sqrt(s);
}

// PointCloud.cpp:65-81: Adam on the minibatch's points, bias corrections at timestep + 1
void PointCloud::Descend_Adam_Stochastic(float alpha, float gradient_norm, float beta1, float beta2, float epsilon, int timestep,
                                         const std::vector<int>& minibatch_indices)
{
    const float c1 = 1.0f - std::pow(beta1, (float)(timestep + 1)), c2 = 1.0f - std::pow(beta2, (float)(timestep + 1));
    for (int p : minibatch_indices) {
        MaterialPoint& mp = points[p];
        for (int i = 0; i < 9; ++i) {
            const float g = mp.dLdF.d[i] / gradient_norm;
            mp.momentum.d[i] = beta1 * mp.momentum.d[i] + (1.0f - beta1) * g;
            mp.vector.d[i] = beta2 * mp.vector.d[i] + (1.0f - beta2) * g * g;
            mp.dFc.d[i] -= alpha * ((mp.momentum.d[i] / c1) / (std::sqrt(mp.vector.d[i] / c2) + epsilon));
        }
    }
}

void PointCloud::RemovePoint(size_t point_index) { points.erase(points.begin() + point_index); }

bool PointCloud::IsEqualToOtherPointCloud(const PointCloud& other_pc)
{
    if (other_pc.points.size() != points.size()) return false;
    for (size_t i = 0; i < points.size(); ++i)
        if (!points[i].IsEqualToOtherPoint(other_pc.points[i])) return false;
    return true;
}

std::vector<Vec3> PointCloud::GetPointPositions() const
{
    std::vector<Vec3> r;
    for (const MaterialPoint& mp : points) r.push_back(mp.x);
    return r;
}
std::vector<float> PointCloud::GetPointMasses() const
{
    std::vector<float> r;
    for (const MaterialPoint& mp : points) r.push_back(mp.m);
    return r;
}
std::vector<float> PointCloud::GetPointVolumes() const
{
    std::vector<float> r;
    for (const MaterialPoint& mp : points) r.push_back(mp.vol);
    return r;
}
std::vector<Mat3> PointCloud::GetPointDefGrads() const
{
    std::vector<Mat3> r;
    for (const MaterialPoint& mp : points) r.push_back(mp.F);
    return r;
}
std::vector<float> PointCloud::GetPointDeterminants() const
{
    std::vector<float> r;
    for (const MaterialPoint& mp : points) r.push_back((mp.F + mp.dFc).determinant());
    return r;
}
// all points in one batched GPU evaluation (FP64, rounded to float)
std::vector<float> PointCloud::GetPointElasticEnergies() const
{
    const size_t n = points.size();
    std::vector<double> F(9 * n), lam(n), mu(n), psi(n);
    for (size_t p = 0; p < n; ++p) {
        const Mat3 Fe = points[p].F + points[p].dFc;
        for (int i = 0; i < 9; ++i) F[9 * p + i] = Fe.d[i];
        lam[p] = points[p].lam;
        mu[p] = points[p].mu;
    }
    if (n) check(dmpm_eval_energy((int32_t)n, F.data(), lam.data(), mu.data(), psi.data()), "dmpm_eval_energy");
    return std::vector<float>(psi.begin(), psi.end());
}

// .mpmbin: cereal's binary archive of PointCloud (PointCloud.cpp:334-362,
// MaterialPoint.h:52-74, cereal_eigen.h:10-38): a uint64 point count, then per
// point x, v as (int32 rows, 3 floats), F, C, P as (int32 rows, int32 cols,
// 9 floats row by row), m, vol, lam, mu, dFc, dLdx, dLdv, dLdF, momentum,
// vector, vector_max; little-endian.
namespace {
void put_vec(std::ostream& o, const Vec3& v)
{
    const int32_t r = 3;
    o.write((const char*)&r, 4);
    o.write((const char*)v.d, 12);
}
void put_mat(std::ostream& o, const Mat3& m)
{
    const int32_t r = 3, c = 3;
    o.write((const char*)&r, 4);
    o.write((const char*)&c, 4);
    o.write((const char*)m.d, 36);
}
void put_f(std::ostream& o, float f) { o.write((const char*)&f, 4); }
bool get_vec(std::istream& i, Vec3& v)
{
    int32_t r = 0;
    i.read((char*)&r, 4);
    if (r != 3) return false;
    i.read((char*)v.d, 12);
    return (bool)i;
}
bool get_mat(std::istream& i, Mat3& m)
{
    int32_t r = 0, c = 0;
    i.read((char*)&r, 4);
    i.read((char*)&c, 4);
    if (r != 3 || c != 3) return false;
    i.read((char*)m.d, 36);
    return (bool)i;
}
bool get_f(std::istream& i, float& f)
{
    i.read((char*)&f, 4);
    return (bool)i;
}
}  // namespace

void PointCloud::WriteEntirePointCloudToBinaryFile(std::string file_path)
{
    std::ofstream o(file_path, std::ios::binary);
    if (!o.good()) return;
    const uint64_t n = points.size();
    o.write((const char*)&n, 8);
    for (const MaterialPoint& mp : points) {
        put_vec(o, mp.x);
        put_vec(o, mp.v);
        put_mat(o, mp.F);
        put_mat(o, mp.C);
        put_mat(o, mp.P);
        put_f(o, mp.m);
        put_f(o, mp.vol);
        put_f(o, mp.lam);
        put_f(o, mp.mu);
        put_mat(o, mp.dFc);
        put_vec(o, mp.dLdx);
        put_vec(o, mp.dLdv);
        put_mat(o, mp.dLdF);
        put_mat(o, mp.momentum);
        put_mat(o, mp.vector);
        put_mat(o, mp.vector_max);
    }
}

bool PointCloud::ReadEntirePointCloudFromBinaryFile(std::string file_path)
{
    std::ifstream i(file_path, std::ios::binary);
    if (!i.good()) {
        std::cout << "couldn't read file: " << file_path << std::endl;
        return false;
    }
    uint64_t n = 0;
    i.read((char*)&n, 8);
    if (!i) return false;
    std::vector<MaterialPoint> pts(n);
    for (MaterialPoint& mp : pts) {
        if (!(get_vec(i, mp.x) && get_vec(i, mp.v) && get_mat(i, mp.F) && get_mat(i, mp.C) && get_mat(i, mp.P) &&
              get_f(i, mp.m) && get_f(i, mp.vol) && get_f(i, mp.lam) && get_f(i, mp.mu) && get_mat(i, mp.dFc) &&
              get_vec(i, mp.dLdx) && get_vec(i, mp.dLdv) && get_mat(i, mp.dLdF) && get_mat(i, mp.momentum) &&
              get_mat(i, mp.vector) && get_mat(i, mp.vector_max)))
            return false;
    }
    points.swap(pts);
    return true;
}

// ------------------------------------------------------------------ Grid ----
void GridNode::ResetGradients()
{
    dLdx.setZero();
    dLdv.setZero();
    dLdp.setZero();
    dLdm = 0.f;
}

// Grid.cpp:3-28 (node x = min + (i,j,k) dx)
Grid::Grid(int _dim_x, int _dim_y, int _dim_z, float _dx, Vec3 _min_point)
    : dim_x(_dim_x), dim_y(_dim_y), dim_z(_dim_z), dx(_dx), min_point(_min_point)
{
    nodes.assign(dim_x, std::vector<std::vector<GridNode>>(dim_y, std::vector<GridNode>(dim_z)));
    for (int i = 0; i < dim_x; ++i)
        for (int j = 0; j < dim_y; ++j)
            for (int k = 0; k < dim_z; ++k) nodes[i][j][k].x = min_point + Vec3((float)i, (float)j, (float)k) * dx;
}

// the stencil of Grid.cpp:40-81 (the base clamped to +-2^20 cells before the
// integer conversion, so that a far-away or non-finite point selects no node)
template <typename Node, typename G>
static std::vector<std::reference_wrapper<Node>> query_stencil(G& g, Vec3 point, std::vector<std::array<int, 3>>* indices)
{
    std::vector<std::reference_wrapper<Node>> ret;
    int base[3];
    const int dims[3] = {g.dim_x, g.dim_y, g.dim_z};
    for (int a = 0; a < 3; ++a) {
        double u = std::floor(((double)point[a] - (double)g.min_point[a]) / (double)g.dx);
        if (!(u >= -1048576.0)) u = -1048576.0;  // also NaN
        if (u > 1048576.0) u = 1048576.0;
        base[a] = (int)u - 1;
    }
    for (int i = 0; i <= 3; ++i)
        for (int j = 0; j <= 3; ++j)
            for (int k = 0; k <= 3; ++k) {
                const int ix[3] = {base[0] + i, base[1] + j, base[2] + k};
                if (ix[0] < 0 || ix[0] >= dims[0] || ix[1] < 0 || ix[1] >= dims[1] || ix[2] < 0 || ix[2] >= dims[2]) continue;
                ret.push_back(std::ref(g.nodes[ix[0]][ix[1]][ix[2]]));
                if (indices) indices->push_back({ix[0], ix[1], ix[2]});
            }
    return ret;
}

std::vector<std::reference_wrapper<GridNode>> Grid::QueryPoint_CubicBSpline(Vec3 point, std::vector<std::array<int, 3>>* indices)
{
    return query_stencil<GridNode>(*this, point, indices);
}

std::vector<std::reference_wrapper<const GridNode>> Grid::QueryPointConst_CubicBSpline(Vec3 point,
                                                                                      std::vector<std::array<int, 3>>* indices) const
{
    return query_stencil<const GridNode>(*this, point, indices);
}

std::vector<GridNode> Grid::flatten() const
{
    std::vector<GridNode> r;
    r.reserve((size_t)dim_x * dim_y * dim_z);
    for (int i = 0; i < dim_x; ++i)
        for (int j = 0; j < dim_y; ++j)
            for (int k = 0; k < dim_z; ++k) r.push_back(nodes[i][j][k]);
    return r;
}

void Grid::unflatten(const std::vector<GridNode>& flat)
{
    size_t q = 0;
    for (int i = 0; i < dim_x; ++i)
        for (int j = 0; j < dim_y; ++j)
            for (int k = 0; k < dim_z; ++k) nodes[i][j][k] = flat[q++];
}

std::vector<Vec3> Grid::GetNodePositions() const
{
    std::vector<Vec3> r;
    for (const GridNode& n : flatten()) r.push_back(n.x);
    return r;
}
std::vector<float> Grid::GetNodeMasses() const
{
    std::vector<float> r;
    for (const GridNode& n : flatten()) r.push_back(n.m);
    return r;
}
std::vector<Vec3> Grid::GetNodeVelocities() const
{
    std::vector<Vec3> r;
    for (const GridNode& n : flatten()) r.push_back(n.v);
    return r;
}

// Grid.cpp:337-361
void Grid::ResetGradients()
{
    for (auto& a : nodes)
        for (auto& b : a)
            for (GridNode& n : b) n.ResetGradients();
}
void Grid::ResetValues()
{
    for (auto& a : nodes)
        for (auto& b : a)
            for (GridNode& n : b) {
                n.m = 0.f;
                n.v.setZero();
                n.p.setZero();
            }
}

// ------------------------------------------------------------ Elasticity ----
// Elasticity.cpp:74-86 on the GPU (FP64)
float FixedCorotatedElasticity(const Mat3& F, float lam, float mu)
{
    double f[9], l = lam, m = mu, psi = 0.0;
    for (int i = 0; i < 9; ++i) f[i] = F.d[i];
    check(dmpm_eval_energy(1, f, &l, &m, &psi), "dmpm_eval_energy");
    return (float)psi;
}

// Elasticity.cpp:88-98 on the GPU (FP64)
Mat3 PK_FixedCorotatedElasticity(const Mat3& F, float lam, float mu)
{
    double f[9], l = lam, m = mu, P[9];
    for (int i = 0; i < 9; ++i) f[i] = F.d[i];
    check(dmpm_eval_pk1(1, f, &l, &m, P), "dmpm_eval_pk1");
    Mat3 r;
    for (int i = 0; i < 9; ++i) r.d[i] = (float)P[i];
    return r;
}

// Elasticity.cpp:100-178 on the GPU (FP64, the reference's formula structure)
Mat3 d2_FCE_psi_dF2_mult_by_dF(const Mat3& F, float lam, float mu, const Mat3& dF)
{
    double f[9], d[9], l = lam, m = mu, out[9];
    for (int i = 0; i < 9; ++i) {
        f[i] = F.d[i];
        d[i] = dF.d[i];
    }
    check(dmpm_eval_hessian(1, f, &l, &m, d, out, 0), "dmpm_eval_hessian");
    Mat3 r;
    for (int i = 0; i < 9; ++i) r.d[i] = (float)out[i];
    return r;
}

// Elasticity.cpp:231-235
void CalculateLameParameters(float young_mod, float poisson, float& lam, float& mu)
{
    lam = young_mod * poisson / ((1.f + poisson) * (1.f - 2.f * poisson));
    mu = young_mod / (2.f + 2.f * poisson);
}

// ------------------------------------------------------ ForwardSimulation ----
namespace SingleThreadMPM {

// ForwardSimulation.cpp:136-157: one step curr -> next on the GPU; the grid
// receives m, v after G_op and p = m (v - dt f_ext).
void ForwardTimeStep(PointCloud& next, PointCloud& curr, Grid& grid, float smoothing_factor, float dt, float drag, Vec3 f_ext)
{
    const size_t n = curr.points.size();
    Ctx ctx(make_config(n, grid, 2, dt, drag, smoothing_factor, f_ext, false));
    upload_particles(ctx.c, curr);
    const std::vector<double> ctrl = controls_of(curr);
    check(dmpm_set_control(ctx.c, 0, ctrl.data()), "dmpm_set_control");
    check(dmpm_forward_step(ctx.c, 0), "dmpm_forward_step");
    if (next.points.size() != n) next.points = curr.points;
    download_layer(ctx.c, 1, next);
    download_grid(ctx.c, 0, grid, dt, f_ext);
}

// ForwardSimulation.cpp:126-134 (no smoothing)
void ForwardTimeStep(PointCloud& next, PointCloud& curr, Grid& grid, float dt, float drag, Vec3 f_ext)
{
    ForwardTimeStep(next, curr, grid, 0.f, dt, drag, f_ext);
}

// ForwardSimulation.cpp:262-290
void CalculatePointCloudVolumes(PointCloud& pc, Grid& grid)
{
    const size_t n = pc.points.size();
    Ctx ctx(make_config(n, grid, 2, 0.f, 0.f, 0.f, Vec3(), false));
    upload_particles(ctx.c, pc);
    check(dmpm_compute_volumes(ctx.c), "dmpm_compute_volumes");
    std::vector<double> vol(n);
    check(dmpm_get_volumes(ctx.c, vol.data()), "dmpm_get_volumes");
    for (size_t p = 0; p < n; ++p) pc.points[p].vol = (float)vol[p];
}

// ----------------------------------------------------------- phases -------
// Each phase is one GPU call (dmpm_phase_*, FP64) on the containers' values.
namespace {

dmpm_grid_desc desc_of(const Grid& g)
{
    dmpm_grid_desc d{};
    d.dim[0] = g.dim_x;
    d.dim[1] = g.dim_y;
    d.dim[2] = g.dim_z;
    d.dx = g.dx;
    for (int a = 0; a < 3; ++a) d.min_point[a] = g.min_point[a];
    return d;
}

// node fields in node order (i dy dz + j dz + k): m [G], p and v [G][3]
struct GridArrays {
    std::vector<double> m, p, v;
};

GridArrays grid_arrays(const Grid& g)
{
    GridArrays r;
    const size_t G = (size_t)g.dim_x * g.dim_y * g.dim_z;
    r.m.resize(G);
    r.p.resize(3 * G);
    r.v.resize(3 * G);
    size_t q = 0;
    for (int i = 0; i < g.dim_x; ++i)
        for (int j = 0; j < g.dim_y; ++j)
            for (int k = 0; k < g.dim_z; ++k, ++q) {
                const GridNode& nd = g.nodes[i][j][k];
                r.m[q] = nd.m;
                for (int a = 0; a < 3; ++a) {
                    r.p[3 * q + a] = nd.p[a];
                    r.v[3 * q + a] = nd.v[a];
                }
            }
    return r;
}

// write node fields back (null: leave that field)
void put_grid(Grid& g, const std::vector<double>* m, const std::vector<double>* p, const std::vector<double>* v)
{
    size_t q = 0;
    for (int i = 0; i < g.dim_x; ++i)
        for (int j = 0; j < g.dim_y; ++j)
            for (int k = 0; k < g.dim_z; ++k, ++q) {
                GridNode& nd = g.nodes[i][j][k];
                if (m) nd.m = (float)(*m)[q];
                for (int a = 0; a < 3; ++a) {
                    if (p) nd.p[a] = (float)(*p)[3 * q + a];
                    if (v) nd.v[a] = (float)(*v)[3 * q + a];
                }
            }
}

PointCloud cloud_of(const MaterialPoint& mp)
{
    PointCloud pc;
    pc.points.push_back(mp);
    return pc;
}

}  // namespace

// ForwardSimulation.cpp:6-9,159-167: P = PK1(F + dFc, lam, mu) per particle
void P_op_1(PointCloud& pc)
{
    const size_t n = pc.points.size();
    if (n == 0) return;
    std::vector<double> Fc(9 * n), lam(n), mu(n), P(9 * n);
    for (size_t p = 0; p < n; ++p) {
        const MaterialPoint& mp = pc.points[p];
        for (int i = 0; i < 9; ++i) Fc[9 * p + i] = (double)mp.F.d[i] + (double)mp.dFc.d[i];
        lam[p] = mp.lam;
        mu[p] = mp.mu;
    }
    check(dmpm_eval_pk1((int32_t)n, Fc.data(), lam.data(), mu.data(), P.data()), "dmpm_eval_pk1");
    for (size_t p = 0; p < n; ++p)
        for (int i = 0; i < 9; ++i) pc.points[p].P.d[i] = (float)P[9 * p + i];
}

void SingleParticle_op_1(MaterialPoint& mp) { mp.P = PK_FixedCorotatedElasticity(mp.F + mp.dFc, mp.lam, mp.mu); }

// ForwardSimulation.cpp:169-172 -> Grid.cpp:349-361
void G_Reset(Grid& grid) { grid.ResetValues(); }

// ForwardSimulation.cpp:11-51,174-182: accumulates m, p into the grid
void P2G(const PointCloud& pc, Grid& grid, float dt, float drag)
{
    const size_t n = pc.points.size();
    if (n == 0) return;
    const Staging s = stage(pc);
    std::vector<double> P(9 * n);
    for (size_t p = 0; p < n; ++p)
        for (int i = 0; i < 9; ++i) P[9 * p + i] = pc.points[p].P.d[i];
    GridArrays ga = grid_arrays(grid);
    const dmpm_grid_desc d = desc_of(grid);
    check(dmpm_phase_p2g(&d, (int32_t)n, s.x.data(), s.v.data(), s.C.data(), s.F.data(), s.dFc.data(), P.data(), s.m.data(),
                         s.vol.data(), dt, drag, ga.m.data(), ga.p.data()),
          "dmpm_phase_p2g");
    put_grid(grid, &ga.m, &ga.p, nullptr);
}

void SingleParticle_to_grid(const MaterialPoint& mp, Grid& grid, float dt, float drag) { P2G(cloud_of(mp), grid, dt, drag); }

// ForwardSimulation.cpp:53-58,184-200
void G_op(Grid& grid, float dt, Vec3 f_ext)
{
    GridArrays ga = grid_arrays(grid);
    const dmpm_grid_desc d = desc_of(grid);
    const double f[3] = {f_ext[0], f_ext[1], f_ext[2]};
    check(dmpm_phase_grid_op(&d, ga.m.data(), ga.p.data(), dt, f, ga.v.data()), "dmpm_phase_grid_op");
    put_grid(grid, nullptr, nullptr, &ga.v);
}

// ForwardSimulation.cpp:202-224
void G_op(Grid& grid, float dt, Vec3 gravity_point, float gravity_mag)
{
    GridArrays ga = grid_arrays(grid);
    const dmpm_grid_desc d = desc_of(grid);
    const double gp[3] = {gravity_point[0], gravity_point[1], gravity_point[2]};
    check(dmpm_phase_grid_op_point(&d, ga.m.data(), ga.p.data(), dt, gp, gravity_mag, ga.v.data()), "dmpm_phase_grid_op_point");
    put_grid(grid, nullptr, nullptr, &ga.v);
}

// ForwardSimulation.cpp:53-58 on one node (a one-node grid at the node's position)
void SingleNode_op(GridNode& node, float dt, Vec3 f_ext)
{
    dmpm_grid_desc d{};
    d.dim[0] = d.dim[1] = d.dim[2] = 1;
    d.dx = 1.0;
    for (int a = 0; a < 3; ++a) d.min_point[a] = node.x[a];
    double m = node.m, p[3], v[3];
    const double f[3] = {f_ext[0], f_ext[1], f_ext[2]};
    for (int a = 0; a < 3; ++a) {
        p[a] = node.p[a];
        v[a] = node.v[a];
    }
    check(dmpm_phase_grid_op(&d, &m, p, dt, f, v), "dmpm_phase_grid_op");
    for (int a = 0; a < 3; ++a) node.v[a] = (float)v[a];
}

// ForwardSimulation.cpp:60-106,226-236: next's v and C from the grid velocities
static void g2p_impl(PointCloud& next, const PointCloud& curr, const Grid& grid)
{
    const size_t n = curr.points.size();
    if (next.points.size() != n) next.points = curr.points;
    if (n == 0) return;
    std::vector<double> x(3 * n), vn(3 * n), Cn(9 * n);
    for (size_t p = 0; p < n; ++p)
        for (int a = 0; a < 3; ++a) x[3 * p + a] = curr.points[p].x[a];
    const GridArrays ga = grid_arrays(grid);
    const dmpm_grid_desc d = desc_of(grid);
    check(dmpm_phase_g2p(&d, (int32_t)n, x.data(), ga.v.data(), vn.data(), Cn.data()), "dmpm_phase_g2p");
    for (size_t p = 0; p < n; ++p) {
        MaterialPoint& mp = next.points[p];
        for (int a = 0; a < 3; ++a) mp.v[a] = (float)vn[3 * p + a];
        for (int i = 0; i < 9; ++i) mp.C.d[i] = (float)Cn[9 * p + i];
    }
}

void G2P(PointCloud& next, const PointCloud& curr, Grid& grid) { g2p_impl(next, curr, grid); }

void Grid_to_SingleParticle(MaterialPoint& next_mp, const MaterialPoint& curr_mp, const Grid& grid)
{
    PointCloud nx = cloud_of(next_mp);
    g2p_impl(nx, cloud_of(curr_mp), grid);
    next_mp.v = nx.points[0].v;
    next_mp.C = nx.points[0].C;
}

// ForwardSimulation.cpp:108-124,238-259: next's F and x from its v and C
void P_op_2(PointCloud& next, const PointCloud& curr, float smoothing_factor, float dt)
{
    const size_t n = curr.points.size();
    if (next.points.size() != n) throw std::invalid_argument("P_op_2: next and curr differ in size");
    if (n == 0) return;
    const Staging s = stage(curr);
    std::vector<double> vn(3 * n), Cn(9 * n), xn(3 * n), Fn(9 * n);
    for (size_t p = 0; p < n; ++p) {
        for (int a = 0; a < 3; ++a) vn[3 * p + a] = next.points[p].v[a];
        for (int i = 0; i < 9; ++i) Cn[9 * p + i] = next.points[p].C.d[i];
    }
    check(dmpm_phase_p_op2((int32_t)n, s.x.data(), s.F.data(), s.dFc.data(), vn.data(), Cn.data(), smoothing_factor, dt, xn.data(),
                           Fn.data()),
          "dmpm_phase_p_op2");
    for (size_t p = 0; p < n; ++p) {
        MaterialPoint& mp = next.points[p];
        for (int a = 0; a < 3; ++a) mp.x[a] = (float)xn[3 * p + a];
        for (int i = 0; i < 9; ++i) mp.F.d[i] = (float)Fn[9 * p + i];
    }
}

void P_op_2(PointCloud& next, const PointCloud& curr, float dt) { P_op_2(next, curr, 0.f, dt); }

void SingleParticle_op_2(MaterialPoint& next_mp, const MaterialPoint& curr_mp, float smoothing_factor, float dt)
{
    PointCloud nx = cloud_of(next_mp);
    P_op_2(nx, cloud_of(curr_mp), smoothing_factor, dt);
    next_mp.x = nx.points[0].x;
    next_mp.F = nx.points[0].F;
}

void SingleParticle_op_2(MaterialPoint& next_mp, const MaterialPoint& curr_mp, float dt) { SingleParticle_op_2(next_mp, curr_mp, 0.f, dt); }

// ForwardSimulation.cpp:114-116
void smooth_deformation_gradient(MaterialPoint& next_mp, const MaterialPoint& curr_mp, float smoothing_factor)
{
    next_mp.F = next_mp.F * (1.0f - smoothing_factor) + curr_mp.F * smoothing_factor;
}

// ForwardSimulation.cpp:292-327: grid.m <- sum w mp_m over the points (the
// other node fields are left as they are)
void P2G_Mass(const std::vector<Vec3> points, Grid& grid, float mp_m)
{
    const size_t n = points.size();
    std::vector<double> x(3 * n), m(n, (double)mp_m);
    for (size_t p = 0; p < n; ++p)
        for (int a = 0; a < 3; ++a) x[3 * p + a] = points[p][a];
    std::vector<double> gm((size_t)grid.dim_x * grid.dim_y * grid.dim_z, 0.0);
    const dmpm_grid_desc d = desc_of(grid);
    check(dmpm_phase_p2g_mass(&d, (int32_t)n, x.data(), m.data(), gm.data()), "dmpm_phase_p2g_mass");
    put_grid(grid, &gm, nullptr, nullptr);
}

// ForwardSimulation.cpp:329-367: grid_2.m <- masses of grid_1 gathered at grid_2's nodes
void G2G_Mass(Grid& grid_1, Grid& grid_2)
{
    const GridArrays g1 = grid_arrays(grid_1);
    std::vector<double> m2((size_t)grid_2.dim_x * grid_2.dim_y * grid_2.dim_z);
    const dmpm_grid_desc d1 = desc_of(grid_1), d2 = desc_of(grid_2);
    check(dmpm_phase_g2g_mass(&d1, g1.m.data(), &d2, m2.data()), "dmpm_phase_g2g_mass");
    put_grid(grid_2, &m2, nullptr, nullptr);
}

// grid.m <- P2G mass of the point cloud's own masses, the other node fields cleared
// (MPMPolyscope.cpp:156-160 in one call)
void P2G_PointCloudMass(const PointCloud& pc, Grid& grid)
{
    const size_t n = pc.points.size();
    std::vector<double> x(3 * n), m(n);
    for (size_t p = 0; p < n; ++p) {
        for (int a = 0; a < 3; ++a) x[3 * p + a] = pc.points[p].x[a];
        m[p] = pc.points[p].m;
    }
    grid.ResetValues();
    std::vector<double> gm((size_t)grid.dim_x * grid.dim_y * grid.dim_z, 0.0);
    const dmpm_grid_desc d = desc_of(grid);
    check(dmpm_phase_p2g_mass(&d, (int32_t)n, x.data(), m.data(), gm.data()), "dmpm_phase_p2g_mass");
    put_grid(grid, &gm, nullptr, nullptr);
}

}  // namespace SingleThreadMPM

namespace MultiThreadMPM {

// MultiThreadForwardSimulation.cpp:7-225: P_op_1, P2G, G_op, G2P and P_op_2
// without smoothing; next may be curr (the device keeps both layers)
void ForwardTimeStep(PointCloud& next, PointCloud& curr, Grid& main_grid, std::vector<std::shared_ptr<Grid>>& proxy_grids,
                     float dt, float drag, Vec3 f_ext)
{
    if (proxy_grids.empty()) throw std::invalid_argument("MultiThreadMPM::ForwardTimeStep: proxy_grids is empty");
    SingleThreadMPM::ForwardTimeStep(next, curr, main_grid, 0.f, dt, drag, f_ext);
}

}  // namespace MultiThreadMPM

// ------------------------------------------------------- BackPropagation ----
// BackPropagation.cpp:6-125: gradients of layer n from those of layer n+1.
// The step n -> n+1 is recomputed on the GPU from layer n (its grid is the
// adjoint's input); layer n+1's dLdx, dLdv, dLdF, dLdC seed the adjoint.
void Back_Timestep(CompGraphLayer& n1, CompGraphLayer& n0, float drag, float dt, float smoothing_factor)
{
    PointCloud& pc0 = *n0.point_cloud;
    const PointCloud& pc1 = *n1.point_cloud;
    const size_t n = pc0.points.size();
    Ctx ctx(make_config(n, *n0.grid, 2, dt, drag, smoothing_factor, Vec3(), false));
    upload_particles(ctx.c, pc0);
    const std::vector<double> ctrl = controls_of(pc0);
    check(dmpm_set_control(ctx.c, 0, ctrl.data()), "dmpm_set_control");
    check(dmpm_forward_step(ctx.c, 0), "dmpm_forward_step");
    std::vector<double> gx(3 * n), gv(3 * n), gF(9 * n), gC(9 * n);
    for (size_t p = 0; p < n; ++p) {
        const MaterialPoint& mp = pc1.points[p];
        for (int a = 0; a < 3; ++a) {
            gx[3 * p + a] = mp.dLdx[a];
            gv[3 * p + a] = mp.dLdv[a];
        }
        for (int i = 0; i < 9; ++i) {
            gF[9 * p + i] = mp.dLdF.d[i];
            gC[9 * p + i] = mp.dLdC.d[i];
        }
    }
    check(dmpm_set_layer_grads(ctx.c, 1, gx.data(), gv.data(), gF.data(), gC.data()), "dmpm_set_layer_grads");
    check(dmpm_backward_step(ctx.c, 0), "dmpm_backward_step");
    download_grads(ctx.c, 0, pc0);
}

// ------------------------------------------------------------- CompGraph ----
// CompGraph.cpp:269-278
CompGraph::CompGraph(std::shared_ptr<PointCloud> initial_point_cloud, std::shared_ptr<Grid> grid,
                     std::shared_ptr<const Grid> _target_grid)
{
    layers.clear();
    layers.resize(1);
    layers[0].point_cloud = std::make_shared<PointCloud>(*initial_point_cloud);
    layers[0].grid = std::make_shared<Grid>(*grid);
    target_grid = _target_grid;
}

CompGraph::~CompGraph() { ReleaseDevice(); }

void CompGraph::ReleaseDevice()
{
    if (ctx) dmpm_destroy(ctx);
    ctx = nullptr;
    ctx_layers = 0;
}

void CompGraph::SetSimulationParameters(float _dt, float _drag, Vec3 _f_ext, float _s)
{
    dt = _dt;
    drag = _drag;
    f_ext = _f_ext;
    smoothing_factor = _s;
}

// CompGraph.cpp:7-17: every layer a deep copy of layer 0
void CompGraph::SetUpCompGraph(int num_layers)
{
    if (num_layers <= 0) throw std::invalid_argument("SetUpCompGraph: num_layers must be > 0");
    layers.resize(num_layers);
    for (int i = 1; i < num_layers; ++i) {
        layers[i].point_cloud = std::make_shared<PointCloud>(*layers.front().point_cloud);
        layers[i].grid = std::make_shared<Grid>(*layers.front().grid);
    }
    ReleaseDevice();
}

// One context per layer count and parameter set; a new one starts from the host layer 0.
void CompGraph::EnsureDevice()
{
    const std::array<double, 6> prm = {dt, drag, smoothing_factor, f_ext[0], f_ext[1], f_ext[2]};
    if (ctx && ctx_layers == (int)layers.size() && prm == ctx_params) return;
    ReleaseDevice();
    if (layers.size() < 2) throw std::invalid_argument("CompGraph: SetUpCompGraph with >= 2 layers first");
    const PointCloud& pc0 = *layers.front().point_cloud;
    const dmpm_config cfg =
        make_config(pc0.points.size(), *layers.front().grid, (int)layers.size(), dt, drag, smoothing_factor, f_ext, true);
    check(dmpm_create(&cfg, &ctx), "dmpm_create");
    ctx_layers = (int)layers.size();
    ctx_params = prm;
    upload_particles(ctx, pc0);
    const std::vector<double> tm = grid_masses(*target_grid);
    check(dmpm_set_target_grid(ctx, tm.data()), "dmpm_set_target_grid");
    ctx_fresh = true;
}

// CompGraph.cpp:246-256.  The host layers' controls from start_layer on, and
// the state of start_layer, are written to the device first; the computed
// layers come back into the host point clouds.
void CompGraph::ComputeForwardPass(size_t start_layer)
{
    EnsureDevice();
    const int L = (int)layers.size();
    int start = (int)start_layer;
    if (ctx_fresh) {
        start = 0;  // the context holds layer 0 only
    } else if (start > 0) {
        const PointCloud& pc = *layers[start].point_cloud;
        Staging s = stage(pc);
        check(dmpm_set_layer_state(ctx, start, s.x.data(), s.v.data(), s.F.data(), s.C.data()), "dmpm_set_layer_state");
    } else {
        const PointCloud& pc = *layers[0].point_cloud;
        Staging s = stage(pc);
        check(dmpm_set_layer_state(ctx, 0, s.x.data(), s.v.data(), s.F.data(), s.C.data()), "dmpm_set_layer_state");
    }
    for (int i = start; i < L - 1; ++i) {
        const std::vector<double> c = controls_of(*layers[i].point_cloud);
        check(dmpm_set_control(ctx, i, c.data()), "dmpm_set_control");
    }
    check(dmpm_forward_pass(ctx, start), "dmpm_forward_pass");
    ctx_fresh = false;
    for (int i = start + 1; i < L; ++i) download_layer(ctx, i, *layers[i].point_cloud);
}

// CompGraph.cpp:110-244
float CompGraph::EndLayerMassLoss()
{
    EnsureDevice();
    if (ctx_fresh) ComputeForwardPass(0);
    double loss = 0.0;
    check(dmpm_end_layer_mass_loss(ctx, &loss), "dmpm_end_layer_mass_loss");
    download_grads(ctx, (int)layers.size() - 1, *layers.back().point_cloud);
    return (float)loss;
}

// CompGraph.cpp:258-267: gradients of layer control_layer (all four) and the
// control gradient dLdF of every layer above it.
void CompGraph::ComputeBackwardPass(size_t control_layer)
{
    EnsureDevice();
    const int L = (int)layers.size();
    const int c = (int)control_layer;
    check(dmpm_backward_pass(ctx, c), "dmpm_backward_pass");
    const size_t n = layers[0].point_cloud->points.size();
    std::vector<double> g(9 * n);
    for (int i = L - 2; i > c; --i) {
        PointCloud& pc = *layers[i].point_cloud;
        pc.ResetGradients();
        check(dmpm_get_control_grad(ctx, i, g.data()), "dmpm_get_control_grad");
        for (size_t p = 0; p < n; ++p)
            for (int k = 0; k < 9; ++k) pc.points[p].dLdF.d[k] = (float)g[9 * p + k];
    }
    layers[c].point_cloud->ResetGradients();
    download_grads(ctx, c, *layers[c].point_cloud);
}

// CompGraph.cpp:280-454: control-stride sweeps of gradient descent with a
// line search (Adam then Lion on the same momentum; the "undo" applies the
// negated step; the gradient norm is read from layer 0).
void CompGraph::OptimizeDefGradControlSequence(int num_steps, float _dt, float _drag, Vec3 _f_ext, int control_stride,
                                               int max_gd_iters, int max_line_search_iters, float initial_alpha,
                                               float gd_tol, float _smoothing_factor)
{
    SetSimulationParameters(_dt, _drag, _f_ext, _smoothing_factor);
    SetUpCompGraph(num_steps);
    ComputeForwardPass(0);
    const float initial_loss = EndLayerMassLoss();
    ComputeBackwardPass(0);
    const float initial_norm = layers.front().point_cloud->Compute_dLdF_Norm();
    const int totalTemporalIterations = 5;
    const float beta1 = 0.9f, beta2 = 0.999f, epsilon = 1e-5f;
    float gradient_norm = 0.f;
    for (int t = 0; t < totalTemporalIterations; ++t) {
        for (int cf = 0; cf < (int)layers.size() - 1; cf += control_stride) {
            float alpha = initial_alpha;
            for (int gd = 0; gd < max_gd_iters; ++gd) {
                ComputeForwardPass(cf);
                const float gd_loss = EndLayerMassLoss();
                ComputeBackwardPass(cf);
                gradient_norm = layers.front().point_cloud->Compute_dLdF_Norm();
                gradientNormHistory.push_back(gradient_norm);
                if (gradient_norm < gd_tol * initial_norm) break;
                bool found = false;
                for (int ls = 0; ls < max_line_search_iters; ++ls) {
                    layers[cf].point_cloud->Descend_Adam(alpha, gradient_norm, beta1, beta2, epsilon, timestep);
                    layers[cf].point_cloud->Descend_Lion(alpha, beta1, gradient_norm);
                    ComputeForwardPass(cf);
                    const float ls_loss = EndLayerMassLoss();
                    if (ls_loss < gd_loss) {
                        found = true;
                        break;
                    }
                    layers[cf].point_cloud->Descend_Adam(-alpha, gradient_norm, beta1, beta2, epsilon, timestep);
                    layers[cf].point_cloud->Descend_Lion(-alpha, beta1, gradient_norm);
                    alpha /= 2.f;
                }
                if (!found) break;
                timestep += 1;
            }
        }
    }
    const float final_loss = EndLayerMassLoss();
    std::cout << "Initial loss = " << initial_loss << "\nFinal loss = " << final_loss << std::endl;
}

// CompGraph.cpp:457-504 with a usable step: central differences of the loss in
// dFc(i,j) of one particle of layer 0 (delta 1e-4; the reference's 1e-9 is
// below float resolution).
void CompGraph::FiniteDifferencesGradientTest(int num_steps, size_t particle_id)
{
    SetUpCompGraph(num_steps);
    ComputeForwardPass(0);
    const float initial_loss = EndLayerMassLoss();
    ComputeBackwardPass(0);
    MaterialPoint& mp = layers.front().point_cloud->points[particle_id];
    const Mat3 analytic = mp.dLdF;
    Mat3 fd;
    const float delta = 1e-4f;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const float orig = mp.dFc(i, j);
            mp.dFc(i, j) = orig + delta;
            ComputeForwardPass(0);
            const float l1 = EndLayerMassLoss();
            mp.dFc(i, j) = orig - delta;
            ComputeForwardPass(0);
            const float l2 = EndLayerMassLoss();
            fd(i, j) = (l1 - l2) / (2.f * delta);
            mp.dFc(i, j) = orig;
        }
    const float diff = (fd - analytic).norm();
    std::cout << "initial loss = " << initial_loss << "\ngradient difference = " << diff
              << "\ngradient difference percentage = " << 100.f * diff / fd.norm() << std::endl;
}

}  // namespace DiffMPMLib3D
