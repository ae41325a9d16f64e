// C++ host mirror of the reference's MLS-MPM API over the C ABI.
//
// Containers (MaterialPoint, PointCloud, Grid, CompGraph layers) keep host
// data as the reference does; the step, the passes, the loss, the adjoint and
// the constitutive model run on the GPU through include/dmpm.h.  Reference
// files cited per function are relative to the reference root.
#include "DiffMPMLib3D/DiffMPMLib3D.h"

#include "dmpm.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

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

// MPMPolyscope.cpp:156-160: grid.m <- P2G mass of the point cloud
void P2G_Mass(const PointCloud& pc, Grid& grid)
{
    const size_t n = pc.points.size();
    Ctx ctx(make_config(1, grid, 2, 0.f, 0.f, 0.f, Vec3(), false));
    std::vector<double> x(3 * n), m(n);
    for (size_t p = 0; p < n; ++p) {
        for (int a = 0; a < 3; ++a) x[3 * p + a] = pc.points[p].x[a];
        m[p] = pc.points[p].m;
    }
    check(dmpm_target_from_points(ctx.c, x.data(), m.data(), (int32_t)n), "dmpm_target_from_points");
    std::vector<double> gm((size_t)grid.dim_x * grid.dim_y * grid.dim_z);
    check(dmpm_get_target_grid(ctx.c, gm.data()), "dmpm_get_target_grid");
    grid.ResetValues();
    size_t k = 0;
    for (int i = 0; i < grid.dim_x; ++i)
        for (int j = 0; j < grid.dim_y; ++j)
            for (int l = 0; l < grid.dim_z; ++l) grid.nodes[i][j][l].m = (float)gm[k++];
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
