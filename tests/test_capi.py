"""C-ABI library checks that need no GPU: it loads and exports every symbol
declared in include/dmpm.h; the Python binding covers all of them; the C++
drop-in headers compile and link with the reference's signatures."""
import ctypes
import os
import re

from conftest import PKG_DIR, ROOT, load_pkg


def _declared():
    text = open(os.path.join(ROOT, "include", "dmpm.h")).read()
    return sorted(set(re.findall(r"\b(dmpm_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_exported():
    lib = ctypes.CDLL(os.path.join(PKG_DIR, "lib", "libdmpm.so"))
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    pkg = load_pkg()
    assert sorted(pkg.EXPORTS) == _declared()


def test_config_layout_matches_header():
    pkg = load_pkg()
    lib = pkg.load_library()
    cfg = pkg.Config()
    assert lib.dmpm_config_init(ctypes.byref(cfg)) == 0
    assert cfg.dt == 1.0 / 120.0 and cfg.drag == 0.5 and cfg.smoothing == 0.1
    assert cfg.num_layers == 2 and cfg.dx == 1.0
    assert ctypes.sizeof(pkg.Config) == 4 * 6 + 8 * 10 + 4 * 6  # int32 x6, double x10, int32 x6


def test_create_without_gpu_fails_cleanly():
    """No GPU in this container: dmpm_create returns an error, it does not fall back to the CPU."""
    pkg = load_pkg()
    import torch
    if torch.cuda.is_available():
        return
    try:
        pkg.Morph(10, (4, 4, 4), 0.25, 3)
    except pkg.DmpmError as e:
        assert "dmpm error" in str(e)
    else:
        raise AssertionError("context created without a GPU")


REFERENCE_HEADERS = ["BackPropagation.h", "CompGraph.h", "Elasticity.h", "ForwardSimulation.h", "Grid.h", "GridNode.h",
                     "MaterialPoint.h", "MultiThreadForwardSimulation.h", "PointCloud.h", "pch.h"]


def test_dropin_headers_compile_and_link(tmp_path):
    """Every header name of the reference's DiffMPMLib3D that callers include compiles; every
    function the reference declares for the path (ForwardSimulation.h:10-43 with the phase
    API, BackPropagation.h:7, Elasticity.h:8-23 without the Tensor3x3x3x3 test aids, the
    CompGraph and PointCloud members) is taken as a pointer of its exact reference type and
    resolves in libdmpm_cpp.so; and the reference callers of the phase API compile as written
    (MPMPolyscope.cpp:158-160, MPMPointCloudVisualization.cpp:11-12, DiffMPMAnimator3D.cpp:501-505,
    CompGraph.cpp:134-135)."""
    import subprocess

    inc = os.path.join(ROOT, "include")
    lib = os.path.join(PKG_DIR, "lib")
    src = tmp_path / "t.cpp"
    src.write_text("".join(f'#include "{h}"\n' for h in REFERENCE_HEADERS) + """
#include <memory>
#include <vector>
using namespace DiffMPMLib3D;
template <typename T> bool ok(T p) { return p != nullptr; }
int main()
{
    namespace S = SingleThreadMPM;
    bool all = true;
    // ForwardSimulation.h
    all &= ok<void (*)(MaterialPoint&)>(&S::SingleParticle_op_1);
    all &= ok<void (*)(const MaterialPoint&, Grid&, float, float)>(&S::SingleParticle_to_grid);
    all &= ok<void (*)(GridNode&, float, Vec3)>(&S::SingleNode_op);
    all &= ok<void (*)(MaterialPoint&, const MaterialPoint&, const Grid&)>(&S::Grid_to_SingleParticle);
    all &= ok<void (*)(MaterialPoint&, const MaterialPoint&, float)>(&S::SingleParticle_op_2);
    all &= ok<void (*)(MaterialPoint&, const MaterialPoint&, float, float)>(&S::SingleParticle_op_2);
    all &= ok<void (*)(PointCloud&, PointCloud&, Grid&, float, float, Vec3)>(&S::ForwardTimeStep);
    all &= ok<void (*)(PointCloud&, PointCloud&, Grid&, float, float, float, Vec3)>(&S::ForwardTimeStep);
    all &= ok<void (*)(MaterialPoint&, const MaterialPoint&, float)>(&S::smooth_deformation_gradient);
    all &= ok<void (*)(PointCloud&)>(&S::P_op_1);
    all &= ok<void (*)(Grid&)>(&S::G_Reset);
    all &= ok<void (*)(const PointCloud&, Grid&, float, float)>(&S::P2G);
    all &= ok<void (*)(Grid&, float, Vec3)>(&S::G_op);
    all &= ok<void (*)(Grid&, float, Vec3, float)>(&S::G_op);
    all &= ok<void (*)(PointCloud&, const PointCloud&, Grid&)>(&S::G2P);
    all &= ok<void (*)(PointCloud&, const PointCloud&, float)>(&S::P_op_2);
    all &= ok<void (*)(PointCloud&, const PointCloud&, float, float)>(&S::P_op_2);
    all &= ok<void (*)(PointCloud&, Grid&)>(&S::CalculatePointCloudVolumes);
    all &= ok<void (*)(const std::vector<Vec3>, Grid&, float)>(&S::P2G_Mass);
    all &= ok<void (*)(Grid&, Grid&)>(&S::G2G_Mass);
    all &= ok<void (*)(PointCloud&, PointCloud&, Grid&, std::vector<std::shared_ptr<Grid>>&, float, float, Vec3)>(
        &MultiThreadMPM::ForwardTimeStep);
    // BackPropagation.h, Elasticity.h
    all &= ok<void (*)(CompGraphLayer&, CompGraphLayer&, float, float, float)>(&Back_Timestep);
    all &= ok<float (*)(const Mat3&, float, float)>(&FixedCorotatedElasticity);
    all &= ok<Mat3 (*)(const Mat3&, float, float)>(&PK_FixedCorotatedElasticity);
    all &= ok<Mat3 (*)(const Mat3&, float, float, const Mat3&)>(&d2_FCE_psi_dF2_mult_by_dF);
    all &= ok<void (*)(float, float, float&, float&)>(&CalculateLameParameters);
    // CompGraph.h members
    all &= ok<void (CompGraph::*)(int)>(&CompGraph::SetUpCompGraph);
    all &= ok<float (CompGraph::*)()>(&CompGraph::EndLayerMassLoss);
    all &= ok<void (CompGraph::*)(size_t)>(&CompGraph::ComputeForwardPass);
    all &= ok<void (CompGraph::*)(size_t)>(&CompGraph::ComputeBackwardPass);
    all &= ok<void (CompGraph::*)(int, float, float, Vec3, int, int, int, float, float, float)>(
        &CompGraph::OptimizeDefGradControlSequence);
    all &= ok<void (CompGraph::*)(int, size_t)>(&CompGraph::FiniteDifferencesGradientTest);
    // PointCloud.h members on the path (optimisers, norms, .mpmbin)
    all &= ok<void (PointCloud::*)()>(&PointCloud::ResetGradients);
    all &= ok<float (PointCloud::*)()>(&PointCloud::Compute_dLdF_Norm);
    all &= ok<void (PointCloud::*)(float, float)>(&PointCloud::Descend_dLdF);
    all &= ok<void (PointCloud::*)(float, float, float, float, float, int)>(&PointCloud::Descend_Adam);
    all &= ok<void (PointCloud::*)(float, float, float, float, float, int)>(&PointCloud::Descend_AMSGrad);
    all &= ok<void (PointCloud::*)(float, float, float)>(&PointCloud::Descend_Lion);
    all &= ok<void (PointCloud::*)(float, float)>(&PointCloud::Descend_GradientDescent);
    all &= ok<std::vector<int> (PointCloud::*)(int, int)>(&PointCloud::generateMinibatchIndices);
    all &= ok<float (PointCloud::*)(const std::vector<int>&)>(&PointCloud::Compute_dLdF_Norm_Stochastic);
    all &= ok<void (PointCloud::*)(float, float, float, float, float, int, const std::vector<int>&)>(
        &PointCloud::Descend_Adam_Stochastic);
    all &= ok<void (PointCloud::*)(size_t)>(&PointCloud::RemovePoint);
    all &= ok<void (PointCloud::*)(std::string)>(&PointCloud::WriteEntirePointCloudToBinaryFile);
    all &= ok<bool (PointCloud::*)(std::string)>(&PointCloud::ReadEntirePointCloudFromBinaryFile);
    all &= ok<bool (PointCloud::*)(const PointCloud&)>(&PointCloud::IsEqualToOtherPointCloud);
    all &= ok<void (Grid::*)()>(&Grid::ResetValues);
    all &= ok<void (Grid::*)()>(&Grid::ResetGradients);
    all &= ok<std::vector<std::reference_wrapper<GridNode>> (Grid::*)(Vec3, std::vector<std::array<int, 3>>*)>(
        &Grid::QueryPoint_CubicBSpline);
    all &= ok<std::vector<std::reference_wrapper<const GridNode>> (Grid::*)(Vec3, std::vector<std::array<int, 3>>*) const>(
        &Grid::QueryPointConst_CubicBSpline);
    if (!all) return 1;

    // Grid.cpp:40-81 on the host: an interior point sees 4^3 nodes from floor - 1, x-major;
    // a point in the corner cell only the in-grid 3^3; a far point none
    Grid q(8, 8, 8, 0.125f, Vec3(0.f, 0.f, 0.f));
    std::vector<std::array<int, 3>> idx;
    auto nodes = q.QueryPoint_CubicBSpline(Vec3(0.52f, 0.30f, 0.40f), &idx);
    if (nodes.size() != 64 || idx.size() != 64) return 2;
    if (idx[0] != std::array<int, 3>{3, 1, 2} || idx[1] != std::array<int, 3>{3, 1, 3} || idx[63] != std::array<int, 3>{6, 4, 5}) return 3;
    if (&nodes[5].get() != &q.nodes[idx[5][0]][idx[5][1]][idx[5][2]]) return 4;
    if (q.QueryPointConst_CubicBSpline(Vec3(0.01f, 0.01f, 0.01f)).size() != 27) return 5;
    if (!q.QueryPoint_CubicBSpline(Vec3(1e30f, 0.5f, 0.5f)).empty()) return 6;

    // the reference's callers of the phase API, as written there (not run: no GPU here)
    volatile bool run = false;
    if (run) {
        auto grid = std::make_shared<Grid>(8, 8, 8, 0.125f, Vec3(0.f, 0.f, 0.f));
        auto target_point_cloud = std::make_shared<PointCloud>();
        auto target_grid = std::make_shared<Grid>(*grid);              // MPMPolyscope.cpp:158-160
        target_grid->ResetValues();
        SingleThreadMPM::P2G(*target_point_cloud, *target_grid, 0.0, 0.0);
        PointCloud mpm_point_cloud;
        Grid g(*grid);
        g.ResetValues();                                               // MPMPointCloudVisualization.cpp:11-12
        SingleThreadMPM::P2G(mpm_point_cloud, g, 0.0, 0.0);
        std::vector<Vec3> _points = mpm_point_cloud.GetPointPositions();
        auto mpm_grid = std::make_shared<Grid>(*grid);
        auto high_res_mpm_grid = std::make_shared<Grid>(16, 16, 16, 0.0625f, Vec3(0.f, 0.f, 0.f));
        SingleThreadMPM::P2G_Mass(_points, *mpm_grid, 1.0);            // DiffMPMAnimator3D.cpp:501-505
        SingleThreadMPM::G2G_Mass(*mpm_grid, *high_res_mpm_grid);
        PointCloud& point_cloud = mpm_point_cloud;                       // CompGraph.cpp:125-135
        Grid& grid_ = *grid;
        SingleThreadMPM::G_Reset(grid_);
        SingleThreadMPM::P2G(point_cloud, grid_, 0.f, 0.f);
    }
    return 0;
}
""")
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-I", os.path.join(inc, "DiffMPMLib3D"), "-I", inc, str(src), "-L", lib,
                    "-ldmpm_cpp", "-ldmpm", f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True, timeout=300)
    subprocess.run([str(exe)], check=True, timeout=60)


def _shell_morph(pkg, cls, n=4, G=8, B=None):
    """A Morph whose library handle is never reached: the length checks of the binding
    raise before any C call, so these run without a GPU."""
    mo = object.__new__(cls)
    mo.n, mo.G, mo._h, mo._L = n, G, None, None
    if B is not None:
        mo.B = B
    return mo


def test_binding_rejects_wrong_lengths():
    """Arrays whose length differs from what the C side reads (n*3, n*9, n, G, len(m)*3)
    raise ValueError instead of handing a short buffer to the library."""
    import numpy as np
    import pytest

    pkg = load_pkg()
    mo = _shell_morph(pkg, pkg.Morph)
    n = 4
    good = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), F=np.zeros((n, 3, 3)), C=np.zeros((n, 3, 3)), m=np.ones(n),
                vol=np.ones(n), lam=np.ones(n), mu=np.ones(n))
    for k, short in [("x", np.zeros((n - 1, 3))), ("v", np.zeros(5)), ("F", np.zeros((n, 3))), ("C", np.zeros((n, 9, 2))),
                     ("m", np.ones(n + 1)), ("vol", np.ones(2)), ("lam", np.ones(1)), ("mu", np.ones(n * 3))]:
        with pytest.raises(ValueError, match=k):
            mo.set_particles(**{**good, k: short})
    with pytest.raises(ValueError, match="dFc"):
        mo.set_control(0, np.zeros((n, 3)))
    with pytest.raises(ValueError, match="target mass"):
        mo.set_target_grid(np.zeros(7))
    with pytest.raises(ValueError, match="x"):
        mo.target_from_points(np.zeros((3, 3)), np.ones(4))
    with pytest.raises(ValueError, match="F"):
        mo.set_layer_state(1, F=np.zeros((n - 1, 3, 3)))
    with pytest.raises(ValueError, match="dLdv"):
        mo.set_layer_grads(1, dLdv=np.zeros((n, 2)))
    with pytest.raises(ValueError, match="dF"):
        pkg.eval_hessian(np.zeros((5, 3, 3)), 1.0, 1.0, np.zeros((4, 3, 3)))
    mb = _shell_morph(pkg, pkg.MorphBatch, n=6, G=16, B=2)
    with pytest.raises(ValueError, match="x"):
        mb.target_from_points([np.zeros((2, 3)), np.zeros((2, 3))], [np.ones(2), np.ones(3)])
    with pytest.raises(ValueError, match="instances"):
        mb.target_from_points([np.zeros((2, 3))], [np.ones(2)])
