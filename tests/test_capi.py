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
    """Every header name of the reference's DiffMPMLib3D that callers include compiles, and the
    free functions of the step keep the reference's signatures (taken as function pointers of
    those exact types) and resolve in libdmpm_cpp.so."""
    import subprocess

    inc = os.path.join(ROOT, "include")
    lib = os.path.join(PKG_DIR, "lib")
    src = tmp_path / "t.cpp"
    src.write_text("".join(f'#include "{h}"\n' for h in REFERENCE_HEADERS) + """
using namespace DiffMPMLib3D;
int main()
{
    void (*mt)(PointCloud&, PointCloud&, Grid&, std::vector<std::shared_ptr<Grid>>&, float, float, Vec3) =
        &MultiThreadMPM::ForwardTimeStep;
    void (*st)(PointCloud&, PointCloud&, Grid&, float, float, float, Vec3) = &SingleThreadMPM::ForwardTimeStep;
    void (*st0)(PointCloud&, PointCloud&, Grid&, float, float, Vec3) = &SingleThreadMPM::ForwardTimeStep;
    void (*bt)(CompGraphLayer&, CompGraphLayer&, float, float, float) = &Back_Timestep;
    Mat3 (*pk)(const Mat3&, float, float) = &PK_FixedCorotatedElasticity;
    return (mt && st && st0 && bt && pk) ? 0 : 1;
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
