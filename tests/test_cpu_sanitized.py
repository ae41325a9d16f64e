"""The FP64 kernels and the runtime built for the host under AddressSanitizer and
UndefinedBehaviorSanitizer (``tests/cpu``): every launch runs its work-items one
after another on the host, device buffers are heap blocks, so an out-of-bounds
access of any kernel and any signed overflow of the index arithmetic stop the
run (``-fno-sanitize-recover``).  Three drivers run in a subprocess with the
sanitizer runtimes preloaded:

* the sphere episode, against the FP64 oracle (1e-12 loss, 1e-10 gradient: the
  host build is the FP64 path's arithmetic run serially);
* diverged particles (1e12 cells out, +-inf, NaN) beside the sphere, whose stencil
  bases must be clamped before their integer conversion;
* the C++ mirror's selftest: every mirror call of its ``run`` mode (passes,
  loss, Back_Timestep, both ForwardTimeStep forms, getters), then
  ``OptimizeDefGradControlSequence`` on the same case its GPU run once faulted
  on (its trial episodes diverge).
"""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

CPU = os.path.join(ROOT, "tests", "cpu")
BUILD = os.path.join(CPU, "_build")
CSRC = os.path.join(ROOT, "a-differentiable-material-point-method-for-shape-morphing_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(BUILD, "libdmpm_host_san.so")
CPPLIB = os.path.join(BUILD, "libdmpm_cpp_host_san.so")
SELFTEST = os.path.join(BUILD, "compgraph_selftest_san")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
FLAGS = ["-std=c++17", "-O1", "-g", *SAN]


def _stale(out, srcs):
    return not os.path.exists(out) or any(os.path.getmtime(s) > os.path.getmtime(out) for s in srcs)


def _build():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    os.makedirs(BUILD, exist_ok=True)
    rt = os.path.join(CSRC, "dmpm_runtime.hip")
    srcs = [rt, os.path.join(CPU, "host_kernels.cpp"), os.path.join(CPU, "hip", "hip_runtime.h"),
            os.path.join(CSRC, "dmpm_kernels.hip"), os.path.join(CSRC, "dmpm_phases.hip"), os.path.join(CSRC, "dmpm_kernels.h"),
            os.path.join(CSRC, "dmpm_math.h"), os.path.join(INCLUDE, "dmpm.h")]
    if _stale(LIB, srcs):
        subprocess.run([gxx, *FLAGS, "-fPIC", "-shared", "-I", CPU, "-I", CSRC, "-I", INCLUDE, "-x", "c++", rt, "-x", "none",
                        os.path.join(CPU, "host_kernels.cpp"), "-o", LIB], check=True, timeout=600)
    host = os.path.join(ROOT, "a-differentiable-material-point-method-for-shape-morphing_amd", "host", "dmpm_host.cpp")
    if _stale(CPPLIB, [host, LIB, os.path.join(INCLUDE, "DiffMPMLib3D", "DiffMPMLib3D.h")]):
        subprocess.run([gxx, *FLAGS, "-fPIC", "-shared", "-I", INCLUDE, host, "-L", BUILD, "-ldmpm_host_san",
                        f"-Wl,-rpath,{BUILD}", "-o", CPPLIB], check=True, timeout=600)
    st = os.path.join(ROOT, "tests", "cpp", "compgraph_selftest.cpp")
    if _stale(SELFTEST, [st, CPPLIB]):
        subprocess.run([gxx, *FLAGS, "-I", INCLUDE, st, "-L", BUILD, "-ldmpm_cpp_host_san", "-ldmpm_host_san",
                        f"-Wl,-rpath,{BUILD}", "-o", SELFTEST], check=True, timeout=600)


def _env():
    pre = []
    for lib in ("libasan.so", "libubsan.so"):
        p = subprocess.run(["g++", f"-print-file-name={lib}"], capture_output=True, text=True).stdout.strip()
        if not os.path.isabs(p):
            pytest.skip(f"{lib} not found")
        pre.append(p)
    env = dict(os.environ)
    env["LD_PRELOAD"] = ":".join(pre)
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    return env


DRIVER = r"""
import sys
sys.path.insert(0, {tests!r}); sys.path.insert(0, {root!r})
import numpy as np
from conftest import load_pkg
pkg = load_pkg()
pkg.load_library({lib!r})
from _scenes import gpu_morph, make_case, orc, rel
import test_gpu_parity as t
case = make_case(L=3)
mo = gpu_morph(case)
mo.forward_pass(0)
loss = mo.end_layer_mass_loss()
mo.backward_pass(0)
ref_loss, layers, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)
assert abs(loss - ref_loss) <= 1e-12 * abs(ref_loss), (loss, ref_loss)
assert rel(mo.control_grad(0), gd[0]) < 1e-10
mo.close()
# the checkpoint tape (segments of 3 steps in two alternating tapes, the segment below
# recomputed while the one above steps back), control layer 4: per-layer gradients
case = make_case(L=11, per_layer_ctrl=True)
mo = gpu_morph(case, tape=pkg.TAPE_CHECKPOINT, ckpt_stride=3)
mo.forward_pass(0)
loss = mo.end_layer_mass_loss()
mo.backward_pass(4)
ref_loss, layers, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L, control_layer=4)
assert abs(loss - ref_loss) <= 1e-12 * abs(ref_loss), (loss, ref_loss)
for i in range(4, case.L - 1):
    assert rel(mo.control_grad(i), gd[i]) < 1e-10, i
x5 = mo.layer_state(5)[0]
assert rel(x5, layers[5].x) < 1e-12
mo.backward_pass(4)  # again, from the tapes the first pass left
assert rel(mo.control_grad(4), gd[4]) < 1e-10
mo.close()
t.test_diverged_particles_touch_no_node(pkg.PREC_F64)
print("host build ok")
"""


def test_fp64_kernels_under_sanitizers(tmp_path):
    _build()
    env = _env()
    code = DRIVER.format(tests=os.path.join(ROOT, "tests"), root=ROOT, lib=LIB)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "runtime error" not in r.stderr, r.stdout[-2000:] + r.stderr[-4000:]
    assert "host build ok" in r.stdout
    # the mirror's calls, then its optimisation loop on the case its GPU run once faulted on
    import test_cpp_mirror
    inputs = test_cpp_mirror._write_case(tmp_path)
    for mode in ("run", "optimize"):
        r = subprocess.run([SELFTEST, mode, str(tmp_path)], env={**env, "LD_PRELOAD": ""}, capture_output=True, text=True,
                           timeout=900)
        assert r.returncode == 0 and "runtime error" not in r.stderr, mode + r.stdout[-2000:] + r.stderr[-4000:]
        if mode == "run":  # the same checks as the GPU test of the mirror
            test_cpp_mirror.check_run_outputs(tmp_path, *inputs)
    assert "Final loss" in r.stdout


NOMEM_DRIVER = r"""
import os, sys
sys.path.insert(0, {tests!r}); sys.path.insert(0, {root!r})
import numpy as np
from conftest import load_pkg
pkg = load_pkg()
pkg.load_library({lib!r})
from _scenes import gpu_morph, make_case
case = make_case(L=3)
mo = gpu_morph(case)
mo.forward_pass(0)
l0 = mo.end_layer_mass_loss()
os.environ["DMPM_HOST_MALLOC_LIMIT"] = "100000"
for call in (lambda: mo.target_from_points(np.full((20000, 3), 0.5), np.ones(20000)),
             lambda: pkg.eval_pk1(np.broadcast_to(np.eye(3), (2000, 3, 3)), 1.0, 1.0)):
    try:
        call()
    except pkg.DmpmError as e:
        assert "dmpm error -4:" in str(e) and "out of memory" in str(e), str(e)
    else:
        raise AssertionError("allocation failure not reported")
del os.environ["DMPM_HOST_MALLOC_LIMIT"]
# no stale HIP error reaches the next call, and the failed target call left the old target
mo.forward_pass(0)
l1 = mo.end_layer_mass_loss()
assert l1 == l0, (l0, l1)
assert pkg.eval_pk1(np.eye(3)[None], 1.0, 1.0).shape == (1, 3, 3)
print("nomem ok")
"""


def test_allocation_failures_report_nomem():
    """A failed device allocation (the host stand-in's hipMalloc refuses requests above
    DMPM_HOST_MALLOC_LIMIT and leaves the error set, as HIP does) returns DMPM_ERR_NOMEM
    from the target-point staging and the batched evaluators, leaves the previous target
    grid in place, and no stale error fails the next forward pass."""
    _build()
    code = NOMEM_DRIVER.format(tests=os.path.join(ROOT, "tests"), root=ROOT, lib=LIB)
    r = subprocess.run([sys.executable, "-c", code], env=_env(), capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "runtime error" not in r.stderr, r.stdout[-2000:] + r.stderr[-4000:]
    assert "nomem ok" in r.stdout
