"""The C++ layer (include/hymi.hpp) compiles against include/hymi.h with g++, links against libhymi.so and runs
its host-only checks (no GPU): mesh counts, owned-DoF count, a stencil row sum, a level error as an exception.
Every class and method of the layer is compiled and linked (the device branch of the program names them all)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_layer_compiles_links_and_runs(tmp_path):
    import hyteg_amd
    lib_dir = os.path.dirname(hyteg_amd._lib.LIB_PATH)
    assert os.path.exists(hyteg_amd._lib.LIB_PATH)
    exe = str(tmp_path / "hymi_cpp_check")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "hymi_cpp_check.cpp"), "-o", exe, "-L", lib_dir, "-lhymi",
           f"-Wl,-rpath,{lib_dir}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    assert "48 cells" in run.stdout


import pytest  # noqa: E402


@pytest.mark.gpu
def test_cpp_layer_device_run(tmp_path):
    """The same program on the GPU: 5 graph-replayed V(3,3) cycles through hymi::GeometricMultigridSolver reduce
    the residual below 1e-2 of the right-hand side, and the refined FP32 solve reaches 1e-10."""
    import hyteg_amd
    lib_dir = os.path.dirname(hyteg_amd._lib.LIB_PATH)
    exe = str(tmp_path / "hymi_cpp_check")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "hymi_cpp_check.cpp"), "-o", exe, "-L", lib_dir, "-lhymi",
           f"-Wl,-rpath,{lib_dir}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    print("[cpp layer] device run", flush=True)
    run = subprocess.run([exe, "0"], capture_output=True, text=True, timeout=120)
    print(run.stdout, flush=True)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
