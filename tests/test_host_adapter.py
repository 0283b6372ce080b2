"""GLAC's Action / Flow / Correlator / System interfaces (glac_amd/host/
glac_gpu.hpp) driven from C++ the way GLAC's System drives them, with a
stand-in of GLAC's Lattice<SU3> type (tests/host_adapter_check.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_bin", "host_adapter_check")
SRC = os.path.join(ROOT, "tests", "host_adapter_check.cpp")


def build_check():
    """g++ build of the check program against the in-tree libglacx.so and
    liboracle.so (rpath relative to the binary, so the tree can move)."""
    deps = [SRC, os.path.join(ROOT, "glac_amd", "host", "glac_gpu.hpp"), os.path.join(ROOT, "include", "glacx.h")]
    if os.path.exists(EXE) and all(os.path.getmtime(EXE) >= os.path.getmtime(d) for d in deps):
        return EXE
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    import oracle
    oracle.lib()  # builds liboracle.so if missing
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", SRC, "-o", EXE,
                    os.path.join(ROOT, "glac_amd", "libglacx.so"), os.path.join(ROOT, "oracle", "liboracle.so"),
                    "-Wl,-rpath,$ORIGIN/../../glac_amd:$ORIGIN/../../oracle"], check=True)
    return EXE


def test_delta_action_through_adapter():
    exe = build_check()
    out = subprocess.run([exe, "cpu"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "PASS dS = -36 on the unit field" in out.stdout
    assert "PASS dS = -36 through GLAC's Action interface" in out.stdout


@pytest.mark.gpu
def test_flow_measure_metropolis_through_adapter(tmp_path):
    exe = build_check()
    print("\n[gpu case] host adapter flow/measure/metropolis/System loop 4x4x4x8 fp64", flush=True)
    out = subprocess.run([exe, "gpu", str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("PASS") == 6 and "PASS System run loop" in out.stdout
