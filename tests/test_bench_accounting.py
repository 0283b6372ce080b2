"""bench.py's algorithmic-byte accounting, the numerator of its roofline
figure: per RK3 step each of the three stage launches reads and writes 4V
links, and the su(3) accumulator is written (stage 0), read and written
(stage 1) and read (stage 2); a Metropolis sub-sweep reads all 4V links and
writes V/2."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

V = 32 ** 4


def test_link_and_accumulator_sizes():
    assert bench.link_bytes(True, 2) == 48 and bench.link_bytes(True, 3) == 72 and bench.link_bytes(False) == 144
    assert bench.alg_bytes(True) == 32 and bench.alg_bytes(False) == 64


def test_stage_bytes_per_step_and_per_launch():
    sL, sZ = 48, 32
    per_step = (24 * sL + 16 * sZ) * V
    assert bench.flow_stage_bytes_per_launch(V, True, 2) * 3 == per_step
    # stages 1 and 2 (timed alone when stage 0 measures): 8 s_L + 6 s_Z per site
    assert bench.flow_stage_bytes_per_launch(V, True, 2, (1, 2)) == (8 * sL + 6 * sZ) * V
    assert bench.flow_stage_bytes_per_launch(V, True, 2, (0,)) == (8 * sL + 4 * sZ) * V
    assert bench.flow_stage_bytes_per_launch(V, True, 2, (1, 2)) == 603979776


def test_metropolis_bytes():
    assert bench.metropolis_bytes_per_launch(V, True, 2) == 4.5 * V * 48


def test_more_than_one_gpu_is_refused():
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
