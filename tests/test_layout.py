"""The device lattice layout, checked on the host: bijections between
storage index, coordinates, GLAC's lexicographic index, the per-parity and
per-time-slice enumerations, and in-range neighbours, for lattices with
extents that are multiples of 4, of 2 only, and mixed.  This is what keeps
every kernel's indexing inside its buffers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

LATTICES = [(2, 2, 2, 2), (4, 4, 4, 4), (4, 4, 4, 8), (6, 4, 8, 6), (6, 6, 6, 6), (2, 6, 4, 10),
            (8, 8, 8, 16), (12, 8, 4, 6), (24, 24, 24, 48), (32, 32, 32, 32),
            # runs of whole 32-site chunks with rows of 2, 6 and 64 entries
            (4, 16, 4, 4), (12, 16, 4, 8), (128, 2, 4, 4)]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_layout_bijections(tmp_path):
    exe = str(tmp_path / "layout_check")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                    os.path.join(ROOT, "tests", "layout_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    args = [str(v) for n in LATTICES for v in n]
    out = subprocess.run([exe] + args, capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("ok ") == len(LATTICES)
