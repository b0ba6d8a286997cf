"""The batched deep marches' frame geometry, on the CPU: a host-only program
(tests/native/frame_geometry.hip) built against the library checks, for many
channel shapes, chunk widths and depths, that each pass's perimeter waves and
the interior waves tile the channel and that every pass reads only what the
one before it wrote."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "memory-aware-lbm_amd")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.skipif(not os.path.exists(os.path.join(PKG, "liblbm_hip.so")),
                    reason="library not built")
def test_batch_frame_tiles_each_channel(tmp_path):
    exe = str(tmp_path / "frame_geometry")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++20", "-w",
                    "-I" + os.path.join(PKG, "csrc"), "-I" + os.path.join(ROOT, "include"),
                    "-o", exe, os.path.join(ROOT, "tests", "native", "frame_geometry.hip"),
                    "-L" + PKG, "-llbm_hip", "-Wl,-rpath," + PKG],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:]
    last = out.stdout.strip().splitlines()[-1]
    ok, bad = int(last.split()[1]), int(last.split()[3])
    assert bad == 0 and ok > 1000, last
