"""The per-tile splat cull is conservative: the region test (bounding sphere of a tile's samples, the arc
swept by the rotating splat centre over the tile's emission times, a margin for FP32 rounding) never
drops a splat that the per-sample support test keeps. A host program built from the device headers
(tests/host/cull_check.cpp) checks 200,000 random regions in both precisions."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is needed to build the host check")
def test_cull_never_drops_a_contributing_splat(tmp_path):
    exe = str(tmp_path / "cull_check")
    src = os.path.join(HERE, "host", "cull_check.cpp")
    subprocess.run([HIPCC, "-O1", "--offload-arch=gfx950", "-x", "hip", "-w", "-o", exe, src], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert "wrongly culled 0" in out.stdout
