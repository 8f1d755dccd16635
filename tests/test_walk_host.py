"""Host build of the FP32 fast voxel walk against the general walks (no GPU): index bounds of
every walk-grid read, tau agreement, cut semantics. See tests/host/walk_check.cpp."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    return None


@pytest.mark.skipif(_hipcc() is None, reason="hipcc not available")
def test_fast_walk_host_check():
    src = os.path.join(HERE, "host", "walk_check.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "walk_check")
        base = [_hipcc(), "-x", "hip", "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", src, "-o", exe]
        asan = subprocess.run(base[:1] + ["-Xarch_host", "-fsanitize=address"] + base[1:], capture_output=True, text=True)
        if asan.returncode != 0:
            subprocess.run(base, check=True, capture_output=True, text=True)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
        print(r.stdout, r.stderr)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "OK" in r.stdout
