"""The serving loop's config deep copy (host code, no GPU): after sstmc_submit returns the caller
may free every array its config points into. See tests/host/config_copy_check.cpp."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "temporalspecklescode_amd")
LIB = os.path.join(PKG, "libsstmc.so")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libsstmc.so is not built")
def test_config_copy_survives_freed_caller_arrays():
    cxx = shutil.which("g++") or "g++"
    src = os.path.join(HERE, "host", "config_copy_check.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "config_copy_check")
        base = [cxx, "-O1", "-g", "-std=c++17", src, "-L" + PKG, "-lsstmc", "-Wl,-rpath," + PKG, "-o", exe]
        r = subprocess.run(base[:1] + ["-fsanitize=address", "-fno-omit-frame-pointer"] + base[1:],
                           capture_output=True, text=True)
        asan = r.returncode == 0
        if not asan:
            subprocess.run(base, check=True, capture_output=True, text=True)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
        print("asan" if asan else "no asan", r.stdout, r.stderr)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "OK" in r.stdout
