"""The device-memory block cache's lifecycle on the CPU, under sanitizers.

tests/native/block_cache_harness.cpp runs the cache code of the library
(manifolds-rs_amd/csrc/block_cache.hpp) against a simulated multi-stream
runtime in the fit server's pattern: prep streams allocate and fill a
request's blocks, a batcher stream keeps using them for several ticks and
returns them to the prep worker's cache at once from its own stream.  Every
block is checked by every queued use; the harness is built with
AddressSanitizer + UBSan and with ThreadSanitizer (host code only, g++), and
both must finish clean with no corrupted block, no live event and no leaked
block.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "native" / "block_cache_harness.cpp"


@pytest.mark.parametrize("sanitizer", ["address,undefined", "thread"])
def test_block_cache_lifecycle_under_sanitizer(tmp_path, sanitizer):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "harness"
    build = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", f"-fsanitize={sanitizer}",
             "-pthread", str(SRC), "-o", str(exe)]
    subprocess.run(build, check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe), "400"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "HARNESS_OK" in r.stdout and "ThreadSanitizer" not in r.stderr
    assert "AddressSanitizer" not in r.stderr
