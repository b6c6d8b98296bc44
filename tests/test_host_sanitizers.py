"""Host code of libscarfhip under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

The library is rebuilt into a temporary directory with the sanitizers on its host code
only (``-Xarch_host``; the device code is compiled at -O0, since no kernel runs), and
``host_sanitizer_driver.c`` calls every workspace-size function over a grid of sizes and
every entry point with arguments its validation must reject.  Any out-of-bounds access,
use after free, leak or undefined arithmetic in that host code fails the test.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "scarf_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang"
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-omit-frame-pointer"]


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                    reason="ROCm toolchain not available")
@pytest.mark.timeout(900)
def test_c_abi_host_code_under_sanitizers(tmp_path):
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))

    def compile_one(f):
        out = tmp_path / (f[:-4] + ".o")
        cmd = [HIPCC, "--offload-arch=gfx950", "-O1", "-Xarch_device", "-O0", "-g0", "-fPIC",
               "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               *SAN, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, f), "-o", str(out)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, f"{f}: {r.stderr[-2000:]}"
        return str(out)

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, srcs))
    lib = tmp_path / "libscarfhip_san.so"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *SAN, "-o", str(lib),
                        *objs], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    drv = tmp_path / "driver"
    r = subprocess.run([CLANG, "-fsanitize=address,undefined", "-g", "-I", INCLUDE,
                        os.path.join(HERE, "host_sanitizer_driver.c"), "-L", str(tmp_path),
                        "-lscarfhip_san", f"-Wl,-rpath,{tmp_path}", "-o", str(drv)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(drv)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert r.stdout.startswith("ok ") and int(r.stdout.split()[1]) > 1000
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
