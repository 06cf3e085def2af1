"""The library's host code (operator assembly and CSR conversion,
mt_host_ops.cpp; receiver rows, mt_sites.cpp) and the oracle (sp2_oracle.c)
built with AddressSanitizer and UndefinedBehaviorSanitizer and run on a small
grid by tests/sanitize/host_check.cpp (CPU only; the GPU kernels are not
built here)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="needs gcc/g++")
def test_host_code_under_asan_ubsan(tmp_path):
    csrc = os.path.join(ROOT, "modem-gpu_amd", "csrc")
    inc = ["-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", f"-I{csrc}",
           f"-I{os.path.join(ROOT, 'oracle')}"]
    objs = []
    for src in (os.path.join(csrc, "mt_host_ops.cpp"), os.path.join(csrc, "mt_sites.cpp"),
                os.path.join(ROOT, "tests", "sanitize", "host_check.cpp")):
        o = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run(["g++", "-std=c++17", *SAN, *inc, "-c", src, "-o", o], check=True)
        objs.append(o)
    o = str(tmp_path / "sp2_oracle.o")
    subprocess.run(["gcc", "-std=c11", "-ffp-contract=off", "-fno-fast-math", *SAN, "-c",
                    os.path.join(ROOT, "oracle", "sp2_oracle.c"), "-o", o], check=True)
    objs.append(o)
    exe = str(tmp_path / "host_check")
    subprocess.run(["g++", "-fsanitize=address,undefined", "-o", exe, *objs, "-lm"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host sanitizer check: ok" in r.stdout
