"""The library's host C++ under AddressSanitizer on the CPU: every entry
point's argument validation and the size/table helpers run in a host-only
build (moljax_amd/csrc/host_check.cpp, sanitizer on host code only) and must
finish with no sanitizer report and no failed check."""
import os
import shutil
import subprocess

import pytest


def test_host_code_under_address_sanitizer():
    if not shutil.which(os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")) and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.fail("hipcc not found: the host sanitizer build needs the ROCm toolchain")
    from moljax_amd.csrc.build import build_host_check
    exe = build_host_check(device=False)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    res = subprocess.run([str(exe), "--cpu"], capture_output=True, text=True, env=env, timeout=120)
    out = res.stdout + res.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert res.returncode == 0 and "0 failure(s)" in out, out[-4000:]
