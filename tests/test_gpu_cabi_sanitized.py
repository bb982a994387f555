"""The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer,
on the GPU.

illum_amd/_asan/cabi_check (built by `make -C tests/host_sanitize cabi`, which
build() runs) links a copy of the library whose host code is compiled with
-fsanitize=address,undefined (device code as usual) and calls every family of
entry points on a 5,120-facet mesh, then the error paths (null pointers, bad
sizes, an unknown method, an exact zero pivot, a missing file), each of which
must return its status code.  Any out-of-bounds host access, use after free or
undefined operation ends the run with a report.  The driver runs in its own
process, so this test loads no library itself.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "illum_amd", "_asan", "cabi_check")


@pytest.mark.skipif(not os.path.exists(EXE),
                    reason="sanitized build absent (make -C tests/host_sanitize cabi)")
def test_c_abi_host_code_under_asan_ubsan():
    # protect_shadow_gap=0: the GPU runtime maps memory where ASan would
    # otherwise protect its shadow gap; leaks of the runtime are not ours.
    env = dict(os.environ,
               ASAN_OPTIONS="protect_shadow_gap=0:detect_leaks=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([EXE], capture_output=True, text=True, env=env, timeout=240)
    out = p.stdout + p.stderr
    print(out[-2000:])
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert p.returncode == 0, out[-4000:]
    assert " 0 failures" in p.stdout, out[-4000:]
