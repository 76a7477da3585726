"""The failure-injection preload library (tests/native/hip_fault_inject.c)
on the CPU: loaded with LD_PRELOAD into a process that reaches the HIP
runtime the way the tests do (a library opened with RTLD_LOCAL by ctypes),
unarmed calls must reach the runtime (which reports no device on a host
without a GPU), the
armed k-th call of a family must fail with hipErrorOutOfMemory without
reaching it, and only successful creations count as live."""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP = "/opt/rocm"

CALLER = r"""
#include <hip/hip_runtime_api.h>
int call_malloc(void) { void* p = 0; return (int)hipMalloc(&p, 64); }
int call_event(void) { hipEvent_t e; return (int)hipEventCreateWithFlags(&e, 0); }
int call_stream(void) { hipStream_t s; return (int)hipStreamCreateWithFlags(&s, 1); }
"""

CHILD = r"""
import ctypes, sys
c = ctypes.CDLL(sys.argv[1])            # RTLD_LOCAL, as ctypes loads the library under test
inj = ctypes.CDLL(sys.argv[2])          # the preloaded copy
inj.dvgpu_inject_live.restype = ctypes.c_long
out = []
for fam, fn in ((0, c.call_malloc), (1, c.call_event), (2, c.call_stream)):
    out.append(fn())                    # unarmed: the runtime's answer
    inj.dvgpu_inject_arm(fam, 2)
    out.append(fn())                    # first call after arming: still the runtime
    out.append(fn())                    # second: injected
    out.append(fn())                    # spent: the runtime again
out.append(inj.dvgpu_inject_fired())
out += [inj.dvgpu_inject_live(f) for f in range(4)]
print(" ".join(map(str, out)))
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(HIP, "lib", "libamdhip64.so")), reason="no HIP runtime")
def test_injection_through_a_locally_loaded_runtime():
    d = tempfile.mkdtemp(prefix="fault_inject_")
    inj = os.path.join(d, "libhip_fault_inject.so")
    subprocess.run(["gcc", "-O2", "-Wall", "-Werror", "-shared", "-fPIC", "-o", inj,
                    os.path.join(ROOT, "tests", "native", "hip_fault_inject.c"), "-ldl", "-lpthread"], check=True)
    src = os.path.join(d, "caller.c")
    with open(src, "w") as f:
        f.write(CALLER)
    caller = os.path.join(d, "libcaller.so")
    subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(HIP, "include"), "-shared", "-fPIC", src,
                    "-o", caller, "-L", os.path.join(HIP, "lib"), "-lamdhip64",
                    "-Wl,-rpath," + os.path.join(HIP, "lib")], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = ":".join(x for x in (env.get("LD_PRELOAD", ""), inj) if x)
    proc = subprocess.run([sys.executable, "-c", CHILD, caller, inj], env=env, capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 0, proc.stderr
    v = [int(x) for x in proc.stdout.split()]
    for fam in range(3):
        unarmed, first, injected, spent = v[4 * fam:4 * fam + 4]
        assert injected == 2                              # hipErrorOutOfMemory, injected
        assert unarmed == first == spent and unarmed != 2  # the runtime's own answer (no device here)
        assert v[13 + fam] == (3 if unarmed == 0 else 0)   # live only for creations that succeeded
    assert v[12] == 3                                     # one injection per family
    assert v[16] == 0
