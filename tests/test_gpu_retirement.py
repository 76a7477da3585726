"""Owner retirement after a device-side resource failure, forced from the
host.

A child process runs with tests/native/libhip_fault_inject.so preloaded: the
library makes the k-th hipMalloc, hipEventCreate* or hipStreamCreate* after
arming fail without calling the runtime (no device work is involved) and
counts live device allocations, pinned allocations, events and streams.  In
each case (one caller, or four threads on one KZG SRS) the failing call
returns -2, the next call on the same owner -4
("retired"), a healthy owner created earlier and a fresh owner created
afterwards both compute correct MSMs / NTTs, and once every owner is closed
no allocation, event or stream is left alive.

Reference: native/tests/fail_cuda_alloc.cpp (the preload fixture) and
crates/gpu/tests/icicle_backend.rs:39-62 (a failed context is retired, a new
backend recovers, no device memory retained).
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INJECT = os.path.join(ROOT, "tests", "native", "libhip_fault_inject.so")
MEMORY, EVENT, STREAM, PINNED = 0, 1, 2, 3
N = 4096


def _child(case):
    import ctypes
    sys.path[:0] = [ROOT, os.path.join(ROOT, "zkp-dynaverse_amd")]
    from dvgpu import BackendError, HipBackend
    from dvgpu.batching import DeviceBatch, NttRows, ntt_batch_device
    from dvgpu.resident import FR, ResidentBackend
    from oracle import bn254 as o

    inj = ctypes.CDLL(INJECT)                       # the preloaded copy
    inj.dvgpu_inject_live.restype = ctypes.c_long
    inj.dvgpu_inject_fired.restype = ctypes.c_uint
    report = {"case": case}
    points = o.g1_to_bytes(o.G1_GEN) * N
    ones = (1).to_bytes(32, "little") * N
    want = o.g1_to_bytes(o.g1_mul(N))

    def msm_ok(b):
        return b.msm_encoded(points, ones, N) == want

    def status(fn):
        try:
            fn()
            return 0, ""
        except BackendError as e:
            return e.status, e.message

    def step(msg):
        print(f"child {case}: {msg}", flush=True)

    healthy = HipBackend()
    report["healthy_before"] = msm_ok(healthy)
    if case == "memory":
        a = HipBackend()
        inj.dvgpu_inject_arm(MEMORY, 3)             # the MSM's third device allocation
        step("MSM with the third hipMalloc failing")
        report["first"] = status(lambda: a.msm_encoded(points, ones, N))
        report["second"] = status(lambda: a.msm_encoded(points[:64], ones[:32], 1))
        a.close()
    elif case == "event":
        a = HipBackend()
        d_in, d_out = a.alloc(32 * 256), a.alloc(32 * 256)
        inj.dvgpu_inject_arm(EVENT, 2)              # the timed NTT batch's second event
        step("timed NTT batch with the second event failing")
        report["first"] = status(lambda: ntt_batch_device(a, [NttRows(d_in, d_out, 256, 1)], timed=True))
        report["second"] = status(lambda: a.msm_encoded(points[:64], ones[:32], 1))
        a.close()
        b = HipBackend()
        inj.dvgpu_inject_arm(EVENT, 3)              # the resident engine's third event
        step("resident configuration with the third event failing")
        report["resident_first"] = status(lambda: ResidentBackend(b))
        report["resident_second"] = status(lambda: b.msm_encoded(points[:64], ones[:32], 1))
        b.close()
    elif case == "stream":
        inj.dvgpu_inject_arm(STREAM, 1)             # the owner's own stream
        step("owner creation with its stream failing")
        report["create"] = status(lambda: HipBackend())
        a = HipBackend()
        d_in, d_out = a.alloc(32 * 256), a.alloc(32 * 256)
        inj.dvgpu_inject_arm(STREAM, 1)             # the batch's side stream
        step("batch creation with its side stream failing")
        report["first"] = status(lambda: DeviceBatch(a, ntt=[NttRows(d_in, d_out, 256, 1)], use_graph=False))
        report["second"] = status(lambda: a.msm_encoded(points[:64], ones[:32], 1))
        a.close()
    elif case == "kzg_threads":
        import threading
        from dvgpu.kzg import KzgSrs
        a = HipBackend()
        z = KzgSrs(a, o.kzg_srs(7, 256))
        small = [3, 1, 4, 1, 5, 9, 2, 6]
        report["warm_ok"] = z.precompute_single_openings(small, 8) == o.kzg_openings(o.kzg_srs(7, 256), small, 8)
        coeffs = [(i * i + 7) % o.R for i in range(200)]
        results = []
        gate = threading.Barrier(4)

        def worker():
            gate.wait()
            results.append(status(lambda: z.precompute_single_openings(coeffs, 256)))

        inj.dvgpu_inject_arm(MEMORY, 1)             # the first new buffer of the larger shape
        step("four threads on one SRS, the first device allocation failing")
        threads = [threading.Thread(target=worker) for _ in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        report["threads"] = sorted(r[0] for r in results)
        report["second"] = status(lambda: z.precompute_single_openings(small, 8))
        z.close()
        a.close()
    report["fired"] = inj.dvgpu_inject_fired()
    step("fresh owner")
    fresh = HipBackend()
    report["fresh_ok"] = msm_ok(fresh)
    r = ResidentBackend(fresh)
    h = r.alloc(FR, 8)
    r.write_fields(h, 0, list(range(8)))
    report["fresh_resident_ok"] = r.read_fields(h, 0, 8) == list(range(8))
    report["healthy_after"] = msm_ok(healthy)
    fresh.close()
    healthy.close()
    report["live"] = [inj.dvgpu_inject_live(f) for f in (MEMORY, EVENT, STREAM, PINNED)]
    print("REPORT " + json.dumps(report), flush=True)


def _run(case):
    if not os.path.exists(INJECT):
        pytest.fail("tests/native/libhip_fault_inject.so is not built (run __graft_entry__.build())")
    env = dict(os.environ)
    env["LD_PRELOAD"] = ":".join(x for x in (env.get("LD_PRELOAD", ""), INJECT) if x)
    print(f"case: injected {case} failure (child process)", flush=True)
    proc = subprocess.run([sys.executable, os.path.abspath(__file__), case], env=env, cwd=ROOT,
                          capture_output=True, text=True, timeout=300)
    print(proc.stdout[-4000:], proc.stderr[-4000:], flush=True)
    assert proc.returncode == 0
    line = [x for x in proc.stdout.splitlines() if x.startswith("REPORT ")][-1]
    return json.loads(line[len("REPORT "):])


def _retired(rep, first, second):
    assert rep[first][0] == -2, rep
    assert rep[second][0] == -4 and "retired" in rep[second][1], rep


def _recovered(rep, fired):
    assert rep["fired"] == fired, rep
    assert rep["healthy_before"] and rep["healthy_after"] and rep["fresh_ok"] and rep["fresh_resident_ok"], rep
    assert rep["live"] == [0, 0, 0, 0], rep       # device memory, events, streams, pinned memory


@pytest.mark.gpu
def test_device_allocation_failure_retires_owner():
    rep = _run("memory")
    _retired(rep, "first", "second")
    assert "hipMalloc" in rep["first"][1]
    _recovered(rep, 1)


@pytest.mark.gpu
def test_event_failure_retires_owner():
    rep = _run("event")
    _retired(rep, "first", "second")
    _retired(rep, "resident_first", "resident_second")
    _recovered(rep, 2)


@pytest.mark.gpu
def test_kzg_failure_on_one_thread_retires_safely():
    """Several threads call one SRS; the first call's device allocation
    fails.  It returns -2, retirement waits for the calls in
    progress, the threads still waiting for the SRS return -4, and nothing
    runs on released state (the process completes and balances)."""
    rep = _run("kzg_threads")
    assert rep["warm_ok"], rep
    statuses = rep["threads"]
    assert statuses.count(-2) == 1 and all(s in (0, -2, -4) for s in statuses), rep
    assert rep["second"][0] == -4 and "retired" in rep["second"][1], rep
    _recovered(rep, 1)


@pytest.mark.gpu
def test_stream_failure_retires_owner():
    rep = _run("stream")
    assert rep["create"][0] == -2, rep              # no owner is published
    _retired(rep, "first", "second")
    _recovered(rep, 2)


if __name__ == "__main__":
    _child(sys.argv[1])
