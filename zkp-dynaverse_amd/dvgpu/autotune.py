"""Autotuner for the launch shapes: MSM window width c per log2 size, the
bucket-task shape (points per accumulation task, buckets per reduction
thread) and the NTT radix (butterfly stages per LDS round trip) per log2 length.

The sweep runs on the device it tunes, over device-resident inputs shaped like
the batched workloads: at every log2 size L it times a batch of
max(1, 2^(18-L)) MSMs of 2^L points (so small sizes are measured as they run
in batches, not launch-bound) and NTT row batches of 2^20 elements in total.
Every candidate computes the same group elements; the sweep also checks that
the MSM results of all candidates agree before it accepts a choice.

    python -m dvgpu.autotune --out tuning.json      # sweep, write the table
    apply(backend, load(path))                      # use a table

The table applies through ``dvgpu_ext_set_tuning``; a table that is never
applied leaves the built-in choices (which are this sweep's results on
MI355X, compiled into the library).
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time
from typing import Dict, Optional

import numpy as np

from . import _lib
from .backend import BackendError, HipBackend
from .batching import MsmItem, msm_batch_device
from .rng import ChaCha20Stream

SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")


def get(backend: HipBackend) -> dict:
    t = _lib.DvgpuTuning()
    if backend.lib.dvgpu_ext_get_tuning(backend.handle, ctypes.byref(t)) != 0:
        raise BackendError(-1, "tuning unavailable")
    return {"msm_window_bits": {i: t.msm_window_bits[i] for i in range(25) if t.msm_window_bits[i]},
            "ntt_pass_stages": {i: t.ntt_pass_stages[i] for i in range(21) if t.ntt_pass_stages[i]},
            "msm_task_points": t.msm_task_points, "msm_chunk": t.msm_chunk,
            "msm_device_finalize_min": t.msm_device_finalize_min,
            "msm_chunk_by_log2": {i: t.msm_chunk_by_log2[i] for i in range(25) if t.msm_chunk_by_log2[i]}}


def window_for(backend: HipBackend, n: int) -> int:
    """The window width the owner uses for an n-point MSM."""
    c = backend.lib.dvgpu_ext_msm_window(backend.handle, n)
    if c <= 0:
        raise BackendError(c, "window query failed")
    return c


def apply(backend: HipBackend, table: dict) -> None:
    """Install a table (keys as written by ``tune``; missing keys = built-in)."""
    t = _lib.DvgpuTuning()
    for k, v in table.get("msm_window_bits", {}).items():
        t.msm_window_bits[int(k)] = int(v)
    for k, v in table.get("ntt_pass_stages", {}).items():
        t.ntt_pass_stages[int(k)] = int(v)
    t.msm_task_points = int(table.get("msm_task_points", 0))
    t.msm_chunk = int(table.get("msm_chunk", 0))
    t.msm_device_finalize_min = int(table.get("msm_device_finalize_min", 0))
    for k, v in table.get("msm_chunk_by_log2", {}).items():
        t.msm_chunk_by_log2[int(k)] = int(v)
    if backend.lib.dvgpu_ext_set_tuning(backend.handle, ctypes.byref(t)) != 0:
        raise BackendError(-1, "invalid tuning table")


def load(path: str = SHIPPED) -> dict:
    with open(path) as f:
        return json.load(f)


def _median_ms(fn, reps: int) -> float:
    fn()
    ts = []
    for _ in range(reps):
        a = time.perf_counter()
        fn()
        ts.append((time.perf_counter() - a) * 1e3)
    return statistics.median(ts)


class _Inputs:
    def __init__(self, backend: HipBackend, log_max: int):
        n = 1 << log_max
        rng = ChaCha20Stream(77)
        sc = rng.uniform_fr_array(n).reshape(-1)
        self.d_sc = backend.to_device(sc)
        self.points = backend.alloc(64 * n)
        backend.fixed_base_device(self.d_sc.ptr, n, self.points.ptr)   # [s_i] g1, resident
        self.d_sc2 = backend.to_device(ChaCha20Stream(78).uniform_fr_array(n).reshape(-1))
        self.n = n


def _msm_items(L: int, log_max: int, c: int):
    k = max(1, 1 << max(0, 18 - L))
    n = 1 << L
    span = (1 << log_max) - n
    return [MsmItem(n, (j * 7919 * 64) % (span + 1), (j * 104729 * 32) % (span + 1), c) for j in range(k)]


def tune_msm(backend, inp: _Inputs, logs, reps: int, log=print) -> Dict[int, dict]:
    out = {}
    log_max = inp.n.bit_length() - 1
    for L in logs:
        cands = [c for c in range(max(4, L - 9), min(20, L - 1) + 1)]
        res, times, ref = {}, {}, None
        for c in cands:
            items = _msm_items(L, log_max, c)
            got = msm_batch_device(backend, inp.points, inp.d_sc2, items)
            if ref is None:
                ref = got
            elif got != ref:
                raise AssertionError(f"MSM results differ between windows at 2^{L}, c={c}")
            times[c] = _median_ms(lambda: msm_batch_device(backend, inp.points, inp.d_sc2, items), reps)
        best = min(times, key=times.get)
        k = len(_msm_items(L, log_max, best))
        out[L] = {"c": best, "ms": {str(c): round(t, 4) for c, t in times.items()}, "msms_per_batch": k,
                  "mpoints_s": round(k * (1 << L) / (times[best] * 1e-3) / 1e6, 2)}
        log(f"msm 2^{L}: best c={best} {times[best]:.3f} ms for {k} MSMs "
            f"({out[L]['mpoints_s']} Mpoints/s); " + " ".join(f"{c}:{t:.2f}" for c, t in times.items()))
    return out


def tune_chunk(backend, inp: _Inputs, windows: Dict[int, int], logs, reps: int, log=print) -> Dict[int, int]:
    """Buckets per reduction thread, per log2 size, at the chosen window."""
    log_max = inp.n.bit_length() - 1
    base = get(backend)
    out = {}
    for L in logs:
        items = _msm_items(L, log_max, windows[L])
        times = {}
        for ch in (1, 2, 4, 8):
            apply(backend, dict(base, msm_chunk_by_log2={L: ch}))
            times[ch] = _median_ms(lambda: msm_batch_device(backend, inp.points, inp.d_sc2, items), reps)
        out[L] = min(times, key=times.get)
        log(f"chunk 2^{L}: best {out[L]}; " + " ".join(f"{c}:{t:.3f}" for c, t in times.items()))
    apply(backend, base)
    return out


def tune_shape(backend, inp: _Inputs, windows: Dict[int, int], logs, reps: int, log=print) -> dict:
    """Task points x chunk, scored by the summed time over `logs`."""
    log_max = inp.n.bit_length() - 1
    scores = {}
    base = get(backend)
    for tp in (16, 32, 64, 128):
        for ch in (2, 4, 8):
            apply(backend, dict(base, msm_task_points=tp, msm_chunk=ch))
            tot = 0.0
            for L in logs:
                items = _msm_items(L, log_max, windows[L])
                tot += _median_ms(lambda: msm_batch_device(backend, inp.points, inp.d_sc2, items), reps)
            scores[(tp, ch)] = tot
            log(f"shape task_points={tp} chunk={ch}: {tot:.3f} ms")
    apply(backend, base)
    tp, ch = min(scores, key=scores.get)
    return {"task_points": tp, "chunk": ch, "ms": {f"{a}x{b}": round(v, 4) for (a, b), v in scores.items()}}


def tune_ntt(backend, logs, reps: int, log=print) -> Dict[int, dict]:
    """Per length 2^L (in rows of 2^20 elements in total), butterfly stages
    per LDS round trip: 1 (radix-2 rounds) against 2 (radix-4 rounds); both
    must give identical results."""
    out = {}
    base = get(backend)
    vals = ChaCha20Stream(79).uniform_fr_array(1 << 20).reshape(-1)
    d_in = backend.to_device(vals)
    d_out = backend.alloc(32 << 20)
    for L in logs:
        rows = 1 << (20 - L)
        times, ref = {}, None
        for s in (1, 2):
            apply(backend, dict(base, ntt_pass_stages={L: s}))
            backend.ntt_device(d_in.ptr, d_out.ptr, 1 << L, rows, False)
            got = d_out.download().tobytes()
            if ref is None:
                ref = got
            elif got != ref:
                raise AssertionError(f"NTT results differ between radices at 2^{L}")
            ms = []
            for _ in range(reps):
                ms.append(backend.ntt_device(d_in.ptr, d_out.ptr, 1 << L, rows, False, timed=True))
            times[s] = statistics.median(ms)
        best = min(times, key=times.get)
        out[L] = {"stages": best, "rows": rows, "ms": {str(s): round(t, 4) for s, t in times.items()}}
        log(f"ntt 2^{L} x {rows}: best stages/round={best} {times[best]:.4f} ms; "
            + " ".join(f"{s}:{t:.4f}" for s, t in times.items()))
    apply(backend, base)
    d_in.free()
    d_out.free()
    return out


def tune(backend: Optional[HipBackend] = None, log_max: int = 20, reps: int = 5, log=print) -> dict:
    own = backend is None
    b = backend or HipBackend()
    try:
        inp = _Inputs(b, log_max)
        logs = list(range(10, log_max + 1))
        msm = tune_msm(b, inp, logs, reps, log)
        windows = {L: v["c"] for L, v in msm.items()}
        apply(b, {"msm_window_bits": windows})
        shape = tune_shape(b, inp, windows, [12, 16, log_max], reps, log)
        apply(b, {"msm_window_bits": windows, "msm_task_points": shape["task_points"]})
        chunks = tune_chunk(b, inp, windows, logs, reps, log)
        ntt = tune_ntt(b, logs, reps, log)
        table = {"device": b.device_identity(), "msm_window_bits": {str(k): v for k, v in windows.items()},
                 "msm_task_points": shape["task_points"], "msm_chunk": shape["chunk"],
                 "msm_chunk_by_log2": {str(k): v for k, v in chunks.items()},
                 "ntt_pass_stages": {str(k): v["stages"] for k, v in ntt.items()},
                 "sweep": {"msm": {str(k): v for k, v in msm.items()}, "shape": shape,
                           "ntt": {str(k): v for k, v in ntt.items()}}}
        return table
    finally:
        if own:
            b.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--log-max", type=int, default=20)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args(argv)
    table = tune(log_max=a.log_max, reps=a.reps, log=lambda s: print(s, flush=True))
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    print(f"wrote {a.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
