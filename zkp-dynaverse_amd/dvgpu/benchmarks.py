"""Measured workloads beside the single-MSM headline: the batched launch
(64 uneven G1 MSMs, 8 G2 MSMs, 32 NTT rows, 4096 G2 checks), the incremental
n = 2^16 session and the served mixed-n update stream against the
one-call-at-a-time path.  Everything runs on seeded synthetic data; every
served proof is compared with the plain path's bytes.
"""
import gc
import statistics
import time
from typing import Dict, List

import numpy as np

from .backend import HipBackend
from .batching import DeviceBatch, MsmItem, NttRows, msm_batch_device
from .codec import FR_MODULUS
from .rng import ChaCha20Stream
from .serving import (PlainSession, Server, SessionSpec, SessionStore, merge_update, percentile, request_stream,
                      session_pool, update_batch)
from .workload import tau_powers

PEAK_MAC_PER_S = 256 * 64 * 2.4e9   # v_mad_u64_u32 issue rate: CUs x lanes x clock


def msm_macs(n: int, fq2: bool = False) -> float:
    """136 MACs per Montgomery product times the Pippenger product count at
    the best window (madd 10, add 14, dbl 9 products; Fq2 x3)."""
    best = None
    for c in range(4, 21):
        w = -(-255 // c)
        f = n * w * 10 + w * (1 << c) * 14 + w * (9 * c + 14)
        best = f if best is None else min(best, f)
    return 136.0 * best * (3 if fq2 else 1)


def ntt_macs(n: int) -> float:
    return 136.0 * (n // 2) * max(1, n.bit_length() - 1)


def batched_launch(backend: HipBackend, replays: int = 10, log=print) -> Dict:
    """The batched configuration as one hipGraph: batches/s and MAC rate."""
    n_srs = 1 << 18
    tau = ChaCha20Stream(4).nonzero_fr()
    pw = tau_powers(tau, n_srs)
    d_pw = backend.to_device(pw.reshape(-1))
    srs = backend.alloc(64 * n_srs)
    backend.fixed_base_device(d_pw.ptr, n_srs, srs.ptr)
    srs2 = backend.alloc(128 * 1024)
    backend.fixed_base_device(d_pw.ptr, 1024, srs2.ptr, g2=True)
    d_pw.free()
    sc = backend.to_device(ChaCha20Stream(40).uniform_fr_array(n_srs).reshape(-1))
    u = ChaCha20Stream(41)
    items = []
    for _ in range(64):
        m = int(round(2 ** (10 + 8 * u.below(1 << 20) / (1 << 20))))
        m = max(1 << 10, min(n_srs, m))
        items.append(MsmItem(m, u.below(n_srs - m + 1), u.below(n_srs - m + 1)))
    items2 = [MsmItem(1024, 0, 4096 * k) for k in range(8)]
    v12 = backend.to_device(ChaCha20Stream(42).uniform_fr_array(16 << 12).reshape(-1))
    v16 = backend.to_device(ChaCha20Stream(43).uniform_fr_array(16 << 16).reshape(-1))
    o12, o16 = backend.alloc(32 << 16), backend.alloc(32 << 20)
    ntt = [NttRows(v12, o12, 1 << 12, 16), NttRows(v16, o16, 1 << 16, 16)]
    # 4096 G2 points: [s_i] g2 on the device; odd ones get a cofactor-order
    # twist point added (on the device), which leaves the subgroup.
    g2sc = ChaCha20Stream(44).uniform_fr_array(4096).reshape(-1)
    d_g2sc = backend.to_device(g2sc)
    g2pts = backend.alloc(128 * 4096)
    backend.fixed_base_device(d_g2sc.ptr, 4096, g2pts.ptr, g2=True)
    wire = _coset_wire(backend, g2pts, 4096)
    d_wire = backend.to_device(wire)
    masks = backend.alloc(4 * 4096)
    log("# batched workload ready")
    eager = DeviceBatch(backend, g1=(srs, sc, items), g2=(srs2, sc, items2), ntt=ntt, g2_wire=d_wire,
                        g2_wire_n=4096, g2_algorithm=3, g2_masks=masks, use_graph=False)
    ref = eager.run()
    eager.close()
    graph = DeviceBatch(backend, g1=(srs, sc, items), g2=(srs2, sc, items2), ntt=ntt, g2_wire=d_wire,
                        g2_wire_n=4096, g2_algorithm=3, g2_masks=masks, use_graph=True)
    ms, same = [], True
    t0 = time.perf_counter()
    for _ in range(replays):
        r = graph.run()
        same &= r[0] == ref[0] and r[1] == ref[1] and r[2] == ref[2]
        ms.append(r[3])
    wall = (time.perf_counter() - t0) / replays
    graph.close()
    invalid = ref[2]["invalid_points"]
    macs = sum(msm_macs(it.n) for it in items) + sum(msm_macs(it.n, True) for it in items2) \
        + 16 * ntt_macs(1 << 12) + 16 * ntt_macs(1 << 16) + 4096 * 3 * 136 * (63 * 9 + 28 * 14 + 4 * 2 + 4 * 14 + 4)
    dev = float(np.median(ms))
    return {"batches_per_s": round(1e3 / (wall * 1e3), 2), "graph_device_ms": round(dev, 3),
            "wall_ms": round(wall * 1e3, 3), "mac_frac": round(macs / (dev * 1e-3) / PEAK_MAC_PER_S, 4),
            "g1_points": sum(it.n for it in items), "g2_invalid": invalid, "replays_match_eager": bool(same)}


# A point of the twist E'(Fq2) of cofactor order: [r] T0 for the twist point
# T0 with x = 1 (canonical wire bytes x.c0 || x.c1 || y.c0 || y.c1).  Adding
# it to a subgroup point leaves the subgroup.
COSET_POINT = bytes.fromhex(
    "44f2242359f77b012b0f75db89f427f32aa85fe8c5695d6c2b326ea8ae37ef26"
    "0732a8d2aaa02f947fe48530695cbf84901fc238b71a00a77a5e6ee19077f00d"
    "57814d44f915fb58fd3f8b83232865f12d1479cee506486278c309f00836571e"
    "c2ff043bf048e69c6e8eeafa83566132e733d5648c342bc63843ffd3c912e61c")


def _coset_wire(backend: HipBackend, g2pts, count: int) -> bytes:
    """Wire bytes of the device points g2pts (Montgomery), odd ones plus
    COSET_POINT: one batch of G2 MSMs ([P_i] or [P_i, T] with scalars 1)."""
    pts = backend.alloc(128 * (count + 1))
    backend.device_copy(pts.ptr, g2pts.ptr, 128 * count)
    t_canon = backend.to_device(COSET_POINT)
    backend.points_to_mont_device(t_canon.ptr, pts.ptr + 128 * count, 1, g2=True)
    idx, items = [], []
    for i in range(count):
        items.append(MsmItem(1 + (i & 1), len(idx), 0))
        idx.extend([i, count] if i & 1 else [i])
    d_idx = backend.to_device(np.array(idx, dtype=np.uint32).view(np.uint8))
    ones = backend.to_device((1).to_bytes(32, "little") * 2)
    out = msm_batch_device(backend, pts, ones, items, g2=True, index=d_idx)
    for a in (pts, t_canon, d_idx, ones):
        a.free()
    return b"".join(out)


def small_sequence_gpu(backend: HipBackend, workload: str = "hotspot", seed: int = 1, n: int = 256,
                       updates: int = 8, k: int = 4, t: int = 3, timed: bool = False) -> Dict:
    """The plain n = 256 update sequence through the per-call drop-in.

    SRS [tau^i] g1 (i < n), g2 and [tau] g2 from the seeded tau; f uniform.
    Every update rewrites k coefficients (workload.small_sequence), then
    recomputes the commitment C = MSM(P, f) (dvgpu_bn254_msm), the
    evaluations NTT(f) and the coefficients back by INTT (dvgpu_bn254_ntt),
    and verifies: C against the incremental form C_prev + sum d_j P_i_j (a
    sparse MSM with the cached point, recompute-and-compare), INTT(NTT(f)) ==
    f, and the KZG opening at z = w_n^t, W = [q(tau)] g1 with
    q = (f - f(z)) / (X - z), y = NTT(f)[t], by its pairing check
    e(C - [y] g1, g2) == e(W, [tau - z] g2) (dvgpu_bn254_kzg_verify_batch).
    Returns per-update records (C, W, y bytes, the check flags) and, with
    `timed`, the per-update wall-clock ms."""
    from .codec import encode_fr, root_of_unity_2_20
    from .pairing import verify_update_proofs
    from .workload import small_sequence
    R = FR_MODULUS
    tau, coeffs, ups = small_sequence(seed, n, updates, k, workload)
    P = backend.fixed_base_encoded(tau_powers(tau, n).tobytes(), n, False)
    z = pow(pow(root_of_unity_2_20(), (1 << 20) // n, R), t, R)
    divisor = backend.fixed_base_encoded(encode_fr((tau - z) % R), 1, True)
    f = [int.from_bytes(coeffs[i].tobytes(), "little") for i in range(n)]
    enc = lambda vals: b"".join(v.to_bytes(32, "little") for v in vals)
    C = backend.msm_encoded(P, enc(f), n)
    records, ms = [], []
    for up in ups:
        a = time.perf_counter()
        deltas = {}
        for i, v in up:
            deltas[i] = (deltas.get(i, 0) + v - f[i]) % R
            f[i] = v
        idx = sorted(deltas)
        C_new = backend.msm_encoded(P, enc(f), n)
        # incremental form: the cached point with scalar 1 plus the deltas
        inc = backend.msm_encoded(C + b"".join(P[64 * i:64 * i + 64] for i in idx),
                                  encode_fr(1) + enc(deltas[i] for i in idx), len(idx) + 1)
        evals = backend.ntt_encoded(enc(f), n, False)
        back = backend.ntt_encoded(evals, n, True)
        y = int.from_bytes(evals[32 * t:32 * t + 32], "little")
        q, acc = [0] * (n - 1), 0          # synthetic division by (X - z)
        for j in range(n - 1, 0, -1):
            acc = (acc * z + f[j]) % R
            q[j - 1] = acc
        W = backend.msm_encoded(P[:64 * (n - 1)], enc(q), n - 1)
        ok = verify_update_proofs(backend, [C_new + W + y.to_bytes(32, "little")], [divisor])[0]
        ms.append((time.perf_counter() - a) * 1e3)
        records.append({"C": C_new, "W": W, "y": y, "evals": evals, "incremental_ok": inc == C_new,
                        "roundtrip_ok": back == enc(f), "verified": bool(ok)})
        C = C_new
    out = {"tau": tau, "z": z, "f": f, "records": records}
    if timed:
        out["ms"] = ms
    return out


def incremental_updates(n: int = 1 << 16, count: int = 16, seed: int = 50):
    """The incremental configuration's updates: k cycling 1, 16, 256, 4096;
    even updates hotspot (uniform over a fixed seed-chosen subset of 1% of
    the indices, 655 at n = 2^16), odd ones spreading (uniform over n); deltas
    uniform Fr.  Each as merge_update's (indices, deltas bytes, deltas)."""
    rng = ChaCha20Stream(seed)
    hot = [rng.below(n) for _ in range(max(1, n // 100))]
    ups = []
    for j in range(count):
        k = (1, 16, 256, 4096)[j % 4]
        idx = [hot[rng.below(len(hot))] if j % 2 == 0 else rng.below(n) for _ in range(k)]
        ups.append(merge_update(idx, rng.uniform_fr(k)))
    return ups


def incremental(backend: HipBackend, log=print) -> Dict:
    """n = 2^16 resident session, 16 updates with k cycling {1, 16, 256, 4096},
    hotspot (a fixed 1% index subset) and spreading alternating; served path
    (one update per call, the state on the device) against the plain path.
    Every update's proof is pairing-verified on the device as the reference
    verifies each update (dynaverse.rs:447-451): separately timed per update
    (`*_verified_*` figures: update + its own verification call) and once for
    all 16 proofs together."""
    from .pairing import verify_update_proofs
    spec = SessionSpec(0, 1 << 16, 5)
    store = SessionStore(backend, 1 << 30)
    s = store.admit(spec)
    plain = PlainSession(backend, spec)
    ups = incremental_updates(spec.n)
    served_ms, plain_ms, verify_ms, same, proofs = [], [], [], True, []
    for idx, raw, dv in ups:
        a = time.perf_counter()
        p1 = update_batch(backend, store, [s.update_struct(idx, raw)])[0]
        b = time.perf_counter()
        ok1 = verify_update_proofs(backend, [p1], [s.divisor])[0]
        c = time.perf_counter()
        p2 = plain.apply(backend, idx, raw, dv)
        d = time.perf_counter()
        served_ms.append((b - a) * 1e3)
        verify_ms.append((c - b) * 1e3)
        plain_ms.append((d - c) * 1e3)
        same &= p1 == p2 and ok1
        proofs.append(p1)
    same &= store.evaluations(s) == plain.evals
    a = time.perf_counter()
    all_ok = all(verify_update_proofs(backend, proofs, [s.divisor] * len(proofs)))
    batch_verify_ms = (time.perf_counter() - a) * 1e3
    store.close()
    sv = [x + y for x, y in zip(served_ms, verify_ms)]
    pv = [x + y for x, y in zip(plain_ms, verify_ms)]
    return {"served_updates_per_s": round(16 / (sum(served_ms) * 1e-3), 1),
            "served_p50_ms": round(percentile(served_ms, 50), 3), "served_p99_ms": round(percentile(served_ms, 99), 3),
            "plain_updates_per_s": round(16 / (sum(plain_ms) * 1e-3), 1),
            "plain_p50_ms": round(percentile(plain_ms, 50), 3), "plain_p99_ms": round(percentile(plain_ms, 99), 3),
            "verify_call_p50_ms": round(percentile(verify_ms, 50), 3),
            "served_verified_updates_per_s": round(16 / (sum(sv) * 1e-3), 1),
            "served_verified_p50_ms": round(percentile(sv, 50), 3),
            "served_verified_p99_ms": round(percentile(sv, 99), 3),
            "plain_verified_updates_per_s": round(16 / (sum(pv) * 1e-3), 1),
            "all_16_verified_one_call_ms": round(batch_verify_ms, 3),
            "verified_all": bool(all_ok), "identical": bool(same)}


def _drive(srv: Server, stream, requests_prepared):
    """Open-loop driver: submit each request at its arrival offset.  The
    cyclic garbage collector is paused meanwhile (its pauses would land in
    the tail latency of whichever requests are in flight).  Latency runs to
    the request's verified result (its proof, when the server does not
    verify)."""
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter() + 0.01
        reqs = []
        for (t, sid, _, _), (mi, raw) in zip(stream, requests_prepared):
            now = time.perf_counter()
            if t0 + t > now + 2e-4:
                time.sleep(t0 + t - now)
            reqs.append(srv.submit(sid, mi, raw, arrival=t0 + t))
        for r in reqs:
            r.future.result(timeout=300)
    finally:
        gc.enable()
    lat = [(r.finished - r.arrival) * 1e3 for r in reqs]
    span = max(r.finished for r in reqs) - t0
    return reqs, lat, span


def _parts(reqs) -> Dict:
    """p50/p99 of the latency components: submission lag behind the
    scheduled arrival, queueing until a batch starts, the batch's service
    time, and the wait for the verification call after the proof."""
    lag = [(r.submitted - r.arrival) * 1e3 for r in reqs]
    wait = [(r.started - r.submitted) * 1e3 for r in reqs]
    svc = [(r.done - r.started) * 1e3 for r in reqs]
    ver = [(r.finished - r.done) * 1e3 for r in reqs]
    proof = [(r.done - r.arrival) * 1e3 for r in reqs]
    return {"p99_submit_lag_ms": round(percentile(lag, 99), 3), "p99_queue_ms": round(percentile(wait, 99), 3),
            "p99_batch_ms": round(percentile(svc, 99), 3), "p50_batch_ms": round(percentile(svc, 50), 3),
            "p50_verify_wait_ms": round(percentile(ver, 50), 3), "p99_verify_wait_ms": round(percentile(ver, 99), 3),
            "p50_proof_ms": round(percentile(proof, 50), 3), "p99_proof_ms": round(percentile(proof, 99), 3)}


def _plain_replay(plain, stream, prepared, backend) -> List[bytes]:
    """The plain path over a stream in arrival order (the order the server
    applies each session's requests in)."""
    return [plain[sid].apply(backend, mi, raw, dv) for (_, sid, _, _), (mi, raw, dv) in zip(stream, prepared)]


def served_stream(backend: HipBackend, counts=(96, 72, 48, 24), count: int = 4000, qps_points=(),
                  plain_verified_count: int = 24, lanes: int = 2, log=print) -> Dict:
    """Poisson mixed-n update stream: served (continuous batching, resident
    state) at several arrival rates and saturated, against the plain path
    run back to back (its maximum rate), without and with its per-update
    verification.  After every served run (server closed) all of its proofs
    are pairing-verified on the device in one call on this owner
    (serving.verify_served); the verified rate counts that call's time
    (count / (serving span + verification)).  Every served proof of every
    run is also compared with the plain path's bytes.  Sessions per size
    follow the request mix (a session's updates apply in order, so the pool,
    not the batch limit, bounds how many requests can share a batch)."""
    from .pairing import verify_update_proofs
    from .serving import verify_served
    specs = session_pool(counts=counts, seed=6)
    store = SessionStore(backend, 32 << 30)
    t_setup = time.perf_counter()
    for sp in specs:
        store.admit(sp)
    log(f"# served: {len(specs)} sessions resident in {time.perf_counter() - t_setup:.1f}s")
    base_stream = request_stream(specs, count, 1.0, seed=6)
    prepared = [merge_update(idx, ds) for _, _, idx, ds in base_stream]
    plain = {sp.sid: PlainSession(backend, sp) for sp in specs}
    # Plain path, closed loop, over the whole stream (the reference's per-call
    # operator path: msm_batch + ntt), unverified.
    p_lat = []
    p_out = []
    t0 = time.perf_counter()
    for (t, sid, _, _), (mi, raw, dv) in zip(base_stream, prepared):
        a = time.perf_counter()
        p_out.append(plain[sid].apply(backend, mi, raw, dv))
        p_lat.append((time.perf_counter() - a) * 1e3)
    plain_rate = count / (time.perf_counter() - t0)
    # Plain path with the reference's per-update verification (one pairing
    # call per update), on fresh copies of the first requests' sessions.
    fresh: Dict[int, PlainSession] = {}
    pv_lat = []
    pv_ok = True
    for (t, sid, _, _), (mi, raw, dv) in zip(base_stream[:plain_verified_count], prepared[:plain_verified_count]):
        if sid not in fresh:
            fresh[sid] = PlainSession(backend, specs[sid])
        a = time.perf_counter()
        proof = fresh[sid].apply(backend, mi, raw, dv)
        pv_ok &= verify_update_proofs(backend, [proof], [fresh[sid].divisor])[0]
        pv_lat.append((time.perf_counter() - a) * 1e3)
    plain_verified_rate = len(pv_lat) / (sum(pv_lat) * 1e-3)
    log(f"# plain path: {plain_rate:.0f} updates/s, {plain_verified_rate:.1f}/s with per-update verification")
    out = {"sessions": len(specs), "requests_per_run": count,
           "plain_updates_per_s": round(plain_rate, 1),
           "plain_p50_ms": round(percentile(p_lat, 50), 3), "plain_p99_ms": round(percentile(p_lat, 99), 3),
           "plain_verified_updates_per_s": round(plain_verified_rate, 1),
           "plain_verified_p50_ms": round(percentile(pv_lat, 50), 3),
           "plain_verified_p99_ms": round(percentile(pv_lat, 99), 3),
           "plain_verified_sample": plain_verified_count}
    identical, verified, checked = True, 0, 0
    verify_ms = []

    def verify(reqs):
        nonlocal verified, checked
        ok, ms = verify_served(backend, store, reqs)
        verified += sum(ok)
        checked += len(ok)
        verify_ms.append(ms)
        return ms

    # Served, saturated: every request queued at t = 0 (throughput ceiling).
    # Sessions start from the same anchors as the plain ones, so every proof
    # must equal the plain run's.
    srv = Server(store, lanes=lanes, max_batch=256)
    sat_stream = [(0.0, sid, idx, ds) for _, sid, idx, ds in base_stream]
    reqs, lat, span = _drive(srv, sat_stream, [(m, r) for m, r, _ in prepared])
    mean_batch = float(np.mean(srv.batch_sizes))
    srv.close()
    identical &= [r.future.result() for r in reqs] == p_out
    vms = verify(reqs)
    sat_rate = count / span
    sat_verified_rate = count / (span + vms * 1e-3)
    out.update({"served_saturated_updates_per_s": round(sat_rate, 1),
                "served_saturated_p50_ms": round(percentile(lat, 50), 3),
                "served_saturated_p99_ms": round(percentile(lat, 99), 3),
                "mean_batch": round(mean_batch, 1),
                "verify_call_ms": round(vms, 3), "verify_proofs_per_s": round(count / (vms * 1e-3), 1),
                "served_saturated_verified_updates_per_s": round(sat_verified_rate, 1),
                "speedup_vs_plain": round(sat_rate / plain_rate, 2),
                "speedup_vs_plain_both_verified": round(sat_verified_rate / plain_verified_rate, 1)})
    log(f"# served saturated: {sat_rate:.0f} updates/s ({sat_verified_rate:.0f}/s with the verification call), "
        f"mean batch {mean_batch:.1f}")
    # Open loop at set arrival rates (continuing the same sessions' states;
    # the plain sessions replay the same requests in arrival order).
    pts = []
    for q in qps_points or (0.5 * sat_rate, 0.8 * sat_rate, plain_rate):
        stream = request_stream(specs, count, q, seed=int(q) + 7)
        prep = [merge_update(idx, ds) for _, _, idx, ds in stream]
        srv = Server(store, lanes=lanes, max_batch=256)
        reqs, lat, span = _drive(srv, stream, [(m, r) for m, r, _ in prep])
        mean_batch = float(np.mean(srv.batch_sizes))
        srv.close()
        identical &= [r.future.result() for r in reqs] == _plain_replay(plain, stream, prep, backend)
        vms = verify(reqs)
        pts.append({"qps": round(q, 1), "updates_per_s": round(count / span, 1),
                    "p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3),
                    "mean_batch": round(mean_batch, 1), "verify_call_ms": round(vms, 3), **_parts(reqs)})
        log(f"# served at {q:.0f} QPS: p50 {pts[-1]['p50_ms']} ms, p99 {pts[-1]['p99_ms']} ms")
    out["open_loop"] = pts
    out.update({"verified_all": bool(verified == checked == count * (1 + len(pts)) and pv_ok),
                "proofs_verified": verified, "proofs_checked": checked,
                "identical_to_plain": bool(identical), "proofs_compared": count * (1 + len(pts))})
    store.close()
    return out


# ---- opening-cache precompute (crates/protocol/src/kzg.rs:217-275) ----------

GLV_PRODUCTS_UNIFORM = 126 * 9 + 0.75 * 126 * 14 + 15   # shared twiddle: adds only on set digits
GLV_PRODUCTS_LANE = 126 * (9 + 14) + 15                  # per-lane scalar: every step adds


def _gfft_products(L: int, rows: int) -> float:
    """Montgomery products of one rows x 2^L G1 transform under the stage
    plan of kzg.hip: 2 additions per butterfly, a GLV scalar product for
    every non-unit twiddle (stage twiddles wave-shared, the diagonal per lane)."""
    if L == 0:
        return 0.0
    k = 0 if rows >= 64 else (6 if L >= 12 else L // 2)
    m = L - k
    half = rows << (L - 1)
    prods = 0.0
    for s in range(L):
        lo = m if s >= m else 0
        nontrivial = half * (1 - 2.0 ** -(s - lo))
        prods += half * 28 + nontrivial * GLV_PRODUCTS_UNIFORM
    if k and m:
        prods += (rows << L) * (1 - 2.0 ** -k) * (1 - 2.0 ** -m) * GLV_PRODUCTS_LANE
    return prods


def opening_cache_macs(log_n: int, rows: int, with_basis: bool) -> float:
    n = 1 << log_n
    d = n - 1
    LM = max(0, (2 * d - 2).bit_length()) if d > 1 else 0
    prods = _gfft_products(LM, rows) + _gfft_products(log_n, rows) + (rows << LM) * GLV_PRODUCTS_LANE
    if with_basis:
        prods += _gfft_products(LM, 1)
    return 136.0 * prods


def opening_cache(backend: HipBackend, log_n: int = 16, rows: int = 3, reps: int = 3, log=print) -> Dict:
    """The anchor opening caches q4, q5, q6 (dynaverse.rs:303-314): three
    polynomials of n coefficients over the n-point domain, one call.  The
    first call builds the retained SRS transform; later calls reuse it.
    Sampled openings are checked against the trapdoor form
    [(f(tau) - f(w^i)) / (tau - w^i)] g1 computed with the fixed-base kernel."""
    from .codec import FR_MODULUS, encode_fr, root_of_unity_2_20
    from .kzg import KzgSrs
    from .workload import poly_eval, srs_scalars
    n = 1 << log_n
    r = FR_MODULUS
    tau, powers = srs_scalars(2024, n)
    srs = b"".join(backend.fixed_base_encoded(powers[i:i + 4096].tobytes(), min(4096, n - i), False)
                   for i in range(0, n, 4096))
    coeffs = ChaCha20Stream(2025).uniform_fr_array(rows * n)
    data = coeffs.tobytes()
    with KzgSrs(backend, encoded=srs) as z:
        a = time.perf_counter()
        raw = z.precompute_bytes(data, n, rows, n)
        cold_ms = (time.perf_counter() - a) * 1e3
        cold_t = dict(z.last_timings)
        warm, phases = [], []
        for _ in range(reps):
            a = time.perf_counter()
            again = z.precompute_bytes(data, n, rows, n)
            warm.append((time.perf_counter() - a) * 1e3)
            phases.append(dict(z.last_timings))
        same = again == raw
    w = pow(root_of_unity_2_20(), 1 << (20 - log_n), r)
    ok = True
    for row in range(rows):
        f = coeffs[row * n:(row + 1) * n]
        ft = poly_eval(f, tau)
        for i in (0, 1, n // 3, n - 1):
            z_i = pow(w, i, r)
            q = (ft - poly_eval(f, z_i)) * pow((tau - z_i) % r, r - 2, r) % r
            exp = backend.fixed_base_encoded(encode_fr(q), 1, False)
            ok &= raw[64 * (row * n + i):64 * (row * n + i + 1)] == exp
    warm_ms = sorted(warm)[len(warm) // 2]
    dev_warm = min(sum(p.values()) for p in phases)
    macs = opening_cache_macs(log_n, rows, False)
    out = {"log_n": log_n, "caches": rows, "cold_ms": round(cold_ms, 3), "warm_ms": round(warm_ms, 3),
           "warm_device_ms": round(dev_warm, 3), "cold_phases_ms": {k: round(v, 3) for k, v in cold_t.items()},
           "warm_phases_ms": {k: round(v, 3) for k, v in min(phases, key=lambda p: sum(p.values())).items()},
           "caches_per_s": round(rows / (warm_ms * 1e-3), 2),
           "mac_frac": round(macs / (dev_warm * 1e-3) / PEAK_MAC_PER_S, 4),
           "verified_sampled_against_trapdoor": bool(ok), "repeat_identical": bool(same)}
    log(f"opening cache 2^{log_n} x{rows}: cold {cold_ms:.1f} ms, warm {warm_ms:.1f} ms "
        f"(device {dev_warm:.1f} ms, {out['mac_frac']:.3f} of MAC peak), ok={ok}")
    return out


# ---- opening-cache validation (rebuilding/machine.rs:357-382) ---------------

def pairing_cache_check(backend: HipBackend, log_n: int = 16, caches: int = 3, reps: int = 3, log=print) -> Dict:
    """validate_initial_caches (machine.rs:357-382) at n = 2^log_n, end to end
    on the device: three caches of one anchor computed by the opening-cache
    kernels, their commitments [f(tau)] g1, the witness values f(w^i) by an Fr
    NTT, then every (wire, i) checked by e(C - [v]g1, g2) == e(pi, [tau - z]g2)
    with the divisors formed on the device once per point.  A second call
    plants one wrong cache point per 97 checks, all of which must be found."""
    from .codec import FR_MODULUS, encode_fr, root_of_unity_2_20
    from .kzg import KzgSrs
    from .pairing import opening_cache_check_encoded
    from .workload import poly_eval, srs_scalars
    r = FR_MODULUS
    n = 1 << log_n
    tau, powers = srs_scalars(2026, n)
    srs = b"".join(backend.fixed_base_encoded(powers[i:i + 4096].tobytes(), min(4096, n - i), False)
                   for i in range(0, n, 4096))
    coeffs = ChaCha20Stream(2027).uniform_fr_array(caches * n)
    with KzgSrs(backend, encoded=srs) as z:
        cache_bytes = z.precompute_bytes(coeffs.tobytes(), n, caches, n)
    values = b"".join(backend.ntt_encoded(coeffs[w * n:(w + 1) * n].tobytes(), n, False) for w in range(caches))
    commits = b"".join(backend.fixed_base_encoded(encode_fr(poly_eval(coeffs[w * n:(w + 1) * n], tau)), 1, False)
                       for w in range(caches))
    w_n = pow(root_of_unity_2_20(), 1 << (20 - log_n), r)
    pts, x = bytearray(), 1
    for _ in range(n):
        pts += x.to_bytes(32, "little")
        x = x * w_n % r
    g1_0 = srs[:64]
    g2_0, _ = backend.fixed_base_g2_checked_encoded(encode_fr(1), 1, 3)
    g2_1, _ = backend.fixed_base_g2_checked_encoded(encode_fr(tau), 1, 3)
    args = (commits, caches, g1_0, g2_0, g2_1, bytes(pts), n, values)
    times, walls = [], []
    for _ in range(reps):
        a = time.perf_counter()
        res, ms = opening_cache_check_encoded(backend, *args, cache_bytes)
        walls.append((time.perf_counter() - a) * 1e3)
        times.append(ms)
    all_pass = res == b"\x01" * (caches * n)
    bad = bytearray(cache_bytes)
    planted = list(range(5, caches * n, 97))
    for t in planted:                      # pi -> identity (g1 where pi is the identity)
        bad[64 * t:64 * t + 64] = (1).to_bytes(32, "little") + (2).to_bytes(32, "little") \
            if cache_bytes[64 * t:64 * t + 64] == bytes(64) else bytes(64)
    res_bad, _ = opening_cache_check_encoded(backend, *args, bytes(bad))
    found = [t for t in range(caches * n) if res_bad[t] != 1]
    kms = sorted(times)[len(times) // 2]
    out = {"log_n": log_n, "wires": caches, "checks": caches * n, "kernel_ms": round(kms, 3),
           "wall_ms": round(sorted(walls)[len(walls) // 2], 3), "checks_per_s": round(caches * n / (kms * 1e-3)),
           "all_valid_accepted": bool(all_pass), "planted_rejections": len(planted),
           "planted_found_exactly": found == planted}
    log(f"opening-cache check n=2^{log_n} x{caches}: kernel {kms:.1f} ms, wall {out['wall_ms']:.1f} ms "
        f"({out['checks_per_s']:.0f} checks/s), valid={all_pass}, planted found={found == planted}")
    return out


def kzg_verify_latency(backend: HipBackend, counts=(1, 16, 256, 4096, 16384), reps: int = 5, log=print) -> Dict:
    """dvgpu_bn254_kzg_verify_batch (the pairing check of an update proof,
    dynaverse.rs:447-451) by call size: proofs e(C - [y] g1, g2) == e(W, D)
    built from scalars (C = [a] g1, W = [(a - y)/(tau - z)] g1,
    D = [tau - z] g2, points by the fixed-base kernels) over 64 divisors,
    every third proof with a changed value (rejected).  Calls of up to 8192
    proofs run one wave per proof, larger ones one lane per proof; the
    divisors' prepared lines are cached after the first call (the first call
    of a size is not timed)."""
    from .codec import FR_MODULUS, encode_fr
    from .pairing import kzg_verify_encoded
    r, tau, distinct = FR_MODULUS, 0x5DEECE66D1234567, 64
    z = [(977 * i + 13) % r for i in range(distinct)]
    a = [(7919 * i + 5) % r for i in range(distinct)]
    y = [(104729 * i + 1) % r for i in range(distinct)]
    w = [(a[i] - y[i]) * pow(tau - z[i], r - 2, r) % r for i in range(distinct)]
    enc = lambda vals: b"".join(encode_fr(v) for v in vals)  # noqa: E731
    d = backend.fixed_base_encoded(enc([(tau - zz) % r for zz in z]), distinct, True)
    c = backend.fixed_base_encoded(enc(a), distinct, False)
    wp = backend.fixed_base_encoded(enc(w), distinct, False)
    out = {}
    for count in counts:
        idx = [i % distinct for i in range(count)]
        vals = enc([(y[j] + (1 if i % 3 == 2 else 0)) % r for i, j in enumerate(idx)])
        want = bytes(0 if i % 3 == 2 else 1 for i in range(count))
        cs = b"".join(c[64 * j:64 * j + 64] for j in idx)
        ws = b"".join(wp[64 * j:64 * j + 64] for j in idx)
        ds = b"".join(d[128 * j:128 * j + 128] for j in idx)
        kzg_verify_encoded(backend, cs, vals, ws, ds, count)  # lines of the divisors prepared
        kms, walls, ok = [], [], True
        for _ in range(reps):
            t0 = time.perf_counter()
            res, ms = kzg_verify_encoded(backend, cs, vals, ws, ds, count)
            walls.append((time.perf_counter() - t0) * 1e3)
            kms.append(ms)
            ok = ok and res == want
        km = statistics.median(kms)
        out[str(count)] = {"kernel_ms": round(km, 3), "wall_ms": round(statistics.median(walls), 3),
                           "proofs_per_s": round(count / (km * 1e-3), 1), "decisions_exact": ok,
                           "path": "one wave per proof" if count <= 8192 else "one lane per proof"}
        log(f"KZG verification, {count} proofs: {km:.3f} ms ({out[str(count)]['path']}), decisions exact={ok}")
    return out


def pairing_equations_mix(backend: HipBackend, distinct: int = 8, equations: int = 65536, reps: int = 3,
                          log=print) -> Dict:
    """The verifiers' pairing equations (verification.rs PairingChecks::equal)
    through dvgpu_bn254_pairing_equations in both modes: the shapes of
    dynamo.rs:423-429 (5 and 2 pairs), dynaverse.rs:1056-1058 and
    consistency.rs:101 (3 pairs), dynaverse.rs:1083 and machine.rs:335
    (2 pairs); `distinct` instances of each, built to hold from scalars
    (ab == sum c_j d_j) with G1/G2 points from the fixed-base kernels, each
    with a failing twin (a + 1), the set tiled to about `equations`
    equations in one call (one device thread per equation)."""
    from .codec import FR_MODULUS, encode_fr
    from .verification import VerificationMode, pairing_equations_encoded
    r = FR_MODULUS
    rng = ChaCha20Stream(2031)

    def rnd():
        return int.from_bytes(rng.uniform_fr_array(1)[0].tobytes(), "little") or 1

    shapes = [(4, False), (1, False), (2, False), (1, True), (2, True), (1, True)]   # (right pairs, b = 1)
    g1s, g2s, sizes, expect = [], [], [], []
    for right, unit_b in shapes:
        for _ in range(distinct):
            b = 1 if unit_b else rnd()
            cd = [(rnd(), rnd()) for _ in range(right)]
            a = sum(c * d for c, d in cd) % r * pow(b, -1, r) % r
            for holds in (True, False):
                g1s += [a if holds else (a + 1) % r] + [c for c, _ in cd]
                g2s += [b] + [d for _, d in cd]
                sizes.append(1 + right)
                expect.append(holds)
    g1 = backend.fixed_base_encoded(b"".join(encode_fr(v) for v in g1s), len(g1s), False)
    g2 = backend.fixed_base_encoded(b"".join(encode_fr(v) for v in g2s), len(g2s), True)
    tile = -(-equations // len(sizes))
    offsets, pairs = [0], 0
    for _ in range(tile):
        for k in sizes:
            pairs += k
            offsets.append(pairs)
    g1, g2, expect = g1 * tile, g2 * tile, expect * tile
    want = bytes(int(e) for e in expect)
    count = len(expect)
    out = {"equations": count, "pairs": pairs, "shapes_pairs": [1 + s for s, _ in shapes]}
    for mode in (VerificationMode.PRODUCT, VerificationMode.REFERENCE):
        times, walls, ok = [], [], True
        for _ in range(reps):
            a = time.perf_counter()
            res, ms = pairing_equations_encoded(backend, g1, g2, offsets, mode)
            walls.append((time.perf_counter() - a) * 1e3)
            times.append(ms)
            ok = ok and res == want
        kms = sorted(times)[len(times) // 2]
        name = mode.name.lower()
        out[name] = {"kernel_ms": round(kms, 3), "wall_ms": round(sorted(walls)[len(walls) // 2], 3),
                     "equations_per_s": round(count / (kms * 1e-3)), "decisions_exact": bool(ok)}
        log(f"pairing equations ({name} mode): {count} equations / {pairs} pairs in {kms:.1f} ms "
            f"({out[name]['equations_per_s']} eq/s), decisions exact={ok}")
    # One verification's worth (the first 16 equations: 5-pair shapes and
    # their twins) in one call: one wave per equation, the G2 lines prepared
    # by an untimed first call.
    small = offsets[16]
    pairing_equations_encoded(backend, g1[:64 * small], g2[:128 * small], offsets[:17], VerificationMode.PRODUCT)
    times = []
    for _ in range(reps):
        res, ms = pairing_equations_encoded(backend, g1[:64 * small], g2[:128 * small], offsets[:17],
                                            VerificationMode.PRODUCT)
        times.append(ms)
    out["small_call_16_equations"] = {"pairs": small, "kernel_ms": round(sorted(times)[len(times) // 2], 3),
                                      "decisions_exact": res == want[:16]}
    log(f"pairing equations, one call of 16 five-pair equations: {out['small_call_16_equations']['kernel_ms']} ms")
    return out


# ---- polynomial tools (domain/fast.rs) ---------------------------------------

def poly_tree(backend: HipBackend, log_k: int = 12, reps: int = 3, log=print) -> Dict:
    """SubproductTree over 2^log_k random points: build, multipoint evaluation
    and interpolation of a degree 2^log_k - 1 polynomial, Cauchy sums; values
    checked against Horner at sampled points and the interpolation round trip.
    Beside it, the reference's call pattern (one product call per tree
    product, fast.rs:204-234) for the products of a 2^10 tree, and trees of
    16 and 256 points (their small rounds are formed on the host)."""
    from .codec import FR_MODULUS
    from .poly import SubproductTree, fft_multiply
    r = FR_MODULUS
    k = 1 << log_k
    rng = ChaCha20Stream(2030)
    points = list(dict.fromkeys(rng.uniform_fr(k + 8)))[:k]
    poly = rng.uniform_fr(k)
    weights = rng.uniform_fr(k)

    def med(fn):
        ts = []
        out = None
        for _ in range(reps):
            a = time.perf_counter()
            out = fn()
            ts.append((time.perf_counter() - a) * 1e3)
        return sorted(ts)[len(ts) // 2], out

    t_build, tree = med(lambda: SubproductTree(backend, points))
    build_stats = tree.stats()
    t_eval, vals = med(lambda: tree.multipoint_eval(poly))
    eval_stats = tree.stats()
    t_interp, back = med(lambda: tree.interpolate(vals))
    t_cauchy, _ = med(lambda: tree.cauchy_sums(weights))

    def horner(c, x):
        acc = 0
        for v in reversed(c):
            acc = (acc * x + v) % r
        return acc
    ok = all(vals[i] == horner(poly, points[i]) for i in (0, 1, k // 2, k - 1)) and back == poly
    tree.close()
    # the per-product path: one fft_multiply call per product, level by level
    small = points[:1024]
    a = time.perf_counter()
    level = [[(-x) % r, 1] for x in small]
    calls = 0
    while len(level) > 1:
        nxt = [fft_multiply(backend, level[i], level[i + 1]) for i in range(0, len(level) - 1, 2)]
        calls += len(nxt)
        if len(level) % 2:
            nxt.append(level[-1])
        level = nxt
    t_plain = (time.perf_counter() - a) * 1e3
    a = time.perf_counter()
    with SubproductTree(backend, small) as t2:
        batched_products = t2.stats()
    t_batched = (time.perf_counter() - a) * 1e3
    small_sizes = {}
    for ks in (16, 256):
        sp = points[:ks]
        t_b, st = med(lambda: SubproductTree(backend, sp))
        t_e, sv = med(lambda: st.multipoint_eval(poly[:ks]))
        t_i, sb = med(lambda: st.interpolate(sv))
        small_sizes[str(ks)] = {"build_ms": round(t_b, 3), "eval_ms": round(t_e, 3), "interpolate_ms": round(t_i, 3),
                                "round_trip_ok": sb == poly[:ks]}
        st.close()
    out = {"points": k, "build_ms": round(t_build, 2), "build_rounds": build_stats["rounds"],
           "build_products": build_stats["products"], "eval_ms": round(t_eval, 2),
           "eval_rounds": eval_stats["rounds"], "interpolate_ms": round(t_interp, 2),
           "cauchy_sums_ms": round(t_cauchy, 2), "values_and_round_trip_ok": bool(ok),
           "small_trees": small_sizes,
           "per_product_path_1024": {"products_only_ms": round(t_plain, 2), "calls": calls,
                                     "batched_full_build_ms": round(t_batched, 2),
                                     "batched_rounds": batched_products["rounds"]}}
    log(f"poly tree 2^{log_k}: build {t_build:.1f} ms ({build_stats['rounds']} rounds), eval {t_eval:.1f} ms, "
        f"interpolate {t_interp:.1f} ms, cauchy {t_cauchy:.1f} ms, ok={ok}; 1024-point products one call each "
        f"{t_plain:.1f} ms vs batched build {t_batched:.1f} ms")
    return out


# ---- per-row figures of the single-call entry points -----------------------

def row_figures(backend: HipBackend, reps: int = 5, log=print) -> Dict:
    """One figure per remaining drop-in row, each on device-made synthetic data
    and checked: the G2 MSM over 2^16 points (against [f(tau)] g2 from the
    fixed-base kernel), G2 wire validation of 4096 subgroup points (all
    accepted), and through the resident engine: an ECNTT over 2^12 G1 points
    (forward then inverse restores the input), a resident G1 MSM over 2^16
    bases, a 64-point G1 read normalised on the device, and the rebuild's
    four-step cursor over resident calls."""
    from .codec import FR_MODULUS, encode_fr
    from .resident import FR, G1, ResidentBackend
    from .workload import poly_eval, srs_scalars
    r = FR_MODULUS
    out = {}

    def med(fn):
        ts, res = [], None
        for _ in range(reps):
            a = time.perf_counter()
            res = fn()
            ts.append((time.perf_counter() - a) * 1e3)
        return sorted(ts)[len(ts) // 2], res

    # G2 MSM, 2^16 points, the device-resident entry point
    n = 1 << 16
    tau, powers = srs_scalars(2031, n)
    g2pts = backend.to_device(powers)           # scalars for the fixed-base G2 kernel
    g2_mont = backend.alloc(128 * n)
    backend.fixed_base_device(g2pts.ptr, n, g2_mont.ptr, g2=True)
    sc = ChaCha20Stream(2032).uniform_fr_array(n)
    d_sc = backend.to_device(sc)
    t, res = med(lambda: backend.msm_device(g2_mont.ptr, d_sc.ptr, n, g2=True))
    exp, _ = backend.fixed_base_g2_checked_encoded(encode_fr(poly_eval(sc, tau)), 1, 3)
    out["g2_msm_2p16"] = {"ms": round(t, 3), "mpoints_s": round(n / (t * 1e-3) / 1e6, 2), "ok": res == exp}
    g2pts.free(); g2_mont.free(); d_sc.free()
    # G2 wire validation of 4096 subgroup points (fixed-base outputs as wire bytes)
    m = 4096
    wire, _ = backend.fixed_base_g2_checked_encoded(powers[:m].tobytes(), m, 3)
    t, (masks, rep) = med(lambda: backend.inspect_g2_validation(wire, 3))
    out["g2_validation_4096"] = {"ms": round(t, 3), "points_s": round(m / (t * 1e-3)),
                                 "ok": rep["invalid_points"] == 0 and rep["checked_points"] == m}
    # resident engine
    rb = ResidentBackend(backend)
    k = 1 << 12
    pts_raw = backend.fixed_base_encoded(powers[:k].tobytes(), k, False)
    from .codec import decode_g1
    pts = [decode_g1(pts_raw[64 * i:64 * i + 64]) for i in range(k)]
    src, mid, dst = rb.alloc(G1, k), rb.alloc(G1, k), rb.alloc(G1, k)
    rb.write_points(src, 0, pts)

    def ecntt_round():
        rb.transform_rows(G1, src, mid, k, 0, 1, False)
        rb.transform_rows(G1, mid, dst, k, 0, 1, True)
    t, _ = med(ecntt_round)
    ok = rb.read_points(dst, 0, 8) == pts[:8] and rb.read_points(dst, k - 8, 8) == pts[k - 8:]
    out["resident_ecntt_2p12_fwd_inv"] = {"ms": round(t, 3), "ok": bool(ok)}
    t64, got = med(lambda: rb.read_points(src, 100, 64, device_affine=True))
    out["resident_read_64_points"] = {"ms": round(t64, 4), "ok": got == pts[100:164]}
    nb = 1 << 16
    bases_raw = b"".join(backend.fixed_base_encoded(powers[i:i + 4096].tobytes(), 4096, False)
                         for i in range(0, nb, 4096))
    bases = [decode_g1(bases_raw[64 * i:64 * i + 64]) for i in range(nb)]
    hb = rb.alloc_bases(bases)
    fr = rb.alloc(FR, nb)
    scal = ChaCha20Stream(2033).uniform_fr(nb)
    rb.write_fields(fr, 0, scal)
    t, res = med(lambda: rb.msm_g1(hb, 0, fr, 0, nb))
    f_t = 0
    for c in reversed(scal):
        f_t = (f_t * tau + c) % r
    exp = decode_g1(backend.fixed_base_encoded(encode_fr(f_t), 1, False))
    out["resident_msm_g1_2p16"] = {"ms": round(t, 3), "ok": res == exp}
    # G1 rows at the rebuild's size (n = 65536, program.rs:278-315): one
    # 2^16-point transform forward and back, the round trip checked
    s16, m16, d16 = rb.alloc(G1, nb), rb.alloc(G1, nb), rb.alloc(G1, nb)
    rb.write_points(s16, 0, bases)

    def ecntt16():
        rb.transform_rows(G1, s16, m16, nb, 0, 1, False)
        rb.transform_rows(G1, m16, d16, nb, 0, 1, True)
    t, _ = med(ecntt16)
    ok16 = rb.read_points(d16, 0, 8) == bases[:8] and rb.read_points(d16, nb - 8, 8) == bases[nb - 8:]
    out["resident_ecntt_2p16_fwd_inv"] = {"ms": round(t, 3), "ok": bool(ok16)}
    # The rebuild's four-step cursor (operators.rs:505-597) as the session
    # schedules it (program.rs:84-104): Fr at 2n = 2^17 with tile 512 and
    # row batch 16 (816 resident calls), its initialization plan (12 calls),
    # and G1 at 2^16 with tile 256 and row batch 16; each forward then
    # inverse, the round trip checked.
    from .resident import FourStepPlan
    n17 = 1 << 17
    fv = ChaCha20Stream(2034).uniform_fr(n17)
    fa, fs, fc, fd = (rb.alloc(FR, n17) for _ in range(4))
    rb.write_fields(fa, 0, fv)

    def cursor(kind, a, s, c, d, n, tile, batch, init=False):
        calls = 0
        for src, dst, inv in ((a, c, False), (c, d, True)):
            plan = FourStepPlan(kind, n, src, s, dst, tile, batch, inv)
            calls += (plan.initialization_plan() if init else plan).run(rb)
        return calls
    for name, init in (("resident_four_step_fr_2p17_row16", False), ("resident_four_step_fr_2p17_init", True)):
        t, calls = med(lambda: cursor(FR, fa, fs, fc, fd, n17, 512, 16, init))
        ok = rb.read_fields(fd, 0, 8) == fv[:8] and rb.read_fields(fd, n17 - 8, 8) == fv[n17 - 8:]
        out[name] = {"ms": round(t, 3), "calls": calls, "ok": bool(ok)}
    e16 = rb.alloc(G1, nb)
    t, calls = med(lambda: cursor(G1, s16, m16, d16, e16, nb, 256, 16))
    ok = rb.read_points(e16, 0, 8) == bases[:8] and rb.read_points(e16, nb - 8, 8) == bases[nb - 8:]
    out["resident_four_step_g1_2p16_row16"] = {"ms": round(t, 3), "calls": calls, "ok": bool(ok)}
    log("row figures: " + ", ".join(f"{key} {v['ms']} ms ok={v['ok']}" for key, v in out.items()))
    return out


def main(argv=None):
    """One workload by name, for profiling: python -m dvgpu.benchmarks opening_cache --log-n 16"""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("workload", choices=["opening_cache", "pairing_cache_check", "pairing_equations", "poly_tree",
                                         "row_figures", "batched_launch", "incremental", "served_stream",
                                         "kzg_verify_latency"])
    ap.add_argument("--log-n", type=int, default=16)
    ap.add_argument("--rows", type=int, default=3)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args(argv)
    b = HipBackend()
    try:
        if a.workload == "opening_cache":
            res = opening_cache(b, a.log_n, a.rows, a.reps)
        elif a.workload == "pairing_cache_check":
            res = pairing_cache_check(b, a.log_n, a.rows, a.reps)
        elif a.workload == "pairing_equations":
            res = pairing_equations_mix(b, equations=1 << a.log_n, reps=a.reps)
        elif a.workload == "poly_tree":
            res = poly_tree(b, a.log_n, a.reps)
        elif a.workload == "row_figures":
            res = row_figures(b, a.reps)
        else:
            res = globals()[a.workload](b)
        print(json.dumps(res), flush=True)
    finally:
        b.close()


if __name__ == "__main__":
    main()
