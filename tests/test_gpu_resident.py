"""Resident operator engine on the MI355X, against the CPU resident
semantics of crates/protocol/src/incremental_rebuild/operators.rs:387-590
(restated here with the oracle's field and group arithmetic) and the
reference's native rules (native/src/icicle_resident.cu)."""
import copy

import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.resident import ADD, FR, G1, MUL, SUB, FourStepPlan, ResidentBackend
from dvgpu.rng import ChaCha20Stream
from oracle import bn254 as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rb():
    b = HipBackend()
    r = ResidentBackend(b)
    yield r
    b.close()


def rand_fr(seed, n):
    return ChaCha20Stream(seed).uniform_fr(n)


def rand_pts(seed, n):
    ks = ChaCha20Stream(seed).uniform_fr(n)
    return [None if i % 5 == 3 else o.g1_mul(k) for i, k in enumerate(ks)], ks


def cpu_transpose(values, rows, cols, start, length, twiddle=None, points=False):
    out = {}
    for i in range(start, start + length):
        row, col = i // cols, i % cols
        v = values[i]
        if twiddle is not None:
            f = pow(twiddle, row * col, o.R)
            v = o.mul(v, f) if points else v * f % o.R
        out[col * rows + row] = v
    return out


def test_configuration_rules():
    b = HipBackend()
    root = o.fr_to_bytes(o.ROOT_2_20)
    import ctypes
    err = ctypes.create_string_buffer(256)
    bad_root = o.fr_to_bytes(o.fr_root_of_unity(1 << 19))
    cp = lambda x: ctypes.cast(ctypes.c_char_p(x), ctypes.c_void_p)
    assert b.lib.dvgpu_resident_configure(b.handle, 4 << 30, 64 << 20, cp(bad_root), err, 256) == -1
    # quota too small for the domain + workspace reserve: -1 and retry allowed
    assert b.lib.dvgpu_resident_configure(b.handle, 32 << 20, 0, cp(root), err, 256) == -1
    assert b.lib.dvgpu_resident_configure(b.handle, 1 << 30, 1 << 20, cp(root), err, 256) == 0
    assert b.lib.dvgpu_resident_configure(b.handle, 1 << 30, 1 << 20, cp(root), err, 256) == -1
    b.close()


def test_alloc_write_read_copy_reverse(rb):
    vals = rand_fr(1, 20)
    h = rb.alloc(FR, 32)
    assert rb.buffer_len(h) == 32 and rb.buffer_kind(h) == FR
    assert rb.read_fields(h, 0, 32) == [0] * 32
    rb.write_fields(h, 5, vals)
    assert rb.read_fields(h, 5, 20) == vals
    h2 = rb.alloc(FR, 32)
    rb.copy(FR, h, 5, h2, 0, 20)
    assert rb.read_fields(h2, 0, 20) == vals
    rb.reverse(FR, h, 5, h2, 10, 20)
    assert rb.read_fields(h2, 10, 20) == vals[::-1]
    with pytest.raises(BackendError) as e:
        rb.copy(FR, h, 0, h, 1, 5)            # overlapping ranges
    assert e.value.status == -1
    with pytest.raises(BackendError):
        rb.read_fields(h, 30, 5)              # out of bounds
    pts, _ = rand_pts(2, 9)
    g = rb.alloc(G1, 9)
    assert rb.read_points(g, 0, 9) == [None] * 9
    rb.write_points(g, 0, pts)
    assert rb.read_points(g, 0, 9) == pts
    assert rb.read_points(g, 0, 9, device_affine=True) == pts
    g2 = rb.alloc(G1, 9)
    rb.reverse(G1, g, 0, g2, 0, 9)
    assert rb.read_points(g2, 0, 9) == pts[::-1]


def test_handles_are_owner_tagged(rb):
    other = HipBackend()
    orb = ResidentBackend(other, limit=1 << 30, journal=1 << 20)
    h = orb.alloc(FR, 4)
    with pytest.raises(BackendError) as e:
        rb.read_fields(h, 0, 4)
    assert e.value.status == -1
    other.close()


def test_zip_and_scale(rb):
    a, b_ = rand_fr(3, 17), rand_fr(4, 17)
    ha, hb, hc = rb.alloc(FR, 17), rb.alloc(FR, 17), rb.alloc(FR, 17)
    rb.write_fields(ha, 0, a)
    rb.write_fields(hb, 0, b_)
    for op, f in ((ADD, lambda x, y: (x + y) % o.R), (SUB, lambda x, y: (x - y) % o.R), (MUL, lambda x, y: x * y % o.R)):
        rb.field_zip(op, ha, 0, hb, 0, hc, 0, 17)
        assert rb.read_fields(hc, 0, 17) == [f(x, y) for x, y in zip(a, b_)]
    pts, _ = rand_pts(5, 6)
    hp, hs, ho = rb.alloc(G1, 6), rb.alloc(FR, 6), rb.alloc(G1, 6)
    sc = rand_fr(6, 6)
    rb.write_points(hp, 0, pts)
    rb.write_fields(hs, 0, sc)
    rb.scale_points(hp, 0, hs, 0, ho, 0, 6)
    assert rb.read_points(ho, 0, 6) == [o.mul(p, s) for p, s in zip(pts, sc)]


@pytest.mark.parametrize("twiddle", [None, "root"])
def test_transpose_fr_and_g1(rb, twiddle):
    rows, cols = 4, 8
    vals = rand_fr(7, rows * cols)
    tw = o.fr_root_of_unity(32) if twiddle else None
    src, dst = rb.alloc(FR, 32), rb.alloc(FR, 32)
    rb.write_fields(src, 0, vals)
    rb.transpose(FR, src, dst, rows, cols, 3, 20, tw)     # partial range
    got = rb.read_fields(dst, 0, 32)
    exp = cpu_transpose(vals, rows, cols, 3, 20, tw)
    for k, v in exp.items():
        assert got[k] == v, k
    pts, _ = rand_pts(8, 8)
    ps, pd = rb.alloc(G1, 8), rb.alloc(G1, 8)
    rb.write_points(ps, 0, pts)
    rb.transpose(G1, ps, pd, 2, 4, 0, 8, tw)
    got = rb.read_points(pd, 0, 8)
    exp = cpu_transpose(pts, 2, 4, 0, 8, tw, points=True)
    assert [got[k] for k in sorted(exp)] == [exp[k] for k in sorted(exp)]


def test_transform_rows_fr_and_g1(rb):
    vals = rand_fr(9, 64)
    src, dst = rb.alloc(FR, 64), rb.alloc(FR, 64)
    rb.write_fields(src, 0, vals)
    for inverse in (False, True):
        rb.transform_rows(FR, src, dst, 16, 1, 3, inverse)     # rows 1..3 of length 16
        got = rb.read_fields(dst, 16, 48)
        exp = sum((o.ntt(vals[16 * r:16 * r + 16], inverse) for r in range(1, 4)), [])
        assert got == exp
    pts, _ = rand_pts(10, 16)
    ps, pd = rb.alloc(G1, 16), rb.alloc(G1, 16)
    rb.write_points(ps, 0, pts)
    for inverse in (False, True):
        rb.transform_rows(G1, ps, pd, 8, 0, 2, inverse)
        got = rb.read_points(pd, 0, 16)
        assert got == o.ecntt(pts[:8], inverse) + o.ecntt(pts[8:], inverse)


@pytest.mark.parametrize("n", [16, 64, 4096])
def test_four_step_packet_fr_equals_full_ntt(rb, n):
    vals = rand_fr(11 + n, n)
    a, s, c = rb.alloc(FR, n), rb.alloc(FR, n), rb.alloc(FR, n)
    rb.write_fields(a, 0, vals)
    for inverse in (False, True):
        rb.transform_four_step(FR, a, s, c, n, inverse)
        assert rb.read_fields(c, 0, n) == o.ntt(vals, inverse), (n, inverse)


def test_four_step_packet_g1(rb):
    pts, _ = rand_pts(12, 16)
    a, s, c = rb.alloc(G1, 16), rb.alloc(G1, 16), rb.alloc(G1, 16)
    rb.write_points(a, 0, pts)
    rb.transform_four_step(G1, a, s, c, 16, False)
    fwd = rb.read_points(c, 0, 16)
    assert fwd == o.ecntt(pts, False)
    rb.write_points(a, 0, fwd)
    rb.transform_four_step(G1, a, s, c, 16, True)
    assert rb.read_points(c, 0, 16) == pts


def drain(plan, rb):
    steps = 0
    while not plan.next_step(rb):
        steps += 1
        assert steps < 100_000
    return steps + 1


def test_four_step_cursor_session_65536_rebuild_length(rb):
    """The rebuild's 2^17 transform of an n = 65536 session as the reference
    schedules it (program.rs:84-104: tile 512, row batch 16; 816 resident
    calls) against the oracle's NTT, and back."""
    n = 1 << 17
    vals = rand_fr(40, n)
    a, s, c, d = rb.alloc(FR, n), rb.alloc(FR, n), rb.alloc(FR, n), rb.alloc(FR, n)
    rb.write_fields(a, 0, vals)
    fwd = FourStepPlan(FR, n, a, s, c, 512, 16, False)
    assert drain(fwd, rb) == fwd.total_steps == 816
    got = rb.read_fields(c, 0, n)
    assert got == o.ntt(vals)
    inv = FourStepPlan(FR, n, c, s, d, 512, 16, True)
    drain(inv, rb)
    assert rb.read_fields(d, 0, n) == vals


@pytest.mark.parametrize("tile,row_batch", [(64, 1), (100, 4), (4096, 16), (4096, 128)])
def test_four_step_cursor_equals_packet(rb, tile, row_batch):
    n = 4096
    vals = rand_fr(41, n)
    a, s, c, p = rb.alloc(FR, n), rb.alloc(FR, n), rb.alloc(FR, n), rb.alloc(FR, n)
    rb.write_fields(a, 0, vals)
    for inverse in (False, True):
        FourStepPlan(FR, n, a, s, p, tile, row_batch, inverse).execute_packet(rb)
        drain(FourStepPlan(FR, n, a, s, c, tile, row_batch, inverse), rb)
        assert rb.read_fields(c, 0, n) == rb.read_fields(p, 0, n), (tile, row_batch, inverse)


def test_four_step_cursor_g1(rb):
    """Group transform by the cursor (every 17th point the identity, as in
    the reference's packet probe, transform_packet.rs:218) against the
    oracle's group FFT, forward and inverse."""
    n = 64
    ks = rand_fr(42, n)
    pts = [None if i % 17 == 0 else o.g1_mul(k) for i, k in enumerate(ks)]
    a, s, c = rb.alloc(G1, n), rb.alloc(G1, n), rb.alloc(G1, n)
    rb.write_points(a, 0, pts)
    for inverse in (False, True):
        plan = FourStepPlan(G1, n, a, s, c, 8, 16, inverse)
        assert drain(plan, rb) == plan.total_steps
        assert rb.read_points(c, 0, n) == o.ecntt_fft(pts, inverse), inverse


def test_four_step_committed_cursor_retried_after_rollback():
    """operators_tests.rs:280-303 on the device: 15 steps of a tentative
    copy inside a transaction, rollback restores the zeroed output, and the
    committed cursor then runs to the natural-order transform."""
    b = HipBackend()
    r = ResidentBackend(b, limit=1 << 30, journal=1 << 20)
    a, s, c = r.alloc(FR, 32), r.alloc(FR, 32, True), r.alloc(FR, 32, True)
    vals = [(i * i + 3 * i + 11) % o.R for i in range(32)]
    r.write_fields(a, 0, vals)
    r.seal()
    committed = FourStepPlan(FR, 32, a, s, c, 3, 2, False)
    r.begin_transaction()
    tentative = copy.copy(committed)
    for _ in range(15):
        assert not tentative.next_step(r)
    with pytest.raises(BackendError):
        r.write_fields(c, 32, [1])                 # out of bounds: the call fails
    r.rollback()
    assert r.read_fields(c, 0, 32) == [0] * 32
    r.begin_transaction()
    drain(committed, r)
    r.commit_transaction()
    assert r.read_fields(c, 0, 32) == o.ntt(vals)
    b.close()


def test_resident_msms(rb):
    pts, ks = rand_pts(13, 40)
    sc = rand_fr(14, 40)
    exp = o.g1_mul(sum(s * k for i, (s, k) in enumerate(zip(sc, ks)) if i % 5 != 3) % o.R)
    hp, hs = rb.alloc(G1, 40), rb.alloc(FR, 40)
    rb.write_points(hp, 0, pts)
    rb.write_fields(hs, 0, sc)
    assert rb.msm_g1(hp, 0, hs, 0, 40) == exp
    hb = rb.alloc_bases(pts)
    assert rb.buffer_kind(hb) == 2
    assert rb.msm_g1(hb, 0, hs, 0, 40) == exp
    assert rb.read_points(hb, 0, 40) == pts
    with pytest.raises(BackendError):
        rb.write_points(hb, 0, pts[:1])      # packed bases are immutable
    g2 = [o.g2_mul(k) for k in ks[:3]]
    assert rb.msm_g2(g2, hs, 0) == o.straight_msm(g2, sc[:3], o.FQ2)
    st = rb.stats()
    assert st["msm_g1_points"] >= 80 and st["msm_g2_points"] >= 3


def test_transactions_rollback_commit_and_seal():
    b = HipBackend()
    rb = ResidentBackend(b, limit=1 << 30, journal=4096)
    h = rb.alloc(FR, 100, writable_after_seal=True)
    frozen = rb.alloc(FR, 10)
    base = rand_fr(15, 100)
    rb.write_fields(h, 0, base)
    rb.seal()
    with pytest.raises(BackendError):
        rb.alloc(FR, 4)                      # allocation after seal
    with pytest.raises(BackendError):
        rb.write_fields(frozen, 0, [1])      # immutable after seal
    rb.begin_transaction()
    rb.write_fields(h, 10, [7] * 20)
    rb.write_fields(h, 15, [9] * 20)         # overlapping second write
    rb.rollback()
    assert rb.read_fields(h, 0, 100) == base
    rb.begin_transaction()
    rb.write_fields(h, 0, [1] * 30)
    rb.commit_transaction()
    assert rb.read_fields(h, 0, 30) == [1] * 30
    rb.begin_transaction()
    with pytest.raises(BackendError) as e:
        rb.write_fields(h, 0, [2] * 100)     # 3200 B of journal needed > what is left? fits 4096
        rb.write_fields(h, 0, [3] * 100)     # already saved: no new journal bytes
        h3 = rb.alloc(FR, 1)                 # allocation inside a transaction
    assert e.value.status == -1
    rb.rollback()
    assert rb.read_fields(h, 0, 30) == [1] * 30
    d = rb.diagnostics()
    assert d["journal_saved_bytes"] > 0 and d["device_operations"] > 0
    b.close()


def test_journal_exhaustion_rejects_before_mutation():
    b = HipBackend()
    rb = ResidentBackend(b, limit=1 << 30, journal=64)   # two Fr elements
    h = rb.alloc(FR, 8)
    rb.write_fields(h, 0, list(range(8)))
    rb.begin_transaction()
    with pytest.raises(BackendError) as e:
        rb.write_fields(h, 0, [5] * 3)
    assert e.value.status == -1
    assert rb.read_fields(h, 0, 8) == list(range(8))
    rb.rollback()
    b.close()


def test_profiled_call(rb):
    h = rb.alloc(FR, 8)
    _, prof = rb.call(6, kind=FR, a=h, length=0, profile=True)
    assert prof["native_wall_ns"] > 0


def test_transaction_refusals_are_whole_operations():
    """A transpose whose destination runs need more journal than is left, and
    a packet whose scratch would fit but whose output would not, are refused
    before anything is saved or written (reference: transform_packet.rs:131,
    whole-packet preflight)."""
    b = HipBackend()
    rb = ResidentBackend(b, limit=1 << 30, journal=32 * 24)      # 24 Fr elements
    vals = rand_fr(21, 32)
    src, dst = rb.alloc(FR, 32), rb.alloc(FR, 32)
    rb.write_fields(src, 0, vals)
    rb.begin_transaction()
    d0 = rb.diagnostics()
    print("case: transpose refused by the journal", flush=True)
    with pytest.raises(BackendError) as e:
        rb.transpose(FR, src, dst, 4, 8, 0, 32)                   # 32 destinations > 24
    assert e.value.status == -1 and "journal" in e.value.message
    d1 = rb.diagnostics()
    assert (d1["journal_saved_bytes"], d1["journal_saved_ranges"]) == (d0["journal_saved_bytes"],
                                                                     d0["journal_saved_ranges"])
    assert rb.read_fields(dst, 0, 32) == [0] * 32
    rb.transpose(FR, src, dst, 4, 8, 3, 20)                       # 20 destinations fit
    d2 = rb.diagnostics()
    assert d2["journal_saved_bytes"] - d1["journal_saved_bytes"] == 20 * 32
    assert d2["journal_saved_ranges"] - d1["journal_saved_ranges"] == 8   # one run per column
    got = rb.read_fields(dst, 0, 32)
    for k, v in cpu_transpose(vals, 4, 8, 3, 20).items():
        assert got[k] == v
    rb.rollback()
    assert rb.read_fields(dst, 0, 32) == [0] * 32
    a, s, c = rb.alloc(FR, 16), rb.alloc(FR, 16), rb.alloc(FR, 16)
    rb.write_fields(a, 0, vals[:16])
    rb.begin_transaction()
    d3 = rb.diagnostics()
    print("case: packet refused by the journal", flush=True)
    with pytest.raises(BackendError) as e:
        rb.transform_four_step(FR, a, s, c, 16, False)           # 16 + 16 elements > 24
    assert e.value.status == -1
    assert rb.diagnostics()["journal_saved_bytes"] == d3["journal_saved_bytes"]
    assert rb.read_fields(s, 0, 16) == [0] * 16 and rb.read_fields(c, 0, 16) == [0] * 16
    rb.rollback()
    rb.transform_four_step(FR, a, s, c, 16, False)               # no transaction: no journal
    assert rb.read_fields(c, 0, 16) == o.ntt(vals[:16], False)
    b.close()


def test_quota_admission_keeps_owner_healthy():
    """Configuration charges the twiddle tables and the journal arena; an
    allocation whose bytes plus the workspace of operations over it exceed
    the quota is refused with -1 and a message naming the quota, and the
    owner keeps working (reference: resident_configuration.rs:118,
    setup_tables.rs:58)."""
    b = HipBackend()
    rb = ResidentBackend(b, limit=256 << 20, journal=1 << 20)
    assert rb.stats()["bytes_allocated"] >= (32 << 20) + (1 << 20)
    h = rb.alloc(FR, 1 << 16)
    print("case: resident allocation over the quota", flush=True)
    with pytest.raises(BackendError) as e:
        rb.alloc(G1, 1 << 20)
    assert e.value.status == -1 and "quota" in e.value.message
    assert rb.buffer_len(h) == 1 << 16
    rb.write_fields(h, 7, [5])
    assert rb.read_fields(h, 7, 1) == [5]
    b.close()


def test_profiled_write_splits_journal_and_work():
    import ctypes
    from dvgpu.codec import encode_fr
    from dvgpu.resident import WRITE, ZIP
    b = HipBackend()
    rb = ResidentBackend(b, limit=1 << 30, journal=1 << 16)
    h = rb.alloc(FR, 64)
    buf = ctypes.create_string_buffer(b"".join(encode_fr(v) for v in rand_fr(22, 64)))
    rb.begin_transaction()
    _, prof = rb.call(WRITE, kind=FR, a=h, length=64, input=ctypes.addressof(buf), profile=True)
    assert prof["event_mask"] == 7                                # journal, work and whole-call events
    assert prof["h2d_bytes"] == 64 * 32 and prof["d2d_bytes"] == 64 * 32
    assert prof["event_total_ns"] > 0 and prof["native_wall_ns"] >= prof["wait_host_ns"]
    _, prof = rb.call(WRITE, kind=FR, a=h, length=64, input=ctypes.addressof(buf), profile=True)
    assert prof["event_mask"] == 6 and prof["d2d_bytes"] == 0    # already saved: no journal work
    rb.rollback()
    assert rb.read_fields(h, 0, 64) == [0] * 64
    with pytest.raises(BackendError):
        rb.call(ZIP, kind=FR, a=h, b=h, c=h, length=0, profile=True)
    b.close()
