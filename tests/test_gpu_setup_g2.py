"""Fixed-base SRS setup and G2 subgroup validation on the MI355X.

Cases follow crates/gpu/tests/setup_tables.rs (lengths 1/127/128/129/4097,
scalars 0, 1, -1, all three setup modes), g2_setup_acceptance.rs and
crates/gpu/src/g2_validation/tests.rs (acceptance corpus, predicate
equivalence over subgroup / non-subgroup points, rejection reports)."""
import pytest

from dvgpu import BackendError, HipBackend
from dvgpu.rng import ChaCha20Stream
from oracle import bn254 as o
from tests.g2corpus import corpus, encoding_cases, twist_cofactor_point

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def backend():
    b = HipBackend()
    yield b
    b.close()


def test_fixed_base_g1_special_scalars_and_modes(backend):
    specials = [0, 1, o.R - 1, 2, 12345]
    exp = [o.g1_mul(s) for s in specials]
    for mode in (0, 1, 2):
        print(f"case: g1 mode={mode}", flush=True)
        backend.set_setup_mode(mode)
        assert backend.fixed_base_g1(specials) == exp, mode
    backend.set_setup_mode(2)
    with pytest.raises(BackendError):
        backend.set_setup_mode(3)


@pytest.mark.parametrize("n", [1, 127, 128, 129, 4097])
def test_fixed_base_g1_lengths(backend, n):
    print(f"case: g1 fixed base n={n}", flush=True)
    rng = ChaCha20Stream(300 + n)
    sc = rng.uniform_fr(n)
    got = backend.fixed_base_g1(sc)
    assert len(got) == n
    # Oracle on a spread of indices; every point decoded with the on-curve check.
    for i in sorted({0, n - 1, n // 2, 127 % n, 128 % n, 4096 % n}):
        assert got[i] == o.g1_mul(sc[i]), i
    # Linear closed form over all points: sum c_i [s_i]G = [sum c_i s_i] G.
    cs = rng.uniform_fr(n)
    assert backend.msm(got, cs) == o.g1_mul(sum(c * s for c, s in zip(cs, sc)) % o.R)


def test_fixed_base_g2_matches_oracle_and_modes(backend):
    rng = ChaCha20Stream(41)
    sc = [0, 1, o.R - 1] + rng.uniform_fr(6)
    exp = [o.g2_mul(s) for s in sc]
    for mode in (0, 1, 2):
        print(f"case: g2 mode={mode}", flush=True)
        backend.set_setup_mode(mode)
        for alg in (1, 2, 3):
            assert backend.fixed_base_g2(sc, algorithm=alg) == exp, (mode, alg)
    backend.set_setup_mode(2)
    st = backend.setup_stats()
    assert st["table_builds"] == [1, 1] and st["points"][1] >= len(sc)


def test_fixed_base_g2_checked_rejects_corruption(backend):
    rng = ChaCha20Stream(42)
    sc = rng.uniform_fr(10)
    raw, rep = backend.fixed_base_g2_checked_encoded(b"".join(o.fr_to_bytes(s) for s in sc), 10, 3)
    assert rep == {"checked_points": 10, "invalid_points": 0, "first_invalid": (1 << 64) - 1,
                   "failure_mask": 0, "algorithm": 3}
    assert backend.g2_test_corrupt_next(7) == 0
    raw, rep = backend.fixed_base_g2_checked_encoded(b"".join(o.fr_to_bytes(s) for s in sc), 10, 3)
    assert raw is None
    assert rep["invalid_points"] == 1 and rep["first_invalid"] == 7 and rep["failure_mask"] in (1, 2)
    # The owner stays healthy after a content rejection.
    assert backend.fixed_base_g2(sc[:2]) == [o.g2_mul(s) for s in sc[:2]]
    assert backend.g2_test_corrupt_next(4096) == -1


def test_g2_encoding_and_curve_corpus(backend):
    cases = encoding_cases()
    wire = b"".join(w for w, _ in cases)
    exp = [m for _, m in cases]
    for alg in (1, 2, 3):
        masks, rep = backend.inspect_g2_validation(wire, alg)
        assert masks == exp, alg
        ref_masks, ref_rep = o.g2_masks(wire, alg)
        assert masks == ref_masks
        assert rep == ref_rep


def test_g2_subgroup_predicates_agree_on_coset_corpus(backend):
    rng = ChaCha20Stream(4)
    sc = rng.uniform_fr(24)
    wire, exp = corpus(sc)
    for alg in (1, 2, 3):
        print(f"case: g2 subgroup alg={alg}", flush=True)
        masks, rep = backend.inspect_g2_validation(wire, alg)
        assert masks == exp, alg
        assert rep["invalid_points"] == sum(1 for m in exp if m)
        assert rep["first_invalid"] == 1 and rep["failure_mask"] == 4
    # Oracle predicates on a subset (each is a full Fq2 scalar chain).
    ref, _ = o.g2_masks(wire[:6 * 128], 3)
    assert ref == exp[:6]


def test_g2_full_config_corpus_4096(backend):
    # 4096 seeded points, half pushed into a non-subgroup coset (config 3).
    rng = ChaCha20Stream(4096)
    sc = rng.uniform_fr(4096)
    pts = backend.fixed_base_g2(sc)        # device-validated subgroup points
    T = twist_cofactor_point()
    wire = b"".join(o.g2_to_bytes(o.add(p, T, o.FQ2) if i % 2 else p) for i, p in enumerate(pts))
    masks, rep = backend.inspect_g2_validation(wire, 3)
    assert masks == [4 if i % 2 else 0 for i in range(4096)]
    assert rep["invalid_points"] == 2048 and rep["first_invalid"] == 1
    m2, rep2 = backend.inspect_g2_validation_batched(wire + wire[:128 * 3], 3)
    assert rep2["checked_points"] == 4099 and rep2["invalid_points"] == 2049
    stats = backend.g2_validation_stats()
    assert stats["checked_points"] >= 4096


def test_g2_config_resources_and_psi_probe(backend):
    with pytest.raises(BackendError):
        backend.set_g2_kernel_config(threads=96)
    backend.set_g2_kernel_config(threads=64)
    res = backend.g2_kernel_resources(3)
    assert res["block_threads"] == 64 and res["registers"] > 0
    backend.set_g2_kernel_config(threads=128)
    # psi on homogeneous inputs: psi(P) then psi^2, psi^3.
    g = o.G2_GEN
    hom = b"".join(v.to_bytes(32, "little") for v in (g[0][0], g[0][1], g[1][0], g[1][1], 1, 0))
    out = backend.g2_psi_probe(hom, 1)
    p = g
    for j in range(3):
        p = o.psi(p)
        rec = out[192 * j:192 * j + 192]
        X = (int.from_bytes(rec[0:32], "little"), int.from_bytes(rec[32:64], "little"))
        Y = (int.from_bytes(rec[64:96], "little"), int.from_bytes(rec[96:128], "little"))
        Z = (int.from_bytes(rec[128:160], "little"), int.from_bytes(rec[160:192], "little"))
        zi = o.f2_inv(Z)
        assert (o.f2_mul(X, zi), o.f2_mul(Y, zi)) == p
    with pytest.raises(BackendError):
        backend.g2_psi_probe(bytes(191) + b"\x01", 1)


@pytest.mark.parametrize("schedule,layout,fused", [(s, l, f) for s in (0, 1) for l in (0, 1) for f in (0, 1)])
def test_g2_kernel_config_variants(backend, schedule, layout, fused):
    """Every kernel configuration (the reference's schedule / layout / fused
    knobs, icicle_g2_validation.cu:117-142) gives the same masks and reports:
    encoding and curve corpus, subgroup / coset corpus, and a rejected
    checked fixed-base chunk; the repack is counted when layout is 1."""
    threads = (64, 128, 256)[(schedule + 2 * layout + fused) % 3]
    print(f"case: g2 config threads={threads} schedule={schedule} layout={layout} fused={fused}", flush=True)
    backend.set_g2_kernel_config(threads=threads, schedule=schedule, layout=layout, fused=fused)
    try:
        cases = encoding_cases()
        wire = b"".join(w for w, _ in cases)
        before = backend.g2_experiment_stats()["logical_repack_bytes"]
        masks, rep = backend.inspect_g2_validation(wire, 3)
        assert masks == [m for _, m in cases]
        ref_masks, ref_rep = o.g2_masks(wire, 3)
        assert masks == ref_masks and rep == ref_rep
        after = backend.g2_experiment_stats()["logical_repack_bytes"]
        assert after - before == (len(cases) * 256 if layout else 0)
        sc = ChaCha20Stream(40 + schedule).uniform_fr(12)
        wire, exp = corpus(sc)
        for alg in (1, 2, 3):
            masks, rep = backend.inspect_g2_validation(wire, alg)
            assert masks == exp, alg
        assert backend.g2_test_corrupt_next(3) == 0
        raw, rep = backend.fixed_base_g2_checked_encoded(b"".join(o.fr_to_bytes(s) for s in sc), 12, 3)
        assert raw is None and rep["invalid_points"] == 1 and rep["first_invalid"] == 3
        raw, rep = backend.fixed_base_g2_checked_encoded(b"".join(o.fr_to_bytes(s) for s in sc[:5]), 5, 3)
        assert rep["invalid_points"] == 0
        assert [raw[128 * i:128 * i + 128] for i in range(5)] == [o.g2_to_bytes(o.g2_mul(s)) for s in sc[:5]]
    finally:
        backend.set_g2_kernel_config(threads=128)
