"""The CPU oracle of the row-pair join (orc_hash_join) pinned without the
reference: numpy brute force over every (probe, build) pair for all four
join types, and the reference's own known answers for it (bench_calibrate.c
cal_measure_join: build keys row % n, probe keys row % n, every probe row
matches exactly one build row; bench_optimizer.c opt_run_key_join's
order-free pair checksum)."""
import numpy as np
import pytest

BUILD_NULL = (1 << 64) - 1


def brute(pk, bk, jt, prow=None, brow=None):
    """Every pair by f64 comparison, in probe then build position order."""
    pv = np.asarray(pk, dtype=np.float64)[prow if prow is not None else slice(None)]
    bv = np.asarray(bk, dtype=np.float64)[brow if brow is not None else slice(None)]
    eq = pv[:, None] == bv[None, :]
    bmap = np.asarray(brow, dtype=np.uint64) if brow is not None else np.arange(len(bv), dtype=np.uint64)
    if jt in ("semi", "anti"):
        hit = eq.any(axis=1)
        return np.nonzero(hit if jt == "semi" else ~hit)[0].astype(np.uint64)
    out = []
    for p in range(len(pv)):
        bs = np.nonzero(eq[p])[0]
        if len(bs):
            out += [(p, int(bmap[b])) for b in bs]
        elif jt == "left":
            out.append((p, BUILD_NULL))
    return np.array(out, dtype=np.uint64).reshape(-1, 2)


def pair_checksum(pairs):
    """bench_optimizer.c:1882-1886: sum of mixed ^ (mixed >> 29) over the
    pairs, mixed = probe * 0x9E3779B97F4A7C15 ^ build * 0xC2B2AE3D27D4EB4F
    (u64 arithmetic), independent of pair order."""
    p = pairs[:, 0].astype(np.uint64)
    b = pairs[:, 1].astype(np.uint64)
    with np.errstate(over="ignore"):
        mixed = (p * np.uint64(0x9E3779B97F4A7C15)) ^ (b * np.uint64(0xC2B2AE3D27D4EB4F))
        return int(np.sum(mixed ^ (mixed >> np.uint64(29)), dtype=np.uint64))


@pytest.mark.parametrize("jt", ["inner", "left", "semi", "anti"])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_against_brute_force(orc, jt, seed):
    r = np.random.default_rng(seed)
    pk = r.integers(0, 300, 1500).astype(np.uint32)
    bk = r.integers(0, 400, 700).astype(np.uint32)   # duplicates fan out, some probe keys absent
    got = orc.hash_join(pk, bk, jt)
    np.testing.assert_array_equal(got, brute(pk, bk, jt))


@pytest.mark.parametrize("jt", ["inner", "left", "semi", "anti"])
def test_row_lists_and_mixed_types(orc, jt):
    r = np.random.default_rng(7)
    pk = r.integers(-50, 50, 900).astype(np.int32)
    bk = r.integers(-60, 60, 500).astype(np.float64)
    prow = r.permutation(900)[:600].astype(np.uint64)
    brow = np.sort(r.permutation(500)[:300]).astype(np.uint64)[::-1].copy()   # not ascending
    got = orc.hash_join(pk, bk, jt, probe_rows=prow, build_rows=brow)
    np.testing.assert_array_equal(got, brute(pk, bk, jt, prow, brow))


def test_f64_key_semantics(orc):
    """-0.0 meets 0.0, NaN meets nothing, 2^53 + 1 as u64 meets 2^53 (both
    convert to the same f64, as in the reference's conversion)."""
    pk = np.array([0.0, -0.0, np.nan, 1.5, 2.0 ** 53], dtype=np.float64)
    bk = np.array([-0.0, np.nan, 1.5, 0.0], dtype=np.float64)
    got = orc.hash_join(pk, bk, "left")
    np.testing.assert_array_equal(got, brute(pk, bk, "left"))
    assert got.tolist() == [[0, 0], [0, 3], [1, 0], [1, 3], [2, BUILD_NULL], [3, 2], [4, BUILD_NULL]]
    big = np.array([2 ** 53 + 1], dtype=np.uint64)
    np.testing.assert_array_equal(orc.hash_join(big, np.array([2.0 ** 53]), "inner"), [[0, 0]])


@pytest.mark.parametrize("nb,np_", [(1, 10), (1000, 5000), (65537, 200000)])
def test_calibrate_known_answer(orc, nb, np_):
    """cal_build_join_table: key = row % modulus; build modulus = build rows,
    probe modulus = build rows, so probe row p pairs with build row p % nb
    and the join returns exactly np rows (bench_calibrate.c:296-360)."""
    bk = (np.arange(nb) % nb).astype(np.uint32)
    pk = (np.arange(np_) % nb).astype(np.uint32)
    got = orc.hash_join(pk, bk, "inner")
    assert len(got) == np_
    want = np.stack([np.arange(np_), np.arange(np_) % nb], axis=1).astype(np.uint64)
    np.testing.assert_array_equal(got, want)
    assert pair_checksum(got) == pair_checksum(want[::-1].copy())


def test_empty_sides(orc):
    e = np.array([], dtype=np.uint32)
    k = np.array([1, 2], dtype=np.uint32)
    assert orc.hash_join(e, k, "inner").shape == (0, 2)
    assert orc.hash_join(k, e, "inner").shape == (0, 2)
    assert orc.hash_join(k, e, "left").tolist() == [[0, BUILD_NULL], [1, BUILD_NULL]]
    assert orc.hash_join(k, e, "anti").tolist() == [0, 1]
    assert orc.hash_join(k, e, "semi").tolist() == []
