"""Pins the CPU restatement (the oracle) against the reference's own known
answers and against what the mathematics fixes.  CPU only."""
import numpy as np
import pytest

import helpers
from helpers import DOC_LAPTOP, blocking, oracle_dot, oracle_mtm, oracle_mtv, oracle_vtm, rand100


# ----------------------------------------------------------- blocking ----
@pytest.mark.parametrize("elem", [4, 8])
def test_blocking_doc_tables(elem):
    """The report's tuning tables for its laptop, float and double
    (tests/golden/blocking_doc_laptop.json cites the doc lines)."""
    want = helpers.golden("blocking_doc_laptop")["tables"][str(elem)]
    got = blocking(DOC_LAPTOP, elem, 0).as_dict()
    assert {k: got[k] for k in want} == want


def test_ct_sqrt_quirk():
    """include/utils.hpp:185-201 moves the lower bound to mid+1, so ct_sqrt(72)
    lands above sqrt(72); MR = ceil(9.15625/8)*8 = 16 depends on it."""
    o = helpers.oracle()
    assert o.oracle_ct_sqrt(72.0) == 9.15625
    assert o.oracle_ct_sqrt(36.0) == 6.375


def _getconf(name: str) -> int:
    """One cache parameter from getconf(1) (0 when the system does not know it)."""
    import subprocess
    out = subprocess.run(["getconf", name], capture_output=True, text=True)
    try:
        return int(out.stdout.strip())
    except ValueError:
        return 0


def _int8(v: int) -> int:
    """The reference stores associativity as int8_t (include/cache_manager.hpp:30-31)."""
    return ((v + 128) % 256) - 128


def test_cache_probe_matches_this_host():
    """oracle_cache_probe against the host's cache parameters as getconf(1)
    reports them, with the reference's Linux slot assignment
    (include/cache_manager.hpp:107-150): sizes and line sizes of L1d/L2/L3 in
    their own slots, the L1 associativity in slot 0, the L3 associativity in
    the L2 slot (:133-135, overwriting the L2 value when L3's is known), and
    slot 2 left at the default associativity 1.  Unknown (<= 0) values keep
    the defaults: size and line 0, associativity 1."""
    import ctypes
    raw = {lv: (_getconf(f"LEVEL{lv}_CACHE_SIZE" if lv > 1 else "LEVEL1_DCACHE_SIZE"),
                _getconf(f"LEVEL{lv}_CACHE_ASSOC" if lv > 1 else "LEVEL1_DCACHE_ASSOC"),
                _getconf(f"LEVEL{lv}_CACHE_LINESIZE" if lv > 1 else "LEVEL1_DCACHE_LINESIZE"))
           for lv in (1, 2, 3)}
    g = helpers.CacheGeom()
    helpers.oracle().oracle_cache_probe(ctypes.byref(g))
    got = [(g.l[i].size, g.l[i].assoc, g.l[i].line) for i in range(3)]
    pos = lambda v, d: v if v > 0 else d
    want_assoc = [_int8(pos(raw[1][1], 1)),
                  _int8(raw[3][1]) if raw[3][1] > 0 else _int8(pos(raw[2][1], 1)),
                  1]
    want = [(pos(raw[lv][0], 0), want_assoc[lv - 1], pos(raw[lv][2], 0)) for lv in (1, 2, 3)]
    assert got == want, (got, raw)
    # On a host whose L2 and L3 associativities differ, the slot bug is
    # visible: the L2 slot holds L3's value, not L2's.
    if raw[2][1] > 0 and raw[3][1] > 0 and raw[2][1] != raw[3][1]:
        assert got[1][1] == raw[3][1] != raw[2][1]
    # The blocking the oracle uses on this host is the one its probe gives.
    if all(v[0] > 0 and v[2] > 0 for v in raw.values()):
        assert helpers.host_blocking().as_dict() == blocking(got, 4, 0).as_dict()


def test_blocking_linux_l3_assoc_slot():
    """include/cache_manager.hpp:133-135 writes the L3 associativity into the L2
    slot; with this container's geometry that gives MC = 464 and NC = 23520."""
    levels = ((49152, 12, 64), (2097152, 15, 64), (110100480, 1, 64))  # as probed
    b = blocking(levels, 4, 0).as_dict()
    assert (b["kc"], b["mc"], b["nc"]) == (512, 464, 23520)
    assert (b["mtv_kb"], b["mtv_mb"]) == (704, 688)


# ---------------------------------------------------------------- pack ----
@pytest.mark.parametrize("which", ["pack", "pack_trans"])
def test_pack_known_answer(which):
    """test/test.pack.cpp:8-44 (pack) and :46-83 (transposed pack): row panels
    of MR = 2 over a 3x3 column-major tensor of 1..9."""
    g = helpers.golden("pack_3x3")
    o = helpers.oracle()
    fn = o.oracle_pack_f32 if which == "pack" else o.oracle_pack_trans_f32
    rows, cols = g["extents"]
    inp = np.array(g["input"], dtype=np.float32)
    out = np.array(g["out_init"], dtype=np.float32)
    w = (helpers.ctypes.c_int64 * 2)(*g["strides"])
    step = g["strides"][0] if which == "pack" else g["strides"][1]
    for i in range(0, rows, g["mr"]):
        ib = min(g["mr"], rows - i)
        fn(out[i * cols:].ctypes.data, ib, inp[i * step:].ctypes.data, w, ib, cols)
    for i_in, i_out in g[which]["pairs"]:
        assert inp[i_in] == out[i_out]
    assert out.tolist() == g[which]["out"]


# ----------------------------------------------------------------- mtm ----
LAYOUTS = ["FFF", "FFL", "FLF", "LFF", "FLL", "LFL", "LLF", "LLL"]  # (C, A, B)


def _order(ch):
    return "F" if ch == "F" else "C"


@pytest.mark.parametrize("combo", LAYOUTS)
def test_mtm_integer_exact_reference_cases(combo):
    """test/test.mtm.cpp:30-460: square sizes 2..31, rand()%100 inputs, C = 0.
    Every product and partial sum is an integer < 2^24, so the result is exact."""
    oc, oa, ob = (_order(ch) for ch in combo)
    for sz in range(2, 32):
        a = rand100((sz, sz), oa)
        b = rand100((sz, sz), ob)
        c = np.zeros((sz, sz), np.float32, order=oc)
        oracle_mtm(c, a, b)
        want = a.astype(np.int64) @ b.astype(np.int64)
        assert np.array_equal(c.astype(np.int64), want), (combo, sz)


@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 3, 5), (33, 17, 65), (100, 1, 3),
                                   (3, 200, 1), (64, 96, 1000)])
def test_mtm_integer_exact_ragged(shape):
    M, N, K = shape
    a = rand100((M, K), "F")
    b = rand100((K, N), "C")
    c = rand100((M, N), "C")
    c0 = c.astype(np.int64)
    oracle_mtm(c, a, b)
    assert np.array_equal(c.astype(np.int64), c0 + a.astype(np.int64) @ b.astype(np.int64))


def test_mtm_leading_dimension_views():
    """Operands as views into larger buffers (ld > extent)."""
    big_a = rand100((50, 70), "C")
    big_b = rand100((60, 40), "F")
    a = big_a[3:33, 5:25]      # 30 x 20, row stride 70
    b = big_b[2:22, 1:31]      # 20 x 30, col stride 60
    c = np.zeros((30, 30), np.float32)
    oracle_mtm(c, a, b)
    assert np.array_equal(c.astype(np.int64), a.astype(np.int64) @ b.astype(np.int64))


def test_mtm_row_slab_is_bitwise_identical():
    """Each C element depends only on its row of A, column of B and KC, so a
    slab of rows computed alone equals the same rows of the full product --
    the property the large-size error measurement relies on."""
    M, N, K = 300, 70, 1500
    a = helpers.rng_uniform(M * K, 5, 0).reshape(M, K)
    b = helpers.rng_uniform(K * N, 5, 1).reshape(K, N)
    full = np.zeros((M, N), np.float32)
    oracle_mtm(full, a, b)
    slab = np.zeros((40, N), np.float32)
    oracle_mtm(slab, np.ascontiguousarray(a[100:140]), b)
    assert np.array_equal(slab, full[100:140])


def test_mtm_error_within_summation_bound():
    """|C - C64| <= gamma_K |A||B| componentwise (Higham, Thm 3.5); a wrong
    index or dropped term breaks it, a valid summation order cannot."""
    M, N, K = 64, 48, 3000
    a = helpers.rng_uniform(M * K, 9, 0).reshape(M, K)
    b = helpers.rng_uniform(K * N, 9, 1).reshape(K, N)
    c = np.zeros((M, N), np.float32)
    oracle_mtm(c, a, b)
    truth = a.astype(np.float64) @ b.astype(np.float64)
    u = 2.0 ** -24
    gamma = K * u / (1 - K * u)
    bound = gamma * (np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64))
    assert np.all(np.abs(c - truth) <= bound)


def test_mtm_kc_blocking_changes_rounding_only():
    """Two KC values give results that differ (the accumulation order is the
    blocking's) yet both stay within the summation bound."""
    M, N, K = 32, 32, 2048
    a = helpers.rng_uniform(M * K, 11, 0).reshape(M, K)
    b = helpers.rng_uniform(K * N, 11, 1).reshape(K, N)
    c1 = np.zeros((M, N), np.float32)
    c2 = np.zeros((M, N), np.float32)
    oracle_mtm(c1, a, b, blocking(DOC_LAPTOP))                       # KC = 384
    oracle_mtm(c2, a, b, blocking(((49152, 12, 64), (2097152, 15, 64),
                                   (110100480, 1, 64))))             # KC = 512
    assert not np.array_equal(c1, c2)
    truth = a.astype(np.float64) @ b.astype(np.float64)
    for c in (c1, c2):
        assert helpers.rel_errors(c, truth)[1] < 1e-5


# ----------------------------------------------------------- mtv / vtm ----
MTV_SIZES = list(range(2, 40)) + list(range(40, 512, 23)) + [511]


@pytest.mark.parametrize("order", ["F", "C"])
def test_mtv_integer_exact(order):
    """test/test.mtv.cpp:29-126: first order accumulates (y += Ax, y = 0 here),
    last order overwrites."""
    for sz in MTV_SIZES:
        a = rand100((sz, sz), order)
        x = rand100((sz,))
        y = np.zeros(sz, np.float32) if order == "F" else np.full(sz, 12345.0, np.float32)
        oracle_mtv(y, a, x)
        assert np.array_equal(y.astype(np.int64), a.astype(np.int64) @ x.astype(np.int64)), sz


def test_mtv_first_order_accumulates():
    a = rand100((37, 29), "F")
    x = rand100((29,))
    y = rand100((37,))
    y0 = y.astype(np.int64)
    oracle_mtv(y, a, x)
    assert np.array_equal(y.astype(np.int64), y0 + a.astype(np.int64) @ x.astype(np.int64))


@pytest.mark.parametrize("order", ["F", "C"])
def test_vtm_integer_exact_true_transpose(order):
    """test/test.vtm.cpp:29-125 checks y = A^T x; rectangular shapes too (the
    reference's vtm is only right for square A, which its tests use)."""
    for (m, n) in [(sz, sz) for sz in MTV_SIZES[::3]] + [(5, 9), (40, 3), (1, 17), (64, 200)]:
        a = rand100((m, n), order)
        x = rand100((m,))
        y = np.zeros(n, np.float32)
        oracle_vtm(y, a, x)
        assert np.array_equal(y.astype(np.int64), a.astype(np.int64).T @ x.astype(np.int64))


def test_mtv_blocked_paths_large():
    """Row counts past one MB block and columns past one KB block (laptop
    blocking: KB 448, MB 96) against FP64."""
    blk = blocking(DOC_LAPTOP)
    M, N = 1000, 3001
    a = np.asfortranarray(helpers.rng_uniform(M * N, 3, 0).reshape(M, N))
    x = helpers.rng_uniform(N, 3, 1)
    y = np.zeros(M, np.float32)
    oracle_mtv(y, a, x, blk)
    truth = a.astype(np.float64) @ x.astype(np.float64)
    u = 2.0 ** -24
    bound = N * u / (1 - N * u) * (np.abs(a).astype(np.float64) @ np.abs(x).astype(np.float64))
    assert np.all(np.abs(y - truth) <= bound)


# ----------------------------------------------------------------- dot ----
DOT_SIZES = list(range(2, 300)) + list(range(300, (1 << 17) + 2, 997)) + [(1 << 17) + 1]


def test_dot_closed_form():
    """test/test.dot_prod.cpp:10-28: a = 1..s, a.a = s(s+1)(2s+1)/6.  The FP32
    result must lie within the summation bound gamma_s * sum a_i^2 of the
    exact value (exact below 2^24)."""
    blk = blocking(DOC_LAPTOP)  # sections 1/2/3 at 16K / 4M elements
    u = 2.0 ** -24
    for s in DOT_SIZES:
        a = np.arange(1, s + 1, dtype=np.float32)
        got = oracle_dot(a, a, blk)
        exact = s * (s + 1) * (2 * s + 1) // 6
        if exact < (1 << 24):
            assert got == exact, s
        else:
            assert abs(got - exact) <= (s * u / (1 - s * u)) * exact, s


def test_dot_sections_agree_with_fp64():
    """All three regimes of include/dot.hpp:143-174 (serial, parallel-for,
    blocked) on U[-1,1) data against FP64, within the summation bound."""
    blk = blocking(DOC_LAPTOP)
    for n in (1000, 20000, 5_000_000):
        a = helpers.rng_uniform(n, 21, 0)
        b = helpers.rng_uniform(n, 21, 1)
        got = oracle_dot(a, b, blk)
        t = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
        u = 2.0 ** -24
        bound = n * u / (1 - n * u) * float(np.dot(np.abs(a).astype(np.float64), np.abs(b)))
        assert abs(got - t) <= bound, n


# ------------------------------------------------------------ generator ----
def test_splitmix64_known_answers():
    """splitmix64 with state 0 yields 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4,
    0x06c45d188009454f (Steele/Lea/Flood reference implementation)."""
    o = helpers.oracle()
    assert [o.omb_splitmix64_at(0, i) for i in range(3)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_generator_host_twins_agree():
    o = helpers.oracle()
    n = 100_003
    c = np.empty(n, np.float32)
    o.omb_rng_fill_uniform(c.ctypes.data, n, 7, 3, 11)
    py = helpers.rng_uniform(n, 7, 3, offset=11)
    assert np.array_equal(c, py)
    assert c.min() >= -1.0 and c.max() < 1.0
    # exact 24-bit lattice: (x + 1) * 2^23 is an integer
    assert np.all(np.mod((c.astype(np.float64) + 1.0) * 2 ** 23, 1.0) == 0.0)


# ----------------------------------------------------- outer / transpose --
@pytest.mark.parametrize("order", ["F", "C"])
def test_outer_integer_exact(order):
    """test/test.outer_prod.cpp:29-182: C += a b^T, sizes 2..31 (square) plus
    rectangular shapes, both storage orders; brute force in int64."""
    shapes = [(s, s) for s in range(2, 32)] + [(3, 17), (40, 5), (1, 9), (300, 257)]
    for m, n in shapes:
        a, b = rand100((m,)), rand100((n,))
        c = rand100((m, n), order)
        want = c.astype(np.int64) + a.astype(np.int64)[:, None] * b.astype(np.int64)[None, :]
        helpers.oracle_outer(c, a, b)
        assert np.array_equal(c.astype(np.int64), want), (order, m, n)


def test_outer_float_is_one_rounding_per_element():
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, 300).astype(np.float32)
    b = rng.uniform(-1, 1, 200).astype(np.float32)
    c0 = rng.uniform(-1, 1, (300, 200)).astype(np.float32)
    c = c0.copy()
    helpers.oracle_outer(c, a, b)
    exact = c0.astype(np.float64) + a.astype(np.float64)[:, None] * b.astype(np.float64)[None, :]
    # one fused multiply-add, or a product and a sum: within 1.5 ulp of the exact value
    assert np.all(np.abs(c - exact) <= 1.5 * np.spacing(np.abs(exact).astype(np.float32)))


def test_transpose_block_edge_formula():
    """transpose_partition::block() = sqrt_pow_of_two(L1 / (2 sizeof(float)))
    (include/trans.hpp:11-21, include/utils.hpp:49-58), worked by hand."""
    o = helpers.oracle()
    assert o.oracle_transpose_block(32768) == 64     # 4096 -> 2^(12 >> 1)
    assert o.oracle_transpose_block(16384) == 32     # 2048 -> 2^(11 >> 1)
    assert o.oracle_transpose_block(131072) == 128   # 16384 -> 2^(14 >> 1)
    assert o.oracle_transpose_block(49152) == 64     # 6144 -> 2^(12 >> 1)


@pytest.mark.parametrize("ao,co", [("C", "C"), ("C", "F"), ("F", "C"), ("F", "F")])
def test_transpose_out_of_place_exact(ao, co):
    """test/test.trans.cpp:11-42: C = A^T, sizes 2..31 plus rectangular and
    multi-block shapes, every storage-order pair; small blocks exercise the
    edges."""
    shapes = [(s, s) for s in range(2, 32)] + [(3, 70), (70, 3), (129, 65)]
    for m, n in shapes:
        a = rand100((m, n), ao)
        c = np.zeros((n, m), np.float32, order=co)
        helpers.oracle_transpose(c, a, block=8)
        assert np.array_equal(c, a.T), (ao, co, m, n)


def test_transpose_in_place_exact():
    """test/test.trans.cpp:44-74: square in place, sizes 2..31 and multi-block."""
    for n in list(range(2, 32)) + [65, 100]:
        a = rand100((n, n), "F")
        want = a.T.copy()
        helpers.oracle_transpose_inplace(a, block=8)
        assert np.array_equal(a, want), n
