"""Pins of the CPU oracle (test infrastructure) against the reference's own
fixtures and known answers, and against closed forms / brute force."""
import os

import numpy as np
import pytest

from oracle import slaph_oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


# --- gamma: tests/gtest/gamma.cpp:18-104 -----------------------------------
def _dense(sp):
    rows, vals = sp
    m = np.zeros((4, 4), complex)
    for k in range(4):
        m[rows[k], k] = vals[k]
    return m


def test_gamma_basis_and_products():
    I = 1j
    g = [np.array(x, complex) for x in (
        [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]],
        [[0, 0, 0, I], [0, 0, I, 0], [0, -I, 0, 0], [-I, 0, 0, 0]],
        [[0, 0, 0, 1], [0, 0, -1, 0], [0, -1, 0, 0], [1, 0, 0, 0]],
        [[0, 0, I, 0], [0, 0, 0, -I], [-I, 0, 0, 0], [0, I, 0, 0]])]
    g5 = g[0] @ g[1] @ g[2] @ g[3]
    want = [g[0], g[1], g[2], g[3], np.eye(4), g5, g[0] @ g5, g[1] @ g5, g[2] @ g5, g[3] @ g5,
            g[0] @ g[1], g[0] @ g[2], g[0] @ g[3], g[0] @ g[1] @ g5, g[0] @ g[2] @ g5, g[0] @ g[3] @ g5]
    for k in range(16):
        assert np.array_equal(_dense(o.GAMMA[k]), want[k]), k
    # gamma_5 is diag(1,1,-1,-1) in this basis (sign rule of the Q2 gamma_5 trick)
    assert np.array_equal(g5, np.diag([1, 1, -1, -1]).astype(complex))


# --- Kahan: tests/gtest/KahanAccumulatorTest.cpp:9-59 ----------------------
def test_kahan_known_answers():
    acc = o.KahanAccumulator(0.0)
    s = 0.0
    for i in range(1, 10):
        acc.add(float(i))
        s += i
    assert abs(acc.value() - s) <= 1e-20
    eps = np.finfo(float).eps
    acc = o.KahanAccumulator(0.0)
    acc.add(1.0)
    s = 1.0
    summand = eps / 100
    for _ in range(10000):
        acc.add(summand)
        s += summand
    exact = 1.0 + 10000 * summand
    assert acc.value() == exact
    assert s != exact
    acc = o.KahanAccumulator(0.0)
    summand = eps * 1e-15
    acc.add(1.0)
    acc.add(summand)
    acc.add(-1.0)
    assert acc.value() == summand


def kahan_trace_cases():
    """Trace vectors built from the gtest known answers
    (KahanAccumulatorTest.cpp:22-59): (values, masks) whose compensated and
    plain sums differ.  Returns (name, [trace vectors], kahan value, plain value)."""
    eps = float(np.finfo(float).eps)
    hard = [(1.0 + 0.5j, 0)] + [(complex(eps / 100, 0.25 * eps / 100), 0)] * 10000
    exact_hard = complex(1.0 + 10000 * (eps / 100), 0.5 + 10000 * (0.25 * eps / 100))
    tiny = eps * 1e-15
    prec = [(1.0, 1), (tiny, 2), (-1.0, 4)]
    one = [(1.0, 0)]
    return [
        # resolve_request1: mean of one trace vector
        ("mean-hard", [hard], o.cdiv(exact_hard, len(hard)), o.cdiv(1.0 + 0.5j, len(hard))),
        ("mean-precision", [prec], tiny / 3, 0.0),
        # inner_product 2-way: right_sum carries the compensation
        ("inner2-precision", [one, prec], tiny / 3, 0.0),
        # inner_product 3-way: right_sum of the third vector
        ("inner3-precision", [one, one, prec], tiny / 3, 0.0),
        # result carries the compensation: left values differ, right sums are 1
        ("inner2-result", [prec, one], tiny / 3, 0.0),
    ]


@pytest.mark.parametrize("case", kahan_trace_cases(), ids=lambda c: c[0])
def test_kahan_inner_products_known_answers(case):
    """SLAPH_KAHAN build: resolve_request1's mean and inner_product's
    right_sum / result are Kahan accumulators (Diagram.cpp:24-25,
    DilutedTrace.cpp:13-33, 51-76); the native build sums plainly."""
    name, vecs, want_kahan, want_plain = case

    def run(kahan):
        if len(vecs) == 1:
            return o.trace_mean(vecs[0], kahan)
        if len(vecs) == 2:
            return o.inner_product2(vecs[0], vecs[1], kahan)
        return o.inner_product3(vecs[0], vecs[1], vecs[2], kahan)

    assert run(True) == want_kahan
    assert run(False) == want_plain
    assert want_kahan != want_plain


def test_kahan_contract_sums_slice_pairs_plainly():
    """The reference's per-slice-pair accumulator holds one value with zero
    compensation, and merging it (KahanAccumulator.hpp:40-44) adds sum and
    compensation separately: across slice pairs the SLAPH_KAHAN build sums
    plainly (Diagram.cpp:96-114)."""
    acc = o.KahanAccumulator(0.0)
    for x in (1.0, np.finfo(float).eps * 1e-15, -1.0):
        one = o.KahanAccumulator(0.0)
        one.add(x)
        assert one.c == 0.0
        acc.merge(one)
    assert acc.value() == 0.0  # a compensated sum would keep the middle term


# --- dilution iterator: tests/dilution-iterator/control.txt ----------------
def test_dilution_iterator_control_text():
    txt = (o.dilution_iterator_text(6, 2, False) + o.dilution_iterator_text(6, 2, True)
           + o.dilution_iterator_text(48, 4, False) + o.dilution_iterator_text(48, 4, True)
           + o.dilution_iterator_text(48, 24, False))
    with open(os.path.join(GOLD, "dilution_iterator_control.txt")) as fh:
        want = fh.read()
    assert txt == want


def test_time_delta_is_source_minus_sink():
    for Lt in (4, 6, 48):
        for a in range(Lt):
            for b in range(Lt):
                assert o.time_delta(a, b, Lt) == (a - b) % Lt


# --- momenta: tests/gtest/create_momenta.cpp:5-20 + closed form -------------
def test_momenta_old_new_and_closed_form():
    Lx, Ly, Lz = 14, 26, 34
    p = (2, 3, 5)
    a = o.create_momenta(Lx, Ly, Lz, p)
    b = o.create_momenta_new(Lx, Ly, Lz, p)
    assert np.max(np.abs(a - b)) < 1e-14
    x, y, z = 3, 7, 11
    site = (x * Ly + y) * Lz + z
    want = np.exp(-2j * np.pi * (p[0] * x / Lx + p[1] * y / Ly + p[2] * z / Lz))
    assert abs(a[site] - want) < 1e-14


def test_momentum_enumeration_order():
    assert o.momentum_combinations(1) == [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 1), (0, 1, 0), (1, 0, 0)]
    assert len(o.momentum_combinations(2)) == 12
    assert len(o.momentum_combinations(3)) == 8
    assert len(o.momentum_combinations(4)) == 6


# --- VdaggerV: invariants and brute force -------------------------------------
def _orthonormal(K, N, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N)))
    return q


def test_vdaggerv_invariants():
    L = 4
    V = _orthonormal(3 * L ** 3, 8, 1)
    one = o.vdaggerv(V, o.create_momenta(L, L, L, (0, 0, 0)))
    assert np.max(np.abs(one - np.eye(8))) < 1e-13
    p = (1, 0, -1)
    a = o.vdaggerv(V, o.create_momenta(L, L, L, p))
    b = o.vdaggerv(V, o.create_momenta(L, L, L, tuple(-x for x in p)))
    assert np.max(np.abs(a - b.conj().T)) < 1e-13


def test_vdaggerv_brute_force_triple_loop():
    Lx, Ly, Lz, N = 2, 3, 2, 3
    rng = np.random.default_rng(5)
    V = rng.standard_normal((3 * Lx * Ly * Lz, N)) + 1j * rng.standard_normal((3 * Lx * Ly * Lz, N))
    p = (1, 2, 1)
    got = o.vdaggerv(V, o.create_momenta(Lx, Ly, Lz, p))
    want = np.zeros((N, N), complex)
    for x in range(Lx):
        for y in range(Ly):
            for z in range(Lz):
                ph = np.exp(-2j * np.pi * (p[0] * x / Lx + p[1] * y / Ly + p[2] * z / Lz))
                s = (x * Ly + y) * Lz + z
                for c in range(3):
                    for i in range(N):
                        for j in range(N):
                            want[i, j] += np.conj(V[3 * s + c, i]) * ph * V[3 * s + c, j]
    assert np.max(np.abs(got - want)) < 1e-13


# --- factors: a planted mistake is caught ------------------------------------
def test_perambulator_resort_matches_index_formula():
    Lt, N, dilT, dilE, dilD = 4, 3, 2, 2, 4
    nb = Lt // dilT
    ninv = nb * dilE * dilD
    raw = np.arange(4 * N * Lt * ninv).astype(complex)
    P = o.resort_perambulator(raw, Lt, N, dilT, dilE, dilD)
    for t in range(Lt):
        for v in range(N):
            for d in range(4):
                for b in range(nb):
                    for e in range(dilE):
                        for dd in range(dilD):
                            row_f = 4 * N * t + 4 * v + d
                            col_f = dilD * dilE * b + dilD * e + dd
                            assert P[4 * N * t + N * d + v, dilE * dilD * b + dilE * dd + e] == raw[row_f * ninv + col_f]


# --- end to end: the 18 golden files (tests/integration-L4) ---------------------
@pytest.fixture(scope="module")
def l4_result():
    return o.run_infile(os.path.join(GOLD, "l4", "test_all.ini"))


def test_l4_golden_all_datasets(l4_result):
    gold = np.load(os.path.join(GOLD, "l4_correlators.npz"))
    mine = {d + "/" + n: v for d, lst in l4_result.items() for n, v in lst}
    assert set(mine) == set(gold.files)
    for k in gold.files:
        # the reference's gate: h5diff -r -d 1e-12 (tests/integration-L4/compare:16)
        assert np.max(np.abs(mine[k] - gold[k])) <= 1e-12, k


def test_l4_planted_sign_error_is_caught():
    """A transposed operand (wrong dagger flag) must break C2c parity."""
    cfg = o.parse_infile(open(os.path.join(GOLD, "l4", "test_all.ini")).read())
    cfg["correlator_list"] = [c for c in cfg["correlator_list"] if c["type"] == "C2c"]
    lut = o.init_lookup_tables(cfg)
    perams, rnds, ev = o.read_inputs(cfg, os.path.join(GOLD, "l4"), 1000)
    vdv = o.build_vdaggerv(cfg, lut, ev)
    gold = np.load(os.path.join(GOLD, "l4_correlators.npz"))
    good = o.contract(cfg, lut, vdv, perams, rnds)
    err_good = max(np.max(np.abs(v - gold["C2c/" + n])) for n, v in good["C2c"])
    assert err_good < 1e-12
    # plant: a wrong dagger flag on every Q2 operator (VdV(p) instead of VdV(-p))
    bad_lut = dict(lut)
    bad_lut["ql_lookup"] = dict(lut["ql_lookup"])
    bad_lut["ql_lookup"]["Q2"] = [(e[0], not e[1], e[2], e[3]) for e in lut["ql_lookup"]["Q2"]]
    bad = o.contract(cfg, bad_lut, vdv, perams, rnds)
    err_bad = max(np.max(np.abs(v - gold["C2c/" + n])) for n, v in bad["C2c"])
    assert err_bad > 1e-6
