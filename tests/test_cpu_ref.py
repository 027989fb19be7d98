"""The C++ / OpenMP CPU restatement (oracle/cpu_ref.cpp), the timed CPU
baseline of bench.py, checked against the reference's golden L4 files (the
reference's gate, h5diff -d 1e-12), against the numpy oracle on the pipi
workload shape (C2c / C4cD / C4cB / C4cC / C6cD, D = 16) with native and Kahan
sums and with several threads, and its VdaggerV loop nest against numpy."""
import os

import numpy as np
import pytest

from oracle import cpu_ref
from oracle import slaph_oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")

PIPI = """
Lt = 8
Lx = 4
Ly = 4
Lz = 4
number_of_eigen_vec = 8
[quarks]
quark = u:4:TB:2:EI:4:DF:4:synthetic
[operator_lists]
operator_list = g5.d0.p0,1
[correlator_lists]
correlator_list = C2c:Q0:Op0:Q0:Op0:P0,1
correlator_list = C4cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1
correlator_list = C4cB:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0
correlator_list = C4cC:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0
correlator_list = C6cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0
"""


@pytest.fixture(scope="module")
def l4():
    cfg = o.parse_infile(open(os.path.join(GOLD, "l4", "test_all.ini")).read())
    lut = o.init_lookup_tables(cfg)
    perams, rnds, ev = o.read_inputs(cfg, os.path.join(GOLD, "l4"), 1000)
    vdv = o.build_vdaggerv(cfg, lut, ev)
    return cfg, lut, vdv, perams, rnds


@pytest.mark.parametrize("threads", [1, 3])
def test_l4_golden_all_datasets(l4, threads):
    cfg, lut, vdv, perams, rnds = l4
    res, secs = cpu_ref.CpuRef(cfg, lut, vdv, perams, rnds).contract(threads=threads)
    gold = np.load(os.path.join(GOLD, "l4_correlators.npz"))
    mine = {d + "/" + n: v for d, lst in res.items() for n, v in lst}
    assert set(mine) == set(gold.files)
    for k in gold.files:
        assert np.max(np.abs(mine[k] - gold[k])) <= 1e-12, k
    assert secs > 0


def test_l4_planted_dagger_error_is_caught(l4):
    cfg, lut, vdv, perams, rnds = l4
    bad = dict(lut)
    bad["ql_lookup"] = dict(lut["ql_lookup"])
    bad["ql_lookup"]["Q2"] = [(e[0], not e[1], e[2], e[3]) for e in lut["ql_lookup"]["Q2"]]
    res, _ = cpu_ref.CpuRef(cfg, bad, vdv, perams, rnds).contract()
    gold = np.load(os.path.join(GOLD, "l4_correlators.npz"))
    assert max(np.max(np.abs(v - gold["C2c/" + n])) for n, v in res["C2c"]) > 1e-6


@pytest.fixture(scope="module")
def pipi():
    cfg = o.parse_infile(PIPI)
    lut = o.init_lookup_tables(cfg)
    rng = np.random.default_rng(3)
    Lt, N, K = 8, 8, 192
    ev = [np.linalg.qr(rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N)))[0] for _ in range(Lt)]
    perams = [rng.standard_normal((4 * N * Lt, 4 * 4 * 4)) + 1j * rng.standard_normal((4 * N * Lt, 4 * 4 * 4))
              for _ in range(4)]
    rnds = [((rng.integers(0, 2, Lt * N * 4) * 2 - 1) + 1j * (rng.integers(0, 2, Lt * N * 4) * 2 - 1)) / np.sqrt(2)
            for _ in range(4)]
    vdv = o.build_vdaggerv(cfg, lut, ev)
    return cfg, lut, vdv, perams, rnds


_WANT = {}  # oracle correlators per Kahan setting


@pytest.mark.parametrize("kahan,threads", [(False, 1), (False, 4), (True, 2)])
def test_pipi_shape_matches_oracle(pipi, kahan, threads):
    cfg, lut, vdv, perams, rnds = pipi
    if kahan not in _WANT:
        _WANT[kahan] = o.contract(cfg, lut, vdv, perams, rnds, kahan=kahan)
    want = _WANT[kahan]
    got, _ = cpu_ref.CpuRef(cfg, lut, vdv, perams, rnds).contract(threads=threads, kahan=kahan)
    assert set(got) == set(want)
    for d, lst in want.items():
        g = dict(got[d])
        for name, vals in lst:
            scale = max(np.max(np.abs(vals)), 1e-300)
            assert np.max(np.abs(g[name] - vals)) / scale <= 1e-12, (d, name)


def test_single_block_pair_and_sample_limits(pipi):
    cfg, lut, vdv, perams, rnds = pipi
    want = o.contract(cfg, lut, vdv, perams, rnds, single_block_pair=2)
    got, _ = cpu_ref.CpuRef(cfg, lut, vdv, perams, rnds).contract(first_block_pair=2, block_pairs=1)
    for d, lst in want.items():
        g = dict(got[d])
        for name, vals in lst:
            assert np.max(np.abs(g[name] - vals)) <= 1e-12 * max(np.max(np.abs(vals)), 1e-300)
    part, _ = cpu_ref.CpuRef(cfg, lut, vdv, perams, rnds).contract(first_block_pair=2, block_pairs=1, slice_pairs=3)
    assert any(np.max(np.abs(dict(part[d])[n] - v)) > 0 for d, lst in want.items() for n, v in lst)


def test_vdaggerv_loop_nest_matches_numpy():
    rng = np.random.default_rng(4)
    K, N = 3 * 125, 17
    V = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    ph = o.create_momenta(5, 5, 5, (1, -1, 2))
    got, secs = cpu_ref.vdaggerv(V, ph)
    want = o.vdaggerv(V, ph)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    assert secs >= 0


def test_factor_key_count_matches_the_keys_built(pipi):
    """bench.py's extrapolation counts the factor keys of a block pair with
    cpu_ref.factor_keys; a full block pair builds exactly those."""
    cfg, lut, vdv, perams, rnds = pipi
    ref = cpu_ref.CpuRef(cfg, lut, vdv, perams, rnds)
    for b, (a, c) in enumerate(o.block_pairs(8, 2)):
        ref.contract(first_block_pair=b, block_pairs=1)
        st = ref.last_stats()
        keys = cpu_ref.factor_keys(lut, 2, o.slice_pairs(a, c, 8, 4))
        for q in ("Q0", "Q1", "Q2"):
            assert st[q][1] == sum(1 for k in keys if k[0] == q), (b, q)
