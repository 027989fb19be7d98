"""GPU parity of the HIP path, called through the C ABI, against the oracle
and the reference's golden L4 correlators."""
import os
import sys

import numpy as np
import pytest

from oracle import slaph_oracle as o
from test_oracle import kahan_trace_cases

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
L4 = os.path.join(GOLD, "l4")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def slaph():
    import slaph as s
    return s


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("M,N,K,phase", [(16, 16, 4, False), (220, 220, 3 * 512, True), (16, 220, 220, False),
                                         (7, 5, 13, True), (64, 129, 100, True)])
def test_ahb_kernel_matches_numpy(slaph, M, N, K, phase):
    """C = A^H diag(phase[k/3]) B on FP64 MFMA (checks the MFMA lane layout too)."""
    say("ahb", M, N, K, phase)
    rng = np.random.default_rng(M * 1000 + N + K)
    A = _rand(rng, K, M)
    Bm = _rand(rng, K, N)
    ph = np.exp(-1j * rng.uniform(0, 2 * np.pi, size=(K + 2) // 3))
    ctx = slaph.Context(0)
    C = ctx.ahb_host(A, Bm, ph if phase else None)
    want = A.conj().T @ ((np.repeat(ph, 3)[:K, None] if phase else 1.0) * Bm)
    assert np.max(np.abs(C - want)) <= 1e-12 * max(1.0, np.max(np.abs(want))) * 4


@pytest.fixture(scope="module")
def l4_ctx(slaph):
    say("l4: configure + read inputs")
    ctx = slaph.Context(0)
    ctx.configure(open(os.path.join(L4, "test_all.ini")).read())
    ctx.read_inputs(L4, 1000)
    ops = slaph.OperatorFactory(ctx)
    ops.create_operators()
    return ctx, ops


@pytest.fixture(scope="module")
def l4_oracle():
    cfg = o.parse_infile(open(os.path.join(L4, "test_all.ini")).read())
    lut = o.init_lookup_tables(cfg)
    perams, rnds, ev = o.read_inputs(cfg, L4, 1000)
    vdv = o.build_vdaggerv(cfg, lut, ev)
    return cfg, lut, perams, rnds, vdv


def test_l4_vdaggerv_matches_oracle(l4_ctx, l4_oracle):
    ctx, ops = l4_ctx
    cfg, lut, perams, rnds, vdv = l4_oracle
    say("l4: vdaggerv")
    for vid, mom, _ in lut["vdv_lookup"]:
        for t in range(cfg["Lt"]):
            got = ops.return_vdaggerv(vid, t)
            assert np.max(np.abs(got - vdv[(vid, t)])) < 1e-13, (vid, t)


def test_l4_diluted_factors_match_oracle(slaph, l4_ctx, l4_oracle):
    ctx, ops = l4_ctx
    cfg, lut, perams, rnds, vdv = l4_oracle
    say("l4: factors")
    fac = o.Factories(cfg, lut, vdv, perams, rnds, cfg["quarks"][0]["dilT"], cfg["quarks"][0]["dilE"])
    cases = [("Q0", (1,)), ("Q0", (3,)), ("Q1", (2, 0)), ("Q1", (0, 1)), ("Q2", (1, 2, 0)), ("Q2", (0, 0, 0))]
    for qt, key in cases:
        dev = slaph.DilutedFactorFactory(ctx, qt)
        want = fac.get(qt, key)
        for op in range(len(lut["ql_lookup"][qt])):
            got = dev.get(key, op)
            assert len(got) == len(want[op])
            for (gm, gric), (wm, wric, _) in zip(got, want[op]):
                assert gric == wric
                assert np.max(np.abs(gm - wm)) < 1e-12, (qt, key, op)


def test_l4_diluted_trace_factories_match_oracle(slaph, l4_ctx, l4_oracle):
    """DilutedTraceFactory::get through slc_diluted_trace for all eight trace
    types, at three slice pairs (same block, different blocks both ways):
    every trace id's vector of (value, used-index mask) against the oracle's
    trace factories, masks exactly and values at 1e-12 of the vector's scale."""
    ctx, ops = l4_ctx
    cfg, lut, perams, rnds, vdv = l4_oracle
    q = cfg["quarks"][0]
    fac = o.Factories(cfg, lut, vdv, perams, rnds, q["dilT"], q["dilE"])
    locs_of = {}
    for reqs in lut["requests"].values():
        for trs, _ in reqs:
            for name, _, locs in trs:
                locs_of.setdefault(name, locs)
    assert len(locs_of) == 8
    for name, locs in sorted(locs_of.items()):
        dev = slaph.DilutedTraceFactory(ctx, name)
        for t_so, t_si in ((0, 1), (1, 2), (3, 0)):
            say("l4 trace factory", name, locs, t_so, t_si)
            tr = o.Traces(fac, lut, q["dilT"])
            key = [t_so if l == "source" else t_si for l in locs]
            got = dev.get((t_so, t_si), locs)
            assert len(got) == len(lut["trace_indices"][name])
            for tid, vec in got.items():
                want = tr.get(name, key, tid)
                assert [m for _, m in vec] == [m for _, m in want], (name, tid)
                scale = max(abs(v) for v, _ in want)
                assert max(abs(g - w) for (g, _), (w, _) in zip(vec, want)) <= 1e-12 * scale, (name, tid)
            fac.clear()


@pytest.mark.parametrize("case", kahan_trace_cases(), ids=lambda c: c[0])
def test_device_inner_products_kahan_known_answers(slaph, case):
    """The device resolve kernel on trace vectors built from the reference's
    Kahan known answers (KahanAccumulatorTest.cpp:22-59): with kahan the
    compensated value of the SLAPH_KAHAN build, without it the plain sum,
    both bit-equal to the oracle."""
    name, vecs, want_kahan, want_plain = case
    say("inner product", name)
    ctx = slaph.Context(0)
    assert slaph.inner_product(ctx, *vecs, kahan=True) == want_kahan
    assert slaph.inner_product(ctx, *vecs, kahan=False) == want_plain


def test_device_inner_product_empty_result_raises(slaph):
    ctx = slaph.Context(0)
    with pytest.raises(slaph.SlaphError, match="has an empty result"):
        slaph.inner_product(ctx, [(1.0, 3)], [(2.0, 1)])


@pytest.mark.parametrize("kahan", [False, True])
def test_l4_golden_correlators(slaph, l4_ctx, kahan):
    """All 18 diagrams, 2,307 datasets, against the reference's golden files at
    1e-12 absolute (the reference gate, tests/integration-L4/compare:16)."""
    ctx, ops = l4_ctx
    say("l4: contract kahan=%d" % kahan)
    slaph.contract(ctx, kahan=kahan)
    gold = np.load(os.path.join(GOLD, "l4_correlators.npz"))
    seen = set()
    worst = 0.0
    for d in ctx.diagrams():
        for name, vals in ctx.correlators(d).items():
            key = d + "/" + name
            seen.add(key)
            err = np.max(np.abs(vals - gold[key]))
            worst = max(worst, err)
            assert err <= 1e-12, (key, err)
    assert seen == set(gold.files)
    say("l4 worst abs error", worst)


def test_l4_golden_with_eigenvectors_streamed_from_files(slaph):
    """Timeslice streaming with the eigenvectors read file by file inside
    create_operators (as the reference reads them, OperatorFactory.cpp:263-277):
    the same 2,307 golden datasets at 1e-12, the full V never resident, and
    get_input reading V(t) back from its file."""
    say("l4: streamed eigenvector files")
    ctx = slaph.Context(0)
    ctx.set_timeslice_streaming(True)
    ctx.configure(open(os.path.join(L4, "test_all.ini")).read())
    ctx.read_inputs(L4, 1000)
    slaph.OperatorFactory(ctx).create_operators()
    slaph.contract(ctx)
    gold = np.load(os.path.join(GOLD, "l4_correlators.npz"))
    seen = set()
    for d in ctx.diagrams():
        for name, vals in ctx.correlators(d).items():
            seen.add(d + "/" + name)
            assert np.max(np.abs(vals - gold[d + "/" + name])) <= 1e-12, (d, name)
    assert seen == set(gold.files)
    cfg = o.parse_infile(open(os.path.join(L4, "test_all.ini")).read())
    _, _, ev = o.read_inputs(cfg, L4, 1000)
    assert np.array_equal(ctx.get_input("eigenvectors", 2), ev[2])
    with pytest.raises(slaph.SlaphError, match="streaming"):
        ctx.set_eigenvectors(0, ev[0])


def test_l4_single_block_pair_and_chunking(slaph, l4_ctx, l4_oracle):
    """single_time_slice_combination and a tiny chunk budget (one slice pair
    per launch) give the oracle's numbers."""
    ctx, ops = l4_ctx
    cfg, lut, perams, rnds, vdv = l4_oracle
    say("l4: single block pair 1, chunk budget 1 byte")
    ctx.set_chunk_budget(1)
    try:
        slaph.contract(ctx, single_time_slice_combination=1)
    finally:
        ctx.set_chunk_budget(0)
    want = o.contract(cfg, lut, vdv, perams, rnds, single_block_pair=1)
    for d, lst in want.items():
        got = ctx.correlators(d)
        for name, vals in lst:
            assert np.max(np.abs(got[name] - vals)) <= 1e-12, (d, name)


SMALL = """
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


WIDE = SMALL.replace("C4cB:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0\n", "C4cB:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1\n").replace(
    "C4cC:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0\n", "C4cC:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1\n")


@pytest.mark.parametrize("infile,kernels", [(SMALL, "lstat+mfma"), (WIDE, "lstat+mfma"), (WIDE, "groups+generic"),
                                            (WIDE, "blocks+items"), (WIDE, "lstat+mfma+window1"),
                                            (WIDE, "lstat+mfma+crossed_direct"), (WIDE, "lstat+mfma+crossed_lstat"),
                                            (WIDE, "lstat+mfma+sym_split")],
                         ids=["small", "wide", "wide-alt-kernels", "wide-block-traces", "wide-one-round-windows",
                              "wide-crossed-direct", "wide-crossed-lstat", "wide-crossed-sym-split"])
def test_host_inputs_small_pipi_matches_oracle(slaph, infile, kernels, monkeypatch):
    """Config-1-like lattice (4^3 x 8, N=8, dilE 4, 4 random vectors) with
    seeded host inputs: every correlator against the oracle at 1e-12
    norm-wise.  The wide variant (C4cB / C4cC up to |P|^2 = 1) runs the
    non-separable traces and the D = 16 products through both kernels each
    (left-stationary / one wave per group; MFMA / generic product), the
    masked inner products both by Moebius inversion and by direct group loops,
    and VdaggerV both by the sitewise Fourier form and directly.  The
    block-traces variant runs the left-stationary traces in blocks of chunks
    instead of the queued form, and the slice-pair tiles item by item instead
    of grouped by right operand.  The one-round-windows variant reloads the
    queued form's task descriptors at every round of its ring (the largest
    units of the 32^3 x 64 workload fit one window, so only this variant
    runs that reload) and runs the slice-pair tiles grouped by right operand
    instead of as GEMMs.  The crossed traces (C4cC) run in the summed-product
    form (products summed over the middle random vector, the coincident
    chains removed by inclusion and exclusion) except in the crossed-direct
    variant (one product per combination) and the alternative-kernels one
    (one wave per group, which the summed form does not use); their traces
    run as symmetric products per trace-graph component, in the crossed-lstat
    variant on the left-stationary kernel, in the sym-split variant with
    every component of more than four entries cut into chunks of four (pieces
    of two chunks, rows from one and columns from the other)."""
    say("small pipi", "wide" if infile is WIDE else "", kernels)
    for var, on in (("SLAPH_TRACE_GROUPS", "groups" in kernels), ("SLAPH_PROD_GENERIC", "generic" in kernels),
                    ("SLAPH_RESOLVE_GROUPS", "groups" in kernels), ("SLAPH_VDV_DIRECT", "groups" in kernels),
                    ("SLAPH_LSTAT_BLOCKS", "blocks" in kernels), ("SLAPH_GRID_ITEMS", "items" in kernels),
                    ("SLAPH_LSTAT_WINDOW", "window1" in kernels), ("SLAPH_GRID_GROUPS", "window1" in kernels),
                    ("SLAPH_XFORM_OFF", "crossed_direct" in kernels), ("SLAPH_SYM_OFF", "crossed_lstat" in kernels),
                    ("SLAPH_SYM_ENTRIES", "sym_split" in kernels)):
        if on:
            monkeypatch.setenv(var, "4" if var == "SLAPH_SYM_ENTRIES" else "1")
        else:
            monkeypatch.delenv(var, raising=False)
    cfg = o.parse_infile(infile)
    lut = o.init_lookup_tables(cfg)
    rng = np.random.default_rng(11)
    Lt, N, K = cfg["Lt"], cfg["number_of_eigen_vec"], 3 * 64
    ev = [np.linalg.qr(_rand(rng, K, N))[0] for _ in range(Lt)]
    nrnd = 4
    nb = Lt // 2
    perams = [_rand(rng, 4 * N * Lt, nb * 4 * 4) for _ in range(nrnd)]
    rnds = [((rng.integers(0, 2, Lt * N * 4) * 2 - 1) + 1j * (rng.integers(0, 2, Lt * N * 4) * 2 - 1)) / np.sqrt(2)
            for _ in range(nrnd)]
    ctx = slaph.Context(0)
    ctx.configure(infile)
    for t in range(Lt):
        ctx.set_eigenvectors(t, ev[t])
    for r in range(nrnd):
        ctx.set_perambulator(r, perams[r])
        ctx.set_random_vector(r, rnds[r])
    slaph.OperatorFactory(ctx).create_operators()
    slaph.contract(ctx)
    if infile not in _SMALL_WANT:
        vdv = o.build_vdaggerv(cfg, lut, ev)
        _SMALL_WANT[infile] = o.contract(cfg, lut, vdv, perams, rnds)
    want = _SMALL_WANT[infile]
    for d, lst in want.items():
        got = ctx.correlators(d)
        for name, vals in lst:
            scale = max(np.max(np.abs(vals)), 1e-300)
            assert np.max(np.abs(got[name] - vals)) / scale <= 1e-12, (d, name)


_SMALL_WANT = {}  # oracle correlators per input file (same seeded inputs)


def dilution_infile(dilE, N):
    """4^3 x 8 lattice with eigenvector interlace dilE (D = 4 dilE), N a
    ragged multiple of it, 4 random vectors, C2c / C3c / C4cD / C4cB / C4cC up
    to |P|^2 = 1 (C3c reads Q1 through trQ1Q0Q2)."""
    return """
Lt = 8
Lx = 4
Ly = 4
Lz = 4
number_of_eigen_vec = {N}
[quarks]
quark = u:4:TB:2:EI:{dilE}:DF:4:synthetic
[operator_lists]
operator_list = g5.d0.p0,1
[correlator_lists]
correlator_list = C2c:Q0:Op0:Q0:Op0:P0,1
correlator_list = C3c:Q0:Op0:Q0:Op0:Q0:Op0:P0,1
correlator_list = C4cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1
correlator_list = C4cB:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1
correlator_list = C4cC:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1
""".format(N=N, dilE=dilE)


_DIL_WANT = {}  # (dilE, N) -> (oracle correlators, inputs)

KERNEL_SWITCHES = ("SLAPH_TRACE_GROUPS", "SLAPH_PROD_GENERIC", "SLAPH_RESOLVE_GROUPS", "SLAPH_VDV_DIRECT",
                   "SLAPH_LSTAT_BLOCKS", "SLAPH_GRID_ITEMS", "SLAPH_LSTAT_WINDOW", "SLAPH_GRID_GROUPS",
                   "SLAPH_XFORM_OFF", "SLAPH_SYM_OFF", "SLAPH_SYM_ENTRIES")


@pytest.mark.parametrize("dilE,N", [(6, 14), (6, 16), (8, 20)], ids=["EI6-D24", "EI6-D24-N16", "EI8-D32"])
@pytest.mark.parametrize("kernels", ["production", "blocks+items", "window1+gridgroups", "groups+generic"])
def test_dilution_d24_d32_matches_oracle(slaph, dilE, N, kernels, monkeypatch):
    """D = 24 and D = 32 (the EI6 of the reference's infiles/pipi_I2.in and the
    EI8 of the 24^3 x 48 production workload) with N not a multiple of dilE:
    Q0 / Q1 / Q2 at dilE 6 and 8 (the MFMA Q2 stage when N is a multiple of
    4, the generic one otherwise), the MFMA and the generic product kernels,
    and the left-stationary trace kernels at D = 24 / 32 (queued and block
    form, one-round descriptor windows) against the oracle at 1e-12
    norm-wise, on device-generated inputs downloaded for the oracle."""
    say("dilution D=%d N=%d" % (4 * dilE, N), kernels)
    on = {"production": (), "blocks+items": ("SLAPH_LSTAT_BLOCKS", "SLAPH_GRID_ITEMS"),
          "window1+gridgroups": ("SLAPH_LSTAT_WINDOW", "SLAPH_GRID_GROUPS"),
          "groups+generic": ("SLAPH_TRACE_GROUPS", "SLAPH_PROD_GENERIC", "SLAPH_RESOLVE_GROUPS",
                             "SLAPH_VDV_DIRECT")}[kernels]
    for var in KERNEL_SWITCHES:
        if var in on:
            monkeypatch.setenv(var, "1")
        else:
            monkeypatch.delenv(var, raising=False)
    text = dilution_infile(dilE, N)
    ctx = slaph.Context(0)
    ctx.configure(text)
    assert ctx.query("dilE") == dilE
    ctx.synthetic_inputs(4000 + dilE, True)
    slaph.OperatorFactory(ctx).create_operators()
    slaph.contract(ctx)
    if (dilE, N) not in _DIL_WANT:
        cfg = o.parse_infile(text)
        lut = o.init_lookup_tables(cfg)
        ev = [ctx.get_input("eigenvectors", t) for t in range(cfg["Lt"])]
        perams = [ctx.get_input("perambulator", r) for r in range(4)]
        rnds = [ctx.get_input("random_vector", r) for r in range(4)]
        _DIL_WANT[(dilE, N)] = o.contract(cfg, lut, o.build_vdaggerv(cfg, lut, ev), perams, rnds)
    want = _DIL_WANT[(dilE, N)]
    assert set(want) == set(ctx.diagrams()) == {"C2c", "C3c", "C4cD", "C4cB", "C4cC"}
    worst = 0.0
    for d, lst in want.items():
        got = ctx.correlators(d)
        assert len(got) == len(lst)
        for name, vals in lst:
            err = np.max(np.abs(got[name] - vals)) / max(np.max(np.abs(vals)), 1e-300)
            worst = max(worst, err)
            assert err <= 1e-12, (d, name, err)
    say("  worst relative error %.2e" % worst)


TWO_FLAVOURS = """
Lt = 8
Lx = 4
Ly = 4
Lz = 4
number_of_eigen_vec = 8
[quarks]
quark = u:4:TB:2:EI:4:DF:4:synthetic
quark = s:2:TB:2:EI:4:DF:4:synthetic
[operator_lists]
operator_list = g5.d0.p0,1
operator_list = g2.d0.p0,1
[correlator_lists]
correlator_list = C4cC:Q0:Op0:Q1:Op0:Q0:Op0:Q1:Op0:P0,1
correlator_list = C4cC:Q0:Op0:Q0:Op1:Q0:Op0:Q0:Op1:P0
correlator_list = C2c:Q0:Op0:Q1:Op0:P0,1
"""
_TWO_WANT = {}


@pytest.mark.parametrize("form", ["summed", "summed-lstat", "summed-split", "per-combination"])
def test_crossed_traces_two_flavours_mixed_gammas(slaph, form, monkeypatch):
    """D = 16 crossed traces (C4cC) with two quark flavours (u u / s s random
    vector pairs of different sets, so the products' and traces' admitted
    chains differ per factor) and a gamma_2 operator beside gamma_5 (Q0's
    nonzero Dirac blocks off the diagonal): the summed-product form (its
    traces as symmetric products, also with components cut into chunks of
    four entries, and on the left-stationary kernel) and one product per
    combination against the oracle at 1e-12 norm-wise."""
    say("two flavours, mixed gammas", form)
    for var in KERNEL_SWITCHES:
        monkeypatch.delenv(var, raising=False)
    if form == "per-combination":
        monkeypatch.setenv("SLAPH_XFORM_OFF", "1")
    elif form == "summed-lstat":
        monkeypatch.setenv("SLAPH_SYM_OFF", "1")
    elif form == "summed-split":
        monkeypatch.setenv("SLAPH_SYM_ENTRIES", "4")
    ctx = slaph.Context(0)
    ctx.configure(TWO_FLAVOURS)
    ctx.synthetic_inputs(4242, True)
    slaph.OperatorFactory(ctx).create_operators()
    slaph.contract(ctx)
    if not _TWO_WANT:
        cfg = o.parse_infile(TWO_FLAVOURS)
        lut = o.init_lookup_tables(cfg)
        ev = [ctx.get_input("eigenvectors", t) for t in range(cfg["Lt"])]
        perams = [ctx.get_input("perambulator", r) for r in range(6)]
        rnds = [ctx.get_input("random_vector", r) for r in range(6)]
        _TWO_WANT.update(o.contract(cfg, lut, o.build_vdaggerv(cfg, lut, ev), perams, rnds))
    assert set(_TWO_WANT) == set(ctx.diagrams()) == {"C2c", "C4cC"}
    worst = 0.0
    for d, lst in _TWO_WANT.items():
        got = ctx.correlators(d)
        assert len(got) == len(lst)
        for name, vals in lst:
            err = np.max(np.abs(got[name] - vals)) / max(np.max(np.abs(vals)), 1e-300)
            worst = max(worst, err)
            assert err <= 1e-12, (d, name, err)
    say("  %d C4cC correlators, worst relative error %.2e" % (len(_TWO_WANT["C4cC"]), worst))


SIX_RANDOM = """
Lt = 4
Lx = 4
Ly = 4
Lz = 4
number_of_eigen_vec = 8
[quarks]
quark = u:6:TB:2:EI:4:DF:4:synthetic
[operator_lists]
operator_list = g5.d0.p0,1
[correlator_lists]
correlator_list = C4cC:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0
"""


def test_crossed_traces_six_random_vectors(slaph):
    """Six random vectors: a Q2 entry has 30 random-vector pairs, more than the
    summed-product kernel stages in LDS, so the crossed traces fall back to one
    product per combination instead of failing; against the oracle at 1e-12."""
    say("crossed traces, six random vectors")
    ctx = slaph.Context(0)
    ctx.configure(SIX_RANDOM)
    ctx.synthetic_inputs(606, True)
    slaph.OperatorFactory(ctx).create_operators()
    slaph.contract(ctx)
    cfg = o.parse_infile(SIX_RANDOM)
    lut = o.init_lookup_tables(cfg)
    ev = [ctx.get_input("eigenvectors", t) for t in range(cfg["Lt"])]
    perams = [ctx.get_input("perambulator", r) for r in range(6)]
    rnds = [ctx.get_input("random_vector", r) for r in range(6)]
    want = o.contract(cfg, lut, o.build_vdaggerv(cfg, lut, ev), perams, rnds)
    got = ctx.correlators("C4cC")
    assert len(got) == len(want["C4cC"]) > 0
    for name, vals in want["C4cC"]:
        err = np.max(np.abs(got[name] - vals)) / max(np.max(np.abs(vals)), 1e-300)
        assert err <= 1e-12, (name, err)


def _rel_errors(got, want):
    """Per-correlator norm-wise relative error max_t |C - C_ref| / max_t |C_ref|."""
    out = []
    for d, vals in want.items():
        for name, v in vals.items():
            out.append(np.max(np.abs(got[d][name] - v)) / max(np.max(np.abs(v)), 1e-300))
    return np.array(out)


VDV_ODD = """
Lt = 2
Lx = 4
Ly = 6
Lz = 5
number_of_eigen_vec = 37
momentum_cutoff_0 = 12
[quarks]
quark = u:4:TB:2:EI:4:DF:4:synthetic
[operator_lists]
operator_list = g5.d0.p0,1,2,3,4,5,6
[correlator_lists]
correlator_list = C2c:Q0:Op0:Q0:Op0:P0,1,2,3,4
correlator_list = C4cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0
"""


@pytest.mark.parametrize("direct", [False, True], ids=["fourier", "direct"])
def test_vdaggerv_odd_lattice_matches_oracle(slaph, direct, monkeypatch):
    """VdaggerV on a 4 x 6 x 5 lattice with N = 37 (three 16-column tiles, the
    last one ragged) and all momenta up to |p|^2 = 6 (components up to 2, so
    the Fourier form needs several passes over (py, pz) combinations with
    two nonzero |pz|): every operator at every timeslice against the oracle
    at 1e-13 relative to the largest entry."""
    say("vdaggerv odd lattice", "direct" if direct else "fourier")
    if direct:
        monkeypatch.setenv("SLAPH_VDV_DIRECT", "1")
    else:
        monkeypatch.delenv("SLAPH_VDV_DIRECT", raising=False)
    cfg = o.parse_infile(VDV_ODD)
    lut = o.init_lookup_tables(cfg)
    ctx = slaph.Context(0)
    ctx.configure(VDV_ODD)
    ctx.synthetic_inputs(21, False)
    ev = [ctx.get_input("eigenvectors", t) for t in range(cfg["Lt"])]
    ops = slaph.OperatorFactory(ctx)
    ops.create_operators()
    want = o.build_vdaggerv(cfg, lut, ev)
    assert len(lut["vdv_lookup"]) > 20
    for vid, mom, _ in lut["vdv_lookup"]:
        for t in range(cfg["Lt"]):
            w = want[(vid, t)]
            got = ops.return_vdaggerv(vid, t)
            assert np.max(np.abs(got - w)) <= 1e-13 * np.max(np.abs(w)), (vid, tuple(mom), t)


def test_precision_forms_against_fp64(slaph):
    """The precision forms run the same contraction with FP32 pieces: the
    single-pass FP32-MFMA VdaggerV and FP32 storage of factors and products.
    Every correlator stays within FP32 rounding of the FP64 path (the bound
    here is 1e-4; the FP32 unit roundoff is 6e-8), and switching back to FP64
    reproduces the FP64 result bit for bit."""
    say("precision forms")
    ctx = slaph.Context(0)
    ctx.configure(WIDE)
    ctx.synthetic_inputs(99, True)
    ops = slaph.OperatorFactory(ctx)

    def run(vdv, fac):
        ctx.set_precision(vdv, fac)
        ops.create_operators()
        slaph.contract(ctx)
        return {d: ctx.correlators(d) for d in ctx.diagrams()}

    ref = run("fp64", "fp64")
    for vdv, fac in (("fp32", "fp64"), ("fp64", "fp32"), ("fp32", "fp32")):
        err = _rel_errors(run(vdv, fac), ref)
        say("  vdaggerv %s factors %s: max %.2e median %.2e" % (vdv, fac, err.max(), np.median(err)))
        assert err.max() < 1e-4, (vdv, fac, err.max())
        assert err.max() > 0.0, (vdv, fac)  # the FP32 pieces did run
    again = run("fp64", "fp64")
    assert np.all(_rel_errors(again, ref) == 0.0)


@pytest.mark.parametrize("form,vdv_tol,corr_tol", [("bf16x3", 1e-5, 1e-4), ("ozaki", 1e-13, 1e-12)])
def test_split_vdaggerv_forms_against_fp64(slaph, form, vdv_tol, corr_tol):
    """VdaggerV on the low-precision matrix cores: three-term split BF16
    (FP32 accumulation; about FP32 accuracy) and the Ozaki int8 split (7
    slices of 7 bits, exact slice products, FP64 combination; FP64-level
    accuracy).  Every VdaggerV (p, t) against the FP64 path, relative to the
    largest entry, and every correlator norm-wise; the p = 0 operator stays
    the exact identity."""
    say("split vdaggerv", form)
    cfg = o.parse_infile(WIDE)
    lut = o.init_lookup_tables(cfg)
    ctx = slaph.Context(0)
    ctx.configure(WIDE)
    ctx.synthetic_inputs(77, True)
    ops = slaph.OperatorFactory(ctx)

    def run(vdv):
        ctx.set_precision(vdv, "fp64")
        ops.create_operators()
        slaph.contract(ctx)
        mats = {(vid, t): ops.return_vdaggerv(vid, t) for vid, _, _ in lut["vdv_lookup"] for t in range(cfg["Lt"])}
        return mats, {d: ctx.correlators(d) for d in ctx.diagrams()}

    ref_m, ref_c = run("fp64")
    got_m, got_c = run(form)
    worst = 0.0
    for key, w in ref_m.items():
        e = np.max(np.abs(got_m[key] - w)) / np.max(np.abs(w))
        worst = max(worst, e)
        assert e <= vdv_tol, (form, key, e)
    assert np.array_equal(got_m[(lut["index_of_unity"], 0)], np.eye(cfg["number_of_eigen_vec"]))
    err = _rel_errors(got_c, ref_c)
    say("  vdaggerv max %.2e; correlators max %.2e median %.2e" % (worst, err.max(), np.median(err)))
    assert err.max() <= corr_tol
    assert worst > 0.0  # the split form did run
    ctx.set_precision("fp64", "fp64")


@pytest.mark.parametrize("infile", ["wide", "EI6-D24", "EI8-D32"])
def test_fp48_storage_form_within_1e_10(slaph, infile):
    """Factor slabs and product tables in the 48-bit form (36 mantissa bits,
    relative rounding error <= 2^-37 per stored part; every product, trace and
    sum in FP64): every correlator within 1e-10 of the FP64 path (norm-wise,
    per correlator), on D = 16, 24 and 32."""
    text = {"wide": WIDE, "EI6-D24": dilution_infile(6, 14), "EI8-D32": dilution_infile(8, 20)}[infile]
    say("fp48 storage", infile)
    ctx = slaph.Context(0)
    ctx.configure(text)
    ctx.synthetic_inputs(314, True)
    ops = slaph.OperatorFactory(ctx)
    ops.create_operators()

    def run(fac):
        ctx.set_precision("fp64", fac)
        slaph.contract(ctx)
        return {d: ctx.correlators(d) for d in ctx.diagrams()}

    ref = run("fp64")
    err = _rel_errors(run("fp48"), ref)
    say("  max %.2e median %.2e over %d correlators" % (err.max(), np.median(err), err.size))
    assert err.max() <= 1e-10
    assert err.max() > 0.0  # the rounded storage did run
    assert np.all(_rel_errors(run("fp64"), ref) == 0.0)


@pytest.mark.parametrize("kahan", [False, True])
def test_graph_replays_match_direct_launches(slaph, kahan):
    """contract() with HIP graphs: the second call of a plan captures the
    launch sequence, later calls replay it.  Every replay -- with eager
    VdaggerV work, host synchronisations and new inputs between them -- gives
    correlators bit-identical to the directly launched contraction of a
    context without graphs on the same inputs."""
    say("graph replays", "kahan" if kahan else "")

    def context(graphs):
        ctx = slaph.Context(0)
        ctx.configure(WIDE)
        ctx.set_graphs(graphs)
        return ctx

    def run(ctx, seed):
        ctx.synthetic_inputs(seed, True)
        slaph.OperatorFactory(ctx).create_operators()
        slaph.contract(ctx, kahan=kahan)
        return {d: ctx.correlators(d) for d in ctx.diagrams()}

    eager = context(False)
    graphed = context(True)
    for call, seed in enumerate((5, 5, 5, 6, 5, 7)):
        want = run(eager, seed)
        got = run(graphed, seed)
        for d in want:
            for name, v in want[d].items():
                assert np.array_equal(got[d][name], v), (call, seed, d, name)


def test_autotune_keeps_correlators_and_tuning_text(slaph):
    """The autotuner times every candidate on the current inputs; whatever it
    picks (and every single non-default choice set by text) gives the
    correlators of the default schedule within 1e-12 norm-wise.  The choice
    is kept for the input file across configure(); bad tuning text raises."""
    say("autotune")
    ctx = slaph.Context(0)
    ctx.configure(WIDE)
    ctx.synthetic_inputs(21, True)
    ops = slaph.OperatorFactory(ctx)
    ops.create_operators()
    slaph.contract(ctx)
    ref = {d: ctx.correlators(d) for d in ctx.diagrams()}
    default = ctx.tuning()
    assert "lstat_wg_per_cu=2" in default and "vdaggerv_direct=0" in default
    rep = ctx.autotune(1)
    knobs = {k["knob"] for k in rep["knobs"]}
    assert knobs == {"vdaggerv_direct", "lstat_wg_per_cu", "gemm_wg_per_cu", "prod_generic", "resolve_groups",
                     "crossed_direct", "crossed_lstat"}
    assert all(len(k["ms"]) >= 2 and all(v > 0 for v in k["ms"].values()) for k in rep["knobs"])
    assert rep["chosen"] == ctx.tuning()
    slaph.contract(ctx)
    assert _rel_errors({d: ctx.correlators(d) for d in ctx.diagrams()}, ref).max() <= 1e-12
    for text in ("vdaggerv_direct=1", "lstat_wg_per_cu=1", "lstat_wg_per_cu=3", "gemm_wg_per_cu=8",
                 "prod_generic=1", "resolve_groups=1", "crossed_direct=1", "crossed_lstat=1"):
        ctx.set_tuning(default)
        ctx.set_tuning(text)
        ops.create_operators()
        slaph.contract(ctx)
        err = _rel_errors({d: ctx.correlators(d) for d in ctx.diagrams()}, ref).max()
        assert err <= 1e-12, (text, err)
    ctx.set_tuning("gemm_wg_per_cu=32")
    kept = ctx.tuning()
    ctx.configure(WIDE)
    assert ctx.tuning() == kept
    with pytest.raises(slaph.SlaphError, match="unknown key"):
        ctx.set_tuning("no_such_knob=1")
    with pytest.raises(slaph.SlaphError, match="bad value"):
        ctx.set_tuning("lstat_wg_per_cu=x")


def test_plan_reuse_is_deterministic(slaph):
    """The contraction plan is built on the first call and reused: a second
    call on new operators gives bit-identical correlators, a budget change
    rebuilds the plan (one-row chunks) and agrees to rounding."""
    say("plan reuse")
    ctx = slaph.Context(0)
    ctx.configure(WIDE)
    ctx.synthetic_inputs(77, True)
    ops = slaph.OperatorFactory(ctx)
    ops.create_operators()
    slaph.contract(ctx)
    first = {d: ctx.correlators(d) for d in ctx.diagrams()}
    ops.create_operators()
    slaph.contract(ctx)
    for d, vals in first.items():
        again = ctx.correlators(d)
        for name, v in vals.items():
            assert np.array_equal(again[name], v), (d, name)
    ctx.set_chunk_budget(1)
    slaph.contract(ctx)
    for d, vals in first.items():
        small = ctx.correlators(d)
        for name, v in vals.items():
            scale = max(np.max(np.abs(v)), 1e-300)
            assert np.max(np.abs(small[name] - v)) / scale <= 1e-13, (d, name)


def test_synthetic_inputs_match_oracle(slaph):
    """Device-generated inputs: eigenvectors orthonormal (V^H V = 1), random
    vectors in {(+-1 +-i)/sqrt 2}, and every correlator of the small pipi
    workload equal to the oracle's on the downloaded inputs."""
    say("synthetic inputs")
    cfg = o.parse_infile(SMALL)
    lut = o.init_lookup_tables(cfg)
    ctx = slaph.Context(0)
    ctx.configure(SMALL)
    ctx.synthetic_inputs(1234, True)
    Lt, nrnd = cfg["Lt"], 4
    ev = [ctx.get_input("eigenvectors", t) for t in range(Lt)]
    perams = [ctx.get_input("perambulator", r) for r in range(nrnd)]
    rnds = [ctx.get_input("random_vector", r) for r in range(nrnd)]
    for V in ev:
        assert np.max(np.abs(V.conj().T @ V - np.eye(V.shape[1]))) < 1e-13
    for rho in rnds:
        assert np.allclose(np.abs(rho.real), 2 ** -0.5) and np.allclose(np.abs(rho.imag), 2 ** -0.5)
    slaph.OperatorFactory(ctx).create_operators()
    slaph.contract(ctx)
    vdv = o.build_vdaggerv(cfg, lut, ev)
    want = o.contract(cfg, lut, vdv, perams, rnds)
    for d, lst in want.items():
        got = ctx.correlators(d)
        for name, vals in lst:
            scale = max(np.max(np.abs(vals)), 1e-300)
            assert np.max(np.abs(got[name] - vals)) / scale <= 1e-12, (d, name)


def test_not_enough_random_vectors_raises(slaph):
    ctx = slaph.Context(0)
    with pytest.raises(slaph.SlaphError, match="not enough random vectors"):
        ctx.configure(SMALL.replace("u:4:", "u:1:"))
