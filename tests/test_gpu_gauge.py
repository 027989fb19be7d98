"""Displaced operators, gauge links, stout smearing and timeslice streaming
on the device against the oracle (displace, stout_step, build_vdaggerv with
links; OperatorFactory.cpp:33-55, 220-326, GaugeField.cpp:228-274, 543-560).

Tolerances: gauge links and VdaggerV 1e-13 absolute (unitary links, unit
columns); correlators 1e-12 relative to the largest magnitude of each
correlator, the bar of the undisplaced tests.
"""
import os
import sys

import numpy as np
import pytest

from oracle import slaph_oracle as o

pytestmark = pytest.mark.gpu

DISP = """
Lt = 4
Lx = 4
Ly = 4
Lz = 4
number_of_eigen_vec = 8
[quarks]
quark = u:4:TB:2:EI:4:DF:4:synthetic
[operator_lists]
operator_list = g5.d0.p0:g5.d>x,<y.p0
operator_list = g5.d>z|<x.p1
[correlator_lists]
correlator_list = C2c:Q0:Op0:Q0:Op0:P0
correlator_list = C2c:Q0:Op1:Q0:Op1:P1
correlator_list = C4cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0
"""
L = (4, 4, 4)
SITES = 64


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def slaph():
    import slaph as s
    return s


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _su3(rng, n, eps):
    import scipy.linalg
    x = _rand(rng, n, 3, 3)
    h = 0.5 * (x + x.conj().swapaxes(1, 2))
    h -= (np.trace(h, axis1=1, axis2=2) / 3.0)[:, None, None] * np.eye(3)
    return np.stack([scipy.linalg.expm(1j * eps * m) for m in h])


@pytest.fixture(scope="module")
def inputs():
    cfg = o.parse_infile(DISP)
    lut = o.init_lookup_tables(cfg)
    rng = np.random.default_rng(21)
    Lt, N = cfg["Lt"], cfg["number_of_eigen_vec"]
    ev = [np.linalg.qr(_rand(rng, 3 * SITES, N))[0] for _ in range(Lt)]
    links = [_su3(rng, SITES * 3, 0.5).reshape(SITES, 3, 3, 3) for _ in range(Lt)]
    perams = [_rand(rng, 4 * N * Lt, 2 * 4 * 4) for _ in range(4)]
    rnds = [((rng.integers(0, 2, Lt * N * 4) * 2 - 1) + 1j * (rng.integers(0, 2, Lt * N * 4) * 2 - 1)) / np.sqrt(2)
            for _ in range(4)]
    return cfg, lut, ev, links, perams, rnds


def _context(slaph, inputs, links=True):
    cfg, lut, ev, U, perams, rnds = inputs
    ctx = slaph.Context(0)
    ctx.configure(DISP)
    for t in range(cfg["Lt"]):
        ctx.set_eigenvectors(t, ev[t])
        if links:
            ctx.set_gauge_links(t, U[t])
    for r in range(4):
        ctx.set_perambulator(r, perams[r])
        ctx.set_random_vector(r, rnds[r])
    return ctx


def _check_correlators(ctx, want, tol=1e-12):
    for d, lst in want.items():
        got = ctx.correlators(d)
        for name, vals in lst:
            scale = max(np.max(np.abs(vals)), 1e-300)
            assert np.max(np.abs(got[name] - vals)) / scale <= tol, (d, name)


def test_lookup_has_displaced_operators(inputs):
    cfg, lut = inputs[0], inputs[1]
    disps = [d for _, _, d in lut["vdv_lookup"]]
    assert ((">", "x"),) in disps and (("<", "y"),) in disps and ((">", "z"), ("<", "x")) in disps


def test_synthetic_gauge_links_are_su3_and_seeded(slaph):
    say("synthetic gauge")
    ctx = slaph.Context(0)
    ctx.configure(DISP)
    ctx.synthetic_gauge(77, 0.3)
    a = [ctx.get_gauge_links(t) for t in range(4)]
    ctx.synthetic_gauge(77, 0.3)
    assert all(np.array_equal(x, ctx.get_gauge_links(t)) for t, x in enumerate(a))
    m = np.concatenate(a).reshape(-1, 3, 3)
    assert np.max(np.abs(m @ m.conj().swapaxes(1, 2) - np.eye(3))) < 1e-13
    assert np.max(np.abs(np.linalg.det(m) - 1.0)) < 1e-13
    assert np.max(np.abs(m - np.eye(3))) > 0.1  # not trivial
    ctx.synthetic_gauge(78, 0.3)
    assert not np.array_equal(a[0], ctx.get_gauge_links(0))


def test_stout_smearing_matches_oracle(slaph, inputs):
    say("stout smearing")
    U = inputs[3]
    ctx = slaph.Context(0)
    ctx.configure(DISP)
    for t in range(4):
        ctx.set_gauge_links(t, U[t])
    ctx.smear_stout(0.0, 1)
    assert max(np.max(np.abs(ctx.get_gauge_links(t) - U[t])) for t in range(4)) < 1e-15
    ctx.smear_stout(0.1, 3)
    for t in range(4):
        want = U[t]
        for _ in range(3):
            want = o.stout_step(want, 0.1, L)
        got = ctx.get_gauge_links(t)
        assert np.max(np.abs(got - want)) < 1e-13, t
    with pytest.raises(slaph.SlaphError):
        ctx.smear_stout(0.1, -1)


def test_hyp_smearing_matches_oracle(slaph, inputs):
    """Two HYP steps (alpha1 = 0.75, alpha2 = 0.6) of every timeslice against
    the oracle's hyp_step (SVD polar factor; the device uses the Newton
    iteration), and zero parameters as the identity."""
    say("hyp smearing")
    U = inputs[3]
    ctx = slaph.Context(0)
    ctx.configure(DISP)
    for t in range(4):
        ctx.set_gauge_links(t, U[t])
    ctx.smear_hyp(0.0, 0.0, 1)
    assert max(np.max(np.abs(ctx.get_gauge_links(t) - U[t])) for t in range(4)) < 1e-14
    ctx.smear_hyp(0.75, 0.6, 2)
    for t in range(4):
        want = o.hyp_step(o.hyp_step(U[t], 0.75, 0.6, L), 0.75, 0.6, L)
        assert np.max(np.abs(ctx.get_gauge_links(t) - want)) < 1e-13, t


def test_infile_hyp_parameters_smear_a_copy_for_the_displacements(slaph, inputs):
    """alpha1 / alpha2 / iterations in the input file: create_operators smears
    a copy of the links (OperatorFactory.cpp:249-260) and displaces with it;
    the stored links stay as set."""
    say("infile hyp")
    cfg, lut, ev, U = inputs[:4]
    text = DISP.replace("number_of_eigen_vec = 8", "number_of_eigen_vec = 8\nalpha1 = 0.75\nalpha2 = 0.6\niterations = 2")
    ctx = slaph.Context(0)
    ctx.configure(text)
    for t in range(cfg["Lt"]):
        ctx.set_eigenvectors(t, ev[t])
        ctx.set_gauge_links(t, U[t])
    for r in range(4):
        ctx.set_perambulator(r, inputs[4][r])
        ctx.set_random_vector(r, inputs[5][r])
    of = slaph.OperatorFactory(ctx)
    of.create_operators()
    smeared = [o.hyp_step(o.hyp_step(u, 0.75, 0.6, L), 0.75, 0.6, L) for u in U]
    want = o.build_vdaggerv(cfg, lut, ev, links=smeared)
    for (vid, t), W in want.items():
        assert np.max(np.abs(of.return_vdaggerv(vid, t) - W)) < 1e-13, (vid, t)
    assert all(np.array_equal(ctx.get_gauge_links(t), U[t]) for t in range(4))


def test_displaced_vdaggerv_matches_oracle(slaph, inputs):
    say("displaced vdaggerv")
    cfg, lut, ev, U = inputs[:4]
    ctx = _context(slaph, inputs)
    slaph.OperatorFactory(ctx).create_operators()
    want = o.build_vdaggerv(cfg, lut, ev, links=U)
    for (vid, t), W in want.items():
        got = slaph.OperatorFactory(ctx).return_vdaggerv(vid, t)
        assert np.max(np.abs(got - W)) < 1e-13, (vid, t, lut["vdv_lookup"][vid])


@pytest.mark.parametrize("form,tol", [("bf16x3", 1e-5), ("ozaki", 1e-13)])
def test_displaced_vdaggerv_split_forms(slaph, inputs, form, tol):
    """Displaced operators through the split-BF16 and Ozaki forms (right
    operand W = D V with its own column exponents) against the oracle."""
    say("displaced vdaggerv", form)
    cfg, lut, ev, U = inputs[:4]
    ctx = _context(slaph, inputs)
    ctx.set_precision(form, "fp64")
    slaph.OperatorFactory(ctx).create_operators()
    want = o.build_vdaggerv(cfg, lut, ev, links=U)
    for (vid, t), W in want.items():
        got = slaph.OperatorFactory(ctx).return_vdaggerv(vid, t)
        assert np.max(np.abs(got - W)) <= tol * np.max(np.abs(W)), (form, vid, t)


def test_displaced_contraction_matches_oracle(slaph, inputs):
    say("displaced contraction")
    cfg, lut, ev, U, perams, rnds = inputs
    ctx = _context(slaph, inputs)
    slaph.OperatorFactory(ctx).create_operators()
    slaph.contract(ctx)
    vdv = o.build_vdaggerv(cfg, lut, ev, links=U)
    _check_correlators(ctx, o.contract(cfg, lut, vdv, perams, rnds))


def test_gauge_invariance_on_device(slaph, inputs):
    """V -> g V, U_mu(x) -> g(x) U_mu(x) g(x + mu)^+ leaves every VdaggerV
    (displaced or not) unchanged."""
    say("gauge invariance")
    cfg, lut, ev, U, perams, rnds = inputs
    rng = np.random.default_rng(5)
    ctx = _context(slaph, inputs)
    slaph.OperatorFactory(ctx).create_operators()
    of = slaph.OperatorFactory(ctx)
    ref = {(vid, t): of.return_vdaggerv(vid, t) for vid, _, _ in lut["vdv_lookup"] for t in range(4)}
    for t in range(4):
        g = _su3(rng, SITES, 1.0)
        ctx.set_eigenvectors(t, (g @ ev[t].reshape(SITES, 3, -1)).reshape(ev[t].shape))
        Ug = np.empty_like(U[t])
        for mu in range(3):
            Ug[:, mu] = g @ U[t][:, mu] @ g[o.neighbour(L, mu, 1)].conj().swapaxes(1, 2)
        ctx.set_gauge_links(t, Ug)
    slaph.OperatorFactory(ctx).create_operators()
    for key, W in ref.items():
        assert np.max(np.abs(of.return_vdaggerv(*key) - W)) < 1e-12, key


def test_streaming_equals_resident_with_smeared_synthetic_links(slaph):
    """One eigenvector timeslice resident (generated per timeslice inside
    create_operators) against all timeslices resident, on the same seeded
    inputs and the same device-generated, stout-smeared links; and both
    against the oracle on the inputs read back."""
    say("streaming vs resident")
    out = []
    for streaming in (False, True):
        ctx = slaph.Context(0)
        ctx.configure(DISP)
        if streaming:
            ctx.set_timeslice_streaming(True)
        ctx.synthetic_inputs(99, True)
        ctx.synthetic_gauge(5, 0.2)
        ctx.smear_stout(0.1, 3)
        of = slaph.OperatorFactory(ctx)
        of.create_operators()
        slaph.contract(ctx)
        vdv = {(vid, t): of.return_vdaggerv(vid, t) for vid in range(ctx.query("n_vdaggerv")) for t in range(4)}
        corr = {d: ctx.correlators(d) for d in ctx.diagrams()}
        out.append((ctx, vdv, corr))
    (ctx_r, vdv_r, corr_r), (ctx_s, vdv_s, corr_s) = out
    for key in vdv_r:
        assert np.max(np.abs(vdv_r[key] - vdv_s[key])) < 1e-14, key
    for d in corr_r:
        for name in corr_r[d]:
            a, b = corr_r[d][name], corr_s[d][name]
            assert np.max(np.abs(a - b)) <= 1e-13 * max(np.max(np.abs(a)), 1e-300), (d, name)
    cfg = o.parse_infile(DISP)
    lut = o.init_lookup_tables(cfg)
    ev = [ctx_r.get_input("eigenvectors", t) for t in range(4)]
    ev_s = [ctx_s.get_input("eigenvectors", t) for t in range(4)]
    assert all(np.array_equal(a, b) for a, b in zip(ev, ev_s))
    U = [ctx_r.get_gauge_links(t) for t in range(4)]
    perams = [ctx_r.get_input("perambulator", r) for r in range(4)]
    rnds = [ctx_r.get_input("random_vector", r) for r in range(4)]
    vdv = o.build_vdaggerv(cfg, lut, ev, links=U)
    _check_correlators(ctx_s, o.contract(cfg, lut, vdv, perams, rnds))


def test_displaced_operators_without_links_raise(slaph, inputs):
    ctx = _context(slaph, inputs, links=False)
    with pytest.raises(slaph.SlaphError, match="gauge links"):
        slaph.OperatorFactory(ctx).create_operators()


GAUGE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "l4", "gauge")


def test_read_gauge_field_and_plaquette(slaph, tmp_path):
    """read_gauge_field of the reference's lime file (tests/integration-L4/
    gauge/conf.1000): the device links equal the oracle's reading bit for bit,
    plaque_ts on the device matches the oracle's per timeslice (1e-14), a
    partial slice range reads those timeslices only, and a file that fails
    its checksum leaves the links unset."""
    import shutil
    say("read_gauge_field + plaquette")
    want, _ = o.read_lime_gauge(os.path.join(GAUGE_DIR, "conf.1000"), L, 4)
    ctx = slaph.Context(0)
    ctx.configure(DISP.replace("Lz = 4", "Lz = 4\npath_config = " + GAUGE_DIR))
    ctx.read_gauge_field(1000)
    for t in range(4):
        assert np.array_equal(ctx.get_gauge_links(t), want[t]), t
        assert abs(ctx.plaquette(t) - o.plaquette_ts(want[t], L)) < 1e-14, t
    ctx.set_gauge_links(2, np.broadcast_to(np.eye(3), (SITES, 3, 3, 3)))
    assert abs(ctx.plaquette(2) - 1.0) < 1e-15
    ctx.read_gauge_field(1000, slice_i=2, slice_f=2, config_path=GAUGE_DIR)
    assert np.array_equal(ctx.get_gauge_links(2), want[2])
    # a corrupted copy fails its SciDAC checksum
    data = bytearray(open(os.path.join(GAUGE_DIR, "conf.1000"), "rb").read())
    data[data.index(b"ildg-binary-data") + 128 + 5000] ^= 0x10  # payload = type string + 128 B
    (tmp_path / "conf.1000").write_bytes(bytes(data))
    with pytest.raises(slaph.SlaphError, match="checksum mismatch"):
        ctx.read_gauge_field(1000, config_path=str(tmp_path))
    with pytest.raises(slaph.SlaphError, match="gauge links are not set"):
        ctx.plaquette(0)
    shutil.copy(os.path.join(GAUGE_DIR, "conf.1000"), str(tmp_path / "conf.1001"))
    ctx.read_gauge_field(1001, config_path=str(tmp_path))
    assert np.array_equal(ctx.get_gauge_links(3), want[3])
