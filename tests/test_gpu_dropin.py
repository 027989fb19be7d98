"""The drop-in surface on the GPU: the `contract` program run as the
reference's integration test runs its own (tests/integration-L4/
run-integration-test:14-16 then compare:16: every output file against
correlators-reference with `h5diff -r -d 1e-12`), and the VdaggerV cache
files of handling_vdaggerv = write / read / liuming
(OperatorFactory.cpp:303-313, 328-493) round-tripped through the device."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
L4 = os.path.join(HERE, "golden", "l4")
REF = os.path.join(L4, "correlators-reference")
CONTRACT = os.path.join(ROOT, "slaph-contractions_amd", "contract")
H5DIFF = "/opt/conda/bin/h5diff"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def slaph():
    import slaph as s
    return s


def _ini(**settings):
    text = open(os.path.join(L4, "test_all.ini")).read()
    for k, v in settings.items():
        if re.search(r"^%s\s*=" % k, text, flags=re.M):
            text = re.sub(r"^%s\s*=.*$" % k, "%s = %s" % (k, v), text, flags=re.M)
        else:
            text = "%s = %s\n" % (k, v) + text
    return text


def _workdir(tmp_path, text):
    """The integration test's directory: the inputs of the golden lattice
    (linked) and the infile; correlators are written below it."""
    wd = tmp_path / "run"
    wd.mkdir()
    for d in ("eigensystem", "light", "strange"):
        os.symlink(os.path.join(L4, d), str(wd / d))
    (wd / "test_all.ini").write_text(text)
    return str(wd)


def _compare(outdir):
    """tests/integration-L4/compare: every reference file against ours."""
    files = sorted(f for f in os.listdir(REF) if f.endswith(".h5"))
    assert len(files) == 18
    for fn in files:
        r = subprocess.run([H5DIFF, "-r", "-d", "1e-12", os.path.join(REF, fn), os.path.join(outdir, fn)],
                           capture_output=True, text=True)
        assert r.returncode == 0, (fn, r.stdout[-3000:], r.stderr[-3000:])


@pytest.mark.skipif(not os.path.exists(CONTRACT), reason="contract program is not built (no HDF5)")
@pytest.mark.parametrize("extra", [[], ["--kahan", "--delta_config", "5"]], ids=["native", "kahan"])
def test_contract_program_integration_l4(tmp_path, extra):
    wd = _workdir(tmp_path, _ini())
    cmd = [CONTRACT, "-i", "test_all.ini", "--start_config", "1000", "--end_config", "1000"] + extra
    say("contract program", " ".join(cmd[1:]))
    r = subprocess.run(cmd, cwd=wd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "processing configuration: 1000" in r.stdout
    _compare(os.path.join(wd, "correlators"))
    # a second run into the same files keeps the existing datasets
    r = subprocess.run(cmd, cwd=wd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "because it exists" in r.stdout
    _compare(os.path.join(wd, "correlators"))


@pytest.mark.skipif(not os.path.exists(CONTRACT), reason="contract program is not built (no HDF5)")
def test_contract_program_rejects_bad_delta(tmp_path):
    wd = _workdir(tmp_path, _ini())
    r = subprocess.run([CONTRACT, "-i", "test_all.ini", "--delta_config", "0"], cwd=wd, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "non-positive value" in r.stderr


def _run(slaph, text, config=1000):
    ctx = slaph.Context(0)
    ctx.configure(text)
    ctx.read_inputs(L4, config)
    ops = slaph.OperatorFactory(ctx)
    ops.create_operators()
    slaph.contract(ctx)
    return ctx, ops, {d: ctx.correlators(d) for d in ctx.diagrams()}


def test_vdaggerv_files_write_read_liuming(slaph, tmp_path):
    """write stores every non-trivial operator under the reference's name
    (column-major, bit-equal to the device matrix); read loads them back from
    the reference's read names, liuming from one row-major file per operator
    (the -p file holding the adjoint); each gives bit-identical operators and
    correlators, and the golden files still pass."""
    vpath = str(tmp_path / "vdv")
    say("vdaggerv write")
    ctx, ops, base = _run(slaph, _ini(handling_vdaggerv="write", path_vdaggerv=vpath))
    Lt, N = ctx.query("Lt"), ctx.query("nev")
    import oracle.slaph_oracle as o
    cfg = o.parse_infile(_ini())
    lut = o.init_lookup_tables(cfg)
    unity = lut["index_of_unity"]
    built = {}
    for vid, mom, disp in lut["vdv_lookup"]:
        for t in range(Lt):
            built[(vid, t)] = ops.return_vdaggerv(vid, t)
            if vid == unity:
                continue
            name = slaph.vdaggerv_filename("write", vpath, 1000, mom, "", t)
            m = np.fromfile(name, dtype=np.complex128).reshape(N, N, order="F")
            assert np.array_equal(m, built[(vid, t)]), name
            os.rename(name, slaph.vdaggerv_filename("read", vpath, 1000, mom, "", t))
    ctx.close()
    say("vdaggerv read")
    ctx, ops, got = _run(slaph, _ini(handling_vdaggerv="read", path_vdaggerv=vpath))
    for (vid, t), m in built.items():
        assert np.array_equal(ops.return_vdaggerv(vid, t), m), (vid, t)
    for d, vals in base.items():
        for name, v in vals.items():
            assert np.array_equal(got[d][name], v), (d, name)
    ctx.close()
    say("vdaggerv liuming")
    lpath = str(tmp_path / "liuming")
    os.makedirs(lpath)
    for k, (vid, mom, disp) in enumerate(lut["vdv_lookup"]):
        if vid == unity:
            continue
        neg = k % 2 == 0  # half the operators in -p files, half in +p files
        name = slaph.vdaggerv_filename("liuming", lpath, 1000, [-c for c in mom] if neg else mom)
        with open("/" + name.lstrip("/"), "wb") as fh:
            for t in range(Lt):
                m = built[(vid, t)]
                (m.conj().T if neg else m).astype(np.complex128).tofile(fh)  # row-major: C order of (r, c)
    ctx, ops, got = _run(slaph, _ini(handling_vdaggerv="liuming", path_vdaggerv=lpath))
    for (vid, t), m in built.items():
        assert np.array_equal(ops.return_vdaggerv(vid, t), m), (vid, t)
    for d, vals in base.items():
        for name, v in vals.items():
            assert np.array_equal(got[d][name], v), (d, name)
    ctx.close()


def test_vdaggerv_read_missing_file_raises(slaph, tmp_path):
    ctx = slaph.Context(0)
    ctx.configure(_ini(handling_vdaggerv="read", path_vdaggerv=str(tmp_path / "none")))
    ctx.read_inputs(L4, 1000)
    with pytest.raises(slaph.SlaphError, match="failed to open file"):
        slaph.OperatorFactory(ctx).create_operators()


@pytest.mark.skipif(not os.path.exists(CONTRACT), reason="HDF5 writer is not built")
def test_class_layer_program_l4(tmp_path):
    """The C++ caller of tests/cpp/class_layer_l4.cpp (reference class names
    over the C ABI): its VdaggerV, Q0 / Q2 factors, trQ0Q2 trace vector and
    inner products against the oracle, and its correlator files through
    WriteHDF5Correlator against the golden files with h5diff."""
    from test_host import _build_class_layer_program
    import oracle.slaph_oracle as o
    prog = str(tmp_path / "prog")
    r = _build_class_layer_program(prog)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "out"
    out.mkdir()
    say("class layer program")
    r = subprocess.run([prog, L4, str(out)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    _compare(str(out / "correlators"))
    cfg = o.parse_infile(open(os.path.join(L4, "test_all.ini")).read())
    lut = o.init_lookup_tables(cfg)
    perams, rnds, ev = o.read_inputs(cfg, L4, 1000)
    vdv = o.build_vdaggerv(cfg, lut, ev)
    got = np.fromfile(str(out / "vdv_1_2.bin"), dtype=np.complex128).reshape(16, 16, order="F")
    assert np.max(np.abs(got - vdv[(1, 2)])) < 1e-13
    q = cfg["quarks"][0]
    fac = o.Factories(cfg, lut, vdv, perams, rnds, q["dilT"], q["dilE"])
    rics = [ln.split() for ln in open(str(out / "ric.txt"))]
    D = 4 * q["dilE"]
    for qt, key, op, fn in (("Q0", (1,), 0, "q0_t1_op0.bin"), ("Q2", (1, 2, 0), 3, "q2_1_2_0_op3.bin")):
        want = fac.get(qt, key)[op]
        mats = np.fromfile(str(out / fn), dtype=np.complex128).reshape(len(want), D, D).transpose(0, 2, 1)
        assert [tuple(map(int, r[1:])) for r in rics if r[0] == qt] == [w[1] for w in want]
        for gm, (wm, _, _) in zip(mats, want):
            assert np.max(np.abs(gm - wm)) < 1e-12
    tr = o.Traces(fac, lut, q["dilT"])
    locs = None
    for reqs in lut["requests"].values():
        for trs, _ in reqs:
            for name, _, l in trs:
                if name == "trQ0Q2" and locs is None:
                    locs = l
    key = [0 if l == "source" else 1 for l in locs]
    w0, w1 = tr.get("trQ0Q2", key, 0), tr.get("trQ0Q2", key, 1)
    t0 = np.fromfile(str(out / "trQ0Q2_0_1_id0.bin"), dtype=np.complex128)
    assert [int(x) for x in open(str(out / "used.txt"))] == [m for _, m in w0]
    assert np.max(np.abs(t0 - np.array([v for v, _ in w0]))) <= 1e-12 * max(abs(v) for v, _ in w0)
    ip = np.fromfile(str(out / "inner_0_1.bin"), dtype=np.complex128)
    for val, kahan in zip(ip, (False, True)):
        want = o.inner_product2(w0, w1, kahan)
        assert abs(val - want) <= 1e-12 * max(abs(want), 1e-300)


GAUGE = os.path.join(L4, "gauge")


def _write_expected(outdir, correlators, config):
    """The oracle's correlators as the reference's HDF5 files, through the
    library's WriteHDF5Correlator (libslaph_h5.so)."""
    import ctypes as c
    h5 = c.CDLL(os.path.join(ROOT, "slaph-contractions_amd", "libslaph_h5.so"))
    h5.slh5_open.argtypes = [c.c_char_p, c.c_char_p, c.c_char_p, c.POINTER(c.c_void_p)]
    h5.slh5_write.argtypes = [c.c_void_p, c.c_char_p, c.POINTER(c.c_double), c.c_int]
    h5.slh5_close.argtypes = [c.c_void_p]
    h5.slh5_filename_ending.argtypes = [c.c_int, c.c_char_p, c.c_int]
    ending = c.create_string_buffer(64)
    h5.slh5_filename_ending(config, ending, 64)
    for diagram, lst in correlators.items():
        f = c.c_void_p()
        assert h5.slh5_open(outdir.encode(), diagram.encode(), ending.value, c.byref(f)) == 0
        for name, vals in lst:
            v = np.ascontiguousarray(vals, dtype=np.complex128)
            assert h5.slh5_write(f, name.encode(), v.ctypes.data_as(c.POINTER(c.c_double)), len(v)) == 0
        h5.slh5_close(f)


@pytest.mark.skipif(not os.path.exists(CONTRACT), reason="contract program is not built (no HDF5)")
def test_contract_program_displaced_operators_from_lime_gauge(tmp_path):
    """Displaced operators end to end through the `contract` program: the
    gauge configuration read from <path_config>/conf.1000 (the reference's
    lime file), HYP-smeared with the infile's alpha1 / alpha2 / iterations
    (OperatorFactory.cpp:241-261), V^+ (D V) for '>x,<y' and '>z|<x' on the
    golden L4 eigenvectors and perambulators; every written dataset against
    the oracle's (read_lime_gauge, hyp_step, build_vdaggerv with links,
    contract) with h5diff -d 1e-12, the reference's gate."""
    import oracle.slaph_oracle as o
    text = open(os.path.join(L4, "test_all.ini")).read().split("[operator_lists]")[0]
    text = re.sub(r"^path_config\s*=.*$", "path_config = gauge\nalpha1 = 0.75\nalpha2 = 0.6\niterations = 2", text,
                  flags=re.M)
    text += """[operator_lists]
operator_list = g5.d0.p0:g5.d>x,<y.p0
operator_list = g5.d>z|<x.p1
[correlator_lists]
correlator_list = C2c:Q0:Op0:Q0:Op0:P0
correlator_list = C2c:Q0:Op1:Q0:Op1:P1
correlator_list = C4cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0
"""
    wd = _workdir(tmp_path, text)
    os.symlink(GAUGE, os.path.join(wd, "gauge"))
    cmd = [CONTRACT, "-i", "test_all.ini", "--start_config", "1000", "--end_config", "1000"]
    say("contract program, displaced operators from the lime file")
    r = subprocess.run(cmd, cwd=wd, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    cfg = o.parse_infile(text)
    lut = o.init_lookup_tables(cfg)
    assert any(d for _, _, d in lut["vdv_lookup"])
    perams, rnds, ev = o.read_inputs(cfg, L4, 1000)
    U, _ = o.read_lime_gauge(os.path.join(GAUGE, "conf.1000"), (4, 4, 4), 4)
    smeared = [o.hyp_step(o.hyp_step(u, 0.75, 0.6, (4, 4, 4)), 0.75, 0.6, (4, 4, 4)) for u in U]
    vdv = o.build_vdaggerv(cfg, lut, ev, links=smeared)
    want = o.contract(cfg, lut, vdv, perams, rnds)
    exp = str(tmp_path / "expected")
    _write_expected(exp, want, 1000)
    files = sorted(os.listdir(exp))
    assert files == sorted(f for f in os.listdir(os.path.join(wd, "correlators")))
    for fn in files:
        r = subprocess.run([H5DIFF, "-r", "-d", "1e-12", os.path.join(exp, fn), os.path.join(wd, "correlators", fn)],
                           capture_output=True, text=True)
        assert r.returncode == 0, (fn, r.stdout[-3000:], r.stderr[-3000:])
    # without the gauge file the program stops with the reader's error
    os.unlink(os.path.join(wd, "gauge"))
    r = subprocess.run(cmd, cwd=wd, capture_output=True, text=True, timeout=240)
    assert r.returncode != 0 and "Could not open file gauge/conf.1000" in (r.stdout + r.stderr)
