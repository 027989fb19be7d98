"""The HDF5 correlator files (include/slaph_h5.h, libslaph_h5.so) against the
reference's own golden files: the golden correlator values written through
the writer must pass the reference's gate, `h5diff -r -d 1e-12`
(tests/integration-L4/compare:16), against correlators-reference/*.h5 -- same
file names, dataset names, compound type {re, im} and values.  Existing
datasets are kept, as WriteHDF5Correlator::write does (h5-wrapper.hpp:145-150).
Host only (no device)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
REF = os.path.join(GOLD, "l4", "correlators-reference")
H5LIB = os.path.join(ROOT, "slaph-contractions_amd", "libslaph_h5.so")
H5DIFF = "/opt/conda/bin/h5diff"

pytestmark = pytest.mark.skipif(not os.path.exists(H5LIB), reason="libslaph_h5.so is not built")


@pytest.fixture(scope="module")
def h5():
    lib = ctypes.CDLL(H5LIB)
    lib.slh5_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.slh5_write.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.slh5_close.argtypes = [ctypes.c_void_p]
    lib.slh5_filename_ending.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.slh5_last_error.restype = ctypes.c_char_p
    return lib


def write_diagrams(h5, outdir, config, diagrams):
    """diagrams: {diagram: [(dataset, complex values)]} -> one file per diagram."""
    ending = ctypes.create_string_buffer(64)
    h5.slh5_filename_ending(config, ending, 64)
    status = []
    for d, lst in diagrams.items():
        f = ctypes.c_void_p()
        assert h5.slh5_open(outdir.encode(), d.encode(), ending.value, ctypes.byref(f)) == 0, h5.slh5_last_error()
        for name, vals in lst:
            v = np.ascontiguousarray(vals, dtype=np.complex128).view(np.float64)
            status.append(h5.slh5_write(f, name.encode(), v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        len(vals)))
        h5.slh5_close(f)
    return status


def golden_by_diagram():
    gold = np.load(os.path.join(GOLD, "l4_correlators.npz"))
    out = {}
    for key in gold.files:
        d, name = key.split("/", 1)
        out.setdefault(d, []).append((name, gold[key]))
    return out


def test_filename_ending(h5):
    buf = ctypes.create_string_buffer(64)
    assert h5.slh5_filename_ending(1000, buf, 64) == len("_cnfg1000.h5") + 1
    assert buf.value == b"_cnfg1000.h5"
    h5.slh5_filename_ending(7, buf, 64)
    assert buf.value == b"_cnfg0007.h5"


@pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(H5DIFF)), reason="golden files or h5diff missing")
def test_golden_values_through_writer_pass_h5diff(h5, tmp_path):
    out = str(tmp_path / "correlators" / "nested")  # missing directories are created
    status = write_diagrams(h5, out, 1000, golden_by_diagram())
    assert all(s == 0 for s in status)
    files = sorted(os.listdir(REF))
    assert len(files) == 18 and sorted(os.listdir(out)) == files
    for fn in files:
        r = subprocess.run([H5DIFF, "-r", "-d", "1e-12", os.path.join(REF, fn), os.path.join(out, fn)],
                           capture_output=True, text=True)
        assert r.returncode == 0, (fn, r.stdout[-2000:], r.stderr[-2000:])


@pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(H5DIFF)), reason="golden files or h5diff missing")
def test_existing_datasets_are_kept_and_changes_are_detected(h5, tmp_path):
    out = str(tmp_path)
    gold = golden_by_diagram()
    c2c = {"C2c": gold["C2c"]}
    write_diagrams(h5, out, 1000, c2c)
    # a second write of different values into the existing file keeps the old ones
    bumped = {"C2c": [(n, v + 1.0) for n, v in gold["C2c"]]}
    status = write_diagrams(h5, out, 1000, bumped)
    assert all(s == 1 for s in status)
    ref = os.path.join(REF, "C2c_cnfg1000.h5")
    mine = os.path.join(out, "C2c_cnfg1000.h5")
    assert subprocess.run([H5DIFF, "-r", "-d", "1e-12", ref, mine], capture_output=True).returncode == 0
    # and the gate does see a wrong value
    write_diagrams(h5, str(tmp_path / "bad"), 1000, bumped)
    bad = os.path.join(str(tmp_path / "bad"), "C2c_cnfg1000.h5")
    assert subprocess.run([H5DIFF, "-r", "-d", "1e-12", ref, bad], capture_output=True).returncode != 0
