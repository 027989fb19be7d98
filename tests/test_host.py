"""Host-side checks that need no GPU: the C-ABI library loads and exports every
symbol the header declares; the C++ lookup tables and dilution scheme match the
oracle and the reference's fixtures bit for bit."""
import os
import re

import numpy as np
import pytest

from oracle import slaph_oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def slaph():
    import slaph as s
    return s


def test_library_exports_every_declared_symbol(slaph):
    header = open(os.path.join(ROOT, "include", "slaph_amd.h")).read()
    declared = set(re.findall(r"\b(slc_[a-z_0-9]+)\s*\(", header))
    assert declared, "no declarations found"
    import ctypes
    lib = ctypes.CDLL(slaph.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(slaph.exported_symbols())


def test_h5_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "slaph_h5.h")).read()
    declared = set(re.findall(r"\b(slh5_[a-z_0-9]+)\s*\(", header))
    assert declared == {"slh5_open", "slh5_write", "slh5_close", "slh5_filename_ending", "slh5_last_error"}
    path = os.path.join(ROOT, "slaph-contractions_amd", "libslaph_h5.so")
    if not os.path.exists(path):
        pytest.skip("libslaph_h5.so is not built (no HDF5)")
    import ctypes
    lib = ctypes.CDLL(path)
    for name in declared:
        assert hasattr(lib, name), name


def test_vdaggerv_file_names_follow_the_reference(slaph):
    """The format strings of OperatorFactory.cpp: written operators
    "/%s/cnfg%04d/" + "operators.%04d.p_%s.d_%s.t_%03d" (:224-226, 306-311),
    read ones ".p_" + digits + "_.t_%03d" (:331-333, 352-356), Liuming's
    "/%s/VdaggerV." + "p<x>p<y>p<z>.conf" + "%04d" (:399-427)."""
    f = slaph.vdaggerv_filename
    assert f("write", "vdv", 1000, (1, 0, -1), "", 3) == "/vdv/cnfg1000/operators.1000.p_10-1.d_000.t_003"
    assert f("write", "/abs/v", 7, (0, 0, 2), ">x<y", 12) == "//abs/v/cnfg0007/operators.0007.p_002.d_>x<y.t_012"
    assert f("read", "vdv", 1000, (1, 0, -1), "", 3) == "/vdv/cnfg1000/operators.1000.p_10-1_.t_003"
    assert f("liuming", "vdv", 7, (0, -1, 2)) == "/vdv/VdaggerV.p0p-1p2.conf0007"


def test_dilution_text_matches_control(slaph):
    txt = (slaph.dilution_text(6, 2, False) + slaph.dilution_text(6, 2, True)
           + slaph.dilution_text(48, 4, False) + slaph.dilution_text(48, 4, True)
           + slaph.dilution_text(48, 24, False))
    assert txt == open(os.path.join(GOLD, "dilution_iterator_control.txt")).read()


def oracle_dump(lut):
    """The oracle's lookup tables in the text format of slc_lookup_dump."""
    lines = ["unity %d" % lut["index_of_unity"]]
    for vid, mom, disp in lut["vdv_lookup"]:
        d = "000" if not disp else "".join(a + b for a, b in disp)
        lines.append("vdv %d %d %d %d %s" % (vid, mom[0], mom[1], mom[2], d))
    for q in ("Q0", "Q1", "Q2"):
        for i, (vid, dag, gam, rnd) in enumerate(lut["ql_lookup"][q]):
            lines.append("%s %d %d %d g%s r%s" % (q, i, vid, int(dag), "".join("%d," % g for g in gam),
                                                 "".join("%d:%d," % p for p in rnd)))
    for name in sorted(lut["trace_indices"]):
        for i, ids in enumerate(lut["trace_indices"][name]):
            lines.append("%s %d %s" % (name, i, " ".join(str(c) for c in ids)))
    for d, reqs in lut["requests"].items():
        for trs, dsn in reqs:
            parts = ["%s:%d:%s" % (n, t, "".join("so" if l == "source" else "si" for l in locs)) for n, t, locs in trs]
            lines.append("%s %s %s" % (d, dsn, " ".join(parts)))
    return "\n".join(lines) + "\n"


def test_lookup_tables_match_oracle_l4(slaph):
    text = open(os.path.join(GOLD, "l4", "test_all.ini")).read()
    got = slaph.lookup_dump(text)
    want = oracle_dump(o.init_lookup_tables(o.parse_infile(text)))
    assert got == want


def test_dataset_names_match_golden(slaph):
    text = open(os.path.join(GOLD, "l4", "test_all.ini")).read()
    got = slaph.lookup_dump(text)
    names = set()
    for line in got.splitlines():
        parts = line.split(" ")
        if parts[0].startswith("C") and len(parts) > 1 and parts[1].startswith(parts[0] + "_"):
            names.add(parts[0] + "/" + parts[1])
    gold = np.load(os.path.join(GOLD, "l4_correlators.npz"))
    assert names == set(gold.files)


PIPI = """
Lt = 16
Lx = 8
Ly = 8
Lz = 8
number_of_eigen_vec = 12
[quarks]
quark = u:5:TB:4:EI:4:DF:4:light
[operator_lists]
operator_list = g5.d0.p0,1,2,3
[correlator_lists]
correlator_list = C2c:Q0:Op0:Q0:Op0:P0,1,2,3,4
correlator_list = C4cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1,2,3,4
correlator_list = C4cB:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1,2,3,4
correlator_list = C4cC:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0,1,2,3,4
correlator_list = C6cD:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:Q0:Op0:P0
"""


def test_lookup_tables_match_oracle_pipi(slaph):
    got = slaph.lookup_dump(PIPI)
    want = oracle_dump(o.init_lookup_tables(o.parse_infile(PIPI)))
    assert got == want
    counts = {}
    for line in got.splitlines():
        d = line.split(" ")[0]
        if d.startswith("C"):
            counts[d] = counts.get(d, 0) + 1
    assert counts["C2c"] == 27


def test_momentum_cutoff_error_is_reported(slaph):
    bad = PIPI.replace("P0,1,2,3,4", "P5")
    with pytest.raises(slaph.SlaphError):
        slaph.lookup_dump(bad)


def test_not_enough_random_vectors_is_reported(slaph):
    bad = PIPI.replace("u:5:", "u:1:")
    with pytest.raises(slaph.SlaphError, match="not enough random vectors"):
        slaph.lookup_dump(bad)


def _build_class_layer_program(out):
    lib = os.path.join(ROOT, "slaph-contractions_amd")
    src = os.path.join(HERE, "cpp", "class_layer_l4.cpp")
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), src, "-L" + lib,
           "-lslaph_amd", "-lslaph_h5", "-Wl,-rpath," + lib, "-o", out]
    import subprocess
    return subprocess.run(cmd, capture_output=True, text=True)


def test_class_layer_header_compiles_and_links(tmp_path):
    """include/slaph_amd.hpp + slaph_h5.hpp: a caller written against the
    reference's class names compiles without warnings and links against the
    two libraries (run on the GPU by test_gpu_dropin.py)."""
    if not os.path.exists(os.path.join(ROOT, "slaph-contractions_amd", "libslaph_h5.so")):
        pytest.skip("libslaph_h5.so is not built (no HDF5)")
    r = _build_class_layer_program(str(tmp_path / "prog"))
    assert r.returncode == 0, r.stderr
