#!/usr/bin/env python3
"""Build the committed L4 fixture set from the reference's integration test.

Copies, as data, what the reference's own integration test holds
(tests/integration-L4 in pittlerf/sLapH-contractions, GPL-3.0):

* the inputs: eigenvectors (4 timeslices), perambulators and random vectors
  (2 flavours x 5 random vectors), and the run configuration test_all.ini;
* the expected outputs: the 18 golden HDF5 correlator files, kept as-is (so a
  drop-in run can be checked with `h5diff -r -d 1e-12`, the reference's gate in
  tests/integration-L4/compare:16) and also flattened into
  `l4_correlators.npz` (dataset name -> complex128[Lt]) so tests need no HDF5
  library;
* tests/dilution-iterator/control.txt, the expected enumeration of
  DilutionScheme/BlockIterator.

Run once in a container that has the reference mounted; the outputs are
committed.  Requires h5dump (HDF5 1.10) on PATH or in /opt/conda/bin.
"""
import os
import re
import shutil
import subprocess
import sys

import numpy as np

REF = "/root/reference/tests"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "l4")


def h5dump_exe():
    for cand in ("h5dump", "/opt/conda/bin/h5dump"):
        if shutil.which(cand):
            return shutil.which(cand)
    raise SystemExit("h5dump not found")


_DS = re.compile(r'^\s*DATASET "(.*)" \{')
_NUM = re.compile(r"^\s*(-?[0-9.eE+\-]+|nan|inf|-inf),?\s*$")


def read_h5(path):
    """Parse `h5dump -m %.17g` output of a file of 1-D {re,im} datasets."""
    txt = subprocess.run([h5dump_exe(), "-m", "%.17g", path], check=True,
                         capture_output=True, text=True).stdout
    out = {}
    name = None
    vals = []
    in_data = False
    for line in txt.splitlines():
        m = _DS.match(line)
        if m:
            if name is not None:
                out[name] = vals
            name, vals, in_data = m.group(1), [], False
            continue
        if "DATA {" in line:
            in_data = True
            continue
        if in_data:
            n = _NUM.match(line)
            if n:
                vals.append(float(n.group(1)))
    if name is not None:
        out[name] = vals
    res = {}
    for k, v in out.items():
        a = np.asarray(v, dtype=np.float64)
        assert a.size % 2 == 0, (path, k)
        res[k] = a[0::2] + 1j * a[1::2]
    return res


def main():
    os.makedirs(OUT, exist_ok=True)
    l4 = os.path.join(REF, "integration-L4")
    # inputs
    for sub in ("eigensystem",):
        d = os.path.join(OUT, sub)
        os.makedirs(d, exist_ok=True)
        for f in sorted(os.listdir(os.path.join(l4, sub))):
            if f.startswith("eigenvectors"):
                shutil.copyfile(os.path.join(l4, sub, f), os.path.join(d, f))
    for flav in ("light", "strange"):
        for r in range(5):
            src = os.path.join(l4, flav, "cnfg1000", "rnd_vec_%02d" % r)
            dst = os.path.join(OUT, flav, "cnfg1000", "rnd_vec_%02d" % r)
            os.makedirs(dst, exist_ok=True)
            for f in os.listdir(src):
                if f.startswith("perambulator") or f.startswith("randomvector"):
                    shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
    shutil.copyfile(os.path.join(l4, "test_all.ini"), os.path.join(OUT, "test_all.ini"))
    # expected outputs
    gold = os.path.join(OUT, "correlators-reference")
    os.makedirs(gold, exist_ok=True)
    flat = {}
    for f in sorted(os.listdir(os.path.join(l4, "correlators-reference"))):
        if not f.endswith(".h5"):
            continue
        p = os.path.join(l4, "correlators-reference", f)
        shutil.copyfile(p, os.path.join(gold, f))
        diagram = f.split("_cnfg")[0]
        for k, v in read_h5(p).items():
            flat[diagram + "/" + k] = v
    np.savez_compressed(os.path.join(HERE, "l4_correlators.npz"), **flat)
    shutil.copyfile(os.path.join(REF, "dilution-iterator", "control.txt"),
                    os.path.join(HERE, "dilution_iterator_control.txt"))
    print("datasets:", len(flat))


if __name__ == "__main__":
    sys.exit(main())
