"""Gauge configurations from lime / ILDG files (GaugeField::read_gauge_field,
GaugeField.cpp:762-928, map_timeslice_to_eigen :111-149) and the spatial
plaquette (plaque_ts, :722-760), on the CPU.

The oracle's reader is pinned by the reference's own gauge file
(tests/integration-L4/gauge/conf.1000, copied to tests/golden/l4/gauge):
its SciDAC checksum record and the plaquette of its xlf-info record, which
fail for a wrong byte order, site order, direction order or matrix
orientation (planted below).  The library's host reader
(slc_read_lime_gauge, no device needed) must then give the oracle's links
bit for bit, on the fixture and on synthetic files of non-cubic shape, and
raise the reference's errors (and its own checksum / format errors)."""
import os
import re
import struct
import sys

import numpy as np
import pytest

from oracle import slaph_oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
CONF = os.path.join(HERE, "golden", "l4", "gauge", "conf.1000")
L4 = (4, 4, 4)
SITE_BYTES = 72 * 8


@pytest.fixture(scope="module")
def records():
    return o.lime_records(open(CONF, "rb").read())


@pytest.fixture(scope="module")
def slaph():
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "slaph-contractions_amd"))
    import slaph as s
    return s


def test_fixture_checksum_matches_its_record(records):
    rec = records["scidac-checksum"].decode()
    want = (int(re.search(r"<suma>(\w+)</suma>", rec).group(1), 16),
            int(re.search(r"<sumb>(\w+)</sumb>", rec).group(1), 16))
    assert o.scidac_checksum(records["ildg-binary-data"], SITE_BYTES) == want
    assert want == (0xEF78DFB3, 0xBA50A396)


def test_fixture_plaquette_matches_xlf_info(records):
    want = float(re.search(r"plaquette\s*=\s*([0-9.eE+-]+)", records["xlf-info"].decode()).group(1))
    full, _ = o.ildg_links(records["ildg-binary-data"], L4, 4)
    assert abs(o.plaquette_4d(full) - want) < 1e-12
    # the links are SU(3)
    u = full.reshape(-1, 3, 3)
    assert np.max(np.abs(u @ u.conj().swapaxes(1, 2) - np.eye(3))) < 1e-12
    assert np.max(np.abs(np.linalg.det(u) - 1)) < 1e-12


def test_planted_reading_mistakes_change_the_plaquette(records):
    """A transposed matrix, a little-endian read, or x and z swapped for the
    sites but not for the directions: each misses the recorded plaquette."""
    full, _ = o.ildg_links(records["ildg-binary-data"], L4, 4)
    want = 0.604379109408
    assert abs(o.plaquette_4d(np.swapaxes(full, -1, -2)) - want) > 1e-3
    assert abs(o.plaquette_4d(np.swapaxes(full, 1, 3)) - want) > 1e-3
    le = np.frombuffer(records["ildg-binary-data"], dtype="<f8").reshape(4, 4, 4, 4, 4, 3, 3, 2)
    with np.errstate(all="ignore"):
        bad = o.plaquette_4d(le[..., 0] + 1j * le[..., 1])
    assert not abs(bad - want) < 1e-3


def test_spatial_plaquette_per_timeslice(records):
    """plaque_ts, averaged over the timeslices, is the 4D plaquette over the
    three spatial planes (two independent neighbour constructions); it is
    gauge invariant and 1 on unit links."""
    full, U = o.ildg_links(records["ildg-binary-data"], L4, 4)
    per_t = [o.plaquette_ts(U[t], L4) for t in range(4)]
    assert abs(np.mean(per_t) - o.plaquette_4d(full, planes=[(0, 1), (0, 2), (1, 2)])) < 1e-14
    rng = np.random.default_rng(5)
    g = np.linalg.qr(rng.standard_normal((64, 3, 3)) + 1j * rng.standard_normal((64, 3, 3)))[0]
    Ug = np.empty_like(U[0])
    for mu in range(3):
        nb = o.neighbour(L4, mu, 1)
        Ug[:, mu] = g @ U[0][:, mu] @ g[nb].conj().swapaxes(1, 2)
    assert abs(o.plaquette_ts(Ug, L4) - per_t[0]) < 1e-13
    ones = np.broadcast_to(np.eye(3), (64, 3, 3, 3)).astype(complex)
    assert abs(o.plaquette_ts(ones, L4) - 1.0) < 1e-15


def test_library_reader_matches_oracle(slaph):
    want, _ = o.read_lime_gauge(CONF, L4, 4)
    got = slaph.read_lime_gauge(CONF, L4, 4)
    assert got.shape == want.shape and np.array_equal(got, want)
    part = slaph.read_lime_gauge(CONF, L4, 4, slice_i=1, slice_f=2)
    assert np.array_equal(part, want[1:3])


def _lime(records):
    """A lime file from (type, payload) pairs."""
    out = b""
    for i, (typ, data) in enumerate(records):
        flags = (1 << 15 if i == 0 else 0) | (1 << 14 if i == len(records) - 1 else 0)
        out += struct.pack(">IHHQ", 0x456789AB, 1, flags, len(data)) + typ.encode().ljust(128, b"\0")
        out += data + b"\0" * ((-len(data)) % 8)
    return out


def _fmt(L, Lt, prec=64):
    return ("<?xml version=\"1.0\"?><ildgFormat><field>su3gauge</field><precision>%d</precision>"
            "<lx>%d</lx><ly>%d</ly><lz>%d</lz><lt>%d</lt></ildgFormat>" % (prec, L[0], L[1], L[2], Lt)).encode()


def _cks(binary):
    a, b = o.scidac_checksum(binary, SITE_BYTES)
    return ("<scidacChecksum><version>1.0</version><suma>%x</suma><sumb>%x</sumb></scidacChecksum>" % (a, b)).encode()


def _random_binary(L, Lt, seed):
    rng = np.random.default_rng(seed)
    n = L[0] * L[1] * L[2] * Lt * 72
    return rng.standard_normal(n).astype(">f8").tobytes()


@pytest.mark.parametrize("L,Lt", [((2, 3, 5), 3), ((6, 4, 2), 2)])
def test_library_reader_non_cubic_lattices(slaph, tmp_path, L, Lt):
    """On Lx != Ly != Lz the file's U_mu(x, y, z, t) lands at site
    (x Ly + y) Lz + z, direction mu (the reference's reader needs
    Lx = Ly = Lz); data need not be SU(3) for the byte layout."""
    b = _random_binary(L, Lt, 7)
    path = tmp_path / "conf.0001"
    path.write_bytes(_lime([("ildg-format", _fmt(L, Lt)), ("ildg-binary-data", b), ("scidac-checksum", _cks(b))]))
    want = o.ildg_links(b, L, Lt)[1]
    got = slaph.read_lime_gauge(str(path), L, Lt)
    assert np.array_equal(got, want)
    # no format or checksum record: read as is
    path.write_bytes(_lime([("ildg-binary-data", b)]))
    assert np.array_equal(slaph.read_lime_gauge(str(path), L, Lt, 1, Lt - 1), want[1:])


def test_library_reader_errors(slaph, tmp_path):
    L, Lt = (2, 2, 2), 2
    b = _random_binary(L, Lt, 8)
    p = tmp_path / "conf"

    def err(data, L_=L, Lt_=Lt, si=0, sf=None):
        p.write_bytes(data)
        with pytest.raises(slaph.SlaphError) as e:
            slaph.read_lime_gauge(str(p), L_, Lt_, si, Lt_ - 1 if sf is None else sf)
        return str(e.value)

    good = _lime([("ildg-format", _fmt(L, Lt)), ("ildg-binary-data", b), ("scidac-checksum", _cks(b))])
    p.write_bytes(good)
    slaph.read_lime_gauge(str(p), L, Lt)
    bad = bytearray(b)
    bad[100] ^= 1
    assert "checksum mismatch" in err(_lime([("ildg-binary-data", bytes(bad)), ("scidac-checksum", _cks(b))]))
    assert err(_lime([("ildg-binary-data", b[:len(b) // 2])])) == "Single precision read!"
    assert err(_lime([("ildg-format", _fmt(L, Lt, 32)), ("ildg-binary-data", b)])) == "Single precision read!"
    assert "Probably wrong lattice size" in err(_lime([("ildg-binary-data", b[:len(b) // 3 // 8 * 8])]))
    assert "no ildg-binary-data record" in err(_lime([("xlf-info", b"plaquette = 1")]))
    assert "differs from Lx Ly Lz Lt" in err(_lime([("ildg-format", _fmt((2, 2, 4), 1)), ("ildg-binary-data", b)]))
    assert "not a lime file" in err(b"\0" * 200)
    assert "timeslice range" in err(good, si=1, sf=0)
    with pytest.raises(slaph.SlaphError, match="Could not open file"):
        slaph.read_lime_gauge(str(tmp_path / "conf.9999"), L, Lt)
