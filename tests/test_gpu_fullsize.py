"""Full-size consistency of the trace kernels on the 32^3 x 64 workload
(N_ev = 220, 5 random vectors, TB16/EI4/DF4, C2c / C4cD / C4cB / C4cC at
|P|^2 <= 4 and C6cD at P = 0): the correlators of the production kernels
(queued left-stationary traces, slice-pair tiles grouped by right operand)
against those of the block-form and item-wise kernels on the same
device-generated inputs.  Both sets sum the same traces, in different orders,
so every correlator must agree to rounding (1e-12 relative); the oracle cannot
run at this size."""
import sys

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

SEED = 20261203


@pytest.fixture(scope="module")
def slaph():
    import slaph as s
    return s


def _correlators(slaph, monkeypatch, alternative):
    for var in ("SLAPH_LSTAT_BLOCKS", "SLAPH_GRID_ITEMS"):
        if alternative:
            monkeypatch.setenv(var, "1")
        else:
            monkeypatch.delenv(var, raising=False)
    print("pipi32", "block-form / item-wise traces" if alternative else "production traces", flush=True)
    sys.stdout.flush()
    ctx = slaph.Context(0)
    try:
        ctx.configure(bench.infile(bench.WORKLOADS["pipi32"]))
        ctx.synthetic_inputs(SEED, True)
        slaph.OperatorFactory(ctx).create_operators()
        slaph.contract(ctx)
        return {d: ctx.correlators(d) for d in ctx.diagrams()}
    finally:
        ctx.close()


def test_pipi32_trace_kernels_agree(slaph, monkeypatch):
    prod = _correlators(slaph, monkeypatch, False)
    alt = _correlators(slaph, monkeypatch, True)
    assert prod.keys() == alt.keys()
    worst, n = 0.0, 0
    for d, vals in prod.items():
        assert vals.keys() == alt[d].keys(), d
        for name, v in vals.items():
            ref = alt[d][name]
            scale = max(np.max(np.abs(ref)), 1e-300)
            err = np.max(np.abs(v - ref)) / scale
            assert err <= 1e-12, (d, name, err)
            worst = max(worst, err)
            n += 1
    assert n == 58489
    print("pipi32: %d correlators, largest relative difference %.2e" % (n, worst), flush=True)
