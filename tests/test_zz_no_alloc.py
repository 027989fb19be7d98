"""The allocation-free A^H diag(phase) B entry (slc_ahb_device_noalloc)
against the allocating one (slc_ahb_device) on the same device buffers: two
calls in a row, the second smaller, must give bit-identical products (same
kernel, same job descriptors); a batch beyond the fixed job table is rejected."""
import sys

import numpy as np
import pytest

from devmem import DeviceArray

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def slaph():
    import slaph as s
    return s


def _product(ctx, M, N, K, batch, seed, noalloc):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((batch, M, K)) + 1j * rng.standard_normal((batch, M, K))  # K x M column-major, lda = K
    B = rng.standard_normal((batch, N, K)) + 1j * rng.standard_normal((batch, N, K))  # K x N column-major, ldb = K
    ph = np.exp(-1j * rng.uniform(0, 2 * np.pi, (K + 2) // 3))
    dA, dB, dP = DeviceArray(A), DeviceArray(B), DeviceArray(ph)
    dC = DeviceArray(np.zeros((batch, N, M), dtype=np.complex128))  # M x N column-major, ldc = M
    f = ctx.ahb_device_noalloc if noalloc else ctx.ahb_device
    f(dA.ptr, dB.ptr, dC.ptr, dP.ptr, M, N, K, K, K, M, batch, M * K, N * K, N * M)
    out = dC.get()
    want = np.stack([A[i].conj() @ (np.repeat(ph, 3)[:K, None] * B[i].T) for i in range(batch)])  # M x N
    return out, want


def test_noalloc_product_equals_allocating_entry_two_calls(slaph):
    """Two calls in a row through the C ABI, the second with a smaller batch
    and smaller matrices: each bit-equal to the allocating entry's product."""
    ctx = slaph.Context(0)
    for M, N, K, batch in ((70, 45, 301, 9), (20, 17, 50, 3)):
        say("ahb noalloc", M, N, K, batch)
        ref, want = _product(ctx, M, N, K, batch, 5 + M, False)
        got, _ = _product(ctx, M, N, K, batch, 5 + M, True)
        assert np.array_equal(got, ref)
        # C is (batch, N, M) row-major = M x N column-major per product
        assert np.max(np.abs(got.transpose(0, 2, 1) - want)) <= 1e-12 * np.max(np.abs(want))


def test_noalloc_rejects_batch_beyond_table(slaph):
    ctx = slaph.Context(0)
    d = DeviceArray(np.zeros(4, dtype=np.complex128))
    with pytest.raises(slaph.SlaphError, match="between 1 and 1024"):
        ctx.ahb_device_noalloc(d.ptr, d.ptr, d.ptr, None, 1, 1, 1, 1, 1, 1, 1025, 0, 0, 0)
