"""hci_sigma_noalloc (the allocation-free product, whose two work vectors sit
in the library's fixed device array) against hci_sigma on the same device
vectors: FULL storage bitwise equal (its product is deterministic), TILED
within 1e-12 relative (row sums in LDS atomics, order varies), on two calls
in a row with a smaller matrix second, several columns, and every sigma
mode; a matrix above the array's 2^21 determinants is refused."""
import numpy as np
import pytest

from helpers import hubbard_space, mol_space
from risq_amd.hci import (MEM_DEVICE, PREC_F16, SIGMA_F16, SIGMA_F32, SIGMA_FULL, STORE_FULL, STORE_TILED,
                          DeviceBuffer, HciError)

pytestmark = pytest.mark.gpu


def _products(ctx, h, X, mode):
    n, k = X.shape[0], X.shape[1]
    xf = np.asfortranarray(X)
    dx = DeviceBuffer.from_array(ctx, xf.ravel(order="F"))
    dy1 = DeviceBuffer(ctx, 8 * n * k)
    dy2 = DeviceBuffer(ctx, 8 * n * k)
    ctx._check(ctx.lib.hci_sigma(h.handle, k, dx.ptr, dy1.ptr, mode, MEM_DEVICE))
    ctx._check(ctx.lib.hci_sigma_noalloc(h.handle, k, dx.ptr, dy2.ptr, mode))
    a = dy1.to_array(np.float64, n * k).reshape(k, n).T
    b = dy2.to_array(np.float64, n * k).reshape(k, n).T
    for d in (dx, dy1, dy2):
        d.free()
    return a, b


@pytest.mark.parametrize("storage", [STORE_FULL, STORE_TILED])
def test_noalloc_sigma_matches_sigma(gpu_ctx, storage):
    for n_det, ncol in ((20000, 3), (3000, 1)):
        ints, dets = mol_space(16, 32, 0, 4, 4, n_det)
        gpu_ctx.set_integrals(ints)
        diag = gpu_ctx.ham_diag(dets)
        prec = PREC_F16 if storage == STORE_TILED else 32
        h = gpu_ctx.build_h(dets, diag, value_prec=prec, storage=storage)
        X = np.random.default_rng(n_det).standard_normal((dets.shape[0], ncol))
        modes = [SIGMA_FULL, SIGMA_F32] + ([SIGMA_F16] if storage == STORE_TILED else [])
        for mode in modes:
            a, b = _products(gpu_ctx, h, X, mode)
            print(f"n={dets.shape[0]} storage={storage} mode={mode} max|diff|={np.max(np.abs(a - b)):.2e}",
                  flush=True)
            if storage == STORE_FULL:
                assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
            else:
                assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
            # and against the host-vector product
            if mode == SIGMA_FULL:
                c = h.matrix_vector_prod(X)
                assert np.max(np.abs(c - b)) <= 1e-12 * np.max(np.abs(c))
        h.free()


def test_noalloc_sigma_rejects_large_matrix(gpu_ctx):
    ints, dets = hubbard_space(5, 4, 6, 6, (1 << 21) + 1000)
    gpu_ctx.set_integrals(ints)
    diag = gpu_ctx.ham_diag(dets)
    h = gpu_ctx.build_h(dets, diag, storage=STORE_FULL)
    n = dets.shape[0]
    dx = DeviceBuffer(gpu_ctx, 8 * n)
    dy = DeviceBuffer(gpu_ctx, 8 * n)
    with pytest.raises(HciError, match="2097152"):
        gpu_ctx._check(gpu_ctx.lib.hci_sigma_noalloc(h.handle, 1, dx.ptr, dy.ptr, SIGMA_FULL))
    dx.free()
    dy.free()
    h.free()
