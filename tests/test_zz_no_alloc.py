"""phf_invert_batched_noalloc (the entry point that allocates no device memory)
against phf_invert_batched, through the C ABI: bit-identical inverses,
determinants and info on two calls in a row, the second with a smaller input
(smaller n and batch, a singular matrix and a permutation among them), and a
call above the 64 MiB staging limit refused before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mats(rng, n, b):
    return rng.normal(size=(b, n, n)) + 1j * rng.normal(size=(b, n, n))


def test_noalloc_inverse_equals_allocating_entry_point():
    from phf_pcc_amd import _lib
    ctx = _lib.Context(0)
    try:
        rng = np.random.default_rng(4242)
        first = _mats(rng, 48, 300)
        second = _mats(rng, 20, 7)
        second[2] = 0.0
        second[2][:, :19] = rng.normal(size=(20, 19))          # exactly singular: last column zero
        second[4] = np.eye(20)[rng.permutation(20)]           # every pivot step swaps
        for m in (first, second):
            inv0, det0, info0 = ctx.invert_batched(m)
            inv1, det1, info1 = ctx.invert_batched_noalloc(m)
            assert np.array_equal(info1, info0)
            assert np.array_equal(det1, det0)
            assert np.array_equal(inv1, inv0)
        assert info1[2] != 0 and det1[2] == 0 and not inv1[2].any()
        assert np.array_equal(inv1[4], second[4].T)
        ok = info1 == 0
        assert np.abs(inv1[ok] @ second[ok] - np.eye(20)).max() < 1e-10
        big = np.zeros((2000, 96, 96), np.complex128)           # 2 * 16 * 96^2 * 2000 bytes > 64 MiB
        with pytest.raises(_lib.PhfError) as err:
            ctx.invert_batched_noalloc(big)
        assert err.value.code == _lib.PHF_ERR_UNSUPPORTED
    finally:
        ctx.close()
