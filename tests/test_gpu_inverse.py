"""GPU parity of the batched overlap inverse + determinant (InvertC and
determinantC, Wrappers.f90:195-231, 295-335) through phf_invert_batched, on
every kernel size class (one wave per matrix for n <= 16, 32, 64; LDS for
n <= 96), against numpy.linalg (LAPACK zgetrf/zgetri):

* random complex matrices: inverse and determinant to 1e-10 relative;
* permutation matrices, where every pivot step interchanges rows: the inverse
  is exactly the transpose and det is exactly the permutation's sign;
* a column with a tie in |Re|+|Im| and a matrix that interchanges rows at
  every step;
* singular matrices: info = the first exactly-zero pivot, zero outputs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from phf_pcc_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n", [1, 5, 16, 17, 31, 32, 33, 48, 64, 65, 96])
def test_random_matrices(ctx, n):
    rng = np.random.default_rng(n)
    b = 37
    m = rng.normal(size=(b, n, n)) + 1j * rng.normal(size=(b, n, n))
    print(f"inverse n={n} batch={b}", flush=True)
    inv, det, info = ctx.invert_batched(m)
    assert not info.any()
    want = np.linalg.inv(m)
    for k in range(b):
        assert _rel(inv[k], want[k]) <= 1e-10
    assert np.all(np.abs(det - np.linalg.det(m)) <= 1e-10 * np.abs(np.linalg.det(m)))


@pytest.mark.parametrize("n", [7, 32, 40, 64, 80])
def test_permutation_matrices(ctx, n):
    rng = np.random.default_rng(100 + n)
    perms = [rng.permutation(n) for _ in range(5)] + [np.arange(n)[::-1].copy()]
    m = np.stack([np.eye(n)[p] for p in perms]).astype(complex)
    inv, det, info = ctx.invert_batched(m)
    assert not info.any()
    assert np.array_equal(inv, np.transpose(m, (0, 2, 1)))
    assert np.array_equal(det, np.linalg.det(m.real).round())


@pytest.mark.parametrize("n", [3, 20, 50])
def test_pivot_order(ctx, n):
    """A first column with a tie in |Re|+|Im| (rows 0 and n-1), and an
    anti-diagonal-dominant matrix that interchanges rows at every step of
    its first half: inverse and det against LAPACK."""
    rng = np.random.default_rng(7 + n)
    u = np.triu(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)), 1) + np.eye(n) * (2 + 1j)
    m = u.copy()
    m[n - 1, 0] = complex(0.0, abs(m[0, 0].real) + abs(m[0, 0].imag))
    inv, det, info = ctx.invert_batched(m[None])
    assert info[0] == 0
    assert _rel(inv[0], np.linalg.inv(m)) <= 1e-10
    assert abs(det[0] - np.linalg.det(m)) <= 1e-10 * abs(np.linalg.det(m))
    # anti-diagonal dominant: an interchange at every step of the first half
    a = np.fliplr(np.eye(n)) * 10 + 0.1 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    inv, det, info = ctx.invert_batched(a[None])
    assert _rel(inv[0], np.linalg.inv(a)) <= 1e-10
    assert abs(det[0] - np.linalg.det(a)) <= 1e-10 * abs(np.linalg.det(a))


@pytest.mark.parametrize("n", [4, 24, 40, 70])
def test_singular_info(ctx, n):
    m = np.eye(n, dtype=complex)
    m[:, 2] = 0.0            # column 2 exactly zero: the third pivot is zero
    good = np.eye(n, dtype=complex) * (1 + 1j)
    inv, det, info = ctx.invert_batched(np.stack([good, m, good]))
    assert list(info) == [0, 3, 0]
    assert not inv[1].any() and det[1] == 0
    assert np.allclose(inv[0], np.eye(n) / (1 + 1j), rtol=0, atol=1e-15)
