"""Matrix-core dense products (the solver's dense base-grid, paddle-grid and
mixed-parity transforms) against numpy in float64.

The kernels' MFMA lane layouts, the K-chunk permutation, the masked edges of
ragged shapes and the batched strides all show up here; the solver's parity
tests then cover them inside the right-hand side."""
import numpy as np
import pytest

from hosnwt import lib

SHAPES = [
    # m, n, k, batch
    (1, 1, 1, 1),
    (16, 16, 4, 1),
    (17, 5, 3, 1),
    (9, 7, 8, 1),          # k a multiple of 4 below one chunk
    (33, 1024, 33, 1),     # paddle grid z transforms (n3a x 4 n2)
    (64, 512, 64, 1),
    (65, 256, 256, 2),     # paddle grid y transforms, batch of two
    (1024, 256, 1024, 1),  # base grid x transform of 1024 x 256
    (1024, 256, 256, 1),
    (100, 37, 61, 3),
    (100, 36, 60, 3),
    (511, 2, 515, 1),      # mixed-parity map of a 2-D tank
]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("a_kmajor,b_kmajor", [(True, True), (False, True), (True, False), (False, False)])
@pytest.mark.parametrize("tile", [0, 16, 32])
def test_dense_product_matches_numpy(precision, a_kmajor, b_kmajor, tile):
    rng = np.random.default_rng(7 + 3 * precision + 2 * int(a_kmajor) + int(b_kmajor) + tile)
    for (m, n, k, batch) in SHAPES:
        print(f"dense product m={m} n={n} k={k} batch={batch} a_kmajor={a_kmajor} b_kmajor={b_kmajor} "
              f"tile={tile} precision={precision}", flush=True)
        a = rng.standard_normal((batch, m, k))
        b = rng.standard_normal((k, n))
        got = lib.dense_product(a, b, a_kmajor=a_kmajor, b_kmajor=b_kmajor, precision=precision, tile=tile)
        bo = b
        if precision == 1:
            a32 = a.astype(np.float32).astype(np.float64)
            bo = bo.astype(np.float32).astype(np.float64)
            want = a32 @ bo
        else:
            want = a @ bo
        # per-entry bound: rounding of a k-term sum of products
        scale = np.abs(a if precision == 0 else a32) @ np.abs(bo)
        eps = 2.0**-52 if precision == 0 else 2.0**-23
        bound = (k + 4) * eps * scale + 1e-300
        assert np.all(np.abs(got - want) <= bound), (m, n, k, batch, float(np.max(np.abs(got - want) / bound)))


@pytest.mark.gpu
def test_dense_product_batched_operands():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((4, 40, 24))
    b = rng.standard_normal((4, 24, 9))
    for ak in (False, True):
        for bk in (False, True):
            got = lib.dense_product(a, b, a_kmajor=ak, b_kmajor=bk)
            np.testing.assert_allclose(got, a @ b, rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_dense_product_is_deterministic():
    rng = np.random.default_rng(11)
    a = rng.standard_normal((300, 777))
    b = rng.standard_normal((777, 130))
    first = lib.dense_product(a, b)
    for _ in range(3):
        assert np.array_equal(lib.dense_product(a, b), first)
