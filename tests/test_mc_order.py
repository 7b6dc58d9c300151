"""The tests' own ordering helpers: the vectorised red-black order and
permutation used at full grid sizes agree with the greedy colouring loop and
the row-by-row permutation on small grids."""
import numpy as np
import pytest

from mc_order import greedy_order, parity_order, permute, permute_fast
from oracle import ppo


@pytest.mark.parametrize("dims", [(7, 6, 1), (5, 4, 3), (6, 6, 6), (1, 9, 1)])
def test_parity_order_is_the_greedy_order(dims):
    A = ppo.laplacian(*dims)
    order, ncol = greedy_order(A)
    assert ncol == (2 if len(A[0]) > 2 else 1)
    np.testing.assert_array_equal(order, parity_order(*dims))
    U = ppo.upwind(*dims, 20.0, 10.0, 5.0)  # same graph, other values
    np.testing.assert_array_equal(greedy_order(U)[0], parity_order(*dims))


def test_permute_fast_matches_permute():
    A = ppo.upwind(5, 4, 3, 20.0, 10.0, 5.0)
    rng = np.random.default_rng(1)
    for order in (parity_order(5, 4, 3), rng.permutation(60).astype(np.uint32)):
        a, b = permute(A, order), permute_fast(A, order)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
