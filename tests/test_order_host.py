"""Hand-out orders built on the host side of the boundary (no GPU needed):
runtime.density_order and runtime.cost_order return permutations of the
cells, densest or costliest first, ties in index order."""
import numpy as np
import torch

from dengo_amd import runtime


def test_density_order_sorts_by_species_total():
    rng = np.random.default_rng(11)
    x = rng.lognormal(size=(500, 11))
    x[:, 10] = rng.uniform(1e10, 1e14, 500)      # the energy column: large, excluded
    got = runtime.density_order(torch.from_numpy(x), 10).numpy()
    assert got.dtype == np.int64 and np.array_equal(np.sort(got), np.arange(500))
    tot = x[:, :10].sum(axis=1)
    assert np.all(np.diff(tot[got]) <= 0)
    # with every column counted the energy decides the order
    allc = runtime.density_order(torch.from_numpy(x)).numpy()
    assert np.array_equal(allc, np.argsort(-x.sum(axis=1), kind="stable"))


def test_orders_keep_index_order_on_ties():
    x = torch.ones(7, 3, dtype=torch.float64)
    x[4] = 2.0
    assert runtime.density_order(x).tolist() == [4, 0, 1, 2, 3, 5, 6]
    sw = torch.tensor([3, 9, 3, 1, 9, 0], dtype=torch.int32)
    assert runtime.cost_order(sw).tolist() == [1, 4, 0, 2, 3, 5]


def test_density_order_puts_nan_totals_last():
    """A NaN total counts as -inf, so it sorts last (tied with a total of
    -inf), ties in index order: the patch server's rule for its dense
    patches."""
    x = torch.tensor([[1.0], [float("nan")], [3.0], [-float("inf")], [float("nan")], [3.0]],
                     dtype=torch.float64)
    assert runtime.density_order(x).tolist() == [2, 5, 0, 1, 3, 4]
