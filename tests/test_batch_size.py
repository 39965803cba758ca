"""Batch-size search (optimal_batch_size_handler.cu:196-435): the oracle pinned by
hand-worked closed-form cases, the library's search against the oracle."""

import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import batch_size as O  # noqa: E402


def _lib_search():
    from cuopt_amd import c_api

    try:
        c_api.lib()
    except c_api.LibraryNotBuilt:
        pytest.skip("library not built")
    return c_api.batch_size_search


# (cost, max_batch, expected) worked by hand from the steps of the search.
CASES = [
    (lambda b: 1.0 / b, 10000, 2048),   # falls forever: 256, 512, 1024, 2048 (four steps), middle 1536 is worse
    (lambda b: 1.0 / b, 300, 300),      # doubling from 256 hits the cap, which is better
    (lambda b: 1.0 / b, 256, 256),      # the double of 128 is the cap
    (lambda b: float(b), 10000, 8),     # rises forever: 64, 32, 16, 8 (four steps), middle 12 is worse
    (lambda b: float(b), 3, 1),         # start 2, its half is 1 and better
    (lambda b: abs(math.log2(b) - 5) + 1, 10000, 32),   # V at 32: 64, 32, then 16 is worse; middle 24 is worse
    (lambda b: abs(math.log2(b) - 9) + 1, 10000, 512),  # V at 512: 256, 512, then 1024 is worse; middle 768 worse
    (lambda b: abs(math.log2(b) - 7) + 1, 10000, 128),  # V at the start
    (lambda b: 1.0, 10000, 128),        # flat: neither neighbour is strictly better
    (lambda b: 5.0, 1, 1),              # a batch of one
]


@pytest.mark.parametrize("cost,max_batch,expected", CASES)
def test_oracle_closed_forms(cost, max_batch, expected):
    assert O.search(max_batch, cost) == expected


def test_oracle_middle_node_wins():
    # V at 48 on a fine scale: 64 beats 128, 32 is worse than 64, and the middle
    # (2*32 + 32) // 2 = 48 beats 64.
    cost = lambda b: abs(b - 48) + 1.0  # noqa: E731
    assert O.search(10000, cost) == 48


@pytest.mark.parametrize("cost,max_batch,expected", CASES)
def test_library_closed_forms(cost, max_batch, expected):
    assert _lib_search()(max_batch, cost) == expected


def test_library_matches_oracle_on_random_costs():
    search = _lib_search()
    rng = np.random.default_rng(0)
    for trial in range(300):
        table = {}
        max_batch = int(rng.choice([1, 2, 3, 7, 100, 128, 129, 255, 256, 300, 1000, 5000, 100000]))
        shape = trial % 3

        def cost(b, table=table, shape=shape):
            if b not in table:
                if shape == 0:
                    table[b] = float(rng.random())
                elif shape == 1:
                    table[b] = abs(math.log2(b) - rng.uniform(0, 12)) + rng.random() * 0.1
                else:
                    table[b] = 1.0 / b + rng.random() * 1e-3
            return table[b]

        got = search(max_batch, cost)
        assert got == O.search(max_batch, cost), (trial, max_batch)
        assert 1 <= got <= max_batch


def test_library_rejects_bad_arguments():
    search = _lib_search()
    from cuopt_amd import c_api

    with pytest.raises(c_api.CuOptError):
        search(0, lambda b: 1.0)
