"""cuOptAmdSpmvNoAlloc, the SpMV entry point that allocates no device memory,
against cuOptAmdSpmv (same kernels and plan): bitwise equal results on two calls
in a row through the C ABI, the second with a smaller input, in both precisions.
The refusal of an input larger than the fixed array needs no GPU."""

import numpy as np
import pytest

from cuopt_amd import c_api
from cuopt_amd import generators as G


def _long_row_case():
    # rows longer than a tile (chunks and the merge launch), empty rows, a row of one
    rng = np.random.default_rng(5)
    lens = np.array([0, 3000, 1, 0, 2049, 7, 0, 1024, 5000, 2])
    n = 9000
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in lens if k]).astype(np.int32)
    return ptr, idx, rng.standard_normal(idx.size), n


def test_refuses_input_beyond_the_fixed_array():
    size = c_api.spmv_scratch_bytes()
    assert size == 64 << 20
    n = size // 8 + 1  # x alone exceeds it in FP64
    ptr = np.array([0, 1], np.int32)
    idx = np.array([0], np.int32)
    with pytest.raises(c_api.CuOptError) as e:
        c_api.spmv_noalloc(ptr, idx, np.ones(1), np.zeros(n), n, "fp64")
    assert e.value.code == c_api.CUOPT_INVALID_ARGUMENT


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_no_alloc_spmv_matches_spmv_on_two_calls(precision):
    big = G.skewed_lp(m=20000, n=40000, seed=2)
    ptr, idx, val, n = _long_row_case()
    small = G.random_lp(m=300, n=700, seed=3)
    rng = np.random.default_rng(0)
    cases = [("skewed 20000 x 40000", big.row_offsets, big.col_indices, big.values, big.n),
             ("long rows 10 x 9000", ptr, idx, val, n),
             ("random 300 x 700", small.row_offsets, small.col_indices, small.values, small.n)]
    for name, p, i, v, cols in cases:  # each call smaller than the one before
        print(f"no-alloc spmv {precision} {name}", flush=True)
        x = rng.standard_normal(cols)
        ref = c_api.spmv(p, i, v, x, cols, precision)
        got = c_api.spmv_noalloc(p, i, v, x, cols, precision)
        assert np.array_equal(got, ref), name
