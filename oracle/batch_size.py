"""CPU restatement of the reference's batch-size search (test infrastructure only).

Checks cuOptAmdBatchSizeSearch / cuOptAmdUnrolledOptimalBatchSize; nothing in the
product imports this module.

optimal_batch_size_handler.cu:196-435 (called from solve.cu:788): the cost of a
batch size b is the mean time of the batch's SpMMs (A and A^T) over five runs,
divided by b.  The search starts at the largest power of two <= min(128,
max_batch) (:212-218), evaluates it and its half and double (:250-262), and
  - if the half is cheaper (:274-346): unless it is 1, keeps halving while the
    cost falls, returning 1 as soon as 1 is reached and better, for at most
    three more sizes (`current_step < max_steps`, :320); then evaluates
    (2 b + b) / 2 for the last size b tried and returns it if it beats the best
    found (:322-345);
  - else if the double is cheaper (:348-423): unless it is max_batch, keeps
    doubling (capped at max_batch) the same way, returning max_batch as soon as
    it is reached and better (:387); then evaluates min((b / 2 + b) / 2,
    max_batch) (:401-423);
  - else returns the start (:425-431).

Pinned in tests/test_batch_size.py by closed-form cases worked by hand from the
steps above (a strictly decreasing cost, a strictly increasing cost, V-shaped
costs with their minimum at a power of two, max_batch = 1 and caps).
"""

from __future__ import annotations


def search(max_batch: int, cost) -> int:
    if max_batch <= 1:
        return 1
    start = 1
    while start * 2 <= min(128, max_batch):
        start *= 2
    lo, hi = max(1, start // 2), min(2 * start, max_batch)
    c_start, c_lo, c_hi = cost(start), cost(lo), cost(hi)
    steps = 1
    if c_lo < c_start:
        if lo == 1:
            return 1
        best, b, c_best = lo, lo, c_lo
        while True:
            b = max(1, b // 2)
            c = cost(b)
            if not c < c_best:
                break
            if b == 1:
                return 1
            steps += 1
            best, c_best = b, c
            if steps >= 4:
                break
        mid = max(1, (2 * b + b) // 2)
        return mid if cost(mid) < c_best else best
    if c_hi < c_start:
        if hi == max_batch:
            return hi
        best, b, c_best = hi, hi, c_hi
        while True:
            b = min(2 * b, max_batch)
            c = cost(b)
            if not c < c_best:
                break
            if b == max_batch:
                return b
            steps += 1
            best, c_best = b, c
            if steps >= 4:
                break
        mid = min(max(1, (b // 2 + b) // 2), max_batch)
        return mid if cost(mid) < c_best else best
    return start
