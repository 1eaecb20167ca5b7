"""CPU check of the non-uniform queue protocol the persistent kernels follow
(kb_vchol_impl.h): the counting sort (histogram of orders, exclusive scan in
descending order with the count of orders above the split, scatter) yields a
permutation of the batch sorted by order, largest first; the workgroup phase
takes tickets [0, counter[1]) from counter[0], the wavefront phase tickets
[counter[1], batch) from counter[2]; under any interleaving of the workers'
ticket draws every matrix is taken exactly once and every worker's loop ends.
The protocol is restated here in Python (host emulation); the kernels are
exercised by the GPU tests."""
import numpy as np
import pytest

BINS = 4096


def sort_queue(orders, split):
    hist = np.zeros(BINS, np.int64)
    b = np.clip(orders, 0, BINS - 1)
    np.add.at(hist, b, 1)
    offs = np.zeros(BINS, np.int64)
    run = 0
    counter1 = None
    for bin_ in range(BINS - 1, -1, -1):       # descending
        offs[bin_] = run
        if bin_ == split:
            counter1 = run                     # orders above the split
        run += hist[bin_]
    order = np.empty(len(orders), np.int64)
    for i in np.random.default_rng(0).permutation(len(orders)):   # atomics in any order
        pos = offs[b[i]]
        offs[b[i]] += 1
        order[pos] = i
    return order, counter1


def run_phases(orders, split, n_wg, n_wave, seed):
    """interleave the draws of n_wg workgroups (tickets below counter[1]) and
    n_wave wavefronts (tickets from counter[1]); returns who took what"""
    order, nbig = sort_queue(orders, split)
    batch = len(orders)
    rng = np.random.default_rng(seed)
    c0, c2 = 0, nbig
    taken = []
    live_wg, live_wave = set(range(n_wg)), set(range(n_wave))
    steps = 0
    # the workgroup kernel runs first on the stream, then the wavefront kernel
    while live_wg:
        w = rng.choice(sorted(live_wg))
        t = c0
        c0 += 1
        if t >= nbig:
            live_wg.discard(w)
        else:
            taken.append(("wg", int(order[t])))
        steps += 1
        assert steps < 10 * (batch + n_wg + n_wave)
    while live_wave:
        w = rng.choice(sorted(live_wave))
        t = c2
        c2 += 1
        if t >= batch:
            live_wave.discard(w)
        else:
            taken.append(("wave", int(order[t])))
        steps += 1
        assert steps < 10 * (batch + n_wg + n_wave)
    return order, nbig, taken


@pytest.mark.parametrize("split", [0, 16, 48, 64])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_queue_sorted_and_every_matrix_taken_once(split, seed):
    rng = np.random.default_rng(seed)
    orders = rng.integers(0, 300, 777)
    orders[:5] = [0, 1, split, split + 1, 5000]
    order, nbig, taken = run_phases(orders, split, n_wg=13, n_wave=29, seed=seed)
    assert sorted(order.tolist()) == list(range(len(orders)))
    so = np.clip(orders[order], 0, BINS - 1)
    assert np.all(np.diff(so) <= 0)                      # largest first
    assert nbig == int(np.sum(np.clip(orders, 0, BINS - 1) > split))
    mats = [m for _, m in taken]
    assert sorted(mats) == list(range(len(orders)))      # each exactly once
    for who, m in taken:
        assert (orders[m] > split) == (who == "wg"), (who, m, orders[m])


def test_empty_phases_end():
    # no order above the split: the workgroup loop ends on its first ticket
    order, nbig, taken = run_phases(np.full(10, 3), 64, n_wg=4, n_wave=4, seed=0)
    assert nbig == 0 and len(taken) == 10
    # every order above the split: the wavefront loop ends at once
    order, nbig, taken = run_phases(np.full(10, 300), 64, n_wg=4, n_wave=4, seed=0)
    assert nbig == 10 and all(w == "wg" for w, _ in taken)
