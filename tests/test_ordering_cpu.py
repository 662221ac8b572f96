"""CPU checks that the pair kernel's results cannot depend on the order of its work items.

Two parts of the device code decide this, restated here in numpy:

* the fixed-point sums (``fx_split``/``fx_value`` in csrc/nbabfs_internal.h): every force and energy
  contribution becomes two int64 words, hi = rint(v / 2^16) and lo = rint((v - hi 2^16) 2^32), added with
  integer atomics.  Integer addition is associative, so any order of the items gives the same words, and the
  words convert to the same double.
* the longest-first item order (``k_item_order`` in csrc/nbabfs_build.hip): a histogram of the costs kept in
  four copies (wave w adds into copy w % 4), an exclusive scan over (bin, copy) and a scatter.  It must be a
  permutation of the items with non-increasing cost, whatever the assignment of items to waves.

The plain FP64 path does not read work items at all (one thread per atom over sorted CSR rows), so its results
are independent of the item order by construction.
"""

import numpy as np

HI_UNIT = 2.0 ** 16
LO_SCALE = 2.0 ** 32


def fx_split(v):
    v = np.clip(v, -2.0 ** 66, 2.0 ** 66)
    hi = np.rint(v / HI_UNIT)
    r = v - hi * HI_UNIT                     # exact (Sterbenz: |r| <= 2^15)
    lo = np.rint(r * LO_SCALE)
    return hi.astype(np.int64), lo.astype(np.int64)


def fx_value(hi, lo):
    return float(hi) * HI_UNIT + float(lo) / LO_SCALE


def test_fixed_point_sums_do_not_depend_on_item_order():
    rng = np.random.default_rng(3)
    # contributions of many items to one gradient component: ordinary forces and a few overlaps
    v = rng.normal(0.0, 50.0, 20000) * np.where(rng.random(20000) < 0.01, 1e9, 1.0)
    hi, lo = fx_split(v)
    ref = (int(hi.sum()), int(lo.sum()))
    for seed in range(5):
        p = np.random.default_rng(seed).permutation(v.size)
        h2, l2 = fx_split(v[p])
        assert (int(h2.sum()), int(l2.sum())) == ref
    # each contribution is represented to half a unit of 2^-32
    back = hi.astype(np.float64) * HI_UNIT + lo.astype(np.float64) / LO_SCALE
    small = np.abs(v) < 2.0 ** 20
    assert np.max(np.abs(back[small] - v[small])) <= 2.0 ** -32


def item_order(cost, maxcost, threads=1024, copies=4, unroll=8):
    """k_item_order's scatter for one assignment of items to threads (thread t takes items t, t + threads, ...)."""
    n = cost.size
    bins = maxcost + 1
    hist = np.zeros((bins, copies), np.int64)
    owner = (np.arange(n) % threads) // 64 % copies          # copy of the wave that holds item k
    np.add.at(hist, (maxcost - cost, owner), 1)
    flat = hist.reshape(-1)
    start = np.concatenate([[0], np.cumsum(flat)[:-1]]).reshape(bins, copies)
    out = np.full(n, -1, np.int64)
    cursor = start.copy()
    # any interleaving of the atomics inside one (bin, copy) only permutes equal-cost items
    order = np.random.default_rng(1).permutation(n)
    for k in order:
        b, c = maxcost - cost[k], owner[k]
        out[cursor[b, c]] = k
        cursor[b, c] += 1
    return out


def test_item_order_is_a_longest_first_permutation():
    rng = np.random.default_rng(7)
    for n, per_item in [(13855, 48), (79436, 16), (1000, 12), (5, 8)]:
        maxcost = 8 * per_item
        cost = np.minimum(rng.binomial(maxcost, 0.6, n), maxcost)
        out = item_order(cost, maxcost)
        assert np.array_equal(np.sort(out), np.arange(n))
        assert np.all(np.diff(cost[out]) <= 0)
