"""The data-parallel form of the reference's hole-walk PARTITION
(src/treecode3d_pb.f:1677-1759) that the GPU tree build implements
(k_part_classify / k_part_move in tabi-pb_amd/csrc/tree_gpu.hip), checked
against a literal restatement of the sequential hole walk on random inputs
with many ties.

Closed form, for a segment a[b..e] split at val, with "small" meaning <= val
and the first element set aside:
  c  = number of smalls in a[b+1..e];  H = b + c  (final hole);
  misplaced smalls: smalls at positions >= H, ranked from the right s_1 > s_2 > ...
  misplaced bigs:   bigs at positions b+1..H, ranked from the left b_1 < b_2 < ...
  s_1 -> b;  s_(j+1) -> b_j;  b_j -> s_j;  the set-aside element -> H;
  midind = H if the set-aside element <= val else H - 1.
"""
import numpy as np
import pytest


def hole_walk(a, ind, b, e, val):
    """Literal restatement of the sequential PARTITION (0-based, inclusive)."""
    a = list(a)
    ind = list(ind)
    if b < e:
        ta, tind = a[b], ind[b]
        a[b] = val
        upper, lower = b, e
        while upper != lower:
            while upper < lower and val < a[lower]:
                lower -= 1
            if upper != lower:
                a[upper], ind[upper] = a[lower], ind[lower]
            while upper < lower and val >= a[upper]:
                upper += 1
            if upper != lower:
                a[lower], ind[lower] = a[upper], ind[upper]
        mid = upper - 1 if ta > val else upper
        a[upper], ind[upper] = ta, tind
        return a, ind, mid
    if b == e:
        return a, ind, (b if a[b] <= val else b - 1)
    return a, ind, b - 1


def parallel_form(a, ind, b, e, val):
    """The same permutation from prefix counts only (what the kernels compute)."""
    a = np.asarray(a, dtype=np.float64)
    ind = np.asarray(ind)
    n = e - b + 1
    if n <= 1:
        if n == 1:
            return list(a), list(ind), (b if a[b] <= val else b - 1)
        return list(a), list(ind), b - 1
    pos = np.arange(b + 1, e + 1)
    small = a[b + 1:e + 1] <= val
    spre = np.cumsum(small)                      # smalls in [b+1, i]
    c = int(spre[-1])
    H = b + c
    dest = np.arange(len(a))
    # rank of each small from the right, of each big from the left
    s_rank = c - spre + 1                        # valid where small
    b_rank = (pos - b) - spre                    # valid where big
    mis_small = small & (pos >= H)
    mis_big = (~small) & (pos <= H)
    small_at = {int(r): int(p) for r, p in zip(s_rank[mis_small], pos[mis_small])}
    big_at = {int(r): int(p) for r, p in zip(b_rank[mis_big], pos[mis_big])}
    for r, p in small_at.items():
        dest[p] = b if r == 1 else big_at[r - 1]
    for r, p in big_at.items():
        dest[p] = small_at[r]
    dest[b] = H
    out_a = np.empty_like(a)
    out_i = np.empty_like(ind)
    out_a[dest] = a
    out_i[dest] = ind
    mid = H if a[b] <= val else H - 1
    return list(out_a), list(out_i), mid


@pytest.mark.parametrize("seed", range(300))
def test_parallel_form_equals_hole_walk(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 40))
    levels = int(rng.integers(1, 6))
    a = rng.integers(0, levels, size=n + 4).astype(np.float64)  # many ties
    ind = np.arange(n + 4)
    b = int(rng.integers(0, 3))
    e = b + n - 1
    val = float(rng.integers(-1, levels + 1)) - 0.5 * int(rng.integers(0, 2))
    ra, ri, rm = hole_walk(a, ind, b, e, val)
    pa, pi, pm = parallel_form(a, ind, b, e, val)
    assert rm == pm
    assert list(ri) == list(pi)
    assert list(ra) == list(pa)
    # the partition property itself
    assert all(v <= val for v in ra[b:rm + 1]) and all(v > val for v in ra[rm + 1:e + 1])
