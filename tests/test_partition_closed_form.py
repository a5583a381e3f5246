"""The closed form of the reference's swap partition (bvh_builder.cpp:123-137)
that the device builder evaluates in parallel, checked against the sequential
loop on every left/right pattern up to 12 elements and on random ones."""
import itertools

import numpy as np


def swap_partition(a, left):
    """The reference's loop: i from the front, j from the back; an element of
    the right side is swapped with a[j] and j decremented."""
    a = list(a)
    i, j = 0, len(a) - 1
    while i <= j:
        if left[a[i]]:
            i += 1
        else:
            a[i], a[j] = a[j], a[i]
            j -= 1
    return a, i


def closed_form(a, left):
    n = len(a)
    c = [bool(left[x]) for x in a]
    L = sum(c)
    fr = [p for p in range(L) if not c[p]]                 # right elements in [0, L), ascending
    bl = [q for q in range(n - 1, L - 1, -1) if c[q]]       # left elements in [L, n), from the back
    rf = len(fr)
    assert rf == len(bl)
    out = [None] * n
    for p in range(n):
        if p < L and c[p]:
            d = p
        elif c[p]:
            d = fr[bl.index(p)]
        elif p < L:
            m = fr.index(p) + 1
            d = n - 1 if m == 1 else bl[m - 2] - 1
        else:
            d = p - 1 if p > L else (n - 1 if rf == 0 else bl[rf - 1] - 1)
        assert out[d] is None
        out[d] = a[p]
    return out, L


def test_closed_form_exhaustive():
    for n in range(1, 13):
        for bits in itertools.product([False, True], repeat=n):
            a = list(range(n))
            assert closed_form(a, bits) == swap_partition(a, bits), bits


def test_closed_form_random():
    rng = np.random.default_rng(5)
    for _ in range(3000):
        n = int(rng.integers(1, 400))
        a = [int(x) for x in rng.permutation(2000)[:n]]
        p = rng.random()
        left = {x: bool(rng.random() < p) for x in a}
        assert closed_form(a, left) == swap_partition(a, left)
