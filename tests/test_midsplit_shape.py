"""The closed form the Morton builders' equal-key kernel uses for a middle-split
subtree (radix.h:455-465 splits a range of equal keys at (begin+end)/2 until at
most T prims remain), checked against the recursion itself.

Closed form (cubql_amd/csrc/radix_builder.hip, MidShape / k_karras_equal):
  - D = min{d : m >> d <= T}; every node above depth D is inner; at depth D the
    inner nodes are those of T + 1 prims, which exist only when q = m >> D == T,
    and there are r = m mod 2^D of them;
  - the node at depth d with path bits k (most significant first, 1 = right)
    holds [b, b + s) where each step halves s to floor(s/2) (left) or
    s - floor(s/2) (right), b advancing by floor(s/2) on a right step;
  - a depth-D inner node's rank among them is b - k q.
"""
import pytest


def recursive_inner(m, T):
    """(depth, begin, size) of every inner node of the recursive middle split."""
    out = []

    def rec(b, e, d):
        if e - b <= T:
            return
        out.append((d, b, e - b))
        mid = (b + e) // 2
        rec(b, mid, d + 1)
        rec(mid, e, d + 1)
    rec(0, m, 0)
    return out


def shape(m, T):
    D, q = 0, m
    while q > T:
        D += 1
        q = m >> D
    r = m & ((1 << D) - 1) if D else 0
    return D, q, r


def closed_form_inner(m, T):
    D, q, r = shape(m, T)
    full = (1 << D) - 1
    items = full + (full + 1 if q == T else 0)
    nodes, ranks = [], []
    for w in range(items):
        if w < full:
            d = (w + 1).bit_length() - 1
            k = w + 1 - (1 << d)
        else:
            d, k = D, w - full
        b, s = 0, m
        for j in range(d - 1, -1, -1):
            h = s >> 1
            if (k >> j) & 1:
                b, s = b + h, s - h
            else:
                s = h
        if s <= T:
            continue
        nodes.append((d, b, s))
        if d == D:
            ranks.append(b - k * q)
    return nodes, ranks, full + (r if q == T else 0)


@pytest.mark.parametrize("T", [1, 2, 3, 4, 5, 8, 16])
def test_middle_split_closed_form(T):
    for m in list(range(T + 1, 600)) + [1000, 1023, 1024, 1025, 4097, 65537]:
        ref = recursive_inner(m, T)
        nodes, ranks, count = closed_form_inner(m, T)
        assert sorted(nodes) == sorted(ref), (m, T)
        assert count == len(ref), (m, T)
        # depth-D ranks number those nodes 0, 1, ... in position order
        assert ranks == list(range(len(ranks))), (m, T)
