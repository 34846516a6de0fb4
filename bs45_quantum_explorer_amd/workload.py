"""Workload definitions shared by the bench and the tests: the signature classes of an n and
a seeded cell sample.

classes(n) enumerates the class representatives the reference's tools enumerate
(tools/test_orbit_q.py:26-35): a <= b, c <= d, all nonnegative, a, b of the parity of n+1,
c, d of the parity of n, and a^2 + b^2 + c^2 + d^2 = 4n + 2 (the norm check of
src/solver/wz_match.cpp:1358-1362).
"""

MASK64 = (1 << 64) - 1


def classes(n):
    out = []
    for a in range(n + 2):
        for b in range(a, n + 2):
            for c in range(n + 1):
                for d in range(c, n + 1):
                    if (a - n - 1) % 2 or (b - n - 1) % 2 or (c - n) % 2 or (d - n) % 2:
                        continue
                    if a * a + b * b + c * c + d * d == 4 * n + 2:
                        out.append((a, b, c, d))
    return out


def splitmix64(state):
    """The splitmix64 generator (Steele, Lea, Flood 2014): an infinite stream of 64-bit words."""
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def sample_cells(live, k, seed):
    """k distinct cells of `live`, drawn without replacement by a partial Fisher-Yates shuffle
    driven by splitmix64(seed); returned in ascending list order."""
    live = list(live)
    rng = splitmix64(seed)
    k = min(k, len(live))
    for i in range(k):
        j = i + next(rng) % (len(live) - i)
        live[i], live[j] = live[j], live[i]
    return sorted(live[:k])


def _pack(v):
    return sum(1 << i for i, x in enumerate(v) if x < 0)


def pair_variants(X, Y):
    """Images of a pair (X, Y) that keep N_X(s) + N_Y(s) at every shift s: negating either
    sequence, reversing either one, and swapping them. Distinct packed (X, Y) words, in a fixed
    order."""
    out = []
    for neg_x, neg_y, rev_x, rev_y, swap in ((0, 0, 0, 0, 0), (1, 0, 0, 0, 0), (0, 1, 0, 0, 0), (1, 1, 0, 0, 0),
                                              (0, 0, 1, 0, 0), (0, 0, 0, 1, 0), (0, 0, 0, 0, 1), (1, 0, 1, 0, 1)):
        x = [-v for v in X] if neg_x else list(X)
        y = [-v for v in Y] if neg_y else list(Y)
        x = x[::-1] if rev_x else x
        y = y[::-1] if rev_y else y
        if swap:
            x, y = y, x
        w = (_pack(x), _pack(y))
        if w not in out:
            out.append(w)
    return out


def join_workload(n, solution, nrand, seed):
    """The join at table scale: nrand random C,D pairs (n elements) and nrand random A,B pairs
    (n+1 elements) -- their autocorrelation vectors meet with vanishing probability -- with the
    images of one known solution planted at seeded positions on both sides. Every planted A,B
    image then hits the one table entry of the planted C,D images, and the solutions are exactly
    the planted A,B images times the planted C,D images. Returns (cd, ab, planted_cd, planted_ab)
    with cd, ab as uint64 arrays of shape (k, 2)."""
    import numpy as np
    A, B, C, D = solution
    rng = np.random.default_rng(seed)
    pcd, pab = pair_variants(C, D), pair_variants(A, B)
    cd = rng.integers(0, 1 << n, size=(nrand + len(pcd), 2), dtype=np.uint64)
    ab = rng.integers(0, 1 << (n + 1), size=(nrand + len(pab), 2), dtype=np.uint64)
    for arr, planted in ((cd, pcd), (ab, pab)):
        at = rng.choice(len(arr), size=len(planted), replace=False)
        for i, (x, y) in zip(at, planted):
            arr[i] = (x, y)
    return cd, ab, pcd, pab
