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
