"""Shared pieces of the lookahead tests (not collected by pytest).

The k-step Op-BHA lookahead (lattice_api.cpp:116-168) scores each candidate c by the L1
distance of its 2^(m+1) group masses from uniform; group bit i says whether a state is
contained in the i-th selected experiment (bit m: in c). Candidates whose scores agree in
exact arithmetic are frequent (whenever both halves of every earlier group lie on one side
of the target, the score does not depend on c at all), and the reference then picks by its
summation rounding. The tie rule checked here: the lowest public candidate whose score lies
within 2 * tau of the lowest score, tau bounding one score's FP64 rounding.
"""
import numpy as np


def tau(n: int, m: int) -> float:
    """Rounding bound of one score over 2^(m+1) groups of n-subject states."""
    g = 1 << (m + 1)
    return g * (g * (n + 1) + 3) * 2.0 ** -53


def tie_pick(scores: np.ndarray, n: int, m: int) -> int:
    """Lowest public candidate (index >= 1) within 2 * tau of the lowest score."""
    s = scores[1:]
    limit = s.min() + 2 * tau(n, m)
    return 1 + int(np.argmax(s <= limit))


def near_tie(scores: np.ndarray, n: int, m: int) -> bool:
    """True when more than one candidate lies within 2 * tau of the lowest score."""
    s = scores[1:]
    return int(np.count_nonzero(s <= s.min() + 2 * tau(n, m))) > 1


def public_order(n: int) -> np.ndarray:
    """perm[internal] = public index (bit reversal over n bits)."""
    v = np.arange(1 << n, dtype=np.uint64)
    r = np.zeros_like(v)
    for i in range(n):
        r |= ((v >> np.uint64(i)) & np.uint64(1)) << np.uint64(n - 1 - i)
    return r.astype(np.int64)


def transform_scores(post_internal: np.ndarray, n: int, selected_public) -> np.ndarray:
    """Every public candidate's score by a second formulation: down-set sums
    D(X) = sum_{s <= X} x(s), the mass of the states inside every experiment of a set K as
    D(intersection of K), and each group's mass by inclusion-exclusion over supersets.
    Index 0 (the empty experiment) is +inf."""
    perm = public_order(n)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)  # inv[public] = internal
    D = np.array(post_internal, dtype=np.float64).copy()
    for b in range(n):
        v = D.reshape(-1, 2, 1 << b)
        v[:, 1, :] += v[:, 0, :]
    m = len(selected_public)
    sel = [int(inv[e]) for e in selected_public]
    full = (1 << n) - 1
    groups = 1 << (m + 1)
    ci = np.arange(1 << n)
    g = np.zeros((groups, 1 << n))
    for k in range(1 << m):
        x = full
        for i in range(m):
            if (k >> i) & 1:
                x &= sel[i]
        g[k] = D[x]
        g[k | (1 << m)] = D[x & ci]
    for b in range(m + 1):
        for j in range(groups):
            if not (j >> b) & 1:
                g[j] -= g[j | (1 << b)]
    sc = np.abs(g - 1.0 / groups).sum(0)
    out = np.empty(1 << n)
    out[perm] = sc
    out[0] = np.inf
    return out
