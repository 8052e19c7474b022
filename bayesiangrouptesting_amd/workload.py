"""Seeded synthetic inputs for the benchmark configurations.

The reference has no random-number generator, so inputs are specified here
portably: splitmix64 seeded by ``seed``; a uniform is ``(x >> 11) * 2**-53``;
a prior is ``lo + (hi - lo) * u``. True-state atom bits are drawn as 0 with
probability ``pi[a]``; each response bit is drawn as 1 with the likelihood
of the drawn true state, from its own stream.
"""
from __future__ import annotations

from dataclasses import dataclass

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def uniform_range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.uniform()

    def below(self, n: int) -> int:
        return self.next_u64() % n


def priors(seed: int, atoms: int, lo: float, hi: float) -> list[float]:
    g = SplitMix64(seed)
    return [g.uniform_range(lo, hi) for _ in range(atoms)]


def true_state(seed: int, pi: list[float]) -> int:
    """Draw a state: bit a is 0 (positive) with probability pi[a]."""
    g = SplitMix64(seed ^ 0x5DEECE66D)
    s = 0
    for a, p in enumerate(pi):
        if g.uniform() >= p:
            s |= 1 << a
    return s


NEGATIVE_RESPONSE = 0.99


def response_bit_probability(dilution_rows, subjects: int, experiment: int, variant_state: int) -> float:
    """P(r_v = 1) for one variant (replicated_lattice.cpp:6-28)."""
    if dilution_rows is None:
        return NEGATIVE_RESPONSE if (experiment & variant_state) == experiment else 1.0 - NEGATIVE_RESPONSE
    size = bin(experiment).count("1")
    if size == 0:
        return 0.0
    missing = size - bin(experiment & variant_state).count("1")
    return dilution_rows[(size - 1) * (subjects + 1) + missing]


class ResponseSampler:
    """Draws responses for experiments against a fixed true state."""

    def __init__(self, seed: int, subjects: int, variants: int, state: int, dilution_rows=None):
        self.g = SplitMix64(seed ^ 0xA5A5A5A5DEADBEEF)
        self.n = subjects
        self.k = variants
        self.state = state
        self.dil = dilution_rows

    def draw(self, experiment: int) -> int:
        r = 0
        mask = (1 << self.n) - 1
        for v in range(self.k):
            vs = (self.state >> (v * self.n)) & mask
            p1 = response_bit_probability(self.dil, self.n, experiment, vs)
            if self.g.uniform() < p1:
                r |= 1 << v
        return r


def dilution_table(subjects: int, alpha: float, h: float) -> list[float]:
    """Flat subjects x (subjects+1) table, row s-1 for pool size s (dilution.cpp:44-67)."""
    out = [0.0] * (subjects * (subjects + 1))
    for rk in range(1, subjects + 1):
        base = (rk - 1) * (subjects + 1)
        out[base] = alpha
        for r in range(1, rk + 1):
            k = rk - r
            out[base + r] = 1 - alpha * r / (k * h + r)
    return out


@dataclass(frozen=True)
class Config:
    name: str
    subjects: int
    variants: int
    dilution: bool
    prior_lo: float
    prior_hi: float
    seed: int
    alpha: float = 0.05
    h: float = 0.9


# The five benchmark workloads (sizes as stated by the benchmark definition).
C1 = Config("binary-n12-k1-nondilution", 12, 1, False, 0.01, 0.15, 1)
C2 = Config("binary-n28-k1-nondilution", 28, 1, False, 0.005, 0.05, 2)
C3 = Config("multinomial-n8-k3-dilution", 8, 3, True, 0.01, 0.10, 3)
C4 = Config("tree-n16-k1-dilution-depth6", 16, 1, True, 0.01, 0.10, 4)
