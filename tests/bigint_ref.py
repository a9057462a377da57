"""Closed-form big-integer references used to pin the oracle (pure Python, small sizes).

Each function states the mathematical definition, not the oracle's algorithm: CRT values,
exact rounding of integer quotients, Horner evaluation, schoolbook negacyclic products.
"""
from fractions import Fraction
import math


def bitrev(i, bits):
    r = 0
    for b in range(bits):
        r |= ((i >> b) & 1) << (bits - 1 - b)
    return r


def crt(residues, primes):
    """The integer in [0, prod(primes)) with the given residues."""
    Q = 1
    for p in primes:
        Q *= p
    x = 0
    for r, p in zip(residues, primes):
        qh = Q // p
        x += int(r) * qh * pow(qh, -1, p)
    return x % Q, Q


def horner(coeffs, x, q):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + int(c)) % q
    return acc


def negacyclic(a, b, q):
    n = len(a)
    c = [0] * n
    for i in range(n):
        ai = int(a[i])
        if ai == 0:
            continue
        for j in range(n):
            k = i + j
            p = ai * int(b[j])
            if k < n:
                c[k] = (c[k] + p) % q
            else:
                c[k - n] = (c[k - n] - p) % q
    return c


def mix64(z):
    M = (1 << 64) - 1
    z &= M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def uniform_word(seed, stream, index, q):
    """Definition of the seeded input generator shared by oracle and device."""
    M = (1 << 64) - 1
    key = mix64(seed ^ mix64((stream + 0x9E3779B97F4A7C15) & M))
    zone = (M // q) * q
    x = 0
    for a in range(64):
        x = mix64((key + ((index * 64 + a) * 0x9E3779B97F4A7C15)) & M)
        if x < zone:
            break
    return x % q


def f64_overflow(ys, qs, fused=False):
    """floor(sum_j f64(y_j) * f64(1/q_j)) evaluated with IEEE-754 round-to-nearest products
    and sums in source order starting from 0.0 (Python floats are IEEE doubles).
    fused=True models an FMA-contracted sum (one rounding of f + y*w), the mistake the
    near-tie inputs are built to expose."""
    f = 0.0
    for y, q in zip(ys, qs):
        w = 1.0 / float(q)
        if fused:
            f = float(Fraction(f) + Fraction(float(y)) * Fraction(w))
        else:
            f = f + float(y) * w
    return math.floor(f)
