"""CPU checks of the C ABI: the library loads, exports every symbol include/ckh.h declares
(and the binding declares the same set), and its host-only entry points (prime hand-out,
Galois elements) agree with the oracle. No device calls."""
import os
import re

import pytest

import ckh
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ckh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ckh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ckh.lib()
    syms = header_symbols()
    assert len(syms) > 50
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert syms == ckh.declared_symbols()


def test_params_primes_match_oracle_assignment():
    for log_n, qb, pb in [(16, [50] * 24, [60] * 6), (15, [50] * 20, [60] * 4),
                          (15, [60] + [40] * 9, [60]), (12, [58, 40, 58], [58])]:
        q, p = ckh.params_primes(log_n, qb, pb)
        ref = O.Params(log_n, qb, pb, 1 if len(qb) % 2 else 1)
        assert q == ref.q and p == ref.p


def test_params_primes_errors():
    with pytest.raises(ckh.InvalidParams):
        ckh.params_primes(18, [50], [60])
    with pytest.raises(ckh.InvalidParams):
        ckh.params_primes(16, [61], [60])
    with pytest.raises(ckh.InvalidParams):
        ckh.params_primes(16, [17] * 5, [60])


def test_galois_elements_match_oracle():
    for log_n in (10, 15, 16):
        n = 1 << log_n
        for step in (0, 1, 3, 5, -1, -7, n // 2, n // 2 + 3, 12345):
            assert ckh.galois_elt(log_n, step) == O.galois_elt(n, step)
        assert ckh.conjugation_elt(log_n) == O.conjugation_elt(n)
