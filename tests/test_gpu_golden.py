"""The device path against the committed fixtures of tests/golden (the oracle's outputs frozen by
tests/golden/make_golden.py): the prime hand-out, key switches at full, middle and low levels,
and relinearizations, all word-exact (SHA-256 of the output limbs)."""
import hashlib
import json
import os

import numpy as np
import pytest

import ckh

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "key_switch_n1024.json")
DOC = json.load(open(FIX))
BITS = {"alpha1": ([60, 40, 40, 40], [60]), "alpha2_k2": ([60, 40, 40, 40, 40, 40], [60, 60]),
        "shape_24q_6p": ([50] * 24, [60] * 6)}


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype="<u8").tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name", sorted(DOC["cases"]))
def test_device_matches_fixture(name):
    print("golden", name, flush=True)
    case = DOC["cases"][name]
    qb, pb = BITS[name]
    q, p = ckh.params_primes(case["log_n"], qb, pb)
    assert [str(x) for x in q] == case["q"] and [str(x) for x in p] == case["p"]
    ctx = ckh.Context(case["log_n"], q, p, case["dnum"])
    L = case["key"]["max_level"]
    key = ctx.ksk_uniform(L, seed=case["key"]["seed"])
    relin = ctx.ksk_uniform(L, seed=case["relin_key"]["seed"])
    ev = ckh.Evaluator(ctx, relin, [])
    for ks in case["key_switch"]:
        kb, ka = ctx.key_switch(ctx.uniform(ks["level"], seed=ks["input_seed"]), key)
        assert digest(kb.download(), ka.download()) == ks["sha256"], ks["level"]
    for rl in case["relinearize"]:
        ct = ctx.ct_from_polys([ctx.uniform(rl["level"], seed=s) for s in rl["input_seeds"]], 2.0 ** 40)
        out = ev.relinearize(ct)
        assert digest(out.component(0).download(), out.component(1).download()) == rl["sha256"], rl["level"]
    ctx.synchronize()
