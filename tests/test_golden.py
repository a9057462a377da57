"""The committed fixtures of tests/golden (written by tests/golden/make_golden.py) reproduced
by the oracle: the prime hand-out, key switches at full, middle and low levels (clipped
digits), and relinearizations. CPU only; the device path is checked against the same fixtures
in test_gpu_parity.py."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O

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
def test_oracle_reproduces_fixture(name):
    case = DOC["cases"][name]
    qb, pb = BITS[name]
    ref = O.Params(case["log_n"], qb, pb, case["dnum"])
    assert [str(q) for q in ref.q] == case["q"] and [str(p) for p in ref.p] == case["p"]
    L = case["key"]["max_level"]
    key = ref.uniform_key(L, case["key"]["seed"])
    relin = ref.uniform_key(L, case["relin_key"]["seed"])
    for ks in case["key_switch"]:
        lv = ks["level"]
        a = O.uniform_poly(ref.q[: lv + 1], ref.n, ks["input_seed"])
        kb, ka = ref.key_switch(a.ravel(), lv, key, L)
        assert digest(kb, ka) == ks["sha256"], lv
        assert [str(int(kb.ravel()[0])), str(int(kb.ravel()[-1]))] == ks["kb_first_last"]
    for rl in case["relinearize"]:
        lv = rl["level"]
        c = [O.uniform_poly(ref.q[: lv + 1], ref.n, s).ravel() for s in rl["input_seeds"]]
        o0, o1 = ref.relinearize(c[0], c[1], c[2], lv, relin, L)
        assert digest(o0, o1) == rl["sha256"], lv


def test_fixture_detects_a_planted_change():
    """A one-word change of an oracle output changes its fixture digest (the comparison is
    word-exact, not a tolerance)."""
    case = DOC["cases"]["alpha1"]
    ref = O.Params(case["log_n"], *BITS["alpha1"], case["dnum"])
    L = case["key"]["max_level"]
    ks = case["key_switch"][0]
    a = O.uniform_poly(ref.q[: ks["level"] + 1], ref.n, ks["input_seed"])
    kb, ka = ref.key_switch(a.ravel(), ks["level"], ref.uniform_key(L, case["key"]["seed"]), L)
    kb = kb.copy()
    kb[0, 0] = (int(kb[0, 0]) + 1) % ref.q[0]
    assert digest(kb, ka) != ks["sha256"]
