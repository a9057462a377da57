"""Writes key_switch_n1024.json: small key-switching fixtures (inputs named by their seeds, since
the seeded residue generator is fixed and checked against its definition; outputs as SHA-256 of
the little-endian u64 words plus a few words in the clear).

The expected outputs come from the oracle (oracle/), which the closed-form tests of
tests/test_oracle.py pin; the fixture freezes them so both the oracle and the device path are
checked against committed data. Run from the repository root:

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

# (name, log_n, q_bits, p_bits, dnum): one-prime digits / two-prime digits with two special
# primes / the 24 + 6 prime shape of the N = 2^16 workload, all at N = 2^10
CASES = [("alpha1", 10, [60, 40, 40, 40], [60], 4),
         ("alpha2_k2", 10, [60, 40, 40, 40, 40, 40], [60, 60], 3),
         ("shape_24q_6p", 10, [50] * 24, [60] * 6, 4)]
KEY_SEED = 901
REL_SEED = 902


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype="<u8").tobytes())
    return h.hexdigest()


def sample(a):
    a = np.asarray(a, dtype=np.uint64)
    return [str(int(a.ravel()[0])), str(int(a.ravel()[-1]))]


def levels(nq):
    L = nq - 1
    return sorted({L, L // 2, 1, 0}, reverse=True)


def build_case(name, log_n, qb, pb, dnum):
    ref = O.Params(log_n, qb, pb, dnum)
    L = len(ref.q) - 1
    key = ref.uniform_key(L, KEY_SEED)
    relin = ref.uniform_key(L, REL_SEED)
    out = {"log_n": log_n, "q": [str(q) for q in ref.q], "p": [str(p) for p in ref.p], "dnum": dnum,
           "key": {"max_level": L, "seed": KEY_SEED}, "relin_key": {"max_level": L, "seed": REL_SEED},
           "key_switch": [], "relinearize": []}
    for lv in levels(len(ref.q)):
        seed = 1000 + lv
        a = O.uniform_poly(ref.q[: lv + 1], ref.n, seed)
        kb, ka = ref.key_switch(a.ravel(), lv, key, L)
        out["key_switch"].append({"level": lv, "input_seed": seed, "sha256": digest(kb, ka),
                                  "kb_first_last": sample(kb), "ka_first_last": sample(ka)})
        if lv >= 1:
            c = [O.uniform_poly(ref.q[: lv + 1], ref.n, 2000 + 10 * lv + j) for j in range(3)]
            o0, o1 = ref.relinearize(c[0].ravel(), c[1].ravel(), c[2].ravel(), lv, relin, L)
            out["relinearize"].append({"level": lv, "input_seeds": [2000 + 10 * lv + j for j in range(3)],
                                       "sha256": digest(o0, o1)})
    return out


def main():
    doc = {"about": "oracle key switches and relinearizations at N = 2^10; inputs: generator seeds "
                    "(stream0 0 for inputs, 1000 (j + 1) for key polynomial j); outputs: SHA-256 "
                    "of (kb, ka) or (c0', c1') as little-endian u64 limbs",
           "cases": {name: build_case(name, lg, qb, pb, dn) for name, lg, qb, pb, dn in CASES}}
    with open(os.path.join(HERE, "key_switch_n1024.json"), "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    main()
