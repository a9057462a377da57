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


STRUCTS = [("ckh_params", "Params"), ("ckh_op_counter", "OpCounter"), ("ckh_bsgs_shape", "BsgsShape"),
           ("ckh_linear_stage", "LinearStageC"), ("ckh_server_config", "ServerConfig"),
           ("ckh_request", "Request"), ("ckh_server_stats", "ServerStats")]


def test_binding_structs_match_the_header_layout(tmp_path):
    """Every struct the binding mirrors has the size and field offsets the C compiler gives the
    header's struct (a mismatch would pass wrong values through the C ABI silently)."""
    import ctypes as C
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "ckh.h"', "int main(void) {"]
    for cname, pname in STRUCTS:
        cls = getattr(ckh, pname)
        lines.append(f'  printf("%zu", sizeof({cname}));')
        lines += [f'  printf(" %zu", offsetof({cname}, {f}));' for f, _ in cls._fields_]
        lines.append('  printf("\\n");')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    for (cname, pname), line in zip(STRUCTS, out):
        cls = getattr(ckh, pname)
        got = [C.sizeof(cls)] + [getattr(cls, f).offset for f, _ in cls._fields_]
        assert got == [int(v) for v in line.split()], cname


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
