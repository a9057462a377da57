"""CPU replay of the register-radix NTT kernels (tests/emu/ntt_emu.cpp): the pass code the gfx950
kernels run (ckks-train_amd/csrc/ntt_passes.h and arith.h, compiled for the host), replayed one
simulated 256-thread workgroup at a time with every barrier, LDS exchange and global access
checked, every f64 and Shoup modular product verified against 128-bit integers, the words
between the passes held to the bounds the arithmetic relies on, and the outputs compared with
the oracle. Mirrors the reference's kernel replay (crates/ckks-train-cuda/src/emu.rs:1-8,72-87).
The planted mistakes (an index map that is not one-to-one in LDS or in memory, a wrong twiddle, a
twiddle whose quotient disagrees) must each be caught."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "emu", "ntt_emu.cpp")
OUT = os.path.join(ROOT, "tests", "emu", "_build", "ntt_emu")
ORACLE_DIR = os.path.join(ROOT, "oracle", "_build")


@pytest.fixture(scope="module")
def emu_bin():
    if not os.path.exists(os.path.join(ORACLE_DIR, "libckks_oracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.run(["g++", "-std=c++20", "-O2", "-pthread", "-Wall", "-Wno-unknown-pragmas",
                    "-I", os.path.join(ROOT, "ckks-train_amd", "csrc"), SRC, "-o", OUT,
                    "-L", ORACLE_DIR, "-lckks_oracle", "-Wl,-rpath," + ORACLE_DIR], check=True)
    return OUT


def test_replay_of_the_ntt_passes_is_exact_and_race_free(emu_bin):
    r = subprocess.run([emu_bin, "1", "8"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "EMU OK" in r.stdout
    lines = [l for l in r.stdout.splitlines() if l.startswith("N=2^")]
    assert len(lines) == 10  # N = 2^12..2^16, 50- and 60-bit primes (M = 6, 7, 8 in both passes)


@pytest.mark.parametrize("plant,expect", [
    (1, "LDS slot written twice"),
    (2, "does not touch every word exactly once"),
    (3, "forward output differs from the oracle"),
    (4, "fmulmod result not congruent"),
])
def test_replay_catches_planted_mistakes(emu_bin, plant, expect):
    r = subprocess.run([emu_bin, "0", "0", str(plant)], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert expect in r.stderr
