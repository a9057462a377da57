"""CPU replay of the device kernels' code, compiled for the host with a checking prelude
(tests/emu/emu_prelude.h). Mirrors the reference's kernel replay
(crates/ckks-train-cuda/src/emu.rs:1-8,72-87).

* The register-radix NTT (tests/emu/ntt_emu.cpp): the pass code the gfx950 kernels run
  (ckks-train_amd/csrc/ntt_passes.h and arith.h), replayed one simulated 256-thread workgroup at a
  time with every barrier, LDS exchange and global access checked, every f64 and Shoup modular
  product verified against 128-bit integers, the words between the passes held to the bounds the
  arithmetic relies on (also under an adversarial search of the column pass's inputs, and with the
  row pass fed words at the bound and checked against the integer-arithmetic pass), and the
  outputs compared with the oracle; the inverse transform in a conversion source's y form; the row
  pass with mod_down's / rescale's finish (k_fwd_rows<M, true>); then the fused row pass + key
  inner product + mod_down finish of the batched key switch (k_rows_ip_fin's body), every input
  word read once and every output written once, against the oracle's transforms and products.
* The base conversion (tests/emu/conv_emu.cpp): the per-coefficient code of k_conv and the block
  code of the matrix-core k_conv_mf (ckks-train_amd/csrc/conv_arith.h; its wave-wide int8 matrix
  product and lane exchange modelled on 64 host threads per wave) over the plans a context builds
  (conv_tables.h, A tiles included) for the key-switch prime shapes and wide digits, with
  exactly-once coverage of every output word, every f64 product checked, the int32 accumulators
  checked for overflow, and the outputs compared with the oracle's convert_exact on random,
  extreme and near-tie inputs of the f64 overflow estimate; and the key inner product
  (ckks-train_amd/csrc/ip_arith.h, k_inner_product) over 1..16 digits on 50- and 60-bit primes,
  with and without the Galois permutation, against 128-bit sums.

The planted mistakes (an index map that is not one-to-one in LDS or in memory, a wrong twiddle, a
twiddle whose quotient disagrees, a reciprocal that moves the overflow estimate, a product
constant whose quotient is wrong, two rows of a matrix-core operand tile exchanged) must each be
caught."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
ORACLE_DIR = os.path.join(ROOT, "oracle", "_build")


def _build(name):
    if not os.path.exists(os.path.join(ORACLE_DIR, "libckks_oracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    out = os.path.join(EMU, "_build", name)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    # no FMA contraction: the f64 steps must round as on the device, where it is off too
    subprocess.run(["g++", "-std=c++20", "-O2", "-pthread", "-Wall", "-Wno-unknown-pragmas", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "ckks-train_amd", "csrc"), os.path.join(EMU, name + ".cpp"), "-o", out,
                    "-L", ORACLE_DIR, "-lckks_oracle", "-Wl,-rpath," + ORACLE_DIR], check=True)
    return out


@pytest.fixture(scope="module")
def emu_bin():
    return _build("ntt_emu")


@pytest.fixture(scope="module")
def conv_bin():
    return _build("conv_emu")


def test_replay_of_the_ntt_passes_is_exact_and_race_free(emu_bin):
    r = subprocess.run([emu_bin, "1", "8"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "EMU OK" in r.stdout
    lines = [l for l in r.stdout.splitlines() if "limbs forward and inverse exact, inverse in y form exact" in l]
    assert len(lines) == 10  # N = 2^12..2^16, 50- and 60-bit primes (M = 6, 7, 8 in both passes)
    fused = [l for l in r.stdout.splitlines() if "fused row pass + inner product + finish exact" in l]
    assert len(fused) == 5  # N = 2^12..2^16
    fin = [l for l in r.stdout.splitlines() if "row pass + finish (a - NTT(x)) inv (+ add) exact" in l]
    assert len(fin) == 5  # N = 2^12..2^16
    adv = [l for l in r.stdout.splitlines() if l.startswith("adversarial f64 inputs")]
    assert len(adv) == 2 and all("row pass on words up to 2.600 q equals the integer pass" in l for l in adv)


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


def test_replay_of_the_conversion_is_exact_and_covers_every_word(conv_bin):
    r = subprocess.run([conv_bin], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "EMU OK" in r.stdout
    lines = [l for l in r.stdout.splitlines() if "output words equal the oracle" in l]
    assert len(lines) == 5  # the five prime shapes
    # every shape has plans with a matrix form (at most 8 sources), replayed on the matrix-core kernel,
    # and plans replayed on y-form inputs (the inverse NTT's folded first step)
    assert all(int(l.split("(")[1].split()[0]) > 0 for l in lines)
    assert all(int(l.split(" on y-form inputs")[0].split()[-1]) > 0 for l in lines)
    assert "inner product: 524288 output words equal the 128-bit sums" in r.stdout
    # the near-tie inputs reach the boundary where the f64 estimate's floor is off by one
    assert all(int(l.split("off by one, as in the oracle: ")[1].split()[0]) > 0 for l in lines)


@pytest.mark.parametrize("plant,expect", [
    (5, "conversion output differs from the oracle"),
    (6, "fmulmod result"),
    (7, "conversion output differs from the oracle"),
])
def test_conversion_replay_catches_planted_mistakes(conv_bin, plant, expect):
    r = subprocess.run([conv_bin, str(plant)], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert expect in r.stderr
