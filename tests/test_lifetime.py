"""The free-ordering rule of the device library (ckks-train_amd/csrc/lifetime.h), exercised on a CPU
model of streams and events (tests/lifetime_model.cpp): with the ledger of every stream that used a
buffer, a free never runs before a queued use in any order the streams allow; ordering the free only
after the releasing thread's stream (the rule it replaced) lets a free overtake a use queued on a
third stream."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("lifetime") / "lifetime_model")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-o", exe, os.path.join(HERE, "lifetime_model.cpp")],
                   check=True)
    return exe


def _run(exe, n):
    out = subprocess.run([exe, str(n)], check=True, capture_output=True, text=True).stdout.split("\n")
    res = {}
    for line in out:
        if line.strip():
            name, total, viol, dead = line.split()
            res[name] = (int(total), int(viol), int(dead))
    return res


def test_ledger_orders_every_use_before_the_free(model):
    res = _run(model, 20000)
    total, viol, dead = res["ledger"]
    assert total == 20000 and dead == 0
    assert viol == 0


def test_releasing_stream_only_misses_third_stream_uses(model):
    res = _run(model, 20000)
    total, viol, dead = res["releasing_stream_only"]
    assert dead == 0
    assert viol > 0  # the gap the ledger closes
