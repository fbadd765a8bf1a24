"""Mutation runs of the oracle, CPU only: the oracle is rebuilt with one
planted mistake at a time (ORC_MUTATION in oracle/oracle.c) and the pin tests
are run against each build.  Every mistake the pins are meant to catch must
make at least one of them fail; the regression digest of the oracle's own
output (test_c1_regression_fixture) is left out, because it would flag any
change at all and pins nothing.

Planted mistakes:
  1  formTreelet puts the expanded node's children in each other's slots
  2  formTreelet expands the smallest-area internal leaf, not the largest
  3  the TRBVH DP's collapse term uses each subset's own triangle count
  4  the treelet threshold grows n, 2n, 3n, ... instead of n, 2n, 4n, ...
  5  ATRBVH does not refresh the distances to a merged cluster
  6  the TRBVH DP skips the first partition of each subset
  7  a TRBVH right child gets its own cost (the reference's quirk removed)
  8  a rewritten TRBVH root's nodesSah is updated (the quirk removed)
  9  an ATRBVH merged node's SAH drops its ci * area term

1, 7 and 8 change only the orientation of children or costs fed to later
treelets; no closed form fixes those, so they stay pinned by the restatement
alone (the GPU kernels agree with it bit for bit), and this test records that
the pins do not catch them.
"""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CAUGHT = (2, 3, 4, 5, 6, 9)
RESTATEMENT_ONLY = (1, 7, 8)
PIN_TESTS = ["tests/test_oracle_composition.py", "tests/test_oracle_pins.py"]


def _build(k: int, out: str) -> str:
    so = os.path.join(out, f"liboracle_m{k}.so")
    o1 = os.path.join(out, f"oracle_m{k}.o")
    o2 = os.path.join(out, "sbvh.o")
    flags = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC"]
    subprocess.run(["gcc", *flags, "-std=c11", f"-DORC_MUTATION={k}", "-c",
                    os.path.join(ROOT, "oracle", "oracle.c"), "-o", o1], check=True)
    if not os.path.exists(o2):
        subprocess.run(["g++", *flags, "-std=c++17", "-pthread", "-c",
                        os.path.join(ROOT, "oracle", "sbvh.cpp"), "-o", o2], check=True)
    subprocess.run(["g++", "-shared", "-pthread", "-o", so, o1, o2, "-lm"], check=True)
    return so


def _run_pins(lib: str, stop_at_first: bool):
    env = dict(os.environ, ATRBVH_ORACLE_LIB=lib)
    cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
           "--deselect", "tests/test_oracle_pins.py::test_c1_regression_fixture", *PIN_TESTS]
    if stop_at_first:
        cmd.insert(3, "-x")
    return subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


@pytest.mark.parametrize("k", CAUGHT)
def test_pins_catch_planted_mistake(mutant_dir, k):
    r = _run_pins(_build(k, mutant_dir), stop_at_first=True)
    assert r.returncode != 0, f"mutation {k} passed every pin:\n{r.stdout[-2000:]}"


@pytest.mark.parametrize("k", RESTATEMENT_ONLY)
def test_restatement_only_mistakes_pass_the_pins(mutant_dir, k):
    r = _run_pins(_build(k, mutant_dir), stop_at_first=False)
    assert r.returncode == 0, (f"mutation {k} is now caught: move it to CAUGHT\n"
                               f"{r.stdout[-2000:]}")
