"""Mutation check of the oracle's pins: each test builds the oracle from its
source with one plausible mistake planted (a dropped or doubled term, a
wrong sign, index or time step), and requires the pins of
test_oracle_adjoint_pins.py to reject that build.  The mistakes are planted
in a temporary copy of the oracle source, on the CPU; the product is not
involved."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from test_oracle_adjoint_pins import (check_adjoint_closed_form, check_gradient_swap, check_interval_invariance,
                                      check_reciprocity_records)

ORACLE_DIR = os.path.dirname(os.path.abspath(oracle.__file__))
IMPL = "fwi2d_oracle_impl.h"

# name -> (exact source text of the oracle, its mutated form)
MUTANTS = {
    "imaging without the nt_interval factor":
        ("src[ip] * p2[ip] * kk * dt;\n                else", "src[ip] * p2[ip] * dt;\n                else"),
    "imaging with the nt_interval factor twice":
        ("src[ip] * p2[ip] * kk * dt;\n                else", "src[ip] * p2[ip] * kk * kk * dt;\n                else"),
    "imaging reads the next store slot":
        ("c->psave + (long)(it / kk) * c->ntp", "c->psave + (long)((it / kk + 1) % c->nsave) * c->ntp"),
    "source without v^2":
        ("* dt * dt * c->velp[s] * c->velp[s];", "* dt * dt;"),
    "source with v instead of v^2":
        ("* dt * dt * c->velp[s] * c->velp[s];", "* dt * dt * c->velp[s];"),
    "residual injected without v^2":
        ("* dt * dt * c->velp[r] * c->velp[r];", "* dt * dt;"),
    "stored second derivative with a wrong sign":
        ("(p2[ip] - 2 * p1[ip] + p0[ip])", "(p2[ip] - 2 * p1[ip] - p0[ip])"),
    "imaging condition with the wrong sign":
        ("grad[gi] += -2.0 / OFWI_POW3", "grad[gi] += 2.0 / OFWI_POW3"),
    "imaging condition with 1/v^4":
        ("-2.0 / OFWI_POW3(v[ip]) * src", "-2.0 / OFWI_POW3(v[ip]) / v[ip] * src"),
    "run_fwi=0 stores the previous field":
        ("dst[ip] = p2[ip];", "dst[ip] = p1[ip];"),
    "run_fwi=0 imaging without the nt_interval factor":
        ("grad[gi] += src[ip] * p2[ip] * kk * dt;", "grad[gi] += src[ip] * p2[ip] * dt;"),
    "run_fwi=0 injects -obs":
        ("else c.record[i] = obs[i];", "else c.record[i] = -obs[i];"),
}

CHECKS = [
    ("reciprocity", lambda: check_reciprocity_records(oracle)),
    ("closed form run_fwi=1", lambda: check_adjoint_closed_form(oracle, 1)),
    ("closed form run_fwi=0", lambda: check_adjoint_closed_form(oracle, 0)),
    ("closed form run_fwi=1 k=2", lambda: check_adjoint_closed_form(oracle, 1, k=2)),
    ("closed form run_fwi=0 k=2", lambda: check_adjoint_closed_form(oracle, 0, k=2)),
    ("swap run_fwi=1", lambda: check_gradient_swap(oracle, 1)),
    ("swap run_fwi=0", lambda: check_gradient_swap(oracle, 0)),
    ("interval run_fwi=1", lambda: check_interval_invariance(oracle, 1)),
    ("interval run_fwi=0", lambda: check_interval_invariance(oracle, 0)),
]


def build_mutant(tmp_path, old, new):
    for name in ("fwi2d_oracle.c", "fwi2d_oracle.h", IMPL):
        shutil.copy(os.path.join(ORACLE_DIR, name), tmp_path / name)
    src = (tmp_path / IMPL).read_text()
    assert src.count(old) == 1, f"mutation site not unique or missing: {old!r}"
    (tmp_path / IMPL).write_text(src.replace(old, new))
    lib = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=gnu11", "-DOFWI_ROW_MAJOR",
                    "-shared", "-o", str(lib), str(tmp_path / "fwi2d_oracle.c"), "-lm"], check=True)
    return str(lib)


def failed_checks():
    failed = []
    for name, fn in CHECKS:
        try:
            fn()
        except AssertionError:
            failed.append(name)
    return failed


@pytest.fixture()
def restore_oracle(oracle_mod):
    prev = oracle.use_library(oracle._active)
    yield
    oracle.use_library(prev)


def test_unmutated_build_passes_every_check(tmp_path, restore_oracle):
    # the same build route with nothing planted
    oracle.use_library(build_mutant(tmp_path, "dst[ip] = p2[ip];", "dst[ip] = p2[ip];"))
    assert failed_checks() == []


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_mutant_is_rejected(tmp_path, restore_oracle, mutant):
    old, new = MUTANTS[mutant]
    oracle.use_library(build_mutant(tmp_path, old, new))
    failed = failed_checks()
    print(f"{mutant}: rejected by {failed}")
    assert failed, f"no pin rejects the mutant '{mutant}'"
