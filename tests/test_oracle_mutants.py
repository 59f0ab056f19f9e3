"""The oracle pins must fail on a wrong oracle (CPU only).

For each planted mistake of orc_core.h (ORC_MUT_*), the oracle is built
once more with that mistake (`make -C oracle mutant`, output in a pytest
temporary directory) and the pin written for the function must then fail:
a check that also passes on a plausible restatement error pins nothing.
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import pin_checks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# ORC_MUT_* value -> the pin checks that must fail on that build
MUTANTS = {
    1: ("ssp-rk3 weights (1/2, 1/2)", ["rk3_2d", "rk3_3d"]),
    2: ("ppm C2 = 1.0", ["ppm_smooth_extremum"]),
    3: ("ppm extremum test && for ||", ["ppm_smooth_extremum"]),
    4: ("1d jacobian J[2,3] factor 2 dropped", ["jacobian1_fd"]),
    5: ("3x3 solve adjugate entry transposed", ["solve3_exact"]),
    6: ("1d newton update before the norm test", ["newton1_break_before_update"]),
    7: ("1dw f' term dropped", ["f1dw_derivative_fd"]),
}

CHECKS = {
    "rk3_2d": lambda O: pin_checks.rk3_temporal_order(O, 2),
    "rk3_3d": lambda O: pin_checks.rk3_temporal_order(O, 3),
    "ppm_smooth_extremum": pin_checks.ppm_smooth_extremum,
    "jacobian1_fd": pin_checks.jacobian1_fd,
    "solve3_exact": pin_checks.solve3_exact,
    "newton1_break_before_update": pin_checks.newton1_break_before_update,
    "f1dw_derivative_fd": pin_checks.f1dw_derivative_fd,
}


@pytest.fixture(scope="session")
def mutant_libs(tmp_path_factory):
    import pyoracle
    out = tmp_path_factory.mktemp("oracle_mutants")

    def build(k):
        path = str(out / f"liboracle_mut{k}.so")
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mutant", f"MUT={k}", f"OUT={path}"],
                       check=True, capture_output=True)
        return k, path

    with ThreadPoolExecutor(max_workers=4) as ex:
        paths = dict(ex.map(build, MUTANTS))
    return {k: pyoracle.load(p) for k, p in paths.items()}


@pytest.mark.parametrize("k,check", [(k, c) for k, (_, cs) in MUTANTS.items() for c in cs])
def test_pin_fails_on_planted_mistake(oracle, mutant_libs, k, check):
    with oracle.use(mutant_libs[k]):
        with pytest.raises(AssertionError):
            CHECKS[check](oracle)


@pytest.mark.parametrize("check", sorted(CHECKS))
def test_pin_passes_on_real_oracle(oracle, check):
    CHECKS[check](oracle)
