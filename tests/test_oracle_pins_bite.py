"""The oracle's pins catch plausible mistakes: each planted mistake of
oracle/lbmc_oracle.c (ORC_MUTANT = k, compiled into a scratch directory on
the CPU) must make the pin written for that function fail, run in a child
process with LBMC_ORACLE_LIB pointing at the mutant."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "lbmc_oracle.c")

# mutant -> (what is planted, the pin that must catch it)
MUTANTS = {
    1: ("equilibrium without the second-order term", "test_oracle_operators.py::test_equilibrium_moments"),
    2: ("Guo source factor 9 -> 3", "test_oracle_operators.py::test_guo_source_moments"),
    3: ("Smagorinsky factor 2 sqrt(2) -> sqrt(2)", "test_oracle_operators.py::test_smagorinsky_closure"),
    4: ("D2Q9 Zou-He sign of (f2 - f4)", "test_oracle_operators.py::test_zou_he_d2q9_closure"),
    5: ("bounceback without reflection", "test_oracle_operators.py::test_bounceback_and_ntpor_limits"),
    6: ("Guo velocity shift F/rho instead of F/2rho", "test_oracle_operators.py::test_guo_source_moments"),
    7: ("D2Q9 moment equilibrium sign of rho", "test_oracle_operators.py::test_mrt_conserves_and_reduces_to_bgk"),
    8: ("index shift x + t e instead of x - t e", "test_oracle_lattice.py::test_streaming_moves_population_by_e"),
    9: ("D3Q15 Zou-He face x: transverse signs of two diagonals swapped",
        "test_oracle_operators.py::test_zou_he_d3q15_reference_defects"),
    10: ("D3Q15 Zou-He face Z: rho-u terms from u[2] instead of the reference's u[0]",
         "test_oracle_operators.py::test_zou_he_d3q15_reference_defects"),
    11: ("D3Q15 Zou-He face X: one diagonal's second transverse sign flipped",
         "test_oracle_operators.py::test_zou_he_d3q15_reference_defects"),
}


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_pin_fails_on_planted_mistake(k, tmp_path):
    what, pin = MUTANTS[k]
    lib = tmp_path / f"liblbmc_oracle_m{k}.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fopenmp", "-shared",
                    f"-DORC_MUTANT={k}", "-o", str(lib), SRC, "-lm"], check=True)
    env = dict(os.environ, LBMC_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", pin)], capture_output=True, text=True, env=env,
                       cwd=ROOT, timeout=300)
    assert r.returncode == 1, f"mutant {k} ({what}) not caught by {pin}:\n{r.stdout[-2000:]}"
    # and the unmutated oracle passes the same pin
    r0 = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                         os.path.join(ROOT, "tests", pin)], capture_output=True, text=True,
                        env=dict(os.environ, OMP_NUM_THREADS="2"), cwd=ROOT, timeout=300)
    assert r0.returncode == 0, r0.stdout[-2000:]
