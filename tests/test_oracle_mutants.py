"""Mutation tests of the oracle pins.

Each case plants one plausible mistake -- a swapped weight, a flipped sign,
a wrong index or offset -- in a private copy of oracle/rh3t_oracle.c,
compiles it (same flags as oracle/Makefile), routes the oracle bridge
through that build, and asserts that the pin named for it in
tests/test_oracle_pins.py fails.  The unmutated oracle passes every one of
those pins (test_oracle_pins.py itself).  Nothing here touches the GPU.
"""
import os
import shutil
import subprocess

import pytest

import test_oracle_pins as pins

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
CFLAGS = ["-std=c99", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared", "-w"]

# name: (text in rh3t_oracle.c, planted replacement, pin that must fail)
MUTANTS = {
    "rk3_stage2_weights": ("dt / 4.0, 0.75, 0.25, 2", "dt / 4.0, 0.25, 0.75, 2",
                           pins.check_rk3_exchange_amplification),
    "rk3_stage3_weights": ("2.0 * dt / 3.0, 1.0 / 3.0, 2.0 / 3.0", "2.0 * dt / 3.0, 2.0 / 3.0, 1.0 / 3.0",
                           pins.check_rk3_exchange_amplification),
    "farfield_incoming_sign": ("rhs[m] = lam[m] < 0.0 ? vinf[m] : v[m];", "rhs[m] = lam[m] > 0.0 ? vinf[m] : v[m];",
                               pins.check_farfield_characteristics),
    "farfield_derivatives_kept": ("rhs[m] = lam[m] < 0.0 ? 0.0 : v[6 * k + m];", "rhs[m] = v[6 * k + m];",
                                  pins.check_farfield_characteristics),
    "outflow_relaxation_sign": ("(p0 - p_inf)", "(p_inf - p0)", pins.check_outflow_lodi),
    "outflow_mach_factor": ("(1.0 - powi(m_local, 2))", "(1.0 + powi(m_local, 2))", pins.check_outflow_lodi),
    "ilw_row1_tangential_sign": ("L1[2] = gt * u3 / (6.0 * u1);", "L1[2] = -gt * u3 / (6.0 * u1);",
                                 pins.check_wall_ilw),
    "ilw_row0_column": ("L0[c] = c == 1 ? 1.0 : 0.0;", "L0[c] = c == 2 ? 1.0 : 0.0;", pins.check_wall_ilw),
    "taylor_coefficient": ("coef *= dist / (double)k;", "coef *= dist / (double)(k + 1);",
                           pins.check_taylor_rotation),
    "rotation_sign": ("out[1] = mom_n * n[0] - mom_t * n[1];", "out[1] = mom_n * n[0] + mom_t * n[1];",
                      pins.check_taylor_rotation),
    "dt_exchange_denominator": ("double sd = 2.0 * A * powi(CVE, 4) * CVI * rho;",
                                "double sd = 2.0 * A * powi(CVE, 3) * CVI * rho;", pins.check_local_dt_exchange),
    "dt_exchange_discriminant": ("4.0 * A * OEI * OER * powi(CVE, 4) * CVI *",
                                 "4.0 * A * OEI * OER * powi(CVE, 4) * CVE *", pins.check_local_dt_exchange),
    "s_table_offset": ("kSOff[6] = {0, 1, 17, 98, 354, 979};", "kSOff[6] = {0, 1, 17, 98, 353, 979};",
                       pins.check_precomputed_matches_per_call),
    "td_stage3_time": ("ps.time = t0 + 0.5 * dt;", "ps.time = t0 + dt;", pins.check_time_dependent_front),
    "td_front_side": ("pb->g_p0x[g] <= x_front ?", "pb->g_p0x[g] >= x_front ?", pins.check_time_dependent_front),
    "w_table_offset": ("kWOff[6] = {0, 1, 9, 36, 100, 225};", "kWOff[6] = {0, 1, 9, 36, 99, 225};",
                       pins.check_precomputed_matches_per_call),
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    shutil.copy(os.path.join(ORACLE_DIR, "rh3t_oracle.h"), d)
    return d


def _build(mutant_dir, name, old, new):
    src = open(os.path.join(ORACLE_DIR, "rh3t_oracle.c")).read()
    assert src.count(old) == 1, f"{name}: the mutated text must occur exactly once"
    c = os.path.join(mutant_dir, f"{name}.c")
    with open(c, "w") as fh:
        fh.write(src.replace(old, new))
    so = os.path.join(mutant_dir, f"lib{name}.so")
    subprocess.run([os.environ.get("CC", "gcc"), *CFLAGS, "-I", str(mutant_dir), "-o", so, c, "-lm"], check=True)
    return so


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pin_catches_mutant(name, mutant_dir, rh3t, workloads, oracle):
    old, new, check = MUTANTS[name]
    lib = oracle.load_variant(_build(mutant_dir, name, old, new))
    with oracle.using(lib):
        with pytest.raises((AssertionError, oracle.OracleError)):
            check(rh3t, workloads, oracle)
