"""The oracle's pinning checks catch plausible mistakes. CPU only.

Each case plants one mistake in a copy of oracle/felis_oracle.c (a text
substitution: a dropped or doubled factor, a flipped sign, swapped
coefficients, a quirk of the reference "fixed"), builds it with the oracle's
own flags into a temporary directory and runs the pinning test that must
catch it against that build. The planted mistake lives only in the temporary
copy; the oracle itself is untouched.
"""
from __future__ import annotations

import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle"


def _flags() -> list[str]:
    """ORCFLAGS from oracle/Makefile (continuation lines joined)."""
    text = (ORACLE / "Makefile").read_text().replace("\\\n", " ")
    m = re.search(r"^ORCFLAGS\s*=\s*(.*)$", text, re.M)
    return m.group(1).split()


# name: (text in felis_oracle.c, replacement, the pinning test that must fail[, which
# occurrence of the text: 0 when omitted])
MUTATIONS = {
    "adi_last_row_fixed": ("CCv[nx - 2] = H - alP[nx - 2] - al[nx - 2] - 1.0 / CCv[nx - 3];",
                           "CCv[nx - 2] = H - alP[nx - 2] - al[nx - 2] - al[nx - 2] / CCv[nx - 3];",
                           "test_adi_equals_dense_solves_of_the_row_equations"),
    "adi_explicit_sign": ("(H + beP[j] + be[j]) * Un[idx]", "(H - beP[j] - be[j]) * Un[idx]",
                          "test_adi_gaussian_beam_diffraction"),
    "adi_alpha_for_alphaP": ("dd[j] = -alP[i] * Uc[idx + 1]", "dd[j] = -al[i] * Uc[idx + 1]",
                             "test_adi_equals_dense_solves_of_the_row_equations"),
    "adi_source_once": ("S[slot * N + (size_t)j * nx + i] * cf;", "0.0 * cf;",
                        "test_deposit_and_push_conserve_energy"),
    "deposit_push_coupling": ("fx = sign * (J1 - K0_alpha * J2);", "fx = sign * (K0_alpha * J1 - J2);",
                              "test_deposit_and_push_conserve_energy"),
    "deposit_coef_half": ("(4.0 * ks * K_EMASS", "(2.0 * ks * K_EMASS", "test_deposit_single_particle_corners"),
    "deposit_phase": ("cexp(-I * dbH * (theta + Phi * 0.5)) / gam", "cexp(-I * (dbH * theta + Phi * 0.5)) / gam",
                      "test_deposit_single_particle_corners"),
    "deposit_even_unswapped": (
        "fx = sign * xi2 * dbH * 0.5 * (expM * (J1 - K0_alpha * J2) + expP * (J2 - K0_alpha * J3));",
        "fx = sign * xi2 * dbH * 0.5 * (expP * (J1 - K0_alpha * J2) + expM * (J2 - K0_alpha * J3));",
        "test_deposit_and_push_conserve_energy"),
    "push_energy_factor": ("sumG += creal(t) / (2.0 * gam);", "sumG += creal(t) / gam;",
                           "test_push_madey_low_gain_curve"),
    "push_k2_radius": ("(1.0 + ku * ku * 0.5 * r2)", "(1.0 + ku * ku * r2)", "test_push_zero_field_closed_form"),
    "push_rk4_weights": ("double dth = dz / 6.0 * (kth[1] + 2.0 * kth[2] + 2.0 * kth[3] + kth[4]);",
                         "double dth = dz / 6.0 * (kth[1] + 2.0 * kth[2] + kth[3] + 2.0 * kth[4]);",
                         "test_push_rk4_step_against_adaptive_integration"),
    "push_phase_term_sign": ("sumTh += creal(I * t) / (1.0 * H);", "sumTh -= creal(I * t) / (1.0 * H);",
                             "test_push_madey_low_gain_curve"),
    "push_axis_em_weighted": (
        "Em[l] += CC_(st->Ez)[l * es + (size_t)slot * nr * F + nr * f + idxR + jj];",
        "Em[l] += 0.5 * CC_(st->Ez)[l * es + (size_t)slot * nr * F + nr * f + idxR + jj];",
        "test_push_rk4_step_against_adaptive_integration"),
    "ez_slot0_read_fixed": ("(d[j] - a[j] * ez[f * nr + j + 1]) / bb[j]",
                            "(d[j] - a[j] * ez[(size_t)slot * F * nr + f * nr + j + 1]) / bb[j]",
                            "test_radial_solve_against_dense_solve"),
    "ez_axis_row": ("bb[0] = -(1.0 + (l + 1.0) * (l + 1.0) * A / 8.0);",
                    "bb[0] = -(1.0 + (l + 1.0) * (l + 1.0) * A / 4.0);", "test_radial_solve_against_dense_solve"),
    "transverse_coupling_sign": ("ky[m] = py / gam + aa * (K0_alpha * x - y) / (gam * gam);",
                                 "ky[m] = py / gam + aa * (K0_alpha * x + y) / (gam * gam);",
                                 "test_coupled_transverse_step_is_the_taylor_polynomial"),
    "transverse_dd": ("double dd = ku * K0 * K0 / (4.0 * gam * gam) * ue;",
                      "double dd = ku * K0 * K0 / (2.0 * gam * gam) * ue;",
                      "test_coupled_transverse_step_is_the_taylor_polynomial"),
    "drift_particle_gamma": ("double t = ku - ks / (2.0 * gamR * gamR) * (K2 + sumU2);",
                             "double t = ku - ks / (2.0 * st->gamma[n] * st->gamma[n]) * (K2 + sumU2);",
                             "test_drift_closed_forms"),
    "drift_wake_on_gamma": ("st->theta[n] -= dz * wakeE;", "st->gamma[n] -= dz * wakeE;", "test_drift_closed_forms"),
    "abc_profile_shift": ("sx[i] = abcSig * (i - nx + L) * (i - nx + L) / (1.0 * L * L);",
                          "sx[i] = abcSig * (i - nx + L + 1) * (i - nx + L + 1) / (1.0 * L * L);",
                          "test_absorbing_layer_profile_closed_form"),
    "em_reset_per_slice": ("if (seg == 0 || nt > 1)", "if (1)", "test_push_carries_em_past_the_radial_grid"),
    "em_not_carried": ("ny - 1)) continue;\n        double r = sqrt(r2);",
                       "ny - 1)) continue;\n        for (int l = 0; l < L; ++l) Em[l] = 0;\n        double r = sqrt(r2);",
                       "test_push_carries_em_past_the_radial_grid"),
    "abc_amplifying": ("cx[i] = 1.0 + I * sx[i];", "cx[i] = 1.0 - I * sx[i];",
                       "test_absorbing_layer_absorbs_a_beam_and_never_amplifies"),
    # the 1D push and source (the second occurrence of a line the 3D path shares)
    "push1d_energy_factor": ("sumG += creal(t) / (2.0 * gam);", "sumG += creal(t) / gam;",
                             "test_oracle_1d.py::test_push_madey_low_gain_curve_1d", 1),
    "push1d_phase_term_sign": ("sumTh += creal(I * t) / (1.0 * H);", "sumTh -= creal(I * t) / (1.0 * H);",
                               "test_oracle_1d.py::test_push_madey_low_gain_curve_1d", 1),
    "push1d_ez_factor": ("sumEz += 2.0 * creal(Em[l] * cexp(I * (l + 1.0) * th));",
                         "sumEz += creal(Em[l] * cexp(I * (l + 1.0) * th));",
                         "test_oracle_1d.py::test_space_charge_plasma_oscillation_1d", 1),
    "source1d_coupling": ("fx = sign * (J1 - K0_alpha * J2);", "fx = sign * (K0_alpha * J1 - J2);",
                          "test_oracle_1d.py::test_source_and_push_conserve_energy_1d", 1),
    "source1d_coef_half": ("(2.0 * ks * K_EMASS * (p->lambda0 * p->num_slice) * area)",
                           "(4.0 * ks * K_EMASS * (p->lambda0 * p->num_slice) * area)",
                           "test_oracle_1d.py::test_source_and_push_conserve_energy_1d"),
    "ez1d_sign": ("+= -1.0 * I * coef / (1.0 + l) * macro", "+= 1.0 * I * coef / (1.0 + l) * macro",
                  "test_oracle_1d.py::test_space_charge_plasma_oscillation_1d"),
}


def _replace_nth(src: str, old: str, new: str, nth: int) -> str:
    """src with the nth (0-based) occurrence of old replaced"""
    at = -1
    for _ in range(nth + 1):
        at = src.find(old, at + 1)
        assert at >= 0, f"occurrence {nth} of the text to mutate is gone from felis_oracle.c"
    return src[:at] + new + src[at + len(old):]

_RUNNER = """
import sys
from pathlib import Path
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
from oracle import pyoracle as po
po.LIB_PATH = Path({lib!r})
import pytest
sys.exit(pytest.main(["-q", "-x", "-p", "no:cacheprovider", "-W", "ignore", {test!r}]))
"""


def _run_pinning_test(lib: Path, test: str) -> subprocess.CompletedProcess:
    """test: a test of test_oracle_pinning.py, or file::test for another file"""
    node = f"{ROOT / 'tests' / test}" if "::" in test else f"{ROOT / 'tests' / 'test_oracle_pinning.py'}::{test}"
    code = _RUNNER.format(root=str(ROOT), tests=str(ROOT / "tests"), lib=str(lib), test=node)
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=str(ROOT), timeout=300)


def _build(d: Path) -> Path:
    """oracle/Makefile's recipe: compile with ORCFLAGS, link without them."""
    obj, lib = d / "felis_oracle.o", d / "liboracle.so"
    subprocess.run(["gcc", *_flags(), "-c", "-o", str(obj), str(d / "felis_oracle.c")], check=True)
    subprocess.run(["gcc", "-shared", "-fopenmp", "-o", str(lib), str(obj), "-lm"], check=True)
    return lib


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_pinning_catches_planted_mistake(name, tmp_path):
    old, new, test, *nth = MUTATIONS[name]
    src = (ORACLE / "felis_oracle.c").read_text()
    assert old in src, f"{name}: the text to mutate is gone from felis_oracle.c"
    (tmp_path / "felis_oracle.c").write_text(_replace_nth(src, old, new, nth[0] if nth else 0))
    (tmp_path / "felis_oracle.h").write_text((ORACLE / "felis_oracle.h").read_text())
    lib = _build(tmp_path)
    r = _run_pinning_test(lib, test)
    assert r.returncode == 1, f"{name}: {test} did not fail\n{r.stdout[-2000:]}"


def test_unmutated_copy_passes(tmp_path):
    """The same build recipe on the unchanged source passes every check the
    mutations are run against."""
    (tmp_path / "felis_oracle.c").write_text((ORACLE / "felis_oracle.c").read_text())
    (tmp_path / "felis_oracle.h").write_text((ORACLE / "felis_oracle.h").read_text())
    lib = _build(tmp_path)
    for test in sorted({m[2] for m in MUTATIONS.values()}):
        r = _run_pinning_test(lib, test)
        assert r.returncode == 0, f"{test} fails on the unmutated oracle\n{r.stdout[-2000:]}"
