"""The oracle's pins catch plausible mistakes.

Each case below plants one plausible mistake in a copy of the oracle's source
(oracle/grasph_oracle.cpp), compiles that copy, and runs the closed-form
functor pins (test_oracle_functors.py) against it in a child process: every
mutant must make at least one pin fail.  The unmodified source must pass the
same run, so a failure here is the planted mistake and nothing else.

The mistakes are the ones a restatement of src/PairwisePhysics.jl and
src/PairwiseFunctors.jl is likely to make: the textbook mean-density
viscosity denominator instead of the reference's product 0.5·ρi·ρj
(:29-34), h² instead of 0.01·h², one particle's sound speed instead of the
mean, a transposed Voigt index in the Cauchy force (:125-151) or the strain
rate (:169-195), a flipped spin sign (:210-220), a dropped factor or sign in
δ-SPH (:238-243), a missing |p| in the surface tension (:286-290), a wrong
Lennard-Jones scale (:258-263), and the dynamic wall's β cap, mass and
partner density (PairwiseFunctors.jl:222-242).
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "grasph_oracle.cpp"

# (name, text in the source, replacement, occurrences replaced: the first n, or
# None for all)
MUTANTS = [
    ("visc_mean_density", "(0.5 * ri * rj)", "(0.5 * (ri + rj))", 1),
    ("visc_softening_h2", "h * h * 0.01", "h * h", 1),
    ("visc_sound_speed_ci", "double mc = 0.5 * (ci + cj);", "double mc = ci;", 1),
    ("visc_no_min", "jmin(dotv(dx, dv, D), 0.0)", "dotv(dx, dv, D)", 1),
    ("pressure_s2_index", "-(pi / (ri * ri) + pj / (rj * rj))",
     "-(pi / (ri * ri) + pj / (ri * rj))", 1),
    ("continuity_s1_ratio", "return S == 1 ? ri / rj : 1.0;", "return S == 1 ? rj / ri : 1.0;", 1),
    ("delta_no_factor_2", "return 2.0 * delta * c * h", "return delta * c * h", 1),
    ("delta_sign", "(ri - rj) * dotv(dx, gx, D) / rr", "(rj - ri) * dotv(dx, gx, D) / rr", 1),
    ("delta_own_density", "P.Aw->drhodt[i] += m * (dr * cdc(ri, rj, f.sigma) + psi / rj);",
     "P.Aw->drhodt[i] += m * (dr * cdc(ri, rj, f.sigma) + psi / ri);", 1),
    ("surface_tension_no_abs", "pfc(std::fabs(pi), std::fabs(pj), ri, rj, 1)",
     "pfc(pi, pj, ri, rj, 1)", 1),
    ("lj_no_001", "return 0.01 * c * c * f;", "return c * c * f;", 1),
    ("lj_over_r", "jl_pow_int(cutoff / r, p2)) / rr;", "jl_pow_int(cutoff / r, p2)) / r;", 1),
    ("cauchy3_index", "out[1] = (si[2] * g[0] + si[1] * g[1]) * ci",
     "out[1] = (si[1] * g[0] + si[2] * g[1]) * ci", 1),
    ("cauchy4_index", "out[0] = (si[0] * g[0] + si[3] * g[1]) * ci",
     "out[0] = (si[0] * g[0] + si[2] * g[1]) * ci", 1),
    ("cauchy6_index", "out[2] = (si[4] * g[0] + si[5] * g[1] + si[2] * g[2]) * ci",
     "out[2] = (si[5] * g[0] + si[4] * g[1] + si[2] * g[2]) * ci", 1),
    ("strain_rate_xy", "out[3] = 0.5 * (dv[0] * g[1] + dv[1] * g[0]);",
     "out[3] = (dv[0] * g[1] + dv[1] * g[0]);", 1),
    ("spin_3d_sign", "out[1] = 0.5 * (dv[0] * g[2] - dv[2] * g[0]);",
     "out[1] = 0.5 * (dv[2] * g[0] - dv[0] * g[2]);", 1),
    ("dynamic_no_beta_cap", "-jmin(P.beta, std::fabs(db / da))", "-std::fabs(db / da)", None),
    ("dynamic_wall_mass", "double rj = ri, pj = pi, m = P.mass_a;",
     "double rj = ri, pj = pi, m = P.B.mass;", 1),
]


def _build(src_text, out):
    cpp = out.with_suffix(".cpp")
    cpp.write_text(src_text)
    # the oracle's header is included by a path relative to the source
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
           f"-I{ROOT / 'oracle'}", f"-I{ROOT / 'include'}", "-shared", "-o", str(out), str(cpp)]
    subprocess.run(cmd, check=True, capture_output=True)


def _pins_pass(lib):
    env = dict(os.environ, GRASPH_ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        str(ROOT / "tests" / "test_oracle_functors.py")],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode in (0, 1), r.stdout[-2000:] + r.stderr[-2000:]
    return r.returncode == 0


@pytest.fixture(scope="module")
def source():
    text = SRC.read_text()
    # a relative include of the header must still resolve from a temporary copy
    return text.replace('#include "grasph_oracle.h"', f'#include "{ROOT / "oracle" / "grasph_oracle.h"}"')


def test_unmodified_copy_passes(source, tmp_path):
    lib = tmp_path / "orc_plain.so"
    _build(source, lib)
    assert _pins_pass(lib)


@pytest.mark.parametrize("name,old,new,count", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_fails_a_pin(source, tmp_path, name, old, new, count):
    n = source.count(old)
    assert n >= (count or 1), f"{name}: source text changed ({n} matches)"
    lib = tmp_path / f"orc_{name}.so"
    _build(source.replace(old, new, -1 if count is None else count), lib)
    assert not _pins_pass(lib), f"planted mistake {name} passed every pin"
