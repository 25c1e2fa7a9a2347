"""The particle oracle's pins can tell: each plausible mistake planted in a
copy of oracle/particles_oracle.c (built into a temporary directory, CPU
only) makes tests/test_particles_oracle.py fail, while the unmodified copy
passes."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

MUTATIONS = {
    # Bouzidi q < 1/2: the two interpolation points swapped
    "bouzidi_below_half_swapped": (
        "gval = (2.0 * q) * fi + (1.0 - 2.0 * q) * fb - 6.0 * w * cu;",
        "gval = (2.0 * q) * fb + (1.0 - 2.0 * q) * fi - 6.0 * w * cu;"),
    # Bouzidi q >= 1/2: the blend weights swapped
    "bouzidi_above_half_swapped": (
        "gval = (1.0 / (2.0 * q)) * fi + ((2.0 * q - 1.0) / (2.0 * q)) * fib - (3.0 / q) * w * cu;",
        "gval = (1.0 / (2.0 * q)) * fib + ((2.0 * q - 1.0) / (2.0 * q)) * fi - (3.0 / q) * w * cu;"),
    # fallback: moving-wall term with the wrong sign
    "fallback_wall_sign": ("gval = fi - 6.0 * w * cu;", "gval = fi + 6.0 * w * cu;"),
    # Wen et al.: the wall-velocity terms exchanged
    "wen_sign": ("const double px = ff * (cx - uwx) + gf * (cx + uwx);",
                 "const double px = ff * (cx + uwx) + gf * (cx - uwx);"),
    # Feng & Michaelides: the overlap stiffness taken from the wrong constant
    "contact_stiffness": ("m = m + stf1 * ((-h) / zeta);", "m = m + stf0 * ((-h) / zeta);"),
    # lubrication: the pair radius R/2 used for the wall
    "lubrication_wall_radius": ("lub_coef(dlo - g->R, g->R, g)", "lub_coef(dlo - g->R, 0.5 * g->R, g)"),
    # refill: a wrong neighbour count
    "refill_density": ("const double rho = cnt > 0 ? rsum / (double)cnt : 0.0;",
                       "const double rho = cnt > 0 ? rsum / (double)(cnt + 1) : 0.0;"),
    # integration: no two-step force average
    "no_force_average": ("const double fh = first ? s->FH[k] : 0.5 * (s->FH[k] + s->FHprev[k]);",
                         "const double fh = s->FH[k];"),
}


def _build(tmp, src):
    for name in ("lbm_oracle.c", "lbm_oracle.h", "particles_oracle.h"):
        shutil.copy(os.path.join(ORACLE, name), tmp)
    with open(os.path.join(tmp, "particles_oracle.c"), "w") as fh:
        fh.write(src)
    lib = os.path.join(tmp, "liblbm_oracle.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c99", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
                    os.path.join(tmp, "lbm_oracle.c"), os.path.join(tmp, "particles_oracle.c"), "-o", lib, "-lm"],
                   check=True)
    return lib


def _pins_pass(lib):
    env = dict(os.environ, D3Q19_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_particles_oracle.py"),
                        "-q", "-x", "-p", "no:cacheprovider"], env=env, cwd=ROOT, capture_output=True, text=True)
    return r.returncode == 0, r.stdout[-400:]


@pytest.mark.parametrize("name", ["unmodified"] + sorted(MUTATIONS))
def test_planted_mistake_fails_a_pin(name, tmp_path):
    src = open(os.path.join(ORACLE, "particles_oracle.c")).read()
    if name != "unmodified":
        old, new = MUTATIONS[name]
        assert src.count(old) == 1, f"{name}: the mutated line is no longer in the oracle"
        src = src.replace(old, new)
    ok, tail = _pins_pass(_build(str(tmp_path), src))
    assert ok == (name == "unmodified"), tail
