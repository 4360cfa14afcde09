"""The boundary pins catch plausible restatement mistakes.

Each case plants one mistake in a temporary copy of oracle/wake_oracle.c
(CPU only), builds it as the FP64 oracle, and checks that at least one pin of
tests/test_oracle_boundary_pins.py fails against it: swapped populations,
a wrong neighbour, a wrong wall coefficient or sign in the curved-wall rule;
a sign, population or radius-filter mistake in the 3D momentum exchange;
swapped weights or the wrong upstream column in the convective outlet; a
dropped off-diagonal stress term in the D2Q9 Smagorinsky tau.
"""
import os
import shutil
import subprocess

import pytest

import oracle_bind as ob
import test_oracle_boundary_pins as pins

SRC = os.path.join(ob.ORACLE_DIR, "wake_oracle.c")

MUTANTS = {
    "bfl_swap_populations": ("reflected = inv_2delta * f_i_xf + (R(1) - inv_2delta) * f_opp_xf - mom;",
                             "reflected = inv_2delta * f_opp_xf + (R(1) - inv_2delta) * f_i_xf - mom;"),
    "bfl_wrong_neighbour": ("int x_ff = (x_f - cx + nx) % nx;\n      int y_ff = y_f - cy, z_ff = z_f - cz;",
                            "int x_ff = (x_f + cx + nx) % nx;\n      int y_ff = y_f + cy, z_ff = z_f + cz;"),
    "bfl_wall_coefficient": ("real mom = (R(1) / delta) * L->w[q]", "real mom = R(2) * L->w[q]"),
    "bfl_wall_sign": ("reflected = R(2) * delta * f_i_xf + (R(1) - R(2) * delta) * f_i_xff - mom;",
                      "reflected = R(2) * delta * f_i_xf + (R(1) - R(2) * delta) * f_i_xff + mom;"),
    "mem3d_sign": ("            sfz += (double)(R(2) * fv * (real)cz);",
                   "            sfz -= (double)(R(2) * fv * (real)cz);"),
    "mem3d_opposite_population": (
        "real fv = fq[idx];\n            sfx += (double)(R(2) * fv * (real)cx);\n"
        "            sfy += (double)(R(2) * fv * (real)cy);\n            sfz",
        "real fv = f[L->opp[q] * total + idx];\n            sfx += (double)(R(2) * fv * (real)cx);\n"
        "            sfy += (double)(R(2) * fv * (real)cy);\n            sfz"),
    "mem3d_no_radius_filter": (
        "if (max_radius <= R(0) || (rx * rx + ry * ry + rz * rz <= r_max_sq)) {\n            real fv",
        "if (1) {\n            real fv"),
    "convective_swapped_weights": ("beta * f_prev[q * total + io] + alpha * f[q * total + ii]",
                                   "alpha * f_prev[q * total + io] + beta * f[q * total + ii]"),
    "convective_upstream_column": ("ii = (size_t)y * nx + (nx - 2);", "ii = (size_t)y * nx + (nx - 1);"),
    "les2d_dropped_shear_term": ("real tensor_sum = pxx * pxx + R(2) * pxy * pxy + pyy * pyy;",
                                 "real tensor_sum = pxx * pxx + pxy * pxy + pyy * pyy;"),
}

PIN_CALLS = [
    *[(pins.test_bfl_reproduces_uniform_equilibrium_of_the_wall_velocity, (q, d))
      for q in pins.LINK_QS for d in pins.DELTAS],
    *[(pins.test_bfl_interpolates_a_linear_field_at_the_departure_point, (q, d))
      for q in pins.LINK_QS for d in pins.DELTAS],
    (pins.test_bfl_limits_and_upstream_neighbour, ()),
    (pins.test_mem3d_single_population_impulses, ()),
    (pins.test_convective_outlet_is_implicit_upwind_advection, ()),
    (pins.test_d2q9_smagorinsky_tau_equilibrium_and_brute_force, ()),
]


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_planted_mistake(name, tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is needed to build the mutant oracle")
    before, after = MUTANTS[name]
    src = open(SRC).read()
    assert before in src, f"mutation anchor for {name} not found in the oracle"
    (tmp_path / "wake_oracle.c").write_text(src.replace(before, after))
    shutil.copy(os.path.join(ob.ORACLE_DIR, "wake_oracle.h"), tmp_path / "wake_oracle.h")
    so = tmp_path / "liboracle64.so"
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c11", "-DORACLE_REAL=double",
                    "-DORACLE_DOUBLE", "-shared", "-o", str(so), str(tmp_path / "wake_oracle.c"), "-lm"], check=True)
    saved_dir, saved_lib = ob.ORACLE_DIR, ob._libs.pop(64, None)
    try:
        ob.ORACLE_DIR = str(tmp_path)
        ob.lib(64)
        caught = 0
        for fn, args in PIN_CALLS:
            try:
                fn(*args)
            except AssertionError:
                caught += 1
        assert caught > 0, f"no pin fails with the planted mistake {name}"
    finally:
        ob._libs.pop(64, None)
        ob.ORACLE_DIR = saved_dir
        if saved_lib is not None:
            ob._libs[64] = saved_lib
