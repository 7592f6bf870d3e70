"""The oracle's pins catch plausible transcription mistakes.

Each case plants one mistake in a temporary copy of oracle/phd_oracle.c (a
dropped or flipped term, a wrong constant, a swapped operand), builds it with
the oracle's own flags, and runs the pin tests against that build through
PHD_ORACLE_LIB: at least one pin must fail.  The mistakes are the kind a
restatement of riemann.pyx / reconstruction.pyx / mesh.pyx could make; the
first four are the ones an earlier pin set let through (HLLC star energy
term, the Q = 2 switch of the wave estimates, the GIZMO psi2 and lever arms).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "phd_oracle.c")
PINS = ["tests/test_oracle_riemann.py", "tests/test_oracle_cells.py",
        "tests/test_oracle_geometry.py"]

# name: [(text in phd_oracle.c, replacement), ...]  (riemann.pyx / reconstruction.pyx lines)
MUTATIONS = {
    # HLLC star energy, riemann.pyx:649,667
    "hllc_energy_sign": [("(s_contact + _pl/(_dl*(sl - Vnl)))", "(s_contact - _pl/(_dl*(sl - Vnl)))"),
                         ("(s_contact + _pr/(_dr*(sr - Vnr)))", "(s_contact - _pr/(_dr*(sr - Vnr)))")],
    "hllc_energy_sign_left": [("(s_contact + _pl/(_dl*(sl - Vnl)))", "(s_contact - _pl/(_dl*(sl - Vnl)))")],
    # get_waves switch, riemann.pyx:443
    "waves_q4": [("const double Q = 2., c_floor", "const double Q = 4., c_floor")],
    "waves_q_non_strict": [("if (((p_max/p_min) < Q)", "if (((p_max/p_min) <= Q)")],
    # GIZMO limiter, reconstruction.pyx:768, 889-890
    "gizmo_psi2": [("psi1 = 0.5, psi2 = 0.25;\n    double state_l", "psi1 = 0.5, psi2 = 0.4;\n    double state_l")],
    "gizmo_psi1": [("psi1 = 0.5, psi2 = 0.25;\n    double state_l", "psi1 = 0.6, psi2 = 0.25;\n    double state_l")],
    "gizmo_lever_arms": [("state_l[f] + sepi_mag/diff_mag", "state_l[f] + sepj_mag/diff_mag"),
                         ("state_r[f] + sepj_mag/diff_mag", "state_r[f] + sepi_mag/diff_mag")],
    "gizmo_branch": [("if (prim[f][i] < prim[f][j]) {", "if (prim[f][i] > prim[f][j]) {")],
    # wave estimates, riemann.pyx:437-502
    "tsrs_sign": [("p_star = (gl*pl + gr*pr - (ur - ul))/(gl + gr);",
                   "p_star = (gl*pl + gr*pr + (ur - ul))/(gl + gr);")],
    "trrs_exponent": [("double plr = pow(pl/pr, z);", "double plr = pow(pl/pr, 1./z);")],
    "pvrs_average": [("double c_avg = .5*(cl + cr);", "double c_avg = (cl + cr);")],
    "shock_speed_factor": [("_sl = ul - cl*sqrt(1.+((gamma+1.)/(2.*gamma))",
                            "_sl = ul - cl*sqrt(1.+((gamma-1.)/(2.*gamma))")],
    "contact_speed_sign": [("*sc = (pr - pl + dl*ul", "*sc = (pl - pr + dl*ul")],
    # HLLC momentum and energy, riemann.pyx:640-672
    "hllc_pressure_term": [("double factor_2 = factor_1*(sl - wn)*(s_contact - Vnl) + _pl;",
                            "double factor_2 = factor_1*(sl - wn)*(s_contact - Vnl) - _pl;")],
    "hllc_energy_velocity": [("(sl - wn)*factor_1*(el/_dl + (s_contact - Vnl)*",
                              "(sl - wn)*factor_1*(el/_dl + (s_contact + Vnl)*")],
    # gradients, reconstruction.pyx:677-681
    "gradient_half": [("area*(d_dif*cfx[k] - 0.5*d_sum*dr[k])", "area*(d_dif*cfx[k] - d_sum*dr[k])")],
    # temporal predictor, reconstruction.pyx:1040-1060
    "temporal_gamma": [("pl[dim+1][m] -= dt*(gamma*prim[dim+1][i]", "pl[dim+1][m] -= dt*(prim[dim+1][i]"),
                       ("pr[dim+1][m] -= dt*(gamma*prim[dim+1][j]", "pr[dim+1][m] -= dt*(prim[dim+1][j]")],
    # update, mesh.pyx:529-544
    "update_momentum_sign": [("cons[1+k][j] += dt*a*flux[1+k][n];", "cons[1+k][j] -= dt*a*flux[1+k][n];")],
    # exact solver fan, riemann.pyx:800-830
    "exact_fan": [("double c = (2.0/(gamma + 1.0))*(cl + 0.5*(gamma - 1.0)*vnl);",
                   "double c = (2.0/(gamma + 1.0))*(cl - 0.5*(gamma - 1.0)*vnl);")],
    # static-mesh epilogue, boundary.pyx:336-348
    "ghost_refresh_momentum": [("if (x[k][i] < lo[k]) { prim[1+k][i] *= -1; cons[1+k][i] *= -1; }",
                                "if (x[k][i] < lo[k]) { prim[1+k][i] *= -1; }")],
    # face velocities, mesh.pyx:470-477
    "face_velocity_sign": [("fv[k][m] = 0.5*(wx[k][i] + wx[k][j]) + factor*(x[k][j] - x[k][i]);",
                            "fv[k][m] = 0.5*(wx[k][i] + wx[k][j]) - factor*(x[k][j] - x[k][i]);")],
    # geometry extraction, tess.cpp:313-355, tess3.cpp:452-559, boundary.pyx
    "circumcentre_3d_operands": [("p[k] = (aa*bc[k] + bb*ca[k] + cc*ab[k])*r;",
                                  "p[k] = (aa*bc[k] + bb*ab[k] + cc*ca[k])*r;")],
    "circumcentre_2d_sign": [("p[1] = (bb*a[0] - aa*b[0])*r;", "p[1] = (aa*b[0] - bb*a[0])*r;")],
    "vertex_mean_divisor": [("for (k = 0; k < 3; k++) c[k] *= 1.0/m;",
                             "for (k = 0; k < 3; k++) c[k] *= 1.0/(m + 1);")],
    "fan_centroid_third": [("s[k] += (c[k] + (v1[k] + v2[k])*third)*a3;",
                            "s[k] += (c[k] + (v1[k] + v2[k])*fourth)*a3;")],
    "face_area_half": [("area[f] = 0.5*area1;", "area[f] = area1;")],
    "tetra_centroid_fourth": [("mo[k] += (fc[k] + (v1[k] + v2[k] + cvo[k])*fourth)*v4;",
                               "mo[k] += (fc[k] + (v1[k] + v2[k] + cvo[k])*third)*v4;")],
    "other_side_frame": [("mj[k] += ((fc[k] + (v1[k] + v2[k] + cvj[k])*fourth) - a[k])*v4;",
                          "mj[k] += (fc[k] + (v1[k] + v2[k] + cvj[k])*fourth)*v4;")],
    "other_side_apex": [("cvj[k] = a[k] - fc[k];", "cvj[k] = 0.0 - fc[k];")],
    "volume_sixth": [("volume[i] = v*(1.0/6.0);", "volume[i] = v*(1.0/3.0);")],
    "triangle_centroid_2d": [("ws[(size_t)f*2*W + 1 + k] = wt*(2.0*fm[k]/3.0);",
                              "ws[(size_t)f*2*W + 1 + k] = wt*(fm[k]/3.0);")],
    "other_side_2d": [("ws[(size_t)f*2*W + W + 1 + k] = wt*(2.0*(fm[k] - a[k])/3.0);",
                       "ws[(size_t)f*2*W + W + 1 + k] = wt*(2.0*(fm[k] + a[k])/3.0);")],
    "side_pick": [("(pair_i[fe] == i ? 0 : W)", "(pair_i[fe] == i ? W : 0)")],
    "periodic_shift_digit": [("return xk + (double)(idx % 3 - 1)*(hi[a] - lo[a]);",
                              "return xk + (double)(idx % 3 - 1)*(hi[0] - lo[0]);")],
    "periodic_shift_base": [("for (b = 0; b < a; b++) idx /= 3;", "for (b = 0; b < a; b++) idx /= 2;")],
    "mirror_wall": [("const double edge = ((code - 1) & 1) ? hi[a] : lo[a];",
                     "const double edge = ((code - 1) & 1) ? lo[a] : hi[a];")],
}


def _build(tmp, name, edits):
    src = open(SRC).read()
    for old, new in edits:
        assert src.count(old) >= 1, (name, old)
        src = src.replace(old, new)
    d = os.path.join(tmp, name)
    os.makedirs(d)
    c = os.path.join(d, "phd_oracle.c")
    open(c, "w").write(src)
    so = os.path.join(d, "libphd_oracle.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
                    "-shared", c, "-o", so, "-lm"], check=True)
    return so


def _run_pins(so):
    env = dict(os.environ, PHD_ORACLE_LIB=so, OMP_NUM_THREADS="1")
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + PINS,
                       cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return p.returncode, p.stdout.strip().splitlines()[-1] if p.stdout.strip() else ""


def test_pins_catch_planted_mistakes(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    sos = {n: _build(str(tmp_path), n, e) for n, e in MUTATIONS.items()}
    with ThreadPoolExecutor(max_workers=8) as ex:
        res = dict(zip(sos, ex.map(_run_pins, sos.values())))
    survived = {n: r[1] for n, r in res.items() if r[0] == 0}
    assert not survived, "mistakes no pin catches: %r" % survived
    # the unmutated source passes the same pins through the same route
    rc, last = _run_pins(_build(str(tmp_path), "unmutated", []))
    assert rc == 0, last
