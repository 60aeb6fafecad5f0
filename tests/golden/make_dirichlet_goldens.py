"""Extract the reference's stored Dirichlet-wall accuracy goldens into dirichlet_goldens.json.

Reads (in this container only) the reference's own result files
  staggered_grid/exec/stokes_preconditioner/test_gmres/run_result/Accuracy_tests/
      {2d,3d}_dis_accuracy_dirichlet.dat
made by main_driver.f90 (test_type = 0) with no-slip walls on every side carrying the exact
solution's velocity as inhomogeneous boundary values (set_inhomogeneous_vel_bcs):
  * "discretization error" = ||rhs_exact - A x_exact||_inf per component, A the apply_matrix
    with vel_bc_n / vel_bc_t (main_driver.f90:339-386);
  * "global error" = ||e||_inf with e the homogeneous GMRES solution of A e = (tau_u, 0)
    (main_driver.f90:388-400).
Sections: 3-D prob_sol 31 on the shifted domain [0.2, 1.2]^3 and the Taylor vortex
(prob_sol 1, 2-D and quasi-2-D 3-D) on [0, 1]^dim, all with prob_coeff 1 and theta_alpha 0.
The variable-coefficient 3-D section holds no numbers and yields no record.
"""
import json
import os
import re
import sys

REF = "/root/reference/staggered_grid/exec/stokes_preconditioner/test_gmres/run_result/Accuracy_tests"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dirichlet_goldens.json")


def parse(fname, dim):
    recs = []
    coeff = sol = visc = None
    lo = 0.0
    cur = None
    for ln, line in enumerate(open(os.path.join(REF, fname)), start=1):
        if re.search(r"prob_lo\s*=\s*0\.2", line):
            lo = 0.2
        if "[0, 1]" in line:
            lo = 0.0
        m = re.search(r"prob_coeff\s*=\s*(\d+)", line)
        if m:
            coeff = int(m.group(1))
        m = re.search(r"prob_sol\s*=\s*(\d+)", line)
        if m:
            sol = int(m.group(1))
        m = re.match(r"\s*visc_type\s*=\s*(-?\d)", line)
        if m:
            visc = int(m.group(1))
            continue
        m = re.match(r"\s*(\d+)\*\d+", line)
        if m:
            cur = {"dim": dim, "n": int(m.group(1)), "visc_type": visc, "prob_coeff": coeff, "prob_sol": sol,
                   "prob_lo": lo, "file": f"staggered_grid/exec/stokes_preconditioner/test_gmres/run_result/"
                                          f"Accuracy_tests/{fname}", "lines": []}
            recs.append(cur)
            continue
        m = re.search(r"(discretization|global) error: L0 ([UVWP])\s*=\s*([-+0-9.Ee]+)", line)
        if m and cur is not None:
            key = ("" if m.group(1) == "discretization" else "g") + m.group(2)
            cur[key] = float(m.group(3))
            cur["lines"].append(ln)
    return [r for r in recs if r["prob_coeff"] == 1 and "U" in r]


def main():
    recs = parse("3d_dis_accuracy_dirichlet.dat", 3) + parse("2d_dis_accuracy_dirichlet.dat", 2)
    json.dump(recs, open(OUT, "w"), indent=1)
    print(f"wrote {len(recs)} records to {OUT}", file=sys.stderr)


if __name__ == "__main__":
    main()
