"""Extract the reference's stored truncation-error goldens into accuracy_goldens.json.

Reads (in this container only) the reference's own result files
  staggered_grid/exec/stokes_preconditioner/test_gmres/run_result/Accuracy_tests/{2d,3d}_dis_accuracy.dat
which hold ||rhs_exact - A x_exact||_inf per component of apply_matrix on manufactured solutions
(main_driver.f90 test_type = 0, [0,1]^dim, periodic, theta_alpha = 0).  Only the sections
produced by apply_matrix ("new code") are kept; the 2-D "old code" section (viscous term only)
and the Taylor-vortex (prob_sol = 1) rows are skipped.
"""
import json
import os
import re
import sys

REF = "/root/reference/staggered_grid/exec/stokes_preconditioner/test_gmres/run_result/Accuracy_tests"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "accuracy_goldens.json")


def parse(fname, dim):
    recs = []
    coeff = sol = visc = n = None
    cur = None
    path = os.path.join(REF, fname)
    for ln, line in enumerate(open(path), start=1):
        if "old code" in line:
            break
        m = re.search(r"prob_coeff\s*=\s*(\d+),\s*prob_sol\s*=\s*(\d+)", line)
        if m:
            coeff, sol = int(m.group(1)), int(m.group(2))
            continue
        m = re.match(r"\s*visc_type\s*=\s*(-?\d)", line)
        if m:
            visc = int(m.group(1))
            if coeff == 5 and visc > 0:
                visc = -visc
            continue
        m = re.match(r"\s*(\d+)\*\d+", line)
        if m:
            n = int(m.group(1))
            cur = {"dim": dim, "n": n, "visc_type": visc, "prob_coeff": coeff, "prob_sol": sol,
                   "file": f"staggered_grid/exec/stokes_preconditioner/test_gmres/run_result/Accuracy_tests/{fname}",
                   "lines": []}
            recs.append(cur)
            continue
        m = re.search(r"discretization error: L0 ([UVWP])\s*=\s*([-+0-9.Ee]+)", line)
        if m and cur is not None:
            cur[m.group(1)] = float(m.group(2))
            cur["lines"].append(ln)
    return [r for r in recs if r["prob_sol"] == 31 and "U" in r]


def main():
    recs = parse("3d_dis_accuracy.dat", 3) + parse("2d_dis_accuracy.dat", 2)
    json.dump(recs, open(OUT, "w"), indent=1)
    print(f"wrote {len(recs)} records to {OUT}", file=sys.stderr)


if __name__ == "__main__":
    main()
