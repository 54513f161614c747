"""Single-core timing of the reference against its C restatement on the
configuration-2 workload -- CHECKER ONLY, run in the build container (where
oracle/_ref is built from the reference's sources; it is never shipped).

Both run the full problem (unit cube, Q2, 64^3 cells, 5 levels, SUPG eps 1e-3,
b = (1, 0.5, 0.25), f = 1; FGMRES(20), Jacobi V-cycle omega 0.5 with 3 + 3
sweeps, relative residual 1e-10) one after the other on the same core.  The
reference build has no LAPACK, so both use the SSOR coarse smoother (10
sweeps or a reduction of 0.1); the restatement is also timed with the exact
coarse solve (what the benchmark runs on the GPU host).  Wall clock
(std::chrono in the reference driver, time.perf_counter here), setup of mesh,
spaces and matrix structures excluded.

usage: python oracle/time_reference.py [--core 2] [--levels 5] [--out profiles/cpu_reference_vs_port_c2_64.json]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def reference_dat(levels):
    data = os.path.join(HERE, "data")
    return dict(GEOFILE=os.path.join(data, "unitcube.GEO"), BNDFILE=os.path.join(data, "unitcube.PRM"), MESH_TYPE=0,
                OUTFILE="ref.out", LOGFILE="ref.log", MEASURE_ERRORS=0, WRITE_VTK=0, WRITE_PS=0,
                ANSATZ_ORDER=2, LEVELS=levels, UNIFORM_STEPS=2, DISCTYPE=2, CELL_MEASURE=4, SDFEM_TYPE=2,
                DELTA0=1.0, P0=1e-3, P1=1.0, P2=0.5, P3=0.25, P4=0, P5=0, P6=1.0,
                SOLVER_TYPE=1, SC_MG_TYPE_SCALAR=0, SC_SOLVER_SCALAR=16, SC_PRECONDITIONER_SCALAR=5,
                SC_GMRES_RESTART=20, SC_LIN_MAXIT_SCALAR=400, SC_LIN_RED_FACTOR_SCALAR=1e-10,
                SC_LIN_RES_NORM_MIN_SCALAR=1e-30, SC_MG_CYCLE_SCALAR=1, SC_SMOOTHER_SCALAR=1,
                SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3, SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5,
                SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5, SC_COARSE_SMOOTHER_SCALAR=3, SC_COARSE_MAXIT_SCALAR=10,
                SC_COARSE_RED_FACTOR_SCALAR=0.1, SC_GMG_DAMP_FACTOR_SCALAR=1.0,
                SC_STEP_LENGTH_CONTROL_FINE_SCALAR=0, SC_STEP_LENGTH_CONTROL_ALL_SCALAR=0, SC_AMG_PREC_IT=1,
                SC_MINIT=0, SC_VERBOSE=0, SC_FLEXIBLE_KRYLOV_SPACE_SOLVER=1)


def time_reference(levels, core):
    exe = os.path.join(HERE, "_ref", "cd3d_ref")
    with tempfile.TemporaryDirectory() as td:
        dat = os.path.join(td, "c2.dat")
        with open(dat, "w") as f:
            for k, v in reference_dat(levels).items():
                f.write(f"{k}: {v}\n")
        p = subprocess.run(["taskset", "-c", str(core), exe, dat, "-", "cd"], cwd=td, capture_output=True,
                           text=True, check=True)
    t = [l for l in p.stdout.splitlines() if l.startswith("TIME ")][-1].split()
    s = [l for l in p.stdout.splitlines() if l.startswith("SOLVE ")][-1].split()
    return dict(assemble_s=float(t[6]), solve_s=float(t[8]), iterations=int(s[2]), rel_residual=float(s[8]))


def time_port(levels, core, coarse_smoother):
    code = ("import json,sys,time; sys.path.insert(0, %r)\n"
            "import numpy as np\n"
            "from oracle import pyoracle as po\n"
            "from oracle.cpu_baseline import C2_COEFFS, C2_SOLVER\n"
            "H = po.Hierarchy(2, 1, 2, %d, box=(1.0, 1.0, 1.0))\n"
            "t = time.perf_counter(); H.assemble(po.coeffs(**C2_COEFFS)); ta = time.perf_counter() - t\n"
            "b = H.rhs(%d)\n"
            "sv = po.solver(**dict(C2_SOLVER, coarse_smoother=%d))\n"
            "t = time.perf_counter(); x, its, _ = H.fgmres(sv, b); ts = time.perf_counter() - t\n"
            "_, r0 = H.defect(%d, np.zeros_like(b), b); _, rf = H.defect(%d, x, b)\n"
            "print(json.dumps(dict(assemble_s=ta, solve_s=ts, iterations=int(its), rel_residual=rf / r0)))\n"
            % (REPO, levels, levels - 1, coarse_smoother, levels - 1, levels - 1))
    p = subprocess.run(["taskset", "-c", str(core), sys.executable, "-c", code], capture_output=True, text=True,
                       check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
    return json.loads(p.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--core", type=int, default=2)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(REPO, "profiles", "cpu_reference_vs_port_c2_64.json"))
    a = ap.parse_args()
    ref = time_reference(a.levels, a.core)
    print("reference", ref, flush=True)
    port_ssor = time_port(a.levels, a.core, 3)
    print("port (SSOR coarse)", port_ssor, flush=True)
    port_exact = time_port(a.levels, a.core, 17)
    print("port (exact coarse)", port_exact, flush=True)
    model = "unknown"
    for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
        if line.startswith("Model name:"):
            model = line.split(":", 1)[1].strip()
    tot = lambda r: r["assemble_s"] + r["solve_s"]
    rec = dict(workload=f"Q2 SUPG unit cube, {2 ** (a.levels + 1)}^3 cells, {a.levels} levels, FGMRES(20) to 1e-10",
               cpu_model=model, core=a.core, compiler="g++/gcc 11.4 -O2, one thread",
               reference_ssor_coarse=ref, port_ssor_coarse=port_ssor, port_exact_coarse=port_exact,
               port_over_reference=round(tot(port_ssor) / tot(ref), 4),
               command="python oracle/time_reference.py --core %d --levels %d" % (a.core, a.levels))
    with open(a.out, "w") as f:
        json.dump(rec, f, indent=1)
        f.write("\n")
    print(json.dumps(rec))
    return 0


if __name__ == "__main__":
    sys.exit(main())
