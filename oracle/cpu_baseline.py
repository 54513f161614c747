"""CPU baseline leg of the benchmark -- TEST INFRASTRUCTURE ONLY.

Times the C restatement of the reference (oracle/cd3d_oracle.c, one thread per
process) on the configuration-2 problem: unit cube, Q2, SUPG eps 1e-3,
b = (1, 0.5, 0.25), f = 1, 5 levels (64^3 cells on the finest), FGMRES(20)
with the damped Jacobi V-cycle (omega 0.5, 3 + 3) and the exact coarse solve,
to a relative residual of 1e-10.  Mesh, DOF maps and CSR patterns are built
before the clock starts (as on the GPU side, where they are resident); the
timed part is the assembly of every level plus the solve.

As a worker (``python -m oracle.cpu_baseline --worker ...``) it runs one
problem and prints one JSON line (times, iterations, optionally the solution
written to an .npy file).  ``run_parallel`` starts ``nproc`` such workers at
once (throughput mode: independent single-thread solves on as many host
cores) and reports solves/s = nproc / wall.
"""
import argparse
import json
import os
import queue
import subprocess
import sys
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)

C2 = dict(order=2, m=1, k0=2, levels=5)
C2_COEFFS = dict(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1)
C2_SOLVER = dict(omega=0.5, pre=3, post=3, coarse_smoother=17, restart=20, maxit=400, rel_tol=1e-10,
                 red_factor=0.0)


def worker(order, m, k0, levels, sol_out=None):
    sys.path.insert(0, REPO)
    import numpy as np

    from oracle import pyoracle as po

    t = time.perf_counter()
    H = po.Hierarchy(order, m, k0, levels, box=(1.0, 1.0, 1.0))
    t_setup = time.perf_counter() - t
    print(json.dumps(dict(event="setup", s=round(t_setup, 3))), flush=True)
    t = time.perf_counter()
    H.assemble(po.coeffs(**C2_COEFFS))
    t_asm = time.perf_counter() - t
    b = H.rhs(levels - 1)
    t = time.perf_counter()
    x, its, hist = H.fgmres(po.solver(**C2_SOLVER), b)
    t_solve = time.perf_counter() - t
    _, r0 = H.defect(levels - 1, np.zeros_like(b), b)
    _, rf = H.defect(levels - 1, x, b)
    if sol_out:
        np.save(sol_out, x)
    print(json.dumps(dict(event="done", setup_s=t_setup, assemble_s=t_asm, solve_s=t_solve, iterations=int(its),
                          rel_residual=rf / r0, ndof=int(len(b)))), flush=True)


def host_cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def host_cores(cap=16):
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, cap))


def run_parallel(nproc, wl=C2, sol_out=None, log=None, timeout=900):
    """nproc independent single-thread workers at once; returns a dict with
    the wall time, per-worker times and solves/s = nproc / (last finish -
    start of the timed parts).  The first worker writes its solution to
    sol_out (.npy) when given."""
    env = dict(os.environ, OMP_NUM_THREADS="1")
    cmd = [sys.executable, "-u", "-m", "oracle.cpu_baseline", "--worker", "--order", str(wl["order"]), "--m",
           str(wl["m"]), "--k0", str(wl["k0"]), "--levels", str(wl["levels"])]
    procs = []
    for i in range(nproc):
        extra = ["--sol-out", sol_out] if (sol_out and i == 0) else []
        procs.append(subprocess.Popen(cmd + extra, cwd=REPO, env=env, stdout=subprocess.PIPE, text=True))
    t0 = time.perf_counter()
    events = queue.Queue()

    def reader(p):
        for line in p.stdout:
            events.put(json.loads(line))
        events.put(None)

    threads = [threading.Thread(target=reader, args=(p,), daemon=True) for p in procs]
    for th in threads:
        th.start()
    results, setups, ended = [], 0, 0
    while ended < nproc:
        ev = events.get(timeout=timeout)
        if ev is None:
            ended += 1
        elif ev["event"] == "setup":
            setups += 1
            if log and setups == nproc:
                log(f"[cpu] {nproc} workers set up ({time.perf_counter() - t0:.0f} s); assembling and solving")
        else:
            results.append(ev)
            if log:
                log(f"[cpu] worker done after {time.perf_counter() - t0:.0f} s: assembly {ev['assemble_s']:.1f} s, "
                    f"solve {ev['solve_s']:.1f} s, {ev['iterations']} its")
    for p in procs:
        p.wait(timeout=timeout)
    wall = time.perf_counter() - t0
    if len(results) != nproc:
        raise RuntimeError("a CPU baseline worker failed")
    # the timed part of every worker starts after its setup: measure the
    # throughput over the longest timed part (all workers run concurrently)
    timed = [r["assemble_s"] + r["solve_s"] for r in results]
    return dict(nproc=nproc, wall_s=wall, max_timed_s=max(timed), per_worker_timed_s=timed,
                assemble_s=[r["assemble_s"] for r in results], solve_s=[r["solve_s"] for r in results],
                iterations=results[0]["iterations"], rel_residual=results[0]["rel_residual"],
                ndof=results[0]["ndof"], solves_per_s=nproc / max(timed))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--worker", action="store_true")
    ap.add_argument("--order", type=int, default=C2["order"])
    ap.add_argument("--m", type=int, default=C2["m"])
    ap.add_argument("--k0", type=int, default=C2["k0"])
    ap.add_argument("--levels", type=int, default=C2["levels"])
    ap.add_argument("--sol-out")
    ap.add_argument("--nproc", type=int, default=0)
    a = ap.parse_args()
    if a.worker:
        worker(a.order, a.m, a.k0, a.levels, a.sol_out)
        return 0
    wl = dict(order=a.order, m=a.m, k0=a.k0, levels=a.levels)
    r = run_parallel(a.nproc or host_cores(), wl, log=lambda s: print(s, file=sys.stderr, flush=True))
    r["cpu_model"] = host_cpu_model()
    print(json.dumps(r))
    return 0


if __name__ == "__main__":
    sys.exit(main())
