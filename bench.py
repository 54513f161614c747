#!/usr/bin/env python3
"""Benchmark of the CD3D hot path on one MI355X.

Workload (one "step"): the configuration-2 problem -- unit cube, 64^3 hexahedra,
Q2 (2 146 689 DOFs), SUPG, eps = 1e-3, b = (1, 0.5, 0.25), c = 0, f = 1,
Dirichlet u = 0 -- assembled on all 5 multigrid levels (4^3 ... 64^3) and
solved by GMG-preconditioned FGMRES(20) to a relative residual of 1e-10.
Inputs (mesh, DOF maps, CSR patterns, transfers, Krylov workspace) are resident
in HBM before the timed region; each step re-assembles and re-solves.

Solver: V-cycle, damped Jacobi (omega 0.5, 3 pre + 3 post), exact coarse
solve, FGMRES(20).  (With ParMooN's configuration-1 smoother, omega 0.8 and
2 + 2 sweeps, the reference itself does not converge on this problem at 32^3.)

value = solves/s of the mixed-precision form (FP32 V-cycle, FP64 FGMRES,
residual, coarse solve and SUPG tau); the plain FP64 form, p50/p99 latency,
the fine-level kernels against their rooflines (HBM bandwidth; FP64 flops for
the element kernel), the reference CPU path (its C restatement on the same
problem, timed on this host's cores) and the error of both GPU forms against
that FP64 solution are reported beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "CD3D solves/sec at rel. residual 1e-10, mixed vs FP64; p50/p99 latency; SpMV GB/s"
HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(order=2, m=1, k0=2, levels=5, box=(1.0, 1.0, 1.0))
COEFFS = dict(eps=1e-3, b=(1.0, 0.5, 0.25), c=0.0, f_val=1.0, supg=1)
SOLVER = dict(omega=0.5, pre=3, post=3, coarse_smoother=17, restart=20, maxit=400, rel_tol=1e-10)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


KERNEL_NAMES = {0: "csr_stream<kSpmv,f64>",
                1: "lat_quad_gk<2,kLJacobi> f32 (compact stencils, 2 vertex/midpoint pairs on 4 lines per thread)",
                2: "lat64_quad_gk<1,kL64Spmv> f64 (1 vertex/midpoint pair on 4 lines per thread)",
                3: "lat_restrict_pair_k f32 (coarse vertex/midpoint pairs)",
                4: "lat_prolong_k<f32>", 5: "csr_fill_k",
                6: "copy_k (f64 vector copy, traffic calibration)",
                7: "axpy_dot2_k f64 (fused Gram-Schmidt step, 16-byte pairs)",
                8: "assemble_k<2> + gather_entries_k f64 (cell-by-cell SUPG element rows, then CSR entries gathered)",
                9: "ls_dots_k + ls_coeffs_k + ls_update2_k f64 (low-synchronisation Gram-Schmidt step, i = 10)"}
FP64_PEAK_TFLOPS = 78.6
FLOP_KERNELS = {8}


def kernel_table(P, reps=50):
    """Every kernel of the path on the finest level, timed with HIP events on
    the handle's stream by the library (cd3d_kernel_bench): mean duration and
    algorithmic bytes per launch."""
    l = P.levels - 1
    out = {}
    for which, name in KERNEL_NAMES.items():
        ms, amount = P.kernel_bench(which, l, reps if which not in FLOP_KERNELS else max(3, reps // 10))
        if which in FLOP_KERNELS:  # FP64 flops per launch instead of bytes
            out[name] = dict(us_per_launch=round(ms * 1e3, 2), algorithmic_flops=int(amount),
                             achieved_tflops=round(amount / (ms * 1e-3) / 1e12, 3))
        else:
            out[name] = dict(us_per_launch=round(ms * 1e3, 2), algorithmic_bytes=int(amount),
                             achieved_gbs=round(amount / (ms * 1e-3) / 1e9, 1))
    return out


# HBM bytes per launch of the fine-level kernels from the rocprofv3 counter
# runs (FETCH_SIZE x 2.0 + WRITE_SIZE, calibrated on a copy kernel)
TRAFFIC_FILE = os.path.join(REPO, "profiles", "cd3d_q2_64_fine_level_traffic.json")
TRAFFIC_KERNEL = {0: "void csr_stream<0, double, double, true>", 1: "void lat_quad_gk<2, 0, 256>",
                  2: "void lat64_quad_gk<1, 0>", 3: "lat_restrict_pair_k", 4: "void lat_prolong_k<2>",
                  5: "csr_fill2_k", 6: "copy_k", 7: "void axpy_dot2_k<true>", 8: "assemble_k + gather_entries_k"}


def measured_traffic(which):
    try:
        with open(TRAFFIC_FILE) as fh:
            k = json.load(fh)["kernels"].get(TRAFFIC_KERNEL[which])
        return int(k["traffic_bytes"]) if k else None
    except (OSError, ValueError, KeyError):
        return None


def roofline_entry(table, which, kernel_note):
    k = table[KERNEL_NAMES[which]]
    return dict(bound="hbm", achieved=k["achieved_gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(k["achieved_gbs"] / HBM_PEAK_GBS, 4), traffic=measured_traffic(which),
                kernel=KERNEL_NAMES[which], note=kernel_note, us_per_launch=k["us_per_launch"],
                algorithmic_bytes=k["algorithmic_bytes"])


def flop_roofline_entry(table, which, kernel_note):
    k = table[KERNEL_NAMES[which]]
    return dict(bound="fp64", achieved=k["achieved_tflops"], peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                frac=round(k["achieved_tflops"] / FP64_PEAK_TFLOPS, 4), traffic=measured_traffic(which),
                kernel=KERNEL_NAMES[which], note=kernel_note, us_per_launch=k["us_per_launch"],
                algorithmic_flops=k["algorithmic_flops"])


def run_steps(pm, P, sp, co, warmup, steps):
    lat, its, res = [], [], []
    for k in range(warmup + steps):
        t = time.perf_counter()
        P.assemble(co)
        st = P.solve_device(sp)
        dt = time.perf_counter() - t
        if st.status != 0:
            raise RuntimeError(f"solve status {st.status}")
        if k >= warmup:
            lat.append(dt)
            its.append(st.iterations)
            res.append(st.rfinal / st.r0)
        log(f"  step {k} ({'warmup' if k < warmup else 'timed'}): {dt * 1e3:.1f} ms, {st.iterations} its, "
            f"rel res {st.rfinal / st.r0:.2e}")
    return np.array(lat), its, res


def wavy_vertices(m):
    """Macro vertices of the unit cube split into m^3 hexahedra with every
    interior vertex moved by up to 0.15 macro cells per axis (general,
    non-parallelepiped hexahedra), x fastest."""
    g = np.arange(m + 1, dtype=np.float64)
    k, j, i = np.meshgrid(g, g, g, indexing="ij")
    inner = (i > 0) & (i < m) & (j > 0) & (j < m) & (k > 0) & (k < m)
    d = 0.15 * np.stack([np.sin(2.1 * i + 1.3 * j + 0.7 * k), np.sin(0.9 * i + 2.3 * j + 1.1 * k + 1.0),
                         np.sin(1.7 * i + 0.5 * j + 2.9 * k + 2.0)], -1) * inner[..., None]
    return ((np.stack([i, j, k], -1) + d) / m).reshape(-1, 3)


def general_hexahedra(pm, steps, warmup):
    """The headline problem on 64^3 general hexahedra (4^3 deformed macro
    cells refined 4 times, 5 levels): the trilinear element kernel's FP64
    rate on the finest level and the assemble + solve rate of both forms."""
    P = pm.Problem(2, 4, 0, 5, vertices=wavy_vertices(4))
    co = pm.coeffs(**COEFFS)
    out = {}
    for name, prec in (("fp64", pm.FP64), ("mixed", pm.MIXED)):
        lat, its, rr = run_steps(pm, P, pm.solver_params(**SOLVER, precision=prec), co, warmup, steps)
        out[name] = dict(solves_per_s=steps / lat.sum(), ms_per_step=1e3 * lat.mean(),
                         iterations=int(np.median(its)), max_rel_residual=float(max(rr)))
    ms, flops = P.kernel_bench(8, 4, 5)
    out["assembly_trilinear"] = dict(
        bound="fp64", achieved=round(flops / (ms * 1e-3) / 1e12, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
        frac=round(flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4), us_per_launch=round(ms * 1e3, 2),
        algorithmic_flops=int(flops),
        kernel="assemble_geo_k<2,false> + gather_entries_k f64 (trilinear map at every quadrature point)")
    P.close()
    return out


def large_q2_128(pm, steps, warmup):
    """The headline problem at 128^3 cells (Q2, 16 974 593 DOFs, 1.06e9
    nonzeros, 5 levels 8^3 .. 128^3): assemble + solve rate of both forms,
    device memory, and the fine-level sweep and SpMV rates at that size."""
    t = time.perf_counter()
    P = pm.Problem(2, 1, 3, 5, box=(1.0, 1.0, 1.0))
    setup_s = time.perf_counter() - t
    co = pm.coeffs(**COEFFS)
    out = dict(ndof=P.info(4)["ndof"], nnz=P.info(4)["nnz"], setup_s=round(setup_s, 2))
    for name, prec in (("mixed", pm.MIXED), ("fp64", pm.FP64)):
        lat, its, rr = run_steps(pm, P, pm.solver_params(**SOLVER, precision=prec), co, warmup, steps)
        out[name] = dict(solves_per_s=steps / lat.sum(), ms_per_step=1e3 * lat.mean(),
                         iterations=int(np.median(its)), max_rel_residual=float(max(rr)))
    out["device_gib"] = round(P.L.cd3d_device_bytes(P.h) / 2**30, 2)
    out["mixed_vs_fp64_speedup"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    for which, key, reps in ((1, "sweep", 50), (0, "spmv_fp64_csr", 10)):
        ms, b = P.kernel_bench(which, 4, reps)
        out[key] = dict(kernel=KERNEL_NAMES[which], us_per_launch=round(ms * 1e3, 2), algorithmic_bytes=int(b),
                        achieved_gbs=round(b / ms / 1e6, 1), frac=round(b / ms / 1e6 / HBM_PEAK_GBS, 4))
    P.close()
    return out


CPU_RECORD = os.path.join(REPO, "profiles", "cpu_reference_vs_port_c2_64.json")


def cpu_baseline(nproc=None, sol_out=None):
    """The reference CPU path on this host: the C restatement of the
    reference (oracle/cd3d_oracle.c, -O2, one thread per process) on the full
    workload -- the same 64^3 Q2 problem, assembly of all 5 levels plus the
    FGMRES solve to 1e-10, mesh and patterns built before the clock starts --
    as nproc independent solves at once on nproc host cores (throughput mode).
    Beside it: the restatement's and the reference's single-thread times on
    one core of the build container (profiles/cpu_reference_vs_port_c2_64.json)."""
    from oracle import cpu_baseline as cb

    n = nproc or cb.host_cores()
    r = cb.run_parallel(n, sol_out=sol_out, log=log)
    model = cb.host_cpu_model()
    rec = None
    try:
        with open(CPU_RECORD) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        pass
    per = float(np.median(r["per_worker_timed_s"]))
    out = dict(value=round(r["solves_per_s"], 6), unit="solves/s", cores=n, kind="port",
               sample=f"the full workload (Q2 64^3, {r['ndof']} DOFs, assembly of 5 levels + FGMRES to 1e-10, "
                      f"{r['iterations']} its): {n} independent single-thread solves of the C restatement at once "
                      f"on {n} cores of {model}; median {per:.1f} s per solve (assembly "
                      f"{np.median(r['assemble_s']):.1f} s + solve {np.median(r['solve_s']):.1f} s), "
                      f"value = {n} / slowest ({r['max_timed_s']:.1f} s)",
               cpu_model=model, per_solve_s_median=round(per, 3), iterations=r["iterations"],
               rel_residual=r["rel_residual"])
    if rec:
        out["single_core_container"] = rec
    return out


# ---------------------------------------------------------------- configurations 3-5

def mesh_for_n(n):
    """Nearest n = m 2^k with 2 <= m <= 7 (coarsest level m^3 cells); returns
    (n, m, levels), preferring more levels."""
    best = None
    for m in range(2, 8):
        for k in range(0, 7):
            v = m * 2 ** k
            key = (abs(v - n), -k)
            if best is None or key < best[0]:
                best = (key, (v, m, k + 1))
    return best[1]


def random_coeffs(rng, supg=1, conv_kind=0):
    """eps log-uniform in [1e-6, 1], b a random unit direction times U(0, 1),
    f = 1 (keyword arguments of coeffs())."""
    eps = 10 ** rng.uniform(-6, 0)
    d = rng.standard_normal(3)
    b = d / np.linalg.norm(d) * rng.uniform(0, 1)
    return dict(eps=eps, b=tuple(float(x) for x in b), f_val=1.0, supg=supg, conv_kind=conv_kind)


def c4_requests(count=256, seed=20261202, n_range=(8, 48)):
    """Configuration 4: `count` Q1 SUPG problems, n log-uniform in n_range
    (snapped to m 2^k, coarsest level m^3 cells), random eps and b.  Each
    request: dict(order, m, k0, levels, n, co)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        n, m, levels = mesh_for_n(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        reqs.append(dict(order=1, m=m, k0=0, levels=levels, n=n, co=random_coeffs(rng)))
    return reqs


def c5_arrivals(rate, seed, span, n_range=(16, 96), tcd_share=0.2):
    """Configuration 5: Poisson arrivals at `rate` per second over `span`
    seconds; n log-uniform in n_range (snapped), Q1 or Q2 with equal
    probability, a share of TCD3D requests (Galerkin, rotating field), the
    others CD3D SUPG with random eps and b.  Each arrival: dict(t, order, m,
    k0, levels, n, co, tcd)."""
    rng = np.random.default_rng(seed)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > span:
            break
        n, m, levels = mesh_for_n(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        order = 1 if rng.uniform() < 0.5 else 2
        tcd = bool(rng.uniform() < tcd_share)
        co = random_coeffs(rng, supg=0 if tcd else 1, conv_kind=1 if tcd else 0)
        arrivals.append(dict(t=t, order=order, m=m, k0=0, levels=levels, n=n, co=co, tcd=tcd))
    return arrivals


# TCD3D steps stop at a relative residual of 1e-10 against the step's own
# initial residual (the reference's restart-relative SC_LIN_RED_FACTOR_SCALAR
# would demand 1e-10 again after every restart)
TCD_STEPS = 10


def tcd_solver(pm, precision, maxit=1000):
    return pm.solver_params(**dict(SOLVER, maxit=maxit), precision=precision)


def tcd_run(pm, precision, steps, n=128, levels=5):
    """Configuration 3: Q1 128^3, eps 1e-4, rotating field, Crank-Nicolson
    dt 1e-3, convection matrix re-assembled every step on every level."""
    P = pm.Problem(1, 1, int(np.log2(n)) - levels + 1, levels, box=(1.0, 1.0, 1.0))
    co = pm.coeffs(eps=1e-4, conv_kind=1, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5), f_val=0.0)
    sp = tcd_solver(pm, precision)
    u0 = P.interp_pulse(levels - 1, co)
    P.tcd_init(co, 0.0, u0)
    P.tcd_step(co, sp, 1e-3, want_state=False)  # warm-up step (allocations)
    P.tcd_init(co, 0.0, u0)
    its = 0
    P.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        _, st = P.tcd_step(co, sp, 1e-3, want_state=False)
        its += st.iterations
    dt = time.perf_counter() - t
    u = P.tcd_state()
    P.close()
    return dict(steps=steps, steps_per_s=steps / dt, ms_per_step=1e3 * dt / steps, iterations=its), u


def failure_record(req, r):
    """What names a request that did not converge."""
    co = req["co"]
    return dict(kind="tcd" if req.get("tcd") else "cd", order=req["order"], m=req["m"], levels=req["levels"],
                n=req["n"], eps=co["eps"], b=list(co["b"]), status=int(r.status), iterations=int(r.iterations),
                rel_residual=float(r.rel_residual), steps_done=int(r.steps_done))


def batched_uneven(pm, precision, count=256, seed=20261202, passes=2, inflight=16, max_batch=None, workers=0):
    """Configuration 4: `count` independent Q1 SUPG problems, n log-uniform in
    [8, 48] (snapped to m 2^k), eps log-uniform in [1e-6, 1], b a random unit
    direction times U(0, 1), f = 1; all submitted at once to the server.  The
    first pass builds the resident hierarchies, the last one is reported;
    solves/s counts converged requests only."""
    reqs = c4_requests(count, seed)
    sp = pm.solver_params(**dict(SOLVER, maxit=1000), precision=precision)
    srv = pm.Server(max_inflight=inflight, max_batch=max_batch, workers=workers)
    out = None
    for _ in range(passes):
        t = time.perf_counter()
        ids = [srv.submit(pm.mesh_desc(q["order"], q["m"], q["k0"], q["levels"]), pm.coeffs(**q["co"]), sp)
               for q in reqs]
        res = [srv.wait(i)[0] for i in ids]
        wall = time.perf_counter() - t
        ok = [r for r in res if r.status == 0]
        lat = np.array([r.latency_ms for r in ok])
        out = dict(problems=count, solves_per_s=len(ok) / wall, wall_s=wall, p50_ms=float(np.percentile(lat, 50)),
                   p99_ms=float(np.percentile(lat, 99)), failed=count - len(ok),
                   failures=[failure_record(q, r) for q, r in zip(reqs, res) if r.status != 0],
                   max_rel_residual=float(max(r.rel_residual for r in ok)),
                   median_iterations=int(np.median([r.iterations for r in ok])),
                   reused=sum(r.reused_hierarchy for r in res),
                   mean_batch=float(np.mean([r.batch_size for r in res])))
    st = srv.stats()
    out["peak_device_gib"] = st.peak_bytes / 2**30
    srv.close()
    return out


def c4_form(pm, precision, args):
    """Configuration 4 in one precision form as same-level batches (all
    problems with the same level count and solver in lockstep,
    cd3d_solve_batch: 256 slots, 4 worker threads, up to --c4-batch per
    batch; the mixed form's batches run an FP32 V-cycle over the CSR
    levels), or with --c4-batch 1 one request per slot."""
    if args.c4_batch > 1:
        return batched_uneven(pm, precision, args.c4_problems, inflight=256, max_batch=args.c4_batch, workers=4)
    return batched_uneven(pm, precision, args.c4_problems, inflight=args.c4_inflight, max_batch=1)


def serving_stream(pm, precision, rates, duration_s=4.0, seed=7, tcd_share=0.2, max_requests=400):
    """Configuration 5: Poisson arrivals of CD3D / TCD3D requests, n in
    [16, 96] (snapped), Q1 or Q2 with equal probability, eps and b as in
    configuration 4; TCD3D requests run 10 Crank-Nicolson steps of the
    rotating field.  Per rate: offered load, completed solves/s, p50/p99.
    Arrivals span duration_s, or max_requests / rate seconds when shorter (a
    rate above the server's capacity then measures the capacity: solves/s
    from the first arrival to the last completion).  Requests that end
    unconverged are listed under `failures` and not counted as solves."""
    import threading
    srv = pm.Server(max_inflight=16)
    out = []
    for ri, rate in enumerate(rates):
        span = min(duration_s, max_requests / rate)
        arrivals = c5_arrivals(rate, seed + ri, span, tcd_share=tcd_share)
        sp_cd = pm.solver_params(**dict(SOLVER, maxit=1000), precision=precision)
        sp_tcd = tcd_solver(pm, precision)

        def submit_all(stream, ids, t0):
            for a in stream:
                d = t0 + a["t"] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                mesh = pm.mesh_desc(a["order"], a["m"], a["k0"], a["levels"])
                if a["tcd"]:
                    ids.append(srv.submit(mesh, pm.coeffs(**a["co"]), sp_tcd, kind=pm.REQ_TCD, tau=1e-3,
                                          nsteps=TCD_STEPS))
                else:
                    ids.append(srv.submit(mesh, pm.coeffs(**a["co"]), sp_cd))

        for phase in ("warm", "timed"):
            ids = []
            t0 = time.perf_counter()
            th = threading.Thread(target=submit_all, args=(arrivals, ids, t0))
            th.start()
            th.join()
            res = [srv.wait(i)[0] for i in ids]
            wall = time.perf_counter() - t0
        ok = [r for r in res if r.status == 0]
        lat = np.array([r.latency_ms for r in ok])
        out.append(dict(rate_per_s=rate, requests=len(res), duration_s=span, wall_s=wall,
                        solves_per_s=len(ok) / wall, p50_ms=float(np.percentile(lat, 50)),
                        p99_ms=float(np.percentile(lat, 99)), failed=len(res) - len(ok),
                        escalated=sum(int(r.escalated) for r in res),
                        median_iterations=int(np.median([r.iterations for r in res])),
                        max_iterations=int(max(r.iterations for r in res)),
                        failures=[failure_record(a, r) for a, r in zip(arrivals, res) if r.status != 0],
                        tcd_requests=sum(a["tcd"] for a in arrivals)))
        log(f"[bench]   rate {rate}/s: {len(res)} requests ({out[-1]['failed']} failed), p50 {out[-1]['p50_ms']:.1f} ms, "
            f"p99 {out[-1]['p99_ms']:.1f} ms, {out[-1]['solves_per_s']:.1f} solves/s")
    srv.close()
    return out


def r5(d):
    if isinstance(d, dict):
        return {k: r5(v) for k, v in d.items()}
    if isinstance(d, list):
        return [r5(v) for v in d]
    if isinstance(d, float):
        return float(f"{d:.5g}")
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=0, help="CPU baseline processes (default: host cores, at most 16)")
    ap.add_argument("--no-extra", action="store_true", help="only the headline configuration")
    ap.add_argument("--no-large", action="store_true", help="skip the 128^3-cell Q2 record")
    ap.add_argument("--c3-steps", type=int, default=100)
    ap.add_argument("--c4-problems", type=int, default=256)
    ap.add_argument("--c4-inflight", type=int, default=16, help="server slots for configuration 4 (unbatched)")
    ap.add_argument("--c4-batch", type=int, default=64,
                    help="configuration 4, FP64 form: requests per same-level batch (with 256 slots, 4 workers)")
    ap.add_argument("--c5-rates", type=str, default="10,50,400")
    ap.add_argument("--c5-seconds", type=float, default=4.0)
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both",
                    help="run one form only (profiling)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="assemble once and time the path's kernels (for rocprofv3 counter runs)")
    ap.add_argument("--only", choices=["tcd", "batched", "serving"],
                    help="run one of configurations 3-5 alone (profiling) and print its result")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every MI355X box here has one GPU)",
              file=sys.stderr)
        return 2
    import parmoon_cmg_amd as pm

    if args.only:
        forms = [f for f in (("fp64", pm.FP64), ("mixed", pm.MIXED)) if args.precision in ("both", f[0])]
        out = {}
        for name, prec in forms:
            if args.only == "tcd":
                out[name] = tcd_run(pm, prec, args.c3_steps)[0]
            elif args.only == "batched":
                out[name] = c4_form(pm, prec, args)
            else:
                out[name] = serving_stream(pm, prec, [float(x) for x in args.c5_rates.split(",") if x],
                                           args.c5_seconds)
        print(json.dumps(r5(out)))
        return 0

    log("[bench] building the 64^3 Q2 hierarchy (5 levels)")
    t = time.perf_counter()
    P = pm.Problem(**WORKLOAD)
    log(f"[bench] setup {time.perf_counter() - t:.1f} s, device memory {P.device_bytes() / 2**30:.2f} GiB")
    co = pm.coeffs(**COEFFS)
    if args.kernels_only:
        P.assemble(co)
        print(json.dumps(kernel_table(P)))
        P.close()
        return 0
    res = {}
    sols = {}
    forms = [f for f in (("fp64", pm.FP64), ("mixed", pm.MIXED)) if args.precision in ("both", f[0])]
    for name, prec in forms:
        sp = pm.solver_params(**SOLVER, precision=prec)
        log(f"[bench] {name}: {args.warmup} warmup + {args.steps} timed steps")
        lat, its, rr = run_steps(pm, P, sp, co, args.warmup, args.steps)
        sols[name] = P.solution()
        res[name] = dict(solves_per_s=args.steps / lat.sum(), ms_per_step=1e3 * lat.mean(),
                         p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                         iterations=int(np.median(its)), max_rel_residual=float(max(rr)))
    if args.precision != "both":
        print(json.dumps({k: r5(v) for k, v in res.items()}))
        P.close()
        return 0
    err = float(np.linalg.norm(sols["mixed"] - sols["fp64"]) / np.linalg.norm(sols["fp64"]))
    log("[bench] kernel table (HIP events on the library stream)")
    table = kernel_table(P)
    roof = roofline_entry(table, 1, "dominant kernel of the mixed step (with the Gram-Schmidt passes, "
                                    "lowsync_gs_step): the fine-level FP32 Jacobi sweep (algorithmic bytes: x, f "
                                    "and the result once, 12 bytes per node)")
    mgs = roofline_entry(table, 7, "fused modified Gram-Schmidt step of the FP64 form's FGMRES on the fine level "
                         "(32 bytes per unknown; the reference's Gram-Schmidt, one launch per basis vector)")
    lowsync = roofline_entry(table, 9, "the mixed form's low-synchronisation Gram-Schmidt step at i = 10 on the "
                             "fine level: one pass for the dots, one for w -= V h and |w|^2 ((2 i + 5) 8 bytes "
                             "per unknown)")
    asm = flop_roofline_entry(table, 8, "per-element quadrature (Gauss4, 64 points) of the Q2 SUPG element "
                              "matrices on the fine level: element rows written out, then every CSR entry "
                              "gathered from its contributions; algorithmic flops n_cells n_q 8 n_b^2 per "
                              "assembly (the path for variable coefficients and general hexahedra)")
    spmv = roofline_entry(table, 0, "FP64 CSR SpMV of the fine level (12 nnz + 20 N bytes); the FP64 form's "
                                    "smoother, defect and FGMRES product")
    extra = {}
    if not args.no_extra:
        log("[bench] 64^3 Q2 SUPG on general hexahedra (deformed 4^3 macro mesh)")
        extra["general_hexahedra_q2_64"] = general_hexahedra(pm, 3, 1)
        if not args.no_large:
            log("[bench] the headline problem at 128^3 cells (16 974 593 DOFs)")
            extra["large_q2_128"] = large_q2_128(pm, 3, 1)
        for name, prec in (("fp64", pm.FP64), ("mixed", pm.MIXED)):
            log(f"[bench] configuration 3 (TCD3D Q1 128^3, {args.c3_steps} CN steps), {name}")
            r, u = tcd_run(pm, prec, args.c3_steps)
            extra.setdefault("tcd3d_q1_128", {})[name] = r
            extra["tcd3d_q1_128"].setdefault("_u", {})[name] = u
        u = extra["tcd3d_q1_128"].pop("_u")
        extra["tcd3d_q1_128"]["mixed_state_rel_err_vs_fp64"] = float(
            np.linalg.norm(u["mixed"] - u["fp64"]) / np.linalg.norm(u["fp64"]))
        for name, prec in (("fp64", pm.FP64), ("mixed", pm.MIXED)):
            log(f"[bench] configuration 4 ({args.c4_problems} batched uneven Q1 problems), {name}")
            extra.setdefault("batched_uneven_q1", {})[name] = c4_form(pm, prec, args)
        for name, prec in (("fp64", pm.FP64), ("mixed", pm.MIXED)):
            log(f"[bench] configuration 4, {name} one request per slot (no batching)")
            extra["batched_uneven_q1"][f"{name}_unbatched"] = batched_uneven(pm, prec, args.c4_problems,
                                                                             inflight=args.c4_inflight, max_batch=1)
        rates = [float(x) for x in args.c5_rates.split(",") if x]
        for name, prec in (("fp64", pm.FP64), ("mixed", pm.MIXED)):
            log(f"[bench] configuration 5 (Poisson serving stream), {name}")
            extra.setdefault("serving_stream", {})[name] = serving_stream(pm, prec, rates, args.c5_seconds)
        ss = extra["serving_stream"]
        # at the highest offered rate both forms run at capacity
        ss["mixed_vs_fp64_at_max_rate"] = round(ss["mixed"][-1]["solves_per_s"] / ss["fp64"][-1]["solves_per_s"], 3)
    cpu, vs_oracle = None, {}
    if not args.no_cpu:
        log("[bench] reference CPU path (C restatement) on the full workload, one solve per host core")
        with tempfile.TemporaryDirectory() as td:
            ofile = os.path.join(td, "oracle_sol.npy")
            cpu = cpu_baseline(args.cpu_procs or None, sol_out=ofile)
            xo = np.load(ofile)
        # error of each GPU form against the FP64 oracle's solution of the same problem
        for name in sols:
            vs_oracle[name] = float(np.linalg.norm(sols[name] - xo) / np.linalg.norm(xo))
        log(f"[bench] solution error vs the FP64 oracle: " + ", ".join(f"{k} {v:.2e}" for k, v in vs_oracle.items()))
    m = res["mixed"]
    out = {
        "metric": METRIC, "value": round(m["solves_per_s"], 4), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(m["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 V-cycle / f64 FGMRES",
        "data": "synthetic (configuration-2 coefficients)",
        "config": {"workload": "CD3D Q2 SUPG, unit cube 64^3 hexahedra, 2146689 DOFs, 5-level GMG-FGMRES(20) to 1e-10",
                   "eps": 1e-3, "b": [1.0, 0.5, 0.25], "smoother": "Jacobi omega 0.5, 3+3, V-cycle, exact coarse",
                   "parallelism": "1 GPU"},
        "mixed": {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in m.items()},
        "fp64": {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in res["fp64"].items()},
        "mixed_vs_fp64_speedup": round(m["solves_per_s"] / res["fp64"]["solves_per_s"], 3),
        "mixed_solution_rel_err_vs_fp64": err,
        "solution_rel_err_vs_fp64_oracle": vs_oracle or None,
        "p50_ms": round(m["p50_ms"], 3), "p99_ms": round(m["p99_ms"], 3),
        "spmv_gbs": spmv["achieved"],
        "roofline": roof,
        "spmv_fp64_csr": spmv,
        "mgs_step": mgs,
        "lowsync_gs_step": lowsync,
        "assembly_cellwise": asm,
        "kernels_fine_level": table,
        "cpu_baseline": cpu,
    }
    out.update(r5(extra))
    print(json.dumps(out))
    P.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
