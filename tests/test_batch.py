"""Same-level batches (cd3d_solve_batch): problems of different sizes, element
orders and coefficients with the same number of levels, solved together in
lockstep.  Each problem must end as it does alone: the same iteration count
as the FP64 oracle (exact coarse solve, relative residual 1e-10) and as a
lone cd3d_solve, solution within 1e-10 of the oracle's; its own status when
it starts converged, stops at maxit, or diverges.  In the mixed form (FP32
V-cycle): relative residual 1e-10 and the solution within 1e-9 of the FP64
oracle's."""
import sys

import numpy as np
import pytest

import parmoon_cmg_amd as pm
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SOLVER = dict(omega=0.5, pre=3, post=3, coarse_smoother=17, rel_tol=1e-10, maxit=300)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# (order, m, k0, levels, coefficients): 3 levels each, 4^3 .. 20^3 cells
MIX = [
    (1, 2, 0, 3, dict(eps=1e-2, b=(0.3, 0.4, -0.5), f_val=1.0, supg=1)),
    (1, 5, 0, 3, dict(eps=1e-4, b=(-0.6, 0.2, 0.7), f_val=1.0, supg=1)),
    (2, 2, 0, 3, dict(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1)),
    (1, 3, 0, 3, dict(eps=1.0, f_kind=1)),
    (2, 3, 0, 3, dict(eps=0.05, b=(0.2, -0.1, 0.3), c=0.5, f_val=1.0, supg=1)),
    (1, 5, 1, 3, dict(eps=3e-6, b=(0.5, 0.5, 0.1), f_val=1.0, supg=1)),
]


def oracle_solve(order, m, k0, levels, co):
    H = po.Hierarchy(order, m, k0, levels, box=(1.0, 1.0, 1.0))
    H.assemble(po.coeffs(**co))
    b = H.rhs(levels - 1)
    _, r0 = H.defect(levels - 1, np.zeros_like(b), b)
    x, its, _ = H.fgmres(po.solver(**dict(SOLVER, res_norm_min=1e-10 * r0, red_factor=0.0)), b)
    return x, its


def make(order, m, k0, levels, co):
    P = pm.Problem(order, m, k0, levels, box=(1.0, 1.0, 1.0))
    P.assemble(pm.coeffs(**co))
    return P


def test_batch_matches_oracle_and_lone_solves():
    sp = pm.solver_params(**SOLVER)
    probs = [make(*c) for c in MIX]
    say(f"[gpu] batch of {len(probs)}: " + ", ".join(str(p.info(2)['ndof']) for p in probs))
    stats = pm.solve_batch(probs, sp)
    xs = [p.solution() for p in probs]
    for c, P, st, x in zip(MIX, probs, stats, xs):
        xo, its_o = oracle_solve(*c)
        _, lone = P.solve(sp)
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        say(f"[gpu] batch {c[:4]}: its {st.iterations} (oracle {its_o}, alone {lone.iterations}), err {err:.1e}, "
            f"rel res {st.rfinal / st.r0:.1e}")
        assert st.status == 0 and st.iterations == its_o == lone.iterations
        assert err < 1e-10 and st.rfinal <= 1.0001e-10 * st.r0
    for P in probs:
        P.close()


def test_mixed_batch_matches_fp64_oracle():
    # the mixed form in a batch: FP32 V-cycle over the CSR levels inside FP64
    # FGMRES.  Each member reaches relative residual 1e-10 (recomputed here
    # in FP64 from the CSR matrix), its solution is within 1e-9 of the FP64
    # oracle's, and it needs at most 2 iterations more than the oracle
    import scipy.sparse as sps
    sp = pm.solver_params(**SOLVER, precision=pm.MIXED)
    probs = [make(*c) for c in MIX]
    stats = pm.solve_batch(probs, sp)
    for c, P, st in zip(MIX, probs, stats):
        top = c[3] - 1
        x = P.solution()
        xo, its_o = oracle_solve(*c)
        rp, kc, e = P.csr(top)
        n, na = len(rp) - 1, P.info(top)["nact"]
        A = sps.csr_matrix((e, kc, rp), shape=(n, n))
        b = P.rhs(top)
        x0 = np.zeros(n)
        x0[na:] = b[na:]
        rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b - A @ x0)
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        say(f"[gpu] mixed batch {c[:4]}: its {st.iterations} (fp64 oracle {its_o}), rel res {rel:.1e}, err {err:.1e}")
        assert st.status == 0 and st.iterations <= its_o + 2
        assert rel <= 1.001e-10 and err < 1e-9
    for P in probs:
        P.close()



def test_mixed_batch_low_sync_and_mgs_agree(monkeypatch):
    # the mixed batch orthogonalises with the low-synchronisation
    # Gram-Schmidt (two passes over the basis per Arnoldi step, each member
    # over its own number of 2048-element chunks: 1 .. 34 here);
    # CD3D_FGMRES_MGS=1 takes MGS vector by vector.  The two are the same
    # algorithm in exact arithmetic: iteration counts within one of each
    # other, solutions within 2e-9.  Bitwise equality over all members would
    # mean the two settings ran the same code
    sp = pm.solver_params(**SOLVER, precision=pm.MIXED)
    probs = [make(*c) for c in MIX]
    monkeypatch.delenv("CD3D_FGMRES_MGS", raising=False)
    st_ls = pm.solve_batch(probs, sp)
    x_ls = [P.solution() for P in probs]
    monkeypatch.setenv("CD3D_FGMRES_MGS", "1")
    st_mgs = pm.solve_batch(probs, sp)
    x_mgs = [P.solution() for P in probs]
    monkeypatch.delenv("CD3D_FGMRES_MGS")
    for c, a, b, xa, xb in zip(MIX, st_ls, st_mgs, x_ls, x_mgs):
        d = np.linalg.norm(xa - xb) / np.linalg.norm(xb)
        say(f"[gpu] mixed batch {c[:4]}: its low-sync {a.iterations}, MGS {b.iterations}, rel diff {d:.1e}")
        assert a.status == 0 and b.status == 0
        assert abs(a.iterations - b.iterations) <= 1 and d < 2e-9
    assert not all(np.array_equal(xa, xb) for xa, xb in zip(x_ls, x_mgs))
    for P in probs:
        P.close()

def test_batch_statuses_of_its_members():
    # a zero rhs (converged before the first iteration), a stop at maxit, a
    # divergence test that trips, and a normal member, all in one batch
    zero = make(1, 2, 0, 3, dict(eps=1e-2, b=(0.3, 0.4, -0.5), f_val=0.0, supg=1))
    norm = make(1, 3, 0, 3, dict(eps=1.0, f_kind=1))
    hard = make(1, 5, 0, 3, dict(eps=1e-6, b=(0.48, -0.36, 0.64), f_val=1.0, supg=1))
    sp = pm.solver_params(**dict(SOLVER, maxit=4))
    st = pm.solve_batch([zero, norm, hard], sp)
    say(f"[gpu] batch statuses: {[s.status for s in st]}, its {[s.iterations for s in st]}")
    assert st[0].status == 0 and st[0].iterations == 0
    assert st[2].status == pm.CD3D_NOT_CONVERGED and st[2].iterations == 4
    _, lone = norm.solve(sp, check=False)
    assert st[1].status == lone.status and st[1].iterations == lone.iterations
    st = pm.solve_batch([norm, hard], pm.solver_params(**dict(SOLVER, div_factor=1e-6)))
    assert [s.status for s in st] == [pm.CD3D_DIVERGED] * 2 and [s.iterations for s in st] == [1, 1]
    for P in (zero, norm, hard):
        P.close()


def test_batch_refuses_mismatched_levels_and_duplicates():
    a = make(1, 2, 0, 3, dict(eps=1.0, f_val=1.0))
    b = make(1, 2, 0, 2, dict(eps=1.0, f_val=1.0))
    with pytest.raises(pm.CD3DError):
        pm.solve_batch([a, b], pm.solver_params(**SOLVER))
    with pytest.raises(pm.CD3DError):
        pm.solve_batch([a, a], pm.solver_params(**SOLVER))
    a.close()
    b.close()
