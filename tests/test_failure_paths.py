"""Failure paths of the C ABI, forced from the host so that every kernel
still runs to its normal end: the reference prints and calls exit(4711)
where the library returns a status (FgmresIte.C:509-512 divergence,
SystemCD3D.C:176 bad input); here each status is checked where it arises."""
import numpy as np
import pytest

import parmoon_cmg_amd as pm

CO = dict(eps=1e-2, b=(0.3, 0.4, -0.5), f_val=1.0, supg=1)


def small():
    return pm.Problem(1, 2, 1, 2, box=(1.0, 1.0, 1.0))


@pytest.mark.gpu
def test_nonfinite_rhs_is_reported_before_any_iteration():
    P = small()
    P.assemble(pm.coeffs(**CO))
    b = P.rhs(1)
    b[len(b) // 3] = np.nan
    x, st = P.solve(pm.solver_params(), rhs=b, check=False)
    assert st.status == pm.CD3D_NONFINITE and st.iterations == 0
    # the handle stays usable
    x, st = P.solve(pm.solver_params())
    assert st.status == pm.CD3D_OK and st.rfinal <= 1e-10 * st.r0
    P.close()


@pytest.mark.gpu
def test_divergence_test_reports_diverged():
    # SC_DIV_FACTOR below 1: the first residual estimate already exceeds it
    P = small()
    P.assemble(pm.coeffs(**CO))
    _, st = P.solve(pm.solver_params(div_factor=1e-6), check=False)
    assert st.status == pm.CD3D_DIVERGED and st.iterations == 1
    P.close()


@pytest.mark.gpu
def test_maxit_reports_not_converged():
    P = small()
    P.assemble(pm.coeffs(**CO))
    _, st = P.solve(pm.solver_params(maxit=3, rel_tol=1e-14), check=False)
    assert st.status == pm.CD3D_NOT_CONVERGED and st.iterations == 3
    assert 0 < st.rfinal < st.r0
    P.close()


@pytest.mark.gpu
def test_out_of_order_and_invalid_calls():
    P = small()
    with pytest.raises(pm.CD3DError) as e:
        P.solve(pm.solver_params())
    assert e.value.status == pm.CD3D_ERR_STATE  # solve before assemble
    with pytest.raises(pm.CD3DError) as e:
        P.assemble(pm.coeffs(eps=float("nan")))
    assert e.value.status == pm.CD3D_ERR_ARG
    with pytest.raises(pm.CD3DError) as e:
        P.assemble(pm.coeffs(**dict(CO, conv_kind=7)))
    assert e.value.status == pm.CD3D_ERR_ARG
    P.assemble(pm.coeffs(**CO))
    with pytest.raises(pm.CD3DError) as e:
        P.solve(pm.solver_params(smoother=5))
    assert e.value.status == pm.CD3D_ERR_ARG
    P.close()


@pytest.mark.gpu
def test_server_reports_a_failed_request_and_serves_the_next():
    srv = pm.Server(max_inflight=2)
    mesh = pm.mesh_desc(1, 2, 1, 2)
    sp = pm.solver_params(rel_tol=1e-10)
    bad = srv.submit(mesh, pm.coeffs(**dict(CO, eps=float("inf"))), sp)
    short = srv.submit(mesh, pm.coeffs(**CO), pm.solver_params(maxit=2, rel_tol=1e-14))
    good = srv.submit(mesh, pm.coeffs(**CO), sp)
    rb, _ = srv.wait(bad, check=False)
    rs, _ = srv.wait(short, check=False)
    rg, _ = srv.wait(good)
    assert rb.status == pm.CD3D_ERR_ARG
    assert rs.status == pm.CD3D_NOT_CONVERGED and rs.iterations == 2
    assert rg.status == pm.CD3D_OK and rg.rel_residual <= 1e-10
    srv.close()


@pytest.mark.gpu
def test_server_wait_times_out_and_the_request_stays():
    srv = pm.Server(max_inflight=1)
    rid = srv.submit(pm.mesh_desc(2, 1, 1, 4), pm.coeffs(**CO), pm.solver_params(rel_tol=1e-10))
    with pytest.raises(pm.CD3DError) as e:
        srv.wait(rid, timeout_ms=0.0)  # the 32^3 Q2 hierarchy is still being built
    assert e.value.status == pm.CD3D_ERR_STATE
    r, _ = srv.wait(rid)
    assert r.status == pm.CD3D_OK and r.rel_residual <= 1e-10
    srv.close()
