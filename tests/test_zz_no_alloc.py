"""cd3d_matvec_noalloc (staged through the library's per-device array, no
device allocation) against cd3d_matvec (staging buffers allocated per call):
bitwise the same products, on two calls in a row with the second on a
smaller level, and a level too large for the array is refused."""
import sys

import numpy as np
import pytest

import parmoon_cmg_amd as pm

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_matvec_equals_matvec():
    say("[gpu] no-alloc matvec: Q2 SUPG, 8^3 cells, 3 levels")
    P = pm.Problem(2, 1, 1, 3, box=(1.0, 1.0, 1.0))
    P.assemble(pm.coeffs(eps=1e-3, b=(1.0, 0.5, 0.25), f_val=1.0, supg=1))
    rng = np.random.default_rng(5)
    for l in (2, 0):  # the second call on a smaller level
        n = P.info(l)["ndof"]
        x = rng.standard_normal(n)
        y_ref = P.matvec(l, x)
        y = P.matvec_noalloc(l, x)
        say(f"[gpu] level {l}: {n} unknowns, max |diff| {np.abs(y - y_ref).max():.1e}")
        assert np.array_equal(y, y_ref)
    P.close()


def test_noalloc_matvec_refuses_a_level_beyond_the_array():
    # Q1, 176^3 cells on the finest of 5 levels (11^3 macro bricks):
    # 177^3 = 5 545 233 unknowns > 4 194 304
    say("[gpu] no-alloc matvec: Q1 176^3 level (too large for the staging array)")
    P = pm.Problem(1, 11, 0, 5, box=(1.0, 1.0, 1.0))
    x = np.zeros(P.info(4)["ndof"])
    with pytest.raises(pm.CD3DError) as e:
        P.matvec_noalloc(4, x)
    assert e.value.status == pm.CD3D_ERR_ARG
    P.close()
