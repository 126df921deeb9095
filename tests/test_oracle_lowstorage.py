"""The low-storage integrators of the oracle (oracle.ls_step, the GR_Z4c
IntegrateZ4c sequence of src/task_list/gr/task_list_gr_z4c.cpp:170-194 over
the tableaus of src/task_list/task_list_integrators_low_storage.cpp:17-204),
pinned by what the methods are:

* the stability polynomial on y' = lambda y is the Taylor polynomial of
  exp(z) to the method's order, with degree = number of stages (vl2 / rk2:
  1 + z + z^2/2; rk3: + z^3/6; rk4 (Ketcheson RK4()4[2S]): + z^4/24; rk1:
  1 + z) -- every delta, gamma and beta enters these coefficients;
* on the nonlinear y' = y^2, y(0) = 1 (y = 1/(1 - t)) the error falls by
  2^order per halving of the step, which also checks the nonlinear order
  conditions the stability polynomial does not see;
* the Z4c step keeps ghosts out of the update (WeightedAve and AddZ4cRHS
  act on interior vertices) and a 2S method never reads u2.
"""
import math

import numpy as np
import pytest

import oracle as orc

ORDER = {"rk1": 1, "vl2": 2, "rk2": 2, "rk3": 3, "rk4": 4}


def _scalar_geom():
    return orc.Geom.block(1, 1.0, 0.0)  # 2 interior vertices per axis


def _run(name, y0, dt, nsteps, f):
    g = _scalar_geom()
    u = np.full((1,) + g.shape, y0, dtype=np.float64)
    for _ in range(nsteps):
        orc.ls_step(None, g, u, dt, name, rhs_fn=lambda v: f(v), constrain=False)
    return u[0, 4, 4, 4], u


@pytest.mark.parametrize("name", sorted(ORDER))
def test_stability_polynomial(name):
    st = orc.LS_TABLEAUS[name]
    m = len(st)
    zs = np.linspace(-0.6, 0.6, m + 1)
    R = np.array([_run(name, 1.0, 1.0, 1, lambda v, z=z: z * v)[0] for z in zs])
    coef = np.linalg.solve(np.vander(zs, m + 1, increasing=True), R)
    p = ORDER[name]
    taylor = np.array([1.0 / math.factorial(k) for k in range(p + 1)])
    np.testing.assert_allclose(coef[:p + 1], taylor, rtol=0, atol=1e-13)
    assert len(coef) == m + 1


@pytest.mark.parametrize("name", sorted(ORDER))
def test_nonlinear_convergence_order(name):
    T = 0.5
    errs = []
    for n in (20, 40, 80):
        y, _ = _run(name, 1.0, T / n, n, lambda v: v * v)
        errs.append(abs(y - 1.0 / (1.0 - T)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    p = ORDER[name]
    assert np.all(np.abs(rates - p) < 0.25), rates


def test_rk4_is_not_classical_rk4():
    """rk4 (Ketcheson 2S) and bt_rk4 (classical) share the stability
    polynomial but not the nonlinear error: one step of y' = y^2 differs."""
    y_ls, _ = _run("rk4", 1.0, 0.1, 1, lambda v: v * v)
    A, B = orc.BT_RK4_A, orc.BT_RK4_B
    k = []
    for s in range(4):
        y = 1.0 + 0.1 * sum(A[s][j] * k[j] for j in range(s))
        k.append(y * y)
    y_bt = 1.0 + 0.1 * sum(B[j] * k[j] for j in range(4))
    assert abs(y_ls - y_bt) > 1e-9


def test_ghosts_untouched_and_interior_updated():
    g = _scalar_geom()
    rng = np.random.default_rng(3)
    u = rng.standard_normal((1,) + g.shape)
    u0 = u.copy()
    orc.ls_step(None, g, u, 0.1, "rk4", rhs_fn=lambda v: -v, constrain=False)
    inner = (slice(None), slice(4, 6), slice(4, 6), slice(4, 6))
    ghost = np.ones(u.shape, bool)
    ghost[inner] = False
    np.testing.assert_array_equal(u[ghost], u0[ghost])
    R = 1 - 0.1 + 0.01 / 2 - 0.001 / 6 + 0.0001 / 24
    np.testing.assert_allclose(u[inner], R * u0[inner], rtol=1e-14)
