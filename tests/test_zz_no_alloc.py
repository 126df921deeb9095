"""z4c_history_noalloc (the history sums with the per-tile partials in the
library's fixed device scratch) against z4c_history, through the C-ABI:
bit-identical columns on two calls in a row, the second on a smaller block
list (fewer tiles than the first left in the scratch), and a third call with
an r_max_con cut."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from z4c_amd import Params  # noqa: E402
from z4c_amd.z4c import Z4C_NCON, MeshBlockGeom, Z4c  # noqa: E402


def _random_inputs(z, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    u = torch.randn(22 * z.total_points, dtype=torch.float64, device="cuda", generator=gen)
    con = torch.randn(Z4C_NCON * z.total_points, dtype=torch.float64, device="cuda",
                      generator=gen)
    return u, con


def _both(z, u, con):
    out = torch.full((8,), float("nan"), dtype=torch.float64, device="cuda")
    z.history_noalloc(u, con, out)
    ref = z.history(u, con)  # waits for the stream
    return out.cpu().numpy(), np.array(ref)


def test_history_noalloc_equals_history():
    big = [MeshBlockGeom.cube(16, 0.1, (0.0, 0.0, 0.0)),
           MeshBlockGeom.cube(24, 0.05, (-1.0, 0.5, 0.0)),
           MeshBlockGeom.cube(32, 0.02, (0.3, -0.2, 0.1)),
           MeshBlockGeom.cube(12, 0.2, (-2.0, -2.0, -2.0))]
    small = [MeshBlockGeom.cube(8, 0.25, (0.5, 0.5, 0.5))]
    p = Params()
    za, zb = Z4c(big, p), Z4c(small, p)
    print(f"no-alloc history: {za.total_points} then {zb.total_points} padded points", flush=True)
    ua, ca = _random_inputs(za, 1)
    ub, cb = _random_inputs(zb, 2)
    got, ref = _both(za, ua, ca)
    assert np.array_equal(got, ref), (got, ref)
    assert np.all(ref > 0)
    got, ref = _both(zb, ub, cb)  # second call, smaller input
    assert np.array_equal(got, ref), (got, ref)
    assert np.all(ref > 0)
    zc = Z4c(big, Params(r_max_con=1.0))  # the cut drops part of the cells
    got, ref = _both(zc, ua, ca)
    assert np.array_equal(got, ref), (got, ref)
    full = np.array(za.history(ua, ca))
    assert np.all(ref < full)
    for z in (za, zb, zc):
        z.close()


def test_history_noalloc_rejects_null_output():
    g = [MeshBlockGeom.cube(8, 0.25, (0.0, 0.0, 0.0))]
    z = Z4c(g, Params())
    u, con = _random_inputs(z, 3)
    rc = z.L.z4c_history_noalloc(z._ctx, C.byref(z.params), C.c_void_p(u.data_ptr()),
                                 C.c_void_p(con.data_ptr()), None, None)
    assert rc == -1
    z.close()
