"""The oracle's finite-difference Jacobian (reference src/retrievalmodel.c:201-298).

The reference ships no kernel output, so the restatement is pinned by what
the definition fixes: each column is (F(x + h e_j) - F(x)) / h for formod on
an atmosphere perturbed here, independently of the oracle's state-vector
code, with the reference's perturbation sizes; the columns add up to the
response to a joint perturbation to first order; a ray's rows are exactly zero
for levels its line of sight never reaches; and with hydrostatic balance on,
pressure columns vanish except at the reference level.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes
from helpers import read_example
from pyjurassic import jtypes as J

# temperature everywhere, H2O below 30 km, pressure 10-20 km
KEYS = ("RETT_ZMIN", 0, "RETT_ZMAX", 80, "RETQ_ZMIN[1]", 0, "RETQ_ZMAX[1]", 30,
        "RETP_ZMIN", 10, "RETP_ZMAX", 20)


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def jacobian(olib, ctl, atm, obs, aero):
    r = C.byref
    n = olib.oracle_atm2x(r(ctl), r(atm), r(aero), None, None, None)
    # sizes after the unperturbed formod (hydrostatic may not change them)
    o = J.obs_t()
    C.memmove(C.addressof(o), C.addressof(obs), C.sizeof(obs))
    a = J.atm_t()
    C.memmove(C.addressof(a), C.addressof(atm), C.sizeof(atm))
    olib.oracle_formod(r(ctl), r(a), r(o), r(aero))
    m = olib.oracle_obs2y(r(ctl), r(o), None, None, None)
    k = np.full((m, n), np.nan)
    olib.oracle_kernel(r(ctl), r(atm), r(obs), r(aero), _dp(k), m, n, n)
    return k


def forward(olib, ctl, atm, obs, aero):
    """Finite radiances of formod on copies of atm / obs, ray-major."""
    a, o = J.atm_t(), J.obs_t()
    C.memmove(C.addressof(a), C.addressof(atm), C.sizeof(atm))
    C.memmove(C.addressof(o), C.addressof(obs), C.sizeof(obs))
    olib.oracle_formod(C.byref(ctl), C.byref(a), C.byref(o), C.byref(aero))
    y = np.array([[o.rad[id][ir] for id in range(ctl.nd)] for ir in range(o.nr)]).ravel()
    return y[np.isfinite(y)], o


def test_state_vector_layout(olib, in_examples):
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", KEYS)
    n = olib.oracle_atm2x(C.byref(ctl), C.byref(atm), C.byref(aero), None, None, None)
    x = np.zeros(n)
    iqa, ipa = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    olib.oracle_atm2x(C.byref(ctl), C.byref(atm), C.byref(aero), _dp(x), ip(iqa), ip(ipa))
    z = np.array(atm.z[:atm.np])
    lev_p = np.where((z >= 10) & (z <= 20))[0]
    lev_t = np.where(z <= 80)[0]
    lev_q = np.where(z <= 30)[0]
    assert n == len(lev_p) + len(lev_t) + len(lev_q)
    # pressure (0), temperature (1), then H2O as emitter 1 (index 2 + 1)
    assert list(iqa) == [0] * len(lev_p) + [1] * len(lev_t) + [3] * len(lev_q)
    assert list(ipa) == list(lev_p) + list(lev_t) + list(lev_q)
    assert x[0] == atm.p[lev_p[0]] and x[len(lev_p)] == atm.t[0]
    assert x[-1] == atm.q[1][lev_q[-1]]
    # x2atm inverts atm2x
    a2 = J.atm_t()
    C.memmove(C.addressof(a2), C.addressof(atm), C.sizeof(atm))
    x2 = x + 1.0
    olib.oracle_x2atm(C.byref(ctl), _dp(x2), C.byref(a2), C.byref(aero))
    assert a2.t[5] == atm.t[5] + 1 and a2.p[lev_p[0]] == atm.p[lev_p[0]] + 1
    assert a2.q[1][40] == atm.q[1][40] and a2.q[0][3] == atm.q[0][3]


def test_perturbation_sizes(olib, in_examples):
    ctl, *_ = read_example(olib, "oracle_", "clear", KEYS)
    h = lambda iq, x: olib.oracle_perturbation(C.byref(ctl), iq, x)
    ng, nw = ctl.ng, ctl.nw
    assert h(0, 500.0) == 5.0 and h(0, 1e-6) == 1e-7          # p: 1 %, floor 1e-7 hPa
    assert h(1, 250.0) == 1.0                                    # T: 1 K
    assert h(2, 4e-4) == pytest.approx(4e-6, rel=1e-15) and h(2 + ng - 1, 0.0) == 1e-15
    assert h(2 + ng, 0.3) == 1e-4                               # extinction
    assert h(2 + ng + nw + 1, 10.0) == 5.0 and h(2 + ng + nw + 1, 0.01) == 0.1   # N
    assert h(2 + ng + nw + 2, 2.0) == 0.5 and h(2 + ng + nw + 3, 1.5) == pytest.approx(0.15)
    assert h(2 + ng + nw + 4, 1.0) < 0                           # no such quantity


def test_columns_match_independent_perturbation(olib, in_examples):
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", KEYS)
    k = jacobian(olib, ctl, atm, obs, aero)
    y0, _ = forward(olib, ctl, atm, obs, aero)
    z = np.array(atm.z[:atm.np])
    lev_p = np.where((z >= 10) & (z <= 20))[0]
    np_, nt = len(lev_p), atm.np
    cases = [(0, "p", lev_p[0], 0.01 * atm.p[lev_p[0]]),
             (np_ + 8, "t", 8, 1.0),
             (np_ + 20, "t", 20, 1.0),
             (np_ + nt + 12, "q", 12, 0.01 * atm.q[1][12])]
    for j, what, lev, h in cases:
        a = J.atm_t()
        C.memmove(C.addressof(a), C.addressof(atm), C.sizeof(atm))
        if what == "p":
            a.p[lev] += h
        elif what == "t":
            a.t[lev] += h
        else:
            a.q[1][lev] += h
        y1, _ = forward(olib, ctl, a, obs, aero)
        np.testing.assert_array_equal(k[:, j], (y1 - y0) / h, err_msg="column %d" % j)
        assert np.any(k[:, j] != 0)


def test_first_order_linearity(olib, in_examples):
    """sum_j K_ij h_j over all temperature levels matches F(T + 1 K) - F(T)."""
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", ("RETT_ZMIN", 0, "RETT_ZMAX", 80))
    k = jacobian(olib, ctl, atm, obs, aero)
    y0, _ = forward(olib, ctl, atm, obs, aero)
    a = J.atm_t()
    C.memmove(C.addressof(a), C.addressof(atm), C.sizeof(atm))
    for ip in range(a.np):
        a.t[ip] += 1.0
    y1, _ = forward(olib, ctl, a, obs, aero)
    lin = k.sum(axis=1)
    np.testing.assert_allclose(lin, y1 - y0, rtol=0.05)
    assert np.all(np.abs(lin) > 0)


def test_rows_zero_below_tangent(olib, in_examples):
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", ("RETT_ZMIN", 0, "RETT_ZMAX", 80))
    k = jacobian(olib, ctl, atm, obs, aero)
    _, o = forward(olib, ctl, atm, obs, aero)
    z = np.array(atm.z[:atm.np])
    for ir in range(o.nr):
        below = z < o.tpz[ir] - 1.01      # levels that bracket no LOS point
        above = (z > o.tpz[ir] + 1) & (z < 60)
        assert np.all(k[ir, below] == 0), ir
        assert np.all(k[ir, above] != 0), ir


def test_hydrostatic_pressure_columns(olib, in_examples):
    """With HYDZ the perturbed pressure is rebuilt from the reference level:
    only that level's column is non-zero."""
    keys = ("HYDZ", 15, "RETP_ZMIN", 10, "RETP_ZMAX", 20)
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", keys)
    k = jacobian(olib, ctl, atm, obs, aero)
    z = np.array(atm.z[:atm.np])
    lev = np.where((z >= 10) & (z <= 20))[0]
    ref = list(lev).index(int(np.argmin(np.abs(z - 15))))
    nz = [j for j in range(len(lev)) if np.any(k[:, j] != 0)]
    assert nz == [ref]


def test_size_mismatch_exits(olib, in_examples, tmp_path):
    import subprocess
    import sys
    code = (
        "import ctypes as C, numpy as np, sys; sys.path[:0] = %r\n"
        "import oracle_ctypes; from helpers import read_example\n"
        "lib = oracle_ctypes.load()\n"
        "ctl, obs, atm, aero = read_example(lib, 'oracle_', 'clear', ('RETT_ZMIN', 0, 'RETT_ZMAX', 5))\n"
        "k = np.zeros((11, 3))\n"
        "lib.oracle_kernel(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero),"
        " k.ctypes.data_as(C.POINTER(C.c_double)), 11, 3, 3)\n" % (sys.path[:4],))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode != 0 and "Matrix size" in r.stdout


def test_column_subset(olib, in_examples):
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", ("RETT_ZMIN", 0, "RETT_ZMAX", 40))
    k = jacobian(olib, ctl, atm, obs, aero)
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", ("RETT_ZMIN", 0, "RETT_ZMAX", 40))
    m, n = k.shape
    cols = np.array([0, 7, 19, n - 1], dtype=np.int32)
    ks = np.full((m, n), np.nan)
    olib.oracle_kernel_columns(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero), _dp(ks), m, n, n,
                               cols.ctypes.data_as(C.POINTER(C.c_int)), len(cols))
    np.testing.assert_array_equal(ks[:, cols], k[:, cols])
    rest = np.setdiff1d(np.arange(n), cols)
    assert np.all(ks[:, rest] == 0)
