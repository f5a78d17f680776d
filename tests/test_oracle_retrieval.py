"""The oracle's optimal estimation (reference src/retrievalmodel.c:407-706).

The reference ships no retrieval output, so the restatement is pinned by what
the method fixes, computed here with numpy's own linear algebra on the
oracle's forward model and Jacobian:
  * the a priori covariance: sigma_i sigma_j exp(-|dz|/cz - d_h/ch) within a
    quantity (d_h the chord between the two points at the surface), zero
    across quantities, percent errors for pressure and volume mixing ratios;
  * the measurement errors: noise and a percentage of |y|, added in squares;
  * the cost function chi^2/m = (dy^T S_eps^-1 dy + dx^T S_a^-1 dx) / m;
  * one Levenberg-Marquardt step from the a priori:
    x_1 = x_a + ((1 + 0.001) S_a^-1 + K^T S_eps^-1 K)^-1 K^T S_eps^-1 dy;
  * the retrieval errors: sqrt diag (S_a^-1 + K^T S_eps^-1 K)^-1;
  * a nearly linear problem converges to the maximum a posteriori solution.
A wrong sign, a dropped (1 + lmpar), a transposed product or a missing a
priori term fails one of these.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes
from helpers import read_example
from pyjurassic import jtypes as J

RE = 6367.421  # Earth radius of the reference's geometry [km]
TKEYS = ("RETT_ZMIN", 5, "RETT_ZMAX", 30, "ERR_TEMP", 5, "ERR_TEMP_CZ", 3, "ERR_TEMP_CH", 500,
         "ERR_NOISE[0]", 1e-6, "ERR_FORMOD[0]", 0.5, "CONV_DMIN", 1e-6)


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _copy(s):
    d = type(s)()
    C.memmove(C.addressof(d), C.addressof(s), C.sizeof(s))
    return d


def problem(olib, keys, bump=3.0):
    """The clear example with a temperature bump around 18 km as the truth:
    ctl, ret, a priori atmosphere, measurements (oracle formod of the truth)."""
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", keys)
    ret = J.ret_t()
    argc, argv = J.argv("clear-air.ctl", *keys)
    olib.oracle_read_ret(argc, argv, C.byref(ctl), C.byref(ret))
    tru = _copy(atm)
    for ip in range(atm.np):
        tru.t[ip] += bump * np.exp(-((atm.z[ip] - 18) / 5) ** 2)
    meas = _copy(obs)
    olib.oracle_formod(C.byref(ctl), C.byref(tru), C.byref(meas), C.byref(aero))
    return ctl, ret, atm, meas, aero


def sizes(olib, ctl, atm, obs, aero):
    r = C.byref
    return (olib.oracle_obs2y(r(ctl), r(obs), None, None, None),
            olib.oracle_atm2x(r(ctl), r(atm), r(aero), None, None, None))


def y_of(olib, ctl, obs, m):
    y = np.zeros(m)
    olib.oracle_obs2y(C.byref(ctl), C.byref(obs), _dp(y), None, None)
    return y


def x_of(olib, ctl, atm, aero, n):
    x = np.zeros(n)
    olib.oracle_atm2x(C.byref(ctl), C.byref(atm), C.byref(aero), _dp(x), None, None)
    return x


def forward(olib, ctl, atm, obs, aero, m):
    a, o = _copy(atm), _copy(obs)
    olib.oracle_formod(C.byref(ctl), C.byref(a), C.byref(o), C.byref(aero))
    return y_of(olib, ctl, o, m)


def jacobian(olib, ctl, atm, obs, aero, m, n):
    k = np.zeros((m, n))
    olib.oracle_kernel(C.byref(ctl), C.byref(_copy(atm)), C.byref(_copy(obs)), C.byref(aero), _dp(k), m, n, n)
    return k


def covariances(olib, ret, ctl, atm, meas, aero, m, n):
    s_a = np.zeros((n, n))
    olib.oracle_cov_apr(C.byref(ret), C.byref(ctl), C.byref(atm), C.byref(aero), _dp(s_a))
    sn, sf, se = np.zeros(m), np.zeros(m), np.zeros(m)
    olib.oracle_cov_meas(C.byref(ret), C.byref(ctl), C.byref(meas), _dp(sn), _dp(sf), _dp(se))
    return s_a, sn, sf, se


def retrieve(olib, ret, ctl, meas, atm, aero, n):
    obs_i, atm_i, aero_i = J.obs_t(), J.atm_t(), J.aero_t()
    chisq = np.zeros(1 + 20 * max(ret.conv_itmax, 0))
    nc = C.c_int(0)
    err = np.zeros(n)
    olib.oracle_optimal_estimation(C.byref(ret), C.byref(ctl), C.byref(meas), C.byref(obs_i), C.byref(atm),
                                   C.byref(atm_i), C.byref(aero), C.byref(aero_i), _dp(chisq), C.byref(nc),
                                   _dp(err))
    return atm_i, obs_i, chisq[:nc.value], err


def test_read_ret_defaults(olib, in_examples):
    ctl, *_ = read_example(olib, "oracle_", "clear")
    ret = J.ret_t()
    argc, argv = J.argv("clear-air.ctl")
    olib.oracle_read_ret(argc, argv, C.byref(ctl), C.byref(ret))
    assert (ret.kernel_recomp, ret.conv_itmax, ret.conv_dmin, ret.resmax, ret.err_ana) == (1, 20, 0.1, -999, 1)
    assert ret.err_temp == 0 and ret.err_temp_cz == -999 and ret.err_q_ch[2] == -999


def test_cov_apr_closed_form(olib, in_examples):
    keys = ("RETT_ZMIN", 5, "RETT_ZMAX", 12, "RETQ_ZMIN[1]", 0, "RETQ_ZMAX[1]", 4, "RETP_ZMIN", 20,
            "RETP_ZMAX", 22, "ERR_TEMP", 2, "ERR_TEMP_CZ", 3, "ERR_TEMP_CH", 400, "ERR_Q[1]", 50,
            "ERR_PRESS", 10, "ERR_PRESS_CZ", 5, "ERR_PRESS_CH", -999)
    ctl, obs, atm, aero = read_example(olib, "oracle_", "clear", keys)
    ret = J.ret_t()
    argc, argv = J.argv("clear-air.ctl", *keys)
    olib.oracle_read_ret(argc, argv, C.byref(ctl), C.byref(ret))
    # move level 9 one degree north: the correlation then also decays with
    # the chord 2 RE sin(0.5 deg) between the two surface points
    atm.lat[9] += 1.0
    m, n = sizes(olib, ctl, atm, obs, aero)
    s_a = np.zeros((n, n))
    olib.oracle_cov_apr(C.byref(ret), C.byref(ctl), C.byref(atm), C.byref(aero), _dp(s_a))
    z = np.array(atm.z[:atm.np])
    lp, lt, lq = np.where((z >= 20) & (z <= 22))[0], np.where((z >= 5) & (z <= 12))[0], np.where(z <= 4)[0]
    ip, it, iq = slice(0, 3), slice(3, 3 + len(lt)), slice(3 + len(lt), n)
    assert n == len(lp) + len(lt) + len(lq)
    # pressure: 10 % errors, no correlation (no horizontal length given)
    p = np.array(atm.p[:atm.np])[lp]
    np.testing.assert_allclose(s_a[ip, ip], np.diag((0.1 * p) ** 2), rtol=1e-15)
    # temperature: 2 K, exp(-|dz| / 3 km) and the horizontal chord for level 9
    chord = 2 * RE * np.sin(np.radians(0.5))
    zt = z[lt]
    want = 4.0 * np.exp(-np.abs(zt[:, None] - zt[None, :]) / 3)
    k9 = list(lt).index(9)
    want[k9, :] *= np.exp(-chord / 400)
    want[:, k9] *= np.exp(-chord / 400)
    want[k9, k9] = 4.0
    np.testing.assert_allclose(s_a[it, it], want, rtol=1e-12)
    # H2O: 50 % errors, uncorrelated; no covariance across quantities
    q = np.array(atm.q[1][:atm.np])[lq]
    np.testing.assert_allclose(s_a[iq, iq], np.diag((0.5 * q) ** 2), rtol=1e-15)
    assert np.all(s_a[ip, it] == 0) and np.all(s_a[it, iq] == 0) and np.all(s_a[ip, iq] == 0)


def test_cov_meas_closed_form(olib, in_examples):
    ctl, ret, atm, meas, aero = problem(olib, TKEYS)
    meas.rad[0][4] = float("nan")  # a masked measurement leaves the vector
    m, n = sizes(olib, ctl, atm, meas, aero)
    assert m == meas.nr - 1
    _, sn, sf, se = covariances(olib, ret, ctl, atm, meas, aero, m, n)
    y = y_of(olib, ctl, meas, m)
    np.testing.assert_array_equal(sn, 1e-6)
    np.testing.assert_allclose(sf, 0.005 * np.abs(y), rtol=1e-15)
    np.testing.assert_allclose(se, 1 / np.hypot(sn, sf), rtol=1e-14)


def test_first_step_and_costs(olib, in_examples):
    keys = TKEYS + ("CONV_ITMAX", 1)
    ctl, ret, atm, meas, aero = problem(olib, keys)
    m, n = sizes(olib, ctl, atm, meas, aero)
    s_a, _, _, se = covariances(olib, ret, ctl, atm, meas, aero, m, n)
    y_m = y_of(olib, ctl, meas, m)
    x_a = x_of(olib, ctl, atm, aero, n)
    dy0 = y_m - forward(olib, ctl, atm, meas, aero, m)
    k = jacobian(olib, ctl, atm, meas, aero, m, n)
    atm_i, _, chisq, err = retrieve(olib, ret, ctl, meas, atm, aero, n)
    # the initial cost has no a priori term
    np.testing.assert_allclose(chisq[0], np.sum((dy0 * se) ** 2) / m, rtol=1e-12)
    sai = np.linalg.inv(s_a)
    w = se ** 2
    a = 1.001 * sai + k.T @ (w[:, None] * k)
    x1 = x_a + np.linalg.solve(a, k.T @ (w * dy0))
    assert len(chisq) == 2 and chisq[1] < chisq[0]  # the step is taken
    x_i = x_of(olib, ctl, atm_i, aero, n)
    np.testing.assert_allclose(x_i - x_a, x1 - x_a, rtol=1e-8, atol=1e-10 * np.max(np.abs(x1 - x_a)))
    # cost at x_1 from a fresh forward run
    a1 = _copy(atm)
    for j, ip in enumerate(np.where((np.array(atm.z[:atm.np]) >= 5) & (np.array(atm.z[:atm.np]) <= 30))[0]):
        a1.t[ip] = x1[j]
    dy1 = y_m - forward(olib, ctl, a1, meas, aero, m)
    dx1 = x1 - x_a
    np.testing.assert_allclose(chisq[1], (np.sum((dy1 * se) ** 2) + dx1 @ sai @ dx1) / m, rtol=1e-8)
    # error analysis with the Jacobian of the a priori state
    cov = np.linalg.inv(sai + k.T @ (w[:, None] * k))
    np.testing.assert_allclose(err, np.sqrt(np.diag(cov)), rtol=1e-10)


def test_rejected_step_raises_damping(olib, in_examples):
    """A step that raises the cost is undone and retried with ten times the
    damping: with a wrong-signed measurement the first tries fail."""
    keys = TKEYS + ("CONV_ITMAX", 1)
    ctl, ret, atm, meas, aero = problem(olib, keys, bump=60.0)
    m, n = sizes(olib, ctl, atm, meas, aero)
    _, chisq, _ = retrieve(olib, ret, ctl, meas, atm, aero, n)[1:]
    assert len(chisq) >= 2
    # every try but the last raised the cost above the starting value
    assert np.all(chisq[1:-1] > chisq[0]) and chisq[-1] <= chisq[0]


def test_converged_state_is_stationary(olib, in_examples):
    """At convergence the gradient of the cost,
    K^T S_eps^-1 (y - F(x)) - S_a^-1 (x - x_a), vanishes (with K the
    Jacobian at the converged state) against its size at the a priori."""
    # a priori of 1 K against a 3 K truth: both terms of the gradient count
    keys = list(TKEYS)
    keys[keys.index("ERR_TEMP") + 1] = 1
    ctl, ret, atm, meas, aero = problem(olib, tuple(keys))
    assert ret.err_temp == 1
    m, n = sizes(olib, ctl, atm, meas, aero)
    s_a, _, _, se = covariances(olib, ret, ctl, atm, meas, aero, m, n)
    sai, w = np.linalg.inv(s_a), se ** 2
    x_a = x_of(olib, ctl, atm, aero, n)
    y_m = y_of(olib, ctl, meas, m)
    g0 = jacobian(olib, ctl, atm, meas, aero, m, n).T @ (w * (y_m - forward(olib, ctl, atm, meas, aero, m)))
    atm_i, _, chisq, _ = retrieve(olib, ret, ctl, meas, atm, aero, n)
    assert len(chisq) < 1 + 20 * ret.conv_itmax and chisq[-1] < 0.05 * chisq[0]
    x_i = x_of(olib, ctl, atm_i, aero, n)
    dy = y_m - forward(olib, ctl, atm_i, meas, aero, m)
    g = jacobian(olib, ctl, atm_i, meas, aero, m, n).T @ (w * dy) - sai @ (x_i - x_a)
    rel = np.linalg.norm(g) / np.linalg.norm(g0)
    print("gradient at convergence %.2e of the initial one after %d costs" % (rel, len(chisq)))
    assert rel < 1e-4
