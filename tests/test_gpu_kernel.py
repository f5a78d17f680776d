"""kernel(): the finite-difference Jacobian (reference src/retrievalmodel.c:201-298)
on the GPU against the oracle's restatement, on the reference examples with
retrieval keys: temperature, pressure, all three emitters, extinction-layer
aerosol parameters (N, r, sigma: fresh Mie optics per column), single
scattering, hydrostatic balance with brightness temperatures, and a field of
view (convolution within each perturbed copy's rays).

The perturbed paths are integrated in FP64 (option kernel_fp64, default):
a 1 % perturbation of one level can change a radiance by less than FP32
resolves.  Bar per case: every entry within `tol` of its row's largest |K|
(a Jacobian row mixes columns of very different size) and the rows' dominant
entries (>= 10 % of the row maximum) within 10 tol relative.  The FP64 path
follows the reference's operation order on FP64 copies of everything the
reference keeps in double, so clear-air cases agree to ~1e-9; the
scattering case carries the FP32 angular integral of the source.
"""
import ctypes as C
import subprocess
import sys

import numpy as np
import pytest

import oracle_ctypes
from helpers import read_example
from pyjurassic import lib as L

pytestmark = pytest.mark.gpu

TQ = ("RETT_ZMIN", 0, "RETT_ZMAX", 80, "RETQ_ZMIN[0]", 0, "RETQ_ZMAX[0]", 80,
      "RETQ_ZMIN[1]", 0, "RETQ_ZMAX[1]", 30, "RETQ_ZMIN[2]", 10, "RETQ_ZMAX[2]", 50,
      "RETP_ZMIN", 10, "RETP_ZMAX", 20)
CASES = {
    "clear_t_p_q": ("clear", TQ, 1e-8),
    "clear_hydz_bbt": ("clear", ("HYDZ", 15, "WRITE_BBT", 1, "RETP_ZMIN", 10, "RETP_ZMAX", 20,
                                 "RETT_ZMIN", 0, "RETT_ZMAX", 40, "RETQ_ZMIN[1]", 0, "RETQ_ZMAX[1]", 20), 1e-5),
    "ext1_aerosol": ("ext1", ("RETT_ZMIN", 0, "RETT_ZMAX", 30, "RETNN", 1, "RETRR", 1, "RETSS", 1), 1e-6),
    "aero0_scattering": ("aero0", ("RETT_ZMIN", 0, "RETT_ZMAX", 20), 5e-5),
}


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def jacobian(lib, prefix, case, keys):
    ctl, obs, atm, aero = read_example(lib, prefix, case, keys)
    r = C.byref
    if prefix:
        n = lib.oracle_atm2x(r(ctl), r(atm), r(aero), None, None, None)
        m = lib.oracle_obs2y(r(ctl), r(obs), None, None, None)
        k = np.full((m, n), np.nan)
        lib.oracle_kernel(r(ctl), r(atm), r(obs), r(aero), _dp(k), m, n, n)
    else:
        n = lib.jur_state_size(r(ctl), r(atm), r(aero))
        m = lib.jur_meas_size(r(ctl), r(obs))
        k = np.full((m, n), np.nan)
        lib.jur_kernel(r(ctl), r(atm), r(obs), r(aero), _dp(k), m, n, n)
    rad = np.array([[obs.rad[id][ir] for id in range(ctl.nd)] for ir in range(obs.nr)])
    p = np.array(atm.p[:atm.np])
    return k, rad, p


def compare(kg, ko, label, tol):
    assert kg.shape == ko.shape and np.all(np.isfinite(kg))
    scale = np.max(np.abs(ko), axis=1, keepdims=True)
    scale[scale == 0] = 1
    err = np.abs(kg - ko) / scale
    big = np.abs(ko) >= 0.1 * scale
    rel = np.abs(kg[big] / ko[big] - 1)
    print("%s: %s, max |dK|/rowmax %.2e, dominant entries max rel %.2e" %
          (label, kg.shape, err.max(), rel.max()), flush=True)
    assert err.max() <= tol
    assert rel.max() <= 10 * tol
    # structural zeros agree exactly (levels a ray never reaches)
    assert np.array_equal(kg == 0, ko == 0) or np.sum((kg == 0) != (ko == 0)) <= 0.01 * kg.size


@pytest.mark.parametrize("name", sorted(CASES))
def test_kernel_vs_oracle(glib, olib, in_examples, name):
    case, keys, tol = CASES[name]
    print("case", name, flush=True)
    kg, rg, pg = jacobian(glib, "", case, keys)
    ko, ro, po = jacobian(olib, "oracle_", case, keys)
    compare(kg, ko, name, tol)
    # side effects of the unperturbed formod: radiances and hydrostatic p
    if "WRITE_BBT" in keys:
        np.testing.assert_allclose(rg, ro, atol=0.01)
    else:
        np.testing.assert_allclose(rg, ro, rtol=1e-4)
    np.testing.assert_allclose(pg, po, rtol=1e-12)


@pytest.mark.parametrize("case,tol", [("clear", 1e-11), ("ext1", 1e-10), ("aero0", 2e-7)])
def test_fp64_formod_vs_oracle(glib, olib, in_examples, case, tol):
    """The FP64 path integration (option fp64) against the oracle: the same
    operation order on the same data (FP64 layer optics included); what
    remains are the device's transcendental functions and, with scattering,
    the source J, which the angular integral forms in FP32."""
    ctl, obs, atm, aero = read_example(glib, "", case)
    glib.jur_set_option(b"fp64", 1)
    try:
        glib.formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
    finally:
        glib.jur_set_option(b"fp64", 0)
    octl, oobs, oatm, oaero = read_example(olib, "oracle_", case)
    olib.oracle_formod(C.byref(octl), C.byref(oatm), C.byref(oobs), C.byref(oaero))
    g = np.array([obs.rad[0][i] for i in range(obs.nr)])
    o = np.array([oobs.rad[0][i] for i in range(oobs.nr)])
    rel = np.max(np.abs(g / o - 1))
    print("%s fp64 max rel %.2e" % (case, rel), flush=True)
    assert rel <= tol


def test_kernel_with_fov(glib, olib, in_examples, tmp_path):
    fov = tmp_path / "fov.tab"
    dz = np.linspace(-2, 2, 41)
    fov.write_text("".join("%g %g\n" % (z, np.exp(-0.5 * (z / 0.8) ** 2)) for z in dz))
    keys = ("FOV", str(fov), "RETT_ZMIN", 0, "RETT_ZMAX", 40)
    kg, _, _ = jacobian(glib, "", "clear", keys)
    ko, _, _ = jacobian(olib, "oracle_", "clear", keys)
    compare(kg, ko, "fov", 1e-8)


def test_grouped_calls_match(glib, in_examples):
    """Perturbed copies split over many batched calls give the same matrix."""
    k1, _, _ = jacobian(glib, "", "clear", TQ)
    glib.jur_set_option(b"kernel_rays", 40)
    try:
        k2, _, _ = jacobian(glib, "", "clear", TQ)
    finally:
        glib.jur_set_option(b"kernel_rays", 1 << 17)
    d = np.abs(k1 - k2).max() / np.abs(k1).max()
    print("grouped vs single call: max diff / max |K| %.2e" % d, flush=True)
    np.testing.assert_array_equal(k1, k2)


def test_gsl_matrix_entry_point(glib, in_examples):
    ctl, obs, atm, aero = read_example(glib, "", "clear", ("RETT_ZMIN", 0, "RETT_ZMAX", 40))
    r = C.byref
    n = glib.jur_state_size(r(ctl), r(atm), r(aero))
    m = glib.jur_meas_size(r(ctl), r(obs))
    tda = n + 3  # a row stride wider than the matrix, as for a gsl submatrix view
    buf = np.full((m, tda), -7.0)
    mat = L.jur_gsl_matrix(m, n, tda, buf.ctypes.data, None, 0)
    glib.kernel(r(ctl), r(atm), r(obs), r(aero), r(mat))
    k, _, _ = jacobian(glib, "", "clear", ("RETT_ZMIN", 0, "RETT_ZMAX", 40))
    np.testing.assert_array_equal(buf[:, :n], k)
    assert np.all(buf[:, n:] == -7.0)


def test_size_mismatch_exits(in_examples):
    code = (
        "import ctypes as C, numpy as np, sys; sys.path[:0] = %r\n"
        "from pyjurassic import lib as L; from helpers import read_example\n"
        "lib = L.load()\n"
        "ctl, obs, atm, aero = read_example(lib, '', 'clear', ('RETT_ZMIN', 0, 'RETT_ZMAX', 5))\n"
        "k = np.zeros((11, 3))\n"
        "lib.jur_kernel(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero),"
        " k.ctypes.data_as(C.POINTER(C.c_double)), 11, 3, 3)\n" % (sys.path[:4],))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "Matrix size" in r.stdout


@pytest.mark.parametrize("case", ["clear", "ext1", "aero0"])
def test_fp64_emitter_lanes_bit_identical(glib, in_examples, case):
    """k_path64e (a ray's emitters on separate lanes, chosen for small FP64
    calls) performs k_path64's arithmetic in its order: FP64 radiances and
    transmittances bit-identical, clear air, extinction and scattering."""
    out = []
    glib.jur_set_option(b"fp64", 1)
    try:
        for split in (0, 1):
            glib.jur_set_option(b"fp64_split", split)
            ctl, obs, atm, aero = read_example(glib, "", case)
            glib.formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
            out.append(np.array([[obs.rad[id][ir], obs.tau[id][ir]] for id in range(ctl.nd)
                                 for ir in range(obs.nr)]))
    finally:
        glib.jur_set_option(b"fp64_split", -1)
        glib.jur_set_option(b"fp64", 0)
    print("%s: %d values, max |diff| %.3g" % (case, out[0].size, np.max(np.abs(out[0] - out[1]))), flush=True)
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("group", [1, 0])
@pytest.mark.parametrize("name", ["clear_t_p_q", "clear_hydz_bbt", "ext1_aerosol"])
def test_prefix_reuse_bit_identical(glib, in_examples, name, group):
    """A perturbed copy's ray repeats copy 0's up to its first differing LOS
    point (FP64 records compared bit for bit) and is integrated from there,
    starting from copy 0's path state; each emitter's factor 1 - eps and the
    point's clear-air optics come from copy 0 up to the first point whose
    inputs to them differ.  The Jacobian is bit-identical to the one
    integrated from the observer (option kernel_prefix 0), with the copies
    grouped by the emitters they compute or in plain order."""
    case, keys, _ = CASES[name]
    ks = []
    glib.jur_set_option(b"reuse_group", group)
    try:
        for pre in (0, 1):
            glib.jur_set_option(b"kernel_prefix", pre)
            k, _, _ = jacobian(glib, "", case, keys)
            ks.append((k, glib.jur_stat(b"integrated_points"), glib.jur_stat(b"emitter_points")))
    finally:
        glib.jur_set_option(b"kernel_prefix", 1)
        glib.jur_set_option(b"reuse_group", 1)
    ng = read_example(glib, "", case)[0].ng
    print("%s: LOS points integrated %d (reuse) vs %d, emitter evaluations %g vs %g"
          % (name, ks[1][1], ks[0][1], ks[1][2], ng * ks[0][1]), flush=True)
    np.testing.assert_array_equal(ks[1][0], ks[0][0])
    # the statistics are of the last batched call: an aerosol column (its own
    # call, no copies to reuse) in ext1_aerosol
    assert np.isnan(ks[0][2])
    if name != "ext1_aerosol":
        assert ks[1][1] < ks[0][1]
        assert ks[1][2] <= ng * ks[1][1]
    if name == "clear_t_p_q":
        # O3 copies take the CO2 and H2O chains from copy 0
        assert ks[1][2] < ng * ks[1][1]
