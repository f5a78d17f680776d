"""Every variant of the FP32 path kernel gives the same radiances, bit for bit.

The variants differ in work mapping (channels per wave, waves per block,
occupancy target) and, for the staged kernel (variants 7-10), in how the
table brackets are found (grids in LDS, per-corner hints, the four corners
searched in lockstep; variant 10: k_pack copies each 64-ray group's point
records lane-contiguously and computes the (p, T) brackets and weights once
per LOS point and emitter); the arithmetic is the same, so radiances and
transmittances must be identical to the default fast kernel's (variant 6)
on every ray and channel.  Cases: a
4,096-ray slice of the 65,536-ray batch (synthetic tables, 5 emitters, 64
channels), the reference's clear-air example (tables with 134-239 column
densities per column, 40 pressures x 29 temperatures) and single scattering
on the aerosol configuration (the source term branch).
"""
import ctypes as C

import numpy as np
import pytest

from helpers import gpu_batch, read_example
from pyjurassic import lib as L
from pyjurassic import synth

pytestmark = pytest.mark.gpu

VARIANTS = list(range(11))


@pytest.fixture(scope="module")
def glib():
    return L.load()


def _with_variant(glib, v, wpb, fn):
    glib.jur_set_option(b"variant", float(v))
    glib.jur_set_option(b"wpb", float(wpb))
    try:
        return fn()
    finally:
        glib.jur_set_option(b"variant", -1.0)
        glib.jur_set_option(b"wpb", -1.0)


def _same(a, b, what):
    for x, y in zip(a, b):
        x, y = np.asarray(x), np.asarray(y)
        if not np.array_equal(x, y, equal_nan=True):
            d = np.nanmax(np.abs(x - y))
            raise AssertionError("%s: outputs differ (max abs difference %.3e)" % (what, d))


@pytest.mark.parametrize("workload,nray", [("clear_batch_65536", 4096), ("aerosol_limb_8192", 48)])
def test_variants_bit_identical_batch(glib, tmp_path, workload, nray):
    w = synth.make(workload, tmp_path, nray=nray)
    ref = _with_variant(glib, 6, 4, lambda: gpu_batch(glib, w, tmp_path / "g.ctl"))
    assert np.all(np.isfinite(ref[0]))
    for v in VARIANTS:
        for wpb in (1, 2, 4):
            print("%s[:%d] variant %d, %d waves per block" % (workload, nray, v, wpb), flush=True)
            got = _with_variant(glib, v, wpb, lambda: gpu_batch(glib, w, tmp_path / "g.ctl"))
            _same(got[:2], ref[:2], "variant %d wpb %d" % (v, wpb))


def test_variants_bit_identical_reference_tables(glib, in_examples):
    """The reference's example tables: columns of different lengths, larger
    grids staged in LDS."""
    def run():
        ctl, obs, atm, aero = read_example(glib, "", "clear")
        glib.formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
        return (np.array([[obs.rad[i][j] for j in range(obs.nr)] for i in range(ctl.nd)]),
                np.array([[obs.tau[i][j] for j in range(obs.nr)] for i in range(ctl.nd)]))
    ref = _with_variant(glib, 6, 4, run)
    for v in VARIANTS:
        print("clear example, variant %d" % v, flush=True)
        _same(_with_variant(glib, v, 4, run), ref, "variant %d" % v)
