"""The device-resident scattering call: a later call of the same shape runs
from the sizes of the first (eager) call with a margin, keeping every count
on the device (stat sca_mode 1), and with option "graph" is recorded once and
replayed (sca_mode 2).  Both give the eager call's radiances bit for bit, on
the inputs they were planned with and on new inputs in the same buffers; a
call with more scattering points than planned is detected and run eagerly
(sca_mode 0, stat plan_overflows)."""
import numpy as np
import pytest

from helpers import gpu_batch
from pyjurassic import lib as L
from pyjurassic import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def glib():
    return L.load()


def _crossing(w, top, n_limb, n_nadir):
    limb = np.where((w.geom[4] > 0) & (w.geom[4] < top))[0][:n_limb]
    nadir = np.where((w.geom[4] == 0) & (w.geom[6] != 0))[0][:n_nadir]
    return np.concatenate([limb, nadir])


def same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def call(glib, w, path, graph=0, plan=1):
    glib.jur_set_option(b"graph", graph)
    glib.jur_set_option(b"sca_plan", plan)
    try:
        out = gpu_batch(glib, w, path)
    finally:
        glib.jur_set_option(b"graph", 0)
        glib.jur_set_option(b"sca_plan", 1)
    return out, int(glib.jur_stat(b"sca_mode")), glib.jur_stat(b"pencils")


def overflows(glib):
    v = glib.jur_stat(b"plan_overflows")
    return 0 if np.isnan(v) else int(v)


@pytest.mark.parametrize("workload,top,nl,nn,nd,ng,fan", [("aerosol_limb_8192", 20.0, 48, 0, 8, None, 16),
                                                          ("cirrus_4096", 12.0, 1, 1, 2, 2, 28)])
def test_planned_and_replayed_calls_match_eager(glib, tmp_path, workload, top, nl, nn, nd, ng, fan):
    w = synth.make(workload, tmp_path, nray=512)
    w = w.select(rays=_crossing(w, top, nl, nn), nd=nd, ng=ng)
    print("%s: %d rays, fan %d" % (workload, w.nray, fan), flush=True)
    glib.jur_set_option(b"fan", fan)
    try:
        ref, mode, pen = call(glib, w, tmp_path / "a.ctl", plan=0)
        assert mode == 0
        e1, m1, p1 = call(glib, w, tmp_path / "b.ctl")           # eager, makes the plan
        p2, m2, pp2 = call(glib, w, tmp_path / "c.ctl")          # planned
        g1, mg1, _ = call(glib, w, tmp_path / "d.ctl", graph=1)  # planned + recorded
        g2, mg2, pg2 = call(glib, w, tmp_path / "e.ctl", graph=1)  # replayed
        print("modes", m1, m2, mg1, mg2, "pencils", pen, pp2, pg2, flush=True)
        assert m1 in (0, 1) and (m2, mg1, mg2) == (1, 1, 2)
        assert pen == p1 == pp2 == pg2
        for o in (e1, p2, g1, g2):
            same(o, ref)
        # new inputs in the same buffers: replayed against eager
        w2 = w.select()
        w2.t = w.t + 1.5
        ref2, _, _ = call(glib, w2, tmp_path / "f.ctl", plan=0)
        g3, mg3, _ = call(glib, w2, tmp_path / "g.ctl", graph=1)
        assert mg3 == 2
        same(g3, ref2)
        assert not np.array_equal(g3[0], ref[0])
    finally:
        glib.jur_set_option(b"fan", 28)


def test_planned_call_outgrowing_its_plan_runs_eagerly(glib, tmp_path):
    """Planned on rays that barely touch the aerosol layer, then called with
    as many rays through its middle: more points than planned."""
    w = synth.make("aerosol_limb_8192", tmp_path, nray=2048)
    th = w.geom[4]
    graze = np.where((th > 19.5) & (th < 22.0))[0][:32]
    deep = np.where((th > 8.0) & (th < 14.0))[0][:32]
    assert len(graze) == len(deep) == 32
    wa, wb = w.select(rays=graze, nd=4), w.select(rays=deep, nd=4)
    glib.jur_set_option(b"fan", 16)
    try:
        call(glib, wa, tmp_path / "a.ctl")
        call(glib, wa, tmp_path / "b.ctl")
        before = overflows(glib)
        got, mode, _ = call(glib, wb, tmp_path / "c.ctl")
        assert mode == 0 and overflows(glib) == before + 1
        ref, _, _ = call(glib, wb, tmp_path / "d.ctl", plan=0)
        same(got, ref)
    finally:
        glib.jur_set_option(b"fan", 28)
