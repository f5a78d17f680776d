"""The clear-air forward call replayed from a captured hipGraph (option
"graph") gives the eager call's radiances bit for bit: on the call that
captures, on replays with new inputs in the same buffers (with eager work
in between), and when new rays overflow the captured LOS capacity (the
replay is detected and the call re-run eagerly)."""
import numpy as np
import pytest

from helpers import gpu_batch
from pyjurassic import lib as L
from pyjurassic import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def glib():
    return L.load()


def eager(glib, w, path):
    glib.jur_set_option(b"graph", 0)
    return gpu_batch(glib, w, path)


def graphed(glib, w, path):
    glib.jur_set_option(b"graph", 1)
    out = gpu_batch(glib, w, path)
    glib.jur_set_option(b"graph", 0)
    return out


def same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_graph_replays_match_eager(glib, tmp_path):
    w = synth.make("clear_batch_65536", tmp_path, nray=4096)
    ref = eager(glib, w, tmp_path / "e.ctl")
    same(graphed(glib, w, tmp_path / "g.ctl"), ref)   # eager + capture
    for k in range(3):
        w2 = w.select()
        w2.t = w.t + 0.5 * (k + 1)                      # new inputs, same shapes
        ref2 = eager(glib, w2, tmp_path / "e2.ctl")     # eager work in between
        out = graphed(glib, w2, tmp_path / "g2.ctl")    # replay
        assert np.isnan(glib.jur_stat(b"path_ms")) and glib.jur_stat(b"graph_ms") > 0
        same(out, ref2)
        assert not np.array_equal(out[0], ref[0])


def test_graph_capacity_overflow_falls_back(glib, tmp_path):
    w = synth.make("clear_batch_65536", tmp_path, nray=4096)
    nadir = np.where(w.geom[4] == 0)[0][:512]
    limb = np.where(w.geom[4] > 0)[0][:512]
    glib.jur_set_option(b"cap", 64)
    short = w.select(rays=nadir)
    ref_s = eager(glib, short, tmp_path / "e.ctl")
    same(graphed(glib, short, tmp_path / "g.ctl"), ref_s)  # captured with a nadir-sized capacity
    long_ = w.select(rays=limb)
    before = glib.jur_stat(b"graph_overflows")
    before = 0 if np.isnan(before) else before
    out_l = graphed(glib, long_, tmp_path / "g2.ctl")      # replay overflows -> eager call
    assert glib.jur_stat(b"graph_overflows") == before + 1
    same(eager(glib, long_, tmp_path / "e2.ctl"), out_l)
