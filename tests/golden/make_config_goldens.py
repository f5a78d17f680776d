"""Oracle fixtures of the scattering configurations at sizes the oracle
needs minutes for (too long to run beside the GPU tests).

Each case selects rays of a seeded synthetic workload (pyjurassic.synth),
runs the FP64 oracle's batched forward model over them and writes
tests/golden/config/<case>.npz: rad, tau [nray][nd], tp [3][nray], npts
[nray], and `digest`, a SHA-256 of the inputs (profiles and ray geometry),
which the GPU tests recompute so that a change of the generator cannot go
unnoticed.  Run: python tests/golden/make_config_goldens.py [case ...]
"""
import hashlib
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path[:0] = [str(ROOT / "jurassic-scatter_amd"), str(ROOT / "oracle"), str(ROOT / "tests")]

from pyjurassic import synth  # noqa: E402


def _crossing(w, top, n_limb, n_nadir):
    limb = np.where((w.geom[4] > 0) & (w.geom[4] < top))[0][:n_limb]
    nadir = np.where((w.geom[4] == 0) & (w.geom[6] != 0))[0][:n_nadir]
    return np.concatenate([limb, nadir])


# case -> (workload, rays generated, selector, channels, emitters, fan, options)
# options are (key, value) pairs set on both libraries ("mie_grid": the
# configuration's radius / angle grid)
CASES = {
    # aerosol layer, single scattering, the configuration's 16 pencils per fan
    "aerosol_limb_64": ("aerosol_limb_8192", 256, lambda w: _crossing(w, 20.0, 64, 0), None, None, 16, ()),
    # cirrus, scattering depth 2, the reference's 28 pencils per fan
    "cirrus_8": ("cirrus_4096", 256, lambda w: _crossing(w, 12.0, 4, 4), 8, None, 28, ()),
    # the configurations' own parameters: the aerosol layer's Mie grid (64
    # log-spaced radii, 180 phase angles) and the cirrus's 32 pencils per fan
    "aerosol_limb_64_cfg": ("aerosol_limb_8192", 256, lambda w: _crossing(w, 20.0, 64, 0), None, None, 16,
                            (("mie_grid", 1),)),
    "cirrus_8_cfg": ("cirrus_4096", 256, lambda w: _crossing(w, 12.0, 4, 4), 8, None, 32, ()),
}


def case_workload(name, outdir):
    """The workload of fixture `name` (tables and files written to outdir)."""
    wl, n, sel, nd, ng, fan, opts = CASES[name]
    w = synth.make(wl, outdir, nray=n)
    return w.select(rays=sel(w), nd=nd, ng=ng), fan, opts


def digest(w):
    h = hashlib.sha256()
    for a in (w.z, w.p, w.t, w.q, w.geom, np.asarray(w.nu, dtype=np.float64)):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def make(name, tmp):
    import oracle_ctypes
    from helpers import oracle_batch
    olib = oracle_ctypes.load()
    w, fan, opts = case_workload(name, tmp)
    olib.oracle_set_fan(fan)
    for k, v in opts:
        getattr(olib, "oracle_set_" + k)(v)
    t0 = time.time()
    try:
        rad, tau, tp, npts = oracle_batch(olib, w, tmp / "o.ctl")
    finally:
        olib.oracle_set_fan(28)
        for k, _ in opts:
            getattr(olib, "oracle_set_" + k)(0)
    out = HERE / "config" / ("%s.npz" % name)
    out.parent.mkdir(exist_ok=True)
    np.savez_compressed(out, rad=rad, tau=tau, tp=tp, npts=npts, digest=np.array(digest(w)))
    print("%s: %d rays x %d channels, oracle %.1f s -> %s" % (name, w.nray, len(w.nu), time.time() - t0, out),
          flush=True)


if __name__ == "__main__":
    import tempfile
    names = sys.argv[1:] or list(CASES)
    for nm in names:
        with tempfile.TemporaryDirectory() as d:
            make(nm, Path(d))
