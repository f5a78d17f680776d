"""jur_formod_batch_device_noalloc (work buffers in a fixed device array, no
allocation) against jur_formod_batch_device on the same device inputs: two
no-allocation calls in a row, the second with a smaller batch (the array
then holds the first call's contents), for clear air and for aerosol
extinction (layer-edge points, scattering depth 0).  Every output element
must equal the allocating call's (same kernels on the same inputs).
"""
import ctypes as C

import numpy as np
import pytest

from helpers import read_workload
from pyjurassic import lib as L
from pyjurassic import synth

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself
    runs on (already loaded with it), for device inputs and outputs."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.live.append(p)
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def device_batch(hip, w, nd, rays):
    """jur_batch_t over device copies of the rays `rays` of w; the outputs
    start as NaN (and -1 LOS points)."""
    sel = w.select(rays=rays)
    n = sel.nray
    b = L.jur_batch_t()
    b.nray, b.nlev, b.natm = n, sel.nlev, sel.natm
    b.atm_of_ray = None
    f = lambda a: np.asarray(a, dtype=np.float64)
    b.z, b.p, b.t, b.q = [hip.put(f(x)) for x in (sel.z, sel.p, sel.t, sel.q)]
    b.k = None
    (b.time, b.obsz, b.obslon, b.obslat, b.vpz, b.vplon, b.vplat) = [hip.put(f(sel.geom[i])) for i in range(7)]
    outs = {"rad": np.full((nd, n), np.nan, np.float32), "tau": np.full((nd, n), np.nan, np.float32),
            "tpz": np.full(n, np.nan), "tplon": np.full(n, np.nan), "tplat": np.full(n, np.nan),
            "npts": np.full(n, -1, np.int32)}
    ptrs = {k: hip.put(v) for k, v in outs.items()}
    for k, v in ptrs.items():
        setattr(b, k, v)
    return b, ptrs, outs


def run(hip, fn, ctl, aero, w, nd, rays):
    b, ptrs, outs = device_batch(hip, w, nd, rays)
    rc = fn(C.byref(ctl), C.byref(b), C.byref(aero) if aero else None, 0, None)
    assert rc == 0
    got = {k: hip.get(ptrs[k], outs[k]) for k in outs}
    hip.free()
    return got


@pytest.mark.parametrize("workload,nbig,nsmall", [("clear_batch_65536", 3000, 700),
                                                  ("aerosol_limb_8192", 2048, 300)])
def test_noalloc_matches_allocating_entry(tmp_path, workload, nbig, nsmall):
    glib = L.load()
    hip = Hip()
    w = synth.make(workload, tmp_path, nray=nbig)
    ctl, aero = read_workload(glib, "", w, tmp_path / "c.ctl")
    nd = ctl.nd
    big = np.arange(nbig)
    small = np.arange(nbig - 1, 0, -(nbig // nsmall))[:nsmall]  # other rays, other order
    print("%s: %d then %d rays, %d channels" % (workload, len(big), len(small), nd), flush=True)
    # the allocating entry point first (it also makes the tables resident)
    ref_big = run(hip, glib.jur_formod_batch_device, ctl, aero, w, nd, big)
    ref_small = run(hip, glib.jur_formod_batch_device, ctl, aero, w, nd, small)
    got_big = run(hip, glib.jur_formod_batch_device_noalloc, ctl, aero, w, nd, big)
    got_small = run(hip, glib.jur_formod_batch_device_noalloc, ctl, aero, w, nd, small)
    for ref, got in ((ref_big, got_big), (ref_small, got_small)):
        assert np.all(np.isfinite(got["rad"])) and np.all(got["npts"] > 0)
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    assert not np.array_equal(got_small["rad"], got_big["rad"][:, : len(small)])
