"""Train the scattering-source surrogate on pairs from the plain method.

The plain GPU forward model runs over a workload's rays in chunks with
option "collect", which keeps (24 features, thermal source J) of every
(primary scattering point, channel); the network is then trained on
ln(J / B(T)) (B(T): feature 12) with Adam on the GPU, one captured hipGraph
per step, and saved with its normalisation and training hull.

    python -m pyjurassic.surrogate --workload cirrus_train_8192 --out net.bin

writes net.bin and net.json (pairs, loss history summary, timings).
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

F = C.POINTER(C.c_float)

# matrix-core FLOPs of one training step per sample: forward GEMMs
# (32-wide padded input), weight-gradient GEMMs, activation-gradient GEMMs
# of layers 2..4, and the 256-wide output layer forward and backward
TRAIN_FLOP_PER_SAMPLE = 2 * (2 * (32 * 256 + 3 * 256 * 256) + 3 * 256 * 256 + 3 * 256)
# algorithmic FLOPs of one inference (24 inputs)
INFER_FLOP_PER_SAMPLE = 2 * (24 * 256 + 3 * 256 * 256 + 256)


def _batch(w):
    from pyjurassic import lib as L
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    arrs = dict(z=f(w.z), p=f(w.p), t=f(w.t), q=f(w.q), g=[f(w.geom[i]) for i in range(7)])
    b = L.jur_batch_t()
    b.nray, b.nlev, b.natm = w.nray, w.nlev, w.natm
    b.z, b.p, b.t, b.q, b.k = (arrs["z"].ctypes.data, arrs["p"].ctypes.data, arrs["t"].ctypes.data,
                               arrs["q"].ctypes.data, None)
    (b.time, b.obsz, b.obslon, b.obslat, b.vpz, b.vplon, b.vplat) = [x.ctypes.data for x in arrs["g"]]
    return b, arrs


def read_ctl_aero(lib, w, ctlpath):
    from pyjurassic import jtypes as J
    w.write_ctl(ctlpath)
    ctl = J.ctl_t()
    argc, argv = J.argv(ctlpath)
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)  # the control parser echoes every key
    try:
        lib.read_ctl(argc, argv, C.byref(ctl))
        aero = J.aero_t()
        lib.read_aero(None, w.aerofile.encode(), C.byref(ctl), C.byref(aero))
        lib.get_opt_prop(C.byref(ctl), C.byref(aero))
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
    return ctl, aero


def collect(lib, w, ctlpath, chunk=256, log=print):
    """(features [n][24], J [n]) of the plain method over all rays of w."""
    ctl, aero = read_ctl_aero(lib, w, ctlpath)
    nd = ctl.nd
    feats, js = [], []
    lib.jur_set_option(b"surrogate", 0)
    lib.jur_set_option(b"collect", 1)
    try:
        for r0 in range(0, w.nray, chunk):
            sub = w.select(rays=np.arange(r0, min(r0 + chunk, w.nray)))
            b, keep = _batch(sub)
            rad = np.zeros((nd, sub.nray), dtype=np.float32)
            tau = np.zeros((nd, sub.nray), dtype=np.float32)
            b.rad, b.tau = rad.ctypes.data, tau.ctypes.data
            b.tpz = b.tplon = b.tplat = b.npts = None
            t0 = time.perf_counter()
            lib.jur_formod_batch(C.byref(ctl), C.byref(b), C.byref(aero), ctl.sca_mult)
            n = lib.jur_sca_collect(None, None, 0)
            f = np.zeros((n, 24), dtype=np.float32)
            j = np.zeros(n, dtype=np.float32)
            lib.jur_sca_collect(f.ctypes.data_as(F), j.ctypes.data_as(F), n)
            feats.append(f)
            js.append(j)
            log("rays %d-%d: %d pairs, %d pencils, %.2f s" % (r0, r0 + sub.nray, n, lib.jur_stat(b"pencils"),
                                                            time.perf_counter() - t0))
    finally:
        lib.jur_set_option(b"collect", 0)
    return np.concatenate(feats), np.concatenate(js)


def fit(lib, feat, jth, steps=20000, batch=8192, lr=1e-3, seed=1, log=print):
    """Train on ln(J / B(T)) with a step schedule (75 % at lr, 20 % at lr/10,
    5 % at lr/100); returns a summary dict."""
    y = (np.log(jth) - feat[:, 12]).astype(np.float32)
    ok = np.isfinite(y) & np.all(np.isfinite(feat), axis=1)
    feat, y = np.ascontiguousarray(feat[ok]), np.ascontiguousarray(y[ok])
    lib.jur_mlp_init(seed)
    lib.jur_mlp_set_data(feat.ctypes.data_as(F), y.ctypes.data_as(F), len(y), 1, seed)
    hist, ms = [], 0.0
    for frac, f_lr in [(0.75, lr), (0.20, lr / 10), (0.05, lr / 100)]:
        n = max(1, int(round(steps * frac)))
        h = np.zeros(n, dtype=np.float32)
        lib.jur_mlp_train(n, batch, C.c_float(f_lr), h.ctypes.data_as(F))
        ms += lib.jur_stat(b"train_ms")
        hist.append(h)
        log("%d steps at lr %.1e: loss %.5f -> %.5f" % (n, f_lr, h[:50].mean(), h[-50:].mean()))
    hist = np.concatenate(hist)
    per_step = ms / len(hist)
    return {"pairs": int(len(y)), "steps": int(len(hist)), "batch": batch, "lr": lr,
            "loss_first": float(hist[:100].mean()), "loss_last": float(hist[-100:].mean()),
            "ms_per_step": per_step,
            "train_tflops": TRAIN_FLOP_PER_SAMPLE * batch / (per_step * 1e-3) / 1e12}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="cirrus_train_8192")
    ap.add_argument("--nray", type=int, default=None)
    ap.add_argument("--chunk", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    from pyjurassic import lib as L, synth
    lib = L.load()
    tmp = Path(tempfile.mkdtemp(prefix="jur_sur_"))
    w = synth.make(args.workload, tmp, nray=args.nray)
    log = lambda m: print(m, flush=True)
    t0 = time.perf_counter()
    feat, jth = collect(lib, w, tmp / "c.ctl", args.chunk, log)
    t_collect = time.perf_counter() - t0
    log("collected %d pairs from %d rays in %.1f s" % (len(jth), w.nray, t_collect))
    summary = fit(lib, feat, jth, args.steps, args.batch, args.lr, log=log)
    summary.update({"workload": args.workload, "rays": w.nray, "collect_s": t_collect})
    lib.jur_mlp_save(args.out.encode())
    Path(args.out).with_suffix(".json").write_text(json.dumps(summary, indent=1) + "\n")
    log(json.dumps(summary))


if __name__ == "__main__":
    main()
