#!/usr/bin/env python3
"""Benchmark: ray x channel radiances per second of the batched forward model.

Workload (default): 65,536 rays (60 % limb, 40 % nadir), each with its own
perturbed atmosphere on 121 levels, 5 emitters, 64 channels 700-1400 cm^-1
("clear_batch_65536", see jurassic-scatter_amd/pyjurassic/synth.py).  One
step = one forward call over all rays with inputs resident in GPU memory:
FP64 ray tracing, sort by path length, FP32 EGA path integration.

Prints ONE JSON line.  `roofline` is for the dominant kernel (path
integration), timed with HIP events recorded by the library on its own
stream; `cpu_baseline` times the FP64 oracle on a bounded sample of the same
rays on this host's cores, and `accuracy` compares the GPU radiances of that
sample with the oracle's.
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

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "jurassic-scatter_amd"))

HBM_PEAK_GBS = 8000.0

# Rays of the CPU oracle sample per workload (about 10-30 s on 16 cores).
CPU_SAMPLE = {"clear_limb_17": 17, "clear_batch_65536": 4096, "aerosol_limb_8192": 192, "cirrus_4096": 2}

# HBM bytes per launch of the path kernel measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE, each scaled by the same counter on a plain copy
# kernel of known size (bin/copy_probe); keyed by (workload, rays, variant,
# waves per block).  The profile summary named beside each figure holds the
# commands and the raw counter values.
MEASURED_TRAFFIC = {
    ("clear_batch_65536", 65536, 2, 4): (2.1814e11, "profiles/k_path_fast_clear_batch_65536.md"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="clear_batch_65536")
    ap.add_argument("--nray", type=int, default=None, help="truncate the workload's rays")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="rays of the CPU oracle sample (default: per workload, ~10-30 s of CPU work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", type=int, default=None,
                    help="path-kernel variant (default: autotuned on the first call)")
    ap.add_argument("--wpb", type=int, default=None, help="waves per block (default: autotuned)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has a single MI355X "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyjurassic import jtypes as J, lib as L, synth

    lib = L.load()
    tmp = Path(tempfile.mkdtemp(prefix="jur_bench_"))
    w = synth.make(args.workload, tmp, nray=args.nray)
    w.write_ctl(tmp / "bench.ctl")
    ctl = J.ctl_t()
    argc, argv = J.argv(tmp / "bench.ctl")
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)  # the control parser echoes every key
    try:
        lib.read_ctl(argc, argv, C.byref(ctl))
    finally:
        os.dup2(saved, 1)
    aero, aero_p, scattering = None, None, 0
    if w.aerofile:
        # layer optics by the library's own Mie (get_opt_prop) before timing
        aero = J.aero_t()
        lib.read_aero(None, w.aerofile.encode(), C.byref(ctl), C.byref(aero))
        lib.get_opt_prop(C.byref(ctl), C.byref(aero))
        aero_p = C.byref(aero)
        scattering = ctl.sca_mult
    if args.variant is not None:
        lib.jur_set_option(b"variant", args.variant)
    if args.wpb is not None:
        lib.jur_set_option(b"wpb", args.wpb)
    nd, ng, nray = ctl.nd, ctl.ng, w.nray
    dev = torch.device("cuda:0")

    def d(a, dt=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)

    z, p, t, q = d(w.z), d(w.p), d(w.t), d(w.q)
    geom = [d(w.geom[i]) for i in range(7)]
    rad = torch.empty((nd, nray), dtype=torch.float32, device=dev)
    tau = torch.empty((nd, nray), dtype=torch.float32, device=dev)
    tpz = torch.empty(nray, dtype=torch.float64, device=dev)
    npts = torch.empty(nray, dtype=torch.int32, device=dev)
    b = L.jur_batch_t()
    b.nray, b.nlev, b.natm = nray, w.nlev, w.natm
    b.atm_of_ray = None
    b.z, b.p, b.t, b.q, b.k = z.data_ptr(), p.data_ptr(), t.data_ptr(), q.data_ptr(), None
    (b.time, b.obsz, b.obslon, b.obslat, b.vpz, b.vplon, b.vplat) = [g.data_ptr() for g in geom]
    b.rad, b.tau, b.tpz, b.tplon, b.tplat, b.npts = rad.data_ptr(), tau.data_ptr(), tpz.data_ptr(), None, None, npts.data_ptr()
    torch.cuda.synchronize()

    def step():
        lib.jur_formod_batch_device(C.byref(ctl), C.byref(b), aero_p, scattering, None)
        lib.jur_sync()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    path_ms, trace_ms, ipts = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        path_ms.append(lib.jur_stat(b"path_ms"))
        trace_ms.append(lib.jur_stat(b"trace_ms"))
        ipts.append(lib.jur_stat(b"integrated_points"))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    value = nray * nd * args.steps / dt

    npts_h = npts.cpu().numpy().astype(np.int64)
    pts = int(npts_h.sum())
    pencils = int(lib.jur_stat(b"pencils"))
    # Algorithmic bytes of the path-integration kernel per call: per
    # (LOS point x channel x emitter) 128 B of table nodes (4 (p,T) corners x
    # 2 lookups x 2 bracketing nodes x (u, eps) x 4 B), per (point x channel)
    # 8 B of Planck-table nodes, per (pencil x channel) 8 B of rad + tau out;
    # summed over every pencil integrated (primaries, secondary fans, solar
    # pencils), with the LOS points the library counted.
    ipts_call = int(ipts[-1])
    alg_bytes = ipts_call * nd * (128 * ng + 8) + pencils * nd * 8
    path_avg = float(np.mean(path_ms))
    achieved = alg_bytes / (path_avg * 1e-3) / 1e9
    variant, wpb = lib.jur_stat(b"tuned_variant"), lib.jur_stat(b"tuned_wpb")
    traffic, traffic_src = MEASURED_TRAFFIC.get((args.workload, nray, int(variant), int(wpb)), (None, None))

    line = {
        "metric": "ray x channel radiances/s (FP32 EGA + Kahan, FP64 geometry)",
        "value": value,
        "unit": "ray*channel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (EGA, path integration), f64 (ray geometry)",
        "data": "synthetic (seeded generator, reference file formats)",
        "config": {"workload": args.workload, "rays": nray, "channels": nd, "emitters": ng,
                   "levels": w.nlev, "los_points": pts, "scattering_depth": scattering,
                   "pencils": pencils, "integrated_points": ipts_call, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_path_fast", "kernel_ms": path_avg,
                     "trace_ms": float(np.mean(trace_ms)) if not scattering else None,
                     "launches_per_call": int(lib.jur_stat(b"path_launches")),
                     "alg_bytes_per_call": alg_bytes, "variant": variant, "wpb": wpb},
    }

    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle_ctypes
        olib = oracle_ctypes.load()
        octl = J.ctl_t()
        os.dup2(devnull, 1)
        try:
            olib.oracle_read_ctl(argc, argv, C.byref(octl))
        finally:
            os.dup2(saved, 1)
        oaero = None
        if w.aerofile:
            oaero = J.aero_t()
            olib.oracle_read_aero(None, w.aerofile.encode(), C.byref(octl), C.byref(oaero))
            olib.oracle_get_opt_prop(C.byref(octl), C.byref(oaero))
        ns = min(args.cpu_sample or CPU_SAMPLE[args.workload], nray)
        # every (nray/ns)-th ray: the sample spans limb and nadir geometries
        sel = np.linspace(0, nray - 1, ns).astype(np.int64)
        arr = lambda a: np.ascontiguousarray(a)
        zs, ps, ts, qs = arr(w.z[sel]), arr(w.p[sel]), arr(w.t[sel]), arr(w.q[sel])
        gs = [arr(w.geom[i][sel]) for i in range(7)]
        orad = np.zeros((ns, nd))
        otau = np.zeros((ns, nd))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        ncores = os.cpu_count() or 1
        try:
            ncores = len(os.sched_getaffinity(0))
        except AttributeError:
            pass
        olib.oracle_formod_batch(C.byref(octl), ns, w.nlev, dp(zs), dp(ps), dp(ts), dp(qs), None,
                                 dp(gs[0]), dp(gs[1]), dp(gs[2]), dp(gs[3]), dp(gs[4]), dp(gs[5]), dp(gs[6]),
                                 0, 1, ncores, dp(orad), dp(otau), None, None, None, None,
                                 C.byref(oaero) if oaero else None, scattering)
        c0 = time.perf_counter()
        olib.oracle_formod_batch(C.byref(octl), ns, w.nlev, dp(zs), dp(ps), dp(ts), dp(qs), None,
                                 dp(gs[0]), dp(gs[1]), dp(gs[2]), dp(gs[3]), dp(gs[4]), dp(gs[5]), dp(gs[6]),
                                 0, ns, ncores, dp(orad), dp(otau), None, None, None, None,
                                 C.byref(oaero) if oaero else None, scattering)
        cdt = time.perf_counter() - c0
        line["cpu_baseline"] = {"value": ns * nd / cdt, "unit": "ray*channel/s", "cores": ncores,
                                "kind": "port",
                                "sample": "%d of the %d rays (evenly spaced), all %d channels, FP64 oracle, "
                                          "%.1f s wall" % (ns, nray, nd, cdt)}
        grad = rad.cpu().numpy().T[sel].astype(np.float64)
        gtau = tau.cpu().numpy().T[sel].astype(np.float64)
        rel = np.abs(grad / orad - 1.0)
        nu = np.array([ctl.nu[i] for i in range(nd)])
        c1, c2 = 1.19104259e-8, 1.43877506
        bt = lambda r: c2 * nu[None, :] / np.log1p(c1 * nu[None, :] ** 3 / r)
        dbt = np.abs(bt(grad) - bt(orad))
        line["accuracy"] = {"max_rel_rad": float(np.max(rel)), "mean_rel_rad": float(np.mean(rel)),
                            "max_bt_K": float(np.max(dbt)), "max_abs_tau": float(np.max(np.abs(gtau - otau))),
                            "sample_rays": ns}
    os.close(devnull)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
