#!/usr/bin/env python3
"""Benchmark: ray x channel radiances per second of the batched forward model.

Workload (default): 65,536 rays (60 % limb, 40 % nadir), each with its own
perturbed atmosphere on 121 levels, 5 emitters, 64 channels 700-1400 cm^-1
("clear_batch_65536", see jurassic-scatter_amd/pyjurassic/synth.py).  One
step = one forward call over all rays with inputs resident in GPU memory:
FP64 ray tracing, sort by path length, FP32 EGA path integration.  `value`
comes from hipGraph replays of that call (--graph 1, the default; the
scattering workloads' calls are recorded from their device-resident form,
which runs from the sizes of the first call); the same number of calls
without a graph is timed before them (`eager_ms_per_step`), and the
path kernel's time (`roofline.kernel_ms`) comes from the HIP events the
library records around its launches in those eager calls (events recorded
inside a captured graph cannot be timed on this runtime).

Prints ONE JSON line.  `roofline` is for the dominant kernel (path
integration), timed with HIP events recorded by the library on its own
stream; `cpu_baseline` times the FP64 oracle on a bounded sample of the same
rays on this host's cores, and `accuracy` compares the GPU radiances of that
sample with the oracle's.

Other workloads (--workload): the synthetic generator's other cases
(clear_limb_17, aerosol_limb_8192, cirrus_4096: scattering), cirrus_16384
(the learned scattering source against the plain method), jacobian_limb_17
(kernel(): the finite-difference Jacobian), retrieval_limb_17
(optimal_estimation) and aerosol3d_steep_16 (srcfunc_sca_3d's direct fans
over a 2-D atmosphere); each prints its own JSON line of the same form.
Library options for experiments: --option KEY=VALUE (jur_set_option).
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

# The configurations' own scattering parameters (--ref-params: the
# reference's 28 pencils per fan, 170 Gauss-Hermite radii, 181 angles): the
# aerosol configuration's 16 pencils per fan and Mie grid (64 log-spaced
# radii 0.01-20 um, 180 phase angles), the cirrus configuration's 32 pencils
# (a grid of this project's: the reference's construction leaves 0-180 deg
# at 32).
CONFIG_PARAMS = {"aerosol_limb_8192": {"fan": 16, "mie_grid": 1}, "cirrus_4096": {"fan": 32, "mie_grid": 0}}
REF_PARAMS = {"fan": 28, "mie_grid": 0}

# Rays of the CPU oracle sample per workload (about 10-30 s on 16 cores; the
# clear-air batch: all of its rays, ~31 s, so that `accuracy` covers every ray) ...
CPU_SAMPLE = {"clear_limb_17": 17, "clear_batch_65536": 65536, "aerosol_limb_8192": 320, "cirrus_4096": 1}
# ... and its channels where all of them would take longer (depth 2: a
# nadir ray through the cloud, 8 of 32 channels, ~20 s)
CPU_SAMPLE_CHANNELS = {"cirrus_4096": 8}
CPU_SAMPLE_NADIR = ("cirrus_4096",)

# HBM bytes per launch of the path kernel measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE, each scaled by the same counter on a plain copy
# kernel of known size (bin/copy_probe); keyed by (workload, rays, variant,
# waves per block).  The profile summary named beside each figure holds the
# commands and the raw counter values.
MEASURED_TRAFFIC = {
    ("clear_batch_65536", 65536, 6, 2): (1.9978e11, "profiles/k_path_fast_clear_batch_65536.md"),
    ("clear_batch_65536", 65536, 6, 4): (1.9911e+11, "profiles/k_path_fast_clear_batch_65536.md"),
    ("clear_batch_65536", 65536, 7, 1): (1.9045e11, "profiles/k_path_stage_clear_batch_65536.md"),
    ("clear_batch_65536", 65536, 7, 2): (1.8952e11, "profiles/k_path_stage_clear_batch_65536.md"),
    ("clear_batch_65536", 65536, 7, 4): (1.8896e11, "profiles/k_path_stage_clear_batch_65536.md"),
    ("clear_batch_65536", 65536, 9, 1): (1.1773e10, "profiles/k_path_stage_clear_batch_65536.md"),
    ("clear_batch_65536", 65536, 9, 2): (1.5405e10, "profiles/k_path_stage_clear_batch_65536.md"),
    ("clear_batch_65536", 65536, 9, 4): (2.4268e10, "profiles/k_path_stage_clear_batch_65536.md"),
    # variant 10: k_pack and k_path_stage together (the launch the HIP events time)
    ("clear_batch_65536", 65536, 10, 1): (3.8417e10, "profiles/k_path_stage_clear_batch_65536.md"),
}
# What binds the path kernel, from rocprofv3 counters of the same
# configuration (the profile named beside each entry holds the passes):
# VALU issue (SQ_INSTS_VALU x 4 cycles / (SIMDs x kernel cycles)), the
# vector-memory data-return path (TD_TD_BUSY_sum / (CUs x cycles)), the
# address path (TA) and the L2 hit rate.
MEASURED_BINDING = {
    ("clear_batch_65536", 65536, 10, 1): ({"valu_issue": 1.00, "td_busy": 0.96, "ta_busy": 0.77, "l2_hit": 0.78},
                                          "profiles/k_path_stage_clear_batch_65536.md"),
    ("clear_batch_65536", 65536, 9, 1): ({"valu_issue": 0.88, "td_busy": 0.92, "ta_busy": 0.66, "l2_hit": 0.96},
                                         "profiles/k_path_stage_clear_batch_65536.md"),
}
# The FP32 path kernel of each autotuner variant (7-9: the staged kernel,
# grids in LDS, lockstep bracket searches, 9 with the tiled block mapping;
# 10: the same on the packed stream of k_pack (lane-contiguous records, (p, T)
# brackets precomputed per point and emitter), whose time is included; 0-6: the fast kernel).
PATH_KERNEL = {7: "k_path_stage", 8: "k_path_stage", 9: "k_path_stage", 10: "k_pack + k_path_stage"}
# The same for the FP64 path kernel of kernel() on the Jacobian workload.
# HBM bytes per kernel() call of the path kernels (FETCH_SIZE / WRITE_SIZE
# passes): with base-ray reuse (the default), and integrating every copy
# from the observer (option kernel_prefix 0)
MEASURED_TRAFFIC_JACOBIAN = {1: (1.900e+10, "profiles/k_path64_jacobian_limb_17.md"),
                             0: (7.323e+10, "profiles/k_path64_jacobian_limb_17.md")}


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
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="library option (jur_set_option) before the run; repeatable")
    ap.add_argument("--variant", type=int, default=None,
                    help="path-kernel variant (default: autotuned on the first call)")
    ap.add_argument("--wpb", type=int, default=None, help="waves per block (default: autotuned)")
    ap.add_argument("--graph", type=int, default=1,
                    help="1 (default): `value` from hipGraph replays of the clear-air call (eager calls are "
                         "timed beside them); 0: eager calls only")
    ap.add_argument("--weights", default=None,
                    help="trained surrogate for --workload cirrus_16384 (default: the packaged one)")
    ap.add_argument("--fan", type=int, default=None,
                    help="secondary pencils per scattering point (default: the configuration's, else 28)")
    ap.add_argument("--ref-params", action="store_true",
                    help="scattering workloads at the reference's fan (28) and Mie grid instead of the "
                         "configuration's")
    ap.add_argument("--chunk", type=int, default=None,
                    help="rays per call: one step = consecutive calls over blocks of this many rays, filling "
                         "one output array (progress printed per call)")
    ap.add_argument("--plain-rays", type=int, default=1024,
                    help="rays of the plain-method comparison for --workload cirrus_16384")
    return ap.parse_args()


_JSON_OUT = None


def emit(line):
    """The one JSON line, on the process's stdout (everything else the
    library or this script prints goes to stderr, see main)."""
    out = _JSON_OUT or sys.stdout
    out.write(json.dumps(line) + "\n")
    out.flush()


def quiet(fn):
    """Run fn with stdout silenced (the control parser echoes every key)."""
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        return fn()
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)


def setup(lib, workload, nray):
    """Workload files, ctl and (for aerosol workloads) aero with the layer
    optics from the library's own Mie (get_opt_prop), before any timing."""
    from pyjurassic import jtypes as J, synth
    tmp = Path(tempfile.mkdtemp(prefix="jur_bench_"))
    w = synth.make(workload, tmp, nray=nray)
    w.write_ctl(tmp / "bench.ctl")
    ctl = J.ctl_t()
    argc, argv = J.argv(tmp / "bench.ctl")
    quiet(lambda: lib.read_ctl(argc, argv, C.byref(ctl)))
    aero, scattering = None, 0
    if w.aerofile:
        aero = J.aero_t()
        lib.read_aero(None, w.aerofile.encode(), C.byref(ctl), C.byref(aero))
        lib.get_opt_prop(C.byref(ctl), C.byref(aero))
        scattering = ctl.sca_mult
    return w, ctl, aero, scattering, (argc, argv)


def device_batch(w, nd, rays=None):
    """jur_batch_t over device copies (torch) of w's rays (a subset with
    `rays`); returns (batch, outputs, tensors kept alive)."""
    import torch
    from pyjurassic import lib as L
    if rays is not None:
        w = w.select(rays=rays)
    dev = torch.device("cuda:0")
    d = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    z, p, t, q = d(w.z), d(w.p), d(w.t), d(w.q)
    geom = [d(w.geom[i]) for i in range(7)]
    rad = torch.empty((nd, w.nray), dtype=torch.float32, device=dev)
    tau = torch.empty((nd, w.nray), dtype=torch.float32, device=dev)
    tpz = torch.empty(w.nray, dtype=torch.float64, device=dev)
    npts = torch.empty(w.nray, dtype=torch.int32, device=dev)
    b = L.jur_batch_t()
    b.nray, b.nlev, b.natm = w.nray, w.nlev, w.natm
    b.atm_of_ray = None
    b.z, b.p, b.t, b.q, b.k = z.data_ptr(), p.data_ptr(), t.data_ptr(), q.data_ptr(), None
    (b.time, b.obsz, b.obslon, b.obslat, b.vpz, b.vplon, b.vplat) = [g.data_ptr() for g in geom]
    b.rad, b.tau, b.tpz, b.tplon, b.tplat, b.npts = (rad.data_ptr(), tau.data_ptr(), tpz.data_ptr(), None, None,
                                                     npts.data_ptr())
    torch.cuda.synchronize()
    return b, (rad, tau, tpz, npts), (z, p, t, q, geom)


def ncores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_oracle(w, argc, argv, sel, scattering, fan=28, nd=None, mie_grid=0):
    """The FP64 oracle on rays `sel` of w (the first nd channels, default
    all) with all host cores: (rad, tau [ns][nd], seconds of the timed call,
    cores)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_ctypes
    from pyjurassic import jtypes as J
    olib = oracle_ctypes.load()
    olib.oracle_set_fan(fan)
    olib.oracle_set_mie_grid(mie_grid)
    octl = J.ctl_t()
    quiet(lambda: olib.oracle_read_ctl(argc, argv, C.byref(octl)))
    if nd is not None:
        octl.nd = nd
    oaero = None
    if w.aerofile:
        oaero = J.aero_t()
        olib.oracle_read_aero(None, w.aerofile.encode(), C.byref(octl), C.byref(oaero))
        olib.oracle_get_opt_prop(C.byref(octl), C.byref(oaero))
    ns, nd = len(sel), octl.nd
    arr = lambda a: np.ascontiguousarray(a)
    zs, ps, ts, qs = arr(w.z[sel]), arr(w.p[sel]), arr(w.t[sel]), arr(w.q[sel])
    gs = [arr(w.geom[i][sel]) for i in range(7)]
    orad = np.zeros((ns, nd))
    otau = np.zeros((ns, nd))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    nc = ncores()
    call = lambda n: olib.oracle_formod_batch(
        C.byref(octl), ns, w.nlev, dp(zs), dp(ps), dp(ts), dp(qs), None, *[dp(g) for g in gs], 0, n, nc,
        dp(orad), dp(otau), None, None, None, None, C.byref(oaero) if oaero else None, scattering)
    call(1)  # tables and optics loaded outside the timed call
    c0 = time.perf_counter()
    call(ns)
    return orad, otau, time.perf_counter() - c0, nc


def accuracy(grad, gtau, orad, otau, ctl):
    """Relative radiance, brightness temperature and transmittance errors."""
    grad, gtau = grad.astype(np.float64), gtau.astype(np.float64)
    nd = orad.shape[1]
    rel = np.abs(grad / orad - 1.0)
    nu = np.array([ctl.nu[i] for i in range(nd)])
    c1, c2 = 1.19104259e-8, 1.43877506
    bt = lambda r: c2 * nu[None, :] / np.log1p(c1 * nu[None, :] ** 3 / r)
    dbt = np.abs(bt(grad) - bt(orad))
    return {"max_rel_rad": float(np.max(rel)), "mean_rel_rad": float(np.mean(rel)),
            "max_bt_K": float(np.max(dbt)), "max_abs_tau": float(np.max(np.abs(gtau - otau)))}


SURROGATE_WEIGHTS = ROOT / "jurassic-scatter_amd" / "data" / "surrogate_cirrus.bin"
MFMA_BF16_PEAK_TFLOPS = 2500.0


def bench_surrogate(args, lib):
    """Learned scattering source on the held-out cirrus set (16,384 rays x
    32 channels, depth 2, other atmospheres than the training set): the
    network (trained on the plain method over cirrus_train_8192 by
    pyjurassic.surrogate) gives J at the primaries' scattering points, the
    fused matrix-core kernel evaluates it; the plain method on the first
    --plain-rays rays is the reference it is compared with (radiance difference,
    speedup of the call)."""
    import torch
    from pyjurassic import surrogate as S
    weights = Path(args.weights) if args.weights else SURROGATE_WEIGHTS
    if not weights.exists():
        sys.exit("no trained surrogate at %s: python -m pyjurassic.surrogate --workload cirrus_train_8192 --out %s"
                 % (weights, weights))
    lib.jur_mlp_load(str(weights).encode())
    info = weights.with_suffix(".json")
    train = json.loads(info.read_text()) if info.exists() else None
    w, ctl, aero, scattering, _ = setup(lib, "cirrus_16384", args.nray)
    nd, nray = ctl.nd, w.nray
    b, (rad, tau, _, _), keep = device_batch(w, nd)

    def call(bb, sur):
        lib.jur_set_option(b"surrogate", sur)
        lib.jur_formod_batch_device(C.byref(ctl), C.byref(bb), C.byref(aero), scattering, None)
        lib.jur_sync()

    for _ in range(args.warmup):
        call(b, 1)
    torch.cuda.synchronize()
    mlp_ms, samples = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call(b, 1)
        mlp_ms.append(lib.jur_stat(b"mlp_ms"))
        samples.append(lib.jur_stat(b"mlp_samples"))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nsur, nfb = lib.jur_stat(b"surrogate_points"), lib.jur_stat(b"fallback_points")
    pencils = lib.jur_stat(b"pencils")
    # plain method against the surrogate on a subset (second call of each timed)
    sub = np.arange(min(args.plain_rays, nray))
    bs, (rs, ts, _, _), keep2 = device_batch(w, nd, rays=sub)
    res = {}
    for sur in (0, 1):
        call(bs, sur)
        t1 = time.perf_counter()
        call(bs, sur)
        res[sur] = (time.perf_counter() - t1, rs.cpu().numpy().copy(), lib.jur_stat(b"pencils"))
    rel = np.abs(res[1][1] / res[0][1] - 1.0)
    mlp_avg = float(np.mean(mlp_ms))
    achieved = float(np.mean(samples)) * S.INFER_FLOP_PER_SAMPLE / (mlp_avg * 1e-3) / 1e12
    line = {
        "metric": "ray x channel radiances/s, learned scattering source (depth 2)",
        "value": nray * nd * args.steps / dt,
        "unit": "ray*channel/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (network, fp32 accumulation), f32 (path integration), f64 (geometry)",
        "data": "synthetic (seeded generator; held-out atmospheres and rays)",
        "config": {"workload": "cirrus_16384", "rays": nray, "channels": nd, "scattering_depth": scattering,
                   "pencils": pencils, "surrogate_points": nsur, "fallback_points": nfb,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS, "traffic": None, "kernel": "k_mlp_infer",
                     "kernel_ms": mlp_avg, "samples": float(np.mean(samples)),
                     "flop_per_sample": S.INFER_FLOP_PER_SAMPLE},
        "surrogate": {"plain_rays": len(sub), "plain_s": res[0][0], "surrogate_s": res[1][0],
                      "speedup": res[0][0] / res[1][0], "plain_pencils": res[0][2],
                      "surrogate_pencils": res[1][2], "max_rel_rad": float(rel.max()),
                      "mean_rel_rad": float(rel.mean()), "training": train},
    }
    emit(line)


JACOBIAN_TH = np.arange(6.0, 30.01, 1.5)  # 17 limb tangent heights [km]


def jacobian_inputs(lib, prefix, w, ctlpath, keys):
    """ctl / atm / obs of the Jacobian workload read or filled for `lib`."""
    from pyjurassic import jtypes as J, synth
    ctl = J.ctl_t()
    argc, argv = J.argv(ctlpath, *keys)
    quiet(lambda: getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl)))
    atm, obs = J.atm_t(), J.obs_t()
    atm.np = w.nlev
    for l in range(w.nlev):
        atm.z[l], atm.p[l], atm.t[l] = w.z[0, l], w.p[0, l], w.t[0, l]
        for g in range(ctl.ng):
            atm.q[g][l] = w.q[0, g, l]
    g = synth._rays_limb(JACOBIAN_TH)
    obs.nr = g.shape[1]
    for ir in range(obs.nr):
        (obs.time[ir], obs.obsz[ir], obs.obslon[ir], obs.obslat[ir], obs.vpz[ir], obs.vplon[ir],
         obs.vplat[ir]) = g[:, ir]
    return ctl, atm, obs


def bench_jacobian(args, lib):
    """kernel(), the finite-difference Jacobian (reference
    src/retrievalmodel.c:201-298): a 17-ray limb scan (tangent heights 6-30 km,
    observer 800 km) through one atmosphere of the 65,536-ray workload (121
    levels, 5 emitters, 64 channels); state vector temperature and the five
    volume mixing ratios on all 121 levels (726 elements: 727 forward models
    of 17 rays, integrated in FP64).  The CPU leg runs the oracle's kernel()
    on the same inputs; `accuracy` compares the two matrices."""
    import torch
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyjurassic import synth
    tmp = Path(tempfile.mkdtemp(prefix="jur_jac_"))
    w = synth.make("clear_batch_65536", tmp, nray=1)
    ctlpath = tmp / "jac.ctl"
    w.write_ctl(ctlpath)
    ng = len(w.emitters)
    keys = ["RETT_ZMIN", 0, "RETT_ZMAX", 120]
    for g in range(ng):
        keys += ["RETQ_ZMIN[%d]" % g, 0, "RETQ_ZMAX[%d]" % g, 120]
    ctl, atm, obs = jacobian_inputs(lib, "", w, ctlpath, keys)
    r = C.byref
    n = lib.jur_state_size(r(ctl), r(atm), None)
    m = lib.jur_meas_size(r(ctl), r(obs))
    k = np.zeros((m, n))
    kp = k.ctypes.data_as(C.POINTER(C.c_double))

    def step():
        lib.jur_kernel(r(ctl), r(atm), r(obs), None, kp, m, n, n)

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    steps = args.steps
    path_ms, ipts = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        path_ms.append(lib.jur_stat(b"path_ms"))
        ipts.append(lib.jur_stat(b"integrated_points"))
    dt = time.perf_counter() - t0
    nd, nr = ctl.nd, obs.nr
    rays = (n + 1) * nr
    # the last batched call holds every perturbed copy (one group); with
    # base-ray reuse, the (point, emitter) evaluations actually made
    ept = lib.jur_stat(b"emitter_points")
    reuse = bool(np.isfinite(ept))
    evals = ept if reuse else ipts[-1] * ng
    pencils = lib.jur_stat(b"pencils")
    alg_bytes = evals * nd * 128 + ipts[-1] * nd * 8 + pencils * nd * 8
    path_avg = float(np.mean(path_ms))
    achieved = alg_bytes / (path_avg * 1e-3) / 1e9
    # the bytes every copy integrated from the observer would move, over the
    # same time: what the reuse buys
    full_pts = lib.jur_stat(b"reuse_all_points") if reuse else ipts[-1]
    full_bytes = full_pts * nd * (128 * ng + 8) + pencils * nd * 8
    traffic = MEASURED_TRAFFIC_JACOBIAN[1 if reuse else 0]
    line = {
        "metric": "Jacobian: ray x channel radiances/s over the n+1 forward models of kernel()",
        "value": rays * nd * steps / dt,
        "unit": "ray*channel/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (path integration of the perturbed states), f64 (ray geometry)",
        "data": "synthetic (seeded generator, reference file formats)",
        "config": {"workload": "jacobian_limb_17", "rays": nr, "channels": nd, "emitters": ng,
                   "levels": w.nlev, "state_elements": n, "measurements": m,
                   "forward_rays_per_kernel": rays, "columns_per_s": n * steps / dt,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic[0], "traffic_source": traffic[1],
                     "kernel": ("k_path64 (reuse pass) + k_path64e (unperturbed rays)" if reuse
                                else "k_path64"),
                     "kernel_ms": path_avg, "alg_bytes_per_call": alg_bytes,
                     "full_work_bytes_per_call": full_bytes,
                     "full_work_frac": full_bytes / (path_avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "integrated_points": ipts[-1], "emitter_points": evals,
                     "reuse_wave_iterations": {k: lib.jur_stat(("reuse_wave_" + k).encode())
                                               for k in ("points", "emitters", "optics")},
                     "reuse_waves_by_emitters": [[lib.jur_stat(b"reuse_waves_e%d" % e),
                                                  lib.jur_stat(b"reuse_wave_points_e%d" % e)]
                                                 for e in range(ng + 1)]},
    }
    if not args.no_cpu:
        import oracle_ctypes
        olib = oracle_ctypes.load()
        octl, oatm, oobs = jacobian_inputs(olib, "oracle_", w, ctlpath, keys)
        # bounded sample: every 4th column (the unperturbed forward model and
        # 182 perturbed ones)
        cols = np.arange(0, n, 4, dtype=np.int32)
        ko = np.zeros((m, n))
        c0 = time.perf_counter()
        quiet(lambda: olib.oracle_kernel_columns(r(octl), r(oatm), r(oobs), None,
                                                 ko.ctypes.data_as(C.POINTER(C.c_double)), m, n, n,
                                                 cols.ctypes.data_as(C.POINTER(C.c_int)), len(cols)))
        cdt = time.perf_counter() - c0
        ks, kos = k[:, cols], ko[:, cols]
        # perturbation sizes and unperturbed radiances (oracle): |dK| h / |y|
        # is the disagreement of the perturbed forward models themselves
        x = np.zeros(n)
        iqa = np.zeros(n, dtype=np.int32)
        olib.oracle_atm2x(r(octl), r(oatm), None, x.ctypes.data_as(C.POINTER(C.c_double)),
                          iqa.ctypes.data_as(C.POINTER(C.c_int)), None)
        h = np.array([olib.oracle_perturbation(r(octl), int(iqa[j]), x[j]) for j in cols])
        y = np.array([[oobs.rad[id][ir] for id in range(nd)] for ir in range(nr)]).ravel()
        dy_err = np.abs(ks - kos) * h[None, :] / np.abs(y)[:, None]
        dy_sig = np.abs(kos) * h[None, :] / np.abs(y)[:, None]
        res = dy_sig >= 1e-8
        line["cpu_baseline"] = {"value": (len(cols) + 1) * nr * nd / cdt, "unit": "ray*channel/s",
                                "cores": ncores(), "kind": "port",
                                "sample": "%d of the %d Jacobian columns (every 4th) plus the unperturbed "
                                          "forward model, %d rays x %d channels each, FP64 oracle kernel(), "
                                          "%.1f s wall" % (len(cols), n, nr, nd, cdt)}
        line["accuracy"] = {
            "max_rel_perturbed_radiance": float(dy_err.max()),
            "max_rel_K_resolved": float(np.max(np.abs(ks[res] / kos[res] - 1))),
            "resolved_entries": int(res.sum()), "entries_compared": int(res.size),
            "resolved": "entries whose radiance change exceeds 1e-8 of the radiance; below that the "
                        "finite difference approaches FP64 rounding in the reference too",
            "columns_compared": len(cols)}
    emit(line)


def retrieval_truth(lib, prefix, ctl, atm, obs):
    """Measurements of a perturbed truth (temperature +2 K around 20 km,
    H2O x1.2 below 15 km, O3 x0.9) by `lib`'s forward model in FP64."""
    from pyjurassic import jtypes as J
    tru, meas = J.atm_t(), J.obs_t()
    C.memmove(C.addressof(tru), C.addressof(atm), C.sizeof(atm))
    C.memmove(C.addressof(meas), C.addressof(obs), C.sizeof(obs))
    for l in range(atm.np):
        z = atm.z[l]
        tru.t[l] += 2.0 * np.exp(-((z - 20) / 6) ** 2)
        if z < 15:
            tru.q[1][l] *= 1.2
        tru.q[2][l] *= 0.9
    if prefix:
        quiet(lambda: lib.oracle_formod(C.byref(ctl), C.byref(tru), C.byref(meas), None))
    else:
        lib.jur_set_option(b"fp64", 1)
        quiet(lambda: lib.formod(C.byref(ctl), C.byref(tru), C.byref(meas), None))
        lib.jur_set_option(b"fp64", 0)
    return meas


def bench_retrieval(args, lib):
    """optimal_estimation (reference src/retrievalmodel.c:407-706): the
    17-ray, 64-channel limb scan of the Jacobian workload; state vector
    temperature, H2O and O3 on all 121 levels (363 elements, 1088
    measurements), a priori errors 5 K / 50 % / 30 % with 5 km vertical
    correlation, 1 % measurement errors; measurements of a perturbed truth.
    One step is one whole retrieval (Levenberg-Marquardt iterations with a
    Jacobian each, error analysis, output files).  The CPU leg runs the
    oracle's restatement on the same inputs; `accuracy` compares the
    retrieved states."""
    import torch
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyjurassic import jtypes as J, synth
    tmp = Path(tempfile.mkdtemp(prefix="jur_ret_"))
    w = synth.make("clear_batch_65536", tmp, nray=1)
    ctlpath = tmp / "ret.ctl"
    w.write_ctl(ctlpath)
    keys = ["RETT_ZMIN", 0, "RETT_ZMAX", 120, "RETQ_ZMIN[1]", 0, "RETQ_ZMAX[1]", 120,
            "RETQ_ZMIN[2]", 0, "RETQ_ZMAX[2]", 120, "ERR_TEMP", 5, "ERR_TEMP_CZ", 5, "ERR_TEMP_CH", 500,
            "ERR_Q[1]", 50, "ERR_Q_CZ[1]", 5, "ERR_Q_CH[1]", 500, "ERR_Q[2]", 30, "ERR_Q_CZ[2]", 5,
            "ERR_Q_CH[2]", 500]
    for d in range(len(w.nu)):
        keys += ["ERR_FORMOD[%d]" % d, 1]

    def inputs(L, prefix):
        ctl, atm, obs = jacobian_inputs(L, prefix, w, ctlpath, keys)
        ret = J.ret_t()
        argc, argv = J.argv(ctlpath, *keys)
        quiet(lambda: getattr(L, prefix + "read_ret")(argc, argv, C.byref(ctl), C.byref(ret)))
        return ctl, atm, obs, ret

    ctl, atm, obs, ret = inputs(lib, "")
    ret.dir = str(tmp).encode()
    meas = retrieval_truth(lib, "", ctl, atm, obs)
    r = C.byref
    n = lib.jur_state_size(r(ctl), r(atm), None)
    m = lib.jur_meas_size(r(ctl), r(meas))
    aero_apr, aero_i = J.aero_t(), J.aero_t()
    obs_i, atm_i = J.obs_t(), J.atm_t()

    def step():
        quiet(lambda: lib.optimal_estimation(r(ret), r(ctl), r(meas), r(obs_i), r(atm), r(atm_i), r(aero_apr),
                                             r(aero_i)))

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    steps = args.steps
    tot = lambda: np.array([lib.jur_stat(k) for k in (b"calls_total", b"path_ms_total",
                                                      b"integrated_points_total", b"pencils_total",
                                                      b"emitter_points_total")])
    s0 = tot()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    dt = time.perf_counter() - t0
    calls, path_ms, ipts, pencils, evals = (tot() - s0) / steps
    costs = [ln.split() for ln in (tmp / "costs.tab").read_text().splitlines() if ln and ln[0].isdigit()]
    nd, ng = ctl.nd, ctl.ng
    # (point, emitter) evaluations made (the Jacobians reuse the unperturbed
    # rays' emitter chains)
    alg_bytes = evals * nd * 128 + ipts * nd * 8 + pencils * nd * 8
    achieved = alg_bytes / (path_ms * 1e-3) / 1e9
    line = {
        "metric": "retrievals/s (optimal_estimation: forward models, Jacobians, linear algebra, outputs)",
        "value": steps / dt,
        "unit": "retrieval/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, reference file formats)",
        "config": {"workload": "retrieval_limb_17", "rays": obs.nr, "channels": nd, "emitters": ng,
                   "levels": w.nlev, "state_elements": n, "measurements": m,
                   "cost_evaluations": len(costs), "forward_calls": calls,
                   "forward_rays": pencils, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_path64",
                     "kernel_ms": path_ms, "kernel_share": path_ms / (1e3 * dt / steps),
                     "alg_bytes_per_retrieval": alg_bytes,
                     "note": "path kernel time and algorithmic bytes summed over the retrieval's calls"},
    }
    # wall time of the last timed retrieval's phases (library timers)
    fm, km, tm = (lib.jur_stat(k) for k in (b"oe_formod_ms", b"oe_kernel_ms", b"oe_total_ms"))
    line["phases_ms"] = {"forward_models": fm, "jacobians": km, "algebra_and_files": tm - fm - km, "total": tm}
    if not args.no_cpu:
        import oracle_ctypes
        olib = oracle_ctypes.load()
        octl, oatm, oobs, oret = inputs(olib, "oracle_")
        ometa = retrieval_truth(olib, "oracle_", octl, oatm, oobs)
        chisq = np.zeros(1 + 20 * oret.conv_itmax)
        nc = C.c_int(0)
        o_obs, o_atm = J.obs_t(), J.atm_t()
        c0 = time.perf_counter()
        quiet(lambda: olib.oracle_optimal_estimation(r(oret), r(octl), r(ometa), r(o_obs), r(oatm), r(o_atm),
                                                     None, None, chisq.ctypes.data_as(C.POINTER(C.c_double)),
                                                     r(nc), None))
        cdt = time.perf_counter() - c0
        line["cpu_baseline"] = {"value": 1 / cdt, "unit": "retrieval/s", "cores": ncores(), "kind": "port",
                                "sample": "one whole retrieval of the same inputs by the FP64 oracle (its "
                                          "error analysis skipped), %.1f s wall" % cdt}
        tg, to = np.array(atm_i.t[:atm_i.np]), np.array(o_atm.t[:o_atm.np])
        qg = np.array([atm_i.q[g][:atm_i.np] for g in (1, 2)])
        qo = np.array([o_atm.q[g][:o_atm.np] for g in (1, 2)])
        line["accuracy"] = {"max_abs_T_K": float(np.max(np.abs(tg - to))),
                            "max_rel_q": float(np.max(np.abs(qg / qo - 1))),
                            "cost_evaluations_gpu": len(costs), "cost_evaluations_oracle": nc.value,
                            "final_chi2_gpu": float(costs[-1][1]), "final_chi2_oracle": float(chisq[nc.value - 1])}
    emit(line)


def scatter3d_inputs(lib, prefix, w, ctlpath):
    """ctl, aero (with optics), the two-profile atmosphere and 16 steep rays
    of the direct-fan workload, read with `lib`'s readers."""
    from pyjurassic import jtypes as J
    ctl = J.ctl_t()
    argc, argv = J.argv(ctlpath, "IP", 2)
    quiet(lambda: getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl)))
    aero = J.aero_t()
    quiet(lambda: getattr(lib, prefix + "read_aero")(None, w.aerofile.encode(), C.byref(ctl), C.byref(aero)))
    quiet(lambda: getattr(lib, prefix + "get_opt_prop")(C.byref(ctl), C.byref(aero)))
    atm = J.atm_t()
    n = 0
    for lat, dt in ((0.0, 0.0), (5.0, 6.0)):
        for l in range(w.nlev):
            atm.z[n], atm.lon[n], atm.lat[n], atm.p[n], atm.t[n] = w.z[0, l], 0.0, lat, w.p[0, l], w.t[0, l] + dt
            for g in range(ctl.ng):
                atm.q[g][n] = w.q[0, g, l]
            n += 1
    atm.np = n
    obs = J.obs_t()
    obs.nr = 16
    for i in range(obs.nr):
        lat0 = 0.5 + 0.25 * i
        obs.obsz[i], obs.obslon[i], obs.obslat[i] = 30.0, 0.0, lat0
        obs.vpz[i], obs.vplon[i], obs.vplat[i] = 0.0, 0.0, lat0 + 0.3
    return ctl, aero, atm, obs


def bench_scatter3d(args, lib):
    """srcfunc_sca_3d (reference src/scatter.c:811-895), the scattering
    source for 2-D / 3-D atmospheres: 16 steep rays from 30 km through the
    15-20 km aerosol layer of the aerosol workload (8 of its channels, its
    five emitters), single scattering, IP 2 over two profiles 5 degrees
    apart.  Every scattering point spawns 32,400 direct pencils (a 180 x 180
    grid of directions), so the work is the pencils: value = pencil x
    channel radiances per second.  The CPU leg times the oracle's
    srcfunc_sca_3d at one scattering point of the layer (its 32,400
    pencils)."""
    import torch
    sys.path.insert(0, str(ROOT / "oracle"))
    from pyjurassic import synth
    tmp = Path(tempfile.mkdtemp(prefix="jur_s3d_"))
    w = synth.make("aerosol_limb_8192", tmp, nray=1).select(nd=8)
    ctlpath = tmp / "s3d.ctl"
    w.write_ctl(ctlpath)
    ctl, aero, atm, obs = scatter3d_inputs(lib, "", w, ctlpath)
    r = C.byref
    from pyjurassic import jtypes as J
    o = J.obs_t()

    def step():
        C.memmove(C.addressof(o), C.addressof(obs), C.sizeof(obs))
        lib.formod(r(ctl), r(atm), r(o), r(aero))

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    steps = args.steps
    tot = lambda: np.array([lib.jur_stat(k) for k in (b"path_ms_total", b"integrated_points_total",
                                                      b"pencils_total")])
    s0 = tot()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    dt = time.perf_counter() - t0
    path_ms, ipts, pencils = (tot() - s0) / steps
    nd, ng = ctl.nd, ctl.ng
    alg_bytes = ipts * nd * (128 * ng + 8) + pencils * nd * 8
    achieved = alg_bytes / (path_ms * 1e-3) / 1e9
    npts = (pencils - obs.nr) / 32400
    line = {
        "metric": "direct-fan pencil x channel radiances/s (srcfunc_sca_3d, IP 2)",
        "value": pencils * nd * steps / dt,
        "unit": "pencil*channel/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (path integration), f64 (ray geometry, source sums)",
        "data": "synthetic (seeded generator, reference file formats)",
        "config": {"workload": "aerosol3d_steep_16", "rays": obs.nr, "channels": nd, "emitters": ng,
                   "profiles": 2, "levels": w.nlev, "scattering_points": npts, "pencils": pencils,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": PATH_KERNEL.get(int(lib.jur_stat(b"tuned_variant")), "k_path_fast")
                     if np.isfinite(lib.jur_stat(b"tuned_variant")) else "k_path_fast",
                     "kernel_ms": path_ms, "kernel_share": path_ms / (1e3 * dt / steps),
                     "alg_bytes_per_call": alg_bytes,
                     "note": "path kernel time and algorithmic bytes summed over the call's launches"},
    }
    if not args.no_cpu:
        import oracle_ctypes
        olib = oracle_ctypes.load()
        octl, oaero, oatm, _ = scatter3d_inputs(olib, "oracle_", w, ctlpath)
        x = (C.c_double * 3)()
        dx = (C.c_double * 3)()
        # a point at 17 km in the layer, LOS tangent slanting down (an
        # exactly vertical tangent leaves bascoord's basis undefined)
        xs = np.array([np.cos(np.radians(2.0)), 0.0, np.sin(np.radians(2.0))]) * (6367.421 + 17.0)
        ds = -xs / np.linalg.norm(xs) + np.array([0.0, 0.0, 0.3])
        for i in range(3):
            x[i], dx[i] = xs[i], ds[i]
        src = np.zeros(nd)
        c0 = time.perf_counter()
        quiet(lambda: olib.oracle_srcfunc_sca(r(octl), r(oatm), r(oaero), 0.0, x, dx, 0,
                                              src.ctypes.data_as(C.POINTER(C.c_double)), 1))
        cdt = time.perf_counter() - c0
        line["cpu_baseline"] = {"value": 32400 * nd / cdt, "unit": "pencil*channel/s", "cores": ncores(),
                                "kind": "port",
                                "sample": "the FP64 oracle's srcfunc_sca_3d at one scattering point of the layer "
                                          "(32,400 pencils x %d channels), %.1f s wall" % (nd, cdt)}
    emit(line)


def main():
    global _JSON_OUT
    args = parse()
    # stdout carries only the JSON line: fd 1 now points at stderr (the
    # library prints progress messages such as its table initialisation)
    sys.stdout.flush()
    _JSON_OUT = os.fdopen(os.dup(1), "w")
    os.dup2(2, 1)
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has a single MI355X "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyjurassic import jtypes as J, lib as L

    lib = L.load()
    for kv in args.option:
        key, val = kv.split("=", 1)
        lib.jur_set_option(key.encode(), float(val))
    if args.workload == "cirrus_16384":
        return bench_surrogate(args, lib)
    if args.workload == "jacobian_limb_17":
        return bench_jacobian(args, lib)
    if args.workload == "retrieval_limb_17":
        return bench_retrieval(args, lib)
    if args.workload == "aerosol3d_steep_16":
        return bench_scatter3d(args, lib)
    params = dict(REF_PARAMS if args.ref_params else CONFIG_PARAMS.get(args.workload, REF_PARAMS))
    if args.fan:
        params["fan"] = args.fan
    fan, mie_grid = params["fan"], params["mie_grid"]
    lib.jur_set_option(b"fan", fan)
    lib.jur_set_option(b"mie_grid", mie_grid)  # before the layer optics (setup)
    w, ctl, aero, scattering, (argc, argv) = setup(lib, args.workload, args.nray)
    aero_p = C.byref(aero) if aero is not None else None
    if args.variant is not None:
        lib.jur_set_option(b"variant", args.variant)
    if args.wpb is not None:
        lib.jur_set_option(b"wpb", args.wpb)
    nd, ng, nray = ctl.nd, ctl.ng, w.nray
    b, (rad, tau, tpz, npts), keep = device_batch(w, nd)
    # --chunk: one step is several calls over consecutive blocks of rays,
    # each writing its columns of the same outputs (jur_batch_t.out_ld)
    chunks = [b]
    if args.chunk and args.chunk < nray:
        chunks = []
        for r0 in range(0, nray, args.chunk):
            m = min(args.chunk, nray - r0)
            c = L.jur_batch_t()
            C.memmove(C.byref(c), C.byref(b), C.sizeof(b))
            c.nray, c.out_ld = m, nray
            if w.natm == nray:
                c.natm = m
                c.z, c.p, c.t = [x + 8 * r0 * w.nlev for x in (b.z, b.p, b.t)]
                c.q = b.q + 8 * r0 * ng * w.nlev
            for f in ("time", "obsz", "obslon", "obslat", "vpz", "vplon", "vplat", "tpz"):
                setattr(c, f, getattr(b, f) + 8 * r0)
            c.rad, c.tau, c.npts = b.rad + 4 * r0, b.tau + 4 * r0, b.npts + 4 * r0
            chunks.append(c)

    # per-step statistics, summed over the step's calls
    SUMMED = ("path_ms", "trace_ms", "integrated_points", "pencils", "path_launches", "graph_ms")
    last = {}

    def step():
        t_call = time.perf_counter()
        acc = dict.fromkeys(SUMMED, 0.0)
        for i, c in enumerate(chunks):
            lib.jur_formod_batch_device(C.byref(ctl), C.byref(c), aero_p, scattering, None)
            lib.jur_sync()
            for k in SUMMED:
                acc[k] += lib.jur_stat(k.encode())
            if len(chunks) > 1:
                print("chunk %d/%d: %d rays, %.1f s" % (i + 1, len(chunks), c.nray, time.perf_counter() - t_call),
                      file=sys.stderr, flush=True)
        last.update(acc)

    # at least one untimed call of a clear-air shape (the autotuner's run)
    for _ in range(args.warmup if scattering else max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    # eager calls: path-kernel and ray-trace times from the library's events
    path_ms, trace_ms, ipts, launches = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        path_ms.append(last["path_ms"])
        trace_ms.append(last["trace_ms"])
        ipts.append(last["integrated_points"])
        launches.append(last["path_launches"])
    torch.cuda.synchronize()
    dt_eager = time.perf_counter() - t0
    # hipGraph replays of the same call (one call per step: the recorded
    # graph holds the call's buffer addresses)
    graph = bool(args.graph) and len(chunks) == 1
    gpath, gcall = [], []
    if graph:
        lib.jur_set_option(b"graph", 1)
        step()  # eager call + capture
        for _ in range(max(args.warmup, 1)):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            gpath.append(last["path_ms"])
            gcall.append(last["graph_ms"])
            ipts.append(last["integrated_points"])
        torch.cuda.synchronize()
        dt_graph = time.perf_counter() - t0
        lib.jur_set_option(b"graph", 0)
    dt = dt_graph if graph else dt_eager
    ms_per_step = 1e3 * dt / args.steps
    value = nray * nd * args.steps / dt
    path_timed = gpath if graph and np.all(np.isfinite(gpath)) else path_ms

    npts_h = npts.cpu().numpy().astype(np.int64)
    pts = int(npts_h.sum())
    pencils = int(last["pencils"])
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
    as_int = lambda v: int(v) if np.isfinite(v) else None
    variant = args.variant if args.variant is not None else as_int(lib.jur_stat(b"tuned_variant"))
    wpb = args.wpb if args.wpb is not None else as_int(lib.jur_stat(b"tuned_wpb"))
    traffic, traffic_src = MEASURED_TRAFFIC.get((args.workload, nray, variant, wpb), (None, None))
    binding, binding_src = MEASURED_BINDING.get((args.workload, nray, variant, wpb), (None, None))

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
                   "fan": fan if scattering else None, "mie_grid": mie_grid if scattering else None,
                   "pencils": pencils, "integrated_points": ipts_call, "hip_graph": graph,
                   "parallelism": "single GPU"},
        "eager_ms_per_step": 1e3 * dt_eager / args.steps,
        "graph_ms_per_step": ms_per_step if graph else None,
        "graph_call_ms": float(np.mean(gcall)) if graph else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": PATH_KERNEL.get(variant, "k_path_fast"), "kernel_ms": path_avg,
                     "kernel_ms_source": "events inside the replayed graph" if path_timed is gpath
                     else "events around the eager launches",
                     "kernel_ms_eager": float(np.mean(path_ms)),
                     "trace_ms": float(np.mean(trace_ms)) if not scattering else None,
                     "launches_per_call": int(launches[-1]),
                     "alg_bytes_per_call": alg_bytes, "variant": variant, "wpb": wpb,
                     "binding": binding, "binding_source": binding_src},
    }

    if not args.no_cpu:
        ns = min(args.cpu_sample or CPU_SAMPLE[args.workload], nray)
        ncs = min(CPU_SAMPLE_CHANNELS.get(args.workload, nd), nd)
        if args.workload in CPU_SAMPLE_NADIR:
            # nadir rays through the cloud (a depth-2 limb ray costs the
            # oracle several minutes)
            sel = np.where((w.geom[4] == 0) & (w.geom[6] != 0))[0][:ns]
            how = "the first %d nadir rays" % len(sel)
        else:
            # every (nray/ns)-th ray: the sample spans limb and nadir geometries
            sel = np.linspace(0, nray - 1, ns).astype(np.int64)
            how = "%d of the %d rays (evenly spaced)" % (ns, nray)
        orad, otau, cdt, nc = cpu_oracle(w, argc, argv, sel, scattering, fan, ncs, mie_grid)
        line["cpu_baseline"] = {"value": len(sel) * ncs / cdt, "unit": "ray*channel/s", "cores": nc,
                                "kind": "port",
                                "sample": "%s, %s %d channels, FP64 oracle, %.1f s wall" %
                                          (how, "all" if ncs == nd else "the first", ncs, cdt)}
        line["accuracy"] = accuracy(rad.cpu().numpy().T[sel][:, :ncs], tau.cpu().numpy().T[sel][:, :ncs], orad,
                                    otau, ctl)
        line["accuracy"]["sample_rays"] = int(len(sel))
        line["accuracy"]["sample_channels"] = ncs
    emit(line)


if __name__ == "__main__":
    main()
