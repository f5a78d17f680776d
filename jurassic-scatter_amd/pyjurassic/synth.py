"""Seeded synthetic inputs for the benchmark configurations.

Writes reference-format emissivity tables (.bin, src/forwardmodel.c:636-651
layout) and filter functions (.filt), an aerosol model file and refractive
index files (src/scatter.c:444-453, 609-624), and returns per-ray
atmospheres and ray geometry as numpy arrays for the batched entry point.

Workloads (names used by bench.py and the tests):

* ``clear_limb_17``: 81 levels 0-80 km; p = 1013.25 exp(-z/7 km) hPa; a
  US-standard-like piecewise temperature profile plus seeded N(0, 1 K) noise
  smoothed over 3 levels; CO2 400 ppmv; H2O 1e-2 exp(-z/2 km) floored at
  4 ppmv; emitters CO2 and H2O; 4 channels 792-832 cm^-1; observer at
  800 km, 17 tangent heights 6-30 km (1.5 km steps).
* ``clear_batch_65536``: 65,536 rays, 60 % limb (tangent height U(3, 60) km),
  40 % nadir (ground zenith angle U(0, 60) deg, observer at 800 km), one
  atmosphere per ray on 121 levels 0-120 km (T noise sigma 3 K, H2O log-normal
  multiplier sigma 0.3); emitters CO2, H2O, O3 (Gaussian 8 ppmv at 30 km,
  sigma 7 km), HNO3 (Gaussian 5 ppbv at 24 km, sigma 6 km), F11 (260 pptv
  below 12 km, scale height 6 km above); 64 channels 700-1400 cm^-1.
* ``aerosol_limb_8192``: the same generator with an aerosol layer at 15-20 km
  (log-normal, median radius U(0.1, 1) um, width U(1.5, 2), number density
  U(1, 100) cm^-3, refractive index n = 1.4 + U(0, 0.1), k = U(1e-3, 0.5) per
  channel), 8,192 limb rays (tangent height 8-30 km), 32 channels
  780-1250 cm^-1, single scattering.
* ``cirrus_4096``: ice layer at 9-12 km (median radius U(5, 50) um, width
  1.6, number density U(0.01, 1) cm^-3, n = U(1.1, 1.5), k = U(0.01, 0.5)),
  4,096 limb and nadir rays, 32 channels, scattering depth 2.
* ``cirrus_16384``: the same cloud and channels (same layer and refractive
  indices), 16,384 other rays and atmospheres (own seed): the held-out set
  of the scattering-source surrogate.
* ``cirrus_train_8192``: the same cloud and channels, 8,192 further rays and
  atmospheres (own seed): the surrogate's training set (about 4 M
  (point, channel) pairs of the plain method).

The scattering workloads observe at 2000-01-01 12:00 UTC (time 43200 s), so
the solar source term is active; the layer optics use the reference's fixed
parameters (170 Gauss-Hermite radius nodes, 181 phase angles, 28 secondary
pencils per scattering point).

Emissivity tables: per emitter x channel, 20 log-spaced pressures
(0.01-1013 hPa) x 12 temperatures (180-320 K) x 60 log-spaced column
densities (1e14-1e24 molecules/cm^2), eps = 1 - exp(-(k u)^a) with
k = 10^U(-24,-19) (p/1013)^0.3 (T/250)^b, a = U(0.5, 1), b = U(-2, 2); each
(p, T) column is cut after its last strictly increasing FP32 emissivity, so
no column holds a saturated plateau.
"""
import struct
from pathlib import Path

import numpy as np

RE = 6367.421
SEED0 = 0x4A555241

WORKLOADS = {
    "clear_limb_17": 0,
    "clear_batch_65536": 1,
    "aerosol_limb_8192": 2,
    "cirrus_4096": 3,
    "cirrus_16384": 4,
    "cirrus_train_8192": 5,
}


def temperature_profile(z):
    """Piecewise US-standard-like temperature [K] at altitudes z [km]:
    288 K at the ground, -6.5 K/km to 11 km, 216.65 K to 20 km, +1 K/km to
    32 km, +2.8 K/km to 47 km, isothermal to 51 km, -2.8 K/km to 71 km,
    -2 K/km above."""
    zn = np.array([0.0, 11.0, 20.0, 32.0, 47.0, 51.0, 71.0])
    tn = np.array([288.0, 216.65, 216.65, 228.65, 270.65, 270.65, 214.65])
    t = np.interp(z, zn, tn)
    above = z > 71
    t[above] = tn[-1] - 2.0 * (z[above] - 71)
    return t


def smooth3(x):
    y = x.copy()
    y[1:-1] = (x[:-2] + x[1:-1] + x[2:]) / 3.0
    return y


def gas_profiles(z, names):
    out = []
    for g in names:
        if g == "CO2":
            q = np.full_like(z, 400e-6)
        elif g == "H2O":
            q = np.maximum(1e-2 * np.exp(-z / 2.0), 4e-6)
        elif g == "O3":
            q = 8e-6 * np.exp(-0.5 * ((z - 30.0) / 7.0) ** 2)
        elif g == "HNO3":
            q = 5e-9 * np.exp(-0.5 * ((z - 24.0) / 6.0) ** 2)
        elif g == "F11":
            q = np.where(z <= 12.0, 260e-12, 260e-12 * np.exp(-(z - 12.0) / 6.0))
        else:
            raise ValueError(g)
        out.append(q)
    return np.array(out)


def write_table(path, rng):
    """One synthetic emissivity table in the reference's binary layout."""
    p = np.logspace(np.log10(0.01), np.log10(1013.0), 20)
    t = np.linspace(180.0, 320.0, 12)
    u = np.logspace(14, 24, 60)
    lk0 = rng.uniform(-24, -19)
    a = rng.uniform(0.5, 1.0)
    b = rng.uniform(-2.0, 2.0)
    parts = [struct.pack("<i", len(p)), p.astype("<f8").tobytes(),
             np.full(len(p), len(t), "<i4").tobytes()]
    for pi in p:
        cols = []
        for ti in t:
            k = 10.0 ** lk0 * (pi / 1013.0) ** 0.3 * (ti / 250.0) ** b
            eps = (1.0 - np.exp(-(k * u) ** a)).astype(np.float32)
            n = 1
            while n < len(eps) and eps[n] > eps[n - 1]:
                n += 1
            n = max(n, 2)
            cols.append((u[:n].astype(np.float32), eps[:n]))
        parts.append(t.astype("<f8").tobytes())
        parts.append(np.array([len(c[0]) for c in cols], "<i4").tobytes())
        for cu, ce in cols:
            parts.append(cu.astype("<f4").tobytes())
            parts.append(ce.astype("<f4").tobytes())
    Path(path).write_bytes(b"".join(parts))


def write_filter(path, nu, half_width=1.0, n=201):
    x = np.linspace(nu - half_width, nu + half_width, n)
    lines = ["# $1 = wavenumber [cm^-1]\n# $2 = filter function\n\n"]
    lines += ["%.6f %g\n" % (xi, 1.0) for xi in x]
    Path(path).write_text("".join(lines))


class Workload:
    """Inputs of one workload; arrays are ray-major, profiles [natm][...]."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def ctl_lines(self):
        lines = ["NG = %d" % len(self.emitters)]
        lines += ["EMITTER[%d] = %s" % (i, e) for i, e in enumerate(self.emitters)]
        lines += ["ND = %d" % len(self.nu)]
        lines += ["NU[%d] = %.4f" % (i, v) for i, v in enumerate(self.nu)]
        lines += ["TBLBASE = %s" % self.tblbase, "RAYDS = 10", "REFRAC = 1"]
        lines += ["%s = %s" % kv for kv in self.extra_ctl.items()]
        return lines

    def write_ctl(self, path):
        Path(path).write_text("\n".join(self.ctl_lines()) + "\n")

    def select(self, rays=None, nd=None, ng=None):
        """A smaller case of the same workload: the rays at indices `rays`
        (with their atmospheres), the first nd channels and the first ng
        emitters (their tables are already written)."""
        kw = dict(self.__dict__)
        if rays is not None:
            rays = np.asarray(rays)
            kw["geom"] = np.ascontiguousarray(self.geom[:, rays])
            kw["nray"] = len(rays)
            if self.natm > 1:
                for f in ("z", "p", "t", "q"):
                    kw[f] = np.ascontiguousarray(getattr(self, f)[rays])
                kw["natm"] = len(rays)
        if nd is not None:
            kw["nu"] = self.nu[:nd]
        if ng is not None:
            kw["emitters"] = self.emitters[:ng]
            kw["q"] = np.ascontiguousarray(kw["q"][:, :ng])
        return Workload(**kw)


def _rays_limb(th, obsz=800.0):
    n = len(th)
    g = np.zeros((7, n))
    g[1] = obsz
    g[3] = np.degrees(np.arccos((RE + th) / (RE + obsz)))
    g[4] = th
    return g  # time, obsz, obslon, obslat, vpz, vplon, vplat


def _rays_nadir(zen_deg, obsz=800.0):
    """Ground view points seen from (obsz, 0, 0) under the local zenith
    angle zen_deg at the ground point."""
    th = np.radians(zen_deg)
    eta = np.arcsin(RE * np.sin(th) / (RE + obsz))
    gamma = th - eta
    n = len(th)
    g = np.zeros((7, n))
    g[1] = obsz
    g[4] = 0.0
    g[6] = np.degrees(gamma)
    return g


def make(name, outdir, nray=None, seed=None, nu=None):
    """Generate workload `name` into outdir (tables, filters, aerosol files).
    nray overrides the ray count (the same seeded sequence, truncated); nu
    replaces the channel wavenumbers (tables are generated for them)."""
    cfg = WORKLOADS[name]
    outdir = Path(outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(SEED0 + cfg if seed is None else seed)
    nu_arg = nu
    if cfg == 0:
        emitters = ["CO2", "H2O"]
        nu = np.linspace(792.0, 832.0, 4)
        z = np.arange(0.0, 81.0)
    elif cfg == 1:
        emitters = ["CO2", "H2O", "O3", "HNO3", "F11"]
        nu = np.linspace(700.0, 1400.0, 64)
        z = np.arange(0.0, 121.0)
    else:
        emitters = ["CO2", "H2O", "O3", "HNO3", "F11"]
        nu = np.linspace(780.0, 1250.0, 32)
        z = np.arange(0.0, 121.0)
    if nu_arg is not None:
        nu = np.asarray(nu_arg, dtype=float)
    tblbase = str(outdir / "syn")
    # the held-out cirrus set shares the training set's emissivity tables
    trng = np.random.default_rng(SEED0 + 100 + (3 if cfg in (4, 5) else cfg))
    for e in emitters:
        for v in nu:
            write_table("%s_%.4f_%s.bin" % (tblbase, v, e), trng)
    for v in nu:
        write_filter("%s_%.4f.filt" % (tblbase, v), v)

    nlev = len(z)
    p0 = 1013.25 * np.exp(-z / 7.0)
    tbase = temperature_profile(z)
    qbase = gas_profiles(z, emitters)
    extra = {}
    aero = None
    if cfg == 0:
        total = 17
        th = 6.0 + 1.5 * np.arange(total)
        geom = _rays_limb(th)
        natm = 1
        t = smooth3(tbase + rng.normal(0.0, 1.0, nlev))[None, :]
        q = qbase[None]
    else:
        total = {1: 65536, 2: 8192, 3: 4096, 4: 16384, 5: 8192}[cfg]
        n = total if nray is None else min(nray, total)
        if cfg == 1:
            nlimb = int(round(0.6 * total))
            kind = np.zeros(total, int)
            kind[nlimb:] = 1
            rng.shuffle(kind)
            th = rng.uniform(3.0, 60.0, total)
            zen = rng.uniform(0.0, 60.0, total)
        elif cfg == 2:
            kind = np.zeros(total, int)
            th = rng.uniform(8.0, 30.0, total)
            zen = np.zeros(total)
        else:  # cirrus: half limb, half nadir
            kind = (rng.uniform(size=total) < 0.5).astype(int)
            th = rng.uniform(5.0, 30.0, total)
            zen = rng.uniform(0.0, 60.0, total)
        tnoise = rng.normal(0.0, 3.0, (total, nlev))
        mult = np.exp(rng.normal(0.0, 0.3, total))
        kind, th, zen, tnoise, mult = kind[:n], th[:n], zen[:n], tnoise[:n], mult[:n]
        geom = np.where(kind[None, :] == 0, _rays_limb(th), _rays_nadir(zen))
        natm = n
        t = tbase[None, :] + np.apply_along_axis(smooth3, 1, tnoise)
        q = np.repeat(qbase[None], n, axis=0)
        ih = emitters.index("H2O")
        q[:, ih, :] *= mult[:, None]
        total = n
        if cfg in (2, 3, 4, 5):
            # the cloud of the held-out cirrus set is the training set's
            arng = np.random.default_rng(SEED0 + 200 + min(cfg, 3))
            if cfg == 2:
                top, bot, rr, ss = 20.0, 15.0, arng.uniform(0.1, 1.0), arng.uniform(1.5, 2.0)
                nn = arng.uniform(1.0, 100.0)
                nr_ = 1.4 + arng.uniform(0, 0.1, len(nu))
                ni_ = arng.uniform(1e-3, 0.5, len(nu))
            else:
                top, bot, rr, ss = 12.0, 9.0, arng.uniform(5.0, 50.0), 1.6
                nn = arng.uniform(0.01, 1.0)
                nr_ = arng.uniform(1.1, 1.5, len(nu))
                ni_ = arng.uniform(0.01, 0.5, len(nu))
            # refractive index file, descending wavenumber like the reference data
            rows = sorted(zip(nu, nr_, ni_), reverse=True)
            rows = [(rows[0][0] + 100.0, rows[0][1], rows[0][2])] + rows + \
                   [(rows[-1][0] - 100.0, rows[-1][1], rows[-1][2])]
            refrac = outdir / "refrac.dat"
            refrac.write_text("# wn[cm-1] real imag\n" +
                              "".join("%.6f %.6f %.6e\n" % r for r in rows))
            aerofile = outdir / "aero.tab"
            aerofile.write_text("%g %g %g MIE %s %g %g %g\n" % (top, bot, 0.1, refrac, nn, rr, ss))
            aero = str(aerofile)
            extra = {"SCA_N": "1", "SCA_MULT": "1" if cfg == 2 else "2"}
            # local noon at longitude 0 on 2000-01-01: the solar term is on
            geom[0] = 43200.0
    w = Workload(
        name=name, cfg=cfg, emitters=emitters, nu=nu, tblbase=tblbase, nray=total, nlev=nlev,
        natm=natm, z=np.repeat(z[None], natm, 0).copy(), p=np.repeat(p0[None], natm, 0).copy(),
        t=np.ascontiguousarray(t), q=np.ascontiguousarray(q), geom=np.ascontiguousarray(geom),
        aerofile=aero, extra_ctl=extra, outdir=outdir)
    return w


def write_atm_file(path, w, ia=0):
    """One profile of a workload in atm.tab format (for the obs/atm tools)."""
    lines = []
    for l in range(w.nlev):
        row = ["%.2f" % 0.0, "%.17g" % w.z[ia, l], "0", "0", "%.17g" % w.p[ia, l], "%.17g" % w.t[ia, l]]
        row += ["%.17g" % w.q[ia, g, l] for g in range(len(w.emitters))] + ["0"]
        lines.append(" ".join(row))
    Path(path).write_text("\n".join(lines) + "\n")
