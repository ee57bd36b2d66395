"""Synthetic workloads (namelist values, grids, forcing, initial spectra).

The reference has no initializer and reads grids and forcing from NetCDF
files; these builders produce the same quantities in memory:

  c1  64x48 regional all-sea grid, dx = dy = 10 km, depth 4000 m,
      om = 37 (0.0313-2 Hz), pm = 36, uniform 15 m/s wind toward 0 rad
      (from 270 deg, mathematical convention of docs/umwm_manual_v2.tex),
      JONSWAP alpha = 0.01, gamma = 3.3, fp = 0.3 Hz, cos^2 spreading; 1 h.
  c2  720x360 global lat-lon grid at 0.5 deg; land from a seeded Gaussian
      random field (1000 km correlation) thresholded to 30 % land; depth
      log-normal (median 3000 m, sigma 0.8, 300 km correlation) clipped to
      10-6000 m; u10, v10 seeded Gaussian random fields (500 km), speed
      rescaled to 0-25 m/s; JONSWAP from the local wind; 6 h.
  c3  64 uneven domains (100x100 .. 600x400, om 25-49, pm 24-72) with
      their own land fraction, depth and uniform wind; 3 h.
  c5  hurricane: 1500x1200 grid at 2 km, om = 49, pm = 72, linear shelf
      5 m -> 3000 m over 200 km from a seeded coastline, translating
      Holland (1980) vortex (Rmax 30 km, Vmax 60 m/s, B 1.5, 5 m/s);
      ensemble members perturb heading (+-10 deg) and Vmax (+-10 %).

Gaussian random fields use random Fourier features of a Gaussian
covariance exp(-r^2 / 2L^2) evaluated at 3-D positions on the sphere
(chord distance), so they are seeded, grid-independent and periodic in
longitude.
"""
import math

import numpy as np

from . import abi

G = 9.80665
R_EARTH = 6.371009e6


class Workload:
    """Everything needed to create a domain and start it."""

    def __init__(self, name, params, grid=None, forcing=None, jonswap=None, hours=1,
                 snapshots=None, note="", storm=None):
        self.name = name
        self.params = params
        self.grid = grid          # dict(lon=, lat=, z=, exclude=) of (mm, nm) arrays
        self.forcing = forcing    # dict(uw=, vw=, ...) of (mm, nm) arrays at t = 0
        self.jonswap = jonswap
        self.hours = hours
        self.snapshots = snapshots  # callable(hour) -> forcing dict at hour
        self.note = note
        self.storm = storm          # callable(hour) -> abi.Holland of that snapshot, or None

    def grid_input(self):
        if not self.grid:
            return abi.make_grid_input()
        return abi.make_grid_input(**self.grid)

    def forcing_input(self, hour=0):
        f = self.forcing if hour == 0 or self.snapshots is None else self.snapshots(hour)
        if not f:
            return None
        return abi.make_forcing_input(**f)


# ---------------------------------------------------------------------------
# random fields
# ---------------------------------------------------------------------------
def sphere_xyz(lon_deg, lat_deg):
    lo, la = np.radians(lon_deg), np.radians(lat_deg)
    return np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)], -1) * R_EARTH


def grf(xyz, corr_m, rng, nfeat=256, chunk=1 << 16):
    """Zero-mean, unit-variance Gaussian random field with covariance
    exp(-r^2 / (2 corr^2)) at positions xyz [..., 3] (metres)."""
    shp = xyz.shape[:-1]
    x = xyz.reshape(-1, xyz.shape[-1])
    k = rng.normal(0.0, 1.0 / corr_m, size=(nfeat, x.shape[1]))
    ph = rng.uniform(0, 2 * np.pi, size=nfeat)
    out = np.empty(x.shape[0])
    for a in range(0, x.shape[0], chunk):
        out[a:a + chunk] = np.cos(x[a:a + chunk] @ k.T + ph).sum(1)
    out *= math.sqrt(2.0 / nfeat)
    return out.reshape(shp)


def planar_xyz(mm, nm, dx, dy):
    x = (np.arange(mm) * dx)[:, None] * np.ones((1, nm))
    y = np.ones((mm, 1)) * (np.arange(nm) * dy)[None, :]
    return np.stack([x, y], -1)


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
def c1(mm=64, nm=48):
    p = abi.make_params(mm=mm, nm=nm, om=37, pm=36, fmin=0.0313, fmax=2.0, fprog=2.0,
                        dtg=3600.0, delx=1e4, dely=1e4, dpt=4000.0, wspd0=15.0, wdir0=0.0,
                        uc0=0.0, vc0=0.0, rhoa0=1.2, rhow0=1030.0, fice0=0.0, stokes=1)
    js = abi.make_jonswap(alpha=0.01, gamma=3.3, sigma_a=0.07, sigma_b=0.09, fp=0.3,
                          fp_from_wind=False, dir_from_wind=True)
    return Workload("c1", p, jonswap=js, hours=1,
                    note="64x48 regional, 10 km, 4000 m, 15 m/s toward 0 rad, JONSWAP fp=0.3 Hz")


def c1_seeded(seed, mm=64, nm=48, hours=1, vary=True):
    """The config-1 grid (64x48, 10 km, 4000 m, om = 37, pm = 36) with a
    seeded wind (mean speed 10-20 m/s, any direction) and JONSWAP start
    (fp 0.2-0.4 Hz, alpha 0.005-0.02): the sample source of the learned
    surrogate, whose evaluation takes a seed the training never saw.  With
    `vary` the wind is a forcing field whose speed varies by +-25 % and
    direction by +-20 deg over the grid (Gaussian random fields of 150 km
    correlation), so that the points of one run differ; else it is uniform
    as in config 1."""
    rng = np.random.default_rng(10_000 + seed)
    w = float(rng.uniform(10.0, 20.0))
    wd = float(rng.uniform(-np.pi, np.pi))
    alpha = float(rng.uniform(0.005, 0.02))
    fp = float(rng.uniform(0.2, 0.4))
    kw = dict(mm=mm, nm=nm, om=37, pm=36, fmin=0.0313, fmax=2.0, fprog=2.0, dtg=3600.0,
              delx=1e4, dely=1e4, dpt=4000.0, wspd0=w, wdir0=wd, uc0=0.0, vc0=0.0, rhoa0=1.2,
              rhow0=1030.0, fice0=0.0, stokes=1)
    forcing = None
    if vary:
        xy = planar_xyz(mm, nm, 1e4, 1e4)
        sp = w * (1 + 0.25 * np.tanh(grf(xy, 1.5e5, rng, nfeat=128)))
        dr = wd + np.radians(20.0) * np.tanh(grf(xy, 1.5e5, rng, nfeat=128))
        forcing = dict(uw=sp * np.cos(dr), vw=sp * np.sin(dr))
        kw.update(winds=1)
    p = abi.make_params(**kw)
    js = abi.make_jonswap(alpha=alpha, gamma=3.3, sigma_a=0.07, sigma_b=0.09, fp=fp,
                          fp_from_wind=False, dir_from_wind=True)
    return Workload(f"c1_s{seed}", p, forcing=forcing, jonswap=js, hours=hours,
                    snapshots=(lambda h: forcing) if forcing else None,
                    note=f"config-1 grid, seed {seed}: {w:.1f} m/s toward {wd:.2f} rad")


def c2(seed=2, mm=720, nm=360, om=37, pm=36, hours=6, land_frac=0.30):
    rng = np.random.default_rng(seed)
    dlon, dlat = 360.0 / mm, 180.0 / nm
    lon = (dlon / 2 + dlon * np.arange(mm))[:, None] * np.ones((1, nm))
    lat = np.ones((mm, 1)) * (-90 + dlat / 2 + dlat * np.arange(nm))[None, :]
    xyz = sphere_xyz(lon, lat)
    land = grf(xyz, 1.0e6, rng)
    thr = np.quantile(land, 1.0 - land_frac)
    logd = grf(xyz, 3.0e5, rng)
    depth = np.clip(3000.0 * np.exp(0.8 * logd), 10.0, 6000.0)
    z = np.where(land >= thr, 1.0, -depth)  # z >= 0 is land
    u = grf(xyz, 5.0e5, rng)
    v = grf(xyz, 5.0e5, rng)
    spd = np.hypot(u, v)
    spd_new = 25.0 * (spd - spd.min()) / max(spd.max() - spd.min(), 1e-12)
    scale = np.where(spd > 0, spd_new / np.maximum(spd, 1e-12), 0.0)
    uw, vw = u * scale, v * scale
    p = abi.make_params(mm=mm, nm=nm, om=om, pm=pm, fmin=0.0313, fmax=2.0, fprog=2.0,
                        dtg=3600.0, isglobal=1, gridfromfile=1, topofromfile=1, winds=1,
                        dmin=10.0, stokes=1)
    js = abi.make_jonswap(alpha=0.01, gamma=3.3, fp_from_wind=True, dir_from_wind=True)
    forcing = dict(uw=uw, vw=vw)
    return Workload(f"c2_s{seed}", p, grid=dict(lon=lon, lat=lat, z=z), forcing=forcing,
                    jonswap=js, hours=hours, snapshots=lambda h: forcing,
                    note=f"{mm}x{nm} global lat-lon {dlon} deg, GRF land {land_frac:.0%}, "
                         f"log-normal depth, GRF winds 0-25 m/s, seed {seed}")


def regional_random(seed, mm, nm, om, pm, dx=1e4, land_frac=0.2, shelf=False, wspd=15.0,
                    wdir=None, hours=1, corr=None):
    """A regional domain with its own land mask, depth and uniform wind."""
    rng = np.random.default_rng(seed)
    xy = planar_xyz(mm, nm, dx, dx)
    corr = corr if corr is not None else 0.15 * min(mm, nm) * dx
    land = grf(xy, corr, rng, nfeat=256)
    if land_frac > 0:
        thr = np.quantile(land, 1.0 - land_frac)
        isl = land >= thr
    else:
        isl = np.zeros((mm, nm), bool)
    if shelf:
        dmax = rng.uniform(100.0, 4000.0)
        depth = 10.0 + (dmax - 10.0) * (np.arange(mm)[:, None] / max(mm - 1, 1)) * np.ones((1, nm))
    else:
        depth = np.full((mm, nm), rng.uniform(10.0, 4000.0))
    z = np.where(isl, 1.0, -depth)
    wd = rng.uniform(-np.pi, np.pi) if wdir is None else wdir
    p = abi.make_params(mm=mm, nm=nm, om=om, pm=pm, fmin=0.0313, fmax=2.0, fprog=2.0,
                        dtg=3600.0, delx=dx, dely=dx, topofromfile=1, dmin=10.0,
                        wspd0=wspd, wdir0=wd, stokes=1)
    js = abi.make_jonswap(alpha=0.01, gamma=3.3, fp_from_wind=True, dir_from_wind=True)
    return Workload(f"reg_s{seed}_{mm}x{nm}_{om}x{pm}", p, grid=dict(z=z), jonswap=js,
                    hours=hours)


def c3(seed=3, n=64, hours=3, scale=1.0):
    """64 uneven domains; `scale` < 1 shrinks the grids for quick runs."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        mm = max(8, int(rng.integers(100, 601) * scale))
        nm = max(8, int(rng.integers(100, 401) * scale))
        om = int(rng.integers(25, 50))
        pm = int(rng.choice(np.arange(24, 73, 4)))
        frac = float(rng.uniform(0.0, 0.4))
        shelf = bool(rng.integers(0, 2))
        w = float(rng.uniform(5.0, 30.0))
        out.append(regional_random(int(rng.integers(1 << 30)), mm, nm, om, pm, dx=1e4,
                                   land_frac=frac, shelf=shelf, wspd=w, hours=hours))
    return out


def holland_wind(x, y, xc, yc, vmax, rmax, B=1.5, rho=1.15, dp=None, trans=(0.0, 0.0),
                 lat_deg=20.0):
    """Holland (1980) gradient wind with inflow rotation and half the
    translation speed added; northern-hemisphere cyclonic rotation."""
    f = 2 * 7.292e-5 * math.sin(math.radians(lat_deg))
    if dp is None:
        dp = rho * math.e * vmax ** 2 / B
    r = np.hypot(x - xc, y - yc) + 1.0
    a = (rmax / r) ** B
    v = np.sqrt(B * dp / rho * a * np.exp(-a) + (r * f / 2) ** 2) - r * f / 2
    th = np.arctan2(y - yc, x - xc) + np.pi / 2 + np.radians(20.0)  # 20 deg inflow
    u = v * np.cos(th) + 0.5 * trans[0]
    w = v * np.sin(th) + 0.5 * trans[1]
    return u, w


def c5(member=0, seed=5, mm=1500, nm=1200, om=49, pm=72, hours=24, dx=2000.0):
    rng = np.random.default_rng(seed)
    rm = np.random.default_rng(seed * 1000 + member)
    heading = math.radians(90.0 + (rm.uniform(-10, 10) if member else 0.0))
    vmax = 60.0 * (1 + (rm.uniform(-0.1, 0.1) if member else 0.0))
    xy = planar_xyz(mm, nm, dx, dx)
    # seeded coastline near the southern edge, shelf deepening northward
    wob = grf(xy[:, :1, :], 5e4, rng, nfeat=128)[:, 0] * 2e4
    ycoast = 0.12 * nm * dx + wob
    dist = xy[..., 1] - ycoast[:, None]
    depth = 5.0 + (3000.0 - 5.0) * np.clip(dist / 2e5, 0, 1)
    z = np.where(dist <= 0, 1.0, -depth)
    p = abi.make_params(mm=mm, nm=nm, om=om, pm=pm, fmin=0.0313, fmax=2.0, fprog=2.0,
                        dtg=3600.0, delx=dx, dely=dx, topofromfile=1, dmin=5.0, winds=1,
                        stokes=1)
    x0, y0 = 0.5 * mm * dx, 0.8 * nm * dx
    sp = 5.0
    vx, vy = sp * math.cos(heading + math.pi), sp * math.sin(heading + math.pi)

    def snap(hour):
        xc, yc = x0 + vx * 3600 * hour, y0 + vy * 3600 * hour
        u, v = holland_wind(xy[..., 0], xy[..., 1], xc, yc, vmax, 3.0e4, 1.5, trans=(vx, vy))
        return dict(uw=u, vw=v)

    def storm(hour):
        """The same snapshot's storm, for generation on the device."""
        return abi.make_holland(x0 + vx * 3600 * hour, y0 + vy * 3600 * hour, vmax, 3.0e4, 1.5,
                                trans=(vx, vy))

    js = abi.make_jonswap(alpha=0.01, gamma=3.3, fp_from_wind=True, dir_from_wind=True)
    return Workload(f"c5_m{member}", p, grid=dict(z=z), forcing=snap(0), jonswap=js,
                    hours=hours, snapshots=snap, storm=storm,
                    note=f"hurricane member {member}: heading {math.degrees(heading):.1f} deg, "
                         f"vmax {vmax:.1f} m/s")


def tiny(seed=0, mm=14, nm=11, om=25, pm=24, land=True, shelf=True, isglobal=False,
         winds=False, currents=False, wspd=12.0, wdir=0.3, ice=False, density=False,
         dmin=10.0, dshelf=20.0):
    """Small domains for parity tests: optional land, sloping depth (from
    dshelf at the shallow edge), periodic lat-lon grid, time-varying wind
    and currents, sea ice (a field spread over [0, 1], so that points lie
    below fice_lth, between the two thresholds where Kohout attenuation
    acts, and above fice_uth where advection is switched off) and air/water
    density fields."""
    rng = np.random.default_rng(seed)
    kw = dict(mm=mm, nm=nm, om=om, pm=pm, fmin=0.0313, fmax=2.0, fprog=2.0, dtg=3600.0,
              delx=2e4, dely=1.5e4, wspd0=wspd, wdir0=wdir, stokes=1, dmin=dmin)
    grid = {}
    forcing = None
    if isglobal:
        dlon, dlat = 360.0 / mm, 120.0 / nm
        lon = (dlon / 2 + dlon * np.arange(mm))[:, None] * np.ones((1, nm))
        lat = np.ones((mm, 1)) * (-60 + dlat / 2 + dlat * np.arange(nm))[None, :]
        grid.update(lon=lon, lat=lat)
        kw.update(isglobal=1, gridfromfile=1)
    if land or shelf or isglobal:
        depth = np.full((mm, nm), 800.0)
        if shelf:
            depth = dshelf + 1500.0 * (np.arange(mm)[:, None] / (mm - 1)) ** 2 * np.ones((1, nm))
            depth = depth * (1 + 0.3 * rng.uniform(-1, 1, size=(mm, nm)))
        z = -depth
        if land:
            ml = rng.uniform(size=(mm, nm)) < 0.12
            z = np.where(ml, 1.0, z)
        grid["z"] = z
        kw.update(topofromfile=1)
    if winds or currents or ice or density:
        forcing = {}
    if winds:
        kw.update(winds=1)
        ang = wdir + 0.4 * rng.uniform(-1, 1, size=(mm, nm))
        spd = wspd * (1 + 0.3 * rng.uniform(-1, 1, size=(mm, nm)))
        forcing.update(uw=spd * np.cos(ang), vw=spd * np.sin(ang))
    if currents:
        kw.update(currents=1)
        forcing.update(uc=0.3 * rng.uniform(-1, 1, size=(mm, nm)),
                       vc=0.3 * rng.uniform(-1, 1, size=(mm, nm)))
    if ice:
        kw.update(seaice=1)
        forcing.update(fice=rng.uniform(0.0, 1.0, size=(mm, nm)))
    if density:
        kw.update(air_density=1, water_density=1)
        forcing.update(rhoa=1.2 * (1 + 0.05 * rng.uniform(-1, 1, size=(mm, nm))),
                       rhow=1025.0 * (1 + 0.01 * rng.uniform(-1, 1, size=(mm, nm))))
    p = abi.make_params(**kw)
    js = abi.make_jonswap(alpha=0.01, gamma=3.3, fp=0.25, fp_from_wind=False,
                          dir_from_wind=True)
    snaps = None
    if forcing is not None:
        f0 = forcing

        def snaps(h):
            r2 = np.random.default_rng(seed + 1000 * h)
            out = {}
            for key, val in f0.items():
                out[key] = val * (1 + 0.2 * r2.uniform(-1, 1, size=val.shape))
            if "fice" in out:
                out["fice"] = np.clip(out["fice"], 0.0, 1.0)
            return out
    return Workload(f"tiny_s{seed}", p, grid=grid or None, forcing=forcing, jonswap=js,
                    hours=1, snapshots=snaps)
