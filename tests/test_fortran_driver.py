"""The Fortran host (umwm_amd/fortran): the driver program umwm_gpu reads a
reference-format namelist, runs the inner loop operator by operator through
the same-named subroutines (or as fused substeps) over the C-ABI, and its
hourly swh / mwp / ustar match the FP64 oracle run of the same namelist."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_py
from umwm_amd import abi, build, rawfile

EXE = os.path.join(build.HERE, "bin", "umwm_gpu")

NAMELIST = """&DOMAIN
  isGlobal = .false.
  mm = 14
  nm = 11
  om = 25
  pm = 24
  fmin = 0.0313
  fmax = 2.0
  fprog = 2.0
  startTimeStr = '{start}'
  stopTimeStr  = '2012-01-01 02:00:00'
  dtg = 3600
  restart = {restart}
/
&PHYSICS
  g = 9.80665
  nu_air = 1.56E-5
  nu_water = 0.90E-6
  sfct = 0.07
  kappa = 0.4
  z = 10.
  gustiness = 0.0
  dmin = 10.
  explim = 0.9
  sin_fac = 0.11
  sin_diss1 = 0.10
  sin_diss2 = 0.001
  sds_fac = 42.
  sds_power = 2.4
  mss_fac = 360
  snl_fac = 5.0
  sdt_fac = 0.002
  sbf_fac = 0.003
  sbp_fac = 0.003
/
&GRID
  gridFromFile = .F.
  delx = 20000
  dely = 15000
  topoFromFile = .F.
  dpt = 800
  fillEstuaries = .F.
  fillLakes = .F.
/
&FORCING
  winds = .F.
  currents = .F.
  air_density = .F.
  water_density = .F.
  seaice = .F.
/
&FORCING_CONSTANT
  wspd0 = 12
  wdir0 = 0.3
  uc0 = 0
  vc0 = 0
  rhoa0 = 1.2
  rhow0 = 1030
  fice0 = 0.0
  fice_lth = 0.30
  fice_uth = 0.75
/
&OUTPUT
  outgrid = 1
  outspec = 0
  outrst = 6
  xpl = 7
  ypl = 5
  stokes = .t.
/
&STOKES
  depths = 0.1 0.2 0.5 1.0 2.0 5.0 10 20 50 100
/
&GPU
  precision = 8
  mode = '{mode}'
  jonswap = .t.
  js_fp = 0.25
  out_file = '{out}'
/
"""


def read_records(path, im):
    raw = open(path, "rb").read()
    rec = 8 + 3 * 8 * im
    out = []
    for k in range(len(raw) // rec):
        b = raw[k * rec:(k + 1) * rec]
        n, hour = np.frombuffer(b[:8], dtype=np.int32)
        assert n == im
        a = np.frombuffer(b[8:], dtype=np.float64).reshape(3, im)
        out.append((int(hour), a[0], a[1], a[2]))
    return out


def run_driver(tmp_path, mode, start="2012-01-01 00:00:00", restart=".false.", tag="run",
               template=NAMELIST):
    """Runs umwm_gpu in tmp_path (which gets the reference's output/ and
    restart/ folders); returns (record file path, completed process)."""
    out = tmp_path / f"{tag}.bin"
    nml = tmp_path / f"{tag}.nml"
    nml.write_text(template.format(mode=mode, out=str(out), start=start, restart=restart))
    for sub in ("output", "restart"):
        (tmp_path / sub).mkdir(exist_ok=True)
    r = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300,
                       cwd=str(tmp_path))
    print(r.stdout)
    return out, r


def oracle_hours():
    p = abi.make_params(mm=14, nm=11, om=25, pm=24, delx=20000.0, dely=15000.0, dpt=800.0,
                        wspd0=12.0, wdir0=0.3, stokes=1,
                        depths=[0.1, 0.2, 0.5, 1.0, 2.0, 5.0, 10, 20, 50, 100])
    o = oracle_py.Oracle(p, None, None, threads=2)
    o.init_jonswap(abi.make_jonswap(alpha=0.01, gamma=3.3, fp=0.25, fp_from_wind=False,
                                    dir_from_wind=True))
    o.stokes_drift(True)
    hours = []
    for _ in range(2):
        o.run_dtg()
        hours.append((o.get("ht").copy(), o.get("mwp").copy(), o.get("ustar").copy()))
    return o.im, hours


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["operators", "fused"])
def test_driver_matches_oracle(mode, tmp_path):
    assert os.path.exists(EXE), "Fortran driver not built (python -m umwm_amd.build)"
    out, r = run_driver(tmp_path, mode)
    assert r.returncode == 0, r.stdout + r.stderr
    im, ref = oracle_hours()
    recs = read_records(str(out), im)
    assert [h for h, *_ in recs] == [1, 2]
    for (hour, hs, tm, ust), (hs0, tm0, ust0) in zip(recs, ref):
        for a, b, name in ((hs, hs0, "swh"), (tm, tm0, "mwp"), (ust, ust0, "ustar")):
            err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
            assert err < 1e-10, f"{mode} hour {hour} {name} rel err {err:.3e}"
    # gridded output every hour (outgrid = 1): swh and ust are the record's
    # fields in single precision at the sea points
    for hour, hs, tm, ust in recs:
        f = rawfile.read(tmp_path / "output" / f"umwmout_2012-01-01_0{hour}:00:00.bin")
        sea = ~np.isnan(f["swh"].data[:, :, 0])
        assert sea.sum() == im
        assert np.array_equal(np.sort(f["swh"].data[:, :, 0][sea]), np.sort(hs.astype(np.float32)))
        assert np.array_equal(np.sort(f["ust"].data[:, :, 0][sea]), np.sort(ust.astype(np.float32)))
        assert np.all(np.isfinite(f["taux_ocn"].data[:, :, 0][sea]))
        assert np.any(f["taux_ocn"].data[:, :, 0][sea] != 0)


@pytest.mark.gpu
def test_driver_restart_continues_the_run(tmp_path):
    """outrst = 1 writes restart/umwmrst_<time>.bin each hour.  A second run
    with restart = .true. from 01:00 reads it into the state that
    umwm_initialize built (so, as in the reference, the surface Stokes drift
    of its first hour is still that of the initial spectrum) and takes no
    cold substep; its hour matches the FP64 oracle started the same way from
    the file's single-precision spectrum and ustar."""
    assert os.path.exists(EXE), "Fortran driver not built (python -m umwm_amd.build)"
    template = NAMELIST.replace("outrst = 6", "outrst = 1")
    out, r = run_driver(tmp_path, "fused", tag="full", template=template)
    assert r.returncode == 0, r.stdout + r.stderr
    rst = tmp_path / "restart" / "umwmrst_2012-01-01_01:00:00.bin"
    assert rst.exists()
    out2, r2 = run_driver(tmp_path, "fused", start="2012-01-01 01:00:00", restart=".true.",
                          tag="cont", template=template)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    f = rawfile.read(rst)
    p = abi.make_params(mm=14, nm=11, om=25, pm=24, delx=20000.0, dely=15000.0, dpt=800.0,
                        wspd0=12.0, wdir0=0.3, stokes=1, restart=1,
                        depths=[0.1, 0.2, 0.5, 1.0, 2.0, 5.0, 10, 20, 50, 100])
    o = oracle_py.Oracle(p, None, None, threads=2)
    o.init_jonswap(abi.make_jonswap(alpha=0.01, gamma=3.3, fp=0.25, fp_from_wind=False,
                                    dir_from_wind=True))
    o.stokes_drift(True)
    o.get("e")[:] = f["F"].data.T
    o.get("ustar")[:] = f["ust"].data
    o.run_dtg()
    im = o.im
    (hour, hs, tm, ust), = read_records(str(out2), im)
    assert hour == 1
    for a, b, name in ((hs, o.get("ht"), "swh"), (tm, o.get("mwp"), "mwp"),
                       (ust, o.get("ustar"), "ustar")):
        err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        assert err < 1e-10, f"restart {name} rel err {err:.3e}"
    # and it is a continuation: close to the uninterrupted run's 02:00
    full = read_records(str(out), im)
    assert float(np.max(np.abs(hs - full[1][1])) / np.max(full[1][1])) < 1e-3


def test_driver_rejects_bad_namelist(tmp_path):
    """A namelist the reference would stop on (pm not a multiple of 4) ends
    the driver with a message and a non-zero status, before any GPU work."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built")
    nml = tmp_path / "main.nml"
    nml.write_text(NAMELIST.format(mode="fused", out=str(tmp_path / "o.bin"),
                                   start="2012-01-01 00:00:00", restart=".false.").replace(
        "pm = 24", "pm = 30"))
    r = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "pm must be divisible by 4" in r.stdout
