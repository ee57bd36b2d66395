"""The Fortran host (umwm_amd/fortran): the driver program umwm_gpu reads a
reference-format namelist, runs the inner loop operator by operator through
the same-named subroutines (or as fused substeps) over the C-ABI, and its
hourly swh / mwp / ustar match the FP64 oracle run of the same namelist."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_py
from umwm_amd import abi, build

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
  startTimeStr = '2012-01-01 00:00:00'
  stopTimeStr  = '2012-01-01 02:00:00'
  dtg = 3600
  restart = .false.
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
    out = tmp_path / "out.bin"
    nml = tmp_path / "main.nml"
    nml.write_text(NAMELIST.format(mode=mode, out=str(out)))
    r = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    im, ref = oracle_hours()
    recs = read_records(str(out), im)
    assert [h for h, *_ in recs] == [1, 2]
    for (hour, hs, tm, ust), (hs0, tm0, ust0) in zip(recs, ref):
        for a, b, name in ((hs, hs0, "swh"), (tm, tm0, "mwp"), (ust, ust0, "ustar")):
            err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
            assert err < 1e-10, f"{mode} hour {hour} {name} rel err {err:.3e}"


def test_driver_rejects_bad_namelist(tmp_path):
    """A namelist the reference would stop on (pm not a multiple of 4) ends
    the driver with a message and a non-zero status, before any GPU work."""
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built")
    nml = tmp_path / "main.nml"
    nml.write_text(NAMELIST.format(mode="fused", out=str(tmp_path / "o.bin")).replace(
        "pm = 24", "pm = 30"))
    r = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "pm must be divisible by 4" in r.stdout
