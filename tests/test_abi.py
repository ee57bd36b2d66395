"""The C-ABI library without a GPU: it loads, exports every function the
header declares, its structs have the layout the ctypes mirrors assume, and
the calls that need no device report errors the documented way.  Also the
host-side logic (workload builders, tile candidates)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from umwm_amd import abi, autotune, configs, hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "umwm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(umwm_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_the_python_exports():
    names = declared_functions()
    assert len(names) >= 30
    assert sorted(set(hip.EXPORTS)) == names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(hip.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and nothing else with the public prefix (no undeclared entry points)
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted(set(re.findall(r"\b(umwm_hip_[a-z0-9_]+)$", out, flags=re.M)))
    assert exported == declared_functions()


def _c_layout():
    """Offsets and sizes of the header's structs as the C compiler sees them."""
    src = ["#include <stddef.h>", "#include <stdio.h>", '#include "umwm_hip.h"', "int main(void){"]
    for cname, py in (("umwm_params", abi.Params), ("umwm_grid_input", abi.GridInput),
                      ("umwm_forcing_input", abi.ForcingInput), ("umwm_jonswap", abi.Jonswap),
                      ("umwm_surrogate_config", abi.SurrogateConfig),
                      ("umwm_surrogate_policy", abi.SurrogatePolicy)):
        src.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    src.append("return 0;}")
    return "\n".join(src)


def test_struct_layout_matches_ctypes(tmp_path):
    c = tmp_path / "layout.c"
    c.write_text(_c_layout())
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(c), "-o",
                    str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    table = {}
    for ln in lines:
        if ln.strip():
            s, f, v = ln.split()
            table[(s, f)] = int(v)
    for cname, py in (("umwm_params", abi.Params), ("umwm_grid_input", abi.GridInput),
                      ("umwm_forcing_input", abi.ForcingInput), ("umwm_jonswap", abi.Jonswap),
                      ("umwm_surrogate_config", abi.SurrogateConfig),
                      ("umwm_surrogate_policy", abi.SurrogatePolicy)):
        assert table[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert table[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_error_strings_and_null_arguments():
    lib = hip.lib()
    for code in (0, -1, -2, -3, -4, -5, -6):
        s = lib.umwm_hip_error_string(code)
        assert s and len(s) >= 2
    # every domain entry point rejects a NULL domain with UMWM_ERR_ARGUMENT
    for name in ("umwm_hip_sin_d12", "umwm_hip_sds_d12", "umwm_hip_snl_d12", "umwm_hip_s_ice",
                 "umwm_hip_source", "umwm_hip_propagation", "umwm_hip_refraction",
                 "umwm_hip_update", "umwm_hip_diag", "umwm_hip_forcinginterpolate",
                 "umwm_hip_synchronize", "umwm_hip_end_first"):
        fn = getattr(lib, name)
        fn.argtypes = [C.c_void_p]
        fn.restype = C.c_int
        assert fn(None) == -1, name
    assert lib.umwm_hip_destroy(None) == 0  # destroying nothing is a no-op


def test_create_validates_params_before_touching_a_device():
    """Bad namelist values are rejected with UMWM_ERR_ARGUMENT (the
    reference stops in src/umwm_init.F90:103-112 for pm % 4 != 0)."""
    lib = hip.lib()
    out = C.c_void_p()
    for bad in (dict(pm=30), dict(om=2), dict(mm=2), dict(fmin=-1.0), dict(fmin=3.0), dict(dtg=0.0)):
        p = abi.make_params(**bad)
        rc = lib.umwm_hip_create(C.byref(p), None, None, 8, C.byref(out))
        assert rc == -1, bad
        assert lib.umwm_hip_last_error()
    p = abi.make_params()
    assert lib.umwm_hip_create(C.byref(p), None, None, 3, C.byref(out)) == -1  # precision


def test_params_defaults_follow_the_namelist():
    p = abi.make_params()
    assert (p.om, p.pm, p.mm, p.nm) == (37, 36, 21, 11)
    assert p.lm == 42
    assert p.sds_power == 2.4 and p.mss_fac == 360.0 and p.explim == 0.9
    with pytest.raises(KeyError):
        abi.make_params(nonsense=1)


def test_c2_workload_statistics():
    wl = configs.c2(seed=2, mm=180, nm=90, hours=1)
    z = wl.grid["z"]
    land = np.mean(z >= 0)
    assert 0.27 < land < 0.33
    depth = -z[z < 0]
    assert depth.min() >= 10.0 and depth.max() <= 6000.0
    spd = np.hypot(wl.forcing["uw"], wl.forcing["vw"])
    assert spd.min() >= 0.0 and spd.max() <= 25.0 + 1e-9
    assert wl.params.isglobal == 1 and wl.params.om == 37 and wl.params.pm == 36


def test_c3_domains_cover_the_stated_ranges():
    doms = configs.c3(n=64, seed=0)
    assert len(doms) == 64
    for wl in doms:
        p = wl.params
        assert 100 <= p.mm <= 600 and 100 <= p.nm <= 400
        assert 25 <= p.om <= 49 and 24 <= p.pm <= 72 and p.pm % 4 == 0


def test_tile_candidates():
    assert autotune.candidates(36, (6, 9, 18, 36)) == [(6, 6), (9, 4), (18, 2), (36, 1)]
    assert autotune.candidates(72, (6, 9)) == [(9, 8)]
    assert all(w <= 8 for _, w in autotune.candidates(24, autotune.INTEGRATE_D))
