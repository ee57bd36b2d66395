"""umwm_hip_substeps_noalloc: the fused substeps with their scratch in the
library's shared device array equal umwm_hip_substeps bit for bit, on two
calls in a row (the second domain smaller than the first, so it runs on
what the first left in the array), and a domain whose scratch exceeds the
array is refused with a size error."""
import numpy as np
import pytest

from umwm_amd import configs, hip

pytestmark = pytest.mark.gpu


def fresh(wl, precision):
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=precision)
    d.init_jonswap(wl.jonswap)
    d.stokes_drift(True)
    if wl.snapshots is not None:
        d.forcinginput(wl.forcing_input(1))
    return d


@pytest.mark.parametrize("precision", [8, 4])
def test_noalloc_equals_substeps(precision):
    cases = [configs.c1(), configs.tiny(seed=3, winds=True, currents=True)]
    assert cases[0].params.mm * cases[0].params.nm > cases[1].params.mm * cases[1].params.nm
    for wl in cases:  # larger domain first, then the smaller one
        a, b = fresh(wl, precision), fresh(wl, precision)
        try:
            for n in (1, 3, 5):  # the cold substep, then runs of substeps
                ra = a.substeps(n)
                rb = b.substeps_noalloc(n)
                assert ra == rb, (wl.name, n, ra, rb)
                for name in ("e", "ustar", "cd", "taux", "tauy", "ht"):
                    assert np.array_equal(a.get(name), b.get(name)), (wl.name, n, name)
        finally:
            a.close()
            b.close()


def test_noalloc_refuses_a_domain_beyond_the_array():
    # 200 x 150 regional at om = 49: about 29,000 sea points need
    # 2 x 4 x 49 x ldi x 8 bytes, more than the 32 MiB array
    wl = configs.regional_random(41, 200, 150, 49, 24, land_frac=0.0, hours=1)
    d = fresh(wl, 8)
    try:
        need = 2 * 4 * d.om * ((d.im + 1 + 63) // 64 * 64) * 8
        assert need > 32 << 20
        with pytest.raises(hip.UmwmError, match="scratch"):
            d.substeps_noalloc(1)
        # the domain is untouched and still runs on the other entry point
        assert d.scalar("nsub") == 0
        d.substeps(1)
    finally:
        d.close()
