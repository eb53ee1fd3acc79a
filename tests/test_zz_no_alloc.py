"""The allocation-free Rg read (fmd_get_rg_noalloc) against fmd_get_rg through
the C ABI: the same bits for the mean and every chain, on two calls in a row,
the second on a system with fewer chains (the scratch array keeps the first
call's values past the second's chain count), and the over-limit error."""
import ctypes as C

import numpy as np
import pytest

from conftest import relax_kg

pytestmark = pytest.mark.gpu


def _read(lib, h, sid, fn, n):
    rg = C.c_double()
    nc = C.c_int32()
    per = np.zeros(max(1, n), np.float64)
    rc = fn(h, sid, C.byref(rg), per.ctypes.data_as(C.POINTER(C.c_double)), n, C.byref(nc))
    return rc, rg.value, per[:n], nc.value


def test_rg_noalloc_equals_rg(fmd):
    lib = fmd.api.lib()
    e = fmd.Engine(0)
    big = relax_kg(fmd, e, 60, 40, seed=90)
    small = relax_kg(fmd, e, 7, 30, seed=91)
    sids = [e.add(big), e.add(small)]
    e.step(sids, 50)
    for sid, n in zip(sids, (60, 7)):
        a = _read(lib, e.h, sid, lib.fmd_get_rg, n)
        b = _read(lib, e.h, sid, lib.fmd_get_rg_noalloc, n)
        assert a[0] == 0 and b[0] == 0 and a[3] == b[3] == n
        assert a[1] == b[1] and np.array_equal(a[2], b[2])
        assert 1.0 < b[1] < 5.0
        # the mean alone (no per-chain array)
        rg = C.c_double()
        assert lib.fmd_get_rg_noalloc(e.h, sid, C.byref(rg), None, 0, None) == 0 and rg.value == a[1]
    # one chain over the limit: a clear error, and the per-chain read still works
    n = 65537
    g = np.arange(n)
    side = int(np.ceil(n ** (1.0 / 3.0)))
    xyz = np.stack([g % side, (g // side) % side, g // side ** 2], 1) * 1.2 + 0.1
    from conftest import KG_LJ, pos4_of
    sid = e.add(fmd.Desc(pos4_of(xyz), float(side * 1.2), KG_LJ, mol=g))
    rg = C.c_double()
    nc = C.c_int32()
    assert lib.fmd_get_rg_noalloc(e.h, sid, C.byref(rg), None, 0, C.byref(nc)) == fmd.ERR_ARG and nc.value == n
    assert "65536" in lib.fmd_last_error(e.h).decode()
    assert lib.fmd_get_rg(e.h, sid, C.byref(rg), None, 0, None) == 0 and rg.value == 0.0
    e.close()
