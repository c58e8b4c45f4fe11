"""pmpic_par2grid_noalloc (par2grid of device arrays with a static scratch area, no allocation)
against pmpic_par2grid on the same context and values: bit-identical, on consecutive calls
whose inputs shrink (so stale scratch contents from a larger call must not leak into a smaller
one), up to the size cap, and refused past it."""
import ctypes

import numpy as np
import pytest

from tests.cases import box, thermal_case

pytestmark = pytest.mark.gpu


def _context(shape, ic):
    from pmpic_amd.lib import Context
    b = box(*shape)
    g = Context(*shape, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"],
                2 * len(ic["x"]) + 64)
    g.set_parcels(ic)
    return g


def _hip():
    """The HIP runtime the library itself runs on (its soname is already loaded by then)."""
    from pmpic_amd import lib
    lib.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _noalloc(g, var):
    """pmpic_par2grid_noalloc on device buffers the test allocates; returns (status, grid)."""
    from pmpic_amd import lib
    L = lib.load()
    hip = _hip()
    var = np.ascontiguousarray(var, dtype=np.float64)
    out = np.full(g.gshape, np.nan)
    dvar, dgrid = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dvar), max(var.nbytes, 8)) == 0
    assert hip.hipMalloc(ctypes.byref(dgrid), out.nbytes) == 0
    try:
        assert hip.hipMemcpy(dvar, var.ctypes.data, var.nbytes, 1) == 0  # host to device
        assert hip.hipMemcpy(dgrid, out.ctypes.data, out.nbytes, 1) == 0  # NaN fill
        st = L.pmpic_par2grid_noalloc(g.h, dvar, dgrid)
        assert hip.hipMemcpy(out.ctypes.data, dgrid, out.nbytes, 2) == 0  # device to host
    finally:
        hip.hipFree(dvar)
        hip.hipFree(dgrid)
    return st, out


def _check(g, var):
    want = g.par2grid(var)
    st, got = _noalloc(g, var)
    assert st == 0, st
    np.testing.assert_array_equal(got, want)
    return got


def test_noalloc_matches_par2grid_shrinking_calls():
    big = thermal_case(20, 18, 16, 3, seed=11, vort=1.0)
    g = _context((20, 18, 16), big)
    rng = np.random.default_rng(7)
    _check(g, big["b"])
    _check(g, rng.standard_normal(len(big["x"])))
    # the same context with a tenth of its parcels
    sub = {k: v[::10].copy() for k, v in big.items()}
    g.set_parcels(sub)
    _check(g, sub["h"])
    g.close()
    # a smaller grid after the larger one
    small = thermal_case(12, 10, 8, 2, seed=12, vort=1.0)
    g = _context((12, 10, 8), small)
    _check(g, small["p"])
    g.close()


def test_noalloc_at_the_size_cap_and_past_it():
    from pmpic_amd.lib import PMPIC_NOALLOC_MAX_CELLS
    # 32^3 cells (the cap) with 27 parcels per cell: 884,736 parcels
    ic = thermal_case(32, 32, 32, 3, seed=13, vort=1.0)
    assert 32 ** 3 == PMPIC_NOALLOC_MAX_CELLS
    g = _context((32, 32, 32), ic)
    _check(g, ic["q"])
    g.close()
    small = thermal_case(10, 12, 5, 2, seed=14)
    g = _context((10, 12, 5), small)
    _check(g, small["b"])
    g.close()
    # one cell layer past the cap is refused before anything runs
    over = thermal_case(32, 32, 33, 1, seed=15)
    g = _context((32, 32, 33), over)
    st, _ = _noalloc(g, over["b"])
    assert st == 10  # PMPIC_ERR_ARG
    g.close()
