"""The Fortran host side (pmpic_amd/fortran): the model-core stand-in driver registers the five
components' descriptors, whose hooks take `type(model_state_type)` as the reference's do and
reach libpmpic_hip.so through the context module (pmpic_gpu_mod) over iso_c_binding.  After the
host-mirror sync writebinary needs, its parcels and writebinary's ten grids are the same, bit for
bit, as the same steps driven through ctypes."""
import os
import subprocess

import numpy as np
import pytest

from tests.cases import box, thermal_case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "pmpic_amd", "pmpic_driver")


def test_fortran_driver_matches_ctypes(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (amdflang absent at build time)")
    from pmpic_amd.dumpio import read_parcels, write_parcels
    from pmpic_amd.lib import Context
    shape = (12, 10, 8)
    ic = thermal_case(*shape, 2, vort=2.0, shrink=0.03, thin=0.05, near_split=0.03)
    b = box(*shape)
    cap = 3 * len(ic["x"])
    fin, fout = str(tmp_path / "parcels_in.dat"), str(tmp_path / "parcels_out.dat")
    write_parcels(fin, 0.0, [b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"]], ic)
    args = [DRIVER, fin, fout, "3", str(shape[0]), str(shape[1]), str(shape[2]),
            repr(b["xmax"]), repr(b["ymax"]), repr(b["zmax"]), str(cap), "0.01", "0.01", "0.5"]
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    t, _, got = read_parcels(fout)
    g = Context(*shape, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"], cap)
    g.set_parcels(ic)
    g.step(3)
    want = g.get_parcels()
    assert abs(t - g.time) < 1e-15
    for k in want:
        np.testing.assert_array_equal(got[k], want[k])
    # writebinary's grids from the Fortran state's u .. vol, each (nz+1, ny+4, nx+4) z fastest
    from pmpic_amd.dumpio import GRID_FIELDS
    g.writebinary_grids()
    raw = np.fromfile(fout + ".grids", dtype="<f8")
    shp = (shape[0] + 4, shape[1] + 4, shape[2] + 1)
    assert raw.size == len(GRID_FIELDS) * np.prod(shp)
    for i, name in enumerate(GRID_FIELDS):
        got_g = raw[i * np.prod(shp):(i + 1) * np.prod(shp)].reshape(shp)
        np.testing.assert_array_equal(got_g, g.grid(name), err_msg=name)
