"""Host-side pieces that need no GPU: config parsing, the workload generator, the C-ABI library
exports, and the status codes shared by the library and the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

from pmpic_amd import config
from pmpic_amd.initial import splitmix64, thermal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fortran_real_semantics():
    # FP32 read then 1e-8 rounding (conversions.F90:111-116, 445-453)
    assert config.fortran_real("6.28318531") == 6.28318548
    assert config.fortran_real("3.14159265") == 3.14159274
    # only the first 10 characters are read by (f10.0)
    assert config.fortran_real("6.283185307179586") == config.fortran_real("6.28318530")
    assert config.fortran_real("0.01") == 0.01
    assert config.fortran_real("2.0e-3") == 0.002
    assert config.infer("32") == 32
    assert config.infer(".true.") is True
    assert config.infer('"time"') == "time"


def test_parse_mcf_with_include_and_overrides(tmp_path):
    (tmp_path / "global_config").write_text("dtm=0.001\nnx=8\nlogging=3  #INFO\n")
    (tmp_path / "config.mcf").write_text(
        "global_configuration=global_config\nnx=32\nxmax=6.28318531\ne_values=0.3, -0.4 ,0.5\n"
        "group_names= start, integrator\nparcel_splitting_enabled=.true.\n")
    opts = config.parse_mcf(str(tmp_path / "config.mcf"))
    assert opts["nx"] == 32 and opts["dtm"] == 0.001 and opts["logging"] == 3
    assert opts["xmax"] == 6.28318548
    assert opts["e_values"] == [0.30000001, -0.40000001, 0.5]  # FP32 read, 1e-8 rounding
    assert opts["parcel_splitting_enabled"] is True
    config.apply_cli(opts, ["--nx=64"])
    assert opts["nx"] == 64


def test_splitmix64_known_values():
    # SplitMix64 reference outputs for seed 0 (Steele, Lea and Flood 2014; the generator of
    # java.util.SplittableRandom): first outputs 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4
    out = splitmix64(0, 2)
    assert int(out[0]) == 0xE220A8397B1DCDAF
    assert int(out[1]) == 0x6E789E6AA1B965F4


def test_thermal_workload_shape():
    lx, lz = config.fortran_real("6.28318531"), config.fortran_real("3.14159265")
    ic = thermal(4, 4, 4, lx, lx, lz, 3)
    assert len(ic["x"]) == 4 * 4 * 4 * 27
    assert np.all((ic["x"] >= 0) & (ic["x"] < lx)) and np.all((ic["z"] > 0) & (ic["z"] < lz))
    np.testing.assert_allclose(ic["vol"].sum(), lx * lx * lz, rtol=1e-12)
    # jitter stays inside the parcel's own cell
    dx = lx / 4
    cell = (np.arange(len(ic["x"])) // 27) % 4
    assert np.all(np.floor(ic["x"] / dx) == cell)
    again = thermal(4, 4, 4, lx, lx, lz, 3)
    np.testing.assert_array_equal(ic["b"], again["b"])


def test_status_codes_shared_with_oracle():
    from pmpic_amd.lib import STATUS
    text = open(os.path.join(ROOT, "include", "pmpic_status.h")).read()
    codes = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"(PMPIC_\w+) = (\d+)", text))
    for value, name in STATUS.items():
        assert codes[name] == value


def test_library_exports_every_declared_symbol():
    from pmpic_amd import lib
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libpmpic_hip.so not built")
    header = open(os.path.join(ROOT, "include", "pmpic_hip.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(pmpic_\w+)\(", header, re.M))
    assert declared == set(lib.EXPORTS)
    L = ctypes.CDLL(lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert L.pmpic_status_message is not None


def test_dump_roundtrip(tmp_path):
    from pmpic_amd.dumpio import FIELDS, read_parcels, write_parcels
    rng = np.random.default_rng(0)
    fields = {k: rng.standard_normal(37) for k in FIELDS}
    path = str(tmp_path / "parcels_00000_00001.dat")
    write_parcels(path, 1.25, [0, 1, 2, 3, 4, 5], fields)
    # 8 (time) + 48 (bounds) + 8 (count) + 17 arrays
    assert os.path.getsize(path) == 8 + 48 + 8 + 17 * 37 * 8
    t, b, back = read_parcels(path)
    assert t == 1.25 and list(b) == [0, 1, 2, 3, 4, 5]
    for k in FIELDS:
        np.testing.assert_array_equal(back[k], fields[k])
