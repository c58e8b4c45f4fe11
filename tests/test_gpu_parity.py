"""GPU parity: every HIP operator and the full timestep against the oracle on the same inputs.

Bars (written per test):
* operations whose order the GPU keeps from the reference (cell location, grid2par, the
  pair-swap derivatives, the filter, the tridiagonal solves): bit-identical;
* everything else, from one par2grid or FFT to whole timesteps: 1e-12 relative to the field's
  maximum (the bar for the plain FP64 form), with identical parcel counts and identical split,
  removed and added counts.  The only differences from the oracle are the grouping of
  par2grid's sums across cells and the FFT algorithm, each about 1e-16 relative per use.
Each test records its achieved errors as test properties (pytest --junitxml).
"""
import numpy as np
import pytest

from tests.cases import box, crowded_case, rel_err, thermal_case

pytestmark = pytest.mark.gpu

SHAPES = [(12, 10, 8), (16, 16, 6), (10, 12, 5)]


def make_pair(oracle_mod, shape, ppc_dir=2, cap_factor=3, ic=None, **kw):
    from pmpic_amd.lib import Context
    nx, ny, nz = shape
    if ic is None:
        ic = thermal_case(nx, ny, nz, ppc_dir, **kw)
    b = box(nx, ny, nz)
    cap = cap_factor * len(ic["x"]) + 64
    o = oracle_mod.Oracle(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                          b["zmax"], cap, 0.01, 0.01, 0.5)
    o.set_parcels(ic)
    g = Context(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"], cap,
                dtm=0.01, dtmax=0.01, timestep_prefactor=0.5)
    g.set_parcels(ic)
    return o, g, ic


def interior(a):
    return a[2:-2, 2:-2, :]


@pytest.mark.parametrize("shape", SHAPES)
def test_cache_weights_bitwise(oracle_mod, shape):
    o, g, _ = make_pair(oracle_mod, shape)
    o.cache_weights()
    (oi, oj, ok), (ox, oy, oz) = o.cache()
    (gi, gj, gk), (gx, gy, gz) = g.cache_weights()
    for a, b in ((oi, gi), (oj, gj), (ok, gk)):
        np.testing.assert_array_equal(a, b)
    for a, b in ((ox, gx), (oy, gy), (oz, gz)):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("shape", SHAPES)
def test_grid2par_bitwise(oracle_mod, shape):
    o, g, _ = make_pair(oracle_mod, shape)
    rng = np.random.default_rng(5)
    grid = rng.standard_normal(o.gshape)
    o.cache_weights()
    want, _ = o.grid2par(grid)
    got = g.grid2par(grid)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("shape", SHAPES)
def test_par2grid(oracle_mod, shape, bar):
    o, g, ic = make_pair(oracle_mod, shape)
    o.cache_weights()
    for name, var in (("b", ic["b"]), ("h", ic["h"]), ("sin_x_cos_z", np.sin(ic["x"]) * np.cos(ic["z"]))):
        want = o.par2grid(var)
        got = g.par2grid(var)
        bar(name, rel_err(interior(got), interior(want)), 1e-12)


@pytest.mark.parametrize("shape", SHAPES[:2])
def test_writebinary_grids(oracle_mod, shape, bar):
    """The grids writebinary dumps (writebinary.F90:217-233) after a step: p, q, r, hg by
    par2grid, hgliq = max(0, hg - exp(-z)), b = par2grid(b) + 12.5 hgliq; 1e-12 relative (the
    par2grid bar); hgliq must not be all zero here (the thermal's humidity exceeds exp(-z))."""
    o, g, _ = make_pair(oracle_mod, shape, vort=0.5)
    o.step()
    g.step()
    o.writebinary_grids()
    g.writebinary_grids()
    for name in ("p", "q", "r", "b", "hg", "hgliq"):
        want = interior(o.grid(name))
        bar(name, rel_err(interior(g.grid(name)), want), 1e-12)
    assert np.max(interior(o.grid("hgliq"))) > 0


def test_par2grid_and_steps_crowded_cells(oracle_mod, bar):
    """Cells holding 32, 33, 44 and 708 parcels: the binning's per-cell capacity edges and a
    cell with more parcels than one workgroup stages in LDS."""
    shape = (12, 10, 8)
    o, g, ic = make_pair(oracle_mod, shape, ic=crowded_case(*shape))
    o.cache_weights()
    for name, var in (("b", ic["b"]), ("p", ic["p"]), ("sin_x_cos_z", np.sin(ic["x"]) * np.cos(ic["z"]))):
        want = o.par2grid(var)
        got = g.par2grid(var)
        bar("par2grid_" + name, rel_err(interior(got), interior(want)), 1e-12)
    for step in range(2):
        o.step()
        g.step()
        assert g.n == o.n and g.last_counts() == o.last_counts(), step
    gp, op = g.get_parcels(), o.get_parcels()
    for k in ("x", "y", "z", "p", "q", "r", "vol", "b", "h"):
        bar(k, rel_err(gp[k], op[k]), 1e-12)


@pytest.mark.parametrize("shape", SHAPES)
def test_fft_roundtrip_vs_oracle(oracle_mod, shape, bar):
    o, g, _ = make_pair(oracle_mod, shape)
    rng = np.random.default_rng(7)
    grid = rng.standard_normal(o.gshape)
    want = o.fft_forward(grid)
    got = g.fft_forward(grid)
    bar("forward", rel_err(got, want), 1e-12)
    spec = rng.standard_normal(o.sshape)
    wb = o.fft_backward(spec)
    gb = g.fft_backward(spec)
    bar("backward", rel_err(interior(gb), interior(wb)), 1e-12)


@pytest.mark.parametrize("shape", SHAPES)
def test_spectral_ops_bitwise(oracle_mod, shape):
    o, g, _ = make_pair(oracle_mod, shape)
    rng = np.random.default_rng(11)
    s = rng.standard_normal(o.sshape)
    np.testing.assert_array_equal(g.diffx(s), o.diffx(s))
    np.testing.assert_array_equal(g.diffy(s), o.diffy(s))
    np.testing.assert_array_equal(g.filter(s), o.filter(s))
    np.testing.assert_array_equal(g.diffz(s), o.diffz(s))
    bot = rng.standard_normal(o.sshape[:2])
    top = rng.standard_normal(o.sshape[:2])
    np.testing.assert_array_equal(g.diffz(s, bot, top), o.diffz(s, bot, top))
    for df0 in (False, True):
        np.testing.assert_array_equal(g.laplinv(s, df0), o.laplinv(s, df0))


@pytest.mark.parametrize("shape", SHAPES)
def test_one_stage_components(oracle_mod, shape, bar):
    """vort2vel + vorticity_tendency + rk4 stage 1 from the same state."""
    o, g, _ = make_pair(oracle_mod, shape, vort=0.5)
    o.vort2vel()
    o.vorticity_tendency()
    g.vort2vel()
    g.vorticity_tendency()
    for name in ("p", "q", "r", "u", "v", "w", "b", "dp", "dq", "dr"):
        bar("grid_" + name, rel_err(interior(g.grid(name)), interior(o.grid(name))), 1e-12)
    for name in ("u_s", "v_s", "w_s"):
        bar(name, rel_err(g.spectral(name), o.spec(name)), 1e-12)
    o.rk4_integrator()
    g.rk4_integrator()
    assert abs(g.dtm - o.dtm) <= 1e-12 * o.dtm
    gp = g.get_parcels()
    op = o.get_parcels()
    for k in ("x", "y", "z", "p", "q", "r", "dxdt", "dydt", "dzdt", "dpdt", "dqdt", "drdt",
              "vol", "b", "h", "tag"):
        bar(k, rel_err(gp[k], op[k]), 1e-12)


def test_haloswap_permutation_matches(oracle_mod):
    """Parcels pushed across every face and corner land in the same slots as the reference's
    one-rank send-to-self unpack order."""
    o, g, ic = make_pair(oracle_mod, (12, 10, 8))
    n = len(ic["x"])
    rng = np.random.default_rng(3)
    L = ic["x"].max() + 0.1
    moved = dict(ic)
    moved["x"] = ic["x"] + rng.choice([-1.0, 0.0, 1.0], n) * rng.uniform(0, 0.3, n)
    moved["y"] = ic["y"] + rng.choice([-1.0, 0.0, 1.0], n) * rng.uniform(0, 0.3, n)
    moved["z"] = ic["z"] + rng.uniform(-0.2, 0.2, n)
    o.set_parcels(moved)
    g.set_parcels(moved)
    o.parcel_haloswap()
    g.parcel_haloswap()
    gp, op = g.get_parcels(), o.get_parcels()
    for k in ("x", "y", "z", "tag"):
        np.testing.assert_array_equal(gp[k], op[k])
    assert L > 0


@pytest.mark.parametrize("shape,kw", [((12, 10, 8), dict(vort=3.0, shrink=0.05, thin=0.1, near_split=0.05)),
                                      ((16, 16, 6), dict(vort=2.0, shrink=0.02, thin=0.05, near_split=0.03))])
def test_steps_match_oracle(oracle_mod, shape, kw, bar):
    o, g, _ = make_pair(oracle_mod, shape, **kw)
    for step in range(4):
        o.step()
        g.step()
        assert g.n == o.n, f"step {step}: parcel count {g.n} vs {o.n}"
        assert g.last_counts() == o.last_counts(), step
        gp, op = g.get_parcels(), o.get_parcels()
        for k in ("x", "y", "z", "p", "q", "r", "vol", "b", "h", "stretch"):
            bar(f"step{step + 1}_{k}", rel_err(gp[k], op[k]), 1e-12)
        assert abs(g.time - o.time) <= 1e-12


def test_twenty_steps_counts_match_oracle(oracle_mod, bar):
    """The count bar: identical parcel counts after split/merge for the first 20
    steps (and the split/removed/added counts of every step), with the fields still within
    1e-12 relative at the end."""
    o, g, _ = make_pair(oracle_mod, (12, 10, 8), vort=3.0, shrink=0.03, thin=0.08, near_split=0.04)
    for step in range(20):
        o.step()
        g.step()
        assert g.n == o.n, f"step {step}: parcel count {g.n} vs {o.n}"
        assert g.last_counts() == o.last_counts(), step
    gp, op = g.get_parcels(), o.get_parcels()
    for k in ("x", "y", "z", "p", "q", "r", "vol", "b", "h"):
        bar(k, rel_err(gp[k], op[k]), 1e-12)
    for name in ("u", "v", "w"):
        bar("grid_" + name, rel_err(interior(g.grid(name)), interior(o.grid(name))), 1e-12)


@pytest.mark.parametrize("shape", [(8, 6, 16), (8, 6, 32), (6, 8, 64), (6, 4, 128), (4, 4, 6),
                                   (4, 6, 40), (8, 6, 48), (4, 4, 51), (6, 4, 96), (4, 4, 127),
                                   (4, 4, 200), (4, 6, 256), (6, 4, 263), (4, 4, 300)])
def test_column_solvers_bitwise_all_forms(oracle_mod, shape):
    """diffz and laplinv bit-identical to the oracle for column lengths served by the
    register-resident solvers (17, 33, 65, 129 z nodes), by the four-lanes-per-column form of
    other lengths up to 129 (7, 41, 49 and its longest, 52, with 13 levels a lane; 97 and 128
    with 33), by the eight-lanes-per-column form of long columns (201, 257 and its longest,
    264) and by the LDS form beyond (301)."""
    from pmpic_amd.lib import Context
    nx, ny, nz = shape
    b = box(nx, ny, nz)
    o = oracle_mod.Oracle(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                          b["zmax"], 64, 0.01, 0.01, 0.5)
    g = Context(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"], 64)
    rng = np.random.default_rng(nz)
    s = rng.standard_normal(o.sshape)
    bot = rng.standard_normal(o.sshape[:2])
    top = rng.standard_normal(o.sshape[:2])
    np.testing.assert_array_equal(g.diffz(s), o.diffz(s))
    np.testing.assert_array_equal(g.diffz(s, bot, top), o.diffz(s, bot, top))
    for df0 in (False, True):
        np.testing.assert_array_equal(g.laplinv(s, df0), o.laplinv(s, df0))


@pytest.mark.parametrize("vort", [4.0, 2.0])
def test_adaptive_timestep_steps_match_oracle(oracle_mod, vort, bar):
    """With dtmax large the CFL or vorticity bound sets every step's dt: the GPU's device-side
    selection gives the oracle's dt sequence (and time), counts and fields over 5 steps."""
    from pmpic_amd.lib import Context
    shape = (12, 10, 8)
    nx, ny, nz = shape
    ic = thermal_case(nx, ny, nz, 2, vort=vort, shrink=0.02)
    b = box(nx, ny, nz)
    cap = 3 * len(ic["x"]) + 64
    o = oracle_mod.Oracle(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                          b["zmax"], cap, 0.01, 10.0, 0.5)
    o.set_parcels(ic)
    g = Context(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"], cap,
                dtm=0.01, dtmax=10.0, timestep_prefactor=0.5)
    g.set_parcels(ic)
    for step in range(5):
        o.step()
        g.step()
        bar(f"step{step + 1}_dtm", abs(g.dtm - o.dtm) / o.dtm, 1e-12)
        assert g.n == o.n, step
    bar("time", abs(g.time - o.time) / o.time, 1e-12)
    gp, op = g.get_parcels(), o.get_parcels()
    for k in ("x", "y", "z", "p", "q", "r", "vol", "b"):
        bar(k, rel_err(gp[k], op[k]), 1e-12)


def test_lean_mixing_bitwise(oracle_mod, monkeypatch):
    """A context without room for mixing's residual partial buffer (forced here with
    PMPIC_LEAN_MIXING=1) sums the kept and the residual parcels in two passes over one buffer:
    every parcel field bit-identical to the one-pass form over steps with splitting and mixing,
    and both equal to the oracle's counts."""
    from pmpic_amd.lib import Context
    shape = (12, 10, 8)
    ic = thermal_case(*shape, 2, vort=3.0, shrink=0.05, thin=0.1, near_split=0.05)
    b = box(*shape)
    cap = 3 * len(ic["x"]) + 64
    ctx = {}
    for lean in ("0", "1"):
        monkeypatch.setenv("PMPIC_LEAN_MIXING", lean)
        ctx[lean] = Context(*shape, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                            b["zmax"], cap)
        ctx[lean].set_parcels(ic)
    o = oracle_mod.Oracle(*shape, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                          b["zmax"], cap, 0.01, 0.01, 0.5)
    o.set_parcels(ic)
    mixed = 0
    for step in range(3):
        o.step()
        for g in ctx.values():
            g.step()
            assert g.n == o.n and g.last_counts() == o.last_counts(), step
        mixed += o.last_counts()[1] + o.last_counts()[2]
    assert mixed > 0
    a, c = ctx["0"].get_parcels(), ctx["1"].get_parcels()
    for k in a:
        np.testing.assert_array_equal(a[k], c[k])
    for name in ("vol", "b", "hg", "p", "q", "r"):
        np.testing.assert_array_equal(ctx["0"].grid(name), ctx["1"].grid(name))

