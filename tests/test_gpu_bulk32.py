"""FP32-bulk form (precision 1): the plain form's step with each position held as a cell index
plus three FP32 offsets, FP32 parcel fields, FP32 par2grid sums and interpolation.

It is checked against the plain FP64 GPU form on the same inputs; the FP64 form is itself pinned
to the oracle by test_gpu_parity.py.  Bars (written per test): the parcel transfer round trip to
FP32 rounding of the offsets; grids and velocities within 1e-5 relative (the FP32-bulk
form's velocity bar) after the steps run; identical split/mix counts on cases
where no parcel sits within FP32 rounding of a threshold; volume conserved to FP32 rounding.
"""
import numpy as np
import pytest

from tests.cases import LX, LZ, box, crowded_case, rel_err, thermal_case

pytestmark = pytest.mark.gpu


def make(shape, precision, ic, cap_factor=3):
    from pmpic_amd.lib import Context
    nx, ny, nz = shape
    b = box(nx, ny, nz)
    cap = cap_factor * len(ic["x"]) + 64
    g = Context(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"], cap,
                dtm=0.01, dtmax=0.01, timestep_prefactor=0.5, precision=precision)
    g.set_parcels(ic)
    return g


def test_transfer_roundtrip():
    shape = (12, 10, 8)
    ic = thermal_case(*shape, 2, vort=1.0)
    g = make(shape, 1, ic)
    got = g.get_parcels()
    assert g.n == len(ic["x"])
    d = (LX / shape[0], LX / shape[1], LZ / shape[2])
    for k, dk in zip(("x", "y", "z"), d):
        # offsets carry 24 bits inside the cell
        assert np.max(np.abs(got[k] - ic[k])) <= dk * 2.0 ** -23, k
    for k in ("p", "q", "r", "b", "h", "vol", "stretch"):
        want = ic.get(k, np.zeros(len(ic["x"])))
        np.testing.assert_array_equal(got[k], want.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(got["tag"], ic["tag"])


def test_fp64_only_entry_points_refuse():
    from pmpic_amd.lib import PmpicError
    shape = (12, 10, 8)
    ic = thermal_case(*shape, 2)
    g = make(shape, 1, ic)
    with pytest.raises(PmpicError):
        g.par2grid(ic["b"])
    with pytest.raises(PmpicError):
        g.parcel_haloswap()


@pytest.mark.parametrize("shape,kw,steps", [
    ((12, 10, 8), dict(vort=1.0), 1),
    ((16, 16, 8), dict(vort=2.0, shrink=0.05, thin=0.1), 6),
    ((10, 12, 6), dict(vort=0.5, thin=0.2), 10),
    # the FP32 column solvers of 17, 33 and 65 z nodes (two lanes a column), 129, 201 and 257
    ((8, 6, 16), dict(vort=1.0), 2),
    ((6, 8, 32), dict(vort=1.0), 2),
    ((8, 8, 64), dict(vort=1.0), 2),
    ((8, 6, 128), dict(vort=1.0), 2),
    ((6, 8, 200), dict(vort=1.0), 2),
    ((8, 8, 256), dict(vort=1.0), 2),
])
def test_steps_close_to_fp64(shape, kw, steps):
    ic = thermal_case(*shape, 2, **kw)
    g64 = make(shape, 0, ic)
    g32 = make(shape, 1, ic)
    for step in range(steps):
        g64.step()
        g32.step()
        assert g32.n == g64.n, f"step {step}: parcel count {g32.n} vs {g64.n}"
        assert g32.last_counts() == g64.last_counts(), step
        assert abs(g32.time - g64.time) <= 1e-12
    for name in ("u", "v", "w"):
        assert rel_err(g32.grid(name), g64.grid(name)) < 1e-5, name
    # after a step the last writer of these is mixing: its kept grids (vol, b, hg, p, q, r)
    for name in ("p", "q", "r", "b", "vol", "hg"):
        assert rel_err(g32.grid(name), g64.grid(name)) < 1e-5, name


@pytest.mark.parametrize("capacity", ["16", "32"])
def test_27_parcels_per_cell_at_either_capacity(monkeypatch, capacity):
    """27 parcels per cell through the slot table at capacity 16 (every cell past it, so all in
    the dense overflow region) and at 32 (the capacity the mean picks), against the FP64 form."""
    monkeypatch.setenv("PMPIC_SLOT_CAPACITY", capacity)
    shape = (12, 10, 8)
    ic = thermal_case(*shape, 3, vort=1.0, thin=0.1)
    g64 = make(shape, 0, ic)
    g32 = make(shape, 1, ic)
    for step in range(2):
        g64.step()
        g32.step()
        assert g32.n == g64.n and g32.last_counts() == g64.last_counts(), step
    for name in ("u", "v", "w", "p", "q", "r", "b"):
        assert rel_err(g32.grid(name), g64.grid(name)) < 1e-5, name


@pytest.mark.parametrize("slots", ["auto", "16", "32", "dense"])
def test_crowded_cells_close_to_fp64(monkeypatch, slots):
    """Cells of 32, 33, 44 and 708 parcels through the FP32 binning: with the per-cell slot
    table at the capacity its mean picks (16 here, about 9 parcels per cell), fixed at 16 or
    32 (PMPIC_SLOT_CAPACITY; cells past the capacity in the dense overflow region, the
    708-parcel cell past the LDS rows of a workgroup), and without it (PMPIC_SLOT_CAPACITY=0,
    every cell dense, the layout of grids too large for the table).  Same bars as the steps
    above."""
    if slots != "auto":
        monkeypatch.setenv("PMPIC_SLOT_CAPACITY", "0" if slots == "dense" else slots)
    shape = (12, 10, 8)
    ic = crowded_case(*shape)
    g64 = make(shape, 0, ic)
    g32 = make(shape, 1, ic)
    for step in range(2):
        g64.step()
        g32.step()
        assert g32.n == g64.n and g32.last_counts() == g64.last_counts(), step
    for name in ("u", "v", "w", "p", "q", "r", "b"):
        assert rel_err(g32.grid(name), g64.grid(name)) < 1e-5, name


def test_volume_conserved():
    shape = (16, 16, 8)
    ic = thermal_case(*shape, 2, vort=3.0, shrink=0.05, thin=0.1, near_split=0.05)
    g = make(shape, 1, ic)
    s0 = g.conservation_sums()
    for _ in range(3):
        g.step()
    s1 = g.conservation_sums()
    # split halves volumes exactly; mixing moves volume between parcels in FP32
    assert abs(s1[0] - s0[0]) <= 1e-6 * s0[0]


def test_ensemble_concurrent_matches_sequential():
    """Members of different grids and both forms stepped together (pmpic_step_concurrent) end
    in exactly the state of the same members stepped one after another."""
    from pmpic_amd.lib import step_concurrent
    members = [((12, 10, 8), 0, 1), ((16, 16, 8), 1, 2), ((10, 12, 6), 0, 3), ((8, 8, 8), 1, 4)]
    seq, con = [], []
    for shape, prec, seed in members:
        ic = thermal_case(*shape, 2, seed=seed, vort=1.5, shrink=0.02, thin=0.05)
        seq.append(make(shape, prec, ic))
        con.append(make(shape, prec, ic))
    for g in seq:
        g.step(3)
    step_concurrent(con, 3)
    for a, b in zip(seq, con):
        assert a.n == b.n
        pa, pb = a.get_parcels(), b.get_parcels()
        for k in ("x", "y", "z", "p", "q", "r", "vol", "b", "h", "stretch"):
            np.testing.assert_array_equal(pa[k], pb[k])


@pytest.mark.parametrize("precision", [0, 1])
def test_timers_kernels_only_and_all(precision):
    """Timing mode 2 keeps only the two hot kernels' timers (taken by their dispatches); mode 1
    adds the component timers.  Four accumulations and four RK stage launches per step."""
    shape = (12, 10, 8)
    ic = thermal_case(*shape, 2, vort=1.0)
    g = make(shape, precision, ic)
    g.timing_enable(True, kernels_only=True)
    g.step(2)
    t = g.timer_table()
    assert t["par2grid_cells"][1] == 8 and t["grid2par_rk_stage"][1] == 8
    assert t["par2grid_cells"][0] > 0 and t["grid2par_rk_stage"][0] > 0
    assert t["vort2vel"][1] == 0 and t["RK4_all_steps"][1] == 0
    g.timing_enable(True)
    g.step(1)
    t = g.timer_table()
    assert t["par2grid_cells"][1] == 4 and t["vort2vel"][1] == 4 and t["RK4_all_steps"][1] == 4
    assert t["vort2vel"][0] > 0


@pytest.mark.parametrize("case", ["8ppc", "27ppc", "crowded300"])
def test_slot_capacity_changes_no_result(monkeypatch, case):
    """8 and 27 parcels per cell, and a crowded case with cells of 32, 33 and 300 parcels (300 is
    past the 256 LDS rows of the 16-slot accumulation, so that capacity sums the cell on its
    serial big-cell path while 32 slots sum it in LDS): two steps with the slot capacity fixed
    at 16, fixed at 32 and tuned (PMPIC_SLOT_CAPACITY=tune: the first par2grid times the
    accumulation at both and keeps the faster).  Every cell's parcels are summed in ascending
    parcel order with the same expressions whatever the capacity and path, so counts, parcels
    and grids are bit-identical across the three."""
    shape = (12, 10, 8)
    if case == "crowded300":
        ic = crowded_case(*shape, extra=(24, 25, 292))
    else:
        ic = thermal_case(*shape, 2 if case == "8ppc" else 3, vort=1.0, thin=0.1)
    names = ("u", "v", "w", "p", "q", "r", "b")
    out = {}
    for mode in ("16", "32", "tune"):
        monkeypatch.setenv("PMPIC_SLOT_CAPACITY", mode)
        g = make(shape, 1, ic)
        for _ in range(2):
            g.step()
        out[mode] = (g.n, g.last_counts(), g.get_parcels(), {k: g.grid(k) for k in names})
        g.close()
    n16, c16, p16, g16 = out["16"]
    for mode in ("32", "tune"):
        n, cnt, p, gr = out[mode]
        assert n == n16 and cnt == c16, mode
        for k in p16:
            assert np.array_equal(p[k], p16[k]), (mode, k)
        for k in names:
            assert np.array_equal(gr[k], g16[k]), (mode, k)


@pytest.mark.parametrize("case", ["8ppc", "27ppc", "crowded", "dense"])
def test_fused_binning_changes_no_result(monkeypatch, case):
    """RK stages 1-3 binning the cells they write (PMPIC_FUSED_BINNING=1, the default on grids
    of 2^21 cells or more) against the separate binning pass (=0): the same cell lists, summed
    in ascending parcel order, so three steps of splitting and mixing end bit-identical in
    counts, parcels and grids; at 8 and 27 parcels per cell, with crowded cells (the overflow
    list) and without the slot table (the dense ranks)."""
    shape = (16, 12, 8)
    if case == "crowded":
        ic = crowded_case(*shape)
    else:
        ic = thermal_case(*shape, 3 if case == "27ppc" else 2, vort=2.0, shrink=0.03, thin=0.05,
                          near_split=0.03)
    if case == "dense":
        monkeypatch.setenv("PMPIC_SLOT_CAPACITY", "0")
    names = ("u", "v", "w", "p", "q", "r", "b", "vol", "hg")
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("PMPIC_FUSED_BINNING", mode)
        g = make(shape, 1, ic)
        counts = []
        for _ in range(3):
            g.step()
            counts.append((g.n, g.last_counts()))
        out[mode] = (counts, g.get_parcels(), {k: g.grid(k) for k in names})
        g.close()
    c0, p0, g0 = out["0"]
    c1, p1, g1 = out["1"]
    assert c1 == c0
    for k in p0:
        np.testing.assert_array_equal(p1[k], p0[k], err_msg=k)
    for k in names:
        np.testing.assert_array_equal(g1[k], g0[k], err_msg=k)


@pytest.mark.parametrize("case,shape", [("8ppc", (32, 16, 8)), ("27ppc", (16, 12, 8)),
                                        ("crowded", (32, 12, 8)), ("dense", (16, 12, 8))])
def test_mixing_sums_reused_by_next_par2grid_change_no_result(monkeypatch, case, shape):
    """The next step's first par2grid taking the mixing's kept sums on every tile mixing left
    unchanged and summing only the changed ones (the default on x-row grids) against summing
    every tile (PMPIC_P2G_REUSE=0): four steps of splitting and mixing, removed, new, absorbing
    and moved parcels, end bit-identical in counts, parcels and grids after every step; several
    tiles per x row (the periodic boundary between the last tile and the first), 8 and 27
    parcels per cell, crowded cells and the dense ranks."""
    if case == "crowded":
        ic = crowded_case(*shape)
    else:
        ic = thermal_case(*shape, 3 if case == "27ppc" else 2, vort=3.0, shrink=0.05, thin=0.1,
                          near_split=0.05)
    if case == "dense":
        monkeypatch.setenv("PMPIC_SLOT_CAPACITY", "0")
    names = ("u", "v", "w", "p", "q", "r", "b", "vol", "hg")
    ctx = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("PMPIC_P2G_REUSE", mode)
        ctx[mode] = make(shape, 1, ic)
    removed = added = 0
    for s in range(4):
        for mode in ("0", "1"):
            ctx[mode].step()
        a, b = ctx["0"], ctx["1"]
        assert (a.n, a.last_counts()) == (b.n, b.last_counts()), s
        removed += a.last_counts()[1]
        added += a.last_counts()[2]
        pa, pb = a.get_parcels(), b.get_parcels()
        for k in pa:
            np.testing.assert_array_equal(pb[k], pa[k], err_msg=f"step {s}: {k}")
        for k in names:
            np.testing.assert_array_equal(b.grid(k), a.grid(k), err_msg=f"step {s}: {k}")
    assert removed > 0 or case == "crowded"


def test_writebinary_grids_close_to_fp64():
    """writebinary's grids (p, q, r, b + 12.5 hgliq, hg, hgliq) from the FP32 parcels after two
    steps, within 1e-5 of the FP64 form's (which test_gpu_parity.py holds to the oracle)."""
    shape = (16, 12, 8)
    ic = thermal_case(*shape, 2, vort=1.0)
    g64 = make(shape, 0, ic)
    g32 = make(shape, 1, ic)
    for g in (g64, g32):
        g.step(2)
        g.writebinary_grids()
    for name in ("p", "q", "r", "b", "hg", "hgliq"):
        assert rel_err(g32.grid(name), g64.grid(name)) < 1e-5, name
