"""Host-side formats and set-up: the grid dump's byte layout, the dump file names, the multi-file
restart reader (range filtering, counts, failures), the plume initial condition and the timer
summary table.  CPU only."""
import math
import os

import numpy as np
import pytest

from pmpic_amd import dumpio, plume
from pmpic_amd.control import timer_summary


def _grid_set(nx, ny, nz, seed=0):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal((nx + 4, ny + 4, nz + 1)) for k in dumpio.GRID_FIELDS}


def test_file_names_follow_the_fortran_formats():
    assert dumpio.parcel_file(3, 50) == "parcels_00003_00050.dat"
    assert dumpio.grid_file(0, 7) == "grids_00000_00007.dat"
    assert len(dumpio.parcel_file(0, 0)) == 23  # character (len=23) :: filename
    assert len(dumpio.grid_file(0, 0)) == 21


def test_grid_dump_byte_layout(tmp_path):
    nx, ny, nz = 6, 4, 5
    grids = _grid_set(nx, ny, nz)
    x, y, z = dumpio.coordinates(nx, ny, nz, 0.0, 6.0, 0.0, 2.0, 0.0, 1.0)
    bounds = dumpio.grid_bounds(x, y, z, nx, ny, nz)
    assert bounds == [0.0, 5.0, 0.0, 1.5, 0.0, 1.0]  # x_coords(start), x_coords(end), ...
    path = tmp_path / dumpio.grid_file(0, 1)
    dumpio.write_grids(path, 1.25, bounds, grids, nx, ny, nz)
    raw = path.read_bytes()
    n = nx * ny * (nz + 1)
    assert len(raw) == 8 + 6 * 8 + 3 * 4 + 10 * n * 8
    assert np.frombuffer(raw[56:68], dtype="<i4").tolist() == [nx, ny, nz + 1]
    # the first field's first values: z fastest at x = y = first interior node
    first = np.frombuffer(raw[68:68 + 8 * (nz + 1)], dtype="<f8")
    np.testing.assert_array_equal(first, grids["u"][2, 2, :])
    t, b, dims, back = dumpio.read_grids(path)
    assert t == 1.25 and dims == (nx, ny, nz + 1) and list(b) == bounds
    for k in dumpio.GRID_FIELDS:
        np.testing.assert_array_equal(back[k], grids[k][2:-2, 2:-2, :])


def test_parcel_dump_bounds():
    nx, ny, nz = 8, 6, 4
    x, y, z = dumpio.coordinates(nx, ny, nz, 0.0, 8.0, 0.0, 3.0, 0.0, 2.0)
    # x_coords(start), x_coords(end + 1): the whole periodic extent
    assert dumpio.parcel_bounds(x, y, z, nx, ny, nz) == [0.0, 8.0, 0.0, 3.0, 0.0, 2.0]


def _random_parcels(n, lx, ly, lz, seed):
    rng = np.random.default_rng(seed)
    f = {k: rng.standard_normal(n) for k in dumpio.FIELDS}
    f["x"] = rng.uniform(0, lx, n)
    f["y"] = rng.uniform(0, ly, n)
    f["z"] = rng.uniform(0, lz, n)
    f["tag"] = np.arange(n, dtype=np.float64)
    return f


def _write_ranks(directory, num, parcels, lx, ly, lz, nrank):
    """Split a parcel set into nrank x-slabs, each its own dump file (slot order kept)."""
    edges = [lx * r / nrank for r in range(nrank + 1)]
    for r in range(nrank):
        m = (parcels["x"] >= edges[r]) & (parcels["x"] < edges[r + 1])
        dumpio.write_parcels(os.path.join(directory, dumpio.parcel_file(r, num)), 0.5 + r,
                             [edges[r], edges[r + 1], 0.0, ly, 0.0, lz],
                             {k: v[m] for k, v in parcels.items()})
    return edges


def test_restart_reads_every_rank_file_in_order(tmp_path):
    p = _random_parcels(5000, 4.0, 2.0, 1.0, 1)
    _write_ranks(tmp_path, 12, p, 4.0, 2.0, 1.0, 3)
    log = []
    t, f = dumpio.read_restart(tmp_path, 12, 0.0, 4.0, 0.0, 2.0, log=log.append)
    assert t == 2.5  # the time of the last file read
    order = np.argsort(np.floor(p["x"] / (4.0 / 3)), kind="stable")
    for k in dumpio.FIELDS:
        np.testing.assert_array_equal(f[k], p[k][order])
    assert "All files read in successfully" in " ".join(log)


def test_restart_failures(tmp_path):
    with pytest.raises(dumpio.DumpError, match="no file"):
        dumpio.read_restart(tmp_path, 3, 0.0, 4.0, 0.0, 2.0, log=lambda s: None)
    p = _random_parcels(1000, 4.0, 2.0, 1.0, 2)
    _write_ranks(tmp_path, 3, p, 4.0, 2.0, 1.0, 2)
    # a domain covering half the box keeps half the parcels: the count check stops the run
    with pytest.raises(dumpio.DumpError, match="incorrect parcel count"):
        dumpio.read_restart(tmp_path, 3, 0.0, 2.0, 0.0, 2.0, log=lambda s: None)
    # a parcel exactly on the upper x bound is outside [xmin, xmax)
    q = _random_parcels(10, 4.0, 2.0, 1.0, 3)
    q["x"][4] = 4.0
    dumpio.write_parcels(tmp_path / dumpio.parcel_file(0, 9), 0.0, [0, 4.0, 0, 2.0, 0, 1.0], q)
    with pytest.raises(dumpio.DumpError):
        dumpio.read_restart(tmp_path, 9, 0.0, 4.0, 0.0, 2.0, log=lambda s: None)


# ---------------------------------------------------------------- plume ----

REF_PLUME = dict(lam=1.0, H=0.80000001, z_c=2.5, mu=0.89999998, z_d=4.0, z_m=5.0,
                 r_plume=0.80000001, e_values=[0.30000001, -0.40000001, 0.5])  # config.mcf


def test_plume_parameter_checks():
    ok = dict(REF_PLUME)
    plume.plume_parameters(**ok, log=lambda s: None)
    for bad, msg in ((dict(H=1.2), "less than 1"), (dict(mu=0.7), "between H and 1"),
                     (dict(mu=1.1), "between H and 1"), (dict(r_plume=1.5), "too big")):
        with pytest.raises(plume.PlumeError, match=msg):
            plume.plume_parameters(**dict(ok, **bad), log=lambda s: None)


def test_plume_derived_quantities():
    P = plume.plume_parameters(**REF_PLUME, log=lambda s: None)
    h_pl = math.exp(-2.5)
    assert P["h_pl"] == h_pl
    assert P["h_bg"] == 0.89999998 * h_pl
    assert P["z_b"] == math.log(0.80000001 / P["h_bg"])
    # glat = g L h0 / (cp theta_l0) with the FP32 literal 0.01: not exactly 12.5
    assert plume.GLAT != 12.5 and abs(plume.GLAT - 12.5) < 1e-6
    assert P["dbdz"] == plume.GLAT * (h_pl - math.exp(-5.0)) / (5.0 - 4.0)
    assert P["b_pl"] == P["dbdz"] * (4.0 - P["z_b"])


def test_plume_parcels_geometry_and_fields():
    nx, ny, nz = 16, 12, 16
    lx, ly, lz = 6.28000021, 6.28000021, 6.28000021
    P = plume.plume_parameters(**REF_PLUME, log=lambda s: None)
    f, n_pl = plume.plume_parcels(nx, ny, nz, 0.0, lx, 0.0, ly, 0.0, lz, P)
    n = len(f["x"])
    r = P["r_plume"]
    xc, yc, zc = lx / 2, ly / 2, r
    d2 = (f["x"] - xc) ** 2 + (f["y"] - yc) ** 2 + (f["z"] - zc) ** 2
    assert np.all(d2[:n_pl] <= r * r * (1 + 1e-12)) and np.all(d2[n_pl:] > r * r * (1 - 1e-12))
    # counts from the integer lattices: the plume lattice's points inside the sphere and the
    # background lattice's points outside it
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    i, j, k = np.meshgrid(np.arange(4 * nx), np.arange(4 * ny),
                          np.arange(math.ceil(2 * r / (dz / 4))), indexing="ij")
    px, py, pz = (i + 0.5) * dx / 4 - xc, (j + 0.5) * dy / 4 - yc, (k + 0.5) * dz / 4 - zc
    assert n_pl == int(np.count_nonzero(px ** 2 + py ** 2 + pz ** 2 <= r * r))
    i, j, k = np.meshgrid(np.arange(2 * nx), np.arange(2 * ny), np.arange(2 * nz), indexing="ij")
    bx, by, bz = (i + 0.5) * dx / 2 - xc, (j + 0.5) * dy / 2 - yc, (k + 0.5) * dz / 2 - zc
    assert n - n_pl == int(np.count_nonzero(bx ** 2 + by ** 2 + bz ** 2 > r * r))
    # inside: h = h_pl and b = b_pl (1 + e1 xp yp + e2 xp zp + e3 yp zp): its mean over the
    # mirror-symmetric lattice is b_pl
    assert np.all(f["h"][:n_pl] == P["h_pl"])
    assert abs(np.mean(f["b"][:n_pl]) - P["b_pl"]) <= 1e-9 * abs(P["b_pl"])
    assert np.ptp(f["b"][:n_pl]) > 0
    # outside: mixed layer below z_b, stratified above
    zo, bo, ho = f["z"][n_pl:], f["b"][n_pl:], f["h"][n_pl:]
    low = zo < P["z_b"]
    assert np.all(bo[low] == 0) and np.all(ho[low] == P["h_bg"])
    np.testing.assert_allclose(bo[~low], P["dbdz"] * (zo[~low] - P["z_b"]), rtol=1e-14)
    np.testing.assert_allclose(ho[~low], 0.80000001 * np.exp(-zo[~low]), rtol=1e-14)
    # the two lattices tile the box up to the sphere's staircase
    assert abs(np.sum(f["vol"]) - lx * ly * lz) < 0.02 * lx * ly * lz
    for key in ("p", "q", "r", "stretch", "tag"):
        assert np.all(f[key] == 0)


def test_plume_from_config_options(tmp_path):
    """The plume keys as the reference's config.mcf writes them (e_values a real array)."""
    from pmpic_amd.config import parse_mcf
    mcf = tmp_path / "plume.mcf"
    mcf.write_text("lambda=1\nH=0.8\nz_c=2.5\nmu=0.9\nz_d=4.\nz_m=5.\nr_plume=0.8\n"
                   "e_values=0.3, -0.4 ,0.5\nmax_parcels=1000000\n")
    o = parse_mcf(str(mcf))
    f, n_pl = plume.plume_from_options(o, 8, 8, 8, 0.0, 6.28000021, 0.0, 6.28000021, 0.0,
                                       6.28000021, log=lambda s: None)
    P = plume.plume_parameters(**REF_PLUME, log=lambda s: None)
    g, m_pl = plume.plume_parcels(8, 8, 8, 0.0, 6.28000021, 0.0, 6.28000021, 0.0, 6.28000021, P)
    assert n_pl == m_pl > 0
    for k in f:
        np.testing.assert_array_equal(f[k], g[k])
    o["max_parcels"] = 100
    with pytest.raises(plume.PlumeError, match="Maxparcels"):
        plume.plume_from_options(o, 8, 8, 8, 0.0, 6.28000021, 0.0, 6.28000021, 0.0, 6.28000021,
                                 log=lambda s: None)


# ---------------------------------------------------------------- timers ----

def test_timer_summary_format():
    lines = []
    rows = [("MONC", 1, 10.0, 10.0, 100.0), ("vort2vel", 4, 2.0, 0.8, 1.1),
            ("mixing", 1, 0.5, 0.5, 0.25), ("splitting", 0, 0.0, 0.0, 0.0)]
    timer_summary(rows, log=lines.append)
    body = [s for s in lines if s.startswith("|") and "Name" not in s and "Function" not in s
            and "----" not in s]
    assert [b.split("|")[1].strip() for b in body] == ["MONC", "vort2vel", "mixing", "splitting"]
    # '|',a20,'|',i8,'|',f9.3,'s','| ',f6.2,'%','|',f8.3,'s|',f8.3,'s|',f7.2,'%|'
    # mean 0.5 s, stddev sqrt(1.1 / 4 - 0.25) = 0.158 s: 31.62 %
    assert body[1] == ("|vort2vel            |       4|    2.000s|  20.00%|   0.500s|   0.800s|"
                       "  31.62%|")
    assert all(len(b) == 80 for b in body)
