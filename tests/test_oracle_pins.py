"""Independent pins of the oracle functions whose composition the reference defines: each check
comes from physics, a closed form or an independently assembled system, chosen so that a
plausible slip in the restatement (a dropped term, a wrong sign, a swapped derivative, a wrong
coefficient or threshold, an exact literal where the reference has an FP32 one) fails it.

* vorticity_tendency (vorticity_tendency.F90:121-265): the buoyancy torque of a buoyancy field
  alone (DP = db/dy, DQ = -db/dx, DR = 0) against numpy.fft derivatives; the total buoyancy
  grid against a brute-force numpy scatter of b + 12.5 max(0, h - exp(-z)); the stretching and
  tilting terms against omega . grad(u) + (b_y, -b_x, 0) with grad(u) taken by numpy (spectral
  in x, y; fourth-order differences in z), which does not use the reference's substitutions
  u_z = eta + w_x, v_z = w_y - xi, w_z = -u_x - v_y;
* rk4_integrator (rk4_integrator.F90:108-309): tendencies lambda * Y at every stage give the
  fourth-order Taylor polynomial of exp(lambda dt) exactly;
* vort2vel quirks: z-level 1 of the corrected z-vorticity zeroed for every mode (vort2vel.F90:
  203-205), u_s / v_s / w_s the filtered spectra of the unfiltered U, V, W (:271-320), P, Q, R
  filtered (:208-226);
* the filter threshold (fft_ops.F90:115-134) against modes just inside and outside
  (2/9)((nx/2/Lx)^2 + (ny/2/Ly)^2);
* parcel_mixing counts (parcel_mixing.F90:155-212) against a numpy count of sparse cells, and
  the new parcels at those cells' centres with volume 2 vmin;
* the diffz / laplinv column systems (fft_ops.F90:419-601) assembled in numpy with float32
  literals and solved by scipy's banded solver;
* the split geometry with the fl32(1/3) exponent (parcel_splitting.F90:64-131) to 1e-14;
* writebinary's hgliq = max(0, hg - exp(-z)) and b + 12.5 hgliq (writebinary.F90:224-233).
"""
import numpy as np
import pytest
import scipy.linalg

from tests.cases import LX, LZ, box, rel_err, thermal_case

F32_THIRD = float(np.float32(1.0) / np.float32(3.0))


def make(oracle_mod, nx, ny, nz, cap=400000, **kw):
    b = box(nx, ny, nz)
    return oracle_mod.Oracle(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                             b["zmax"], cap, kw.get("dtm", 0.01), kw.get("dtmax", 0.01), 0.5)


def inner(a):
    """interior nodes of a reference-layout grid, as [x, y, z]"""
    return a[2:-2, 2:-2, :]


def ddx_fft(f, d):
    k = 2 * np.pi * np.fft.fftfreq(f.shape[0], d)
    return np.real(np.fft.ifft(1j * k[:, None, None] * np.fft.fft(f, axis=0), axis=0))


def ddy_fft(f, d):
    k = 2 * np.pi * np.fft.fftfreq(f.shape[1], d)
    return np.real(np.fft.ifft(1j * k[None, :, None] * np.fft.fft(f, axis=1), axis=1))


def ddz_fd4(f, d):
    """fourth-order central differences in z (levels 2 .. nz-2; the ends are left at 0)"""
    out = np.zeros_like(f)
    out[:, :, 2:-2] = (f[:, :, :-4] - 8 * f[:, :, 1:-3] + 8 * f[:, :, 3:-1] - f[:, :, 4:]) / (12 * d)
    return out


def brute_par2grid(X, Y, Z, vol, var, nx, ny, nz, dx, dy, dz):
    """trilinear volume-weighted scatter, periodic in x/y, normalised (numpy.add.at)"""
    fx, fy, fz = X / dx, Y / dy, Z / dz
    i, j, k = (np.floor(a).astype(int) for a in (fx, fy, fz))
    ax, ay, az = fx - i, fy - j, fz - k
    W = np.zeros((nx, ny, nz + 1))
    D = np.zeros((nx, ny, nz + 1))
    for bx in (0, 1):
        for by in (0, 1):
            for bz in (0, 1):
                w = (ax if bx else 1 - ax) * (ay if by else 1 - ay) * (az if bz else 1 - az) * vol
                ii, jj, kk = (i + bx) % nx, (j + by) % ny, k + bz
                np.add.at(W, (ii, jj, kk), w)
                np.add.at(D, (ii, jj, kk), w * var)
    W[W == 0] = 1
    return D / W


def numpy_filter_mask(nx, ny, lx, ly):
    """the keep (1) / drop (0) mask over the semi-spectral (ii, jj) layout: wavenumbers
    floor(ii/2)/L, dropped where kx^2 + ky^2 > (2/9)((nx/2/Lx)^2 + (ny/2/Ly)^2)"""
    sx, sy = (nx // 2 + 1) * 2, (ny // 2 + 1) * 2
    kx = (np.arange(sx) // 2) / lx
    ky = (np.arange(sy) // 2) / ly
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    thr = 2.0 / 9.0 * ((nx / 2 / lx) ** 2 + (ny / 2 / ly) ** 2)
    return (k2 <= thr).astype(float), k2, thr


# ------------------------------------------------------ vorticity_tendency ----

def test_buoyancy_torque_alone(oracle_mod):
    """Zero vorticity and h = 0: every velocity and stretching term vanishes, so DP = dB/dy,
    DQ = -dB/dx and DR = 0 for the gridded buoyancy B (vorticity_tendency.F90:155-265)."""
    nx, ny, nz = 16, 12, 8
    o = make(oracle_mod, nx, ny, nz)
    g = o.geometry()
    ic = thermal_case(nx, ny, nz, 2, seed=21)
    n = len(ic["x"])
    Lx, Ly = nx * g["dx"], ny * g["dy"]
    b = 0.7 * np.sin(2 * np.pi * ic["y"] / Ly) + 0.4 * np.cos(4 * np.pi * ic["x"] / Lx + 0.3) \
        + 0.1 * ic["z"]
    o.set_parcels({"x": ic["x"], "y": ic["y"], "z": ic["z"], "vol": ic["vol"], "b": b,
                   "h": np.zeros(n)})
    o.vort2vel()
    o.vorticity_tendency()
    B = inner(o.grid("b"))
    DP, DQ, DR = (inner(o.grid(k)) for k in ("dp", "dq", "dr"))
    by, bx = ddy_fft(B, g["dy"]), ddx_fft(B, g["dx"])
    assert rel_err(DP, by) < 1e-12
    assert rel_err(DQ, -bx) < 1e-12
    assert np.max(np.abs(DR)) == 0.0
    for k in ("u", "v", "w"):
        assert np.max(np.abs(o.grid(k))) == 0.0


def test_total_buoyancy_grid(oracle_mod):
    """B = par2grid(b + 12.5 max(0, h - exp(-z))) (vorticity_tendency.F90:121-129) against a
    brute-force numpy scatter, with about half the parcels above saturation."""
    nx, ny, nz = 12, 10, 8
    o = make(oracle_mod, nx, ny, nz)
    g = o.geometry()
    ic = thermal_case(nx, ny, nz, 2, seed=22)
    rng = np.random.default_rng(22)
    n = len(ic["x"])
    h = np.exp(-ic["z"]) * rng.uniform(0.5, 1.5, n)
    b = rng.uniform(-0.2, 0.2, n)
    o.set_parcels({"x": ic["x"], "y": ic["y"], "z": ic["z"], "vol": ic["vol"], "b": b, "h": h})
    o.vort2vel()
    o.vorticity_tendency()
    btot = b + 12.5 * np.maximum(0.0, h - np.exp(-ic["z"]))
    assert np.mean(h > np.exp(-ic["z"])) > 0.3
    want = brute_par2grid(ic["x"], ic["y"], ic["z"], ic["vol"], btot, nx, ny, nz, g["dx"],
                          g["dy"], g["dz"])
    assert rel_err(inner(o.grid("b")), want) < 1e-12


def test_stretching_and_tilting(oracle_mod):
    """DP, DQ, DR = omega . grad(u) + (dB/dy, -dB/dx, 0), the Boussinesq vorticity tendency,
    with grad(u) from the filtered velocity spectra by numpy (spectral x, y derivatives;
    fourth-order z differences), away from the lids.  The bar, 1.2 % of the largest tendency,
    is set by how closely the compact-scheme velocity's curl reproduces the parcel-sampled
    vorticity grids (about 0.9 % at grid-scale sampling noise, independent of nz); each term
    checked below is several times larger than the bar, so dropping it or flipping its sign
    fails."""
    nx, ny, nz = 16, 16, 32
    o = make(oracle_mod, nx, ny, nz, cap=1200000)
    g = o.geometry()
    ic = thermal_case(nx, ny, nz, 2, seed=23)
    n = len(ic["x"])
    Lx, Ly, Lz = nx * g["dx"], ny * g["dy"], nz * g["dz"]
    x, y, z = ic["x"], ic["y"], ic["z"]
    sx, sy, sz = 2 * np.pi * x / Lx, 2 * np.pi * y / Ly, np.pi * z / Lz
    fields = {"x": x, "y": y, "z": z, "vol": ic["vol"], "h": np.zeros(n),
              "p": 0.8 * np.sin(sy) * np.sin(sz) + 0.3 * np.cos(sx),
              "q": 0.6 * np.cos(sx) * np.sin(sz) - 0.2 * np.sin(sy + 0.4),
              "r": 0.5 * np.cos(sx + sy) * np.cos(sz) + 0.25,
              "b": 0.3 * np.sin(sx) * np.cos(sy) + 0.1 * z}
    o.set_parcels(fields)
    o.vort2vel()
    o.vorticity_tendency()
    P, Q, R, B = (inner(o.grid(k)) for k in ("p", "q", "r", "b"))
    DP, DQ, DR = (inner(o.grid(k)) for k in ("dp", "dq", "dr"))

    # the velocity grids the tendency differentiates: the inverse transforms of u_s, v_s, w_s,
    # taken by numpy from the oracle's spectral layout
    def phys(s):
        S = np.transpose(s, (2, 1, 0))            # [k, jj, ii]
        Xc = S[:, :, 0::2] + 1j * S[:, :, 1::2]   # x-transform of each real y-row
        rows = np.fft.irfft(Xc, n=nx, axis=2)     # [k, jj, x]
        Yc = rows[:, 0::2, :] + 1j * rows[:, 1::2, :]
        f = np.fft.irfft(Yc, n=ny, axis=1)        # [k, y, x]
        return np.transpose(f, (2, 1, 0))

    U, V, W = (phys(o.spec(k)) for k in ("u_s", "v_s", "w_s"))
    dx, dy, dz = g["dx"], g["dy"], g["dz"]
    ux, uy, uz = ddx_fft(U, dx), ddy_fft(U, dy), ddz_fd4(U, dz)
    vx, vy, vz = ddx_fft(V, dx), ddy_fft(V, dy), ddz_fd4(V, dz)
    wx, wy, wz = ddx_fft(W, dx), ddy_fft(W, dy), ddz_fd4(W, dz)
    want_p = P * ux + Q * uy + R * uz + ddy_fft(B, dy)
    want_q = P * vx + Q * vy + R * vz - ddx_fft(B, dx)
    want_r = P * wx + Q * wy + R * wz
    sl = (slice(None), slice(None), slice(8, -8))
    terms = {"dp": (P * ux, Q * uy, R * uz, R * Q, ddy_fft(B, dy)),
             "dq": (P * vx, Q * vy, R * vz, R * P, ddx_fft(B, dx)),
             "dr": (P * wx, Q * wy, R * wz, R * ux, R * vy)}
    for name, got, want in (("dp", DP, want_p), ("dq", DQ, want_q), ("dr", DR, want_r)):
        scale = np.max(np.abs(want[sl]))
        assert scale > 0.05
        assert np.max(np.abs(got[sl] - want[sl])) < 1.2e-2 * scale, name
        for t in terms[name]:
            assert np.max(np.abs(t[sl])) > 0.05 * scale, name


# ------------------------------------------------------------------ RK4 ----

def test_rk4_coefficients_exponential(oracle_mod):
    """Tendencies lambda_v * Y_v evaluated at every stage's Y: the four stages must give
    Y = Yo * sum_{m<=4} (lambda dt)^m / m! (the classical RK4 weights 1/6, 1/3, 1/3, 1/6 and
    the half / half / full stage points), to 1e-14.  Parcels stay inside the box."""
    o = make(oracle_mod, 8, 8, 8, dtm=0.01)
    n = 50
    rng = np.random.default_rng(31)
    Y0 = {"x": rng.uniform(2.0, 4.0, n), "y": rng.uniform(2.0, 4.0, n),
          "z": rng.uniform(1.0, 2.0, n), "p": rng.uniform(-1, 1, n), "q": rng.uniform(-1, 1, n),
          "r": rng.uniform(-1, 1, n)}
    lam = {"x": 1.3, "y": -0.7, "z": 0.4, "p": 2.1, "q": -1.9, "r": 0.9}
    tend = {"x": "dxdt", "y": "dydt", "z": "dzdt", "p": "dpdt", "q": "dqdt", "r": "drdt"}
    o.set_parcels(dict(Y0, vol=np.ones(n)))
    dt = o.dtm
    for _ in range(4):
        for k in Y0:
            o.field(tend[k])[:n] = lam[k] * o.field(k)[:n]
        o.rk4_integrator()
    for k in Y0:
        a = lam[k] * dt
        want = Y0[k] * (1 + a + a ** 2 / 2 + a ** 3 / 6 + a ** 4 / 24)
        np.testing.assert_allclose(o.field(k)[:n], want, rtol=1e-14, atol=0)


# ------------------------------------------------------------ vort2vel ----

def test_vort2vel_filter_and_level1_quirks(oracle_mod):
    """vort2vel.F90:203-205 zeroes z-level 1 of the corrected z-vorticity for every mode, so the
    gridded R is exactly 0 there although r is not; u_s, v_s, w_s are the filtered spectra of
    U, V, W (the numpy transform of the exported grid, masked), while U, V, W themselves keep
    their filtered-out modes and the gridded P, Q, R have none (:208-226, 262-320)."""
    nx, ny, nz = 16, 12, 8
    o = make(oracle_mod, nx, ny, nz)
    g = o.geometry()
    ic = thermal_case(nx, ny, nz, 2, seed=41, vort=1.0)  # random vorticity: every mode present
    o.set_parcels(ic)
    o.vort2vel()
    R = inner(o.grid("r"))
    assert np.max(np.abs(R[:, :, 0])) == 0.0
    assert np.max(np.abs(R[:, :, 1])) > 0.1
    mask, _, _ = numpy_filter_mask(nx, ny, nx * g["dx"], ny * g["dy"])
    drop = (1.0 - mask)[:, :, None]
    for grid, spec in (("u", "u_s"), ("v", "v_s"), ("w", "w_s")):
        S = o.fft_forward(o.grid(grid))
        s = o.spec(spec)
        assert rel_err(s, S * mask[:, :, None]) < 1e-12, spec
        assert np.max(np.abs(S * drop)) > 1e-3 * np.max(np.abs(S)), grid  # U unfiltered
    for grid in ("p", "q", "r"):
        S = o.fft_forward(o.grid(grid))
        assert np.max(np.abs(S * drop)) < 1e-12 * np.max(np.abs(S)), grid


@pytest.mark.parametrize("nx,ny", [(30, 20), (20, 30), (24, 24)])
def test_filter_threshold_both_sides(oracle_mod, nx, ny):
    """fft_ops.F90:115-134: drop where (kx^2 + ky^2)/(2pi)^2 > (2/9)((nx/2/Lx)^2 +
    (ny/2/Ly)^2).  With Lx = Ly these grids put modes just inside and just outside the
    threshold (for 30 x 20: l^2 + m^2 = 72 kept, 73 dropped against 72.2), so a different
    fraction or a missing factor of the threshold fails."""
    o = make(oracle_mod, nx, ny, 4)
    k = o.wavenumbers()  # [which, ii, jj]
    g = o.geometry()
    mask, k2, thr = numpy_filter_mask(nx, ny, nx * g["dx"], ny * g["dy"])
    near = np.abs(k2 - thr) < 0.02 * thr
    assert np.any(near & (mask == 1)) and np.any(near & (mask == 0))
    tie = np.abs(k2 - thr) < 1e-9 * thr  # exactly on the threshold: rounding decides
    np.testing.assert_array_equal(k[3][~tie], mask[~tie])
    np.testing.assert_allclose(k[2], (2 * np.pi) ** 2 * k2, rtol=1e-14)


# -------------------------------------------------------------- mixing ----

def test_mixing_counts_and_new_parcels(oracle_mod):
    """parcel_mixing.F90:155-212: removed = #(vol < vmin); added = the number of interior cells
    holding fewer than 3 kept parcels (a numpy count over the parcels' cells); each new parcel
    sits at the centre of one such cell with volume 2 vmin."""
    nx, ny, nz = 12, 10, 8
    ic = thermal_case(nx, ny, nz, 2, seed=51, vort=1.0, shrink=0.06, thin=0.15)
    o = make(oracle_mod, nx, ny, nz, cap=3 * len(ic["x"]))
    o.set_parcels(ic)
    g = o.geometry()
    dx, dy, dz = g["dx"], g["dy"], g["dz"]
    vmin = dx * dy * dz / 6.0 / 6.0 / 6.0
    kept = ic["vol"] >= vmin
    ci = np.floor(ic["x"] / dx).astype(int)
    cj = np.floor(ic["y"] / dy).astype(int)
    ck = np.floor(ic["z"] / dz).astype(int)
    counts = np.zeros((nx, ny, nz), dtype=int)
    np.add.at(counts, (ci[kept], cj[kept], ck[kept]), 1)
    sparse = np.argwhere(counts < 3)
    for _ in range(4):
        o.parcel_mixing()
    removed, added = o.last_counts()[1:]
    assert removed == int(np.sum(~kept)) and removed > 0
    assert added == len(sparse) and added > 0
    c = o.get_parcels()
    new = c["vol"] == 2 * vmin
    assert int(np.sum(new)) == added
    centres = np.stack([(sparse[:, 0] + 0.5) * dx, (sparse[:, 1] + 0.5) * dy,
                        (sparse[:, 2] + 0.5) * dz], axis=1)
    got = np.stack([c["x"][new], c["y"][new], c["z"][new]], axis=1)
    key = lambda a: np.lexsort((a[:, 2], a[:, 1], a[:, 0]))  # noqa: E731
    np.testing.assert_allclose(got[key(got)], centres[key(centres)], rtol=0, atol=1e-13)


# ------------------------------------------------------ column systems ----

def _banded(a, b, c):
    n = len(b)
    ab = np.zeros((3, n))
    ab[0, 1:] = c[:-1]
    ab[1] = b
    ab[2, :-1] = a[1:]
    return ab


def test_column_systems_against_scipy(oracle_mod):
    """diffz (fft_ops.F90:419-483, with and without bot/top) and laplinv (:505-601, both boundary
    conditions) assembled here with float32 literals 1/6, 2/3, 1/3, 1/12, 5/6 and the
    reference's boundary rows, solved by scipy.linalg.solve_banded: the oracle's columns agree to
    1e-13.  The same systems with exact literals differ from the oracle by more than that."""
    nx, ny, nz = 8, 6, 24
    o = make(oracle_mod, nx, ny, nz)
    g = o.geometry()
    dz = g["dz"]
    nzn = nz + 1
    rng = np.random.default_rng(61)
    s = rng.standard_normal(o.sshape)
    bot, top = rng.standard_normal(o.sshape[:2]), rng.standard_normal(o.sshape[:2])
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    f16, f23, f13, f112, f56 = (f32(np.float32(1) / np.float32(d)) for d in (6, 1.5, 3, 12, 1.2))
    assert f23 == f32(np.float32(2) / np.float32(3)) and f56 == f32(np.float32(5) / np.float32(6))
    k2 = o.wavenumbers()[2]

    def diffz_cols(lit, bt):
        a1, b1, c1, cl = lit["16"], lit["23"], lit["16"], lit["13"]
        out = np.zeros_like(s)
        for ii in range(s.shape[0]):
            for jj in range(s.shape[1]):
                f = s[ii, jj]
                a = np.full(nzn, a1); b = np.full(nzn, b1); c = np.full(nzn, c1)
                a[0] = b[0] - b[0]; b[0] = 1.0; c[0] = cl if bt else 0.0
                a[-1] = cl if bt else 0.0; b[-1] = 1.0; c[-1] = 0.0
                d = np.zeros(nzn)
                d[1:-1] = (f[2:] - f[:-2]) / 2.0 / dz
                if bt:
                    d[0] = f[1] / dz - a1 * bot[ii, jj] * dz
                    d[-1] = -f[-2] / dz + a1 * top[ii, jj] * dz
                out[ii, jj] = scipy.linalg.solve_banded((1, 1), _banded(a, b, c), d)
        return out

    def laplinv_cols(lit, df0):
        out = s.copy()
        idz2 = 1.0 / dz / dz
        for ii in range(s.shape[0]):
            for jj in range(s.shape[1]):
                kk = k2[ii, jj]
                if kk == 0:
                    continue
                f = s[ii, jj]
                a = np.full(nzn, idz2 - lit["112"] * kk)
                c = a.copy()
                b = np.full(nzn, -2.0 / dz / dz - lit["56"] * kk)
                d = np.zeros(nzn)
                d[1:-1] = lit["112"] * f[:-2] + lit["56"] * f[1:-1] + lit["112"] * f[2:]
                if df0:
                    b[0] = -idz2 - lit["13"] * kk; c[0] = idz2 - lit["16"] * kk
                    d[0] = lit["13"] * f[0] + lit["16"] * f[1]
                    a[-1] = idz2 - lit["16"] * kk; b[-1] = -idz2 - lit["13"] * kk
                    d[-1] = lit["13"] * f[-1] + lit["16"] * f[-2]
                else:
                    b[0] = b[-1] = 1.0; c[0] = a[-1] = 0.0
                a[0] = 0.0; c[-1] = 0.0
                out[ii, jj] = scipy.linalg.solve_banded((1, 1), _banded(a, b, c), d)
        return out

    lit32 = {"16": f16, "23": f23, "13": f13, "112": f112, "56": f56}
    exact = {"16": 1 / 6, "23": 2 / 3, "13": 1 / 3, "112": 1 / 12, "56": 5 / 6}
    cases = [(o.diffz(s), lambda L: diffz_cols(L, False)),
             (o.diffz(s, bot, top), lambda L: diffz_cols(L, True)),
             (o.laplinv(s, False), lambda L: laplinv_cols(L, False)),
             (o.laplinv(s, True), lambda L: laplinv_cols(L, True))]
    for got, build in cases:
        assert rel_err(got, build(lit32)) < 1e-13
        assert rel_err(got, build(exact)) > 1e-10


# ------------------------------------------------------------ splitting ----

def test_split_closed_form_fp32_third(oracle_mod):
    """parcel_splitting.F90:64-131 on parcels over the edge of the stretch limit: stretch grows
    by |p dp + q dq + r dr|^fl32(1/3) dt, the radius is (vol/4/pi)^fl32(1/3), parent and child sit
    at x -/+ r omega/|omega| with half the volume; to 1e-14 (an exact 1/3 exponent misses)."""
    o = make(oracle_mod, 8, 8, 8)
    g = o.geometry()
    vol = g["dx"] * g["dy"] * g["dz"] / 8
    p = np.array([0.6, 0.3]); q = np.array([-0.8, 0.4]); r = np.array([0.0, -0.5])
    dp = np.array([2.0, 1.5]); dq = np.array([1.0, -0.5]); dr = np.array([0.0, 2.5])
    s0 = np.array([3.999, 3.9995])
    X = np.array([1.0, 2.5]); Y = np.array([1.1, 2.0]); Z = np.array([1.2, 1.5])
    o.set_parcels({"x": X, "y": Y, "z": Z, "vol": np.full(2, vol), "p": p, "q": q, "r": r,
                   "dpdt": dp, "dqdt": dq, "drdt": dr, "stretch": s0, "tag": np.array([7.0, 8.0])})
    for _ in range(4):
        o.parcel_splitting()
    assert o.n == 4
    ds = np.abs(p * dp + q * dq + r * dr) ** F32_THIRD * o.dtm
    assert np.all(s0 + ds > 4)
    rad = (vol / 4 / np.pi) ** F32_THIRD
    assert abs(rad - (vol / 4 / np.pi) ** (1.0 / 3.0)) > 1e-9 * rad
    w = np.sqrt(p * p + q * q + r * r)
    c = o.get_parcels()
    for key, pos, comp in (("x", X, p), ("y", Y, q), ("z", Z, r)):
        np.testing.assert_allclose(c[key][:2], pos - rad * comp / w, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(c[key][2:], pos + rad * comp / w, rtol=1e-14, atol=1e-15)
    np.testing.assert_array_equal(c["vol"], np.full(4, vol / 2))
    np.testing.assert_array_equal(c["stretch"], np.zeros(4))
    np.testing.assert_array_equal(c["tag"], [7.0, 8.0, 7.0, 8.0])


# ---------------------------------------------------------- writebinary ----

def test_writebinary_hgliq_closed_form(oracle_mod):
    """writebinary.F90:217-233 with constant h0 and b0 on every parcel: hg = h0 at every node, so
    hgliq(z) = max(0, h0 - exp(-z)) exactly as a function of the node height and
    b = b0 + 12.5 hgliq(z)."""
    nx, ny, nz = 10, 8, 12
    o = make(oracle_mod, nx, ny, nz)
    g = o.geometry()
    ic = thermal_case(nx, ny, nz, 2, seed=71)
    n = len(ic["x"])
    h0, b0 = 0.5, 0.1
    o.set_parcels({"x": ic["x"], "y": ic["y"], "z": ic["z"], "vol": ic["vol"],
                   "h": np.full(n, h0), "b": np.full(n, b0)})
    o.writebinary_grids()
    z = np.arange(nz + 1) * g["dz"]
    liq = np.maximum(0.0, h0 - np.exp(-z))
    assert np.sum(liq > 0) > 2 and np.sum(liq == 0) > 2
    np.testing.assert_allclose(inner(o.grid("hg")), h0, rtol=1e-14)
    np.testing.assert_allclose(inner(o.grid("hgliq")), np.broadcast_to(liq, (nx, ny, nz + 1)),
                               rtol=0, atol=1e-14)
    np.testing.assert_allclose(inner(o.grid("b")), np.broadcast_to(b0 + 12.5 * liq,
                                                                   (nx, ny, nz + 1)),
                               rtol=0, atol=1e-13)
