"""Pins of the oracle (CPU FP64 restatement) against independent ground truth.

Every check here is chosen so that a plausible slip in the restatement (a dropped term, a wrong
sign, index or coefficient, a transposed layout) fails it:
* FFT conventions against numpy.fft and against the reference's own FFTE built from its
  sources into oracle/_ref (model_core/src/fft/ffte/*.f, ffte_mod.F90);
* the semi-spectral layout against a numpy construction of pencilfft.F90's definition;
* spectral derivatives against exact derivatives of single Fourier modes;
* the Thomas solver against scipy's banded solver; diffz / laplinv against manufactured
  solutions with their 4th-order convergence;
* grid2par / par2grid against the linear-field known answers of grid2partest.F90 and a
  brute-force numpy scatter with the periodic fold;
* vort2vel against curl(u) = corrected vorticity and div(u) = 0 on the grid;
* RK4 + mean flow against the exact shear flow of a uniform y-vorticity;
* haloswap against an independent model of the one-rank send-to-self unpack;
* splitting against its closed form; mixing against volume-integral conservation.
"""
import os

import numpy as np
import pytest
import scipy.linalg

from tests.cases import LX, LZ, box, rel_err, thermal_case


def make(oracle_mod, nx, ny, nz, cap=100000, **kw):
    b = box(nx, ny, nz)
    return oracle_mod.Oracle(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                             b["zmax"], cap, kw.get("dtm", 0.01), kw.get("dtmax", 0.01), 0.5)


# ---------------------------------------------------------------- FFT ----

@pytest.mark.parametrize("n", [8, 12, 15, 16, 30, 45, 48, 96, 128])
def test_r2c_c2r_match_numpy(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    np.testing.assert_allclose(oracle_mod.r2c(x), np.fft.rfft(x), rtol=0, atol=1e-12 * n)
    z = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    # c2r keeps only the real part of the Hermitian completion (ffte_mod.F90:109-122), so the
    # imaginary parts of the DC and (even n) Nyquist bins drop out exactly as in numpy's irfft
    np.testing.assert_allclose(oracle_mod.c2r(z, n), np.fft.irfft(z, n), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [8, 12, 15, 30, 48, 128])
def test_r2c_c2r_match_reference_ffte(oracle_mod, n):
    if not oracle_mod.RefFFTE.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    ref = oracle_mod.RefFFTE()
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n)
    np.testing.assert_allclose(oracle_mod.r2c(x), ref.r2c(x), rtol=0, atol=1e-13 * n)
    z = ref.r2c(x)
    np.testing.assert_allclose(oracle_mod.c2r(z, n), ref.c2r(z, n), rtol=0, atol=1e-14 * n)


def semispectral_numpy(f):
    """pencilfft.F90's forward transform of one field f[k, y, x] (numpy)."""
    nz, ny, nx = f.shape
    Y = np.fft.rfft(f, axis=1)                      # along y: [k, m, x]
    my = Y.shape[1]
    rb = np.empty((nz, 2 * my, nx))
    rb[:, 0::2, :] = Y.real
    rb[:, 1::2, :] = Y.imag
    X = np.fft.rfft(rb, axis=2)                     # along x of each real y-row: [k, jj, l]
    S = np.empty((nz, 2 * my, 2 * X.shape[2]))
    S[:, :, 0::2] = X.real
    S[:, :, 1::2] = X.imag
    return S


def to_oracle_grid(f, nx, ny):
    """[k, y, x] interior -> reference layout (x+4, y+4, z) with zero halos."""
    g = np.zeros((nx + 4, ny + 4, f.shape[0]))
    g[2:-2, 2:-2, :] = np.transpose(f, (2, 1, 0))
    return g


@pytest.mark.parametrize("shape", [(12, 10, 5), (16, 16, 4), (15, 6, 3)])
def test_semispectral_layout(oracle_mod, shape):
    nx, ny, nz = shape
    o = make(oracle_mod, nx, ny, nz)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((nz + 1, ny, nx))
    S = o.fft_forward(to_oracle_grid(f, nx, ny))
    want = semispectral_numpy(f)
    got = np.transpose(S, (2, 1, 0))
    assert rel_err(got, want) < 1e-13
    back = o.fft_backward(S)
    assert rel_err(back[2:-2, 2:-2, :], np.transpose(f, (2, 1, 0))) < 1e-13


# ---------------------------------------------------- spectral operators ----

@pytest.mark.parametrize("l,m", [(1, 0), (0, 2), (3, 1), (2, 4)])
def test_diffx_diffy_exact_on_modes(oracle_mod, l, m):
    nx, ny, nz = 16, 12, 4
    o = make(oracle_mod, nx, ny, nz)
    g = o.geometry()
    x = np.arange(nx) * g["dx"]
    y = np.arange(ny) * g["dy"]
    Lx, Ly = nx * g["dx"], ny * g["dy"]
    X, Y = np.meshgrid(x, y)                      # [y, x]
    ph = 2 * np.pi * (l * X / Lx + m * Y / Ly)
    f = np.broadcast_to(np.sin(ph), (nz + 1, ny, nx))
    dfdx = np.broadcast_to(2 * np.pi * l / Lx * np.cos(ph), (nz + 1, ny, nx))
    dfdy = np.broadcast_to(2 * np.pi * m / Ly * np.cos(ph), (nz + 1, ny, nx))
    S = o.fft_forward(to_oracle_grid(f, nx, ny))
    gx = o.fft_backward(o.diffx(S))[2:-2, 2:-2, :]
    gy = o.fft_backward(o.diffy(S))[2:-2, 2:-2, :]
    np.testing.assert_allclose(gx, np.transpose(dfdx, (2, 1, 0)), atol=1e-12)
    np.testing.assert_allclose(gy, np.transpose(dfdy, (2, 1, 0)), atol=1e-12)


def test_filter_keeps_low_drops_high(oracle_mod):
    nx, ny, nz = 24, 24, 3
    o = make(oracle_mod, nx, ny, nz)
    k = o.wavenumbers()  # [which, ii, jj]
    filt = k[3]
    kx, ky = k[0], k[1]
    kmax = np.pi * nx / LX
    kk = np.sqrt(kx ** 2 + ky ** 2)
    # the rule keeps everything below 1/3 of the isotropic cutoff and drops the corner modes
    assert np.all(filt[kk < kmax / 3] == 1)
    assert np.all(filt[kk > 0.95 * kmax] == 0)
    assert set(np.unique(filt)) <= {0.0, 1.0}
    # pairs (Re, Im) share their wavenumbers and mask
    assert np.array_equal(filt[0::2], filt[1::2])
    np.testing.assert_allclose(kx[2 * 3], 2 * np.pi * 3 / LX)


def test_tridiagonal_matches_scipy(oracle_mod):
    rng = np.random.default_rng(4)
    n = 40
    a = rng.uniform(0.1, 1, n)
    c = rng.uniform(0.1, 1, n)
    b = a + c + rng.uniform(1, 2, n)
    a[0] = 0
    c[-1] = 0
    d = rng.standard_normal(n)
    ab = np.zeros((3, n))
    ab[0, 1:] = c[:-1]
    ab[1] = b
    ab[2, :-1] = a[1:]
    np.testing.assert_allclose(oracle_mod.tridiagonal(a, b, c, d),
                               scipy.linalg.solve_banded((1, 1), ab, d), rtol=1e-12)
    with pytest.raises(ZeroDivisionError):
        oracle_mod.tridiagonal(np.zeros(3), np.array([1.0, 0.0, 1.0]), np.zeros(3), np.ones(3))


def column_spec(o, col_vals, jj=3, ii=5):
    """a spectral array that is zero except one (jj, ii) column."""
    s = np.zeros(o.sshape)
    s[ii, jj, :] = col_vals
    return s


def zgrid(o):
    g = o.geometry()
    return np.arange(o.nz + 1) * g["dz"], o.nz * g["dz"]


def test_diffz_fourth_order(oracle_mod):
    errs = []
    for nz in (16, 32):
        o = make(oracle_mod, 8, 8, nz)
        z, L = zgrid(o)
        f = np.cos(np.pi * z / L) + np.cos(3 * np.pi * z / L)   # f' = 0 at both ends
        d = o.diffz(column_spec(o, f))[5, 3, :]
        exact = -np.pi / L * np.sin(np.pi * z / L) - 3 * np.pi / L * np.sin(3 * np.pi * z / L)
        errs.append(np.max(np.abs(d - exact)))
    assert errs[1] < 1e-3 and errs[0] / errs[1] > 12.0


def test_diffz_with_boundary_values_reference_rows(oracle_mod):
    """diffz with bot/top (fft_ops.F90:452-464) uses b=1 in its boundary rows where the comment
    says 2/3.  For smooth f with f=0 at the ends and bot/top = f'' there, the Pade interior makes
    the boundary error e_1 = -f'_0/(3+lam) with lam = sqrt(3)-2 the decay factor of the interior
    recurrence, so the computed boundary slope tends to (1 - 1/(1+sqrt 3)) f'_0 and the error
    decays into the column as lam^k; far from the ends the scheme is 4th order."""
    lam = np.sqrt(3.0) - 2.0
    far = []
    for nz in (32, 64):
        o = make(oracle_mod, 8, 8, nz)
        z, L = zgrid(o)
        f = np.sin(np.pi * z / L) + z * (L - z)
        bot = np.zeros(o.sshape[:2])
        top = np.zeros(o.sshape[:2])
        bot[5, 3], top[5, 3] = -2.0, -2.0              # f'' at both ends
        d = o.diffz(column_spec(o, f), bot, top)[5, 3, :]
        exact = np.pi / L * np.cos(np.pi * z / L) + L - 2 * z
        ratio = 1.0 - 1.0 / (1.0 + np.sqrt(3.0))
        np.testing.assert_allclose(d[0] / exact[0], ratio, rtol=2e-3)
        np.testing.assert_allclose(d[-1] / exact[-1], ratio, rtol=2e-3)
        e = d - exact
        np.testing.assert_allclose(e[1:4] / e[0:3], lam, rtol=0.05)
        far.append(np.max(np.abs(e[nz // 4: 3 * nz // 4])))
    assert far[0] / far[1] > 12.0


@pytest.mark.parametrize("df0", [False, True])
def test_laplinv_convergence(oracle_mod, df0):
    """f = 0 rows keep the compact scheme 4th order; the dA/dz = 0 rows of fft_ops.F90:579-586
    are 3rd-order accurate, which sets the global order with that boundary condition."""
    errs = []
    for nz in (16, 32):
        o = make(oracle_mod, 8, 8, nz)
        z, L = zgrid(o)
        k2 = o.wavenumbers()[2][5, 3]
        A = np.cos(2 * np.pi * z / L) if df0 else np.sin(2 * np.pi * z / L)
        B = -(2 * np.pi / L) ** 2 * A - k2 * A
        got = o.laplinv(column_spec(o, B), df0)[5, 3, :]
        errs.append(np.max(np.abs(got - A)))
    if df0:
        assert errs[1] < 1e-3 and 7.0 < errs[0] / errs[1] < 9.5
    else:
        assert errs[1] < 1e-4 and errs[0] / errs[1] > 14.0


def test_laplinv_skips_mean_columns(oracle_mod):
    o = make(oracle_mod, 8, 8, 8)
    rng = np.random.default_rng(2)
    s = rng.standard_normal(o.sshape)
    out = o.laplinv(s, True)
    # the 2x2 block of zero wavenumber (fft_ops.F90:567) is returned unchanged
    np.testing.assert_array_equal(out[:2, :2, :], s[:2, :2, :])
    assert not np.array_equal(out[2:, :, :], s[2:, :, :])


# ------------------------------------------------- interpolation ----

def lattice(nx, ny, nz, n, dx, dy, dz):
    """parcels at the centres of an n^3 sub-lattice per cell (no jitter)."""
    a = (np.arange(n) + 0.5) / n
    xs = (np.arange(nx)[:, None] + a[None, :]).ravel() * dx
    ys = (np.arange(ny)[:, None] + a[None, :]).ravel() * dy
    zs = (np.arange(nz)[:, None] + a[None, :]).ravel() * dz
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def test_grid2par_linear_fields_known_answer(oracle_mod):
    """grid2partest.F90:105-201: u=x, v=y, w=z grids interpolate to x, y, z (tol 1e-9) away
    from the periodic seams."""
    nx, ny, nz = 12, 10, 8
    o = make(oracle_mod, nx, ny, nz)
    g = o.geometry()
    X, Y, Z = lattice(nx, ny, nz, 2, g["dx"], g["dy"], g["dz"])
    n = len(X)
    o.set_parcels({"x": X, "y": Y, "z": Z, "vol": np.full(n, g["dx"] * g["dy"] * g["dz"] / 8)})
    o.cache_weights()
    xc = g["xmin_h"] + np.arange(nx + 4) * g["dx"]
    yc = g["ymin_h"] + np.arange(ny + 4) * g["dy"]
    zc = np.arange(nz + 1) * g["dz"]
    gx = np.broadcast_to(xc[:, None, None], o.gshape)
    gy = np.broadcast_to(yc[None, :, None], o.gshape)
    gz = np.broadcast_to(zc[None, None, :], o.gshape)
    inner = (X > g["dx"]) & (X < (nx - 1) * g["dx"]) & (Y > g["dy"]) & (Y < (ny - 1) * g["dy"])
    for grid, ref in ((gx, X), (gy, Y), (gz, Z)):
        got, _ = o.grid2par(grid.copy())
        assert np.max(np.abs(got - ref)[inner]) < 1e-9


def test_par2grid_linear_fields_known_answer(oracle_mod):
    """grid2partest.F90:204-265: p=x parcels on a symmetric lattice grid to x at interior nodes."""
    nx, ny, nz = 12, 10, 8
    o = make(oracle_mod, nx, ny, nz)
    g = o.geometry()
    X, Y, Z = lattice(nx, ny, nz, 2, g["dx"], g["dy"], g["dz"])
    n = len(X)
    o.set_parcels({"x": X, "y": Y, "z": Z, "vol": np.full(n, 1.0)})
    o.cache_weights()
    xc = np.arange(-2, nx + 2) * g["dx"]
    yc = np.arange(-2, ny + 2) * g["dy"]
    zc = np.arange(nz + 1) * g["dz"]
    for var, coord, axis in ((X, xc, 0), (Y, yc, 1), (Z, zc, 2)):
        grid = o.par2grid(var)
        sl = [slice(3, -3), slice(3, -3), slice(1, -1)]
        want = np.broadcast_to(coord.reshape([-1 if a == axis else 1 for a in range(3)]), o.gshape)
        assert np.max(np.abs(grid[tuple(sl)] - want[tuple(sl)])) < 1e-9


def brute_par2grid(X, Y, Z, vol, var, nx, ny, nz, dx, dy, dz):
    """trilinear volume-weighted scatter with periodic x/y, normalised (numpy.add.at)."""
    fx, fy, fz = X / dx, Y / dy, Z / dz
    i = np.floor(fx).astype(int)
    j = np.floor(fy).astype(int)
    k = np.floor(fz).astype(int)
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


def test_par2grid_matches_brute_force(oracle_mod):
    nx, ny, nz = 12, 10, 8
    ic = thermal_case(nx, ny, nz, 2, seed=3)
    o = make(oracle_mod, nx, ny, nz)
    o.set_parcels(ic)
    o.cache_weights()
    g = o.geometry()
    for var in (ic["b"], np.sin(ic["x"]) + ic["z"]):
        got = o.par2grid(var)[2:-2, 2:-2, :]
        want = brute_par2grid(ic["x"], ic["y"], ic["z"], ic["vol"], var, nx, ny, nz,
                              g["dx"], g["dy"], g["dz"])
        assert rel_err(got, want) < 1e-12


def test_par2grid_constant_and_grid2par_roundtrip(oracle_mod):
    nx, ny, nz = 12, 10, 8
    ic = thermal_case(nx, ny, nz, 2, seed=4)
    o = make(oracle_mod, nx, ny, nz)
    o.set_parcels(ic)
    o.cache_weights()
    n = len(ic["x"])
    grid = o.par2grid(np.full(n, 2.5))
    np.testing.assert_allclose(grid[2:-2, 2:-2, :], 2.5, rtol=1e-14)
    back, _ = o.grid2par(grid)
    np.testing.assert_allclose(back, 2.5, rtol=1e-14)


def test_cache_out_of_box_z(oracle_mod):
    o = make(oracle_mod, 8, 8, 8)
    g = o.geometry()
    o.set_parcels({"x": np.array([0.1]), "y": np.array([0.1]), "z": np.array([g["zmax_h"]]),
                   "vol": np.array([1.0])})
    with pytest.raises(RuntimeError, match="status 3"):
        o.cache_weights()


# ------------------------------------------------------ composition ----

def test_vort2vel_velocity_is_curl_inverse(oracle_mod):
    """u = -curl(A) with laplacian(A) = vorticity: the gridded velocity is divergence free and
    its curl reproduces the corrected vorticity away from the rigid lids (the boundary rows of
    diffz perturb the first few levels, decaying by sqrt(3)-2 per level)."""
    nx, ny, nz = 16, 16, 32
    o = make(oracle_mod, nx, ny, nz, cap=200000)
    g = o.geometry()
    ic = thermal_case(nx, ny, nz, 2, seed=9)
    Lx, Lz = nx * g["dx"], nz * g["dz"]
    x, z = ic["x"], ic["z"]
    ic["q"] = np.sin(2 * np.pi * x / Lx) * np.sin(np.pi * z / Lz)   # a smooth y-vorticity
    o.set_parcels(ic)
    o.vort2vel()
    U, V, W = (o.grid(k)[2:-2, 2:-2, :] for k in ("u", "v", "w"))
    P, Q, R = (o.grid(k)[2:-2, 2:-2, :] for k in ("p", "q", "r"))
    kx = 2 * np.pi * np.fft.fftfreq(nx, g["dx"])
    ky = 2 * np.pi * np.fft.fftfreq(ny, g["dy"])

    def ddx(f):
        return np.real(np.fft.ifft(1j * kx[:, None, None] * np.fft.fft(f, axis=0), axis=0))

    def ddy(f):
        return np.real(np.fft.ifft(1j * ky[None, :, None] * np.fft.fft(f, axis=1), axis=1))

    def ddz(f):
        return np.gradient(f, g["dz"], axis=2, edge_order=2)

    sl = (slice(None), slice(None), slice(7, -7))
    div = ddx(U) + ddy(V) + ddz(W)
    scale = np.max(np.abs(U)) / g["dz"]
    assert np.max(np.abs(div[sl])) < 0.02 * scale
    curl_y = ddz(U) - ddx(W)
    assert rel_err(curl_y[sl], Q[sl]) < 0.02
    curl_x = ddy(W) - ddz(V)
    assert np.max(np.abs(curl_x[sl])) < 0.02 * np.max(np.abs(Q))
    # the divergence cleaning only removes the parcel-sampling noise of q(x, z)
    assert np.max(np.abs(P)) < 1e-2 * np.max(np.abs(Q)) and np.max(np.abs(R)) < 1e-2 * np.max(np.abs(Q))


def test_uniform_y_vorticity_gives_linear_shear(oracle_mod):
    """A uniform y-vorticity q0 carries no buoyancy or stretching, so u(z) = q0 z + const from
    the mean-flow integral and every parcel moves by dt*u(z) in x over one RK4 step."""
    nx, ny, nz = 8, 8, 16
    o = make(oracle_mod, nx, ny, nz, cap=200000)
    ic = thermal_case(nx, ny, nz, 2, seed=5)
    n = len(ic["x"])
    q0 = 0.3
    fields = {"x": ic["x"], "y": ic["y"], "z": ic["z"], "vol": ic["vol"], "q": np.full(n, q0),
              "h": np.zeros(n), "b": np.zeros(n), "tag": np.arange(n, dtype=float)}
    o.set_parcels(fields)
    dt = o.dtm
    for _ in range(4):
        o.vort2vel()
        o.vorticity_tendency()
        o.rk4_integrator()
    after = o.get_parcels(["x", "y", "z", "tag"])
    order = np.argsort(after["tag"])
    x1, z1 = after["x"][order], after["z"][order]
    L = o.geometry()["xshift_w"]
    dx = (x1 - ic["x"] + L / 2) % L - L / 2
    # linear in z with slope dt*q0
    slope, icpt = np.polyfit(ic["z"], dx, 1)
    assert abs(slope - dt * q0) < 1e-9
    np.testing.assert_allclose(dx, slope * ic["z"] + icpt, atol=1e-12)
    np.testing.assert_allclose(z1, ic["z"], atol=1e-13)


# ------------------------------------------------------ parcel events ----

def haloswap_model(fields, n, geom):
    """one-rank parcel_haloswap (haloswap.F90:109-289) as plain list operations."""
    X, Y = fields["x"].copy(), fields["y"].copy()
    xmin, xmax, ymin, ymax = geom["hs_xmin"], geom["hs_xmax"], geom["hs_ymin"], geom["hs_ymax"]

    def dir_of(x, y):
        if y >= ymax and xmin <= x < xmax: return 1
        if x >= xmax and y >= ymax: return 2
        if x >= xmax and ymin <= y < ymax: return 3
        if x >= xmax and y < ymin: return 4
        if y < ymin and xmin <= x < xmax: return 5
        if x < xmin and y < ymin: return 6
        if x < xmin and ymin <= y < ymax: return 7
        if x < xmin and y >= ymax: return 8
        return 0

    dirs = [dir_of(X[i], Y[i]) for i in range(n)]
    messages = []
    for d in range(1, 9):
        xs = geom["xshift_e"] if d in (2, 3, 4) else geom["xshift_w"] if d in (6, 7, 8) else 0.0
        ys = geom["yshift_n"] if d in (1, 2, 8) else geom["yshift_s"] if d in (4, 5, 6) else 0.0
        messages += [(i, xs, ys) for i in range(n) if dirs[i] == d]
    holes = [i for i in range(n) if dirs[i]]
    out = {k: v.copy() for k, v in fields.items()}
    for slot, (i, xs, ys) in zip(holes, messages):
        for k in fields:
            out[k][slot] = fields[k][i]
        out["x"][slot] = fields["x"][i] + xs
        out["y"][slot] = fields["y"][i] + ys
    return out


def test_haloswap_matches_independent_model(oracle_mod):
    nx, ny, nz = 12, 10, 8
    o = make(oracle_mod, nx, ny, nz)
    ic = thermal_case(nx, ny, nz, 2, seed=6)
    n = len(ic["x"])
    rng = np.random.default_rng(6)
    moved = {k: ic[k].copy() for k in ("x", "y", "z", "vol", "tag")}
    moved["x"] += rng.choice([-1.0, 0.0, 1.0], n) * rng.uniform(0, 0.3, n)
    moved["y"] += rng.choice([-1.0, 0.0, 1.0], n) * rng.uniform(0, 0.3, n)
    o.set_parcels(moved)
    o.parcel_haloswap()
    got = o.get_parcels(["x", "y", "tag"])
    want = haloswap_model({k: moved[k] for k in ("x", "y", "tag")}, n, o.geometry())
    for k in ("x", "y", "tag"):
        np.testing.assert_array_equal(got[k], want[k])
    g = o.geometry()
    assert np.all((got["x"] >= g["hs_xmin"]) & (got["x"] < g["hs_xmax"]))
    np.testing.assert_allclose(g["xshift_w"], LX, rtol=1e-15)


def test_reflection_closed_form(oracle_mod):
    o = make(oracle_mod, 8, 8, 8)
    g = o.geometry()
    z = np.array([-0.05, 0.2, g["zmax_h"] + 0.07])
    o.set_parcels({"x": np.full(3, 1.0), "y": np.full(3, 1.0), "z": z, "vol": np.ones(3)})
    o.parcel_haloswap()
    np.testing.assert_allclose(o.field("z")[:3], [0.05, 0.2, g["zmax_h"] - 0.07], rtol=1e-14)


def test_split_closed_form(oracle_mod):
    """parcel_splitting.F90:64-131 on one parcel over the edge of the stretch limit."""
    o = make(oracle_mod, 8, 8, 8)
    g = o.geometry()
    vol = g["dx"] * g["dy"] * g["dz"] / 8
    p, q, r = 0.6, -0.8, 0.0
    f = {"x": np.array([1.0]), "y": np.array([1.0]), "z": np.array([1.0]), "vol": np.array([vol]),
         "p": np.array([p]), "q": np.array([q]), "r": np.array([r]),
         "dpdt": np.array([2.0]), "dqdt": np.array([1.0]), "drdt": np.array([0.0]),
         "stretch": np.array([3.999]), "tag": np.array([7.0])}
    o.set_parcels(f)
    for _ in range(3):
        o.parcel_splitting()            # stages 1-3: nothing
    assert o.n == 1
    o.parcel_splitting()
    assert o.n == 2
    ds = abs(p * 2.0 + q * 1.0) ** (1.0 / 3.0) * o.dtm
    assert 3.999 + ds > 4
    rad = (vol / 4 / np.pi) ** (1.0 / 3.0)
    c = o.get_parcels()
    np.testing.assert_allclose(c["x"], [1.0 - rad * 0.6, 1.0 + rad * 0.6], rtol=1e-6)
    np.testing.assert_allclose(c["y"], [1.0 + rad * 0.8, 1.0 - rad * 0.8], rtol=1e-6)
    np.testing.assert_allclose(c["vol"], [vol / 2, vol / 2], rtol=1e-15)
    np.testing.assert_array_equal(c["stretch"], [0.0, 0.0])
    np.testing.assert_array_equal(c["tag"], [7.0, 7.0])


def test_mixing_conserves_volume_integrals(oracle_mod):
    """parcel_mixing removes small parcels and fills sparse cells while conserving the
    volume-integrated quantities (report.tex:264: ~1e-13)."""
    nx, ny, nz = 12, 10, 8
    ic = thermal_case(nx, ny, nz, 2, seed=8, vort=1.0, shrink=0.05, thin=0.1)
    o = make(oracle_mod, nx, ny, nz, cap=3 * len(ic["x"]))
    o.set_parcels(ic)

    def integrals():
        c = o.get_parcels()
        v = c["vol"]
        return np.array([v.sum(), (v * c["b"]).sum(), (v * c["h"]).sum(), (v * c["p"]).sum(),
                         (v * c["q"]).sum(), (v * c["r"]).sum()])

    before = integrals()
    n0 = o.n
    for _ in range(4):
        o.parcel_mixing()
    removed, added = o.last_counts()[1:]
    assert removed > 0 and added > 0
    assert o.n == n0 - removed + added
    after = integrals()
    assert np.max(np.abs(after - before) / np.abs(before)) < 1e-12
    vmin = o.geometry()["dx"] * o.geometry()["dy"] * o.geometry()["dz"] / 216
    assert np.all(o.get_parcels()["vol"] >= vmin)


def test_step_is_deterministic(oracle_mod):
    nx, ny, nz = 8, 8, 6
    ic = thermal_case(nx, ny, nz, 2, seed=2, vort=2.0, shrink=0.02, thin=0.05, near_split=0.03)
    res = []
    for _ in range(2):
        o = make(oracle_mod, nx, ny, nz, cap=3 * len(ic["x"]))
        o.set_parcels(ic)
        o.step()
        o.step()
        res.append(o.get_parcels())
    for k in res[0]:
        np.testing.assert_array_equal(res[0][k], res[1][k])


def test_threads_change_no_result(oracle_mod):
    """The threaded oracle on 1, 2 and 5 threads: identical parcels, grids and counts over three
    steps with splits, removals, additions and haloswap crossings (each grid node receives its
    terms in ascending parcel order whatever the thread count)."""
    nx, ny, nz = 20, 15, 8
    ic = thermal_case(nx, ny, nz, 3, seed=5, vort=3.0, shrink=0.03, thin=0.05, near_split=0.03)
    res = []
    try:
        for thr in (1, 2, 5):
            oracle_mod.set_threads(thr)
            o = make(oracle_mod, nx, ny, nz, cap=3 * len(ic["x"]))
            o.set_parcels(ic)
            counts = []
            for _ in range(3):
                o.step()
                counts.append(o.last_counts())
            res.append((o.get_parcels(), {g: o.grid(g).copy() for g in oracle_mod.GRIDS},
                        counts, o.dtm))
    finally:
        oracle_mod.set_threads(os.cpu_count() or 1)
    assert all(c > 0 for c in res[0][2][0]), res[0][2]  # splits, removals and additions ran
    for p, g, counts, dtm in res[1:]:
        assert counts == res[0][2] and dtm == res[0][3]
        for k in p:
            np.testing.assert_array_equal(p[k], res[0][0][k], err_msg=k)
        for k in g:
            np.testing.assert_array_equal(g[k], res[0][1][k], err_msg=k)


@pytest.mark.parametrize("vort,dtmax", [(4.0, 10.0), (0.2, 10.0), (4.0, 0.01)])
def test_adaptive_timestep_rule(oracle_mod, vort, dtmax):
    """The stage-1 timestep is min(dtmax, Lx/umax, Ly/vmax, Lz'/wmax, prefactor/omax)
    (vort2vel.F90:329-381, vorticity_tendency.F90:277-321): recomputed here in numpy from the
    oracle's own exported grids (interior maxima, Lz' = (nz+1) dz), with dtmax large enough for
    the CFL or vorticity bound to bind and small enough for it not to."""
    nx, ny, nz = 12, 10, 8
    b = box(nx, ny, nz)
    ic = thermal_case(nx, ny, nz, 2, vort=vort)
    o = oracle_mod.Oracle(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                          b["zmax"], 3 * len(ic["x"]), 0.01, dtmax, 0.5)
    o.set_parcels(ic)
    o.vort2vel()
    o.vorticity_tendency()
    inner = lambda a: a[2:-2, 2:-2, :]
    u, v, w = (inner(o.grid(k)) for k in ("u", "v", "w"))
    p, q, r = (inner(o.grid(k)) for k in ("p", "q", "r"))
    dx, dy, dz = LX / nx, LX / ny, LZ / nz
    vmax = [max(float(np.sqrt(np.max(a ** 2))), 1e-15) for a in (u, v, w)]
    cfl = min(nx * dx / vmax[0], ny * dy / vmax[1], (nz + 1) * dz / vmax[2])
    dt = dtmax if dtmax <= cfl else cfl
    omax = float(np.sqrt(np.max(p ** 2 + q ** 2 + r ** 2)))
    if omax > 0:
        dt = min(dt, 0.5 / omax)
    assert o.dtm == pytest.approx(dt, rel=1e-13, abs=0)
    if dtmax == 10.0:
        assert o.dtm < dtmax  # an adaptive bound did bind


def test_mean_flow_mode(oracle_mod):
    """The horizontal mean of the velocity is the zero-wavenumber mode of vort2vel.F90:221-247:
    ubar(k) = ubar(k-1) + dz/2 (e(k-1) + e(k)), vbar with -d, from the (1,1) columns of the
    y- and x-vorticity spectra, minus the reference's average
    fl32(1/nz') * sum_{k<nz'} (ubar(k) + ubar(nz')/2).  Recomputed in numpy from the
    horizontal means of the oracle's own vorticity grids (the mean mode is not touched by the
    divergence correction or the filter), and checked against its velocity grids."""
    nx, ny, nz = 12, 10, 8
    b = box(nx, ny, nz)
    ic = thermal_case(nx, ny, nz, 2)
    rng = np.random.default_rng(3)
    n = len(ic["x"])
    # vorticity with a horizontal mean that varies in z
    ic["p"] = 0.7 * np.cos(ic["z"]) + 0.1 * rng.uniform(-1, 1, n)
    ic["q"] = 0.5 + 0.3 * ic["z"] + 0.1 * rng.uniform(-1, 1, n)
    ic["r"] = 0.1 * rng.uniform(-1, 1, n)
    o = oracle_mod.Oracle(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                          b["zmax"], 3 * n, 0.01, 0.01, 0.5)
    o.set_parcels(ic)
    o.vort2vel()
    inner = lambda a: a[2:-2, 2:-2, :]
    mean_xy = lambda a: inner(a).mean(axis=(0, 1))
    dz = LZ / nz
    nzn = nz + 1
    e = mean_xy(o.grid("q"))  # (1,1) column of the y-vorticity spectrum / (nx ny)
    d = mean_xy(o.grid("p"))
    ub = np.zeros(nzn)
    vb = np.zeros(nzn)
    for k in range(1, nzn):
        ub[k] = ub[k - 1] + dz / 2.0 * (e[k - 1] + e[k])
        vb[k] = vb[k - 1] - dz / 2.0 * (d[k - 1] + d[k])
    rnz = float(np.float32(1.0) / np.float32(nzn))
    ref_avg = lambda a: rnz * sum(a[k - 1] + 0.5 * a[nzn - 1] for k in range(1, nzn))
    uref, vref = ub - ref_avg(ub), vb - ref_avg(vb)
    np.testing.assert_allclose(mean_xy(o.grid("u")), uref, rtol=0, atol=1e-12 * np.abs(uref).max())
    np.testing.assert_allclose(mean_xy(o.grid("v")), vref, rtol=0, atol=1e-12 * np.abs(vref).max())
    # removing the true vertical mean (trapezoid) instead would differ visibly
    ualt = ub - np.trapezoid(ub, dx=1.0) / (nzn - 1)
    assert np.abs(ualt - uref).max() > 1e-6 * np.abs(uref).max()
