"""The plume initial condition of components/plume_parcelsetup/src/plume_parcelsetup.F90, for one
rank holding the whole domain.

A spherical thermal (Dritschel et al. 2018) of radius r_plume resting on the ground, centred in
x and y, over a mixed layer below z_b and a stratified layer above, all lengths in units of the
scale height 1/lambda (plume_parcelsetup.F90:80-145):

* h_pl = exp(-z_c) (humidity in the plume), h_bg = mu h_pl (in the background),
  z_b = log(H / h_bg) (base of the stratified layer),
  dbdz = glat (h_pl - exp(-z_m)) / (z_m - z_d), b_pl = dbdz (z_d - z_b);
  glat = gravity * latent with latent = 0.125, fdthetal0 = 0.01 (a default-real literal, so
  FP32-rounded: :30-43) and gravity = 1 / fdthetal0;
* then z_c, z_b, z_d, z_m and r_plume are multiplied by lambda and e_values divided by
  r_plume^2;
* plume parcels (:183-220): a 4 x 4 x 4 lattice per cell over x and y and the lowest
  ceiling(2 r_plume / dz_plume) levels, x outermost and z innermost, kept inside the sphere
  (zp^2 + xp^2 + yp^2 <= r_plume^2), with b = b_pl (1 + e1 xp yp + e2 xp zp + e3 yp zp),
  h = h_pl and the lattice volume;
* background parcels (:224-282): a 2 x 2 x 2 lattice per cell over the whole box, kept
  outside the sphere, with b = 0, h = h_bg below z_b and b = dbdz / lambda (z - z_b),
  h = H exp(-z / lambda) above;
* lattice coordinates are running sums (x = x + dx_plume), as the reference forms them.

The reference leaves p, q, r, stretch and tag unset; they are zero here.  Parameter checks stop
the run as the reference does (PlumeError).
"""
import math

import numpy as np

from .dumpio import coordinates

LATENT = 0.125
FDTHETAL0 = float(np.float32(0.01))
GRAVITY = 1.0 / FDTHETAL0
GLAT = GRAVITY * LATENT
N_PER_CELL_DIR_PLUME = 4
N_PER_CELL_DIR_BG = 2


class PlumeError(RuntimeError):
    pass


def _running(start, step, count):
    """start, start + step, (start + step) + step, ...: the reference's accumulated lattice."""
    out = np.empty(count)
    v = start
    for i in range(count):
        out[i] = v
        v = v + step
    return out


def plume_parameters(lam, H, z_c, mu, z_d, z_m, r_plume, e_values, log=print):
    """The derived plume quantities, scaled by lambda (plume_parcelsetup.F90:80-145)."""
    log(f"Scale height= {lam:7.2f}m")
    if H > 1:
        log(f"Error: Relative humidity fraction must be less than 1. The selected value is {H:6.3f}")
        raise PlumeError("relative humidity fraction H must be less than 1")
    h_pl = math.exp(-z_c)
    log(f"Humidity inside the plume is {h_pl:6.3f}")
    if mu > 1 or mu <= H:
        log(f"Error: mu must be between H and 1. The selected value is {mu:6.3f}")
        raise PlumeError("mu must be between H and 1")
    h_bg = mu * h_pl
    log(f"Background humidity is {h_bg:6.3f}")
    z_b = math.log(H / h_bg)
    log(f"Base of mixed layer is {z_b:6.3f}")
    dbdz = GLAT * (h_pl - math.exp(-z_m)) / (z_m - z_d)
    log(f"The buoyancy frequency in the stratified zone is {math.sqrt(dbdz):6.3f}")
    b_pl = dbdz * (z_d - z_b)
    if 2. * r_plume > z_b:
        log(f"Error: Plume radius is too big. At most it can be {z_b / 2.:7.5f}")
        raise PlumeError("plume radius too big")
    e = np.asarray(e_values, dtype=np.float64)
    r = r_plume * lam
    return {"h_pl": h_pl, "h_bg": h_bg, "z_b": z_b * lam, "dbdz": dbdz, "b_pl": b_pl,
            "z_c": z_c * lam, "z_d": z_d * lam, "z_m": z_m * lam, "r_plume": r,
            "e_values": e / r / r, "lambda": lam, "H": H}


def plume_parcels(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, params, max_parcels=None):
    """(parcel arrays, number of plume parcels) of the plume initial condition on an
    nx x ny x nz-cell box (one rank); the plume parcels come first."""
    P = params
    xc_, yc_, zc_ = coordinates(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax)
    # this rank's box (plume_parcelsetup.F90:168-177): x_coords(halo + 1),
    # x_coords(2 halo + nx - 1), z_coords(1), z_coords(nz + 1)
    xmin_l, xmax_l = xc_[2], xc_[nx + 2]
    ymin_l, ymax_l = yc_[2], yc_[ny + 2]
    zmin_l = zc_[0]
    dx = (xmax - xmin) / nx
    dy = (ymax - ymin) / ny
    dz = (zmax - zmin) / nz
    top_x, top_y = xmax - dx, ymax - dy  # setup_grid.F90:91-93
    x_c = (top_x + dx - xmin) / 2 + xmin
    y_c = (top_y + dy - ymin) / 2 + ymin
    r = P["r_plume"]
    z_c = r
    e1, e2, e3 = P["e_values"]
    out = {k: [] for k in ("x", "y", "z", "b", "h", "vol")}

    # plume lattice
    dxp, dyp, dzp = dx / N_PER_CELL_DIR_PLUME, dy / N_PER_CELL_DIR_PLUME, dz / N_PER_CELL_DIR_PLUME
    volp = dxp * dyp * dzp
    if xmax_l > x_c - r and xmin_l < x_c + r and ymax_l > y_c - r and ymin_l < y_c + r:
        xs = _running(xmin_l + dxp / 2., dxp, nx * N_PER_CELL_DIR_PLUME)
        ys = _running(ymin_l + dyp / 2., dyp, ny * N_PER_CELL_DIR_PLUME)
        zs = _running(zmin_l + dzp / 2., dzp, math.ceil(2. * r / dzp))
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")  # x outermost, z innermost
        xp, yp, zp = X - x_c, Y - y_c, Z - z_c
        inside = (zp * zp + xp * xp + yp * yp) <= r * r
        xp, yp, zp = xp[inside], yp[inside], zp[inside]
        out["x"].append(X[inside])
        out["y"].append(Y[inside])
        out["z"].append(Z[inside])
        out["b"].append(P["b_pl"] * (1. + e1 * xp * yp + e2 * xp * zp + e3 * yp * zp))
        out["h"].append(np.full(xp.size, P["h_pl"]))
        out["vol"].append(np.full(xp.size, volp))

    # background lattice
    dxb, dyb, dzb = dx / N_PER_CELL_DIR_BG, dy / N_PER_CELL_DIR_BG, dz / N_PER_CELL_DIR_BG
    volb = dxb * dyb * dzb
    xs = _running(xmin_l + dxb / 2., dxb, nx * N_PER_CELL_DIR_BG)
    ys = _running(ymin_l + dyb / 2., dyb, ny * N_PER_CELL_DIR_BG)
    zs = _running(zmin_l + dzb / 2., dzb, nz * N_PER_CELL_DIR_BG)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    xp, yp, zp = X - x_c, Y - y_c, Z - z_c
    outside = (zp * zp + yp * yp + xp * xp) > r * r
    Zo = Z[outside]
    mixed = Zo < P["z_b"]
    lam = P["lambda"]
    out["x"].append(X[outside])
    out["y"].append(Y[outside])
    out["z"].append(Zo)
    out["b"].append(np.where(mixed, 0., P["dbdz"] / lam * (Zo - P["z_b"])))
    out["h"].append(np.where(mixed, P["h_bg"], P["H"] * np.exp(-Zo / lam)))
    out["vol"].append(np.full(Zo.size, volb))
    fields = {k: np.concatenate(v) for k, v in out.items()}
    n = len(fields["x"])
    if max_parcels is not None and n > max_parcels:
        raise PlumeError("Maxparcels reached")
    for k in ("p", "q", "r", "stretch", "tag"):
        fields[k] = np.zeros(n)
    return fields, n - int(np.count_nonzero(outside))


def plume_from_options(options, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, log=print):
    """plume_parcelsetup's initialisation from the config options (lambda, H, z_c, mu, z_d, z_m,
    r_plume, e_values; plume_parcelsetup.F90:83-135), reals parsed the reference's way;
    returns (parcel arrays, number of plume parcels)."""
    from .config import get_real
    e = options.get("e_values")
    if isinstance(e, str):
        from .config import fortran_real
        e = [fortran_real(v) for v in e.split(",")]
    params = plume_parameters(get_real(options, "lambda"), get_real(options, "H"),
                              get_real(options, "z_c"), get_real(options, "mu"),
                              get_real(options, "z_d"), get_real(options, "z_m"),
                              get_real(options, "r_plume"), e, log=log)
    maxp = options.get("max_parcels")
    return plume_parcels(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, params,
                         int(maxp) if maxp is not None else None)
