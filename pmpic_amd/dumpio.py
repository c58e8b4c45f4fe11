"""Dump files in the reference's binary formats, and the multi-file restart reader.

Parcel dumps (components/writebinary/src/writebinary.F90:149-209), one unformatted stream file
per rank named ``parcels_RRRRR_NNNNN.dat`` (rank and dump number, five digits each): time
(FP64), the rank's x/y/z bounds xmin xmax ymin ymax zmin zmax (6 x FP64), the parcel count
(int64), then the 17 parcel arrays x .. tag (FP64, n each) and the qvalues (none on this path).

Grid dumps (writebinary.F90:211-299), ``grids_RRRRR_NNNNN.dat``: time (FP64), the bounds
x_coords(first), x_coords(last), y..., z... of the rank's interior nodes (6 x FP64), the node
counts nx, ny, nz+1 (3 x int32), then the interior of u, v, w, p, q, r, b, hg, hgliq, vol in the
grids' own (z, y, x) order, z fastest.

The restart reader follows components/readfromfile_parcelsetup/src/readfromfile_parcelsetup.F90:
26-150 and 183-316: it counts the files ``parcels_00000_NNNNN.dat``, ``parcels_00001_...`` of dump
number NNNNN (restart_num) until one is missing, adds up their counts, reads every file whose x/y
range overlaps this domain's, keeps the parcels with x in [xmin, xmax) and y in [ymin, ymax) in
file order, takes the time of the last file read, and fails if the parcels kept are not all the
parcels the files hold.
"""
import os

import numpy as np

FIELDS = ["x", "y", "z", "p", "q", "r", "dxdt", "dydt", "dzdt", "dpdt", "dqdt", "drdt",
          "h", "b", "vol", "stretch", "tag"]
GRID_FIELDS = ["u", "v", "w", "p", "q", "r", "b", "hg", "hgliq", "vol"]
HALO = 2  # x/y halo of the reference's grids (decomposition.F90:449-451)


class DumpError(RuntimeError):
    pass


def parcel_file(rank, num):
    """writebinary.F90:161 and readfromfile_parcelsetup.F90:58: "(A8,i5.5,A1,I5.5,A4)"."""
    return f"parcels_{rank:05d}_{num:05d}.dat"


def grid_file(rank, num):
    """writebinary.F90:262: "(A,i5.5,A1,I5.5,A4)"."""
    return f"grids_{rank:05d}_{num:05d}.dat"


def coordinates(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax):
    """x_coords, y_coords, z_coords of one rank holding the whole domain, as 0-based arrays
    (element i is the reference's coordinate i+1): parcel_interpolation.F90:93-135 with the
    grid spacings of setup_grid.F90:73-75.  The x/y arrays cover the halo of 2 on either side;
    the bottom is added twice, as there (zero for the boxes used here)."""
    dx = (xmax - xmin) / nx
    dy = (ymax - ymin) / ny
    dz = (zmax - zmin) / nz  # (nz_nodes - 1) in the reference's own count
    xs = (-HALO - 1 + 1) * dx + xmin  # (xstart - 1) dx + bottom with xstart = 1 - halo
    ys = (-HALO - 1 + 1) * dy + ymin
    zs = 0 * dz + zmin
    x = np.array([xs + n * dx + xmin for n in range(nx + 2 * HALO)])
    y = np.array([ys + n * dy + ymin for n in range(ny + 2 * HALO)])
    z = np.array([zs + n * dz + zmin for n in range(nz + 1)])
    return x, y, z


def parcel_bounds(x, y, z, nx, ny, nz):
    """The bounds writebinary puts in a parcel dump: x_coords(start), x_coords(end + 1), ...,
    z_coords(start), z_coords(end) (writebinary.F90:170-175), start = halo + 1."""
    return [x[HALO], x[HALO + nx], y[HALO], y[HALO + ny], z[0], z[nz]]


def grid_bounds(x, y, z, nx, ny, nz):
    """The bounds of a grid dump: x_coords(start), x_coords(end), ... (writebinary.F90:280-285)."""
    return [x[HALO], x[HALO + nx - 1], y[HALO], y[HALO + ny - 1], z[0], z[nz]]


def write_parcels(path, time, bounds, fields):
    n = len(fields["x"])
    with open(path, "wb") as fh:
        np.array([time], dtype="<f8").tofile(fh)
        np.asarray(bounds, dtype="<f8").reshape(6).tofile(fh)
        np.array([n], dtype="<i8").tofile(fh)
        for k in FIELDS:
            np.ascontiguousarray(fields.get(k, np.zeros(n)), dtype="<f8").tofile(fh)


def read_parcels(path):
    with open(path, "rb") as fh:
        time = float(np.fromfile(fh, dtype="<f8", count=1)[0])
        bounds = np.fromfile(fh, dtype="<f8", count=6)
        n = int(np.fromfile(fh, dtype="<i8", count=1)[0])
        fields = {k: np.fromfile(fh, dtype="<f8", count=n) for k in FIELDS}
    return time, bounds, fields


def write_grids(path, time, bounds, grids, nx, ny, nz):
    """`grids[name]`: the reference layout data(z, y, x) with the x/y halo, i.e. a C-order array
    of shape (nx + 4, ny + 4, nz + 1); the interior is written z fastest."""
    with open(path, "wb") as fh:
        np.array([time], dtype="<f8").tofile(fh)
        np.asarray(bounds, dtype="<f8").reshape(6).tofile(fh)
        np.array([nx, ny, nz + 1], dtype="<i4").tofile(fh)
        for k in GRID_FIELDS:
            g = np.asarray(grids[k], dtype="<f8")
            assert g.shape == (nx + 2 * HALO, ny + 2 * HALO, nz + 1), (k, g.shape)
            np.ascontiguousarray(g[HALO:HALO + nx, HALO:HALO + ny, :]).tofile(fh)


def read_grids(path):
    """(time, bounds, (nx, ny, nz+1), {name: array of shape (nx, ny, nz+1)}) of a grid dump."""
    with open(path, "rb") as fh:
        time = float(np.fromfile(fh, dtype="<f8", count=1)[0])
        bounds = np.fromfile(fh, dtype="<f8", count=6)
        dims = tuple(int(v) for v in np.fromfile(fh, dtype="<i4", count=3))
        n = dims[0] * dims[1] * dims[2]
        grids = {k: np.fromfile(fh, dtype="<f8", count=n).reshape(dims) for k in GRID_FIELDS}
    return time, bounds, dims, grids


def dump(ctx, directory, num, rank=0, parcels=True, grids=True):
    """writebinary's dump of a context (writebinary.F90:149-299): the parcels, and the grids
    after the grid computation of write_grids_to_file (pmpic_writebinary_grids).  Returns the
    paths written."""
    b = ctx.box
    x, y, z = coordinates(ctx.nx, ctx.ny, ctx.nz, *b)
    out = []
    if parcels:
        path = os.path.join(directory, parcel_file(rank, num))
        write_parcels(path, ctx.time, parcel_bounds(x, y, z, ctx.nx, ctx.ny, ctx.nz),
                      ctx.get_parcels())
        out.append(path)
    if grids:
        ctx.writebinary_grids()
        path = os.path.join(directory, grid_file(rank, num))
        write_grids(path, ctx.time, grid_bounds(x, y, z, ctx.nx, ctx.ny, ctx.nz),
                    {k: ctx.grid(k) for k in GRID_FIELDS}, ctx.nx, ctx.ny, ctx.nz)
        out.append(path)
    return out


def read_restart(directory, num, xmin, xmax, ymin, ymax, log=print):
    """The parcels of dump number `num` that lie in x in [xmin, xmax), y in [ymin, ymax)
    (this domain: x_coords(start) .. x_coords(end + 1)), from every rank's file of that dump.
    Returns (time, fields).  Raises DumpError where the reference stops."""
    log(f" restart from file number {num}")
    nfiles = 0
    while os.path.exists(os.path.join(directory, parcel_file(nfiles, num))):
        nfiles += 1
    if nfiles == 0:
        log(" there is no file to restart from")
        raise DumpError("there is no file to restart from")
    log(f" Found {nfiles} files")
    ranges, total = [], 0
    for rank in range(nfiles):
        with open(os.path.join(directory, parcel_file(rank, num)), "rb") as fh:
            np.fromfile(fh, dtype="<f8", count=1)
            ranges.append(np.fromfile(fh, dtype="<f8", count=6))
            total += int(np.fromfile(fh, dtype="<i8", count=1)[0])
    kept = {k: [] for k in FIELDS}
    time = None
    nread = 0
    for rank in range(nfiles):
        r = ranges[rank]
        # check_ranges (readfromfile_parcelsetup.F90:155-179): the file's x/y range overlaps
        if not (r[0] < xmax and r[1] > xmin and r[2] < ymax and r[3] > ymin):
            continue
        path = os.path.join(directory, parcel_file(rank, num))
        log(f" rank 0 opening file {parcel_file(rank, num)}")
        t, _, fields = read_parcels(path)
        time = t
        mask = ((fields["x"] < xmax) & (fields["x"] >= xmin) & (fields["y"] < ymax) &
                (fields["y"] >= ymin))
        for k in FIELDS:
            kept[k].append(fields[k][mask])
        nread += int(np.count_nonzero(mask))
    if nread != total:
        log(" Error - incorrect parcel count read in")
        raise DumpError(f"incorrect parcel count read in: {nread} of {total}")
    log(" All files read in successfully")
    log(f" Total parcel count read in = {nread}")
    return time, {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in kept.items()}
