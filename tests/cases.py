"""Seeded small workloads shared by the CPU and GPU parity tests."""
import numpy as np

from pmpic_amd.config import fortran_real
from pmpic_amd.initial import thermal

# the box of the thermal workloads as config.mcf strings, parsed the reference's way
LX = fortran_real("6.28318531")
LZ = fortran_real("3.14159265")


def box(nx, ny, nz):
    return dict(nx=nx, ny=ny, nz=nz, xmin=0.0, xmax=LX, ymin=0.0, ymax=LX, zmin=0.0, zmax=LZ)


def thermal_case(nx, ny, nz, ppc_dir, seed=1, vort=0.0, shrink=0.0, thin=0.0, n2=1.0,
                 near_split=0.0):
    """Jittered-lattice thermal; optional random vorticity (amplitude `vort`), a fraction
    `shrink` of parcels shrunk below the mixing threshold, a fraction `thin` of cells
    emptied to two parcels and a fraction `near_split` of parcels given a stretch just below
    the split threshold, so splitting and both mixing paths are exercised."""
    ic = thermal(nx, ny, nz, LX, LX, LZ, ppc_dir, seed=0x504D504943000000 + seed, n2=n2)
    n = len(ic["x"])
    rng = np.random.default_rng(seed)
    if vort:
        for k in ("p", "q", "r"):
            ic[k] = vort * rng.uniform(-1, 1, n)
    ppc = ppc_dir ** 3
    if thin:
        ncell = nx * ny * nz
        cells = rng.choice(ncell, size=max(1, int(thin * ncell)), replace=False)
        drop = np.zeros(n, dtype=bool)
        for c in cells:
            drop[c * ppc + 2:(c + 1) * ppc] = True
        for k in list(ic):
            ic[k] = ic[k][~drop]
        n = len(ic["x"])
    if shrink:
        dx, dy, dz = LX / nx, LX / ny, LZ / nz
        vmin = dx * dy * dz / 6. / 6. / 6.
        idx = rng.choice(n, size=max(1, int(shrink * n)), replace=False)
        ic["vol"] = ic["vol"].copy()
        ic["vol"][idx] = 0.5 * vmin
    if near_split:
        idx = rng.choice(n, size=max(1, int(near_split * n)), replace=False)
        ic["stretch"] = np.zeros(n)
        ic["stretch"][idx] = rng.uniform(3.95, 4.0, len(idx))
    ic["tag"] = np.arange(n, dtype=np.float64)
    return ic


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


def crowded_case(nx, ny, nz, seed=3, extra=(24, 25, 36, 700)):
    """The vortical thermal (8 parcels per cell) with `extra` parcels appended in as many random
    cells, so those cells hold 8 + extra parcels: exactly 32, one past 32, a few past 32, and
    more than one workgroup's LDS rows (512), the cell-binning capacity edges."""
    ic = thermal_case(nx, ny, nz, 2, seed=seed, vort=1.0)
    rng = np.random.default_rng(seed)
    dx, dy, dz = LX / nx, LX / ny, LZ / nz
    cells = rng.choice(nx * ny * nz, size=len(extra), replace=False)
    add = {k: [] for k in ic}
    for m, c in zip(extra, cells):
        i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
        add["x"].append((i + rng.uniform(0.02, 0.98, m)) * dx)
        add["y"].append((j + rng.uniform(0.02, 0.98, m)) * dy)
        add["z"].append((k + rng.uniform(0.02, 0.98, m)) * dz)
        for f in ("p", "q", "r"):
            add[f].append(rng.uniform(-1, 1, m))
        add["b"].append(rng.uniform(0.0, 0.5, m))
        add["h"].append(rng.uniform(0.0, 0.4, m))
        add["vol"].append(np.full(m, dx * dy * dz / 8.0))
    n0 = len(ic["x"])
    for k in ic:
        if k == "tag":
            continue
        ic[k] = np.concatenate([ic[k]] + add[k])
    ic["tag"] = np.arange(len(ic["x"]), dtype=np.float64)
    assert len(ic["x"]) == n0 + sum(extra)
    return ic
