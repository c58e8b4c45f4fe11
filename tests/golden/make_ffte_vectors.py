"""Writes ffte_semispectral.npz: inputs and outputs of the reference's own FFTE.

The FFTE library and its ffte_mod wrappers are built from the reference's sources into
oracle/_ref (`make -C oracle ref`, needs the reference tree); this script runs only where that
build exists.  It applies the 1-D transforms the way pencilfft.F90:151-212 composes them for one
field f[k, y, x]: r2c along y (ffte_mod.F90:66-95), the (Re, Im) parts interleaved as real rows
(pencilfft.F90:838-899), then r2c along x of each real row.  The inverse fixture applies c2r
along x, then rebuilds the complex y-spectra and applies c2r along y (ffte_mod.F90:97-124).

Stored per case (shape nx, ny, nz): the real field f [nz+1, ny, nx], its semi-spectral array
S [nz+1, (ny/2+1)*2, (nx/2+1)*2], a random semi-spectral array T of the same shape, and the
field FFTE's inverse makes of T.  Run: python tests/golden/make_ffte_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle  # noqa: E402

CASES = [(12, 10, 4), (15, 6, 3), (16, 16, 2)]


def forward(ref, f):
    nz1, ny, nx = f.shape
    my, mx = ny // 2 + 1, nx // 2 + 1
    S = np.empty((nz1, 2 * my, 2 * mx))
    for k in range(nz1):
        rows = np.empty((2 * my, nx))
        for x in range(nx):
            z = ref.r2c(f[k, :, x])
            rows[0::2, x] = z.real
            rows[1::2, x] = z.imag
        for jj in range(2 * my):
            z = ref.r2c(rows[jj])
            S[k, jj, 0::2] = z.real
            S[k, jj, 1::2] = z.imag
    return S


def backward(ref, S, nx, ny):
    nz1 = S.shape[0]
    my = ny // 2 + 1
    f = np.empty((nz1, ny, nx))
    for k in range(nz1):
        rows = np.empty((2 * my, nx))
        for jj in range(2 * my):
            rows[jj] = ref.c2r(S[k, jj, 0::2] + 1j * S[k, jj, 1::2], nx)
        for x in range(nx):
            f[k, :, x] = ref.c2r(rows[0::2, x] + 1j * rows[1::2, x], ny)
    return f


def main():
    if not oracle.RefFFTE.available():
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` with the reference tree present")
    ref = oracle.RefFFTE()
    out = {}
    rng = np.random.default_rng(20260)
    for nx, ny, nz in CASES:
        tag = f"{nx}x{ny}x{nz}"
        f = rng.standard_normal((nz + 1, ny, nx))
        T = rng.standard_normal((nz + 1, (ny // 2 + 1) * 2, (nx // 2 + 1) * 2))
        out[f"f_{tag}"] = f
        out[f"S_{tag}"] = forward(ref, f)
        out[f"T_{tag}"] = T
        out[f"g_{tag}"] = backward(ref, T, nx, ny)
    np.savez_compressed(os.path.join(HERE, "ffte_semispectral.npz"), **out)
    print("wrote", os.path.join(HERE, "ffte_semispectral.npz"), sorted(out))


if __name__ == "__main__":
    main()
