"""The semi-spectral transforms against golden vectors of the reference's own FFTE.

tests/golden/ffte_semispectral.npz holds, for three grids (radix 2, 3 and 5 lengths), a field
and the semi-spectral array the reference's FFTE (built from its sources into oracle/_ref)
composes from it the way pencilfft.F90:151-212 does, and a random semi-spectral array with the
field FFTE's inverse makes of it (tests/golden/make_ffte_vectors.py wrote them).  The oracle is
held to them on the CPU, the HIP transforms on the GPU; neither needs oracle/_ref at run time.
"""
import os

import numpy as np
import pytest

from tests.cases import box, rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "ffte_semispectral.npz")
CASES = [(12, 10, 4), (15, 6, 3), (16, 16, 2)]


def _case(shape):
    nx, ny, nz = shape
    d = np.load(GOLDEN)
    tag = f"{nx}x{ny}x{nz}"
    return d[f"f_{tag}"], d[f"S_{tag}"], d[f"T_{tag}"], d[f"g_{tag}"]


def _grid(f, nx, ny):
    """[k, y, x] interior -> (x+4, y+4, z) with zero halos."""
    g = np.zeros((nx + 4, ny + 4, f.shape[0]))
    g[2:-2, 2:-2, :] = np.transpose(f, (2, 1, 0))
    return g


def _spec(S):
    """[k, jj, ii] -> (ii, jj, k)."""
    return np.ascontiguousarray(np.transpose(S, (2, 1, 0)))


@pytest.mark.parametrize("shape", CASES)
def test_oracle_matches_reference_ffte_vectors(oracle_mod, shape):
    nx, ny, nz = shape
    f, S, T, g = _case(shape)
    b = box(nx, ny, nz)
    o = oracle_mod.Oracle(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"],
                          b["zmax"], 1000, 0.01, 0.01, 0.5)
    assert rel_err(o.fft_forward(_grid(f, nx, ny)), _spec(S)) < 1e-13
    back = o.fft_backward(_spec(T))[2:-2, 2:-2, :]
    assert rel_err(back, np.transpose(g, (2, 1, 0))) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("shape", CASES)
def test_hip_transforms_match_reference_ffte_vectors(shape):
    from pmpic_amd.lib import Context
    nx, ny, nz = shape
    f, S, T, g = _case(shape)
    b = box(nx, ny, nz)
    ctx = Context(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"],
                  1000)
    got = ctx.fft_forward(_grid(f, nx, ny))
    back = ctx.fft_backward(_spec(T))[2:-2, 2:-2, :]
    ctx.close()
    assert rel_err(got, _spec(S)) < 1e-12
    assert rel_err(back, np.transpose(g, (2, 1, 0))) < 1e-12
