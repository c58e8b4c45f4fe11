"""Error paths of the C-ABI on the GPU, in both numeric forms: each condition the reference
stops on (parcel_interpolation.F90:213-238 out of box, parcel_splitting.F90:133-137 maxparcels)
comes back as its status code, forced from the host; an empty parcel set steps cleanly."""
import numpy as np
import pytest

from tests.cases import LZ, box, thermal_case

pytestmark = pytest.mark.gpu


def ctx(shape, cap, precision):
    from pmpic_amd.lib import Context
    b = box(*shape)
    return Context(*shape, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"], cap,
                   precision=precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_parcel_above_the_box_is_reported(precision):
    from pmpic_amd.lib import PmpicError
    shape = (8, 8, 6)
    ic = thermal_case(*shape, 2)
    ic["z"] = ic["z"].copy()
    ic["z"][5] = LZ + 0.5  # above zmax by more than a cell
    g = ctx(shape, 3 * len(ic["x"]), precision)
    with pytest.raises(PmpicError) as e:
        g.set_parcels(ic)
        g.step()
    assert e.value.status == 3  # PMPIC_ERR_OUT_OF_BOX_Z


@pytest.mark.parametrize("precision", [0, 1])
def test_split_beyond_capacity_is_reported(precision):
    from pmpic_amd.lib import PmpicError
    shape = (8, 8, 6)
    ic = thermal_case(*shape, 2, vort=2.0)
    n = len(ic["x"])
    ic["stretch"] = np.full(n, 4.5)  # every parcel splits at the first stage 4
    g = ctx(shape, n + 10, precision)
    g.set_parcels(ic)
    with pytest.raises(PmpicError) as e:
        g.step()
    assert e.value.status == 5  # PMPIC_ERR_MAXPARCELS


@pytest.mark.parametrize("precision", [0, 1])
def test_empty_parcel_set_steps(precision):
    from pmpic_amd.lib import Context
    shape = (8, 8, 6)
    ncell = 8 * 8 * 6
    b = box(*shape)
    empty = {k: np.zeros(0) for k in ("x", "y", "z")}
    # without mixing nothing happens: zero grids, no parcels
    g = Context(*shape, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"],
                2 * ncell, mixing=False, precision=precision)
    g.set_parcels(empty)
    g.step(3)
    assert g.n == 0
    assert np.all(g.grid("u") == 0.0)
    # with mixing every cell is sparse, so the first stage 4 adds one parcel per cell
    # (parcel_mixing.F90:160-200)
    g = ctx(shape, 2 * ncell, precision)
    g.set_parcels(empty)
    g.step(1)
    assert g.n == ncell
    assert g.last_counts() == (0, 0, ncell)


@pytest.mark.parametrize("nx,ny", [(14, 8), (8, 1024)])
def test_unsupported_transform_length_is_refused(nx, ny):
    """A horizontal size with a prime factor other than 2, 3, 5 (the reference's FFTE check,
    pencilfft.F90:114-117) or longer than the 512 points a transform workgroup stages is
    refused at creation with PMPIC_ERR_FFT_SIZE, before anything is allocated."""
    from pmpic_amd.lib import PmpicError
    with pytest.raises(PmpicError) as e:
        ctx((nx, ny, 4), 64, 0)
    assert e.value.status == 7  # PMPIC_ERR_FFT_SIZE
