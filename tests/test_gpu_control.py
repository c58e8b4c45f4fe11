"""GPU: the host run loop around the library step (model_core/src/monc.F90:211-230).

run_model over nn_timesteps=N performs exactly N library steps (the integration runs on the
terminating iteration too) and leaves the model time where the reference leaves it: one dtm
behind the time after N steps, because the final iteration does not advance it.
"""
import numpy as np
import pytest

from pmpic_amd import control
from tests.cases import box, thermal_case

pytestmark = pytest.mark.gpu


def _ctx(ic, shape, precision):
    from pmpic_amd.lib import Context
    nx, ny, nz = shape
    b = box(nx, ny, nz)
    g = Context(nx, ny, nz, b["xmin"], b["xmax"], b["ymin"], b["ymax"], b["zmin"], b["zmax"],
                3 * len(ic["x"]) + 64, dtm=0.01, dtmax=0.01, timestep_prefactor=0.5,
                precision=precision)
    g.set_parcels(ic)
    return g


@pytest.mark.parametrize("precision", [0, 1])
def test_run_model_matches_plain_steps(tmp_path, precision):
    shape = (12, 10, 8)
    ic = thermal_case(*shape, 2)
    a = _ctx(ic, shape, precision)
    b = _ctx(ic, shape, precision)
    logs = []
    st = control.run_model(a, {"nn_timesteps": 3, "display_synopsis_frequency": 1},
                           log=logs.append, parcels_dat=str(tmp_path / "parcels.dat"))
    times = []
    for _ in range(3):
        times.append(b.time)
        b.step(1)
    assert st.timestep == 3
    assert a.n == b.n
    pa, pb = a.get_parcels(), b.get_parcels()
    for k in pa:
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    # the time after the last step is not kept: the run ends at the start time of step 3
    assert a.time == times[2]
    assert st.time == times[2]
    counts = np.loadtxt(tmp_path / "parcels.dat")
    assert counts.shape == (3,)
    assert int(counts[-1]) == b.n
    assert any("Model run complete" in s for s in logs)
    a.close()
    b.close()
