"""Launch-configuration autotuning of the FP32-bulk form (pmpic_autotune).

Every configuration the tuner tries (slot capacity 16 or 32, binning fused into the RK stages
or separate, RK stages 1-3 as a HIP graph or eager, RK workgroups of 128 or 256) gives the same
bits, so the tuning steps are ordinary model steps: a tuned context ends bit-identical to an
untuned one stepped to the same model time (parcel count, every parcel field, the grids), and
stays so over further steps, through splitting and mixing.  A second context of the same grid
size and occupancy class takes the stored choice without stepping."""
import numpy as np
import pytest

from tests.cases import thermal_case
from tests.test_gpu_bulk32 import make

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "p", "q", "r", "b", "h", "vol", "stretch", "tag")
GRIDS = ("u", "v", "w", "p", "q", "r", "b", "vol", "hg")


def _same(a, b, what):
    assert a.n == b.n, what
    assert a.time == b.time, what
    pa, pb = a.get_parcels(list(FIELDS)), b.get_parcels(list(FIELDS))
    for k in FIELDS:
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=f"{what}: {k}")
    for name in GRIDS:
        np.testing.assert_array_equal(a.grid(name), b.grid(name), err_msg=f"{what}: {name}")


def test_autotune_steps_match_untuned_and_choice_is_kept(monkeypatch):
    monkeypatch.delenv("PMPIC_GRAPH", raising=False)
    monkeypatch.delenv("PMPIC_SLOT_CAPACITY", raising=False)
    monkeypatch.delenv("PMPIC_FUSED_BINNING", raising=False)
    shape = (24, 20, 12)  # a size no other test tunes
    ic = thermal_case(*shape, 2, vort=3.0, shrink=0.05, thin=0.1, near_split=0.05)
    tuned = make(shape, 1, ic)
    plain = make(shape, 1, ic)
    before = tuned.launch_config()
    assert before["rk_workgroup"] == 256 and not before["graph"]
    got = tuned.autotune(steps=1)
    assert got["measured"] and got["ms_per_step"] > 0
    assert got["slot_capacity"] in (16, 32) and got["rk_workgroup"] in (128, 256)
    cfg = {k: got[k] for k in ("slot_capacity", "fused_binning", "graph", "rk_workgroup")}
    assert tuned.launch_config() == cfg
    # 16 configurations (two capacities), each run twice, of one settling and one timed step
    steps = 0
    while plain.time < tuned.time and steps < 128:
        plain.step()
        steps += 1
    assert steps == 64
    _same(tuned, plain, "after tuning")
    for s in range(3):
        tuned.step()
        plain.step()
        _same(tuned, plain, f"tuned step {s}")

    again = make(shape, 1, ic)
    t0 = again.time
    got2 = again.autotune(steps=1)
    assert not got2["measured"] and again.time == t0
    assert {k: got2[k] for k in cfg} == cfg
    assert again.launch_config() == cfg


def test_autotune_refuses_fp64_and_empty():
    from pmpic_amd.lib import PmpicError
    shape = (8, 8, 8)
    ic = thermal_case(*shape, 2)
    g64 = make(shape, 0, ic)
    with pytest.raises(PmpicError):
        g64.autotune(steps=1)
    g32 = make(shape, 1, ic)
    with pytest.raises(PmpicError):
        g32.autotune(steps=0)


def test_set_launch_config_and_refusals():
    from pmpic_amd.lib import PmpicError
    shape = (16, 16, 8)
    g = make(shape, 1, thermal_case(*shape, 2))
    g.set_launch_config(32, fused_binning=True, graph=False, rk_workgroup=128)
    assert g.launch_config() == {"slot_capacity": 32, "fused_binning": True, "graph": False,
                                 "rk_workgroup": 128}
    g.set_launch_config(0)  # the mean rule: 16 entries at 8 parcels per cell
    assert g.launch_config()["slot_capacity"] == 16
    for cap, wg in ((8, 256), (64, 256), (16, 512), (16, 64)):
        with pytest.raises(PmpicError):
            g.set_launch_config(cap, rk_workgroup=wg)


def test_tune_together_matches_untuned_members():
    from pmpic_amd.ensemble import tune_together
    from pmpic_amd.lib import step_concurrent
    cases = [((12, 10, 8), 2, 1), ((16, 16, 8), 2, 2), ((8, 8, 8), 3, 3)]
    ics = [thermal_case(*sh, ppcd, seed=sd, vort=1.5, shrink=0.02, thin=0.05)
           for sh, ppcd, sd in cases]
    tuned = [make(sh, 1, ic) for (sh, _, _), ic in zip(cases, ics)]
    plain = [make(sh, 1, ic) for (sh, _, _), ic in zip(cases, ics)]
    got = tune_together(tuned, steps=1)
    assert set(got["by_slot_capacity"]) == {16, 32}
    for t, p in zip(tuned, plain):
        while p.time < t.time:
            p.step()
        _same(t, p, "after tuning")
    step_concurrent(tuned, 2)
    for p in plain:
        p.step(2)
    for m, (t, p) in enumerate(zip(tuned, plain)):
        _same(t, p, f"member {m}")
