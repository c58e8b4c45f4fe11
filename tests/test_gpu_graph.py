"""HIP-graph replay of RK stages 1-3 in the FP32-bulk form's pmpic_step.

A context whose steps launch stages 1-3 as one captured graph (PMPIC_GRAPH=1 at creation) ends
every step bit-identical to a context launching every kernel eagerly (the default): parcel count, time, every
parcel field and the grids, compared after each step, with the eager stage-4 work (splitting,
mixing) and host transfers between the replays.  Cases: splitting and mixing changing the count
from step to step, a re-capture when the count drops out of the captured range, the slot
capacities 16 and 32, the tuned capacity, the dense binning, the binning done by the RK stages,
and members stepped concurrently.
The graph is shown in use by its capture and launch counts."""
import numpy as np
import pytest

from tests.cases import crowded_case, thermal_case
from tests.test_gpu_bulk32 import make

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "p", "q", "r", "b", "h", "vol", "stretch", "tag")
GRIDS = ("u", "v", "w", "p", "q", "r", "b", "vol", "hg")


def _pair(monkeypatch, shape, ic, cap_factor=3):
    monkeypatch.delenv("PMPIC_GRAPH", raising=False)
    eager = make(shape, 1, ic, cap_factor)
    monkeypatch.setenv("PMPIC_GRAPH", "1")
    graph = make(shape, 1, ic, cap_factor)
    monkeypatch.delenv("PMPIC_GRAPH")
    return eager, graph


def _same(a, b, what):
    assert a.n == b.n, what
    assert a.time == b.time, what
    assert a.last_counts() == b.last_counts(), what
    pa, pb = a.get_parcels(list(FIELDS)), b.get_parcels(list(FIELDS))
    for k in FIELDS:
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=f"{what}: {k}")
    for name in GRIDS:
        np.testing.assert_array_equal(a.grid(name), b.grid(name), err_msg=f"{what}: {name}")


def test_graph_steps_match_eager_through_splits_and_mixing(monkeypatch):
    shape = (16, 16, 8)
    ic = thermal_case(*shape, 2, vort=3.0, shrink=0.05, thin=0.1, near_split=0.05)
    eager, graph = _pair(monkeypatch, shape, ic)
    steps = 8
    counts = set()
    for s in range(steps):
        eager.step()
        graph.step()
        _same(eager, graph, f"step {s}")
        counts.add(graph.n)
    assert len(counts) > 1  # the count changed between replays
    assert eager.graph_stats() == (0, 0)
    captures, launches = graph.graph_stats()
    # the first step runs eagerly; every later one launches the graph
    assert launches == steps - 1 and 1 <= captures <= launches


def test_graph_recaptured_when_count_drops(monkeypatch):
    shape = (32, 32, 16)  # 131072 parcels: the captured bound is mostly the count itself
    ic = thermal_case(*shape, 2, vort=1.5, shrink=0.02, thin=0.05)
    eager, graph = _pair(monkeypatch, shape, ic)
    for s in range(3):
        eager.step()
        graph.step()
        _same(eager, graph, f"step {s}")
    cap0, _ = graph.graph_stats()
    # keep every third parcel: the count falls far below the captured bound
    p = graph.get_parcels()
    sub = {k: v[::3].copy() for k, v in p.items()}
    eager.set_parcels(sub)
    graph.set_parcels(sub)
    eager.set_time(graph.time, graph.dtm)
    graph.set_time(graph.time, graph.dtm)
    for s in range(3):
        eager.step()
        graph.step()
        _same(eager, graph, f"after the cut, step {s}")
    cap1, launches = graph.graph_stats()
    assert cap1 == cap0 + 1 and launches == 5


@pytest.mark.parametrize("slots", ["16", "32", "tune", "dense", "fused"])
def test_graph_matches_eager_at_each_binning(monkeypatch, slots):
    if slots == "fused":  # RK stages 1-3 bin for the next par2grid (as on large grids)
        monkeypatch.setenv("PMPIC_FUSED_BINNING", "1")
    else:
        monkeypatch.setenv("PMPIC_SLOT_CAPACITY", "0" if slots == "dense" else slots)
    shape = (12, 10, 8)
    ic = crowded_case(*shape)
    eager, graph = _pair(monkeypatch, shape, ic)
    for s in range(4):
        eager.step()
        graph.step()
        _same(eager, graph, f"{slots}, step {s}")
    assert graph.graph_stats()[1] == 3


def test_graph_members_stepped_concurrently(monkeypatch):
    from pmpic_amd.lib import step_concurrent
    shapes = [(12, 10, 8), (16, 16, 8), (8, 8, 8)]
    pairs = [_pair(monkeypatch, sh, thermal_case(*sh, 2, seed=m + 1, vort=1.5, shrink=0.02,
                                                 thin=0.05))
             for m, sh in enumerate(shapes)]
    for s in range(3):
        step_concurrent([g for _, g in pairs], 1)
        for e, _ in pairs:
            e.step()
        for m, (e, g) in enumerate(pairs):
            _same(e, g, f"member {m}, step {s}")
    assert all(g.graph_stats()[1] == 2 for _, g in pairs)
