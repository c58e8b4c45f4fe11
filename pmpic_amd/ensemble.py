"""The batched-ensemble workload: many independent moist thermals on uneven grids.

Member m of an ensemble of M takes the grid sizes[m % 4] cubed from (32, 48, 64, 96) cells, 8
parcels per cell for the first half of each run of eight members and 27 for the second (a
jittered 2x2x2 or 3x3x3 lattice), a thermal radius drawn from U(0.5, 1.0) and a buoyancy
amplitude from U(0.5, 2.0) (numpy's PCG64 seeded with 0x504D5049, two draws per member in
member order), and its own SplitMix64 seed 0x504D504943000000 + 3 + m for the lattice jitter and
noise.  The box is the thermal's 2pi x 2pi x pi with the reals parsed the reference's way.
"""
import numpy as np

from .config import fortran_real
from .initial import SEED_BASE, thermal

SIZES = (32, 48, 64, 96)


def member_specs(count):
    """[(n, ppc_dir, radius, amplitude, seed)] for an ensemble of `count` members."""
    rng = np.random.default_rng(0x504D5049)
    out = []
    for m in range(count):
        n = SIZES[m % len(SIZES)]
        ppcd = 2 if (m // len(SIZES)) % 2 == 0 else 3
        radius = float(rng.uniform(0.5, 1.0))
        amplitude = float(rng.uniform(0.5, 2.0))
        out.append((n, ppcd, radius, amplitude, SEED_BASE + 3 + m))
    return out


def box():
    return fortran_real("6.28318531"), fortran_real("3.14159265")


def member_parcels(spec):
    """The initial parcels of one member."""
    n, ppcd, radius, amplitude, seed = spec
    lx, lz = box()
    return thermal(n, n, n, lx, lx, lz, ppcd, seed=seed, radius=radius, amplitude=amplitude)


def member_context(spec, ic=None, cap_factor=2.0):
    """An FP32-bulk context holding member `spec`'s initial parcels."""
    from .lib import Context
    n = spec[0]
    lx, lz = box()
    if ic is None:
        ic = member_parcels(spec)
    c = Context(n, n, n, 0.0, lx, 0.0, lx, 0.0, lz, int(cap_factor * len(ic["x"])) + 64,
                precision=1)
    c.set_parcels(ic)
    return c


def tune_together(ctxs, steps=2):
    """Pick launch configurations for members stepped together (pmpic_step_concurrent).

    pmpic_autotune times one context alone, which does not predict the members' speed when
    they share the GPU; this tuner times the whole ensemble instead.  The candidates are the
    binning fused into the RK stages or not, RK stages 1-3 as a graph or not, and RK workgroups
    of 128 or 256, with each member's slot capacity left to the mean-occupancy rule.  First one
    configuration for every member, then, from that start, one per occupancy class (the
    capacity the rule gives) while the other classes keep theirs.  Each trial runs one settling
    step and `steps` timed steps of the whole ensemble, and each list of trials runs forward and
    then backward, a trial scored by the mean of its two timings, so that the drift of the step
    time with the model state cancels.  The steps are model steps: tune copies of the members
    where their states matter.  Returns {class: configuration} and the ensemble ms per step of
    the choice."""
    import itertools
    import time

    from .lib import step_concurrent
    cands = [dict(fused_binning=f, graph=g, rk_workgroup=w)
             for f, g, w in itertools.product((False, True), (False, True), (128, 256))]
    classes = {}
    for c in ctxs:
        classes.setdefault(c.launch_config()["slot_capacity"], []).append(c)

    def run(assign):
        for cls, members in classes.items():
            for c in members:
                c.set_launch_config(0, **assign[cls])
        step_concurrent(ctxs, 1)
        t0 = time.perf_counter()
        step_concurrent(ctxs, steps)
        return (time.perf_counter() - t0) * 1e3 / steps

    def best_of(trials):
        ms = [0.0] * len(trials)
        for i in list(range(len(trials))) + list(reversed(range(len(trials)))):
            ms[i] += run(trials[i]) / 2
        i = min(range(len(trials)), key=ms.__getitem__)
        return trials[i], ms[i]

    choice, best_ms = best_of([{cls: cand for cls in classes} for cand in cands])
    if len(classes) > 1:
        for cls in classes:
            trials = []
            for cand in cands:
                trial = dict(choice)
                trial[cls] = cand
                trials.append(trial)
            choice, best_ms = best_of(trials)
    for cls, members in classes.items():
        for c in members:
            c.set_launch_config(0, **choice[cls])
    return {"by_slot_capacity": choice, "ms_per_step": best_ms}
