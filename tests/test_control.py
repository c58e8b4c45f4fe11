"""Host-side run control (termination check, model synopsis, the timestep loop) against the
reference's rules, on a stand-in context (no GPU).  The step-number rule is the one
components/terminationcheck/test/test_terminationcheck.F90:11-30 checks: the run continues
while mod(timestep, nn_timesteps) /= 0."""
import random

import pytest

from pmpic_amd import control


class FakeCtx:
    """Steps advance time by dtm and grow the parcel count by one."""

    def __init__(self, dtm=0.25, n=100):
        self.time, self.dtm, self.n, self.steps = 0.0, dtm, n, 0

    def step(self, k=1):
        for _ in range(k):
            self.steps += 1
            self.n += 1
            self.time += self.dtm

    def set_time(self, t, dtm):
        self.time, self.dtm = t, dtm


def test_timestep_rule_random():
    rng = random.Random(7)
    for _ in range(1000):
        ts = rng.randrange(1, 10001)
        nn = rng.randrange(1, 10)
        st = control.RunState(timestep=ts, time=0.0)
        control.TerminationCheck({"nn_timesteps": nn, "termination_time": 1e9})(st)
        assert st.continue_timestep == (ts % nn != 0)


def test_nn_timesteps_loop(tmp_path):
    logs = []
    ctx = FakeCtx(dtm=0.1)
    st = control.run_model(ctx, {"nn_timesteps": 5, "termination_time": 100.0,
                                 "display_synopsis_frequency": 2},
                           log=logs.append, parcels_dat=str(tmp_path / "parcels.dat"))
    # five integrations, the time advanced after the first four only
    assert ctx.steps == 5 and st.timestep == 5
    assert ctx.time == pytest.approx(0.4) and st.time == pytest.approx(0.4)
    assert st.termination_reason == control.TIMESTEP_TERMINATION_REASON
    assert "[INFO] Number of completed timesteps 2" in logs
    assert "[INFO] Number of completed timesteps 4" in logs
    assert logs[-1] == "[INFO] Model run complete due to timestep completion, model time is 0.40"
    lines = (tmp_path / "parcels.dat").read_text().split("\n")
    assert [int(x) for x in lines if x.strip()] == [101, 102, 103, 104, 105]


def test_termination_time_loop(tmp_path):
    logs = []
    ctx = FakeCtx(dtm=0.25)
    st = control.run_model(ctx, {"termination_time": 1.0, "display_synopsis_frequency": 10},
                           log=logs.append, parcels_dat=None)
    assert ctx.steps == 5 and st.time == 1.0 and ctx.time == 1.0
    assert st.termination_reason == control.TIME_TERMINATION_REASON
    assert logs[-1] == ("[INFO] Model run complete due to model time 1.0 exceeding limit of "
                        "1.0")


def test_synopsis_lines():
    logs = []
    clock = iter([0.0, 0.0, 0.5, 0.5, 1.25, 1.25, 2.0, 2.0] + [2.0] * 20)
    ctx = FakeCtx(dtm=0.01)
    control.run_model(ctx, {"nn_timesteps": 4, "display_synopsis_frequency": 2,
                            "termination_time": 9.0}, log=logs.append, parcels_dat=None,
                      wtime=lambda: next(clock))
    i = logs.index("[INFO] Number of completed timesteps 2")
    assert logs[i + 1] == "[INFO] Completed 2 timesteps in 500ms"
    j = logs.index("[INFO] Number of completed timesteps 4")
    assert logs[j + 1] == "[INFO] Completed 3 timesteps in 750ms"  # (4 - 2) + 1, as the reference counts
    assert logs[i + 2] == "[INFO] Model time 0.010 seconds; dtm=0.010"


def test_messages_file(tmp_path):
    msg = tmp_path / "messages"
    msg.write_text("terminate\n")
    ctx = FakeCtx()
    st = control.run_model(ctx, {"check_msg_frequency": 3, "msg_filename": str(msg),
                                 "termination_time": 99.0}, log=lambda s: None,
                           parcels_dat=None)
    assert st.timestep == 3 and ctx.steps == 3
    assert st.termination_reason == control.MESSAGE_TERMINATION_REASON


def test_walltime():
    t = [0.0]

    def wtime():
        t[0] += 10.0
        return t[0]

    ctx = FakeCtx()
    st = control.run_model(ctx, {"walltime_limit": "00:01:00", "check_walltime_frequency": 2,
                                 "termination_time": 99.0}, log=lambda s: None,
                           parcels_dat=None, wtime=wtime)
    assert st.termination_reason == control.WALLTIME_TERMINATION_REASON
    assert st.timestep % 2 == 0


@pytest.mark.parametrize("text,secs,warn", [("01:30:00", 5400, False), ("1:75:80", 3600, True),
                                            ("10", None, True), ("none", None, False),
                                            ("0:0:59", 59, False)])
def test_walltime_parse(text, secs, warn):
    logs = []
    assert control.parse_walltime(text, logs.append) == secs
    assert bool(logs) == warn


@pytest.mark.parametrize("x,dp,want", [(0.1, 5, "0.10"), (0.01, 5, "0.010"), (1.0, 5, "1.0"),
                                       (1.23456789, 5, "1.23456"), (12.5, 2, "12.50"),
                                       (1e-7, 5, "0.00000"), (2.0 / 3.0, 2, "0.66")])
def test_real_strings(x, dp, want):
    assert control.real_str(x, dp) == want


class DumpCtx(FakeCtx):
    """A stand-in that also answers what a dump reads."""

    nx, ny, nz = 4, 4, 2
    box = (0.0, 4.0, 0.0, 4.0, 0.0, 2.0)

    def get_parcels(self, names=None):
        import numpy as np
        return {"x": np.array([0.5, 1.5]), "y": np.array([0.5, 0.5]), "z": np.array([1.0, 1.0])}

    def writebinary_grids(self):
        pass

    def grid(self, name):
        import numpy as np
        return np.zeros((self.nx + 4, self.ny + 4, self.nz + 1))


@pytest.mark.parametrize("mode,opts,parcels,grids", [
    ("step", {"parcel_step_write_frequency": 2, "grid_step_write_frequency": 3},
     [0, 2, 4], [0, 3]),
    ("time", {"parcel_time_write_frequency": 0.5, "grid_time_write_frequency": 0.25},
     [0, 2, 4], [0, 1, 2, 3, 4]),
    ("none", {}, [], []),
])
def test_writebinary_schedule(tmp_path, mode, opts, parcels, grids):
    """writebinary.F90:84-135 in the start group: the dump number is the iteration count; step
    mode writes every n iterations, time mode once the model time reaches the next multiple."""
    import os
    logs = []
    w = control.WriteBinary(dict(opts, writebinary_mode=f'"{mode}"'), str(tmp_path), logs.append)
    ctx = DumpCtx(dtm=0.25)
    control.run_model(ctx, {"nn_timesteps": 5, "termination_time": 100.0,
                            "display_synopsis_frequency": 100}, log=logs.append,
                      parcels_dat=None, writer=w)
    files = sorted(os.listdir(tmp_path))
    assert [f for f in files if f.startswith("parcels_")] == \
        [f"parcels_00000_{n:05d}.dat" for n in parcels]
    assert [f for f in files if f.startswith("grids_")] == [f"grids_00000_{n:05d}.dat" for n in grids]
    assert f" Written {len(parcels)} parcel dumps" in logs


def test_writebinary_rejects_unknown_mode():
    with pytest.raises(ValueError):
        control.WriteBinary({"writebinary_mode": "often"}, ".", lambda s: None)
