"""Host-side run control: the reference's termination check and model synopsis around the
timestep loop.

* ``TerminationCheck`` follows components/terminationcheck/src/terminationcheck.F90:38-121:
  ``nn_timesteps`` (stop when the step number is a multiple of it, if positive),
  ``termination_time`` (stop once the model time reaches it), ``walltime_limit``
  (``hh:mm:ss``, checked every ``check_walltime_frequency`` steps) and the messages file
  ``msg_filename`` (checked every ``check_msg_frequency`` steps; a first line ``terminate``
  stops the run).
* ``ModelSynopsis`` follows components/modelsynopsis/src/modelsynopsis.F90:38-111: every
  ``display_synopsis_frequency`` steps the step count, the wall time of the interval and the
  model time and dtm are logged; each step appends the parcel count to ``parcels.dat``; the
  finalisation line names the termination reason.
* ``run_model`` is the loop of model_core/src/monc.F90:211-230 with the groups of
  global_config:30-43: the start group (termination check), the integrator group (one full
  step of the library: the four RK stages of vort2vel, vorticity_tendency, rk4_integrator,
  parcel_splitting, parcel_mixing) and the last group (synopsis) run every iteration, the
  integration included on the iteration whose check ends the run; the step number and the
  model time advance only when the run continues.

Strings follow the reference's conversions (conversions.F90:193-199, 238-330) and logging
preambles (logging.F90:16-19, 74-92).
"""
import time as _time

TIME_TERMINATION_REASON = 1
TIMESTEP_TERMINATION_REASON = 2
MESSAGE_TERMINATION_REASON = 3
WALLTIME_TERMINATION_REASON = 4


def int_str(n):
    """integer_to_string: (i15), left-adjusted and trimmed."""
    return str(int(n))


def real_str(x, decimal_places=None):
    """real_double_to_string with optional decimal places: (f30.10) (es30.10 if it does not
    fit), trailing zeros after the point cut to one (trim_trailing_zeros with 2 retained keeps
    one character past the last non-zero digit), then the decimals truncated (not rounded) to
    decimal_places."""
    s = f"{x:30.10f}"
    if len(s) > 30:
        s = f"{x:30.10E}"
    s = s.strip()
    # trim_trailing_zeros(s, 2)
    dp = s.find(".")
    if dp >= 0 and "E" not in s:
        i = len(s) - 1
        zeros = 0
        while i >= dp and s[i] == "0":
            zeros += 1
            i -= 1
        if zeros > 2:
            s = s[:i + 2]
    if decimal_places is not None and dp >= 0:
        e = s.find("E")
        if e < 0:
            s = s[:dp + decimal_places + 1]
        else:
            s = s[:dp + decimal_places + 1] + s[e:]
    return s.strip()


def parse_walltime(text, log):
    """Walltime limit in seconds from ``hh:mm:ss`` (terminationcheck.F90:48-80), or None for
    no limit.  Fewer than two colons: a warning and no limit.  Minutes or seconds outside
    0..59 are set to 0 with a warning."""
    if text.strip() == "none":
        return None
    pidx = 0
    parts = []
    for i in range(2):
        idx = text[pidx:].find(":")
        if idx < 0:
            log("[WARN] Walltime limit of  `" + text.strip() +
                "` does not contains hh:mm:ss, defaulting to no limit")
            return None
        parts.append(int(text[pidx:pidx + idx]))
        pidx += idx + 1
    hours, mins = parts
    secs = int(text[pidx:])
    mangled = False
    if mins < 0 or mins > 59:
        mins = 0
        mangled = True
    if secs < 0 or secs > 59:
        secs = 0
        mangled = True
    if mangled:
        log("[WARN] Walltime limit of `" + text.strip() + "` mangled, defaulting to " +
            f"{hours}:{mins}:{secs}")
    return hours * 3600 + mins * 60 + secs


class RunState:
    """The parts of model_state_type the control components read and write (state.F90:39-54)."""

    def __init__(self, timestep=1, time=0.0, dtm=0.0):
        self.timestep = timestep
        self.time = time
        self.dtm = dtm
        self.continue_timestep = True
        self.termination_reason = 0
        self.numparcels = 0
        self.model_start_wtime = 0.0
        self.iterations = 0  # state.F90:53; restart_num after a restart


class TerminationCheck:
    def __init__(self, options, log=print, wtime=_time.time):
        self.max_timesteps = int(options.get("nn_timesteps", -1))
        self.termination_time = float(options.get("termination_time", 1.0))
        self.check_msg_frequency = int(options.get("check_msg_frequency", 500))
        self.msg_filename = str(options.get("msg_filename", "messages"))
        self.check_walltime_frequency = int(options.get("check_walltime_frequency", 200))
        self.max_walltime_secs = parse_walltime(str(options.get("walltime_limit", "none")), log)
        self.wtime = wtime

    def messages_say_terminate(self):
        try:
            with open(self.msg_filename) as f:
                line = f.readline()
        except OSError:
            return False
        return line.rstrip("\n").rstrip() == "terminate"

    def __call__(self, st):
        if self.max_timesteps > 0:
            st.continue_timestep = st.timestep % self.max_timesteps != 0
            if not st.continue_timestep:
                st.termination_reason = TIMESTEP_TERMINATION_REASON
        else:
            st.continue_timestep = True
        if st.continue_timestep:
            st.continue_timestep = st.time < self.termination_time
            if not st.continue_timestep:
                st.termination_reason = TIME_TERMINATION_REASON
        if (st.continue_timestep and self.max_walltime_secs is not None and
                st.timestep % self.check_walltime_frequency == 0):
            st.continue_timestep = int(self.wtime() - st.model_start_wtime) < self.max_walltime_secs
            if not st.continue_timestep:
                st.termination_reason = WALLTIME_TERMINATION_REASON
        if st.continue_timestep and st.timestep % self.check_msg_frequency == 0:
            if self.messages_say_terminate():
                st.continue_timestep = False
            if not st.continue_timestep:
                st.termination_reason = MESSAGE_TERMINATION_REASON


class ModelSynopsis:
    def __init__(self, options, st, log=print, parcels_dat="parcels.dat", wtime=_time.time):
        self.reporting_frequency = int(options.get("display_synopsis_frequency", 10))
        self.options = options
        self.previous_ts = st.timestep
        self.wtime = wtime
        self.start_time = wtime()
        self.log = log
        self.dat = open(parcels_dat, "w") if parcels_dat else None

    def timestep(self, st):
        if st.timestep % self.reporting_frequency == 0:
            end_time = self.wtime()
            self.log("")
            self.log("[INFO] Number of completed timesteps " + int_str(st.timestep))
            self.log("[INFO] Completed " + int_str(st.timestep - self.previous_ts + 1) +
                     " timesteps in " + int_str(int((end_time - self.start_time) * 1000)) + "ms")
            self.log("[INFO] Model time " + real_str(st.time, 5) + " seconds; dtm=" +
                     real_str(st.dtm, 5))
            self.previous_ts = st.timestep
            self.start_time = self.wtime()
        if self.dat:
            self.dat.write(f" {st.numparcels:11d}\n")

    def finalise(self, st):
        self.log("")
        r = st.termination_reason
        if r == TIME_TERMINATION_REASON:
            msg = ("Model run complete due to model time " + real_str(st.time, 2) +
                   " exceeding limit of " +
                   real_str(float(self.options.get("termination_time", 1.0)), 2))
        elif r == TIMESTEP_TERMINATION_REASON:
            msg = "Model run complete due to timestep completion, model time is " + \
                real_str(st.time, 2)
        elif r == MESSAGE_TERMINATION_REASON:
            msg = ("Model run complete due to messages file containing termination command, "
                   "model time is " + real_str(st.time, 2))
        elif r == WALLTIME_TERMINATION_REASON:
            msg = ("Model run complete due to walltime limit of '" +
                   str(self.options.get("walltime_limit", "none")).strip() +
                   "' reached, model time is " + real_str(st.time, 2))
        else:
            msg = "Model run complete due to unknown reason, model time is " + real_str(st.time, 2)
        self.log("[INFO] " + msg)
        if self.dat:
            self.dat.close()


class WriteBinary:
    """writebinary's timestep callback (components/writebinary/src/writebinary.F90:38-135), run
    in the start group before the termination check.  Mode "step": the parcels every
    parcel_step_write_frequency iterations, then the grids every grid_step_write_frequency;
    mode "time": the grids once the model time reaches the next multiple of
    grid_time_write_frequency, then the parcels likewise; "none" writes nothing.  The dump
    number is the iteration count; files go to `directory`."""

    def __init__(self, options, directory=".", log=print):
        self.mode = str(options.get("writebinary_mode", "none")).strip().strip('"')
        self.directory = directory
        self.log = log
        self.pwritten = self.gwritten = 0
        self.ppersteps = int(options.get("parcel_step_write_frequency", 0))
        self.gpersteps = int(options.get("grid_step_write_frequency", 0))
        self.dtparcels = float(options.get("parcel_time_write_frequency", 0.0))
        self.dtgrids = float(options.get("grid_time_write_frequency", 0.0))
        self.tnextparcels = self.tnextgrids = 0.0
        if self.mode not in ("step", "time", "none"):
            raise ValueError(f"Error: mode '{self.mode}' not recognised.")
        log(f" Selected writing mode: {self.mode}")

    def _parcels(self, ctx, num):
        from .dumpio import dump, parcel_file
        self.log(f" Writing parcels to '{parcel_file(0, num).replace('00000', '[rank]')}'")
        dump(ctx, self.directory, num, grids=False)
        self.pwritten += 1

    def _grids(self, ctx, num):
        from .dumpio import dump, grid_file
        self.log(f" Writing grids to '{grid_file(0, num).replace('00000', '[rank]')}'")
        dump(ctx, self.directory, num, parcels=False)
        self.gwritten += 1

    def timestep(self, ctx, st):
        num = st.iterations
        if self.mode == "step":
            if self.ppersteps != 0 and num % self.ppersteps == 0:
                self._parcels(ctx, num)
            if self.gpersteps != 0 and num % self.gpersteps == 0:
                self._grids(ctx, num)
        elif self.mode == "time":
            if self.dtgrids != 0. and st.time >= self.tnextgrids:
                self._grids(ctx, num)
                self.tnextgrids += self.dtgrids
            if self.dtparcels != 0. and st.time >= self.tnextparcels:
                self._parcels(ctx, num)
                self.tnextparcels += self.dtparcels

    def finalise(self):
        self.log(f" Written {self.pwritten} parcel dumps")
        self.log(f" Written {self.gwritten} grid dumps")


def timer_summary(rows, log=print):
    """The master rank's timer summary of finalize_timing (timer.F90:247-303).

    rows: [(name, calls, total_s, max_s, sum of squared intervals in s^2)], the first being the
    run's own timer (``MONC``, monc.F90:155) whose total the percentages are of.  Rows are listed
    from the longest total to the shortest (the reference's bubble sort: stable for equal
    totals); a timer's stddev % is sqrt(sum dt^2 / n - mean^2) / mean."""
    if not rows:
        return
    mxtime = rows[0][2]
    order = list(range(len(rows)))
    vals = [r[2] for r in rows]
    swapped = True
    while swapped:  # bubble sort, descending, as sort_values (timer.F90:412-443)
        swapped = False
        for i in range(len(vals) - 1):
            if vals[i] < vals[i + 1]:
                vals[i], vals[i + 1] = vals[i + 1], vals[i]
                order[i], order[i + 1] = order[i + 1], order[i]
                swapped = True
    log("")
    log(" Summary of times for the master rank (sorted from longest to shortest time):")
    log(" ------------------------------------------------------------------------------ ")
    log("|     Function       |   #    |  Total   |  % of  |  Mean   |  Max    | stddev |")
    log("|       Name         | Calls  |  Time    |  Time  |Duration |Duration |    %   |")
    log("|------------------------------------------------------------------------------|")
    for n in order:
        name, calls, tot, mx, sq = rows[n]
        if calls > 0:
            mean = tot / calls
            var = sq / calls - mean * mean
            stddev = var ** 0.5 if var > 0 else 0.0
            imbalance = stddev / mean * 100 if mean > 0 else 0.0
        else:
            mean = stddev = imbalance = 0.0
        pct = tot / mxtime * 100 if mxtime > 0 else 0.0
        log(f"|{name[:20]:<20}|{calls:8d}|{tot:9.3f}s| {pct:6.2f}%|{mean:8.3f}s|{mx:8.3f}s|"
            f"{imbalance:7.2f}%|")
    log(" ------------------------------------------------------------------------------  ")


def run_model(ctx, options, log=print, parcels_dat="parcels.dat", wtime=_time.time,
              timers=False, writer=None, iterations=0):
    """The timestep loop (monc.F90:211-230) over a library context; returns the RunState.
    With timers, the library's component timers run and the reference's timer summary is
    logged at the end, under the run's own MONC timer.  `writer` (a WriteBinary) runs first in
    the start group; `iterations` is the iteration count to start from (restart_num after a
    restart)."""
    st = RunState(timestep=int(options.get("start_timestep", 1)), time=ctx.time, dtm=ctx.dtm)
    st.iterations = iterations
    st.model_start_wtime = wtime()
    if timers:
        ctx.timing_enable(True)
    term = TerminationCheck(options, log, wtime)
    syn = ModelSynopsis(options, st, log, parcels_dat, wtime)
    while st.continue_timestep:
        # start group: writebinary, then the termination check (global_config:36)
        if writer is not None:
            writer.timestep(ctx, st)
        term(st)
        # integrator group: one whole step of the path
        t_before = ctx.time
        ctx.step(1)
        st.dtm = ctx.dtm
        st.numparcels = ctx.n
        # last group
        syn.timestep(st)
        if st.continue_timestep:
            st.timestep += 1
            st.time = ctx.time
            st.iterations += 1
        else:
            # the reference does not advance the model time on the final iteration
            ctx.set_time(t_before, ctx.dtm)
            st.time = t_before
    syn.finalise(st)
    if writer is not None:
        writer.finalise()
    if timers:
        total = wtime() - st.model_start_wtime
        rows = [("MONC", 1, total, total, total * total)]
        for name, (t, n, mx, sq) in ctx.timer_stats().items():
            rows.append((name, n, t / 1e3, mx / 1e3, sq / 1e6))
        timer_summary(rows, log)
    return st
