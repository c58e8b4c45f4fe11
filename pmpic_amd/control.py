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


def run_model(ctx, options, log=print, parcels_dat="parcels.dat", wtime=_time.time):
    """The timestep loop (monc.F90:211-230) over a library context; returns the RunState."""
    st = RunState(timestep=int(options.get("start_timestep", 1)), time=ctx.time, dtm=ctx.dtm)
    st.model_start_wtime = wtime()
    term = TerminationCheck(options, log, wtime)
    syn = ModelSynopsis(options, st, log, parcels_dat, wtime)
    while st.continue_timestep:
        # start group
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
        else:
            # the reference does not advance the model time on the final iteration
            ctx.set_time(t_before, ctx.dtm)
            st.time = t_before
    syn.finalise(st)
    return st
