"""ctypes binding of libpmpic_hip.so (the C ABI declared in include/pmpic_hip.h).

The library is built in-tree (``pmpic_amd/libpmpic_hip.so``) by ``pmpic_amd.build``; importing
this module never builds or falls back to anything else: a missing library is an error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmpic_hip.so")

PARCEL_FIELDS = ["x", "y", "z", "p", "q", "r", "dxdt", "dydt", "dzdt", "dpdt", "dqdt", "drdt",
                 "h", "b", "vol", "stretch", "tag"]
RK_FIELDS = ["xo", "yo", "zo", "xf", "yf", "zf", "po", "qo", "ro", "pf", "qf", "rf"]
GRID_IDS = {"u": 0, "v": 1, "w": 2, "p": 3, "q": 4, "r": 5, "b": 6, "hg": 7, "hgliq": 8,
            "vol": 9, "dp": 10, "dq": 11, "dr": 12}
SPEC_IDS = {"u_s": 0, "v_s": 1, "w_s": 2}

# size caps of pmpic_par2grid_noalloc (its static scratch area)
PMPIC_NOALLOC_MAX_PARCELS = 1 << 20
PMPIC_NOALLOC_MAX_CELLS = 1 << 15

STATUS = {
    0: "PMPIC_OK", 1: "PMPIC_ERR_OUT_OF_BOX_X", 2: "PMPIC_ERR_OUT_OF_BOX_Y",
    3: "PMPIC_ERR_OUT_OF_BOX_Z", 4: "PMPIC_ERR_DELTA", 5: "PMPIC_ERR_MAXPARCELS",
    6: "PMPIC_ERR_ZERO_PIVOT", 7: "PMPIC_ERR_FFT_SIZE", 8: "PMPIC_ERR_COUNT",
    9: "PMPIC_ERR_BACKFILL", 10: "PMPIC_ERR_ARG", 11: "PMPIC_ERR_HIP", 12: "PMPIC_ERR_NOMEM",
    13: "PMPIC_ERR_BC",
}

# every symbol include/pmpic_hip.h declares
EXPORTS = [
    "pmpic_create", "pmpic_destroy", "pmpic_status_message", "pmpic_set_parcels",
    "pmpic_get_parcels", "pmpic_set_rk_arrays", "pmpic_get_rk_arrays", "pmpic_get_state",
    "pmpic_set_time", "pmpic_vort2vel", "pmpic_vorticity_tendency", "pmpic_rk4_integrator",
    "pmpic_parcel_splitting", "pmpic_parcel_mixing", "pmpic_step",
    "pmpic_cache_parcel_interp_weights", "pmpic_par2grid", "pmpic_grid2par",
    "pmpic_parcel_haloswap", "pmpic_get_grid", "pmpic_get_spectral", "pmpic_fft_forward",
    "pmpic_fft_backward", "pmpic_diffx", "pmpic_diffy", "pmpic_spectral_filter", "pmpic_diffz",
    "pmpic_laplinv", "pmpic_conservation_sums", "pmpic_last_counts", "pmpic_timing_enable",
    "pmpic_timing_read", "pmpic_step_concurrent", "pmpic_timing_name", "pmpic_hbm_triad",
    "pmpic_writebinary_grids", "pmpic_par2grid_noalloc", "pmpic_timing_stats",
    "pmpic_graph_stats", "pmpic_autotune", "pmpic_launch_config", "pmpic_set_launch_config",
]


class PmpicError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)} ({message(status)})")


class Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
                ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
                ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
                ("max_parcels", ctypes.c_int64),
                ("dtm", ctypes.c_double), ("dtmax", ctypes.c_double),
                ("timestep_prefactor", ctypes.c_double),
                ("precision", ctypes.c_int32), ("splitting_enabled", ctypes.c_int32),
                ("mixing_enabled", ctypes.c_int32), ("device", ctypes.c_int32)]


_lib = None


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -c 'import pmpic_amd.build as b; b.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        pp = ctypes.POINTER(ctypes.c_void_p)
        for name in EXPORTS:
            getattr(L, name).restype = ctypes.c_int
        L.pmpic_status_message.restype = ctypes.c_char_p
        L.pmpic_status_message.argtypes = [ctypes.c_int]
        L.pmpic_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
        L.pmpic_destroy.argtypes = [vp]
        L.pmpic_set_parcels.argtypes = [vp, i64, pp]
        L.pmpic_get_parcels.argtypes = [vp, ctypes.POINTER(i64), pp]
        L.pmpic_set_rk_arrays.argtypes = [vp, pp]
        L.pmpic_get_rk_arrays.argtypes = [vp, pp]
        L.pmpic_get_state.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(d), ctypes.POINTER(d)]
        L.pmpic_set_time.argtypes = [vp, d, d]
        for name in ("pmpic_vort2vel", "pmpic_vorticity_tendency", "pmpic_rk4_integrator",
                     "pmpic_parcel_splitting", "pmpic_parcel_mixing", "pmpic_parcel_haloswap",
                     "pmpic_writebinary_grids"):
            getattr(L, name).argtypes = [vp]
        L.pmpic_step.argtypes = [vp, i32]
        L.pmpic_step_concurrent.argtypes = [pp, i32, i32]
        L.pmpic_cache_parcel_interp_weights.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.pmpic_par2grid.argtypes = [vp, vp, vp]
        L.pmpic_grid2par.argtypes = [vp, vp, vp]
        L.pmpic_par2grid_noalloc.argtypes = [vp, vp, vp]
        L.pmpic_get_grid.argtypes = [vp, i32, vp]
        L.pmpic_get_spectral.argtypes = [vp, i32, vp]
        L.pmpic_fft_forward.argtypes = [vp, vp, vp]
        L.pmpic_fft_backward.argtypes = [vp, vp, vp]
        L.pmpic_diffx.argtypes = [vp, vp, vp]
        L.pmpic_diffy.argtypes = [vp, vp, vp]
        L.pmpic_spectral_filter.argtypes = [vp, vp, vp]
        L.pmpic_diffz.argtypes = [vp, vp, vp, vp, vp]
        L.pmpic_laplinv.argtypes = [vp, vp, i32]
        L.pmpic_conservation_sums.argtypes = [vp, vp]
        L.pmpic_last_counts.argtypes = [vp, vp]
        L.pmpic_timing_enable.argtypes = [vp, i32]
        L.pmpic_timing_read.argtypes = [vp, i32, ctypes.POINTER(d), ctypes.POINTER(i64)]
        L.pmpic_timing_stats.argtypes = [vp, i32, ctypes.POINTER(d), ctypes.POINTER(i64),
                                         ctypes.POINTER(d), ctypes.POINTER(d)]
        L.pmpic_graph_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.pmpic_autotune.argtypes = [vp, i32, vp, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(i32)]
        L.pmpic_launch_config.argtypes = [vp, vp]
        L.pmpic_set_launch_config.argtypes = [vp, vp]
        L.pmpic_timing_name.restype = ctypes.c_char_p
        L.pmpic_timing_name.argtypes = [i32]
        L.pmpic_hbm_triad.argtypes = [i32, i64, i32, ctypes.POINTER(d), ctypes.POINTER(d)]
        _lib = L
    return _lib


def message(status):
    try:
        return load().pmpic_status_message(int(status)).decode()
    except Exception:  # message lookup must never mask the original error
        return "?"


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ptrs(arrays):
    arr = (ctypes.c_void_p * len(arrays))()
    for i, a in enumerate(arrays):
        arr[i] = None if a is None else a.ctypes.data
    return arr


def step_concurrent(contexts, nsteps=1):
    """Advance an ensemble of distinct contexts by nsteps timesteps at once (one host thread
    and HIP stream per member, pmpic_step_concurrent)."""
    L = load()
    arr = (ctypes.c_void_p * len(contexts))()
    for i, cx in enumerate(contexts):
        arr[i] = cx.h.value
    st = L.pmpic_step_concurrent(arr, len(contexts), int(nsteps))
    if st != 0:
        raise PmpicError(st, "pmpic_step_concurrent")


def hbm_triad(bytes_per_array=1 << 30, reps=5, device=0):
    """STREAM-style triad bandwidth of the device (pmpic_hbm_triad): (best ms, GB/s)."""
    ms, gbps = ctypes.c_double(), ctypes.c_double()
    st = load().pmpic_hbm_triad(int(device), int(bytes_per_array), int(reps), ctypes.byref(ms),
                                ctypes.byref(gbps))
    if st != 0:
        raise PmpicError(st, "pmpic_hbm_triad")
    return ms.value, gbps.value


class Context:
    """One device-resident PMPIC state (a pmpic_ctx)."""

    def __init__(self, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, max_parcels,
                 dtm=0.01, dtmax=0.01, timestep_prefactor=0.5, splitting=True, mixing=True,
                 device=0, precision=0):
        """precision 0 = plain FP64 form; 1 = FP32-bulk form (cell index + FP32 offsets, FP32
        parcel fields and interpolation; components, step, parcel transfer and conservation
        sums only)."""
        self.L = load()
        cfg = Config(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, int(max_parcels), dtm, dtmax,
                     timestep_prefactor, int(precision), int(bool(splitting)), int(bool(mixing)),
                     device)
        h = ctypes.c_void_p()
        self._chk(self.L.pmpic_create(ctypes.byref(cfg), ctypes.byref(h)), "pmpic_create")
        self.h = h
        self.precision = int(precision)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.box = (xmin, xmax, ymin, ymax, zmin, zmax)
        self.max_parcels = int(max_parcels)
        self.gshape = (nx + 4, ny + 4, nz + 1)
        self.sshape = ((nx // 2 + 1) * 2, (ny // 2 + 1) * 2, nz + 1)

    def close(self):
        if getattr(self, "h", None):
            self.L.pmpic_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def _chk(st, what):
        if st != 0:
            raise PmpicError(st, what)

    # ---- state ----
    def set_parcels(self, fields):
        n = len(fields["x"])
        zero = None  # one shared array of zeros for every field not given
        arrs = []
        for k in PARCEL_FIELDS:
            if k in fields:
                arrs.append(np.ascontiguousarray(fields[k], dtype=np.float64))
            else:
                if zero is None:
                    zero = np.zeros(n)
                arrs.append(zero)
        self._chk(self.L.pmpic_set_parcels(self.h, n, _ptrs(arrs)), "pmpic_set_parcels")

    def get_parcels(self, names=None):
        n = self.n
        names = names or PARCEL_FIELDS
        out = {k: np.zeros(n) for k in names}
        arrs = [out.get(k) for k in PARCEL_FIELDS]
        cnt = ctypes.c_int64()
        self._chk(self.L.pmpic_get_parcels(self.h, ctypes.byref(cnt), _ptrs(arrs)), "pmpic_get_parcels")
        return out

    def set_rk_arrays(self, fields):
        n = self.n
        arrs = [np.ascontiguousarray(fields.get(k, np.zeros(n)), dtype=np.float64) for k in RK_FIELDS]
        self._chk(self.L.pmpic_set_rk_arrays(self.h, _ptrs(arrs)), "pmpic_set_rk_arrays")

    def get_rk_arrays(self):
        n = self.n
        out = {k: np.zeros(n) for k in RK_FIELDS}
        self._chk(self.L.pmpic_get_rk_arrays(self.h, _ptrs([out[k] for k in RK_FIELDS])),
                  "pmpic_get_rk_arrays")
        return out

    def state(self):
        n, t, dtm = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.pmpic_get_state(self.h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(dtm)),
                  "pmpic_get_state")
        return n.value, t.value, dtm.value

    @property
    def n(self):
        return self.state()[0]

    @property
    def time(self):
        return self.state()[1]

    @property
    def dtm(self):
        return self.state()[2]

    def set_time(self, time, dtm):
        self._chk(self.L.pmpic_set_time(self.h, float(time), float(dtm)), "pmpic_set_time")

    # ---- components ----
    def vort2vel(self):
        self._chk(self.L.pmpic_vort2vel(self.h), "pmpic_vort2vel")

    def vorticity_tendency(self):
        self._chk(self.L.pmpic_vorticity_tendency(self.h), "pmpic_vorticity_tendency")

    def rk4_integrator(self):
        self._chk(self.L.pmpic_rk4_integrator(self.h), "pmpic_rk4_integrator")

    def parcel_splitting(self):
        self._chk(self.L.pmpic_parcel_splitting(self.h), "pmpic_parcel_splitting")

    def parcel_mixing(self):
        self._chk(self.L.pmpic_parcel_mixing(self.h), "pmpic_parcel_mixing")

    def parcel_haloswap(self):
        self._chk(self.L.pmpic_parcel_haloswap(self.h), "pmpic_parcel_haloswap")

    def step(self, nsteps=1):
        self._chk(self.L.pmpic_step(self.h, int(nsteps)), "pmpic_step")

    def step_status(self, nsteps=1):
        return int(self.L.pmpic_step(self.h, int(nsteps)))

    def last_counts(self):
        out = np.zeros(3, dtype=np.int64)
        self._chk(self.L.pmpic_last_counts(self.h, _p(out)), "pmpic_last_counts")
        return tuple(int(v) for v in out)

    def conservation_sums(self):
        out = np.zeros(6)
        self._chk(self.L.pmpic_conservation_sums(self.h, _p(out)), "pmpic_conservation_sums")
        return out

    # ---- parcel_interpolation_mod ----
    def cache_weights(self):
        n = self.n
        ijk = [np.zeros(n, dtype=np.int32) for _ in range(3)]
        dl = [np.zeros(n) for _ in range(3)]
        self._chk(self.L.pmpic_cache_parcel_interp_weights(self.h, *[_p(a) for a in ijk + dl]),
                  "pmpic_cache_parcel_interp_weights")
        return ijk, dl

    def par2grid(self, var):
        var = np.ascontiguousarray(var, dtype=np.float64)
        out = np.zeros(self.gshape)
        self._chk(self.L.pmpic_par2grid(self.h, _p(var), _p(out)), "pmpic_par2grid")
        return out

    def grid2par(self, grid):
        grid = np.ascontiguousarray(grid, dtype=np.float64)
        out = np.zeros(self.n)
        self._chk(self.L.pmpic_grid2par(self.h, _p(grid), _p(out)), "pmpic_grid2par")
        return out

    def grid(self, name):
        out = np.zeros(self.gshape)
        self._chk(self.L.pmpic_get_grid(self.h, GRID_IDS[name], _p(out)), "pmpic_get_grid")
        return out

    def writebinary_grids(self):
        """Compute the grids writebinary dumps (p, q, r, b, hg, hgliq) from the current parcels."""
        self._chk(self.L.pmpic_writebinary_grids(self.h), "pmpic_writebinary_grids")

    def spectral(self, name):
        out = np.zeros(self.sshape)
        self._chk(self.L.pmpic_get_spectral(self.h, SPEC_IDS[name], _p(out)), "pmpic_get_spectral")
        return out

    # ---- standalone spectral operators ----
    def fft_forward(self, grid):
        grid = np.ascontiguousarray(grid, dtype=np.float64)
        out = np.zeros(self.sshape)
        self._chk(self.L.pmpic_fft_forward(self.h, _p(grid), _p(out)), "pmpic_fft_forward")
        return out

    def fft_backward(self, spec):
        spec = np.ascontiguousarray(spec, dtype=np.float64)
        out = np.zeros(self.gshape)
        self._chk(self.L.pmpic_fft_backward(self.h, _p(spec), _p(out)), "pmpic_fft_backward")
        return out

    def _unary(self, fn, s, what):
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.zeros(self.sshape)
        self._chk(fn(self.h, _p(s), _p(out)), what)
        return out

    def diffx(self, s):
        return self._unary(self.L.pmpic_diffx, s, "pmpic_diffx")

    def diffy(self, s):
        return self._unary(self.L.pmpic_diffy, s, "pmpic_diffy")

    def filter(self, s):
        return self._unary(self.L.pmpic_spectral_filter, s, "pmpic_spectral_filter")

    def diffz(self, s, bot=None, top=None):
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.zeros(self.sshape)
        if bot is None:
            st = self.L.pmpic_diffz(self.h, _p(s), _p(out), None, None)
        else:
            bot = np.ascontiguousarray(bot, dtype=np.float64)
            top = np.ascontiguousarray(top, dtype=np.float64)
            st = self.L.pmpic_diffz(self.h, _p(s), _p(out), _p(bot), _p(top))
        self._chk(st, "pmpic_diffz")
        return out

    def laplinv(self, s, df_zero_on_boundary):
        f = np.ascontiguousarray(s, dtype=np.float64).copy()
        self._chk(self.L.pmpic_laplinv(self.h, _p(f), 1 if df_zero_on_boundary else 0), "pmpic_laplinv")
        return f

    # ---- timing ----
    def timing_enable(self, on=True, kernels_only=False):
        """Timers on (every timer) or off; kernels_only keeps just the two hot kernels' timers,
        which are taken by the kernel dispatches themselves and add nothing between kernels."""
        mode = (2 if kernels_only else 1) if on else 0
        self._chk(self.L.pmpic_timing_enable(self.h, mode), "pmpic_timing_enable")

    def timing_read(self, which):
        ms, calls = ctypes.c_double(), ctypes.c_int64()
        self._chk(self.L.pmpic_timing_read(self.h, int(which), ctypes.byref(ms), ctypes.byref(calls)),
                  "pmpic_timing_read")
        return ms.value, calls.value

    def timing_stats(self, which):
        """(total ms, intervals, longest ms, sum of squared ms) of timer `which`."""
        t, n, mx, sq = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.pmpic_timing_stats(self.h, int(which), ctypes.byref(t), ctypes.byref(n),
                                            ctypes.byref(mx), ctypes.byref(sq)),
                  "pmpic_timing_stats")
        return t.value, n.value, mx.value, sq.value

    def autotune(self, steps=2):
        """Pick this FP32-bulk context's launch configuration by timing each over `steps` model
        steps (or take the one picked for its grid size earlier in the process).  Returns the
        configuration (see launch_config), the chosen ms per step and whether it was measured
        here."""
        cfg = (ctypes.c_int32 * 4)()
        ms, meas = ctypes.c_float(), ctypes.c_int32()
        self._chk(self.L.pmpic_autotune(self.h, int(steps), cfg, ctypes.byref(ms),
                                        ctypes.byref(meas)), "pmpic_autotune")
        out = self._config(cfg)
        out.update(ms_per_step=ms.value, measured=bool(meas.value))
        return out

    @staticmethod
    def _config(cfg):
        return {"slot_capacity": cfg[0], "fused_binning": bool(cfg[1]), "graph": bool(cfg[2]),
                "rk_workgroup": cfg[3]}

    def launch_config(self):
        """The FP32-bulk launch configuration in use: slot capacity, binning fused into the RK
        stages, stages 1-3 as a HIP graph, RK workgroup size."""
        cfg = (ctypes.c_int32 * 4)()
        self._chk(self.L.pmpic_launch_config(self.h, cfg), "pmpic_launch_config")
        return self._config(cfg)

    def set_launch_config(self, slot_capacity=0, fused_binning=False, graph=False,
                          rk_workgroup=256):
        """Set the FP32-bulk launch configuration (slot_capacity 0: the mean-occupancy rule)."""
        cfg = (ctypes.c_int32 * 4)(int(slot_capacity), int(bool(fused_binning)), int(bool(graph)),
                                   int(rk_workgroup))
        self._chk(self.L.pmpic_set_launch_config(self.h, cfg), "pmpic_set_launch_config")

    def graph_stats(self):
        """(graphs captured, graph launches) of pmpic_step in the FP32-bulk form."""
        cap, run = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self.L.pmpic_graph_stats(self.h, ctypes.byref(cap), ctypes.byref(run)),
                  "pmpic_graph_stats")
        return cap.value, run.value

    def timer_stats(self):
        """{region name: (total ms, intervals, longest ms, sum of squared ms)} for every timer."""
        out = {}
        w = 0
        while True:
            name = self.L.pmpic_timing_name(w)
            if name is None:
                return out
            out[name.decode()] = self.timing_stats(w)
            w += 1

    def timer_table(self):
        """{region name: (total ms, count)} for every timer (timing must be enabled)."""
        out = {}
        w = 0
        while True:
            name = self.L.pmpic_timing_name(w)
            if name is None:
                return out
            out[name.decode()] = self.timing_read(w)
            w += 1
