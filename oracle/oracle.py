"""ctypes front end of liboracle.so, the C FP64 restatement of the PMPIC path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  Arrays are exposed as numpy views over the
oracle's own buffers; grids keep the reference layout data(z,y,x) (z fastest), which numpy sees
as shape (nx+4, ny+4, nz+1) in C order, i.e. index [i, j, k].
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_FFTE_PATH = os.path.join(HERE, "_ref", "libffte_ref.so")

FIELDS = ["x", "y", "z", "p", "q", "r", "dxdt", "dydt", "dzdt", "dpdt", "dqdt", "drdt",
          "h", "b", "vol", "stretch", "tag",
          "xo", "yo", "zo", "xf", "yf", "zf", "po", "qo", "ro", "pf", "qf", "rf", "btot"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
GRIDS = ["u", "v", "w", "p", "q", "r", "b", "hg", "hgliq", "vol", "dp", "dq", "dr"]
GRID_ID = {n: i for i, n in enumerate(GRIDS)}
SPECS = ["u_s", "v_s", "w_s"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d, i, i64, vp = ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, i, i, d, d, d, d, d, d, i64, d, d, d]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_threads.argtypes = [i]
        L.orc_max_threads.restype = i
        L.orc_max_threads.argtypes = []
        for name in ("orc_field", "orc_grid", "orc_spec"):
            getattr(L, name).restype = ctypes.POINTER(ctypes.c_double)
            getattr(L, name).argtypes = [vp, i]
        L.orc_cache_ijk.restype = ctypes.POINTER(ctypes.c_int32)
        L.orc_cache_ijk.argtypes = [vp, i]
        L.orc_cache_del.restype = ctypes.POINTER(ctypes.c_double)
        L.orc_cache_del.argtypes = [vp, i]
        L.orc_get_n.restype = i64
        L.orc_get_n.argtypes = [vp]
        L.orc_set_n.argtypes = [vp, i64]
        L.orc_get_dtm.restype = d
        L.orc_get_dtm.argtypes = [vp]
        L.orc_set_dtm.argtypes = [vp, d]
        L.orc_get_time.restype = d
        L.orc_get_time.argtypes = [vp]
        L.orc_geometry.argtypes = [vp, vp]
        L.orc_wavenumber.restype = d
        L.orc_wavenumber.argtypes = [vp, i, i, i]
        L.orc_last_counts.restype = i64
        L.orc_last_counts.argtypes = [vp, i]
        for name in ("orc_cache_weights", "orc_vort2vel", "orc_vorticity_tendency",
                     "orc_rk4_integrator", "orc_parcel_splitting", "orc_parcel_mixing",
                     "orc_parcel_haloswap", "orc_step", "orc_writebinary_grids"):
            getattr(L, name).restype = i
            getattr(L, name).argtypes = [vp]
        L.orc_par2grid.restype = i
        L.orc_par2grid.argtypes = [vp, vp, vp]
        L.orc_grid2par.restype = i
        L.orc_grid2par.argtypes = [vp, vp, vp]
        L.orc_fold.argtypes = [vp, vp]
        L.orc_fill.argtypes = [vp, vp]
        L.orc_fft_forward.argtypes = [vp, vp, vp]
        L.orc_fft_backward.argtypes = [vp, vp, vp]
        L.orc_diffx.argtypes = [vp, vp, vp]
        L.orc_diffy.argtypes = [vp, vp, vp]
        L.orc_filter.argtypes = [vp, vp, vp]
        L.orc_diffz.restype = i
        L.orc_diffz.argtypes = [vp, vp, vp, vp, vp]
        L.orc_laplinv.restype = i
        L.orc_laplinv.argtypes = [vp, vp, i]
        L.orc_tridiagonal.restype = i
        L.orc_tridiagonal.argtypes = [i, vp, vp, vp, vp, vp]
        L.orc_r2c.argtypes = [i, vp, vp]
        L.orc_c2r.argtypes = [i, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    """OpenMP threads of the oracle's later calls; its results do not depend on the count."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


class Oracle:
    """One oracle model instance: grid geometry, parcel SoA of capacity maxparcels, grids."""

    def __init__(self, nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, maxparcels,
                 dtm=0.01, dtmax=0.01, prefactor=0.5):
        self.L = lib()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.maxparcels = int(maxparcels)
        self.h = self.L.orc_create(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax,
                                   self.maxparcels, dtm, dtmax, prefactor)
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        self.gshape = (nx + 4, ny + 4, nz + 1)
        self.sshape = ((nx // 2 + 1) * 2, (ny // 2 + 1) * 2, nz + 1)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    # ---- views ----
    def field(self, name):
        ptr = self.L.orc_field(self.h, FIELD_ID[name])
        return np.ctypeslib.as_array(ptr, shape=(self.maxparcels,))

    def grid(self, name):
        ptr = self.L.orc_grid(self.h, GRID_ID[name])
        return np.ctypeslib.as_array(ptr, shape=self.gshape)

    def spec(self, name):
        ptr = self.L.orc_spec(self.h, SPECS.index(name))
        return np.ctypeslib.as_array(ptr, shape=self.sshape)

    def cache(self):
        n = self.n
        ijk = [np.ctypeslib.as_array(self.L.orc_cache_ijk(self.h, a), shape=(self.maxparcels,))[:n].copy()
               for a in range(3)]
        dl = [np.ctypeslib.as_array(self.L.orc_cache_del(self.h, a), shape=(self.maxparcels,))[:n].copy()
              for a in range(3)]
        return ijk, dl

    @property
    def n(self):
        return int(self.L.orc_get_n(self.h))

    @n.setter
    def n(self, v):
        self.L.orc_set_n(self.h, int(v))

    @property
    def dtm(self):
        return float(self.L.orc_get_dtm(self.h))

    @dtm.setter
    def dtm(self, v):
        self.L.orc_set_dtm(self.h, float(v))

    @property
    def time(self):
        return float(self.L.orc_get_time(self.h))

    def geometry(self):
        out = np.zeros(16)
        self.L.orc_geometry(self.h, _p(out))
        keys = ["dx", "dy", "dz", "xmin_h", "ymin_h", "zmin_h", "zmax_h", "hs_xmin", "hs_xmax",
                "hs_ymin", "hs_ymax", "xshift_e", "xshift_w", "yshift_n", "yshift_s", "dtm"]
        return dict(zip(keys, out.tolist()))

    def set_parcels(self, fields):
        """fields: dict name -> 1-D array (all the same length); unspecified fields are zeroed."""
        n = len(next(iter(fields.values())))
        if n > self.maxparcels:
            raise ValueError("too many parcels")
        for name in FIELDS:
            a = self.field(name)
            a[:] = 0.0
            if name in fields:
                a[:n] = fields[name]
        self.n = n

    def get_parcels(self, names=None):
        n = self.n
        names = names or FIELDS[:17]
        return {k: self.field(k)[:n].copy() for k in names}

    # ---- operations (status codes raised as errors) ----
    def _chk(self, st, what):
        if st:
            raise RuntimeError(f"oracle {what} failed with status {st}")

    def cache_weights(self):
        self._chk(self.L.orc_cache_weights(self.h), "cache_weights")

    def par2grid(self, var):
        var = np.ascontiguousarray(var, dtype=np.float64)
        out = np.zeros(self.gshape)
        self._chk(self.L.orc_par2grid(self.h, _p(var), _p(out)), "par2grid")
        return out

    def grid2par(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64).copy()
        out = np.zeros(self.n)
        self._chk(self.L.orc_grid2par(self.h, _p(g), _p(out)), "grid2par")
        return out, g

    def fft_forward(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64)
        out = np.zeros(self.sshape)
        self.L.orc_fft_forward(self.h, _p(g), _p(out))
        return out

    def fft_backward(self, spec, grid=None):
        s = np.ascontiguousarray(spec, dtype=np.float64)
        out = np.zeros(self.gshape) if grid is None else grid
        self.L.orc_fft_backward(self.h, _p(s), _p(out))
        return out

    def diffx(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.zeros(self.sshape)
        self.L.orc_diffx(self.h, _p(s), _p(out))
        return out

    def diffy(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.zeros(self.sshape)
        self.L.orc_diffy(self.h, _p(s), _p(out))
        return out

    def filter(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.zeros(self.sshape)
        self.L.orc_filter(self.h, _p(s), _p(out))
        return out

    def diffz(self, s, bot=None, top=None):
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.zeros(self.sshape)
        if bot is None:
            st = self.L.orc_diffz(self.h, _p(s), _p(out), None, None)
        else:
            bot = np.ascontiguousarray(bot, dtype=np.float64)
            top = np.ascontiguousarray(top, dtype=np.float64)
            st = self.L.orc_diffz(self.h, _p(s), _p(out), _p(bot), _p(top))
        self._chk(st, "diffz")
        return out

    def laplinv(self, s, df_zero_on_boundary):
        out = np.ascontiguousarray(s, dtype=np.float64).copy()
        self._chk(self.L.orc_laplinv(self.h, _p(out), 1 if df_zero_on_boundary else 0), "laplinv")
        return out

    def wavenumbers(self):
        sx, sy, _ = self.sshape
        k = np.zeros((4, sx, sy))
        for w in range(4):
            for ii in range(sx):
                for jj in range(sy):
                    k[w, ii, jj] = self.L.orc_wavenumber(self.h, w, jj + 1, ii + 1)
        return k

    def vort2vel(self):
        self._chk(self.L.orc_vort2vel(self.h), "vort2vel")

    def vorticity_tendency(self):
        self._chk(self.L.orc_vorticity_tendency(self.h), "vorticity_tendency")

    def rk4_integrator(self):
        self._chk(self.L.orc_rk4_integrator(self.h), "rk4_integrator")

    def parcel_splitting(self):
        self._chk(self.L.orc_parcel_splitting(self.h), "parcel_splitting")

    def parcel_mixing(self):
        self._chk(self.L.orc_parcel_mixing(self.h), "parcel_mixing")

    def parcel_haloswap(self):
        self._chk(self.L.orc_parcel_haloswap(self.h), "parcel_haloswap")

    def writebinary_grids(self):
        self._chk(self.L.orc_writebinary_grids(self.h), "writebinary_grids")

    def step(self):
        st = self.L.orc_step(self.h)
        self._chk(st, "step")

    def step_status(self):
        return int(self.L.orc_step(self.h))

    def last_counts(self):
        return tuple(int(self.L.orc_last_counts(self.h, w)) for w in range(3))


def tridiagonal(a, b, c, d):
    a, b, c, d = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d))
    f = np.zeros(len(d))
    st = lib().orc_tridiagonal(len(d), _p(a), _p(b), _p(c), _p(d), _p(f))
    if st:
        raise ZeroDivisionError("zero pivot")
    return f


def r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    out = np.zeros((n // 2 + 1) * 2)
    lib().orc_r2c(n, _p(x), _p(out))
    return out[0::2] + 1j * out[1::2]


def c2r(z, n):
    z = np.asarray(z, dtype=np.complex128)
    inter = np.zeros((n // 2 + 1) * 2)
    inter[0::2] = z.real
    inter[1::2] = z.imag
    out = np.zeros(n)
    lib().orc_c2r(n, _p(inter), _p(out))
    return out


class RefFFTE:
    """The reference's own FFTE + ffte_mod wrappers, built from its sources into oracle/_ref."""

    def __init__(self):
        self.L = ctypes.CDLL(REF_FFTE_PATH)
        self.init = self.L._QMffte_modPffte_init
        self.fin = self.L._QMffte_modPffte_finalise
        self._r2c = self.L._QMffte_modPffte_r2c
        self._c2r = self.L._QMffte_modPffte_c2r

    @staticmethod
    def available():
        return os.path.exists(REF_FFTE_PATH)

    def r2c(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = ctypes.c_int(len(x))
        out = np.zeros(len(x) // 2 + 1, dtype=np.complex128)
        self.init(ctypes.byref(n))
        self._r2c(_p(x), _p(out), ctypes.byref(n))
        self.fin()
        return out

    def c2r(self, z, n):
        z = np.ascontiguousarray(z, dtype=np.complex128)
        nn = ctypes.c_int(n)
        out = np.zeros(n)
        self.init(ctypes.byref(nn))
        self._c2r(_p(z), _p(out), ctypes.byref(nn))
        self.fin()
        return out
