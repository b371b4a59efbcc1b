"""ctypes binding of the engine's C ABI (include/fwi2d.h, lib/libfwi2d.so).

Used by the Python tests and the benchmark; the product host side is the
C++ executables under ``fwi2d_amd/host``.  Loading fails loudly when the
library has not been built: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libfwi2d.so")

FWI_OK = 0
FWI_ERR_ARG = 1
FWI_ERR_STATE = 2
FWI_ERR_HIP = 3
FWI_ERR_NOMEM = 4
FWI_ERR_NUMERIC = 5
FWI_ERR_INTERNAL = 6

ABC = 0
RBC = 1

RESET_IMAGE = 1
RESET_GRADIENT = 2
RESET_ILLUM = 4


class FwiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[fwi2d error {code}] {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("nz", C.c_int32), ("nx", C.c_int32), ("dh", C.c_float),
        ("nt", C.c_int32), ("dt", C.c_float),
        ("abc_nb", C.c_int32), ("abc_factor", C.c_float),
        ("rbc_nb", C.c_int32), ("rbc_varVp", C.c_float), ("rbc_varVp_rel", C.c_float),
        ("rbc_seed", C.c_uint64), ("max_batch", C.c_int32),
        ("rbc_kind", C.c_int32), ("rbc_ratio", C.c_float),
    ]


RBC_UNIFORM, RBC_GAUSSIAN = 0, 1


def rb_blobs(nz, nx, dh, rbc_nb, varVp, ratio, seed, shot, iteration=0):
    """Blobs of the FWI_RBC_GAUSSIAN boundary of one shot and iteration, in
    sample order, as an [n, 5] array (x, z, 16 r^2, 0.5 / r^2, amplitude);
    host only (fwi_rb_blobs)."""
    cfg = Config(nz, nx, dh, 2, 1e-3, 0, 0.0, rbc_nb, varVp, 0.0, seed, 1, RBC_GAUSSIAN, ratio)
    n = C.c_int64()
    _check(lib().fwi_rb_blobs(C.byref(cfg), shot, iteration, np.zeros(5, np.float32), 0,
                              C.byref(n)))
    out = np.zeros(5 * max(n.value, 1), np.float32)
    _check(lib().fwi_rb_blobs(C.byref(cfg), shot, iteration, out, n.value, C.byref(n)))
    return out[:5 * n.value].reshape(n.value, 5)


_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_ctx = C.c_void_p

# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("fwi_last_error", C.c_char_p, []),
    ("fwi_version", C.c_char_p, []),
    ("fwi_create", C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    ("fwi_destroy", C.c_int, [_ctx]),
    ("fwi_ricker_wavelet", C.c_int, [C.c_int32, C.c_float, C.c_float, _f32p,
                                     C.POINTER(C.c_int32)]),
    ("fwi_set_wavelet", C.c_int, [_ctx, _f32p, C.c_int32]),
    ("fwi_set_geometry", C.c_int, [_ctx, C.c_int32, _f32p, _f32p, C.c_int32, _f32p, _f32p]),
    ("fwi_set_model", C.c_int, [_ctx, _f32p]),
    ("fwi_set_observed", C.c_int, [_ctx, C.c_int32, C.c_int32, _f32p]),
    ("fwi_modeling", C.c_int, [_ctx, C.c_int32, C.c_int32, C.c_int32, _vp]),
    ("fwi_migration", C.c_int, [_ctx, C.c_int32, C.c_int32]),
    ("fwi_gradient", C.c_int, [_ctx, C.c_int32, C.c_int32, C.c_uint32,
                               C.POINTER(C.c_double)]),
    ("fwi_misfit", C.c_int, [_ctx, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    ("fwi_reset", C.c_int, [_ctx, C.c_int32]),
    ("fwi_get_image", C.c_int, [_ctx, _vp, _vp, _vp]),
    ("fwi_get_gradient", C.c_int, [_ctx, C.c_float, _vp, _vp, _vp, C.POINTER(C.c_double)]),
    ("fwi_parabola_step", C.c_int, [_f64p, _f64p, C.POINTER(C.c_double)]),
    ("fwi_guarded_step", C.c_int, [_f64p, _f64p, _f32p, _f32p, C.c_int64, C.c_double,
                                   C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    ("fwi_rb_velocity", C.c_int, [_ctx, C.c_int32, C.c_uint32, _f32p]),
    ("fwi_rb_velocity_noalloc", C.c_int, [_ctx, C.c_int32, C.c_uint32, _f32p]),
    ("fwi_rb_blobs", C.c_int, [C.POINTER(Config), C.c_int32, C.c_uint32, _f32p, C.c_int64,
                               C.POINTER(C.c_int64)]),
    ("fwi_get_accumulator", C.c_int, [_ctx, C.c_int32, _f32p]),
    ("fwi_set_accumulator", C.c_int, [_ctx, C.c_int32, _f32p]),
    ("fwi_get_config", C.c_int, [_ctx, C.POINTER(Config)]),
    ("fwi_geometry_size", C.c_int, [_ctx, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("fwi_grid_shape", C.c_int, [_ctx, C.c_int32, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("fwi_synchronize", C.c_int, [_ctx]),
    ("fwi_set_tuning", C.c_int, [_ctx, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    ("fwi_get_tuning", C.c_int, [_ctx, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("fwi_autotune", C.c_int, [_ctx, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]),
    ("fwi_stream_copy", C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    ("fwi_profile_enable", C.c_int, [_ctx, C.c_int32]),
    ("fwi_profile_read", C.c_int, [_ctx, C.c_int32, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
]

NN_SIGNATURES = [
    ("fwi_nn_create", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                C.POINTER(C.c_void_p)]),
    ("fwi_nn_destroy", C.c_int, [_ctx]),
    ("fwi_nn_get_shape", C.c_int, [_ctx, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]),
    ("fwi_nn_num_params", C.c_int, [_ctx, C.POINTER(C.c_int64)]),
    ("fwi_nn_get_params", C.c_int, [_ctx, _f32p]),
    ("fwi_nn_set_params", C.c_int, [_ctx, _f32p]),
    ("fwi_nn_get_optimizer", C.c_int, [_ctx, _f32p, _f32p, C.POINTER(C.c_int64)]),
    ("fwi_nn_set_optimizer", C.c_int, [_ctx, _f32p, _f32p, C.c_int64]),
    ("fwi_nn_forward", C.c_int, [_ctx, C.c_int32, _f32p, _f32p]),
    ("fwi_nn_gradient", C.c_int, [_ctx, C.c_int32, _f32p, _f32p, C.POINTER(C.c_float), _vp]),
    ("fwi_nn_train_step", C.c_int, [_ctx, C.c_int32, _f32p, _f32p, C.c_float,
                                    C.POINTER(C.c_float)]),
    ("fwi_nn_mfma_probe", C.c_int, [_f32p, _f32p, _f32p]),
]

TRIAL_NONE, TRIAL_PLAIN, TRIAL_NETWORK, TRIAL_GIVEN = 0, 1, 2, 3


class UpdateOpts(C.Structure):
    """fwi_update_opts (include/fwi2d_update.h)."""
    _fields_ = [("alpha1", C.c_double), ("alpha2", C.c_double), ("m_min", C.c_double),
                ("zmask", C.c_float), ("trial", C.c_int32), ("net", C.c_void_p),
                ("direction", C.c_void_p), ("nmultiples", C.c_int32),
                ("multiples", C.c_double * 4)]


class UpdateInfo(C.Structure):
    """fwi_update_info (include/fwi2d_update.h)."""
    _fields_ = [("f", C.c_double * 3), ("step", C.c_double), ("rule", C.c_int32),
                ("plain_misfit", C.c_double), ("trial_misfit", C.c_double),
                ("trial_multiple", C.c_double), ("taken", C.c_int32), ("passes", C.c_int32)]

    def as_dict(self):
        return {"f": list(self.f), "step": self.step, "rule": self.rule,
                "plain_misfit": self.plain_misfit, "trial_misfit": self.trial_misfit,
                "trial_multiple": self.trial_multiple, "taken": self.taken,
                "passes": self.passes}


UPDATE_SIGNATURES = [
    ("fwi_update_model", C.c_int, [_ctx, C.POINTER(UpdateOpts), C.c_double, _f32p, _vp, _vp,
                                   _f32p, C.POINTER(UpdateInfo)]),
    ("fwi_trial_direction", C.c_int, [_ctx, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                      _f32p, _f32p, _f32p, _f32p, _f32p]),
]

_lib = None


def lib():
    """Loads lib/libfwi2d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` or `make -C fwi2d_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES + NN_SIGNATURES + UPDATE_SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != FWI_OK:
        raise FwiError(rc, lib().fwi_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def ricker_wavelet(nt, dt, fmax):
    """Host wavelet (FP64 DFT rounded to FP32) and its tlag."""
    w = np.zeros(nt, np.float32)
    tl = C.c_int32(0)
    _check(lib().fwi_ricker_wavelet(nt, dt, fmax, w, C.byref(tl)))
    return w, tl.value


def stream_copy(nbytes, reps=20):
    """Device copy bandwidth in GB/s (read + write) of a plain float4 copy."""
    g = C.c_double()
    _check(lib().fwi_stream_copy(int(nbytes), reps, C.byref(g)))
    return g.value


def parabola_step(alpha, f):
    s = C.c_double(0.0)
    _check(lib().fwi_parabola_step(np.ascontiguousarray(alpha, np.float64),
                                   np.ascontiguousarray(f, np.float64), C.byref(s)))
    return s.value


def guarded_step(alpha, f, m, g, m_min=0.0):
    """(step, rule) of fwi_guarded_step for the update m - step * g."""
    m = np.ascontiguousarray(m, np.float32).ravel()
    g = np.ascontiguousarray(g, np.float32).ravel()
    if m.size != g.size:
        raise ValueError("m and g differ in size")
    s, r = C.c_double(0.0), C.c_int32(0)
    _check(lib().fwi_guarded_step(np.ascontiguousarray(alpha, np.float64),
                                  np.ascontiguousarray(f, np.float64), m, g, m.size,
                                  float(m_min), C.byref(s), C.byref(r)))
    return s.value, r.value


class Engine:
    """One survey context on the GPU (thin owner of an ``fwi_ctx``)."""

    def __init__(self, nz, nx, dh, nt, dt, abc_nb=0, abc_factor=0.0, rbc_nb=0,
                 rbc_varVp=0.0, rbc_varVp_rel=0.0, rbc_seed=0, max_batch=1,
                 rbc_kind=RBC_UNIFORM, rbc_ratio=0.0):
        cfg = Config(nz, nx, dh, nt, dt, abc_nb, abc_factor, rbc_nb, rbc_varVp,
                     rbc_varVp_rel, rbc_seed, max_batch, rbc_kind, rbc_ratio)
        h = C.c_void_p()
        _check(lib().fwi_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.nz, self.nx, self.nt = nz, nx, nt
        self.nsrc = 0
        self.nrec = 0

    def close(self):
        if self._h:
            lib().fwi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_wavelet(self, w, tlag):
        _check(lib().fwi_set_wavelet(self._h, np.ascontiguousarray(w, np.float32), tlag))

    def set_geometry(self, xsrc, zsrc, xrec, zrec):
        f = lambda a: np.ascontiguousarray(a, np.float32)
        xs, zs, xr, zr = f(xsrc), f(zsrc), f(xrec), f(zrec)
        _check(lib().fwi_set_geometry(self._h, xs.size, xs, zs, xr.size, xr, zr))
        self.nsrc, self.nrec = xs.size, xr.size

    def set_model(self, vp):
        """vp: [nx, nz] float32 (z fastest)."""
        _check(lib().fwi_set_model(self._h, np.ascontiguousarray(vp, np.float32)))

    def set_observed(self, data, shot0=0):
        data = np.ascontiguousarray(data, np.float32)
        n = data.shape[0]
        _check(lib().fwi_set_observed(self._h, shot0, n, data))

    def modeling(self, boundary, shot0=0, nshots=None, record=True):
        nshots = self.nsrc - shot0 if nshots is None else nshots
        out = np.zeros((nshots, self.nrec, self.nt), np.float32) if record else None
        _check(lib().fwi_modeling(self._h, boundary, shot0, nshots, _p(out)))
        return out

    def migration(self, shot0=0, nshots=None):
        nshots = self.nsrc - shot0 if nshots is None else nshots
        _check(lib().fwi_migration(self._h, shot0, nshots))

    def gradient(self, shot0=0, nshots=None, iteration=0):
        nshots = self.nsrc - shot0 if nshots is None else nshots
        s = C.c_double(0.0)
        _check(lib().fwi_gradient(self._h, shot0, nshots, iteration, C.byref(s)))
        return s.value

    def misfit(self, shot0=0, nshots=None):
        nshots = self.nsrc - shot0 if nshots is None else nshots
        s = C.c_double(0.0)
        _check(lib().fwi_misfit(self._h, shot0, nshots, C.byref(s)))
        return s.value

    def reset(self, what=RESET_IMAGE | RESET_GRADIENT | RESET_ILLUM):
        _check(lib().fwi_reset(self._h, what))

    def get_image(self):
        """Returns (image, raw_image, raw_illum), each [nx, nz]."""
        shp = (self.nx, self.nz)
        img, raw, ill = (np.zeros(shp, np.float32) for _ in range(3))
        _check(lib().fwi_get_image(self._h, _p(img), _p(raw), _p(ill)))
        return img, raw, ill

    def get_gradient(self, zmask):
        """Returns (direction, raw_grad, raw_illum, gmax)."""
        shp = (self.nx, self.nz)
        g, raw, ill = (np.zeros(shp, np.float32) for _ in range(3))
        gm = C.c_double(0.0)
        _check(lib().fwi_get_gradient(self._h, zmask, _p(g), _p(raw), _p(ill), C.byref(gm)))
        return g, raw, ill, gm.value

    def get_accumulator(self, what):
        """Shot-sum image, gradient or illumination (what = RESET_*), [nxx, nzz]."""
        nzz, nxx, _ = self.grid_shape(RBC)
        out = np.zeros((nxx, nzz), np.float32)
        _check(lib().fwi_get_accumulator(self._h, what, out))
        return out

    def set_accumulator(self, what, values):
        _check(lib().fwi_set_accumulator(self._h, what, np.ascontiguousarray(values, np.float32)))

    def rb_velocity(self, shot, iteration=0, noalloc=False):
        """Random-boundary velocity [nxx, nzz] of a shot's pass; noalloc uses
        the entry point that stages through the library's static device
        buffer instead of allocating one."""
        nzz, nxx, _ = self.grid_shape(RBC)
        out = np.zeros((nxx, nzz), np.float32)
        fn = lib().fwi_rb_velocity_noalloc if noalloc else lib().fwi_rb_velocity
        _check(fn(self._h, shot, iteration, out))
        return out

    def grid_shape(self, boundary):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().fwi_grid_shape(self._h, boundary, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def synchronize(self):
        _check(lib().fwi_synchronize(self._h))

    def set_tuning(self, vz=0, xw=0, use_graph=-1, kernel=-1, steps=0):
        _check(lib().fwi_set_tuning(self._h, kernel, steps, vz, xw, use_graph))

    def get_tuning(self, kernel=0):
        """(steps per launch, rows per lane, chunk width, graph on)."""
        st, a, b, c = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().fwi_get_tuning(self._h, kernel, C.byref(st), C.byref(a), C.byref(b),
                                    C.byref(c)))
        return st.value, a.value, b.value, c.value

    def get_config(self):
        cfg = Config()
        _check(lib().fwi_get_config(self._h, C.byref(cfg)))
        return cfg

    def update_model(self, m, g, f0, alpha1, alpha2, m_min, zmask=0.0, trial=TRIAL_NONE,
                     net=None, direction=None, multiples=(2.0, 4.0), raw=None, illum=None):
        """fwi_update_model: the line search, the step and the update of
        m (slowness squared, [nx, nz]; a new array is returned) with an
        optional trial direction.  Returns (new m, info dict)."""
        m2 = np.array(m, np.float32, order="C", copy=True)
        g = np.ascontiguousarray(g, np.float32)
        raw = None if raw is None else np.ascontiguousarray(raw, np.float32)
        illum = None if illum is None else np.ascontiguousarray(illum, np.float32)
        d = None if direction is None else np.ascontiguousarray(direction, np.float32)
        for a in (g, raw, illum, d):
            if a is not None and a.shape != m2.shape:
                raise ValueError("m, g, raw gradient, illumination and direction must have "
                                 "one shape")
        o = UpdateOpts()
        o.alpha1, o.alpha2, o.m_min, o.zmask, o.trial = alpha1, alpha2, m_min, zmask, trial
        o.net = net._h if net is not None else None
        o.direction = _p(d)
        o.nmultiples = len(multiples)
        for k, v in enumerate(multiples):
            o.multiples[k] = v
        info = UpdateInfo()
        _check(lib().fwi_update_model(self._h, C.byref(o), float(f0), g, _p(raw), _p(illum),
                                      m2, C.byref(info)))
        return m2, info.as_dict()

    def autotune(self, nslots, reps=10):
        f, r = C.c_double(), C.c_double()
        _check(lib().fwi_autotune(self._h, nslots, reps, C.byref(f), C.byref(r)))
        return f.value, r.value

    def profile_enable(self, on=True):
        _check(lib().fwi_profile_enable(self._h, 1 if on else 0))

    def profile_read(self, kind):
        ms, n, cells = C.c_double(), C.c_int64(), C.c_double()
        _check(lib().fwi_profile_read(self._h, kind, C.byref(ms), C.byref(n), C.byref(cells)))
        return ms.value, n.value, cells.value


def mfma_probe(a, b):
    """d = a @ b through one bf16 MFMA (a: 16x32, b: 32x16)."""
    d = np.zeros((16, 16), np.float32)
    _check(lib().fwi_nn_mfma_probe(np.ascontiguousarray(a, np.float32),
                                   np.ascontiguousarray(b, np.float32), d))
    return d


class Net:
    """The learned preconditioner network on the GPU (owner of an ``fwi_nn``)."""

    def __init__(self, batch, height, width, seed=0):
        h = C.c_void_p()
        _check(lib().fwi_nn_create(batch, height, width, seed, C.byref(h)))
        self._h = h
        self.batch, self.height, self.width = batch, height, width
        n = C.c_int64()
        _check(lib().fwi_nn_num_params(h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if self._h:
            lib().fwi_nn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def direction(self, nx, nz, dh, zmask, raw, illum, m, g):
        """fwi_trial_direction: the network's search direction [nx, nz]."""
        d = np.zeros((nx, nz), np.float32)
        arr = [np.ascontiguousarray(a, np.float32) for a in (raw, illum, m, g)]
        if any(a.shape != (nx, nz) for a in arr):
            raise ValueError(f"inputs must be {nx} x {nz}")
        _check(lib().fwi_trial_direction(self._h, nx, nz, dh, zmask, *arr, d))
        return d

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        _check(lib().fwi_nn_get_params(self._h, p))
        return p

    def set_params(self, p):
        _check(lib().fwi_nn_set_params(self._h, np.ascontiguousarray(p, np.float32)))

    def get_optimizer(self):
        """Adam state (m, v, steps taken) in the flat parameter layout."""
        m = np.zeros(self.nparams, np.float32)
        v = np.zeros(self.nparams, np.float32)
        t = C.c_int64()
        _check(lib().fwi_nn_get_optimizer(self._h, m, v, C.byref(t)))
        return m, v, t.value

    def set_optimizer(self, m, v, steps):
        m = np.ascontiguousarray(m, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        if m.size != self.nparams or v.size != self.nparams:
            raise ValueError(f"Adam moments need {self.nparams} values each")
        _check(lib().fwi_nn_set_optimizer(self._h, m, v, int(steps)))

    def forward(self, x):
        """x: [nb, 3, H, W] -> y: [nb, H, W]."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.height, self.width), np.float32)
        _check(lib().fwi_nn_forward(self._h, x.shape[0], x, y))
        return y

    def gradient(self, x, target):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        _check(lib().fwi_nn_gradient(self._h, x.shape[0], x, t, C.byref(loss), _p(g)))
        return loss.value, g

    def train_step(self, x, target, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_float()
        _check(lib().fwi_nn_train_step(self._h, x.shape[0], x, t, lr, C.byref(loss)))
        return loss.value
