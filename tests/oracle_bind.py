"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests, ``__graft_entry__.smoke`` and
the CPU-baseline leg of ``bench.py``, always as the checker.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.or_wavelet.restype = C.c_int
        L.or_wavelet.argtypes = [C.c_int, C.c_double, C.c_double, _f64p]
        L.or_dft_r2c.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.or_dft_c2r.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.or_abc_b1d.argtypes = [C.c_int, C.c_double, _f64p]
        L.or_damper.restype = C.c_double
        L.or_damper.argtypes = [_f64p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_expand.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, _f32p]
        L.or_reduce.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, _f32p]
        L.or_grid_index.restype = C.c_int
        L.or_grid_index.argtypes = [C.c_float, C.c_float, C.c_int]
        L.or_image_post.argtypes = [_f64p, _f64p, C.c_int, C.c_int, C.c_double, _f64p]
        L.or_gradient_post.argtypes = [_f64p, _f64p, C.c_int, C.c_int, C.c_double,
                                       C.c_double, _f64p]
        L.or_parabola_step.restype = C.c_double
        L.or_parabola_step.argtypes = [_f64p, _f64p]
        L.or_rb_mix64.restype = C.c_uint64
        L.or_rb_mix64.argtypes = [C.c_uint64]
        L.or_rb_uniform.restype = C.c_float
        L.or_rb_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]
        L.or_rb_velocity.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                     C.c_float, C.c_float, C.c_float, C.c_uint64,
                                     C.c_uint64, _f32p]
        L.or_rb_blobs.restype = C.c_long
        L.or_rb_blobs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                  C.c_uint64, C.c_uint64, C.c_void_p, C.c_long]
        L.or_rb_normal.restype = C.c_float
        L.or_rb_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]
        L.or_rb_expneg.restype = C.c_float
        L.or_rb_expneg.argtypes = [C.c_float]
        L.or_rb_velocity_gauss.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                           C.c_float, C.c_float, C.c_float, C.c_float,
                                           C.c_uint64, C.c_uint64, _f32p]
        for sfx, fp in (("f64", _f64p), ("f32", _f32p)):
            fwd = getattr(L, "or_forward_" + sfx)
            fwd.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                            C.c_int, _f32p, fp, C.c_int, C.c_int, C.c_int, _i32p, _i32p,
                            C.c_int, _f64p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
            lap = getattr(L, "or_laplacian_" + sfx)
            lap.argtypes = [C.c_int, C.c_int, C.c_double, fp, fp]
            rev = getattr(L, "or_reverse_" + sfx)
            rev.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                            _f32p, fp, fp, C.c_int, _i32p, _i32p, fp, C.c_int, fp, fp,
                            C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def wavelet(nt, dt, fmax):
    w = np.zeros(nt, np.float64)
    tlag = lib().or_wavelet(nt, dt, fmax, w)
    return w, tlag


def abc_b1d(nb, factor):
    b = np.zeros(max(nb, 1), np.float64)
    lib().or_abc_b1d(nb, factor, b)
    return b


def expand(v, nb):
    """v: [nx, nz] float32 (z fastest)."""
    nx, nz = v.shape
    out = np.zeros((nx + 2 * nb, nz + 2 * nb), np.float32)
    lib().or_expand(np.ascontiguousarray(v, np.float32), nz, nx, nb, out)
    return out


def reduce(vpad, nb):
    nxx, nzz = vpad.shape
    nx, nz = nxx - 2 * nb, nzz - 2 * nb
    out = np.zeros((nx, nz), np.float32)
    lib().or_reduce(np.ascontiguousarray(vpad, np.float32), nz, nx, nb, out)
    return out


def rb_velocity(vpad, nb, amp_abs, amp_rel, vlo, vhi, seed, stream):
    nxx, nzz = vpad.shape
    out = np.zeros_like(vpad, dtype=np.float32)
    lib().or_rb_velocity(np.ascontiguousarray(vpad, np.float32), nzz, nxx, nb,
                         amp_abs, amp_rel, vlo, vhi, seed, stream, out)
    return out


def rb_blobs(nzz, nxx, nb, dh, ratio, amp, seed, stream):
    """Blobs of the reference-style boundary of one stream: [n, 5] (x, z,
    16 r^2, 0.5 / r^2, amplitude) in sample order."""
    n = lib().or_rb_blobs(nzz, nxx, nb, dh, ratio, amp, seed, stream, None, 0)
    out = np.zeros(5 * max(n, 1), np.float32)
    lib().or_rb_blobs(nzz, nxx, nb, dh, ratio, amp, seed, stream,
                      out.ctypes.data_as(C.c_void_p), n)
    return out[:5 * n].reshape(n, 5)


def rb_velocity_gauss(vpad, nb, dh, varVp, ratio, vlo, vhi, seed, stream):
    nxx, nzz = vpad.shape
    out = np.zeros_like(vpad, dtype=np.float32)
    lib().or_rb_velocity_gauss(np.ascontiguousarray(vpad, np.float32), nzz, nxx, nb, dh, varVp,
                               ratio, vlo, vhi, seed, stream, out)
    return out


def forward(vpad, dh, dt, nt, tlag, w, siz, six, riz, rix, abc=False, b1d=None,
            nb=0, precision="f64", want_fields=True, nthreads=0):
    """Full forward pass; returns (seis[nrec, nt], P=u_T, Q=u_{T-1})."""
    nxx, nzz = vpad.shape
    dtype = np.float64 if precision == "f64" else np.float32
    riz = np.ascontiguousarray(riz, np.int32)
    rix = np.ascontiguousarray(rix, np.int32)
    nrec = riz.size
    seis = np.zeros((nrec, nt), dtype)
    P = np.zeros((nxx, nzz), dtype) if want_fields else None
    Q = np.zeros((nxx, nzz), dtype) if want_fields else None
    b = np.ascontiguousarray(b1d if b1d is not None else np.zeros(1), np.float64)
    fn = getattr(lib(), "or_forward_" + precision)
    fn(nzz, nxx, nb, dh, dt, nt, tlag, np.ascontiguousarray(vpad, np.float32),
       np.ascontiguousarray(w, dtype), siz, six, nrec, riz, rix, 1 if abc else 0, b,
       _ptr(seis), _ptr(P), _ptr(Q), nthreads)
    return seis, P, Q


def reverse(vpad, dh, dt, nt, tlag, Ps, Psold, riz, rix, data, mode,
            acc=None, illum=None, precision="f64", nthreads=0):
    """Reverse pass; Ps = u_T, Psold = u_{T-1} (copied).  Returns (acc, illum)."""
    nxx, nzz = vpad.shape
    dtype = np.float64 if precision == "f64" else np.float32
    if acc is None:
        acc = np.zeros((nxx, nzz), dtype)
    if illum is None:
        illum = np.zeros((nxx, nzz), dtype)
    riz = np.ascontiguousarray(riz, np.int32)
    rix = np.ascontiguousarray(rix, np.int32)
    fn = getattr(lib(), "or_reverse_" + precision)
    fn(nzz, nxx, dh, dt, nt, tlag, np.ascontiguousarray(vpad, np.float32),
       np.array(Ps, dtype=dtype, copy=True), np.array(Psold, dtype=dtype, copy=True),
       riz.size, riz, rix, np.ascontiguousarray(data, dtype), mode, acc, illum, nthreads)
    return acc, illum


def laplacian(P, dh, precision="f64"):
    """P: [nxx, nzz]; returns the oracle's Laplacian (0 on the 4-cell rim)."""
    dtype = np.float64 if precision == "f64" else np.float32
    P = np.ascontiguousarray(P, dtype)
    out = np.zeros_like(P)
    getattr(lib(), "or_laplacian_" + precision)(P.shape[1], P.shape[0], dh, P, out)
    return out


def image_post(image, illum, dh):
    """image, illum: cropped [nx, nz]."""
    nx, nz = image.shape
    out = np.zeros((nx, nz), np.float64)
    lib().or_image_post(np.ascontiguousarray(image, np.float64),
                        np.ascontiguousarray(illum, np.float64), nz, nx, dh, out)
    return out


def gradient_post(grad, illum, dh, zmask):
    nx, nz = grad.shape
    out = np.zeros((nx, nz), np.float64)
    lib().or_gradient_post(np.ascontiguousarray(grad, np.float64),
                           np.ascontiguousarray(illum, np.float64), nz, nx, dh, zmask, out)
    return out


def parabola_step(a, f):
    return lib().or_parabola_step(np.asarray(a, np.float64), np.asarray(f, np.float64))


def grid_index(x, dh, nb):
    return lib().or_grid_index(x, dh, nb)
