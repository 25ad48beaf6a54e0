"""ctypes wrapper of oracle/_ref/libsdft_ref.so (the reference's own sources
compiled by oracle/build_ref.sh). TEST INFRASTRUCTURE ONLY: used in this
container to pin the oracle and to write tests/golden fixtures. The library
is absent on machines without /root/reference; callers skip then."""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libsdft_ref.so")
_LIB = None
_dp = ct.POINTER(ct.c_double)
CB = ct.CFUNCTYPE(None, ct.c_void_p, ct.c_int, _dp, ct.c_long, _dp, _dp)


def available() -> bool:
    return os.path.exists(PATH)


def lib():
    global _LIB
    if _LIB is None:
        _LIB = ct.CDLL(PATH)
        _LIB.ref_last_error.restype = ct.c_char_p
        _LIB.ref_kinetic_preconditioner.restype = ct.c_double
    return _LIB


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def _check(rc):
    if rc:
        raise RuntimeError(lib().ref_last_error().decode())


def enumerate_g(reciprocal, cutoff_ev, kfrac, ev_per_ha=27.211386245988):
    r = np.ascontiguousarray(reciprocal, dtype=np.float64)
    k = np.ascontiguousarray(kfrac, dtype=np.float64)
    count = ct.c_long()
    _check(lib().ref_enumerate_g(_p(r), ct.c_double(cutoff_ev), _p(k), ct.c_double(ev_per_ha), ct.c_long(0), None,
                                 None, ct.byref(count)))
    n = count.value
    hkl = np.zeros((n, 3), dtype=np.int32)
    kin = np.zeros(n)
    _check(lib().ref_enumerate_g(_p(r), ct.c_double(cutoff_ev), _p(k), ct.c_double(ev_per_ha), ct.c_long(n),
                                 _p(hkl, ct.POINTER(ct.c_int)), _p(kin), ct.byref(count)))
    return hkl, kin


def kinetic_preconditioner(t, wk):
    return lib().ref_kinetic_preconditioner(ct.c_double(t), ct.c_double(wk))


def compensated_product(aop, a, b):
    a = np.asfortranarray(a, dtype=np.complex128)
    b = np.asfortranarray(b, dtype=np.complex128)
    rows = a.shape[1] if aop else a.shape[0]
    inner = a.shape[0] if aop else a.shape[1]
    out = np.zeros((rows, b.shape[1]), dtype=np.complex128, order="F")
    _check(lib().ref_compensated_product(ct.c_int(aop), ct.c_long(rows), ct.c_long(b.shape[1]), ct.c_long(inner),
                                         _p(a), ct.c_long(a.shape[0]), _p(b), ct.c_long(b.shape[0]), _p(out),
                                         ct.c_long(rows), ct.c_int(0)))
    return out


def davidson(oracle_op, roots, max_iterations=256, tolerance=1e-12, initial=None, band_group_size=0):
    """The reference Davidson (or, with band_group_size > 0, the reference's
    bounded band-group solver) driving the oracle operator through callbacks."""
    n = oracle_op.n

    def cb(user, which, x, cols, hy, sy):
        xa = np.ctypeslib.as_array(x, shape=(2 * n * cols,)).view(np.complex128).reshape(cols, n).T
        if which == 2:
            h, s = oracle_op.apply(xa, "pair")
            np.ctypeslib.as_array(hy, shape=(2 * n * cols,))[:] = np.asfortranarray(h).ravel(order="F").view(np.float64)
            np.ctypeslib.as_array(sy, shape=(2 * n * cols,))[:] = np.asfortranarray(s).ravel(order="F").view(np.float64)
        else:
            y = oracle_op.apply(xa, "h" if which == 0 else "s")
            out = hy if which == 0 else sy
            np.ctypeslib.as_array(out, shape=(2 * n * cols,))[:] = np.asfortranarray(y).ravel(order="F").view(np.float64)

    cbf = CB(cb)
    ev = np.zeros(roots)
    res = np.zeros(roots)
    vec = np.zeros((n, roots), dtype=np.complex128, order="F")
    it = ct.c_long()
    conv = ct.c_int()
    init = None if initial is None else np.asfortranarray(initial[:, :roots], dtype=np.complex128)
    _check(lib().ref_davidson_callbacks(ct.c_long(n), _p(oracle_op.h_diag), _p(oracle_op.s_diag),
                                        _p(oracle_op.kinetic), cbf, None, ct.c_long(roots),
                                        ct.c_long(max_iterations), ct.c_double(tolerance), _p(init), _p(ev),
                                        _p(vec), _p(res), ct.byref(it), ct.byref(conv),
                                        ct.c_long(band_group_size)))
    return dict(eigenvalues=ev, eigenvectors=vec, residual_norms=res, iterations=it.value, converged=bool(conv.value))
