"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() (as the
checker) and bench.py's CPU-baseline leg.  The product library never loads it.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
_lp = ct.POINTER(ct.c_long)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ct.CDLL(path)
        _LIB.orc_last_error.restype = ct.c_char_p
    return _LIB


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def _check(rc):
    if rc:
        msg = lib().orc_last_error().decode()
        raise (ValueError if rc == 2 else RuntimeError)(msg)


def enumerate_g(lattice, cutoff_ev, kfrac, ev_per_ha=27.211386245988):
    L = lib()
    lat = np.ascontiguousarray(lattice, dtype=np.float64)
    kf = np.ascontiguousarray(kfrac, dtype=np.float64)
    count = ct.c_long(0)
    _check(L.orc_enumerate_g(_p(lat), ct.c_double(cutoff_ev), _p(kf), ct.c_double(ev_per_ha), ct.c_long(0),
                             None, None, None, ct.byref(count)))
    n = count.value
    hkl = np.zeros((n, 3), dtype=np.int32)
    kin = np.zeros(n)
    cart = np.zeros((n, 3))
    _check(L.orc_enumerate_g(_p(lat), ct.c_double(cutoff_ev), _p(kf), ct.c_double(ev_per_ha), ct.c_long(n),
                             _p(hkl, _ip), _p(kin), _p(cart), ct.byref(count)))
    return hkl, kin, cart


def numpy_openblas_path():
    """numpy's bundled ILP64 OpenBLAS (CBLAS + LAPACKE), or None."""
    import glob
    d = os.path.join(os.path.dirname(os.path.dirname(np.__file__)), "numpy.libs")
    hits = sorted(glob.glob(os.path.join(d, "libscipy_openblas64_*.so")))
    return hits[0] if hits else None


def bind_blas(path):
    """Route the oracle's GEMMs and dense eigensolves through CBLAS/LAPACKE
    (the reference's CPU configuration with BLAS/LAPACK); None restores the
    plain loops."""
    _check(lib().orc_bind_blas(None if path is None else path.encode()))


def set_threads(n):
    """OpenMP threads of the oracle and of the bound BLAS."""
    _check(lib().orc_set_threads(ct.c_int(int(n))))


def fast_fft_length(m):
    return lib().orc_fast_fft_length(ct.c_int(m))


def fft3d(grid, sign):
    g = np.ascontiguousarray(grid, dtype=np.complex128).copy()
    shape = np.array(g.shape, dtype=np.int32)
    _check(lib().orc_fft3d(_p(shape, _ip), _p(g), ct.c_int(sign)))
    return g


class Operator:
    """The reference plane-wave operator restated on the CPU (FP64)."""

    def __init__(self, prob, shape=None):
        L = lib()
        self.n = prob.n
        hkl = np.ascontiguousarray(prob.hkl, dtype=np.int32)
        kin = np.ascontiguousarray(prob.kinetic_ev, dtype=np.float64)
        vh = np.ascontiguousarray(prob.v_hkl, dtype=np.int32)
        vv = np.ascontiguousarray(np.stack([prob.v_values.real, prob.v_values.imag], axis=1))
        rows = np.ascontiguousarray(prob.atom_rows, dtype=np.int64)
        amp = np.ascontiguousarray(np.concatenate([a.ravel() for a in prob.amplitudes]))
        d = np.ascontiguousarray(np.concatenate([a.ravel() for a in prob.d_blocks]))
        q = np.ascontiguousarray(np.concatenate([a.ravel() for a in prob.q_blocks]))
        shp = np.array(shape if shape is not None else prob.shape, dtype=np.int32)
        h = ct.c_void_p()
        _check(L.orc_operator_create(ct.c_long(prob.n), _p(hkl, _ip), _p(kin), ct.c_double(27.211386245988),
                                     ct.c_long(len(vh)), _p(vh, _ip), _p(vv), ct.c_long(len(rows)), _p(rows, _lp),
                                     _p(amp), _p(d), _p(q),
                                     ct.c_double(prob.regularization), _p(shp, _ip), ct.byref(h)))
        self.handle = h
        shape_out = np.zeros(3, dtype=np.int32)
        count = ct.c_long()
        npj = ct.c_long()
        _check(L.orc_operator_info(h, _p(shape_out, _ip), ct.byref(count), ct.byref(npj), None, None, None, None,
                                   None, None))
        self.shape = tuple(int(s) for s in shape_out)
        self.count = count.value
        self.np = npj.value
        self.indices = np.zeros(self.n, dtype=np.uint64)
        self.kinetic = np.zeros(self.n)
        self.local_real = np.zeros(self.count)
        self.h_diag = np.zeros(self.n)
        self.s_diag = np.zeros(self.n)
        _check(L.orc_operator_info(h, _p(shape_out, _ip), ct.byref(count), ct.byref(npj),
                                   self.indices.ctypes.data_as(ct.POINTER(ct.c_ulonglong)), _p(self.kinetic),
                                   _p(self.local_real), _p(self.h_diag), _p(self.s_diag), None))

    def __del__(self):
        try:
            lib().orc_operator_destroy(self.handle)
        except Exception:
            pass

    def apply(self, x, which):
        """which: 'h', 's' or 'pair'. x: n x m complex (any order)."""
        x = np.asfortranarray(x, dtype=np.complex128)
        m = x.shape[1]
        hy = np.zeros((self.n, m), dtype=np.complex128, order="F")
        sy = np.zeros((self.n, m), dtype=np.complex128, order="F")
        code = {"h": 0, "s": 1, "pair": 2}[which]
        _check(lib().orc_operator_apply(self.handle, ct.c_int(code), _p(x), ct.c_long(m),
                                        _p(hy), _p(sy)))
        return {"h": hy, "s": sy, "pair": (hy, sy)}[which]

    def davidson(self, roots, max_iterations=256, tolerance=1e-12, initial=None, subspace_factor=8):
        ev = np.zeros(roots)
        res = np.zeros(roots)
        vec = np.zeros((self.n, roots), dtype=np.complex128, order="F")
        it = ct.c_long()
        conv = ct.c_int()
        hcols = ct.c_long()
        init = None if initial is None else np.asfortranarray(initial[:, :roots], dtype=np.complex128)
        _check(lib().orc_davidson(self.handle, ct.c_long(roots), ct.c_long(max_iterations), ct.c_double(tolerance),
                                  None if init is None else _p(init), ct.c_long(subspace_factor),
                                  _p(ev), _p(vec), _p(res), ct.byref(it), ct.byref(conv),
                                  ct.byref(hcols)))
        return dict(eigenvalues=ev, eigenvectors=vec, residual_norms=res, iterations=it.value,
                    converged=bool(conv.value), h_columns=hcols.value)


def zgemm(aop, bop, a, b, c=None, add=False):
    a = np.asfortranarray(a, dtype=np.complex128)
    b = np.asfortranarray(b, dtype=np.complex128)
    m = a.shape[1] if aop else a.shape[0]
    k = a.shape[0] if aop else a.shape[1]
    n = b.shape[0] if bop else b.shape[1]
    out = np.zeros((m, n), dtype=np.complex128, order="F") if c is None else np.asfortranarray(c).copy()
    _check(lib().orc_zgemm(ct.c_int(aop), ct.c_int(bop), ct.c_long(m), ct.c_long(n), ct.c_long(k),
                           _p(a), ct.c_long(a.shape[0]), _p(b),
                           ct.c_long(b.shape[0]), _p(out), ct.c_long(m), ct.c_int(int(add))))
    return out


def compensated_product(aop, a, b, c=None, add=False):
    a = np.asfortranarray(a, dtype=np.complex128)
    b = np.asfortranarray(b, dtype=np.complex128)
    rows = a.shape[1] if aop else a.shape[0]
    inner = a.shape[0] if aop else a.shape[1]
    cols = b.shape[1]
    out = np.zeros((rows, cols), dtype=np.complex128, order="F") if c is None else np.asfortranarray(c).copy()
    _check(lib().orc_compensated_product(ct.c_int(aop), ct.c_long(rows), ct.c_long(cols), ct.c_long(inner),
                                         _p(a), ct.c_long(a.shape[0]), _p(b),
                                         ct.c_long(b.shape[0]), _p(out), ct.c_long(rows),
                                         ct.c_int(int(add))))
    return out


def kinetic_preconditioner(t, wk):
    out = ct.c_double()
    _check(lib().orc_kinetic_preconditioner(ct.c_double(t), ct.c_double(wk), ct.byref(out)))
    return out.value


def generalized_eigen(a, b):
    a = np.asfortranarray(a, dtype=np.complex128)
    b = np.asfortranarray(b, dtype=np.complex128)
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n), dtype=np.complex128, order="F")
    _check(lib().orc_hermitian_generalized_eigen(ct.c_long(n), _p(a), _p(b),
                                                 _p(w), _p(v)))
    return w, v


def density_fft(hkl, coefficients, occupations, volume, support):
    """plane_wave_density_fft (src/density_functionals.cpp:406-460) with
    numpy's FFT on the reference's grid max(3, 2 m + 1) per axis:
    rho(r) = sum_b occ_b / volume |backward FFT of c_b|^2 (unnormalised), and
    the forward FFT / N at the support Miller triples."""
    hkl = np.asarray(hkl, dtype=np.int64)
    support = np.asarray(support, dtype=np.int64)
    c = np.asarray(coefficients, dtype=np.complex128)
    m = np.maximum(np.abs(support).max(axis=0), hkl.max(axis=0) - hkl.min(axis=0))
    shape = tuple(int(max(3, 2 * v + 1)) for v in m)
    count = int(np.prod(shape))
    idx = tuple(np.mod(hkl[:, a], shape[a]) for a in range(3))
    rho = np.zeros(shape)
    for b in range(c.shape[1]):
        if occupations[b] == 0:
            continue
        grid = np.zeros(shape, dtype=np.complex128)
        grid[idx] = c[:, b]
        psi = np.fft.ifftn(grid) * count  # unnormalised backward transform (e^{+i})
        rho += occupations[b] / volume * (psi.real ** 2 + psi.imag ** 2)
    f = np.fft.fftn(rho) / count
    return f[tuple(np.mod(support[:, a], shape[a]) for a in range(3))]


def density_brute_force(hkl, coefficients, occupations, volume, support):
    """rho(G) = sum_b occ_b / volume sum_{G1 - G2 = G} c_b(G1) conj(c_b(G2))
    (small bases only)."""
    hkl = [tuple(int(v) for v in row) for row in np.asarray(hkl)]
    c = np.asarray(coefficients, dtype=np.complex128)
    acc = {}
    for b in range(c.shape[1]):
        w = occupations[b] / volume
        for i, g1 in enumerate(hkl):
            for j, g2 in enumerate(hkl):
                key = (g1[0] - g2[0], g1[1] - g2[1], g1[2] - g2[2])
                acc[key] = acc.get(key, 0.0) + w * c[i, b] * np.conj(c[j, b])
    return np.array([acc.get(tuple(int(v) for v in k), 0.0) for k in np.asarray(support)])
