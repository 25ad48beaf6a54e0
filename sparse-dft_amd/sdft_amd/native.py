"""ctypes binding of libsdft_hip.so (the C ABI in include/sdft_hip.h).

The product path: every call goes to the HIP library. There is no CPU
fallback; a missing or unbuilt library raises at import of this module's
first call.
"""
from __future__ import annotations

import atexit
import ctypes as ct
import os
import subprocess
import threading
import weakref

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsdft_hip.so")
_LIB = None

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
_lp = ct.POINTER(ct.c_long)
_u64p = ct.POINTER(ct.c_uint64)


class OperatorDesc(ct.Structure):
    _fields_ = [("n", ct.c_size_t), ("hkl", _ip), ("kinetic_ev", _dp), ("ev_per_hartree", ct.c_double),
                ("nv", ct.c_size_t), ("v_hkl", _ip), ("v_values", _dp), ("atoms", ct.c_size_t), ("rows", _lp),
                ("amplitudes", _dp), ("d", _dp), ("q", _dp), ("regularization", ct.c_double),
                ("shape", ct.c_int * 3)]


class DavidsonParams(ct.Structure):
    _fields_ = [("roots", ct.c_long), ("max_iterations", ct.c_long), ("tolerance", ct.c_double),
                ("subspace_factor", ct.c_long), ("search_precision", ct.c_int), ("search_handoff", ct.c_double),
                ("stagnation_window", ct.c_int), ("handoff_restart", ct.c_int), ("verbose", ct.c_int),
                ("timing", ct.c_int), ("band_group_size", ct.c_long), ("refinement", ct.c_int),
                ("inner_iterations", ct.c_int), ("refinement_guards", ct.c_long), ("device_pointers", ct.c_int),
                ("bf16_projectors", ct.c_int), ("band_group_sweep", ct.c_int)]


class DavidsonResult(ct.Structure):
    _fields_ = [("eigenvalues", _dp), ("eigenvectors", _dp), ("residual_norms", _dp), ("iterations", ct.c_long),
                ("converged", ct.c_int), ("h_columns_search", ct.c_long), ("h_columns_exact", ct.c_long),
                ("handoff_iteration", ct.c_long), ("seconds_total", ct.c_double), ("seconds_apply", ct.c_double),
                ("seconds_gemm", ct.c_double), ("seconds_ritz", ct.c_double), ("seconds_ortho", ct.c_double),
                ("cholesky_blocks", ct.c_long), ("rr_cholesky", ct.c_long), ("rr_stedc", ct.c_long),
                ("rr_barrier_fallbacks", ct.c_long), ("polish_accepted", ct.c_long), ("polish_rejected", ct.c_long),
                ("manifold_accepted", ct.c_long), ("manifold_rejected", ct.c_long), ("search_rejected", ct.c_long),
                ("refinement_steps", ct.c_long)]


class SnapshotInfo(ct.Structure):
    _fields_ = [("format", ct.c_int), ("n", ct.c_size_t), ("count", ct.c_size_t), ("shape", ct.c_size_t * 3),
                ("projectors", ct.c_size_t), ("atoms", ct.c_size_t), ("roots", ct.c_size_t),
                ("iterations", ct.c_size_t), ("initial_columns", ct.c_size_t), ("regularization", ct.c_double),
                ("tolerance", ct.c_double)]


class HipError(RuntimeError):
    pass


# Handles still open, closed in a fixed order when the interpreter exits
# (services first, so no worker thread is inside a call, then operators, then
# contexts), before any module or library state is torn down. Nothing relies on
# finalisers: __del__ only closes what a caller dropped without closing.
_OPEN = {"services": weakref.WeakSet(), "operators": weakref.WeakSet(), "contexts": weakref.WeakSet()}
_OPEN_LOCK = threading.Lock()


def _track(kind, obj):
    with _OPEN_LOCK:
        _OPEN[kind].add(obj)


def close_all():
    """Close every open solve service, operator and context (in that order)."""
    for kind in ("services", "operators", "contexts"):
        with _OPEN_LOCK:
            live = list(_OPEN[kind])
        for obj in live:
            try:
                obj.close()
            except Exception:  # noqa: BLE001 - teardown continues with the rest
                pass


atexit.register(close_all)


def build():
    subprocess.run(["make", "-s", "-j8", "-C", PKG_ROOT], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HipError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        _LIB = ct.CDLL(LIB_PATH)
        _LIB.sdft_hip_last_error.restype = ct.c_char_p
    return _LIB


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def _check(rc):
    if rc:
        msg = lib().sdft_hip_last_error().decode()
        if rc == 2:
            raise ValueError(msg)
        if rc == 3:
            raise OverflowError(msg)
        raise HipError(msg)


def gemm_real_products(precision=0) -> int:
    """Real MFMA products per complex multiply-add of the GEMM kernels (3 or 4)."""
    v = ct.c_int()
    _check(lib().sdft_hip_gemm_real_products(ct.c_int(precision), ct.byref(v)))
    return v.value


def capabilities() -> str:
    buf = ct.create_string_buffer(1024)
    _check(lib().sdft_hip_capabilities(buf, ct.c_size_t(1024)))
    return buf.value.decode()


def enumerate_g_vectors(lattice, cutoff_ev, kfrac, ev_per_ha=27.211386245988):
    L = lib()
    lat = np.ascontiguousarray(lattice, dtype=np.float64)
    kf = np.ascontiguousarray(kfrac, dtype=np.float64)
    count = ct.c_long()
    _check(L.sdft_hip_enumerate_g_vectors(_p(lat), ct.c_double(cutoff_ev), _p(kf), ct.c_double(ev_per_ha),
                                          ct.c_long(0), None, None, None, ct.byref(count)))
    n = count.value
    hkl = np.zeros((n, 3), dtype=np.int32)
    kin = np.zeros(n)
    cart = np.zeros((n, 3))
    _check(L.sdft_hip_enumerate_g_vectors(_p(lat), ct.c_double(cutoff_ev), _p(kf), ct.c_double(ev_per_ha),
                                          ct.c_long(n), _p(hkl, _ip), _p(kin), _p(cart), ct.byref(count)))
    return hkl, kin, cart


def grid_indices(shape, hkl):
    shp = np.ascontiguousarray(shape, dtype=np.int32)
    h = np.ascontiguousarray(hkl, dtype=np.int32)
    out = np.zeros(len(h), dtype=np.uint64)
    _check(lib().sdft_hip_grid_indices(_p(shp, _ip), ct.c_size_t(len(h)), _p(h, _ip), _p(out, _u64p)))
    return out


def snapshot_info(path):
    """Header and validation of a SDFTPWO1/2/3 snapshot (reads the whole file; no GPU)."""
    info = SnapshotInfo()
    _check(lib().sdft_hip_snapshot_info_read(os.fsencode(path), ct.byref(info)))
    return {f: (tuple(getattr(info, f)) if f == "shape" else getattr(info, f)) for f, _ in SnapshotInfo._fields_}


class DeviceBlock:
    """A rows x cols complex128 column-major block in HBM, from the HIP runtime
    libsdft_hip.so links (hipMalloc / hipMemcpy / hipFree), for callers that
    keep their blocks on the device (`Operator.davidson(device_blocks=...)`,
    `Operator.apply_device`)."""

    _rt = None
    _H2D, _D2H = 1, 2

    @classmethod
    def runtime(cls):
        if cls._rt is None:
            lib()
            rt = ct.CDLL("libamdhip64.so.7")
            rt.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
            rt.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
            rt.hipFree.argtypes = [ct.c_void_p]
            cls._rt = rt
        return cls._rt

    def __init__(self, rows, cols):
        self.shape = (int(rows), int(cols))
        self.nbytes = 16 * self.shape[0] * self.shape[1]
        p = ct.c_void_p()
        if self.runtime().hipMalloc(ct.byref(p), ct.c_size_t(self.nbytes)) != 0:
            raise HipError(f"hipMalloc of {self.nbytes} bytes failed")
        self.ptr = p.value

    def upload(self, a):
        a = np.asfortranarray(a, dtype=np.complex128)
        if a.shape != self.shape:
            raise ValueError(f"block shape {a.shape} != {self.shape}")
        if self.runtime().hipMemcpy(ct.c_void_p(self.ptr), a.ctypes.data_as(ct.c_void_p), ct.c_size_t(a.nbytes),
                                    self._H2D) != 0:
            raise HipError("hipMemcpy to the device failed")
        return self

    def download(self):
        out = np.empty(self.shape, dtype=np.complex128, order="F")
        if self.runtime().hipMemcpy(out.ctypes.data_as(ct.c_void_p), ct.c_void_p(self.ptr), ct.c_size_t(self.nbytes),
                                    self._D2H) != 0:
            raise HipError("hipMemcpy from the device failed")
        return out

    def close(self):
        if self.ptr:
            self.runtime().hipFree(ct.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One HIP stream + rocBLAS/rocSOLVER handle, owned by one host thread."""

    def __init__(self, device: int = 0):
        self.handle = ct.c_void_p()
        _check(lib().sdft_hip_context_create(ct.c_int(device), ct.byref(self.handle)))
        _track("contexts", self)

    def close(self):
        if self.handle:
            lib().sdft_hip_context_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_name(self) -> str:
        buf = ct.create_string_buffer(256)
        _check(lib().sdft_hip_device_name(self.handle, buf, ct.c_size_t(256)))
        return buf.value.decode()

    def synchronize(self):
        _check(lib().sdft_hip_synchronize(self.handle))

    FAULTS = {"barrier_expiry": 0, "nonfinite_apply": 1}

    def inject_fault(self, kind, count=1):
        """Make the next `count` events of `kind` misbehave on purpose (tests
        of failure handling): 'barrier_expiry' (the Rayleigh-Ritz reduction's
        grid barrier expires; zheevd redoes the solve) or 'nonfinite_apply'
        (an operator application writes a NaN)."""
        _check(lib().sdft_hip_inject_fault(self.handle, ct.c_int(self.FAULTS[kind]), ct.c_long(count)))

    def hermitian_eigen(self, a, k=None):
        """All eigenvalues (ascending) of the Hermitian matrix a and the lowest k
        eigenvectors, through the device Rayleigh-Ritz eigensolver; returns
        (values, vectors, path): path 1 = one-launch tridiagonalisation with
        multisection and inverse iteration, 2 = one-launch tridiagonalisation
        with divide and conquer, 0 = rocSOLVER zheevd."""
        a = np.asfortranarray(a, dtype=np.complex128)
        n = a.shape[0]
        k = n if k is None else int(k)
        w = np.zeros(n)
        v = np.zeros((n, k), dtype=np.complex128, order="F")
        fused = ct.c_int(0)
        _check(lib().sdft_hip_hermitian_eigen(self.handle, ct.c_size_t(n), _p(a), ct.c_size_t(k), _p(w),
                                              _p(v) if k else None, ct.byref(fused)))
        return w, v, int(fused.value)

    def tridiagonal_eigen(self, d, e, k):
        """Device tridiagonal stage: (all eigenvalues, n x k real inverse-iteration
        vectors before orthonormalisation, status)."""
        d = np.ascontiguousarray(d, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        n = d.size
        w = np.zeros(n)
        v = np.zeros((n, k), order="F")
        st = ct.c_int(0)
        _check(lib().sdft_hip_tridiagonal_eigen(self.handle, ct.c_size_t(n), _p(d), _p(e) if n > 1 else None,
                                                ct.c_size_t(k), _p(w), _p(v) if k else None, ct.byref(st)))
        return w, v, st.value

    CLOCK_KINDS = {"gemm_fp64": 0, "gemm_fp32": 1, "local_fp64": 2, "local_fp32": 3, "gemm_bf16": 4}

    def kernel_clock(self, enable):
        """Start (True) or stop (False) the per-launch HIP-event clock; both clear the sums."""
        _check(lib().sdft_hip_kernel_clock(self.handle, ct.c_int(int(bool(enable)))))

    def kernel_clock_stats(self):
        """{kind: {launches, ms, work, bytes}}: event-timed sums since the clock was
        started (work in flops for GEMM kinds, algorithmic bytes for local-H kinds;
        bytes: algorithmic HBM bytes)."""
        out = {}
        for name, kind in self.CLOCK_KINDS.items():
            n, ms, work, nbytes = ct.c_long(), ct.c_double(), ct.c_double(), ct.c_double()
            _check(lib().sdft_hip_kernel_clock_stats(self.handle, ct.c_int(kind), ct.byref(n), ct.byref(ms),
                                                     ct.byref(work)))
            _check(lib().sdft_hip_kernel_clock_bytes(self.handle, ct.c_int(kind), ct.byref(nbytes)))
            out[name] = {"launches": n.value, "ms": ms.value, "work": work.value, "bytes": nbytes.value}
        return out

    # ---- Accelerator mirror (host pointers)
    def zgemm(self, aop, bop, a, b, c=None, add=False, left_identity=0, precision=0, shape=None):
        a = np.asfortranarray(a, dtype=np.complex128)
        b = np.asfortranarray(b, dtype=np.complex128)
        m = a.shape[1] if aop else a.shape[0]
        k = a.shape[0] if aop else a.shape[1]
        n = b.shape[0] if bop else b.shape[1]
        out = np.zeros((m, n), dtype=np.complex128, order="F") if c is None else c
        _check(lib().sdft_hip_zgemm(self.handle, ct.c_int(aop), ct.c_int(bop), ct.c_size_t(m), ct.c_size_t(n),
                                    ct.c_size_t(k), _p(a), ct.c_size_t(a.shape[0]), _p(b), ct.c_size_t(b.shape[0]),
                                    _p(out), ct.c_size_t(out.shape[0]), ct.c_int(int(add)),
                                    ct.c_uint64(left_identity), ct.c_int(precision)))
        return out

    def zgemm_strided(self, aop, bop, m, n, k, a, lda, b, ldb, c, ldc, add=False, left_identity=0, precision=0):
        """Raw call on flat complex128 buffers (tests of padded leading dimensions)."""
        _check(lib().sdft_hip_zgemm(self.handle, ct.c_int(aop), ct.c_int(bop), ct.c_size_t(m), ct.c_size_t(n),
                                    ct.c_size_t(k), _p(a), ct.c_size_t(lda), _p(b), ct.c_size_t(ldb), _p(c),
                                    ct.c_size_t(ldc), ct.c_int(int(add)), ct.c_uint64(left_identity),
                                    ct.c_int(precision)))

    def zgemm_triple_right(self, a_list, b, precision=0):
        a_list = [np.asfortranarray(a, dtype=np.complex128) for a in a_list]
        b = np.asfortranarray(b, dtype=np.complex128)
        m, k = a_list[0].shape
        n = b.shape[1]
        outs = [np.zeros((m, n), dtype=np.complex128, order="F") for _ in range(3)]
        aa = (_dp * 3)(*[_p(a) for a in a_list])
        cc = (_dp * 3)(*[_p(o) for o in outs])
        _check(lib().sdft_hip_zgemm_triple_right(self.handle, ct.c_size_t(m), ct.c_size_t(n), ct.c_size_t(k), aa,
                                                 ct.c_size_t(m), _p(b), ct.c_size_t(k), cc, ct.c_size_t(m),
                                                 ct.c_int(precision)))
        return outs

    def local_hamiltonian(self, shape, indices, kinetic, potential, x, operator_identity=0, precision=0):
        shp = np.ascontiguousarray(shape, dtype=np.int32)
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        kin = np.ascontiguousarray(kinetic, dtype=np.float64)
        pot = np.ascontiguousarray(potential, dtype=np.float64)
        x = np.asfortranarray(x, dtype=np.complex128)
        out = np.zeros_like(x, order="F")
        _check(lib().sdft_hip_local_hamiltonian(self.handle, _p(shp, _ip), _p(idx, _u64p), ct.c_size_t(len(idx)),
                                                _p(kin), _p(pot), _p(x), ct.c_size_t(x.shape[1]), _p(out),
                                                ct.c_uint64(operator_identity), ct.c_int(precision)))
        return out

    def paw_pair(self, identity, b, d, q, x, projector_identity=0, precision=0):
        b = np.asfortranarray(b, dtype=np.complex128)
        d = np.asfortranarray(d, dtype=np.complex128)
        q = np.asfortranarray(q, dtype=np.complex128)
        x = np.asfortranarray(x, dtype=np.complex128)
        npj, npw = b.shape
        h = np.zeros((npw, x.shape[1]), dtype=np.complex128, order="F")
        s = np.zeros_like(h)
        handled = ct.c_int()
        _check(lib().sdft_hip_paw_pair(self.handle, ct.c_uint64(identity), ct.c_size_t(npw), ct.c_size_t(npj),
                                       ct.c_size_t(x.shape[1]), _p(b), _p(d), _p(q), _p(x), _p(h), _p(s),
                                       ct.c_uint64(projector_identity), ct.c_int(precision), ct.byref(handled)))
        return h, s

    def onsite_blocks(self, rows, d_blocks, q_blocks, p, precision=0):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        d = np.ascontiguousarray(np.concatenate([np.asfortranarray(b).ravel(order="F") for b in d_blocks]))
        q = np.ascontiguousarray(np.concatenate([np.asfortranarray(b).ravel(order="F") for b in q_blocks]))
        p = np.asfortranarray(p, dtype=np.complex128)
        npj, cols = p.shape
        w = np.zeros((npj, 2 * cols), dtype=np.complex128, order="F")
        _check(lib().sdft_hip_onsite_blocks(self.handle, ct.c_size_t(len(rows)), _p(rows, _lp), _p(d), _p(q), _p(p),
                                            ct.c_size_t(cols), _p(w), ct.c_int(precision)))
        return w


def pack_problem(prob):
    """The operator description of a Problem as contiguous arrays (the layout
    of sdft_hip_operator_desc): Miller indices, kinetic energies (eV), the
    Fourier potential, projector rows per atom, row-major projectors and
    onsite blocks, plus the regularisation and grid shape."""
    return dict(
        hkl=np.ascontiguousarray(prob.hkl, dtype=np.int32),
        kin=np.ascontiguousarray(prob.kinetic_ev, dtype=np.float64),
        vh=np.ascontiguousarray(prob.v_hkl, dtype=np.int32).reshape(-1, 3),
        vv=np.ascontiguousarray(np.stack([np.real(prob.v_values), np.imag(prob.v_values)], axis=1)),
        rows=np.ascontiguousarray(prob.atom_rows, dtype=np.int64),
        amp=np.ascontiguousarray(np.concatenate([np.ravel(a) for a in prob.amplitudes]) if len(prob.amplitudes)
                                 else np.zeros(0, dtype=np.complex128)),
        d=np.ascontiguousarray(np.concatenate([np.ravel(a) for a in prob.d_blocks]) if len(prob.d_blocks)
                               else np.zeros(0, dtype=np.complex128)),
        q=np.ascontiguousarray(np.concatenate([np.ravel(a) for a in prob.q_blocks]) if len(prob.q_blocks)
                               else np.zeros(0, dtype=np.complex128)),
        regularization=float(prob.regularization),
        shape=tuple(int(v) for v in prob.shape),
    )


class Operator:
    """Device-resident plane-wave operator built from a synthetic Problem (or
    its packed description, pack_problem)."""

    def __init__(self, ctx: Context, prob, shape=None, _handle=None, packed=None):
        self.ctx = ctx
        if _handle is not None:
            self.handle = _handle
            _track("operators", self)
            self.n = None
            self._describe()
            return
        self._keep = packed if packed is not None else pack_problem(prob)
        k = self._keep
        self.n = int(k["hkl"].shape[0])
        desc = OperatorDesc()
        desc.n = self.n
        desc.hkl = _p(k["hkl"], _ip)
        desc.kinetic_ev = _p(k["kin"])
        desc.ev_per_hartree = 27.211386245988
        desc.nv = len(k["vh"])
        desc.v_hkl = _p(k["vh"], _ip)
        desc.v_values = _p(k["vv"])
        desc.atoms = len(k["rows"])
        desc.rows = _p(k["rows"], _lp)
        desc.amplitudes = _p(k["amp"])
        desc.d = _p(k["d"])
        desc.q = _p(k["q"])
        desc.regularization = k["regularization"]
        shp = shape if shape is not None else k["shape"]
        for a in range(3):
            desc.shape[a] = int(shp[a])
        self.handle = ct.c_void_p()
        _check(lib().sdft_hip_operator_create(ctx.handle, ct.byref(desc), ct.byref(self.handle)))
        _track("operators", self)
        self._keep = None
        self._describe()

    @classmethod
    def from_snapshot(cls, ctx: Context, path):
        """(operator, warm block or None) from a SDFTPWO1/2/3 snapshot."""
        info = snapshot_info(path)
        initial = None
        if info["initial_columns"]:
            initial = np.zeros((info["n"], info["initial_columns"]), dtype=np.complex128, order="F")
        h = ct.c_void_p()
        _check(lib().sdft_hip_snapshot_load(ctx.handle, os.fsencode(path), _p(initial), ct.byref(h)))
        op = cls(ctx, None, _handle=h)
        op.n = info["n"]
        return op, initial

    def project(self, x, precision=0):
        """P = B X (projectors x m), src/plane_wave_operator.cpp:375-380."""
        x = np.asfortranarray(x, dtype=np.complex128)
        p = np.zeros((self.np, x.shape[1]), dtype=np.complex128, order="F")
        _check(lib().sdft_hip_operator_project(self.handle, _p(x), ct.c_size_t(x.shape[1]), _p(p),
                                               ct.c_int(precision)))
        return p

    def projector_metric(self, p):
        """Q P with the per-atom onsite metric, src/plane_wave_operator.cpp:381-389."""
        p = np.asfortranarray(p, dtype=np.complex128)
        qp = np.zeros_like(p, order="F")
        _check(lib().sdft_hip_operator_projector_metric(self.handle, _p(p), ct.c_size_t(p.shape[1]), _p(qp)))
        return qp

    def real_space_search(self, factor=1.3, lattice=None, fractional=None, radius=None):
        """Enable (factor in [1, 2]) or disable (0) the real-space projector
        search for FP32 search applications. With geometry (lattice rows in
        bohr, fractional atoms x 3, radius per atom) atoms of radius > 0 keep
        the points of their sphere; otherwise the amplitude threshold is used.
        Returns (retained points, retained projector power fraction)."""
        kept = ct.c_size_t()
        power = ct.c_double()
        geo = [None, None, None]
        if lattice is not None:
            geo = [np.ascontiguousarray(lattice, dtype=np.float64).reshape(9),
                   np.ascontiguousarray(fractional, dtype=np.float64).reshape(-1),
                   np.ascontiguousarray(radius, dtype=np.float64).reshape(-1)]
        _check(lib().sdft_hip_operator_real_space_search(
            self.handle, ct.c_double(factor), *[None if g is None else _p(g) for g in geo], ct.byref(kept),
            ct.byref(power)))
        return kept.value, power.value

    def apply_real_space(self, x, which, precision=0):
        x = np.asfortranarray(x, dtype=np.complex128)
        y = np.zeros_like(x, order="F")
        _check(lib().sdft_hip_operator_apply_real_space(self.handle, ct.c_int({"h": 0, "s": 1}[which]), _p(x),
                                                        ct.c_size_t(x.shape[1]), _p(y), ct.c_int(precision)))
        return y

    def real_space_points(self, atoms):
        """(natural grid indices of the retained points, atom by atom; count per atom)."""
        counts = np.zeros(atoms, dtype=np.int64)
        _check(lib().sdft_hip_operator_real_space_points(self.handle, None, _p(counts, _lp)))
        natural = np.zeros(int(counts.sum()), dtype=np.int64)
        _check(lib().sdft_hip_operator_real_space_points(self.handle, _p(natural, _lp), _p(counts, _lp)))
        return natural, counts

    def density(self, coefficients, occupations, volume, support_hkl):
        """Smooth density Fourier coefficients at support_hkl (m x 3 ints),
        src/density_functionals.cpp:406-460."""
        c = np.asfortranarray(coefficients, dtype=np.complex128)
        occ = np.ascontiguousarray(occupations, dtype=np.float64)
        keys = np.ascontiguousarray(support_hkl, dtype=np.int32)
        out = np.zeros(len(keys), dtype=np.complex128)
        _check(lib().sdft_hip_operator_density(self.handle, _p(c), ct.c_size_t(c.shape[1]), _p(occ),
                                               ct.c_double(volume), ct.c_size_t(len(keys)), _p(keys, _ip), _p(out)))
        return out

    def save_snapshot(self, path, roots, iterations, tolerance, initial=None):
        init = None if initial is None else np.asfortranarray(initial[:, :roots], dtype=np.complex128)
        _check(lib().sdft_hip_snapshot_save(self.handle, os.fsencode(path), ct.c_size_t(roots),
                                            ct.c_size_t(iterations), ct.c_double(tolerance), _p(init)))

    def _describe(self):
        shape_out = (ct.c_int * 3)()
        count = ct.c_long()
        npj = ct.c_long()
        sticks = ct.c_long()
        _check(lib().sdft_hip_operator_info(self.handle, shape_out, ct.byref(count), ct.byref(npj), ct.byref(sticks),
                                            None, None, None, None, None))
        self.shape = tuple(shape_out)
        self.count = count.value
        self.np = npj.value
        self.sticks = sticks.value

    def davidson_seeds(self, roots, tolerance=1e-12, initial=None):
        """The device Davidson seed block (n x columns complex) for `roots` roots."""
        cols_max = 3 * int(roots)
        block = np.zeros((self.n, cols_max), dtype=np.complex128, order="F")
        init = None if initial is None else np.asfortranarray(initial[:, :roots], dtype=np.complex128)
        cols = ct.c_size_t(0)
        _check(lib().sdft_hip_davidson_seeds(self.handle, ct.c_size_t(roots), ct.c_double(tolerance),
                                             _p(init) if init is not None else None, _p(block), ct.byref(cols)))
        return block[:, :cols.value]

    def device_bytes(self):
        """Device memory held by the operator (arrays, work buffers, kept Davidson buffers)."""
        v = ct.c_size_t()
        _check(lib().sdft_hip_operator_device_bytes(self.handle, ct.byref(v)))
        return v.value

    def strict_step(self, kind, x, direction=None, tolerance=1e-12):
        """One strict certificate-path step on the block x (n x roots) from its
        fresh FP64 certificate: kind 'polish' (tangent step along `direction`)
        or 'manifold' (re-Ritz inside the block). Returns (accepted, x,
        energies, residual norms) of the kept block."""
        x = np.asfortranarray(x, dtype=np.complex128).copy(order="F")
        roots = x.shape[1]
        d = None if direction is None else np.asfortranarray(direction, dtype=np.complex128)
        e = np.zeros(roots)
        r = np.zeros(roots)
        acc = ct.c_int()
        _check(lib().sdft_hip_strict_step(self.handle, ct.c_int({"polish": 0, "manifold": 1}[kind]),
                                          ct.c_size_t(roots), ct.c_double(tolerance), _p(x), _p(d), _p(e), _p(r),
                                          ct.byref(acc)))
        return bool(acc.value), x, e, r

    def close(self):
        if self.handle:
            lib().sdft_hip_operator_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        idx = np.zeros(self.n, dtype=np.uint64)
        kin = np.zeros(self.n)
        vr = np.zeros(self.count)
        hd = np.zeros(self.n)
        sd = np.zeros(self.n)
        shape_out = (ct.c_int * 3)()
        _check(lib().sdft_hip_operator_info(self.handle, shape_out, None, None, None, _p(idx, _u64p), _p(kin), _p(vr),
                                            _p(hd), _p(sd)))
        return dict(indices=idx, kinetic=kin, local_real=vr, h_diag=hd, s_diag=sd)

    def apply(self, x, which="pair", precision=0):
        x = np.asfortranarray(x, dtype=np.complex128)
        m = x.shape[1]
        hy = np.zeros((self.n, m), dtype=np.complex128, order="F")
        sy = np.zeros((self.n, m), dtype=np.complex128, order="F")
        code = {"h": 0, "s": 1, "pair": 2}[which]
        _check(lib().sdft_hip_operator_apply(self.handle, ct.c_int(code), _p(x), ct.c_size_t(m),
                                             _p(hy) if code != 1 else None, _p(sy) if code != 0 else None,
                                             ct.c_int(precision), ct.c_int(0)))
        return {"h": hy, "s": sy, "pair": (hy, sy)}[which]

    def apply_device(self, x_ptr, columns, hy_ptr, sy_ptr, which="pair", precision=0, allocate=True):
        """H x and/or S x on device buffers (integer device addresses of n x
        columns complex column-major blocks), returning when they are written.
        allocate=False takes the allocation-free entry point, which stages
        everything in the library's fixed per-GPU scratch array."""
        code = {"h": 0, "s": 1, "pair": 2}[which]
        hy = ct.c_void_p(hy_ptr) if code != 1 else None
        sy = ct.c_void_p(sy_ptr) if code != 0 else None
        if allocate:
            rc = lib().sdft_hip_operator_apply(self.handle, ct.c_int(code), ct.c_void_p(x_ptr), ct.c_size_t(columns),
                                               hy, sy, ct.c_int(precision), ct.c_int(1))
        else:
            rc = lib().sdft_hip_operator_apply_noalloc(self.handle, ct.c_int(code), ct.c_void_p(x_ptr),
                                                       ct.c_size_t(columns), hy, sy, ct.c_int(precision))
        _check(rc)

    def apply_shifted(self, x, theta, precision=0):
        """(H - theta_j S) x_j per column (precision 0 FP64, 1 FP32, 2 FP32 with
        BF16-MFMA projector GEMMs): the refinement's correction operator."""
        x = np.asfortranarray(x, dtype=np.complex128)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        if x.shape[0] != self.n or th.shape != (x.shape[1],):
            raise ValueError("shifted application: x must be n x columns with one shift per column")
        y = np.zeros_like(x)
        _check(lib().sdft_hip_operator_apply_shifted(self.handle, _p(x), ct.c_size_t(x.shape[1]), _p(th), _p(y),
                                                     ct.c_int(precision)))
        return y

    def benchmark(self, columns, which="pair", precision=0, repetitions=10):
        ms = np.zeros(3)
        code = {"h": 0, "s": 1, "pair": 2}[which]
        _check(lib().sdft_hip_operator_benchmark(self.handle, ct.c_int(code), ct.c_size_t(columns),
                                                 ct.c_int(precision), ct.c_int(repetitions), _p(ms)))
        return dict(local_ms=ms[0], paw_ms=ms[1], apply_ms=ms[2])

    def davidson(self, roots, max_iterations=256, tolerance=1e-12, initial=None, subspace_factor=8,
                 search_precision=0, search_handoff=1e-8, stagnation_window=3, handoff_restart=1, verbose=0,
                 timing=0, want_vectors=True, band_group_size=0, refinement=0, inner_iterations=0,
                 refinement_guards=0, device_blocks=None, bf16_projectors=0, band_group_sweep=0):
        """device_blocks = (initial, eigenvectors): device addresses of n x
        roots complex128 column-major blocks resident in HBM (initial 0 for a
        cold start); the solve then moves no block over PCIe and returns
        eigenvectors=None (they are in the device block)."""
        p = DavidsonParams(roots, max_iterations, tolerance, subspace_factor, search_precision, search_handoff,
                           stagnation_window, handoff_restart, verbose, timing, band_group_size, refinement,
                           inner_iterations, refinement_guards, 1 if device_blocks is not None else 0,
                           bf16_projectors, band_group_sweep)
        ev = np.zeros(roots)
        res = np.zeros(roots)
        r = DavidsonResult()
        r.eigenvalues = _p(ev)
        r.residual_norms = _p(res)
        if device_blocks is not None:
            if initial is not None:
                raise ValueError("device_blocks carries the initial block")
            init_ptr, vec_ptr = device_blocks
            if not vec_ptr:
                raise ValueError("device_blocks needs a device eigenvector block")
            vec = None
            r.eigenvectors = ct.cast(ct.c_void_p(vec_ptr), _dp)
            init = ct.cast(ct.c_void_p(init_ptr), _dp) if init_ptr else None
        else:
            vec = np.zeros((self.n, roots), dtype=np.complex128, order="F") if want_vectors else None
            r.eigenvectors = _p(vec) if vec is not None else None
            host_init = None if initial is None else np.asfortranarray(initial[:, :roots], dtype=np.complex128)
            init = _p(host_init)
        _check(lib().sdft_hip_davidson(self.handle, ct.byref(p), init, ct.byref(r)))
        return dict(eigenvalues=ev, eigenvectors=vec, residual_norms=res, iterations=r.iterations,
                    converged=bool(r.converged), h_columns_search=r.h_columns_search,
                    h_columns_exact=r.h_columns_exact, handoff_iteration=r.handoff_iteration,
                    seconds_total=r.seconds_total, seconds_apply=r.seconds_apply, seconds_gemm=r.seconds_gemm,
                    seconds_ritz=r.seconds_ritz, seconds_ortho=r.seconds_ortho,
                    cholesky_blocks=r.cholesky_blocks, rr_cholesky=r.rr_cholesky, rr_stedc=r.rr_stedc,
                    rr_barrier_fallbacks=r.rr_barrier_fallbacks, polish_accepted=r.polish_accepted,
                    polish_rejected=r.polish_rejected, manifold_accepted=r.manifold_accepted,
                    manifold_rejected=r.manifold_rejected, search_rejected=r.search_rejected,
                    refinement_steps=r.refinement_steps)


class Batch:
    """H psi and S psi of many device operators of one context that share an
    FFT grid (the k-points of one crystal) as one persistent kernel launch
    replayed from a HIP graph; the launch shape is autotuned on creation and
    cached per problem shape."""

    def __init__(self, ctx: Context, operators, columns, precision=0):
        self.ctx = ctx
        self.operators = list(operators)  # kept alive while the batch exists
        arr = (ct.c_void_p * len(self.operators))(*[op.handle.value for op in self.operators])
        self.handle = ct.c_void_p()
        self.columns = int(columns)
        _check(lib().sdft_hip_batch_create(ctx.handle, ct.c_size_t(len(self.operators)), arr,
                                           ct.c_size_t(self.columns), ct.c_int(precision), ct.byref(self.handle)))
        _track("operators", self)
        off = np.zeros(len(self.operators) + 1, dtype=np.int64)
        _check(lib().sdft_hip_batch_offsets(self.handle, _p(off, _lp)))
        self.offsets = off

    def apply(self, blocks):
        """blocks: one n_k x columns array per operator; returns (H blocks, S blocks)."""
        x = np.zeros(int(self.offsets[-1]), dtype=np.complex128)
        for k, b in enumerate(blocks):
            x[self.offsets[k]:self.offsets[k + 1]] = np.asfortranarray(b, dtype=np.complex128).ravel(order="F")
        hy = np.zeros_like(x)
        sy = np.zeros_like(x)
        _check(lib().sdft_hip_batch_apply(self.handle, _p(x), _p(hy), _p(sy)))
        split = lambda a: [a[self.offsets[k]:self.offsets[k + 1]].reshape(self.columns, -1).T
                           for k in range(len(self.operators))]
        return split(hy), split(sy)

    def benchmark(self, replays=20):
        ms = ct.c_double()
        _check(lib().sdft_hip_batch_benchmark(self.handle, ct.c_int(replays), ct.byref(ms)))
        return ms.value

    def plan(self):
        t, p, g, c = ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int()
        ms = ct.c_double()
        _check(lib().sdft_hip_batch_plan(self.handle, ct.byref(t), ct.byref(p), ct.byref(g), ct.byref(ms),
                                         ct.byref(c)))
        return dict(threads=t.value, planes=p.value, groups=g.value, tuned_ms=ms.value, from_cache=bool(c.value))

    def close(self):
        if self.handle:
            lib().sdft_hip_batch_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
