"""CPU restatement of the reference's solver snapshot format (SDFTPWO1/2/3).

TEST INFRASTRUCTURE ONLY: an independent numpy writer/reader of
src/plane_wave_operator.cpp:520-653 used by tests/ to pin the product's C++
reader and writer (byte layout, validation order and messages).

Layout: 8-byte magic; uint64 header {n, count, s0, s1, s2, projectors, roots,
iterations, initial_size}; float64 {regularization, tolerance}; PWO3 only:
uint64 atoms and uint64 channels[atoms]; uint64 indices[n]; float64 kinetic,
h_diagonal, s_diagonal [n]; float64 local_real[count]; complex128 B
(projectors x n, column-major); onsite part (PWO3: per atom D then Q,
channels^2 each; PWO2: D, Q projectors^2; PWO1: DB, QB projectors x n);
complex128 initial[initial_size]. Exact length 96 + 8 (1 + atoms) (PWO3) +
8 * doubles.
"""
from __future__ import annotations

import os

import numpy as np

MAGIC = {1: b"SDFTPWO1", 2: b"SDFTPWO2", 3: b"SDFTPWO3"}
LIMIT = 1 << 40
INT_MAX = 2 ** 31 - 1


class SnapshotError(RuntimeError):
    pass


def _mul(a, b):
    if b and a > LIMIT // b:
        raise SnapshotError("solver snapshot dimension overflow")
    return a * b


def _hermitian(m):
    return np.linalg.norm(m - m.conj().T) <= 1e-12 * max(1.0, np.linalg.norm(m))


def write(path, s):
    """s: dict with form, n, count, shape, projectors, roots, iterations,
    regularization, tolerance, indices, kinetic, h_diagonal, s_diagonal,
    local_real, all_b (projectors x n), onsite_d / onsite_q (list of blocks for
    form 3, one matrix for forms 1 and 2), channels (form 3), initial (n x roots
    or None)."""
    if os.path.exists(path) or os.path.exists(path + ".tmp"):
        raise SnapshotError("refusing to overwrite solver snapshot: " + path)
    form = s["form"]
    initial = s.get("initial")
    ini = np.zeros(0, np.complex128) if initial is None else np.asarray(initial, np.complex128).ravel(order="F")
    parts = [MAGIC[form],
             np.array([s["n"], s["count"], *s["shape"], s["projectors"], s["roots"], s["iterations"], ini.size],
                      dtype="<u8").tobytes(),
             np.array([s["regularization"], s["tolerance"]], dtype="<f8").tobytes()]
    if form == 3:
        parts.append(np.array([len(s["channels"]), *s["channels"]], dtype="<u8").tobytes())
    parts.append(np.asarray(s["indices"], dtype="<u8").tobytes())
    for key in ("kinetic", "h_diagonal", "s_diagonal", "local_real"):
        parts.append(np.asarray(s[key], dtype="<f8").tobytes())
    parts.append(np.asarray(s["all_b"], np.complex128).ravel(order="F").tobytes())
    if form == 3:
        for d, q in zip(s["onsite_d"], s["onsite_q"]):
            parts.append(np.asarray(d, np.complex128).ravel(order="F").tobytes())
            parts.append(np.asarray(q, np.complex128).ravel(order="F").tobytes())
    else:
        parts.append(np.asarray(s["onsite_d"], np.complex128).ravel(order="F").tobytes())
        parts.append(np.asarray(s["onsite_q"], np.complex128).ravel(order="F").tobytes())
    parts.append(ini.tobytes())
    with open(path + ".tmp", "wb") as f:
        f.write(b"".join(parts))
    os.rename(path + ".tmp", path)


def read(path):
    data = open(path, "rb").read()
    length = len(data)
    pos = 0

    def take(dtype, count):
        nonlocal pos
        size = np.dtype(dtype).itemsize * count
        if pos + size > length:
            raise SnapshotError("truncated solver snapshot")
        out = np.frombuffer(data, dtype=dtype, count=count, offset=pos).copy()
        pos += size
        return out

    if length < 8:
        raise SnapshotError("truncated solver snapshot")
    magic = data[:8]
    pos = 8
    forms = {v: k for k, v in MAGIC.items()}
    if magic not in forms:
        raise SnapshotError("unsupported solver snapshot format")
    form = forms[magic]
    n, count, s0, s1, s2, projectors, roots, iterations, ni = (int(v) for v in take("<u8", 9))
    ok = n and count and n <= count and roots and roots <= n and iterations and iterations <= INT_MAX
    ok = ok and all(3 <= x <= INT_MAX for x in (s0, s1, s2))
    if ok:
        ok = _mul(_mul(s0, s1), s2) == count and (not ni or ni == _mul(n, roots))
    if not ok:
        raise SnapshotError("invalid solver snapshot dimensions")
    reg, tol = (float(v) for v in take("<f8", 2))
    channels, block_square, extra = [], 0, 0
    if form == 3:
        atoms = int(take("<u8", 1)[0])
        if not atoms or atoms > projectors or atoms > length // 8:
            raise SnapshotError("invalid solver snapshot atom count")
        channels = [int(c) for c in take("<u8", atoms)]
        spanned = 0
        for c in channels:
            if not c or c > projectors - spanned:
                raise SnapshotError("invalid solver snapshot atom boundaries")
            spanned += c
            sq = _mul(c, c)
            if sq > LIMIT - block_square:
                raise SnapshotError("solver snapshot atom storage overflow")
            block_square += sq
        if spanned != projectors:
            raise SnapshotError("solver snapshot atom boundaries do not span projectors")
        extra = 8 * (1 + atoms)
    if form == 1:
        paw = _mul(6, _mul(n, projectors))
    else:
        paw = _mul(2, _mul(n, projectors)) + _mul(4, block_square if form == 3 else _mul(projectors, projectors))
    doubles = _mul(4, n) + count + paw + _mul(2, ni)
    if doubles > LIMIT // 8 or length != 96 + extra + _mul(8, doubles):
        raise SnapshotError("solver snapshot size mismatch")
    if not np.isfinite(reg) or reg < 0 or not np.isfinite(tol) or tol <= 0:
        raise SnapshotError("invalid solver snapshot scalars")
    indices = take("<u8", n)
    srt = np.sort(indices)
    if srt[-1] >= count or np.any(srt[1:] == srt[:-1]):
        raise SnapshotError("invalid solver snapshot FFT indices")
    arrays = {}
    for key, size in (("kinetic", n), ("h_diagonal", n), ("s_diagonal", n), ("local_real", count)):
        arrays[key] = take("<f8", size)
        if not np.all(np.isfinite(arrays[key])):
            raise SnapshotError("nonfinite solver snapshot array")
    if np.any(arrays["s_diagonal"] <= 0):
        raise SnapshotError("invalid solver snapshot overlap diagonal")
    all_b = take("<c16", projectors * n).reshape((projectors, n), order="F")
    if not np.all(np.isfinite(all_b)):
        raise SnapshotError("nonfinite solver snapshot projector array")

    def onsite(rows, cols, square):
        m = take("<c16", rows * cols).reshape((rows, cols), order="F")
        if not np.all(np.isfinite(m)):
            raise SnapshotError("nonfinite solver snapshot PAW array")
        if square and not _hermitian(m):
            raise SnapshotError("non-Hermitian solver snapshot onsite matrix")
        return m

    if form == 3:
        od, oq = [], []
        for c in channels:
            od.append(onsite(c, c, True))
            oq.append(onsite(c, c, True))
    elif form == 2:
        od, oq = onsite(projectors, projectors, True), onsite(projectors, projectors, True)
    else:
        od, oq = onsite(projectors, n, False), onsite(projectors, n, False)
    initial = take("<c16", ni)
    if not np.all(np.isfinite(initial)):
        raise SnapshotError("nonfinite solver snapshot warm vector")
    return dict(form=form, n=n, count=count, shape=(s0, s1, s2), projectors=projectors, roots=roots,
                iterations=iterations, regularization=reg, tolerance=tol, channels=channels, indices=indices,
                all_b=all_b, onsite_d=od, onsite_q=oq,
                initial=initial.reshape((n, roots), order="F") if ni else None, **arrays)


def from_problem(prob, op, roots, iterations, tolerance, initial=None):
    """PWO3 arrays of the oracle operator `op` built from the synthetic problem
    `prob` (onsite blocks Hermitised as the operator constructor does,
    src/plane_wave_operator.cpp:494-499)."""
    herm = [0.5 * (d + d.conj().T) for d in prob.d_blocks]
    hermq = [0.5 * (q + q.conj().T) for q in prob.q_blocks]
    all_b = np.concatenate([np.asarray(a) for a in prob.amplitudes], axis=0)
    return dict(form=3, n=op.n, count=op.count, shape=op.shape, projectors=int(all_b.shape[0]), roots=roots,
                iterations=iterations, regularization=prob.regularization, tolerance=tolerance,
                channels=[int(r) for r in prob.atom_rows], indices=op.indices, kinetic=op.kinetic,
                h_diagonal=op.h_diag, s_diagonal=op.s_diag, local_real=op.local_real, all_b=all_b,
                onsite_d=herm, onsite_q=hermq, initial=initial)


def legacy(s, form):
    """The same operator as PWO2 (block-diagonal dense D/Q) or PWO1 (folded
    D B, Q B), built from PWO3 arrays as the reference smoke test builds its
    legacy fixtures (tests/native/plane_wave_operator_smoke.cpp:180-205)."""
    npj = s["projectors"]
    d = np.zeros((npj, npj), np.complex128)
    q = np.zeros((npj, npj), np.complex128)
    o = 0
    for db, qb in zip(s["onsite_d"], s["onsite_q"]):
        r = db.shape[0]
        d[o:o + r, o:o + r] = db
        q[o:o + r, o:o + r] = qb
        o += r
    out = dict(s)
    out["form"] = form
    out["channels"] = []
    if form == 2:
        out["onsite_d"], out["onsite_q"] = d, q
    else:
        out["onsite_d"], out["onsite_q"] = d @ s["all_b"], q @ s["all_b"]
    return out
