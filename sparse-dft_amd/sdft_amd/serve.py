"""Concurrent certified solves on one GPU.

Independent eigenproblems (the k-points of one crystal, or solve requests of
a service) are the natural unit of parallelism of the path: the reference
runs k-point teams concurrently, each thread with its own accelerator stream
and handle (src/compute_backend.cpp:63-64, src/scf_core.cpp:3051-3061). Here
each worker thread owns one HIP context (stream + rocBLAS/rocSOLVER handle);
the C ABI releases the GIL, so solves of different workers overlap on the
device. A request is (problem, roots, warm block); its result carries the
certified eigenpairs and the arrival, start and finish times.
"""
from __future__ import annotations

import collections
import queue
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import native


class SolveService:
    def __init__(self, workers=4, device=0, davidson=None):
        if workers < 1:
            raise ValueError("at least one worker")
        self._davidson = dict(davidson or {})
        self._contexts = [native.Context(device) for _ in range(workers)]
        self._queues = [queue.Queue() for _ in range(workers)]
        self._pending = [0] * workers
        self._lock = threading.Lock()
        self._threads = [threading.Thread(target=self._run, args=(i,), daemon=True) for i in range(workers)]
        self._closed = False
        for t in self._threads:
            t.start()
        native._track("services", self)

    def context(self, worker):
        return self._contexts[worker]

    def submit(self, prob, roots, initial=None, operator=None, arrival=None, **davidson):
        """Queue one solve. A pre-built `operator` runs on the worker owning its
        context (a context's stream and handle serve one thread); otherwise the
        least-loaded worker builds the operator from `prob`."""
        fut: Future = Future()
        kw = dict(self._davidson)
        kw.update(davidson)
        with self._lock:
            if operator is not None:
                worker = next(i for i, c in enumerate(self._contexts) if c is operator.ctx)
            else:
                worker = min(range(len(self._queues)), key=lambda i: self._pending[i])
            self._pending[worker] += 1
        self._queues[worker].put((fut, prob, roots, initial, operator, kw,
                                  time.perf_counter() if arrival is None else arrival))
        return fut

    def _run(self, worker):
        ctx = self._contexts[worker]
        while True:
            job = self._queues[worker].get()
            if job is None:
                return
            fut, prob, roots, initial, operator, kw, arrival = job
            op = None
            try:
                if not fut.set_running_or_notify_cancel():
                    continue
                start = time.perf_counter()
                op = operator if operator is not None else native.Operator(ctx, prob)
                r = op.davidson(roots, kw.pop("max_iterations", 256), kw.pop("tolerance", 1e-12), initial, **kw)
                ctx.synchronize()
                r.update(arrival=arrival, start=start, finish=time.perf_counter(), n=op.n, worker=worker)
                fut.set_result(r)
            except Exception as e:  # noqa: BLE001 - delivered to the caller through the future
                fut.set_exception(e)
            finally:
                # an operator built here is released whether or not its solve succeeded
                if operator is None and op is not None:
                    op.close()
                with self._lock:
                    self._pending[worker] -= 1

    def close(self):
        """Stop the workers after the queued solves and release the contexts
        (operators a caller built on them keep their context alive until they
        are closed themselves)."""
        if self._closed:
            return
        self._closed = True
        for q in self._queues:
            q.put(None)
        for t in self._threads:
            t.join()
        for ctx in self._contexts:
            ctx.close()


# ---------------------------------------------------------------- serving engine

MiB = 1 << 20


def _fast_fft_length(minimum):
    """Smallest 2,3,5,7-smooth length >= minimum (src/fft_planning.hpp:16-23)."""
    n = max(1, int(minimum))
    while True:
        rem = n
        for p in (2, 3, 5, 7):
            while rem % p == 0:
                rem //= p
        if rem == 1:
            return n
        n += 1


def grid_shape(prob):
    """The FFT grid the device operator builds for `prob`: the given shape, or
    per axis the reference rule max(3, 2 max(R_V, span) + 1) rounded up to a
    2,3,5,7-smooth length (src/plane_wave_operator.cpp:400-457)."""
    hkl = np.asarray(prob.hkl)
    span = hkl.max(axis=0) - hkl.min(axis=0)
    vk = np.abs(np.asarray(prob.v_hkl)).reshape(-1, 3)
    kept = vk[(vk <= span).all(axis=1)] if len(vk) else vk
    radius = kept.max(axis=0) if len(kept) else np.zeros(3, dtype=int)
    return tuple(int(prob.shape[a]) if prob.shape[a] > 0 else
                 _fast_fft_length(max(3, 2 * max(int(radius[a]), int(span[a])) + 1)) for a in range(3))


def estimate_request_bytes(prob, roots, subspace_factor=8):
    """Upper bound of the device bytes one solve request holds: the operator's
    arrays (V(r) FP64 + FP32, kinetic, projectors FP64 + FP32, onsite blocks,
    stick plan), its stick and projection work buffers for the largest block
    a solve applies at once (the subspace capacity, at most the operator's
    apply chunk), and the Davidson buffers (six n x capacity blocks, seven
    n x roots blocks, the projected matrices and the Rayleigh-Ritz
    eigensolver), with capacity = min(n, max(factor roots, 64)) as the solver
    sizes it, plus 1 MiB for the small control buffers."""
    n = int(prob.n)
    shape = grid_shape(prob)
    count = int(np.prod(shape))
    hkl = np.asarray(prob.hkl)
    sticks = len({(int(h) % shape[0], int(k) % shape[1]) for h, k in hkl[:, :2]})
    np_ = int(prob.np_total)
    cap = min(n, max(subspace_factor * roots, 64))
    per_col = sticks * shape[2] * 16
    chunk = max(1, min(512, (512 * MiB) // max(per_col, 1)))
    cols = min(max(cap, 2 * roots), chunk)
    operator = (count * 12 + n * 8 + np_ * n * 24 + 2 * 16 * sum(r * r for r in prob.atom_rows) +
                sticks * (shape[2] + 1) * 4 + 2 * cols * per_col + 3 * 16 * np_ * cols)
    davidson = 16 * (6 * n * cap + 7 * n * roots + 12 * cap * cap + 4 * roots * roots + 8 * (2 * roots) ** 2)
    return operator + davidson + MiB


class RequestRejected(MemoryError):
    """A request whose device memory cannot fit the serving budget."""


# ---- shared-memory handoff between the front end and the worker processes

def _to_shared(arrays):
    """Copy named arrays into one new shared-memory segment (64-byte aligned
    slots); returns (segment, layout [(name, dtype, shape, offset)])."""
    from multiprocessing import shared_memory
    layout, off = [], 0
    for name, a in arrays.items():
        layout.append((name, a.dtype.str, a.shape, off))
        off += (a.nbytes + 63) // 64 * 64
    seg = shared_memory.SharedMemory(create=True, size=max(off, 64))
    for (name, dt, shape, o), a in zip(layout, arrays.values()):
        np.ndarray(shape, dtype=dt, buffer=seg.buf, offset=o)[...] = a
    return seg, layout


def _attach(name):
    """Attach to a segment the front end owns and unlinks (the spawned slots
    share the front end's resource tracker, which keeps one record per
    segment, so the front end's unlink is the one that clears it)."""
    from multiprocessing import shared_memory
    return shared_memory.SharedMemory(name=name)


def _views(seg, layout):
    return {name: np.ndarray(shape, dtype=dt, buffer=seg.buf, offset=o) for name, dt, shape, o in layout}


def _worker_main(index, device, jobs, results):
    """One serving slot: its own process and HIP context (stream, rocBLAS and
    rocSOLVER handles); every HIP call of the process is made by this thread."""
    ctx = native.Context(device)
    try:
        while True:
            job = jobs.get()
            if job is None:
                return
            out = dict(id=job["id"], worker=index)
            seg = vec_seg = None
            op = None
            try:
                seg = _attach(job["segment"])
                arrays = _views(seg, job["layout"])
                packed = {k: arrays[k] for k in ("hkl", "kin", "vh", "vv", "rows", "amp", "d", "q")}
                packed.update(regularization=job["regularization"], shape=job["shape"])
                start = time.perf_counter()
                op = native.Operator(ctx, None, packed=packed)
                if job["fault"] == "nonfinite":
                    ctx.inject_fault("nonfinite_apply", 1)
                kw = dict(job["davidson"])
                init = arrays.get("initial")
                r = op.davidson(job["roots"], kw.pop("max_iterations", 256), kw.pop("tolerance", 1e-12), init,
                                want_vectors=bool(job["vectors"]), **kw)
                ctx.synchronize()
                finish = time.perf_counter()
                if job["vectors"]:
                    vec_seg = _attach(job["vectors"])
                    np.ndarray(r["eigenvectors"].shape, dtype=np.complex128, buffer=vec_seg.buf, order="F")[...] = \
                        r["eigenvectors"]
                r["eigenvectors"] = None
                r.update(start=start, finish=finish, n=op.n, device_bytes=op.device_bytes())
                out["result"] = r
            except Exception as e:  # noqa: BLE001 - delivered to the front end, the slot continues
                out["error"] = (type(e).__name__, str(e))
            finally:
                ctx.inject_fault("nonfinite_apply", 0)
                if op is not None:
                    op.close()
                arrays = packed = init = None
                for sg in (seg, vec_seg):
                    if sg is not None:
                        try:
                            sg.close()
                        except BufferError:  # a view still alive: the mapping goes with the process
                            pass
            results.put(out)
    finally:
        ctx.close()


class ServingEngine:
    """Continuous batching of certified block solves on one GPU.

    `workers` serving slots, each its own process with its own HIP context
    (stream, rocBLAS/rocSOLVER handles): the kernels of every in-flight
    request run on their own stream and are co-scheduled on the device, each
    Davidson iteration of one request beside the iterations of the others. A
    request is admitted when a slot is free and the estimated device bytes of
    the in-flight requests plus its own stay within `hbm_budget`
    (estimate_request_bytes); otherwise it waits, in arrival order, and a
    request that fits is not held back by a larger one ahead of it. A request
    larger than the whole budget fails alone with RequestRejected. The front
    end hands a request's operator description and warm block to its slot
    through shared memory; the slot builds the operator on the device (the
    potential and projectors stay resident there with the Davidson blocks
    until the certified result, then are released) and returns the result,
    its eigenvectors through shared memory.

    Per-request fault injection for tests of failure handling:
    fault="nonfinite" makes one H application of that request write a NaN (its
    solve fails with a runtime error); fault="budget" makes its admission
    estimate exceed the budget (it is rejected). Either fails that request
    alone; the slot goes on serving.
    """

    def __init__(self, workers=4, hbm_budget=None, device=0, davidson=None):
        import multiprocessing as mp
        if workers < 1:
            raise ValueError("at least one worker")
        ctx = mp.get_context("spawn")
        self._davidson = dict(davidson or {})
        self._budget = hbm_budget
        self._lock = threading.Condition()
        self._jobs = [ctx.Queue() for _ in range(workers)]
        self._results = ctx.Queue()
        self._procs = [ctx.Process(target=_worker_main, args=(i, device, self._jobs[i], self._results), daemon=True)
                       for i in range(workers)]
        for p in self._procs:
            p.start()
        self._free = list(range(workers))
        self._waiting = collections.deque()
        self._pending = {}
        self._next_id = 0
        self._in_flight = 0
        self._running = 0
        self.max_in_flight_bytes = 0
        self.max_concurrent = 0
        self._closed = False
        self._collector = threading.Thread(target=self._collect, daemon=True)
        self._collector.start()
        native._track("services", self)

    def submit(self, prob, roots, initial=None, arrival=None, fault=None, want_vectors=True, **davidson):
        fut = Future()
        kw = dict(self._davidson)
        kw.update(davidson)
        need = estimate_request_bytes(prob, roots, kw.get("subspace_factor", 8))
        if fault == "budget":
            need = (self._budget or 0) + 1
        if self._budget is not None and need > self._budget:
            fut.set_exception(RequestRejected(f"request needs {need / MiB:.0f} MiB of device memory, above the "
                                              f"serving budget of {self._budget / MiB:.0f} MiB"))
            return fut
        packed = native.pack_problem(prob)
        arrays = {k: packed[k] for k in ("hkl", "kin", "vh", "vv", "rows", "amp", "d", "q")}
        if initial is not None:
            arrays["initial"] = np.asfortranarray(initial[:, :roots], dtype=np.complex128)
        seg, layout = _to_shared(arrays)
        vec = None
        if want_vectors:
            from multiprocessing import shared_memory
            vec = shared_memory.SharedMemory(create=True, size=16 * int(prob.n) * int(roots))
        with self._lock:
            if self._closed:
                seg.close()
                seg.unlink()
                raise RuntimeError("serving engine closed")
            job = dict(id=self._next_id, segment=seg.name, layout=layout, regularization=packed["regularization"],
                       shape=packed["shape"], roots=int(roots), davidson=kw, fault=fault,
                       vectors=vec.name if vec is not None else None)
            self._pending[self._next_id] = dict(fut=fut, seg=seg, vec=vec, bytes=need, n=int(prob.n), roots=int(roots),
                                                arrival=time.perf_counter() if arrival is None else arrival)
            self._next_id += 1
            self._waiting.append(job)
            self._dispatch()
        return fut

    def _dispatch(self):
        # with the lock held: admit waiting requests in arrival order while
        # slots are free; one that does not fit yet is passed over
        i = 0
        while self._free and i < len(self._waiting):
            job = self._waiting[i]
            meta = self._pending[job["id"]]
            if self._budget is not None and self._in_flight + meta["bytes"] > self._budget:
                i += 1
                continue
            del self._waiting[i]
            worker = self._free.pop()
            self._in_flight += meta["bytes"]
            self._running += 1
            self.max_in_flight_bytes = max(self.max_in_flight_bytes, self._in_flight)
            self.max_concurrent = max(self.max_concurrent, self._running)
            meta["admitted"] = time.perf_counter()
            meta["worker"] = worker
            self._jobs[worker].put(job)

    def _collect(self):
        while True:
            out = self._results.get()
            if out is None:
                return
            with self._lock:
                meta = self._pending.pop(out["id"])
                self._in_flight -= meta["bytes"]
                self._running -= 1
                self._free.append(meta["worker"])
                self._dispatch()
                self._lock.notify_all()
            fut = meta["fut"]
            try:
                if "error" in out:
                    name, msg = out["error"]
                    exc = {"HipError": native.HipError, "ValueError": ValueError,
                           "OverflowError": OverflowError}.get(name, RuntimeError)
                    fut.set_exception(exc(msg))
                    continue
                r = out["result"]
                if meta["vec"] is not None:
                    r["eigenvectors"] = np.ndarray((meta["n"], meta["roots"]), dtype=np.complex128,
                                                   buffer=meta["vec"].buf, order="F").copy(order="F")
                r.update(arrival=meta["arrival"], admitted=meta["admitted"], worker=out["worker"],
                         estimated_bytes=meta["bytes"])
                fut.set_result(r)
            finally:
                for sg in (meta["seg"], meta["vec"]):
                    if sg is not None:
                        sg.close()
                        sg.unlink()

    def close(self):
        """Finish the admitted and waiting requests, then stop the slots."""
        with self._lock:
            if self._closed:
                return
            while self._waiting or self._running:
                self._lock.wait()
            self._closed = True
        for q in self._jobs:
            q.put(None)
        for p in self._procs:
            p.join(timeout=120)
        self._results.put(None)
        self._collector.join()
