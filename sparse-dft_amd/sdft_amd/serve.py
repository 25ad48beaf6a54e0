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

import queue
import threading
import time
from concurrent.futures import Future

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
