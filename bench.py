"""Benchmark: certified eigenpairs/s on the Si64-shaped Gamma problem.

Workload: diamond Si 2x2x2 supercell (a = 20.52 bohr, 64 atoms), Ecut 15 Ha
(npw 23847), 72^3 grid, 512 PAW projectors, 160 bands; synthetic inputs of
the recipe in sparse-dft_amd/sdft_amd/synthetic.py. One step = one complete
block-Davidson solve of the 160 lowest eigenpairs from a seeded warm start,
every pair FP64-certified to ||H x - e S x|| <= 1e-12 Ha on the original
operators. The headline value is the FP32-search form; the plain FP64 path is
timed beside it. Operators and initial vectors are resident in HBM before the
timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "sparse-dft_amd"))

PEAK_HBM_GBS = 8000.0
ROOTS = 160


def algorithmic_local_bytes(npw, count, m, c, r):
    """B_loc = m (2 npw c + 4 N c) + N r + npw (4 + r): each 3-D transform reads
    and writes the grid once, V, indices and T read once per launch."""
    return m * (2 * npw * c + 4 * count * c) + count * r + npw * (4 + r)


def host_cpu():
    """CPUs this process may run on, their physical cores and the CPU model
    (from the affinity mask, /sys topology and /proc/cpuinfo)."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id") as f:
                core = f.read().strip()
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id") as f:
                pkg = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add((None, c))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(cpus=len(cpus), physical_cores=len(cores), model=model)


def cpu_baseline(prob):
    """Oracle (FP64 C++/OpenMP restatement of the reference CPU path; GEMMs
    and dense eigensolves through CBLAS/LAPACKE as the reference's CPU build;
    FFTs its own mixed radix, FFTW being absent) on this host, threads bound
    close: one complete certified Si64 solve (the headline workload) and one
    Si8 solve, timed whole; the per-column H psi / S psi time of Si64 on one
    thread and on all, and of the two-species slab on all."""
    cpu = host_cpu()
    threads = min(cpu["cpus"], int(os.environ.get("OMP_NUM_THREADS", cpu["cpus"])))
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    from sdft_amd import synthetic as syn
    blas = pyoracle.numpy_openblas_path()
    if blas:
        pyoracle.bind_blas(blas)
    pyoracle.set_threads(threads)
    op = pyoracle.Operator(prob)
    init = syn.initial_psi(prob, ROOTS, 1)
    t0 = time.perf_counter()
    r = op.davidson(ROOTS, 256, 1e-12, init, subspace_factor=8)
    t_si64 = time.perf_counter() - t0
    if not r["converged"]:
        raise RuntimeError("the CPU baseline solve did not certify")
    x = syn.initial_psi(prob, 16, 1, extra=9)
    t0 = time.perf_counter()
    op.apply(x, "pair")
    col_all = (time.perf_counter() - t0) / 16
    pyoracle.set_threads(1)
    t0 = time.perf_counter()
    op.apply(x[:, :2], "pair")
    col_one = (time.perf_counter() - t0) / 2
    pyoracle.set_threads(threads)
    p0 = syn.config_problem(0)
    op0 = pyoracle.Operator(p0)
    t0 = time.perf_counter()
    r0 = op0.davidson(16, 256, 1e-12, syn.initial_psi(p0, 16, 0))
    t_si8 = time.perf_counter() - t0
    p3 = syn.config_problem(3)
    op3 = pyoracle.Operator(p3)
    x3 = syn.initial_psi(p3, 16, 3)
    t0 = time.perf_counter()
    op3.apply(x3, "pair")
    col_slab = (time.perf_counter() - t0) / 16
    return {"value": ROOTS / t_si64, "unit": "certified eigenpairs/s", "cores": threads, "kind": "port",
            "sample": (f"oracle FP64 Davidson (reference structure, subspace factor 8, in-repo mixed-radix FFT, "
                       f"{'OpenBLAS zgemm/zhegvd' if blas else 'plain loops'}), one complete certified Si64 solve "
                       f"of the same problem and warm block: {t_si64:.1f} s, {r['iterations']} iterations"),
            "seconds_sample": t_si64, "iterations": r["iterations"],
            "host": dict(cpu, threads=threads, omp_proc_bind=os.environ.get("OMP_PROC_BIND")),
            "si8_certified_eigenpairs_per_s": 16 / t_si8, "si8_iterations": r0["iterations"],
            "si64_pair_apply_ms_per_column": {"threads": 1e3 * col_all, "one_thread": 1e3 * col_one},
            "slab_pair_apply_ms_per_column": 1e3 * col_slab}


def kpoint_batch_leg(ctx):
    """Config 2: the Si8 4x4x4 Monkhorst-Pack mesh (64 k-points, 1622-1653
    plane waves, one 32^3 grid, 32 bands each) applied as one graph replay of
    the fused persistent kernel, FP64 and FP32, against the 64 per-k-point
    device applications. Work per launch: the FFT count 2 x 5 N log2 N per
    column (both transforms) and the HBM model of the local-H roofline
    B_loc = m (2 npw c + 4 N c) + N r + npw (4 + r) per k-point block."""
    from sdft_amd import native
    from sdft_amd import synthetic as syn
    ks = syn.mp_kpoints(4)
    probs = [syn.config_problem(2, kfrac=tuple(k)) for k in ks]
    ops = [native.Operator(ctx, p) for p in probs]
    cols = 32
    count = 32 ** 3
    out = {"kpoints": len(ops), "columns_per_kpoint": cols, "npw": [min(p.n for p in probs), max(p.n for p in probs)]}
    flops = 2 * 5 * count * np.log2(count) * cols * len(ops)
    try:
        for prec, name, c, r, peak in ((0, "fp64", 16, 8, 78.6), (1, "fp32", 8, 4, 157.3)):
            b = native.Batch(ctx, ops, cols, precision=prec)
            ms = b.benchmark(20)
            seq = sum(op.benchmark(cols, "pair", precision=prec, repetitions=3)["apply_ms"] for op in ops)
            model = sum(algorithmic_local_bytes(p.n, count, cols, c, r) for p in probs)
            out[name] = {"replay_ms": ms, "plan": b.plan(), "per_k_applies_ms": seq, "speedup_vs_per_k": seq / ms,
                         "fft_tflops": flops / (ms * 1e-3) / 1e12, "fft_frac_of_vector_peak": flops / (ms * 1e-3) / 1e12 / peak,
                         "local_model_gbs": model / (ms * 1e-3) / 1e9, "local_model_frac_of_hbm": model / (ms * 1e-3) / 1e9
                         / PEAK_HBM_GBS}
            b.close()
    finally:
        for op in ops:
            op.close()
    return out


TRAFFIC_FILE = os.path.join(REPO, "profiles", "si64_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per launch by kernel kind, from the committed PMC summary
    (lines `traffic_bytes_per_launch <kind> <bytes>`); empty if absent."""
    out = {}
    if os.path.exists(TRAFFIC_FILE):
        for line in open(TRAFFIC_FILE):
            f = line.strip().strip("`").split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch":
                out[f[1]] = float(f[2])
    return out


def copy_calibration(mib):
    """Eight device-to-device copies of `mib` MiB through the HIP runtime's copy
    kernel: the reference workload for scaling the FETCH_SIZE / WRITE_SIZE
    counters to bytes (each copy reads and writes exactly mib MiB)."""
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so")
    n = ct.c_size_t(mib << 20)
    src, dst = ct.c_void_p(), ct.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ct.byref(p), n):
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, n):
        raise RuntimeError("hipMemset failed")
    t0 = time.perf_counter()
    for _ in range(8):
        if hip.hipMemcpy(dst, src, n, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    if hip.hipDeviceSynchronize():
        raise RuntimeError("hipDeviceSynchronize failed")
    dt = time.perf_counter() - t0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_mib": mib, "copies": 8, "seconds": dt}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--verbose", type=int, default=0)
    ap.add_argument("--timing", type=int, default=0, help="synchronise and time Davidson phases")
    ap.add_argument("--subspace-factor", type=int, default=5,
                    help="Davidson capacity factor (the reference's SDFT_DAVIDSON_SUBSPACE_FACTOR, 4..16; "
                         "5 measured fastest for this workload, the reference defaults to 8)")
    ap.add_argument("--stagnation-window", type=int, default=1,
                    help="FP32 search hands off to FP64 after this many iterations without halving the max "
                         "residual below 1e-3 (1 measured fastest: the hand-off then meets a full subspace)")
    ap.add_argument("--handoff-restart", type=int, default=1,
                    help="1: thick-restart to the 2 x roots Ritz vectors at the FP32 -> FP64 hand-off; "
                         "0: refresh the whole subspace in FP64")
    ap.add_argument("--refinement", type=int, default=1,
                    help="FP32 search after the hand-off: 1 = block Jacobi-Davidson refinement (FP32 CG correction "
                         "solves, FP64 images and Rayleigh-Ritz); 0 = the reference's FP64 Davidson iterations")
    ap.add_argument("--inner-iterations", type=int, default=6,
                    help="CG steps per refinement correction solve (with 60 guards and BF16 projector products: "
                         "4 / 5 / 6 / 7 / 8 / 10 gave 1137 / 1172 / 1199-1201 / 1167 / 1173 / 1086 certified "
                         "eigenpairs/s)")
    ap.add_argument("--refinement-guards", type=int, default=60,
                    help="guard Ritz vectors beside the 160 roots in the refinement block (with 6 CG steps: "
                         "50 / 60 / 70 gave 1168 / 1199 / 1136; with 8: 40 / 60 / 80 gave 1146 / 1173 / 1148; "
                         "0 = roots / 2)")
    ap.add_argument("--bf16-projectors", type=int, default=1,
                    help="1: the refinement's correction solves run their projector GEMMs on BF16 MFMA with FP32 "
                         "accumulation; 0: FP32 MFMA (1 / 0 gave 1056 / 1013 certified eigenpairs/s, 12 iterations "
                         "either way)")
    ap.add_argument("--search-handoff", type=float, default=1.0,
                    help="max residual at which the FP32 search hands off (the reference hands off at 1e-8). With the "
                         "Jacobi-Davidson refinement taking its images by rotation, the earliest hand-off measured "
                         "fastest: 1 (after the second Rayleigh-Ritz) / 3e-1 / 1e-1 / 3e-2 / 3e-3 gave 714-716 / 703 / "
                         "667 / 692 / 647 certified eigenpairs/s, and 10 (after the first) 685")
    ap.add_argument("--serving", type=int, default=60,
                    help="serving leg: N requests (Si8..Si64 mix, Poisson arrivals at 0.7 x saturation) through the "
                         "serving engine, p50/p99 of arrival -> certified result (0: skip)")
    ap.add_argument("--serving-workers", type=int, default=4)
    ap.add_argument("--serving-budget-gib", type=float, default=16.0, help="HBM admission budget of the serving leg")
    ap.add_argument("--kpoint-batch", type=int, default=1,
                    help="also time the batched Si8 4x4x4 k-point application (one graph replay, 64 k x 32 bands)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 8 device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    ap.add_argument("--real-space-factor", type=float, default=1.3,
                    help="also time a solve with the real-space projector search (0: skip)")
    args = ap.parse_args()
    os.environ.setdefault("OMP_PROC_BIND", "close")  # before the oracle's OpenMP runtime loads
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards by k-point and by "
              f"request across GPUs with no exchange between them", file=sys.stderr)
        sys.exit(2)

    from sdft_amd import native
    from sdft_amd import synthetic as syn

    t_setup = time.perf_counter()
    prob = syn.config_problem(1)
    init = syn.initial_psi(prob, ROOTS, 1)
    ctx = native.Context(0)
    op = native.Operator(ctx, prob)
    # the initial block resident in HBM before the timed region, and a device
    # block for the eigenvectors: the timed solves move nothing over PCIe
    init_dev = native.DeviceBlock(prob.n, ROOTS).upload(init[:, :ROOTS])
    vec_dev = native.DeviceBlock(prob.n, ROOTS)
    ctx.synchronize()
    setup_s = time.perf_counter() - t_setup

    def solve(precision, host_buffers=False):
        ctx.synchronize()
        t0 = time.perf_counter()
        blocks = dict(initial=init) if host_buffers else dict(device_blocks=(init_dev.ptr, vec_dev.ptr))
        r = op.davidson(ROOTS, 256, 1e-12, search_precision=precision, verbose=args.verbose, **blocks,
                        timing=args.timing, subspace_factor=args.subspace_factor,
                        stagnation_window=args.stagnation_window, handoff_restart=args.handoff_restart,
                        refinement=args.refinement if precision else 0, inner_iterations=args.inner_iterations,
                        refinement_guards=args.refinement_guards, bf16_projectors=args.bf16_projectors,
                        search_handoff=args.search_handoff)
        ctx.synchronize()
        r["wall"] = time.perf_counter() - t0
        if not r["converged"]:
            raise RuntimeError(f"solve did not certify: max residual {r['residual_norms'].max():.3e}")
        return r

    for _ in range(args.warmup):
        solve(1)
    lat, runs = [], []
    ctx.kernel_clock(True)  # HIP events around every GEMM / local-H launch of the timed steps
    for _ in range(args.steps):
        r = solve(1)
        lat.append(r["wall"])
        runs.append(r)
    clock = ctx.kernel_clock_stats()
    ctx.kernel_clock(False)
    # the same solve taking and returning host buffers, as the reference's
    # call does: the initial block up and the eigenvectors down (2 x 61 MB)
    # through the library's pinned staging inside the timed call
    host = [solve(1, host_buffers=True) for _ in range(3)]
    host_wall = float(np.median([r["wall"] for r in host]))
    solve(0)  # FP64 warm-up: the FP64 path's buffers are allocated outside its timed solves too
    fp64 = [solve(0) for _ in range(max(1, args.fp64_steps))]
    fp64_wall = float(np.median([r["wall"] for r in fp64]))
    total = float(np.sum(lat))
    value = ROOTS * args.steps / total
    ev_err = float(np.abs(runs[-1]["eigenvalues"] - fp64[-1]["eigenvalues"]).max())
    max_res = float(max(r["residual_norms"].max() for r in runs + fp64 + host))

    # Roofline of the kernel with the largest event-timed share of the timed steps.
    npw, count = prob.n, int(np.prod(prob.shape))
    peaks = {"gemm_fp64": (78.6, "TFLOP/s", "mfma", 1e12), "gemm_fp32": (157.3, "TFLOP/s", "mfma", 1e12),
             "gemm_bf16": (2500.0, "TFLOP/s", "mfma", 1e12),
             "local_fp64": (PEAK_HBM_GBS, "GB/s", "hbm", 1e9), "local_fp32": (PEAK_HBM_GBS, "GB/s", "hbm", 1e9)}
    names = {"gemm_fp64": "sdft_hip::k_zgemm<double, ...> (FP64 MFMA complex GEMM, all shapes of the solve)",
             "gemm_fp32": "sdft_hip::k_zgemm<float, ...> (FP32 MFMA complex GEMM, all shapes of the solve)",
             "gemm_bf16": "sdft_hip::k_zgemm_bf16 (BF16 MFMA projector GEMMs of the correction solves)",
             "local_fp64": "local-H pipeline FP64 (k_sticks_backward + k_planes + k_sticks_forward)",
             "local_fp32": "local-H pipeline FP32 (k_sticks_backward + k_planes + k_sticks_forward)"}
    kernels = {}
    for k, st in clock.items():
        if not st["launches"]:
            continue
        peak, unit, bound, scale = peaks[k]
        ach = st["work"] / (st["ms"] * 1e-3) / scale
        kernels[k] = {"launches": st["launches"], "ms_total": st["ms"], "ms_per_launch": st["ms"] / st["launches"],
                      "work_per_launch": st["work"] / st["launches"], "achieved": ach, "unit": unit,
                      "frac": ach / peak, "share_of_solve_wall": st["ms"] * 1e-3 / total,
                      "algorithmic_bytes_per_launch": st["bytes"] / st["launches"]}
    dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
    peak, unit, bound, _ = peaks[dom]
    traffic = measured_traffic().get(dom)
    # the GEMM kinds count 8 m n k flops; the FP64 kernels execute 2 x 3 m n k
    # real MFMA flops when they use the Gauss (3M) form
    products = {"gemm_fp64": native.gemm_real_products(0), "gemm_fp32": native.gemm_real_products(1),
                "gemm_bf16": 4}
    for k in kernels:
        if k in products:
            kernels[k]["real_products_per_complex_mac"] = products[k]
            kernels[k]["mfma_executed_frac"] = kernels[k]["frac"] * products[k] / 4
        else:
            # inside the solve the local-H pipeline runs on a second stream beside
            # the projector GEMMs, so its event time includes the CUs it shares;
            # isolated_local_h times it alone
            kernels[k]["concurrent_with"] = "projector GEMM chain on the main stream"
    roofline = {"bound": bound, "achieved": kernels[dom]["achieved"], "peak": peak, "unit": unit,
                "frac": kernels[dom]["frac"], "traffic": traffic, "kernel": names[dom],
                "work_per_launch": kernels[dom]["work_per_launch"],
                "work_unit": "flops (8 m n k)" if bound == "mfma" else "algorithmic bytes",
                "ms_per_launch": kernels[dom]["ms_per_launch"], "launches": kernels[dom]["launches"],
                "algorithmic_bytes_per_launch": kernels[dom]["algorithmic_bytes_per_launch"],
                "mfma_executed_frac": kernels[dom].get("mfma_executed_frac"),
                "traffic_source": "profiles/si64_hbm_traffic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
                if traffic is not None else None,
                "measured": "HIP events on the launch stream over the timed steps"}
    # isolated 160-column local-H launches (FP32 and FP64) for reference
    bench = op.benchmark(ROOTS, which="h", precision=1, repetitions=20)
    bench64 = op.benchmark(ROOTS, which="h", precision=0, repetitions=10)
    isolated = {"local_fp32_160": {"ms": bench["local_ms"], "GB/s": algorithmic_local_bytes(npw, count, ROOTS, 8, 4)
                                   / (bench["local_ms"] * 1e-3) / 1e9},
                "local_fp64_160": {"ms": bench64["local_ms"], "GB/s": algorithmic_local_bytes(npw, count, ROOTS, 16, 8)
                                   / (bench64["local_ms"] * 1e-3) / 1e9}}

    # "block size 32": the bounded band-group solver with groups of 32 roots
    # (the reference's SDFT_BAND_GROUP_SOLVER=1, SDFT_BAND_GROUP_SIZE=32), FP64
    ctx.synchronize()
    t0 = time.perf_counter()
    rbg = op.davidson(ROOTS, 256, 1e-12, init, band_group_size=32, subspace_factor=args.subspace_factor)
    ctx.synchronize()
    t_bg = time.perf_counter() - t0
    band_groups = {"group_size": 32, "value": ROOTS / t_bg, "latency_s": t_bg, "converged": rbg["converged"],
                   "iterations": rbg["iterations"],
                   "max_eigenvalue_error_vs_fp64_ha": float(np.abs(rbg["eigenvalues"] - fp64[-1]["eigenvalues"]).max()),
                   "max_certified_residual_ha": float(rbg["residual_norms"].max())}

    kpoint_batch = kpoint_batch_leg(ctx) if args.kpoint_batch else None

    real_space = None
    if args.real_space_factor > 0:
        # Real-space projector search as the search operator (the certificate
        # stays FP64 reciprocal). The synthetic projectors are Gaussians of
        # width rc (amplitude exp(-r^2 / 2 rc^2)); their support radius is
        # taken as 3 rc (amplitude 1.1e-2), scaled by the factor.
        lat_rows = prob.cell.lattice
        frac = prob.cell.positions @ np.linalg.inv(lat_rows)
        radius = np.array([3.0 * sp.rc for sp in prob.cell.species])
        recip = op.benchmark(ROOTS, which="pair", precision=1, repetitions=10)
        points, retained = op.real_space_search(args.real_space_factor, lat_rows, frac, radius)
        rs_apply = op.benchmark(ROOTS, which="pair", precision=1, repetitions=10)
        rs_runs = [solve(1) for _ in range(max(1, args.fp64_steps))]
        op.real_space_search(0)
        rs_wall = float(np.median([r["wall"] for r in rs_runs]))
        real_space = {"factor": args.real_space_factor, "radius": "3 rc per species", "points": points,
                      "retained_power": retained,
                      "cost_ratio_vs_dense_projection": points / (len(prob.atom_rows) * npw),
                      "pair_apply_160_ms": rs_apply["apply_ms"], "reciprocal_fp32_pair_apply_160_ms": recip["apply_ms"],
                      "value": ROOTS / rs_wall, "latency_s": rs_wall, "iterations": rs_runs[-1]["iterations"],
                      "max_eigenvalue_error_vs_fp64_ha": float(np.abs(rs_runs[-1]["eigenvalues"]
                                                                      - fp64[-1]["eigenvalues"]).max()),
                      "max_certified_residual_ha": float(max(r["residual_norms"].max() for r in rs_runs))}

    out = {
        "metric": "certified eigenpairs/s, Si64 Γ (FP32-search vs FP64); p50/p99 solve latency",
        "value": value,
        "unit": "certified eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-search/fp64-certified",
        "data": "synthetic",
        "config": {"workload": "Si64 Gamma block Davidson, 160 bands, FP64-certified <= 1e-12 Ha",
                   "npw": npw, "grid": list(prob.shape), "projectors": prob.np_total, "roots": ROOTS,
                   "ecut_ha": 15.0, "subspace_factor": args.subspace_factor, "parallelism": "single GPU"},
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "pcie_inclusive": {"value": ROOTS / host_wall, "latency_s": host_wall, "steps": len(host),
                           "blocks": "initial block and eigenvectors as host buffers, copied inside the timed call"},
        "fp64_path": {"value": ROOTS / fp64_wall, "latency_s": fp64_wall, "iterations": fp64[-1]["iterations"]},
        "fp32_search_vs_fp64": (ROOTS / (total / args.steps)) / (ROOTS / fp64_wall),
        "iterations": runs[-1]["iterations"],
        "handoff_iteration": runs[-1]["handoff_iteration"],
        "max_eigenvalue_error_vs_fp64_ha": ev_err,
        "max_certified_residual_ha": max_res,
        "setup_s": setup_s,
        "roofline": roofline,
        "kernels": kernels,
        "isolated_local_h": isolated,
    }
    out["band_groups_32"] = band_groups
    if kpoint_batch is not None:
        out["kpoint_batch"] = kpoint_batch
    if real_space is not None:
        out["real_space_search"] = real_space
    if args.timing:
        keys = ("seconds_total", "seconds_apply", "seconds_gemm", "seconds_ritz", "seconds_ortho",
                "h_columns_search", "h_columns_exact")
        out["phases"] = {"fp32_search": {k: runs[-1][k] for k in keys}, "fp64": {k: fp64[-1][k] for k in keys}}
    init_dev.close()
    vec_dev.close()
    if args.serving:
        from sdft_amd import serving_bench
        op.close()
        out["serving"] = serving_bench.serving(args.serving, workers=args.serving_workers,
                                               budget_gib=args.serving_budget_gib,
                                               davidson=dict(search_precision=1, subspace_factor=args.subspace_factor,
                                                             stagnation_window=args.stagnation_window,
                                                             refinement=args.refinement,
                                                             inner_iterations=args.inner_iterations,
                                                             refinement_guards=0,
                                                             bf16_projectors=args.bf16_projectors,
                                                             search_handoff=args.search_handoff))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
