#!/usr/bin/env python3
"""H-matrix product benchmark on one MI355X.

Workload (the headline configuration): N = 1e5 points uniform on the unit
sphere (seeded), Laplace kernel 1/(4 pi r) with 1/(4 pi h) on the diagonal,
h = sqrt(4 pi / N), eta = 10, epsilon = 1e-6, leaf size (minclustersize) 32,
partial-pivot ACA. One step = one H-matvec (mu = 1) of the mixed-precision path
(FP32 leaves, FP32 arithmetic inside a leaf tile, FP64 accumulation over leaves)
on device-resident x and y. The plain FP64 path and the 32-RHS products are
timed in the same run and reported beside it.

value = algorithmic bytes of one product / time, in GB/s (every stored entry
once, x read and y written once, the permutations).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
C_void = ctypes.c_void_p
RECOMPRESS_TOL = 2e-6  # per-leaf relative Frobenius tail dropped by the FP64 recompression
LOWPREC_TOL = 4e-6     # per-leaf relative error budget of the FP16-stored trailing terms
HELM_RECOMPRESS_TOL = 3e-6  # the complex Helmholtz workload (epsilon 1e-5): recompression tolerance
HELM_LOWPREC_TOL = 5e-6     # and the 16-bit trailing terms' budget


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-instance batched product")
    ap.add_argument("--no-helmholtz", action="store_true", help="skip the complex Helmholtz N=2e5 workload")
    ap.add_argument("--no-assembly-5e5", action="store_true",
                    help="skip the N=5e5 assembly with and without the learned rank predictor")
    ap.add_argument("--cpu-reps", type=int, default=20)
    ap.add_argument("--matmat-counters-run", action="store_true",
                    help="short run for rocprofv3 --pmc on the 32-RHS products: three products per storage "
                         "(FP64, FP32, FP32+FP16) of the headline H-matrices; prints nothing else")
    ap.add_argument("--helmholtz-counters-run", action="store_true",
                    help="short run for rocprofv3 on the complex 64-RHS products of the Helmholtz N=2e5 workload: "
                         "one product per storage (FP64, FP32, FP32+FP16, FP32+BF16); prints nothing else")
    ap.add_argument("--probe-only", action="store_true",
                    help="run only the device bandwidth probe (htg_stream_probe) and print its JSON")
    ap.add_argument("--counters-run", action="store_true",
                    help="short run for rocprofv3 --pmc: a 256 MiB device copy of known size, then a few "
                         "mixed-precision and FP64 products; prints nothing else")
    return ap.parse_args()


def timed(fn, steps, warmup, stream, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def leaf_bytes(leaves):
    """Entry counts read by each product kernel."""
    import numpy as np
    d = leaves[leaves[:, 4] < 0]
    lr = leaves[leaves[:, 4] > 0]
    dense = int((d[:, 1].astype(np.int64) * d[:, 3]).sum())
    u = int((lr[:, 1].astype(np.int64) * lr[:, 4]).sum())
    v = int((lr[:, 3].astype(np.int64) * lr[:, 4]).sum())
    return dense, u, v


def host_cpu():
    """CPU model, logical CPUs of the host and those this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"model": model, "cpu_count": os.cpu_count(), "affinity": allowed,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(pts, diag, x, leaves_entries, args):
    """The reference's own OpenMP CPU product (1 MPI rank) when its driver is
    built, else the oracle's single-thread restatement, on the same workload.
    Threads: the CPUs this process may use, at most 16 (the host's share for one
    GPU on the box this runs on; a shared host is not ours to fill)."""
    import numpy as np
    from oracle import refcase
    hc = host_cpu()
    threads = max(1, min(hc["affinity"], 16))
    if refcase.available():
        try:
            t0 = time.time()
            r = refcase.run(pts, x=x, kernel=0, diag=diag, eta=10.0, eps=1e-6, minclustersize=32, compressor=0,
                            nreps=args.cpu_reps, dump_data=False, threads=threads, timeout=900)
            wall = time.time() - t0
            L = r.leaves
            dense = int((L[L[:, 4] < 0, 1].astype(np.int64) * L[L[:, 4] < 0, 3]).sum())
            lr = L[L[:, 4] > 0]
            lre = int((lr[:, 4].astype(np.int64) * (lr[:, 1] + lr[:, 3])).sum())
            n = pts.shape[0]
            b = 8 * (dense + lre) + 16 * n + 8 * n
            t = float(r.info["matvec_s"])
            return {"value": round(b / t / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "reference",
                    "ms_per_matvec": round(t * 1e3, 3), "host": hc,
                    "sample": f"full N={n} workload, {args.cpu_reps} timed FP64 matvecs of the reference HMatrix "
                              f"after 3 warm-ups (OpenMP {threads} threads, MKL sequential, 1 MPI rank) after its "
                              f"own build ({float(r.info['build_s']):.1f}s); driver wall {wall:.1f}s; "
                              f"host {hc['model']}, {hc['cpu_count']} logical CPUs, {hc['affinity']} usable"}
        except Exception as e:  # fall through to the restatement
            print(f"reference baseline unavailable: {e}", file=sys.stderr)
    from oracle import lib as orc
    O = orc.hmat_build(pts, kernel=0, diag=diag, minclustersize=32, eta=10.0, eps=1e-6)
    O.mvprod(x)
    t0 = time.perf_counter()
    reps = max(1, args.cpu_reps // 2)
    for _ in range(reps):
        O.mvprod(x)
    t = (time.perf_counter() - t0) / reps
    b = 8 * leaves_entries + 24 * pts.shape[0]
    return {"value": round(b / t / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "ms_per_matvec": round(t * 1e3, 3), "host": hc,
            "sample": f"oracle FP64 leaf loop, {reps} matvecs on the full workload, one thread"}


def ellipsoid_area(a, b, c):
    """Knud Thomsen's approximation of an ellipsoid's area (relative error < 1.1%)."""
    p = 1.6075
    return 4 * math.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)


def batch_64(torch, log):
    """64 independent Laplace H-matrices, N log-uniform in [2e3, 5e4], each on a
    seeded ellipsoid with semi-axes U(0.5, 2); eta 10, epsilon 1e-6, FP32
    storage. One block-diagonal H-matrix: every instance's leaves in the same
    launches, replayed from a HIP graph, layout picked by the autotuner; against
    the 64 instances' products launched one after another."""
    import numpy as np
    from htool_amd import Generator, HMatrix, geometry
    rng = np.random.default_rng(2024)
    pts, gens = [], []
    for i in range(64):
        n = int(round(math.exp(rng.uniform(math.log(2e3), math.log(5e4)))))
        ax = rng.uniform(0.5, 2.0, 3)
        pts.append(geometry.ellipsoid(n, tuple(ax), seed=1000 + i))
        gens.append(Generator(geometry.KERNEL_LAPLACE, 0.0,
                              geometry.laplace_diag(math.sqrt(ellipsoid_area(*ax) / n))))
    ntot = sum(p.shape[0] for p in pts)

    def make():
        H = HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
        H.set_compression("partialACA")
        return H

    torch.cuda.synchronize()
    t0 = time.time()
    B = make().build_batch(gens, pts)
    torch.cuda.synchronize()
    build_s = time.time() - t0
    B.set_storage("fp32")
    s = torch.cuda.Stream()
    sp = s.cuda_stream
    x = torch.randn(ntot, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    out = {"instances": 64, "N_total": ntot, "N_min": min(p.shape[0] for p in pts),
           "N_max": max(p.shape[0] for p in pts), "storage": "fp32", "build_s": round(build_s, 3)}
    with torch.cuda.stream(s):
        out["eager_default_layout_ms"] = round(timed(lambda: B.mvprod_device(x.data_ptr(), y.data_ptr(), 1, sp),
                                                     50, 5, s, torch), 4)
        lay, tuned_ms, tried = B.autotune(mu=1, reps=20)
        out["autotune"] = {"layout": lay, "ms": round(tuned_ms, 4), "candidates": tried}
        out["eager_tuned_layout_ms"] = round(timed(lambda: B.mvprod_device(x.data_ptr(), y.data_ptr(), 1, sp),
                                                   50, 5, s, torch), 4)
        g = B.capture(x.data_ptr(), y.data_ptr(), 1, sp)
        ms = timed(lambda: g.launch(sp), 50, 5, s, torch)
        out["graph_ms"] = round(ms, 4)
        out["graph_gbps"] = round(B.algorithmic_bytes(1, "fp32") / (ms * 1e-3) / 1e9, 1)
        # the same 64 H-matrices, one product launch sequence each
        singles = []
        for p, gsp in zip(pts, gens):
            H = make().build(gsp, p)
            H.set_storage("fp32")
            singles.append(H)
        xs = [torch.randn(p.shape[0], dtype=torch.float64, device="cuda") for p in pts]
        ys = [torch.zeros_like(v) for v in xs]

        def all_singles():
            for H, a, b in zip(singles, xs, ys):
                H.mvprod_device(a.data_ptr(), b.data_ptr(), 1, sp)
        out["separate_launches_ms"] = round(timed(all_singles, 20, 3, s, torch), 4)
    out["speedup_vs_separate"] = round(out["separate_launches_ms"] / out["graph_ms"], 2)
    log(f"batch of 64: {out}")
    return out


def helmholtz_2e5(torch, log):
    """N=2e5 points on a 4:1 prolate spheroid (semi-axes 1, 1, 4, area-uniform),
    e^{ikr}/(4 pi r) with k = 2 pi / (10 h) (10 points per wavelength),
    h = sqrt(area / N), 1/(4 pi h) on the diagonal, epsilon 1e-5, eta 10, leaf
    size 32 (blocks with M + N >= 16384 on the multi-workgroup ACA). Every
    low-rank term per-term scaled in 16-bit storage; dense leaves
    FP32 (mixed) against the FP64 complex product; 1 and 64 right-hand sides
    CN(0,1). Then a second build with FP64 QR+SVD recompression and only the
    trailing terms in 16 bits ("recompressed")."""
    import numpy as np
    from htool_amd import Generator, HMatrix, geometry
    from oracle import lib as orc
    n = 200000
    pts = geometry.ellipsoid(n, (1.0, 1.0, 4.0), seed=5)
    h = math.sqrt(geometry.prolate_spheroid_area(1.0, 4.0) / n)
    k = 2 * math.pi / (10 * h)
    diag = geometry.laplace_diag(h)
    H = HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.set_lowprec(0.0, all_terms=True)
    torch.cuda.synchronize()
    t0 = time.time()
    H.build(Generator(geometry.KERNEL_HELMHOLTZ, 0.0, diag, k), pts)
    torch.cuda.synchronize()
    i = H.get_info()
    out = {"N": n, "wavenumber": round(k, 4), "h": round(h, 6), "epsilon": 1e-5, "eta": 10.0, "leaf_size": 32,
           "build_s": round(time.time() - t0, 3), "aca_s": round(i.aca_seconds, 3), "aca_wide_blocks": i.n_aca_wide,
           "rank_mean": round(i.rank_mean, 2),
           "rank_max": i.rank_max, "dense_entries": int(i.dense_entries), "lowrank_entries": int(i.lowrank_entries),
           # blocks with M N > 2^31 - 1 that the reference's int test q (M + N) > M N splits and the
           # exact 64-bit test used here compresses (htg_params.ref_int_overflow follows the reference)
           "aca_int_wrap_blocks": i.n_aca_int_wrap}
    rng = np.random.default_rng(9)
    X = rng.standard_normal((n, 64)) + 1j * rng.standard_normal((n, 64))
    dX = torch.from_numpy(np.asfortranarray(X).reshape(-1, order="F").copy()).cuda()
    dY = torch.zeros_like(dX)
    sp = torch.cuda.current_stream().cuda_stream
    ys = {}
    for storage in ("fp64", "fp32", "fp32_fp16", "fp32_bf16"):
        H.set_storage(storage)
        r = {}
        for mu in (1, 64):
            ms = timed(lambda: H.mvprod_device(dX.data_ptr(), dY.data_ptr(), mu, sp), 5 if mu == 1 else 2, 1,
                       torch.cuda.current_stream(), torch)
            r[f"ms_mu{mu}"] = round(ms, 4)
            r[f"gbps_mu{mu}"] = round(H.algorithmic_bytes(mu, storage) / (ms * 1e-3) / 1e9, 1)
        H.mvprod_device(dX.data_ptr(), dY.data_ptr(), 64, sp)
        torch.cuda.synchronize()
        ys[storage] = dY.cpu().numpy().view(np.complex128).reshape(64, n).T.copy()
        out[storage] = r
    del H
    # The same workload with complex FP64 QR+SVD recompression and only the
    # trailing terms in 16 bits (within a per-leaf error budget), so the 16-bit
    # storages keep the error at the ACA's epsilon.
    Hr = HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    Hr.set_compression("partialACA")
    Hr.set_recompression(HELM_RECOMPRESS_TOL)
    Hr.set_lowprec(HELM_LOWPREC_TOL)
    torch.cuda.synchronize()
    t0 = time.time()
    Hr.build(Generator(geometry.KERNEL_HELMHOLTZ, 0.0, diag, k), pts)
    torch.cuda.synchronize()
    ir = Hr.get_info()
    rec = {"recompress_tol": HELM_RECOMPRESS_TOL, "lowprec_budget": HELM_LOWPREC_TOL,
           "build_s": round(time.time() - t0, 3), "rank_mean": round(ir.rank_mean, 2),
           "lowrank_entries": int(ir.lowrank_entries), "lowprec_entries": int(ir.lowprec_entries)}
    for storage in ("fp32_fp16", "fp32_bf16"):
        Hr.set_storage(storage)
        r = {}
        for mu in (1, 64):
            ms = timed(lambda: Hr.mvprod_device(dX.data_ptr(), dY.data_ptr(), mu, sp), 5 if mu == 1 else 2, 1,
                       torch.cuda.current_stream(), torch)
            r[f"ms_mu{mu}"] = round(ms, 4)
            r[f"gbps_mu{mu}"] = round(Hr.algorithmic_bytes(mu, storage) / (ms * 1e-3) / 1e9, 1)
        Hr.mvprod_device(dX.data_ptr(), dY.data_ptr(), 64, sp)
        torch.cuda.synchronize()
        ys["recompressed_" + storage] = dY.cpu().numpy().view(np.complex128).reshape(64, n).T.copy()
        rec[storage] = r
    del Hr
    rows = rng.choice(n, 200, replace=False).astype(np.int32)
    yd = orc.dense_rows_c(pts, None, rows, X[:, 0], kernel=4, wavenumber=k, diag=diag, same=True)
    for storage in ys:
        e = {"rel_err_vs_fp64_hmat": float(np.linalg.norm(ys[storage] - ys["fp64"]) / np.linalg.norm(ys["fp64"])),
             "rel_err_vs_dense_200_rows": float(np.linalg.norm(ys[storage][rows, 0] - yd) / np.linalg.norm(yd))}
        (rec[storage[len("recompressed_"):]] if storage.startswith("recompressed_") else out[storage]).update(e)
    out["recompressed"] = rec
    log(f"Helmholtz N=2e5: {out}")
    return out


def helmholtz_products_only(torch, log):
    """The Helmholtz N=2e5 workload of helmholtz_2e5 (every low-rank term
    16-bit in the mixed storages), one 64-RHS product per storage after a
    256 MiB copy that calibrates the traffic counters."""
    import numpy as np
    from htool_amd import Generator, HMatrix, capi, geometry
    n = 200000
    pts = geometry.ellipsoid(n, (1.0, 1.0, 4.0), seed=5)
    h = math.sqrt(geometry.prolate_spheroid_area(1.0, 4.0) / n)
    H = HMatrix(epsilon=1e-5, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    H.set_lowprec(0.0, all_terms=True)
    H.build(Generator(geometry.KERNEL_HELMHOLTZ, 0.0, geometry.laplace_diag(h), 2 * math.pi / (10 * h)), pts)
    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    capi.check(capi.lib().htg_stream_copy(C_void(a.data_ptr()), C_void(b.data_ptr()), 256 << 20))
    del a, b
    rng = np.random.default_rng(9)
    X = rng.standard_normal((n, 64)) + 1j * rng.standard_normal((n, 64))
    dX = torch.from_numpy(np.asfortranarray(X).reshape(-1, order="F").copy()).cuda()
    dY = torch.zeros_like(dX)
    sp = torch.cuda.current_stream().cuda_stream
    for storage in ("fp64", "fp32", "fp32_fp16", "fp32_bf16"):
        H.set_storage(storage)
        H.mvprod_device(dX.data_ptr(), dY.data_ptr(), 64, sp)
    torch.cuda.synchronize()
    log("helmholtz counters run done")


def assembly_5e5(torch, log):
    """N=5e5 unit sphere, Laplace, FP64 ACA at epsilon 1e-6: build without and
    with the trained rank predictor (trained on other sizes and geometries), the
    workspace each reserved, reruns, the misprediction rate, and the product."""
    import numpy as np
    from htool_amd import Generator, HMatrix, RankPredictor, geometry
    from htool_amd.predictor import misprediction
    n = 500000
    pts = geometry.sphere(n, seed=3)
    diag = geometry.laplace_diag(geometry.sphere_spacing(n))
    P = RankPredictor.load(os.path.join(ROOT, "htool_amd", "data", "rank_predictor.npz"))
    out = {"N": n, "epsilon": 1e-6, "eta": 10.0, "leaf_size": 32}
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    sp = torch.cuda.current_stream().cuda_stream
    for tag, pred in (("fixed_first_workspace", None), ("predictor", P)):
        H = HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
        H.set_compression("partialACA")
        H.set_predictor(pred)
        torch.cuda.synchronize()
        t0 = time.time()
        H.build(Generator(geometry.KERNEL_LAPLACE, 0.0, diag), pts)
        torch.cuda.synchronize()
        build_s = time.time() - t0
        i = H.get_info()
        r = {"build_s": round(build_s, 3), "aca_s": round(i.aca_seconds, 3), "aca_runs": i.n_aca,
             "reruns": i.n_aca_retry, "workspace_entries": int(i.aca_ws_entries),
             "needed_entries": int(i.aca_need_entries)}
        if pred is not None:
            f, t, cap = H.aca_records()
            m = misprediction(P.predict_terms(f), t)
            r["misprediction_rate"] = round(float(np.mean(t > cap)), 5)
            r["mean_abs_log_ratio"] = round(m["mean_abs_log_ratio"], 4)
            ms = timed(lambda: H.mvprod_device(x.data_ptr(), y.data_ptr(), 1, sp), 20, 3,
                       torch.cuda.current_stream(), torch)
            r["matvec_fp64_ms"] = round(ms, 4)
            H.set_storage("fp32")
            r["matvec_fp32_ms"] = round(timed(lambda: H.mvprod_device(x.data_ptr(), y.data_ptr(), 1, sp), 20, 3,
                                              torch.cuda.current_stream(), torch), 4)
        out[tag] = r
        log(f"N=5e5 {tag}: {r}")
        del H
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the H-matrix product is measured on "
                                   "one MI355X; a multi-GPU run is not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    from htool_amd import Generator, HMatrix, geometry

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.probe_only:
        from htool_amd import capi
        rd, cp, bg = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        capi.check(capi.lib().htg_stream_probe(1 << 30, 20, ctypes.byref(rd), ctypes.byref(cp), ctypes.byref(bg)))
        print(json.dumps({"read_gbps": round(rd.value, 1), "copy_gbps": round(cp.value, 1),
                          "read_best_grid": bg.value}), flush=True)
        return
    if args.helmholtz_counters_run:
        helmholtz_products_only(torch, log)
        return
    n = args.n
    pts = geometry.sphere(n, seed=0)
    h = geometry.sphere_spacing(n)
    diag = geometry.laplace_diag(h)
    H = HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
    H.set_compression("partialACA")
    log(f"building H-matrix N={n}")
    t0 = time.time()
    H.build(Generator(geometry.KERNEL_LAPLACE, 0.0, diag), pts)
    build_s = time.time() - t0
    log(f"built in {build_s:.2f}s")
    # the mixed-precision H-matrix: same tree and ACA, then FP64 recompression
    # and FP16 storage of the trailing terms
    Hm = HMatrix(epsilon=1e-6, eta=10.0, minclustersize=32)
    Hm.set_compression("partialACA")
    Hm.set_recompression(RECOMPRESS_TOL)
    Hm.set_lowprec(LOWPREC_TOL)
    t0 = time.time()
    Hm.build(Generator(geometry.KERNEL_LAPLACE, 0.0, diag), pts)
    build_m_s = time.time() - t0
    log(f"mixed H-matrix built in {build_m_s:.2f}s")
    info = H.get_info()
    leaves = H.get_leaves()
    dense_e, u_e, v_e = leaf_bytes(leaves)

    rng = np.random.default_rng(1)
    x = rng.standard_normal(n)
    if args.counters_run:
        from htool_amd import capi
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
        b = torch.empty_like(a)
        capi.check(capi.lib().htg_stream_copy(C_void(a.data_ptr()), C_void(b.data_ptr()), 256 << 20))
        torch.cuda.synchronize()
        dx = torch.from_numpy(x).cuda()
        dy = torch.zeros(n, dtype=torch.float64, device="cuda")
        for M, storage in ((Hm, "fp32_fp16"), (H, "fp32"), (H, "fp64")):
            M.set_storage(storage)
            for _ in range(3):
                M.mvprod_device(dx.data_ptr(), dy.data_ptr(), 1, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        log("counters run done")
        return
    if args.matmat_counters_run:
        from htool_amd import capi
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB: calibrates FETCH/WRITE_SIZE
        b = torch.empty_like(a)
        capi.check(capi.lib().htg_stream_copy(C_void(a.data_ptr()), C_void(b.data_ptr()), 256 << 20))
        del a, b
        dX = torch.from_numpy(rng.standard_normal(n * 32)).cuda()
        dY = torch.zeros(n * 32, dtype=torch.float64, device="cuda")
        for M, storage in ((H, "fp64"), (H, "fp32"), (Hm, "fp32_fp16")):
            M.set_storage(storage)
            for _ in range(3):
                M.mvprod_device(dX.data_ptr(), dY.data_ptr(), 32, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        log("matmat counters run done")
        return
    X = rng.standard_normal((n, 32))
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    dx = torch.from_numpy(x).cuda()
    dy = torch.zeros(n, dtype=torch.float64, device="cuda")
    dX = torch.from_numpy(np.asfortranarray(X).reshape(-1, order="F").copy()).cuda()
    dY = torch.zeros(n * 32, dtype=torch.float64, device="cuda")

    res = {}
    ys = {}
    for name, M, storage in (("fp64", H, "fp64"), ("fp32", H, "fp32"), ("mixed", Hm, "fp32_fp16")):
        M.set_storage(storage)
        layout = M.get_layout()  # the library default for the build (row groups of 32 with a 16-bit class)
        ms1 = timed(lambda: M.mvprod_device(dx.data_ptr(), dy.data_ptr(), 1, sp), args.steps, args.warmup, stream, torch)
        ys[name] = dy.cpu().numpy().copy()
        ms32 = timed(lambda: M.mvprod_device(dX.data_ptr(), dY.data_ptr(), 32, sp), max(5, args.steps // 5),
                     2, stream, torch)
        prof = M.profile(dx.data_ptr(), dy.data_ptr(), 1, 20, sp)
        prof32 = M.profile(dX.data_ptr(), dY.data_ptr(), 32, 5, sp)
        log(f"{name}: {ms1:.4f} ms/matvec, {ms32:.4f} ms/32-RHS product "
            f"(V.x {prof32['vx_ms']:.4f} ms, rows {prof32['row_apply_ms']:.4f} ms)")
        res[name] = {"layout": layout, "ms": ms1, "ms_mu32": ms32, "bytes": M.algorithmic_bytes(1, storage),
                     "bytes_mu32": M.algorithmic_bytes(32, storage), "flops_mu32": M.flops(32), "prof": prof,
                     "prof_mu32": prof32}

    # accuracy: against dense FP64 rows and against the FP64 H product
    from oracle import lib as orc
    rows = np.random.default_rng(2).choice(n, 2000, replace=False).astype(np.int32)
    yd = orc.dense_rows(pts, None, rows, x, kernel=0, diag=diag, same=True)
    err = {s: float(np.linalg.norm(ys[s][rows] - yd) / np.linalg.norm(yd)) for s in ys}
    err_vs_fp64 = {s: float(np.linalg.norm(ys[s] - ys["fp64"]) / np.linalg.norm(ys["fp64"])) for s in ys}

    # The same mixed product from host buffers (x copied in and y out over PCIe
    # by the library, pageable memory): reported beside the value, never as it.
    from htool_amd import capi
    xh = np.ascontiguousarray(x)
    yh = np.zeros(n)
    for _ in range(3):
        capi.check(capi.lib().htg_hmatrix_mvprod(Hm._handle(), capi.ptr(xh), capi.ptr(yh), 1))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        capi.check(capi.lib().htg_hmatrix_mvprod(Hm._handle(), capi.ptr(xh), capi.ptr(yh), 1))
    host_ms = (time.perf_counter() - t0) / args.steps * 1e3
    host = {"ms": round(host_ms, 4), "gbps": round(res["mixed"]["bytes"] / (host_ms * 1e-3) / 1e9, 2),
            "rel_diff_vs_device": float(np.linalg.norm(yh - ys["mixed"]) / np.linalg.norm(ys["mixed"]))}
    log(f"mixed from host buffers: {host}")

    mixed = res["mixed"]
    value = mixed["bytes"] / (mixed["ms"] * 1e-3) / 1e9
    # dominant kernel of the headline step and its own achieved bandwidth
    p = mixed["prof"]
    mi = Hm.get_info()
    u_hi = mi.lowrank_u_entries - mi.lowprec_u_entries
    v_hi = (mi.lowrank_entries - mi.lowrank_u_entries) - (mi.lowprec_entries - mi.lowprec_u_entries)
    kern = {"row_apply": (p["row_apply_ms"], 4 * (mi.dense_entries + u_hi) + 2 * mi.lowprec_u_entries + 12 * n),
            "lr_vx": (p["vx_ms"], 4 * v_hi + 2 * (mi.lowprec_entries - mi.lowprec_u_entries) + 4 * n),
            "gather_x": (p["gather_ms"], 8 * n + 4 * n + 4 * n)}
    dom = max(kern, key=lambda k: kern[k][0])
    achieved = kern[dom][1] / (kern[dom][0] * 1e-3) / 1e9

    # the device's achievable streaming rate, measured with the library's own
    # read and copy kernels (htg_stream_probe, 1 GiB, 20 launches per grid size)
    from htool_amd import capi
    rd, cp, bg = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    capi.check(capi.lib().htg_stream_probe(1 << 30, 20, ctypes.byref(rd), ctypes.byref(cp), ctypes.byref(bg)))
    measured = {"read_gbps": round(rd.value, 1), "copy_gbps": round(cp.value, 1), "read_best_grid": bg.value,
                "kernel": "htg_stream_probe (k_stream_read, k_stream_copy), 1 GiB"}
    log(f"stream probe: {measured}")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.isfile(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom)

    asm = None if args.no_assembly_5e5 else assembly_5e5(torch, log)
    batch = None if args.no_batch else batch_64(torch, log)
    helm = None if args.no_helmholtz else helmholtz_2e5(torch, log)

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline")
        cpu = cpu_baseline(pts, diag, x, dense_e + u_e + v_e, args)

    def entry(name):
        r = res[name]
        return {"layout": r["layout"], "ms": round(r["ms"], 4),
                "gbps": round(r["bytes"] / (r["ms"] * 1e-3) / 1e9, 2),
                "bytes": r["bytes"], "ms_mu32": round(r["ms_mu32"], 4),
                "tflops_mu32": round(r["flops_mu32"] / (r["ms_mu32"] * 1e-3) / 1e12, 3),
                "kernels_ms_mu32": {k: round(v, 5) for k, v in r["prof_mu32"].items()},
                "rel_err_vs_dense": err[name], "rel_err_vs_fp64_hmat": err_vs_fp64[name]}

    line = {
        "metric": "hmatvec_hbm_gbps",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "hmatvec_laplace_sphere", "N": n, "eta": 10.0, "epsilon": 1e-6, "leaf_size": 32,
                   "compressor": "partialACA", "mu": 1,
                   "storage": f"mixed: FP64 recompression (tol {RECOMPRESS_TOL}), FP32 leaves and leading terms, "
                              f"FP16 trailing terms (budget {LOWPREC_TOL}), FP32 tile arithmetic, FP64 accumulation"},
        "fp64": entry("fp64"),
        "fp32": entry("fp32"),
        "mixed": entry("mixed"),
        "speedup_mixed_vs_fp64": round(res["fp64"]["ms"] / mixed["ms"], 3),
        "speedup_target": 4.0,
        "mixed_host_buffers_pcie": host,
        "build_s": round(build_s, 3),
        "build_mixed_s": round(build_m_s, 3),
        "leaves": {"dense": info.n_dense, "lowrank": info.n_lowrank, "false_positive": info.n_false_positive,
                   "rank_mean": round(info.rank_mean, 3), "rank_max": info.rank_max,
                   "dense_entries": dense_e, "lowrank_entries": u_e + v_e,
                   "mixed_rank_mean": round(mi.rank_mean, 3), "mixed_lowrank_entries": mi.lowrank_entries,
                   "mixed_fp16_entries": mi.lowprec_entries, "aca_int_wrap_blocks": info.n_aca_int_wrap},
        "kernels_ms": {k: round(v, 5) for k, v in p.items()},
        "kernel_algorithmic_bytes": {k: int(v[1]) for k, v in kern.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "measured_peak_gbps": measured["read_gbps"],
                     "frac_of_measured_peak": round(achieved / measured["read_gbps"], 4),
                     "whole_step_frac": round(value / HBM_PEAK_GBS, 4),
                     "whole_step_frac_of_measured_peak": round(value / measured["read_gbps"], 4)},
        "stream_probe": measured,
        "cpu_baseline": cpu,
        "assembly_5e5": asm,
        "batch_64": batch,
        "helmholtz_2e5": helm,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
