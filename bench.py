#!/usr/bin/env python3
"""Benchmark of the device numeric factorization.

Workload: 3D 7-point Laplacian on an 80^3 grid (n = 512,000, diagonal 6,
off-diagonals -1), nested-dissection ordering (the library's own, the
analogue of the reference's default METIS ordering), 16 right-hand sides
uniform in [-1, 1] from seed 0.

One step = one pass of the hot path: numeric factorization + solve of the 16
right-hand sides. The headline (`value`) is the mixed-precision path (FP32
fronts, updates accumulated in FP32 on MFMA) whose solve runs FP64 iterative
refinement to scaled backward error <= 1e-15; the plain FP64 path (factor +
solve) is measured in the same run. Times are HIP-event device times on the
library's stream (values and right-hand sides already in HBM: the
host->device copy of the values is reported separately as upload_ms).
GFLOP/s use the reference's flop model (sum over fronts of
sum_{r=m-n+1}^{m} r^2).

The CPU baseline is the oracle (a restatement of the reference's multifrontal
algorithm, OpenMP over the host cores) timed on a bounded sample of the same
tree: the nodes of the lowest levels, up to about --cpu-seconds of work.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # MI355X dense TFLOP/s
HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_path(S, A, akeep, opts, b, precision, steps, warmup, refine_tol):
    """Factorize + solve `warmup + steps` times; returns per-step records."""
    S.set_ext_options(precision=precision, refine=S.REFINE_AUTO, refine_tol=refine_tol,
                      record_timing=0, use_graph=1)
    fkeep = None
    recs = []
    for it in range(warmup + steps):
        x = b.copy(order="F")
        # wall clock around each call, as the reference driver times them
        # (drivers/spldlt_test.F90:133-148): calls are synchronous at return
        t0 = time.perf_counter()
        fkeep, inf = S.spldlt_factorize(True, A.val, akeep, fkeep, options=opts)
        t1 = time.perf_counter()
        if inf.flag < 0:
            raise RuntimeError(f"factorize failed: flag {inf.flag}")
        ft = inf.tail()
        t2 = time.perf_counter()
        sinf = S.spldlt_solve(x, akeep, fkeep, options=opts)
        t3 = time.perf_counter()
        if sinf.flag < 0:
            raise RuntimeError(f"solve failed: flag {sinf.flag}")
        wall = (t1 - t0) + (t3 - t2)
        st = sinf.tail()
        if it >= warmup:
            recs.append(dict(factor_ms=ft.factor_ms, upload_ms=ft.upload_ms, solve_ms=st.solve_ms,
                             refine_iters=st.refine_iters, wall_ms=1e3 * wall, factor_call_ms=1e3 * (t1 - t0),
                             solve_call_ms=1e3 * (t3 - t2), x=x))
        log(f"  prec={precision} step {it}: factor {ft.factor_ms:.2f} ms, solve {st.solve_ms:.2f} ms,"
            f" IR iters {st.refine_iters}, wall {1e3 * wall:.1f} ms")
    return fkeep, recs


def kernel_profile(S, A, akeep, opts, precision, fkeep):
    """One eager factorization with HIP events around every launch."""
    S.set_ext_options(precision=precision, record_timing=1, use_graph=0)
    fkeep, inf = S.spldlt_factorize(True, A.val, akeep, fkeep, options=opts)
    st = S.fkeep_stats(fkeep)
    S.set_ext_options(record_timing=0, use_graph=1)
    kinds = {}
    for k, name in enumerate(S.KIND_NAMES):
        if st.kind_launches[k] == 0:
            continue
        kinds[name] = dict(ms=round(st.kind_ms[k], 4), launches=int(st.kind_launches[k]),
                           gflop=round(st.kind_flops[k] / 1e9, 3),
                           gbytes=round(st.kind_bytes[k] / 1e9, 4))
    return kinds, st.factor_ms, fkeep


def cpu_baseline(sym, A, seconds):
    """Oracle factorization of the lowest levels of the same tree."""
    import oracle
    nn = sym["nnodes"]
    par = sym["sparent"] - 1
    m = np.diff(sym["rptr"]).astype(np.float64)
    n = np.diff(sym["sptr"]).astype(np.float64)
    q = m - n
    flops = (m ** 3 - q ** 3) / 3.0 + (m ** 2 - q ** 2) / 2.0 + (m - q) / 6.0
    h = np.zeros(nn, np.int64)
    for i in range(nn):
        if par[i] < nn:
            h[par[i]] = max(h[par[i]], h[i] + 1)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    # calibrate on the leaves, then extend the sample level by level
    H = 0
    mask = h <= H
    t0 = time.perf_counter()
    st, _, _ = oracle.factor_posdef(sym, A.val, node_mask=mask.astype(np.int8), nthreads=cores)
    t = time.perf_counter() - t0
    rate = flops[mask].sum() / max(t, 1e-9)
    best = (H, t, flops[mask].sum())
    while H + 1 <= h.max():
        nmask = h <= H + 1
        pred = flops[nmask].sum() / rate
        if pred > seconds:
            break
        H += 1
        t0 = time.perf_counter()
        st, _, _ = oracle.factor_posdef(sym, A.val, node_mask=nmask.astype(np.int8), nthreads=cores)
        t = time.perf_counter() - t0
        rate = flops[nmask].sum() / max(t, 1e-9)
        best = (H, t, flops[nmask].sum())
        mask = nmask
    H, t, f = best
    return dict(value=round(f / t / 1e9, 3), unit="GFLOP/s", cores=cores, kind="port",
                sample=(f"oracle multifrontal LL^T (FP64) of the {int(mask.sum())} fronts of levels"
                        f" 0..{H} of the tree ({f / 1e9:.1f} GFLOP of {flops.sum() / 1e9:.1f},"
                        f" {t:.1f} s)"), seconds=round(t, 2))


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of `nbytes` (torch's elementwise copy
    kernel), run `reps` times: under rocprofv3 --pmc its FETCH_SIZE and
    WRITE_SIZE per dispatch give the factor between the counters and bytes."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}))


def bench_batched_fronts(S, M, nfronts=20000, seed=1, cpu_fronts=1500, reps=3, model=None):
    """Batched small fronts: ncol log-uniform in [8, 256], nrow = ncol (1 + r),
    r ~ U[0, 2]; partial Cholesky (POTRF, TRSM, SYRK into the contribution
    block) of all fronts in one launch, FP64 and FP32. Roofline time =
    max(total flops / MFMA peak, total bytes / HBM bandwidth) with the
    lower triangle read and written once per front. The fronts go through
    the launch plan with the learned cost model's per-front variant (or the
    default rule); the single one-workgroup-per-front launch is timed beside."""
    import oracle
    from sylver_amd import costmodel as CM
    from concurrent.futures import ThreadPoolExecutor
    nrow, ncol = M.dense_front_sizes(nfronts, seed=seed)
    m = nrow.astype(np.float64)
    n = ncol.astype(np.float64)
    q = m - n
    flops = float((n ** 3 / 3 + q * n ** 2 + q ** 2 * n).sum())
    off = np.zeros(nfronts, np.int64)
    off[1:] = np.cumsum(nrow.astype(np.int64) ** 2)[:-1]
    total = int(off[-1] + int(nrow[-1]) ** 2)
    out = {"fronts": nfronts, "seed": seed, "gflop": round(flops / 1e9, 3)}
    for dt, name, pk, tier in ((np.float64, "fp64", PEAK["fp64"], 0), (np.float32, "fp32", PEAK["fp32"], 1)):
        w = np.dtype(dt).itemsize
        nbytes = float((2 * w * m * (m + 1) / 2).sum())
        arr = S.DeviceArray.empty(total, dt)
        variant = None
        if model is not None:
            band = CM.in_band(nrow)
            pick = np.zeros(nfronts, bool)
            pick[band] = model.pick_one_wg(CM.features(nrow[band], ncol[band], np.zeros(int(band.sum())), tier))
            variant = CM.choose_variants(nrow, pick)

        def timed(run):
            best, worst = None, 0
            for _ in range(reps):
                if S.batched_front_fill(arr, nrow, off, dt, seed) != 0:
                    raise RuntimeError("front fill failed")
                rc, info, ms = run()
                if rc != 0:
                    raise RuntimeError(f"batched factorization failed: {rc}")
                worst = max(worst, int(np.abs(info).max()))
                best = ms if best is None else min(best, ms)
            return best, worst

        one_ms, _ = timed(lambda: S.batched_front_potrf(arr, nrow, ncol, off, dt))
        best, worst_info = timed(lambda: S.batched_front_potrf_v(arr, nrow, ncol, off, dt, variant))
        t = best / 1e3
        bound = max(flops / (pk * 1e12), nbytes / (HBM_GBS * 1e9))
        rec = {"ms": round(best, 3), "gflops": round(flops / t / 1e9, 1), "gbytes_s": round(nbytes / t / 1e9, 1),
               "roofline_ms": round(bound * 1e3, 3), "frac": round(bound / t, 4), "info_max": worst_info,
               "variants": "learned cost model" if model is not None else "default rule (nrow <= 256)",
               "tiled_fronts": int((variant == 0).sum()) if variant is not None else int((nrow > 256).sum()),
               "one_workgroup_launch_ms": round(one_ms, 3)}
        if name == "fp64":
            # parity and CPU baseline on the first fronts: same values, copied back before factoring
            if S.batched_front_fill(arr, nrow, off, dt, seed) != 0:
                raise RuntimeError("front fill failed")
            k = min(cpu_fronts, nfronts)
            host = [arr.read(off[i], int(nrow[i]) ** 2).reshape(nrow[i], nrow[i], order="F") for i in range(k)]
            S.batched_front_potrf_v(arr, nrow, ncol, off, dt, variant)
            cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1

            def one(i):
                return oracle.front_potrf(host[i], int(ncol[i]))

            t0 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=cores) as ex:
                res = list(ex.map(one, range(k)))
            tc = time.perf_counter() - t0
            err = 0.0
            for i, (inf_o, fo) in enumerate(res[:200]):
                g = arr.read(off[i], int(nrow[i]) ** 2).reshape(nrow[i], nrow[i], order="F")
                c = int(ncol[i])
                err = max(err, float(np.abs(np.tril(g) - np.tril(fo)).max() / max(1.0, np.abs(fo).max())))
            fs = float((n[:k] ** 3 / 3 + q[:k] * n[:k] ** 2 + q[:k] ** 2 * n[:k]).sum())
            rec["parity_max_rel_err_200_fronts"] = err
            out["cpu_baseline"] = {"value": round(fs / tc / 1e9, 3), "unit": "GFLOP/s", "cores": cores,
                                   "kind": "port", "sample": f"oracle front_potrf of the first {k} fronts"}
        arr.free()
        out[name] = rec
    return out


def bench_small_laplacian(S, M, oracle_mod):
    """2D 5-point Laplacian 100 x 100 (n = 10,000), posdef, FP64, 1 RHS."""
    A = M.laplacian_2d(100)
    opts = S.default_options()
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    flops = float(inf.num_flops)
    b = M.random_rhs(A.n, 1, seed=0)
    S.set_ext_options(precision=S.PREC_FP64, use_graph=1, record_timing=0)
    fk = None
    fms, sms = [], []
    for it in range(6):
        fk, finf = S.spldlt_factorize(True, A.val, akeep, fk, options=opts)
        x = b.copy(order="F")
        sinf = S.spldlt_solve(x, akeep, fk, options=opts)
        if it > 0:
            fms.append(finf.tail().factor_ms)
            sms.append(sinf.tail().solve_ms)
    err = M.backward_error(A, x, b)
    sym = S.akeep_symbolic(akeep)
    t0 = time.perf_counter()
    st, off, lcol = oracle_mod.factor_posdef(sym, A.val)
    tc = time.perf_counter() - t0
    fm = float(np.median(fms))
    return {"n": A.n, "gflop": round(flops / 1e9, 4), "factor_ms": round(fm, 3),
            "solve_ms": round(float(np.median(sms)), 3), "factor_gflops": round(flops / fm / 1e6, 2),
            "bwd_err": err, "cpu_factor_ms": round(tc * 1e3, 3),
            "cpu_kind": "port (oracle, all host threads)"}


def bench_lu_fronts(S, oracle_mod, nfronts=20000, seed=5, reps=3, cpu_fronts=300):
    """Dense front LU of the unsymmetric path (the reference's lu_pp_factor /
    lu_tpp_factor, src/kernels/lu_pp.hxx, lu_tpp.hxx), batched: orders m
    log-uniform in [8, 128], every column eliminated, entries U[-1, 1]; one
    launch of sylver_amd_batched_front_lu per mode and precision. Flops
    2 m^3 / 3 per front; bytes: the front read and written once. The first
    fronts are checked bit for bit against the oracle, whose time on one host
    thread over cpu_fronts fronts is the CPU baseline."""
    rng = np.random.default_rng(seed)
    m = np.exp(rng.uniform(np.log(8), np.log(128.999), nfronts)).astype(np.int32)
    off = np.zeros(nfronts, np.int64)
    off[1:] = np.cumsum(m.astype(np.int64) ** 2)[:-1]
    total = int(off[-1] + int(m[-1]) ** 2)
    host = rng.uniform(-1.0, 1.0, total)
    flops = float((2.0 / 3.0 * m.astype(np.float64) ** 3).sum())
    out = {"fronts": nfronts, "seed": seed, "orders": "log-uniform [8, 128]", "gflop": round(flops / 1e9, 3)}
    for mode in ("pp", "tpp"):
        for dt, name in ((np.float64, "fp64"), (np.float32, "fp32")):
            w = np.dtype(dt).itemsize
            nbytes = float((2 * w * m.astype(np.float64) ** 2).sum())
            best = None
            for _ in range(reps):
                arr = S.DeviceArray(host.astype(dt))
                rc, ne, perms = S.batched_front_lu(arr, m, m, off, dt, mode, 0.01)
                ms = float(S.lib().sylver_amd_last_batched_ms())
                if rc != 0:
                    raise RuntimeError(f"batched LU failed: {rc}")
                best = ms if best is None else min(best, ms)
                if _ == 0:
                    # parity on the first fronts, bit for bit
                    res = arr.read(0, int(off[16]))
                    for f in range(16):
                        a = host[off[f]:off[f] + m[f] ** 2].astype(dt).reshape((m[f], m[f]), order="F")
                        ne_o, rp_o, cp_o, f_o = oracle_mod.lu_front(a, int(m[f]), mode, 0.01)
                        got = res[off[f]:off[f] + m[f] ** 2].reshape((m[f], m[f]), order="F")
                        if ne[f] != ne_o or not np.array_equal(perms[f][0], rp_o) or \
                                np.asfortranarray(got).tobytes("F") != np.asfortranarray(f_o).tobytes("F"):
                            raise RuntimeError(f"LU parity failed at front {f} ({mode}, {name})")
                arr.free()
            t = best / 1e3
            out[f"{mode}_{name}"] = {"ms": round(best, 3), "gflops": round(flops / t / 1e9, 1),
                                     "gbytes_s": round(nbytes / t / 1e9, 1), "parity_first16": "bitwise"}
    # CPU baseline: the oracle on one host thread, FP64 partial pivoting
    t0 = time.perf_counter()
    cf = 0.0
    for f in range(cpu_fronts):
        a = host[off[f]:off[f] + m[f] ** 2].reshape((m[f], m[f]), order="F")
        oracle_mod.lu_front(a, int(m[f]), "pp")
        cf += 2.0 / 3.0 * float(m[f]) ** 3
    tc = time.perf_counter() - t0
    out["cpu_baseline"] = {"kind": "port (oracle, one thread)", "fronts": cpu_fronts,
                           "gflops": round(cf / tc / 1e9, 3)}
    return out


def bench_kkt(S, M, oracle_mod, side=48, mrows=27648, sample_side=16):
    """KKT [H A^T; A 0], H the 7-point Laplacian on side^3, A mrows x side^3
    with 3 entries per row at random columns: APP LDL^T (u = 0.01, small =
    1e-20) in FP64, and in the FP32 tier (contribution-block updates on FP32
    MFMA, pivot tests and D in FP64) whose solve runs FP64 GMRES-IR to
    backward error <= 1e-15; the FP64 factor also with the matching-based
    ordering and its scaling. Inertia must be (side^3, mrows, 0). The oracle
    (one thread) and the device are also timed on a sample_side^3 case with
    mrows = sample_side^3 / 4, where the oracle's delay counts are compared."""
    def run(A, reps, precision=None, ordering=1):
        opts = S.default_options()
        opts.ordering = ordering
        akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
        if ordering == 2:
            opts.scaling = 3
        S.set_ext_options(precision=S.PREC_FP64 if precision is None else precision, use_graph=1, record_timing=0,
                          refine=S.REFINE_GMRES, refine_tol=1e-15, refine_max_iter=20)
        fk, fms = None, []
        for it in range(reps):
            fk, finf = S.spldlt_factorize(False, A.val, akeep, fk, options=opts)
            if finf.flag < 0:
                raise RuntimeError(f"indefinite factorization failed: {finf.flag}")
            if it > 0 or reps == 1:
                fms.append(finf.tail().factor_ms)
        return akeep, inf, fk, finf, float(np.median(fms)), opts

    def record(A, akeep, ainf, fk, finf, fms, opts, nrhs=1):
        b = M.random_rhs(A.n, nrhs, seed=0)
        x = b.copy(order="F")
        sinf = S.spldlt_solve(x, akeep, fk, options=opts)
        err = M.backward_error(A, x, b)
        st = sinf.tail()
        rec = {"flops_symbolic": float(ainf.num_flops), "flops_executed": float(finf.num_flops),
               "factor_ms": round(fms, 3), "factor_gflops": round(float(finf.num_flops) / (fms * 1e6), 2),
               "solve_ms": round(st.solve_ms, 3), "bwd_err": err,
               "num_neg": int(finf.num_neg), "inertia_ok": int(finf.num_neg) == mrows and finf.matrix_rank == A.n,
               "num_delay": int(finf.num_delay), "num_two": int(finf.num_two), "maxfront": int(finf.maxfront)}
        if st.refine_iters > 0:
            rec["refine_iters"] = int(st.refine_iters)
            rec["refine_mode"] = {1: "IR", 2: "GMRES-IR"}.get(int(st.refine_mode), int(st.refine_mode))
            rec["factor_plus_solve_ms"] = round(fms + st.solve_ms, 3)
        S.spldlt_free_fkeep(fk)
        S.spldlt_free_akeep(akeep)
        return rec

    A = M.kkt(side, mrows, seed=0)
    nh = side ** 3
    out = {"n": A.n, "nh": nh, "m": mrows}
    out.update(record(A, *run(A, 2)))
    out["fp32_gmres_ir"] = record(A, *run(A, 2, precision=S.PREC_FP32))
    out["matching_order_fp64"] = record(A, *run(A, 1, ordering=2))
    S.set_ext_options(precision=S.PREC_FP64, refine=S.REFINE_AUTO)
    # the bounded sample: device and oracle on the same smaller case
    ms = sample_side ** 3 // 4
    A2 = M.kkt(sample_side, ms, seed=0)
    akeep2, _, fk2, finf2, fms2, opts2 = run(A2, 2)
    st = S.fkeep_indef_stats(fk2)
    sym = S.akeep_symbolic(akeep2)
    t0 = time.perf_counter()
    ref = oracle_mod.IndefFactor(sym, A2.val, nb=64, u=opts2.u, small=opts2.small)
    tc = time.perf_counter() - t0
    out["sample"] = {"side": sample_side, "m": ms, "n": A2.n, "flops_executed": float(finf2.num_flops),
                     "device_factor_ms": round(fms2, 3), "cpu_factor_ms": round(tc * 1e3, 1),
                     "cpu_gflops": round(float(finf2.num_flops) / (tc * 1e9), 3),
                     "cpu_kind": "port (oracle, one thread)",
                     "stats_equal_oracle": all(st[k] == ref.stats[k] for k in st)}
    S.spldlt_free_fkeep(fk2)
    S.set_ext_options(precision=S.PREC_FP64, refine=S.REFINE_AUTO)
    return out


def bench_costmodel(S, M, n_train=None, n_held=None, full=False, seed=5, return_model=False, e2e_max=None):
    """Learned cost model (per-front kernel variant, one-workgroup vs tiled):
    a seeded stream of SPD grid matrices (2D 5-point, 3D 7-point with aspect
    ratios up to 4:1), the autotuner's exhaustive per-shape timings (FP64 and
    FP32) as targets, the device-trained MLP's choices on the held-out
    matrices: per-front regret against the autotuner, and end-to-end factor
    time with the default rule, the autotuner's and the model's variants.
    full=True uses the configured sizes (2D sides 200-1000, 3D 30-90) and
    400 + 100 matrices; the default is a reduced stream."""
    from sylver_amd import costmodel as CM
    if n_train is None:
        n_train = 400 if full else 40
    if n_held is None:
        n_held = 100 if full else 10
    if e2e_max is None:
        e2e_max = 30 if full else n_held
    rng = np.random.default_rng(seed)
    s2, s3 = ((200, 1000), (30, 90)) if full else ((150, 400), (20, 45))

    def gen():
        if rng.uniform() < 0.4:
            nx = int(rng.integers(s2[0], s2[1] + 1))
            ny = int(min(s2[1], max(s2[0], nx * rng.uniform(0.5, 1.0))))
            return f"2d {nx}x{ny}", M.laplacian_2d(nx, ny)
        nx = int(rng.integers(s3[0], s3[1] + 1))
        ar = rng.uniform(1.0, 4.0)
        ny = max(4, int(nx / math.sqrt(ar)))
        nz = max(4, int(nx / ar))
        return f"3d {nx}x{ny}x{nz}", M.laplacian_3d(nx, ny, nz)

    t0 = time.perf_counter()
    tuners = {S.PREC_FP64: CM.Autotuner(S, S.PREC_FP64, reps=2), S.PREC_FP32: CM.Autotuner(S, S.PREC_FP32, reps=2)}
    xs, ts, held = [], [], []
    for k in range(n_train + n_held):
        name, A = gen()
        akeep, _, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=S.default_options())
        m, n, nc = CM.tree_fronts(S, akeep)
        band = CM.in_band(m)
        if k < n_train:
            for prec, tier in ((S.PREC_FP64, 0), (S.PREC_FP32, 1)):
                xs.append(CM.features(m[band], n[band], nc[band], tier))
                ts.append(tuners[prec].measure(m[band], n[band]))
            S.spldlt_free_akeep(akeep)
        else:
            held.append((name, A, akeep, m, n, nc))
        if (k + 1) % 25 == 0:
            log(f"  cost model: {k + 1} matrices analysed and tuned ({time.perf_counter() - t0:.0f} s)")
    t_data = time.perf_counter() - t0
    x, t = np.concatenate(xs), np.concatenate(ts)
    log(f"  cost model: training on {len(x)} records")
    model = CM.CostModel(S, seed=seed).fit(x, t, epochs=300, batch=256 if len(x) < 100_000 else 1024)
    log(f"  cost model: trained, final loss {float(model.loss[-1]):.4f}")
    out = {"train_matrices": n_train, "held_out_matrices": n_held, "records": int(len(x)),
           "shape_buckets_timed": len(tuners[S.PREC_FP64].cache) + len(tuners[S.PREC_FP32].cache),
           "final_train_loss_log2ms": round(float(model.loss[-1]), 4), "data_and_tuning_s": round(t_data, 1),
           "sizes": "configured (2D 200-1000, 3D 30-90)" if full else "reduced (2D 150-400, 3D 20-45)"}
    opts = S.default_options()
    for prec, tier, label in ((S.PREC_FP64, 0, "fp64"), (S.PREC_FP32, 1, "fp32")):
        reg_m = reg_d = best = 0.0
        e2e = {"default": 0.0, "autotuner": 0.0, "model": 0.0}
        S.set_ext_options(precision=prec, use_graph=1, record_timing=0)
        for h, (name, A, akeep, m, n, nc) in enumerate(held):
            band = CM.in_band(m)
            tm = tuners[prec].measure(m, n)
            pick = np.zeros(len(m), bool)
            pick[band] = model.pick_one_wg(CM.features(m[band], n[band], nc[band], tier))
            tb = tm[band]
            best += tb.min(axis=1).sum()
            reg_m += (np.where(pick[band], tb[:, 0], tb[:, 1]) - tb.min(axis=1)).sum()
            reg_d += (np.where(m[band] <= 256, tb[:, 0], tb[:, 1]) - tb.min(axis=1)).sum()
            variants = {"default": None, "autotuner": CM.choose_variants(m, tm[:, 0] < tm[:, 1]),
                        "model": CM.choose_variants(m, pick)}
            if h >= e2e_max:
                log(f"  cost model {label}: held-out {name} (regret only)")
                continue
            for key, v in variants.items():
                S.set_front_variants(akeep, v)
                fk, fms = None, []
                for it in range(3):
                    fk, finf = S.spldlt_factorize(True, A.val, akeep, fk, options=opts)
                    if finf.flag != 0:
                        raise RuntimeError(f"factorization failed: {finf.flag}")
                    if it > 0:
                        fms.append(finf.tail().factor_ms)
                S.spldlt_free_fkeep(fk)
                e2e[key] += float(np.median(fms))
            S.set_front_variants(akeep, None)
            log(f"  cost model {label}: held-out {name} done")
        out[label] = {"end_to_end_matrices": min(e2e_max, len(held)),
                      "front_time_best_ms": round(best, 3), "regret_model_ms": round(reg_m, 3),
                      "regret_default_rule_ms": round(reg_d, 3),
                      "end_to_end_factor_ms": {k: round(v, 3) for k, v in e2e.items()}}
    S.set_ext_options(precision=S.PREC_FP64)
    for _, _, akeep, _, _, _ in held:
        S.spldlt_free_akeep(akeep)
    return (out, model) if return_model else out


def measured_traffic():
    """HBM bytes per launch of the Schur-update kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, calibrated against a
    copy kernel); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_update_fp32_laplacian80_pmc.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=80)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--mixed", default="fp32", choices=["fp32", "mixed"])
    ap.add_argument("--paths", default="both", choices=["both", "mixed", "fp64"],
                    help="run only one precision path (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (scales PMC byte counters)")
    ap.add_argument("--profile-only", action="store_true",  # with --paths fp64: the FP64 factorization
                    help="only the per-launch timed (eager, one stream) factorization of the FP32 path")
    ap.add_argument("--variants", default="model", choices=["model", "default"],
                    help="per-front kernel variants: the learned cost model's (trained first) or the default rule")
    ap.add_argument("--variants-file", default=None,
                    help="per-front variants of both tiers as .npz: loaded when the file exists (no training), "
                         "else saved there after training (profiled runs reuse the same choices)")
    ap.add_argument("--costmodel-full", action="store_true",
                    help="learned cost model on the configured stream (400 + 100 matrices, full sizes)")
    ap.add_argument("--only", default=None, choices=["batched", "kkt", "lu"],
                    help="run one secondary workload alone and print its record (profiling runs)")
    ap.add_argument("--kkt-side", type=int, default=48, help="grid side of H for --only kkt")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the secondary workloads (2D Laplacian, batched fronts, KKT)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import sylver_amd as S
    from sylver_amd import matrices as M

    if S.device_count() < 1:
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(3)
    if args.only == "batched":
        print(json.dumps(bench_batched_fronts(S, M, model=None, cpu_fronts=200), default=float))
        return
    if args.only == "lu":
        import oracle as oracle_mod
        print(json.dumps(bench_lu_fronts(S, oracle_mod), default=float))
        return
    if args.only == "kkt":
        import oracle as oracle_mod
        side = args.kkt_side
        print(json.dumps(bench_kkt(S, M, oracle_mod, side=side, mrows=27648 if side == 48 else side ** 3 // 4),
                         default=float))
        return
    g = args.grid
    A = M.laplacian_3d(g)
    b = M.random_rhs(A.n, args.nrhs, seed=0)
    opts = S.default_options()
    t0 = time.perf_counter()
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    t_an = time.perf_counter() - t0
    if inf.flag < 0:
        raise RuntimeError(f"analyse failed: {inf.flag}")
    flops = float(inf.num_flops)
    log(f"analyse {t_an:.2f} s: n={A.n} nodes={inf.num_sup} maxfront={inf.maxfront}"
        f" flops={flops:.4e}")

    prec = S.PREC_FP32 if args.mixed == "fp32" else S.PREC_MIXED
    # per-front kernel variants proposed by the learned cost model, trained
    # first on its own matrix stream (also the config's report below)
    cm_report, model = None, None
    vcache = {}  # tier -> per-front variant array
    loaded = bool(args.variants_file) and os.path.exists(args.variants_file)
    if loaded:
        with np.load(args.variants_file) as z:
            vcache = {0: z["v0"].astype(np.int8), 1: z["v1"].astype(np.int8)}
    elif args.variants == "model":
        cm_report, model = bench_costmodel(S, M, full=args.costmodel_full, return_model=True)
    from sylver_amd import costmodel as CM
    fm, fn, fc = CM.tree_fronts(S, akeep)

    def variants_for(tier):
        if tier not in vcache:
            if model is None:
                return None
            band = CM.in_band(fm)
            pick = np.zeros(len(fm), bool)
            pick[band] = model.pick_one_wg(CM.features(fm[band], fn[band], fc[band], tier))
            vcache[tier] = CM.choose_variants(fm, pick)
        return vcache[tier]

    def use_variants(tier):
        S.set_front_variants(akeep, variants_for(tier))

    if args.variants_file and not loaded and model is not None:
        np.savez(args.variants_file, v0=variants_for(0), v1=variants_for(1))

    if args.profile_only:  # rocprofv3 runs of the roofline figure: the same eager launches
        pr = S.PREC_FP64 if args.paths == "fp64" else prec
        use_variants(0 if pr == S.PREC_FP64 else 1)
        kinds, fms, _ = kernel_profile(S, A, akeep, opts, pr, None)
        print(json.dumps({"profile_only": True, "kernels": kinds}))
        return
    if args.paths != "both":  # profiling runs: one path, no report
        pr = S.PREC_FP64 if args.paths == "fp64" else prec
        use_variants(0 if pr == S.PREC_FP64 else 1)
        _, recs = run_path(S, A, akeep, opts, b, pr, args.steps, args.warmup, 1e-15)
        print(json.dumps({"paths": args.paths, "factor_ms": [r["factor_ms"] for r in recs],
                          "solve_ms": [r["solve_ms"] for r in recs]}))
        return
    # plain FP64 path
    use_variants(0)
    fk64, r64 = run_path(S, A, akeep, opts, b, S.PREC_FP64, args.steps, args.warmup, 1e-15)
    err64 = M.backward_error(A, r64[-1]["x"], b)
    # mixed path (headline)
    use_variants(1)
    fkmx, rmx = run_path(S, A, akeep, opts, b, prec, args.steps, args.warmup, 1e-15)
    errmx = M.backward_error(A, rmx[-1]["x"], b)
    # the other low-precision tier beside the headline one (FP32 only, or FP32
    # with BF16 updates on fronts of more than 1024 pivots)
    alt_prec = S.PREC_MIXED if prec == S.PREC_FP32 else S.PREC_FP32
    fkalt, ralt = run_path(S, A, akeep, opts, b, alt_prec, args.steps, args.warmup, 1e-15)
    erralt = M.backward_error(A, ralt[-1]["x"], b)
    S.spldlt_free_fkeep(fkalt)
    for name, e in (("fp64", err64), ("headline", errmx), ("other tier", erralt)):
        if not (e <= 1e-15):
            raise RuntimeError(f"{name} path: scaled backward error {e} above 1e-15")

    f64_ms = float(np.mean([r["factor_ms"] for r in r64]))
    s64_ms = float(np.mean([r["solve_ms"] for r in r64]))
    fmx_ms = float(np.mean([r["factor_ms"] for r in rmx]))
    smx_ms = float(np.mean([r["solve_ms"] for r in rmx]))
    step_ms = fmx_ms + smx_ms

    roof = None
    kinds64 = kindsmx = None
    if not args.no_profile:
        use_variants(0)
        kinds64, _, fk64 = kernel_profile(S, A, akeep, opts, S.PREC_FP64, fk64)
        use_variants(1)
        kindsmx, _, fkmx = kernel_profile(S, A, akeep, opts, prec, fkmx)
        up = kindsmx.get("schur_update")
        wide = kindsmx.get("schur_update_wide")
        if wide and up:
            up = {"ms": up["ms"] + wide["ms"], "gflop": up["gflop"] + wide["gflop"],
                  "launches": up["launches"] + wide["launches"]}
        elif wide:
            up = wide
        if up and up["ms"] > 0:
            ach = up["gflop"] / up["ms"]  # GFLOP/ms == TFLOP/s
            pk = PEAK["fp32"]
            name = ("FP32 Schur update, MFMA 16x16x4: k_update_far_f32 (128x128 tiles, eight waves, the far "
                    "look-ahead updates) and k_update<float, 0, 64, 64> (the critical-path launches)"
                    if args.mixed == "fp32"
                    else "k_update_far_f32 + k_update<float, 0, 64, 64> + k_update_bf16 (Schur update of large fronts)")
            roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": pk, "unit": "TFLOP/s",
                    "frac": round(ach / pk, 4), "traffic": measured_traffic(),
                    "kernel": name, "launches": up["launches"],
                    "avg_launch_ms": round(up["ms"] / up["launches"], 5),
                    "gflop_per_launch": round(up["gflop"] / up["launches"], 4)}
    S.set_ext_options(precision=S.PREC_FP64)

    secondary = None
    if not args.no_secondary:
        import oracle as oracle_mod
        secondary = {}
        for key, fn in (("laplacian_2d_100", lambda: bench_small_laplacian(S, M, oracle_mod)),
                        ("batched_small_fronts", lambda: bench_batched_fronts(S, M, model=model)),
                        ("kkt_app_48cubed_27648", lambda: bench_kkt(S, M, oracle_mod)),
                        ("dense_front_lu", lambda: bench_lu_fronts(S, oracle_mod)),
                        ("learned_cost_model",
                         lambda: cm_report if cm_report is not None else bench_costmodel(S, M, full=args.costmodel_full))):
            try:
                secondary[key] = fn()
            except Exception as e:  # a secondary workload never blocks the headline
                secondary[key] = {"error": str(e)}
            log(f"secondary {key}: {secondary[key]}")
        S.set_ext_options(precision=S.PREC_FP64)

    cpu = None
    if not args.no_cpu:
        try:
            sym = S.akeep_symbolic(akeep)
            cpu = cpu_baseline(sym, A, args.cpu_seconds)
        except Exception as e:  # the baseline never blocks the GPU figures
            cpu = {"error": str(e)}

    out = {
        "metric": "factor time (s) & GFLOP/s: mixed+IR vs FP64, 80^3 Laplacian, bwd err <=1e-15",
        "value": round(flops / (step_ms * 1e6), 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64-ir" if args.mixed == "fp32" else "bf16/fp32+fp64-ir",
        "data": "synthetic",
        "config": {"workload": f"3D 7-point Laplacian {g}^3 (n={A.n}), nested dissection,"
                               f" {args.nrhs} RHS U[-1,1] seed 0, posdef LL^T",
                   "n": A.n, "nrhs": args.nrhs, "nodes": int(inf.num_sup),
                   "maxfront": int(inf.maxfront), "flops": flops,
                   "front_variants": ("learned cost model (per front, trained in this run)" if model is not None
                                      else "learned cost model (per front, choices loaded from --variants-file)"
                                      if loaded else "default rule (one workgroup for nrow <= 256)")},
        "mixed": {"factor_s": round(fmx_ms / 1e3, 5), "solve_ir_s": round(smx_ms / 1e3, 5),
                  "factor_gflops": round(flops / (fmx_ms * 1e6), 2),
                  "ir_iters": int(rmx[-1]["refine_iters"]), "bwd_err": errmx},
        "fp64": {"factor_s": round(f64_ms / 1e3, 5), "solve_s": round(s64_ms / 1e3, 5),
                 "factor_gflops": round(flops / (f64_ms * 1e6), 2), "bwd_err": err64,
                 "frac_fp64_peak": round(flops / (f64_ms * 1e6) / 1e3 / PEAK["fp64"], 4)},
        "other_tier": {"tier": "bf16/fp32+fp64-ir (BF16 updates on fronts with > 1024 pivots)"
                       if alt_prec == S.PREC_MIXED else "fp32+fp64-ir",
                       "factor_s": round(float(np.mean([r["factor_ms"] for r in ralt])) / 1e3, 5),
                       "solve_ir_s": round(float(np.mean([r["solve_ms"] for r in ralt])) / 1e3, 5),
                       "ir_iters": int(ralt[-1]["refine_iters"]), "bwd_err": erralt,
                       "step_gflops": round(flops / (float(np.mean([r["factor_ms"] + r["solve_ms"]
                                                                     for r in ralt])) * 1e6), 2)},
        "call_wall_ms": {
            "convention": "wall clock around each spldlt_factorize / spldlt_solve call (synchronous at return;"
                          " values and x handed over from host memory), as drivers/spldlt_test.F90:133-148",
            "headline": {"factorize": round(float(np.mean([r["factor_call_ms"] for r in rmx])), 3),
                         "solve": round(float(np.mean([r["solve_call_ms"] for r in rmx])), 3)},
            "fp64": {"factorize": round(float(np.mean([r["factor_call_ms"] for r in r64])), 3),
                     "solve": round(float(np.mean([r["solve_call_ms"] for r in r64])), 3)},
            "headline_step_gflops": round(flops / (float(np.mean([r["wall_ms"] for r in rmx])) * 1e6), 2),
            "speedup_vs_fp64": round(float(np.mean([r["wall_ms"] for r in r64])) /
                                     float(np.mean([r["wall_ms"] for r in rmx])), 3)},
        "speedup_mixed_vs_fp64": round((f64_ms + s64_ms) / step_ms, 3),
        "upload_ms": round(float(np.mean([r["upload_ms"] for r in rmx])), 3),
        "wall_ms_per_step_pcie_incl": round(float(np.mean([r["wall_ms"] for r in rmx])), 3),
        "roofline": roof,
        "kernels_fp64": kinds64,
        "kernels_mixed": kindsmx,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
