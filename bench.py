"""Benchmark of the HCI variational hot path on one MI355X.

One "step" is one pass of the hot path over one determinant list: the sparse
H build from HBM-resident determinants followed by Davidson to E0 (lowest
root, residual 1e-4 and energy change 1e-8).  Workloads (--config):

  1 (default)  n_orb=40, 6 alpha + 6 beta electrons, synthetic molecular
               integrals (h_pp = -1 + 2p/15, h_pq ~ N(0, 0.05^2),
               (pq|rs) = sum_K L^K_pq L^K_rs, K = 80, L ~ N(0, 0.08^2), seed 1)
               over a 2e6-determinant space from the deterministic
               energy-truncated generator (risq_amd.dets).
  2            6x6 periodic Hubbard lattice, t = 1, U = 4, on-site disorder
               U(-0.1, 0.1) (seed 2), 18 + 18 electrons in the site basis;
               the space is the one the HCI selection (eps_var = 1e-3, library
               heat-bath selection) holds after its first 10 iterations
               (1.08e7 determinants), built before the timed region.
  4            n_orb=80, 10 alpha + 10 beta electrons, the same molecular
               recipe with K = 160 (seed 4), energy-truncated space of
               --n-dets determinants (default 1.2e7: its FP32 hi+lo H, dense
               mirror window pairs stored once, takes 272 GB of the 288 GB;
               1.6e7 does not fit).

Timed path (value): tiled storage with FP32 hi+lo values (plus an FP16 copy of
each value for workloads 1 and 2: --values), Davidson with FP16 sigmas, then
FP32 sigmas from 3x the tolerance on, then FP64 refinement (FP32 then FP64 for
workload 4), thick restarts (workload 1: subspace 16, 4 vectors kept;
others: subspace 11, 2 kept; --max-sub, --keep), and the correction chosen by
--precond (learned for workload 1, diagonal otherwise).  Reported beside it from the same run: the
plain FP64 GPU path (FP64 values, the reference's restart-to-one-vector control
flow, diagonal correction) and its E0, the other correction, the H-build
nonzeros/s, the sigma products' HBM-roofline fraction (algorithmic bytes / HIP
event time of the products inside the timed steps), and the CPU oracle timed
on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# (8,337,145.7 + 1,110,114.9) KiB x 2 x 1024 read + (15,901.0 + 15,834.9) KiB x 1024 written
MEASURED_F16_PRODUCT_BYTES = 19_380_487_270


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def molecular_inputs(n_dets: int, n_orb: int, n_el: int, n_chol: int, seed: int):
    from risq_amd.dets import energy_truncated_space
    from risq_amd.fcidump import synthetic_molecular
    ints = synthetic_molecular(n_orb, n_chol, seed)
    eps = np.array([ints.h1(p, p) for p in range(n_orb)])
    dets = energy_truncated_space(eps, n_el, n_el, n_dets)
    return ints, dets


def hubbard_inputs(ctx, eps_var: float = 1e-3, iterations: int = 10):
    """The determinants the HCI loop holds after `iterations` iterations on the
    6x6 Hubbard lattice (library selection, FP64 builds, reference-flow
    Davidson between selections)."""
    from risq_amd import hci_select as hs
    from risq_amd.fcidump import hubbard_2d
    ints = hubbard_2d(6, 6, 1.0, 4.0, 0.1, 2)
    solver = hs.GpuSolver(ctx, ints)
    r = hs.variational(solver, ints, 18, 18, eps_var, max_iters=iterations, log=log)
    solver.close()
    return ints, hs._words(r.dets)


def make_inputs(config: int, n_dets: int | None, ctx):
    """(integrals, (n, 4) uint64 determinant words, workload description)."""
    if config == 1:
        n = n_dets or 2_000_000
        ints, dets = molecular_inputs(n, 40, 6, 80, 1)
        return ints, dets, ("molecular n_orb=40, 6a6b, K=80, seed 1, energy-truncated space, "
                            f"{dets.shape[0]} dets")
    if config == 2:
        ints, dets = hubbard_inputs(ctx)
        return ints, dets, ("Hubbard 6x6 periodic, t=1, U=4, disorder 0.1 (seed 2), 18a18b, "
                            f"HCI space after 10 iterations at eps_var=1e-3, {dets.shape[0]} dets")
    if config == 4:
        n = n_dets or 12_000_000
        ints, dets = molecular_inputs(n, 80, 10, 160, 4)
        return ints, dets, ("molecular n_orb=80, 10a10b, K=160, seed 4, energy-truncated space, "
                            f"{dets.shape[0]} dets")
    raise ValueError(f"unknown workload {config}")


def hci_space_run(ctx, ints, failures) -> dict:
    """Workload 1's integrals over the space the HCI selection reaches
    (eps_var 2e-3, six iterations; library selection with FP64 FULL builds and
    reference-flow Davidson, untimed), then one untimed pass of the timed path
    (tiled FP16 -> FP32 -> refinement, subspace 16 keeping 4, learned
    correction) and of the plain FP64 path on it."""
    from risq_amd import hci_select as hs
    from risq_amd.hci import PREC_F16, PREC_F64, PRECOND_DPR, PRECOND_MLP, STORE_TILED
    t = time.perf_counter()
    solver = hs.GpuSolver(ctx, ints)
    r = hs.variational(solver, ints, 6, 6, 2e-3, max_iters=6, log=log)
    solver.close()
    t_sel = time.perf_counter() - t
    words = hs._words(r.dets)
    diag = np.asarray(r.diag if hasattr(r, "diag") else ctx.ham_diag(words))
    ctx.set_integrals(ints)
    from risq_amd import precond_train
    precond_train.load(ctx)
    res = {"n_dets": int(words.shape[0]), "selection_s": round(t_sel, 2),
           "space": "HCI selection at eps_var 2e-3, 6 iterations (untimed)"}
    for name, prec, mixed, keep, pc, max_sub in (("fast", PREC_F16, 2, 4, PRECOND_MLP, 16),
                                                 ("fp64_reference_path", PREC_F64, 0, 1, PRECOND_DPR, 11)):
        ts = time.perf_counter()
        h = ctx.build_h(words, diag, value_prec=prec, storage=STORE_TILED)
        tb = time.perf_counter() - ts
        d = ctx.davidson(h, tol=1e-4, etol=1e-8, max_iters=400, mixed=mixed, restart_keep=keep, precond=pc,
                         max_sub=max_sub)
        ctx.sync()
        tt = time.perf_counter() - ts
        nnz, nd = h.nnz, words.shape[0]
        res["nnz_u"] = nnz
        h.free(recycle=True)
        st = d.stats
        n16, n32, nfull = st["n_sigma_f16"], st["n_sigma_f32"], st["n_sigma_full"]
        # full-precision products read 8 bytes per value (FP64, or FP32 hi + lo)
        alg = (n16 * spmv_bytes(nnz, nd, 2) + n32 * spmv_bytes(nnz, nd, 4) + nfull * spmv_bytes(nnz, nd, 8))
        frac = alg / (st["sigma_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS if st["sigma_ms"] > 0 else None
        res[name] = {"s_to_E0": round(tt, 4), "build_s": round(tb, 4), "E0": d.energy, "iters": st["iters"],
                     "products": {"f16": n16, "f32": n32, "full": nfull},
                     "sigma_ms_per_product": round(st["sigma_ms"] / max(1, n16 + n32 + nfull), 3),
                     "sigma_roofline_frac": round(frac, 4) if frac is not None else None,
                     "converged": bool(d.converged)}
        log(f"hci space {name}: {tt:.3f}s build {tb:.3f}s E0={d.energy:.10f} {d.stats}")
        if not d.converged:
            failures.append(f"HCI-selected space, {name}: not converged")
    res["speedup_vs_fp64_reference"] = round(res["fp64_reference_path"]["s_to_E0"] / res["fast"]["s_to_E0"], 3)
    res["abs_dE"] = abs(res["fast"]["E0"] - res["fp64_reference_path"]["E0"])
    if res["abs_dE"] > 1e-6:
        failures.append("HCI-selected space: |E0 - E0(FP64)| exceeds 1 uHa")
    return res


def spmv_bytes(nnz_u: int, n: int, value_bytes: int) -> int:
    # upper-triangle (col + value) + row_ptr + diag + x + y
    return nnz_u * (value_bytes + 4) + 8 * (n + 1) + 8 * n + 8 * n + 8 * n


def host_cpu() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(ints, dets, n_full: int, nnz_full: int, fp64_iters: int | None, full_matrix=None) -> dict:
    """The oracle (C++ restatement of the reference path, one core) on bounded
    prefixes of the same determinant list.  The build is timed at four prefix
    sizes and fitted as t = a N^b (its singles search is all-pairs, so b is
    near 2), then extrapolated to the full list; the product is timed on the
    largest prefix and charged per stored pair; Davidson is timed on that
    prefix.  With `full_matrix` (a callable returning the full list's
    caller-order upper CSR, built on the GPU) the oracle's product is also
    timed at full size, one core.  The full-size time to E0 is extrapolated
    as build(N) + iterations x product(N), with the iteration count of the
    plain FP64 GPU path (the same control flow)."""
    from oracle import pyoracle
    H = pyoracle.OracleHam(ints)
    sizes = [n for n in (5000, 10000, 20000, 30000) if n <= dets.shape[0]]
    builds = []
    S = None
    for n_s in sizes:
        sub = dets[:n_s]
        diag = H.ham_diag(sub)
        t = time.perf_counter()
        S = pyoracle.OracleSparse().build(H, sub, diag)
        builds.append((n_s, time.perf_counter() - t, S.nnz))
        log(f"cpu oracle build: N={n_s} {builds[-1][1]:.2f}s nnz_u={S.nnz}")
    n_s, t_build, nnz = builds[-1]
    x = np.random.default_rng(0).standard_normal(n_s)
    reps = 0
    t = time.perf_counter()
    while reps < 3 or (time.perf_counter() - t < 1.0 and reps < 50):
        S.spmv(x)
        reps += 1
    t_spmv = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    rc, e0, _, iters, nsig = S.davidson(tol=1e-4, etol=1e-8, max_iters=300)
    t_dav = time.perf_counter() - t
    out = {
        "value": round(t_build + t_dav, 4), "unit": "s to E0 (sample)", "cores": 1, "kind": "port",
        "sample": f"first {n_s} determinants of the same list (nnz_u={nnz}); "
                  f"oracle build {t_build:.2f}s + Davidson {t_dav:.2f}s ({iters} iters, {nsig} sigmas)",
        "host": host_cpu(),
        "h_build_nnz_per_s": nnz / t_build,
        "spmv_gbps": spmv_bytes(nnz, n_s, 8) / t_spmv / 1e9,
        "spmv_nnz_per_s": nnz / t_spmv,
        "davidson_converged": rc == 0,
        "build_sweep": [{"n": a, "s": round(b, 4), "nnz_u": c} for a, b, c in builds],
    }
    t_prod_measured = None
    if full_matrix is not None:
        rp, col, val, dg = full_matrix()
        xf = np.random.default_rng(1).standard_normal(n_full)
        yf = np.zeros(n_full)
        L = pyoracle.lib()
        reps = []
        for _ in range(2):
            t = time.perf_counter()
            L.or_spmv_csr(n_full, rp.ctypes.data, col.ctypes.data, val.ctypes.data, dg.ctypes.data, xf.ctypes.data,
                          yf.ctypes.data)
            reps.append(time.perf_counter() - t)
        t_prod_measured = min(reps)
        del rp, col, val, dg
        out["full_size_spmv"] = {
            "n": n_full, "nnz_u": nnz_full, "s_per_product": round(t_prod_measured, 3),
            "pairs_per_s": nnz_full / t_prod_measured,
            "gbps": spmv_bytes(nnz_full, n_full, 8) / t_prod_measured / 1e9,
            "what": "the oracle's upper-triangle product (src/var/sparse.rs:64-78), one core, over the full "
                    "list's matrix (built on the GPU, FULL storage, downloaded in caller order)"}
        log(f"cpu full-size product: {t_prod_measured:.2f}s")
    if len(builds) >= 2:
        lx = np.log([b[0] for b in builds])
        ly = np.log([b[1] for b in builds])
        slope, icpt = np.polyfit(lx, ly, 1)
        t_build_full = float(np.exp(icpt) * n_full ** slope)
        t_prod_full = t_prod_measured if t_prod_measured is not None else nnz_full / out["spmv_nnz_per_s"]
        out["extrapolated"] = {
            "label": "extrapolated, not measured",
            "n": n_full, "build_fit_exponent": round(float(slope), 3),
            "build_s": round(t_build_full, 1),
            "product_s": round(t_prod_full, 3),
            "product_measured_at_full_size": t_prod_measured is not None,
            "davidson_iters_assumed": fp64_iters,
            "s_to_E0": None if fp64_iters is None else round(t_build_full + fp64_iters * t_prod_full, 1),
        }
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, choices=[1, 2, 4], default=1,
                    help="workload: 1 molecular n_orb=40 2e6 dets, 2 Hubbard 6x6 1e7 dets, 4 molecular n_orb=80")
    ap.add_argument("--n-dets", type=int, default=None,
                    help="determinants of the energy-truncated space (workloads 1 and 4)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-cpu-full", action="store_true",
                    help="skip the CPU baseline's full-size product (workload 1: FULL build and download)")
    ap.add_argument("--no-hci-space", action="store_true",
                    help="workload 1: skip the run on the HCI-selected space of the same integrals")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the untimed comparison runs (FP64 reference path, other preconditioner)")
    ap.add_argument("--values", choices=["f16", "f32"], default=None,
                    help="lowest value precision of the timed path: f16 (FP16 copy, FP16 -> FP32 -> FP64 "
                         "sigmas) or f32 (FP32 -> FP64); default f16 for workloads 1 and 2, f32 for 4")
    ap.add_argument("--max-sub", type=int, default=None,
                    help="Davidson subspace size (default 16 for workload 1, 11 otherwise)")
    ap.add_argument("--keep", type=int, default=None,
                    help="vectors kept on restart (default 4 for workload 1, 2 otherwise)")
    ap.add_argument("--precond", choices=["mlp", "dpr"], default=None,
                    help="correction of the timed path (the other one is reported beside it); "
                         "default mlp for workload 1, dpr otherwise")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures a single MI355X (--gpus 1); sharding the path over several GPUs "
              "is a design note, not an implemented or measured path", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    from risq_amd import precond_train
    from risq_amd.hci import (PREC_F16, PREC_F32, PREC_F64, PRECOND_DPR, PRECOND_MLP, STORE_TILED, DeviceBuffer,
                              HciContext)

    t0 = time.perf_counter()
    ctx = HciContext(0)
    ints, dets, workload = make_inputs(args.config, args.n_dets, ctx)
    n = dets.shape[0]
    log(f"inputs ready: {workload} ({time.perf_counter() - t0:.1f}s)")
    ctx.set_integrals(ints)
    precond_train.load(ctx)
    if args.precond is None:
        # the correction is trained on n_orb <= 40 molecules: on the lattice
        # and the n_orb = 80 molecule the diagonal correction converges sooner
        args.precond = "mlp" if args.config == 1 else "dpr"
    if args.max_sub is None:
        args.max_sub = 16 if args.config == 1 else 11
    if args.keep is None:
        args.keep = 4 if args.config == 1 else 2
    precond = PRECOND_MLP if args.precond == "mlp" else PRECOND_DPR
    diag = ctx.ham_diag(dets)
    d_dets = DeviceBuffer.from_array(ctx, dets)
    d_diag = DeviceBuffer.from_array(ctx, diag)
    # one device-to-device copy of 256 MiB: a kernel of known traffic that
    # calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters for this run
    cal_a, cal_b = DeviceBuffer(ctx, 256 << 20), DeviceBuffer(ctx, 256 << 20)
    ctx._check(ctx.lib.hci_memcpy(ctx.handle, cal_b.ptr, cal_a.ptr, 256 << 20, 1, 1))
    cal_a.free()
    cal_b.free()

    def run(prec, mixed, keep, pc, max_sub=11):
        ts = time.perf_counter()
        h = ctx.build_h(d_dets, d_diag, n=n, value_prec=prec, storage=STORE_TILED)
        tb = time.perf_counter() - ts
        res = ctx.davidson(h, tol=1e-4, etol=1e-8, max_iters=400, mixed=mixed, restart_keep=keep, precond=pc,
                           max_sub=max_sub)
        ctx.sync()
        return h, res, tb, time.perf_counter() - ts

    if args.values is None:
        # workload 4's matrix leaves no room for the FP16 copy (the lattice's
        # hopping elements, +-t with t = 1, are exact in FP16)
        args.values = "f32" if args.config == 4 else "f16"
    fast_prec, fast_mixed = (PREC_F16, 2) if args.values == "f16" else (PREC_F32, 1)

    def step():
        return run(fast_prec, fast_mixed, args.keep, precond, args.max_sub)

    failures = []  # steps whose Davidson did not converge

    for w in range(args.warmup):
        h, res, tb, tt = step()
        log(f"warmup {w}: build {tb:.3f}s total {tt:.3f}s E0={res.energy:.10f} iters={res.stats['iters']}")
        if not res.converged:
            failures.append(f"warmup {w} did not converge")
        h.free(recycle=True)
    ctx.sync()
    times, builds, results = [], [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        h, res, tb, tt = step()
        times.append(tt)
        builds.append(tb)
        results.append(res)
        log(f"step {k}: build {tb:.3f}s total {tt:.3f}s E0={res.energy:.10f} {res.stats}")
        if not res.converged:
            failures.append(f"step {k} did not converge")
        if k + 1 < args.steps:
            h.free(recycle=True)
    ctx.sync()
    wall = time.perf_counter() - t_start
    nnz = h.nnz
    h.free(recycle=True)
    last = results[-1].stats
    n16, n32, nfull = last["n_sigma_f16"], last["n_sigma_f32"], last["n_sigma_full"]
    # upper-triangle algorithmic bytes: FP16 sigmas col + FP16 value (6 B per
    # pair), FP32 sigmas col + hi (8 B), refinement sigmas col + hi + lo (12 B)
    alg_bytes = (n16 * spmv_bytes(nnz, n, 2) + n32 * spmv_bytes(nnz, n, 4) + nfull * spmv_bytes(nnz, n, 8))
    achieved = alg_bytes / (last["sigma_ms"] * 1e-3) / 1e9
    out = {
        "metric": "s to E0 within 1µHa (plain FP64 vs FP32+learned precond); H-build nnz/s; SpMV GB/s",
        "value": round(float(np.mean(times)), 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f16 then f32 values" if args.values == "f16" else "f32 values")
                 + " + f64 accumulation, f64 refinement",
        "data": "synthetic (seeded integrals; deterministic determinant space)",
        "config": {"workload": f"{workload}; H build + Davidson to E0 (tol 1e-4, etol 1e-8)",
                   "n_dets": n, "nnz_u": nnz,
                   "storage": "tiled, fp32 hi+lo" + (" + fp16 copy" if args.values == "f16" else ""),
                   "davidson": f"mixed precision, subspace {args.max_sub}, restart keeps {args.keep} vectors, "
                               f"correction={args.precond}",
                   "parallelism": "single GPU"},
        "matrix_hbm_gb": round(h.build_stats["matrix_bytes"] / 1e9, 2),
        "E0": results[-1].energy,
        "davidson": {"iters": last["iters"], "n_sigma_f16": n16, "f16_iter": last["f16_iter"],
                     "n_sigma_f32": n32, "n_sigma_full": nfull,
                     "seconds": last["seconds"], "converged": bool(last["converged"]),
                     "mlp_applied": last["mlp_applied"], "mlp_fallbacks": last["mlp_fallbacks"]},
        "h_build_nnz_per_s": nnz / float(np.mean(builds)),
        "h_build_s": float(np.mean(builds)),
        "spmv_ms": last["sigma_ms"] / max(1, n16 + n32 + nfull),
        "spmv_gbps": achieved,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_spmv_win (sigma product: all its launches in the timed step: "
                               "FP16-value, FP32-value and FP32 hi+lo)"},
    }
    # HBM bytes per FP16-value product measured with rocprofv3 on this exact
    # workload (FETCH_SIZE x2 + WRITE_SIZE of the two k_spmv_win launches,
    # calibrated against a 256 MiB copy); a counter run cannot happen inside
    # the timed process, so other workloads report none
    if args.config == 1 and n == 2_000_000 and args.values == "f16":
        out["roofline"]["traffic"] = MEASURED_F16_PRODUCT_BYTES
        out["roofline"]["traffic_source"] = ("profiles/k_spmv_win_config1_2e6dets.md: bytes per FP16-value "
                                             "product (both triangles at 4 B per entry + FP32 x slices + y)")
    fp64_iters = None
    if not args.no_compare:
        h, ref, tb, tr = run(PREC_F64, False, 1, PRECOND_DPR)
        st = ref.stats
        fp64_iters = st["iters"]
        r_ms = st["sigma_ms"] / max(1, st["n_sigma_full"])
        r_ach = spmv_bytes(h.nnz, n, 8) / (r_ms * 1e-3) / 1e9
        h.free(recycle=True)
        out["fp64_reference_path"] = {
            "s_to_E0": round(tr, 4), "build_s": round(tb, 4), "E0": ref.energy, "iters": st["iters"],
            "converged": bool(st["converged"]), "sigma_ms": r_ms, "sigma_gbps": r_ach,
            "sigma_roofline_frac": round(r_ach / HBM_PEAK_GBS, 4),
            "what": "FP64 values, restart to the Ritz vector (reference control flow), diagonal correction"}
        out["speedup_vs_fp64_reference"] = round(tr / out["value"], 3)
        out["abs_dE_vs_fp64_reference"] = abs(ref.energy - out["E0"])
        log(f"fp64 reference path: {tr:.3f}s E0={ref.energy:.10f} iters={st['iters']} sigma {r_ms:.2f} ms")
        if not ref.converged:
            failures.append("FP64 reference path did not converge")
        elif out["abs_dE_vs_fp64_reference"] > 1e-6:
            failures.append(f"|E0 - E0(FP64)| = {out['abs_dE_vs_fp64_reference']:.2e} Ha exceeds 1 uHa")

        def compare(name, prec, mixed, keep, pc, max_sub, what):
            h, r, tb, tt = run(prec, mixed, keep, pc, max_sub)
            h.free(recycle=True)
            st = r.stats
            out[name] = {"s_to_E0": round(tt, 4), "build_s": round(tb, 4), "E0": r.energy,
                         "abs_dE_vs_fp64_reference": abs(r.energy - ref.energy), "iters": st["iters"],
                         "n_sigma_f16": st["n_sigma_f16"], "n_sigma_f32": st["n_sigma_f32"],
                         "n_sigma_full": st["n_sigma_full"], "mlp_applied": st["mlp_applied"],
                         "mlp_fallbacks": st["mlp_fallbacks"], "converged": bool(st["converged"]), "what": what}
            log(f"{name}: {tt:.3f}s E0={r.energy:.10f} {st}")
            if not st["converged"]:
                failures.append(f"{name} did not converge")
            elif abs(r.energy - ref.energy) > 1e-6:
                failures.append(f"{name}: |E0 - E0(FP64)| exceeds 1 uHa")

        other = PRECOND_DPR if precond == PRECOND_MLP else PRECOND_MLP
        oname = "dpr" if other == PRECOND_DPR else "mlp"
        compare("other_correction", fast_prec, fast_mixed, args.keep, other, args.max_sub,
                f"the timed path with the {oname} correction")
        compare("fp64_same_restart", PREC_F64, 0, args.keep, precond, args.max_sub,
                f"FP64 values with the timed path's restart (subspace {args.max_sub}, keep {args.keep}) and "
                f"correction: separates the restart's gain from the precision ladder's")
        if args.config == 1:
            compare("fp32_learned_path", PREC_F32, 1, args.keep, PRECOND_MLP, args.max_sub,
                    "FP32 values (FP32 -> FP64 refinement, no FP16 phase) with the learned correction")
            if os.path.exists(precond_train.SMALL_RECIPE_PATH):
                precond_train.load(ctx, precond_train.SMALL_RECIPE_PATH)
                compare("learned_small_recipe_network", fast_prec, fast_mixed, args.keep, PRECOND_MLP, args.max_sub,
                        "the timed path with the network trained only on the small systems "
                        "(n_orb 20-28, 4-5 electrons per spin)")
                precond_train.load(ctx)
    if args.config == 1 and n == 2_000_000 and not args.no_hci_space:
        out["hci_selected_space"] = hci_space_run(ctx, ints, failures)
    if rank == 0 and not args.no_cpu:
        log("cpu baseline (oracle, 1 core, bounded sample)...")
        full_matrix = None
        if args.config == 1 and not args.no_cpu_full:
            def full_matrix():
                from risq_amd.hci import STORE_FULL
                hf = ctx.build_h(d_dets, d_diag, n=n, storage=STORE_FULL)
                t = time.perf_counter()
                m = hf.download()
                log(f"full matrix downloaded in caller order ({time.perf_counter() - t:.1f}s)")
                hf.free()
                return m
        out["cpu_baseline"] = cpu_baseline(ints, dets, n, nnz, fp64_iters, full_matrix)
    if failures:
        out["error"] = "; ".join(failures)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
