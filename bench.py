"""Benchmark: the AKM fit at 5,000,000 workers x 8 years (40M rows, 400,000
firms, move probability 0.10) on one MI355X, inputs resident in HBM.

One step is one whole fit of `y ~ x1 + x2 | worker_id + firm_id + year` on
the device: pass 0 (factorize, singletons, row sort), pass 1 (cells), the
explicit S and right-hand side, the block PCG, normalisation, V' M_D V, beta
and pass 2.  `value` is the wall time of the mixed-precision fit (FP32 CG inside
FP64 refinement to the same 1e-10 relative residual); the FP64 fit is timed
beside it, and the coefficients of the two are compared.

Printed: one JSON line.  `roofline` covers the dominant solve kernel (the
SpMM), timed with HIP events on the stream it runs on; `cpu_baseline` is the
CPU oracle (a C/numpy restatement of the reference's FP64 algorithm) timed on a
bounded sample on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FML = "y ~ x1 + x2 | worker_id + firm_id + year"
# The FP64 SpMM of the CG at k = 3 on the 400,008-level S: P (9.6 MB) outgrows
# an XCD's L2, so it runs column-blocked as two launches, timed together.
SPMM_KERNEL = "k_spmm_tiles<double,3,2> + k_fin_small<double,3>"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0     # MI355X dense bf16 matrix peak


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


COLUMNS = ("worker_id", "firm_id", "year", "x1", "x2", "y")


def make_inputs(n_workers, n_firms, seed):
    """The panel's columns in host memory (numpy) and in HBM (torch)."""
    import torch
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(n_workers, n_firms, seed=seed)
    host = {k: np.ascontiguousarray(d[k]) for k in COLUMNS}
    cols = {k: torch.from_numpy(host[k]).cuda() for k in COLUMNS}
    torch.cuda.synchronize()
    return host, cols


def spmm_bytes(nnz, L, k, vbytes):
    """Algorithmic bytes of one SpMM launch with the fused pAp dot: the CSR
    (values + int32 column indices + int32 row pointers), P read once, AP
    written once."""
    return nnz * (vbytes + 4) + (L + 1) * 4 + 2 * L * k * vbytes


def host_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(host, n_workers, n_firms, sample_workers, sample_firms, seed, threads=16):
    """The oracle's FP64 fit (its C/OpenMP S build and PCG, numpy passes) of
    the whole panel on `threads` host threads, unscaled; then on one thread,
    of the same generator's panel with `sample_workers` workers (the whole
    panel on one thread would take minutes), unscaled and labelled so."""
    from hdfe_stream_amd.simulate import akm_panel
    from oracle import oracle as O
    O.build()
    nproc = host_cpus()
    threads = max(1, min(threads, nproc))
    O.set_threads(threads)
    t = time.time()
    O.fit(host)
    dt = time.time() - t
    d = akm_panel(sample_workers, sample_firms, seed=seed)
    O.set_threads(1)
    t = time.time()
    O.fit(d)
    d1 = time.time() - t
    O.set_threads(threads)
    return {"value": round(dt, 2), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"the whole {n_workers:,}-worker x 8-year panel ({n_firms:,} firms), "
                      f"oracle FP64 fit on {threads} of the host's {nproc} CPUs, unscaled",
            "one_thread": {"value": round(d1, 2), "unit": "s", "cores": 1,
                           "sample": f"{sample_workers:,} workers x 8 years ({sample_firms:,} "
                                     "firms), same generator, unscaled"}}


def host_ingest_phase(host, fit_kwargs, steps):
    """The same fit from the columns in ordinary host memory (numpy): every
    column streamed to HBM in 1M-row batches through the page-locked ring,
    then the fit.  PCIe-inclusive, so not `value`: reported beside it."""
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    feols_stream(FML, host, vcov="CRV1:worker_id", **fit_kwargs)          # warm
    times, ingest = [], None
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = feols_stream(FML, host, vcov="CRV1:worker_id", **fit_kwargs)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        ingest = r.diagnostics["ingest"]
        del r
    return {"fit_from_host_s": round(float(np.mean(times)), 4),
            "ingest_s": ingest["seconds"], "ingest_GBps": round(ingest["GBps"], 1),
            "ingest_bytes": ingest["bytes"], "batch_rows": 1_000_000,
            "pack_s": ingest["pack_seconds"],
            "note": "host arrays (pageable numpy) -> page-locked ring -> HBM, then the fit; "
                    "48 B per row over PCIe"}


def jla_phase(n_workers, n_firms, draws):
    """JLA leverages on the KSS panel (observation level: worker, firm and
    year effects): `draws` Rademacher right-hand sides solved as one block in
    mixed precision, beside FP64 single right-hand-side solves."""
    import torch
    from hdfe_stream_amd.estimator import Panel
    from hdfe_stream_amd.leaveout import jla_leverages
    from hdfe_stream_amd.simulate import kss_panel
    d = kss_panel(n_workers, n_firms, seed=3)
    corr = float(np.corrcoef(d["alpha"], d["psi"])[0, 1])
    cols = {k: torch.from_numpy(d[k]).cuda() for k in ("worker_id", "firm_id", "year", "y")}
    P = Panel([cols["worker_id"], cols["firm_id"], cols["year"]], [cols["y"]], sort_vars=True)
    P.system(with_rhs=False)
    out = {"workload": f"{n_workers:,} workers x 8 years, {n_firms:,} firms, p_move 0.15, "
                       f"corr(alpha, psi) {corr:.3f}, observation level, {draws} draws",
           "levels": P.info["L"], "rows": P.info["n"], "S_nnz": P.info["nnz"]}
    jla_leverages(P, n_draws=draws, block=draws, seed=0, precision="mixed")     # warm
    tm = {}
    t0 = time.perf_counter()
    lev = jla_leverages(P, n_draws=draws, block=draws, seed=0, precision="mixed", timings=tm)
    torch.cuda.synchronize()
    out.update(wall_s=round(time.perf_counter() - t0, 4),
               solve_s=round(tm["solve_seconds"], 4), row_pass_s=round(tm["row_pass_seconds"], 4),
               rhs_solves_per_s=round(tm["rhs_solves_per_second"], 1),
               iterations=lev.diagnostics["iterations"], mean_leverage=lev.diagnostics["mean_leverage"])
    tm64 = {}
    lev64 = jla_leverages(P, n_draws=draws, block=draws, seed=0, precision="fp64", timings=tm64)
    out["fp64_block_rhs_solves_per_s"] = round(tm64["rhs_solves_per_second"], 1)
    out["leverage_max_abs_diff_mixed_vs_fp64"] = float(
        (lev.leverage - lev64.leverage).abs().max().item())
    ts = {}
    jla_leverages(P, n_draws=8, block=1, seed=0, precision="fp64", timings=ts)
    out["fp64_single_rhs_solves_per_s"] = round(ts["rhs_solves_per_second"], 1)
    out["speedup_mixed_block_vs_fp64_single"] = round(
        tm["rhs_solves_per_second"] / ts["rhs_solves_per_second"], 2)
    P.close()
    return out


def kss_phase(n_workers, n_firms, draws):
    """KSS end to end on the KSS panel, leaving out a match (the reference's
    default) and an observation, then at match level with standard errors:
    leave-one-out pruning of (worker, firm), the
    fit of y ~ 1 | worker_id + firm_id + year with workers streamed (at match
    level then the collapse and the weighted match fit), `draws` JLA draws
    and `draws` Hutchinson draws (two solves each), all blocks of `draws`
    right-hand sides in mixed precision."""
    import torch
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    d = kss_panel(n_workers, n_firms, seed=3)
    cols = {k: torch.from_numpy(d[k]).cuda() for k in ("worker_id", "firm_id", "year", "y")}
    torch.cuda.synchronize()
    out = {"workload": f"KSS leave-out y ~ 1 | worker_id + firm_id + year, "
                       f"{n_workers:,} workers x 8 years, {n_firms:,} firms, p_move 0.15, "
                       f"{draws} JLA + {draws} Hutchinson draws, mixed precision"}
    # one untimed run first, as the fits have their warm-up: the first call in
    # the process also pays its host allocations (page faults of fresh arrays)
    leave_out_kss("y ~ 1 | worker_id + firm_id + year", cols, n_draws=draws, seed=0,
                  block=draws, precision="mixed", leave_out="match")
    for level in ("match", "observation"):
        tm = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lo = leave_out_kss("y ~ 1 | worker_id + firm_id + year", cols, n_draws=draws, seed=0,
                           block=draws, precision="mixed", timings=tm, leave_out=level)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[level] = {
            "wall_s": round(wall, 4), "person_years": lo.n_obs, "movers": lo.n_movers,
            "matches": lo.diagnostics.get("n_matches"), "pruning": lo.pruning,
            "seconds": {k: round(v, 4) for k, v in tm.items() if k.endswith("seconds")},
            "trace_rhs_solves_per_s": round(2 * draws / tm["trace_seconds"], 1),
            "plug_in": lo.plug_in, "bias": lo.bias, "leave_out": lo.leave_out,
            "max_leverage": lo.diagnostics["max_leverage"]}
    # with standard errors (the reference's default match level): B_ii by
    # `draws` projections and the trace term by `draws` draws, blocks of 64
    tm = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id + year", cols, n_draws=draws, seed=0,
                       block=64, precision="mixed", timings=tm, se=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out["match_se"] = {
        "wall_s": round(wall, 4), "se_seconds": round(tm["se_seconds"], 4),
        "se_draws": draws, "se": lo.se.se, "se_conservative": lo.se.se_conservative,
        "trace_term": lo.se.trace_term, "theta": lo.se.theta, "leave_out": lo.leave_out}
    return out


def adjoint_phase(n_workers, n_firms, fd_rows):
    """The differentiable weighted fit: y ~ x1 + x2 | worker_id + firm_id +
    year on the weighted panel (w ~ U(0.5, 2) float32), forward (the fit and
    the plug-in components) and backward (one adjoint solve) for the loss
    beta_1 + var(psi), timed; then central FP64 finite differences of the
    device fit itself on `fd_rows` random rows, in w and in y."""
    import torch
    from hdfe_stream_amd.autograd import fe_components
    from hdfe_stream_amd.simulate import weighted_panel
    d = weighted_panel(n_workers, n_firms, seed=4)
    cols = {k: torch.from_numpy(d[k]).cuda() for k in ("worker_id", "firm_id", "year", "x1", "x2")}
    y = torch.from_numpy(d["y"]).cuda().requires_grad_(True)
    w = torch.from_numpy(d["w"]).cuda().requires_grad_(True)
    fml = "y ~ x1 + x2 | worker_id + firm_id + year"
    g = torch.tensor([1.0, 0.0, 1.0, 0.0, 0.0], dtype=torch.float64, device="cuda")
    out = fe_components(fml, cols, y=y, weights=w, tol=1e-13)      # warm
    (out * g).sum().backward()
    y.grad = w.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fe_components(fml, cols, y=y, weights=w, tol=1e-13)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    (out * g).sum().backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    rng = np.random.default_rng(11)
    rows = rng.choice(len(d["y"]), fd_rows, replace=False)
    h = 0.05
    pairs = {"w": [], "y": []}
    for i in rows:
        for kind in ("w", "y"):
            vals = []
            for sgn in (1.0, -1.0):
                yy, ww = y.detach(), w.detach().double()
                if kind == "w":
                    ww = ww.clone()
                    ww[i] += sgn * h
                else:
                    yy = yy.clone()
                    yy[i] += sgn * h
                vals.append(float((fe_components(fml, cols, y=yy, weights=ww, tol=1e-13) * g).sum()))
            num = (vals[0] - vals[1]) / (2 * h)
            an = float((w.grad if kind == "w" else y.grad)[i])
            pairs[kind].append((num, an))

    def err(kind):
        # the largest FD-vs-adjoint gap relative to the gradients' RMS over
        # the checked rows (a relative error per row is meaningless where a
        # row's gradient is near zero)
        a = np.array(pairs[kind])
        rms = float(np.sqrt(np.mean(a[:, 1] ** 2)))
        return float(np.max(np.abs(a[:, 0] - a[:, 1])) / rms), rms
    ew, rw = err("w")
    ey, ry = err("y")
    return {"workload": f"weighted AKM y ~ x1 + x2 | worker_id + firm_id + year, "
                        f"{n_workers:,} workers x 8 years, {n_firms:,} firms, w ~ U(0.5, 2) f32; "
                        "loss beta_x1 + var(psi)",
            "forward_s": round(t1 - t0, 4), "backward_s": round(t2 - t1, 4),
            "outputs": [float(v) for v in out.detach().cpu()],
            "fd_rows": int(fd_rows), "fd_step": h,
            "fd_max_err_over_rms_w": ew, "grad_rms_w": rw,
            "fd_max_err_over_rms_y": ey, "grad_rms_y": ry}


def glm_phase(n_models, lo, hi, n_cov):
    """Uneven Poisson models with three fixed effects and `n_cov` covariates,
    all fitted as one batch (fepois_batch) with its default solves (FP64
    through the matrix-free S) timed, mixed-precision solves beside them for
    the time and the slopes' agreement.  The split-BF16 SYRK
    of the covariate Hessian is the MFMA-bound kernel: its algorithmic flops
    per IRLS step are sum_i 2 n_i c^2 (c = n_cov + 1, as the reference's
    V'V), against the BF16 dense peak over three (the split's three
    products)."""
    import torch
    from hdfe_stream_amd.glm import fepois_batch
    from hdfe_stream_amd.simulate import poisson_models_device
    ds = poisson_models_device(n_models, lo, hi, n_cov, seed=5)
    rows = int(sum(len(d["y"]) for d in ds))
    fml = "y ~ X | worker_id + firm_id + year"
    # an untimed batch of two models first: the kernels' code objects load
    # and the library's caches fill (a process's first batch pays ~0.3 s)
    fepois_batch(fml, ds[:2])
    tm = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = fepois_batch(fml, ds, timings=tm)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tmm = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    resm = fepois_batch(fml, ds, precision="mixed", timings=tmm)
    torch.cuda.synchronize()
    wall_mixed = time.perf_counter() - t0
    err = max(float(np.max(np.abs(a.beta - b.beta)) / np.max(np.abs(b.beta)))
              for a, b in zip(resm, res))
    c = n_cov + 1
    flops_step = 2.0 * rows * c * c
    steps = tm["steps"]
    syrk_tflops = flops_step * steps / tm["syrk"] / 1e12 if tm["syrk"] > 0 else float("nan")
    peak3 = BF16_PEAK_TFLOPS / 3.0
    return {"workload": f"{n_models} Poisson models, rows log-uniform in [{lo:,}, {hi:,}] "
                        f"({rows:,} rows), worker/firm/year fixed effects, {n_cov} float32 "
                        "covariates; fepois_batch, default (FP64) solves",
            "wall_s": round(wall, 4), "timings_s": tm, "precision": "fp64",
            "mixed_wall_s": round(wall_mixed, 4), "mixed_timings_s": tmm,
            "irls_steps": steps, "converged": all(r.converged for r in res),
            "max_iterations": int(max(r.iterations for r in res)),
            "beta_rel_err_mixed_vs_fp64": err,
            "syrk": {"kernel": "k_glm_syrk (split-BF16, v_mfma_f32_32x32x16_bf16 x3)",
                     "flops_per_step": flops_step, "achieved_TFLOPs": round(syrk_tflops, 1),
                     "peak_TFLOPs_bf16_over_3": round(peak3, 1),
                     "frac": round(syrk_tflops / peak3, 4)}}


def copy_calibration(nbytes=1 << 30, reps=20):
    """A plain copy kernel of known traffic: `nbytes` read and written per
    launch, timed with HIP events."""
    import ctypes as C
    from hdfe_stream_amd import _lib
    us = C.c_double()
    _lib.check(_lib.load().hdfe_copy_calibrate(nbytes, reps, C.byref(us)), "hdfe_copy_calibrate")
    return {"kernel": "k_copy_calibrate", "bytes_read": nbytes, "bytes_written": nbytes,
            "avg_launch_us": round(us.value, 2),
            "GBps": round(2 * nbytes / (us.value * 1e-6) / 1e9, 1)}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE scaled by the copy kernel's factors),
    or None when the summary is absent or for another kernel or workload."""
    p = ROOT / "profiles" / "spmm_c2_traffic.txt"
    if not p.exists():
        return None
    vals = {}
    for line in p.read_text().splitlines():
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    if vals.get("kernel") != kernel or vals.get("workload") != workload:
        return None
    try:
        return int(float(vals["traffic_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workers", type=int, default=5_000_000)
    ap.add_argument("--firms", type=int, default=400_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-jla", action="store_true")
    ap.add_argument("--jla-workers", type=int, default=1_000_000)
    ap.add_argument("--jla-firms", type=int, default=100_000)
    ap.add_argument("--jla-draws", type=int, default=256)
    ap.add_argument("--no-kss", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-glm", action="store_true")
    ap.add_argument("--glm-models", type=int, default=32)
    ap.add_argument("--glm-lo", type=int, default=50_000)
    ap.add_argument("--glm-hi", type=int, default=2_000_000)
    ap.add_argument("--glm-cov", type=int, default=300)
    ap.add_argument("--adjoint-workers", type=int, default=2_000_000)
    ap.add_argument("--adjoint-firms", type=int, default=200_000)
    ap.add_argument("--adjoint-fd-rows", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is described, not built)", file=sys.stderr)
        sys.exit(2)
    import torch
    from hdfe_stream_amd import _lib
    from hdfe_stream_amd.estimator import feols_stream
    torch.cuda.set_device(0)
    name = _lib.require_device()
    log(f"device {name}; generating {args.workers:,} workers x 8 years")
    t = time.time()
    host, data = make_inputs(args.workers, args.firms, seed=2)
    log(f"inputs in HBM ({time.time() - t:.1f} s)")

    def fit(precision):
        # the reference's benchmark job: the fit with CRV1 standard errors by
        # worker (benchmarks/results/akm.csv)
        return feols_stream(FML, data, precision=precision, tol=1e-10, vcov="CRV1:worker_id")

    # each precision's last result is kept (for the comparison below) and
    # released before that precision's next fit, in the warm-up as in the
    # timed steps, so every timed fit finds the same device memory free
    results = {}
    for w in range(args.warmup):
        for prec in ("fp64", "mixed"):
            results.pop(prec, None)
            results[prec] = fit(prec)
        log(f"warmup {w + 1}/{args.warmup} done")
    times = {"fp64": [], "mixed": []}
    for s in range(args.steps):
        for prec in ("fp64", "mixed"):
            results.pop(prec, None)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fit(prec)
            torch.cuda.synchronize()
            times[prec].append(time.perf_counter() - t0)
            results[prec] = r
        log(f"step {s + 1}/{args.steps}: fp64 {times['fp64'][-1]:.3f} s, "
            f"mixed {times['mixed'][-1]:.3f} s")
    f64, mix = results["fp64"], results["mixed"]
    beta_err = float(np.max(np.abs(mix.beta - f64.beta)) / np.max(np.abs(f64.beta)))
    gy64 = np.concatenate([f64.fe_values[d] for d in f64.fe_names[1:]])
    gymx = np.concatenate([mix.fe_values[d] for d in mix.fe_names[1:]])
    fe_err = float(np.max(np.abs(gymx - gy64)) / np.max(np.abs(gy64)))
    t_mixed = float(np.mean(times["mixed"]))
    t_f64 = float(np.mean(times["fp64"]))
    # roofline of the SpMM: the same kernel and launch as in the FP64 fit's
    # CG, timed with HIP events around a burst of launches on the solver's
    # stream, on the last fit's matrix (still resident)
    L = f64.panel.info["L"]
    nnz = f64.panel.info["nnz"]
    k = 3
    info = f64.solver_info
    M = f64.panel.matrix
    # the other forms of the column-blocked product first, for comparison,
    # then the default (the last launches of its kernels, whose counters the
    # committed traffic summary averages)
    grouped, tiled32 = {}, {}
    for prec in ("fp64", "mixed"):
        M.small_form(k, prec, set=1)
        grouped[prec] = round(M.spmm_time(k, reps=100, precision=prec), 2)
        M.small_form(k, prec, set=3)
        tiled32[prec] = round(M.spmm_time(k, reps=100, precision=prec), 2)
        M.small_form(k, prec, set=0)
    avg32 = M.spmm_time(k, reps=100, precision="mixed") * 1e-6
    avg64 = M.spmm_time(k, reps=100, precision="fp64") * 1e-6
    byt = spmm_bytes(nnz, L, k, 8)
    ach = byt / avg64 / 1e9
    kname = SPMM_KERNEL
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(kname, f"c2:{args.workers}x{args.firms}"),
            "kernel": kname, "bytes_per_launch": int(byt),
            "avg_launch_us": round(avg64 * 1e6, 2),
            "grouped_form_us": grouped["fp64"], "tiled_32bit_index_us": tiled32["fp64"]}
    b32 = spmm_bytes(nnz, L, k, 4)
    line = {
        "metric": "AKM fit wall-s @5M workers (mixed FP32/FP64-refined, 1e-10 rel. residual)",
        "value": round(t_mixed, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_mixed * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"AKM OLS y ~ x1 + x2 | worker_id + firm_id + year, "
                               f"{args.workers:,} workers x 8 years, {args.firms:,} firms, "
                               "p_move 0.10, seed 2; CRV1 standard errors by worker; inputs "
                               "resident in HBM (the fit from host memory, PCIe included, is "
                               "from_host)",
                   "rows": int(f64.panel.info["n_in"]), "levels": L, "S_nnz": nnz},
        "fp64_fit_s": round(t_f64, 4),
        "fit_breakdown_fp64": f64.diagnostics,
        "fit_breakdown_mixed": mix.diagnostics,
        "iterations_fp64": info.get("iterations"),
        "iterations_mixed": mix.solver_info.get("iterations"),
        "refinement_steps": mix.solver_info.get("refinement_steps"),
        "beta_rel_err_mixed_vs_fp64": beta_err,
        "se_crv1_worker": [float(v) for v in np.sqrt(np.diag(mix.vcov_matrix))],
        "fe_rel_err_mixed_vs_fp64": fe_err,
        "roofline": roof,
        "spmm_fp32": {"kernel": "k_spmm_tiles<float,3,1> + k_fin_small<float,3>",
                      "avg_launch_us": round(avg32 * 1e6, 2),
                      "bytes_per_launch": int(b32), "achieved_GBps": round(b32 / avg32 / 1e9, 1),
                      "grouped_form_us": grouped["mixed"],
                      "tiled_32bit_index_us": tiled32["mixed"]},
        "copy_calibration": copy_calibration(),
    }
    if not args.no_host:
        log("fit from host memory")
        line["from_host"] = host_ingest_phase(host, {"precision": "mixed", "tol": 1e-10},
                                              max(1, min(args.steps, 5)))
    if not (args.no_jla and args.no_kss and args.no_adjoint and args.no_glm):
        del results, f64, mix
    if not args.no_jla:
        _lib.release_memory()   # the previous phase's cached blocks back to the driver
        log("JLA phase")
        line["jla"] = jla_phase(args.jla_workers, args.jla_firms, args.jla_draws)
    if not args.no_kss:
        _lib.release_memory()   # the previous phase's cached blocks back to the driver
        log("KSS phase")
        line["kss"] = kss_phase(args.jla_workers, args.jla_firms, args.jla_draws)
    if not args.no_adjoint:
        _lib.release_memory()   # the previous phase's cached blocks back to the driver
        log("adjoint phase")
        line["adjoint"] = adjoint_phase(args.adjoint_workers, args.adjoint_firms,
                                        args.adjoint_fd_rows)
    if not args.no_glm:
        _lib.release_memory()   # the previous phase's cached blocks back to the driver
        log("GLM phase")
        line["glm"] = glm_phase(args.glm_models, args.glm_lo, args.glm_hi, args.glm_cov)
    if not args.no_cpu:
        log("CPU baseline (oracle: the whole panel, then a 1-thread sample)")
        line["cpu_baseline"] = cpu_baseline(host, args.workers, args.firms, args.workers // 8,
                                            args.firms // 8, seed=2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
