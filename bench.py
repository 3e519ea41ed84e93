#!/usr/bin/env python3
"""Benchmark of the Explicit higher-order encoding operator on one MI355X.

Headline: E^H E applications per second at 200x200, 32 coils, 16000-sample
spiral (workload 1), mixed-precision MFMA backend, with the plain FP32 direct
backend, the error of both against each other and FP64 (sampled forward),
the field-coefficient gradient time, the roofline of the dominant kernel
(HIP events in libhoe on its own stream) and the restated CPU reference path
timed on a bounded sample of the same workload.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 1]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "highordermri.jl_amd"))

METRIC = "E^H E applies/s (200^2, 32 coils): FP32 vs mixed MFMA @<=1e-4 vs FP64; grad ms"
PEAK_F16_MFMA_TFLOPS = 2500.0     # MI355X dense FP16 MFMA
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 (the direct kernel's VALU ceiling)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_op(cfg, backend):
    import hoe
    return hoe.HighOrderExplicitOp(cfg["grid"], cfg["kspha"], cfg["times"], fieldmap=cfg["fieldmap"],
                                   csm=cfg["csm"], mask=cfg["mask"], concomitant=cfg.get("concomitant"),
                                   backend=backend, dwell=cfg["dwell"])


def time_normal(op, x_dev, out_dev, steps, warmup):
    """Returns (seconds total, per-step forward-kernel ms list, adjoint-kernel ms list).

    The timed steps are queued back to back (HOE_ASYNC, device vectors), as a
    device-resident solver issues them, and the region is closed by a device
    synchronisation. The kernel times come from separate synchronous calls,
    which record libhoe's HIP events around each contraction launch."""
    import torch
    from hoe import _lib
    lib, h = op._lib, op._handle()
    _lib.check(lib.hoe_set_weights(h, None, 0), h)
    flags = _lib.HOE_DEVICE_PTRS
    run = lambda f: _lib.check(lib.hoe_normal(h, None, _lib.HOE_CURRENT_WEIGHTS, x_dev.data_ptr(),
                                              out_dev.data_ptr(), f), h)
    for _ in range(warmup):
        run(flags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run(flags | _lib.HOE_ASYNC)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    kf, ka = [], []
    for _ in range(min(steps, 10)):
        run(flags)
        kf.append(lib.hoe_last_kernel_ms(h, 0))
        ka.append(lib.hoe_last_kernel_ms(h, 1))
    return t, kf, ka


def _finite(v):
    """NaN/inf (parts skipped with --no-*) as null, so the line is strict JSON."""
    if isinstance(v, dict):
        return {k: _finite(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_finite(x) for x in v]
    if isinstance(v, float) and not np.isfinite(v):
        return None
    return v


def measured_traffic(workload):
    """HBM bytes per k_mixed<f16> launch (forward/adjoint average) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload (copy-calibrated), or None."""
    path = ROOT / "profiles" / f"traffic_workload{workload}.json"
    if not path.exists():
        return None
    try:
        return float(json.loads(path.read_text())["k_mixed_f16_avg_bytes"])
    except (ValueError, KeyError):
        return None


def host_cpu():
    """Cores this process may run on, the machine's logical CPU count and the CPU model."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return usable, os.cpu_count() or usable, model


def cpu_baseline(cfg, threads, budget_s=30.0, n_sub=0):
    """Restated reference CPU Array path (oracle/cpu_ref, OpenMP on `threads` cores) E^H E per
    second. The sample count is the whole readout when 1 warm-up and 3 timed runs fit in
    about `budget_s` seconds (probed on 512 samples), else the largest slice of the readout
    that does (cost is linear in samples: the per-sample work is independent), scaled to the
    full readout. Returns (seconds per full E^H E, samples timed)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    O.lib().oracle_set_threads(threads)
    ns_full = cfg["kspha"].shape[1]

    def model(n):
        conc = None if cfg.get("concomitant") is None else cfg["concomitant"][:, :n]
        return O.make_model(cfg["grid"], cfg["kspha"][:, :n], cfg["times"][:n], cfg["fieldmap"], cfg["csm"],
                            cfg["mask"], concomitant=conc)
    if not n_sub:
        probe = min(512, ns_full)
        mp = model(probe)
        O.cpu_ref_normal(mp, cfg["image"])
        t0 = time.perf_counter()
        O.cpu_ref_normal(mp, cfg["image"])
        per_sample = (time.perf_counter() - t0) / probe
        n_sub = int(min(ns_full, max(probe, budget_s / (4.0 * per_sample))))
    m = model(n_sub)
    O.cpu_ref_normal(m, cfg["image"])           # warm-up
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.cpu_ref_normal(m, cfg["image"])
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    return t * ns_full / n_sub, n_sub


def cpu_object(t_full, n_sub, ns, threads, machine, what):
    usable, logical, model = machine
    return {"value": round(1.0 / t_full, 6), "unit": "E^H E applies/s", "cores": threads, "kind": "port",
            "usable_cores": usable, "logical_cpus": logical, "cpu_model": model,
            "sample": (f"oracle cpu_ref (restated array_operator.jl:241-348, OpenMP, {threads} threads) E^H E of "
                       f"{what}: median of 3 after 1 warm-up on "
                       + ("the whole readout" if n_sub == ns else
                          f"the first {n_sub} of {ns} samples, scaled x{ns / n_sub:.1f} (extrapolated)"))}


def workload_entry(w, steps, with_cpu, with_grad, threads, machine):
    """E^H E rate of the mixed FP16 and FP32 backends on workload w, their contraction
    kernels' roofline fractions, the gradient (workload 3 convention) when asked, and the CPU
    baseline on a bounded sample."""
    import torch
    import hoe
    from hoe import synth
    cfg = synth.make(w)
    ns, nc = cfg["kspha"].shape[1], cfg["csm"].shape[1]
    dev = torch.device("cuda:0")
    x_dev = torch.from_numpy(cfg["image"]).to(dev)
    out = torch.empty(cfg["grid"].n, dtype=torch.complex64, device=dev)
    res = {"workload": synth.describe(w), "n_sam": ns, "n_cha": nc}
    k_big = w == 4
    with make_op(cfg, "mixed_fp16") as op:
        res["n_vox"] = nv = op.nVox
        flops = 8.0 * ns * nv * nc
        for be, k in (("mixed_fp16", steps), ("direct", max(2, steps // 4))):
            op.set_backend(be)
            t, kf, ka = time_normal(op, x_dev, out, k, 1 if k_big else 3)
            kf_avg, ka_avg = float(np.mean(kf)), float(np.mean(ka))
            tflops = flops / (0.5 * (kf_avg + ka_avg) * 1e-3) / 1e12
            peak = PEAK_F16_MFMA_TFLOPS if be == "mixed_fp16" else PEAK_FP32_TFLOPS
            res[be] = {"value": round(k / t, 4), "unit": "E^H E applies/s", "ms_per_step": round(1e3 * t / k, 3),
                       "steps": k, "fwd_kernel_ms": round(kf_avg, 4), "adj_kernel_ms": round(ka_avg, 4),
                       "tflops": round(tflops, 2), "frac_of_peak": round(tflops / peak, 4), "peak_tflops": peak}
            log(f"workload {w} {be}: {k / t:.3f} E^H E/s, kernels {kf_avg:.3f} / {ka_avg:.3f} ms")
        res["mixed_over_fp32"] = round(res["mixed_fp16"]["value"] / res["direct"]["value"], 3)
        if with_grad:
            kp = synth.perturb(cfg)
            with make_op(dict(cfg, kspha=kp), "mixed_fp16") as gop:
                gop.set_coefficients(None, 3e-6)
                y_obs = gop * x_dev
            torch.cuda.synchronize()
            grad = {}
            for be in ("mixed_fp16", "direct"):
                op.set_backend(be)
                gts = []
                for i in range(4):          # 1 warm-up call, then 3 timed
                    t0 = time.perf_counter()
                    g, gtau, loss = op.grad(x_dev, y_obs)
                    torch.cuda.synchronize()
                    if i:
                        gts.append(time.perf_counter() - t0)
                grad[be] = {"ms": round(1e3 * float(np.median(gts)), 3), "kernel_ms": round(op.last_kernel_ms(0), 3),
                            "dL_dtau": gtau}
                log(f"workload {w} grad {be}: {grad[be]['ms']:.1f} ms (kernel {grad[be]['kernel_ms']:.1f} ms)")
            res["grad"] = grad
    if with_cpu:
        # the 3-D volume: a bounded slice of one stack partition (its samples cost as much as any)
        t_cpu, n_sub = cpu_baseline(cfg, threads, budget_s=20.0 if k_big else 15.0)
        what = f"workload {w}" + (" (one partition's first samples, x the readout)" if k_big else "")
        res["cpu_baseline"] = cpu_object(t_cpu, n_sub, ns, threads, machine, what)
        res["gpu_over_cpu"] = round(res["mixed_fp16"]["value"] * t_cpu, 1)
        log(f"workload {w} cpu: {1.0 / t_cpu:.6f} applies/s on {threads} threads ({n_sub} samples)")
    return res


def real_data(steps):
    """The reference's 7 T demo data (in-tree fixtures written by
    tests/golden/make_7t_fixture.py), measured as demo/Recon_Benchmark.jl measures its
    cases: operator setup, the first reconstruction, then the median and minimum of 3
    reconstructions (recon_HOOp, 20 CGNR iterations, L2 1e-9, physical coils); plus the
    E^H E rate, the image's error against the FP32 backend's image and its correlation
    with the reference's own reconstruction of the same data. None when the fixtures
    are absent."""
    import torch
    import hoe
    sys.path.insert(0, str(ROOT / "tests" / "golden"))
    import fixture7t
    out = {}
    for name, stem in (("epi", "7T_2D_EPI_1p0_200_r4"), ("spiral", "7T_2D_Spiral_1p0_200_r4")):
        p = fixture7t.load(name, "measured")
        if p is None:
            continue
        n = p["grid"].nX
        dev = torch.device("cuda:0")
        x = torch.zeros(p["grid"].n, dtype=torch.complex64, device=dev)
        x[torch.from_numpy(p["mask"]).to(dev)] = 1.0
        y = torch.empty_like(x)
        rp = {"reconSize": (n, p["grid"].nY), "reg": hoe.L2Regularization(1e-9), "iterations": 20,
              "solver": hoe.CGNR}
        res = {"dataset": f"{stem} (reference demo data), measured field, recon terms 0111",
               "n_sam": int(p["kspha"].shape[1]), "n_vox": int(p["mask"].sum()), "n_cha": int(p["csm"].shape[1])}
        imgs = {}
        for backend in ("direct", "mixed_fp16"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            op = hoe.HighOrderExplicitOp(p["grid"], p["kspha"], p["times"], fieldmap=p["fieldmap"], csm=p["csm"],
                                         mask=p["mask"], recon_terms=p["recon_terms"], backend=backend)
            torch.cuda.synchronize()
            setup = time.perf_counter() - t0
            try:
                t0 = time.perf_counter()
                img = np.asarray(hoe.recon_HOOp(op, p["data"], p["weight"], rp))
                first = time.perf_counter() - t0
                ts = []
                for _ in range(3):
                    t0 = time.perf_counter()
                    img = np.asarray(hoe.recon_HOOp(op, p["data"], p["weight"], rp))
                    ts.append(time.perf_counter() - t0)
                k = steps if backend == "mixed_fp16" else max(3, steps // 5)
                t, _, _ = time_normal(op, x, y, k, 2)
            finally:
                op.close()
            imgs[backend] = img
            res[backend] = {"setup_s": round(setup, 4), "first_recon_s": round(first, 4),
                            "median_recon_s": round(float(np.median(ts)), 4), "min_recon_s": round(min(ts), 4),
                            "normal_applies_per_s": round(k / t, 3),
                            "corr_with_reference_image": round(fixture7t.corr(fixture7t.rendered(img, n),
                                                                              p["golden"]), 5)}
        res["image_mixed_vs_fp32"] = float(np.linalg.norm(imgs["mixed_fp16"] - imgs["direct"])
                                           / np.linalg.norm(imgs["direct"]))
        out[name] = res
    return out or None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-direct", action="store_true")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=0)
    ap.add_argument("--no-realdata", action="store_true")
    ap.add_argument("--no-workloads", action="store_true", help="skip workloads 0, 2 and 4")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import hoe
    from hoe import synth
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    assert torch.cuda.is_available(), "no GPU visible"
    cfg = synth.make(args.workload)
    ns, nc = cfg["kspha"].shape[1], cfg["csm"].shape[1]
    n_grid = cfg["grid"].n
    log(f"workload {args.workload}: grid {cfg['grid'].matrixSize}, samples {ns}, coils {nc}")
    dev = torch.device("cuda:0")
    x_dev = torch.from_numpy(cfg["image"]).to(dev)
    out_mixed = torch.empty(n_grid, dtype=torch.complex64, device=dev)
    out_direct = torch.empty(n_grid, dtype=torch.complex64, device=dev)

    op = make_op(cfg, "mixed_fp16")
    nv, L = op.nVox, op.nTerm
    # ---- mixed (headline) ----
    t_mixed, kf_m, ka_m = time_normal(op, x_dev, out_mixed, args.steps, args.warmup)
    log(f"mixed: {args.steps} E^H E in {t_mixed:.3f} s")
    # ---- plain FP32 direct ----
    steps_d = max(3, args.steps // 5)
    if args.no_direct:
        t_direct, kf_d, ka_d = float("nan"), [float("nan")], [float("nan")]
    else:
        op.set_backend("direct")
        t_direct, kf_d, ka_d = time_normal(op, x_dev, out_direct, steps_d, max(1, args.warmup // 2))
        log(f"direct: {steps_d} E^H E in {t_direct:.3f} s")
    # ---- accuracy: mixed vs direct (full normal), forward vs FP64 on a sample subset ----
    torch.cuda.synchronize()
    nm = out_mixed.cpu().numpy()
    nd = out_direct.cpu().numpy()
    err_mixed_vs_direct = None if args.no_direct else O.relerr(nm, nd)
    n_chk = 256
    sel = np.linspace(0, ns - 1, n_chk).astype(int)
    sub_k, sub_t = cfg["kspha"][:, sel], cfg["times"][sel]
    msub = O.make_model(cfg["grid"], sub_k, sub_t, cfg["fieldmap"], cfg["csm"], cfg["mask"])
    y64 = O.forward_f64(msub, cfg["image"])
    fwd_err = {}
    for be in ("mixed_fp16", "direct"):
        with hoe.HighOrderExplicitOp(cfg["grid"], sub_k, sub_t, fieldmap=cfg["fieldmap"], csm=cfg["csm"],
                                     backend=be) as so:
            fwd_err[be] = O.relerr(so * cfg["image"], y64)
    log(f"errors: normal mixed vs direct {err_mixed_vs_direct}, fwd vs fp64 {fwd_err}")
    # ---- 30-iteration CGNR (recon_HOOp path, device resident) on noiseless synthetic data ----
    cg = {}
    if not args.no_cg:
        from hoe import _lib as L_
        op.set_backend("mixed_fp16")
        y_syn = op.prod(x_dev)
        torch.cuda.synchronize()
        imgs = {}
        for be in ("mixed_fp16", "direct"):
            op.set_backend(be)
            out = torch.empty(n_grid, dtype=torch.complex64, device=dev)
            log_ = np.zeros(31)
            D = np.ones(ns, np.float32)
            ts = []
            for rep in range(2):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                L_.check(op._lib.hoe_cg(op._handle(), y_syn.data_ptr(), D.ctypes.data, D.size, 30, 0.0, 0.0,
                                        out.data_ptr(), log_.ctypes.data, L_.HOE_DEVICE_PTRS), op._handle())
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            imgs[be] = out.cpu().numpy()
            cg[be] = {"ms": round(1e3 * min(ts), 3), "final_rel_residual": float(log_[-1])}
        cg["image_mixed_vs_fp32"] = O.relerr(imgs["mixed_fp16"], imgs["direct"])
        cg["image_vs_truth_fp32"] = O.relerr(imgs["direct"], cfg["image"])
        log(f"cg: {cg}")
    # ---- gradient (workload 3 convention: y from perturbed coefficients and a 3 us delay) ----
    grad_ms = grad_kernel_ms = float("nan")
    grad = {}
    if not args.no_grad:
        kp = synth.perturb(cfg)
        gop = make_op(dict(cfg, kspha=kp), "direct")
        gop.set_coefficients(None, 3e-6)
        y_obs = gop * x_dev
        gop.close()
        torch.cuda.synchronize()
        gres = {}
        for be in ("mixed_fp16", "direct"):
            op.set_backend(be)
            gts = []
            for i in range(13):          # 3 warm-up calls, then 10 timed
                t0 = time.perf_counter()
                g, gtau, loss = op.grad(x_dev, y_obs)      # device vectors: dL/dh stays on the GPU
                torch.cuda.synchronize()
                if i >= 3:
                    gts.append(time.perf_counter() - t0)
            gres[be] = g.cpu().numpy()
            grad[be] = {"ms": round(1e3 * float(np.median(gts)), 3), "kernel_ms": round(op.last_kernel_ms(0), 3),
                        "loss": loss, "dL_dtau": gtau}
            log(f"grad {be}: {grad[be]['ms']:.2f} ms (kernel {grad[be]['kernel_ms']:.2f} ms), loss {loss:.6e}, "
                f"dL/dtau {gtau:.6e}")
        grad["mixed_vs_fp32"] = O.relerr(gres["mixed_fp16"], gres["direct"])
        grad_ms, grad_kernel_ms = grad["mixed_fp16"]["ms"], grad["mixed_fp16"]["kernel_ms"]
        log(f"grad mixed vs fp32: {grad['mixed_vs_fp32']:.2e}")
    op.close()
    realdata = None if args.no_realdata else real_data(args.steps)
    if realdata:
        log(f"7 T demo data: {realdata}")

    value = args.steps / t_mixed
    fp32_value = steps_d / t_direct
    kf_m_avg = float(np.mean(kf_m))
    ka_m_avg = float(np.mean(ka_m))
    flops_launch = 8.0 * ns * nv * nc          # complex contraction of one product (forward or adjoint)
    achieved = flops_launch / (0.5 * (kf_m_avg + ka_m_avg) * 1e-3) / 1e12
    roofline = {
        "bound": "mfma", "kernel": "k_mixed<f16> (forward and adjoint contraction, avg)",
        "achieved": round(achieved, 2), "peak": PEAK_F16_MFMA_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_F16_MFMA_TFLOPS, 4), "traffic": measured_traffic(args.workload),
        "flops_per_launch": flops_launch, "fwd_kernel_ms": round(kf_m_avg, 4), "adj_kernel_ms": round(ka_m_avg, 4),
    }
    kf_d_avg, ka_d_avg = float(np.mean(kf_d)), float(np.mean(ka_d))
    direct_tf = flops_launch / (0.5 * (kf_d_avg + ka_d_avg) * 1e-3) / 1e12
    machine = host_cpu()
    threads = machine[0]             # every core this process may use
    cpu = None
    if not args.no_cpu:
        t_cpu, n_sub = cpu_baseline(cfg, threads, budget_s=30.0, n_sub=args.cpu_samples)
        cpu = cpu_object(t_cpu, n_sub, ns, threads, machine, f"workload {args.workload} ({nv} voxels x {nc} coils)")
        log(f"cpu baseline: {1.0 / t_cpu:.4f} applies/s on {threads} threads ({n_sub} samples)")
    workloads = None
    if not args.no_workloads:
        workloads = {}
        for w in (0, 2, 4):
            if w == args.workload:
                continue
            workloads[str(w)] = workload_entry(w, 3 if w == 4 else max(10, args.steps // 2), not args.no_cpu,
                                               w == 4 and not args.no_grad, threads, machine)
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "E^H E applies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * t_mixed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16 hi/lo MFMA (f32 acc), f64 phase, f32 sincos",
        "data": "synthetic",
        "config": {"workload": synth.describe(args.workload), "n_vox": nv,
                   "n_sam": ns, "n_cha": nc, "n_term": L, "weights": "unit"},
        "fp32_direct": {"value": round(fp32_value, 3), "unit": "E^H E applies/s",
                        "ms_per_step": round(1e3 * t_direct / steps_d, 4), "steps": steps_d,
                        "tflops_fp32": round(direct_tf, 2),
                        "frac_fp32_peak": round(direct_tf / PEAK_FP32_TFLOPS, 4)},
        "mixed_over_fp32": round(value / fp32_value, 3),
        "errors": {"normal_mixed_vs_fp32": err_mixed_vs_direct,
                   "forward_vs_fp64_sampled": {k: float(v) for k, v in fwd_err.items()}},
        "grad_ms": round(grad_ms, 3), "grad_kernel_ms": round(grad_kernel_ms, 3), "grad": grad,
        "cg30": cg,
        "realdata": realdata,
        "workloads": workloads,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(_finite(line)), flush=True)


if __name__ == "__main__":
    main()
