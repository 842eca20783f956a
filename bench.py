#!/usr/bin/env python3
"""Benchmark of the anchor-selection hot path on one MI355X.

Workload (default ``--config h2o64``): seeded (H2O)_64 cluster in an 18 A box,
cc-pVDZ (1536 spherical AOs), level-3 Becke grid (~2.2e6 points), randomized
quadrature Khatri-Rao selection (rand_sketch) of k = 12,000 anchors with
ell = k + 10% sketch rows, seed 7.  A step is one full selection -- AO
evaluation on grid batches, sketch GEMMs, GPU QRCP -- followed by the FP64
fit on the chosen anchors (anchor-pair Gram, Cholesky, theta for every kept
grid point, pair-space error), with the grid coordinates and weights already
resident in HBM.

Printed: one JSON line (metric, grid-pts/s value, per-stage seconds, the
all-FP64 form for comparison, roofline of the dominant kernel measured with
HIP events, and the CPU baseline of the oracle on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

CONFIGS = {
    # name: (waters, basis, grid level, anchors)
    "h2o8": (8, "cc-pvdz", 3, 1536),
    "h2o32": (32, "cc-pvdz", 3, 6144),
    "h2o64": (64, "cc-pvdz", 3, 12000),
}
PEAK = {"fp64_mfma_tflops": 78.6, "bf16_mfma_tflops": 2500.0, "fp8_mfma_tflops": 5000.0, "int8_mfma_tops": 5000.0,
        "hbm_gbs": 8000.0}
# The INT8 fold splits q into 5 and y into 3 signed base-128 digits and runs
# the 14 digit pairs with p + r <= 5 as INT8 products.
I8_PAIRS = 14
FIT_BATCH = 32768  # grid points per fit batch (measured: 11.8 s fit against 12.1 s at 8192)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--config", default="h2o64", choices=sorted(CONFIGS))
    p.add_argument("--precision", default="bf16", choices=["bf16", "fp8", "fp64"])
    p.add_argument("--no-fp64-form", action="store_true", help="skip the all-FP64 comparison run")
    p.add_argument("--cpu-points", type=int, default=12288, help="grid points in the CPU baseline sample (~10-15 s of work)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--metric-points", type=int, default=262144,
                   help="grid points of the Coulomb-metric leg (0 skips it)")
    p.add_argument("--metric-vectors", type=int, default=1536, help="interpolation vectors of the metric leg")
    p.add_argument("--metric-cpu-points", type=int, default=4096, help="grid points of the metric's CPU sample")
    p.add_argument("--copy-probe-mib", type=int, default=0,
                   help="finish with three device-to-device copies of this many MiB (a known-size kernel for "
                        "calibrating HBM byte counters under rocprofv3 --pmc)")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              f"(the selection path shards over grid points across GPUs by design, but no multi-GPU "
              f"run is part of it)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import mp_isdf_amd
    from mp_isdf_amd import capi, grid as gridmod, molecule

    mp = mp_isdf_amd.mp_isdf
    if not mp.device_enabled:
        raise SystemExit("no HIP device: the benchmark runs on the GPU only")
    nwater, basis, level, k = CONFIGS[args.config]
    oversampling = (k + 9) // 10
    ell = k + oversampling
    mol = molecule.water_cluster(nwater, basis)
    t0 = time.perf_counter()
    xyz, w = gridmod.make_grid(mol, level)
    log(f"grid {len(w)} points in {time.perf_counter() - t0:.1f} s; nao {mol.nao}; k {k}; ell {ell}")
    keep = w != 0
    n = int(keep.sum())
    ctx = mp.LibcintContext(mol.atm, mol.bas, mol.env)
    sel = mp.AnchorSelection(ctx)
    alg = mp.LocalAnchorSelectionAlgorithm.RandomizedQuadratureKhatriRao
    ao = mp.AxisType.AtomicOrbital

    cctx = capi.Context(mol.atm, mol.bas, mol.env)
    xyz_dev = torch.tensor(xyz[keep], device="cuda")
    w_dev = torch.tensor(w[keep], device="cuda")

    def fit(anchors):
        """ISDF fit on the kept grid: Gram, Cholesky and L^-1 of the anchor pairs,
        then theta for every kept grid point (formed batch by batch on the GPU
        and consumed there as theta . w) and the pair-space error."""
        ax = torch.tensor(xyz[anchors], device="cuda")
        aw = np.ascontiguousarray(w[anchors])
        res = np.zeros(8)
        tsum = np.zeros(len(anchors))
        capi.call("mpisdf_isdf_fit_device", cctx.handle, ax, aw, len(anchors), xyz_dev, w_dev, n, FIT_BATCH, tsum, res,
                  None)
        return res

    def step(precision):
        t = time.perf_counter()
        anchors, d = sel.select_with_diagnostics(xyz, w, ao, algorithm=alg, max_rank=k, oversampling=oversampling,
                                                 seed=7, precision=precision)
        t_sel = time.perf_counter() - t
        res = fit(anchors)
        d = dict(d)
        d["seconds_select"] = t_sel
        d["seconds_fit_factor"] = res[4]
        d["seconds_fit_grid"] = res[5]
        d["fit_gemm_seconds"] = res[6]
        d["fit_gemm_flops"] = res[7]
        d["pair_space_relative_error"] = float(res[1])
        d["gram_condition_estimate"] = float(res[2])
        return anchors, d

    def describe(i, label, t_total, d):
        log(f"{label} {i}: {t_total:.2f} s (sketch {d['seconds_sketch']:.2f}, qrcp {d['seconds_qrcp']:.2f}, "
            f"blocks {int(d['blocks'])}, fit factor {d['seconds_fit_factor']:.2f}, fit grid "
            f"{d['seconds_fit_grid']:.2f}; pair-space error {d['pair_space_relative_error']:.4e})")

    for i in range(args.warmup):
        t = time.perf_counter()
        _, d = step(args.precision)
        describe(i, "warmup", time.perf_counter() - t, d)
    torch.cuda.synchronize()
    times, diags, anchors = [], [], None
    for i in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        anchors, d = step(args.precision)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        diags.append(d)
        describe(i, "step", times[-1], d)
    t_step = max(times) if times else float("nan")
    d = diags[-1] if diags else {}

    fp64_form = None
    if not args.no_fp64_form and args.precision != "fp64":
        torch.cuda.synchronize()
        t = time.perf_counter()
        ref, d64 = step("fp64")
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t
        describe(0, "fp64 form", t64, d64)
        overlap = len(set(ref) & set(anchors)) / max(1, len(ref))
        e, e64 = d["pair_space_relative_error"], d64["pair_space_relative_error"]
        fp64_form = {"seconds": round(t64, 4), "seconds_sketch": round(d64["seconds_sketch"], 4),
                     "seconds_qrcp": round(d64["seconds_qrcp"], 4),
                     "seconds_fit": round(d64["seconds_fit_factor"] + d64["seconds_fit_grid"], 4),
                     "blocks": int(d64["blocks"]), "speedup_of_value_form": round(t64 / t_step, 3),
                     "anchor_overlap": round(overlap, 4), "pair_space_relative_error": e64,
                     "pair_space_error_difference": abs(e - e64),
                     "relative_error_difference": abs(e - e64) / e64 if e64 else None}
        log(f"fp64 form: overlap {overlap:.3f}; pair error {e64:.4e} vs {e:.4e}")

    # ---- dominant kernels of the timed step, from its own HIP-event timings ----
    fold_s, fold_f, fold_n = d.get("fold_seconds", 0.0), d.get("fold_flops", 0.0), max(1, int(d.get("fold_launches", 0)))
    fold_kind = int(d.get("fold_kind", 0))
    fold_work = f"2*b*ell*n per fold, b = pivots folded (mean {k / fold_n:.1f}), ell={ell}, n={n}"
    if fold_kind == 2:
        ops = I8_PAIRS * fold_f
        roofline_fold = {"kernel": "fold_i8_32x32_kernel (QRCP fold on INT8 digits of the FP32 sketch)",
                         "bound": "mfma", "achieved": round(ops / fold_s / 1e12, 2) if fold_s else None,
                         "peak": PEAK["int8_mfma_tops"], "unit": "TOP/s",
                         "frac": round(ops / fold_s / 1e12 / PEAK["int8_mfma_tops"], 4) if fold_s else None,
                         "fp64_equivalent_tflops": round(fold_f / fold_s / 1e12, 2) if fold_s else None,
                         "launches": fold_n, "launch_ms": round(fold_s / fold_n * 1e3, 3),
                         "ops_per_launch": ops / fold_n,
                         "work": f"{I8_PAIRS} INT8 digit-pair products of " + fold_work}
    else:
        y_type = "FP64" if fold_kind == 0 else "FP32"
        roofline_fold = {"kernel": f"tn_gemm_kernel fold (QRCP fold on {y_type} Y)", "bound": "mfma",
                         "achieved": round(fold_f / fold_s / 1e12, 2) if fold_s else None,
                         "peak": PEAK["fp64_mfma_tflops"], "unit": "TFLOP/s",
                         "frac": round(fold_f / fold_s / 1e12 / PEAK["fp64_mfma_tflops"], 4) if fold_s else None,
                         "launches": fold_n, "launch_ms": round(fold_s / fold_n * 1e3, 3),
                         "flops_per_launch": fold_f / fold_n, "work": fold_work}
    fit_s, fit_f = d.get("fit_gemm_seconds", 0.0), d.get("fit_gemm_flops", 0.0)
    fit_n = 3 * ((n + FIT_BATCH - 1) // FIT_BATCH)
    roofline_fit = {"kernel": "tn_gemm_kernel (fit grid pass: S = (Phi_a Phi^T)^2, Z = L^-1 S, theta = L^-T Z)",
                    "bound": "mfma", "achieved": round(fit_f / fit_s / 1e12, 2) if fit_s else None,
                    "peak": PEAK["fp64_mfma_tflops"], "unit": "TFLOP/s",
                    "frac": round(fit_f / fit_s / 1e12 / PEAK["fp64_mfma_tflops"], 4) if fit_s else None,
                    "launches": fit_n, "launch_ms": round(fit_s / fit_n * 1e3, 3), "flops_per_launch": fit_f / fit_n,
                    "work": f"(2*k*d + 2*k*(k+1)) per grid point, k={k}, d={mol.nao}, n={n}, {FIT_BATCH}-point batches"}
    sk_s, sk_f, sk_n = d.get("sketch_kernel_seconds", 0.0), d.get("sketch_flops", 0.0), max(1, int(d.get("sketch_launches", 0)))
    peak_sk = PEAK["fp64_mfma_tflops" if args.precision == "fp64" else
                   ("bf16_mfma_tflops" if args.precision == "bf16" else "fp8_mfma_tflops")]
    roofline_sketch = {"kernel": f"kr_sketch ({args.precision})", "bound": "mfma",
                       "achieved": round(sk_f / sk_s / 1e12, 2) if sk_s else None, "peak": peak_sk,
                       "unit": "TFLOP/s", "frac": round(sk_f / sk_s / 1e12 / peak_sk, 4) if sk_s else None,
                       "launches": sk_n, "launch_ms": round(sk_s / sk_n * 1e3, 3),
                       "work": f"4*ell*d*n = 4*{ell}*{mol.nao}*{n} over {sk_n} grid batches"}
    # HBM bytes per launch from the committed counter passes (FETCH_SIZE and
    # WRITE_SIZE, calibrated on a known-size copy) when they cover this workload.
    traffic_file = ROOT / "profiles" / "hbm_traffic_h2o64_rand_sketch.json"
    if traffic_file.exists() and args.config == "h2o64" and args.precision == "bf16":
        measured = json.loads(traffic_file.read_text())
        for key, entry in (("qrcp_fold", roofline_fold), ("fit_grid_gemm", roofline_fit), ("bf16_sketch", roofline_sketch)):
            tag = measured.get(key, {}).get("kernel_tag")
            if tag and tag in entry["kernel"]:
                entry["traffic"] = round(measured[key]["traffic_bytes_per_launch"] / 1e9, 3)
                entry["traffic_unit"] = "GB per launch"
                entry["traffic_source"] = (f"{traffic_file.relative_to(ROOT)} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                                           f"algorithmic {measured[key]['algorithmic_bytes_per_launch'] / 1e9:.3f} "
                                           "GB per launch)")
    for entry in (roofline_fold, roofline_fit, roofline_sketch):
        entry.setdefault("traffic", None)
    # The dominant kernel is the one with the largest summed duration in the step.
    dominant = max(((fold_s, roofline_fold), (fit_s, roofline_fit), (sk_s, roofline_sketch)), key=lambda t: t[0])
    roofline = dict(dominant[1], seconds_in_step=round(dominant[0], 3))
    u_h = np.empty((mol.nao, ell))
    v_h = np.empty((mol.nao, ell))
    capi.call("mpisdf_omega", mol.nao, ell, 7, u_h, v_h)

    # ---- CPU baseline: oracle AO + FP64 sketch on a bounded sample ----------
    cpu = None
    if not args.no_cpu:
        import oracle
        oracle.build()
        m = min(args.cpu_points, n)
        sample = xyz[keep][:m]
        t = time.perf_counter()
        phi_c = oracle.ao_eval(mol.atm, mol.bas, mol.env, sample)
        oracle.sketch(phi_c, u_h, v_h, np.sqrt(np.abs(w[keep][:m])))
        tc = time.perf_counter() - t
        cpu = {"value": round(m / tc, 2), "unit": "grid-pts/s", "cores": oracle.threads(), "kind": "port",
               "sample": f"oracle FP64 AO evaluation + Khatri-Rao sketch of {m} grid points at the workload's "
                         f"shape (d={mol.nao}, ell={ell}) in {tc:.1f} s; excludes the CPU QRCP, which is "
                         f"infeasible at this size (dgeqp3 on {ell}x{n})"}
        log(f"cpu baseline: {m} points in {tc:.1f} s")

    # ---- grid Coulomb metric leg (python/isdf.py:48-67 on the GPU) ------------
    metric_leg = None
    if args.metric_points > 0:
        nm, km = min(args.metric_points, n), args.metric_vectors
        gen = torch.Generator(device="cuda").manual_seed(7)
        theta_m = torch.randn((km, nm), generator=gen, dtype=torch.float64, device="cuda")
        out_m = torch.empty((km, km), dtype=torch.float64, device="cuda")
        xm, wm = xyz_dev[:nm].contiguous(), w_dev[:nm].contiguous()
        torch.cuda.synchronize()
        t = time.perf_counter()
        capi.call("mpisdf_grid_coulomb_metric_device", theta_m, km, nm, theta_m, km, nm, xm, wm, nm, True, out_m, 0,
                  None)
        tm = time.perf_counter() - t
        # Algorithmic work of the symmetric metric: each unordered point pair
        # once (2 k flops per pair) plus the k x k accumulation (2 n k^2).
        flops_m = 1.0 * nm * nm * km + 2.0 * nm * km * km
        log(f"coulomb metric: {nm} points x {km} vectors in {tm:.2f} s ({flops_m / tm / 1e12:.1f} TFLOP/s)")
        metric_leg = {
            "workload": f"grid Coulomb metric of {km} synthetic interpolation vectors over the first {nm} kept grid "
                        f"points (the (H2O)8 cc-pVDZ size: ~2.5e5 points, 8 anchors per AO)",
            "seconds": round(tm, 4),
            "roofline": {"kernel": "coulomb_rows_kernel + tn_gemm_kernel (V_I = sum_{J>=I} K'_IJ Theta_J^T, "
                                   "M += Theta_I V_I)",
                         "bound": "mfma", "achieved": round(flops_m / tm / 1e12, 2),
                         "peak": PEAK["fp64_mfma_tflops"], "unit": "TFLOP/s",
                         "frac": round(flops_m / tm / 1e12 / PEAK["fp64_mfma_tflops"], 4), "traffic": None,
                         "work": f"n^2*k + 2*n*k^2 = {flops_m:.3e} flops (symmetric kernel: each point pair "
                                 f"once), n={nm}, k={km}, one call (wall clock of the synchronous entry point)"},
        }
        del theta_m
        if not args.no_cpu:
            import oracle
            mc = min(args.metric_cpu_points, nm)
            th_c = np.random.default_rng(7).normal(size=(km, mc))
            t = time.perf_counter()
            oracle.grid_coulomb_metric(th_c, xyz[keep][:mc], w[keep][:mc])
            tc = time.perf_counter() - t
            fc = 1.0 * mc * mc * km + 2.0 * mc * km * km  # the same algorithmic count
            metric_leg["cpu_baseline"] = {
                "value": round(fc / tc / 1e9, 2), "unit": "GFLOP/s", "cores": oracle.threads(), "kind": "port",
                "sample": f"oracle grid metric (C loops, FP64) of {km} vectors over {mc} grid points in {tc:.1f} s"}
            log(f"coulomb metric cpu sample: {mc} points in {tc:.1f} s")

    value = n / t_step
    out = {
        "metric": "anchor-select+fit s and grid-pts/s, BF16/FP8 vs FP64 GPU at equal fit error",
        "value": round(value, 1), "unit": "grid-pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic: seeded water cluster, library Becke grid, seed-7 Gaussian sketch",
        "config": {"workload": f"(H2O){nwater}/{basis} level-{level} grid rand_sketch k={k}",
                   "grid_points": n, "nao": mol.nao, "anchors": k, "sketch_rows": ell,
                   "step": "selection + FP64 fit (theta on every kept grid point, pair-space error)",
                   "precision": args.precision},
        "stages": {"seconds_sketch": round(d.get("seconds_sketch", 0), 4),
                   "seconds_qrcp": round(d.get("seconds_qrcp", 0), 4), "qrcp_blocks": int(d.get("blocks", 0)),
                   "seconds_fit_factor": round(d.get("seconds_fit_factor", 0), 4),
                   "seconds_fit_grid": round(d.get("seconds_fit_grid", 0), 4)},
        "fit_quality": {"pair_space_relative_error": d.get("pair_space_relative_error"),
                        "gram_condition_estimate": d.get("gram_condition_estimate")},
        "fp64_form": fp64_form,
        "roofline": roofline,
        "roofline_fold": dict(roofline_fold, seconds_in_step=round(fold_s, 3)),
        "roofline_fit": dict(roofline_fit, seconds_in_step=round(fit_s, 3)),
        "roofline_sketch": dict(roofline_sketch, seconds_in_step=round(sk_s, 3)),
        "cpu_baseline": cpu,
        "coulomb_metric": metric_leg,
    }
    print(json.dumps(out), flush=True)
    if args.copy_probe_mib > 0:
        src = torch.ones(args.copy_probe_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 3 x {args.copy_probe_mib} MiB device-to-device")


if __name__ == "__main__":
    main()
