"""Benchmark of the batched adding-doubling solve on one MI355X.

One step = one forward solve plus one reverse pass (gradients of a weighted
sum of flux_up, flux_down and mean_intensity at every interface with respect
to delta_tau, single_scat_albedo, the phase moments, the level Planck values,
surface and top emission and the surface albedo) over the full column set of
the headline configuration: 65536 wavenumbers x 100 layers, N = 16, solar
mu0 = 0.3 + thermal 300 -> 180 K, albedo 0.15, 32 Rayleigh/HG moments per layer
shared by all wavenumbers, with every input already resident in HBM.  The
workspace holds the whole batch, so the reverse pass reuses the forward's
layer records and recomputes only the doubling steps.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--wavenumbers W] [--form auto]

forms: the step is timed in each form the library has for this workload --
plain FP64 ("f64"), the mixed form ("mixed": FP32 doubling, FP64 adding; layers
with an unstable doubling start and their neighbours in FP64) and the FP64
forward with an FP32 reverse doubling ("mixed_reverse") -- each with its own
warm-up and K timed steps.  Each reduced-precision form is compared with the
FP64 form on the same inputs: fluxes per column relative to the column's
largest value (bar 1e-6), gradients per column norm-wise (bar 1e-4).  value and
ms_per_step are those of the fastest form within both bars on every column
(FP64 when none is); "forms" lists every form's time and differences.

roofline: the dominant kernel of the chosen form's step, with its algorithmic
work from addingdoublingrt_amd/flops.py (forward doubling (2 + 2/3 + 2 + 2 + 2
+ 2) N^3 per step, reverse doubling 20 N^3 per step, counted from the inputs)
over its time from HIP events the library records around each launch on the
solve stream; peak = MI355X FP64 78.6 TFLOP/s (FP32 matrix cores, 157.3, for a
kernel whose products run in FP32).
cpu_baseline: the FP64 CPU oracle (a restatement of the reference solver,
forward only: the reference has no adjoint) on a bounded sample of the same
columns, on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "column-wavenumbers/s fwd+adjoint, mixed vs plain FP64 at ≤1e-6 rel flux err"
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3  # the mixed form's doubling runs on FP32 matrix cores
KINDS = ("prep", "layer_doubling", "adding_interfaces", "reverse_adding_interfaces",
         "reverse_layer_doubling", "reverse_column_terms", "reverse_start_terms")
TRAFFIC_FILES = {form: os.path.join(ROOT, "profiles",
                                    f"fwd_adj_{form}_n16_65536x100_hbm_traffic.md")
                 for form in ("f64", "mixed", "mixed_reverse")}


def profiled_traffic(kernel, W, config, form):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary, for the workload it was measured on, and a note on
    where it comes from.  The summary carries the digest of the HIP sources it
    was measured on; if the sources have changed since, the figure is stale and
    None is returned."""
    from addingdoublingrt_amd.build import source_digest

    path = TRAFFIC_FILES.get(form)
    if (W, config) != (65536, 5) or path is None or not os.path.exists(path):
        return None, "not measured for this workload"
    stamp, val = None, None
    with open(path) as f:
        for line in f:
            if line.startswith("source-digest:"):
                stamp = line.split(":", 1)[1].strip()
            if line.startswith("bench-traffic:"):
                name, v = line[len("bench-traffic:"):].strip().rsplit(None, 1)
                if name == kernel:
                    val = float(v)
    rel = os.path.relpath(path, ROOT)
    if val is None:
        return None, f"{kernel} not in {rel}"
    if stamp != source_digest():
        return None, (f"{rel} was measured on HIP sources {stamp}, these are "
                      f"{source_digest()}: stale, not reported")
    return val, f"{rel}, rocprofv3 FETCH_SIZE + WRITE_SIZE on HIP sources {stamp}"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--wavenumbers", type=int, default=65536)
    p.add_argument("--config", type=int, default=5)
    p.add_argument("--form", choices=("auto", "f64", "mixed", "mixed_reverse"), default="auto",
                   help="auto: time every form, report the fastest within the accuracy bar")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--forward-only", action="store_true",
                   help="time the forward solve alone (any quadrature order; the default "
                        "line is forward + reverse pass)")
    p.add_argument("--workspace-gb", type=float, default=96.0)
    p.add_argument("--summarize-stats", metavar="DIR")
    p.add_argument("--summarize-traffic", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH",
                                                            "COPY_WRITE"))
    p.add_argument("--copy-mib-size", type=float, default=1024.0)
    p.add_argument("--title", default="rocprofv3 summary")
    p.add_argument("--cmd", action="append", default=[])
    p.add_argument("--copy-mib", type=int, default=0,
                   help="only run a plain device copy of this many MiB, three times (the "
                        "reference kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cols, planck_host, seconds):
    """FP64 oracle on a bounded sample of the same columns (forward only), on
    every core this process may use (at most 16, the host share of one GPU),
    and on one thread."""
    import oracle

    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(avail or 1, 16))
    B, Bs, Bt = planck_host

    def run(n, threads=threads):
        sl = slice(0, n)
        mom = cols.phase_moments if cols.phase_moments.ndim == 2 else cols.phase_moments[sl]
        t0 = time.perf_counter()
        oracle.solve_batch(cols.delta_tau[sl], cols.single_scat_albedo[sl], mom, cols.N,
                           planck_levels=None if B is None else B[sl],
                           surface_emission=0.0 if Bs is None else Bs[sl],
                           top_emission=0.0 if Bt is None else Bt[sl],
                           surface_albedo=cols.surface_albedo, solar_flux=cols.solar_flux,
                           solar_mu=cols.solar_mu, use_delta_m=cols.use_delta_m,
                           use_diffusion_lower_bc=cols.use_diffusion_lower_bc,
                           threads=threads)
        return time.perf_counter() - t0

    n = min(cols.W, 4 * threads)
    dt = run(n)
    n2 = int(min(cols.W, max(n, n * seconds / max(dt, 1e-3))))
    dt2 = run(n2)
    # one thread: a quarter of the wall time on a proportional sample
    n1 = max(1, int(n2 / threads * (dt2 > 0) / 4) or 1)
    dt1 = run(n1, threads=1)
    return dict(value=n2 / dt2, unit="column-wavenumbers/s", cores=threads, kind="port",
                nproc=os.cpu_count(), cores_available=avail, cpu_model=cpu_model(),
                single_thread_value=n1 / dt1,
                sample=f"first {n2} of the {cols.W} columns, FP64 forward only (the reference "
                       f"has no adjoint), {dt2:.1f} s on {threads} threads; one thread: first "
                       f"{n1} columns in {dt1:.1f} s",
                note="the FP64 C restatement of the reference column solver in oracle/ "
                     "(naive loops, no Eigen: the reference's Eigen build cannot be made "
                     "offline), one column per task over a thread pool, gcc -O3 -march=x86-64-v3 -ffp-contract=off")


FORM_TEXT = {
    "f64": "plain FP64 forward",
    "mixed": "mixed forward (FP32 doubling with FP64-refined solves, FP64 adding; unstable-start "
             "layers and their neighbours in FP64)",
    "mixed_reverse": "FP64 forward (outputs identical to the plain form)",
}
REV_TEXT = {"f64": " + FP64 reverse pass", "mixed": " + reverse pass with FP32 doubling",
            "mixed_reverse": " + reverse pass with FP32 doubling (routed layers FP64)"}
FLUX_BAR, GRAD_BAR = 1e-6, 1e-4


def accuracy(out, grads, ref_out, ref_grads, W):
    """Differences of one form from the FP64 form on the same inputs.

    fluxes: per column, the largest |x - x_f64| over the interfaces of flux_up,
    flux_down, mean_intensity and flux_direct, relative to that output's largest
    |x_f64| in the column; gradients: per column, ||g - g_f64|| / ||g_f64|| over
    the column's entries of each [W][...] gradient, and for the shared moments
    and the surface albedo (sums over the columns) the relative difference of
    the whole array.  Reports the largest value, the 99th percentile and the
    fraction of columns within the bar (fluxes 1e-6, gradients 1e-4)."""
    import torch

    def summary(e):
        e = e.double().cpu()
        return {"max": float(e.max()), "p99": float(torch.quantile(e, 0.99)) if e.numel() > 1
                else float(e.max()), "median": float(e.median()),
                "frac_within_bar": None}

    res = {}
    col = torch.zeros(W, dtype=torch.float64, device=out["flux_up"].device)
    for k in ("flux_up", "flux_down", "mean_intensity", "flux_direct"):
        scale = ref_out[k].abs().amax(dim=1).clamp_min(1e-300)
        e = ((out[k] - ref_out[k]).abs().amax(dim=1) / scale)
        col = torch.maximum(col, e)
    fl = summary(col)
    fl["frac_within_bar"] = float((col <= FLUX_BAR).double().mean())
    res["fluxes"] = fl
    worst_grad = 0.0
    gcol = torch.zeros(W, dtype=torch.float64, device=col.device)
    for k, v in ref_grads.items():
        m = grads[k]
        if v.dim() >= 1 and v.shape[0] == W and not (k == "surface_albedo"):
            a, b = m.reshape(W, -1), v.reshape(W, -1)
            nb = b.norm(dim=1)
            e = torch.where(nb > 0, (a - b).norm(dim=1) / nb.clamp_min(1e-300),
                            torch.zeros_like(nb))
            gcol = torch.maximum(gcol, e)
            r = summary(e)
            r["frac_within_bar"] = float((e <= GRAD_BAR).double().mean())
        else:  # one array for all columns
            tot_b = v.sum() if k == "surface_albedo" else v
            tot_a = m.sum() if k == "surface_albedo" else m
            e = float((tot_a - tot_b).norm() / tot_b.norm().clamp_min(1e-300))
            r = {"max": e, "p99": e, "median": e, "frac_within_bar": float(e <= GRAD_BAR)}
        res["d_" + k] = r
        worst_grad = max(worst_grad, r["max"])
    both = torch.maximum(col / FLUX_BAR, gcol / GRAD_BAR)
    res["columns_within_both_bars"] = float((both <= 1.0).double().mean())
    res["meets_bar"] = fl["max"] <= FLUX_BAR and worst_grad <= GRAD_BAR
    return res


def summarize(a):
    """Markdown summaries of rocprofv3 output for profiles/, stamped with the
    digest of the HIP sources they were measured on:
      --summarize-stats DIR            kernel statistics (--kernel-trace --stats)
      --summarize-traffic F W CF CW    FETCH_SIZE / WRITE_SIZE per dispatch of the
                                       step (F, W) and of the calibration copy
                                       (CF, CW; bench.py --copy-mib 1024)"""
    import csv
    import glob

    from addingdoublingrt_amd.build import source_digest

    def rows(d, pattern):
        files = glob.glob(os.path.join(d, "**", pattern), recursive=True)
        if not files:
            raise SystemExit(f"no {pattern} under {d}")
        with open(files[0]) as f:
            return list(csv.DictReader(f))

    def short(name):
        name = name.replace("(anonymous namespace)", "{anonymous}")
        return name.split("(")[0].replace("void ", "").strip() or name

    print(f"# {a.title}\n")
    print(f"source-digest: {source_digest()}\n")
    print("Commands (from the repository root, after `cd /tmp && export TMPDIR=/tmp`):\n")
    print("```")
    for c in a.cmd:
        print(c)
    print("```\n")
    if a.summarize_stats:
        r = rows(a.summarize_stats, "*kernel_stats.csv")
        print("| kernel | calls | average ms | total ms | % |")
        print("|---|---:|---:|---:|---:|")
        for x in sorted(r, key=lambda x: -float(x["TotalDurationNs"])):
            print(f"| `{short(x['Name'])}` | {x['Calls']} | {float(x['AverageNs']) / 1e6:.3f} | "
                  f"{float(x['TotalDurationNs']) / 1e6:.2f} | {float(x['Percentage']):.2f} |")
        return
    fd, wd, cfd, cwd = a.summarize_traffic

    def per_dispatch(d, counter):
        acc = {}
        for x in rows(d, "*counter_collection.csv"):
            if x["Counter_Name"] != counter:
                continue
            key = (short(x["Kernel_Name"]), x["Dispatch_Id"])
            acc[key] = acc.get(key, 0.0) + float(x["Counter_Value"])
        per = {}
        for (k, _), v in acc.items():
            per.setdefault(k, []).append(v)
        return {k: sum(v) / len(v) for k, v in per.items()}

    cf = per_dispatch(cfd, "FETCH_SIZE")
    cw = per_dispatch(cwd, "WRITE_SIZE")
    copy = [k for k in cf if "copyBuffer" in k][0]
    copy_kib = float(a.copy_mib_size) * 1024
    ff, fw = copy_kib / cf[copy], copy_kib / cw[copy]
    print(f"Calibration: a {a.copy_mib_size} MiB device copy reads FETCH_SIZE = {cf[copy]:.0f} KiB and "
          f"writes WRITE_SIZE = {cw[copy]:.0f} KiB per dispatch, so bytes = FETCH_SIZE x 1024 x "
          f"{ff:.3f} and WRITE_SIZE x 1024 x {fw:.3f}.\n")
    F = per_dispatch(fd, "FETCH_SIZE")
    Wr = per_dispatch(wd, "WRITE_SIZE")
    print("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read GB | written GB | total GB |")
    print("|---|---:|---:|---:|---:|---:|")
    tot = {}
    for k in sorted(F, key=lambda k: -(F[k] + Wr.get(k, 0.0))):
        if "adrt" not in k:
            continue
        rd, wr = F[k] * 1024 * ff, Wr.get(k, 0.0) * 1024 * fw
        tot[k] = rd + wr
        print(f"| `{k}` | {F[k]:.0f} | {Wr.get(k, 0.0):.0f} | {rd / 1e9:.1f} | {wr / 1e9:.1f} | "
              f"{(rd + wr) / 1e9:.1f} |")
    print("\nPer-launch HBM-side bytes (read + written) that bench.py reports as roofline.traffic:\n")
    for k, v in tot.items():
        name = k.split("::")[-1]
        print(f"bench-traffic: {name} {v:.6e}")


def main():
    a = parse()
    if a.summarize_stats or a.summarize_traffic:
        summarize(a)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); columns are independent, "
              "so more GPUs would each solve their own share of the wavenumbers.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if a.copy_mib:
        src = torch.ones(a.copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8, "copies": 3}))
        return

    from addingdoublingrt_amd import _lib, adjoint_batch, flops, solve_batch, workloads
    from addingdoublingrt_amd.solver import get_workspace

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    cols = workloads.make(a.config, W=a.wavenumbers)
    d = workloads.device_inputs(cols, "cuda")
    cfg = workloads.batch_config(cols)
    W, L = cols.W, cols.L
    ws_bytes = int(a.workspace_gb * (1 << 30))
    g = torch.Generator(device="cuda").manual_seed(5)
    cot = {k: torch.rand((W, L + 1), dtype=torch.float64, device="cuda", generator=g)
           for k in ("flux_up", "flux_down", "mean_intensity")}
    inputs = (d["delta_tau"], d["single_scat_albedo"], d["phase_moments"], cfg)
    kw = dict(planck_levels=d["planck_levels"], surface_emission=d["surface_emission"],
              top_emission=d["top_emission"])
    reverse_ok = cols.N <= 16  # the reduced-precision reverse forms run up to order 16
    npad = 1 << max(cols.N - 1, 1).bit_length()  # the library's padded order
    if a.form == "auto":
        forms = ["f64"] + (["mixed", "mixed_reverse"] if reverse_ok else [])
        if a.forward_only:  # the mixed forward also runs at N = 32
            forms = ["f64"] + (["mixed"] if reverse_ok or npad == 32 else [])
    else:
        forms = [a.form]
    lib = _lib.lib()
    Bs_host = d["surface_emission"].cpu().numpy() if d["surface_emission"] is not None else None
    work = flops.forward_flops(cols.delta_tau, cols.single_scat_albedo, cols.phase_moments,
                               cols.N, cols.use_delta_m, cols.surface_albedo, Bs_host,
                               cols.use_diffusion_lower_bc)

    def run_form(form):
        # reduced precision below N = 16 runs padded to 16 on a reverse-pass workspace
        adjoint_ws = not a.forward_only or (form != "f64" and reverse_ok)
        out = {k: torch.empty((W, L + 1), dtype=torch.float64, device="cuda")
               for k in ("flux_up", "flux_down", "mean_intensity", "flux_direct",
                         "flux_divergence")}
        grads = {}

        def step():
            o = solve_batch(*inputs, **kw, out=out, max_workspace_bytes=ws_bytes,
                            adjoint_workspace=adjoint_ws, precision=form)
            gen = o.pop("_generation", None)
            if not a.forward_only:
                grads.update(adjoint_batch(*inputs, out, cot, **kw, forward_generation=gen,
                                           max_workspace_bytes=ws_bytes, check_status=False,
                                           precision=form))

        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize()
        ws = get_workspace(W, L, cols.N, cols.phase_moments.shape[-1], ws_bytes, adjoint_ws)
        _lib.check(lib.adrt_profile(ws.handle, 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        ms = (C.c_double * _lib.PROFILE_KINDS)()
        nl = (C.c_int * _lib.PROFILE_KINDS)()
        _lib.check(lib.adrt_profile_read(ws.handle, ms, nl))
        _lib.check(lib.adrt_profile(ws.handle, 0))
        if not a.forward_only:
            stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
            _lib.check(lib.adrt_adjoint_status(ws.handle, stream))
        return dict(ms_per_step=elapsed * 1e3 / a.steps, kind_ms=list(ms), launches=list(nl),
                    out=out, grads=grads)

    runs = {f: run_form(f) for f in forms}
    acc = {}
    if "f64" in runs:
        for f in forms:
            if f != "f64":
                acc[f] = accuracy(runs[f]["out"], runs[f]["grads"], runs["f64"]["out"],
                                  runs["f64"]["grads"], W)
    # the value: the fastest form whose fluxes and gradients are within the bar
    # of the FP64 form on every column of this workload (FP64 itself qualifies)
    ok = [f for f in forms if f == "f64" or acc.get(f, {}).get("meets_bar")]
    chosen = min(ok or forms, key=lambda f: runs[f]["ms_per_step"])
    r = runs[chosen]
    ms_per_step = r["ms_per_step"]
    ms, nl = r["kind_ms"], r["launches"]
    per_kind = {k: ms[i] / a.steps for i, k in enumerate(KINDS)}
    # dominant kernel and its algorithmic work
    if chosen == "f64" and cols.N <= 8:  # two layers per wave (pair mode)
        model = {1: ("layer16p_kernel + layer16_kernel<false, true>",
                     "phase setup + doubling (forward), two layers per wave", work["doubling"]),
                 4: ("layer16p_adjoint_kernel + layer16_adjoint_kernel<false, true>",
                     "recomputed doubling + its reverse (reverse counted), two layers per wave",
                     work["doubling_reverse"])}
    elif chosen == "f64" and npad == 32:
        model = {1: ("layer32_kernel<false>", "phase setup + doubling (forward)", work["doubling"]),
                 4: ("layer32_adjoint_kernel",
                     "recomputed doubling + its reverse (reverse counted)",
                     work["doubling_reverse"])}
    elif chosen == "f64":
        model = {1: ("layer16_kernel<false, false>", "phase setup + doubling (forward)", work["doubling"]),
                 4: ("layer16_adjoint_kernel<false, false>",
                     "recomputed doubling + its reverse (reverse counted)",
                     work["doubling_reverse"])}
    elif chosen == "mixed":
        model = {1: ("layer16m_kernel + layer16_kernel<true, false>",
                     "mixed doubling (forward; FP64 for the routed layers), FP64 flop count",
                     work["doubling"]),
                 4: ("layer16m_adjoint_kernel + layer16_adjoint_kernel<true, false>",
                     "FP32 recomputed doubling + reverse (FP64 for the routed layers), "
                     "FP64 flop count", work["doubling_reverse"])}
    else:
        model = {1: ("layer16_kernel<false, false>", "phase setup + doubling (forward)", work["doubling"]),
                 4: ("layer16m_adjoint_kernel + layer16_adjoint_kernel<true, false>",
                     "FP32 recomputed doubling + reverse (FP64 for the routed layers), "
                     "FP64 flop count", work["doubling_reverse"])}
    if a.forward_only:
        names = ({1: "layer32_kernel<false>", 2: "adding32_kernel<false>"} if npad == 32 else
                 {1: "layer16_kernel<false, false>", 2: "adding16_kernel<false, false>"}
                 if npad == 16 or chosen != "f64" else
                 {1: f"layer_kernel<{npad}>", 2: f"adding_kernel<{npad}>"})
        if chosen == "mixed":
            names = ({1: "layer32m_kernel + layer32_kernel<true>", 2: "adding32_kernel<false>"}
                     if npad == 32 else
                     {1: "layer16m_kernel + layer16_kernel<true, false>",
                      2: "adding16_kernel<false, true>"})
        model = {1: (names[1], "phase setup + doubling (forward)", work["doubling"]),
                 2: (names[2], "adding sweeps + interfaces (forward)",
                     work["adding"] + work["interfaces"])}
    dom = max(model, key=lambda k: ms[k])
    kname, name, fl_step = model[dom]
    launches = nl[dom] / a.steps
    avg_ms = ms[dom] / max(nl[dom], 1)
    achieved = (fl_step / launches) / (avg_ms * 1e-3) / 1e12
    fp32_kernel = (dom == 4 and chosen != "f64") or (dom == 1 and chosen == "mixed")
    peak = FP32_MFMA_PEAK_TFLOPS if fp32_kernel else FP64_PEAK_TFLOPS
    vals = [profiled_traffic(part, W, a.config, chosen) for part in kname.split(" + ")]
    traffic = sum(v for v, _ in vals) if all(v is not None for v, _ in vals) else None
    traffic_note = "; ".join(dict.fromkeys(n for _, n in vals))
    res = {
        "metric": METRIC if not a.forward_only else "column-wavenumbers/s forward",
        "value": W / (ms_per_step * 1e-3),
        "unit": "column-wavenumbers/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if chosen == "f64" else "f32+f64",
        "data": "synthetic (numpy PCG64 seed 20261203+k; Planck levels from the device kernel)",
        "config": {
            "workload": f"config {a.config}: {cols.name}, {W} wavenumbers x {L} layers, "
                        f"N={cols.N}",
            "form": chosen + ": " + FORM_TEXT[chosen]
                    + ("" if a.forward_only else REV_TEXT[chosen]),
            "num_wavenumbers": W, "num_layers": L, "num_quadrature": cols.N,
            "algorithmic_gflop_per_step": (work["total"] if a.forward_only else
                                           2 * work["total"] + work["doubling_reverse"]
                                           - work["doubling"]) / 1e9,
            "doubling_steps": work["doubling_steps"], "general_adds": work["general_adds"],
            "kernel_ms_per_step": per_kind,
        },
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "traffic_source": traffic_note,
            "kernel": f"{kname}: {name}",
            "launches_per_step": launches, "avg_launch_ms": avg_ms,
        },
        "forms": {f: {"ms_per_step": runs[f]["ms_per_step"],
                      "value": W / (runs[f]["ms_per_step"] * 1e-3),
                      "speedup_vs_f64": (runs["f64"]["ms_per_step"] / runs[f]["ms_per_step"]
                                         if "f64" in runs else None),
                      "kernel_ms_per_step": {k: runs[f]["kind_ms"][i] / a.steps
                                             for i, k in enumerate(KINDS)},
                      "vs_f64": acc.get(f)}
                  for f in forms},
        "bar": {"fluxes": FLUX_BAR, "gradients": GRAD_BAR,
                "rule": "value is the fastest form whose fluxes (per column, relative to the "
                        "column's largest value) and gradients (per column, norm-wise) are "
                        "within the bar of the FP64 form on every column"},
    }
    if not a.no_cpu:
        ph = (None, None, None)
        if cols.thermal:
            ph = (d["planck_levels"].cpu().numpy(), Bs_host, d["top_emission"].cpu().numpy())
        res["cpu_baseline"] = cpu_baseline(cols, ph, a.cpu_seconds)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
