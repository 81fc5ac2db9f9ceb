#!/usr/bin/env python3
"""Benchmark of the SUS2-SH EFS hot path on one MI355X.

Workload ("effect4096"): Effect L3/K4/R2/O2 model (gate coupling cross, outer
coupling diagonal, body order 4, radial basis 10, warp sharing K, per-element
readout), 2 species drawn 50/50, diamond 8x8x8 conventional cells = 4096
atoms, a = 3.567 A, Gaussian displacements sigma = 0.10 A, cutoff 5.0 A.
Random-initialised coefficients from a seed.

A step = one EFS evaluation of the frame (device neighbour list, gate pass,
outer pass, gather-based force/virial backward, FP64 per-frame reductions)
with inputs resident in HBM.  `value` is FP32 atom-steps/s; the FP64 rate, the
FP32/FP64 speed ratio and the FP32 errors against the FP64 oracle ride along.

Training workload ("train256", reported beside the headline figure): the same
model form, 256 periodic frames with atom counts uniform in [8, 216], random
triclinic cells of density 0.17-0.19 atoms/A^3, sigma = 0.15 A, labels (E, F)
from a teacher model (the same coefficients perturbed with a fixed seed).  A
training step = the objective (weighting "structures", energy weight 1, force
weight 0.01: E/atom MSE + 0.01 force MSE) and its full parameter gradient for
the whole batch (EFS, loss adjoints, forward-over-reverse gradient kernels).

Static workload ("static64"): Static L3/K4/O2, one species, the 64-atom
diamond (2x2x2 cells, sigma 0.10 A), plain FP64 EFS (and FP32) with the
reference's embedded CPU engine timed beside it.

Batched workload ("batch4096"): Effect L4/K6/R2/O3, 3 species, 4096 frames
with N ~ logU[16, 512], half periodic cells and half molecules, all in one
call; atom-steps/s, p50/p99 batch latency, FP64 ratio and the force-loss
gradient rate.  Large-frame workload ("md2m"): Effect L5/K6/R2/O3, diamond
64^3 cells = 2,097,152 atoms, FP32 and FP64 EFS.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "EFS atom-steps/s FP64 vs FP32 @4096 atoms; train grad frames/s; % roofline"


def _timed(model, batch, precision, steps, warmup, profile):
    import torch
    for _ in range(warmup):
        model.efs_device(batch, precision)
    torch.cuda.synchronize()
    model.profile_reset()
    model.profile(profile)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = model.efs_device(batch, precision)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    model.profile(False)
    prof = model.profile_read() if profile else {}
    return t1 - t0, out, prof


def _timed_train(model, batch, labels, precision, steps, warmup, profile):
    import torch
    for _ in range(warmup):
        model.loss_grad_device(batch, *labels, weighting="structures", wE=1.0, wF=0.01, wS=0.0, precision=precision)
    torch.cuda.synchronize()
    model.profile_reset()
    model.profile(profile)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, g = model.loss_grad_device(batch, *labels, weighting="structures", wE=1.0, wF=0.01, wS=0.0,
                                         precision=precision)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    model.profile(False)
    prof = model.profile_read() if profile else {}
    return t1 - t0, loss, g.cpu().numpy(), prof


def labels_host(lab, batch):
    return dict(energy=lab["energy"].cpu().numpy(), forces=lab["forces"].cpu().numpy())


def train_cpu_baseline(path, frames, labels, threads=None, count=1):
    """The FP64 oracle's objective + full gradient on the first `count`
    training frames (hyper-dual numbers: one forward pass per coefficient,
    OpenMP over coefficients).  There is no reference gradient engine that
    builds here, so this is the port's figure, not a reference-equivalent."""
    from oracle import oracle
    _oracle_ready()
    threads = threads or min(16, os.cpu_count() or 1)
    sub = frames[:count]
    n = sum(len(f["types"]) for f in sub)
    lab = dict(energy=labels["energy"][:count], forces=labels["forces"][:n])
    m = oracle.OracleModel(path)
    t0 = time.perf_counter()
    m.loss_grad(sub, lab, mode=1, wE=1.0, wF=0.01, wS=0.0, nthreads=threads)
    dt = time.perf_counter() - t0
    return dict(value=round(count / dt, 4), unit="frame-gradients/s", cores=threads, kind="port",
                sample=f"first {count} training frames ({n} atoms), objective + all {m.offsets()['ncoeff']} "
                       f"gradient entries by hyper-dual forward passes, {dt:.1f} s")


def train_section(path, steps, warmup, profile, no_cpu=False):
    """Loss + parameter gradient over the 256-frame training batch, FP32 and FP64."""
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.flops import kernel_flops, step_grad_per_edge
    from superlinear_amd.model import DeviceBatch, Sus2Model
    frames = W.config_frames("train256")
    nf = len(frames)
    batch = DeviceBatch(frames)
    teacher = Sus2Model(path)
    th = teacher.coeffs()
    rng = np.random.default_rng(3031)
    teacher.set_coeffs(th + 0.05 * np.abs(th) * rng.standard_normal(th.size))
    lab = teacher.efs_device(batch, "fp64")
    labels = (lab["energy"], lab["forces"], None)
    model = Sus2Model(path)
    dims = model.dims()
    print(f"[bench] train256: {nf} frames, {batch.natoms} atoms", file=sys.stderr, flush=True)
    t32, l32, g32, _ = _timed_train(model, batch, labels, "fp32", steps, warmup, False)
    nedges = model.last_edge_count()
    t64, l64, g64, _ = _timed_train(model, batch, labels, "fp64", steps, warmup, False)
    tmx, lmx, gmx, _ = _timed_train(model, batch, labels, "mixed", steps, warmup, False)
    prof = _timed_train(model, batch, labels, "fp32", steps, 0, True)[3] if profile else {}
    out = dict(
        train_grad_frames_per_s=round(nf * steps / t32, 2),
        train_grad_fp64_frames_per_s=round(nf * steps / t64, 2),
        train_fp32_over_fp64=round(t64 / t32, 3),
        train_ms_per_step_fp32=round(1e3 * t32 / steps, 3),
        train_ms_per_step_fp64=round(1e3 * t64 / steps, 3),
        train_config={"workload": "train256", "frames": nf, "atoms": batch.natoms, "directed_edges": nedges,
                      "loss": "weighting structures, wE 1, wF 0.01, wS 0", "coefficients": len(g32)},
        train_errors_fp32_vs_fp64_gpu=dict(loss_rel=float(abs(l32 - l64) / abs(l64)),
                                           grad_rel=float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))),
        train_mixed_frames_per_s=round(nf * steps / tmx, 2),
        train_mixed_over_fp64=round(t64 / tmx, 3),
        train_errors_mixed_vs_fp64_gpu=dict(loss_rel=float(abs(lmx - l64) / abs(l64)),
                                            grad_rel=float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64))),
        train_step_formula_tflops=round(step_grad_per_edge(dims) * nedges * steps / t32 / 1e12, 4),
    )
    if prof:
        out["kernels_train_fp32"] = {k: dict(avg_us=round(1e3 * ms / max(c, 1), 2), launches=c)
                                     for k, (ms, c) in prof.items()}
        grad = {k: v for k, v in prof.items() if k.startswith("grad_") and kernel_flops(k, dims, 1) is not None}
        if grad:
            dom = max(grad.items(), key=lambda kv: kv[1][0])[0]
            avg_s = prof[dom][0] / prof[dom][1] / 1e3
            out["train_roofline"] = _roofline(dom, avg_s, kernel_flops(dom, dims, nedges), "fp32",
                                              measured_traffic(dom, "train256"))
    out["train_cpu_baseline"] = None if no_cpu else train_cpu_baseline(path, frames, labels_host(lab, batch))
    return out


def _call_latencies(fn, steps, warmup):
    """Seconds of each of `steps` synchronised calls of fn (after `warmup`)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    lat = []
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    return np.array(lat)


def _roofline(kernel, avg_s, flops_, precision, traffic=None):
    """Roofline entry of one kernel: algorithmic flops per launch over its
    average launch time, against MI355X's vector peak of the precision."""
    from superlinear_amd.flops import PEAK
    ach = flops_ / avg_s / 1e12
    if precision == "fp32":
        peak, peak_np = PEAK["fp32_tflops"], PEAK["fp32_nonpacked_tflops"]
        basis = ("FP32 vector peak 157.3 TF counts packed two-wide FMA; this build emits no packed FP32, "
                 "so its reachable FP32 rate is the one-FMA-per-lane 78.6 TF (frac_nonpacked)")
    else:
        peak = peak_np = PEAK["fp64_tflops"]
        basis = "FP64 vector peak 78.6 TF"
    return dict(bound="valu", kernel=kernel, achieved=round(ach, 4), peak=peak, unit="TFLOP/s",
                frac=round(ach / peak, 5), frac_nonpacked=round(ach / peak_np, 5), peak_basis=basis,
                traffic=traffic, flops_per_launch=int(flops_), avg_launch_us=round(avg_s * 1e6, 2))


def _roofline_of(prof, dims, nedges, natoms, graph, precision, workload=None):
    """Roofline entry of the kernel with the largest summed time among those
    with an algorithmic flop count (HIP-event profile of the same calls)."""
    from superlinear_amd.flops import kernel_flops
    timed = {k: v for k, v in prof.items() if kernel_flops(k, dims, 1, 1, graph) is not None}
    if not timed:
        return None
    dom = max(timed.items(), key=lambda kv: kv[1][0])[0]
    avg_s = prof[dom][0] / prof[dom][1] / 1e3
    fl = kernel_flops(dom, dims, nedges, natoms, graph)
    return _roofline(dom, avg_s, fl, precision, measured_traffic(dom, workload) if workload else None)


def static_section(steps, warmup, profile, no_cpu):
    """The 64-atom Static L3/K4/O2 diamond frame (1 species, FP64 plain path
    and FP32), errors against the oracle, the reference CPU engine beside it."""
    from oracle import oracle
    from superlinear_amd import workloads as W
    from superlinear_amd.flops import graph_stats, step_efs_per_atom, step_efs_per_edge
    from superlinear_amd.model import DeviceBatch, Sus2Model
    path = W.model_path("static_l3k4o2")
    frames = W.config_frames("static64")
    n = len(frames[0]["types"])
    batch = DeviceBatch(frames)
    model = Sus2Model(path)
    dims = model.dims()
    print(f"[bench] static64: {n} atoms", file=sys.stderr, flush=True)
    t64, o64, _ = _timed(model, batch, "fp64", steps, warmup, False)
    nedges = model.last_edge_count()
    t32, o32, _ = _timed(model, batch, "fp32", steps, warmup, False)
    graph = graph_stats(path)
    out = dict(workload="static64", model="Static L3/K4/O2 body 4, B=10, 1 species", atoms=n, directed_edges=nedges,
               atom_steps_per_s=round(n * steps / t64, 1), ms_per_step=round(1e3 * t64 / steps, 4),
               dtype="f64", fp32_atom_steps_per_s=round(n * steps / t32, 1),
               fp32_ms_per_step=round(1e3 * t32 / steps, 4),
               step_formula_tflops=round((step_efs_per_edge(dims) * nedges + step_efs_per_atom(dims, graph) * n)
                                         * steps / t64 / 1e12, 5))
    if profile:
        prof = _timed(model, batch, "fp64", steps, 1, True)[2]
        out["kernels_fp64"] = {k: dict(avg_us=round(1e3 * ms / max(c, 1), 2), launches=c) for k, (ms, c) in prof.items()}
        out["roofline"] = _roofline_of(prof, dims, nedges, n, graph, "fp64", "static64")
    if not no_cpu:
        _oracle_ready()
        ref = oracle.OracleModel(path).efs(frames)
        e64 = {k: v.cpu().numpy() for k, v in o64.items()}
        e32 = {k: v.cpu().numpy() for k, v in o32.items()}
        out["errors_vs_fp64_oracle"] = dict(
            fp64_energy_rel=float(abs(e64["energy"][0] - ref["energy"][0]) / abs(ref["energy"][0])),
            fp64_force_maxabs=float(np.abs(e64["forces"] - ref["forces"]).max()),
            fp32_force_maxabs=float(np.abs(e32["forces"] - ref["forces"]).max()))
        out["cpu_baseline"] = cpu_baseline(path, frames, "static64")
    return out


def batch_section(steps, warmup, profile):
    """Uneven batched EFS and its force-loss gradient: 4096 frames, the launch
    shape picked by the autotuner, eager calls and the same work captured in
    one HIP graph per step."""
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.flops import graph_stats, step_efs_per_atom, step_efs_per_edge
    from superlinear_amd.model import DeviceBatch, Sus2Model
    path = W.model_path("effect_l4k6r2o3")
    frames = W.config_frames("batch4096")
    nf, n = len(frames), sum(len(f["types"]) for f in frames)
    batch = DeviceBatch(frames)
    model = Sus2Model(path)
    dims = model.dims()
    graph = graph_stats(path)
    print(f"[bench] batch4096: {nf} frames, {n} atoms", file=sys.stderr, flush=True)
    tuned = {}
    for prec in ("fp32", "fp64"):
        best, table = model.autotune(batch, prec, reps=3)
        tuned[prec] = dict(choice=best, median_ms={f"eb{k[0]}_grid{k[1]}": round(v, 3) for k, v in table.items()})
    print(f"[bench] batch4096 tuned: fp32 {tuned['fp32']['choice']} fp64 {tuned['fp64']['choice']}",
          file=sys.stderr, flush=True)
    l32 = _call_latencies(lambda: model.efs_device(batch, "fp32"), steps, warmup)
    nedges = model.last_edge_count()
    l64 = _call_latencies(lambda: model.efs_device(batch, "fp64"), steps, warmup)
    # the same EFS captured once in a HIP graph, one graph launch per step
    g32 = model.graph_efs(batch, "fp32")
    lgr = _call_latencies(g32.launch, steps, warmup)
    assert not g32.overflow()
    per_step = step_efs_per_edge(dims) * nedges + step_efs_per_atom(dims, graph) * n
    out = dict(workload="batch4096", model="Effect L4/K6/R2/O3 body 4, B=10, 3 species", frames=nf, atoms=n,
               directed_edges=nedges, launch_shape=tuned,
               atom_steps_per_s=round(n * steps / l32.sum(), 1),
               frames_per_s=round(nf * steps / l32.sum(), 1), p50_ms=round(1e3 * float(np.percentile(l32, 50)), 3),
               p99_ms=round(1e3 * float(np.percentile(l32, 99)), 3),
               graph_atom_steps_per_s=round(n * steps / lgr.sum(), 1),
               graph_p50_ms=round(1e3 * float(np.percentile(lgr, 50)), 3),
               graph_p99_ms=round(1e3 * float(np.percentile(lgr, 99)), 3),
               fp64_atom_steps_per_s=round(n * steps / l64.sum(), 1), fp32_over_fp64=round(l64.sum() / l32.sum(), 3),
               step_formula_tflops=round(per_step * steps / l32.sum() / 1e12, 4))
    del g32
    # force-loss gradient of the whole batch (labels from a perturbed teacher)
    teacher = Sus2Model(path)
    th = teacher.coeffs()
    teacher.set_coeffs(th + 0.05 * np.abs(th) * np.random.default_rng(4041).standard_normal(th.size))
    lab = teacher.efs_device(batch, "fp64")
    gsteps = max(1, steps // 2)
    lg = _call_latencies(lambda: model.loss_grad_device(batch, lab["energy"], lab["forces"], weighting="structures",
                                                        wE=1.0, wF=0.01, precision="fp32"), gsteps, 1)
    out["grad_frames_per_s"] = round(nf * gsteps / lg.sum(), 1)
    out["grad_ms_per_step"] = round(1e3 * lg.mean(), 3)
    gl = model.graph_loss_grad(batch, lab["energy"], lab["forces"], weighting="structures", wE=1.0, wF=0.01,
                               precision="fp32")
    lgl = _call_latencies(gl.launch, gsteps, 1)
    assert not gl.overflow()
    out["graph_grad_frames_per_s"] = round(nf * gsteps / lgl.sum(), 1)
    out["graph_grad_p50_ms"] = round(1e3 * float(np.percentile(lgl, 50)), 3)
    out["graph_grad_p99_ms"] = round(1e3 * float(np.percentile(lgl, 99)), 3)
    del gl
    if profile:
        model.profile_reset()
        model.profile(True)
        model.efs_device(batch, "fp32")
        torch.cuda.synchronize()
        model.profile(False)
        prof = model.profile_read()
        out["kernels_fp32"] = {k: dict(avg_us=round(1e3 * ms / max(c, 1), 2), launches=c) for k, (ms, c) in prof.items()}
        out["roofline"] = _roofline_of(prof, dims, nedges, n, graph, "fp32", "batch4096")
    return out


def md_section(steps, warmup, profile):
    """One large frame: 2,097,152-atom diamond, Effect L5/K6, FP32 and FP64."""
    from superlinear_amd import workloads as W
    from superlinear_amd.flops import graph_stats, step_efs_per_edge
    from superlinear_amd.model import DeviceBatch, Sus2Model
    path = W.model_path("effect_l5k6r2o3")
    frame = W.config_frames("md2m")[0]
    n = len(frame["types"])
    batch = DeviceBatch([frame])
    model = Sus2Model(path)
    dims = model.dims()
    print(f"[bench] md2m: {n} atoms", file=sys.stderr, flush=True)
    l32 = _call_latencies(lambda: model.efs_device(batch, "fp32"), steps, warmup)
    nedges = model.last_edge_count()
    print(f"[bench] md2m fp32 {1e3 * l32.mean():.1f} ms/step", file=sys.stderr, flush=True)
    l64 = _call_latencies(lambda: model.efs_device(batch, "fp64"), max(1, steps - 1), 1)
    out = dict(workload="md2m", model="Effect L5/K6/R2/O3 body 4, B=10, 2 species", atoms=n, directed_edges=nedges,
               atom_steps_per_s=round(n * steps / l32.sum(), 1), ms_per_step=round(1e3 * l32.mean(), 3),
               fp64_atom_steps_per_s=round(n * len(l64) / l64.sum(), 1),
               fp64_ms_per_step=round(1e3 * l64.mean(), 3), fp32_over_fp64=round(l64.mean() / l32.mean(), 3),
               step_formula_tflops=round(step_efs_per_edge(dims) * nedges / l32.mean() / 1e12, 4))
    if profile:
        model.profile_reset()
        model.profile(True)
        model.efs_device(batch, "fp32")
        import torch
        torch.cuda.synchronize()
        model.profile(False)
        prof = model.profile_read()
        out["kernels_fp32"] = {k: dict(avg_us=round(1e3 * ms / max(c, 1), 2), launches=c) for k, (ms, c) in prof.items()}
        out["roofline"] = _roofline_of(prof, dims, nedges, n, graph_stats(path), "fp32", "md2m")
    return out


def _oracle_ready():
    from oracle import oracle
    try:
        oracle.lib()
    except RuntimeError:
        oracle.build()


def _calibration(workload):
    """reference_over_oracle of a workload from tests/golden/cpu_calibration.json."""
    try:
        cal = json.load(open(os.path.join(ROOT, "tests", "golden", "cpu_calibration.json")))
        return cal["workloads"][workload]["reference_over_oracle"]
    except (OSError, KeyError, ValueError):
        return None


def reference_cpu_baseline(model_path, frames, procs, budget_s=15.0):
    """The reference's embedded EFS engine (built from its own sources into
    oracle/_ref/ref-timing) on the host cores: `procs` independent processes,
    each evaluating the frames repeatedly for about budget_s (throughput mode,
    as independent MPI ranks on copies of the work).  Neighbour lists are
    built before each process's timed region.  None when the binary is absent
    or fails on this host."""
    import subprocess
    import tempfile
    from oracle import oracle
    binary = oracle.ref_timing_binary()
    if binary is None:
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_cpu_calibration import write_frames
    n = sum(len(f["types"]) for f in frames)
    with tempfile.TemporaryDirectory() as tmp:
        fp = os.path.join(tmp, "frames.txt")
        write_frames(fp, frames)
        try:
            one = subprocess.run([str(binary), model_path, fp, "1"], capture_output=True, text=True, check=True,
                                 timeout=300)
            single = json.loads(one.stdout.split("REF_TIMING_JSON ")[1])["seconds_per_rep"]
            reps = max(1, int(budget_s / max(single, 1e-6)))
            ps = [subprocess.Popen([str(binary), model_path, fp, str(reps)], stdout=subprocess.PIPE, text=True)
                  for _ in range(procs)]
            secs = [json.loads(p.communicate(timeout=budget_s * 10 + 300)[0].split("REF_TIMING_JSON ")[1])
                    ["seconds_per_rep"] for p in ps]
        except (OSError, subprocess.SubprocessError, IndexError, ValueError) as e:
            print(f"[bench] reference CPU engine unavailable here: {e}", file=sys.stderr, flush=True)
            return None
    value = sum(n / t for t in secs)
    return dict(value=round(value, 1), unit="atom-steps/s", cores=procs, kind="reference",
                sample=f"{procs} processes x {reps} evaluations of {len(frames)} frame(s) of {n} atoms by the "
                       f"reference embedded engine (interfaces/lammps/ML-SUS2, -O3, direct radial mode, "
                       f"prebuilt edge lists), {max(secs) * reps:.1f} s each")


def oracle_cpu_baseline(model_path, frames, threads, workload, budget_s=15.0):
    """FP64 oracle EFS on the host cores: copies of the frames, one per OpenMP
    thread (frame-parallel), with the reference-equivalent rate from the
    committed single-thread calibration."""
    from oracle import oracle
    _oracle_ready()
    m = oracle.OracleModel(model_path)
    t0 = time.perf_counter()
    m.efs(frames, nthreads=1)
    single = time.perf_counter() - t0
    copies = max(1, int(round(budget_s / max(single, 1e-3))))
    copies = threads * max(1, -(-copies // threads))  # whole rounds of `threads` frames
    t0 = time.perf_counter()
    m.efs(frames * copies, nthreads=threads)
    dt = time.perf_counter() - t0
    n = sum(len(f["types"]) for f in frames)
    out = dict(value=round(n * copies / dt, 1), unit="atom-steps/s", cores=threads, kind="port",
               sample=f"{copies} copies of {len(frames)} frame(s) of {n} atoms, FP64 oracle, OpenMP over copies, "
                      f"{dt:.1f} s")
    r = _calibration(workload)
    if r:
        out["reference_equivalent_value"] = round(out["value"] / r, 1)
        out["calibration"] = f"reference/oracle single-thread time {r:.3f} (tests/golden/cpu_calibration.json)"
    return out


def cpu_baseline(model_path, frames, workload, threads=None):
    """The reference engine on the host cores when its prebuilt binary runs
    here, else the oracle port with its calibrated reference-equivalent rate."""
    threads = threads or min(16, os.cpu_count() or 1)
    ref = reference_cpu_baseline(model_path, frames, threads)
    port = oracle_cpu_baseline(model_path, frames, threads, workload, budget_s=8.0 if ref else 15.0)
    if ref is None:
        return port
    ref["port"] = port
    return ref


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` in `workload` (FETCH_SIZE + WRITE_SIZE,
    each scaled by the copy-kernel calibration) from the committed rocprofv3
    counter summaries (profiles/hbm_traffic_*.md), or None when none covers it."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "hbm_traffic_*.md"))):
        m = re.search(r"```json\n(.*?)\n```", open(path).read(), re.S)
        if not m:
            continue
        v = json.loads(m.group(1)).get(workload, {}).get(kernel)
        if v is not None:
            return int(v)
    return None


def _copy_probe(mib):
    """A plain FP32 copy kernel (y = 2x) over `mib` MiB, three launches: the
    calibration dispatch of the rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    if mib <= 0:
        return
    import torch
    x = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    print(f"[bench] copy probe: {mib} MiB read + {mib} MiB written per launch", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-train", action="store_true", help="skip the training (loss + gradient) workload")
    ap.add_argument("--train-steps", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true", help="skip the 4096-frame uneven batch workload")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--no-md", action="store_true", help="skip the 2,097,152-atom frame workload")
    ap.add_argument("--md-steps", type=int, default=3)
    ap.add_argument("--no-static", action="store_true", help="skip the 64-atom Static FP64 workload")
    ap.add_argument("--static-steps", type=int, default=200)
    ap.add_argument("--only", choices=["effect4096", "static64", "train256", "batch4096", "md2m"],
                    help="run one workload section only (for per-workload profiles); the headline line still "
                         "reports effect4096 unless it is skipped")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed runs, launch a plain FP32 copy kernel (y = 2x) over this many MiB, "
                         "three times: the calibration dispatch for rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path runs on one MI355X); --gpus {args.gpus} is not run",
              file=sys.stderr)
        return 2

    if args.only:
        for k in ("static", "train", "batch", "md"):
            setattr(args, "no_" + k, not args.only.startswith(k))
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.flops import graph_stats, step_efs_per_atom, step_efs_per_edge
    from superlinear_amd.model import DeviceBatch, Sus2Model

    if args.only and args.only != "effect4096":  # one non-headline section (per-workload profiles)
        path = W.model_path("effect_l3k4r2o2")
        sec = {"static64": lambda: static_section(args.static_steps, 20, not args.no_profile, args.no_cpu_baseline),
               "train256": lambda: train_section(path, args.train_steps, 1, not args.no_profile, args.no_cpu_baseline),
               "batch4096": lambda: batch_section(args.batch_steps, 2, not args.no_profile),
               "md2m": lambda: md_section(args.md_steps, 1, not args.no_profile)}[args.only]()
        _copy_probe(args.copy_probe_mib)
        print(json.dumps({"metric": METRIC, "section": args.only, "result": sec}))
        return 0

    path = W.model_path("effect_l3k4r2o2")
    frame = W.config_frames("effect4096")[0]
    n = len(frame["types"])
    model = Sus2Model(path)
    dims = model.dims()
    batch = DeviceBatch([frame])
    print(f"[bench] {n} atoms, model {os.path.basename(path)}, steps={args.steps} warmup={args.warmup}",
          file=sys.stderr, flush=True)

    # launch shape of the per-atom kernels from the autotuner, per precision
    tuned = {}
    for prec in ("fp32", "fp64"):
        best, table = model.autotune(batch, prec, reps=5)
        tuned[prec] = dict(choice=best, median_ms={f"eb{k[0]}_grid{k[1]}": round(v, 4) for k, v in table.items()})
    print(f"[bench] tuned: fp32 {tuned['fp32']['choice']} fp64 {tuned['fp64']['choice']}", file=sys.stderr, flush=True)
    # both precisions timed without the per-kernel timers; a separate profiled
    # run of the same steps gives the kernel averages behind the roofline
    t32, out32, _ = _timed(model, batch, "fp32", args.steps, args.warmup, False)
    nedges = model.last_edge_count()
    t64, out64, _ = _timed(model, batch, "fp64", args.steps, args.warmup, False)
    # the same step captured in a HIP graph: one graph launch per step
    graph_ms = {}
    for prec in ("fp32", "fp64"):
        g = model.graph_efs(batch, prec)
        for _ in range(args.warmup):
            g.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.launch()
        torch.cuda.synchronize()
        graph_ms[prec] = 1e3 * (time.perf_counter() - t0) / args.steps
        assert not g.overflow()
        del g
    prof32 = {} if args.no_profile else _timed(model, batch, "fp32", args.steps, 1, True)[2]
    rate32 = n * args.steps / t32
    rate64 = n * args.steps / t64
    print(f"[bench] fp32 {rate32:.4g} atom-steps/s  fp64 {rate64:.4g} atom-steps/s", file=sys.stderr, flush=True)

    # FP32 / FP64-GPU errors against the FP64 oracle
    e32 = {k: v.cpu().numpy() for k, v in out32.items()}
    e64 = {k: v.cpu().numpy() for k, v in out64.items()}
    errors = {}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        _oracle_ready()
        ref = oracle.OracleModel(path).efs([frame], nthreads=1)
        errors = dict(
            fp64_gpu_energy_rel=float(abs(e64["energy"][0] - ref["energy"][0]) / abs(ref["energy"][0])),
            fp64_gpu_force_maxabs=float(np.abs(e64["forces"] - ref["forces"]).max()),
            fp32_energy_per_atom_abs=float(abs(e32["energy"][0] - ref["energy"][0]) / n),
            fp32_force_maxabs=float(np.abs(e32["forces"] - ref["forces"]).max()),
            fp32_virial_maxabs=float(np.abs(e32["virial"] - ref["virial"]).max()),
            force_scale_maxabs=float(np.abs(ref["forces"]).max()))
        cpu = cpu_baseline(path, [frame], "effect4096")

    graph = graph_stats(path)
    roof = None
    kern = {}
    if prof32:
        for name, (ms, cnt) in prof32.items():
            kern[name] = dict(avg_us=round(1e3 * ms / max(cnt, 1), 2), launches=cnt)
        roof = _roofline_of(prof32, dims, nedges, n, graph, "fp32", "effect4096")
    formula_rate = (step_efs_per_edge(dims) * nedges + step_efs_per_atom(dims, graph) * n) * args.steps / t32 / 1e12
    static64 = None if args.no_static else static_section(args.static_steps, 20, not args.no_profile,
                                                           args.no_cpu_baseline)
    train = {} if args.no_train else train_section(path, args.train_steps, 1, not args.no_profile,
                                                     args.no_cpu_baseline)
    batch4096 = None if args.no_batch else batch_section(args.batch_steps, 2, not args.no_profile)
    md2m = None if args.no_md else md_section(args.md_steps, 1, not args.no_profile)

    _copy_probe(args.copy_probe_mib)

    line = {
        "metric": METRIC,
        "value": round(rate32, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t32 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "effect4096", "atoms": n, "species": 2, "sus2": "Effect L3/K4/R2/O2 body 4, B=10",
                   "cutoff_A": 5.0, "directed_edges": nedges, "neighbors_per_atom": round(nedges / n, 2),
                   "step": "neighbour list + gate pass + outer pass + gather force/virial backward"},
        "launch_shape": tuned,
        "graph_ms_per_step": round(graph_ms["fp32"], 4),
        "graph_atom_steps_per_s": round(n / graph_ms["fp32"] * 1e3, 1),
        "graph_fp64_ms_per_step": round(graph_ms["fp64"], 4),
        "fp64_atom_steps_per_s": round(rate64, 1),
        "fp64_ms_per_step": round(1e3 * t64 / args.steps, 4),
        "fp32_over_fp64": round(rate32 / rate64, 3),
        "train_grad_frames_per_s": train.get("train_grad_frames_per_s"),
        "step_formula_tflops": round(formula_rate, 4),
        "errors_vs_fp64_oracle": errors,
        "kernels_fp32": kern,
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu": torch.cuda.get_device_name(0),
    }
    line.update({k: v for k, v in train.items() if k != "train_grad_frames_per_s"})
    line["static64"] = static64
    line["batch4096"] = batch4096
    line["md2m"] = md2m
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
