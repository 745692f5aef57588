"""Benchmark: backpropagate_3d Rytov reconstruction, 256x256 detector,
360 angles (padded to 512x512), 256^3 output, on one MI355X.

A step is one full backpropagate_3d call (pad, FFT, filter, per-z
propagation and inverse transform, fused prefilter, rotate-accumulate over
all 360 angles) on a complex64 sinogram already resident in HBM, computed
in float32.  Prints one JSON line with, beside the FP32 throughput:
the FP64 GPU path; the rotate-accumulate kernel's HBM roofline; the CPU
baseline (the oracle on all host cores, in full, with a bounded-sample
extrapolation beside it) and the FP32/FP64 error against its complex128
volume; the 64^2 x 100 workload (GPU and CPU); the prologue and epilogue
of the reference's pipeline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                    [--cpu-sample-only] [--no-fp64] [--no-c1]
"""
import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("voxel-angle updates/s (256² det, 360 angles) FP32 vs FP64 GPU "
          "at ≤1e-5 rel L2")
N_DET, N_ANG = 256, 360
RES, NM = 6.0, 1.333
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_source_hash():
    """sha256 (first 16 hex digits) of the HIP sources the rotate-accumulate
    kernel is built from; the committed traffic summary carries the hash of
    the sources it was measured on."""
    import hashlib
    h = hashlib.sha256()
    for name in ("odt_kernels.hip", "odt_device.h", "odt_launch.h"):
        h.update((ROOT / "odtbrain_amd" / "csrc" / name).read_bytes())
    return h.hexdigest()[:16]


def traffic_from_profile():
    """HBM bytes per rotate-accumulate launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), read
    from the committed summary; None when the summary is missing or was
    measured on other kernel sources than the ones built here."""
    f = ROOT / "profiles" / "rotate_acc_c2_traffic.txt"
    if not f.exists():
        return None, "no traffic summary"
    vals = {}
    for line in f.read_text().splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    src = vals.get("kernel_sources_sha256_16")
    if src != kernel_source_hash():
        return None, "traffic summary measured on other kernel sources"
    return float(vals["traffic_bytes_per_launch"]), f"profiles/{f.name}"


def cpu_cores():
    """Host cores available to this process (the CPU share of the GPU host;
    os.cpu_count() and the affinity mask agree there)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(uB, angles, sample_angles, full):
    """Oracle (CPU restatement of the reference, float64, y-chunk process
    pool over all host cores + FFTs threaded over the same cores), timed per
    call including pool and filter set-up as the reference runs: first a
    bounded contiguous sample of ``sample_angles`` angles extrapolated
    linearly, then (``full``) the whole reconstruction, whose output also
    serves as the accuracy reference.  Returns (baseline dict, volume or
    None)."""
    from oracle import odt_oracle
    cores = cpu_cores()
    kw = dict(padval="edge", intp_order=2, num_cores=cores)
    uB = uB.astype(np.complex128)
    V = N_DET ** 3
    stats = {}
    odt_oracle.backpropagate_3d(uB, angles, RES, NM,
                                angle_range=(0, sample_angles), stats=stats,
                                **kw)
    per_angle = stats["loop_s"] / stats["angles"]
    t_sample = stats["setup_s"] + per_angle * len(angles)
    sample = {"value": V * len(angles) / t_sample,
              "angles": stats["angles"], "loop_s": stats["loop_s"],
              "setup_s": stats["setup_s"],
              "extrapolated_s": t_sample}
    ref = None
    if full:
        t0 = time.perf_counter()
        ref = odt_oracle.backpropagate_3d(uB, angles, RES, NM, **kw)
        t_full = time.perf_counter() - t0
        out = {"value": V * len(angles) / t_full,
               "unit": "voxel-angle updates/s", "cores": cores,
               "kind": "port",
               "sample": (f"oracle float64 (reference structure: Pool({cores}) "
                          f"y-chunk rotation, FFTs on {cores} threads), the "
                          f"whole 256^2 x 360 reconstruction: {t_full:.1f} s "
                          f"per call; bounded sample of {stats['angles']} "
                          f"angles extrapolated: {t_sample:.1f} s"),
               "full_s": t_full, "bounded_sample": sample}
    else:
        out = {"value": sample["value"], "unit": "voxel-angle updates/s",
               "cores": cores, "kind": "port",
               "sample": (f"oracle float64, {stats['angles']} of "
                          f"{len(angles)} angles ({stats['loop_s']:.1f} s) + "
                          f"setup {stats['setup_s']:.1f} s, extrapolated: "
                          f"{t_sample:.1f} s per reconstruction")}
    return out, ref


def c1_workload(cpu):
    """The CPU-runnable workload (64^2 detector, 100 angles, centred sphere
    r=16 px n=1.343, Rytov data with noise 0.01, edge padding to 128^2,
    order 2): GPU FP32/FP64 time per call and error against the oracle,
    the oracle run in full on the host cores beside it."""
    import torch
    import odtbrain_amd
    from odtbrain_amd import phantoms
    from oracle import odt_oracle
    uB, angles = phantoms.sphere_sinogram_3d()
    kw = dict(padval="edge", intp_order=2)
    d64 = torch.from_numpy(uB).cuda()
    d32 = d64.to(torch.complex64)
    res = {}
    for name, d, dt in (("fp32", d32, np.float32), ("fp64", d64, np.float64)):
        for _ in range(2):
            out = odtbrain_amd.backpropagate_3d(d, angles, RES, NM, dtype=dt,
                                                **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            out = odtbrain_amd.backpropagate_3d(d, angles, RES, NM, dtype=dt,
                                                **kw)
        torch.cuda.synchronize()
        res[name] = ((time.perf_counter() - t0) / 10, out.cpu().numpy())
    V = 64 ** 3 * len(angles)
    line = {"workload": "backpropagate_3d Rytov, 64x64 detector, 100 angles, "
                        "sphere r=16 n=1.343, padded to 128x128, 64^3",
            "fp32_ms": res["fp32"][0] * 1e3,
            "fp32_value": V / res["fp32"][0],
            "fp64_ms": res["fp64"][0] * 1e3,
            "fp64_value": V / res["fp64"][0]}
    if cpu:
        cores = cpu_cores()
        t0 = time.perf_counter()
        ref = odt_oracle.backpropagate_3d(uB, angles, RES, NM,
                                          num_cores=cores, **kw)
        t_cpu = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": V / t_cpu, "s": t_cpu,
                                "cores": cores, "kind": "port",
                                "sample": "oracle float64, whole call"}
        for name in ("fp32", "fp64"):
            g = res[name][1].astype(np.complex128)
            line[name + "_rel_l2_vs_oracle"] = float(
                np.linalg.norm(g - ref) / np.linalg.norm(ref))
            line[name + "_max_abs_vs_oracle"] = float(np.abs(g - ref).max())
    return line


def batch_bench(steps, warmup, lanes=4):
    """The batched workload (48 sinograms of uneven sizes) replayed as one
    HIP graph; aggregate voxel-angle updates/s (2D: N^2 per angle, 3D: N^3
    per angle), the eager sequential time beside it, and per-instance
    FP32 error against the FP64 GPU path."""
    import torch
    import odtbrain_amd as G
    from odtbrain_amd import phantoms
    fns = {"2d": G.backpropagate_2d, "3d": G.backpropagate_3d}
    log("bench: generating the 48 batched sinograms")
    raw = phantoms.batch_jobs()
    dev = torch.device("cuda:0")
    jobs, work = [], 0
    for kind, s, ang, res, nm in raw:
        d = torch.from_numpy(s).to(dev)
        jobs.append((fns[kind], (d, ang, res, nm), {"dtype": np.float32}))
        N = s.shape[-1]
        work += (N ** 2 if kind == "2d" else N ** 3) * len(ang)
    log(f"bench: capturing the batch graph ({lanes} lanes)")
    b = G.Batch(jobs, lanes=lanes)
    for _ in range(warmup):
        b.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        outs = b.run()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for fn, a, kw in jobs:
        fn(*a, **kw)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    errs = []
    for (fn, a, kw), o in zip(jobs, outs):
        ref = fn(*a, dtype=np.float64)
        errs.append(float(torch.linalg.vector_norm(o.to(ref.dtype) - ref) /
                          torch.linalg.vector_norm(ref)))
    line = {
        "metric": "batched voxel-angle updates/s (48 sinograms, one graph)",
        "value": work / t_graph, "unit": "voxel-angle updates/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": t_graph * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded disk/sphere phantoms, complex64)",
        "config": {"workload": "32 backpropagate_2d (N in 128..1024, "
                               "90-720 angles) + 16 backpropagate_3d "
                               "(N 64-160, 60-300 angles)"},
        "lanes": lanes,
        "eager_sequential_ms": t_eager * 1e3,
        "fp32_rel_l2_vs_fp64_max": max(errs),
        "fp32_rel_l2_vs_fp64_per_instance": errs,
    }
    print(json.dumps(line), flush=True)


def _tv(f, eps=1e-3):
    """Isotropic total variation of a complex volume (forward differences,
    smoothed by eps)."""
    import torch
    d0 = f[1:, :-1, :-1] - f[:-1, :-1, :-1]
    d1 = f[:-1, 1:, :-1] - f[:-1, :-1, :-1]
    d2 = f[:-1, :-1, 1:] - f[:-1, :-1, :-1]
    return torch.sum(torch.sqrt(d0.abs() ** 2 + d1.abs() ** 2 +
                                d2.abs() ** 2 + eps))


def adjoint_bench(steps, warmup, directions=64):
    """Differentiable backpropagation, 128x128 detector, 180 angles, FP32:
    gradient descent on ||BP(u) - f_target||^2 + 1e-3 TV(BP(u)) from u = 0
    with f_target the FP32 reconstruction of a seeded 4-ellipsoid phantom.
    ``steps`` (default 50) descent steps with a fixed step 1/L (L from a
    power iteration on 2 M^H M, untimed) are timed; the gradient is checked
    against FP64 central differences on 64 seeded directions (eps=1e-6) and
    the dot-product test <g, M u> = <M^H g, u> in FP64 and FP32."""
    import torch
    import odtbrain_amd as G
    from odtbrain_amd import autograd as AG, phantoms
    dev = torch.device("cuda:0")
    N, A, lam = 128, 180, 1e-3
    log("bench: generating the 128^2 x 180 four-ellipsoid sinogram")
    sino, angles = phantoms.ellipsoid_sinogram_3d(
        N=N, A=A, count=4, semi=(10, 30), spread=15.0, seed=5)
    d_sino = torch.from_numpy(sino).to(dev)
    f_t = G.backpropagate_3d(d_sino, angles, RES, NM, dtype=np.float32)

    def loss(u, dtype, target):
        f = AG.backpropagate_3d_torch(u, angles, RES, NM, dtype=dtype)
        return torch.sum((f - target).abs() ** 2) + lam * _tv(f)

    # step size from the largest eigenvalue of the data-term Hessian
    v = torch.randn(A, N, N, dtype=torch.complex64, device=dev)
    for _ in range(12):
        v = v / torch.linalg.vector_norm(v)
        Mv = G.backpropagate_3d(v, angles, RES, NM, dtype=np.float32)
        w = 2 * AG.backpropagate_3d_adjoint(Mv, angles, RES, NM,
                                            dtype=np.float32)
        L_est = float(torch.linalg.vector_norm(w))
        v = w
    t = 1.0 / (1.05 * L_est)

    u = torch.zeros(A, N, N, dtype=torch.complex64, device=dev,
                    requires_grad=True)
    losses = []

    def descend(n):
        for _ in range(n):
            L = loss(u, np.float32, f_t)
            u.grad = None
            L.backward()
            with torch.no_grad():
                u.sub_(t * u.grad)
            losses.append(L.detach())

    descend(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    descend(steps)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    losses = [float(x) for x in losses]

    log("bench: FP64 central differences on %d directions" % directions)
    u64 = u.detach().to(torch.complex128)
    f_t64 = f_t.to(torch.complex128)
    x = u64.clone().requires_grad_(True)
    loss(x, np.float64, f_t64).backward()
    g = x.grad.detach()
    gen = torch.Generator(device=dev).manual_seed(7)
    eps, fd_err = 1e-6, []
    with torch.no_grad():
        for _ in range(directions):
            d = torch.complex(
                torch.randn(A, N, N, dtype=torch.float64, device=dev,
                            generator=gen),
                torch.randn(A, N, N, dtype=torch.float64, device=dev,
                            generator=gen))
            lp = float(loss(u64 + eps * d, np.float64, f_t64))
            lm = float(loss(u64 - eps * d, np.float64, f_t64))
            fd = (lp - lm) / (2 * eps)
            an = float(torch.sum(torch.real(torch.conj(g) * d)))
            fd_err.append(abs(fd - an) / max(abs(an), 1e-300))

    dots = {}
    for name, dt_np, ct in (("fp64", np.float64, torch.complex128),
                            ("fp32", np.float32, torch.complex64)):
        a = torch.randn(A, N, N, dtype=ct, device=dev, generator=gen)
        b = torch.randn(N, N, N, dtype=ct, device=dev, generator=gen)
        Ma = G.backpropagate_3d(a, angles, RES, NM, dtype=dt_np)
        MHb = AG.backpropagate_3d_adjoint(b, angles, RES, NM, dtype=dt_np)
        lhs = complex(torch.sum(torch.conj(b).to(torch.complex128) *
                                Ma.to(torch.complex128)))
        rhs = complex(torch.sum(torch.conj(MHb).to(torch.complex128) *
                                a.to(torch.complex128)))
        dots[name] = abs(lhs - rhs) / abs(lhs)

    line = {
        "metric": "gradient-descent steps/s (differentiable backpropagation, "
                  "128² det, 180 angles)",
        "value": 1.0 / dt, "unit": "steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 4-ellipsoid Rytov phantom, complex64)",
        "config": {"workload": "||BP(u)-f_target||^2 + 1e-3 TV(BP(u)), "
                               "128x128 detector, 180 angles, 128^3 volume"},
        "loss_first": losses[0], "loss_last": losses[-1],
        "fd_directions": directions, "fd_eps": eps,
        "fd_rel_err_max": max(fd_err),
        "fd_rel_err_median": float(np.median(fd_err)),
        "dot_test_rel_err": dots,
    }
    print(json.dumps(line), flush=True)


def apple_bench(max_iter):
    """apple.correct (non-negativity, envelope 0.95) of the 256^3 FP32
    reconstruction of the 256^2 x 360 workload, ``max_iter`` iterations
    (min_diff 0 so that all run); time per call and per iteration."""
    import torch
    import odtbrain_amd as G
    from odtbrain_amd import phantoms
    dev = torch.device("cuda:0")
    uB, angles = phantoms.ellipsoid_sinogram_3d(N=N_DET, A=N_ANG)
    f = G.backpropagate_3d(torch.from_numpy(uB).to(dev), angles, RES, NM,
                           dtype=np.float32)
    G.apple.correct(f, RES, NM, max_iter=2, min_diff=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = G.apple.correct(f, RES, NM, max_iter=max_iter, min_diff=0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    line = {
        "metric": "apple-core correction iterations/s (256^3 volume)",
        "value": max_iter / dt, "unit": "iterations/s", "n_gpus": 1,
        "steps": max_iter, "warmup": 2, "ms_per_step": dt / max_iter * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (FP32 reconstruction of the "
                                "3-ellipsoid 256^2 x 360 sinogram)",
        "config": {"workload": "apple.correct method nn, enforce_envelope "
                               "0.95, 256^3 complex, transforms in complex128"},
        "total_ms": dt * 1e3,
        "finite": bool(torch.isfinite(out.real).all()),
    }
    print(json.dumps(line), flush=True)


def tilted_bench(steps, warmup):
    """backpropagate_3d_tilted, 192x192 detector, 240 non-uniform angles over
    1.4 pi, axis tilted 0.2 rad about x, weight_angles, order 2, complex64
    (5 seeded spheres); FP32 and FP64 on the GPU, FP32 error against FP64.
    Voxel-angle updates/s = 192^3 * 240 / time."""
    import torch
    import odtbrain_amd as G
    from odtbrain_amd import _core, phantoms
    dev = torch.device("cuda:0")
    log("bench: generating the tilted 192^2 x 240 sinogram")
    uB, angles, axis = phantoms.tilted_sphere_sinogram()
    d_in = torch.from_numpy(uB).to(dev)

    def step(dtype):
        return G.backpropagate_3d_tilted(d_in, angles, RES, NM,
                                         tilted_axis=axis, dtype=dtype)
    for _ in range(warmup):
        step(np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out32 = step(np.float32)
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / steps
    plan = _core.last_plan
    plan.enable_timing(True)
    step(np.float32)
    tm = plan.last_timing()
    plan.enable_timing(False)
    out64 = step(np.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out64 = step(np.float64)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    err = float(torch.linalg.vector_norm(out32.to(torch.complex128) - out64) /
                torch.linalg.vector_norm(out64))
    V = 192 ** 3 * len(angles)
    line = {
        "metric": "voxel-angle updates/s (tilted axis, 192² det, 240 angles)",
        "value": V / t32, "unit": "voxel-angle updates/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": t32 * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded 5-sphere Rytov phantom, "
                                "complex64)",
        "config": {"workload": "backpropagate_3d_tilted, 192x192 detector, "
                               "240 jittered angles over 1.4 pi, axis tilted "
                               "0.2 rad, weight_angles, order 2"},
        "affine_stage_ms": tm["rotate_ms"],
        "fp64_gpu": {"ms_per_step": t64 * 1e3, "value": V / t64,
                     "fp32_rel_l2_vs_fp64": err},
    }
    print(json.dumps(line), flush=True)


def example_bench(path, steps, warmup):
    """The reference's 3D example pipeline on one of its example archives
    (a local ``.tar.lzma`` file, e.g. fdtd_3d_sino_A180_R6.500.tar.lzma):
    sinogram_as_rytov -> backpropagate_3d (res, nm, lD from the archive) ->
    odt_to_ri, inputs resident on the GPU; FP32 and FP64, the FP32
    refractive index against FP64 and both against the archive's phantom."""
    import torch
    import odtbrain_amd as G
    from odtbrain_amd import example_data
    dev = torch.device("cuda:0")
    log(f"bench: loading {path}")
    sino, angles, phantom, cfg = example_data.load_data(path)
    res, nm, lD = cfg["res"], cfg["nm"], cfg.get("lD", 0.0)
    d64 = torch.from_numpy(np.ascontiguousarray(sino)).to(dev)
    d32 = d64.to(torch.complex64)

    def step(d, dtype):
        uR = G.sinogram_as_rytov(d)
        f = G.backpropagate_3d(uR, angles, res, nm, lD=lD, dtype=dtype)
        return G.odt_to_ri(f, res, nm)
    for _ in range(warmup):
        step(d32, np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ri32 = step(d32, np.float32)
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / steps
    ri64 = step(d64, np.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ri64 = step(d64, np.float64)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    err = float(torch.linalg.vector_norm(ri32.to(torch.complex128) - ri64) /
                torch.linalg.vector_norm(ri64))
    A, N = sino.shape[0], sino.shape[-1]
    vs_phantom = None
    if phantom.shape == tuple(ri32.shape):
        ph = torch.from_numpy(phantom).to(dev)
        vs_phantom = {"fp32": float((ri32.real.double() - ph).abs().mean()),
                      "fp64": float((ri64.real - ph).abs().mean())}
    line = {
        "metric": "example pipeline (sinogram_as_rytov + backpropagate_3d + "
                  "odt_to_ri) time", "value": t32 * 1e3, "unit": "ms",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": t32 * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"example archive {pathlib.Path(path).name} ({A} x {N}^2)",
        "config": {"workload": "reference example pipeline",
                   "res": res, "nm": nm, "lD": lD},
        "voxel_angle_per_s": N ** 3 * A / t32,
        "fp64_gpu": {"ms_per_step": t64 * 1e3,
                     "fp32_ri_rel_l2_vs_fp64": err},
        "ri_vs_phantom_mean_abs": vs_phantom,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-angles", type=int, default=36,
                    help="angles of the bounded CPU-baseline sample")
    ap.add_argument("--cpu-sample-only", action="store_true",
                    help="report the bounded sample only (no full oracle run, "
                         "no error against the oracle)")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 64^2 x 100 workload block")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--batch", action="store_true",
                    help="run the batched 48-sinogram workload instead")
    ap.add_argument("--batch-lanes", type=int, default=4,
                    help="concurrent graph branches of --batch")
    ap.add_argument("--adjoint", action="store_true",
                    help="run the differentiable (gradient-descent) workload "
                         "instead; --steps defaults to 50 there")
    ap.add_argument("--tilted", action="store_true",
                    help="run the tilted-axis workload (192^2, 240 angles)")
    ap.add_argument("--apple", action="store_true",
                    help="run apple.correct on the 256^3 reconstruction "
                         "instead (--steps iterations, default 100)")
    ap.add_argument("--example", metavar="PATH",
                    help="run the reference's 3D example pipeline on a local "
                         "example archive (.tar.lzma) instead")
    ap.add_argument("--fd-directions", type=int, default=64,
                    help="finite-difference directions of --adjoint")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device: a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "multi-GPU angle sharding is a design (DESIGN.md) "
                          "and is not benchmarked"}))
        sys.exit(2)

    if args.example:
        example_bench(args.example, args.steps, args.warmup)
        return
    if args.batch:
        batch_bench(args.steps, args.warmup, args.batch_lanes)
        return
    if args.tilted:
        tilted_bench(args.steps, args.warmup)
        return
    if args.apple:
        apple_bench(100 if args.steps == 5 else args.steps)
        return
    if args.adjoint:
        adjoint_bench(50 if args.steps == 5 else args.steps, args.warmup,
                      args.fd_directions)
        return

    import torch
    import odtbrain_amd
    from odtbrain_amd import _core, phantoms

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(2 ** 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log("bench: generating 256^2 x 360 ellipsoid Rytov sinogram")
    uB, angles = phantoms.ellipsoid_sinogram_3d(N=N_DET, A=N_ANG)
    d_in = torch.from_numpy(uB).to(dev)
    kw = dict(res=RES, nm=NM, padval="edge", intp_order=2)

    def step(dtype):
        return odtbrain_amd.backpropagate_3d(d_in, angles, dtype=dtype, **kw)

    log("bench: warmup")
    for _ in range(args.warmup):
        out32 = step(np.float32)
    torch.cuda.synchronize()
    log("bench: timed steps fp32")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out32 = step(np.float32)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    V = N_DET ** 3
    value = V * N_ANG / t_step

    # Rotate-accumulate time from HIP events on the plan's stream.  The stage
    # runs as several launches (angle groups).  Algorithmic HBM bytes per
    # voxel and angle: the complex64 coefficient sample read once (8 B) plus
    # the final complex64 volume written once per reconstruction, spread
    # over the angles (8/A B); a launch of B angles is credited with
    # B * V * (8 + 8/A) bytes.
    plan = _core.last_plan
    plan.enable_timing(True)
    rot, tot = [], []
    for _ in range(3):
        step(np.float32)
        tm = plan.last_timing()
        rot.append(tm["rotate_ms"])
        tot.append(tm["total_ms"])
    plan.enable_timing(False)
    launches = tm["rotate_launches"]
    rot_stage_ms = float(np.mean(rot))
    rot_ms = rot_stage_ms / launches
    ang_per_launch = N_ANG / launches
    alg_bytes = ang_per_launch * V * (8 + 8 / N_ANG)
    achieved = alg_bytes / (rot_ms * 1e-3) / 1e9
    log(f"bench: fp32 {t_step * 1e3:.1f} ms/step, rotate {rot_stage_ms:.2f} ms "
        f"in {launches} launches, plan total {np.mean(tot):.1f} ms")

    # Prologue before the path: sinogram_as_rytov of the complex field
    # exp(uB) of the same workload (odtbrain/_prepare_sino.py:105-182:
    # phase and log, per-angle 2D unwrapping, alignment), complex64.
    field = torch.exp(d_in)
    odtbrain_amd.sinogram_as_rytov(field)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        ryt = odtbrain_amd.sinogram_as_rytov(field)
    torch.cuda.synchronize()
    prologue = {"function": "sinogram_as_rytov",
                "ms": (time.perf_counter() - t0) / 2 * 1e3,
                "phase_max_abs_vs_uB": float((ryt.imag - d_in.imag).abs().max())}
    del field, ryt

    # Epilogue after the path: odt_to_ri of the 256^3 complex64 volume
    # (odtbrain/_translate_ri.py:5-57), one elementwise HBM stream, timed
    # with events on the stream it runs on; 16 B per voxel (read + write).
    ri = odtbrain_amd.odt_to_ri(out32, RES, NM)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), \
        torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(20):
        ri = odtbrain_amd.odt_to_ri(out32, RES, NM)
    ev1.record()
    torch.cuda.synchronize()
    ri_ms = ev0.elapsed_time(ev1) / 20
    ri_gbs = 16 * V / (ri_ms * 1e-3) / 1e9
    epilogue = {"function": "odt_to_ri", "kernel": "k_odt_to_ri_c",
                "ms": ri_ms, "achieved_GBps": ri_gbs,
                "frac": ri_gbs / HBM_PEAK_GBS,
                "algorithmic_bytes": 16 * V}
    del ri

    fp64, out64 = None, None
    if not args.no_fp64:
        log("bench: fp64 path")
        out64 = step(np.float64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n64 = max(1, args.steps // 2)
        for _ in range(n64):
            out64 = step(np.float64)
        torch.cuda.synchronize()
        t64 = (time.perf_counter() - t0) / n64
        a = out32.to(torch.complex128)
        err = float(torch.linalg.vector_norm(a - out64) /
                    torch.linalg.vector_norm(out64))
        maxabs = float((a - out64).abs().max())
        fp64 = {"value": V * N_ANG / t64, "ms_per_step": t64 * 1e3,
                "fp32_rel_l2_vs_fp64": err, "fp32_max_abs_vs_fp64": maxabs}
        log(f"bench: fp64 {t64 * 1e3:.1f} ms/step, fp32 vs fp64 rel_l2 "
            f"{err:.2e}")

    c1 = None
    if not args.no_c1:
        log("bench: 64^2 x 100 workload")
        c1 = c1_workload(cpu=not args.no_cpu_baseline)

    cpu, accuracy = None, None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle on all host cores)")
        cpu, ref = cpu_baseline(uB, angles, args.cpu_sample_angles,
                                full=not args.cpu_sample_only)
        log(f"bench: cpu {cpu['value']:.3e} voxel-angle/s ({cpu['sample']})")
        if ref is not None:
            accuracy = {}
            for name, g in (("fp32", out32), ("fp64", out64)):
                if g is None:
                    continue
                g = g.cpu().numpy().astype(np.complex128)
                accuracy[name + "_rel_l2_vs_oracle"] = float(
                    np.linalg.norm(g - ref) / np.linalg.norm(ref))
                accuracy[name + "_max_abs_vs_oracle"] = float(
                    np.abs(g - ref).max())
            accuracy["oracle_max_abs"] = float(np.abs(ref).max())
            log(f"bench: vs oracle {accuracy}")
            del ref

    traffic, traffic_src = traffic_from_profile()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-angle updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 3-ellipsoid Rytov phantom, complex64)",
        "config": {"workload": "backpropagate_3d Rytov, 256x256 detector, "
                               "360 angles uniform on [0,2pi), padded to "
                               "512x512, output 256^3 complex64",
                   "res": RES, "nm": NM, "intp_order": 2, "padval": "edge"},
        "roofline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_rotate_lds2", "ms_per_launch": rot_ms,
                     "launches_per_step": launches,
                     "angles_per_launch": ang_per_launch,
                     "bytes_per_voxel_angle": 8 + 8 / N_ANG,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "accuracy_vs_oracle": accuracy,
        "c1": c1,
        "fp64_gpu": fp64,
        "prologue": prologue,
        "epilogue": epilogue,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
