"""Benchmark: Siddon forward projection + matched adjoint at 512^3 x 720 views.

One step = y = A x (ap_project_voxel) and g = A^T y (ap_backproject_matched)
on a 512^3 volume (voxel 0.48 mm, 30 seeded random ellipsoids), 720 views of a
circular cone-beam scan (SID 500 mm, SDD 1000 mm), 768x768 detector at 1 mm.
Inputs are resident in HBM before timing starts.

Reported: GUPS = 2 N^3 V / (t_A + t_AT) / 1e9 (whole job), the roofline of
the dominant kernel against MI355X's 8 TB/s HBM using the per-view streaming
model B_op = V (N^3 s_v + nu nv s_p) bytes per operator, SIRT seconds per
iteration on the 256^3 x 360 x 384^2 workload, and the FP32 OpenMP CPU port of
the reference semantics timed on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--volume-dtype f32|bf16]
    python bench.py --calibrate-copy     # plain 1 GiB device copies, for
                                         # scaling rocprofv3 FETCH/WRITE_SIZE
    python bench.py --analytic-only      # only the analytic FORBILD leg

roofline.traffic (HBM bytes per launch of the dominant kernel) is read from
profiles/hbm_traffic_siddon_512cube_720views.txt, the summary of two
rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) scaled by the same counters
read for --calibrate-copy; it is null when that file is absent.  With it,
roofline.measured_gbs = traffic / kernel time.  roofline.valu_issue_frac is
SQ_INSTS_VALU per launch (profiles/sq_instruction_mix_siddon_512cube_720views.md)
over the VALU issue capacity in the kernel's time (one wave64 VALU
instruction per CU per cycle, 256 CUs at 2.4 GHz), roofline.lds_active_frac
the LDS-active cycles (SQ_LDS_IDX_ACTIVE) over the same cycles, and
roofline.bound names the largest of the three fractions ("hbm", "valu" or
"lds"); with no counter file the bound stays the per-view model's "hbm".
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "GUPS fwd+adjoint proj at 512\u00b3\u00d7720 views"
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "hbm_traffic_siddon_512cube_720views.txt")
SQ_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                       "sq_instruction_mix_siddon_512cube_720views.md")
# VALU issue capacity: each CU issues one wave64 VALU instruction per cycle
# (4 SIMDs, 4 cycles each); MI355X: 256 CUs, 2.4 GHz peak engine clock
VALU_ISSUE_PER_S = 256 * 2.4e9


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_volume(n, w, n_ellipsoids, seed, dtype):
    from aprict_amd import phantoms as ph
    from aprict_amd.volume import FBVolume, Volume

    shapes = ph.random_ellipsoids(n_ellipsoids, n * w / 2, seed)
    v = Volume.zeros((n, n, n), voxwid=w, dtype=np.float32)
    FBVolume(shapes).voxelise(v)
    if dtype == "bf16":
        vb = Volume.zeros((n, n, n), voxwid=w, dtype="bfloat16", fill=False)
        vb.data.copy_(v.data)
        return vb
    return v


FP64_PEAK_TFLOPS = 78.6
FP64_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "fp64_mix_forbild_cfg3.txt")


def read_fp64_flops(kernel: str):
    """FP64 flops per launch of ``kernel`` from the committed SQ-counter summary, or None."""
    try:
        with open(FP64_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "fp64_flops_per_launch" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        return None
    return None


def analytic_leg(reps: int = 3) -> dict:
    """Analytic FORBILD projection (ap_project_forbild) of 40 random clipped
    primitives, 720 views x 512^2 at 0.5 mm: ray-primitive tests per second and,
    with the FP64 instruction counts of the committed profile, the fraction of the
    FP64 vector peak."""
    import torch

    from aprict_amd import phantoms as ph
    from aprict_amd.cb import flatpanel as cbfp
    from aprict_amd.volume import FBVolume

    c = ph.CONFIGS[3]
    V, nu, nv, k = c["views"], c["nu"], c["nv"], c["n_primitives"]  # noqa: N806
    fb = FBVolume(ph.random_primitives(k, c["scale"], c["seed"]))
    pg = ph.circular_geometry(V, nu, nv, c["pitch"])
    ps = cbfp.ProjectionSet(pg, data=torch.zeros((V, nu, nv), dtype=torch.float32, device="cuda"))
    cfg = cbfp.ConfigProject()
    stream = torch.cuda.current_stream()
    cbfp.project(cfg, fb, ps)   # warm-up (uploads the shapes and specs)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        ps.data.zero_()
        ev = time_events(lambda: cbfp.project(cfg, fb, ps), stream)
        torch.cuda.synchronize()
        times.append(ev[0].elapsed_time(ev[1]) / 1e3)
    t = min(times)
    rays = V * nu * nv
    out = {"workload": f"{k} FORBILD primitives (ellipsoids, elliptic cylinders, boxes, 0-3 clip planes), "
                       f"{V} views, {nu}x{nv} detector at {c['pitch']} mm", "ms": t * 1e3,
           "rays_per_s": rays / t, "ray_primitive_tests_per_s": rays * k / t}
    flops = read_fp64_flops("k_project_forbild")
    if flops is not None:
        out["fp64"] = {"achieved": flops / t / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": flops / t / 1e12 / FP64_PEAK_TFLOPS,
                       "model": "(2 FMA_F64 + ADD_F64 + MUL_F64 + TRANS_F64) x 64 per launch from rocprofv3 SQ counters"}
    return out


def time_events(fn, stream):
    import torch

    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    fn()
    e1.record(stream)
    return e0, e1


def fdk_leg(reps: int = 3) -> dict:
    """FDK on the 256^3 x 360 x 384^2 workload: cosine weight + Ram-Lak filter
    (FBPFilterer, FDK mode) and the direct ray-casting backprojection with
    weight (Z/R)^-2, timed separately; error of the central slice against the
    phantom for information."""
    import torch

    from aprict_amd import phantoms as ph
    from aprict_amd.autograd import Projector
    from aprict_amd.cb import flatpanel as cbfp
    from aprict_amd.cb.ramp_filter import FBPFilterer
    from aprict_amd.volume import Volume

    c = ph.CONFIGS[2]
    n, w, V = c["N"], c["voxwid"], c["views"]  # noqa: N806
    R, L = 500.0, 1000.0  # noqa: N806
    pg = ph.circular_geometry(V, c["nu"], c["nv"], c["pitch"], R=R, L=L)
    truth = make_volume(n, w, c["n_ellipsoids"], c["seed"], "f32")
    b = Projector(pg, vol_shape=truth.shape, voxwid=w).A(truth.data)
    filt = FBPFilterer(L=L, physpixwid=c["pitch"], proj_2d_shape=(c["nu"], c["nv"]), _hacky_fdk_mode=True,
                       limit_zeta=0.0)
    cfg = cbfp.ConfigBackproject(fast_approx=True, cyl_mask_rad=n * w / 2, Z_on_val_factor_power=-2,
                                 Z_on_val_factor_val=R, approx_mode_direct_raycast=True,
                                 approx_mode_bilinear_interpolate=True)
    vol = Volume.zeros(truth.shape, voxwid=w)
    stream = torch.cuda.current_stream()
    # the backprojector's launch shape for this problem, timed on a scratch volume
    from aprict_amd import _abi
    from aprict_amd.autotune import Autotuner

    tuner = Autotuner(repeats=2)
    ps_b = cbfp.ProjectionSet(pg, data=b)
    scratch = Volume.zeros(truth.shape, voxwid=w)
    key = ("fdk", n, V, c["nu"], c["nv"])
    tuner.tune(_abi.TUNE_BACKPROJECT, key, lambda: cbfp.backproject(cfg, scratch, ps_b))
    del scratch, ps_b
    tf, tb = [], []
    for _ in range(reps + 1):
        vol.data.zero_()
        holder = {}
        ef = time_events(lambda: holder.update(f=filt.filter_bulk_axis1_concat(b, block=False)[0][0]), stream)
        ps = cbfp.ProjectionSet(pg, data=holder["f"])
        eb = time_events(lambda: cbfp.backproject(cfg, vol, ps), stream)
        torch.cuda.synchronize()
        tf.append(ef[0].elapsed_time(ef[1]) / 1e3)
        tb.append(eb[0].elapsed_time(eb[1]) / 1e3)
    vol.multiply(1.0 / (4.0 * np.pi * R * V))
    t_f, t_b = min(tf[1:]), min(tb[1:])
    mid = n // 2
    got, want = vol.data[:, :, mid].double(), truth.data[:, :, mid].double()
    return {"workload": f"{n}^3 ({w} mm), {V} views, {c['nu']}x{c['nv']}, cosine weight + Ram-Lak + direct "
                        "ray-casting backprojection (Z/R)^-2", "ms_filter": t_f * 1e3, "ms_backproject": t_b * 1e3,
            "ms_total": (t_f + t_b) * 1e3, "gups_backproject": float(n) ** 3 * V / t_b / 1e9,
            "rel_l2_central_slice_vs_phantom": float((got - want).norm() / want.norm()),
            "tuned": {"backproject": {"variant": tuner.best(_abi.TUNE_BACKPROJECT, key),
                                      "times_ms": tuner.times_ms(_abi.TUNE_BACKPROJECT, key)}}}


def sirt_seconds_per_iter(iters: int = 20) -> dict:
    import torch

    from aprict_amd import phantoms as ph
    from aprict_amd.autograd import Projector
    from aprict_amd.cb import flatpanel as cbfp
    from aprict_amd.iterators import SIRT_flatpanel
    from aprict_amd.volume import Volume

    from aprict_amd import _abi
    from aprict_amd.autotune import Autotuner

    c = ph.CONFIGS[2]
    pg = ph.circular_geometry(c["views"], c["nu"], c["nv"], c["pitch"])
    truth = make_volume(c["N"], c["voxwid"], c["n_ellipsoids"], c["seed"], "f32")
    P = Projector(pg, vol_shape=truth.shape, voxwid=c["voxwid"])
    b = P.A(truth.data)
    vol = Volume.zeros(truth.shape, voxwid=c["voxwid"])
    # launch shapes of the forward and the backprojector picked by the autotuner
    # on scratch buffers of this problem
    tuner = Autotuner(repeats=2)
    scratch_p = torch.zeros_like(b)
    scratch_v = Volume.zeros(truth.shape, voxwid=c["voxwid"])
    ps_b = cbfp.ProjectionSet(pg, data=b)
    cfg_bp = cbfp.ConfigBackproject(fast_approx=True, cyl_mask_rad=c["N"] * c["voxwid"] / 2,
                                    soften_det_edges_proportion=0.05)
    key = ("sirt", c["N"], c["views"], c["nu"], c["nv"])
    tuner.tune(_abi.TUNE_PROJECT_COLUMN, key, lambda: P.A(truth.data, out=scratch_p))
    tuner.tune(_abi.TUNE_BACKPROJECT, key, lambda: cbfp.backproject(cfg_bp, scratch_v, ps_b))
    del scratch_p, scratch_v
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sirt = SIRT_flatpanel(SIRT_flatpanel.Config(), vol, cbfp.ProjectionSet(pg, data=b))
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    sirt.iterate()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters - 1):
        sirt.iterate()
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / (iters - 1)
    r = float(torch.linalg.vector_norm(sirt.residual.double()) / torch.linalg.vector_norm(b.double()))
    tuned = {name: {"variant": tuner.best(op, key), "times_ms": tuner.times_ms(op, key)}
             for name, op in (("project_column", _abi.TUNE_PROJECT_COLUMN), ("backproject", _abi.TUNE_BACKPROJECT))}
    for op in (_abi.TUNE_PROJECT_COLUMN, _abi.TUNE_BACKPROJECT):
        tuner.select(op, None)   # back to the defaults
    return {"workload": "256^3 fp32, 360 views, 384x384, 30 ellipsoids, x0=0", "s_per_iter": per,
            "setup_s": t_setup, "iterations": iters, "relative_residual_last": r, "tuned": tuned}


def storage_leg(P, x32, y32, g32, t_fp32: float, steps: int = 3) -> dict:
    """The same forward + adjoint step with 16-bit storage: bf16 volume, fp16
    sinogram, bf16 adjoint output (FP32 arithmetic in the kernels), timed like
    the fp32 step, with its error against the fp32 results of the same input."""
    import torch

    xb = x32.to(torch.bfloat16)
    yh = torch.zeros(P.sino_shape, dtype=torch.float16, device="cuda")
    gb = torch.zeros(P.vol_shape, dtype=torch.bfloat16, device="cuda")
    stream = torch.cuda.current_stream()
    rec = []
    for it in range(steps + 1):
        yh.zero_()
        gb.zero_()
        ea = time_events(lambda: P.A(xb, out=yh), stream)
        eb = time_events(lambda: P.AT(yh, out=gb), stream)
        if it > 0:
            rec.append((ea, eb))
    torch.cuda.synchronize()
    t_a = float(np.mean([ea[0].elapsed_time(ea[1]) for ea, _ in rec])) / 1e3
    t_at = float(np.mean([eb[0].elapsed_time(eb[1]) for _, eb in rec])) / 1e3

    def rel(a, b):
        return float(torch.linalg.vector_norm((a.double() - b.double()).flatten())
                     / torch.linalg.vector_norm(b.double().flatten()))

    return {"workload": "same step, bf16 volume / fp16 sinogram / bf16 adjoint output, fp32 arithmetic",
            "ms_forward": t_a * 1e3, "ms_adjoint": t_at * 1e3,
            "speedup_vs_fp32": t_fp32 / (t_a + t_at),
            "rel_l2_forward_vs_fp32": rel(yh, y32), "rel_l2_adjoint_vs_fp32": rel(gb, g32)}


def batched_leg(iters: int = 10) -> dict:
    """64 uneven tilted-detector SIRT problems (32^3-192^3, 60-360 views,
    64^2-256^2), `iters` iterations each: the plain FP32 GPU form (fp32
    storage, problems one after another, launched eagerly, default launch
    shapes) against the batched form (fp16 storage, FP32 arithmetic,
    autotuned launch shapes, all problems in one HIP graph over 8 streams)."""
    import torch

    from aprict_amd import phantoms as ph
    from aprict_amd.autotune import Autotuner, set_variant
    from aprict_amd.iterators import BatchedSIRT

    c = ph.CONFIGS[5]
    entries = ph.batched_problems(c["problems"], c["seed"])
    for op in range(3):
        set_variant(op, 0)
    plain = BatchedSIRT([ph.batched_problem_data(e, "float32")[:2] for e in entries], streams=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.run_eager(iters)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    tuner = Autotuner()
    t0 = time.perf_counter()
    batched = BatchedSIRT([ph.batched_problem_data(e, "float16")[:2] for e in entries], streams=8, tuner=tuner)
    batched.capture(iters)
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    t0 = time.perf_counter()
    batched.run()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    num = den = 0.0
    for a, b in zip(batched.volumes, plain.volumes, strict=True):
        num += float(torch.linalg.vector_norm((a.data.double() - b.data.double())) ** 2)
        den += float(torch.linalg.vector_norm(b.data.double()) ** 2)
    voxel_views = sum(e["N"] ** 3 * e["views"] for e in entries)
    return {"workload": f"{len(entries)} tilted-detector SIRT problems, {iters} iterations each",
            "s_plain_fp32_eager": t_plain, "s_batched_fp16_graph": t_batched, "speedup": t_plain / t_batched,
            "gups_batched": 2 * voxel_views * iters / t_batched / 1e9,
            "prep_s": t_prep, "rel_l2_fp16_vs_fp32": (num / den) ** 0.5}


def host_cpu() -> dict:
    """The host CPU: model name, logical CPUs, the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count() or 1, "affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(vol_np, pg, n, w, views: int, runs: int = 5) -> dict:
    """FP32 OpenMP port of the reference semantics (oracle.c built with
    -march=native on this host): the reference's snap DDA forward projector and
    its exact voxel-driven backprojector, on `views` evenly spaced views of the
    full workload; one warm-up run, then the median of `runs`.  Threads: every
    CPU this process may run on, capped by OMP_NUM_THREADS when that is set
    (the host's share for this GPU)."""
    import tempfile

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from aprict_amd.cb import flatpanel as cbfp
    from oracle import oracle as orc

    cpu = host_cpu()
    cores = cpu["affinity"]
    if cpu["omp_num_threads"]:
        try:
            cores = max(1, min(cores, int(cpu["omp_num_threads"])))
        except ValueError:
            pass
    build = orc.use_native_port(tempfile.mkdtemp(prefix="oracle_native_"))
    specs_all = pg.to_proj_spec_array()
    idx = np.linspace(0, len(specs_all), views, endpoint=False).astype(int)
    specs = specs_all[idx]
    texpos = -w * n / 2
    chunk = np.array((w, (0, 0, 0), (n, n, n), (texpos,) * 3),
                     dtype=__import__("aprict_amd")._abi.vol_tex_chunk_spec)
    nu, nv = int(specs[0]["dspec"]["nu"]), int(specs[0]["dspec"]["nv"])
    cfg = cbfp.ConfigProject().contents
    bp = cbfp.ConfigBackproject(fast_approx=False).contents
    tf, tb = [], []
    for _ in range(runs + 1):
        t0 = time.perf_counter()
        sino = orc.project_voxel_refdda(cfg, specs, chunk, vol_np, np.zeros((views, nu, nv), np.float32),
                                        nthreads=cores)
        tf.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        orc.backproject(bp, specs, chunk, np.zeros_like(vol_np), sino, nthreads=cores, fp32=True)
        tb.append(time.perf_counter() - t0)
    t_fwd, t_bp = float(np.median(tf[1:])), float(np.median(tb[1:]))
    gups = 2.0 * n**3 * views / (t_fwd + t_bp) / 1e9
    return {"value": gups, "unit": "GUPS", "cores": cores, "kind": "port",
            "host": cpu, "build": build,
            "sample": f"{views} of {len(specs_all)} views evenly spaced, full {n}^3 volume and {nu}x{nv} "
                      f"detector; FP32 OpenMP port of the reference's DDA forward projector and exact "
                      f"voxel-driven backprojector ({build} build, {cores} threads); 1 warm-up run, median of "
                      f"{runs}: fwd {t_fwd:.2f} s, bp {t_bp:.2f} s"}


def tv_gd_leg(P, truth, iters: int = 5, power_iterations: int = 8) -> dict:
    """The differentiable workload at 512^3 x 720 x 768^2: b = A x_true +
    N(0, (0.01 max b)^2) (noise from torch's seeded generator on the device),
    gradient descent on ||A x - b||^2 + 1e-3 TV_eps(x) from x = 0, with x
    stored fp32 and bf16.  Reports seconds per iteration (synchronised wall
    time of `iters` iterations) and the objective before and after."""
    import torch

    from aprict_amd.iterators import TVGradientDescent, power_iteration

    b = P.A(truth)
    gen = torch.Generator(device=b.device)
    gen.manual_seed(4)
    sigma = 0.01 * float(b.max())
    b.add_(torch.randn(b.shape, generator=gen, device=b.device, dtype=torch.float32), alpha=sigma)
    t0 = time.perf_counter()
    lam = power_iteration(P, power_iterations)
    torch.cuda.synchronize()
    t_pow = time.perf_counter() - t0
    out = {"workload": "gradient descent on ||Ax-b||^2 + 1e-3 TV_eps (eps 1e-3), 512^3, 720 x 768^2, "
                       "b = A x_true + 1 % Gaussian noise, x0 = 0, step 1/L",
           "lambda_max": lam, "power_iterations": power_iterations, "s_power_iteration": t_pow}
    for name, dt in (("f32", torch.float32), ("bf16", torch.bfloat16)):
        gd = TVGradientDescent(P, b, beta=1e-3, eps=1e-3, dtype=dt, lam=lam)
        f0 = sum(gd.objective())
        gd.iterate()   # warm-up (workspace, first launches)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(iters):
            gd.iterate()
        torch.cuda.synchronize()
        per = (time.perf_counter() - t0) / iters
        d, t = gd.objective()
        out[name] = {"s_per_iter": per, "objective_start": f0, "objective_end": d + t,
                     "objective_ratio": (d + t) / f0, "iterations": iters + 1, "tau": gd.tau}
        del gd
        torch.cuda.empty_cache()
    return out


def _summary_line(path: str, tag: str, kernel: str, nvals: int):
    """The numbers of the line '<tag> <kernel> v1 .. vn' of a committed
    rocprofv3 summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 2 + nvals and parts[0] == tag and parts[1] == kernel:
                    return [float(v) for v in parts[2:]]
    except OSError:
        return None
    return None


def read_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (fetch + write) from the committed
    rocprofv3 summary (a line 'traffic_bytes_per_launch <kernel> <fetch> <write>')."""
    v = _summary_line(TRAFFIC_FILE, "traffic_bytes_per_launch", kernel, 2)
    return None if v is None else v[0] + v[1]


def read_valu(kernel: str):
    """SQ_INSTS_VALU per launch of `kernel` from the committed rocprofv3
    summary (a line 'sq_insts_valu_per_launch <kernel> <count>')."""
    v = _summary_line(SQ_FILE, "sq_insts_valu_per_launch", kernel, 1)
    return None if v is None else v[0]


def read_lds_active(kernel: str):
    """SQ_LDS_IDX_ACTIVE per launch of `kernel` (LDS-active cycles summed over
    the CUs) from the committed rocprofv3 summary."""
    v = _summary_line(SQ_FILE, "sq_lds_idx_active_per_launch", kernel, 1)
    return None if v is None else v[0]


def calibrate_copy() -> int:
    """Two plain device-to-device copies of 1 GiB each (2 GiB read, 2 GiB
    written), the reference for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = 256 * 1024 * 1024
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each": n * 4, "copies": 2}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--volume-dtype", choices=["f32", "bf16"], default="f32")
    ap.add_argument("--cpu-views", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sirt", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-storage", action="store_true")
    ap.add_argument("--no-analytic", action="store_true")
    ap.add_argument("--no-fdk", action="store_true")
    ap.add_argument("--no-tv", action="store_true")
    ap.add_argument("--analytic-only", action="store_true", help="run only the analytic FORBILD leg")
    ap.add_argument("--calibrate-copy", action="store_true")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.analytic_only:
        log("analytic FORBILD projection, 40 primitives, 720 x 512^2")
        print(json.dumps({"analytic": analytic_leg()}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the projector shards by views across GPUs, which this build does not run)", file=sys.stderr)
        return 2

    import torch

    from aprict_amd import phantoms as ph
    from aprict_amd.autograd import Projector

    c = ph.CONFIGS[4]
    n, w, V, nu, nv = c["N"], c["voxwid"], c["views"], c["nu"], c["nv"]
    pg = ph.circular_geometry(V, nu, nv, c["pitch"])
    log(f"voxelising {n}^3 phantom ({c['n_ellipsoids']} ellipsoids, {args.volume_dtype})")
    vol = make_volume(n, w, c["n_ellipsoids"], c["seed"], args.volume_dtype)
    P = Projector(pg, vol_shape=(n, n, n), voxwid=w)
    x = vol.data
    y = torch.zeros(P.sino_shape, dtype=torch.float32, device="cuda")
    g = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    # the forward's launch shape picked by the autotuner (the adjoint has one)
    from aprict_amd import _abi
    from aprict_amd.autotune import Autotuner

    tuner = Autotuner(repeats=2)
    tkey = ("step", n, V, nu, nv)
    tuner.tune(_abi.TUNE_PROJECT_COLUMN, tkey, lambda: P.A(x, out=y))
    tuned = {"project_column": {"variant": tuner.best(_abi.TUNE_PROJECT_COLUMN, tkey),
                                "times_ms": tuner.times_ms(_abi.TUNE_PROJECT_COLUMN, tkey)}}

    def step(record):
        y.zero_()
        g.zero_()
        ea = time_events(lambda: P.A(x, out=y), stream)
        eb = time_events(lambda: P.AT(y, out=g), stream)
        if record is not None:
            record.append((ea, eb))

    log(f"warmup x{args.warmup}")
    for _ in range(args.warmup):
        step(None)
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    rec = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(rec)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_a = float(np.mean([ea[0].elapsed_time(ea[1]) for ea, _ in rec])) / 1e3
    t_at = float(np.mean([eb[0].elapsed_time(eb[1]) for _, eb in rec])) / 1e3
    log(f"A {t_a * 1e3:.2f} ms, A^T {t_at * 1e3:.2f} ms, wall per step {wall / args.steps * 1e3:.2f} ms")

    # adjoint consistency at full size: <Ax, y> vs <x, A^T y> with y = A x
    lhs = float((y.double() * y.double()).sum())
    rhs = float((x.double() * g.double()).sum())
    dot_rel = abs(lhs - rhs) / abs(lhs)

    voxel_views = float(n) ** 3 * V
    gups = 2 * voxel_views / (t_a + t_at) / 1e9
    s_v = 2 if args.volume_dtype == "bf16" else 4
    b_op = V * (n**3 * s_v + nu * nv * 4)
    dom = "adjoint" if t_at >= t_a else "forward"
    dom_kernel = "k_gather_adjoint" if dom == "adjoint" else "k_project_columns"
    t_dom = max(t_a, t_at)
    achieved = b_op / t_dom / 1e9
    traffic = read_traffic(dom_kernel) if args.volume_dtype == "f32" else None
    valu = read_valu(dom_kernel) if args.volume_dtype == "f32" else None
    lds = read_lds_active(dom_kernel) if args.volume_dtype == "f32" else None
    measured_gbs = None if traffic is None else traffic / t_dom / 1e9
    valu_frac = None if valu is None else valu / (VALU_ISSUE_PER_S * t_dom)
    # LDS-active cycles per CU over the cycles of the kernel's time
    lds_frac = None if lds is None else lds / (VALU_ISSUE_PER_S * t_dom)
    fracs = {"hbm": None if measured_gbs is None else measured_gbs / HBM_PEAK_GBS, "valu": valu_frac, "lds": lds_frac}
    known = {k: v for k, v in fracs.items() if v is not None}
    bound = max(known, key=known.get) if len(known) > 1 else "hbm"
    out = {
        "metric": METRIC,
        "value": gups,
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.volume_dtype == "f32" else "bf16-storage/f32",
        "data": "synthetic (seeded random-ellipsoid phantom)",
        "config": {"workload": "Siddon forward + matched adjoint, 512^3 volume (0.48 mm), 720 views, "
                               "768x768 detector (1 mm), SID 500 / SDD 1000 mm",
                   "N": n, "views": V, "detector": [nu, nv], "volume_dtype": args.volume_dtype},
        "ms_forward": t_a * 1e3,
        "ms_adjoint": t_at * 1e3,
        "gups_forward": voxel_views / t_a / 1e9,
        "gups_adjoint": voxel_views / t_at / 1e9,
        "dot_test_rel": dot_rel,
        "tuned": tuned,
        "roofline": {"bound": bound, "kernel": dom_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "measured_gbs": measured_gbs,
                     "measured_frac": None if measured_gbs is None else measured_gbs / HBM_PEAK_GBS,
                     "valu_issue_frac": valu_frac,
                     "lds_active_frac": lds_frac,
                     "model": "achieved = B_op / t with B_op = V (N^3 s_v + nu nv s_p) bytes per operator "
                              "application and t = the operator call on its stream (HIP events); "
                              "measured_gbs = rocprofv3 FETCH+WRITE bytes per launch / t; valu_issue_frac = "
                              "SQ_INSTS_VALU per launch / (256 CUs x 2.4 GHz x t); lds_active_frac = "
                              "SQ_LDS_IDX_ACTIVE per launch / (256 CUs x 2.4 GHz x t); bound = the largest of "
                              "measured_frac, valu_issue_frac and lds_active_frac"},
    }
    if args.volume_dtype == "f32" and not args.no_storage:
        log("16-bit storage leg")
        out["bf16_storage"] = storage_leg(P, x, y, g, t_a + t_at)
    if args.volume_dtype == "f32" and not args.no_tv:
        log("TV-regularised gradient descent 512^3 (fp32 and bf16 storage)")
        out["tv_gd"] = tv_gd_leg(P, x)
    if not args.no_sirt:
        log("SIRT 256^3 x 360 x 384^2")
        out["sirt"] = sirt_seconds_per_iter()
    if not args.no_batched:
        log("batched SIRT: 64 problems, plain eager vs one graph")
        out["batched"] = batched_leg()
    if not args.no_fdk:
        log("FDK 256^3 x 360 x 384^2")
        out["fdk"] = fdk_leg()
    if not args.no_analytic:
        log("analytic FORBILD projection, 40 primitives, 720 x 512^2")
        out["analytic"] = analytic_leg()
    if not args.no_cpu_baseline:
        log(f"CPU baseline on {args.cpu_views} views")
        vnp = vol.data.float().cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(vnp, pg, n, w, args.cpu_views)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
