"""Benchmark of the NeRF ray-marching hot path on one MI355X.

Workload (plain coarse+fine NeRF, 64 + 128 samples, 8x256 MLPs): one step is a
forward render of one 800x800 view (focal 1111.1, 640,000 rays in 32,768-ray
chunks) of the procedural Gaussian-blob scene; train throughput is measured on
4096-ray batches drawn across 100 seed-placed views.  Inputs (rays, bounds,
ground truth) are resident in HBM before the timed region.

Prints ONE JSON line.  `value` is render rays/s of the whole job; train rays/s,
the roofline of the dominant kernel (the fused fine-model MLP forward, timed
with HIP events on its stream) and the CPU baseline (the oracle restatement on
the host cores, bounded sample) ride along.

Quality half of the metric (`learned_vs_plain`): the plain FP32 form and the
learned BF16 + proposal form are trained for --quality-steps steps (2000) on the
same 4096-ray batches (distinct batches, float64 ground truth) from the same
initial weights, then render --quality-views held-out 800x800 views; PSNR and
max |rgb error| against the float64 ground truth are reported for the final
weights and for the weights averaged over the last --quality-average-last steps.
Their train rays/s are timed over those steps.
"""
from __future__ import annotations

import argparse
import importlib
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/sec render & train (64+128 samples), learned vs plain at ≤0.1 dB PSNR"
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
MLP_FLOP_PER_SAMPLE = 2 * 593408          # both dense stacks of one model, per sample
RENDER_FLOP_PER_RAY = (64 + 192) * MLP_FLOP_PER_SAMPLE


# HBM bytes per fine-model MLP launch (FETCH_SIZE + WRITE_SIZE, each scaled by the
# same counter on a 1 GiB device copy), from the rocprofv3 --pmc runs whose commands
# and figures are in profiles/mlp_fwd_f32r_render800.md and
# profiles/mlp_fwd_bf16ws8_render800.md.
TRAFFIC_BYTES_PER_LAUNCH = {"fp32": 153.8e6, "bf16": 142.5e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--train-steps", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="bound on the CPU baseline's render sample (the rest of it takes about a third more)")
    ap.add_argument("--forms", default="plain_fp32,plain_bf16,learned_bf16",
                    help="comma list; the first is the headline `value`")
    ap.add_argument("--learned-warmup-steps", type=int, default=2000,
                    help="train steps of the learned form before its render is timed")
    ap.add_argument("--no-graphs", action="store_true",
                    help="run the train steps eagerly instead of replaying a captured HIP graph "
                         "(for kernel traces: the profiler does not follow graph replays)")
    ap.add_argument("--recompute-rays", type=int, default=16384,
                    help="train batch of the activation-recomputation measurement (0: skip)")
    ap.add_argument("--recompute-chunk", type=int, default=2048,
                    help="rays per recomputation chunk")
    ap.add_argument("--no-hbm-kernels", action="store_true",
                    help="skip the timing of the HBM-bound sampling / compositing kernels and of "
                         "the train step's MLP stages")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (HBM counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--quality-steps", type=int, default=2000,
                    help="train steps of plain FP32 and learned BF16 on identical batches before their "
                         "PSNR against the float64 ground truth is measured (0: no quality run)")
    ap.add_argument("--quality-views", type=int, default=2, help="held-out 800x800 views rendered for PSNR")
    ap.add_argument("--quality-average-last", type=int, default=500,
                    help="also evaluate the weights averaged over the last N train steps")
    return ap.parse_args()


def under_profiler() -> bool:
    """rocprofv3 starts the program with its tool library named in ROCP_TOOL_LIBRARIES
    and its settings in ROCPROF_* variables."""
    return any(k.startswith("ROCPROF") or k == "ROCP_TOOL_LIBRARIES" for k in os.environ)


def main():
    args = parse()
    if not args.no_graphs and under_profiler():
        # A kernel trace of the learned form with HIP-graph replays ended in a segmentation
        # fault of the profiled process, and a counter run of it stopped printing; the same
        # form ran eagerly under the same profiler without either.  Under a profiler the
        # train steps therefore run eagerly.
        print("bench: running under rocprofv3: train steps run eagerly (no HIP graph capture)",
              file=sys.stderr, flush=True)
        args.no_graphs = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only (--gpus 1): the path shards by ray ranges, "
              "but it is not run or measured on several GPUs.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    nerf = importlib.import_module("nerf-tf2_amd")
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    sc = nerf.scene
    scene = sc.BlobScene.from_seed(0)
    P = nerf.Params()

    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).uniform_()
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---------------- inputs resident in HBM
    H = W = 800
    K = sc.intrinsics(H, W, 1111.1)
    pose = sc.random_poses(1, 1234)[0]
    ro, rd = nerf.ray_utils.get_rays(H, W, K, pose, device=dev)
    n_rays = ro.shape[0]
    near = torch.full((n_rays, 1), sc.NEAR, device=dev)
    far = torch.full((n_rays, 1), sc.FAR, device=dev)
    chunk = 32768

    # training batches: 4096 random rays across 100 seed-placed 800x800 views
    poses = sc.random_poses(100, 7)
    rng = np.random.default_rng(0)
    batches = []
    for b in range(16):
        views = rng.integers(0, 100, size=4096)
        pix = rng.integers(0, H * W, size=4096)
        o_l, d_l = [], []
        for v in np.unique(views):
            o, d = nerf.ray_utils.get_rays_tf(H, W, K.astype(np.float32), poses[v].astype(np.float32), device=dev)
            sel = torch.from_numpy(pix[views == v]).to(dev)
            o_l.append(o[sel])
            d_l.append(d[sel])
        bo = torch.cat(o_l).contiguous()
        bd = torch.cat(d_l).contiguous()
        gt, _, _ = scene.render(bo.double(), bd.double(), n_samples=1024)
        batches.append((bo, bd, torch.full((4096, 1), sc.NEAR, device=dev),
                        torch.full((4096, 1), sc.FAR, device=dev), gt.float().contiguous()))
    torch.cuda.synchronize()

    forms = [f for f in args.forms.split(",") if f]
    quality = None
    if args.quality_steps > 0:
        quality = quality_data(nerf, scene, args, dev)
    results = {}
    for form in forms:
        q = quality if form in ("plain_fp32", "learned_bf16") else None
        results[form] = measure_form(form, args, nerf, P, dev, ro, rd, near, far, chunk, batches, q)
    head = results[forms[0]]
    hbm = None if args.no_hbm_kernels else measure_hbm_kernels(nerf, dev)
    stages = None if args.no_hbm_kernels else measure_train_stages(nerf, dev)
    recompute = None
    if args.recompute_rays > 0:
        recompute = {p: measure_recompute(p, args, nerf, P, dev, batches) for p in ("fp32", "bf16")}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_seconds)

    prec = "fp32" if forms[0].endswith("fp32") else "bf16"
    peak = FP32_PEAK_TFLOPS if prec == "fp32" else BF16_PEAK_TFLOPS
    line = {
        "metric": METRIC, "value": head["render_rays_per_s"], "unit": "rays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["render_ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": prec, "data": "synthetic (procedural Gaussian-blob scene, random-init weights)",
        "config": {"workload": f"{forms[0]}: coarse+fine NeRF render of one 800x800 view (640,000 rays, "
                               "32,768-ray chunks, 64+128 samples, 8x256 MLPs); train on 4096-ray batches",
                   "rays_per_step": n_rays, "chunk": chunk, "train_batch": 4096},
        "train_rays_per_s": head["train_rays_per_s"],
        "train_ms_per_step": head["train_ms_per_step"],
        "render_tflops": round(head["render_rays_per_s"] * RENDER_FLOP_PER_RAY / 1e12, 2),
        "roofline": None if "mlp_tflops" not in head else {
            "bound": "mfma", "kernel": "mlp_fwd (fine model, 32,768 rays x 192 samples per launch)",
            "achieved": head["mlp_tflops"], "peak": peak, "unit": "TFLOP/s",
            "frac": round(head["mlp_tflops"] / peak, 4), "traffic": TRAFFIC_BYTES_PER_LAUNCH[prec],
            "avg_launch_ms": head["mlp_avg_launch_ms"], "flop_per_launch": head["mlp_flop_per_launch"]},
        "cpu_baseline": cpu,
        "forms": results,
        "hbm_kernels": hbm,
        "train_stages": stages,
        "recompute": recompute,
    }
    if "learned_bf16" in results and "plain_fp32" in results and results["plain_fp32"]["train_rays_per_s"]:
        lp, pp = results["learned_bf16"], results["plain_fp32"]
        line["learned_vs_plain"] = {
            "render_speedup": round(lp["render_rays_per_s"] / pp["render_rays_per_s"], 2),
            "train_speedup": round(lp["train_rays_per_s"] / pp["train_rays_per_s"], 2),
        }
        if "quality" in lp and "quality" in pp:
            ql, qp = lp["quality"], pp["quality"]
            line["learned_vs_plain"]["psnr"] = {
                "measured": f"this run: {args.quality_steps} train steps each on identical 4096-ray batches, "
                            f"{args.quality_views} held-out 800x800 views vs the float64 ground truth, seed 0",
                "plain_fp32_db": qp["final"]["psnr_mean"], "learned_bf16_db": ql["final"]["psnr_mean"],
                "learned_minus_plain_fp32_db": round(ql["final"]["psnr_mean"] - qp["final"]["psnr_mean"], 3),
                "max_abs_rgb_err_plain_fp32": qp["final"]["max_abs_rgb_err"],
                "max_abs_rgb_err_learned_bf16": ql["final"]["max_abs_rgb_err"],
                "max_abs_rgb_err_learned_vs_plain": max(ql["final"]["max_abs_rgb_err_vs_plain_fp32"]),
            }
            if "averaged" in ql and "averaged" in qp:
                line["learned_vs_plain"]["psnr"].update({
                    "averaged_last": args.quality_average_last,
                    "plain_fp32_avg_db": qp["averaged"]["psnr_mean"],
                    "learned_bf16_avg_db": ql["averaged"]["psnr_mean"],
                    "learned_minus_plain_fp32_avg_db": round(ql["averaged"]["psnr_mean"] -
                                                             qp["averaged"]["psnr_mean"], 3)})
    print(json.dumps(line))


def quality_data(nerf, scene, args, dev):
    """The quality run's data: args.quality_steps distinct 4096-ray batches drawn
    over 100 seed-placed views with float64 ground truth, and args.quality_views
    held-out 800x800 views with theirs (as nerf-tf2_amd/quality.py)."""
    import torch
    quality = importlib.import_module("nerf-tf2_amd.quality")
    sc = nerf.scene
    t0 = time.perf_counter()
    qb = quality.make_batches(nerf, scene, args.quality_steps, device=dev)
    K = sc.intrinsics(800, 800, 1111.1)
    views = []
    for pose in sc.random_poses(args.quality_views, 999):
        o, d = nerf.ray_utils.get_rays(800, 800, K, pose, device=dev)
        gt, _, _ = scene.render(o.double(), d.double(), n_samples=1024, chunk=4096)
        views.append((o, d, gt))
    torch.cuda.synchronize()
    print(f"bench: quality data ({args.quality_steps} batches, {args.quality_views} views) "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return {"batches": qb, "views": views, "renders": {}}


def evaluate_quality(net, learned, views, ref=None):
    """PSNR (per element, float64) and max |rgb error| of each held-out view
    against its float64 ground truth (and against `ref`'s renders)."""
    import torch
    sc = importlib.import_module("nerf-tf2_amd").scene
    preds = []
    for o, d, _ in views:
        n = o.shape[0]
        nr = torch.full((n, 1), sc.NEAR, device=o.device)
        fr = torch.full((n, 1), sc.FAR, device=o.device)
        out = net.predict(o, d, nr, fr, fallback=True) if learned else net.predict(o, d, nr, fr)
        preds.append(out["pred_rgb"])
    rec = {"psnr": [round(sc.psnr(v[2], p), 3) for p, v in zip(preds, views)],
           "max_abs_rgb_err": max(round(float((p.double() - v[2]).abs().max()), 4) for p, v in zip(preds, views))}
    rec["psnr_mean"] = round(sum(rec["psnr"]) / len(rec["psnr"]), 3)
    if ref is not None:
        rec["max_abs_rgb_err_vs_plain_fp32"] = [round(float((a - b).abs().max()), 4) for a, b in zip(preds, ref)]
    return rec, preds


def measure_form(form, args, nerf, P, dev, ro, rd, near, far, chunk, batches, quality=None):
    """Train (timed, graph-replayed 4096-ray steps) and render (timed, K steps of
    one 800x800 view) rays/s of one form.  With `quality`, the form is trained on
    the quality run's batches (timing every step) and its held-out views are
    evaluated before the render is timed; otherwise the learned form trains its
    warm-up steps first so that its render runs with a trained proposal."""
    import numpy as np
    import torch
    n_rays = ro.shape[0]
    learned = form.startswith("learned")
    prec = "bf16" if form.endswith("bf16") else "fp32"
    net = nerf.ProposalNeRF(P, prec, seed=0, device=dev) if learned else nerf.NeRF(P, prec, seed=0, device=dev)
    out = {"train_rays_per_s": None, "train_ms_per_step": None}
    nb = len(batches)
    if quality is not None:
        out.update(train_for_quality(form, net, learned, args, nerf, dev, quality))
    elif args.train_steps > 0 or learned:
        step = net.train_step if args.no_graphs else nerf.GraphedTrainStep(net, batches[0][0].shape[0])
        pre = args.learned_warmup_steps if learned else 2
        for i in range(pre):
            step(*batches[i % nb])
            if (i + 1) % 500 == 0:
                print(f"bench: {form} warm-up train step {i + 1}/{pre}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        if args.train_steps > 0:
            t0 = time.perf_counter()
            for i in range(args.train_steps):
                step(*batches[(pre + i) % nb])
            torch.cuda.synchronize()
            t_train = time.perf_counter() - t0
            out["train_rays_per_s"] = round(batches[0][0].shape[0] * args.train_steps / t_train, 1)
            out["train_ms_per_step"] = round(1e3 * t_train / args.train_steps, 3)
        del step
    events = []

    def render(timed_kernel):
        fb = 0.0
        for i in range(0, n_rays, chunk):
            sl = slice(i, min(i + chunk, n_rays))
            if learned:
                with torch.no_grad():
                    net.forward(ro[sl], rd[sl], near[sl], far[sl], fallback=True)
                fb += net.last_fallback_fraction * (sl.stop - sl.start)
            else:
                _render_chunk(net, nerf, ro[sl], rd[sl], near[sl], far[sl], timed_kernel, events)
        return fb / n_rays

    for _ in range(args.warmup):
        render(False)
    torch.cuda.synchronize()
    events.clear()
    t0 = time.perf_counter()
    fb = 0.0
    for _ in range(args.steps):
        fb = render(True)
    torch.cuda.synchronize()
    t_render = time.perf_counter() - t0
    out["render_rays_per_s"] = round(n_rays * args.steps / t_render, 1)
    out["render_ms_per_step"] = round(1e3 * t_render / args.steps, 3)
    if learned:
        out["fallback_fraction"] = round(fb, 4)
        out["warmup_train_steps"] = args.quality_steps if quality is not None else args.learned_warmup_steps
    full = [e0.elapsed_time(e1) for e0, e1, m in events if m == chunk * 192]
    if full:
        avg_ms = float(np.mean(full))
        flop = chunk * 192 * MLP_FLOP_PER_SAMPLE
        out["mlp_avg_launch_ms"] = round(avg_ms, 3)
        out["mlp_flop_per_launch"] = flop
        out["mlp_tflops"] = round(flop / (avg_ms * 1e-3) / 1e12, 2)
    del net
    torch.cuda.empty_cache()
    return out


def train_for_quality(form, net, learned, args, nerf, dev, quality):
    """args.quality_steps graph-replayed train steps on the quality batches (all
    timed), then the held-out views with the final weights and with the weights
    averaged over the last args.quality_average_last steps (the final weights are
    restored afterwards)."""
    import torch
    sc = nerf.scene
    qb = quality["batches"]
    n = qb[0][0].shape[0]
    near = torch.full((n, 1), sc.NEAR, device=dev)
    far = torch.full((n, 1), sc.FAR, device=dev)
    torch.manual_seed(1000)                      # the samplers' draws: the same stream for every form
    step = net.train_step if args.no_graphs else nerf.GraphedTrainStep(net, n)
    mods = net.optimizer.models
    avg_n = min(args.quality_average_last, len(qb))
    sums = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, (o, d, gt) in enumerate(qb):
        step(o, d, near, far, gt)
        if avg_n and i >= len(qb) - avg_n:
            if sums is None:
                sums = [m.params.detach().clone() for m in mods]
            else:
                for acc, m in zip(sums, mods):
                    acc.add_(m.params.detach())
        if (i + 1) % 500 == 0:
            print(f"bench: {form} train step {i + 1}/{len(qb)}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del step
    out = {"train_rays_per_s": round(n * len(qb) / t_train, 1), "train_ms_per_step": round(1e3 * t_train / len(qb), 3),
           "train_steps_timed": len(qb)}
    ref = quality["renders"].get("plain_fp32")
    q = {}
    with torch.no_grad():
        q["final"], preds = evaluate_quality(net, learned, quality["views"], ref)
        if learned:
            q["final"]["fallback_fraction"] = round(net.last_fallback_fraction, 4)
        if form == "plain_fp32":
            quality["renders"]["plain_fp32"] = preds
        if sums is not None:
            final = [m.params.detach().clone() for m in mods]
            for acc, m in zip(sums, mods):
                m.params.copy_(acc / float(avg_n))
                m.repack()
            q["averaged"], _ = evaluate_quality(net, learned, quality["views"], None)
            if learned:
                q["averaged"]["fallback_fraction"] = round(net.last_fallback_fraction, 4)
            for f, m in zip(final, mods):
                m.params.copy_(f)
                m.repack()
    print(f"bench: {form} quality {json.dumps(q)}", file=sys.stderr, flush=True)
    out["quality"] = q
    return out


def measure_hbm_kernels(nerf, dev, n=32768, reps=20):
    """The HBM-bound kernels of the path at one 32,768-ray render chunk (64 + 128
    samples), each launched alone through the C ABI and timed with HIP events on
    its stream: average duration, algorithmic bytes per ray (the float32 inputs it
    must read and outputs it must write, as called by the render: no materialized
    xyz, no index outputs) and the achieved rate against the 8 TB/s HBM peak."""
    import torch
    L = nerf._lib
    call, ptr, st = L.call, L.ptr, L.stream(dev)
    Nc, Nf = 64, 128
    S = Nc + Nf
    g = torch.Generator(device=dev).manual_seed(3)
    f = lambda *shape: torch.rand(shape, device=dev, generator=g)
    ro = torch.zeros((n, 3), device=dev)
    rd = torch.nn.functional.normalize(f(n, 3) + 0.1, dim=1)
    near, far = torch.full((n, 1), 2.0, device=dev), torch.full((n, 1), 6.0, device=dev)
    u_c, u_f = f(n, Nc), f(n, Nf)
    edges, widths, t_c = (torch.empty((n, Nc + 1), device=dev), torch.empty((n, Nc), device=dev),
                          torch.empty((n, Nc), device=dev))
    w_c = f(n, Nc) ** 4
    t = torch.empty((n, S), device=dev)
    sigma, rgb = torch.relu(torch.randn((n * S, 1), device=dev, generator=g)) * 5, f(n * S, 3)
    w_out, rgb_out, depth, acc = (torch.empty((n, S), device=dev), torch.empty((n, 3), device=dev),
                                  torch.empty((n,), device=dev), torch.empty((n,), device=dev))
    g_rgb, g_depth, g_acc = f(n, 3), f(n), f(n)
    d_sigma, d_rgb = torch.empty((n * S, 1), device=dev), torch.empty((n * S, 3), device=dev)
    launches = {
        "coarse_kernel (Nc=64)": (8 + 16 * Nc + 4, lambda: call(
            "nerf_create_input_batch_coarse", n, Nc, 1, 1, ptr(ro), ptr(rd), ptr(near), ptr(far), ptr(u_c),
            ptr(edges), ptr(widths), ptr(t_c), None, None, st)),
        "fine128_kernel (Nc=64, Nf=128)": (4 * (3 * Nc + (Nc + 1) + Nf + S), lambda: call(
            "nerf_create_input_batch_fine", n, Nc, Nf, ptr(ro), ptr(rd), ptr(w_c), ptr(edges), ptr(widths),
            ptr(t_c), ptr(u_f), ptr(t), None, None, None, None, 0, None, st)),
        "fine128_kernel compensated (Nc=64, Nf=128)": (4 * (3 * Nc + (Nc + 1) + Nf + S), lambda: call(
            "nerf_create_input_batch_fine", n, Nc, Nf, ptr(ro), ptr(rd), ptr(w_c), ptr(edges), ptr(widths),
            ptr(t_c), ptr(u_f), ptr(t), None, None, None, None, 1, None, st)),
        "composite_fwd_kernel (S=192)": (24 * S + 20, lambda: call(
            "nerf_composite_forward", n, S, 1, ptr(sigma), ptr(rgb), ptr(t), ptr(w_out), ptr(rgb_out), ptr(depth),
            ptr(acc), st)),
        "composite_bwd_kernel (S=192)": (36 * S + 20, lambda: call(
            "nerf_composite_backward", n, S, 1, ptr(sigma), ptr(rgb), ptr(t), ptr(g_rgb), ptr(g_depth), ptr(g_acc),
            None, ptr(d_sigma), ptr(d_rgb), st)),
    }
    out = {"rays_per_launch": n}
    for name, (bpr, fn) in launches.items():
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = 1e3 * e0.elapsed_time(e1) / reps
        gbs = bpr * n / (us * 1e-6) / 1e9
        out[name] = {"avg_us": round(us, 2), "bytes_per_ray": bpr, "achieved_gbs": round(gbs, 1),
                     "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 3)}
    return out


def measure_train_stages(nerf, dev, rays=4096, spr=176, reps=20):
    """The four MLP launches of the BF16 train step at the learned form's fine-model
    shape (4096 rays x (48 + 128) samples = 720,896 samples), each launched alone
    through the C ABI and timed with HIP events on its stream: the forward without
    and with the activation stash, the dX chain and the weight gradients, with the
    algorithmic FLOPs and HBM bytes of each.  Stash and dz rows are bf16 (bytes per
    sample = 2 x the row length in elements)."""
    import ctypes
    import torch
    L = nerf._lib
    M = rays * spr
    mlp = nerf.NerfMLP("fine", nerf.MLPConfig(), "bf16", seed=4, device=dev)
    d = ctypes.byref(mlp._desc)
    g = torch.Generator(device=dev).manual_seed(11)
    xyz = torch.rand((M, 3), device=dev, generator=g) * 6 - 3
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), device=dev, generator=g), dim=1)
    g_rgb = torch.randn((M, 3), device=dev, generator=g)
    g_sig = torch.randn((M, 1), device=dev, generator=g)
    rgb, sigma = torch.empty((M, 3), device=dev), torch.empty((M, 1), device=dev)
    stash = mlp._alloc_stash(M)
    dz = torch.empty((M, mlp.dz_per_sample), dtype=torch.bfloat16, device=dev)
    grad = torch.empty((mlp.n_params,), device=dev)
    ws = torch.empty((int(L.load().nerf_mlp_wgrad_workspace_bytes(d, M)),), dtype=torch.uint8, device=dev)
    st_b, dz_b = 2 * mlp.stash_per_sample, 2 * mlp.dz_per_sample
    stages = {
        "fwd": (MLP_FLOP_PER_SAMPLE, 0, lambda: mlp._launch_forward(
            M, xyz, dirs, None, None, None, rgb, sigma, None)),
        "fwd_stash": (MLP_FLOP_PER_SAMPLE, st_b, lambda: mlp._launch_forward(
            M, xyz, dirs, None, None, None, rgb, sigma, stash)),
        # dX of every layer but the input sides of dense_0, dense_5 (63 x 256 each) and dense_9 (27 x 128)
        "bwd": (2 * (593408 - 35712), dz_b, lambda: L.call(
            "nerf_mlp_backward", d, mlp.prec, L.ptr(mlp.packed), M, L.ptr(g_rgb), L.ptr(g_sig), L.ptr(rgb),
            L.ptr(sigma), L.ptr(stash), L.ptr(dz), L.stream(dev))),
        "wgrad": (MLP_FLOP_PER_SAMPLE, st_b + dz_b, lambda: L.call(
            "nerf_mlp_weight_grads", d, mlp.prec, M, L.ptr(stash), L.ptr(dz), L.ptr(grad), 0, L.ptr(ws),
            ws.numel(), L.stream(dev))),
    }
    out = {"samples": M, "workload": f"fine model of the learned BF16 train step: {rays} rays x {spr} samples"}
    for name, (flop, nbytes, fn) in stages.items():
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = 1e3 * e0.elapsed_time(e1) / reps
        r = {"avg_us": round(us, 1), "flop_per_sample": flop,
             "tflops": round(flop * M / (us * 1e-6) / 1e12, 1)}
        if nbytes:
            gbs = nbytes * M / (us * 1e-6) / 1e9
            r.update({"hbm_bytes_per_sample": nbytes, "achieved_gbs": round(gbs, 1),
                      "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 3)})
        out[name] = r
    del stash, dz, ws
    torch.cuda.empty_cache()
    return out


def measure_recompute(prec, args, nerf, P, dev, batches):
    """Train steps on one large batch (args.recompute_rays rays x 192 samples) with
    and without activation recomputation: peak HBM allocated by the step and
    train rays/s (eager steps, 1 untimed + 3 timed)."""
    import torch
    n = args.recompute_rays
    reps = -(-n // batches[0][0].shape[0])
    big = [torch.cat([batches[i % len(batches)][k] for i in range(reps)])[:n].contiguous() for k in range(5)]
    out = {"rays": n, "samples_per_ray": 192, "chunk_rays": args.recompute_chunk}
    for mode in ("stash", "recompute"):
        net = nerf.NeRF(P, prec, seed=0, device=dev)
        if mode == "recompute":
            net.coarse_model.recompute_chunk = args.recompute_chunk
            net.fine_model.recompute_chunk = args.recompute_chunk
        net.train_step(*big)
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        t0 = time.perf_counter()
        for _ in range(3):
            net.train_step(*big)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 3
        out[mode] = {"train_rays_per_s": round(n / dt, 1), "ms_per_step": round(1e3 * dt, 2),
                     "peak_step_hbm_bytes": int(torch.cuda.max_memory_allocated() - base)}
        print(f"bench: {prec} {n}-ray train step ({mode}) {1e3 * dt:.1f} ms", file=sys.stderr, flush=True)
        del net
        torch.cuda.empty_cache()
    return out


def _render_chunk(net, nerf, ro, rd, near, far, timed_kernel, events):
    """NeRF.forward for one chunk with the fine MLP launch bracketed by HIP events
    on the stream it is launched on (read after the timed region)."""
    import torch
    ru = nerf.ray_utils
    with torch.no_grad():
        dc = ru.create_input_batch_coarse_model(net.params, ro, rd, near, far, materialize=False)
        rgb_c, sig_c = net.coarse_model.forward_rays(ro, rd, dc["t_vals"])
        pc = ru.post_process_model_output(rgb_c, sig_c, dc["t_vals"], net.white_bg)
        df = ru.create_input_batch_fine_model(net.params, ro, rd, pc["weights"], dc["bin_data"], dc["t_vals"],
                                              materialize=False, compensated=net.compensated_cdf)
        if timed_kernel:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        rgb_f, sig_f = net.fine_model.forward_rays(ro, rd, df["t_vals"])
        if timed_kernel:
            e1.record()
            events.append((e0, e1, int(df["t_vals"].numel())))
        pf = ru.post_process_model_output(rgb_f, sig_f, df["t_vals"], net.white_bg)
    return pf


def cpu_baseline(seconds: float):
    """The oracle (PyTorch-CPU float32 restatement of the reference path) on the
    host cores, on the workloads of the measurement plan: config 0 in full (a
    1024-ray forward and one Adam train step on a 100x100 view), one 32,768-ray
    chunk of the 800x800 render (a whole view is 19.53 such chunks) and one
    4096-ray train step.  `seconds` bounds the render sample: the chunk runs in
    4096-ray slices until it is done or `seconds` have passed."""
    import numpy as np
    import torch
    oracle = importlib.import_module("oracle.restatement")
    sc = importlib.import_module("nerf-tf2_amd").scene
    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    spec = oracle.MLPSpec()
    wc = [torch.from_numpy(w) for w in oracle.glorot_weights(spec, 0)]
    wf = [torch.from_numpy(w) for w in oracle.glorot_weights(spec, 1)]
    P = oracle.make_params()
    g = torch.Generator().manual_seed(0)

    def rays(H, W, focal, n, seed):
        ro, rd = oracle.get_rays(H, W, sc.intrinsics(H, W, focal), sc.random_poses(1, seed)[0])
        idx = np.random.default_rng(seed).integers(0, H * W, size=n)
        f = lambda a: torch.from_numpy(np.ascontiguousarray(a[idx], dtype=np.float32))
        return f(ro), f(rd), torch.full((n, 1), 2.0), torch.full((n, 1), 6.0)

    def forward(o, d, nr, fr):
        n = o.shape[0]
        with torch.no_grad():
            oracle.forward(P, spec, wc, wf, o, d, nr, fr, torch.rand(n, 64, generator=g), torch.rand(n, 128, generator=g))

    def train_step(o, d, nr, fr, ws):
        n = o.shape[0]
        _, _, gc, gf, _ = oracle.train_step_grads(P, spec, ws[0], ws[1], o, d, nr, fr, torch.rand(n, 3, generator=g),
                                                  torch.rand(n, 64, generator=g), torch.rand(n, 128, generator=g))
        new = oracle.AdamTF(ws[0] + ws[1]).step(ws[0] + ws[1], gc + gf)
        return [new[:len(ws[0])], new[len(ws[0]):]]

    out = {"unit": "rays/s", "cores": threads, "kind": "port"}
    # config 0: 1024 rays of a 100x100 view, forward and one train step
    c0 = rays(100, 100, 138.9, 1024, 3)
    t0 = time.perf_counter()
    forward(*c0)
    t_c0f = time.perf_counter() - t0
    t0 = time.perf_counter()
    train_step(*c0, [[w.clone() for w in wc], [w.clone() for w in wf]])
    t_c0t = time.perf_counter() - t0
    # one 32,768-ray chunk of the 800x800 render (rays of the bench's view)
    n_chunk = 32768
    r1 = rays(800, 800, 1111.1, n_chunk, 1234)
    n_run = 0
    t0 = time.perf_counter()
    while n_run < n_chunk and (n_run == 0 or time.perf_counter() - t0 < seconds):
        forward(*(x[n_run:n_run + 4096] for x in r1))
        n_run += 4096
    t_chunk = time.perf_counter() - t0
    # one 4096-ray train step on the 800x800 view's rays
    t0 = time.perf_counter()
    train_step(*(x[:4096] for x in r1), [[w.clone() for w in wc], [w.clone() for w in wf]])
    t_train = time.perf_counter() - t0
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    render_rps = n_run / t_chunk
    out.update({
        "value": round(render_rps, 1),
        "sample": (f"{n_run} rays of one 32,768-ray chunk of the 800x800 render (coarse+fine forward in 4096-ray "
                   f"slices, {t_chunk:.1f} s; a whole view = 19.53 chunks = {640000 / render_rps:.0f} s at this "
                   f"rate), torch-CPU float32 restatement of the reference path, {threads} threads, {cpu_model}"),
        "train_rays_per_s": round(4096 / t_train, 1),
        "train_sample": f"one 4096-ray train step (forward, autograd backward, Adam), {t_train:.1f} s",
        "config0": {"forward_s": round(t_c0f, 2), "train_step_s": round(t_c0t, 2),
                    "forward_rays_per_s": round(1024 / t_c0f, 1), "train_rays_per_s": round(1024 / t_c0t, 1),
                    "workload": "1024 rays of a 100x100 view: one forward, one Adam train step"},
    })
    return out


if __name__ == "__main__":
    main()
