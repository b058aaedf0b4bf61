"""Quality and throughput of the three forms on the Gaussian-blob scene.

    python -m nerf-tf2_amd.quality --steps 2000          (run from the repo root)

Trains (graph-replayed steps), for each seed, on identical 4096-ray batches
drawn across 100 seed-placed 800x800 views, (a) the plain FP32 coarse+fine NeRF,
(b) the same in BF16 (precision form) and (c) the BF16 learned-proposal form;
then renders held-out 800x800 views with each and reports per-element PSNR and
max |rgb error| against the float64 ground truth, the errors of (b) and (c)
against (a), render / train rays/s, and the PSNR change against (a) per seed and
averaged.  The control form plain_fp32_redraw is (a) again with another stream
of stratified / fine-sample draws (same initial weights, same batches): its PSNR
change against (a) is the run-to-run spread that sample placement alone causes.
With --shared, the trained weights of (a) are also rendered in BF16 and by the
learned form (its proposal fitted to the fixed coarse + fine pair for the same
number of steps): the PSNR change of the form itself on identical weights.
With --average-last N each run is also evaluated with its weights averaged over
its last N steps, which removes most of the step-to-step jitter of the final
iterate from the comparison.  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import importlib
import json
import math
import sys
import time

import numpy as np
import torch


def _pkg():
    return importlib.import_module("nerf-tf2_amd")


def make_batches(nerf, scene, n_steps, batch=4096, n_views=100, H=800, W=800, focal=1111.1, seed=7,
                 device="cuda"):
    """n_steps batches of `batch` rays drawn uniformly over (view, pixel), with
    float64 ground truth (1024-sample closed-form optical depth)."""
    sc = nerf.scene
    K = sc.intrinsics(H, W, focal).astype(np.float32)
    poses = sc.random_poses(n_views, seed)
    ro_all, rd_all = [], []
    for v in range(n_views):
        o, d = nerf.ray_utils.get_rays_tf(H, W, K, poses[v].astype(np.float32), device=device)
        ro_all.append(o[:1].clone())        # one origin per view
        rd_all.append(d)
    rd_all = torch.stack(rd_all)            # (V, H*W, 3)
    ro_all = torch.cat(ro_all)              # (V, 3)
    g = torch.Generator(device=device).manual_seed(seed)
    out = []
    for step in range(n_steps):
        if step % 500 == 0:
            print(f"quality: ground truth for batch {step}/{n_steps}", file=sys.stderr, flush=True)
        v = torch.randint(0, n_views, (batch,), device=device, generator=g)
        p = torch.randint(0, H * W, (batch,), device=device, generator=g)
        o = ro_all[v].contiguous()
        d = rd_all[v, p].contiguous()
        gt, _, _ = scene.render(o.double(), d.double(), n_samples=1024, chunk=4096)
        out.append((o, d, gt.float().contiguous()))
    return out


class _NoFallback:
    """A learned-form net rendered with its proposal only (no plain-pass fallback)."""

    def __init__(self, net):
        self.net = net

    def predict(self, ro, rd, near, far):
        out = self.net.predict(ro, rd, near, far, fallback=False)
        self.last_fallback_fraction = self.net.last_fallback_fraction   # what the fallback would re-render
        return out


class _Stratified:
    """The fine model alone on n stratified samples per ray (no proposal, no
    resampling): the same number of fine-model evaluations as the learned form."""

    def __init__(self, nerf, fine_model, n, white_bg=True):
        self.ru, self.fine, self.white_bg = nerf.ray_utils, fine_model, white_bg
        self.p = nerf.proposal._P(n, 0, True, white_bg)

    @torch.no_grad()
    def predict(self, ro, rd, near, far, chunk=32768):
        out = []
        for i in range(0, ro.shape[0], chunk):
            sl = slice(i, i + chunk)
            d = self.ru.create_input_batch_coarse_model(self.p, ro[sl], rd[sl], near[sl], far[sl], materialize=False)
            rgb, sig = self.fine.forward_rays(ro[sl], rd[sl], d["t_vals"])
            out.append(self.ru.post_process_model_output(rgb, sig, d["t_vals"], self.white_bg)["pred_rgb"])
        return {"pred_rgb": torch.cat(out)}


def psnr(a, b):
    mse = float(torch.mean((a.double() - b.double()) ** 2))
    return -10.0 * math.log10(max(mse, 1e-30))


def run(steps=2000, eval_views=4, seeds=(0, 1), device="cuda",
        forms_sel=("plain_fp32", "plain_bf16", "learned_bf16"), shared=False, average_last=0):
    nerf = _pkg()
    sc = nerf.scene
    scene = sc.BlobScene.from_seed(0)
    P = nerf.Params()
    t0 = time.perf_counter()
    batches = make_batches(nerf, scene, steps, device=device)
    near = torch.full((4096, 1), sc.NEAR, device=device)
    far = torch.full((4096, 1), sc.FAR, device=device)
    # held-out views
    H = W = 800
    K = sc.intrinsics(H, W, 1111.1)
    views = []
    for pose in sc.random_poses(eval_views, 999):
        ro, rd = nerf.ray_utils.get_rays(H, W, K, pose, device=device)
        gt, _, _ = scene.render(ro.double(), rd.double(), n_samples=1024, chunk=4096)
        views.append((ro, rd, gt))
    t_data = time.perf_counter() - t0
    makers = {
        "plain_fp32": lambda sd: nerf.NeRF(P, "fp32", seed=sd, device=device),
        "plain_bf16": lambda sd: nerf.NeRF(P, "bf16", seed=sd, device=device),
        "learned_bf16": lambda sd: nerf.ProposalNeRF(P, "bf16", seed=sd, device=device),
        "plain_fp32_redraw": lambda sd: nerf.NeRF(P, "fp32", seed=sd, device=device),
    }
    draw_seed = {"plain_fp32_redraw": 5000}
    res = {"steps": steps, "batch": 4096, "eval_views": eval_views, "seeds": list(seeds),
           "data_seconds": round(t_data, 1), "runs": []}
    renders = {}

    def evaluate(net, name, sd, t_train):
        preds, fb = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for ro, rd, _ in views:
            n = ro.shape[0]
            nr = torch.full((n, 1), sc.NEAR, device=device)
            fr = torch.full((n, 1), sc.FAR, device=device)
            preds.append(net.predict(ro, rd, nr, fr)["pred_rgb"])
            fb.append(getattr(net, "last_fallback_fraction", 0.0))
        torch.cuda.synchronize()
        t_render = time.perf_counter() - t0
        renders[(sd, name)] = preds
        rec = {"seed": sd, "form": name,
               "train_rays_per_s": None if t_train is None else round(4096 * steps / t_train, 1),
               "render_rays_per_s": round(sum(v[0].shape[0] for v in views) / t_render, 1),
               "psnr_vs_fp64_gt": [round(psnr(p, v[2]), 3) for p, v in zip(preds, views)],
               "max_abs_rgb_err_vs_fp64_gt": [round(float((p.double() - v[2]).abs().max()), 4)
                                              for p, v in zip(preds, views)]}
        rec["psnr_mean"] = round(float(np.mean(rec["psnr_vs_fp64_gt"])), 3)
        if name.startswith("learned"):
            rec["fallback_fraction"] = [round(x, 4) for x in fb]
        if getattr(net, "scaler", None) is not None:
            rec["loss_scale"] = net.scaler.scale
            rec["skipped_steps"] = net.scaler.skipped
        if name != "plain_fp32" and (sd, "plain_fp32") in renders:
            ref = renders[(sd, "plain_fp32")]
            rec["psnr_vs_plain_fp32"] = [round(psnr(a, b), 3) for a, b in zip(preds, ref)]
            rec["max_abs_rgb_err_vs_plain_fp32"] = [round(float((a - b).abs().max()), 4)
                                                    for a, b in zip(preds, ref)]
        res["runs"].append(rec)
        print(f"quality: {json.dumps(rec)}", file=sys.stderr, flush=True)

    for sd in seeds:
        for name in forms_sel:
            torch.manual_seed(draw_seed.get(name, 1000) + sd)
            net = makers[name](sd)
            step = nerf.GraphedTrainStep(net, 4096)
            mods = net.optimizer.models
            sums = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i in range(steps):
                o, d, gt = batches[i]
                st = step(o, d, near, far, gt)
                if average_last and i >= steps - average_last:
                    if sums is None:
                        sums = [m.params.detach().clone() for m in mods]
                    else:
                        for acc, m in zip(sums, mods):
                            acc.add_(m.params.detach())
                if i % 500 == 0 or i == steps - 1:
                    print(f"quality: seed {sd} {name} step {i} fine_loss {float(st['fine_loss']):.5f}",
                          file=sys.stderr, flush=True)
            torch.cuda.synchronize()
            t_train = time.perf_counter() - t0
            evaluate(net, name, sd, t_train)
            del step
            if sums is not None:
                # the same run, its weights averaged over the last `average_last` steps
                with torch.no_grad():
                    for acc, m in zip(sums, mods):
                        m.params.copy_(acc / float(average_last))
                        m.repack()
                evaluate(net, name + "+avg", sd, None)
                del sums
            if shared and name == "plain_fp32":
                # the same trained weights rendered by the other forms: the error of
                # the form itself, without the spread of separate training runs
                w_c, w_f = net.coarse_model.get_weights(), net.fine_model.get_weights()
                nb = nerf.NeRF(P, "bf16", seed=sd, device=device)
                nb.coarse_model.set_weights(w_c)
                nb.fine_model.set_weights(w_f)
                evaluate(nb, "plain_bf16@plain_fp32_weights", sd, None)
                del nb
                pn = nerf.ProposalNeRF(P, "bf16", seed=sd, device=device)
                pn.coarse_model.set_weights(w_c)
                pn.fine_model.set_weights(w_f)
                # what the trained proposal gains: the untrained one, and the fine model on
                # the same number of stratified samples without any proposal
                evaluate(_NoFallback(pn), "learned_bf16_untrained_nofallback@plain_fp32_weights", sd, None)
                evaluate(_Stratified(nerf, pn.fine_model, pn.n_prop + pn.n_fine), "stratified176_bf16@plain_fp32_weights",
                         sd, None)
                torch.manual_seed(3000 + sd)
                for i in range(steps):
                    o, d, _ = batches[i]
                    st = pn.train_proposal_step(o, d, near, far)
                    if i % 500 == 0 or i == steps - 1:
                        print(f"quality: seed {sd} proposal fit step {i} interlevel {float(st['interlevel']):.5f} "
                              f"fallback {float(st['fallback_fraction']):.4f}", file=sys.stderr, flush=True)
                evaluate(pn, "learned_bf16@plain_fp32_weights", sd, None)
                evaluate(_NoFallback(pn), "learned_bf16_nofallback@plain_fp32_weights", sd, None)
                del pn
            del net
            torch.cuda.empty_cache()
    # PSNR change against plain FP32 (final weights against final weights, averaged
    # against averaged), per seed and averaged over the seeds
    summary = {}
    by = {(r["seed"], r["form"]): r for r in res["runs"]}
    names = sorted({r["form"] for r in res["runs"]}, key=lambda f: [r["form"] for r in res["runs"]].index(f))
    for name in names:
        ref = "plain_fp32+avg" if name.endswith("+avg") else "plain_fp32"
        if name == ref or not all((sd, ref) in by and (sd, name) in by for sd in seeds):
            continue
        deltas = [round(by[(sd, name)]["psnr_mean"] - by[(sd, ref)]["psnr_mean"], 3) for sd in seeds]
        summary[name] = {"against": ref, "psnr_delta_db_per_seed": deltas,
                         "psnr_delta_db_mean": round(float(np.mean(deltas)), 3),
                         "render_speedup": round(np.mean([by[(sd, name)]["render_rays_per_s"] /
                                                          by[(sd, ref)]["render_rays_per_s"] for sd in seeds]), 2)}
        if all(by[(sd, name)]["train_rays_per_s"] and by[(sd, ref)]["train_rays_per_s"] for sd in seeds):
            summary[name]["train_speedup"] = round(np.mean([by[(sd, name)]["train_rays_per_s"] /
                                                            by[(sd, ref)]["train_rays_per_s"]
                                                            for sd in seeds]), 2)
    res["summary"] = summary
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--eval-views", type=int, default=4)
    ap.add_argument("--seeds", default="0,1")
    ap.add_argument("--forms", default="plain_fp32,plain_bf16,learned_bf16")
    ap.add_argument("--shared", action="store_true",
                    help="also render the trained plain FP32 weights with the BF16 forms (the proposal "
                         "fitted to them for the same number of steps)")
    ap.add_argument("--average-last", type=int, default=500,
                    help="also evaluate each run with its weights averaged over its last N steps")
    args = ap.parse_args(argv)
    print(json.dumps(run(args.steps, args.eval_views, tuple(int(x) for x in args.seeds.split(",")),
                         forms_sel=tuple(args.forms.split(",")), shared=args.shared,
                         average_last=args.average_last)))


if __name__ == "__main__":
    sys.exit(main())
