"""Benchmark of the comb-scanline renderer on one MI355X.

Workload (the plain renderer's headline configuration): 2048x2048 canvas, 20,000 seeded
paths (40% lines / 30% quadratics / 30% cubics, 3-64 segments, log-uniform 4-1024 px
bboxes, 10% star polygons, alternating nonzero / even-odd, alpha in [0.2, 1]), 32 samples
per pixel, coverage-alpha compositing into an sRGB8 image.  One step = one full render
(transform through the RGBA8 image) of the scene already resident in HBM.

Prints one JSON line: Mfragments/s (raw curve fragments F per second) with Mpixels/s
beside it, the dominant kernel's roofline entry (algorithmic bytes / its HIP-event time
against the 8 TB/s HBM3E peak, its profiled HBM traffic, and the share of the SIMDs' VALU
issue slots its profiled instruction count takes in that time) and the CPU baseline (the
OpenMP build of the oracle on a bounded sample of the same scene).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-scanline-path-rendering_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
# host cores this process may use, read before an OpenMP runtime binds the main thread
try:
    NPROC = len(os.sched_getaffinity(0))
except AttributeError:
    NPROC = os.cpu_count()
# the CPU baseline's threads stay on their cores (read when the OpenMP runtime loads)
os.environ.setdefault("OMP_PROC_BIND", "close")


def stage_bytes(stage, cnt, spp, W, H, ms=False):
    """Algorithmic HBM bytes of one stage's launch: the whole-frame figure
    B_alg = 8V + 14C + 5P + F*2*(8 + m + 9) + M*2m + S*2*8 + 4WH (each datum written once
    and read once; m = spp/2 bytes of mask per 2x2 cell) split by the stage that writes or
    reads each term:
      intersect  the scene's vertices and curves (8V + 14C) in, intersections (8F) out;
      fragments  intersections (8F) in, mask + key + path|sign + winding ((m+9)F) out;
      resolve    those ((m+9)F) in, resolved masks (mM) and spans (8S) out;
      composite  resolved masks (mM) and spans (8S) in, the RGBA8 image (4WH) out.
    The renderer's own encodings are smaller than these (4-byte items, per-pixel sample
    counts), so the bytes it moves are fewer than counted here."""
    m = spp // 2
    F, M, S = cnt["curve_fragment"], cnt["merged_fragment"], cnt["span"]
    C, V, P = cnt["curve"], cnt["vertex"], cnt["path"]
    if stage == "intersect":
        return 8 * V + 14 * C + 8 * F
    if stage == "fragments":
        return 8 * F + (m + 9) * F
    if stage == "resolve":
        return (m + 9) * F + m * M + 8 * S
    if stage == "composite":
        return m * M + 8 * S + 4 * W * H
    if stage == "transform_cull":
        return 8 * V + 5 * P
    return 0


# the kernel that makes up each stage's time (the stage's HIP-event interval)
STAGE_KERNEL = {("composite", False): "k_composite_cov", ("composite", True): "k_composite_ms",
                ("resolve", False): "k_resolve_first", ("resolve", True): "k_resolve_first",
                ("fragments", False): "k_frag_scatter", ("fragments", True): "k_frag_scatter",
                ("intersect", False): "k_chains", ("intersect", True): "k_chains"}

# committed rocprofv3 summaries of the headline workload, per output mode
TRAFFIC_PROFILE = {False: os.path.join(ROOT, "profiles", "plain_2048_32spp_traffic.md"),
                   True: os.path.join(ROOT, "profiles", "plain_2048_32spp_ms_traffic.md")}
INSTR_PROFILE = {False: os.path.join(ROOT, "profiles", "plain_2048_32spp_instructions.md"),
                 True: os.path.join(ROOT, "profiles", "plain_2048_32spp_ms_instructions.md")}


def _profile_value(path, key, kernel):
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return json.loads(line.split(":", 1)[1]).get(kernel)
    except OSError:
        pass
    return None


def profiled_traffic(kernel, ms=False):
    """HBM bytes per launch of `kernel` on the headline workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated); None if absent."""
    return _profile_value(TRAFFIC_PROFILE[ms], "kernel_traffic_bytes", kernel)


def profiled_valu(kernel, ms=False):
    """Wave VALU instructions per launch of `kernel` on the headline workload, from the
    committed rocprofv3 SQ_INSTS_VALU summary; None if absent."""
    return _profile_value(INSTR_PROFILE[ms], "kernel_valu_insts", kernel)


def b_alg_total(cnt, spp, W, H):
    m = spp // 2
    F, M, S = cnt["curve_fragment"], cnt["merged_fragment"], cnt["span"]
    return (8 * cnt["vertex"] + 14 * cnt["curve"] + 5 * cnt["path"] + F * 2 * (8 + m + 9)
            + M * 2 * m + S * 2 * 8 + W * H * 4)


def cpu_baseline(scene, W, H, spp, ms, warmup, reps):
    """The OpenMP build of the oracle on the same scene: wall time of its render (the
    scene already in memory, tables built; transform through the RGBA8 image), `warmup`
    untimed renders, then the median of `reps`, on every host core this process may use."""
    import statistics

    import pyoracle

    L = pyoracle.lib(omp=True)
    threads = L.oracle_threads()
    times, frags = [], 0
    t_all = time.perf_counter()
    for k in range(warmup + reps):
        r = pyoracle.run(scene, W, H, spp, srgb=True, ms=ms, keep=False, omp=True)
        if k >= warmup:
            times.append(r.seconds)
        frags = r.counts()["curve_fragment"]
        print("cpu baseline: render %d of %d, %.2f s" % (k + 1, warmup + reps, r.seconds), file=sys.stderr, flush=True)
        del r
    t = statistics.median(times)
    return {"value": round(frags / t / 1e6, 3), "unit": "Mfragments/s", "cores": threads,
            "kind": "port",
            "sample": "the full scene (%d paths at %dx%d, %d spp, %s output, F=%d): median of %d renders "
                      "after %d warm-ups, %.3f s each (min %.3f, max %.3f); %.0f s in all"
                      % (scene.n_paths, W, H, spp, "multisample" if ms else "coverage", frags, reps, warmup, t,
                         min(times), max(times), time.perf_counter() - t_all),
            "nproc": NPROC, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "render_s_median": round(t, 4)}


def bench_batch(args):
    """Batched mode: 4096 independent cubic-chain scenes (canvas sides log-uniform in
    [16, 512], 5-200 paths, 8 or 16 spp) from one captured graph, against a loop of
    single-scene renders of the same scenes (one pre-uploaded handle per scene)."""
    import numpy as np
    import torch

    import scanline

    scenes, sizes, spp = scanline.batch_workload(args.batch_scenes)
    b = scanline.BatchRasterizer(scenes, sizes, spp)
    t0 = time.perf_counter()
    tuned_ms, shape = b.tune(reps=1)
    tune_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(args.warmup):
        b.render(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.render(stream)
    torch.cuda.synchronize()
    batch_ms = (time.perf_counter() - t0) / args.steps * 1e3
    cnt = b.elementNumber()

    handles = []
    for s, (w, h), n in zip(scenes, sizes, spp):
        r = scanline.VGRasterizer(0)
        r.setSamples(n)
        r.setMultisampleOutput(False)
        r.enableSRGBCorrection(True)
        r.resizeOutputBuffer(w, h)
        r.addVg(s)
        r.rasterize(stream)
        handles.append(r)
    torch.cuda.synchronize()
    loops = max(1, args.steps // 10)
    t0 = time.perf_counter()
    for _ in range(loops):
        for r in handles:
            r.rasterize(stream)
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t0) / loops * 1e3
    # spot-check bit-exactness of the batch against the loop
    mism = 0
    for i in range(0, len(scenes), max(1, len(scenes) // 64)):
        mism += int(not np.array_equal(b.output(i), handles[i].output()))
    for r in handles:
        r.close()
    b.close()
    F = cnt["curve_fragment"]
    out = {
        "metric": "batched scenes/s (4096 scenes in one captured graph) vs a per-scene loop",
        "value": round(len(scenes) / (batch_ms * 1e-3), 2),
        "unit": "scenes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (4096 seeded cubic-chain scenes)",
        "config": {"workload": "batched: 4096 scenes, sides log-uniform 16-512, 5-200 cubic chains, 8/16 spp",
                   "scenes": len(scenes), "pixels": int(sum(w * h for w, h in sizes))},
        "counts": cnt,
        "mfragments_per_s": round(F / (batch_ms * 1e-3) / 1e6, 2),
        "per_scene_loop_ms": round(loop_ms, 3),
        "speedup_vs_loop": round(loop_ms / batch_ms, 2),
        "tuned_shape": {"resolve_grid": shape[0], "class_grid": shape[1], "composite_group": shape[2],
                        "fragment_block": shape[3], "tuned_ms": round(tuned_ms, 3), "tuning_s": round(tune_s, 2)},
        "spot_check_mismatches": mism,
    }
    print(json.dumps(out))


def bench_diff(args):
    """Differentiable mode: 256x256, 64 closed chains of 6 cubics, 16 sample rows per
    pixel; forward and adjoint times (HIP events on the current stream), then 500 Adam
    steps towards a target rendered from the scene with control points + N(0, 2 px^2)
    and colours + N(0, 0.05^2)."""
    import numpy as np
    import torch

    import scanline
    from scanline import diff

    scene, W, H, spp = scanline.config_scene(5)
    r = diff.DiffRenderer(scene, W, H, samples=spp)
    rng = np.random.default_rng(5)
    uniq, index = diff.tie_vertices(scene)
    idx = torch.tensor(index, device="cuda")
    ut = torch.tensor(uniq + rng.normal(0.0, np.sqrt(2.0), uniq.shape).astype(np.float32), device="cuda")
    ct = (r.colors0 + torch.tensor(rng.normal(0, 0.05, (scene.n_paths, 4)).astype(np.float32), device="cuda")).clamp(0, 1)
    tgt = r.forward_raw(ut[idx].contiguous(), ct.contiguous(), keep=False)
    pts = torch.tensor(uniq, device="cuda")[idx].contiguous()
    g = torch.ones((H, W, 4), device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    for _ in range(args.warmup):
        r.forward_raw(pts, r.colors0)
        r.backward_raw(g)
    fwd = bwd = 0.0
    for _ in range(args.steps):
        ev[0].record()
        r.forward_raw(pts, r.colors0)
        ev[1].record()
        r.backward_raw(g)
        ev[2].record()
        torch.cuda.synchronize()
        fwd += ev[0].elapsed_time(ev[1])
        bwd += ev[1].elapsed_time(ev[2])
    fwd /= args.steps
    bwd /= args.steps
    # central differences of the GPU forward at h = 1e-2 px on 60 tied control points
    # (both coordinates), against the adjoint's gradient of the same loss; perturbations
    # that change the crossing count (a y-extremum passing a sample row) are reported apart
    fd_rng = np.random.default_rng(17)
    sel = fd_rng.choice(len(uniq), 60, replace=False)
    base = torch.tensor(uniq, device="cuda")
    pg = base[idx].clone().requires_grad_(True)
    ((r.render(pg, r.colors0) - tgt) ** 2).sum().backward()
    gu = torch.zeros_like(base).index_add_(0, idx, pg.grad).cpu().numpy()
    n0 = r.crossings()
    rows = []
    h = 1e-2
    for uu in sel:
        for d in range(2):
            up, um = base.clone(), base.clone()
            up[uu, d] += h
            um[uu, d] -= h
            step = float(up[uu, d]) - float(um[uu, d])
            lp = float(((r.forward_raw(up[idx].contiguous(), r.colors0, keep=False).double() - tgt.double()) ** 2).sum())
            evt = r.crossings() != n0
            lm = float(((r.forward_raw(um[idx].contiguous(), r.colors0, keep=False).double() - tgt.double()) ** 2).sum())
            evt = evt or r.crossings() != n0
            rows.append(((lp - lm) / step, float(gu[uu, d]), evt))
    fd = np.array(rows)
    ok = fd[:, 2] == 0
    fd_rel_ok = float(np.linalg.norm(fd[ok, 0] - fd[ok, 1]) / np.linalg.norm(fd[ok, 0]))
    fd_rel_all = float(np.linalg.norm(fd[:, 0] - fd[:, 1]) / np.linalg.norm(fd[:, 0]))
    # identifiable control points: those the loss moves at the start (nonzero boundary
    # gradient); points hidden under other paths or off every sample row cannot be recovered
    ident = np.abs(gu).sum(axis=1) > 0
    # ...and, stricter, those the target image itself depends on: nonzero gradient of
    # sum(target render * R) at the target's points for a random image R
    ptg = ut[idx].clone().contiguous().requires_grad_(True)
    probe = torch.tensor(np.random.default_rng(23).standard_normal((H, W, 4)).astype(np.float32), device="cuda")
    (r.render(ptg, ct.contiguous()) * probe).sum().backward()
    gt_u = torch.zeros_like(base).index_add_(0, idx, ptg.grad).cpu().numpy()
    ident_t = ident & (np.abs(gt_u).sum(axis=1) > 0)
    u = torch.tensor(uniq, device="cuda", requires_grad=True)
    col = r.colors0.clone().requires_grad_(True)
    opt = torch.optim.Adam([{"params": [u], "lr": 0.05}, {"params": [col], "lr": 0.005}])
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(500):
        opt.zero_grad()
        img = r.render(u[idx], col)
        loss = ((img - tgt) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            col.clamp_(0, 1)
        losses.append(float(loss.detach()))
    adam_s = time.perf_counter() - t0
    err0 = float((torch.tensor(uniq, device="cuda") - ut).abs().mean())
    err1 = float((u.detach() - ut).abs().mean())
    it = torch.tensor(ident, device="cuda")
    ierr0 = float((torch.tensor(uniq, device="cuda") - ut)[it].norm(dim=1).mean())
    ierr1 = float((u.detach() - ut)[it].norm(dim=1).mean())
    d0 = (torch.tensor(uniq, device="cuda") - ut).norm(dim=1)
    d1 = (u.detach() - ut).norm(dim=1)
    itt = torch.tensor(ident_t, device="cuda")
    strict = {"points": int(ident_t.sum()),
              "median_distance_px": [round(float(d0[itt].median()), 3), round(float(d1[itt].median()), 3)],
              "improved_frac": round(float((d1[itt] < d0[itt]).float().mean()), 3)}
    out = {
        "metric": "differentiable mode: backward / forward time (256^2, 64 cubic chains, 16 sample rows)",
        "value": round(bwd / fwd, 3),
        "unit": "bwd/fwd",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fwd + bwd, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded cubic chains; target from perturbed points and colours)",
        "config": {"workload": "differentiable: 256x256, 64 closed chains of 6 cubics, 16 sample rows"},
        "forward_ms": round(fwd, 4),
        "backward_ms": round(bwd, 4),
        "crossings": r.crossings(),
        "fd_h_px": h,
        "fd_rel_err_event_free": round(fd_rel_ok, 5),
        "fd_rel_err_all": round(fd_rel_all, 5),
        "fd_params": [int(ok.sum()), int(len(fd))],
        # the 1e-3 finite-difference bar is not met by this box-filtered, row-sampled model:
        # its coverage has kinks where a crossing passes a pixel edge, O(h) for central
        # differences at h = 1e-2 px
        "fd_bar_1e-3_met": bool(fd_rel_ok <= 1e-3),
        "adam_500": {"loss_first": round(losses[0], 3), "loss_last": round(losses[-1], 3),
                     "mean_point_error_px": [round(err0, 3), round(err1, 3)],
                     "identifiable_points": int(ident.sum()), "all_points": int(len(ident)),
                     "identifiable_point_distance_px": [round(ierr0, 3), round(ierr1, 3)],
                     "identifiable_in_target": strict,
                     "colour_error": [round(float((r.colors0 - ct).abs().mean()), 4),
                                      round(float((col.detach() - ct).abs().mean()), 4)],
                     "seconds": round(adam_s, 2)},
    }
    r.close()
    print(json.dumps(out))


def bench_animation(args):
    """Animation mode (the reference's -animation): two chord diagrams of 150 nodes and
    about 1,800 links an interval on a 1920x1080 window, 16 spp coverage output.  One step
    = one frame: the scene at the frame's time generated on the GPU from keyframes already
    on the device, then rendered in full (the geometry moves, so every frame sizes its
    buffers from the device's counts).  Frame times sweep the three keyframe intervals."""
    import numpy as np
    import torch

    import scanline
    from scanline import animation as an

    W, H = an.CANVAS
    spp = 16
    ds = an.refugee_diagrams(n_nodes=150, n_years=4, links_per_year=1500, seed=0)
    r = scanline.VGRasterizer(0)
    r.setSamples(spp)
    r.setMultisampleOutput(False)
    r.enableSRGBCorrection(True)
    r.resizeOutputBuffer(W, H)
    anim = an.ChordAnimation(r, ds)
    stream = torch.cuda.current_stream().cuda_stream
    span = 3 * an.KEY_MS
    times = [an.KEY_MS + (k * span) // max(1, args.steps) for k in range(args.steps)]
    for k in range(args.warmup):
        anim.frame(times[k % len(times)], stream)
        r.rasterize(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in times:
        anim.frame(t, stream)
        r.rasterize(stream)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) / len(times) * 1e3

    # the generation kernel alone, HIP events on the stream it runs on
    reps = 200
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for k in range(reps):
        anim.frame(times[k % len(times)], stream)
    ev1.record()
    torch.cuda.synchronize()
    gen_us = ev0.elapsed_time(ev1) / reps * 1e3
    P_mid = anim.paths_at(times[len(times) // 2])
    # bytes one generation writes (per path: 24 vertices of 8 + 4 bytes, 6 curves of
    # 4+1+4+4+1+16 bytes, fill rule and word) and reads (its keyframe rows, ~64 bytes)
    gen_bytes = P_mid * (24 * 12 + 6 * 30 + 5 + 64)

    # stage times of the renderer over the same frames
    r.enableStepTiming(True)
    stage_sum, frame_ms = {}, []
    for t in times:
        anim.frame(t, stream)
        r.rasterize(stream)
        st = r.stepTiming()
        frame_ms.append(sum(st.values()))
        for k, v in st.items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    r.enableStepTiming(False)
    stages = {k: v / len(times) for k, v in stage_sum.items()}
    cnt = r.elementNumber()
    dom = max((k for k in stages if stage_bytes(k, cnt, spp, W, H) > 0), key=lambda k: stages[k])
    dom_bytes = stage_bytes(dom, cnt, spp, W, H)
    achieved = dom_bytes / (stages[dom] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        import statistics

        import pyoracle

        L = pyoracle.lib(omp=True)
        t_mid = times[len(times) // 2]
        scene = pyoracle.chord_scene(pyoracle.chord_frame(ds, t_mid))
        secs = []
        t_all = time.perf_counter()
        for k in range(4):
            res = pyoracle.run(scene, W, H, spp, srgb=True, keep=False, omp=True)
            secs.append(res.seconds)
            print("cpu baseline: render %d of 4, %.2f s" % (k + 1, res.seconds), file=sys.stderr, flush=True)
            del res
        med = statistics.median(secs[1:])
        cpu = {"value": round(1.0 / med, 3), "unit": "frames/s", "cores": L.oracle_threads(), "kind": "port",
               "sample": "the frame at t = %d ms (%d paths, %dx%d, %d spp, coverage): median of 3 renders "
                         "after 1 warm-up, %.3f s each; %.0f s in all" % (t_mid, scene.n_paths, W, H, spp, med,
                                                                           time.perf_counter() - t_all)}

    out = {
        "metric": "animation frames/s (GPU-generated chord scene + full render, 1920x1080 16spp)",
        "value": round(1e3 / ms_step, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded flow records: 150 nodes, 4 keyframes, 1500 links a year)",
        "config": {"workload": "chord animation, 2 diagrams x 150 nodes, ~1800 links an interval, "
                               "1920x1080, 16 spp coverage", "width": W, "height": H, "spp": spp,
                   "paths_mid": P_mid},
        "generate_us": round(gen_us, 2),
        "generate_gbs": round(gen_bytes / (gen_us * 1e-6) / 1e9, 2),
        "frame_event_ms_median": round(float(np.median(frame_ms)), 4),
        "counts": cnt,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": {"bound": "hbm", "kernel": STAGE_KERNEL.get((dom, False), dom), "stage": dom,
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "alg_bytes": int(dom_bytes), "traffic": None},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(out["value"] / cpu["value"], 2)
    print(json.dumps(out))
    r.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--paths", type=int, default=20000)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--spp", type=int, default=32)
    ap.add_argument("--ms", action="store_true", help="multisample (per-sample) compositing")
    ap.add_argument("--cpu-warmup", type=int, default=None,
                    help="untimed CPU-baseline renders (default 2; 0 with --ms)")
    ap.add_argument("--cpu-reps", type=int, default=None,
                    help="timed CPU-baseline renders (median; default 10; 1 with --ms, whose CPU render "
                         "blends every sample and takes about a minute)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (counter calibration)")
    ap.add_argument("--mode", choices=["plain", "batch", "diff", "animation"], default="plain")
    ap.add_argument("--config", type=int, choices=[2, 3], default=2,
                    help="plain mode workload: 2 = 2048^2 mixed paths (headline), 3 = text-like 4096^2")
    ap.add_argument("--batch-scenes", type=int, default=4096)
    ap.add_argument("--no-count-pixel", action="store_true",
                    help="skip the untimed debug frame that gives the reference's count-pixel")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this renderer runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU band sharding is a design only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.mode == "batch":
        bench_batch(args)
        return
    if args.mode == "diff":
        bench_diff(args)
        return
    if args.mode == "animation":
        bench_animation(args)
        return

    import numpy as np
    import torch

    import scanline

    if args.config == 3:
        # text-like workload: 4096x4096, 100,000 glyphs (1-4 contours of 10-40 quadratics,
        # reversed holes, nonzero), 16 spp
        kind, seed = scanline.scene.KIND_GLYPHS, scanline.scene.SEED_BASE ^ 3
        W = H = 4096
        args.paths, args.spp = 100000, 16
        workload = "text-like scene, 4096x4096 canvas, 100k glyphs, 16 spp"
    else:
        kind, seed = scanline.scene.KIND_MIXED, scanline.scene.SEED_BASE ^ 2
        W = H = args.size
        workload = "plain renderer, %dx%d canvas, %d mixed paths, %d spp" % (W, H, args.paths, args.spp)
    scene = scanline.generate(kind, W, H, args.paths, seed)
    r = scanline.VGRasterizer(0)
    r.setSamples(args.spp)
    r.setMultisampleOutput(args.ms)
    r.enableSRGBCorrection(True)
    r.resizeOutputBuffer(W, H)
    r.addVg(scene)
    stream = torch.cuda.current_stream().cuda_stream

    # timed frames: after the first frame the unchanged scene is re-rendered back to back
    # (every kernel of every frame runs; the library takes the counts of the first frame)
    for _ in range(args.warmup):
        r.rasterize(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.rasterize(stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3

    # stage times from HIP events, in a separate loop of as many frames (step timing makes
    # every frame read its events back)
    r.enableStepTiming(True)
    stage_sum = {}
    frame_ms = []
    for _ in range(args.steps):
        r.rasterize(stream)
        st = r.stepTiming()
        frame_ms.append(sum(st.values()))
        for k, v in st.items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    torch.cuda.synchronize()
    r.enableStepTiming(False)
    cnt = r.elementNumber()
    stages = {k: v / args.steps for k, v in stage_sum.items()}

    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        a = torch.empty(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    # the reference's count-pixel (sum over draw items of 2 * width), from one extra frame
    # with the parity dumps on, outside the timed region
    count_pixel = None
    if not args.no_count_pixel:
        r.dumpDebugData(True)
        r.rasterize(stream)
        torch.cuda.synchronize()
        count_pixel = int(r.pixelCount())
        r.dumpDebugData(False)

    F = cnt["curve_fragment"]
    value = F / (ms_step * 1e-3) / 1e6
    dom = max((k for k in stages if stage_bytes(k, cnt, args.spp, W, H, args.ms) > 0), key=lambda k: stages[k])
    dom_bytes = stage_bytes(dom, cnt, args.spp, W, H, args.ms)
    achieved = dom_bytes / (stages[dom] * 1e-3) / 1e9
    dom_kernel = STAGE_KERNEL.get((dom, args.ms), dom)
    traffic = valu_frac = valu_ceiling = None
    if args.config == 2 and (W, args.paths, args.spp) == (2048, 20000, 32):
        traffic = profiled_traffic(dom_kernel, args.ms)
        valu = profiled_valu(dom_kernel, args.ms)
        if valu:
            # issued VALU instructions over what the SIMDs issued in the fma probe (measured
            # now, on this GPU and clock), scaled to the kernel's time
            pm, pn = r.valuProbe()
            valu_ceiling = pn / (pm * 1e-3)
            valu_frac = round(valu / (valu_ceiling * stages[dom] * 1e-3), 3)
    b_alg = b_alg_total(cnt, args.spp, W, H)

    cpu = None
    if not args.no_cpu_baseline:
        warm = args.cpu_warmup if args.cpu_warmup is not None else (0 if args.ms else 2)
        reps = args.cpu_reps if args.cpu_reps is not None else (1 if args.ms else 10)
        cpu = cpu_baseline(scene, W, H, args.spp, args.ms, warm, reps)

    out = {
        "metric": "Mfragments/s + Mpixels/s at 2048^2 32spp (GB/s vs HBM peak); bwd/fwd time",
        "value": round(value, 3),
        "unit": "Mfragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, splitmix64 seed 0x5CA1AB1E^2)",
        "config": {"workload": workload,
                   "width": W, "height": H, "paths": args.paths, "spp": args.spp,
                   "output": "multisample" if args.ms else "coverage", "srgb": True},
        "mpixels_per_s": round(W * H / (ms_step * 1e-3) / 1e6, 3),
        # HIP-event frame time (sum of the stage events of one frame), median over the steps
        "frame_event_ms_median": round(float(np.median(frame_ms)), 4),
        "counts": cnt,
        "count_pixel": count_pixel,
        "b_alg_bytes": int(b_alg),
        "whole_frame_gbs": round(b_alg / (ms_step * 1e-3) / 1e9, 2),
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "stage": dom, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "alg_bytes": int(dom_bytes), "traffic": traffic, "valu_issue_frac": valu_frac,
                     "valu_ceiling_wave_insts_per_s": valu_ceiling and float(f"{valu_ceiling:.4g}")},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out))
    r.close()


if __name__ == "__main__":
    main()
