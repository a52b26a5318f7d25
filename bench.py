"""Benchmark: wavefront path tracing of the 1.31M-triangle workload on one MI355X.

Workload (one "step" = one full frame): icosphere subdivided to level 8
(1,310,720 triangles) radially displaced by seeded 3-octave value noise of
amplitude 0.05, 64 GGX rough-conductor face groups, rough-dielectric ground
(IOR 1.5), one quad area light; path integrator (NEE + MIS), 1920x1080,
64 spp, max depth 12, SAH BVH with max_prim 4.  The scene is uploaded to HBM
before the timed region; frames render into device buffers.

Prints one JSON line with Mrays/s (closest-hit + any-hit BVH queries per
second), the roofline of the dominant kernel (HIP-event timing on the stream
the kernel runs on) and the CPU baseline (the oracle, a scalar C++
restatement of the reference's algorithm, on a bounded crop).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s on 1 GPU at equal RMSE vs CPU ref; batched speedup; adjoint/primal ratio"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per query: ray in (32) + hit out (16 closest, 4 any-hit)
# + 32 per node visited + 36 per triangle tested
RAY_IN = 32
# traversal stages -> (query class, hit-out bytes) of the rays they trace;
# classes as in rd_render_stats: 0 extension, 1 MIS, 2 shadow
TRACE_STAGES = {"trace_closest": [(0, 16)], "trace_mis": [(1, 16)], "trace_any": [(2, 4)],
                "trace_secondary": [(1, 16), (2, 4)]}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_scene(args):
    from radiantium_amd import scenes
    return scenes.displaced_icosphere_scene(level=args.level, width=args.width, height=args.height,
                                            max_depth=args.max_depth, seed=0)


def cpu_baseline(desc, args, gpu_rate):
    """Times the oracle on a centred crop of the same workload with the box's
    host cores; the crop is sized from a short calibration to ~15 s."""
    from oracle import cpuref
    if not os.path.exists(cpuref.LIB_PATH):
        cpuref.build()
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    o = cpuref.OracleScene(desc)
    log(f"[cpu] oracle BVH build {time.perf_counter() - t0:.1f}s, threads={threads}")
    W, H = args.width, args.height

    def crop(cw, ch):
        x0, y0 = (W - cw) // 2, (H - ch) // 2
        return (x0, y0, x0 + cw, y0 + ch)

    # calibrate on a centred crop, then time the whole frame when it fits the
    # budget (the same workload the GPU renders), else the crop that does
    cw, ch = 192, 108
    _, _, st, secs = o.render(args.spp, seed=1, crop=crop(cw, ch), threads=threads, with_sum2=False)
    full = secs * (W * H) / (cw * ch)
    log(f"[cpu] calibration {cw}x{ch}: {secs:.2f}s, full frame projected {full:.1f}s")
    if full <= 3 * args.cpu_seconds:
        cw, ch = W, H
    else:
        scale = (args.cpu_seconds / max(secs, 1e-3)) ** 0.5
        cw, ch = min(W, max(16, int(cw * scale) // 4 * 4)), min(H, max(16, int(ch * scale) // 4 * 4))
    _, _, st, secs = o.render(args.spp, seed=1, crop=crop(cw, ch), threads=threads, with_sum2=False)
    rays = st["rays_closest"] + st["rays_any"]
    rate = rays / secs / 1e6
    log(f"[cpu] crop {cw}x{ch} spp {args.spp}: {rays} rays in {secs:.2f}s = {rate:.3f} Mrays/s")
    what = "the whole frame" if (cw, ch) == (W, H) else f"a centred {cw}x{ch} crop of the frame"
    return {"value": round(rate, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{what} at {args.spp} spp, seed 1: {rays} BVH queries in {secs:.1f} s"}


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "traffic_icosphere_l8_1080p_64spp.json")


def pmc_traffic(stage, args):
    """HBM bytes per launch of `stage` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (scaled by a plain copy kernel's counters), when this run is
    the profiled workload; None otherwise."""
    if (args.width, args.height, args.spp, args.level, args.max_depth) != (1920, 1080, 64, 8, 12):
        return None
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)["stages"][stage]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def batched_speedup(ctx, args, small=False):
    """Batched workload: 16 vol_path Cornell boxes with a glass sphere of
    homogeneous medium each (sigma_a ~ U[0.01,0.2], sigma_s ~ U[0.5,5],
    g ~ U[-0.3,0.9]), resolutions from {128^2..1024^2}, spp from {16,64,256},
    max depth 64.  Sequential = one rd_render_device per request; batched =
    one rd_render_batch_device over all of them.  Same outputs (bit-identical)."""
    from radiantium_amd import native, scenes
    if small:
        # the same scene family with the small resolutions and spp only: the
        # serving regime where one request alone cannot fill the GPU
        reqs = scenes.batch_requests(args.batch_seed, resolutions=(128, 256), spps=(16, 64))
    else:
        reqs = scenes.batch_requests(args.batch_seed)
    sc = [native.Scene(ctx, d) for d, _, _ in reqs]
    outs = [ctx.alloc(d.camera["width"] * d.camera["height"] * 12) for d, _, _ in reqs]
    paths = sum(d.camera["width"] * d.camera["height"] * spp for d, spp, _ in reqs)
    # warm both paths once at low spp
    for s_, o in zip(sc, outs):
        s_.render_device(1, o, seed=7)
    ctx.render_batch_device([(s_, 1, 7, o, None) for s_, o in zip(sc, outs)])
    ctx.synchronize()
    t0 = time.perf_counter()
    rays_seq = 0
    for s_, (d, spp, seed), o in zip(sc, reqs, outs):
        s_.render_device(spp, o, seed=seed)
        st = ctx.render_stats()
        rays_seq += st["rays_closest"] + st["rays_any"]
    ctx.synchronize()
    seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx.render_batch_device([(s_, spp, seed, o, None) for s_, (d, spp, seed), o in zip(sc, reqs, outs)])
    ctx.synchronize()
    bat = time.perf_counter() - t0
    st = ctx.render_stats()
    rays_bat = st["rays_closest"] + st["rays_any"]
    for o in outs:
        ctx.free(o)
    log(f"[bench] batched: {len(reqs)} requests, {paths / 1e6:.0f} M paths; sequential {seq * 1000:.0f} ms, "
        f"batched {bat * 1000:.0f} ms")
    return {"speedup": round(seq / bat, 3), "sequential_ms": round(seq * 1000, 1), "batched_ms": round(bat * 1000, 1),
            "requests": len(reqs), "paths": paths, "rays": rays_bat, "rays_equal": rays_seq == rays_bat,
            "batched_mrays_s": round(rays_bat / bat / 1e6, 1),
            "workload": f"16 vol_path glass-medium Cornell boxes, seed {args.batch_seed}, "
                        + ("res {128,256}^2, spp {16,64}" if small else "res 128^2-1024^2, spp 16-256") + ", depth 64"}


def adjoint_ratio(ctx, args):
    """Differentiable-rendering workload: the Cornell box with the side walls
    GGX rough plastic carrying 512x512 RGB albedo and roughness textures,
    image 512x512.  Target = 256 spp render of seeded random textures; the
    optimisation starts from uniform 0.5 textures; a step is a 64 spp primal
    plus a 64 spp path-replay adjoint of the L2 loss, then an Adam update.
    The ratio is timed on direct rd_render_device / rd_adjoint_device calls
    (same spp and seed); the optimisation loop runs through the autograd
    Function and reports its loss."""
    import torch
    from radiantium_amd import native, scenes
    from radiantium_amd.autograd import render
    R, T, spp = 512, 512, 64
    tdesc, _ = scenes.diff_cornell(R, R, tex_res=T, seed=args.adj_seed)
    tscene = native.Scene(ctx, tdesc)
    target = torch.empty(R * R * 3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    tscene.render_device(256, target.data_ptr(), seed=77)
    ctx.synchronize()
    tscene.close()
    desc, tex = scenes.diff_cornell(R, R, tex_res=T, init=(0.5, 0.5))
    scene = native.Scene(ctx, desc)
    ids = sorted(tex.values())
    img = torch.empty(R * R * 3, dtype=torch.float32, device="cuda")
    grads = {t: torch.zeros(T * T * 3, dtype=torch.float32, device="cuda") for t in ids}
    ptrs = [None] * len(desc.textures)
    for t in ids:
        ptrs[t] = grads[t].data_ptr()
    scene.render_device(4, img.data_ptr(), seed=100)
    ctx.synchronize()
    dl = (2 * (img - target)).contiguous()
    torch.cuda.synchronize()
    scene.adjoint_device(4, dl.data_ptr(), ptrs, seed=100)
    ctx.synchronize()
    prim, adj, prim_dev, adj_dev, rays_p, rays_a = [], [], [], [], 0, 0
    for i in range(args.adj_reps):
        t0 = time.perf_counter()
        scene.render_device(spp, img.data_ptr(), seed=i)
        ctx.synchronize()
        prim.append(time.perf_counter() - t0)
        st = ctx.render_stats()
        prim_dev.append(st["render_ms"])
        rays_p += st["rays_closest"] + st["rays_any"]
        dl = (2 * (img - target)).contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scene.adjoint_device(spp, dl.data_ptr(), ptrs, seed=i)
        ctx.synchronize()
        adj.append(time.perf_counter() - t0)
        st = ctx.render_stats()
        adj_dev.append(st["render_ms"])
        rays_a += st["rays_closest"] + st["rays_any"]
    p_ms, a_ms = float(np.median(prim)) * 1000, float(np.median(adj)) * 1000
    log(f"[bench] adjoint: primal {p_ms:.1f} ms, adjoint {a_ms:.1f} ms, ratio {a_ms / p_ms:.2f}")
    params = {t: torch.full((T * T * 3,), 0.5, dtype=torch.float32, device="cuda", requires_grad=True) for t in ids}
    rough = {tex[k] for k in tex if k.endswith("rough")}
    opt = torch.optim.Adam(list(params.values()), lr=0.02)
    tgt = target.view(R, R, 3)
    losses, step_ms = [], []
    for k in range(args.adj_steps):
        t0 = time.perf_counter()
        opt.zero_grad()
        im = render(scene, params, spp, seed=1000 + k)
        loss = ((im - tgt) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            for t, x in params.items():
                x.clamp_(0.05, 1.0) if t in rough else x.clamp_(0.0, 1.0)
        losses.append(float(loss.detach()))
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - t0) * 1000)
        log(f"[bench] adjoint opt step {k}: loss {losses[-1]:.2f}, {step_ms[-1]:.0f} ms")
    scene.close()
    return {"ratio": round(a_ms / p_ms, 3), "primal_ms": round(p_ms, 1), "adjoint_ms": round(a_ms, 1),
            "primal_device_ms": round(float(np.median(prim_dev)), 1),
            "adjoint_device_ms": round(float(np.median(adj_dev)), 1),
            "primal_mrays_s": round(rays_p / sum(prim) / 1e6, 1), "adjoint_mrays_s": round(rays_a / sum(adj) / 1e6, 1),
            "opt_steps": args.adj_steps, "opt_step_ms": round(float(np.median(step_ms)), 1) if step_ms else None,
            "loss_first": round(losses[0], 3) if losses else None, "loss_last": round(losses[-1], 3) if losses else None,
            "workload": f"Cornell box, rough-plastic side walls with {T}^2 albedo + roughness textures, "
                        f"{R}x{R}, {spp} spp primal + {spp} spp adjoint, max depth 8, target 256 spp "
                        f"(texture seed {args.adj_seed}), start 0.5"}


def instanced_throughput(ctx, args):
    """Instanced workload: 1024 InstancedPrimitives of 8 superquadrics
    (99,904 triangles each, exponents in [0.5, 2.5]; 102 M instanced
    triangles) with random rotations and scales in [0.5, 2], Disney / plastic
    / rough-metal materials, a 2048 x 1024 log-normal HDR sky with a sun disk
    importance-sampled through Distribution2D; path Mis, 1024 x 1024,
    128 spp, max depth 16.  One frame, timed like the main workload, and the
    oracle on a centred crop of the same frame."""
    from radiantium_amd import native, scenes
    from oracle import cpuref
    t0 = time.perf_counter()
    desc = scenes.instanced_workload()
    scene = native.Scene(ctx, desc)
    info = scene.info()
    log(f"[bench] instanced scene: {len(desc.instances)} instances, built and uploaded in "
        f"{time.perf_counter() - t0:.1f}s")
    W, H, spp = desc.camera["width"], desc.camera["height"], args.inst_spp
    d_rgb = ctx.alloc(W * H * 12)
    scene.render_device(1, d_rgb, seed=99)
    ctx.synchronize()
    t0 = time.perf_counter()
    scene.render_device(spp, d_rgb, seed=1)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    st = ctx.render_stats()
    ctx.free(d_rgb)
    rays = st["rays_closest"] + st["rays_any"]
    rate = rays / secs / 1e6
    log(f"[bench] instanced: {secs * 1000:.0f} ms, {rays / 1e6:.0f} M rays, {rate:.1f} Mrays/s")
    cpu = None
    if not args.no_cpu_baseline:
        if not os.path.exists(cpuref.LIB_PATH):
            cpuref.build()
        threads = max(1, min(16, os.cpu_count() or 1))
        o = cpuref.OracleScene(desc)
        cw = ch = 64
        crop = ((W - cw) // 2, (H - ch) // 2, (W + cw) // 2, (H + ch) // 2)
        _, _, cst, csecs = o.render(spp, seed=1, crop=crop, threads=threads, with_sum2=False)
        crays = cst["rays_closest"] + cst["rays_any"]
        cpu = {"value": round(crays / csecs / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"a centred {cw}x{ch} crop at {spp} spp: {crays} BVH queries in {csecs:.1f} s"}
        log(f"[cpu] instanced crop: {cpu['value']} Mrays/s")
    c = st["class_rays"]
    return {"mrays_s": round(rate, 1), "ms": round(secs * 1000, 1), "rays": rays,
            "nodes_per_ray": [round(n / max(r, 1), 2) for n, r in zip(st["class_nodes"], c)],
            "prims_per_ray": [round(n / max(r, 1), 2) for n, r in zip(st["class_prims"], c)],
            "stack_depth": info["max_depth"], "cpu_baseline": cpu,
            "speedup_vs_cpu": round(rate / cpu["value"], 1) if cpu else None,
            "workload": f"1024 instances of 8 superquadrics (99,904 tris each), 2048x1024 HDR sky + sun; "
                        f"path Mis, {W}x{H}, {spp} spp, max depth 16"}


def copy_calibration():
    """A plain device copy of known size, profiled beside the path kernels so
    FETCH_SIZE / WRITE_SIZE can be scaled to bytes."""
    import torch
    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)   # one elementwise kernel: reads n*4 bytes, writes n*4 bytes
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: 3 x MulFunctor over {n * 4} bytes read + written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--max-depth", type=int, default=12)
    ap.add_argument("--wave-paths", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-seed", type=int, default=0, help="seed of the batched workload (16 vol_path scenes)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-sequential measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/primal measurement")
    ap.add_argument("--no-instanced", action="store_true", help="skip the instanced-workload measurement")
    ap.add_argument("--inst-spp", type=int, default=128, help="spp of the instanced workload frame")
    ap.add_argument("--adj-seed", type=int, default=9, help="texture seed of the adjoint target")
    ap.add_argument("--adj-reps", type=int, default=3, help="timed primal/adjoint pairs")
    ap.add_argument("--adj-steps", type=int, default=6, help="optimisation steps through the autograd Function")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device copy (scale for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the path renders one frame per GPU and no multi-GPU "
              "figure is produced (the frame would shard over samples, with one all-reduce of the "
              "sum and sum2 buffers).", file=sys.stderr)
        sys.exit(2)

    from radiantium_amd import native

    if args.copy_calibration:
        copy_calibration()
    t0 = time.perf_counter()
    desc = make_scene(args)
    ntri = desc.arrays()["indices"].shape[0]
    log(f"[bench] scene generated: {ntri} triangles in {time.perf_counter() - t0:.1f}s")
    ctx = native.Context(0)
    t0 = time.perf_counter()
    scene = native.Scene(ctx, desc)
    info = scene.info()
    log(f"[bench] BVH {info['num_nodes']} nodes, depth {info['max_depth']}, build {info['build_ms']:.0f} ms, "
        f"upload total {time.perf_counter() - t0:.1f}s")
    npix = args.width * args.height
    d_rgb = ctx.alloc(npix * 12)
    for i in range(args.warmup):
        scene.render_device(args.spp, d_rgb, seed=1000 + i, wave_paths=args.wave_paths)
        log(f"[bench] warmup {i} done")
    ctx.synchronize()
    ctx.reset_kernel_timing()
    rays = 0
    agg = {"class_rays": [0, 0, 0], "class_nodes": [0, 0, 0], "class_prims": [0, 0, 0]}
    t_start = time.perf_counter()
    for i in range(args.steps):
        scene.render_device(args.spp, d_rgb, seed=i, wave_paths=args.wave_paths)
        st = ctx.render_stats()
        rays += st["rays_closest"] + st["rays_any"]
        for k in agg:
            agg[k] = [a + b for a, b in zip(agg[k], st[k])]
        log(f"[bench] step {i}: {st['render_ms']:.1f} ms, {(st['rays_closest'] + st['rays_any']) / 1e6:.1f} M rays "
            f"(extension / MIS / shadow: {' / '.join(f'{r / 1e6:.1f}' for r in st['class_rays'])} M, nodes per ray "
            f"{' / '.join(f'{n / max(r, 1):.1f}' for n, r in zip(st['class_nodes'], st['class_rays']))})")
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    value = rays / elapsed / 1e6

    stages = {}
    for name in ("compact", "sort", "trace_closest", "shade", "compact_shadow", "trace_secondary", "trace_any",
                 "trace_mis", "resolve", "accumulate", "finalize"):
        avg, cnt = ctx.kernel_timing(name)
        if not cnt:
            continue
        stages[name] = (avg, cnt)
        log(f"[bench] stage {name:15s} avg {avg * 1000:9.1f} us x {cnt:5d} = {avg * cnt:9.1f} ms")
    dom = max(stages, key=lambda k: stages[k][0] * stages[k][1])
    roofline = None
    traced = [k for k in TRACE_STAGES if k in stages]
    trace_dom = dom if dom in TRACE_STAGES else max(traced, key=lambda k: stages[k][0] * stages[k][1])
    avg_ms, cnt = stages[trace_dom]
    if cnt:
        cls = TRACE_STAGES[trace_dom]
        total_bytes = sum(agg["class_rays"][c] * (RAY_IN + out) + 32 * agg["class_nodes"][c]
                          + 36 * agg["class_prims"][c] for c, out in cls)
        rays_c = sum(agg["class_rays"][c] for c, _ in cls)
        per_launch = total_bytes / cnt
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roofline = {"kernel": trace_dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": pmc_traffic(trace_dom, args),
                    "bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg_ms, 4),
                    "nodes_per_ray": round(sum(agg["class_nodes"][c] for c, _ in cls) / max(rays_c, 1), 2),
                    "prims_per_ray": round(sum(agg["class_prims"][c] for c, _ in cls) / max(rays_c, 1), 2)}
    if dom != trace_dom and roofline is not None:
        roofline["note"] = f"longest stage is {dom}; roofline reported for the longest traversal kernel"

    batched = None
    if not args.no_batch:
        batched = batched_speedup(ctx, args)
        batched["small_requests"] = batched_speedup(ctx, args, small=True)

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_ratio(ctx, args)

    instanced = None
    if not args.no_instanced:
        instanced = instanced_throughput(ctx, args)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(desc, args, value)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1000, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"displaced icosphere level {args.level} ({ntri} triangles) + rough-glass ground "
                               f"+ quad light; path Mis; {args.width}x{args.height}, {args.spp} spp, max depth "
                               f"{args.max_depth}, bvh max_prim 4",
                   "width": args.width, "height": args.height, "spp": args.spp, "max_depth": args.max_depth,
                   "triangles": ntri, "parallelism": "1 GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "adjoint": adjoint,
        "instanced": instanced,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
