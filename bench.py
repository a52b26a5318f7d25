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


def make_scene(args, camera_z=None):
    from radiantium_amd import scenes
    return scenes.displaced_icosphere_scene(level=args.level, width=args.width, height=args.height,
                                            max_depth=args.max_depth, seed=0,
                                            camera_z=args.camera_z if camera_z is None else camera_z)


def host_cores():
    """Cores this process may run on (its CPU affinity; on the GPU box the
    container's cgroup quota equals it)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def primary_coverage(scene, desc, level, nx=192, ny=108):
    """Fraction of primary rays through an nx x ny grid of pixel centres that
    first meet the displaced mesh, the ground, or nothing (pinhole rays with
    the camera's horizontal fov; traced on the GPU).  The mesh's triangles
    are the scene's first 20 * 4^level primitives, the ground's the next two."""
    from radiantium_amd import abi
    c = desc.camera
    o = np.array(c["origin"], np.float64)
    z = np.array(c["target"], np.float64) - o
    z /= np.linalg.norm(z)
    x = np.cross(np.array(c["up"], np.float64), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    tx = np.tan(np.radians(c["fov"]) / 2)
    ty = tx * c["height"] / c["width"]
    U, V = np.meshgrid(((np.arange(nx) + 0.5) / nx * 2 - 1) * tx, ((np.arange(ny) + 0.5) / ny * 2 - 1) * ty)
    d = z + U.reshape(-1, 1) * x - V.reshape(-1, 1) * y
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(d.shape[0], abi.ray_dtype())
    rays["o"] = o.astype(np.float32)
    rays["d"] = d.astype(np.float32)
    rays["tmin"] = 1e-4
    rays["tmax"] = 3.4e38
    prim = scene.trace_closest(rays)["prim"]
    nmesh = 20 * 4 ** level
    return {"mesh": round(float(np.mean((prim >= 0) & (prim < nmesh))), 4),
            "ground": round(float(np.mean((prim >= nmesh) & (prim < nmesh + 2))), 4),
            "miss": round(float(np.mean(prim < 0)), 4)}


def close_framing(ctx, args):
    """The same workload with the camera at z = 4.6 instead of 7.0, where
    about half of the primary rays meet the mesh (more nodes per ray)."""
    from radiantium_amd import native
    desc = make_scene(args, camera_z=4.6)
    scene = native.Scene(ctx, desc)
    cov = primary_coverage(scene, desc, args.level)
    d_rgb = ctx.alloc(args.width * args.height * 12)
    scene.render_device(args.spp, d_rgb, seed=1000)
    ctx.synchronize()
    rays, agg = 0, {"class_rays": [0, 0, 0], "class_nodes": [0, 0, 0]}
    t0 = time.perf_counter()
    for i in range(2):
        scene.render_device(args.spp, d_rgb, seed=i)
        st = ctx.render_stats()
        rays += st["rays_closest"] + st["rays_any"]
        for k in agg:
            agg[k] = [a + b for a, b in zip(agg[k], st[k])]
    ctx.synchronize()
    secs = time.perf_counter() - t0
    ctx.free(d_rgb)
    scene.close()
    rate = rays / secs / 1e6
    log(f"[bench] close framing (camera z 4.6): {rate:.1f} Mrays/s, primary rays {cov}")
    return {"mrays_s": round(rate, 1), "ms_per_frame": round(secs / 2 * 1000, 1), "camera_z": 4.6,
            "primary_rays": cov,
            "nodes_per_ray": [round(n / max(r, 1), 2) for n, r in zip(agg["class_nodes"], agg["class_rays"])]}


def cpu_baseline(desc, args, gpu_rate, scene):
    """Times the oracle on the same workload with every host core this
    process may use: the whole frame when a short calibration says it fits
    ~45 s, else a centred crop sized to ~15 s.  Then checks equal RMSE: the
    GPU renders the same pixels at the same spp with an independent seed,
    and both images are compared with a 16x-spp GPU render of them."""
    from oracle import cpuref
    if not os.path.exists(cpuref.LIB_PATH):
        cpuref.build()
    threads = host_cores()
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
    cr = crop(cw, ch)
    cpu_img, _, st, secs = o.render(args.spp, seed=1, crop=cr, threads=threads, with_sum2=False)
    rays = st["rays_closest"] + st["rays_any"]
    rate = rays / secs / 1e6
    log(f"[cpu] crop {cw}x{ch} spp {args.spp}: {rays} rays in {secs:.2f}s = {rate:.3f} Mrays/s")
    # equal RMSE: the same estimator at the same spp, against a converged image
    gpu_img, _ = scene.render(args.spp, seed=2, crop=cr, with_sum2=False)
    ref, _ = scene.render(args.spp * 16, seed=3, crop=cr, with_sum2=False)
    x, y = np.meshgrid(np.arange(cr[0], cr[2]), np.arange(cr[1], cr[3]), indexing="ij")
    idx = ((x * H + y) * 3).reshape(-1, 1) + np.arange(3)
    rm = lambda a: float(np.sqrt(np.mean((a[idx].astype(np.float64) - ref[idx]) ** 2)))
    rmse_cpu, rmse_gpu = rm(cpu_img), rm(gpu_img)
    log(f"[cpu] RMSE against {args.spp * 16} spp: oracle {rmse_cpu:.5f}, GPU {rmse_gpu:.5f}")
    what = "the whole frame" if (cw, ch) == (W, H) else f"a centred {cw}x{ch} crop of the frame"
    return {"value": round(rate, 4), "unit": "Mrays/s", "cores": threads, "host_cpus": os.cpu_count(),
            "kind": "port",
            "sample": f"{what} at {args.spp} spp, seed 1: {rays} BVH queries in {secs:.1f} s",
            "rmse": {"cpu": round(rmse_cpu, 6), "gpu": round(rmse_gpu, 6), "ratio": round(rmse_gpu / rmse_cpu, 4),
                     "equal": bool(abs(rmse_gpu / rmse_cpu - 1) < 0.05),
                     "against": f"the GPU render of the same pixels at {args.spp * 16} spp (seed 3); "
                                f"CPU seed 1, GPU seed 2, same spp"}}


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "traffic_icosphere_l8_1080p_64spp.json")


def pmc_stage(stage, args):
    """Per-frame counter figures of `stage` from the committed rocprofv3 runs
    of this workload (one frame each): bytes fetched and written at the
    L2 / fabric interface (FETCH_SIZE, WRITE_SIZE scaled by a plain copy
    kernel's counters; the fabric serves them from the Infinity Cache or
    HBM), the L2 hit rate and the occupancy, summed or weighted over every
    launch of the stage in the frame.  None unless this run is the profiled
    workload."""
    if (args.width, args.height, args.spp, args.level, args.max_depth, args.camera_z) != (1920, 1080, 64, 8, 12, 7.0):
        return None
    try:
        with open(TRAFFIC_FILE) as f:
            st = json.load(f)["stages"][stage]
        return st if "fabric_bytes_per_frame" in st else None
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
    outs_seq = [ctx.alloc(d.camera["width"] * d.camera["height"] * 12) for d, _, _ in reqs]
    paths = sum(d.camera["width"] * d.camera["height"] * spp for d, spp, _ in reqs)
    # warm both paths once at low spp
    for s_, o in zip(sc, outs):
        s_.render_device(1, o, seed=7)
    ctx.render_batch_device([(s_, 1, 7, o, None) for s_, o in zip(sc, outs)])
    ctx.synchronize()
    t0 = time.perf_counter()
    rays_seq = 0
    for s_, (d, spp, seed), o in zip(sc, reqs, outs_seq):
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
    same = True
    for (d, _, _), a, b in zip(reqs, outs, outs_seq):
        n = d.camera["width"] * d.camera["height"] * 3
        ia = ctx.d2h(np.empty(n, np.float32), a)
        ib = ctx.d2h(np.empty(n, np.float32), b)
        same = same and bool(np.array_equal(ia.view(np.uint32), ib.view(np.uint32)))
    for o in outs + outs_seq:
        ctx.free(o)
    log(f"[bench] batched: {len(reqs)} requests, {paths / 1e6:.0f} M paths; sequential {seq * 1000:.0f} ms, "
        f"batched {bat * 1000:.0f} ms")
    return {"speedup": round(seq / bat, 3), "sequential_ms": round(seq * 1000, 1), "batched_ms": round(bat * 1000, 1),
            "requests": len(reqs), "paths": paths, "rays": rays_bat, "rays_equal": rays_seq == rays_bat,
            "images_bit_identical": same,
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
    tdesc, ttex = scenes.diff_cornell(R, R, tex_res=T, seed=args.adj_seed)
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
    opt = torch.optim.Adam(list(params.values()), lr=0.03)
    tgt = target.view(R, R, 3)
    # progress measures the render noise does not hide: the mean absolute
    # texel error of each texture against the target's, and the loss of a
    # fixed-seed render
    want = {tex[k]: torch.tensor(tdesc.textures[ttex[k]].image.reshape(-1), device="cuda") for k in tex}

    def texel_err():
        return {k: round(float((params[tex[k]].detach() - want[tex[k]]).abs().mean()), 5) for k in sorted(tex)}

    def eval_loss():
        with torch.no_grad():
            return round(float(((render(scene, params, spp, seed=424242) - tgt) ** 2).sum()), 3)

    err0, eval0 = texel_err(), eval_loss()
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
    err1, eval1 = texel_err(), eval_loss()
    log(f"[bench] adjoint opt: texel error {err0} -> {err1}; fixed-seed loss {eval0} -> {eval1}")
    scene.close()
    return {"ratio": round(a_ms / p_ms, 3), "primal_ms": round(p_ms, 1), "adjoint_ms": round(a_ms, 1),
            "primal_device_ms": round(float(np.median(prim_dev)), 1),
            "adjoint_device_ms": round(float(np.median(adj_dev)), 1),
            "primal_mrays_s": round(rays_p / sum(prim) / 1e6, 1), "adjoint_mrays_s": round(rays_a / sum(adj) / 1e6, 1),
            "opt_steps": args.adj_steps, "opt_step_ms": round(float(np.median(step_ms)), 1) if step_ms else None,
            "loss_first": round(losses[0], 3) if losses else None, "loss_last": round(losses[-1], 3) if losses else None,
            "texel_error_start": err0, "texel_error_end": err1,
            "fixed_seed_loss_start": eval0, "fixed_seed_loss_end": eval1,
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
    ctx.reset_kernel_timing()
    t0 = time.perf_counter()
    scene.render_device(spp, d_rgb, seed=1)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    st = ctx.render_stats()
    ctx.free(d_rgb)
    stages = {}
    for name in ("compact", "sort", "trace_closest", "shade", "compact_shadow", "trace_secondary", "trace_any",
                 "trace_mis", "resolve", "accumulate"):
        avg, cnt = ctx.kernel_timing(name)
        if cnt:
            stages[name] = round(avg * cnt, 1)
            log(f"[bench] instanced stage {name:15s} avg {avg * 1000:9.1f} us x {cnt:4d} = {avg * cnt:8.1f} ms")
    rays = st["rays_closest"] + st["rays_any"]
    rate = rays / secs / 1e6
    log(f"[bench] instanced: {secs * 1000:.0f} ms, {rays / 1e6:.0f} M rays, {rate:.1f} Mrays/s")
    cpu = None
    if not args.no_cpu_baseline:
        if not os.path.exists(cpuref.LIB_PATH):
            cpuref.build()
        threads = host_cores()
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
            "stack_depth": info["max_depth"], "stage_ms": stages, "cpu_baseline": cpu,
            "speedup_vs_cpu": round(rate / cpu["value"], 1) if cpu else None,
            "workload": f"1024 instances of 8 superquadrics (99,904 tris each), 2048x1024 HDR sky + sun; "
                        f"path Mis, {W}x{H}, {spp} spp, max depth 16"}


def light_paths(ctx, args):
    """The light-path integrators on the Cornell box of the plain-path
    workload (config-0 geometry): light tracing (adjoint_particle: 2^24
    particles, 512x512, the reference's default depth rule max_depth -1,
    Russian roulette 0.99 after 3 bounces) and bidirectional path tracing
    (bdpt: 256x256, 16 spp, the reference's default max_depth 7).  One timed
    render each after a warmup, and the oracle on a bounded sample of the
    same work (2^23 particles; the whole frame at 16 spp)."""
    from radiantium_amd import native, scenes
    from oracle import cpuref
    out = {}
    threads = host_cores()
    if not args.no_cpu_baseline and not os.path.exists(cpuref.LIB_PATH):
        cpuref.build()
    # light tracing
    desc = scenes.cornell_box(512, 512)
    desc.integrator.update(type="adjoint_particle", max_depth=-1, min_depth=3)
    scene = native.Scene(ctx, desc)
    n = args.particles
    d_rgb = ctx.alloc(512 * 512 * 12)
    scene.render_particles_device(n // 16, d_rgb, seed=99)
    ctx.synchronize()
    ctx.reset_kernel_timing()
    t0 = time.perf_counter()
    scene.render_particles_device(n, d_rgb, seed=1)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    st = ctx.render_stats()
    rays = st["rays_closest"] + st["rays_any"]
    stages = {}
    for name in ("lt_emit", "trace_closest", "shade", "trace_any", "splat", "finalize"):
        avg, cnt = ctx.kernel_timing(name)
        if cnt:
            stages[name] = round(avg * cnt, 2)
    lt = {"mrays_s": round(rays / secs / 1e6, 1), "ms": round(secs * 1000, 1), "rays": rays, "particles": n,
          "mparticles_s": round(n / secs / 1e6, 1), "stage_ms": stages,
          "workload": "Cornell box 512x512, adjoint_particle, max_depth -1, min_depth 3, rr 0.99"}
    log(f"[bench] light tracing: {n} particles in {secs * 1000:.1f} ms, {lt['mrays_s']} Mrays/s, stages {stages}")
    if not args.no_cpu_baseline:
        cn = 1 << 23
        _, cst, csecs = cpuref.OracleScene(desc).render_particles(cn, seed=1, threads=threads)
        crays = cst["rays_closest"] + cst["rays_any"]
        lt["cpu_baseline"] = {"value": round(crays / csecs / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                              "kind": "port", "sample": f"{cn} particles: {crays} BVH queries in {csecs:.2f} s"}
        lt["speedup_vs_cpu"] = round(lt["mrays_s"] / lt["cpu_baseline"]["value"], 1)
        log(f"[cpu] light tracing: {lt['cpu_baseline']['value']} Mrays/s")
    ctx.free(d_rgb)
    scene.close()
    out["light_tracing"] = lt
    # bidirectional
    desc = scenes.cornell_box(256, 256)
    desc.integrator.update(type="bdpt", max_depth=7, min_depth=5)
    scene = native.Scene(ctx, desc)
    spp = 16
    d_rgb = ctx.alloc(256 * 256 * 12)
    scene.render_device(2, d_rgb, seed=99)
    ctx.synchronize()
    ctx.reset_kernel_timing()
    t0 = time.perf_counter()
    scene.render_device(spp, d_rgb, seed=1)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    st = ctx.render_stats()
    rays = st["rays_closest"] + st["rays_any"]
    stages = {}
    for name in ("bd_start", "trace_closest", "bd_walk", "bd_connect", "bd_chain", "bd_sum", "accumulate"):
        avg, cnt = ctx.kernel_timing(name)
        if cnt:
            stages[name] = round(avg * cnt, 2)
    bd = {"mrays_s": round(rays / secs / 1e6, 1), "ms": round(secs * 1000, 1), "rays": rays,
          "msamples_s": round(256 * 256 * spp / secs / 1e6, 2), "stage_ms": stages,
          "workload": f"Cornell box 256x256, bdpt, {spp} spp, max_depth 7"}
    log(f"[bench] bdpt: {secs * 1000:.1f} ms, {bd['mrays_s']} Mrays/s, stages {stages}")
    if not args.no_cpu_baseline:
        _, _, cst, csecs = cpuref.OracleScene(desc).render_bdpt(spp, seed=1, threads=threads, with_sum2=False)
        crays = cst["rays_closest"] + cst["rays_any"]
        bd["cpu_baseline"] = {"value": round(crays / csecs / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                              "kind": "port",
                              "sample": f"the whole frame at {spp} spp: {crays} BVH queries in {csecs:.2f} s"}
        bd["speedup_vs_cpu"] = round(bd["mrays_s"] / bd["cpu_baseline"]["value"], 1)
        log(f"[cpu] bdpt: {bd['cpu_baseline']['value']} Mrays/s")
    ctx.free(d_rgb)
    scene.close()
    out["bdpt"] = bd
    return out


def octree_leg(ctx, args):
    """The same frame through the loose octree accelerator (the config's
    "octree" accel, Accelerators/Octree.cs, its defaults max_depth 10,
    max_count 10, out_bound 2) at fewer spp: octree build time, Mrays/s,
    box and primitive tests per ray, the closest-hit traversal's
    algorithmic bytes per launch over its launch time, and the oracle's
    octree on a bounded crop.  The octree drops the primitives whose bound
    lies flat on a cell face (here the ground and light quads, on the root's
    faces), as the reference's does, so its frame is not the BVH frame."""
    from radiantium_amd import native
    from oracle import cpuref
    spp = args.oct_spp
    desc = make_scene(args)
    desc.accel = dict(type="octree", max_depth=10, max_count=10, out_bound=2.0)
    t0 = time.perf_counter()
    scene = native.Scene(ctx, desc)
    create_s = time.perf_counter() - t0
    nodes, first, count, prims = scene.export_octree()
    npr = desc.num_prims()
    kept = int(np.unique(prims).size)
    d_rgb = ctx.alloc(args.width * args.height * 12)
    scene.render_device(1, d_rgb, seed=99)
    ctx.synchronize()
    ctx.reset_kernel_timing()
    t0 = time.perf_counter()
    scene.render_device(spp, d_rgb, seed=1)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    st = ctx.render_stats()
    rays = st["rays_closest"] + st["rays_any"]
    nodes_c, prims_c, rays_c = st["class_nodes"], st["class_prims"], st["class_rays"]
    stages = {}
    for name in ("trace_closest", "shade", "trace_any", "trace_mis", "resolve"):
        avg, cnt = ctx.kernel_timing(name)
        if cnt:
            stages[name] = (avg, cnt)
    oc = {"mrays_s": round(rays / secs / 1e6, 1), "ms": round(secs * 1000, 1), "rays": rays, "spp": spp,
          "scene_create_s": round(create_s, 2), "nodes": int(nodes.shape[0]), "leaves": int(first.size),
          "leaf_entries": int(prims.size), "primitives_kept": kept, "primitives": npr,
          "box_tests_per_ray": round(sum(nodes_c) / max(sum(rays_c), 1), 2),
          "prim_tests_per_ray": round(sum(prims_c) / max(sum(rays_c), 1), 2),
          "stage_ms": {k: round(a * c, 2) for k, (a, c) in stages.items()},
          "workload": f"the icosphere frame ({args.width}x{args.height}, max depth {args.max_depth}) at {spp} spp "
                      f"through the octree accel (max_depth 10, max_count 10, out_bound 2)"}
    if "trace_closest" in stages:
        avg, cnt = stages["trace_closest"]
        byts = rays_c[0] * (RAY_IN + 16) + 32 * nodes_c[0] + 36 * prims_c[0]
        ach = byts / cnt / (avg * 1e-3) / 1e9
        oc["roofline"] = {"kernel": "trace_closest (k_trace_oct)", "bound": "hbm", "achieved": round(ach, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                          "bytes_per_launch": int(byts / cnt), "avg_launch_ms": round(avg, 4),
                          "achieved_is": "32 B ray + 16 B hit + 32 B per box test + 36 B per triangle test, "
                                         "counted by the kernel, per launch / launch time"}
    log(f"[bench] octree: {secs * 1000:.1f} ms at {spp} spp, {oc['mrays_s']} Mrays/s, "
        f"{oc['box_tests_per_ray']} boxes / {oc['prim_tests_per_ray']} triangles per ray, "
        f"{kept}/{npr} primitives kept, create {create_s:.2f}s")
    if not args.no_cpu_baseline:
        if not os.path.exists(cpuref.LIB_PATH):
            cpuref.build()
        threads = host_cores()
        o = cpuref.OracleScene(desc)
        cspp = 4 * spp
        _, _, cst, csecs = o.render(cspp, seed=1, threads=threads, with_sum2=False)
        crays = cst["rays_closest"] + cst["rays_any"]
        oc["cpu_baseline"] = {"value": round(crays / csecs / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                              "kind": "port", "sample": f"the whole frame at {cspp} spp: {crays} octree "
                                                        f"queries in {csecs:.2f} s"}
        oc["speedup_vs_cpu"] = round(oc["mrays_s"] / oc["cpu_baseline"]["value"], 1)
        log(f"[cpu] octree: {oc['cpu_baseline']['value']} Mrays/s")
    ctx.free(d_rgb)
    scene.close()
    return oc


def bssrdf_leg(ctx, args):
    """Subsurface scattering through the path integrator's Mis method (the
    BSSRDF branch of PathTracer.cs:244-259): the Cornell box with a Disney
    tall box with a scattering distance and a sphere of the "subsurface"
    material, 512x512, 64 spp, max depth 8.  One timed frame after a warmup,
    the BSSRDF stage's share of it, and the oracle on the whole frame at 16 spp."""
    from radiantium_amd import native, scenes
    from oracle import cpuref
    W = H = 512
    spp = 64
    desc = scenes.sss_cornell(W, H)
    scene = native.Scene(ctx, desc)
    d_rgb = ctx.alloc(W * H * 12)
    scene.render_device(4, d_rgb, seed=99)
    ctx.synchronize()
    ctx.reset_kernel_timing()
    t0 = time.perf_counter()
    scene.render_device(spp, d_rgb, seed=1)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    st = ctx.render_stats()
    rays = st["rays_closest"] + st["rays_any"]
    stages = {}
    for name in ("trace_closest", "shade", "trace_secondary", "resolve", "sss", "tail"):
        avg, cnt = ctx.kernel_timing(name)
        if cnt:
            stages[name] = round(avg * cnt, 2)
    out = {"mrays_s": round(rays / secs / 1e6, 1), "ms": round(secs * 1000, 1), "rays": rays,
           "msamples_s": round(W * H * spp / secs / 1e6, 1), "stage_ms": stages,
           "workload": f"Cornell box with a Disney subsurface box and a subsurface sphere, {W}x{H}, {spp} spp, "
                       f"path Mis, max depth 8"}
    log(f"[bench] bssrdf: {secs * 1000:.1f} ms, {out['mrays_s']} Mrays/s, stages {stages}")
    if not args.no_cpu_baseline:
        if not os.path.exists(cpuref.LIB_PATH):
            cpuref.build()
        threads = host_cores()
        _, _, cst, csecs = cpuref.OracleScene(desc).render(16, seed=1, threads=threads, with_sum2=False)
        crays = cst["rays_closest"] + cst["rays_any"]
        out["cpu_baseline"] = {"value": round(crays / csecs / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                               "kind": "port", "sample": f"the whole frame at 16 spp: {crays} BVH queries in "
                                                         f"{csecs:.2f} s"}
        out["speedup_vs_cpu"] = round(out["mrays_s"] / out["cpu_baseline"]["value"], 1)
        log(f"[cpu] bssrdf: {out['cpu_baseline']['value']} Mrays/s")
    ctx.free(d_rgb)
    scene.close()
    return out


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
    ap.add_argument("--camera-z", type=float, default=7.0, help="camera distance of the main workload")
    ap.add_argument("--no-close", action="store_true", help="skip the close-framing (camera z 4.6) frame")
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
    ap.add_argument("--adj-steps", type=int, default=16, help="optimisation steps through the autograd Function")
    ap.add_argument("--no-light", action="store_true", help="skip the light tracing / bdpt measurement")
    ap.add_argument("--only-light", action="store_true", help="only the light tracing / bdpt measurement")
    ap.add_argument("--particles", type=int, default=1 << 24, help="particles of the light tracing leg")
    ap.add_argument("--no-octree", action="store_true", help="skip the octree-accelerator measurement")
    ap.add_argument("--only-octree", action="store_true", help="only the octree-accelerator measurement")
    ap.add_argument("--oct-spp", type=int, default=4, help="spp of the octree frame")
    ap.add_argument("--no-bssrdf", action="store_true", help="skip the subsurface-scattering measurement")
    ap.add_argument("--only-bssrdf", action="store_true", help="only the subsurface-scattering measurement")
    ap.add_argument("--only-instanced", action="store_true",
                    help="run the instanced-workload frame alone (for profiling); prints its record")
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
    if args.only_instanced:
        args.no_cpu_baseline = True
        print(json.dumps({"instanced": instanced_throughput(native.Context(0), args)}), flush=True)
        return
    if args.only_light:
        print(json.dumps({"light_paths": light_paths(native.Context(0), args)}), flush=True)
        return
    if args.only_octree:
        print(json.dumps({"octree": octree_leg(native.Context(0), args)}), flush=True)
        return
    if args.only_bssrdf:
        print(json.dumps({"bssrdf": bssrdf_leg(native.Context(0), args)}), flush=True)
        return
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
    coverage = primary_coverage(scene, desc, args.level)
    log(f"[bench] primary rays (camera z {args.camera_z}): {coverage}")
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
        launches_per_frame = cnt / args.steps
        pmc = pmc_stage(trace_dom, args)
        # the counters' bytes of the launches this run's stage timer counts (the
        # frame's launches before its device-driven tail) over their number: the
        # same launch partition as the algorithmic bytes per launch
        traffic = None
        if pmc and pmc.get("timed_launches_per_frame"):
            traffic = int(pmc["fabric_bytes_timed_launches"] / pmc["timed_launches_per_frame"])
        roofline = {"kernel": trace_dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "bytes_per_launch": int(per_launch), "bytes_per_frame": int(total_bytes / args.steps),
                    "launches_per_frame": round(launches_per_frame, 2), "avg_launch_ms": round(avg_ms, 4),
                    "nodes_per_ray": round(sum(agg["class_nodes"][c] for c, _ in cls) / max(rays_c, 1), 2),
                    "prims_per_ray": round(sum(agg["class_prims"][c] for c, _ in cls) / max(rays_c, 1), 2),
                    "achieved_is": "algorithmic bytes (32 B ray + hit + 32 B per node + 36 B per triangle, "
                                   "counted by the kernel) per launch / launch time"}
        if pmc:
            fabric_gbs = traffic / (avg_ms * 1e-3) / 1e9
            roofline.update({
                "fabric_achieved": round(fabric_gbs, 1), "fabric_frac": round(fabric_gbs / HBM_PEAK_GBS, 4),
                "fabric_share_of_algorithmic": round(traffic / per_launch, 4),
                "l2_hit_rate": pmc.get("l2_hit_rate"), "occupancy_pct": pmc.get("occupancy_pct"),
                "limiter": "L1 / texture-address path: the bytes that reach the L2-fabric interface "
                           "(Infinity Cache + HBM) are a fifth of the algorithmic bytes, and the kernel runs "
                           "well below HBM bandwidth on them"})
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

    close = None
    if not args.no_close:
        close = close_framing(ctx, args)

    lightp = None
    if not args.no_light:
        lightp = light_paths(ctx, args)

    octree = None
    if not args.no_octree:
        octree = octree_leg(ctx, args)

    bssrdf = None
    if not args.no_bssrdf:
        bssrdf = bssrdf_leg(ctx, args)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(desc, args, value, scene)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1000, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"displaced icosphere level {args.level} ({ntri} triangles) + rough-glass ground "
                               f"+ quad light; path Mis; {args.width}x{args.height}, {args.spp} spp, max depth "
                               f"{args.max_depth}, bvh max_prim 4; camera (0, 2.2, {args.camera_z:g}) -> (0, 0.9, 0), "
                               f"fov 40: primary rays {coverage['mesh']:.0%} mesh, {coverage['ground']:.0%} ground, "
                               f"{coverage['miss']:.0%} miss",
                   "camera_z": args.camera_z, "primary_rays": coverage,
                   "width": args.width, "height": args.height, "spp": args.spp, "max_depth": args.max_depth,
                   "triangles": ntri, "parallelism": "1 GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "adjoint": adjoint,
        "instanced": instanced, "close_framing": close, "light_paths": lightp, "octree": octree,
        "bssrdf": bssrdf,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
