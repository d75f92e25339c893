#!/usr/bin/env python
"""Benchmark of the MI355X ray-tracing hot path; prints ONE JSON line.

Default workload: the full wavefront render (1024x1024 film, 16 spp stratified
streams, MIS direct lighting, a 2 009 780-triangle displaced icosphere, 64
spheres, a ground quad, 4 sphere lights and a 1024x512 env map, all 1024 film
tiles per pass).  One step = one whole render (16 passes).  value = Msamples/s.
The LBVH build (1 048 576 triangles) and closest-hit traversal (4 194 304 rays)
of the build+traversal microbenchmark are measured in the same run and reported
as extra fields.  Inputs are resident in HBM before the timed region.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Msamples/s path-traced + LBVH Mprims/s build; fwd+adjoint vs fwd time"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes: traversal 28 B ray in + 8 B hit out
TRACE_BYTES_PER_RAY = 36.0
LBVH_BYTES_PER_PRIM = 196.0
LBVH_BYTES_PER_NODE = 36.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = ROOT / "profiles" / "render_c2_pmc.txt"
CPU_TILE_STRIDE = 4


def measured_traffic():
    """HBM bytes per k_trace launch of the render, from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


CPU_SINGLE_TILE_STRIDE = 64


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_share():
    """Threads for the CPU baseline: the host-CPU share of one GPU on the box
    (OMP_NUM_THREADS, which the box sets to it), else every visible core."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(share, visible) if share > 0 else visible), visible


def cpu_baseline(scene):
    """The oracle (CPU restatement of the reference path) on every 4th film tile
    with the GPU's CPU share of threads, and on every 64th tile single-threaded."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    ntx = scene.width // 32
    ntiles = ntx * (scene.height // 32)
    threads, visible = cpu_share()

    def run(stride, nthreads):
        tiles = np.arange(0, ntiles, stride, dtype=np.uint32)
        out = po.render(scene, nthreads=nthreads, tiles=tiles)
        secs = float(out["counters"][7]) * 1e-9
        samples = len(tiles) * 32 * 32 * scene.spp
        return samples / secs / 1e6, len(tiles), samples, secs

    v, nt, ns, secs = run(CPU_TILE_STRIDE, threads)
    v1, nt1, ns1, secs1 = run(CPU_SINGLE_TILE_STRIDE, 1)
    return {"value": v, "unit": "Msamples/s", "cores": threads,
            "kind": "port",
            "sample": "every %dth 32x32 tile of the render (%d of %d tiles, %d samples), "
                      "%.2f s wall on %d threads" % (CPU_TILE_STRIDE, nt, ntiles, ns, secs, threads),
            "single_thread": {"value": v1, "unit": "Msamples/s",
                              "sample": "every %dth tile (%d tiles, %d samples), %.2f s wall on 1 thread"
                                        % (CPU_SINGLE_TILE_STRIDE, nt1, ns1, secs1)},
            "cpu_model": cpu_model(), "cpus_visible": visible,
            "threads_rule": "the host-CPU share of one GPU (OMP_NUM_THREADS) when set, else every visible core"}


# wpbr_lbvh_kernel_times intervals (events 0..11 of a build)
KERNEL_NAMES = ["bounds", "centroid_bounds+morton", "sort_pass0", "sort_pass1", "sort_pass2",
                "sort_pass3 | treelets", "slot_gather+treelets", "emit", "upper_sah", "flatten",
                "traversal_layout"]


def bench_microbench(lib, scenes, ctx, steps):
    """LBVH build (1 048 576 triangles) in both key orders, and closest-hit
    traversal of 4 194 304 rays over each tree."""
    tris, rays = scenes.microbench()
    sc = lib.Scene(ctx, np.zeros((0, 24), np.float32), np.zeros(0, np.float32), tris)
    d_rays = ctx.array(rays)
    n = rays.shape[1]
    saved = ctx.trace_config()
    res = {}
    for order, key in ((lib.ORDER_REFERENCE, ""), (lib.ORDER_FULL, "_full")):
        bvh = lib.Lbvh(ctx, order)
        bvh.build(sc)
        build_ms = [bvh.build(sc) for _ in range(max(steps, 3))]
        ktimes = bvh.kernel_times()
        cnt = bvh.counts()
        nodes = cnt["nodes"]
        tuned = lib.autotune_collide(ctx, sc, bvh, d_rays, n)
        trace_ms = []
        for _ in range(max(steps, 3)):
            lib.collide(ctx, sc, bvh, d_rays, n)
            trace_ms.append(ctx.last_ms())
        b = float(np.median(build_ms))
        t = float(np.median(trace_ms))
        layout = ktimes[10] if len(ktimes) > 10 else 0.0
        lbvh_bytes = LBVH_BYTES_PER_PRIM * len(tris) + LBVH_BYTES_PER_NODE * nodes
        res["lbvh_build" + key] = {
            "order": "reference (bits 0..23, what the reference's sort job leaves)" if order == lib.ORDER_REFERENCE
                     else "full 30-bit sort",
            "prims": len(tris), "nodes": nodes, "treelets": cnt["treelets"], "ms": b,
            "mprims_per_s": len(tris) / b / 1e3,
            "hbm_frac": lbvh_bytes / (b * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "ms_without_traversal_layout": b - layout,
            "kernel_ms": {k: round(x, 4) for k, x in zip(KERNEL_NAMES, ktimes)}}
        res["traversal" + key] = {"rays": n, "ms": t, "mrays_per_s": n / t / 1e3, "config": tuned}
        bvh.destroy()
    ctx.set_trace_config(**saved)
    return res


def bench_adjoint(lib, scenes, ctx, steps, tune=True):
    """Differentiable render: forward, L2 loss against a target rendered with
    parameters perturbed by up to +-20 %, and the path-replay backward pass."""
    s = scenes.c4_scene()
    r = lib.DiffRenderer(ctx, s)
    alpha, etak, env = r.params()
    g = np.random.default_rng(2026)
    r.set_params(*[(p * (1 + 0.2 * g.uniform(-1, 1, p.shape))).astype(np.float32)
                   for p in (alpha, etak, env)])
    if tune:
        r.autotune()
    r.forward()
    r.set_target(r.film_rgb().astype(np.float32))
    r.set_params(alpha, etak, env)
    fwd, bwd, tot = [], [], []
    for k in range(max(steps, 3) + 1):
        ctx.sync()
        t0 = time.perf_counter()
        f = r.forward()
        loss = r.loss()
        b = r.backward()
        ctx.sync()
        if k:  # the first iteration builds the adjoint tables
            fwd.append(f)
            bwd.append(b)
            tot.append((time.perf_counter() - t0) * 1e3)
    f, b, t = float(np.median(fwd)), float(np.median(bwd)), float(np.median(tot))
    grads = r.gradients()
    r.destroy()
    return {"adjoint": {"workload": "differentiable render: 512x512 film, 8 spp, 2009850 shapes, "
                                    "%d alpha + %d eta/k + %d env parameters" % (r.n_mat, r.n_etak, r.n_env),
                        "forward_ms": f, "backward_ms": b, "fwd_plus_bwd_wall_ms": t,
                        "ratio_fwd_plus_bwd_over_fwd": (f + b) / f, "loss": loss,
                        "grad_abs_max": [float(np.abs(x).max()) for x in grads]}}


def bench_batch(lib, scenes, ctx, steps, tune=True):
    """Batched many-scene render: 64 scenes (1k-500k triangles, films 128^2 ..
    1024x512, 8 spp): the segmented LBVH build of all scenes, then the batched
    render (one graph replay per step); beside it the same scenes rendered one
    at a time with the single-scene renderer."""
    ss = scenes.c3_scenes()
    samples = sum(s.width * s.height * s.spp for s in ss)
    shapes = sum(s.n_shapes for s in ss)
    br = lib.BatchRenderer(ctx, ss)
    # segmented build of all scenes into the batch's trees, timed alone
    geo = [lib.Scene(ctx, s.sphere_xform, s.sphere_radius, s.tri_verts) for s in ss]
    bv = [lib.Lbvh(ctx) for _ in ss]
    lib.build_segmented(ctx, geo, bv)
    seg_ms = float(np.median([lib.build_segmented(ctx, geo, bv) for _ in range(max(steps, 3))]))
    for b in bv:
        b.destroy()
    del geo
    tuned = br.autotune() if tune else ctx.trace_config()
    br.render()
    ms = []
    for _ in range(max(steps, 3)):
        ctx.sync()
        t0 = time.perf_counter()
        br.render()
        ms.append((time.perf_counter() - t0) * 1e3)
    st = br.stats()
    br.destroy()
    t = float(np.median(ms))
    # one scene at a time (wpbr_render per scene, same traversal configuration)
    single = 0.0
    for s in ss:
        r = lib.Renderer(ctx, s)
        r.render()
        ctx.sync()
        t0 = time.perf_counter()
        r.render()
        single += (time.perf_counter() - t0) * 1e3
        r.destroy()
    return {"batch": {"workload": "batched many-scene render: %d scenes, %d shapes, %d pixels, 8 spp"
                                  % (len(ss), shapes, samples // 8),
                      "segmented_build_ms": seg_ms, "segmented_build_mprims_per_s": shapes / seg_ms / 1e3,
                      "render_ms": t, "msamples_per_s": samples / t / 1e3,
                      "device_ms": st.get("ms"), "secondary_rays": st.get("secondary_rays"),
                      "one_scene_at_a_time_ms": single, "batched_speedup": single / t,
                      "trace_config": tuned}}


# BSDF jobs: per request read frame 36 B, wo 12 B, material 4 B, the queue
# index 4 B and wi 12 B (Compute) or u 8 B (Sample); written: 64 spectral values
# 256 B, pdf 4 B and for Sample wi 12 B
BSDF_BYTES = {"compute": 68 + 260, "sample": 64 + 272}


def bench_bsdf(lib, scenes, ctx, scene, r, steps):
    """The BSDF Compute / Sample jobs over the surface interactions of the
    render's primary rays (one per pixel through the camera's eye): per-kind
    queues and one launch per kind against one fused launch over all requests."""
    W, H = scene.width, scene.height
    rays = np.ascontiguousarray(scenes.pinhole_rays(W, H, (0.0, 1.2, -5.2)), np.float32)
    n = rays.shape[1]
    d_rays, d_surf = ctx.array(rays), ctx.array(np.ones(n, np.uint8))
    d_s, d_t, d_si = r.collide(d_rays, n, d_surf)
    hit, si = d_s.numpy(), d_si.numpy()
    for a in (d_rays, d_surf, d_s, d_t, d_si):
        a.free()
    keep = hit != 0xFFFFFFFF
    hit, si = hit[keep], si[:, keep].astype(np.float64)
    f = lib.SI_FIELDS
    ns = si[f["shading_n"]:f["shading_n"] + 3]
    ss = si[f["shading_dpdu"]:f["shading_dpdu"] + 3]
    ss = ss / np.maximum(np.linalg.norm(ss, axis=0), 1e-30)
    frame = np.concatenate([ss, np.cross(ns.T, ss.T).T, ns]).astype(np.float32)
    wo = si[f["wo"]:f["wo"] + 3].astype(np.float32)
    mat = scene.shape_material[hit].astype(np.uint32)
    ok = mat < len(scene.material_table)
    frame, wo, mat = frame[:, ok], wo[:, ok], mat[ok]
    m = len(mat)
    g = np.random.default_rng(7)
    wi = g.normal(size=(3, m))
    wi = (wi / np.linalg.norm(wi, axis=0, keepdims=True)).astype(np.float32)
    req, arrs = r.bsdf_requests(frame, wo, mat, wi=wi, u=g.random((2, m)).astype(np.float32))
    d_q, counts = r.bsdf_queues(req, m)
    q_us = []
    for _ in range(max(steps, 5)):
        r.bsdf_queues(req, m)
        q_us.append(ctx.last_ms() * 1e3)
    offs = np.concatenate([[0], np.cumsum(counts)])
    subs = []
    for k in range(3):
        sub = ctx.empty(max(counts[k], 1), np.uint32)
        if counts[k]:
            sub.upload(d_q.numpy()[offs[k]:offs[k + 1]])
        subs.append(sub)
    res = {"requests": m, "kinds": {"conductor": counts[0], "dielectric": counts[1], "specular": counts[2]},
           "queues_us": float(np.median(q_us))}
    for sample, tag in ((False, "compute"), (True, "sample")):
        out = r.bsdf(sample, lib.BSDF_ANY, req, m)
        split, fused = [], []
        for _ in range(max(steps, 5)):
            t = 0.0
            for k in range(3):
                r.bsdf(sample, k, req, m, subs[k], counts[k], out=out)
                t += ctx.last_ms()
            split.append(t * 1e3)
            r.bsdf(sample, lib.BSDF_ANY, req, m, out=out)
            fused.append(ctx.last_ms() * 1e3)
        sp, fu = float(np.median(split)), float(np.median(fused))
        gbs = BSDF_BYTES[tag] * m / (sp * 1e-6) / 1e9
        res[tag] = {"queue_split_us": sp, "fused_us": fu, "queue_split_gbs": gbs,
                    "queue_split_hbm_frac": gbs / HBM_PEAK_GBS,
                    "fused_gbs": BSDF_BYTES[tag] * m / (fu * 1e-6) / 1e9, "bytes_per_request": BSDF_BYTES[tag]}
        for a in out:
            if a is not None:
                a.free()
    for a in subs + [d_q] + list(arrs.values()):
        a.free()
    return {"bsdf_jobs": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["render", "c0"], default="render")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-microbench", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-stratified", action="store_true", help="skip the stratified-sampler render")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-bsdf", action="store_true", help="skip the BSDF job timings")
    ap.add_argument("--serial-passes", action="store_true",
                    help="run the timed renders with their sample passes one after another (no two "
                         "kernels share the GPU), for per-kernel counter profiles")
    ap.add_argument("--bsdf-only", action="store_true",
                    help="only the BSDF jobs over the full render's primary interactions, for kernel profiles")
    ap.add_argument("--trace-config", default=None,
                    help="fixed traversal configuration block,stack_depth,mapping,width,refill,chunk "
                         "instead of the autotuner (for profiling one configuration)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run one plain 512 MiB device copy (reference kernel for the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE scale factor)")
    ap.add_argument("--lbvh-only", action="store_true",
                    help="only the LBVH builds (1 048 576 triangles, then the 64-scene segmented "
                         "build), for kernel profiles")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU (--gpus 1); sharding the render over several GPUs "
              "is not implemented", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = 1 << 27
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
        return
    from woodenpbrengine_amd import lib, scenes

    if args.lbvh_only:
        # the LBVH builds alone (for kernel profiles): the 1 048 576-triangle build
        # (k_* kernels) and the 64-scene segmented build (ks_* kernels)
        ctx = lib.Context(0)
        tris, _ = scenes.microbench()
        sc = lib.Scene(ctx, np.zeros((0, 24), np.float32), np.zeros(0, np.float32), tris)
        ss = scenes.c3_scenes()
        geo = [lib.Scene(ctx, s.sphere_xform, s.sphere_radius, s.tri_verts) for s in ss]
        out = {"prims": len(tris), "segmented_shapes": int(sum(s.n_shapes for s in ss))}
        for order, key in ((lib.ORDER_REFERENCE, "reference"), (lib.ORDER_FULL, "full")):
            bvh = lib.Lbvh(ctx, order)
            bvh.build(sc)
            one = [bvh.build(sc) for _ in range(max(args.steps, 3))]
            bv = [lib.Lbvh(ctx, order) for _ in ss]
            lib.build_segmented(ctx, geo, bv)
            seg = [lib.build_segmented(ctx, geo, bv) for _ in range(max(args.steps, 3))]
            out[key] = {"lbvh_build_ms": float(np.median(one)),
                        "kernel_ms": {k: round(x, 4) for k, x in zip(KERNEL_NAMES, bvh.kernel_times())},
                        "segmented_build_ms": float(np.median(seg))}
            log(key, json.dumps(out[key]))
            for b in bv + [bvh]:
                b.destroy()
        print(json.dumps(out), flush=True)
        ctx.close()
        return

    if args.bsdf_only:
        scene = scenes.c2_scene()
        ctx = lib.Context(0)
        r = lib.Renderer(ctx, scene)
        res = bench_bsdf(lib, scenes, ctx, scene, r, args.steps)
        r.destroy()
        print(json.dumps(res), flush=True)
        ctx.close()
        return

    t0 = time.time()
    if args.workload == "render":
        scene = scenes.c2_scene()
        wl = ("full wavefront render: 1024x1024 film, 16 spp stratified, displaced icosphere (2009780 tris), "
              "reference-order LBVH")
    else:
        scene = scenes.c0_scene()
        wl = "plain render: 128x128 film, 4 spp stratified, icosphere (20480 tris), reference-order LBVH"
    # the reference's stratified sampler (JobSamplerStratifiedGenerateSampels1D/2D)
    scene.sampler = lib.SAMPLER_STRATIFIED
    log("scene generated in %.1f s (%d shapes)" % (time.time() - t0, scene.n_shapes))
    ctx = lib.Context(0)
    r = lib.Renderer(ctx, scene)
    log("scene uploaded + LBVH built")
    if args.trace_config:
        vals = [int(x) for x in args.trace_config.split(",")]
        ctx.set_trace_config(**dict(zip(["block", "stack_depth", "mapping", "width", "refill", "chunk"], vals)))
        tuned = ctx.trace_config()
    else:
        tuned = r.autotune()
    log("traversal configuration:", tuned)
    if args.serial_passes:
        ctx.set_render_overlap(0)
    for _ in range(args.warmup):
        r.render()
    ctx.sync()
    dev_ms, trace_ms, shade_ms, rays, launches, prim_ms = [], [], [], [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        dev_ms.append(r.render())
        st = r.stats()
        trace_ms.append(st["trace_ms"])
        shade_ms.append(st["shade_ms"])
        rays.append(st["rays"])
        prim_ms.append(st["primary_trace_ms"])
        launches.append(st["trace_launches"])
    ctx.sync()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / args.steps
    samples = scene.width * scene.height * scene.spp
    value = samples / (ms_per_step * 1e-3) / 1e6
    # rays traced per render: primary (W*H*spp) + compacted secondary (queue lengths)
    rays_per_step = float(np.mean(rays))
    t_ms = float(np.mean(trace_ms))
    n_launch = float(np.mean(launches))
    log("render: %.2f ms/step wall, %.2f ms device, trace %.2f ms (primary %.2f; %d launches, %.0f rays), "
        "shade %.2f ms" % (ms_per_step, np.mean(dev_ms), t_ms, np.mean(prim_ms), n_launch, rays_per_step,
                           np.mean(shade_ms)))
    img = r.film_rgb()
    # film output (tonemap to 8-bit RGB): 16 B read + 3 B written per pixel
    film_ms = []
    for _ in range(max(3, args.steps)):
        lib.film_rgb8(ctx, r.d_film, scene.width, scene.height)
        film_ms.append(ctx.last_ms())
    npix = scene.width * scene.height
    f_ms = float(np.median(film_ms))
    f_gbs = 19.0 * npix / (f_ms * 1e-3) / 1e9
    log("film output: %.1f us per 8-bit image (%.0f GB/s)" % (f_ms * 1e3, f_gbs))
    # the per-pixel sample arrays of this film (16 + spp 1D and 16 + 4 spp 2D values)
    n1, n2 = 16 + scene.spp, 16 + 4 * scene.spp
    d1, d2 = ctx.empty((npix, n1), np.float32), ctx.empty((npix, n2, 2), np.float32)
    sa_ms = []
    for _ in range(max(3, args.steps)):
        ctx.check(ctx.lib.wpbr_sample_arrays(ctx.h, scene.seed, lib.SAMPLER_STRATIFIED_N, 0, npix, n1, n2,
                                             d1.ptr, d2.ptr))
        sa_ms.append(ctx.last_ms())
    d1.free()
    d2.free()
    sa_ms = float(np.median(sa_ms))
    sa_gbs = npix * (4.0 * n1 + 8.0 * n2) / (sa_ms * 1e-3) / 1e9
    log("stratified sample arrays: %.1f us (%.0f GB/s written)" % (sa_ms * 1e3, sa_gbs))
    # scene ingestion: the render's mesh as an OBJ reader hands it over (one vertex
    # per corner, with normals) baked to world space and gathered into the triangle
    # array; then the env map's 8-bit texels decoded to linear
    nt = len(scene.tri_verts)
    d_pos = ctx.array(scene.tri_verts.reshape(-1, 3))
    with_nrm = len(scene.tri_normal) == nt
    d_nrm = ctx.array(scene.tri_normal.reshape(-1, 3)) if with_nrm else None
    d_idx = ctx.array(np.arange(3 * nt, dtype=np.uint32).reshape(-1, 3))
    xf = np.concatenate([np.eye(3, 4), np.eye(3, 4)]).astype(np.float32).reshape(-1)
    xf[3], xf[15] = 0.25, -0.25  # a translation and its inverse
    bake_ms = []
    for _ in range(max(3, args.steps)):
        outs = lib.mesh_bake(ctx, xf, d_pos, 3 * nt, d_idx, nt, d_normals=d_nrm)
        bake_ms.append(ctx.last_ms())
        for a in outs:
            if a is not None:
                a.free()
    for a in (d_pos, d_nrm, d_idx):
        if a is not None:
            a.free()
    bake_ms = float(np.median(bake_ms))
    # positions (and normals) read + written (36 B each way), indices 12 B, triangle array 36 B
    bake_bytes = 192 if with_nrm else 120
    bake_gbs = bake_bytes * nt / (bake_ms * 1e-3) / 1e9
    log("mesh bake: %.1f us for %d triangles (%.0f GB/s)" % (bake_ms * 1e3, nt, bake_gbs))
    env8 = np.random.default_rng(1).integers(0, 256, 3 * 1024 * 512, dtype=np.uint8)
    d_env8 = ctx.array(env8)
    dec_ms = []
    for _ in range(max(3, args.steps)):
        d_lin = lib.srgb8_to_linear(ctx, d_env8, env8.size)
        dec_ms.append(ctx.last_ms())
        d_lin.free()
    d_env8.free()
    dec_ms = float(np.median(dec_ms))
    dec_gbs = 5.0 * env8.size / (dec_ms * 1e-3) / 1e9
    log("texel decode: %.1f us for a 1024x512 RGB image (%.0f GB/s)" % (dec_ms * 1e3, dec_gbs))
    extra = {"film_output": {"us": f_ms * 1e3, "gbs": f_gbs, "hbm_frac": f_gbs / HBM_PEAK_GBS,
                             "pixels": npix},
             "sample_arrays": {"us": sa_ms * 1e3, "gbs": sa_gbs, "hbm_frac": sa_gbs / HBM_PEAK_GBS,
                               "n1": n1, "n2": n2, "pixels": npix},
             "mesh_bake": {"us": bake_ms * 1e3, "gbs": bake_gbs, "hbm_frac": bake_gbs / HBM_PEAK_GBS,
                           "triangles": nt, "bytes_per_triangle": bake_bytes},
             "texel_decode": {"us": dec_ms * 1e3, "gbs": dec_gbs, "hbm_frac": dec_gbs / HBM_PEAK_GBS,
                              "channels": int(env8.size)},
             "image_mean_rgb": [float(x) for x in np.nanmean(img, axis=(0, 1))],
             "device_ms_per_step": float(np.mean(dev_ms)),
             "trace_ms_per_step": t_ms,
             "shade_ms_per_step": float(np.mean(shade_ms)),
             "primary_trace_ms_per_step": float(np.mean(prim_ms)),
             "rays_per_step": rays_per_step,
             "trace_config": tuned,
             "render_mrays_per_s": rays_per_step / (t_ms * 1e-3) / 1e6}
    # Dominant kernel: k_trace.  The same render with its passes one after
    # another (wpbr_set_render_overlap(0): no kernel shares the GPU), its
    # launches (a primary and a secondary-queue launch per pass) timed with HIP
    # events on the stream each is launched on.
    ctx.set_render_overlap(0)
    r.render()
    ser_ms, ser_trace, ser_launch, ser_rays = [], [], [], []
    for _ in range(args.steps):
        ser_ms.append(r.render())
        st = r.stats()
        ser_trace.append(st["trace_ms"])
        ser_launch.append(st["trace_launches"])
        ser_rays.append(st["rays"])
    ctx.set_render_overlap(1)
    s_trace = float(np.mean(ser_trace))
    s_launch = float(np.mean(ser_launch))
    s_rays = float(np.mean(ser_rays))
    extra["render_serial_passes"] = {"device_ms_per_step": float(np.mean(ser_ms)), "trace_ms_per_step": s_trace,
                                     "trace_launches": s_launch, "rays_per_step": s_rays}
    log("render with serial passes: %.2f ms device, trace %.2f ms in %d launches" % (np.mean(ser_ms), s_trace,
                                                                                     s_launch))
    if not args.no_bsdf:
        extra.update(bench_bsdf(lib, scenes, ctx, scene, r, args.steps))
        log("bsdf jobs:", json.dumps(extra["bsdf_jobs"]))
    r.destroy()
    for name, mode in (("render_independent", lib.SAMPLER_INDEPENDENT),
                       ("render_stratified_n", lib.SAMPLER_STRATIFIED_N)):
        if args.no_stratified:
            break
        # the same render with the hash streams / the 1/n-scaled stratified arrays
        scene.sampler = mode
        r2 = lib.Renderer(ctx, scene)
        r2.render()
        ctx.sync()
        t0s = time.perf_counter()
        for _ in range(args.steps):
            r2.render()
        ctx.sync()
        st_ms = (time.perf_counter() - t0s) * 1e3 / args.steps
        r2.destroy()
        extra[name] = {"ms_per_step": st_ms, "msamples_per_s": samples / (st_ms * 1e-3) / 1e6}
        log("%s: %.2f ms/step" % (name, st_ms))
    scene.sampler = lib.SAMPLER_STRATIFIED
    avg_launch_ms = s_trace / s_launch
    bytes_per_launch = TRACE_BYTES_PER_RAY * s_rays / s_launch
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                "kernel": "k_trace in the render, passes serialised (36 B of ray I/O per ray; %.0f rays per launch)"
                          % (s_rays / s_launch),
                "avg_launch_us": avg_launch_ms * 1e3,
                "limiter": "latency/L2: divergence and the latency of node and primitive gathers, not HBM "
                           "bandwidth (profiles/render_c2_pmc.txt: 29 % VALU busy, 33 % of lanes active, "
                           "L1 95 %, L2 57 %, 262 GB/s of HBM-side traffic)"}
    mb = {}
    if not args.no_microbench:
        mb = bench_microbench(lib, scenes, ctx, args.steps)
        log("microbench:", json.dumps(mb))
    if not args.no_adjoint:
        adj = bench_adjoint(lib, scenes, ctx, args.steps, tune=not args.trace_config)
        log("adjoint:", json.dumps(adj))
        mb.update(adj)
    if not args.no_batch:
        bat = bench_batch(lib, scenes, ctx, args.steps, tune=not args.trace_config)
        log("batch:", json.dumps(bat))
        mb.update(bat)
    cpu = None
    if not args.no_cpu_baseline:
        t1 = time.time()
        cpu = cpu_baseline(scene)
        log("cpu baseline in %.1f s: %s" % (time.time() - t1, cpu))
    out = {
        "metric": METRIC, "value": value, "unit": "Msamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene generators, counter-based sample streams)",
        "config": {"workload": wl, "width": scene.width, "height": scene.height,
                   "spp": scene.spp, "shapes": scene.n_shapes,
                   "triangles": int(len(scene.tri_verts)), "spheres": int(len(scene.sphere_radius))},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    out.update(mb)
    out.update(extra)
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
