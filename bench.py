"""Headline benchmark: incoherent closest-hit Mrays/s on one MI355X.

Workload (default): the 3,240 x 3,240-cell fBm heightfield + 64 level-6
icospheres (26,238,080 triangles, one single-level tree); the rays are one
cosine-weighted diffuse bounce from every hit of a 1920x1080 pinhole camera
(the incoherent set).  One step = one trace of the whole bounce buffer
(inputs resident in HBM), timed with HIP events on the trace stream.

Prints one JSON line (see the keys below).  The CPU baseline is the oracle
(C++ restatement of the reference BVHAccel::Intersect over a 4-ary binned-SAH
tree in the reference node layout) timed on this host's cores over a bounded
sample of the same rays.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from luxcore_amd import device as D  # noqa: E402
from luxcore_amd import scenes  # noqa: E402
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX  # noqa: E402

METRIC = "Mrays/s incoherent closest-hit, 10M tris, 1 GPU; p50/p99 wavefront latency"
HBM_PEAK_GBS = 8000.0
RAY_BYTES, HIT_BYTES = 48, 20


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(args, dev):
    t0 = time.time()
    meshes = scenes.heightfield_scene(cells=args.cells, n_spheres=64, seed=3)
    ntris = sum(m.ntris for m in meshes)
    log(f"scene: {ntris} tris in {len(meshes)} meshes ({time.time() - t0:.1f}s)")
    builder = {"sah": D.BUILD_HOST_SAH, "lbvh": D.BUILD_DEVICE_LBVH, "ploc": D.BUILD_DEVICE_PLOC}[args.builder]
    scene = dev.scene_bvh(meshes, layout=args.layout, max_leaf=args.max_leaf, builder=builder)
    info = scene.info()
    log(f"device tree ({args.builder}): {info['nodeCount']} nodes, {info['nodeBytes'] / 2**20:.0f} MiB nodes, "
        f"{info['primBytes'] / 2**20:.0f} MiB tris, build {info['buildSeconds']:.3f}s, "
        f"upload {info['uploadSeconds']:.3f}s, depth {info['maxDepth']}")
    prim = scenes.pinhole_rays(1920, 1080)
    hits = dev.trace_numpy(scene, prim)
    nhit = int((hits["meshIndex"] != NULL_INDEX).sum())
    bounce = scenes.bounce_rays(meshes, prim, hits, seed=5)
    log(f"primaries {prim.shape[0]} ({nhit} hit) -> {bounce.shape[0]} bounce rays")
    return meshes, scene, info, bounce, ntris


def host_info():
    """CPU model, sockets, cores, threads and frequency governor of this host."""
    info = {"threads": os.cpu_count()}
    try:
        with open("/proc/cpuinfo") as f:
            txt = f.read()
        models = [ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ln.startswith("model name")]
        phys = {ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ln.startswith("physical id")}
        cores = {ln.split(":", 1)[1].strip() for ln in txt.splitlines() if ln.startswith("core id")}
        info.update(model=models[0] if models else None, sockets=len(phys) or None,
                    cores_per_socket=len(cores) or None)
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/cpu0/cpufreq/scaling_governor") as f:
            info["governor"] = f.read().strip()
    except OSError:
        info["governor"] = None
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return info


def _timed_passes(ob, sample, threads, target_s, min_passes=5):
    ob.intersect(sample, threads=threads)  # warm-up
    rates, spent = [], 0.0
    while (spent < target_s and len(rates) < 50) or len(rates) < min_passes:
        t0 = time.time()
        ob.intersect(sample, threads=threads)
        dt = time.time() - t0
        spent += dt
        rates.append(sample.shape[0] / dt / 1e6)
    return float(np.median(rates)), len(rates), spent


def cpu_baseline(meshes, rays, target_s):
    """Oracle BVHAccel::Intersect on this host's cores: a warm-up pass, then
    repeated passes over a bounded sample until ~target_s seconds of work
    (at least 5); the median pass rate is reported.  The same sample is also
    traced once with the oracle's counters, for the reference-layout bytes
    per ray (32-byte BVHArrayNode visits, 36-byte triangle gathers), and
    timed over a CLASSIC-built tree for comparison."""
    import oracle  # the checker, used here only as the timed CPU baseline
    threads = os.cpu_count() or 1
    t0 = time.time()
    ob = oracle.OracleBVH(meshes, builder="SAH", tree_type=4, threads=threads)
    log(f"cpu oracle tree built in {time.time() - t0:.1f}s ({threads} threads)")
    probe = rays[: 1 << 15]
    t0 = time.time()
    ob.intersect(probe, threads=threads)
    rate = probe.shape[0] / max(time.time() - t0, 1e-6)
    n = int(min(rays.shape[0], max(1 << 15, rate * target_s / 6)))
    sample = rays[:n]
    value, passes, spent = _timed_passes(ob, sample, threads, target_s)
    _, (visits, tests) = ob.intersect(sample, threads=threads, counters=True)
    # every visited BVHArrayNode (inner or leaf) is 32 bytes; a leaf adds its
    # three 12-byte vertex gathers
    ref_bytes = 32.0 * (visits + tests) / n + 36.0 * tests / n + RAY_BYTES + HIT_BYTES
    del ob
    classic = None
    try:
        t0 = time.time()
        oc = oracle.OracleBVH(meshes, builder="CLASSIC", tree_type=4, threads=threads)
        log(f"cpu CLASSIC tree built in {time.time() - t0:.1f}s")
        classic, _, _ = _timed_passes(oc, sample, threads, target_s / 3, min_passes=3)
        del oc
    except Exception as e:  # the comparison line is optional; the baseline above is not
        log(f"CLASSIC comparison skipped: {e}")
    return {"value": round(value, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of the {rays.shape[0]} bounce rays x {passes} passes ({spent:.1f}s) after a "
                      f"warm-up, oracle BVHAccel::Intersect over a 4-ary binned-SAH tree in the 32-byte "
                      f"BVHArrayNode layout, {threads} std::threads, median pass",
            "classic_tree_value": round(classic, 3) if classic else None,
            "reference_layout_bytes_per_ray": round(ref_bytes, 1),
            "reference_per_ray": {"inner_node_visits": round(visits / n, 3), "leaf_visits": round(tests / n, 3)},
            "host": host_info()}


NODE_READ_BYTES = {0: 64, 1: 128, 2: 96}  # bytes a node visit reads, per layout


def two_level_bytes(c, layout=0):
    """Algorithmic bytes per ray of a two-level trace from the counting kernel:
    Ray + RayHit, node visits (64 bytes 4-wide, 96 read of a 128-byte 8-wide
    node), 48-byte triangle records, and per object entry its 16-byte
    instance record plus the 64-byte matrix (static instance) or the 32-byte
    motion record and the 112 bytes of the segment a sample reads (motion)."""
    rays = max(c["rays"], 1)
    tris = c["prim_tests"] - c["object_entries"]
    b = (NODE_READ_BYTES[layout] * c["node_visits"] + 48 * tris + 16 * c["object_entries"]
         + 64 * c["matrix_entries"] + (32 + 112) * c["motion_entries"])
    return RAY_BYTES + HIT_BYTES + b / rays


def _rate_line(n, wall, lat, steps, b_ray):
    kernel_ms = float(lat.mean())
    achieved = b_ray * n / (kernel_ms * 1e-3) / 1e9
    return {"mrays": round(n * steps / wall / 1e6, 1), "kernel_ms": round(kernel_ms, 4),
            "bytes_per_ray": round(b_ray, 1), "achieved_gbs": round(achieved, 1),
            "frac": round(achieved / HBM_PEAK_GBS, 4)}


def extra_workloads(args, dev):
    """The other configurations in the same run, so their numbers are the
    driver's measurement too: the two-level instance/motion scene (closest
    and any hit), the batched workload (64 buffers over 8 resident scenes in
    one graph replay) and the served workload (continuous batching of
    512x512 path-tracing sessions, next to one plain batch of the same rays).
    Each rate comes with its bytes per ray from the counting kernel and its
    fraction of the 8 TB/s roof."""
    from luxcore_amd import batch as B
    from luxcore_amd import serve
    steps = args.extra_steps
    out = {}
    # -- two level: 1,000 instances of a 50,000-triangle sphere, 200 with 2-key motion
    t0 = time.time()
    isc = scenes.instanced_scene()
    sc = dev.scene_mbvh(isc, layout=2)
    rays = scenes.instanced_rays(1 << 22)
    n = rays.shape[0]
    rb = dev.upload(rays)
    hb = dev.alloc(n * HIT_DTYPE.itemsize)
    tl = {"workload": "1,000 instances of a 50,000-tri displaced sphere (T*R*S, scale 0.5-2, 200^3 box), 200 "
                      "with 2-key motion; 4,194,304 rays uniform in the box, time ~ U[0,1]; 8-wide nodes", "rays": n}
    for name, q in (("closest", 0), ("any", 1)):
        c = dev.trace_counters(sc, rb, hb, n, q)
        wall, lat = _event_timed(dev, lambda: dev.trace(sc, rb, hb, n, q), steps, 3)
        tl[name] = _rate_line(n, wall, lat, steps, two_level_bytes(c, 2))
        tl[name]["per_ray"] = {k: round(c[k] / max(c["rays"], 1), 3)
                               for k in ("node_visits", "prim_tests", "object_entries", "motion_entries")}
    out["two_level"] = tl
    log(f"two level ({time.time() - t0:.1f}s): {tl}")
    rb.free()
    hb.free()
    sc.close()
    # -- batched: 8 resident scenes (10k..5M tris), 64 buffers, one persistent launch replayed from a graph
    t0 = time.time()
    meshes = scenes.batch_scenes()
    bufs = scenes.batch_buffers()
    wl = B.BatchedWorkload(dev, meshes, bufs, layout=2)
    items = wl.items()
    total = wl.total_rays()
    key = B.signature(dev, [m[0].ntris for m in meshes], [c for _, c, _ in bufs], "batched-v3-layout2")
    opts, _ = B.autotune(dev, items, key=key)
    cnt = {"node_visits": 0, "prim_tests": 0, "rays": 0}
    for (s, r, h, cn, q) in items:
        c = dev.trace_counters(s, r, h, cn, q, opts)
        for k in cnt:
            cnt[k] += c[k]
    # 8-wide nodes: 96 bytes read per visit
    b_ray = RAY_BYTES + HIT_BYTES + (96 * cnt["node_visits"] + 48 * cnt["prim_tests"]) / max(cnt["rays"], 1)
    bt = dev.batch(items, opts)
    bt.capture()
    wall, lat = _event_timed(dev, bt.replay, steps, 2)
    bl = {"workload": "8 scenes 10k..5M tris (soups + sphere grids), 64 buffers log-uniform 4Ki..4Mi rays, "
                      "mixed closest/any; 8-wide nodes", "rays": total,
          "shape": {"block": opts.blockSize, "lds": opts.ldsStack, "refill": opts.refillThreshold}}
    bl.update(_rate_line(total, wall, lat, steps, b_ray))
    bl["per_ray"] = {"node_visits": round(cnt["node_visits"] / max(cnt["rays"], 1), 3),
                     "tri_tests": round(cnt["prim_tests"] / max(cnt["rays"], 1), 3)}
    pwall, _ = _event_timed(dev, lambda: [dev.trace(s, r, h, cn, q, opts) for (s, r, h, cn, q) in items], steps, 1)
    bl["plain_per_buffer_mrays"] = round(total * steps / pwall / 1e6, 1)
    bt.close()
    out["batched"] = bl
    log(f"batched ({time.time() - t0:.1f}s): {bl}")
    # -- served: Poisson arrivals of 512x512 sessions over the same 8 scenes
    t0 = time.time()
    scs = wl.scenes
    for rbuf in wl.ray_bufs + wl.hit_bufs:
        rbuf.free()
    sopts = D.trace_opts(256, 8, 0, 12)
    w = serve.Server(dev, scs, opts=sopts)
    w.run(serve.poisson_specs(1000.0, 0.004, len(scs), seed=99))
    w.close()
    # 1200 sessions/s is more than the GPU serves: the admission queue never
    # empties, so every tick carries its full ray budget (the server's
    # throughput at this wavefront-latency bound)
    specs = serve.poisson_specs(1200.0, 0.5, len(scs), seed=1200)
    srv = serve.Server(dev, scs, opts=sopts, max_rays_per_tick=5 << 19)
    swall = srv.run(specs, time_limit=120.0)
    row = serve.served_row(srv, swall, 1200.0, 5 << 19)
    srv.close()
    plain = serve.plain_comparison(dev, scs, sopts, n_sessions=16)
    out["served"] = {
        "workload": "Poisson arrivals (1200 sessions/s for 0.5 s: more than the GPU serves, so the admission "
                    "queue never empties and every tick is full) over the 8 batched scenes; 512x512 camera + <=8 "
                    "bounces (cosine, RR 0.7 after 2) + 1 shadow ray per bounce, GPU shading (lrh_shade), 2 "
                    "cohorts, 2.5Mi rays per tick, each tick's launch on 80% of the persistent grid (the "
                    "server's default), 4 GiB HBM budget",
        "run": row,
        "latency_ms": {"p50": row["p50_ms"], "p99": row["p99_ms"], "samples": row["wavefronts"],
                       "per": "one session wavefront: ready -> completion of the tick that traced it"},
        "busy_mrays": row["busy_mrays"], "trace_mrays": row["trace_mrays"],
        "plain_single_batch": plain,
        "trace_over_plain_batch": round(row["trace_mrays"] / plain["plain_batch_mrays"], 3),
        "busy_over_plain_batch": round(row["busy_mrays"] / plain["plain_batch_mrays"], 3),
        "note": "busy = GPU time of the ticks (batched trace + shading kernels); trace = the batched traces alone; "
                "plain_single_batch = the first 16 sessions' wavefronts captured and traced as ONE lrh_trace_batch "
                "(and as one lrh_trace each); live rays only"}
    log(f"served ({time.time() - t0:.1f}s): {out['served']}")
    for s in scs:
        s.close()
    return out


def run_batched(args, dev):
    """64 ray buffers (log-uniform 4Ki..4Mi rays, mixed closest/any) over 8
    resident scenes (10k..5M tris): plain per-buffer traces vs one persistent
    batched launch replayed from a hipGraph, launch shape from the autotuner."""
    from luxcore_amd import batch as B
    t0 = time.time()
    meshes = scenes.batch_scenes()
    bufs = scenes.batch_buffers()
    wl = B.BatchedWorkload(dev, meshes, bufs, layout=args.layout)
    total = wl.total_rays()
    log(f"batched workload: {sum(m[0].ntris for m in meshes)} tris in 8 scenes, {len(bufs)} buffers, "
        f"{total} rays ({time.time() - t0:.1f}s)")
    build, build_table = {"layout": args.layout, "max_leaf": args.max_leaf}, None
    if not args.layout_given:
        # node layout and leaf size from the autotuner: each candidate build of
        # the 8 scenes timed over the same buffers with its best launch shape
        tune_items = [(si, wl.ray_bufs[i], wl.hit_bufs[i], c, q) for i, (si, c, q) in enumerate(bufs)]
        best, build_table = B.tune_build(dev, meshes, tune_items, builds=((0, 4), (1, 4), (2, 4), (2, 2)),
                                         shapes=[(256, 8, 8), (128, 16, 16), (256, 8, 12), (128, 8, 12)], log=log)
        build = {"layout": best["layout"], "max_leaf": best["max_leaf"]}
        if (build["layout"], build["max_leaf"]) != (args.layout, args.max_leaf):
            for sc in wl.scenes:
                sc.close()
            wl.scenes = [dev.scene_bvh(m, layout=build["layout"], max_leaf=build["max_leaf"]) for m in meshes]
        log(f"autotuned build: layout {build['layout']} max_leaf {build['max_leaf']}")
    items = wl.items()
    key = B.signature(dev, [m[0].ntris for m in meshes], [c for _, c, _ in bufs],
                      f"batched-v3-layout{build['layout']}-leaf{build['max_leaf']}")
    opts, table = B.autotune(dev, items, key=key, log=log)
    log(f"autotuned shape: block {opts.blockSize} lds {opts.ldsStack} refill {opts.refillThreshold}")
    # plain: one lrh_trace per buffer, same launch shape
    for _ in range(2):
        for (sc, rb, hb, c, q) in items:
            dev.trace(sc, rb, hb, c, q, opts)
    dev.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for (sc, rb, hb, c, q) in items:
            dev.trace(sc, rb, hb, c, q, opts)
    dev.sync()
    plain_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    b = dev.batch(items, opts)
    b.capture()
    for _ in range(2):
        b.replay()
    dev.sync()
    evs = [(dev.event(), dev.event()) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record()
        b.replay()
        e1.record()
    dev.sync()
    batched_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    lat = np.array([e0.elapsed_ms(e1) for e0, e1 in evs])
    value = total / batched_ms / 1e3
    plain = total / plain_ms / 1e3
    log(f"batched {value:.1f} Mrays/s ({batched_ms:.2f} ms/launch) vs plain {plain:.1f} Mrays/s "
        f"({plain_ms:.2f} ms for 64 launches)")
    return {"metric": "Mrays/s batched: 64 ray buffers over 8 resident scenes in one persistent hipGraph launch",
            "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": 2,
            "ms_per_step": round(batched_ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "8 scenes 10k..5M tris (soups + sphere grids), 64 buffers log-uniform 4Ki..4Mi "
                                   "rays, mixed closest/any", "rays_per_step": total,
                       "shape": {"block": opts.blockSize, "lds": opts.ldsStack, "refill": opts.refillThreshold,
                                 "blocks": b.blocks},
                       "build": build},
            "build_autotune": build_table,
            "plain_mrays": round(plain, 2), "batched_over_plain": round(value / plain, 3),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                           "p99": round(float(np.percentile(lat, 99)), 4), "per": "one batched launch"}}


def _event_timed(dev, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    dev.sync()
    evs = [(dev.event(), dev.event()) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        fn()
        b.record()
    dev.sync()
    wall = time.perf_counter() - t0
    return wall, np.array([a.elapsed_ms(b) for a, b in evs])


def run_scene(args, dev):
    """A LuxCore scene file read by the product reader and built with the
    DataSet AUTO policy; closest-hit throughput of its camera primaries
    (1920x1080 pinhole at the scene camera) and of incoherent rays (origins
    uniform in the world box, directions uniform, time uniform in [0,1])."""
    import tempfile
    from luxcore_amd import scenefile as SF
    path = args.scene
    if path is None:
        from tests.scene_fixtures import unpack
        path = os.path.join(unpack("lightinstances", tempfile.mkdtemp(prefix="lrh_scene_")), "scene.scn")
    t0 = time.time()
    sf = SF.load_scene(path)
    t_load = time.time() - t0
    t0 = time.time()
    scene = SF.scene_on_device(dev, sf)
    t_build = time.time() - t0
    info = scene.info()
    log(f"scene {path}: {sf.info['nObjects']} objects, {sf.info['nTris']} unique / {sf.info['nWorldTris']} world "
        f"tris, {'MBVH' if sf.two_level else 'BVH'}; read {t_load:.2f}s, build+upload {t_build:.2f}s")
    rng = np.random.default_rng(11)
    lo, hi = sf.world_bounds()
    incoh = scenes.random_rays(args.scene_rays, lo, hi, rng)
    incoh["time"] = rng.random(incoh.shape[0], dtype=np.float32)
    orig, target, up, fov = sf.camera
    sets = {"incoherent": incoh}
    if np.linalg.norm(np.subtract(target, orig)) > 0 and np.linalg.norm(up) > 0:
        prim = scenes.pinhole_rays(1920, 1080, eye=orig, look=target, up=up, vfov_deg=fov)
        prim["time"] = rng.random(prim.shape[0], dtype=np.float32)
        sets["camera"] = prim
    res = {}
    for name, rays in sets.items():
        n = rays.shape[0]
        rb = dev.upload(rays)
        hb = dev.alloc(n * HIT_DTYPE.itemsize)
        wall, lat = _event_timed(dev, lambda: dev.trace(scene, rb, hb, n, 0), args.steps, args.warmup)
        hits = dev.download(hb, HIT_DTYPE, n)
        res[name] = {"rays": int(n), "mrays": round(n * args.steps / wall / 1e6, 2),
                     "kernel_ms": round(float(lat.mean()), 4),
                     "p50_ms": round(float(np.percentile(lat, 50)), 4),
                     "p99_ms": round(float(np.percentile(lat, 99)), 4),
                     "hit_rate": round(float((hits["meshIndex"] != NULL_INDEX).mean()), 4)}
        log(f"{name}: {res[name]}")
        rb.free()
        hb.free()
    main_set = res["incoherent"]
    return {"metric": "Mrays/s closest-hit on a LuxCore scene file (incoherent rays)",
            "value": main_set["mrays"], "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": main_set["kernel_ms"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "scene file + synthetic rays",
            "config": {"workload": os.path.basename(os.path.dirname(path)) + "/" + os.path.basename(path),
                       "objects": sf.info["nObjects"], "unique_tris": sf.info["nTris"],
                       "world_tris": sf.info["nWorldTris"], "two_level": sf.two_level,
                       "node_bytes": info["nodeBytes"], "prim_bytes": info["primBytes"],
                       "read_seconds": round(t_load, 3), "build_upload_seconds": round(t_build, 3)},
            "sets": res}


def _moved(sc, rng, extent, step=5.0):
    """The same objects with new instance matrices and motion keys."""
    import copy
    out = copy.deepcopy(sc)
    for o in range(len(out.obj_kind)):
        if out.obj_kind[o] == 1:
            out.obj_mat[o] = scenes.random_trs(rng, extent)
        elif out.obj_kind[o] == 2:
            M = scenes.random_trs(rng, extent).astype(np.float64)
            for k in range(int(out.key_off[o]), int(out.key_off[o + 1])):
                Mk = M.copy()
                Mk[:3, 3] += rng.uniform(-step, step, size=3)
                out.key_mats[k] = np.linalg.inv(Mk).astype(np.float32)
    return out


def run_refit(args, dev):
    """Edit latency of a two-level scene (1,000 instances of a 50,000-triangle
    sphere, 200 with 2-key motion): every object moves, then either
    lrh_scene_update (root rebuilt on the host, MotionSystem::Bound on the
    host, upload) or lrh_scene_refit (same boxes computed on the GPU, root
    refitted in place).  Also the trace rate after each, since a refitted
    tree keeps the old grouping."""
    rng = np.random.Generator(np.random.PCG64(77))
    sc = scenes.instanced_scene()
    rays = scenes.instanced_rays(1 << 22, seed=78)
    edits = [_moved(sc, rng, 100.0) for _ in range(6)]
    scene = dev.scene_mbvh(sc, layout=args.layout)
    n = rays.shape[0]
    rb = dev.upload(rays)
    hb = dev.alloc(n * HIT_DTYPE.itemsize)
    res = {}
    for how in ("update", "refit"):
        fn = scene.update if how == "update" else scene.refit
        fn(edits[0])  # warm
        times = []
        for e in edits[1:]:
            t0 = time.perf_counter()
            fn(e)
            times.append((time.perf_counter() - t0) * 1e3)
        wall, lat = _event_timed(dev, lambda: dev.trace(scene, rb, hb, n, 0), args.steps, args.warmup)
        res[how] = {"edit_ms_median": round(float(np.median(times)), 3), "edit_ms_min": round(float(min(times)), 3),
                    "trace_mrays_after": round(n * args.steps / wall / 1e6, 1)}
        log(f"{how}: {res[how]}")
    return {"metric": "two-level edit latency (ms): host rebuild vs GPU refit",
            "value": res["refit"]["edit_ms_median"], "unit": "ms", "n_gpus": 1, "steps": len(edits) - 1,
            "warmup": 1, "ms_per_step": res["refit"]["edit_ms_median"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "1,000 instances of a 50,000-tri displaced sphere, 200 with 2-key motion; "
                                   "every object moves per edit", "rays_traced_after": n},
            "update": res["update"], "refit": res["refit"],
            "speedup": round(res["update"]["edit_ms_median"] / res["refit"]["edit_ms_median"], 1)}


def run_build(args, dev):
    """Device BVH builds of the headline scene (26.2M triangles): LBVH (radix
    tree) and PLOC (clustering), each built --steps times (vertices and
    triangles uploaded from host memory each time, timed apart), and the
    closest-hit rate of the headline bounce rays over each tree and over the
    host binned-SAH tree."""
    meshes = scenes.heightfield_scene(cells=args.cells, n_spheres=64, seed=3)
    ntris = sum(m.ntris for m in meshes)
    t0 = time.perf_counter()
    sah = dev.scene_bvh(meshes, layout=args.layout, max_leaf=args.max_leaf)
    sah_s = time.perf_counter() - t0
    prim = scenes.pinhole_rays(1920, 1080)
    rays = scenes.bounce_rays(meshes, prim, dev.trace_numpy(sah, prim), seed=5)
    n = rays.shape[0]
    rb = dev.upload(rays)
    hb = dev.alloc(n * HIT_DTYPE.itemsize)
    trace_steps = 50

    def trace_rate(sc):
        wall, _ = _event_timed(dev, lambda: dev.trace(sc, rb, hb, n, 0), trace_steps, 5)
        vis, tests, cnt = dev.trace_stats(sc, rb, hb, n, 0)
        return round(n * trace_steps / wall / 1e6, 1), round(vis / cnt, 2), round(tests / cnt, 2)
    res = {}
    r, v, t = trace_rate(sah)
    res["host_sah"] = {"build_s": round(sah_s, 3), "trace_mrays": r, "node_visits": v, "tri_tests": t,
                       "nodes": sah.info()["nodeCount"]}
    sah.close()
    log(f"host SAH: {res['host_sah']}")
    for name, b in (("lbvh", D.BUILD_DEVICE_LBVH), ("ploc", D.BUILD_DEVICE_PLOC)):
        builds, uploads, sc = [], [], None
        for _ in range(args.warmup // 5 + 1 + args.steps):
            if sc is not None:
                sc.close()
            sc = dev.scene_bvh(meshes, layout=args.layout, max_leaf=args.max_leaf, builder=b)
            builds.append(sc.info()["buildSeconds"])
            uploads.append(sc.info()["uploadSeconds"])
        builds, uploads = builds[-args.steps:], uploads[-args.steps:]
        r, v, t = trace_rate(sc)
        med = float(np.median(builds))
        res[name] = {"build_s_median": round(med, 4), "build_s_min": round(min(builds), 4),
                     "upload_s_median": round(float(np.median(uploads)), 4),
                     "mtris_per_s": round(ntris / med / 1e6, 1), "trace_mrays": r, "node_visits": v,
                     "tri_tests": t, "nodes": sc.info()["nodeCount"], "depth": sc.info()["maxDepth"]}
        sc.close()
        log(f"{name}: {res[name]}")
    return {"metric": "device BVH build throughput (Mtris/s), LBVH", "value": res["lbvh"]["mtris_per_s"],
            "unit": "Mtris/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup // 5 + 1,
            "ms_per_step": round(res["lbvh"]["build_s_median"] * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "heightfield 3240x3240 fBm + 64 L6 icospheres; build from vertices and "
                                   "triangles resident in HBM (upload timed apart); trace: the headline bounce "
                                   "rays", "triangles": ntris, "rays": n, "layout": args.layout,
                       "max_leaf": args.max_leaf},
            "builders": res}


def run_configs(args, dev):
    """Closest-hit and any-hit rates of every workload the parity tests use,
    HIP-event timed, with the oracle on this host's cores beside each
    closest-hit set (bounded sample, median pass):
      plain        box + level-6 icosphere + 20,000 random triangles (101,932),
                   1,048,576 rays uniform in the box;
      heightfield  the 26.2M-triangle headline scene: 1920x1080 primaries
                   (coherent), one cosine bounce per hit (incoherent, the
                   headline set) and one shadow ray per hit towards a 1x1 area
                   light (any-hit);
      instanced    1,000 instances of a 50,000-triangle sphere, 200 with 2-key
                   motion, 4,194,304 rays with time ~ U[0,1]."""
    threads = os.cpu_count() or 1
    out = {}

    def rate(scene, rays, query):
        n = rays.shape[0]
        rb = dev.upload(rays)
        hb = dev.alloc(n * HIT_DTYPE.itemsize)
        wall, lat = _event_timed(dev, lambda: dev.trace(scene, rb, hb, n, query), args.steps, args.warmup)
        hits = dev.download(hb, HIT_DTYPE, n)
        rb.free()
        hb.free()
        return {"rays": int(n), "mrays": round(n * args.steps / wall / 1e6, 1),
                "kernel_ms": round(float(lat.mean()), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
                "hit_rate": round(float((hits["meshIndex"] != NULL_INDEX).mean()), 4)}

    def cpu(ob, rays, seconds=4.0):
        probe = rays[: 1 << 14]
        t0 = time.time()
        ob.intersect(probe, threads=threads)
        r = probe.shape[0] / max(time.time() - t0, 1e-6)
        sample = rays[: int(min(rays.shape[0], max(1 << 14, r * seconds / 6)))]
        v, passes, spent = _timed_passes(ob, sample, threads, seconds, min_passes=3)
        return {"mrays": round(v, 3), "sample_rays": int(sample.shape[0]), "passes": passes, "threads": threads}

    sets = {x.strip() for x in args.sets.split(",") if x.strip()}
    if not sets <= {"plain", "heightfield", "instanced"}:
        raise SystemExit(f"--sets: unknown workload in {args.sets}")
    if "plain" in sets:
        _configs_plain(args, dev, rate, cpu, out)
    if "heightfield" in sets:
        _configs_heightfield(args, dev, rate, out)
    if "instanced" in sets:
        _configs_instanced(args, dev, rate, cpu, out)
    value = out["heightfield"]["bounce_closest"]["mrays"] if "heightfield" in out else None
    return {"metric": "Mrays/s per workload (closest / any hit)", "value": value,
            "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "plain 101,932 tris / heightfield 26.2M tris / 1,000 instances",
                       "sets": sorted(sets)}, "configs": out}


def _configs_plain(args, dev, rate, cpu, out):
    import oracle  # the checker, timed here only as the CPU baseline
    threads = os.cpu_count() or 1
    meshes = scenes.plain_scene()
    sc = dev.scene_bvh(meshes, layout=args.layout)
    rays = scenes.plain_rays(1 << 20)
    out["plain"] = {"triangles": sum(m.ntris for m in meshes), "closest": rate(sc, rays, 0),
                    "any": rate(sc, rays, 1)}
    if not args.no_cpu:
        out["plain"]["cpu_closest"] = cpu(oracle.OracleBVH(meshes, threads=threads), rays)
    sc.close()
    log(f"plain: {out['plain']}")


def _configs_heightfield(args, dev, rate, out):
    meshes = scenes.heightfield_scene()
    sc = dev.scene_bvh(meshes, layout=args.layout)
    prim = scenes.pinhole_rays(1920, 1080)
    h = dev.trace_numpy(sc, prim)
    bounce = scenes.bounce_rays(meshes, prim, h, seed=5)
    shadow = scenes.shadow_rays(meshes, prim, h, seed=6)
    out["heightfield"] = {"triangles": sum(m.ntris for m in meshes), "camera_closest": rate(sc, prim, 0),
                          "bounce_closest": rate(sc, bounce, 0), "shadow_any": rate(sc, shadow, 1),
                          "shadow_closest": rate(sc, shadow, 0)}
    sc.close()
    log(f"heightfield: {out['heightfield']}")


def _configs_instanced(args, dev, rate, cpu, out):
    import oracle  # the checker, timed here only as the CPU baseline
    threads = os.cpu_count() or 1
    isc = scenes.instanced_scene()
    sc = dev.scene_mbvh(isc, layout=args.layout)
    rays = scenes.instanced_rays(1 << 22)
    out["instanced"] = {"objects": int(len(isc.obj_kind)), "motion_objects": int((isc.obj_kind == 2).sum()),
                        "closest": rate(sc, rays, 0), "any": rate(sc, rays, 1)}
    if not args.no_cpu:
        out["instanced"]["cpu_closest"] = cpu(oracle.OracleMBVH(isc, threads=threads), rays)
    sc.close()
    log(f"instanced: {out['instanced']}")


def run_render(args, dev):
    """Path tracing on the GPU end to end: camera rays (lrh_rays_camera), then
    per vertex a closest-hit trace of the path rays, an any-hit trace of the
    previous vertex's shadow rays and one shading step (lrh_shade: Lambertian,
    triangle-light next-event estimation, Russian roulette), 8 vertices, one
    sample per pixel per pass, 1280x720, on the bundled bigmonkey scene with
    its light made constant (8.0) and its monkey red.  Reports samples
    (paths) per second and rays per second, HIP-event timed over whole passes."""
    import tempfile
    from luxcore_amd import scenefile as SF
    from tests.scene_fixtures import unpack
    d = unpack("bigmonkey", tempfile.mkdtemp(prefix="lrh_render_"))
    path = os.path.join(d, "bigmonkey-lit.scn")
    with open(os.path.join(d, "bigmonkey.scn")) as f, open(path, "w") as g:
        g.write(f.read() + "\nscene.materials.redmat.kd = 0.75 0.0 0.0\n"
                "scene.materials.whitelight.emission = 8.0 8.0 8.0\n"
                "scene.materials.whitelight.emission.gain = 1.0 1.0 1.0\n")
    sf = SF.load_scene(path)
    scene = SF.scene_on_device(dev, sf)
    mats, lights = sf.materials(), sf.light_triangles()
    W, H, depth = 1280, 720, 8
    n = W * H
    orig, target, up, fov = sf.camera
    bufs = {k: dev.alloc(n * 48) for k in ("ray", "next", "shadow")}
    hb, shb = dev.alloc(n * 20), dev.alloc(n * 20)
    pb = dev.alloc(n * D.PATH_DTYPE.itemsize)
    mb, lb = dev.upload(mats), dev.upload(lights)
    counter = dev.alloc(4)
    zero = np.zeros(1, np.uint32)

    def one_pass(seed):
        rb, nb, sb = bufs["ray"], bufs["next"], bufs["shadow"]
        dev.camera_rays(rb, orig, target, up, fov, W, H)
        dev.paths_init(pb, n)
        traced = n
        for k in range(depth):
            dev.trace(scene, rb, hb, n, 0)
            if k:
                dev.trace(scene, sb, shb, n, 1)
            dev.shade(scene, mb, len(mats), lb, len(lights), rb, hb, shb if k else None, pb, nb, sb, n,
                      max_depth=depth, rr_depth=3, rr_survive=0.7, seed=seed, live_counter=counter.ptr)
            rb, nb = nb, rb
        return traced

    dev.upload(zero, counter)
    for w in range(max(1, args.warmup)):
        one_pass(1000 + w)
    dev.sync()
    dev.upload(zero, counter)
    dev.sync()
    passes = args.steps
    wall, lat = _event_timed(dev, lambda: one_pass(7), passes, 0)
    live = int(dev.download(counter, np.uint32, 1)[0])
    rad = dev.download(pb, D.PATH_DTYPE, n)["rad"]
    # camera rays plus the live next-vertex and shadow rays the shading steps wrote (the last
    # step's are written but not traced)
    rays_per_pass = n + live / passes
    out = {"metric": "path tracing on the GPU: samples/s (1280x720, 8 vertices, Lambertian + NEE)",
           "value": round(n * passes / wall / 1e6, 2), "unit": "Msamples/s", "n_gpus": 1, "steps": passes,
           "warmup": args.warmup, "ms_per_step": round(float(lat.mean()), 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "bundled bigmonkey scene",
           "config": {"workload": "bigmonkey.scn (3 objects, 6,320 tris, light emission 8.0), 1280x720, 1 spp "
                                  "per pass, 8 path vertices, RR 0.7 from vertex 3", "paths_per_pass": n},
           "rays_written_per_pass": round(rays_per_pass),
           "mrays_live": round(rays_per_pass * passes / wall / 1e6, 1),
           "mean_radiance": [round(float(x), 4) for x in rad.mean(axis=0)]}
    log(f"render: {out['value']} Msamples/s, {out['mrays_live']} Mrays/s live, {out['ms_per_step']} ms per pass")
    for b in list(bufs.values()) + [hb, shb, pb, mb, lb, counter]:
        b.free()
    scene.close()
    return out


def calibration_copy(args, dev):
    """A device-to-device copy of --calib-copy MiB after the timed work: a
    kernel of known traffic that scales rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    if not args.calib_copy:
        return
    nb = args.calib_copy << 20
    src = dev.alloc(nb)
    dst = dev.alloc(nb)
    D.check(D.lib().lrh_memcpy(dev.h, dst.ptr, src.ptr, nb, D.LRH_D2D, None))
    dev.sync()
    src.free()
    dst.free()
    log(f"calibration copy: {args.calib_copy} MiB device-to-device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", choices=("headline", "batched", "served", "scene", "refit", "configs", "render",
                                       "build"),
                    default="headline",
                    help="headline: incoherent closest-hit on the 26.2M-tri scene (the default JSON line); "
                         "batched: 64 buffers over 8 resident scenes in one launch; served: continuous batching; "
                         "scene: a LuxCore scene file (--scene, default the bundled lightinstances scene); "
                         "refit: two-level edit latency, host rebuild (lrh_scene_update) vs GPU refit; "
                         "configs: closest/any-hit rates of the plain, heightfield and instanced workloads; "
                         "render: path tracing of a scene file with lrh_shade (samples per second); "
                         "build: device BVH builds (LBVH, PLOC) of the headline scene and their trace rates")
    ap.add_argument("--sets", default="plain,heightfield,instanced",
                    help="--mode configs: which workloads (comma list of plain, heightfield, instanced)")
    ap.add_argument("--scene", default=None, help="scene file for --mode scene (.scn/.cfg)")
    ap.add_argument("--scene-rays", type=int, default=1 << 22, help="incoherent rays for --mode scene")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed traces (default 1,000 for the headline, whose p50/p99 wavefront latency is "
                         "taken over them; 20 per set in the other modes)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cells", type=int, default=3240)
    ap.add_argument("--layout", type=int, default=None, choices=(0, 1, 2),
                    help="node layout: 0 4-wide 64B quantized, 1 4-wide 128B fp32, 2 8-wide 128B quantized "
                         "(default 2)")
    ap.add_argument("--max-leaf", type=int, default=4)
    ap.add_argument("--builder", choices=("sah", "lbvh", "ploc"), default="sah",
                    help="headline tree: host binned SAH (default), or a device build: LBVH (radix tree) or "
                         "PLOC (clustering)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lds", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--serve-setting", default=None, metavar="RATE,TICK_MI,COHORTS",
                    help="--mode served: run one setting (sessions per second, Mi rays per tick, cohorts) "
                         "instead of the sweep, e.g. 1200,2.5,2")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline only: skip the two-level, batched and served workloads of the JSON line")
    ap.add_argument("--extra-steps", type=int, default=20, help="timed traces per extra workload")
    ap.add_argument("--traffic-file", default=None,
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE measurement of this workload (bytes per ray); default: "
                         "the profiles/ file of the node layout")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="after the timed region, copy this many MiB device-to-device (a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs on a single "
              "MI355X: the BVH would be replicated per GPU and rays split by range, with no exchange step)",
              file=sys.stderr)
        sys.exit(2)

    if args.steps is None:
        args.steps = 1000 if args.mode == "headline" else 20
    args.layout_given = args.layout is not None
    if args.layout is None:
        args.layout = 2
    dev = D.Device(0)
    log(f"device {dev.name}, {dev.cu_count} CUs")
    modes = {"batched": run_batched, "scene": run_scene, "refit": run_refit, "configs": run_configs,
             "render": run_render, "build": run_build}
    if args.mode == "served":
        from luxcore_amd import serve
        modes["served"] = lambda a, d: serve.bench(a, d, log)
    if args.mode in modes:
        out = modes[args.mode](args, dev)
        calibration_copy(args, dev)
        print(json.dumps(out), flush=True)
        return
    meshes, scene, info, rays, ntris = build_workload(args, dev)
    n = rays.shape[0]
    rb = dev.upload(rays)
    hb = dev.alloc(n * HIT_DTYPE.itemsize)
    opts = D.trace_opts(args.block, args.lds, args.blocks_per_cu, args.refill)
    use_opts = opts if any((args.block, args.lds, args.blocks_per_cu, args.refill)) else None

    # counting pass: node and triangle-record fetches per ray of this traversal
    nodes, prims, counted = dev.trace_stats(scene, rb, hb, n, 0, use_opts)
    # bytes a node visit reads: 64-byte 4-wide record, 128-byte fp32 record,
    # the 96 bytes of a 128-byte 8-wide record that hold its data
    node_bytes = NODE_READ_BYTES[args.layout]
    b_ray = RAY_BYTES + HIT_BYTES + node_bytes * nodes / counted + 48 * prims / counted
    log(f"per ray: {nodes / counted:.2f} node visits, {prims / counted:.2f} triangle tests, "
        f"B_ray = {b_ray:.1f} bytes")

    for _ in range(args.warmup):
        dev.trace(scene, rb, hb, n, 0, use_opts)
    dev.sync()
    evs = [(dev.event(), dev.event()) for _ in range(args.steps)]
    dev.sync()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        dev.trace(scene, rb, hb, n, 0, use_opts)
        b.record()
    dev.sync()
    wall = time.perf_counter() - t0
    lat = np.array([a.elapsed_ms(b) for a, b in evs])
    hits = dev.download(hb, HIT_DTYPE, n)
    value = n * args.steps / wall / 1e6
    kernel_ms = float(lat.mean())
    achieved = b_ray * n / (kernel_ms * 1e-3) / 1e9
    # wavefront latency percentiles over at least 1,000 wavefronts, whatever --steps is
    n_lat = max(1000, args.steps)
    if n_lat > args.steps:
        _, lat_p = _event_timed(dev, lambda: dev.trace(scene, rb, hb, n, 0, use_opts), n_lat, 0)
    else:
        lat_p = lat
    log(f"{value:.1f} Mrays/s wall, kernel {kernel_ms:.3f} ms avg; over {lat_p.size} wavefronts p50 "
        f"{np.percentile(lat_p, 50):.3f} p99 {np.percentile(lat_p, 99):.3f} ms, hit rate "
        f"{(hits['meshIndex'] != NULL_INDEX).mean():.3f}")

    calibration_copy(args, dev)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(meshes, rays, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.3f} Mrays/s on {cpu['cores']} threads")

    traffic, traffic_source = None, None
    if args.traffic_file is None:
        args.traffic_file = os.path.join(ROOT, "profiles", {0: "traffic_heightfield_bounce.json",
                                                            2: "traffic_heightfield_bounce_wide8.json"}
                                         .get(args.layout, "none.json"))
    tr = None
    if args.traffic_file and os.path.exists(args.traffic_file):
        with open(args.traffic_file) as f:
            tr = json.load(f)
    if (tr is not None and args.cells == 3240 and tr.get("layout", 0) == args.layout and args.builder == "sah"
            and args.max_leaf == 4):
        traffic = round(tr["traffic_bytes_per_ray"] * n)
        traffic_source = (f"not measured in this run: {os.path.relpath(args.traffic_file, ROOT)} "
                          f"(rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel on this workload, scaled by a copy "
                          f"kernel of known size) x the rays of this run")

    extra = {} if args.no_extra else extra_workloads(args, dev)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "heightfield 3240x3240 fBm + 64 L6 icospheres, 1920x1080 primaries -> 1 cosine "
                               "bounce per hit (incoherent closest-hit)",
                   "triangles": ntris, "rays_per_step": int(n),
                   "builder": {"sah": "host binned SAH", "lbvh": "device LBVH", "ploc": "device PLOC"}[args.builder],
                   "build_s": round(info["buildSeconds"], 3), "node_layout": {0: "bvh4 64B quantized", 1: "bvh4 128B fp32",
                                                    2: "bvh8 128B quantized"}[args.layout], "max_leaf": args.max_leaf,
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": round(float(np.percentile(lat_p, 50)), 4),
                       "p99": round(float(np.percentile(lat_p, 99)), 4), "samples": int(lat_p.size),
                       "per": "one wavefront = one lrh_trace of the bounce buffer"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_source,
                     "bytes_per_ray": round(b_ray, 1),
                     "per_ray": {"node_visits": round(nodes / counted, 3),
                                 "tri_tests": round(prims / counted, 3)},
                     "kernel_ms": round(kernel_ms, 4),
                     "compulsory_bytes_per_ray": RAY_BYTES + HIT_BYTES,
                     "compulsory_ceiling_mrays": round(HBM_PEAK_GBS * 1e9 / (RAY_BYTES + HIT_BYTES) / 1e6, 1)},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
