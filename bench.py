"""Benchmark of the kd-tree closest-hit hot path on one MI355X.

Workload: 1,311,488 triangles (icosphere level 8 radially displaced by 0.05 x
3-octave value noise, plus 64 random boxes) and 16,777,216 FP32 rays: 8,388,608
coherent primary rays of a 4096x2048 pinhole at (0,0,-3.5) (40 deg) and
8,388,608 incoherent rays (origins uniform in the scene box, directions uniform
on the sphere).  One step = one closest-hit pass over all rays, inputs and
outputs resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies a 16-lane
# SIMD for 4 cycles; 2.4 GHz peak engine clock
N_SIMDS, CLOCK_GHZ, CYCLES_PER_WAVE_INSTR = 1024, 2.4, 4
VALU_PEAK_GINSTR = N_SIMDS * CLOCK_GHZ / CYCLES_PER_WAVE_INSTR  # 614.4 G wave instructions/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-stride", type=int, default=4,
                    help="the CPU baseline traces every n-th ray of each batch")
    ap.add_argument("--cpu-single-stride", type=int, default=64,
                    help="the single-thread CPU figure traces every n-th ray of each batch")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--skip-pt", action="store_true", help="skip the 1080p path-tracing leg")
    ap.add_argument("--skip-adjoint", action="store_true", help="skip the adjoint/primal leg")
    ap.add_argument("--skip-batch", action="store_true", help="skip the batched/sequential leg")
    ap.add_argument("--skip-kd-build", action="store_true",
                    help="skip the device kd-tree build timing")
    ap.add_argument("--skip-mqo", action="store_true",
                    help="skip the Metasequoia scene load timing")
    ap.add_argument("--skip-bdpt", action="store_true",
                    help="skip the bidirectional path-tracing leg")
    ap.add_argument("--skip-idpt", action="store_true",
                    help="skip the importance-driven path-tracing leg")
    ap.add_argument("--pt-spp", type=int, default=256,
                    help="samples per pixel timed in the 1080p leg (the workload renders 256)")
    ap.add_argument("--pt-cpu-rows", type=int, default=128,
                    help="image rows the CPU baseline renders at 1 spp in the 1080p leg")
    ap.add_argument("--pt-cpu-single-rows", type=int, default=8,
                    help="image rows of the single-thread CPU figure in the 1080p leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device-to-device copy (the reference dispatch "
                         "that scales rocprof FETCH_SIZE / WRITE_SIZE readings)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the closest-hit batch would be split across GPUs with a replicated tree).",
              file=sys.stderr)
        sys.exit(2)
    from xyz_amd import build as xbuild
    from xyz_amd import capi, hip, scenes
    if not os.path.exists(xbuild.LIB):
        xbuild.build()
    if hip.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)

    t0 = time.time()
    mesh = scenes.closest_hit_mesh(8, 64, seed=1)
    (oc, dc), (oi, di) = scenes.closest_hit_rays(mesh, 4096, 2048, seed=2)
    n_c, n_i = len(oc), len(oi)
    log(f"[bench] workload generated in {time.time() - t0:.1f}s: {len(mesh['v0'])} tris, "
        f"{n_c}+{n_i} rays")
    scene = capi.Scene(mesh["v0"], mesh["v1"], mesh["v2"])
    info = scene.info()
    log(f"[bench] kd-tree: {info['n_nodes']} nodes, {info['n_leaf_refs']} refs, "
        f"build {info['build_seconds']:.2f}s, device {info['device_bytes'] / 2**20:.1f} MiB")

    o = hip.DeviceArray.from_numpy(np.concatenate([oc, oi]))
    d = hip.DeviceArray.from_numpy(np.concatenate([dc, di]))
    n = o.shape[0]
    prim = hip.DeviceArray(n, np.int32)
    t = hip.DeviceArray(n, np.float32)
    stream = None  # the null stream: kernels and events below are on the same stream

    # exact traversal counts of this ray set (identical to the oracle's traversal)
    cnt = hip.DeviceArray(2 * n, np.uint32)
    scene.closest_hit(o, d, prim, t, counters=cnt, stream=stream)
    c = cnt.numpy().astype(np.int64)
    nodes_sum, tris_sum = int(c[:n].sum()), int(c[n:].sum())
    del cnt
    alg_bytes = 32 * n + 8 * nodes_sum + 40 * tris_sum
    hit_rate = float((prim.numpy() >= 0).mean())
    # work the kernel executes: culling skips leaf clusters that cannot hold a hit
    wk = hip.DeviceArray.zeros(4, np.uint64)
    scene.closest_hit_work(o, d, prim, wk, stream=stream)
    w_nodes, w_ref, w_tris, w_boxes = (int(x) for x in wk.numpy())
    assert w_nodes == nodes_sum and w_ref == tris_sum, "counting launches disagree"
    exec_bytes = 32 * n + 8 * nodes_sum + 40 * w_tris + 32 * w_boxes

    if args.copy_calibration:
        import ctypes
        src = hip.DeviceArray(64 << 20, np.float32)
        dst = hip.DeviceArray(64 << 20, np.float32)
        src.zero()
        hip.synchronize()
        hip.rt().hipMemcpy(dst.ptr, src.ptr, ctypes.c_size_t(src.nbytes), hip.D2D)
        hip.synchronize()
        del src, dst

    def one(oo, dd, pp, tt):
        scene.closest_hit(oo, dd, pp, tt, stream=stream)

    for _ in range(args.warmup):
        one(o, d, prim, t)
    hip.synchronize()
    starts = [hip.Event() for _ in range(args.steps)]
    ends = [hip.Event() for _ in range(args.steps)]
    t_wall0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        one(o, d, prim, t)
        ends[k].record(stream)
    hip.synchronize()
    wall = time.perf_counter() - t_wall0
    ms = [s.elapsed_ms(e) for s, e in zip(starts, ends)]
    ms_avg = float(np.mean(ms))
    mrays = n / (ms_avg * 1e-3) / 1e6
    log(f"[bench] {args.steps} steps: {ms_avg:.3f} ms/launch (min {min(ms):.3f}), "
        f"wall {wall * 1e3 / args.steps:.3f} ms/step, {mrays:.1f} Mrays/s")

    # coherent / incoherent split (same kernel, separate launches)
    split = {}
    for name, (a, b) in (("coherent", (0, n_c)), ("incoherent", (n_c, n))):
        oo, dd, pp, tt = o.rows(a, b), d.rows(a, b), prim.rows(a, b), t.rows(a, b)
        one(oo, dd, pp, tt)
        e0, e1 = hip.Event(), hip.Event()
        reps = max(3, args.steps // 2)
        e0.record(stream)
        for _ in range(reps):
            one(oo, dd, pp, tt)
        e1.record(stream)
        ms_s = e0.elapsed_ms(e1) / reps
        nn = b - a
        split[name] = dict(mrays_per_s=round(nn / (ms_s * 1e-3) / 1e6, 1), ms=round(ms_s, 3),
                           nodes_per_ray=round(float(c[a:b].mean()), 2),
                           tris_per_ray=round(float(c[n + a:n + b].mean()), 2))
    log(f"[bench] split: {split}")

    roofline = closest_hit_roofline(n, ms_avg, alg_bytes, exec_bytes, w_tris, w_boxes,
                                    int(info["device_bytes"]))

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        stride = args.cpu_sample_stride
        so = np.concatenate([oc[::stride], oi[::stride]])
        sd = np.concatenate([dc[::stride], di[::stride]])
        tb = time.time()
        osc = oracle.OracleScene(mesh["v0"], mesh["v1"], mesh["v2"])
        build_s = time.time() - tb
        threads = args.cpu_threads or oracle.max_threads()
        r = osc.closest_hit(so, sd, threads=threads)
        secs = r["seconds"]
        # bit-exact spot check of the same rays on the GPU
        gp = hip.DeviceArray(len(so), np.int32)
        scene.closest_hit(hip.DeviceArray.from_numpy(so), hip.DeviceArray.from_numpy(sd), gp,
                          stream=stream)
        same = bool(np.array_equal(gp.numpy(), r["prim"]))
        s1 = args.cpu_single_stride
        r1 = osc.closest_hit(np.concatenate([oc[::s1], oi[::s1]]),
                             np.concatenate([dc[::s1], di[::s1]]), threads=1)
        n1 = len(r1["prim"])
        cpu = dict(value=round(len(so) / secs / 1e6, 3), unit="Mrays/s", cores=threads,
                   kind="port",
                   sample=f"every {stride}th ray of both batches ({len(so)} rays) on the full "
                          f"mesh; oracle kd build {build_s:.1f}s excluded; "
                          f"GPU ids identical on the sample: {same}",
                   single_thread=dict(value=round(n1 / r1["seconds"] / 1e6, 3), cores=1,
                                      sample=f"every {s1}th ray of both batches ({n1} rays)"))
        log(f"[bench] cpu baseline: {cpu}")

    kdb = None if args.skip_kd_build else bench_kd_build(capi, mesh, info)
    mqo_leg = None if args.skip_mqo else bench_mqo(capi, scenes)
    bdpt = None if args.skip_bdpt else bench_bdpt(args, hip, capi, scenes)
    idpt = None if args.skip_idpt else bench_idpt(args, hip, capi, scenes)
    pt = None if args.skip_pt else bench_pt(args, hip, capi, scenes)
    adj = None if args.skip_adjoint else bench_adjoint(args, hip, capi, scenes)
    bat = None if args.skip_batch else bench_batch(args, hip, capi, scenes)

    out = {
        "metric": "Mrays/s closest-hit",
        "value": round(mrays, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_avg, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "kd-tree closest-hit: displaced icosphere L8 + 64 boxes, "
                        "4096x2048 coherent + 8M incoherent rays",
            "n_tris": int(info["n_tris"]), "n_rays": int(n), "kd_nodes": int(info["n_nodes"]),
            "kd_leaf_refs": int(info["n_leaf_refs"]), "hit_rate": round(hit_rate, 4),
            "nodes_per_ray": round(nodes_sum / n, 2), "tris_per_ray": round(tris_sum / n, 2),
            "parallelism": "single GPU",
        },
        "split": split,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pt_1080p": pt,
        "adjoint": adj,
        "batched": bat,
        "kd_build": kdb,
        "mqo_load": mqo_leg,
        "bdpt_cornell": bdpt,
        "idpt_caustic": idpt,
    }
    print(json.dumps(out), flush=True)


def bench_kd_build(capi, mesh, info):
    """SAH kd-tree build of the workload mesh on the device vs the host builder
    (the scene above was built on the host); the trees must be identical."""
    host = capi.KDTree(mesh["v0"], mesh["v1"], mesh["v2"])
    hn, hp = host.export()
    host_s = host.info()["build_seconds"]
    del host
    secs = []
    same = True
    for _ in range(3):  # the first call also loads the build kernels
        dev = capi.KDTree(mesh["v0"], mesh["v1"], mesh["v2"], device=True)
        secs.append(dev.info()["build_seconds"])
        dn, dp = dev.export()
        leaf = (hn[:, 0] & 3) == 3
        same = same and hn.shape == dn.shape and bool(
            np.array_equal(hn[:, 0], dn[:, 0]) and np.array_equal(hn[leaf, 1], dn[leaf, 1])
            and np.array_equal(hn[~leaf, 1].view(np.float32), dn[~leaf, 1].view(np.float32))
            and np.array_equal(hp, dp))
        del dev
    dev_s = min(secs[1:])
    r = dict(n_tris=int(info["n_tris"]), host_seconds=round(host_s, 4),
             device_seconds=round(dev_s, 4), first_device_call_seconds=round(secs[0], 4),
             speedup=round(host_s / dev_s, 2), identical_trees=same,
             note="host arrays in, host tree out; device time includes upload, root box and "
                  "pre-order layout on the host")
    log(f"[bench] kd build: {r}")
    return r


def mqo_document(v0, v1, v2):
    """Metasequoia text of a triangle soup (centimetres, one Lambert material)."""
    import io
    n = len(v0)
    verts = np.stack([v0, v1, v2], 1).reshape(-1, 3) * 100.0
    buf = io.StringIO()
    buf.write("Metasequoia Document\nFormat Text Ver 1.0\n\nMaterial 1 {\n"
              '\t"mcc-white" shader(3) col(1.000 1.000 1.000 1.000) dif(0.800) amb(0.600) '
              "emi(0.000) spc(0.000) power(5.00)\n}\n")
    buf.write(f'Object "mesh" {{\n\tvisible 15\n\tshading 0\n\tvertex {3 * n} {{\n')
    np.savetxt(buf, verts, fmt="\t\t%.4f %.4f %.4f")
    buf.write(f"\t}}\n\tface {n} {{\n")
    idx = np.arange(3 * n, dtype=np.int64).reshape(-1, 3)
    np.savetxt(buf, idx, fmt="\t\t3 V(%d %d %d) M(0)")
    buf.write("\t}\n}\nEof\n")
    return buf.getvalue()


def bench_mqo(capi, scenes):
    """Metasequoia text -> device scene: parse (xyz_mqo_load) and scene creation
    with the host and the device kd-tree builders, on a displaced icosphere."""
    from xyz_amd import mqo
    m = scenes.closest_hit_mesh(6, 64, seed=3)
    text = mqo_document(m["v0"], m["v1"], m["v2"])
    t0 = time.perf_counter()
    doc = mqo.loads(text)
    t_parse = time.perf_counter() - t0
    out = dict(n_tris=doc.n_tris, megabytes=round(len(text) / 1e6, 2),
               parse_seconds=round(t_parse, 4),
               parse_mb_per_s=round(len(text) / 1e6 / t_parse, 1))
    for name, builder in (("host", capi.KD_HOST), ("device", capi.KD_DEVICE)):
        doc.scene(kd_builder=builder)  # warm-up (first call loads kernels)
        t0 = time.perf_counter()
        sc = doc.scene(kd_builder=builder)
        out[f"scene_create_seconds_{name}_kd"] = round(time.perf_counter() - t0, 4)
        out[f"kd_build_seconds_{name}"] = round(sc.info()["build_seconds"], 4)
        del sc
    log(f"[bench] mqo load: {out}")
    return out


def bench_bdpt(args, hip, capi, scenes):
    """Bidirectional path tracing of the Cornell box at 256x256, depth 8, 64 spp
    (xyz_render_bdpt) vs the oracle's BDPT on the host cores (16 spp) and on one
    thread (2 spp); the own films of the CPU sample are checked bit-identical."""
    sc = scenes.cornell_box()
    g = capi.Scene.from_dict(sc)
    W = H = 256
    D, spp = 8, 64
    warm = hip.DeviceArray.zeros((H, W, 3), np.float32)
    g.render_bdpt(sc["camera"], warm, 1, D, seed=1)
    film = hip.DeviceArray.zeros((H, W, 3), np.float32)
    st = g.render_bdpt(sc["camera"], film, spp, D, seed=1)
    ms = st["seconds"] * 1e3
    gpu = W * H * spp / st["seconds"] / 1e6
    out = dict(msamples_per_s=round(gpu, 2), ms=round(ms, 2), spp=spp, width=W, height=H,
               max_depth=D, queries_per_sample=round(st["queries"] / (W * H * spp), 2))
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        o = orc.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc["tri_bsdf"], bsdfs=sc["bsdfs"])
        threads = args.cpu_threads or orc.max_threads()
        own, spl, cst = o.render_bdpt(sc["camera"], W, H, 16, D, seed=1, threads=threads, stats=True)
        gown = hip.DeviceArray.zeros((H, W, 3), np.float32)
        gspl = hip.DeviceArray.zeros((H, W, 3), np.float32)
        g.render_bdpt(sc["camera"], gown, 16, D, seed=1, splat=gspl)
        same = bool(np.array_equal(gown.numpy().view(np.uint32), own.view(np.uint32)))
        _, _, c1 = o.render_bdpt(sc["camera"], W, H, 2, D, seed=1, threads=1, stats=True)
        cpu = W * H * 16 / cst["seconds"] / 1e6
        out["cpu_baseline"] = dict(value=round(cpu, 3), unit="Msamples/s", cores=threads,
                                   kind="port", sample=f"16 spp of the same frame; own film "
                                   f"identical to the GPU's: {same}",
                                   single_thread=dict(value=round(W * H * 2 / c1["seconds"] / 1e6, 4),
                                                      cores=1, sample="2 spp"))
        out["speedup_vs_cpu"] = round(gpu / cpu, 1)
    out["production"] = bench_bdpt_production(hip, capi, scenes)
    log(f"[bench] bdpt: {out}")
    return out


def bench_bdpt_production(hip, capi, scenes):
    """BDPT on the ~2M-triangle production scene at 480x270, depth 8, 8 spp:
    the queued form of the connections (which xyz_render_bdpt takes for a scene
    this size) against each thread tracing its own queries; the own films of
    the two forms are compared bit for bit."""
    sc = scenes.production_scene()
    g = capi.Scene.from_dict(sc)
    W, H, D, spp = 480, 270, 8, 8
    res = {}
    films = {}
    for name, engine in (("queued", 2), ("per_thread", 1)):
        warm = hip.DeviceArray.zeros((H, W, 3), np.float32)
        g.render_bdpt(sc["camera"], warm, 1, D, seed=1, engine=engine)
        film = hip.DeviceArray.zeros((H, W, 3), np.float32)
        splat = hip.DeviceArray.zeros((H, W, 3), np.float32)
        st = g.render_bdpt(sc["camera"], film, spp, D, seed=1, splat=splat, engine=engine)
        res[name] = round(W * H * spp / st["seconds"] / 1e6, 2)
        films[name] = film.numpy()
    out = dict(msamples_per_s=res["queued"], per_thread_msamples_per_s=res["per_thread"],
               n_tris=int(g.info()["n_tris"]), width=W, height=H, spp=spp, max_depth=D,
               default_engine="queued" if g.info()["n_leaf_refs"] >= 300000 else "per_thread",
               own_films_identical=bool(np.array_equal(films["queued"].view(np.uint32),
                                                       films["per_thread"].view(np.uint32))))
    return out


def bench_idpt(args, hip, capi, scenes):
    """Importance-driven path tracing of the caustic scene at 1920x1080 with the
    reference's settings (depth 16, Russian roulette at every vertex): the
    importon map of 100,000 light paths (device vs the oracle on one thread),
    then 16 spp of xyz_render_idpt against xyz_render on the same frame, and the
    oracle on a band of rows at 1 spp (checked bit-identical)."""
    sc = scenes.caustic_scene(sphere_level=4)
    g = capi.Scene.from_dict(sc)
    W, H, D, spp, n_emit = 1920, 1080, 15, 16, 100000
    g.importon_map(1000, max_depth=D, seed=1)  # warm-up
    builds = [g.importon_map(n_emit, max_depth=D, seed=1) for _ in range(3)]
    m = builds[-1]
    info = m.info()
    build_s = min(b.info()["seconds"] for b in builds)
    film = hip.DeviceArray.zeros((H, W, 3), np.float32)
    g.render_idpt(m, sc["camera"], film, 1, D, seed=2, sample_offset=10**6)  # warm-up
    film.zero()
    st = g.render_idpt(m, sc["camera"], film, spp, D, seed=2)
    pfilm = hip.DeviceArray.zeros((H, W, 3), np.float32)
    pst = g.render(sc["camera"], pfilm, spp, D, seed=2)
    samples = W * H * spp
    rate = samples / st["seconds"] / 1e6
    prate = samples / pst["seconds"] / 1e6
    out = dict(msamples_per_s=round(rate, 2), ms=round(st["seconds"] * 1e3, 2), spp=spp,
               width=W, height=H, max_depth=D, n_tris=int(g.info()["n_tris"]),
               queries_per_sample=round(st["queries"] / samples, 3),
               pt_same_frame_msamples_per_s=round(prate, 2),
               guided_cost=round(prate / rate, 3),
               importons=int(info["n_importons"]), n_emit=n_emit,
               map_build_ms=round(build_s * 1e3, 2), grid=info["grid"])
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        o = orc.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc["tri_bsdf"], bsdfs=sc["bsdfs"],
                            normals=sc["normals"])
        tb = time.time()
        om = o.importon_map(n_emit, max_depth=D, seed=1)
        cpu_build = time.time() - tb
        same_map = bool(np.array_equal(om.export().view(np.uint32), m.export().view(np.uint32)))
        threads = args.cpu_threads or orc.max_threads()
        y0 = H // 2 - 64
        rows = (y0, y0 + 128)
        ofilm, ost = o.render_idpt(om, sc["camera"], W, H, 2, D, seed=2, rows=rows,
                                   threads=threads, stats=True)
        gfilm = hip.DeviceArray.zeros((H, W, 3), np.float32)
        g.render_idpt(m, sc["camera"], gfilm, 2, D, seed=2)
        same = bool(np.array_equal(gfilm.numpy()[rows[0]:rows[1]].view(np.uint32),
                                   ofilm[rows[0]:rows[1]].view(np.uint32)))
        rows1 = (y0, y0 + 8)
        _, ost1 = o.render_idpt(om, sc["camera"], W, H, 1, D, seed=2, rows=rows1, threads=1,
                                stats=True)
        n = W * (rows[1] - rows[0]) * 2
        cpu = n / ost["seconds"] / 1e6
        out["cpu_baseline"] = dict(
            value=round(cpu, 4), unit="Msamples/s", cores=threads, kind="port",
            sample=f"rows {rows[0]}-{rows[1] - 1} of the same frame at 2 spp ({n} samples); "
                   f"GPU film bit-identical on those rows: {same}",
            single_thread=dict(value=round(W * 8 / ost1["seconds"] / 1e6, 4), cores=1,
                               sample=f"rows {rows1[0]}-{rows1[1] - 1} at 1 spp"),
            map_build_s=round(cpu_build, 3), map_build_cores=1, map_identical=same_map)
        out["vs_cpu"] = round(rate / cpu, 1)
    log(f"[bench] idpt: {out}")
    return out


def closest_hit_roofline(n, ms_avg, alg_bytes, exec_bytes, w_tris, w_boxes, device_bytes):
    """The closest-hit kernel's ceilings.  It is bound by VALU issue (divergent
    wave-cooperative traversal), not by HBM: `achieved` is its VALU wave
    instructions per launch (rocprofv3 SQ_INSTS_VALU, profiles/
    closest_hit_pmc.json) over the launch time measured here, against 1024
    SIMDs x 2.4 GHz / 4 cycles.  The HBM fraction (measured FETCH_SIZE +
    WRITE_SIZE bytes over the same time), the L2 hit rate, and the byte counts
    of the reference's work and of the work executed sit beside it."""
    s = ms_avg * 1e-3
    pmc_path = os.path.join(ROOT, "profiles", "closest_hit_pmc.json")
    pmc = {}
    if os.path.exists(pmc_path):
        with open(pmc_path) as fh:
            pmc = json.load(fh)
    valu = pmc.get("sq_insts_valu_per_launch")
    traffic = pmc.get("hbm_bytes_per_launch")
    r = dict(bound="valu", achieved=None, peak=VALU_PEAK_GINSTR, unit="G wave-instr/s", frac=None,
             traffic=traffic, source="profiles/closest_hit_pmc.json")
    if valu:
        a = valu / s / 1e9
        r.update(achieved=round(a, 1), frac=round(a / VALU_PEAK_GINSTR, 4),
                 valu_instr_per_ray=round(valu / n, 1),
                 lanes_active_per_valu_instr=pmc.get("lanes_active_per_valu_instruction"),
                 valu_busy=pmc.get("valu_busy"))
    if traffic:
        r["hbm"] = dict(achieved=round(traffic / s / 1e9, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(traffic / s / 1e9 / HBM_PEAK_GBS, 4))
    if "l2_hit_rate" in pmc:
        r["l2_hit_rate"] = pmc["l2_hit_rate"]
    # the reference's work in bytes (32 per ray, 8 per node, 40 per triangle
    # test): culling executes about a third of those tests, so this rate is
    # not a bound
    r["reference_work"] = dict(bytes_per_launch=int(alg_bytes),
                               gb_per_s=round(alg_bytes / s / 1e9, 1))
    r["executed"] = dict(bytes_per_launch=int(exec_bytes),
                         gb_per_s=round(exec_bytes / s / 1e9, 1),
                         tris_tested_per_ray=round(w_tris / n, 2),
                         boxes_tested_per_ray=round(w_boxes / n, 2))
    # compulsory-only floor: the rays in and the results out (32 B per ray) plus
    # one pass over the scene's device data
    floor_bytes = 32 * n + device_bytes
    r["floor"] = dict(bytes_per_launch=floor_bytes,
                      ms_at_peak=round(floor_bytes / (HBM_PEAK_GBS * 1e9) * 1e3, 4),
                      frac_of_launch=round(floor_bytes / (HBM_PEAK_GBS * 1e9) / s, 4))
    return r


def bench_pt(args, hip, capi, scenes):
    """Path tracing at 1920x1080 on the ~2M-triangle production scene: max depth
    16, Russian roulette after depth 4.  Msamples/s over `--pt-spp` samples per
    pixel of the 256-spp workload; the CPU oracle renders a band of rows at 1 spp."""
    W, H, D, R = 1920, 1080, 16, 4
    sc = scenes.production_scene()
    g = capi.Scene.from_dict(sc)
    info = g.info()
    film = hip.DeviceArray.zeros((H, W, 3), np.float32)
    g.render(sc["camera"], film, 1, D, rr_start=R, seed=1, sample_offset=10**6)  # warm-up
    film.zero()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(None)
    st = g.render(sc["camera"], film, args.pt_spp, D, rr_start=R, seed=1)
    e1.record(None)
    ms = e0.elapsed_ms(e1)
    samples = W * H * args.pt_spp
    rate = samples / (ms * 1e-3) / 1e6
    res = dict(msamples_per_s=round(rate, 2), ms=round(ms, 1), spp=args.pt_spp,
               frame_s=round(ms * 1e-3, 3),
               queries_per_sample=round(st["queries"] / samples, 3),
               mrays_per_s=round(st["queries"] / (ms * 1e-3) / 1e6, 1),
               n_tris=int(info["n_tris"]), config="1920x1080, depth 16, RR after 4")
    res["roofline"] = pt_roofline(rate)
    log(f"[bench] pt 1080p: {res}")
    if not args.no_cpu_baseline:
        from oracle import oracle
        tb = time.time()
        o = oracle.OracleScene(sc["v0"], sc["v1"], sc["v2"], tri_bsdf=sc["tri_bsdf"],
                               bsdfs=sc["bsdfs"])
        build_s = time.time() - tb
        y0 = H // 2
        rows = (y0, y0 + args.pt_cpu_rows)
        threads = args.cpu_threads or oracle.max_threads()
        ofilm, ost = o.render(sc["camera"], W, H, 1, D, rr_start=R, seed=1, rows=rows,
                              threads=threads, stats=True)
        gfilm = hip.DeviceArray.zeros((H, W, 3), np.float32)
        g.render(sc["camera"], gfilm, 1, D, rr_start=R, seed=1)
        same = bool(np.array_equal(gfilm.numpy()[rows[0]:rows[1]].view(np.uint32),
                                   ofilm[rows[0]:rows[1]].view(np.uint32)))
        n = W * args.pt_cpu_rows
        rows1 = (y0, y0 + args.pt_cpu_single_rows)
        _, ost1 = o.render(sc["camera"], W, H, 1, D, rr_start=R, seed=1, rows=rows1,
                           threads=1, stats=True)
        n1 = W * args.pt_cpu_single_rows
        res["cpu_baseline"] = dict(
            value=round(n / ost["seconds"] / 1e6, 4), unit="Msamples/s", cores=threads,
            kind="port", sample=f"rows {rows[0]}-{rows[1] - 1} of the same frame at 1 spp "
                                f"({n} samples); oracle kd build {build_s:.1f}s excluded; "
                                f"GPU film bit-identical on those rows: {same}",
            single_thread=dict(value=round(n1 / ost1["seconds"] / 1e6, 4), cores=1,
                               sample=f"rows {rows1[0]}-{rows1[1] - 1} at 1 spp ({n1} samples)"))
        res["vs_cpu"] = round(rate / res["cpu_baseline"]["value"], 1)
        # the reference's work per sample in bytes, counted by the oracle on its
        # band: 32 per query, 8 per node, 40 per triangle test, 64 per shaded
        # vertex, the film's 12 bytes per pixel over the spp
        bs = (32 * ost["queries"] + 8 * ost["nodes"] + 40 * ost["tris"] + 64 * ost["vertices"]) / n
        bs += 12.0 / args.pt_spp
        res["roofline"]["reference_work"] = dict(
            bytes_per_sample=round(bs, 1), gb_per_s=round(bs * rate * 1e6 / 1e9, 1),
            nodes_per_query=round(ost["nodes"] / ost["queries"], 1),
            tris_per_query=round(ost["tris"] / ost["queries"], 1),
            vertices_per_sample=round(ost["vertices"] / n, 3),
            sample=f"oracle counts on rows {rows[0]}-{rows[1] - 1} at 1 spp")
        log(f"[bench] pt cpu baseline: {res['cpu_baseline']}")
    return res


def pt_roofline(rate_msamples):
    """The path tracer's kernels (extend, shade, shadow) at the frame rate just
    measured: HBM bytes per sample from rocprofv3 FETCH_SIZE + WRITE_SIZE on the
    same frame (profiles/pt_1080p_pmc.json) give the HBM fraction; each
    kernel's VALU busy, lanes per instruction and L2 hit rate come from the same
    file.  None of the three kernels saturates HBM or VALU issue: the traversal
    kernels stall on memory latency (L2 hit rate 0.57-0.74)."""
    path = os.path.join(ROOT, "profiles", "pt_1080p_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        pmc = json.load(fh)
    bps = pmc["hbm_bytes_per_sample"]
    a = bps * rate_msamples * 1e6 / 1e9
    return dict(bound="hbm", achieved=round(a, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(a / HBM_PEAK_GBS, 4), traffic_bytes_per_sample=bps,
                kernels={k: dict(valu_busy=v["valu_busy"],
                                 lanes_active_per_valu_instr=v["lanes_active_per_valu_instruction"],
                                 l2_hit_rate=v["l2_hit_rate"])
                         for k, v in pmc["kernels"].items()},
                source="profiles/pt_1080p_pmc.json")


def bench_adjoint(args, hip, capi, scenes):
    """Cornell box with 64x64 RGB albedo textures on the walls, 512x512, 64 spp,
    depth 6: time of the path-replay backward pass over the primal pass."""
    W = H = 512
    spp, D, R = 64, 6, 3
    sc = scenes.cornell_box(textured=True, tex_res=64)
    g = capi.Scene.from_dict(sc)
    film = hip.DeviceArray.zeros((H, W, 3), np.float32)
    w = hip.DeviceArray.from_numpy(np.random.default_rng(0).uniform(-1, 1, (H, W, 3)).astype(np.float32))
    dt = hip.DeviceArray.zeros(g.n_texels, np.float32)
    de = hip.DeviceArray.zeros(g.n_bsdfs * 3, np.float32)
    g.render(sc["camera"], film, 4, D, rr_start=R, seed=2)  # warm-up
    g.render_backward(sc["camera"], w, dt, de, 4, D, rr_start=R, seed=2)
    e0, e1, e2 = hip.Event(), hip.Event(), hip.Event()
    e0.record(None)
    g.render(sc["camera"], film, spp, D, rr_start=R, seed=3)
    e1.record(None)
    g.render_backward(sc["camera"], w, dt, de, spp, D, rr_start=R, seed=3)
    e2.record(None)
    fwd = e0.elapsed_ms(e1)
    bwd = e1.elapsed_ms(e2)
    res = dict(primal_ms=round(fwd, 2), backward_ms=round(bwd, 2),
               adjoint_over_primal=round(bwd / fwd, 3),
               config="textured Cornell box (61,440 texels), 512x512, 64 spp, depth 6")
    log(f"[bench] adjoint: {res}")
    return res


def bench_batch(args, hip, capi, scenes):
    """64 generated scenes (about 1e3..5e5 triangles, 128^2..1024^2, 16..256 spp,
    depth 8): xyz_render_batch with the autotuner's choice (the shared-wave
    engine, replayed from a captured hipGraph; the first call captures, the
    timed call is one graph launch) against the same renders one after another
    through xyz_render.  `ideal_batched_over_sequential` is the ratio if every
    job ran at the rate of the job with the most samples.
    `bound_smaller_jobs_free` is the ratio if the jobs of at least 2^26 samples
    kept their sequential time and every other job cost nothing: each of those
    jobs fills the GPU on its own (batched together they gain 1 %), so no
    batching of this mix can exceed it."""
    D = 8
    jobs = [scenes.batch_scene(s) for s in range(64)]
    gs = [capi.Scene.from_dict(j) for j in jobs]
    films = [hip.DeviceArray.zeros((j["height"], j["width"], 3), np.float32) for j in jobs]
    cams = [j["camera"] for j in jobs]
    spps = [j["spp"] for j in jobs]
    seeds = list(range(64))
    samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    tb = time.time()
    best = capi.batch_autotune(gs, cams, films, spps, D, rr_start=3, seeds=seeds)
    tune_s = time.time() - tb
    # sequential plain renders
    for f in films:
        f.zero()
    e0, e1 = hip.Event(), hip.Event()
    ev = [hip.Event() for _ in range(len(gs) + 1)]
    e0.record(None)
    ev[0].record(None)
    for k, (s, c, f, sp, sd) in enumerate(zip(gs, cams, films, spps, seeds)):
        s.render(c, f, sp, D, rr_start=3, seed=sd)
        ev[k + 1].record(None)
    e1.record(None)
    seq_ms = e0.elapsed_ms(e1)
    job_ms = [ev[k].elapsed_ms(ev[k + 1]) for k in range(len(gs))]
    job_samples = [j["width"] * j["height"] * j["spp"] for j in jobs]
    big = int(np.argmax(job_samples))
    big_rate = job_samples[big] / (job_ms[big] * 1e-3)
    large = [k for k in range(len(gs)) if job_samples[k] >= (1 << 26)]
    large_ms = sum(job_ms[k] for k in large)
    seq_films = [f.numpy() for f in films]
    # the same sequence again: later passes must not be slower than the first
    # (workspace reuse across calls)
    e0.record(None)
    for s, c, f, sp, sd in zip(gs, cams, films, spps, seeds):
        s.render(c, f, sp, D, rr_start=3, seed=sd)
    e1.record(None)
    seq2_ms = e0.elapsed_ms(e1)
    # batched: one graph launch (first call captures; the timed call replays)
    for f in films:
        f.zero()
    capi.render_batch(gs, cams, films, spps, D, rr_start=3, seeds=seeds, tuning=best)
    same = all(np.array_equal(f.numpy().view(np.uint32), s.view(np.uint32))
               for f, s in zip(films, seq_films))
    for f in films:
        f.zero()
    e0.record(None)
    capi.render_batch(gs, cams, films, spps, D, rr_start=3, seeds=seeds, tuning=best)
    e1.record(None)
    bat_ms = e0.elapsed_ms(e1)
    res = dict(batched_ms=round(bat_ms, 1), sequential_ms=round(seq_ms, 1),
               sequential_second_pass_ms=round(seq2_ms, 1),
               batched_over_sequential=round(seq_ms / bat_ms, 3),
               batched_msamples_per_s=round(samples / (bat_ms * 1e-3) / 1e6, 1),
               sequential_msamples_per_s=round(samples / (seq_ms * 1e-3) / 1e6, 1),
               samples=int(samples), films_bit_identical=bool(same),
               ideal_batched_over_sequential=round(seq_ms / (samples / big_rate * 1e3), 3),
               largest_job=dict(samples=int(job_samples[big]), ms=round(job_ms[big], 1),
                                msamples_per_s=round(big_rate / 1e6, 1)),
               smallest_job_msamples_per_s=round(
                   min(a / (b * 1e-3) for a, b in zip(job_samples, job_ms)) / 1e6, 1),
               bound_smaller_jobs_free=round(seq_ms / large_ms, 3),
               large_jobs=dict(count=len(large),
                               sample_share=round(sum(job_samples[k] for k in large) / samples, 3),
                               sequential_ms=round(large_ms, 1)),
               tuning=dict(engine=["persistent megakernel", "shared wavefront"][best.engine],
                           tile_pixels=best.tile_pixels, blocks_per_cu=best.blocks_per_cu,
                           block_threads=best.block_threads, use_graph=best.use_graph),
               autotune_s=round(tune_s, 1))
    log(f"[bench] batched: {res}")
    return res


if __name__ == "__main__":
    main()
