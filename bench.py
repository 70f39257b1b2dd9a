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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-stride", type=int, default=16,
                    help="the CPU baseline traces every n-th ray of each batch")
    ap.add_argument("--cpu-threads", type=int, default=0)
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

    achieved = alg_bytes / (ms_avg * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None)
    pmc = os.path.join(ROOT, "profiles", "closest_hit_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            roofline["traffic"] = json.load(fh).get("hbm_bytes_per_launch")

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
        cpu = dict(value=round(len(so) / secs / 1e6, 3), unit="Mrays/s", cores=threads,
                   kind="port",
                   sample=f"every {stride}th ray of both batches ({len(so)} rays) on the full "
                          f"mesh; oracle kd build {build_s:.1f}s excluded; "
                          f"GPU ids identical on the sample: {same}")
        log(f"[bench] cpu baseline: {cpu}")

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
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
