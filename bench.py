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
    import torch
    from xyz_amd import build as xbuild
    from xyz_amd import capi, scenes
    if not os.path.exists(xbuild.LIB):
        xbuild.build()
    dev = torch.device("cuda:0")

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

    o = torch.from_numpy(np.concatenate([oc, oi])).to(dev)
    d = torch.from_numpy(np.concatenate([dc, di])).to(dev)
    n = o.shape[0]
    prim = torch.empty(n, dtype=torch.int32, device=dev)
    t = torch.empty(n, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    # exact traversal counts of this ray set (identical to the oracle's traversal)
    cnt = torch.empty(2 * n, dtype=torch.int32, device=dev)
    scene.closest_hit(o, d, prim, t, counters=cnt, stream=stream)
    torch.cuda.synchronize()
    c = cnt.cpu().numpy().view(np.uint32).astype(np.int64)
    nodes_sum, tris_sum = int(c[:n].sum()), int(c[n:].sum())
    del cnt
    alg_bytes = 32 * n + 8 * nodes_sum + 40 * tris_sum
    hit_rate = float((prim.cpu().numpy() >= 0).mean())

    def one(oo, dd, pp, tt):
        scene.closest_hit(oo, dd, pp, tt, stream=stream)

    for _ in range(args.warmup):
        one(o, d, prim, t)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t_wall0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        one(o, d, prim, t)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall0
    ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms_avg = float(np.mean(ms))
    mrays = n / (ms_avg * 1e-3) / 1e6
    log(f"[bench] {args.steps} steps: {ms_avg:.3f} ms/launch (min {min(ms):.3f}), "
        f"wall {wall * 1e3 / args.steps:.3f} ms/step, {mrays:.1f} Mrays/s")

    # coherent / incoherent split (same kernel, separate launches)
    split = {}
    for name, sl in (("coherent", slice(0, n_c)), ("incoherent", slice(n_c, n))):
        oo, dd = o[sl], d[sl]
        pp, tt = prim[sl], t[sl]
        one(oo, dd, pp, tt)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        reps = max(3, args.steps // 2)
        e0.record(stream)
        for _ in range(reps):
            one(oo, dd, pp, tt)
        e1.record(stream)
        torch.cuda.synchronize()
        ms_s = e0.elapsed_time(e1) / reps
        nn = sl.stop - sl.start
        cs = c[sl]
        ci = c[n + sl.start:n + sl.stop]
        split[name] = dict(mrays_per_s=round(nn / (ms_s * 1e-3) / 1e6, 1), ms=round(ms_s, 3),
                           nodes_per_ray=round(float(cs.mean()), 2),
                           tris_per_ray=round(float(ci.mean()), 2))
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
        gp = torch.empty(len(so), dtype=torch.int32, device=dev)
        scene.closest_hit(torch.from_numpy(so).to(dev), torch.from_numpy(sd).to(dev), gp,
                          stream=stream)
        torch.cuda.synchronize()
        same = bool(np.array_equal(gp.cpu().numpy(), r["prim"]))
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
