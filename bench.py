#!/usr/bin/env python3
"""Benchmark of the raw-development path on one MI355X.

Default workload: the plain per-image path on a 24 MP Bayer still
(6000 x 4000 RGGB, 14-bit in uint16, black 512, white 16383, seeded scene of
smooth gradients + 200 rectangles over 12 stops, 2% clipped, Poisson-Gaussian
noise a=2e-4 b=1e-6) through the graph i-raw -> denoise -> demosaic -> colour
-> filmcurv -> llap -> o-pfm, run by the host runtime (module DSOs, one HIP
launch per node).  The DNG is decoded and uploaded once before timing; a
step is one pass of every node over the resident raw (no host transfer).

Prints one JSON line: value is Mpix/s
(CFA pixels after i-raw block rounding per second), roofline is for the
node kernel with the largest share of the step, measured with HIP events
on the graph's stream inside the timed region, cpu_baseline times the
oracle on the host on a bounded sample (the 2048 x 1536 case).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpix/s raw→rec2020, batched-served vs plain per-image; p99 latency @1e-3 err"
HBM_PEAK_GBS = 8000.0


def node_bytes(n):
    """algorithmic HBM bytes of one node launch: each distinct bound buffer
    read once, each output written once; dead stores, unused bindings and
    dummies aliasing another binding count zero.  llap/assemble touches 3
    planes of the fine level (v and the two selected curves), the coarse
    output and 2 planes of the coarse level."""
    if n["module"] == "llap" and n["kernel"] == "assemble":
        fine = n["conns"][3]["buf"]
        lo = n["conns"][2]["buf"]
        pf, pc = fine.wd * fine.ht * 2, lo.wd * lo.ht * 2
        return 4 * pf + 3 * pc
    seen, total = set(), 0
    for c in n["conns"]:
        b = c["buf"]
        if not b.ptr or b.ptr in seen:
            continue
        seen.add(b.ptr)
        total += b.plane_stride * max(1, b.array_length)
    return total


def measured_traffic(tag, node):
    """HBM bytes per launch of `node` from the committed FETCH_SIZE/WRITE_SIZE
    summary of this workload (counters scaled by a copy kernel of known size)"""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{tag}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == node:
            return int(f[2])
    return None


def cpu_baseline(seconds_hint=True):
    """the oracle (FP64 scalar C, one thread) on the 2048 x 1536 Bayer case"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import graphspec as gs
    import oracle_py as op
    from vkdt_amd import synth
    W, H = 2048, 1536
    cfa, _ = synth.mosaic(W, H, 1, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)
    p = gs.DevelopParams(kind="bayer", noise_a=na, noise_b=nb)
    t0 = time.perf_counter()
    out = op.develop(64, cfa, gs.oracle_params(p))
    dt = time.perf_counter() - t0
    assert np.isfinite(out).all()
    return {"value": round(W * H / dt / 1e6, 4), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": f"oracle (FP64 scalar C restatement, 1 thread) on the 2048x1536 Bayer case, {dt:.1f} s"}


def calibrate_copy():
    """a copy kernel of known traffic: 1 GiB read + 1 GiB written per launch"""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--kind", default="bayer", choices=["bayer", "xtrans"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", default="", help="write per-kernel timings as JSON to this path")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB (5 times), the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(images are independent units and would shard across GPUs with no communication)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from vkdt_amd import hip, pipe, synth
    hip.init(0)
    if args.kind == "bayer":
        W, H, workload = 6000, 4000, "plain per-image path, 24 MP Bayer RGGB still 6000x4000"
    else:
        W, H, workload = 6240, 4160, "plain per-image path, X-Trans still 6240x4160 (6240x4158 after block rounding)"
    cfa, _ = synth.mosaic(W, H, 1, args.kind)
    na, nb = synth.noise_params(2e-4, 1e-6)
    tmp = tempfile.mkdtemp(prefix="vkdt_bench_")
    dng = os.path.join(tmp, "still.dng")
    synth.write_dng(dng, cfa, args.kind)
    extra = ["param:llap:01:numgamma:8", "param:llap:01:levels:6"] if args.kind == "xtrans" else []
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb, extra=extra))
    g.export("o-pfm", os.path.join(tmp, "out"))
    g.run(pipe.RUN_ALL & ~pipe.RUN_DOWNLOAD)        # plan, allocate, upload, first pass
    ip = g.img_param("i-raw", "main")
    P = ip["full"][0] * ip["full"][1]
    nodes = {n["id"]: n for n in g.nodes()}
    order = g.exec_nodes()
    run_flags = pipe.RUN_RECORD                      # resident raw, no host traffic
    for _ in range(args.warmup):
        g.run(run_flags)
    hip.sync(g.stream())

    g.set_perf(True)
    per = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.run(run_flags)                             # syncs the stream (per-node events)
        for k, ms in zip(order, g.exec_ms()):
            n = nodes[k]
            if n["module"] in ("i-raw", "o-pfm"):
                continue
            key = n["module"] + "/" + n["kernel"]
            e = per.setdefault(key, {"ms": 0.0, "bytes": 0, "launches": 0})
            e["ms"] += ms
            e["bytes"] += node_bytes(n)
            e["launches"] += 1
    hip.sync(g.stream())
    wall = time.perf_counter() - t0
    g.set_perf(False)
    ms_step = wall * 1e3 / args.steps
    gpu_ms = sum(e["ms"] for e in per.values()) / args.steps
    value = P / (ms_step * 1e-3) / 1e6
    graph_bytes = sum(node_bytes(nodes[k]) for k in order if nodes[k]["module"] not in ("i-raw", "o-pfm"))

    dom = max(per, key=lambda k: per[k]["ms"])
    e = per[dom]
    achieved = e["bytes"] / (e["ms"] * 1e-3) / 1e9      # GB/s over its launches
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(f"plain_24mp_{args.kind}", dom),
            "bytes_per_launch": int(e["bytes"] / e["launches"]),
            "avg_launch_ms": round(e["ms"] / e["launches"], 4)}
    graph_roof = {"algorithmic_bytes": graph_bytes, "bytes_per_px": round(graph_bytes / P, 1),
                  "gpu_ms": round(gpu_ms, 4), "achieved": round(graph_bytes / (gpu_ms * 1e-3) / 1e9, 1),
                  "frac": round(graph_bytes / (gpu_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if args.breakdown:
        with open(args.breakdown, "w") as f:
            json.dump({k: dict(v, ms_per_step=v["ms"] / args.steps,
                               gbs=v["bytes"] / (v["ms"] * 1e-3) / 1e9) for k, v in per.items()}, f, indent=1)
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene + Poisson-Gaussian noise, uncompressed DNG)",
        "config": {"workload": workload, "width": W, "height": H, "cfa_pixels": P,
                   "path": "plain per-image: one HIP launch per node", "storage": "f16 intermediates",
                   "graph": "i-raw>denoise(strength .5)>demosaic>colour>filmcurv>llap(clarity .2)>o-pfm"},
        "roofline": roof, "graph_roofline": graph_roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    g.close()


if __name__ == "__main__":
    main()
