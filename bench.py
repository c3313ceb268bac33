#!/usr/bin/env python3
"""Benchmark of the raw-development path on one MI355X.

--mode batch (default): 64 stills of mixed sizes and CFAs -- 16x 4000x3000,
24x 6000x4000, 16x 8256x5504, 8x 9504x6336, a quarter of them X-Trans by a
seeded shuffle, 14-bit in uint16, per-image seeded scenes (smooth gradients +
200 rectangles over 12 stops, 2% clipped) with Poisson-Gaussian noise
a ~ U[1e-4, 1e-3], b = 1e-6 -- developed through i-raw -> denoise ->
demosaic -> colour -> filmcurv -> llap as one batched hipGraph (one launch
per stage over all 64, mosaics resident in HBM).  A step is one replay of
that graph.  The same 64 images through the plain per-image path (one
per-node graph per image) are reported beside it.

--mode plain: the per-image drop-in path on one 24 MP Bayer still
(6000 x 4000) or X-Trans still (6240 x 4160), run by the host runtime
(module DSOs, one HIP launch per node).  The DNG is decoded and uploaded
once before timing; a step is one pass of every node over the resident raw.

--mode mcraw: i-mcraw's frame codec (prefix + decode nodes) on one
4096 x 2160 frame resident in HBM; its own JSON line (frames in Mpix/s).

Prints one JSON line: value is CFA Mpix/s (after i-raw block rounding);
roofline is for the kernel or stage with the largest share of a step,
timed with HIP events on the stream it runs on inside the measured run;
cpu_baseline times the oracle on the host on a bounded sample (the
2048 x 1536 Bayer case).
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


def cpu_baseline():
    """the oracle on the host cores, on a bounded sample of the workload: the
    2048 x 1536 Bayer case of the CPU configuration.  The float build
    (float32 arithmetic in the reference's operation order, per-pixel row
    loops over all host threads) stands in for the reference's own CPU path,
    vkdt-cli on lavapipe, which this image cannot run; the FP64 build on one
    thread is reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import graphspec as gs
    import oracle_py as op
    from vkdt_amd import synth
    W, H = 2048, 1536
    cfa, _ = synth.mosaic(W, H, 1, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)
    p = gs.oracle_params(gs.DevelopParams(kind="bayer", noise_a=na, noise_b=nb))
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    L32 = op.lib(32)
    L32.or_set_threads(cores)
    op.develop(32, cfa[:256, :256], p)                 # threads up
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < 8.0 or reps < 2:
        out = op.develop(32, cfa, p)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    assert np.isfinite(out).all()
    L64 = op.lib(64)
    L64.or_set_threads(1)
    t1 = time.perf_counter()
    op.develop(64, cfa, p)
    d64 = time.perf_counter() - t1
    L64.or_set_threads(cores)
    return {"value": round(W * H / dt / 1e6, 3), "unit": "Mpix/s", "cores": cores, "kind": "port",
            "sample": f"the oracle's float build (FP32 in the reference's operation order; stand-in for vkdt-cli on "
                      f"lavapipe, which this image lacks) on {cores} host threads, the 2048x1536 Bayer case: "
                      f"{dt:.2f} s per image over {reps} runs; its FP64 build on 1 thread: {d64:.1f} s "
                      f"({W * H / d64 / 1e6:.3f} Mpix/s)"}


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


def calibrate_valu(iters=4096, blocks=8192):
    """issue rates of single VALU instructions: K independent chains of one
    instruction per lane (vkdt_amd/csrc/calib/calib.hip), 8192 work groups of
    256 threads (32 waves per CU resident over 256 CUs), timed with HIP
    events.  Rate = wave-instructions / s, and per CU per clock at the
    MI355X's 2.4 GHz peak engine clock."""
    import ctypes as C
    from vkdt_amd import hip
    hip.init(0)
    L = C.CDLL(os.path.join(ROOT, "vkdt_amd", "lib", "libvkdt_calib.so"))
    L.vkdt_calib_run.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    K = L.vkdt_calib_chains()
    out = hip.DeviceBuffer(blocks * 256, 1, 1, hip.FMT_F32)
    waves = blocks * 256 // 64
    res = {}
    for op, name in enumerate(["v_fma_f32", "v_mul_f32", "v_exp_f32", "v_rcp_f32", "v_sqrt_f32",
                               "v_fma_f32+v_exp_f32", "v_fma_f32+v_add_u32"]):
        hip.check(L.vkdt_calib_run(op, C.c_void_p(out.ptr), blocks, iters, None), name)
        hip.sync()
        e0, e1 = hip.Event(), hip.Event()
        e0.record()
        for _ in range(5):
            hip.check(L.vkdt_calib_run(op, C.c_void_p(out.ptr), blocks, iters, None), name)
        e1.record()
        hip.sync()
        ms = e0.elapsed_ms(e1) / 5
        rate = waves * iters * K / (ms * 1e-3)
        res[name] = {"ms": round(ms, 4), "wave_instr_per_s": float(f"{rate:.4g}"),
                     "per_cu_per_clk": round(rate / (256 * 2.4e9), 3)}
        print(json.dumps({"op": name, **res[name]}), flush=True)
    out.free()
    return res


# i-mcraw: bits per value stored by each 4-bit block count (counts 7, 9 and
# 11..15 alias 8, 10 and 16 bits), i-mcraw/prefix.comp:40-55
MC_WIDTH = [0, 1, 2, 3, 4, 5, 6, 8, 8, 10, 10, 16, 16, 16, 16, 16]


def run_mcraw(args):
    """i-mcraw's frame codec on one 4096 x 2160 frame (the serving workload's
    video frames), resident in HBM: a step is the prefix node (block offsets
    from the 4-bit counts) and the decode node (64-value blocks into the ui16
    mosaic).  Two frames: counts uniform over 2..6 (2-6 bits a value, a
    noisy 12-bit video frame) and all 16 bits (the codec's worst case).  The
    roofline is the decode kernel's: algorithmic bytes = counts (n/2) +
    offsets (4 n) + references (2 n) + the coded data, read, and the frame
    (2 B a sample), written; n = 138240 blocks."""
    import numpy as np
    from vkdt_amd import hip
    hip.init(0)
    owd, oht = 4096, 2160
    nb = owd * oht // 64
    rblock = -(-nb // 64) * 64
    res = {}
    for name, lo, hi in (("counts 2..6", 2, 7), ("16 bit", 11, 12)):
        rng = np.random.default_rng(7)
        cnt = rng.integers(lo, hi, size=rblock).astype(np.uint8)
        cnt[nb:] = 0
        bits = (cnt[0::2] | (cnt[1::2] << 4)).astype(np.uint8)
        refs = rng.integers(0, 4096, size=rblock).astype(np.uint16)
        dlen = int(sum(MC_WIDTH[c] * 8 for c in cnt[:nb]))
        data = rng.integers(0, 256, size=max(dlen, 1), dtype=np.uint8)
        b_bits = hip.DeviceBuffer(bits.size, 1, 1, hip.FMT_UI8)
        b_ref = hip.DeviceBuffer(rblock, 1, 1, hip.FMT_UI16)
        b_dat = hip.DeviceBuffer(data.size, 1, 1, hip.FMT_UI8)
        b_off = hip.DeviceBuffer(rblock, 1, 1, hip.FMT_UI32)
        b_scr = hip.DeviceBuffer((rblock + 511) // 512 + 1, 4, 1, hip.FMT_UI32)
        b_out = hip.DeviceBuffer(owd, oht, 1, hip.FMT_UI16)
        b_bits.upload(bits)
        b_ref.upload(refs)
        b_dat.upload(data)
        pc = np.array([0, 0, owd, oht], np.int32)
        prefix = lambda: hip.launch("i-mcraw", "prefix", None, None, [b_bits, b_off, b_scr], (nb + 511) // 512 * 8, 1)
        decode = lambda: hip.launch("i-mcraw", "decode", pc, None, [b_out, b_bits, b_off, b_ref, b_dat], nb * 8, 1)
        for _ in range(args.warmup):
            prefix()
            decode()
        hip.sync()
        e0, e1, e2 = hip.Event(), hip.Event(), hip.Event()
        e0.record()
        for _ in range(args.steps):
            prefix()
            decode()
        e1.record()
        for _ in range(args.steps):
            decode()
        e2.record()
        hip.sync()
        ms_step = e0.elapsed_ms(e1) / args.steps
        ms_dec = e1.elapsed_ms(e2) / args.steps
        alg = nb // 2 + 4 * nb + 2 * nb + dlen + 2 * owd * oht
        res[name] = {"ms_per_step": round(ms_step, 4), "decode_ms": round(ms_dec, 4),
                     "Mpix_s": round(owd * oht / (ms_step * 1e3), 1), "frames_s": round(1e3 / ms_step, 1),
                     "data_bytes": dlen, "decode_bytes": alg,
                     "decode_GB_s": round(alg / (ms_dec * 1e6), 1)}
        for b in (b_bits, b_ref, b_dat, b_off, b_scr, b_out):
            b.free()
    typ = res["counts 2..6"]
    return {
        "metric": "Mpix/s i-mcraw frame decode (prefix + decode)", "value": typ["Mpix_s"], "unit": "Mpix/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": typ["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16",
        "data": "synthetic (seeded block counts, references and coded bytes)",
        "config": {"workload": "i-mcraw 4096x2160 frame, block counts uniform over 2..6", "frames": res},
        "roofline": {"bound": "hbm", "achieved": typ["decode_GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(typ["decode_GB_s"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "i-mcraw/decode"},
    }


def plain_graph(dng, kind, noise_a, noise_b, out_dir, extra=(), hilite=False):
    from vkdt_amd import pipe
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=noise_a, noise_b=noise_b, extra=list(extra),
                                            hilite=hilite))
    g.export("o-pfm", os.path.join(out_dir, "out"))
    return g


def run_plain(args):
    """the per-image drop-in path on one still, per-node timings"""
    from vkdt_amd import hip, pipe, synth
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
    if args.noise_profile:
        # i-raw -> rawhist -> nprof: the FP64 noise statistics of the frame
        g = pipe.Graph(lines=pipe.noise_profile_cfg(dng))
        workload = workload.replace("plain per-image path", "noise profile (rawhist collect, FP64 moments)")
    else:
        g = plain_graph(dng, args.kind, na, nb, tmp, extra, hilite=args.hilite)
    if args.hilite:
        workload += ", hilite between denoise and demosaic (the stock darkroom chain)"
    g.run(pipe.RUN_ALL & ~pipe.RUN_DOWNLOAD)        # plan, allocate, upload, first pass
    ip = g.img_param("i-raw", "main")
    P = ip["full"][0] * ip["full"][1]
    nodes = {n["id"]: n for n in g.nodes()}
    order = g.exec_nodes()
    run_flags = pipe.RUN_RECORD                      # resident raw, no host traffic
    for _ in range(args.warmup):
        g.run(run_flags)
    hip.sync(g.stream())

    # per-node breakdown: HIP events between the launches (an eager pass)
    g.set_perf(True)
    per = {}
    for _ in range(args.steps):
        g.run(run_flags)                             # syncs the stream (per-node events)
        for k, ms in zip(order, g.exec_ms()):
            n = nodes[k]
            if n["module"] in ("i-raw", "o-pfm", "nprof"):
                continue
            key = n["module"] + "/" + n["kernel"]
            e = per.setdefault(key, {"ms": 0.0, "bytes": 0, "launches": 0})
            e["ms"] += ms
            e["bytes"] += node_bytes(n)
            e["launches"] += 1
    hip.sync(g.stream())
    g.set_perf(False)
    # the timed steps: the node launches recorded once and replayed as one
    # hipGraph per step (dt_graph_set_replay), one sync after the last step
    g.set_replay(True)
    for _ in range(2):
        g.run(run_flags)
    hip.sync(g.stream())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.run(run_flags)
    hip.sync(g.stream())
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    gpu_ms = sum(e["ms"] for e in per.values()) / args.steps
    value = P / (ms_step * 1e-3) / 1e6
    graph_bytes = sum(node_bytes(nodes[k]) for k in order if nodes[k]["module"] not in ("i-raw", "o-pfm", "nprof"))

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
    g.close()
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene + Poisson-Gaussian noise, uncompressed DNG)",
        "config": {"workload": workload, "width": W, "height": H, "cfa_pixels": P,
                   "path": "plain per-image: one HIP launch per node, replayed as one hipGraph per step",
                   "storage": "f16 intermediates",
                   "graph": "i-raw>rawhist>nprof" if args.noise_profile else
                   "i-raw>denoise(strength .5)>" + ("hilite>" if args.hilite else "")
                   + "demosaic>colour>filmcurv>llap(clarity .2)>o-pfm"},
        "roofline": roof, "graph_roofline": graph_roof, "cpu_baseline": cpu,
    }


# 64 stills: 16x 12 MP, 24x 24 MP, 16x 45 MP, 8x 61 MP
BATCH_SIZES = [(4000, 3000, 16), (6000, 4000, 24), (8256, 5504, 16), (9504, 6336, 8)]


def batch_workload(seed=2024):
    """(W, H, kind, a) per image: 25% X-Trans by a seeded shuffle, gaussian
    noise term a ~ U[1e-4, 1e-3], poissonian b = 1e-6"""
    import numpy as np
    rng = np.random.default_rng(seed)
    shapes = [(W, H) for W, H, n in BATCH_SIZES for _ in range(n)]
    xt = np.zeros(len(shapes), bool)
    xt[rng.permutation(len(shapes))[:len(shapes) // 4]] = True
    a = rng.uniform(1e-4, 1e-3, len(shapes))
    return [(W, H, "xtrans" if x else "bayer", float(ai)) for (W, H), x, ai in zip(shapes, xt, a)]


def stage_bytes(d):
    """compulsory HBM bytes of each batch stage group for one image: each
    stored input read once, each stored output written once (the fused
    stages' internal values never reach HBM)"""
    W, H, ng, nl = d.wd, d.ht, d.ll_numgamma, d.ll_levels
    b = 3 if d.filters == 9 else 2
    Q = (W // b) * (H // b)
    lw, lh = [W], [H]
    for _ in range(1, nl):
        lw.append((lw[-1] - 1) // 2 + 1)
        lh.append((lh[-1] - 1) // 2 + 1)
    L = [lw[l] * lh[l] for l in range(nl)]
    pyr = sum((ng + 1) * 2 * (L[l - 1] + L[l]) for l in range(2, nl))            # reduce 2..
    pyr += sum(4 * 2 * L[l] + 3 * 2 * L[l + 1] for l in range(1, nl - 1))        # assemble nl-2..1
    hl = 0
    if d.hl_on:   # half, reduce 1..L, assemble L-1..0, doub (rgba f16 levels of the block grid)
        hq = [Q]
        for _ in range(d.hl_levels):
            hq.append(((W // b - 1) // 2 ** len(hq) + 1) * ((H // b - 1) // 2 ** len(hq) + 1))
        hl = 2 * W * H + 8 * Q + sum(8 * (hq[l - 1] + hq[l]) for l in range(1, len(hq)))
        hl += sum(8 * (2 * hq[l] + hq[l + 1]) for l in range(len(hq) - 1)) + 4 * W * H + 8 * Q
    return [2 * W * H + 8 * Q,                  # half: mosaic -> Q rgba
            4 * 16 * Q,                         # downcov + 3 down
            6 * 8 * Q,                          # assemble: 5 levels -> Q
            4 * W * H + 16 * Q,                 # doub: mosaic + 2 Q -> mosaic f16
            hl,                                 # hilite: mosaic -> mosaic f16 over its wavelet levels
            2 * W * H + 8 * Q,                  # gauss: mosaic f16 -> block covariance
            10 * W * H + 8 * Q,                 # develop: mosaic + cov -> rgba f16
            8 * W * H + (ng + 1) * 2 * L[1],    # pyr1: rgba f16 -> level 1 planes
            pyr,                                # pyramid levels
            24 * W * H + 3 * 2 * L[1]]          # final: rgba f16 + level 1 -> rgba f32


def run_batch(args):
    """64 mixed stills in one batched hipGraph vs the plain per-image path"""
    import numpy as np
    import torch
    from vkdt_amd import batch, hip, pipe, synth, synth_gpu
    hip.init(0)
    torch.cuda.init()
    work = batch_workload()
    tmp = tempfile.mkdtemp(prefix="vkdt_batch_")
    combos = sorted(set((W, H, k) for W, H, k, _ in work))
    dngs = {}
    for ci, (W, H, k) in enumerate(combos):     # one DNG per shape and CFA: metadata + the plain run's input
        cfa, _ = synth_gpu.mosaic(W, H, 900 + ci, k)
        dngs[(W, H, k)] = os.path.join(tmp, f"s{ci}.dng")
        synth.write_dng(dngs[(W, H, k)], cfa.cpu().numpy().view(np.uint16), k)
        del cfa
    print(f"[bench] wrote {len(combos)} DNGs", file=sys.stderr, flush=True)

    # plain per-image path: each shape/CFA timed (median of 3 resident runs)
    plain_ms = {}
    for (W, H, k) in ([] if args.no_plain else combos):
        na, nb = synth.noise_params(5e-4, 1e-6)
        g = plain_graph(dngs[(W, H, k)], k, na, nb, tmp, hilite=args.hilite)
        g.run(pipe.RUN_ALL & ~pipe.RUN_DOWNLOAD)
        g.run(pipe.RUN_RECORD)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            g.run(pipe.RUN_RECORD)
            hip.sync(g.stream())
            ts.append((time.perf_counter() - t0) * 1e3)
        plain_ms[(W, H, k)] = sorted(ts)[1]
        g.close()
    print(f"[bench] plain {dict((f'{W}x{H}{k[0]}', round(v, 2)) for (W, H, k), v in plain_ms.items())}",
          file=sys.stderr, flush=True)

    # the batch: descriptors from each image's planned graph, mosaics on the GPU
    descs, keep = [], []
    for i, (W, H, k, a) in enumerate(work):
        na, nb = synth.noise_params(a, 1e-6)
        g = plain_graph(dngs[(W, H, k)], k, na, nb, tmp, hilite=args.hilite)
        g.plan()
        d = batch.describe(g)
        g.close()
        # the mosaic as i-raw stages it: the block-rounded extent of the DNG
        cfa, _ = synth_gpu.mosaic(d.cfa_wd, d.cfa_ht, i, k, a=a, b=1e-6)
        out = torch.empty((d.ht, d.wd, 4), dtype=torch.float32, device="cuda")
        d.cfa, d.out = cfa.data_ptr(), out.data_ptr()
        descs.append(d)
        keep += [cfa, out]
        if (i + 1) % 16 == 0:
            print(f"[bench] planned and synthesised {i + 1} images", file=sys.stderr, flush=True)
    P = sum(d.wd * d.ht for d in descs)
    nbytes = batch.scratch_bytes(descs)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    print(f"[bench] {len(descs)} images, {P / 1e6:.0f} Mpix, scratch {nbytes / 2**30:.1f} GiB",
          file=sys.stderr, flush=True)
    stream = hip.Stream()
    b = batch.Batch(descs, scratch.data_ptr(), nbytes)
    tune = b.autotune(stream.ptr)
    b.set_timing(True)
    b.run(stream.ptr)
    stage_ms = b.stage_ms()
    b.set_timing(False)
    for _ in range(args.warmup):
        b.run(stream.ptr, graph=True)
    stream.sync()
    e0, e1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    e0.record(stream.ptr)
    for _ in range(args.steps):
        b.run(stream.ptr, graph=True)
    e1.record(stream.ptr)
    stream.sync()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_ms(e1) / args.steps
    ms_step = wall * 1e3 / args.steps
    value = P / (ms_step * 1e-3) / 1e6
    b.close()

    names = batch.stage_names()
    sb = [sum(x) for x in zip(*[stage_bytes(d) for d in descs])]
    dom = max(range(len(names)), key=lambda s: stage_ms[s])
    achieved = sb[dom] / (stage_ms[dom] * 1e-3) / 1e9
    plain_total = sum(plain_ms[(W, H, k)] for W, H, k, _ in work) if plain_ms else None
    plain_value = P / (plain_total * 1e-3) / 1e6 if plain_ms else None
    roof = {"bound": "hbm", "kernel": "batch/" + names[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("batch64", "batch/" + names[dom]),
            "bytes_per_launch": int(sb[dom]), "avg_launch_ms": round(stage_ms[dom], 4)}
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scenes generated on the GPU, Poisson-Gaussian noise)",
        "config": {"workload": "batched: 64 mixed stills in one hipGraph (16x 4000x3000, 24x 6000x4000, "
                               "16x 8256x5504, 8x 9504x6336; 16 X-Trans by seeded shuffle; noise a~U[1e-4,1e-3])",
                   "images": len(descs), "cfa_pixels": P, "storage": "f16 intermediates, rgba f32 out",
                   "graph": "i-raw>denoise(strength .5)>" + ("hilite>" if args.hilite else "")
                            + "demosaic>colour>filmcurv>llap(clarity .2)",
                   "tune": dict(zip(names, tune))},
        "gpu_ms_per_step": round(gpu_ms, 3),
        "plain": None if not plain_ms else {
            "value": round(plain_value, 2), "unit": "Mpix/s", "ms_64": round(plain_total, 2),
            "method": "per-node graph, resident raw; each shape/CFA pair timed (median of 3) "
                      "and summed over the 64 images"},
        "speedup_vs_plain": round(value / plain_value, 3) if plain_ms else None,
        "stages_ms": dict(zip(names, [round(m, 3) for m in stage_ms])),
        "stages_compulsory_gb": dict(zip(names, [round(b / 1e9, 2) for b in sb])),
        "roofline": roof,
        "graph_roofline": {"compulsory_bytes": int(sum(sb)), "bytes_per_px": round(sum(sb) / P, 1),
                           "gpu_ms": round(sum(stage_ms), 3),
                           "achieved": round(sum(sb) / (sum(stage_ms) * 1e-3) / 1e9, 1),
                           "frac": round(sum(sb) / (sum(stage_ms) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }


# serving: 8 raw-video streams (4096x2160 RGGB, 12-bit: black 256, white 4095,
# 30 fps, scenes drifting 2 px per frame) and 24-61 MP 14-bit stills
STREAMS, FRAME_W, FRAME_H, POOL_FRAMES = 8, 4096, 2160, 8
STILLS = [(6000, 4000), (7008, 4672), (8256, 5504), (9504, 6336)]


SERVE_P99_MS = 50.0   # the latency bound the serving capacity is quoted at


def serve_workload(seconds, rate=600.0, p_still=0.2, seed=31):
    """Poisson arrivals at `rate`; each is a still with probability p_still,
    else the next frame of a uniformly chosen stream"""
    import numpy as np
    rng = np.random.default_rng(seed)
    t, out, nxt = 0.0, [], [0] * STREAMS
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        if rng.random() < p_still:
            out.append((("still", int(rng.integers(len(STILLS)))), t))
        else:
            s = int(rng.integers(STREAMS))
            out.append((("frame", s, nxt[s] % POOL_FRAMES), t))
            nxt[s] += 1


def run_serve(args, tune=None):
    """the serving loop on the mixed request stream, beside a FIFO of the
    plain per-image path over the same arrivals (measured per-image times).
    tune None: the built-in launch shape, one tile per work group, so that
    the small batches in flight on the three streams overlap on the device
    (persistent work groups, which the 64-still batch's autotuner may pick,
    fill every CU for one large batch and serialise concurrent ones)"""
    import numpy as np
    import torch
    from vkdt_amd import batch, hip, pipe, serve, synth, synth_gpu
    import gc
    gc.collect()
    torch.cuda.empty_cache()      # the batch run's tensors go back to the driver
    tmp = tempfile.mkdtemp(prefix="vkdt_serve_")
    rng = np.random.default_rng(17)
    profiles, host, plain_ms = {}, {}, {}

    def dng_for(W, H, black, white, seed):
        cfa, _ = synth_gpu.mosaic(W, H, seed, "bayer", black=black, white=white)
        path = os.path.join(tmp, f"p{W}x{H}.dng")
        synth.write_dng(path, cfa.cpu().numpy().view(np.uint16), "bayer", black=black, white=white)
        return path

    def plan(path, a, black, white):
        na, nb = synth.noise_params(a, 1e-6, black, white)
        g = plain_graph(path, "bayer", na, nb, tmp, hilite=args.hilite)
        g.plan()
        d = batch.describe(g)
        g.close()
        return d, na, nb, a

    def time_plain(path, na, nb):
        g = plain_graph(path, "bayer", na, nb, tmp, hilite=args.hilite)
        g.run(pipe.RUN_ALL & ~pipe.RUN_DOWNLOAD)
        ts = []
        for _ in range(4):
            t0 = time.perf_counter()
            g.run(pipe.RUN_RECORD)
            hip.sync(g.stream())
            ts.append((time.perf_counter() - t0) * 1e3)
        g.close()
        return sorted(ts)[1]

    fdng = dng_for(FRAME_W, FRAME_H, 256, 4095, 700)
    for s in range(STREAMS):
        d, na, nb, a = plan(fdng, float(rng.uniform(1e-4, 1e-3)), 256, 4095)
        profiles[("frame", s)] = serve.Profile(d, f"stream{s}")
        sc = None
        for f in range(POOL_FRAMES):
            cfa, sc = synth_gpu.mosaic(d.cfa_wd, d.cfa_ht, 1000 + s, "bayer", a=a, black=256, white=4095,
                                       origin=(2 * f, 0), canvas=(d.cfa_wd + 2 * POOL_FRAMES, d.cfa_ht), scale=sc)
            host[("frame", s, f)] = cfa.cpu().pin_memory()
        if s == 0:
            plain_ms["frame"] = time_plain(fdng, na, nb)
    for i, (W, H) in enumerate(STILLS):
        path = dng_for(W, H, 512, 16383, 710 + i)
        d, na, nb, _ = plan(path, 2e-4, 512, 16383)
        profiles[("still", i)] = serve.Profile(d, f"still{W}x{H}")
        cfa, _ = synth_gpu.mosaic(d.cfa_wd, d.cfa_ht, 2000 + i, "bayer")
        host[("still", i)] = cfa.cpu().pin_memory()
        plain_ms[("still", i)] = time_plain(path, na, nb)
    torch.cuda.synchronize()
    print(f"[bench] serving profiles ready; plain ms {plain_ms}", file=sys.stderr, flush=True)

    def serve_at(rate, seconds):
        work = serve_workload(seconds, rate=rate)
        reqs = []
        for i, (key, t) in enumerate(work):
            prof = profiles[key[:2]]
            reqs.append(serve.Request(i, prof, host[key].data_ptr(), t, tag=key[0]))
        srv = serve.Server(200 << 30, slots=3, max_batch=48, tune=tune)
        try:
            wall = srv.run(reqs)
        finally:
            srv.close()
        steady = [r for r in reqs if r.arrival >= 0.5]
        s = serve.summary(steady, wall - 0.5)
        # the plain path as one FIFO server over the same arrivals
        fin, lat = 0.0, []
        for key, t in work:
            ms = plain_ms["frame"] if key[0] == "frame" else plain_ms[key[:2]]
            fin = max(fin, t) + ms * 1e-3
            if t >= 0.5:
                lat.append((fin - t) * 1e3)
        px = sum(r.profile.pixels for r in reqs)
        busy = sum(plain_ms["frame"] if k[0] == "frame" else plain_ms[k[:2]] for k, _ in work) * 1e-3
        sizes = [b[0] for b in srv.batches]
        return work, reqs, srv, s, lat, px, busy, sizes

    work, reqs, srv, s, lat, px, busy, sizes = serve_at(600.0, args.serve_seconds)
    nst = sum(1 for r in reqs if r.tag == "still")
    # capacity: offered load against p99, arrivals of the same mix at rising rates
    sweep = []
    for rate in [float(r) for r in args.serve_sweep.split(",") if r.strip()]:
        w2, r2, _, s2, lat2, px2, _, sz2 = serve_at(rate, args.sweep_seconds)
        sweep.append({"rate_per_s": rate, "offered_mpix_per_s": round(px2 / args.sweep_seconds / 1e6, 1),
                      "served_mpix_per_s": s2["mpix_per_s"], "p50_ms": s2["p50_ms"], "p99_ms": s2["p99_ms"],
                      "mean_batch": round(float(np.mean(sz2)), 2),
                      "plain_fifo_p99_ms": round(float(np.percentile(lat2, 99)), 3)})
        print(f"[bench] serving {rate:.0f} req/s: {sweep[-1]}", file=sys.stderr, flush=True)
    ok = [e for e in sweep if e["p99_ms"] <= SERVE_P99_MS and e["served_mpix_per_s"] >= 0.97 * e["offered_mpix_per_s"]]
    capacity = {"p99_bound_ms": SERVE_P99_MS,
                "rate_per_s": ok[-1]["rate_per_s"] if ok else None,
                "mpix_per_s": ok[-1]["served_mpix_per_s"] if ok else None,
                "method": f"highest swept rate ({args.sweep_seconds:g} s of arrivals each) whose p99 stays within "
                          f"{SERVE_P99_MS} ms and whose served rate is within 3% of the offered rate"}
    return {
        "workload": f"Poisson {len(work) / args.serve_seconds:.0f} req/s for {args.serve_seconds:g} s: "
                    f"{STREAMS} streams of {FRAME_W}x{FRAME_H} RGGB 12-bit frames + 24-61 MP 14-bit stills"
                    + (", hilite between denoise and demosaic" if args.hilite else ""),
        "requests": len(reqs), "stills": nst, "frames": len(reqs) - nst,
        "offered_mpix_per_s": round(px / args.serve_seconds / 1e6, 1),
        "served": s, "batches": len(sizes), "mean_batch": round(float(np.mean(sizes)), 2),
        "budget_gib": 200, "peak_gib": round(srv.arena.peak / 2**30, 2), "slots": 3,
        "launch_shape": "built-in: one tile per work group" if tune is None else dict(zip(batch.stage_names(), tune)),
        "plain_fifo": {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                       "capacity_mpix_per_s": round(px / busy / 1e6, 1),
                       "method": "FIFO over the same arrivals with the per-node graph's measured time per shape"},
        "sweep": sweep, "capacity": capacity,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: 20 plain, 5 batch)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default: 3 plain, 2 batch)")
    ap.add_argument("--mode", default="batch", choices=["batch", "plain", "mcraw"],
                    help="batch: 64 mixed stills in one batched hipGraph vs the plain path; "
                         "plain: the per-image path on one still; mcraw: i-mcraw's frame codec on one "
                         "4096x2160 frame")
    ap.add_argument("--kind", default="bayer", choices=["bayer", "xtrans"], help="plain mode: the still's CFA")
    ap.add_argument("--hilite", action="store_true",
                    help="insert hilite between denoise and demosaic (the stock darkroom chain)")
    ap.add_argument("--noise-profile", action="store_true",
                    help="plain mode: time the noise-profile graph i-raw > rawhist > nprof instead")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="batch mode: skip timing the plain path beside it")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="batch mode: also run the serving loop on this many seconds of arrivals (0: skip)")
    ap.add_argument("--serve-sweep", default="650,700,750,900,1050,1200",
                    help="batch mode: arrival rates (req/s) of the capacity sweep, comma separated ('' to skip)")
    ap.add_argument("--sweep-seconds", type=float, default=3.0, help="seconds of arrivals per swept rate")
    ap.add_argument("--breakdown", default="", help="plain mode: write per-kernel timings as JSON to this path")
    ap.add_argument("--calibrate-valu", action="store_true",
                    help="only time chains of v_fma/v_mul/v_exp/v_rcp/v_sqrt_f32 (the VALU issue rates)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB (5 times), the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.calibrate_valu:
        calibrate_valu()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(images are independent units and would shard across GPUs with no communication)",
              file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = 5 if args.mode == "batch" else 20
    if args.warmup is None:
        args.warmup = 2 if args.mode == "batch" else 3
    if args.mode == "mcraw":
        line = run_mcraw(args)
    elif args.mode == "plain":
        from vkdt_amd import hip
        hip.init(0)
        line = run_plain(args)
    else:
        line = run_batch(args)
        if args.serve_seconds > 0:
            line["serving"] = run_serve(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
