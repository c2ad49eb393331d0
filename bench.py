#!/usr/bin/env python3
"""MDAS sampler benchmark on one MI355X.

A step is one full MDAS run over one synthetic 1080p image: initial Morton
grid, adaptive passes (error, selection, best-candidate placement, median
split) with the analytic integrand between them, until N >= spp * W * H
(optixMotionBlur.cpp:768-791), then the reconstruction to pixels.

value = final samples / GPU time of the MDAS phases (initial + errors +
adaptive + reconstruction), the reference's "total sampling time"
(SDK/data/Scripts/breakdown_table.py:58-66); the integrand (the OptiX trace
in the reference) is timed separately.  e2e_msamples_per_s = final samples /
wall time of the same steps (integrand and reconstruction included) run
again with kernel timing off (wall_ms_untimed), as the library runs in
production; ms_per_step is the wall time of the timed steps.

Prints one JSON line.  Usage:
    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # 5D motion blur + depth of field, 1920x1080, 32 spp (64 moving spheres)
    "mdas-5d-mbdof-1080p-32spp": dict(dim=5, kind=3, W=1920, H=1080, spp=32, b=1, e=8,
                                      alpha=0.25, cand=4, seed=1),
    # 4D depth of field, 1920x1080, 16 spp (64 static spheres)
    "mdas-4d-dof-1080p-16spp": dict(dim=4, kind=2, W=1920, H=1080, spp=16, b=1, e=8,
                                    alpha=0.25, cand=4, seed=1),
}
DEFAULT = "mdas-5d-mbdof-1080p-32spp"
HBM_PEAK_GBS = 8000.0


def capacity(w):
    target = int(w["spp"] * w["W"] * w["H"])
    l0 = 1 << (w["b"] * w["dim"] + 2 * w["e"])
    return target, int(target * 1.5) + 4 * l0 + 1024


def run_gpu(name, steps, warmup, log=print):
    from pmdas_amd import hip, kdgrid
    w = WORKLOADS[name]
    target, cap = capacity(w)
    g = kdgrid.KDGrid(w["dim"], cap, w["cand"], w["b"], w["e"], w["alpha"], float(w["W"]),
                      float(w["H"]))
    sc = kdgrid.Scene(w["kind"], w["seed"], w["dim"], float(w["W"]), float(w["H"]))
    px = hip.DeviceBuffer(w["W"] * w["H"] * 16)
    tile = None
    for i in range(warmup):
        g.Run(sc, target)
        tile = g.autotune_recon(px.ptr, w["W"], w["H"])
        hip.synchronize()
        log(f"[bench] {name}: warmup {i + 1}/{warmup} done, reconstruction tile {tile}")
    g.set_timing(True)
    tot_mdas = tot_wall = tot_integrand = 0.0
    samples = 0
    kstats = {}
    phases = None
    for i in range(steps):
        hip.synchronize()
        t0 = time.perf_counter()
        g.Run(sc, target)
        g.Integrate(px.ptr, w["W"], w["H"])
        hip.synchronize()
        t1 = time.perf_counter()
        ph = g.phase_times()
        mdas_ms = ph["initial_ms"] + ph["errors_ms"] + ph["adaptive_ms"] + ph["recon_ms"]
        tot_mdas += mdas_ms
        tot_integrand += ph["integrand_ms"]
        tot_wall += (t1 - t0) * 1e3
        samples += g.GetNumberOfSamples()
        for k, (la, ms, by) in g.kernel_stats().items():
            a = kstats.setdefault(k, [0, 0.0, 0.0])
            a[0] += la
            a[1] += ms
            a[2] += by
        phases = ph
        log(f"[bench] {name}: step {i + 1}/{steps}: {mdas_ms:.2f} ms MDAS, "
            f"{ph['integrand_ms']:.2f} ms integrand, {(t1 - t0) * 1e3:.2f} ms wall, "
            f"{ph['passes']} passes, N={g.GetNumberOfSamples()}")
    pstats = g.pass_stats()
    # the same steps with kernel timing off (no per-kernel events), as the
    # library runs in production: the end-to-end rate
    g.set_timing(False)
    hip.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        g.Run(sc, target)
        g.Integrate(px.ptr, w["W"], w["H"])
    hip.synchronize()
    wall_untimed = (time.perf_counter() - t0) * 1e3 / steps
    res = dict(
        workload=name, target=target, samples=samples // steps, recon_tile=tile,
        leaves=g.GetNumberOfNodes(), passes=phases["passes"],
        mdas_ms=tot_mdas / steps, integrand_ms=tot_integrand / steps, wall_ms=tot_wall / steps,
        wall_ms_untimed=wall_untimed,
        msamples_per_s=samples / (tot_mdas * 1e3),
        e2e_msamples_per_s=samples / steps / (wall_untimed * 1e3),
        phases_ms={k: phases[k] for k in ("initial_ms", "errors_ms", "adaptive_ms", "recon_ms",
                                          "sort_ms", "integrand_ms")},
        kernels={k: dict(launches=v[0] // steps, ms=v[1] / steps, bytes=v[2] / steps)
                 for k, v in kstats.items()},
    )
    res["phase_model_bytes"] = phase_model_bytes(w, pstats, res["leaves"], res["samples"],
                                                 res["kernels"].get("k_rs_scatter"))
    px.free()
    g.close()
    sc.close()
    return res


def batch_images(n=32, seed=7):
    """n images with pixel counts log-uniform in [256^2, 2048^2] (16:9-ish),
    extraImgBits chosen per image so the initial grid holds about 4 spp."""
    import math
    import random
    rng = random.Random(seed)
    out = []
    for i in range(n):
        p = 2.0 ** rng.uniform(16.0, 22.0)
        W = max(64, int(round(math.sqrt(p * 16.0 / 9.0))))
        H = max(64, int(round(p / W)))
        e = int(math.floor(0.5 * math.log2(W * H / 32.0) + 0.5))
        out.append(dict(dim=5, kind=3, W=W, H=H, spp=32, b=1, e=e, alpha=0.25, cand=4, seed=100 + i))
    return out


def run_batch(steps, log=print):
    """The batch workload: 32 images of uneven resolution, 5D, 32 spp, advanced as one
    batch replayed from a captured hipGraph, then reconstructed by a second
    captured graph holding the 32 reconstructions, each with the gather tile
    the autotuner picked for its image in the warmup."""
    from pmdas_amd import hip, kdgrid
    imgs = batch_images()
    grids, scenes, targets, bufs = [], [], [], []
    for w in imgs:
        target, cap = capacity(w)
        grids.append(kdgrid.KDGrid(w["dim"], cap, w["cand"], w["b"], w["e"], w["alpha"],
                                   float(w["W"]), float(w["H"])))
        scenes.append(kdgrid.Scene(w["kind"], w["seed"], w["dim"], float(w["W"]), float(w["H"])))
        targets.append(target)
        bufs.append(hip.DeviceBuffer(w["W"] * w["H"] * 16))
    batch = kdgrid.Batch(grids, scenes)
    shape = batch.autotune(targets)
    log(f"[bench] batch of {len(imgs)}: autotuned launch shape {shape} workgroups")
    times = []
    Ws, Hs = [w["W"] for w in imgs], [w["H"] for w in imgs]
    for i in range(steps + 1):
        hip.synchronize()
        t0 = time.perf_counter()
        batch.run(targets, use_graph=True)
        if i == 0:  # warmup: the autotuner picks each image's gather tile
            for g, w, b in zip(grids, imgs, bufs):
                g.autotune_recon(b.ptr, w["W"], w["H"])
        batch.integrate(bufs, Ws, Hs, use_graph=True)
        hip.synchronize()
        t1 = time.perf_counter()
        if i > 0:
            times.append((t1 - t0) * 1e3)
        log(f"[bench] batch of {len(imgs)}: {'warmup' if i == 0 else 'step'} "
            f"{(t1 - t0) * 1e3:.1f} ms")
    n = sum(g.GetNumberOfSamples() for g in grids)
    ms = sum(times) / len(times)
    res = dict(images=len(imgs), pixels=sum(w["W"] * w["H"] for w in imgs), samples=n,
               ms_per_step=round(ms, 2), e2e_msamples_per_s=round(n / (ms * 1e3), 1),
               passes_max=max(g.pass_stats().shape[0] for g in grids), launch_shape=shape,
               note="wall time of the batch run (with the integrand) and the 32 "
                    "reconstructions, both replayed from captured hipGraphs")
    batch.close()
    for b in bufs:
        b.free()
    for g in grids:
        g.close()
    return res


LEARNED = dict(train_seeds=200, train_W=160, train_H=90, spp=32, epochs=4, batch=65536, lr=1e-3,
               val_seeds=(1000, 1001, 1002, 1003), val_W=240, val_H=135,
               kappas=(1.0, 0.5, 0.25, 0.125, 0.0625),
               eval_seeds=(2000, 2001, 2002, 2003), eval_W=1920, eval_H=1080,
               eval_spp=(4, 8, 12, 16, 24, 32, 64), adaptive_spp=8, gt_spp=4096)


def learned_comparison(log=print):
    """Learned proposal vs plain MDAS on the 5D moving-sphere scene: rows
    from plain runs of 200 seeded scenes, GPU training, the score exponent
    chosen on 4 validation scenes, then image MSE against an FP64 4096-spp
    QMC ground truth on 4 held-out 1920x1080 scenes; samples and time the learned form
    needs to reach plain's error at 32 spp, and the same for plain sampling
    with the learned form's effective exponent (the flattening alone)."""
    from pmdas_amd import learned as L
    c = LEARNED
    t0 = time.perf_counter()
    X, Y, n = L.collect(range(c["train_seeds"]), c["train_W"], c["train_H"], c["spp"],
                        cap_rows=1 << 25, log=log)
    Xh, Yh, nh = L.collect(range(1000, 1010), c["train_W"], c["train_H"], c["spp"],
                           cap_rows=1 << 22, log=log)
    m, losses = L.train(X, Y, n, c["epochs"], c["batch"], c["lr"], seed=0, log=log)
    net_loss, const_loss = L.heldout_loss(m, Xh, Yh, nh)
    kappa, ktable = L.tune_kappa(m, c["val_seeds"], c["val_W"], c["val_H"], 32, c["kappas"],
                                 gt_spp=c["gt_spp"], log=log)
    rows = L.compare(m, c["eval_seeds"], c["eval_W"], c["eval_H"], c["eval_spp"],
                     gt_spp=c["gt_spp"], kappa=kappa, log=log)
    arms = ("plain", "learned", "plain_same_exponent")
    # 4 spp is the initial grid alone (the same image for every arm), so the
    # comparisons use the adaptive points only
    lo = c["adaptive_spp"]
    eq = {arm: {k: L.samples_to_equal(rows, 32, k, arm, min_spp=lo) for k in ("mse", "rel_mse")}
          for arm in arms[1:]}
    at_equal = {f"{a}_vs_{r}": {k: L.error_at_equal_samples(rows, a, r, k, min_spp=lo)
                                for k in ("mse", "rel_mse")}
                for a, r in (("learned", "plain"), ("plain_same_exponent", "plain"),
                             ("learned", "plain_same_exponent"))}
    # what the network adds over the flattened exponent alone: the learned
    # form's samples and time to plain's 32-spp error, over those of plain
    # sampling with the same effective exponent
    lm, pm = eq["learned"]["mse"], eq["plain_same_exponent"]["mse"]
    network_gain = None
    if lm and pm:
        network_gain = {"sample_ratio": round(lm["sample_ratio"] / pm["sample_ratio"], 3),
                        "time_ratio": round(lm["time_ratio"] / pm["time_ratio"], 3),
                        "bound": "upper" if "bound" in lm or "bound" in pm else None}
    per_spp = {}
    for spp in c["eval_spp"]:
        sel = [r for r in rows if r["spp"] == spp]
        per_spp[str(spp)] = {k: {"mse": float(np.mean([r[k]["mse"] for r in sel])),
                                 "rel_mse": float(np.mean([r[k]["rel_mse"] for r in sel])),
                                 "samples": float(np.mean([r[k]["samples"] for r in sel])),
                                 "ms": float(np.mean([r[k]["ms"] + r[k]["integrand_ms"]
                                                      for r in sel]))}
                             for k in arms}
    for b in (X, Y, Xh, Yh):
        b.free()
    m.close()
    return {"train_rows": n, "epoch_loss": [round(float(v), 5) for v in losses],
            "heldout_loss": round(net_loss, 5), "heldout_const_loss": round(const_loss, 5),
            "kappa": kappa, "validation_mse_by_kappa": {str(k): v for k, v in ktable.items()},
            "eval": f"{len(c['eval_seeds'])} held-out seeds, {c['eval_W']}x{c['eval_H']} 5D, "
                    f"FP64 QMC {c['gt_spp']} spp ground truth",
            "per_spp": per_spp, "equal_error_at_32spp": eq,
            "learned_vs_plain_same_exponent": network_gain,
            "error_at_equal_samples": at_equal,
            "note": f"sample_ratio / time_ratio: what each arm needs to reach plain MDAS's "
                    f"32-spp MSE (alpha 0.25), relative to plain, over the adaptive points "
                    f"(spp >= {lo}; 'bound: upper' when the arm is below it at its first one); "
                    f"error_at_equal_samples: an arm's error over another's at the other's "
                    f"sample count; plain_same_exponent is plain MDAS with alpha * kappa, the "
                    f"learned form's flattening without the network, so learned_vs_"
                    f"plain_same_exponent < 1 is what the network adds",
            "seconds": round(time.perf_counter() - t0, 1)}


FRAMES = dict(dim=6, kind=4, W=3840, H=2160, spp=64, b=1, alpha=0.25, cand=4, frames=16,
              first_seed=100, train_seeds=100, train_W=192, train_H=108,
              val_seeds=(1000, 1001), val_W=240, val_H=135, kappas=(1.0, 0.5, 0.25, 0.125))


def run_frames(log=print):
    """6D ambient occlusion at 3840x2160, 64 spp, 16 frames with per-frame
    seeds, learned proposal (network trained on plain runs of 100 small AO
    scenes, score exponent chosen on 2 validation scenes) and batching: the
    frames stream through a batch of as many frames as device memory holds
    (each frame's grid needs ~130 GB at this size, so 16 at once cannot be
    resident); each batch runs as one captured hipGraph per block of passes,
    with the network's inference fused into the error pass."""
    from pmdas_amd import hip, kdgrid, learned as L
    c = FRAMES
    t0 = time.perf_counter()
    X, Y, n = L.collect(range(c["train_seeds"]), c["train_W"], c["train_H"], c["spp"], dim=c["dim"],
                        kind=c["kind"], cap_rows=1 << 25, log=log)
    net, losses = L.train(X, Y, n, 4, 65536, 1e-3, seed=0, log=log)
    X.free()
    Y.free()
    kappa, _ = L.tune_kappa(net, c["val_seeds"], c["val_W"], c["val_H"], c["spp"], c["kappas"],
                            dim=c["dim"], kind=c["kind"], log=log)
    t_train = time.perf_counter() - t0
    W, H, D = c["W"], c["H"], c["dim"]
    e = L.extra_bits(W, H, D, c["b"])
    target = c["spp"] * W * H
    cap = int(target * 1.25) + 4 * (1 << (c["b"] * D + 2 * e)) // 8
    # one warm-up frame on the first grid, outside the timed region: it picks
    # the gather tile and, with its reconstruction buffers sized, tells how
    # much device memory a frame's grid really takes, hence how many frames
    # a batch can hold
    px = hip.DeviceBuffer(W * H * 16)
    hip.synchronize()
    free0, _ = hip.mem_info()

    def new_grid():
        g = kdgrid.KDGrid(D, cap, c["cand"], c["b"], e, c["alpha"], float(W), float(H))
        g.set_learned(2, net)
        g.set_learned_exponent(kappa)
        return g

    grids = [new_grid()]
    sc = kdgrid.Scene(c["kind"], c["first_seed"], D, float(W), float(H))
    grids[0].Run(sc, target)
    tile = grids[0].autotune_recon(px.ptr, W, H)
    sc.close()
    hip.synchronize()
    free1, _ = hip.mem_info()
    per_frame = max(1, free0 - free1)
    F = int(max(1, min(c["frames"], 1 + (free1 - (4 << 30)) // per_frame)))
    log(f"[bench] frames: {per_frame / 2**30:.1f} GiB per frame grid, {free0 / 2**30:.0f} GiB free, "
        f"{F} frame(s) per batch, gather tile {tile}")
    grids += [new_grid() for _ in range(F - 1)]
    samples, wall = 0, 0.0
    for f0 in range(0, c["frames"], F):
        fr = list(range(f0, min(c["frames"], f0 + F)))
        scenes = [kdgrid.Scene(c["kind"], c["first_seed"] + f, D, float(W), float(H)) for f in fr]
        batch = kdgrid.Batch(grids[:len(fr)], scenes)
        hip.synchronize()
        t1 = time.perf_counter()
        batch.run([target] * len(fr), use_graph=True)
        for g in grids[:len(fr)]:
            g.set_recon_tile(*tile)
            g.Integrate(px.ptr, W, H)
            samples += g.GetNumberOfSamples()
        hip.synchronize()
        wall += time.perf_counter() - t1
        log(f"[bench] frames {fr[0]}..{fr[-1]}: {samples} samples so far, {wall:.2f} s")
        batch.close()
        for sc in scenes:
            sc.close()
    # MDAS-phase throughput of one frame (same grid, learned, timed per kernel)
    g = grids[0]
    sc = kdgrid.Scene(c["kind"], c["first_seed"], D, float(W), float(H))
    g.set_timing(True)
    g.Run(sc, target)
    g.Integrate(px.ptr, W, H)
    pt = g.phase_times()
    mdas_ms = pt["initial_ms"] + pt["errors_ms"] + pt["adaptive_ms"] + pt["recon_ms"]
    n1 = g.GetNumberOfSamples()
    sc.close()
    px.free()
    for g in grids:
        g.close()
    net.close()
    return {"workload": "mdas-6d-ao-4k-64spp-16frames-learned", "frames": c["frames"],
            "frames_per_batch": F, "samples": samples,
            "e2e_msamples_per_s": round(samples / wall / 1e6, 1),
            "note": "wall time of all frames: batched sampling with the integrand + reconstruction",
            "one_frame_mdas_ms": round(mdas_ms, 3),
            "one_frame_msamples_per_s": round(n1 / (mdas_ms * 1e3), 1),
            "one_frame_integrand_ms": round(pt["integrand_ms"], 3), "kappa": kappa,
            "train_seconds": round(t_train, 1), "recon_tile": tile}


def phase_model_bytes(w, pstats, L, N, scatter):
    """Bytes of each MDAS phase as the reference's kernels move them, from the
    per-pass counters (rows of L, N, selected, split, samples of the selected
    leaves, re-estimated leaves, values they read): initial sampling writes
    4 points, node, box, error and seed per leaf (kdgrid.cu:121-202); the
    error pass reads every error and, per re-estimated leaf, node, box and
    values and writes its error (:204-288); the adaptive pass reads every
    leaf's error, box and seed and writes the seed, per selected leaf its
    node, its samples' points, and writes the new point and node, per split
    a node, two boxes and two errors (:290-481); the reconstruction reads
    node, box and values per leaf and writes every pixel (:483-582); the
    sort reads and writes key and payload per pass (12 B each way)."""
    D = w["dim"]
    L0 = 1 << (w["b"] * D + 2 * w["e"])
    rows = np.asarray(pstats, dtype=np.float64)
    Lp, sel, split, selcnt, stale, svals = (rows[:, 0], rows[:, 2], rows[:, 3], rows[:, 4],
                                            rows[:, 5], rows[:, 6])
    b = {
        "initial": L0 * (16 * D + 16 + 8 * D + 8),
        "errors": float(np.sum(4 * Lp + stale * (16 + 8 * D + 4) + 12 * svals)),
        "adaptive": float(np.sum(Lp * (4 + 8 * D + 8) + sel * (16 + 4 * D + 16) + 4 * D * selcnt
                                 + split * (16 + 16 * D + 8))),
        "recon": float(L * (16 + 8 * D) + 12 * N + 16 * w["W"] * w["H"]),
    }
    if scatter:  # key passes = scatter bytes / 16 (8-byte keys read and written)
        b["recon"] += 24.0 * scatter["bytes"] / 16.0
    return b


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def device_record():
    """What can tell two GPU machines apart when the same build times
    differently: product name, compute units and peak engine clock
    (rocminfo), current clocks and compute / memory partition modes
    (rocm-smi).  Both run as child processes before this one touches the
    GPU; a tool that is missing or fails leaves its fields out."""
    import subprocess
    rec = {}
    try:
        r = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60)
        agent, gpu = {}, None
        for ln in r.stdout.splitlines():
            ln = ln.strip()
            if ln.startswith("Agent "):
                if agent.get("Device Type") == "GPU" and gpu is None:
                    gpu = agent
                agent = {}
            elif ":" in ln:
                k, v = ln.split(":", 1)
                agent.setdefault(k.strip(), v.strip())
        if agent.get("Device Type") == "GPU" and gpu is None:
            gpu = agent
        if gpu:
            for k in ("Marketing Name", "Name", "Compute Unit", "Max Clock Freq. (MHz)",
                      "SIMDs per CU", "Shader Engines"):
                if k in gpu:
                    rec[k] = gpu[k]
    except Exception as e:  # noqa: BLE001 - a record, not a requirement
        rec["rocminfo_error"] = str(e)[:120]
    try:
        r = subprocess.run(["rocm-smi", "--showproductname", "--showclocks", "--showcomputepartition",
                            "--showmemorypartition", "--json"], capture_output=True, text=True,
                           timeout=60)
        js = json.loads(r.stdout[r.stdout.index("{"):])
        card = next((v for k, v in sorted(js.items()) if k.startswith("card")), {})
        rec["rocm_smi"] = {k: v for k, v in card.items() if isinstance(v, str) and len(v) < 80}
    except Exception as e:  # noqa: BLE001
        rec["rocm_smi_error"] = str(e)[:120]
    return rec


def copy_probe():
    """Achievable HBM bandwidth of this GPU: best of 5 plain float4 copies of
    1 GiB (pmdas_amd/bw_probe, read + write bytes / time), run as a child
    process before this one touches the GPU; None if the probe is not built."""
    exe = os.path.join(ROOT, "pmdas_amd", "bw_probe")
    if not os.path.exists(exe):
        return None
    import subprocess
    r = subprocess.run([exe, "1024"], capture_output=True, text=True, timeout=120)
    for ln in r.stdout.splitlines():
        if r.returncode == 0 and ln.startswith("best:"):
            return float(ln.split()[1])
    return None


def roofline(res, traffic_file, copy_gbs=None):
    """Dominant MDAS kernel (integrand excluded): algorithmic bytes / its time."""
    ks = {k: v for k, v in res["kernels"].items() if k != "k_eval"}
    name = max(ks, key=lambda k: ks[k]["ms"])
    k = ks[name]
    achieved = k["bytes"] / (k["ms"] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            t = json.load(f)
        if t.get("kernel") == name:
            traffic = t.get("hbm_bytes_per_launch")
    out = dict(bound="hbm", kernel=name, achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
               unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
               algorithmic_bytes_per_launch=k["bytes"] / max(k["launches"], 1),
               avg_launch_ms=k["ms"] / max(k["launches"], 1), launches=k["launches"],
               copy_probe_GBps=copy_gbs,
               frac_of_copy=round(achieved / copy_gbs, 4) if copy_gbs else None)
    if traffic:
        # the counted HBM traffic (whole 128-byte lines for the scattered
        # accesses) over the same launch time: how close the kernel runs to
        # the bandwidth a plain copy reaches
        tr = traffic / (out["avg_launch_ms"] * 1e-3) / 1e9
        out["traffic_GBps"] = round(tr, 1)
        out["traffic_frac_of_copy"] = round(tr / copy_gbs, 4) if copy_gbs else None
    # the whole run under the reference kernels' per-phase bytes: what those
    # kernels would move, over this build's MDAS phase time
    pb = res.get("phase_model_bytes")
    if pb:
        ph = res["phases_ms"]
        tot = sum(pb.values())
        gbs = tot / (res["mdas_ms"] * 1e-3) / 1e9
        out["whole_run"] = dict(
            bytes=round(tot), ms=round(res["mdas_ms"], 3), achieved=round(gbs, 1),
            frac=round(gbs / HBM_PEAK_GBS, 4),
            frac_of_copy=round(gbs / copy_gbs, 4) if copy_gbs else None,
            phases={k: dict(bytes=round(v), GBps=round(v / (ph[k + "_ms"] * 1e-3) / 1e9, 1))
                    for k, v in pb.items()},
            note="reference kernels' bytes per phase (per-leaf error/box/seed sweeps, "
                 "node/box/value reads) over this build's phase times")
    return out


def cpu_baseline(name, shrink=0, log=print, runs=3):
    """The OpenMP oracle on the host cores, MDAS phases only (integrand
    excluded, as on the GPU), the median of `runs` runs.  shrink = k runs a
    (W/2^k) x (H/2^k) version of the workload with extraImgBits - k, which
    keeps 4 initial samples per initial cell and the same spp."""
    from oracle import pyoracle
    w = dict(WORKLOADS[name])
    w["W"] >>= shrink
    w["H"] >>= shrink
    w["e"] -= shrink
    target, cap = capacity(w)
    lib = pyoracle.load()
    # threads: the CPU share this process is given for its one GPU
    # (OMP_NUM_THREADS, which the GPU machines set to 16 per GPU), else every
    # core it may run on
    visible = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or visible
    threads = max(1, min(threads, visible))
    lib.orc_set_threads(threads)
    params = pyoracle.make_scene(w["kind"], w["seed"], w["dim"], float(w["W"]), float(w["H"]))
    times = []
    for r in range(runs):
        o = pyoracle.OracleGrid(w["dim"], cap, w["cand"], w["b"], w["e"], w["alpha"],
                                float(w["W"]), float(w["H"]))
        t_mdas = 0.0
        t0 = time.perf_counter()
        o.initial_sampling()
        t_mdas += time.perf_counter() - t0
        o.eval_scene(params, 0, o.n_samples)
        passes = 0
        while o.n_samples < target:
            t0 = time.perf_counter()
            o.compute_errors()
            rc = o.adaptive_sampling()
            t_mdas += time.perf_counter() - t0
            if rc:
                raise RuntimeError("oracle capacity")
            o.eval_scene(params, o.n_samples - o.new_samples, o.new_samples)
            passes += 1
        t0 = time.perf_counter()
        o.integrate(w["W"], w["H"])
        t_mdas += time.perf_counter() - t0
        n = o.n_samples
        o.close()
        times.append(t_mdas)
        log(f"[bench] cpu baseline run {r + 1}/{runs}: {n} samples in {t_mdas:.2f} s MDAS time, "
            f"{passes} passes, {threads} threads")
    t_med = sorted(times)[len(times) // 2]
    return dict(value=round(n / t_med / 1e6, 3), unit="Msamples/s", cores=threads, kind="port",
                cpu=cpu_model(), cores_visible=visible,
                threads_from="OMP_NUM_THREADS" if os.environ.get("OMP_NUM_THREADS") else "affinity",
                sample=(f"full workload {name}" if shrink == 0 else
                        f"{w['W']}x{w['H']} version of {name} (extraImgBits {w['e']})")
                + f", {n} samples, {passes} passes, MDAS phases only, median of {runs} runs "
                  f"({', '.join(f'{t:.2f}' for t in times)} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-frames", action="store_true")
    ap.add_argument("--no-device-record", action="store_true",
                    help="skip rocminfo / rocm-smi (e.g. under a profiler whose preload "
                         "would follow them into their child processes)")
    ap.add_argument("--cpu-shrink", type=int, default=0,
                    help="run the CPU baseline on a 2^-k per-axis version of the workload")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the sampler is single-GPU; "
              "sharding by image or screen tile is a design, not built)", file=sys.stderr)
        sys.exit(2)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    # child processes, before this one initialises the GPU
    dev_rec = {} if a.no_device_record else device_record()
    log(f"[bench] device: {dev_rec}")
    copy_gbs = copy_probe()
    log(f"[bench] copy probe: {copy_gbs} GB/s")
    import pmdas_amd
    from pmdas_amd import hip
    pmdas_amd.load()
    if hip.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    res = run_gpu(a.workload, a.steps, a.warmup, log)
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{a.workload}.json")
    line = {
        "metric": "Msamples/s (build+error+sample+recon); time-to-equal-MSE learned vs plain",
        "value": round(res["msamples_per_s"], 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(res["wall_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded analytic scene)",
        "config": {"workload": a.workload, "width": WORKLOADS[a.workload]["W"],
                   "height": WORKLOADS[a.workload]["H"], "dim": WORKLOADS[a.workload]["dim"],
                   "spp": WORKLOADS[a.workload]["spp"],
                   "bits_per_dim": WORKLOADS[a.workload]["b"],
                   "extra_img_bits": WORKLOADS[a.workload]["e"],
                   "alpha": WORKLOADS[a.workload]["alpha"],
                   "candidates_num": WORKLOADS[a.workload]["cand"],
                   "parallelism": "single GPU"},
        "mdas_ms_per_step": round(res["mdas_ms"], 3),
        "recon_tile": res["recon_tile"],
        "integrand_ms_per_step": round(res["integrand_ms"], 3),
        "e2e_msamples_per_s": round(res["e2e_msamples_per_s"], 2),
        "wall_ms_untimed": round(res["wall_ms_untimed"], 3),
        "samples": res["samples"], "leaves": res["leaves"], "passes": res["passes"],
        "phases_ms": {k: round(v, 3) for k, v in res["phases_ms"].items()},
        "roofline": roofline(res, traffic_file, copy_gbs),
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 4),
                        "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)}
                    for k, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["ms"])},
        "device": (dev_rec.get("Marketing Name") or hip.device_name() or
                   f"{dev_rec.get('Name', 'unknown')}, {dev_rec.get('Compute Unit', '?')} CUs, "
                   f"{dev_rec.get('Max Clock Freq. (MHz)', '?')} MHz peak engine clock"),
        "device_record": dev_rec,
    }
    def section(key, fn):
        # a failing secondary line is recorded as such; the headline line
        # is printed regardless
        try:
            line[key] = fn()
        except Exception as e:  # noqa: BLE001 -- reported in the line itself
            log(f"[bench] {key} failed: {type(e).__name__}: {e}")
            line[key] = {"error": f"{type(e).__name__}: {e}"}

    def secondary():
        other = [k for k in WORKLOADS if k != a.workload][0]
        r2 = run_gpu(other, max(1, min(3, a.steps)), 1, log)
        return {"workload": other, "value": round(r2["msamples_per_s"], 2),
                "unit": "Msamples/s", "mdas_ms_per_step": round(r2["mdas_ms"], 3),
                "e2e_msamples_per_s": round(r2["e2e_msamples_per_s"], 2),
                "samples": r2["samples"], "passes": r2["passes"]}

    if not a.no_secondary:
        section("secondary", secondary)
    if not a.no_batch:
        section("batch", lambda: run_batch(max(1, min(2, a.steps)), log))
    if not a.no_learned:
        section("learned", lambda: learned_comparison(log))
    if not a.no_frames:
        section("frames", lambda: run_frames(log))
    if not a.no_cpu_baseline:
        section("cpu_baseline", lambda: cpu_baseline(a.workload, a.cpu_shrink, log))
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
