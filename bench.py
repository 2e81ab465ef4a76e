#!/usr/bin/env python3
"""Headline benchmark: PT+NEE at 1920x1080 on the ~1M-triangle scene.

One step = one 64-spp frame (64 reference frames of the running mean, depth 8)
of the 1,021,120-triangle scene (cupbr_amd.scenes.million_triangles), followed
by bloom (threshold 1, knee 1; the 5-level Gaussian pyramid) and the Reinhard tone map, all on the GPU with
the scene resident in HBM. Prints one JSON line:

  value        Msamples/s = 1920*1080*64 / median step time (samples = camera paths;
               ms_per_step is the mean)
  roofline     the trace kernels (closest hit for extension rays + any hit for shadow rays,
               one launch per bounce), the dominant stage, against the VALU issue ceiling
               they saturate: achieved = VALU instructions of a step's trace launches x 64
               lanes x lane utilisation (rocprofv3 counters SQ_INSTS_VALU,
               SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU), from a child run of the same
               step) / the trace launches' time (HIP events on the renderer's stream during
               the timed steps); peak = 157.3 TFLOP/s FP32 / 4 = 39.3 T lane-instructions/s;
               traffic = HBM bytes per trace launch (FETCH_SIZE + WRITE_SIZE, two more child
               runs, scaled by what the counters read for a copy of known size); the
               algorithmic bytes (nodes x 128 B + triangles x 48 B + queue records, mostly
               served by L2 and the MALL) are reported beside them, and so are the vector
               memory units' busy fractions (TA_TA_BUSY, TD_TD_BUSY), the L1 access rate
               (TCP_TOTAL_CACHE_ACCESSES) and the L1 and L2 hit rates (TCP_TCC_READ_REQ,
               TCC_HIT, TCC_MISS; a fourth child run)
  cpu_baseline the CPU oracle (the reference estimator restated in C++, SAH BVH, FP32)
               timed on this host on a bounded sample of the same workload, on this
               process's CPU share (affinity mask, bounded by a cgroup quota), with the CPU
               model and SMT state
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--cpu-spp", type=int, default=16, help="samples per pixel of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-nrc", action="store_true",
                    help="pt workload: skip the neural-radiance-cache half of the metric (the equal-time rel-MSE "
                         "ratio of --workload nrc, reported in the same JSON line under neural_cache)")
    ap.add_argument("--nrc-spp", type=int, default=4, help="samples per render call of the radiance-cache workload")
    ap.add_argument("--nrc-camera-roughness", default="0.1,inf",
                    help="comma list of camera-hit roughness thresholds of the radiance cache (inf: the camera hit "
                         "is never a cache vertex); the first one gives the reported value, the others are reported "
                         "beside it")
    ap.add_argument("--workload", choices=["pt", "cornell", "grid", "nrc", "gdpt", "volume"], default="pt",
                    help="pt: the headline PT+NEE frame; grid: one gradient evaluation (forward + path-replay "
                         "adjoint) of the 64^3 heterogeneous medium at 256x256, 64 spp; nrc: equal-time rel-MSE of "
                         "1-spp frames with the neural radiance cache against plain PT on the 1080p/1M-triangle scene; "
                         "cornell: PT+NEE of the 34-triangle Cornell box at 256x256, 16 spp, depth 5; "
                         "gdpt: gradient-domain frame (5 paths per sample + 50 reconstruction iterations) of the "
                         "Cornell box at 512x512, 128 spp; volume: the homogeneous-media renderer at 512x512, 64 spp")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the four rocprofv3 child runs that measure VALU, cache and HBM counters (pt: of "
                         "the trace kernels; cornell, gdpt, volume: of all kernels of a render call)")
    ap.add_argument("--counter-dir", default=None, help="directory for the rocprofv3 child runs' output")
    ap.add_argument("--counters-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run three device-to-device copies of MIB MiB (a plain copy kernel of known "
                         "size, for scaling the HBM byte counters of a --pmc run)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this renderer core runs on one GPU; --gpus must be 1 "
                          "(sharding samples over GPUs is not implemented)"}))
        sys.exit(2)

    import numpy as np

    import cupbr_amd as cb
    from cupbr_amd import scenes

    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    if args.counters_child:
        return counters_child(args)
    if args.copy_calibration > 0:
        import torch

        n = args.copy_calibration << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads n bytes, writes n bytes
        torch.cuda.synchronize()
        log(f"copy calibration: 3 elementwise copies of {n} bytes")
        del src, dst
    if args.workload == "grid":
        return bench_grid(args, log)
    if args.workload == "nrc":
        print(json.dumps(bench_nrc(args, log)), flush=True)
        return None
    if args.workload in ("gdpt", "volume", "cornell"):
        return bench_integrator(args, log)
    return bench_pt(args, log)


# --------------------------------------------------------------------------- host CPU description
def host_cpu():
    """The host's CPU model, SMT state and the CPU share of this process: the CPUs in its affinity
    mask, bounded by its cgroup CPU quota where one is set (the share a worker pool should use)."""
    info = {"logical_cpus": os.cpu_count() or 1}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity_cpus"] = info["logical_cpus"]
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    info["cgroup_cpu_quota"] = quota
    model, siblings, cores = None, None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "siblings" and siblings is None:
                    siblings = int(v)
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
    except OSError:
        pass
    info["model"] = model
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        if siblings and cores:
            smt = siblings > cores
    info["smt"] = smt
    info["threads_per_core"] = (siblings // cores) if siblings and cores else None
    info["threads_used"] = min(info["affinity_cpus"], quota) if quota else info["affinity_cpus"]
    return info


def cpu_baseline_record(value, cpu, sample):
    return {"value": round(value, 4), "unit": "Msamples/s", "cores": cpu["threads_used"], "kind": "port",
            "sample": sample, "cpu_model": cpu["model"], "smt": cpu["smt"],
            "threads_per_core": cpu["threads_per_core"], "logical_cpus": cpu["logical_cpus"],
            "affinity_cpus": cpu["affinity_cpus"], "cgroup_cpu_quota": cpu["cgroup_cpu_quota"]}


# --------------------------------------------------------------------------- counters (rocprofv3 children)
VALU_COUNTERS = ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVES", "GRBM_GUI_ACTIVE",
                 "TA_TA_BUSY_sum", "TD_TD_BUSY_sum", "TCP_TOTAL_CACHE_ACCESSES_sum"]
# cache hit rates: L2 (TCC) hits and misses, and the L1's (TCP) requests to L2 against its accesses
CACHE_COUNTERS = ["TCC_HIT_sum", "TCC_MISS_sum", "TCP_TCC_READ_REQ_sum", "TCP_TOTAL_CACHE_ACCESSES_sum",
                  "GRBM_GUI_ACTIVE"]
# rocprofv3 reports GRBM_GUI_ACTIVE summed over the MI355X's 8 XCDs (a dispatch of t seconds reads
# about 8 x 2.4e9 x t), so a dispatch lasts GRBM_GUI_ACTIVE / 8 cycles of each CU's clock
N_XCD = 8
VALU_FMA_ITERS = 20000
COPY_MIB = 256
# VALU issue ceiling from the published FP32 vector peak: 157.3 TFLOP/s counts packed FMAs (2 lanes
# x 2 flops per lane-instruction); this code is built without packed FP32, so one VALU instruction
# on one lane is one op: 157.3 / 4 = 39.3 T lane-instructions/s (256 CUs x 4 SIMDs x 64 lanes per
# 4 cycles x 2.4 GHz)
VALU_PEAK_TOPS = 157.3 / 4


def _read_counter_csv(d):
    """Rows of a rocprofv3 --output-format csv counter collection under directory d, as
    [(dispatch, kernel, counter, value)], in dispatch order."""
    import csv
    import glob

    files = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
    rows = []
    for fn in files:
        with open(fn, newline="") as f:
            for r in csv.DictReader(f):
                disp = r.get("Dispatch_Id") or r.get("Dispatch_ID") or r.get("Correlation_Id") or "0"
                rows.append((int(float(disp)), r.get("Kernel_Name", ""), r.get("Counter_Name", ""),
                             float(r.get("Counter_Value", "nan"))))
    rows.sort(key=lambda x: x[0])
    return rows


def _per_dispatch(rows):
    out = {}
    for disp, kern, cname, val in rows:
        e = out.setdefault(disp, {"kernel": kern})
        e[cname] = e.get(cname, 0.0) + val
    return [out[k] for k in sorted(out)]


def _run_counter_child(args, counters, log, copy=False):
    """rocprofv3 --pmc <counters> over `bench.py --counters-child` (one step of the same workload, the
    VALU microkernel and, with copy, a copy kernel of known size); returns per-dispatch counters."""
    import shutil
    import signal
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        log("counters: rocprofv3 not found")
        return None
    if args.counter_dir:
        os.makedirs(args.counter_dir, exist_ok=True)
    out_dir = tempfile.mkdtemp(prefix="cupbr_pmc_", dir=args.counter_dir or None)
    cmd = [prof, "--pmc", *counters, "--output-format", "csv", "-d", out_dir, "-o", "c", "--", sys.executable,
           os.path.join(ROOT, "bench.py"), "--counters-child", "--spp", str(args.spp), "--width", str(args.width),
           "--height", str(args.height), "--depth", str(args.depth), "--workload", args.workload] + \
        (["--copy-calibration", str(COPY_MIB)] if copy else [])
    env = dict(os.environ, TMPDIR="/tmp")
    t0 = time.time()
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, env=env, start_new_session=True)
    try:
        text = p.communicate(timeout=240)[0].decode(errors="replace")
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        log(f"counters {counters}: rocprofv3 timed out")
        return None
    rows = _read_counter_csv(out_dir)
    log(f"counters {' '.join(counters)}: rocprofv3 exit {p.returncode}, {len(rows)} rows in {time.time() - t0:.0f}s")
    if not rows:
        log(text[-2000:])
        return None
    return _per_dispatch(rows)


def is_trace_kernel(k):
    return "k_pt_camera<" in k or "k_pt_trace<false>" in k


def is_render_kernel(k):
    return "cupbr::" in k and "k_valu_peak" not in k


def measure_counters(args, log, is_trace=is_trace_kernel, n_cu=256):
    """VALU, cache and HBM counters of the dispatches whose kernel name satisfies is_trace (the
    trace kernels by default) from four rocprofv3 runs of one step each."""
    res = {}
    valu = _run_counter_child(args, VALU_COUNTERS, log)
    cache = _run_counter_child(args, CACHE_COUNTERS, log)
    fetch = _run_counter_child(args, ["FETCH_SIZE"], log, copy=True)
    write = _run_counter_child(args, ["WRITE_SIZE"], log, copy=True)
    if cache:
        tc = [d for d in cache if is_trace(d["kernel"])]
        hit, miss = sum(d.get("TCC_HIT_sum", 0.0) for d in tc), sum(d.get("TCC_MISS_sum", 0.0) for d in tc)
        acc = sum(d.get("TCP_TOTAL_CACHE_ACCESSES_sum", 0.0) for d in tc)
        if hit + miss > 0:
            res["l2_hit_rate"] = hit / (hit + miss)
        if acc > 0:
            res["l1_hit_rate"] = 1.0 - sum(d.get("TCP_TCC_READ_REQ_sum", 0.0) for d in tc) / acc
    if valu:
        tr = [d for d in valu if is_trace(d["kernel"])]
        mk = [d for d in valu if "k_valu_peak" in d["kernel"]]
        if tr:
            res["trace_dispatches"] = len(tr)
            res["valu_instructions"] = sum(d.get("SQ_INSTS_VALU", 0.0) for d in tr)
            act = sum(d.get("SQ_ACTIVE_INST_VALU", 0.0) for d in tr)
            res["lane_utilisation"] = sum(d.get("SQ_THREAD_CYCLES_VALU", 0.0) for d in tr) / max(64 * act, 1.0)
            # vector-memory address (TA) and data (TD) units, one of each per CU: busy cycles over the
            # dispatches' cycles
            cyc = sum(d.get("GRBM_GUI_ACTIVE", 0.0) for d in tr) / N_XCD * n_cu
            if cyc > 0 and any("TA_TA_BUSY_sum" in d for d in tr):
                res["ta_busy"] = sum(d.get("TA_TA_BUSY_sum", 0.0) for d in tr) / cyc
                res["td_busy"] = sum(d.get("TD_TD_BUSY_sum", 0.0) for d in tr) / cyc
                # vector L1 (TCP) accesses per CU per cycle; the L1 serves one access (one cache
                # line of one load instruction) per clock
                res["l1_access_rate"] = sum(d.get("TCP_TOTAL_CACHE_ACCESSES_sum", 0.0) for d in tr) / cyc
            res["per_dispatch"] = [{"kernel": d["kernel"].split("(")[0].replace("void ", ""),
                                    "valu_instructions": d.get("SQ_INSTS_VALU", 0.0),
                                    "lane_utilisation": d.get("SQ_THREAD_CYCLES_VALU", 0.0) /
                                    max(64 * d.get("SQ_ACTIVE_INST_VALU", 0.0), 1.0),
                                    "grbm_gui_active": d.get("GRBM_GUI_ACTIVE", 0.0)} for d in tr]
        if mk:
            d = mk[-1]
            res["micro_valu_instructions"] = d.get("SQ_INSTS_VALU", 0.0)
            res["micro_lane_utilisation"] = d.get("SQ_THREAD_CYCLES_VALU", 0.0) / max(64 * d.get("SQ_ACTIVE_INST_VALU", 0.0), 1.0)

    def bytes_of(disp, name):
        if not disp:
            return None, None
        # the copy kernels of --copy-calibration (torch elementwise kernels: a fill, then three
        # copies; the fill reads nothing, so reads are taken from the dispatches that read)
        cal = sorted(d.get(name, 0.0) for d in disp if "elementwise" in d["kernel"] and d.get(name, 0.0) > 1000)
        scale = (COPY_MIB * 1024.0) / cal[len(cal) // 2] if cal else None
        tr = [d.get(name, 0.0) for d in disp if is_trace(d["kernel"])]
        if not tr or scale is None:
            return None, scale
        return sum(tr) * 1024.0 * scale, scale

    fb, res["fetch_scale"] = bytes_of(fetch, "FETCH_SIZE")
    wb, res["write_scale"] = bytes_of(write, "WRITE_SIZE")
    n = max(res.get("trace_dispatches", 0), 1)
    res["fetch_bytes_per_launch"] = None if fb is None else fb / n
    res["write_bytes_per_launch"] = None if wb is None else wb / n
    res["fetch_bytes_total"], res["write_bytes_total"] = fb, wb
    return res


def counters_child(args):
    """The work profiled by measure_counters: the VALU microkernel, an optional copy of known size,
    then one step of the headline workload (no warm-up, no counting pass)."""
    import cupbr_amd as cb
    from cupbr_amd import scenes

    cb.measure_valu_peak(VALU_FMA_ITERS)
    work = None
    if args.workload in ("cornell", "gdpt", "volume"):
        work = integrator_setup(args)  # scene build and upload before the copy and the render call
    if args.copy_calibration > 0:
        import torch

        n = args.copy_calibration << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    if work is not None:
        work["step"]()
        return
    W, H = args.width, args.height
    desc, ca = scenes.million_triangles(W, H)
    scene = cb.Scene.from_desc(desc)
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], W, H)
    r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
    r.set_output_size(W, H)
    r.register_scene(scene)
    r.set_max_trace_depth(args.depth)
    post = cb.PostProcessor()
    post.register_renderer(r)
    post.set_bloom_pyramid(cb.BloomPyramid.GAUSSIAN5)
    tm = cb.ToneMapper(cb.ToneMappingType.REINHARD)
    tm.register_renderer(r)
    r.render(cam, args.spp)
    r.synchronize()
    tm.tone_map_bloom(post, cb.threshold_curve(1.0, 1.0))
    tm.get_render_buffer()


# --------------------------------------------------------------------------- headline workload
def bench_pt(args, log):
    import numpy as np
    import torch

    import cupbr_amd as cb
    from cupbr_amd import scenes

    W, H, SPP = args.width, args.height, args.spp
    t0 = time.time()
    desc, cam_args = scenes.million_triangles(W, H)
    t_gen = time.time() - t0
    scene = cb.Scene.from_desc(desc)
    scene.upload()
    bvh = scene.bvh_stats()
    log(f"scene: {len(desc['tris'])} triangles, {len(desc['types'])} objects; generated in {t_gen:.1f}s, "
        f"BVH4 build {bvh['build_ms']:.0f} ms + upload {bvh['upload_ms']:.0f} ms ({bvh['nodes']} nodes, depth {bvh['depth']})")
    cam = cb.Camera.look_at(cam_args["eye"], cam_args["target"], cam_args["up"], cam_args["vfov"], W, H)

    r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
    r.set_output_size(W, H)
    r.register_scene(scene)
    r.set_max_trace_depth(args.depth)
    r.set_russian_roulette(False)
    post = cb.PostProcessor()
    post.register_renderer(r)
    post.set_bloom_pyramid(cb.BloomPyramid.GAUSSIAN5)  # the 5-level Gaussian pyramid the workload names
    tm = cb.ToneMapper(cb.ToneMappingType.REINHARD)
    tm.register_renderer(r)
    curve = cb.threshold_curve(1.0, 1.0)

    def step():
        r.reset()  # frameIndex = 0: each step is a fresh 64-spp frame (PBRenderer::reset)
        r.render(cam, SPP)
        r.synchronize()
        tm.tone_map_bloom(post, curve)

    # instrumented pass (not timed): traversal visit counts of one step
    r.set_traversal_counting(True)
    step()
    sc = r.stats()
    r.set_traversal_counting(False)
    ext_rays = sc["extension_rays"]
    bytes_per_ext_ray = (sc["extension_node_visits"] * 128 + sc["extension_prim_tests"] * 48 + ext_rays * 40) / max(ext_rays, 1)
    sh_rays = sc["shadow_rays"]
    bytes_per_sh_ray = (sc["shadow_node_visits"] * 128 + sc["shadow_prim_tests"] * 48 + sh_rays * 48) / max(sh_rays, 1)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    r.set_profiling(True)
    stage_ms = np.zeros(4)
    stage_n = np.zeros(4)
    rays = np.zeros(2)
    step_s = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        step_s.append(time.perf_counter() - t0)
        st = r.stats()
        stage_ms += st["stage_ms"]
        stage_n += st["stage_launches"]
        rays += (st["extension_rays"], st["shadow_rays"])
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    r.set_profiling(False)
    tm_buf = tm.get_render_buffer()

    samples_per_step = W * H * SPP
    t_med = float(np.median(step_s))
    value = samples_per_step / t_med / 1e6  # median step (BASELINE: median of the timed runs)
    value_mean = samples_per_step * args.steps / t_total / 1e6
    trace_launch_ms = stage_ms[0] / max(stage_n[0], 1)
    trace_ms_per_step = stage_ms[0] / args.steps
    alg_bytes_per_launch = (bytes_per_ext_ray * rays[0] + bytes_per_sh_ray * rays[1]) / max(stage_n[0], 1)
    alg_gbs = alg_bytes_per_launch / (trace_launch_ms * 1e-3) / 1e9
    img = r.get_output_image()
    log(f"GPU: {value:.1f} Msamples/s (median step {t_med * 1e3:.1f} ms; mean over the timed steps {value_mean:.1f}), "
        f"{(rays.sum()) / t_total / 1e6:.0f} Mrays/s (ext {rays[0] / args.steps / 1e6:.1f}M + shadow "
        f"{rays[1] / args.steps / 1e6:.1f}M per step); stage ms/step trace {trace_ms_per_step:.1f} shade "
        f"{stage_ms[1] / args.steps:.1f} other {stage_ms[3] / args.steps:.1f}; image mean {img.mean():.4f} finite "
        f"{bool(np.isfinite(img).all())}; rgba mean {tm_buf[..., :3].mean():.1f}")
    log(f"extend: {bytes_per_ext_ray:.0f} algorithmic B/ray ({sc['extension_node_visits'] / max(ext_rays, 1):.1f} nodes, "
        f"{sc['extension_prim_tests'] / max(ext_rays, 1):.1f} tris per ray); shadow {bytes_per_sh_ray:.0f} B/ray "
        f"({sc['shadow_node_visits'] / max(sh_rays, 1):.1f} nodes, {sc['shadow_prim_tests'] / max(sh_rays, 1):.1f} tris); "
        f"trace kernel {trace_launch_ms:.3f} ms/launch, {alg_gbs:.0f} GB/s of algorithmic bytes (L2/MALL-served)")

    # not part of `value`: reading the 1080p FP32 image back to the host over PCIe
    t0 = time.perf_counter()
    for _ in range(5):
        r.get_output_image()
    t_read = (time.perf_counter() - t0) / 5
    log(f"host read-back of the {W}x{H} FP32 image: {t_read * 1e3:.2f} ms ({W * H * 12 / t_read / 1e9:.1f} GB/s); "
        f"with one read-back per step the rate would be {samples_per_step / (t_med + t_read) / 1e6:.1f} Msamples/s")

    # VALU ceiling: the microkernel timed here (unprofiled); its instruction count is checked by the
    # counters below
    micro = cb.measure_valu_peak(VALU_FMA_ITERS)
    log(f"VALU microkernel: {micro['wave_instructions'] / 1e9:.2f} G wave64 FMAs in {micro['ms']:.3f} ms = "
        f"{micro['wave_instructions_per_s'] * 64 / 1e12:.2f} T lane-instructions/s")
    del r, post, tm

    roof = {"bound": "valu", "achieved": None, "peak": round(VALU_PEAK_TOPS, 3), "unit": "TOP/s", "frac": None,
            "traffic": None,
            "kernel": "k_pt_camera + k_pt_trace<false> (all trace launches of a step: bounce 0 camera rays made and traced as "
                      "wave packets, then per bounce extension closest hit + previous bounce's shadow any hit, BVH4)",
            "op": "one VALU instruction on one active lane",
            "avg_launch_ms": round(trace_launch_ms, 4), "trace_ms_per_step": round(trace_ms_per_step, 3),
            "micro_valu_tops": round(micro["wave_instructions_per_s"] * 64 / 1e12, 3),
            "algorithmic_bytes_per_launch": int(alg_bytes_per_launch), "algorithmic_gbs": round(alg_gbs, 1),
            "bytes_per_ext_ray": round(bytes_per_ext_ray, 1), "bytes_per_shadow_ray": round(bytes_per_sh_ray, 1)}
    if not args.no_counters:
        c = measure_counters(args, log, n_cu=torch.cuda.get_device_properties(0).multi_processor_count)
        if c.get("valu_instructions"):
            t_tr = trace_ms_per_step * 1e-3  # the child's dispatches are one step's trace launches
            issue = c["valu_instructions"] / t_tr  # wave instructions per second
            lane = c["lane_utilisation"]
            achieved = issue * 64 * lane / 1e12
            roof.update({"achieved": round(achieved, 3), "frac": round(achieved / VALU_PEAK_TOPS, 4),
                         "lane_utilisation": round(lane, 4),
                         "valu_issue_of_peak": round(issue * 64 / 1e12 / VALU_PEAK_TOPS, 4),
                         "valu_instructions_per_step": int(c["valu_instructions"]),
                         "trace_dispatches": c["trace_dispatches"]})
            if c.get("micro_valu_instructions"):
                roof["micro_counted_instructions"] = int(c["micro_valu_instructions"])
                roof["micro_lane_utilisation"] = round(c["micro_lane_utilisation"], 4)
                rate = c["micro_valu_instructions"] / (micro["ms"] * 1e-3)
                roof["micro_valu_tops"] = round(rate * 64 / 1e12, 3)
                roof["valu_issue_of_micro"] = round(issue / rate, 4)
            log(f"trace VALU: {c['valu_instructions'] / 1e9:.1f} G wave instructions per step, lane utilisation "
                f"{lane:.3f}, issue {issue * 64 / 1e12:.2f} T lane-slots/s = {issue * 64 / 1e12 / VALU_PEAK_TOPS:.3f} "
                f"of peak; useful {achieved:.2f} TOP/s = {achieved / VALU_PEAK_TOPS:.3f} of peak")
        if c.get("ta_busy") is not None:
            roof.update({"vmem_address_unit_busy": round(c["ta_busy"], 4), "vmem_data_unit_busy": round(c["td_busy"], 4),
                         "l1_accesses_per_cu_cycle": round(c["l1_access_rate"], 4)})
            log(f"trace vector memory: address unit (TA) busy {c['ta_busy']:.3f}, data unit (TD) busy "
                f"{c['td_busy']:.3f}, L1 accesses {c['l1_access_rate']:.3f} per CU per cycle, of the trace "
                f"dispatches' cycles under the profiler")
        if c.get("l2_hit_rate") is not None:
            roof.update({"l2_hit_rate": round(c["l2_hit_rate"], 4), "l1_hit_rate": round(c.get("l1_hit_rate", 0.0), 4)})
            log(f"trace caches: L2 hit rate {c['l2_hit_rate']:.3f} (TCC_HIT / (TCC_HIT + TCC_MISS)), L1 hit rate "
                f"{c.get('l1_hit_rate', 0.0):.3f} (1 - TCP_TCC_READ_REQ / TCP_TOTAL_CACHE_ACCESSES)")
        if c.get("fetch_bytes_per_launch") is not None and c.get("write_bytes_per_launch") is not None:
            traffic = c["fetch_bytes_per_launch"] + c["write_bytes_per_launch"]
            hbm_gbs = traffic / (trace_launch_ms * 1e-3) / 1e9
            roof.update({"traffic": int(traffic), "hbm_gbs": round(hbm_gbs, 1),
                         "hbm_of_peak": round(hbm_gbs / HBM_PEAK_GBS, 4),
                         "fetch_scale": round(c["fetch_scale"], 4), "write_scale": round(c["write_scale"], 4)})
            log(f"trace HBM traffic: {traffic / 1e9:.2f} GB per launch (FETCH_SIZE x {c['fetch_scale']:.3f} + "
                f"WRITE_SIZE x {c['write_scale']:.3f}) = {hbm_gbs:.0f} GB/s = {hbm_gbs / HBM_PEAK_GBS:.3f} of HBM peak")

    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        hc = host_cpu()
        threads = hc["threads_used"]
        t0 = time.time()
        ref = oracle.OracleScene(desc, precisions=(32,))
        t_ref_build = time.time() - t0
        t0 = time.perf_counter()
        ref.render(cam.as_dict(), W, H, 0, args.cpu_spp, depth=args.depth, precision=32, mode=2, threads=threads)
        t_cpu = time.perf_counter() - t0
        cpu_value = W * H * args.cpu_spp / t_cpu / 1e6
        cpu = cpu_baseline_record(
            cpu_value, hc, f"{args.cpu_spp} spp of the same {W}x{H} frame ({W * H * args.cpu_spp} paths, depth "
                           f"{args.depth}) on the C++ oracle with its own SAH BVH, FP32, {threads} threads (this "
                           f"process's CPU share), {t_cpu:.1f}s (scene build {t_ref_build:.1f}s not timed)")
        log(f"CPU oracle: {cpu_value:.3f} Msamples/s on {threads} threads of {hc['model']} (logical CPUs "
            f"{hc['logical_cpus']}, affinity {hc['affinity_cpus']}, quota {hc['cgroup_cpu_quota']}, SMT {hc['smt']}); "
            f"GPU/CPU = {value / cpu_value:.0f}x")

    out = {
        "metric": "Msamples/s at 1080p/1M-tri",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, cupbr_amd.scenes.million_triangles)",
        "config": {
            "workload": f"PT+NEE {W}x{H}, {SPP} spp, depth {args.depth}, {len(desc['tris'])} triangles "
                        f"({len(desc['types'])} mesh objects), 4 area lights + 512x256 env map, then bloom "
                        f"(5-level Gaussian pyramid) + Reinhard tone map",
            "width": W, "height": H, "spp": SPP, "max_depth": args.depth, "triangles": int(len(desc["tris"])),
            "russian_roulette": False, "rng": "reference TEA<4> seed + LCG",
            "value_from": "median step time", "median_step_ms": round(t_med * 1e3, 3),
            "value_mean_msamples": round(value_mean, 3),
            "mrays_per_s": round(float(rays.sum()) / t_total / 1e6, 1),
            "bvh_build_ms": round(bvh["build_ms"], 1),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if not args.no_nrc:
        # the metric's second half, on the same scene in the same run (after the Msamples/s line's own
        # buffers are released): the neural radiance cache's equal-time rel-MSE against plain PT
        del scene
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        n = bench_nrc(args, log)
        cfg = n["config"]
        out["neural_cache"] = {
            "metric": n["metric"], "value": n["value"], "unit": n["unit"], "higher_is_better": True,
            "workload": cfg["workload"], "cache_vertices": cfg["cache_vertices"],
            "rel_mse_cache": cfg["rel_mse_cache"], "cache_call_ms": cfg["cache_call_ms"],
            "pt_equal_time": cfg["pt_equal_time"], "second_vertex_cache": cfg["second_vertex_cache"],
            "roofline": n["roofline"]}
    print(json.dumps(out), flush=True)


def bench_grid(args, log):
    """One step = forward render of the 64^3 medium at 256x256, 64 spp, and the path-replay adjoint of
    L = sum (I - T)^2 against a target rendered from the perturbed density (untimed). The adjoint
    takes each path's radiance from the forward render (cupbr_grid_adjoint_from_forward_device), as
    the torch.autograd wrapper does, so every path is replayed once in it."""
    import numpy as np
    import torch

    import cupbr_amd as cb
    from cupbr_amd import scenes

    W = H = 256
    SPP, DEPTH, RES = 64, 6, 64
    G = scenes.grid_medium_scene(res=RES, width=W, height=H)
    ca = G["camera"]
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], W, H)
    scene = cb.Scene.from_desc(G["desc"])
    scene.upload()
    R = cb.GridRenderer(scene, cam, cb.GridMedium.from_dict(G["density"].shape, G["medium"]), W, H, spp=SPP,
                        max_depth=DEPTH)
    rho = torch.from_numpy(G["density"]).cuda()
    alb = torch.from_numpy(G["albedo"]).cuda()
    target = R.render(torch.from_numpy(G["target_density"]).cuda(), alb)

    def step():
        img = R.render(rho, alb)
        return R.adjoint(rho, alb, 2 * (img - target), from_forward=True)

    R.set_counting(True)
    step()
    cnt = R.counts()
    R.set_counting(False)
    # per adjoint launch: the grid reads of the replay (the step's reads minus the forward render's)
    R.set_counting(True)
    img = R.render(rho, alb)
    fwd = R.counts()
    R.set_counting(False)
    adj_reads = cnt["grid_reads"] - fwd["grid_reads"]
    adj_bytes = adj_reads * 32 + cnt["atomics"] * 8  # 8 voxels x 4 B per read; read-modify-write per atomic

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    adj_ms = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        img = R.render(rho, alb)
        d = 2 * (img - target)
        ev[k][0].record()
        R.adjoint(rho, alb, d, from_forward=True)
        ev[k][1].record()
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    adj_ms = [a.elapsed_time(b) for a, b in ev]
    adj_launch_ms = float(np.mean(adj_ms))
    samples = W * H * SPP * args.steps
    value = samples / t_total / 1e6
    alg_gbs = adj_bytes / (adj_launch_ms * 1e-3) / 1e9
    atomic_rate = cnt["atomics"] / (adj_launch_ms * 1e-3)  # the adjoint's gradient atomics per second
    log(f"grid: {value:.1f} Msamples/s (forward + adjoint), {t_total / args.steps * 1e3:.1f} ms/step, adjoint "
        f"{adj_launch_ms:.2f} ms; per step {cnt['closest_rays'] / 1e6:.1f}M closest + {cnt['shadow_rays'] / 1e6:.1f}M "
        f"shadow rays, {cnt['grid_reads'] / 1e6:.1f}M grid reads, {cnt['atomics'] / 1e6:.1f}M atomics")
    # the ceiling: FP32 atomics without return to scattered words of 8 per-XCD copies of the grid
    micro = cb.measure_atomic_peak(256, RES ** 3)
    frac = atomic_rate / micro["atomics_per_s"]
    log(f"atomics: adjoint {atomic_rate / 1e9:.2f} G/s; scattered-atomic microkernel {micro['atomics_per_s'] / 1e9:.2f} "
        f"G/s ({micro['atomics'] / 1e6:.0f}M in {micro['ms']:.2f} ms) -> {frac:.3f} of it")

    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        hc = host_cpu()
        threads = hc["threads_used"]
        S = oracle.OracleScene(G["desc"], precisions=(32,))
        sw = sh = 256  # the full frame
        cam_s = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], sw, sh)
        reps = 5  # gradient evaluations with distinct seeds, ~10 s of CPU work on 16 threads
        t0 = time.perf_counter()
        for r in range(reps):
            im = S.grid(cam_s.as_dict(), G["medium"], G["density"], G["albedo"], sw, sh, frame0=r * SPP, spp=SPP,
                        depth=DEPTH, precision=32, threads=threads)
            S.grid(cam_s.as_dict(), G["medium"], G["density"], G["albedo"], sw, sh, frame0=r * SPP, spp=SPP,
                   depth=DEPTH, precision=32, threads=threads, d_image=2 * im)  # cotangent of sum I^2
        t_cpu = time.perf_counter() - t0
        cpu_value = sw * sh * SPP * reps / t_cpu / 1e6
        cpu = cpu_baseline_record(cpu_value, hc, f"{reps} forward + adjoint evaluations of the same medium at "
                                                 f"{sw}x{sh}, {SPP} spp, frames 0..{reps * SPP - 1}, on the C++ oracle "
                                                 f"(FP32), {threads} threads, {t_cpu:.1f}s")
        log(f"CPU oracle: {cpu_value:.3f} Msamples/s on {threads} threads; GPU/CPU = {value / cpu_value:.0f}x")

    out = {
        "metric": "Msamples/s (grid medium forward + adjoint)",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (fBm density, cupbr_amd.scenes.grid_medium_scene)",
        "config": {"workload": f"heterogeneous medium {RES}^3 in a 10-triangle room, {W}x{H}, {SPP} spp, depth "
                               f"{DEPTH}: forward render + path-replay adjoint of an L2 image loss",
                   "width": W, "height": H, "spp": SPP, "grid": RES, "max_depth": DEPTH},
        "roofline": {"bound": "l2_atomics", "achieved": round(atomic_rate / 1e9, 3),
                     "peak": round(micro["atomics_per_s"] / 1e9, 3), "unit": "G atomics/s",
                     "frac": round(frac, 4), "traffic": None,
                     "peak_source": "scattered FP32 atomic adds without return into 8 per-XCD copies of the "
                                    f"{RES}^3 grid (cupbr_measure_atomic_peak), measured in this run",
                     "kernel": "k_grid<true> (path-replay adjoint)", "atomics_per_launch": int(cnt["atomics"]),
                     "algorithmic_bytes_per_launch": int(adj_bytes), "algorithmic_gbs": round(alg_gbs, 1),
                     "avg_launch_ms": round(adj_launch_ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_nrc(args, log):
    """Equal-time comparison on the 1080p / 1M-triangle scene (depth 8).

    reference  plain PT, 4096 spp, FP64-accumulated: each frame f is rendered alone (the running mean
               restarted at frame index f from a zero image, so the image is sample_f / (f + 1)) and
               its samples are summed in FP64
    cache      online training over 300 warm-up frames in calls of s frames (--nrc-spp, default 4;
               2^16 training paths per frame, 16 Adam steps of 2^14 rows after each call), then
               `steps` timed calls of s spp from frame 0 (inference, training rows and the call's
               training included)
    plain PT   timed and rendered at 1..12 spp (one call each); equal time = the cache call's time T:
               value = rel-MSE(plain PT at the smallest spp whose measured time is >= T) /
               rel-MSE(cache), i.e. plain PT is given at least the cache's time; the ratio against
               the largest spp with time <= T is reported beside it
    ceiling    the squared error of the primary vertex's direct light alone (plain PT of depth 1
               against its own FP64 reference, normalised by the full reference as the metric is):
               a cache queried from the second vertex on cannot remove it, so a cached sample has at
               least that error, and the equal-time ratio is at most
               rel-MSE(PT, 1 spp) / cost(PT sample) x cost(cached sample) / error(depth 1, 1 spp)
    roofline   the inference kernel (matrix cores): 2 * 16,576 flops per query / launch time"""
    import numpy as np
    import torch

    import cupbr_amd as cb
    from cupbr_amd import scenes

    W, H, DEPTH = args.width, args.height, args.depth
    S_NRC = max(1, args.nrc_spp)
    desc, ca = scenes.million_triangles(W, H)
    scene = cb.Scene.from_desc(desc)
    scene.upload()
    cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], W, H)

    def renderer(cache=None, depth=DEPTH):
        r = cb.PBRenderer(cb.RenderingMethod.PATHTRACER)
        r.set_output_size(W, H)
        r.register_scene(scene)
        r.set_max_trace_depth(depth)
        if cache is not None:
            r.set_radiance_cache(cache, training_paths=1 << 16, train_steps=16, batch=1 << 14, learning_rate=1e-3)
        return r

    def rel_mse(x, ref):
        return float(np.mean((x.astype(np.float64) - ref) ** 2 / (ref ** 2 + 1e-2)))

    def frame_ms(r, spp):
        r.reset()
        r.synchronize()
        t0 = time.perf_counter()
        r.render(cam, spp)
        r.synchronize()
        return (time.perf_counter() - t0) * 1e3

    def fp64_reference(depth, n):
        """mean of frames 0..n-1 summed in FP64, each frame's samples recovered exactly: from a zero
        image at frame index f the running mean is a * L with a = 1/(f + 1) in FP32 (PathTracer.cu:162-168)"""
        R = renderer(depth=depth)
        zero = np.zeros((H, W, 3), np.float32)
        acc = np.zeros((H, W, 3), np.float64)
        for f in range(n):
            R.resume(zero, f)
            R.render(cam, 1)
            img = R.get_output_image().astype(np.float64)
            acc += img if f == 0 else img / float(np.float32(1.0) / (np.float32(f) + np.float32(1.0)))
        return acc / n

    t0 = time.perf_counter()
    ref = fp64_reference(DEPTH, 4096)
    log(f"reference: 4096 spp, FP64-accumulated frame by frame, in {time.perf_counter() - t0:.1f}s")

    P = renderer()
    frame_ms(P, 8)
    pt_spp = list(range(1, 13))
    pt_ms, pt_err = [], []
    for k in pt_spp:
        pt_ms.append(float(np.median([frame_ms(P, k) for _ in range(3)])))
        pt_err.append(rel_mse(P.get_output_image(), ref))
    log("plain PT: " + ", ".join(f"{k} spp {t:.2f} ms {e:.4f}" for k, t, e in zip(pt_spp, pt_ms, pt_err)))

    warm = 300

    def run_cache(camera_min_roughness):
        cache = cb.RadianceCache((-10.0, -10.0, -10.0), (110.0, 110.0, 110.0), seed=1)
        N = renderer(cache)
        N.set_radiance_cache_camera_hit(camera_min_roughness)
        t0 = time.perf_counter()
        for _ in range(warm // S_NRC):
            N.render(cam, S_NRC)
        N.synchronize()
        t_warm = time.perf_counter() - t0
        times, errs = [], []
        for _ in range(args.steps):
            times.append(frame_ms(N, S_NRC))  # frames 0 .. S_NRC - 1 each time
            errs.append(rel_mse(N.get_output_image(), ref))
        st = N.stats()
        t_c, e_c = float(np.median(times)), float(np.mean(errs))
        # the running mean of `steps` successive calls (frames 0 .. steps * S_NRC - 1): its error is
        # the variance of one call divided by the number of calls plus the cache's bias, which more
        # samples do not remove
        N.reset()
        for _ in range(args.steps):
            N.render(cam, S_NRC)
        e_mean = rel_mse(N.get_output_image(), ref)
        hi = next((i for i, t in enumerate(pt_ms) if t >= t_c), len(pt_ms) - 1)
        lo = max(0, hi - 1) if pt_ms[hi] > t_c else hi
        res = {"camera_min_roughness": camera_min_roughness, "cache_call_ms": t_c, "rel_mse_cache": e_c,
               "rel_mse_running_mean_of_calls": e_mean,
               "queries": int(st["cache_queries"]), "training_rows": int(st["cache_training_rows"]),
               "hi": hi, "lo": lo, "ratio": pt_err[hi] / e_c, "ratio_lo": pt_err[lo] / e_c, "cache": cache}
        log(f"cache (camera hit: {'never' if camera_min_roughness == float('inf') else f'roughness >= {camera_min_roughness}'}): "
            f"{warm} training frames in calls of {S_NRC} in {t_warm:.1f}s; {S_NRC} spp in {t_c:.2f} ms, rel-MSE "
            f"{e_c:.4f} ({res['queries']} queries, {res['training_rows']} training rows); plain PT at {pt_spp[hi]} spp "
            f"({pt_ms[hi]:.2f} ms, at least equal time): {pt_err[hi]:.4f}, ratio {res['ratio']:.2f}; at {pt_spp[lo]} "
            f"spp ({pt_ms[lo]:.2f} ms): {pt_err[lo]:.4f}, ratio {res['ratio_lo']:.2f}; running mean of {args.steps} "
            f"successive calls ({args.steps * S_NRC} spp): rel-MSE {e_mean:.4f}")
        return res

    modes = [float(x) for x in args.nrc_camera_roughness.split(",")]
    runs = [run_cache(m) for m in modes]
    main_run = runs[0]
    second = next((r_ for r_ in runs if r_["camera_min_roughness"] == float("inf")), None) or run_cache(float("inf"))
    cache, st = main_run["cache"], {"cache_queries": main_run["queries"]}
    t_nrc, e_nrc, hi, lo = main_run["cache_call_ms"], main_run["rel_mse_cache"], main_run["hi"], main_run["lo"]
    ratio, ratio_lo = main_run["ratio"], main_run["ratio_lo"]
    b_fit, a_fit = np.polyfit(pt_spp, pt_ms, 1)

    # variance decomposition: the primary vertex's direct light alone (depth 1)
    ref1 = fp64_reference(1, 1024)
    D1 = renderer(depth=1)
    t_d1 = float(np.median([frame_ms(D1, 1) for _ in range(5)]))
    # the direct part's squared error, normalised as the metric normalises the full image's
    e_d1 = float(np.mean((D1.get_output_image().astype(np.float64) - ref1) ** 2 / (ref ** 2 + 1e-2)))
    e_pt1 = pt_err[0]
    c_nrc = second["cache_call_ms"] / S_NRC  # per cached sample of the second-vertex cache, training included
    c_pt = b_fit  # marginal cost of a plain PT sample
    bound = (e_pt1 / c_pt * c_nrc) / e_d1 if e_d1 > 0 else float("nan")
    log(f"decomposition: plain PT 1 spp rel-MSE {e_pt1:.4f}, of which the primary vertex's direct light (depth 1, "
        f"{t_d1:.2f} ms) {e_d1:.4f} ({e_d1 / e_pt1:.0%}); a cache from the second vertex on, at {c_nrc:.2f} ms per "
        f"sample against {c_pt:.2f} ms per plain sample, is bounded by an equal-time ratio of {bound:.2f} (measured "
        f"{second['ratio']:.2f}); only a cache vertex at the camera hit can go past it")

    # inference kernel alone, timed with HIP events on torch's stream
    q = int(st["cache_queries"])
    x = torch.rand(max(q, 1 << 20), 16, device="cuda")
    x[:, 3:6] = torch.nn.functional.normalize(x[:, 3:6] - 0.5, dim=1)
    x[:, 6:9] = torch.nn.functional.normalize(x[:, 6:9] - 0.5, dim=1)
    cache.infer(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        cache.infer(x)
    e1.record()
    torch.cuda.synchronize()
    inf_ms = e0.elapsed_time(e1) / 10
    tflops = 2 * 16576 * x.shape[0] / (inf_ms * 1e-3) / 1e12
    log(f"inference: {x.shape[0]} queries in {inf_ms:.3f} ms = {tflops:.1f} TFLOP/s (BF16 matrix cores)")
    out = {
        "metric": "equal-time rel-MSE ratio, plain PT / neural radiance cache (1080p/1M-tri)",
        "value": round(ratio, 3),
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": round(t_nrc, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded procedural scene, cupbr_amd.scenes.million_triangles)",
        "config": {"workload": f"PT+NEE {W}x{H}, depth {DEPTH}, {S_NRC} spp per call with a 64->4x64->3 radiance "
                               f"cache trained online after each call; plain PT at (at least) equal time, measured; "
                               f"reference 4096 spp FP64-accumulated",
                   "cache_vertices": "the camera hit when Lambertian or GGX with roughness >= "
                                     f"{main_run['camera_min_roughness']}, and every diffuse/glossy vertex after it"
                                     if main_run["camera_min_roughness"] != float("inf") else
                                     "every diffuse/glossy vertex after the camera hit",
                   "cache_spp": S_NRC, "rel_mse_cache": round(e_nrc, 5), "cache_call_ms": round(t_nrc, 3),
                   "pt_equal_time": {"spp": pt_spp[hi], "ms": round(pt_ms[hi], 3), "rel_mse": round(pt_err[hi], 5)},
                   "pt_less_time": {"spp": pt_spp[lo], "ms": round(pt_ms[lo], 3), "rel_mse": round(pt_err[lo], 5),
                                    "ratio": round(ratio_lo, 3)},
                   "pt_measured": {str(k_): [round(t_, 3), round(e_, 5)] for k_, t_, e_ in zip(pt_spp, pt_ms, pt_err)},
                   "pt_ms_fit": [round(float(a_fit), 3), round(float(b_fit), 3)],
                   "direct_only_rel_mse_1spp": round(e_d1, 5), "pt_rel_mse_1spp": round(e_pt1, 5),
                   "second_vertex_cache": {"ratio": round(second["ratio"], 3),
                                           "rel_mse_cache": round(second["rel_mse_cache"], 5),
                                           "cache_call_ms": round(second["cache_call_ms"], 3),
                                           "ceiling_ratio": round(bound, 3)},
                   "modes": [{"camera_min_roughness": None if r_["camera_min_roughness"] == float("inf") else
                              r_["camera_min_roughness"], "ratio": round(r_["ratio"], 3),
                              "rel_mse_cache": round(r_["rel_mse_cache"], 5),
                              "rel_mse_running_mean_of_calls": round(r_["rel_mse_running_mean_of_calls"], 5),
                              "cache_call_ms": round(r_["cache_call_ms"], 3)} for r_ in runs]},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 5), "traffic": None, "kernel": "k_nrc_infer<false>",
                     "avg_launch_ms": round(inf_ms, 4), "queries_per_launch": int(x.shape[0])},
        "cpu_baseline": None,
    }
    return out


def integrator_setup(args):
    """Scene, camera and renderer of the Cornell-box path tracer (256x256, 16 spp, depth 5), the
    gradient-domain workload (Cornell box 512x512, 128 spp, depth 5; every sample traces the base
    path and 4 neighbour paths on the wavefront queues, then 50 reconstruction iterations) or the
    volume workload (homogeneous media: two scattering spheres in a glass-walled box and a scene
    medium, 512x512, 64 spp, depth 8); step() = reset + one render call."""
    import cupbr_amd as cb
    from cupbr_amd import scenes

    W = H = 512
    if args.workload == "cornell":
        W = H = 256
        desc, ca = scenes.cornell_box(W, H)
        cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], W, H)
        method, spp, depth, paths = cb.RenderingMethod.PATHTRACER, 16, 5, 1
    elif args.workload == "gdpt":
        desc, ca = scenes.cornell_box(W, H)
        cam = cb.Camera.look_at(ca["eye"], ca["target"], ca["up"], ca["vfov"], W, H)
        method, spp, depth, paths = cb.RenderingMethod.GRADIENTDOMAIN, 128, 5, 5
    else:
        desc = scenes.volume_spheres()
        cam = cb.Camera(W, H)
        method, spp, depth, paths = cb.RenderingMethod.VOLUME, 64, 8, 1
    scene = cb.Scene.from_desc(desc)
    scene.upload()
    r = cb.PBRenderer(method)
    r.set_output_size(W, H)
    r.register_scene(scene)
    r.set_method(method)
    r.set_max_trace_depth(depth)
    if args.workload == "cornell":
        r.set_rng(cb.RNG_PCG)  # configuration C0 names a seeded PCG generator

    def step():
        r.reset()
        r.render(cam, spp)
        r.synchronize()

    return {"desc": desc, "cam": cam, "scene": scene, "renderer": r, "W": W, "H": H, "spp": spp, "depth": depth,
            "paths": paths, "step": step}


def bench_integrator(args, log):
    """The Cornell-box path tracer, the gradient-domain and the volume workloads (integrator_setup).
    Roofline: VALU issue of all kernels of a render call, from counters of one render call in
    rocprofv3 child runs, over the call's device time (HIP events on the renderer's stream); the
    algorithmic BVH bytes of all traversals (counted in an instrumented pass) are reported beside
    it: these scenes' BVHs stay in L1/L2."""
    import numpy as np
    import torch

    import oracle

    wk = integrator_setup(args)
    desc, cam, r, step = wk["desc"], wk["cam"], wk["renderer"], wk["step"]
    W, H, spp, depth, paths = wk["W"], wk["H"], wk["spp"], wk["depth"], wk["paths"]

    r.set_traversal_counting(True)
    step()
    sc = r.stats()
    r.set_traversal_counting(False)
    bytes_step = (sc["extension_node_visits"] + sc["shadow_node_visits"]) * 128 + \
                 (sc["extension_prim_tests"] + sc["shadow_prim_tests"]) * 48
    for _ in range(args.warmup):
        step()
    dev_ms = 0.0
    step_s = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        step()
        step_s.append(time.perf_counter() - t1)
        dev_ms += r.stats()["milliseconds"]
    t_total = time.perf_counter() - t0
    t_med = float(np.median(step_s))
    value = W * H * spp / t_med / 1e6  # median step
    alg_gbs = bytes_step * args.steps / (dev_ms * 1e-3) / 1e9
    dev_step_ms = dev_ms / args.steps
    rays = sc["extension_rays"] + sc["shadow_rays"]
    log(f"{args.workload}: {value:.1f} Msamples/s ({paths} paths per sample; median step {t_med * 1e3:.2f} ms), "
        f"device {dev_step_ms:.2f} ms/step, {rays / t_med / 1e6:.0f} Mrays/s, "
        f"{bytes_step / max(rays, 1):.0f} algorithmic B/ray")
    kernel = {"gdpt": "gradient-domain wavefront + reconstruction (all kernels of a render call)",
              "volume": "volume wavefront (all kernels of a render call)",
              "cornell": "path tracer wavefront (all kernels of a render call)"}[args.workload]
    roof = {"bound": "valu", "achieved": None, "peak": round(VALU_PEAK_TOPS, 3), "unit": "TOP/s", "frac": None,
            "traffic": None, "kernel": kernel, "op": "one VALU instruction on one active lane",
            "device_ms_per_step": round(dev_step_ms, 3), "algorithmic_bvh_bytes_per_step": int(bytes_step),
            "algorithmic_bvh_gbs": round(alg_gbs, 1)}
    if not args.no_counters:
        c = measure_counters(args, log, is_render_kernel, torch.cuda.get_device_properties(0).multi_processor_count)
        t_dev = dev_step_ms * 1e-3  # the child's dispatches are one render call
        if c.get("valu_instructions"):
            issue = c["valu_instructions"] / t_dev
            lane = c["lane_utilisation"]
            achieved = issue * 64 * lane / 1e12
            roof.update({"achieved": round(achieved, 3), "frac": round(achieved / VALU_PEAK_TOPS, 4),
                         "lane_utilisation": round(lane, 4),
                         "valu_issue_of_peak": round(issue * 64 / 1e12 / VALU_PEAK_TOPS, 4),
                         "valu_instructions_per_step": int(c["valu_instructions"]), "dispatches": c["trace_dispatches"]})
            log(f"{args.workload} VALU: {c['valu_instructions'] / 1e9:.2f} G wave instructions per render call in "
                f"{c['trace_dispatches']} dispatches, lane utilisation {lane:.3f}, issue "
                f"{issue * 64 / 1e12 / VALU_PEAK_TOPS:.3f} of peak; useful {achieved:.2f} TOP/s = "
                f"{achieved / VALU_PEAK_TOPS:.3f} of peak")
        if c.get("ta_busy") is not None:
            roof.update({"vmem_address_unit_busy": round(c["ta_busy"], 4), "vmem_data_unit_busy": round(c["td_busy"], 4),
                         "l1_accesses_per_cu_cycle": round(c["l1_access_rate"], 4)})
            log(f"{args.workload} vector memory: address unit (TA) busy {c['ta_busy']:.3f}, data unit (TD) busy "
                f"{c['td_busy']:.3f}, L1 accesses {c['l1_access_rate']:.3f} per CU per cycle")
        if c.get("l2_hit_rate") is not None:
            roof.update({"l2_hit_rate": round(c["l2_hit_rate"], 4), "l1_hit_rate": round(c.get("l1_hit_rate", 0.0), 4)})
            log(f"{args.workload} caches: L2 hit rate {c['l2_hit_rate']:.3f}, L1 hit rate {c.get('l1_hit_rate', 0.0):.3f}")
        if c.get("fetch_bytes_total") is not None and c.get("write_bytes_total") is not None:
            traffic = c["fetch_bytes_total"] + c["write_bytes_total"]
            hbm_gbs = traffic / t_dev / 1e9
            roof.update({"traffic": int(traffic), "traffic_per": "render call", "hbm_gbs": round(hbm_gbs, 1),
                         "hbm_of_peak": round(hbm_gbs / HBM_PEAK_GBS, 4)})
            log(f"{args.workload} HBM traffic: {traffic / 1e9:.3f} GB per render call = {hbm_gbs:.0f} GB/s = "
                f"{hbm_gbs / HBM_PEAK_GBS:.3f} of HBM peak")
    extra = {}
    if args.workload == "cornell":  # the frame against the FP64 oracle with the same seeds (untimed)
        img = r.get_output_image().astype(np.float64)
        oracle.set_rng(True)  # the oracle's PCG mode, as the renderer's
        try:
            r64 = oracle.OracleScene(desc, precisions=(64,)).render(cam.as_dict(), W, H, 0, spp, depth=depth,
                                                                    precision=64)
        finally:
            oracle.set_rng(False)
        extra["rel_mse_vs_fp64_oracle"] = float(np.mean((img - r64) ** 2 / (r64 ** 2 + 1e-2)))
        log(f"cornell: rel-MSE against the FP64 oracle (same seeds) {extra['rel_mse_vs_fp64_oracle']:.2e}")
    cpu = None
    if not args.no_cpu_baseline:
        hc = host_cpu()
        threads = hc["threads_used"]
        ref = oracle.OracleScene(desc, precisions=(32,))
        cs = {"gdpt": 128, "volume": 640, "cornell": 256}[args.workload]  # ~10 s of CPU work on 16 threads
        t0 = time.perf_counter()
        if args.workload == "cornell":
            oracle.set_rng(True)
            try:
                ref.render(cam.as_dict(), W, H, 0, cs, depth=depth, precision=32, threads=threads)
            finally:
                oracle.set_rng(False)
        elif args.workload == "gdpt":
            bufs = ref.gdpt(cam.as_dict(), W, H, 0, cs, depth=depth, precision=32, threads=threads)
            oracle.gdpt_reconstruct(bufs.astype(np.float32), 50)
        else:
            ref.render(cam.as_dict(), W, H, 0, cs, depth=depth, method=oracle.VOLUME, precision=32, threads=threads)
        t_cpu = time.perf_counter() - t0
        cpu_value = W * H * cs / t_cpu / 1e6
        cpu = cpu_baseline_record(cpu_value, hc, f"{cs} spp of the same {W}x{H} frame on the C++ oracle (FP32), "
                                                 f"{threads} threads, {t_cpu:.1f}s")
        log(f"CPU oracle: {cpu_value:.3f} Msamples/s on {threads} threads; GPU/CPU = {value / cpu_value:.0f}x")
    out = {
        "metric": f"Msamples/s ({args.workload})",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"value_from": "median step time","workload": {"gdpt": f"gradient-domain PT, Cornell box {W}x{H}, {spp} spp, depth {depth}, 5 paths "
                                         f"per sample + 50 reconstruction iterations",
                                "volume": f"volume renderer (homogeneous media), {W}x{H}, {spp} spp, depth {depth}",
                                "cornell": f"PT+NEE, Cornell box (34 triangles), {W}x{H}, {spp} spp, depth {depth}, "
                                           f"PCG random numbers"}[
                                    args.workload],
                   "width": W, "height": H, "spp": spp, "max_depth": depth,
                   "mrays_per_s": round(rays / t_med / 1e6, 1), **extra},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
