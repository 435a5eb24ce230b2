"""Benchmark of the align+stack hot path on one MI355X.

Workload (batched affine ECC, 1080p): 2000 float32 mono 1920x1080 frames of
the seeded planetary-disk generator (disk radius 420 px, texture sigma 3 at
10%, limb darkening 0.6, per-frame affine warp: shift +-20 px, rotation +-1
deg, scale 1+-0.005, shear +-0.002; per-frame gaussian blur sigma U(0.8,2.5);
Poisson noise at 2000 e- peak + 4 DN read noise, stored as DN/65536).
Registration: forward-additive ECC, AFFINE, 4-level pyramid, ecc.scale 1,
eps 0.02 px (doubling per coarser level), <= 50 iterations, translation-first
pass, rho gates at min_rho 0.8.  Stacking: fused remap (bilinear,
REFLECT101) + base_remap mask + weighted average with the per-frame
sharpness weight, FP64 final normalisation.

One step = ssg_stack_batch (weights of the unaligned frames, batched registration and the fused
warp/accumulate, every per-frame decision on the device, replayed as one captured hipGraph) +
compute of the stack, with the frames resident in HBM before the timed region.

Also reported:
  * roofline of the dominant kernel (the kernel family with the largest
    summed time, the translation-first ECC iteration pass on this workload):
    algorithmic bytes / summed kernel time from HIP events on the library's
    stream, over the same steps run eagerly right after the timed graph
    replays (a replayed graph times as one unit); `traffic` is the measured HBM bytes
    per launch of that kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by
    the same counters on a copy of known size), read from
    profiles/pmc_traffic.json when it names the same kernel;
  * plain per-frame GPU form (one frame per launch sequence, host sync per ECC
    iteration), frames/s on a sample of frames;
  * CPU baseline: the FP32 oracle (plain C + OpenMP restatement of the
    reference algorithm) built on the host with -O3 -march=native -ffast-math (the
    reference's own flags), all the process's cores, on the first 100 frames (bounded
    to about 40 s); the same without -ffast-math on the parity frames;
  * parity of the timed batch on its first 16 frames: warps vs the FP64 oracle (and the
    FP32 oracle's own distance), the stack of those frames vs the FP32 oracle with
    bit-exact per-pixel counts.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


ECC_METHODS = {"fa": 0, "ic": 1, "lm": 2, "iclm": 3}
ECC_METHOD_NAMES = {"fa": "forward-additive", "ic": "inverse-compositional", "lm": "Levenberg-Marquardt",
                    "iclm": "inverse-compositional LM"}


def c2_options(S, method="fa"):
    return S.c_image_registration_options(motion_type=S.MOTION_AFFINE, scale=1.0, eps=0.02, ecch_max_level=4,
                                          max_iterations=50, min_rho=0.8, ecc_method=ECC_METHODS[method])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=2000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--plain-frames", type=int, default=24)
    ap.add_argument("--cpu-frames", type=int, default=100, help="upper bound of the CPU sample")
    ap.add_argument("--cpu-seconds", type=float, default=40.0, help="the CPU sample stops after this much time")
    ap.add_argument("--oracle-frames", type=int, default=16,
                    help="frames checked against the FP64 oracle (warps) and the FP32 oracle (stack, counts)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--retune", action="store_true", help="tune launch shapes again instead of using the cache")
    ap.add_argument("--no-graph", action="store_true",
                    help="run the batched steps eagerly instead of replaying the captured hipGraph")
    ap.add_argument("--workload", choices=["affine", "turbulence", "focus", "sweep"], default="affine",
                    help="affine: batched affine ECC + weighted stack of 1080p frames (the headline metric); "
                         "turbulence: phase correlation + eccflow + bicubic remap + average of 1024x1024 frames; "
                         "focus: 8 mixed-size RGB focus sequences, translation ECC + one batched focus stack; "
                         "sweep: each remaining kernel family at its configuration's frame size")
    ap.add_argument("--focus-scale", type=float, default=1.0,
                    help="focus workload: multiply every sequence's frame count by this (1 = 40..200 frames)")
    ap.add_argument("--ecc-method", choices=list(ECC_METHODS), default="fa",
                    help="ECC_ALIGN_METHOD of the affine workload (the headline is forward additive)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="after the timed region, copy a buffer of this size (a known-size reference for "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              f"(the path shards over frames: each GPU registers and stacks its own share, and the "
              f"accumulator planes are summed once at the end)", file=sys.stderr)
        sys.exit(2)

    import torch

    import serstacker_amd as S
    from serstacker_amd import _lib

    if not torch.cuda.is_available():
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)
    lib = _lib.load()
    dev = torch.device("cuda:0")
    if a.workload == "turbulence":
        return bench_turbulence(a, S, _lib, torch)
    if a.workload == "sweep":
        return bench_sweep(a, S, _lib, torch)
    if a.workload == "focus":
        return bench_focus(a, S, _lib, torch)
    n, W, H = a.frames, a.width, a.height

    # ---- inputs resident in HBM (not timed)
    p = S.synth_params(width=W, height=H, disk_radius=420.0 * min(W / 1920, H / 1080))
    frames_t = torch.empty((n, H, W), dtype=torch.float32, device=dev)
    frames = [frames_t[i] for i in range(n)]
    t0 = time.time()
    truth = S.synth_render(p, 0, frames)
    ref = torch.empty((H, W), dtype=torch.float32, device=dev)
    S.synth_render(p, -1, [ref])
    torch.cuda.synchronize()
    t_synth = time.time() - t0
    print(f"[bench] synthesized {n} frames {W}x{H} in {t_synth:.1f}s", file=sys.stderr, flush=True)

    opts = c2_options(S, a.ecc_method)
    reg = S.c_frame_registration(opts)
    assert reg.setup_reference_frame(ref)
    acc = S.c_weigthed_average(W, H, 1)
    views = _lib.images([S.image_view(f) for f in frames])

    # launch shapes and frames per launch: the project's autotuner (cached per device and size class;
    # the shipped choices for MI355X come from serstacker_amd/tuning.json), untimed
    from serstacker_amd.autotune import Autotuner, size_class

    batches = {}

    def chunk_views(k):
        if k not in batches:
            batches[k] = [_lib.images([S.image_view(f) for f in frames[i:i + k]]) for i in range(0, n, k)]
        return batches[k]

    def run_batches(k, graph=True):
        acc.clear()
        sts = []
        for v in chunk_views(k):
            sts.extend(reg.stack_batch(acc, v, use_graph=graph))
        return sts

    # two stages: the accumulate's tile shape and chunking on the fused warp/accumulate alone (one
    # registration of the batch gives its statuses and weights), then the ECC grid and the frames
    # per launch on the whole step
    t_tune = time.perf_counter()
    tuner = Autotuner()
    wkey = "stack_batch_affine_fa" if a.ecc_method == "fa" else f"stack_batch_affine_{a.ecc_method}"
    acc_choice = tuner.lookup(wkey + "_accumulate", size_class(W, H, n))
    if acc_choice is None or a.retune:
        st_t = reg.register_batch(views)
        w_t = S.frame_sharpness(views)
        acc_choice, _, _ = tuner.tune(
            wkey + "_accumulate", size_class(W, H, n), {"acc_rows": [4, 2], "acc_blocks": [4096, 2048, 8192, 16384]},
            lambda c: (acc.clear(), acc.add_registered_batch(views, st_t, S.MOTION_AFFINE, frame_weights=w_t),
                       torch.cuda.synchronize()), repeats=3, retune=a.retune)
    Autotuner.apply(acc_choice)
    tune_space = {"frames_per_launch": [n, max(1, n // 2), max(1, n // 4)], "ecc_grid": [8, 6, 4, 12]}
    choice, tune_times, tune_cached = tuner.tune(
        wkey, size_class(W, H, n), tune_space,
        lambda c: (run_batches(c["frames_per_launch"]), torch.cuda.synchronize()), retune=a.retune)
    choice = {**acc_choice, **choice}
    t_tune = time.perf_counter() - t_tune
    kpl = choice["frames_per_launch"]
    print(f"[bench] tuning ({'cached' if tune_cached else f'{t_tune:.1f} s'}): {choice}", file=sys.stderr, flush=True)

    def step(graph=not a.no_graph):
        # weights + registration + fused warp/accumulate, every per-frame decision on the device
        # (ssg_stack_batch; with a graph: one replay of the captured launch sequence per launch of
        # `kpl` frames), then compute
        st = run_batches(kpl, graph)
        avg, mask = acc.compute()
        return st, avg

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        st, avg = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1e3 / a.steps
    fps = n * a.steps / elapsed
    # the same steps run eagerly with every kernel family bracketed by HIP events on the library's
    # stream (a replayed graph times as one unit): the per-kernel roofline figures
    lib.ssg_registration_set_profiling(reg._h, 1)
    prof_r, prof_a = _lib.Profile(), _lib.Profile()
    lib.ssg_registration_get_profile(reg._h, C.byref(prof_r), 1)
    torch.cuda.synchronize()
    te = time.perf_counter()
    for _ in range(a.steps):
        st, avg = step(graph=False)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - te) * 1e3 / a.steps
    lib.ssg_registration_get_profile(reg._h, C.byref(prof_r), 1)
    lib.ssg_registration_set_profiling(reg._h, 0)
    prof_a = prof_r  # the stack batch accounts its accumulate launches on the registration stream
    nok = int(sum(s.ok for s in st))
    its = np.array([s.iterations for s in st])
    its_t = np.array([s.iterations_translation for s in st])
    print(f"[bench] batched ({'eager' if a.no_graph else 'hipGraph replay'}): {fps:.1f} frames/s ({ms_step:.1f} "
          f"ms/step; profiled eager steps {eager_ms:.1f} ms), {nok}/{n} registered, "
          f"main iterations mean {its.mean():.1f} min {its.min()} max {its.max()}, translation-first "
          f"mean {its_t.mean():.1f} max {its_t.max()}", file=sys.stderr, flush=True)

    kinds = {"ecc_translation": _lib.PROF_ECC_TRANSLATION, "ecc_main": _lib.PROF_ECC_MAIN,
             "ecc_solve": 5, "correlation": _lib.PROF_CORRELATION, "pyramid": _lib.PROF_PYRAMID}
    kernels = {}
    for name, k in kinds.items():
        kernels[name] = {"launches": int(prof_r.launches[k]), "ms": prof_r.ms[k], "bytes": prof_r.bytes[k]}
    kernels["accumulate"] = {"launches": int(prof_a.launches[_lib.PROF_ACCUMULATE]),
                             "ms": prof_a.ms[_lib.PROF_ACCUMULATE], "bytes": prof_a.bytes[_lib.PROF_ACCUMULATE]}
    for k in kernels.values():
        k["avg_us"] = 1e3 * k["ms"] / max(k["launches"], 1)
        k["gbs"] = k["bytes"] / max(k["ms"], 1e-9) / 1e6
        k["frac"] = k["gbs"] / HBM_PEAK_GBS
    dom_name = max(kernels, key=lambda k: kernels[k]["ms"])
    dom = kernels[dom_name]
    print(f"[bench] kernels: " + ", ".join(f"{k}: {v['ms']:.1f} ms {v['gbs']:.0f} GB/s" for k, v in kernels.items()),
          file=sys.stderr, flush=True)

    if a.copy_calib_mib > 0:
        nb = a.copy_calib_mib << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {a.copy_calib_mib} MiB done", file=sys.stderr, flush=True)
        del src, dst

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("kernel_family") == dom_name:
            traffic = tr.get("hbm_bytes_per_launch")

    # ---- plain per-frame GPU form: one frame per launch sequence, host sync per ECC iteration
    npl = min(a.plain_frames, n)
    plain_fps = None
    if npl > 0:
        acc1 = S.c_weigthed_average(W, H, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(npl):
            w1 = S.frame_sharpness([frames[i]])
            reg.register_frame(frames[i])
            s1 = (_lib.RegistrationStatus * 1)(reg.status())
            acc1.add_registered_batch([frames[i]], s1, S.MOTION_AFFINE, frame_weights=w1)
        acc1.compute()
        torch.cuda.synchronize()
        plain_fps = npl / (time.perf_counter() - t0)
        print(f"[bench] plain per-frame GPU: {plain_fps:.1f} frames/s on {npl} frames", file=sys.stderr, flush=True)

    # ---- CPU baseline and parity of the timed batch (the oracle is the checker here, never the thing measured
    # on the GPU side)
    cpu, cpu_plain, warp, parity = None, None, None, None
    if not a.no_cpu and a.cpu_frames > 0:
        import tempfile

        from oracle.oracle import Oracle, build_native, reg_options

        threads = len(os.sched_getaffinity(0))
        cpu_model = "unknown"
        try:
            with open("/proc/cpuinfo") as fh:
                cpu_model = next(ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name"))
        except (OSError, StopIteration):
            pass
        libs = build_native(os.path.join(tempfile.gettempdir(), "ssg_oracle_native"))
        kw = dict(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=4, max_iterations=50, min_rho=0.8,
                  ecc_method=ECC_METHODS[a.ecc_method])
        ref_h = ref.cpu().numpy()

        def run_cpu(library, max_frames, max_s):
            """register + remap + weighted-average update of frames 0.. until max_frames or max_s of CPU time"""
            o32 = Oracle(32, library=library)
            o32.set_threads(threads)
            oreg = o32.registration(reg_options(**kw))
            oreg.setup_reference_frame(ref_h)
            A = np.zeros((H, W), np.float32)
            Wt = np.zeros((H, W), np.float32)
            done, cpu_s, stats = 0, 0.0, []
            for i in range(min(max_frames, n)):
                f = frames[i].cpu().numpy()
                t0 = time.perf_counter()
                wgt = o32.sharpness(f)
                ost = oreg.register_frame(f)
                t = o32.xform(3, list(ost.params))
                rmap = o32.create_map(t, W, H)
                d, m = o32.base_remap(f, rmap, interp=1, border=4)
                wimg = np.full((H, W), np.float32(wgt))
                wr, _ = o32.base_remap(wimg, rmap, interp=1, border=0, want_mask=False)
                o32.weighted_average_update(d, A, Wt, weights=(wr * (m / np.float32(255))).astype(np.float32))
                cpu_s += time.perf_counter() - t0
                done += 1
                stats.append(np.array(ost.matrix[:]))
                if cpu_s >= max_s:
                    break
            return done, cpu_s, stats

        k = min(a.oracle_frames, n)
        if k > 0:
            # the reference-faithful FP32 build (contraction off; also timed, on these frames) and the FP64 oracle
            done32, s32, m32 = run_cpu(libs["native"], k, 1e9)
            cpu_plain = {"value": round(done32 / s32, 4), "unit": "frames/s", "cores": threads,
                         "sample": f"{done32} frames, -O3 -march=native without -ffast-math"}
            o64 = Oracle(64)
            o64.set_threads(threads)
            r64 = o64.registration(reg_options(**kw))
            r64.setup_reference_frame(ref_h)
            c = np.array([[0, 0, 1], [W - 1, 0, 1], [0, H - 1, 1], [W - 1, H - 1, 1]], np.float64).T
            disp = lambda m1, m2: float(np.sqrt((((np.asarray(m1) - np.asarray(m2)).reshape(2, 3) @ c) ** 2).sum(0)).max())
            eg, eo = [], []
            for i in range(k):
                o = r64.register_frame(frames[i].cpu().numpy())
                eg.append(disp(st[i].matrix[:], o.matrix[:]))
                eo.append(disp(m32[i], o.matrix[:]))
            warp = {"frames": k, "gpu_vs_fp64_max_px": round(max(eg), 5), "gpu_vs_fp64_median_px": round(float(np.median(eg)), 5),
                    "within_0.01px": int(sum(e <= 0.01 for e in eg)),
                    "fp32_oracle_vs_fp64_max_px": round(max(eo), 5),
                    "fp32_oracle_vs_fp64_median_px": round(float(np.median(eo)), 5)}
            print(f"[bench] warps vs FP64 over {k} frames: GPU max {max(eg):.5f} median {np.median(eg):.5f} px; "
                  f"FP32 oracle max {max(eo):.5f} median {np.median(eo):.5f} px", file=sys.stderr, flush=True)
            # the stack of those frames with the batch's own parameters and weights: counts bit-exact and the
            # average against the FP32 incremental mean and the FP64 sums
            wts = S.frame_sharpness(frames[:k])
            sub = (_lib.RegistrationStatus * k)(*st[:k])
            acck = S.c_weigthed_average(W, H, 1)
            acck.add_registered_batch(frames[:k], sub, S.MOTION_AFFINE, frame_weights=wts)
            gavg, _ = acck.compute()
            ora = Oracle(32)
            ora.set_threads(threads)
            A32 = np.zeros((H, W), np.float32)
            W32 = np.zeros((H, W), np.float32)
            Ssum = np.zeros((H, W))
            Wsum = np.zeros((H, W))
            cnt = np.zeros((H, W), np.int64)
            for i in range(k):
                if not st[i].ok:
                    continue
                f = frames[i].cpu().numpy()
                rmap = ora.create_map(ora.xform(3, list(st[i].params)), W, H)
                d, m = ora.base_remap(f, rmap, interp=1, border=4)
                wr, _ = ora.base_remap(np.full((H, W), np.float32(wts[i])), rmap, interp=1, border=0, want_mask=False)
                ora.weighted_average_update(d, A32, W32, weights=(wr * (m / np.float32(255))).astype(np.float32))
                mb = m > 0
                Ssum += np.where(mb, np.float64(np.float32(wts[i])) * d, 0)
                Wsum += np.where(mb, np.float64(np.float32(wts[i])), 0)
                cnt += mb
            rel = lambda x, y: float(np.linalg.norm(np.float64(x) - np.float64(y)) / max(np.linalg.norm(np.float64(y)), 1e-300))
            f64avg = np.where(Wsum > 0, Ssum / np.maximum(Wsum, 1e-300), 0)
            parity = {"frames": k, "counts_bitexact": bool(np.array_equal(acck.counts(), cnt)),
                      "accumulated_frames_match": acck.accumulated_frames() == int(sum(s.ok for s in st[:k])),
                      "stack_rel_l2_vs_fp32_oracle": rel(gavg, A32), "stack_rel_l2_vs_fp64_sums": rel(gavg, f64avg)}
            print(f"[bench] stack of {k} frames: counts bit-exact {parity['counts_bitexact']}, rel-L2 vs FP32 oracle "
                  f"{parity['stack_rel_l2_vs_fp32_oracle']:.2e}, vs FP64 {parity['stack_rel_l2_vs_fp64_sums']:.2e}",
                  file=sys.stderr, flush=True)

        # last: loading a -ffast-math library sets flush-to-zero for the process, so the checks above
        # run before it
        done, cpu_s, _ = run_cpu(libs["native_fastmath"], a.cpu_frames, a.cpu_seconds)
        per_frame = cpu_s / done
        cpu = {"value": round(done / cpu_s, 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"{done} of the {n} frames ({W}x{H}, {cpu_s:.1f} s; the whole batch would take "
                         f"{per_frame * n:.0f} s): sharpness, registration (translation-first + affine ECC, 4 levels), "
                         f"base_remap and weighted-average update; the FP32 oracle built -O3 -march=native "
                         f"-ffast-math (the reference's flags) on this host, {threads} OpenMP threads, {cpu_model}"}
        print(f"[bench] cpu oracle (native, fast-math): {cpu['value']} frames/s, {threads} threads, {cpu_model}",
              file=sys.stderr, flush=True)

    out = {
        "metric": "aligned+stacked frames/s at 1080p (batched vs plain GPU vs CPU), % HBM roofline",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched affine ECC + weighted stack, {n} f32 mono frames {W}x{H}, 4-level "
                               f"pyramid, translation-first, eps 0.02 px, <=50 iterations"
                               + ("" if a.ecc_method == "fa" else f", {ECC_METHOD_NAMES[a.ecc_method]} ECC"),
                   "frames": n, "width": W, "height": H},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(dom["gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom["frac"], 4), "traffic": traffic,
                     "avg_launch_us": round(dom["avg_us"], 2), "launches": dom["launches"]},
        "cpu_baseline": cpu,
        "graph": not a.no_graph,
        "tuning": {"choice": choice, "cached": tune_cached, "seconds": round(t_tune, 2),
                   "size_class": size_class(W, H, n)},
        "eager_profiled_ms_per_step": round(eager_ms, 3),
        "plain_gpu": {"value": round(plain_fps, 2), "unit": "frames/s", "frames": npl} if plain_fps else None,
        "speedup_vs_plain": round(fps / plain_fps, 2) if plain_fps else None,
        "speedup_vs_cpu": round(fps / cpu["value"], 1) if cpu else None,
        "registered": nok,
        "main_iterations": {"mean": round(float(its.mean()), 2), "min": int(its.min()), "max": int(its.max())},
        "cpu_baseline_no_fastmath": cpu_plain,
        "warp_vs_fp64": warp,
        "parity": parity,
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    print(json.dumps(out))


def remap_dev(lib, _lib, S, src, rmap, dst, interp):
    sv, mv, dv = S.image_view(src), S.image_view(rmap), S.image_view(dst)
    _lib.check(lib.ssg_remap(C.byref(sv), None, C.byref(mv), None, 0, None, C.byref(dv), None, interp,
                             _lib.BORDER_REFLECT101, 0.0), "ssg_remap")


def bench_turbulence(a, S, _lib, torch):
    """Dense local alignment for turbulence: 500 f32 mono 1024x1024 frames of the disk generator
    (radius 400 px), each the reference warped by a global shift U(-10, 10) px plus a smooth field
    (8x8 control grid of N(0, 1.5) px, bicubic-upsampled), plus gaussian noise.  One step = FFT
    phase-correlation pre-alignment, eccflow (pyramid downscale, 5 levels, support scale 3,
    3 iterations per level) seeded with the translation, and the fused per-pixel bicubic remap +
    average, with the frames resident in HBM."""
    lib = _lib.load()
    dev = torch.device("cuda:0")
    n = a.frames if a.frames != 2000 else 500
    W = H = 1024
    p = S.synth_params(width=W, height=H, disk_radius=400.0)
    ref = torch.empty((H, W), dtype=torch.float32, device=dev)
    S.synth_render(p, -1, [ref])
    g = torch.Generator(device="cpu").manual_seed(20261203)
    yy, xx = torch.meshgrid(torch.arange(H, dtype=torch.float32, device=dev),
                            torch.arange(W, dtype=torch.float32, device=dev), indexing="ij")
    ident = torch.stack([xx, yy], -1)
    frames_t = torch.empty((n, H, W), dtype=torch.float32, device=dev)
    noise = 0.005 * float(ref.max())  # gaussian read noise, 0.5% of the disk peak
    # the flow is defined on the textured disk; the dark background has no structure to track
    disk = ((xx - W / 2) ** 2 + (yy - H / 2) ** 2) < (0.9 * 400.0) ** 2
    true_flow = []
    t0 = time.time()
    for i in range(n):
        ctrl = torch.randn((1, 2, 8, 8), generator=g) * 1.5
        field = torch.nn.functional.interpolate(ctrl.to(dev), size=(H, W), mode="bicubic", align_corners=True)[0]
        shift = (torch.rand(2, generator=g) * 20 - 10).to(dev)
        d = field.permute(1, 2, 0) + shift
        gen_map = (ident - d).contiguous()
        torch.cuda.synchronize()  # the library's stream does not wait for torch's
        remap_dev(lib, _lib, S, ref, gen_map, frames_t[i], _lib.INTER_CUBIC)
        frames_t[i] += torch.randn((H, W), generator=g).to(dev) * noise
        if i < 4:
            true_flow.append(d)
    torch.cuda.synchronize()
    print(f"[bench] synthesized {n} turbulent frames {W}x{H} in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    frames = [frames_t[i] for i in range(n)]
    views = _lib.images([S.image_view(f) for f in frames])
    maps_t = torch.empty((n, H, W, 2), dtype=torch.float32, device=dev)
    mviews = _lib.images([S.image_view(maps_t[i]) for i in range(n)])
    pc = S.phase_correlation(W, H, 1)
    pc.set_reference(ref)
    fopts = dict(downscale=_lib.ECCFLOW_DOWNSCALE_PYRAMID, support_scale=3, max_iterations=3, min_image_size=-1,
                 max_pyramid_level=4, update_multiplier=1.5, noise_level=-1.0)
    flow = S.c_eccflow(**fopts)
    flow.set_reference_image(ref)
    assert flow.num_levels() == 5
    acc = S.c_weigthed_average(W, H, 1)

    def step():
        est = pc.estimate(views)
        T = torch.from_numpy(est[:, :2].astype(np.float32)).to(dev)
        maps_t.copy_(ident.unsqueeze(0) + T[:, None, None, :])
        torch.cuda.synchronize()  # the library's stream does not wait for torch's
        flow.compute_batch(views, mviews, rmaps=mviews)
        acc.clear()
        acc.add_remapped_batch(views, mviews, interpolation=_lib.INTER_CUBIC)
        return est, acc.compute()

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    lib.ssg_eccflow_set_profiling(flow._h, 1)
    lib.ssg_accumulation_set_profiling(acc._h, 1)
    lib.ssg_phasecorr_set_profiling(pc._h, 1)
    pf, pa, pp = _lib.Profile(), _lib.Profile(), _lib.Profile()
    lib.ssg_eccflow_get_profile(flow._h, C.byref(pf), 1)
    lib.ssg_accumulation_get_profile(acc._h, C.byref(pa), 1)
    lib.ssg_phasecorr_get_profile(pc._h, C.byref(pp), 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        est, (avg, mask) = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lib.ssg_eccflow_get_profile(flow._h, C.byref(pf), 1)
    lib.ssg_accumulation_get_profile(acc._h, C.byref(pa), 1)
    lib.ssg_phasecorr_get_profile(pc._h, C.byref(pp), 1)
    ms_step = elapsed * 1e3 / a.steps
    fps = n * a.steps / elapsed
    kernels = {
        "eccflow": (pf, _lib.PROF_ECCFLOW),
        "accumulate": (pa, _lib.PROF_ACCUMULATE),
        "phasecorr": (pp, _lib.PROF_FFT),
    }
    kern = {}
    for name, (pr, k) in kernels.items():
        kern[name] = {"launches": int(pr.launches[k]), "ms": pr.ms[k], "bytes": pr.bytes[k]}
        kern[name]["avg_us"] = 1e3 * kern[name]["ms"] / max(kern[name]["launches"], 1)
        kern[name]["gbs"] = kern[name]["bytes"] / max(kern[name]["ms"], 1e-9) / 1e6
        kern[name]["frac"] = kern[name]["gbs"] / HBM_PEAK_GBS
    dom_name = max(kern, key=lambda k: kern[k]["ms"])
    dom = kern[dom_name]
    print(f"[bench] turbulence: {fps:.1f} frames/s ({ms_step:.1f} ms/step); " +
          ", ".join(f"{k}: {v['ms']:.1f} ms {v['gbs']:.0f} GB/s" for k, v in kern.items()), file=sys.stderr, flush=True)

    # flow endpoint error against the generator's field on a few frames (interior)
    maps = maps_t[:len(true_flow)].clone()
    epe_truth = []
    for i, d in enumerate(true_flow):
        e = torch.linalg.norm(maps[i] - ident - d, dim=-1)[disk]
        epe_truth.append(round(float(e.median()), 4))

    cpu = None
    epe64 = None
    if not a.no_cpu and a.cpu_frames > 0:
        from oracle import eccflow as EF
        from oracle import phasecorr as PC
        from oracle.oracle import Oracle

        threads = min(16, os.cpu_count() or 1)
        o32 = Oracle(32)
        o32.set_threads(threads)
        ref_h = ref.cpu().numpy()
        f32 = EF.EccFlow(o32, **fopts)
        t0 = time.perf_counter()
        f32.set_reference_image(ref_h)
        ref_spec = PC.spectrum(ref_h, 1, True)
        A = np.zeros((H, W), np.float32)
        Wt = np.zeros((H, W), np.float32)
        done = 0
        cpu_s = time.perf_counter() - t0
        fx, fy = np.meshgrid(np.arange(W, dtype=np.float32), np.arange(H, dtype=np.float32))
        for i in range(min(a.cpu_frames, n)):
            f = frames[i].cpu().numpy()
            t0 = time.perf_counter()
            tx, ty, _ = PC.estimate(ref_spec, f, 1)
            init = np.stack([fx + np.float32(tx), fy + np.float32(ty)], -1)
            m = f32.compute(f, init)
            d, dm = o32.base_remap(f, m, interp=2, border=4)
            o32.weighted_average_update(d, A, Wt, mask=dm)
            cpu_s += time.perf_counter() - t0
            done += 1
            if cpu_s >= a.cpu_seconds:
                break
        cpu = {"value": round(done / cpu_s, 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"{done} of the {n} frames ({W}x{H}, {cpu_s:.1f} s incl. the reference pyramid): phase "
                         f"correlation (numpy FFT), eccflow (numpy float32 restatement over the C oracle's pyrDown, "
                         f"differentiate and remap, {threads} OpenMP threads), bicubic base_remap and average update"}
        print(f"[bench] cpu oracle: {cpu['value']} frames/s", file=sys.stderr, flush=True)
        if a.oracle_frames > 0:
            o64 = Oracle(64)
            o64.set_threads(threads)
            f64 = EF.EccFlow(o64, **fopts)
            f64.set_reference_image(ref_h.astype(np.float64))
            errs = []
            dk = disk.cpu().numpy()
            for i in range(min(a.oracle_frames, n)):
                init = np.stack([fx + np.float32(est[i, 0]), fy + np.float32(est[i, 1])], -1)
                m64 = f64.compute(frames[i].cpu().numpy().astype(np.float64), init.astype(np.float64))
                e = np.hypot(*(maps_t[i].cpu().numpy() - m64).transpose(2, 0, 1))[dk]
                errs.append((float(e.mean()), float(np.percentile(e, 99)), float(e.max())))
            epe64 = {"region": "disk, r < 360 px", "mean_px": max(x[0] for x in errs),
                     "p99_px": max(x[1] for x in errs), "max_px": max(x[2] for x in errs), "frames": len(errs)}
            print(f"[bench] flow EPE vs FP64 oracle: {epe64}", file=sys.stderr, flush=True)

    out = {
        "metric": "aligned+stacked frames/s, turbulence workload (phase correlation + eccflow + bicubic remap)",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"phase correlation + eccflow (5 levels, support scale 3, 3 iterations) + bicubic "
                               f"remap + average, {n} f32 mono frames {W}x{H}", "frames": n, "width": W, "height": H},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(dom["gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom["frac"], 4), "traffic": None,
                     "avg_launch_us": round(dom["avg_us"], 2), "launches": dom["launches"]},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(fps / cpu["value"], 1) if cpu else None,
        "flow_epe_vs_truth_median_px_disk": epe_truth,
        "flow_epe_vs_fp64": epe64,
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kern.items()},
    }
    print(json.dumps(out))


FOCUS_SIZES = [(1024, 768), (2048, 1536), (2560, 2048), (3072, 2048)]


def focus_sequences(torch, dev, seed=5, scale=1.0):
    """The focus workload's inputs, built on the GPU (untimed): 8 sequences with sizes drawn from
    FOCUS_SIZES and 40..200 RGB f32 frames each.  The scene is a sum of 6 octaves of value noise
    per channel (lattice values bicubically upsampled, amplitude halving per octave); frame k of a
    sequence is the scene blurred per pixel with sigma = |z(x, y) - f_k| * 1.2 px, z a smooth depth
    map in [0, 8] and f_k a focus sweep over [0, 8] (the blur interpolated between a stack of
    gaussian blurs at sigma 0, 0.6, 1.2, 2.4, 4.8, 9.6), then translated by U(-6, 6) px (bilinear,
    border replicated).  Returns [(w, h, frames (n, h, w, 3) BGR interleaved, true shifts (n, 2))]."""
    import torch.nn.functional as F

    g = torch.Generator(device="cpu").manual_seed(seed)
    out = []
    for q in range(8):
        w, h = FOCUS_SIZES[int(torch.randint(0, 4, (1,), generator=g))]
        n = max(2, int(round(int(torch.randint(40, 201, (1,), generator=g)) * scale)))

        def noise(octaves, base, ch):
            acc = torch.zeros((ch, h, w), device=dev)
            amp = 1.0
            for o in range(octaves):
                gw, gh = max(2, w // (base >> o) + 2), max(2, h // (base >> o) + 2)
                lat = torch.rand((1, ch, gh, gw), generator=g).to(dev)
                acc += amp * F.interpolate(lat, size=(h, w), mode="bicubic", align_corners=False)[0]
                amp *= 0.5
            return acc / (2.0 - 2.0 ** (1 - octaves))

        scene = noise(6, 256, 3).clamp(0, 1)
        z = noise(2, 512, 1)[0]
        z = (z - z.min()) / (z.max() - z.min()).clamp_min(1e-6) * 8.0
        sig = [0.0, 0.6, 1.2, 2.4, 4.8, 9.6]
        stack = [scene]
        for sg in sig[1:]:
            r = int(3 * sg + 0.5)
            t = torch.arange(-r, r + 1, device=dev, dtype=torch.float32)
            k = torch.exp(-t * t / (2 * sg * sg))
            k = k / k.sum()
            x = F.pad(scene[None], (r, r, 0, 0), mode="replicate")
            x = F.conv2d(x, k.view(1, 1, 1, -1).repeat(3, 1, 1, 1), groups=3)
            x = F.pad(x, (0, 0, r, r), mode="replicate")
            x = F.conv2d(x, k.view(1, 1, -1, 1).repeat(3, 1, 1, 1), groups=3)
            stack.append(x[0])
        stack = torch.stack(stack)  # (6, 3, h, w)
        sigt = torch.tensor(sig, device=dev)
        frames = torch.empty((n, h, w, 3), device=dev)
        shifts = (torch.rand((n, 2), generator=g) * 12.0 - 6.0).to(dev)
        yy, xx = torch.meshgrid(torch.arange(h, device=dev, dtype=torch.float32),
                                torch.arange(w, device=dev, dtype=torch.float32), indexing="ij")
        for k in range(n):
            fk = 8.0 * k / max(n - 1, 1)
            s_ = (z - fk).abs() * 1.2
            j = torch.searchsorted(sigt, s_.clamp(max=9.6).contiguous(), right=True).clamp(1, 5)
            lo, hi = sigt[j - 1], sigt[j]
            wgt = ((s_.clamp(max=9.6) - lo) / (hi - lo))[None]
            a = torch.gather(stack, 0, (j - 1)[None, None].expand(1, 3, h, w))[0]
            b = torch.gather(stack, 0, j[None, None].expand(1, 3, h, w))[0]
            img = a * (1 - wgt) + b * wgt
            gx = (xx - shifts[k, 0]) / (w - 1) * 2 - 1
            gy = (yy - shifts[k, 1]) / (h - 1) * 2 - 1
            img = F.grid_sample(img[None], torch.stack([gx, gy], -1)[None], mode="bilinear", padding_mode="border",
                                align_corners=True)[0]
            frames[k] = img.permute(1, 2, 0)
        out.append((w, h, frames, shifts.cpu().numpy()))
        del stack
    torch.cuda.synchronize()
    return out


def bench_focus(a, S, _lib, torch):
    """Focus stacking in mixed-size batches: 8 sequences (sizes from 1024x768 .. 3072x2048, 40..200
    RGB f32 frames each).  One step = translation registration of every sequence (forward-additive
    ECC on the gray channel, 3 levels, against the sequence's first frame) + ONE batched focus stack
    of all sequences (lpg k=6 p=2 weights, gaussian sigma 5, 6-level laplacian fusion) + compute of
    the 8 stacks.  Frames resident in HBM before the timed region."""
    dev = torch.device("cuda:0")
    lib = _lib.load()
    t0 = time.time()
    seqs = focus_sequences(torch, dev, scale=a.focus_scale)
    nf = sum(x[2].shape[0] for x in seqs)
    px = sum(x[0] * x[1] * x[2].shape[0] for x in seqs)
    print(f"[focus] synthesized {nf} frames in {time.time() - t0:.1f}s: " +
          ", ".join(f"{w}x{h}x{f.shape[0]}" for w, h, f, _ in seqs), file=sys.stderr, flush=True)
    regs, views = [], []
    for w, h, f, _ in seqs:
        o = S.c_image_registration_options(motion_type=S.MOTION_TRANSLATION, scale=1.0, eps=0.02, ecch_max_level=3,
                                           max_iterations=50)
        r = S.c_frame_registration(o)
        assert r.setup_reference_frame(f[0])
        regs.append(r)
        views.append(_lib.images([S.image_view(f[k]) for k in range(f.shape[0])]))
    stacks = [S.c_laplacian_pyramid_focus_stacking(w, h, 3, fusing_policy=2, lpg_k=6.0, lpg_p=2.0, lpg_sigma=5.0,
                                                   max_levels=6) for w, h, _, _ in seqs]
    outs = [torch.empty((h, w, 3), device=dev) for w, h, _, _ in seqs]
    # the sequences register concurrently: one host thread per sequence, each handle on its own HIP
    # stream (the library calls release the interpreter lock)
    from concurrent.futures import ThreadPoolExecutor
    pool = ThreadPoolExecutor(max_workers=len(seqs))

    def register(q):
        st = regs[q].register_batch(views[q])
        return np.array([[x.params[0], x.params[1]] for x in st], np.float32)

    times = {"reg": 0.0, "stack": 0.0}

    def step():
        ts = time.perf_counter()
        shifts = list(pool.map(register, range(len(seqs))))
        t1 = time.perf_counter()
        for st in stacks:
            st.clear()
        S.focus_stack_batch(stacks, views, shifts)
        for st, o in zip(stacks, outs):
            st.compute(o)
        torch.cuda.synchronize()
        times["reg"] += t1 - ts
        times["stack"] += time.perf_counter() - t1
        return shifts

    # frames per sequence and chunk of the batched focus stack: the project's autotuner (cached)
    from serstacker_amd.autotune import Autotuner, size_class

    shifts0 = list(pool.map(register, range(len(seqs))))

    def stack_only(_choice):
        for st in stacks:
            st.clear()
        S.focus_stack_batch(stacks, views, shifts0)
        torch.cuda.synchronize()

    t_tune = time.perf_counter()
    choice, _, tune_cached = Autotuner().tune("focus_stack_batch", size_class(int(px / nf), 1, nf),
                                              {"focus_chunk": [32, 16, 8]}, stack_only, retune=a.retune)
    t_tune = time.perf_counter() - t_tune
    print(f"[focus] tuning ({'cached' if tune_cached else f'{t_tune:.1f} s'}): {choice}", file=sys.stderr, flush=True)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    times["reg"] = times["stack"] = 0.0
    t0 = time.perf_counter()
    for it in range(a.steps):
        if it == a.steps - 1:
            lib.ssg_accumulation_set_profiling(stacks[0]._h, 1)
            prof_reset = _lib.Profile()
            lib.ssg_accumulation_get_profile(stacks[0]._h, C.byref(prof_reset), 1)
        shifts = step()
    elapsed = time.perf_counter() - t0
    pool.shutdown()
    t_reg, t_stack = times["reg"], times["stack"]
    prof_on = stacks[0]
    prof = _lib.Profile()
    lib.ssg_accumulation_get_profile(prof_on._h, C.byref(prof), 1)
    fps = nf * a.steps / elapsed
    ms_step = elapsed * 1e3 / a.steps
    kinds = {"focus_prep": _lib.PROF_FOCUS_PREP, "focus_blur": _lib.PROF_FOCUS_BLUR,
             "focus_down": _lib.PROF_FOCUS_DOWN, "focus_fuse": _lib.PROF_FOCUS_FUSE}
    kernels = {}
    for name, k in kinds.items():
        kk = {"launches": int(prof.launches[k]), "ms": prof.ms[k], "bytes": prof.bytes[k]}
        kk["avg_us"] = 1e3 * kk["ms"] / max(kk["launches"], 1)
        kk["gbs"] = kk["bytes"] / max(kk["ms"], 1e-9) / 1e6
        kk["frac"] = kk["gbs"] / HBM_PEAK_GBS
        kernels[name] = kk
    dom_name = max(kernels, key=lambda k: kernels[k]["ms"])
    dom = kernels[dom_name]
    # frame k is the scene moved by d_k, the reference (frame 0) by d_0: the registration finds d_k - d_0
    err = max(float(np.abs(sh - (x[3] - x[3][0])).max()) for sh, x in zip(shifts, seqs))
    print(f"[focus] {fps:.1f} frames/s ({ms_step:.1f} ms/step: registration {1e3 * t_reg / a.steps:.1f} ms, "
          f"stack {1e3 * t_stack / a.steps:.1f} ms); max |shift - true shift| {err:.3f} px", file=sys.stderr,
          flush=True)
    print("[focus] kernels: " + ", ".join(f"{k}: {v['ms']:.1f} ms {v['gbs']:.0f} GB/s x{v['launches']}"
                                          for k, v in kernels.items()), file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        # the FP32 oracle (checker; a bounded sample): translation ECC + remap + lpg-weight focus add of frames
        # of the first 1024x768 sequence (or the smallest), until about cpu_seconds of CPU work
        from oracle import focus as ofocus
        from oracle.oracle import Oracle, reg_options

        threads = min(16, os.cpu_count() or 1)
        o32 = Oracle(32)
        o32.set_threads(threads)
        w, h, f, _ = min(seqs, key=lambda x: x[0] * x[1])
        host = f.cpu().numpy()
        gray = lambda im: ((im[..., 0] * np.float32(0.114) + im[..., 1] * np.float32(0.587)) +
                           im[..., 2] * np.float32(0.299)).astype(np.float32)
        oreg = o32.registration(reg_options(motion_type=0, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=50))
        oreg.setup_reference_frame(gray(host[0]))
        fs = ofocus.FocusStack(o32, ofocus.LPG_WEIGHTS, lpg_k=6.0, lpg_p=2.0, lpg_sigma=5.0, max_levels=6)
        done, cpu_s = 0, 0.0
        for k in range(host.shape[0]):
            ts = time.perf_counter()
            st = oreg.register_frame(gray(host[k]))
            t = o32.xform(0, [st.params[0], st.params[1]])
            rmap = o32.create_map(t, w, h)
            fs.add([o32.remap(np.ascontiguousarray(host[k][..., c]), rmap, interp=1, border=4) for c in range(3)])
            cpu_s += time.perf_counter() - ts
            done += 1
            if cpu_s >= a.cpu_seconds:
                break
        cpu = {"value": round(done / cpu_s, 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"{done} frames of a {w}x{h} RGB sequence ({cpu_s:.1f} s): translation ECC, remap and the "
                         f"lpg-weight focus add, FP32 oracle, {threads} threads (the frame size mix of the workload "
                         f"averages {px / nf / (w * h):.2f}x this size)"}
        print(f"[focus] cpu oracle: {cpu['value']} frames/s of {w}x{h}", file=sys.stderr, flush=True)

    out = {
        "metric": "focus-stacked frames/s (8 mixed-size RGB sequences, one batched launch per step), % HBM roofline",
        "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "focus stacking of 8 mixed-size RGB sequences: translation ECC per sequence + one "
                               "batched lpg(k=6,p=2)/sigma-5/6-level focus stack",
                   "sequences": [f"{w}x{h}x{f.shape[0]}" for w, h, f, _ in seqs], "frames": nf,
                   "megapixels": round(px / 1e6, 1)},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(dom["gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom["frac"], 4), "traffic": None,
                     "avg_launch_us": round(dom["avg_us"], 2), "launches": dom["launches"]},
        "cpu_baseline": cpu,
        "registration_ms_per_step": round(1e3 * t_reg / a.steps, 2),
        "stack_ms_per_step": round(1e3 * t_stack / a.steps, 2),
        "max_shift_error_px": round(err, 4),
        "tuning": {"choice": choice, "cached": tune_cached, "seconds": round(t_tune, 2)},
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
    }
    print(json.dumps(out))


def bench_sweep(a, S, _lib, torch):
    """Per-family throughput of the kernels outside the headline workload, at the frame sizes of
    their configurations, on synthetic device-resident inputs.  Each entry: mean wall time per call
    (the calls return after their stream finished) and algorithmic bytes per call / that time."""
    lib = _lib.load()
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(7)
    reps = max(1, a.steps)
    res = {}

    def timed(name, fn, nbytes, unit_count, unit):
        for _ in range(max(1, a.warmup)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[name] = {"ms_per_call": round(dt * 1e3, 3), f"{unit}_per_s": round(unit_count / dt, 1),
                     "algorithmic_bytes": nbytes, "gbs": round(nbytes / dt / 1e9, 1),
                     "frac": round(nbytes / dt / 1e9 / HBM_PEAK_GBS, 4)}
        print(f"[sweep] {name}: {dt * 1e3:.2f} ms/call, {nbytes / dt / 1e9:.0f} GB/s", file=sys.stderr, flush=True)

    # Bayer average: 64 u16 RGGB frames of 4096x3000 (12-bit), even-step translations + subpixel
    W, H, n = 4096, 3000, 64
    raw = (torch.rand((n, H, W), generator=g) * 4095).to(torch.int16).to(dev)
    st = (_lib.RegistrationStatus * n)()
    for i in range(n):
        st[i].ok, st[i].nparams = 1, 2
        st[i].params[0] = float(2 * int(torch.randint(-7, 8, (1,), generator=g)) + 0.3)
        st[i].params[1] = float(2 * int(torch.randint(-7, 8, (1,), generator=g)) - 0.2)
    views = _lib.images([S.image_view(raw[i]) for i in range(n)])
    bay = S.c_bayer_average(W, H, bpp=12)
    torch.cuda.synchronize()
    # u16 frames read once; the BGR sum, weight and count planes (3 x 12 B/px) read and written once
    timed("bayer_batch_64x4096x3000_u16", lambda: bay.add_registered_batch(views, st, S.MOTION_TRANSLATION),
          n * 2 * W * H + 2 * 36 * W * H, n, "frames")
    # phase correlation on 2x2-binned luminance of the same frames: padded 2048x2048 spectra
    pc = S.phase_correlation(W, H, 2)
    pc.set_reference(raw[0])
    M = 2048 * 2048
    timed("phasecorr_64x4096x3000_binned2", lambda: pc.estimate(views), n * (2 * W * H + 17 * 8 * M), n, "frames")
    del raw, views
    # focus stacking: 16 RGB f32 frames of 2048x1536 per stack (add + compute)
    W, H, n = 2048, 1536, 16
    fr = torch.rand((n, H, W, 3), generator=g).to(dev)
    fused = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    for pol in (_lib.FOCUS_SELECT_MAX_ENERGY, _lib.FOCUS_WEIGHTED_AVERAGE, _lib.FOCUS_LPG_WEIGHTS):
        for batched in (False, True):
            def stack(pol=pol, batched=batched):
                fs = S.c_laplacian_pyramid_focus_stacking(W, H, 3, fusing_policy=pol)
                if batched:
                    fs.add_frames([fr[i] for i in range(n)])
                else:
                    for i in range(n):
                        fs.add(fr[i])
                fs.compute(fused)
            pname = {_lib.FOCUS_SELECT_MAX_ENERGY: "max_energy", _lib.FOCUS_WEIGHTED_AVERAGE: "weighted_average",
                     _lib.FOCUS_LPG_WEIGHTS: "lpg"}[pol]
            timed(f"focus_stack_{pname}{'_batched' if batched else ''}_16x2048x1536_rgb", stack, n * 12 * W * H, n,
                  "frames")
    # input stage: 12-bit RGGB frames of 4096x3000 -> f32 BGR (read_input_frame's conversion with
    # the default NN2 debayer), one launch for the batch; and SER frames from a file to HBM
    Wi, Hi, ni = 4096, 3000, 32
    raw16 = (torch.rand((ni, Hi, Wi), generator=g) * 4095).to(torch.int16).to(dev)
    bgr = torch.empty((ni, Hi, Wi, 3), dtype=torch.float32, device=dev)
    rin = _lib.images([S.image_view(raw16[i]) for i in range(ni)])
    rout = _lib.images([S.image_view(bgr[i]) for i in range(ni)])
    timed(f"input_nn2_{ni}x{Wi}x{Hi}_u16_to_f32bgr",
          lambda: _lib.check(lib.ssg_read_input_frames(ni, rin, _lib.COLORID_BAYER_RGGB, 12, _lib.DEBAYER_NN2, None,
                                                       None, rout), "read_input_frames"),
          ni * (2 + 12) * Wi * Hi, ni, "frames")
    # the same conversion with OpenCV's bilinear (DEBAYER_NN) and edge-aware (DEBAYER_EA) demosaicing
    for algo, aname in ((_lib.DEBAYER_NN, "nn"), (_lib.DEBAYER_EA, "ea")):
        timed(f"input_opencv_{aname}_{ni}x{Wi}x{Hi}_u16_to_f32bgr",
              lambda algo=algo: _lib.check(lib.ssg_read_input_frames(ni, rin, _lib.COLORID_BAYER_RGGB, 12, algo, None,
                                                                     None, rout), "read_input_frames"),
              ni * (2 + 12) * Wi * Hi, ni, "frames")
    del bgr
    import struct
    import tempfile
    ns = 8
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "bench.ser")
        host = raw16[:ns].cpu().numpy().view(np.uint16)
        with open(path, "wb") as fh:  # SER v3 header (178 bytes, little-endian) + frames
            fh.write(struct.pack("<14s7i40s40s40sQQ", b"LUCAM-RECORDER", 0, _lib.COLORID_BAYER_RGGB, 0, Wi, Hi, 12, ns,
                                 b"", b"", b"", 0, 0))
            fh.write(np.ascontiguousarray(host).tobytes())
        rd = S.c_ser_reader(path)

        def ser_read():
            rd.seek(0)
            rd.read_batch([raw16[i] for i in range(ns)])
        timed(f"ser_read_batch_{ns}x{Wi}x{Hi}_u16", ser_read, ns * 2 * Wi * Hi, ns, "frames")
        rd.close()
    del raw16
    # lpg sharpness map of one RGB frame; local-variance weight map of a mono 1920x1080 frame
    out = torch.empty((H, W), dtype=torch.float32, device=dev)
    sv, ov = S.image_view(fr[0]), S.image_view(out)
    sh = C.c_double()
    timed("lpg_2048x1536_rgb", lambda: _lib.check(lib.ssg_lpg(C.byref(sv), C.byref(ov), 6.0, 2.0, C.byref(sh))),
          (12 + 4) * W * H, 1, "frames")
    W, H = 1920, 1080
    mono = torch.rand((H, W), generator=g).to(dev)
    mout = torch.empty((H, W), dtype=torch.float32, device=dev)
    mv, mo = S.image_view(mono), S.image_view(mout)
    q = C.c_double()
    timed("local_variance_map_1920x1080", lambda: _lib.check(lib.ssg_local_variance_map(C.byref(mv), 1, 1,
                                                                                         C.byref(mo), C.byref(q))),
          (4 + 4) * W * H, 1, "frames")
    # master-frame scan: average_bayer_planes + local-variance Q of 32 raw 12-bit RGGB 4096x3000 frames
    Wm, Hm, nm = 4096, 3000, 32
    rawm = (torch.rand((nm, Hm, Wm), generator=g) * 4095).to(torch.int16).to(dev)
    mviews = _lib.images([S.image_view(rawm[i]) for i in range(nm)])
    qs = (C.c_double * nm)()
    timed(f"master_frame_quality_{nm}x{Wm}x{Hm}_u16_rggb",
          lambda: _lib.check(lib.ssg_master_frame_quality(nm, mviews, _lib.COLORID_BAYER_RGGB, 1, 1, qs),
                             "master_frame_quality"), nm * 2 * Wm * Hm, nm, "frames")
    del rawm
    # upscale after align: 64 registration remaps of 1920x1080 to x1.5 (2880x1620)
    nu = 64
    dwu, dhu = S.upscale_size(_lib.UPSCALE_X15, W, H)
    src_maps = torch.rand((nu, H, W, 2), generator=g).to(dev) * 1000
    dst_maps = torch.empty((nu, dhu, dwu, 2), dtype=torch.float32, device=dev)
    uin = _lib.images([S.image_view(src_maps[i]) for i in range(nu)])
    uout = _lib.images([S.image_view(dst_maps[i]) for i in range(nu)])
    timed(f"upscale_remaps_x1.5_{nu}x{W}x{H}",
          lambda: _lib.check(lib.ssg_upscale_remaps(_lib.UPSCALE_X15, nu, uin, uout), "upscale_remaps"),
          nu * 8 * (W * H + dwu * dhu), nu, "maps")
    del src_maps, dst_maps
    # upscale before align: 64 mono f32 frames of 1920x1080 with their 0 / 255 masks to x1.5
    ufr = torch.rand((nu, H, W), generator=g).to(dev)
    umk = torch.where(torch.rand((nu, H, W), generator=g) < 0.01, 0, 255).to(torch.uint8).to(dev)
    ufo = torch.empty((nu, dhu, dwu), dtype=torch.float32, device=dev)
    umo = torch.empty((nu, dhu, dwu), dtype=torch.uint8, device=dev)
    fin = _lib.images([S.image_view(ufr[i]) for i in range(nu)])
    fmi = _lib.images([S.image_view(umk[i]) for i in range(nu)])
    fout = _lib.images([S.image_view(ufo[i]) for i in range(nu)])
    fmo = _lib.images([S.image_view(umo[i]) for i in range(nu)])
    timed(f"upscale_images_x1.5_{nu}x{W}x{H}_with_masks",
          lambda: _lib.check(lib.ssg_upscale_images(_lib.UPSCALE_X15, nu, fin, fmi, fout, fmo), "upscale_images"),
          nu * 5 * (W * H + dwu * dhu), nu, "frames")
    del ufr, umk, ufo, umo
    # finish of a x1.5 stack (2880x1620 RGB) with 5% scattered holes and a band of missing rows
    img = torch.rand((dhu, dwu, 3), generator=g).to(dev)
    hole = torch.rand((dhu, dwu), generator=g) < 0.05
    hole[dhu // 3:dhu // 3 + 40, :] = True
    msk = torch.where(hole, 0, 255).to(torch.uint8).to(dev)
    fimg = torch.empty_like(img)
    fmsk = torch.empty_like(msk)
    iv, mv2, fv, fm = S.image_view(img), S.image_view(msk), S.image_view(fimg), S.image_view(fmsk)
    timed(f"average_pyramid_inpaint_{dwu}x{dhu}_rgb",
          lambda: _lib.check(lib.ssg_average_pyramid_inpaint(C.byref(iv), C.byref(mv2), C.byref(fv), C.byref(fm), 100),
                             "average_pyramid_inpaint"), 2 * 13 * dwu * dhu, 1, "images")
    lin_img = torch.empty_like(img)
    lin_msk = torch.empty_like(msk)
    lv, lm = S.image_view(lin_img), S.image_view(lin_msk)

    def lin():
        lin_img.copy_(img)
        lin_msk.copy_(msk)
        _lib.check(lib.ssg_inpaint_linear(C.byref(lv), C.byref(lm)), "inpaint_linear")
    timed(f"inpaint_linear_{dwu}x{dhu}_rgb", lin, 2 * 13 * dwu * dhu, 1, "images")
    timed(f"unsharp_mask_sigma2_{dwu}x{dhu}_rgb",
          lambda: _lib.check(lib.ssg_unsharp_mask(C.byref(iv), C.byref(fv), 2.0, 0.5), "unsharp_mask"),
          2 * 12 * dwu * dhu, 1, "images")
    out_line = {"metric": "per-family kernel throughput (sweep)", "unit": "GB/s", "n_gpus": 1, "data": "synthetic",
                "dtype": "f32", "config": {"workload": "kernel sweep"}, "families": res}
    print(json.dumps(out_line))


if __name__ == "__main__":
    main()
