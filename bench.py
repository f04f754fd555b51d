"""Throughput of the gfx950 Shack-Hartmann engine: subaperture images + centroids per second.

Default workload: the 80x80 WFS at nfft=32 (D=39 m, 0.28 obstruction, 1280x1280 phase, n=16,
nimg=16, noise off), 64 frames of synthetic von Karman phase per step, resident in HBM before the
timed region. One step = one device-resident engine call over the 64 frames. Prints one JSON line.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fft32|bf16|fp16]
                [--workload wfs80x80_nfft32|wfs40x40_noise|wfs80x80_nyquist_corr|multi_wfs|scao8x8]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp16": 2500.0, "fp16x1": 2500.0}
# engine form codes (DASP_CENT_FORM_*): the compensated DFT forms (three products of hi/lo
# splits) and the single-product FP16 one (x1)
FORM_CODE = {"fft32": 0, "bf16": 1, "fp16": 2, "fp16x1": 4, "general": 3}
FORM_DTYPE = {"fft32": "f32", "bf16": "bf16x3+f32acc", "fp16": "f16x3+f32acc", "fp16x1": "f16x1+f32acc",
              "general": "f32"}
# Vector issue peak: one wave64 VALU instruction per CU per cycle, 256 CUs at the 2.4 GHz peak
# engine clock (FP32 work here issues no packed FP32 instructions, so this is also the FP32 roof).
VALU_PEAK_GINST_S = 256 * 2.4


def algorithmic_bytes(w, nactive, frames, images, noise):
    """HBM bytes one launch must move: phase windows in, slopes out (+ images, + pre-drawn noise),
    plus the per-subaperture pupil read once per launch."""
    per = 4 * w.n * w.n + 8
    if images:
        per += 4 * w.nimg * w.nimg
    if noise:
        per += 2 * 4 * w.nimg * w.nimg
    return frames * nactive * per + nactive * 4 * w.n * w.n


def measured_traffic(workload, form):
    """HBM bytes per subaperture image of the main kernel, from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs recorded under profiles/ (lines 'traffic: <workload> <form> <bytes>'), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 4 and p[0] == "traffic:" and p[1] == workload and p[2] == form:
                return float(p[3])
    return None


def measured_valu(workload, form):
    """Wave-level VALU instructions per subaperture image of the main kernel (rocprofv3
    SQ_INSTS_VALU / images, lines 'valu: <workload> <form> <count>' under profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "valu_counts.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 4 and p[0] == "valu:" and p[1] == workload and p[2] == form:
                return float(p[3])
    return None


def valu_roof(workload, form, images, kavg):
    """Vector-issue roofline: the counted VALU instructions per image at the measured rate against
    VALU_PEAK_GINST_S (the roof that binds the FP32 kernels)."""
    per = measured_valu(workload, form)
    if per is None:
        return None
    ach = per * images / kavg / 1e9
    return {"insts_per_image": per, "achieved": ach, "peak": VALU_PEAK_GINST_S, "unit": "G wave-instr/s",
            "frac": ach / VALU_PEAK_GINST_S, "source": "profiles/valu_counts.md (SQ_INSTS_VALU per image)"}


def dft_mfma_per_image(w, form):
    per = 18 if w.nfft == 32 else 144
    return per // 3 if form.endswith("x1") else per


def dft_form_shape(w):
    """Geometries the DFT-GEMM forms cover."""
    return w.nimg == 16 and ((w.n == 16 and w.nfft == 32) or (16 < w.n <= 32 and w.nfft == 64))


def roofline(form, w, nact, frames, kavg, abytes, noise):
    """Roofline entry of the main kernel: HBM for the FP32 form, MFMA for the DFT-GEMM forms."""
    ach = abytes / kavg / 1e9
    per = measured_traffic(w.name, form)
    traffic = per * nact * frames if per is not None and not noise else None
    tsrc = "profiles/hbm_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE per image, scaled)" if traffic else None
    valu = valu_roof(w.name, form, nact * frames, kavg)
    if form == "fft32":
        return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": traffic, "traffic_source": tsrc, "kernel_avg_ms": kavg * 1e3,
                "algorithmic_bytes": abytes, "valu": valu}
    # DFT-GEMM form, v_mfma_f32_32x32x16 per subaperture image (hi/lo split, 3 products):
    # 18 for n = 16 / nfft = 32, 144 for n <= 32 / nfft = 64; a third of that with one product
    flops = dft_mfma_per_image(w, form) * 2 * 32 * 32 * 16 * nact * frames
    tf = flops / kavg / 1e12
    return {"bound": "mfma", "achieved": tf, "peak": MFMA_PEAK_TFLOPS[form], "unit": "TFLOP/s",
            "frac": tf / MFMA_PEAK_TFLOPS[form], "traffic": traffic, "traffic_source": tsrc,
            "kernel_avg_ms": kavg * 1e3,
            "mfma_flops": flops, "useful_flops": 8 * (w.nfft * w.n ** 2 + w.nfft ** 2 * w.n) * nact * frames,
            "hbm_achieved_gbs": ach, "algorithmic_bytes": abytes, "valu": valu}


def host_cpus():
    """The host CPUs this process may use and how that number was found: the scheduler affinity
    mask, bounded by the cgroup CPU quota and by OMP_NUM_THREADS when the environment sets one
    (the share of the host this process is given). Returns (threads, description dict)."""
    aff = len(os.sched_getaffinity(0))
    n = aff
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = int(q) / int(per)
            n = min(n, max(1, int(quota)))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, {"nproc": os.cpu_count(), "affinity": aff, "cgroup_quota": quota, "omp_num_threads": omp,
               "model": model}


def cpu_baseline(w, phs0, budget_s=10.0, corr=None):
    """CPU leg: the FP32 C restatement of the reference engine (oracle/cref), threaded like the
    reference's pthread pool over every host CPU this process is given (host_cpus), timed on
    repeated frames of the same workload until the budget is spent; then the FP64 oracle on
    frame 0 for the error figure.
    Returns (subap-images/s, threads, sample description, FP64 slopes of frame 0, host dict)."""
    from oracle import centoracle as O
    from oracle.cref import CRef
    from dasp_amd.util.tel import obstructed_pupil
    pup = obstructed_pupil(w.nsubx, w.n, w.obs, 0.5)
    pupfn = pup.perSubap(w.nsubx, vectorAlign=0)
    flag = pup.getSubapFlag(w.nsubx, 0.5)
    frac = (pup.getSubarea(w.nsubx) / (w.n * w.n)).astype(np.float32)
    nact = int(flag.sum())
    nthr, host = host_cpus()
    cr = CRef(w.nsubx, w.n, w.nfft, w.clip, w.nimg, w.cen, pupfn.reshape(-1, w.n, w.n), flag, frac, sig=float(w.sig),
              corrPattern=corr, parabolicFit=w.parabolicFit if corr is not None else 0)
    cr.run(phs0, nthreads=nthr)
    t0 = time.perf_counter()
    frames = 0
    while True:
        cr.run(phs0, nthreads=nthr)
        frames += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    p = O.Params(w.nsubx, w.n, w.nfft, w.clip, w.nimg, w.cen, pupfn, flag, frac, sig=float(w.sig),
                 correlationCentroiding=1 if corr is not None else 0, corrPattern=corr,
                 parabolicFit=w.parabolicFit if corr is not None else 0)
    oc, _, _ = O.run_phase(p, phs0)
    sample = ("%d frames x %d active subapertures of %s, FP32 C restatement of centmodule.c (own radix-2 FFT, "
              "no FFTW), %d threads, %.1f s" % (frames, nact, w.name, nthr, dt))
    return nact * frames / dt, nthr, sample, oc, host


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", default="fft32", choices=sorted(FORM_CODE))
    workloads = {"scao8x8": "SCAO8", "wfs40x40_noise": "NOISE40", "wfs80x80_nfft32": "WFS80",
                 "wfs80x80_nyquist_corr": "CORR80"}   # name -> dasp_amd.workloads attribute
    ap.add_argument("--workload", default="wfs80x80_nfft32", choices=sorted(workloads) + ["multi_wfs", "closed_loop80"])
    ap.add_argument("--tune-cache", default=None,
                    help="JSON file of autotuned multi-WFS launch configurations (read if present, else written)")
    ap.add_argument("--batch", action="store_true",
                    help="run the single WFS through the multi-WFS step (persistent queue kernel + graph)")
    ap.add_argument("--frames", type=int, default=0, help="frames per step (default: the workload's)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="keep the default launch grids (no single-WFS grid autotuning before timing)")
    ap.add_argument("--accurate-trig", action="store_true", help="sincosf instead of the reduced hardware sin/cos")
    ap.add_argument("--no-compare-forms", dest="compare_forms", action="store_false",
                    help="skip timing the other compute forms (geometries the DFT-GEMM forms cover)")
    ap.add_argument("--calib-mib", type=int, default=0,
                    help="after timing, copy a buffer of this many MiB once (known-size kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    from dasp_amd import workloads as W, synth
    from dasp_amd.cmod import cent
    from dasp_amd.util import centcmod

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload == "multi_wfs":
        return bench_multi(args, dev)
    if args.workload == "closed_loop80":
        return bench_loop(args, dev)
    w = getattr(W, workloads[args.workload])
    frames = args.frames or w.nframes
    S = w.nsubx * w.nsubx
    pupil = w.pupil()
    host_phs = np.zeros((w.npup, w.npup), np.float32)
    cents_h = np.zeros((w.nsubx, w.nsubx, 2), np.float32)
    bimg_h = np.zeros((w.nsubx, w.nsubx, w.nimg, w.nimg), np.float32)
    corr = corrimg = None
    if w.correlation:
        from dasp_amd.util.correlation import transformPSF
        corr = transformPSF(W.reference_images(w, pupil))
        corrimg = np.zeros_like(corr)
    cc = centcmod.centcmod(1, w.nsubx, w.cen, w.nfft, w.clip, w.nimg, w.n, w.readnoise, 0., w.addPoisson,
                           w.noiseFloor, float(w.sig), 0., 0, 1., 1, 0, 1, host_phs, pupil, None, cents_h, bimg_h,
                           0.5, 0, None, w.correlation, 0., corr, corrimg, w.threshType, 0, 0, 1, w.parabolicFit,
                           None)
    h = cc.centstruct
    if w.relativeFloor:
        cent.setRelativeFloor(h, True)
    cent.setForm(h, FORM_CODE[args.form])
    cent.setFastTrig(h, not args.accurate_trig)
    nact, npup, csz = cent.info(h)

    # synthetic phase frames, generated and kept in HBM
    phs = synth.phase_screens_torch(frames, w.npup, w.D, w.r0, w.L0, w.seed, dev, wavelength_scale=500. / 640.)
    phs[0].copy_(torch.from_numpy(w.phase(0)).to(dev))       # frame 0 = the numpy screen (error check)
    d_cents = torch.zeros((frames, S, 2), dtype=torch.float32, device=dev)
    noise = bool(w.addPoisson or w.readnoise)
    zp = zr = None
    if noise:
        m2 = S * w.nimg * w.nimg
        zp = torch.from_numpy(np.stack([cent.drawNormals(1, 2 * f, m2) for f in range(frames)])).to(dev)
        zr = torch.from_numpy(np.stack([cent.drawNormals(1, 2 * f + 1, m2) for f in range(frames)])).to(dev)
    d_corr = None
    if w.correlation:
        d_corr = torch.zeros((frames, S, csz, csz), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    wb = None
    if args.batch:
        from dasp_amd.wfsbatch import WFSBatch
        if noise:
            raise SystemExit("--batch: the noise arrays of the multi-WFS step are filled by WFSBatch.draw_noise")
        wb = WFSBatch([cc], frames)
        wb.phs[0].copy_(phs)
        wb.tune(reps=3)

    grids = {}

    def run_form(form, fast_trig=not args.accurate_trig):
        """Time args.steps steps of one compute form; returns (ms per step, kernel avg s, slopes of frame 0)."""
        cent.setForm(h, FORM_CODE[form])
        cent.setFastTrig(h, fast_trig)
        cents = wb.cents[0].view(frames, S, 2) if wb is not None else d_cents
        if args.tune and wb is None:
            # the launch-grid autotuner, before (and outside) the timed steps
            cent.setGrid(h, 0, 0)
            gi, gc, _ = cent.tuneGrid(h, phs.data_ptr(), npup * npup, frames, d_cents.data_ptr(),
                                      zp.data_ptr() if zp is not None else 0, zr.data_ptr() if zr is not None else 0,
                                      reps=3, stream=stream.cuda_stream)
            grids[(form, fast_trig)] = [gi, gc]

        def step():
            if wb is not None:
                wb.run()
                return
            cent.runDevice(h, phs.data_ptr(), npup * npup, frames, d_cents.data_ptr(), 0,
                           d_corr.data_ptr() if d_corr is not None else 0,
                           zp.data_ptr() if zp is not None else 0, zr.data_ptr() if zr is not None else 0,
                           stream.cuda_stream)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        cent.setTiming(h, True)
        cent.kernelTime(h, reset=True)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        kms, klaunch = cent.kernelTime(h, reset=True)
        cent.setTiming(h, False)
        if wb is not None:
            launches, _ = wb.info()
            kms, klaunch = launches[0]["tuned_ms"], 1
        kavg = kms / klaunch / 1e3 if klaunch > 0 else None
        if form == "general":
            kavg = None   # its timed interval is the image-formation kernel alone, not the path
        return ms, kavg, cents[0].cpu().numpy().reshape(-1, 2)

    t0 = time.perf_counter()
    ms_step, kavg, c0 = run_form(args.form)
    wall = time.perf_counter() - t0
    if args.calib_mib:
        # a plain elementwise pass of known size for scaling the FETCH_SIZE/WRITE_SIZE counters:
        # calib_mib MiB read and calib_mib MiB written by one negation kernel
        src = torch.ones(args.calib_mib * (1 << 20) // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    value = nact * frames / (ms_step / 1e3)
    abytes = algorithmic_bytes(w, nact, frames, False, noise)

    result = {
        "metric": "subap images+centroids/s @80x80 nfft=32, plain vs MFMA-DFT; RMS err vs FP64",
        "value": value,
        "unit": "subap-images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": FORM_DTYPE[args.form],
        "data": "synthetic",
        "config": {"workload": w.name, "nsubx": w.nsubx, "n": w.n, "nfft": w.nfft, "nimg": w.nimg,
                   "npup": w.npup, "frames_per_step": frames, "active_subaps": nact, "form": args.form,
                   "noise": noise, "images_returned": False, "fast_trig": not args.accurate_trig, "wall_s": wall},
    }
    if wb is not None:
        result["config"]["path"] = "multi-WFS step (1 WFS): %s" % wb.info()[0][0]
    if (args.form, not args.accurate_trig) in grids:
        # image and correlation kernel grids the autotuner chose (0: the default persistent grid,
        # -1: one workgroup per work unit)
        result["config"]["launch_grid"] = grids[(args.form, not args.accurate_trig)]
    if kavg:
        result["roofline"] = roofline(args.form, w, nact, frames, kavg, abytes, noise)
    oc = None
    if not args.no_cpu:
        # CPU leg: the C restatement timed on this host; the FP64 oracle on frame 0 gives the
        # error figure of every form.
        v, nthr, sample, oc, host = cpu_baseline(w, w.phase(0), corr=corr)
        result["cpu_baseline"] = {"value": v, "unit": "subap-images/s", "cores": nthr, "kind": "port",
                                  "sample": sample, "host": host}
    act = cc.subflag.ravel() == 1
    if oc is not None and not noise:
        d = (c0 - oc.reshape(-1, 2))[act]
        result["rms_err_vs_fp64_px"] = float(np.sqrt((d ** 2).mean()))
        result["max_err_vs_fp64_px"] = float(np.abs(d).max())
    if args.compare_forms and wb is None and dft_form_shape(w):
        # the headline compares the plain FP32 form with the MFMA DFT-GEMM forms on the same input
        # (with noise, on the same pre-drawn normals; no error figure, since a count can step)
        forms = {}
        for form in ("fft32", "fp16", "bf16", "fp16x1"):
            if form == args.form:
                fms, fk, fc = ms_step, kavg, c0
            else:
                fms, fk, fc = run_form(form)
            e = {"value": nact * frames / (fms / 1e3), "ms_per_step": fms,
                 "roofline": roofline(form, w, nact, frames, fk, abytes, noise) if fk else None}
            if (form, not args.accurate_trig) in grids:
                e["launch_grid"] = grids[(form, not args.accurate_trig)]
            if oc is not None and not noise:
                d = (fc - oc)[act]
                e["rms_err_vs_fp64_px"] = float(np.sqrt((d ** 2).mean()))
                e["max_err_vs_fp64_px"] = float(np.abs(d).max())
            forms[form] = e
        # the plain form with the other trig mode (sincosf when the headline uses the reduced
        # hardware sin/cos, and the reverse)
        other = "fft32_sincosf" if not args.accurate_trig else "fft32_fast_trig"
        fms, fk, fc = run_form("fft32", fast_trig=args.accurate_trig)
        e = {"value": nact * frames / (fms / 1e3), "ms_per_step": fms, "fast_trig": args.accurate_trig,
             "roofline": roofline("fft32", w, nact, frames, fk, abytes, noise) if fk else None}
        if e["roofline"] is not None:
            e["roofline"]["valu"] = None     # the counted instructions are the headline trig mode's
        if oc is not None and not noise:
            d = (fc - oc)[act]
            e["rms_err_vs_fp64_px"] = float(np.sqrt((d ** 2).mean()))
            e["max_err_vs_fp64_px"] = float(np.abs(d).max())
        forms[other] = e
        for k, e in forms.items():   # step-time ratio of the plain FP32 form to each form
            e["speedup_vs_fft32"] = forms["fft32"]["ms_per_step"] / e["ms_per_step"]
        cent.setForm(h, FORM_CODE[args.form])
        cent.setFastTrig(h, not args.accurate_trig)
        result["forms"] = forms
    if wb is not None:
        wb.free()
    cc.free()
    print(json.dumps(result), flush=True)


def bench_multi(args, dev):
    """Multi-WFS step: 6 LGS 80x80 (n16/nfft32/nimg16, per-subaperture spot kernels), 3 NGS 8x8
    (n32/nfft64/nimg8) and 1 truth 40x40 (n16/nfft32/nimg16), Poisson + read noise, independent
    seeded phase; every frame of every WFS in one step graph (persistent launches per shape)."""
    import torch
    from dasp_amd import workloads as W, synth
    from dasp_amd.util import centcmod
    from dasp_amd.wfsbatch import WFSBatch
    frames = args.frames or 16
    geo = [(W.LGS80, k) for k in range(6)] + [(W.NGS8, 6 + k) for k in range(3)] + [(W.TRUTH40, 9)]
    ccs, hosts, spot_cache = [], [], {}
    for w, k in geo:
        pupil = w.pupil()
        host_phs = np.zeros((w.npup, w.npup), np.float32)
        cents_h = np.zeros((w.nsubx, w.nsubx, 2), np.float32)
        bimg_h = np.zeros((w.nsubx, w.nsubx, w.nimg, w.nimg), np.float32)
        if w.name not in spot_cache:
            spot_cache[w.name] = w.spots()
        spots = spot_cache[w.name]
        cc = centcmod.centcmod(1, w.nsubx, w.cen, w.nfft, w.clip, w.nimg, w.n, w.readnoise, 0., w.addPoisson,
                               w.noiseFloor, float(w.sig), 0., 0, 1., 1, 0, 1 + k, host_phs, pupil, spots, cents_h,
                               bimg_h, 0.5, 0, None, 0, 0., None, None, w.threshType, 0, 0, 1, 0, None)
        if w.relativeFloor:
            cent_mod().setRelativeFloor(cc.centstruct, True)
        cent_mod().setFastTrig(cc.centstruct, not args.accurate_trig)
        ccs.append(cc)
        hosts.append((host_phs, cents_h, bimg_h))
    wb = WFSBatch(ccs, frames)
    gen = torch.Generator(device=dev)
    for k, (w, seed) in enumerate(geo):
        wb.phs[k].copy_(synth.phase_screens_torch(frames, w.npup, w.D, w.r0, w.L0, 10 + seed, dev,
                                                  wavelength_scale=500. / 640.))
        gen.manual_seed(1000 + seed)
        for z in (wb.zpois[k], wb.zread[k]):
            if z is not None:
                z.copy_(torch.randn(z.shape, generator=gen, device=dev))
    torch.cuda.synchronize()
    launches, others = wb.tune(reps=3, cache=args.tune_cache)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        wb.run()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        wb.run()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    per_step = wb.subap_images_per_step()
    result = {
        "metric": "subap images+centroids/s, multi-WFS step (6 LGS + 3 NGS + 1 truth)",
        "value": per_step / (ms / 1e3), "unit": "subap-images/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "multi_wfs", "frames_per_step": frames, "subap_images_per_step": per_step,
                   "launches": launches, "wfs_outside_persistent_launches": others,
                   "fast_trig": not args.accurate_trig, "wall_s": wall},
    }
    wb.free()
    for cc in ccs:
        cc.free()
    print(json.dumps(result), flush=True)


def bench_loop(args, dev):
    """One closed-loop step of the 80x80 system with every array in HBM: three atmosphere layers
    advanced (makeLayer: one new column and one new row each) and summed into the pupil phase
    (createPupilPhs, float64), the 81x81-actuator spline mirror refitted and added along the line
    of sight straight into the WFS's phase buffer, the 80x80 nfft32 WFS (one frame), and the
    reconstructor step (9264 slopes x 6561 actuators). Phase, slopes and actuators never leave the
    device; the time per step is the loop's latency."""
    import torch
    from dasp_amd import workloads as W
    from dasp_amd.cmod import cent
    from dasp_amd.science.atmos import DeviceAtmos
    from dasp_amd.science.dm import DeviceXinterpDM
    from dasp_amd.science.tomorecon import DeviceRecon
    from dasp_amd.util import centcmod
    w = W.WFS80
    npup, nact = w.npup, w.nsubx + 1
    pupil = w.pupil()
    pfn = pupil.fn.astype(np.float32)
    host_phs = np.zeros((npup, npup), np.float32)
    cents_h = np.zeros((w.nsubx, w.nsubx, 2), np.float32)
    bimg_h = np.zeros((w.nsubx, w.nsubx, w.nimg, w.nimg), np.float32)
    cc = centcmod.centcmod(1, w.nsubx, w.cen, w.nfft, w.clip, w.nimg, w.n, 0., 0., 0, 0., 1., 0., 0, 1., 1, 0, 1,
                           host_phs, pupil, None, cents_h, bimg_h, 0.5, 0, None, 0, 0., None, None, 0, 0, 0, 1, 0, None)
    cent.setFastTrig(cc.centstruct, not args.accurate_trig)
    rng = np.random.default_rng(1)
    side = npup + 64
    layers = [dict(key="L%d" % k, screen=rng.standard_normal((side, side)) * (0.5 / (k + 1)), x0=8 + k, y0=8 + 2 * k,
                   pos_col=0.0, pos_row=0.0, col_add=-1, row_add=-1 if k % 2 == 0 else 1) for k in range(3)]
    da = DeviceAtmos(npup, layers, pupfn=pfn, source_lam=640.)
    dm = DeviceXinterpDM(npup, nact, pfn, actCoupling=0.1)
    ncents = 2 * int(cc.subflag.sum())
    R = torch.randn((ncents, nact * nact), device=dev, dtype=torch.float32) * 1e-3
    rec = DeviceRecon([cc.subflag], R, gains=0.3, decay=0.99)
    phs32 = torch.zeros((npup, npup), dtype=torch.float32, device=dev)
    atm64 = torch.empty((npup, npup), dtype=torch.float64, device=dev)
    new_cols = torch.randn((1, side), dtype=torch.float64, device=dev)
    new_rows = torch.randn((1, side), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    st = stream.cuda_stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
    stage_ms = np.zeros(5)

    def step(k, timed=False):
        icol = {L["key"]: (0.37 * k) % 1.0 for L in layers}
        irow = {L["key"]: (0.21 * k) % 1.0 for L in layers}
        if timed:
            ev[0].record(stream)
        for L in layers:
            da.make_layer(L["key"], new_cols, new_rows)
        if timed:
            ev[1].record(stream)
        da.create_pupil_phs(icol, irow, out64=atm64)
        if timed:
            ev[2].record(stream)
        dm.generateNext(phs32, atmos=atm64, actmap=rec.acts[0].view(nact, nact))
        if timed:
            ev[3].record(stream)
        cent.runDevice(cc.centstruct, phs32.data_ptr(), npup * npup, 1, rec.slope_buffer(0).data_ptr(), stream=st)
        if timed:
            ev[4].record(stream)
        rec.run()
        if timed:
            ev[5].record(stream)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(args.steps):
        step(args.warmup + k)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    for k in range(5):   # per-stage times in separately timed steps (events between the stages)
        step(10_000 + k, timed=True)
        torch.cuda.synchronize()
        stage_ms += [ev[i].elapsed_time(ev[i + 1]) for i in range(5)]
    stage_ms /= 5
    # createPupilPhs: three (npup + 1)^2 float64 windows read, the sum written and read back, the
    # pupil read twice, the float64 phase written
    atm_bytes = 3 * (npup + 1) ** 2 * 8 + 3 * npup * npup * 8 + 2 * npup * npup * 4
    result = {
        "metric": "closed-loop step time, 80x80 WFS + 3-layer atmosphere + 81x81 spline DM + reconstructor, all in HBM",
        "value": ms, "unit": "ms/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64 phase, f32 WFS/DM/recon",
        "data": "synthetic",
        "config": {"workload": "closed_loop80", "npup": npup, "layers": 3, "nact": nact, "ncents": ncents,
                   "nacts": nact * nact, "wall_s": wall},
        "stages_ms": {"make_layer x3": stage_ms[0], "create_pupil_phs": stage_ms[1], "dm_fit_and_line_of_sight": stage_ms[2],
                      "wfs_80x80": stage_ms[3], "recon": stage_ms[4]},
        "atmos_roofline": {"bound": "hbm", "achieved": atm_bytes / (stage_ms[1] / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": atm_bytes / (stage_ms[1] / 1e3) / 1e9 / HBM_PEAK_GBS,
                           "algorithmic_bytes": atm_bytes},
    }
    rec.free()
    cc.free()
    print(json.dumps(result), flush=True)


def cent_mod():
    from dasp_amd.cmod import cent
    return cent


if __name__ == "__main__":
    main()
