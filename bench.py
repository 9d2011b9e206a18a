"""Batched STEM benchmark: probe-slices/s at 1024^2 FP32 on one MI355X.

Workload (batched STEM): SrTiO3 perovskite (a = 3.905 A) 10 x 10 x 40 cells on a
1024 x 1024 grid, 80 slices of 1.9525 A, 300 kV probe with 25 mrad semi-angle,
Cs = 0, defocus 0, 64 x 64 raster, annular detector 60-200 mrad, frozen
phonons with Gaussian u_rms = 0.08 A (seeded).  One step = one frozen-phonon
configuration: displace the atoms, slice and bin them, build the 80 potential
slices, then run all 4096 probes through the fused multislice in batches,
ending in detector sums.  Inputs are synthetic; everything the step touches on
the GPU is resident in HBM before the timed region starts.

Prints one JSON line.  --gpus N > 1 is refused: this build measures one GPU.
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

N_GRID = 1024
CELLS = (10, 10, 40)
A_SRTIO3 = 3.905
DZ = A_SRTIO3 / 2
E0 = 300.0
ALPHA = 25.0
DET = (60.0, 200.0)
SCAN = 64
U_RMS = 0.08
SEED = 1234
HBM_PEAK = 8.0e12
# HBM traffic per probe of one launch of the fused passes at 1024^2, groups of
# 37 probes, band skip 1e-9, row_blocks 768 / col_blocks 4096 (the autotuner's
# usual choice): rocprofv3 --pmc FETCH_SIZE (x 2.000) and WRITE_SIZE (x 1.000),
# factors from a 1 GiB device copy; profiles/batched_stem_1024_hbm_traffic.md
PMC_BYTES_PER_PROBE = {"column pass": (265.74 + 229.18) / 37 * 2**20, "row pass": (230.21 + 222.92) / 37 * 2**20}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--group", type=int, default=0, help="probes kept together per fused pass (0 = auto)")
    ap.add_argument("--cpu-probes", type=int, default=16, help="probes of the CPU baseline (one thread each)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--npos", type=int, default=0, help="scan positions per step (0 = the full 64x64 raster)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--row-blocks", type=int, default=0,
                    help="row-pass workgroup target instead of the default (with --no-autotune)")
    ap.add_argument("--col-blocks", type=int, default=0,
                    help="column-pass workgroup target instead of the default (with --no-autotune)")
    ap.add_argument("--plain-probes", type=int, default=4)
    ap.add_argument("--no-band-skip", action="store_true",
                    help="compute every propagator column (no band-limit skipping)")
    ap.add_argument("--only-stem", action="store_true",
                    help="time the batched STEM steps only (profiling runs): no pass timing, plain path, "
                         "adjoint or CPU baseline")
    ap.add_argument("--grid", type=int, default=N_GRID,
                    help="grid size of the batched STEM steps; other than 1024 only with --only-stem")
    ap.add_argument("--only-ragged", action="store_true",
                    help="run only the ragged mix, three times (profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device-to-device 5 times (known bytes for HBM counter scaling)")
    return ap.parse_args()


def cpu_baseline(at, xs, ys, ngpu_sums, n_probe):
    """FP64 host oracle on a bounded sample of the same workload: n_probe probes
    of phonon configuration 1 through all 80 slices (the slice loop is timed),
    OpenMP over probes.  Also returns the GPU result's error against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import concurrent.futures as cf
    import pyoracle as O

    lam, _, sig = O.constants(E0)
    order = np.argsort(at["z"], kind="stable")
    x, y, z, occ, Z = (at[k][order] for k in ("x", "y", "z", "occ", "Z"))
    bounds, i0, i1 = O.slicing(z, DZ, O.BORDER_MIN)
    types = sorted(set(int(v) for v in Z))
    tab = O.Tables(types)
    tidx = np.searchsorted(np.array(types), Z).astype(np.int32)
    S = len(i0)
    n = N_GRID
    with cf.ThreadPoolExecutor(max_workers=16) as ex:
        Vo = np.stack(list(ex.map(lambda s: O.potential(n, n, at["lx"], at["ly"], i0[s], i1[s], x, y, occ, tidx,
                                                         tab, sig), range(S))))
    Po = O.propagator(n, n, at["lx"], at["ly"], DZ, True, lam)
    lens = {"outer_aper_mrad": ALPHA}
    zmid0 = 0.5 * (bounds[0] + bounds[1])
    g2min = np.array([(np.sin(DET[0] * 1e-3) / lam) ** 2])
    g2max = np.array([(np.sin(DET[1] * 1e-3) / lam) ** 2])
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    ref, used = O.stem(n, n, at["lx"], at["ly"], lam, lens, zmid0, Po, Vo, xs[:n_probe], ys[:n_probe], g2min,
                       g2max, nthreads=threads)
    dt = time.perf_counter() - t0
    rel = float(np.linalg.norm(ngpu_sums[:n_probe] - ref[0]) / np.linalg.norm(ref[0]))
    # the same loop in FP32 (the GPU's arithmetic), timed beside it
    t0 = time.perf_counter()
    r32, _ = O.stem(n, n, at["lx"], at["ly"], lam, lens, zmid0, Po, Vo, xs[:n_probe], ys[:n_probe], g2min,
                    g2max, nthreads=threads, fp32=True)
    dt32 = time.perf_counter() - t0
    fp32 = {"value": n_probe * S / dt32, "rel_l2_vs_fp64": float(np.linalg.norm(r32[0] - ref[0]) / np.linalg.norm(ref[0]))}
    # threads with work: the oracle runs one probe per thread
    return n_probe * S / dt, min(used, n_probe), rel, n_probe * S, fp32


def plain_path(em, torch, at, xs, ys, n_probe):
    """The reference's one-probe-per-launch STEM path (run_STEM_general_simu,
    scan_simulator.cuh:114-183) on this build's kernels: per probe, re-cut
    the atoms around the probe (2 A margin), slice them, rebuild every
    potential slice, build the probe, run the slices, sum the detector and
    copy the result to the host.  Returns probe-slices/s."""
    L = at["lx"]
    plan = em.Plan(N_GRID, N_GRID, L, L, DZ, E0, max_batch=1)
    lens = {"outer_aper_mrad": ALPHA}
    done = 0

    def one(i):
        c = em.specimen.cut_window(at, xs[i] - L / 2, xs[i] + L / 2, ys[i] - L / 2, ys[i] + L / 2, margin=2.0)
        S = plan.set_specimen(c["x"], c["y"], c["z"], c["Z"], c["occ"], border=em.BORDER_MIN)
        zs, ze, _, _ = plan.slice_info(0)
        V = plan.potential_slices()
        psi = plan.build_probes(lens, [0.0], [0.0], defocus_shift=0.5 * (zs + ze))
        plan.multislice_forward(psi, V=V)
        plan.detector_sums(psi, DET[0], DET[1]).cpu()
        return S

    one(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(n_probe):
        done += one(i)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plan.close()
    return done / dt


def copy_bandwidth(torch, nbytes=1 << 30, reps=5):
    """Achievable streaming bandwidth (read + write bytes per second) on 1 GiB
    float32 buffers of random values, timed with events on torch's stream.
    Returns (copy, in_place):
      copy      best of a device copy (b.copy_(a)) and an elementwise kernel
                that writes another buffer (torch.neg(a, out=b));
      in_place  an elementwise kernel that reads and rewrites one buffer
                (a.neg_()), the access pattern of the fused passes."""
    a = torch.randn(nbytes // 4, device="cuda")
    b = torch.empty_like(a)

    def rate(f):
        f()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        best = float("inf")
        for _ in range(reps):
            e0.record()
            f()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return 2 * nbytes / (best * 1e-3)

    copy = max(rate(lambda: b.copy_(a)), rate(lambda: torch.neg(a, out=b)))
    in_place = rate(lambda: a.neg_())
    del a, b
    return copy, in_place


def adjoint_ratio(em, torch, n=512, S=40, scan=16, reps=3):
    """Differentiable inverse multislice workload: 512^2, 40 slices of smooth
    random transmission (phase in [0, 0.5] rad, amplitude 0.95-1), 16 x 16
    probes at 300 kV / 20 mrad, loss = sum of squared far-field intensity
    differences to a seeded target.  Returns (forward s, forward+adjoint s)."""
    g = torch.Generator(device="cuda").manual_seed(5)
    L = 30.0
    k = torch.fft.fftfreq(n, device="cuda")
    lp = torch.exp(-(k[None, :] ** 2 + k[:, None] ** 2) / 2e-3)

    def smooth(lo, hi):
        f = torch.randn((S, n, n), generator=g, device="cuda")
        f = torch.fft.ifft2(torch.fft.fft2(f) * lp).real
        f = (f - f.amin()) / (f.amax() - f.amin())
        return lo + (hi - lo) * f

    t = (smooth(0.95, 1.0) * torch.exp(1j * smooth(0.0, 0.5))).to(torch.complex64).contiguous()
    plan = em.Plan(n, n, L, L, 2.0, 300.0, max_batch=scan * scan)
    xs, ys = em.specimen.raster(scan, scan, L, L)
    psi0 = plan.build_probes({"outer_aper_mrad": 20.0}, xs, ys)
    B = psi0.shape[0]
    target = torch.rand((B, n, n), generator=g, device="cuda")
    ck = torch.empty((S, B, n, n), dtype=torch.complex64, device="cuda")

    def fwd():
        psi = psi0.clone()
        plan.multislice_forward(psi, t=t)
        I = psi.real ** 2 + psi.imag ** 2
        return ((I - target) ** 2).sum()

    def fwd_adj():
        psi = psi0.clone()
        plan.multislice_forward(psi, t=t, ckpt=ck)
        I = psi.real ** 2 + psi.imag ** 2
        gPsi = (4 * (I - target) * psi).contiguous()
        gt = plan.multislice_adjoint(ck, gPsi, t=t)
        return gt

    res = []
    for f in (fwd, fwd_adj):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        res.append((time.perf_counter() - t0) / reps)
    plan.close()
    del ck
    return res[0], res[1]


def ragged_mix(em, torch, n_req=8, seed=7, reps=3):
    """Ragged batch in one launch: CBED/STEM requests on grids of 512^2, 768^2,
    1024^2 and 2048^2 (seeded mix), Si diamond (a = 5.431 A) 50-400 A thick
    (dz = a/4), detector sums every 10 slices, 200 / 300 kV, 10-30 mrad.
    Returns (probe-slices/s, probe-slices) of the ragged launch: the median of
    `reps` timed launches (one launch takes about 60 ms, and single timings
    right after the STEM steps vary by about 10%)."""
    rng = np.random.default_rng(seed)
    a = 5.431
    plans, Vs, psis = [], [], []
    work = 0
    for r in range(n_req):
        n = int(rng.choice([512, 768, 1024, 2048]))
        kv = float(rng.choice([200.0, 300.0]))
        alpha = float(rng.uniform(10.0, 30.0))
        thick = float(rng.uniform(50.0, 400.0))
        nz = max(1, int(round(thick / a)))
        cells = 5
        at = em.specimen.si_diamond(cells, cells, nz, a=a)
        L = at["lx"]
        B = 4 if n <= 1024 else 2
        p = em.Plan(n, n, L, L, a / 4, kv, max_batch=B)
        S = p.set_specimen(at["x"], at["y"], at["z"], at["Z"], at["occ"])
        Vs.append(p.potential_slices())
        xs = rng.uniform(0, L, B)
        ys = rng.uniform(0, L, B)
        zs, ze, _, _ = p.slice_info(0)
        psis.append(p.build_probes({"outer_aper_mrad": alpha}, xs, ys, defocus_shift=0.5 * (zs + ze)))
        plans.append(p)
        work += B * S
    inputs = [q.clone() for q in psis]
    em.ragged(plans, Vs, psis, [10.0, 40.0], [40.0, 150.0], stride=10)  # warm
    dts = []
    for _ in range(reps):
        for q, q0 in zip(psis, inputs):
            q.copy_(q0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        em.ragged(plans, Vs, psis, [10.0, 40.0], [40.0, 150.0], stride=10)
        torch.cuda.synchronize()
        dts.append(time.perf_counter() - t0)
    for p in plans:
        p.close()
    return work / sorted(dts)[len(dts) // 2], work


def main():
    global N_GRID
    args = parse()
    if args.grid != N_GRID and not args.only_stem:
        print(json.dumps({"error": "--grid other than 1024 needs --only-stem (the roofline's traffic "
                                   "constants are measured at 1024^2)"}))
        sys.exit(2)
    N_GRID = args.grid
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    import torch
    import cudaem_amd as em

    torch.cuda.init()
    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": a.numel() * 4, "bytes_written": a.numel() * 4,
                          "launches": 5}))
        return
    if args.only_ragged:
        for _ in range(3):
            rag_ps, rag_work = ragged_mix(em, torch, reps=1)
            print(json.dumps({"metric": "probe-slices/s (ragged mix only)", "value": rag_ps,
                              "probe_slices": rag_work}), flush=True)
        return
    at0 = em.specimen.srtio3(*CELLS, a=A_SRTIO3)
    L = at0["lx"]
    xs, ys = em.specimen.raster(SCAN, SCAN, L, L)
    if args.npos:
        xs, ys = xs[:args.npos], ys[:args.npos]
    npos = len(xs)
    plan = em.Plan(N_GRID, N_GRID, L, L, DZ, E0, max_batch=args.batch)
    plan.set_probe_group(args.group)
    band = plan.set_band_skip(0.0 if args.no_band_skip else 1e-9)
    tune_ms = None
    if args.no_autotune:
        if args.row_blocks:
            plan.set_tuning(row_blocks=args.row_blocks)
        if args.col_blocks:
            plan.set_tuning(col_blocks=args.col_blocks)
    else:
        tuned, tune_ms = plan.autotune(min(args.batch, npos), nslices=8)
        if args.group:
            plan.set_tuning(probe_group=args.group)
    lens = {"outer_aper_mrad": ALPHA}
    out = torch.zeros((1, npos), dtype=torch.float32, device="cuda")

    def step(iconf, w, accumulate):
        at = em.specimen.vibrate(at0, U_RMS, SEED, iconf)
        S = plan.set_specimen(at["x"], at["y"], at["z"], at["Z"], at["occ"], border=em.BORDER_MIN)
        plan.stem(lens, xs, ys, DET[0], DET[1], w=w, out=out, accumulate=accumulate)
        return S

    for i in range(args.warmup):
        step(1 + i, 1.0, False)
    torch.cuda.synchronize()
    # the timed steps are the frozen-phonon configurations 1..steps of one
    # image: each adds its detector sums with weight 1/steps
    w = 1.0 / args.steps
    out.zero_()
    t0 = time.perf_counter()
    S = 0
    sums = None
    for i in range(args.steps):
        S = step(1 + i, w, True)
        if i == 0:
            torch.cuda.synchronize()
            sums = out[0].cpu().numpy().copy() / w  # configuration 1 alone, for the oracle check
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ps = args.steps * npos * S / dt

    if args.only_stem:
        print(json.dumps({"metric": "probe-slices/s (batched STEM steps only)", "value": ps, "steps": args.steps,
                          "grid": N_GRID, "probes_per_step": npos, "probe_group": plan.group_size(args.batch),
                          "tuning": plan.tuning()}), flush=True)
        return

    # Dominant kernels at the real shape: one batch of the same STEM slice
    # loop run eagerly, every interior row / column pass launch bracketed by
    # HIP events on the plan's stream; averages over all those launches.
    N = N_GRID * N_GRID
    tun = plan.tuning()
    nb = min(args.batch, npos)
    G = plan.group_size(nb)
    probes_per_launch = nb / -(-nb // G)
    tmp = torch.zeros((1, nb), dtype=torch.float32, device="cuda")
    plan.set_pass_timing(True)
    plan.stem(lens, xs[:nb], ys[:nb], DET[0], DET[1], w=1.0, out=tmp, accumulate=False)
    row_ms, col_ms, n_row, n_col = plan.pass_times()
    plan.set_pass_timing(False)
    # Algorithmic bytes per launch: one read and one write of the columns the
    # pass must touch (band skipping leaves out whole propagator columns; the
    # column pass skips whole tiles of them), plus the shared table (t: every
    # pixel; P: the kept columns).
    lo, hi = band
    skip_row = hi - lo + 1 if hi >= lo else 0
    tile = tun["col_tile"] or 8
    skip_col = tile * max(0, (hi + 1) // tile - -(-lo // tile)) if hi >= lo else 0
    row_bytes = 16 * N_GRID * (N_GRID - skip_row) * probes_per_launch + 8 * N
    col_bytes = 16 * N_GRID * (N_GRID - skip_col) * probes_per_launch + 8 * N_GRID * (N_GRID - skip_col)
    dom = ("column pass", col_ms, col_bytes) if col_ms >= row_ms else ("row pass", row_ms, row_bytes)
    achieved = dom[2] / (dom[1] * 1e-3)
    traffic = PMC_BYTES_PER_PROBE[dom[0]] * probes_per_launch
    copy_bw, inplace_bw = copy_bandwidth(torch)

    plain_ps = plain_path(em, torch, em.specimen.vibrate(at0, U_RMS, SEED, 1), xs, ys, args.plain_probes)
    t_fwd, t_fwd_adj = adjoint_ratio(em, torch)
    rag_ps, rag_work = ragged_mix(em, torch)

    cpu = None
    if not args.no_cpu:
        at1 = em.specimen.vibrate(at0, U_RMS, SEED, 1)
        v, used, rel, units, fp32 = cpu_baseline(at1, xs, ys, sums, args.cpu_probes)
        cpu = {"value": v, "unit": "probe-slices/s", "cores": used, "kind": "port",
               "sample": f"{args.cpu_probes} probes x {S} slices of phonon configuration 1 "
                         f"({units} probe-slices, FP64 host oracle, OpenMP over probes)",
               "gpu_rel_l2_vs_oracle": rel,
               "fp32_variant": fp32}

    line = {
        "metric": "probe-slices/s at 1024^2 FP32, % HBM roofline; batched/plain and adjoint/fwd ratio",
        "value": ps,
        "unit": "probe-slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "batched STEM, SrTiO3 10x10x40 cells, 1024^2 grid, 80 slices, 300 kV, "
                               "25 mrad, 64x64 scan, 1 frozen-phonon configuration per step, the steps' detector "
                               "sums averaged into one image (weight 1/steps)",
                   "grid": N_GRID, "slices": S, "probes_per_step": npos, "batch": args.batch,
                   "probe_group": G, "tuning": tun, "autotune_ms_per_probe_slice": tune_ms,
                   "phonon_u_rms": U_RMS, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "launch_ms": dom[1], "launches_timed": n_col if dom[0] == "column pass" else n_row,
                     "algorithmic_bytes_per_launch": dom[2], "probes_per_launch": probes_per_launch,
                     "row_pass_ms": row_ms, "col_pass_ms": col_ms,
                     "row_pass_frac": row_bytes / (row_ms * 1e-3) / HBM_PEAK,
                     "copy_kernel_GBps": copy_bw / 1e9, "frac_of_copy": achieved / copy_bw,
                     "in_place_kernel_GBps": inplace_bw / 1e9, "frac_of_in_place": achieved / inplace_bw,
                     "traffic_source": "rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE per launch (copy-kernel calibrated), "
                                       "per probe, scaled to probes_per_launch"},
        # whole slice step against one read + write of the wave per probe-slice
        # (16 B per pixel) plus the shared potential and propagator tables
        "slice_roofline_frac": ps * (16 * N + 12 * N / G) / HBM_PEAK,
        # arithmetic, reported only (HBM bounds the path): two 2D FFT pairs'
        # worth of radix-2 flops per probe-slice, 2 * 5 N log2 N
        "flops": {"per_probe_slice": 10.0 * N * math.log2(N),
                  "achieved_TFLOPs": ps * 10.0 * N * math.log2(N) / 1e12,
                  "fp32_peak_TFLOPs": 157.3},  # published dense FP32 rate
        "band_skip": {"threshold": 0.0 if args.no_band_skip else 1e-9, "skipped_columns": band[1] - band[0] + 1 if band[1] >= band[0] else 0,
                      "of": N_GRID},
        "batched_vs_plain": {"ratio": ps / plain_ps, "plain_probe_slices_per_s": plain_ps,
                             "plain_path": "one probe per launch, atoms re-cut and potentials rebuilt per probe "
                                           f"(general scanner), {args.plain_probes} probes"},
        "ragged_mix": {"probe_slices_per_s": rag_ps, "probe_slices": rag_work,
                       "workload": "8 seeded requests on 512^2/768^2/1024^2/2048^2 grids, Si 50-400 A, "
                                   "200/300 kV, 10-30 mrad, detector sums every 10 slices, one graph launch; "
                                   "median of 3 timed launches"},
        "adjoint_vs_forward": {"ratio": t_fwd_adj / t_fwd, "forward_s": t_fwd, "forward_plus_adjoint_s": t_fwd_adj,
                               "workload": "512^2, 40 slices, 16x16 probes, 300 kV / 20 mrad, "
                                           "checkpoint every slice, default launch shapes"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
