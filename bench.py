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
# HBM traffic of the fused passes at 1024^2 per probe and launch, from
# rocprofv3 --pmc FETCH_SIZE (x 2.000) and WRITE_SIZE (x 1.000) with factors
# from a 1 GiB device copy (profiles/batched_stem_1024_hbm_traffic.md), and
# the launch shapes they were measured at (the plan defaults at 1024^2).  The
# bench reports it only when its own tuning moves the same bytes: the same
# probe groups, chunks, tiles, column exchange form, transmission form and
# pass structure.  Radix sequences load and store the same elements (16.16.4
# and 16.4.16 rows measured 116,548 and 116,542 KiB per launch), so they are
# not compared.
PMC_SHAPES = {"row_blocks": 768, "col_blocks": 4096, "col_tile": 0, "col_variant": 1, "row_variant": 0,
              "persistent": 0, "probe_group": 37}
PMC_BYTES_PER_PROBE = {"column pass": (132279 * 2 + 236486) * 1024 / 37,
                       "row pass": (116542 * 2 + 224695) * 1024 / (256 / 7)}
T_START = time.perf_counter()


def note(msg):
    """One progress line per finished stage, on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench {time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--group", type=int, default=0, help="probes kept together per fused pass (0 = auto)")
    ap.add_argument("--cpu-probes", type=int, default=0,
                    help="probes of the CPU baseline, one per thread (0 = one per usable core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--npos", type=int, default=0, help="scan positions per step (0 = the full 64x64 raster)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--row-blocks", type=int, default=0,
                    help="row-pass workgroup target instead of the default (with --no-autotune)")
    ap.add_argument("--col-blocks", type=int, default=0,
                    help="column-pass workgroup target instead of the default (with --no-autotune)")
    ap.add_argument("--row-radix", type=int, default=-1,
                    help="row-pass radix sequence instead of the default (with --no-autotune)")
    ap.add_argument("--col-radix", type=int, default=-1,
                    help="column-pass radix sequence instead of the default (with --no-autotune)")
    ap.add_argument("--plain-probes", type=int, default=32,
                    help="probes of the one-probe-per-launch side of batched_vs_plain (each grid)")
    ap.add_argument("--persistent", action="store_true",
                    help="run the interior of each probe group's slice loop as one persistent launch "
                         "(with --only-stem)")
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


def cpu_cores():
    """(threads the CPU baseline uses, what nproc reports).  A one-GPU box
    gives its job a 16-core share whatever the machine's count is; the
    baseline runs one probe per core of that share (or of the affinity mask,
    if smaller)."""
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = nproc
    return max(1, min(16, usable)), nproc


def cpu_baseline(at, xs, ys, ngpu_sums, n_probe, threads):
    """FP64 host oracle on a bounded sample of the same workload: n_probe probes
    of phonon configuration 1 through all 80 slices (the slice loop is timed),
    one OpenMP thread per probe.  Also returns the GPU result's error against
    it."""
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
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        Vo = np.stack(list(ex.map(lambda s: O.potential(n, n, at["lx"], at["ly"], i0[s], i1[s], x, y, occ, tidx,
                                                         tab, sig), range(S))))
    Po = O.propagator(n, n, at["lx"], at["ly"], DZ, True, lam)
    lens = {"outer_aper_mrad": ALPHA}
    zmid0 = 0.5 * (bounds[0] + bounds[1])
    g2min = np.array([(np.sin(DET[0] * 1e-3) / lam) ** 2])
    g2max = np.array([(np.sin(DET[1] * 1e-3) / lam) ** 2])
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


def plain_path(em, torch, at, xs, ys, n_probe, n_grid):
    """The reference's one-probe-per-launch STEM path (run_STEM_general_simu,
    scan_simulator.cuh:114-183) on this build's kernels: per probe, re-cut
    the atoms around the probe (2 A margin), slice them, rebuild every
    potential slice, build the probe, run the slices, sum the detector and
    copy the result to the host.  Returns probe-slices/s."""
    L = at["lx"]
    plan = em.Plan(n_grid, n_grid, L, L, DZ, E0, max_batch=1)
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


def batched_steps(em, torch, at0, xs, ys, n_grid, steps=2, batch=256):
    """probe-slices/s of the batched STEM steps on another grid (default
    launch shapes, one frozen-phonon configuration per step)."""
    L = at0["lx"]
    plan = em.Plan(n_grid, n_grid, L, L, DZ, E0, max_batch=batch)
    out = torch.zeros((1, len(xs)), dtype=torch.float32, device="cuda")
    lens = {"outer_aper_mrad": ALPHA}

    def step(i):
        at = em.specimen.vibrate(at0, U_RMS, SEED, i)
        S = plan.set_specimen(at["x"], at["y"], at["z"], at["Z"], at["occ"], border=em.BORDER_MIN)
        plan.stem(lens, xs, ys, DET[0], DET[1], out=out, accumulate=True)
        return S

    step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S = sum(step(2 + i) for i in range(steps))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plan.close()
    return len(xs) * S / dt


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
    Returns (probe-slices/s, probe-slices, first-call s) of the ragged launch:
    the median of `reps` timed calls of the same request set, which re-launch
    the graph the first call captured (one launch takes about 50 ms, and
    single timings right after the STEM steps vary by a few %), and the first
    call's time, capture and instantiation included."""
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
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = em.ragged(plans, Vs, psis, [10.0, 40.0], [40.0, 150.0], stride=10)
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    dts = []
    for _ in range(reps):
        for q, q0 in zip(psis, inputs):
            q.copy_(q0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        em.ragged(plans, Vs, psis, [10.0, 40.0], [40.0, 150.0], stride=10, outs=outs)
        torch.cuda.synchronize()
        dts.append(time.perf_counter() - t0)
    for p in plans:
        p.close()
    return work / sorted(dts)[len(dts) // 2], work, first


def main():
    global N_GRID
    args = parse()
    if args.persistent and not (args.only_stem and args.no_autotune):
        print(json.dumps({"error": "--persistent needs --only-stem --no-autotune"}))
        sys.exit(2)
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
            rag_ps, rag_work, rag_first = ragged_mix(em, torch, reps=3)
            print(json.dumps({"metric": "probe-slices/s (ragged mix only)", "value": rag_ps,
                              "probe_slices": rag_work, "first_call_s": rag_first}), flush=True)
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
        if args.persistent:
            plan.set_tuning(persistent=1)
        if args.row_radix >= 0:
            plan.set_tuning(row_radix=args.row_radix)
        if args.col_radix >= 0:
            plan.set_tuning(col_radix=args.col_radix)
    else:
        tuned, tune_ms = plan.autotune(min(args.batch, npos), nslices=8)
        if args.group:
            plan.set_tuning(probe_group=args.group)
        note(f"autotune done: {plan.tuning()}")
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
    note(f"{args.steps} timed STEM steps: {ps:.0f} probe-slices/s")

    if args.only_stem:
        print(json.dumps({"metric": "probe-slices/s (batched STEM steps only)", "value": ps, "steps": args.steps,
                          "grid": N_GRID, "probes_per_step": npos, "probe_group": plan.group_size(args.batch),
                          "tuning": plan.tuning()}), flush=True)
        return

    # The slice step at the real shape: one batch of the same STEM slice
    # loop run eagerly, every interior row / column pass launch bracketed by
    # HIP events on the plan's stream; averages over all those launches.  One
    # slice step of a probe group is one row pass plus one column pass.
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
    step_ms = row_ms + col_ms
    note(f"event-timed batch: row {row_ms:.4f} ms, column {col_ms:.4f} ms")
    # Algorithmic bytes of one slice step over a group: per probe-slice one
    # read and one write of the wave (16 B per pixel) plus the shared potential
    # (4 B) and propagator (8 B) per pixel amortised over the group.
    slice_bytes = (16 * N + 12 * N / probes_per_launch) * probes_per_launch
    achieved = slice_bytes / (step_ms * 1e-3)
    # Each pass on its own: one read and one write of the columns it must touch
    # (band skipping leaves out whole propagator columns; the column pass skips
    # whole tiles of them), plus its table (t: every pixel; P: kept columns).
    lo, hi = band
    skip_row = hi - lo + 1 if hi >= lo else 0
    tile = tun["col_tile"] or 8
    skip_col = tile * max(0, (hi + 1) // tile - -(-lo // tile)) if hi >= lo else 0
    row_bytes = 16 * N_GRID * (N_GRID - skip_row) * probes_per_launch + 8 * N
    col_bytes = 16 * N_GRID * (N_GRID - skip_col) * probes_per_launch + 8 * N_GRID * (N_GRID - skip_col)
    shapes = {k: tun[k] for k in ("row_blocks", "col_blocks", "col_tile", "col_variant", "row_variant",
                                  "persistent")}
    shapes["col_variant"] = shapes["col_variant"] or 1  # 0 = automatic = the block exchange (1)
    shapes["probe_group"] = G
    profiled = all(PMC_SHAPES[k] == v for k, v in shapes.items())
    pmc = {k: v * probes_per_launch for k, v in PMC_BYTES_PER_PROBE.items()} if profiled else None
    copy_bw, inplace_bw = copy_bandwidth(torch)

    # batched / plain on both grids: the plain side one probe per launch on a
    # seeded sample of raster positions; the batched side the full raster
    at1 = em.specimen.vibrate(at0, U_RMS, SEED, 1)
    rng = np.random.default_rng(SEED)
    sample = rng.choice(len(xs), args.plain_probes, replace=False)
    plain_ps = plain_path(em, torch, at1, xs[sample], ys[sample], args.plain_probes, N_GRID)
    note(f"plain path at {N_GRID}^2: {plain_ps:.0f} probe-slices/s")
    xs32, ys32 = em.specimen.raster(32, 32, L, L)
    sample32 = rng.choice(len(xs32), args.plain_probes, replace=False)
    plain512 = plain_path(em, torch, at0, xs32[sample32], ys32[sample32], args.plain_probes, 512)
    batched512 = batched_steps(em, torch, at0, xs32, ys32, 512)
    note(f"512^2: plain {plain512:.0f}, batched {batched512:.0f} probe-slices/s")
    t_fwd, t_fwd_adj = adjoint_ratio(em, torch)
    note(f"forward+adjoint / forward: {t_fwd_adj / t_fwd:.3f}")
    rag_ps, rag_work, rag_first = ragged_mix(em, torch, reps=5)
    note(f"ragged mix: {rag_ps:.0f} probe-slices/s")

    cpu = None
    if not args.no_cpu:
        threads, nproc = cpu_cores()
        n_cpu = args.cpu_probes or threads
        v, used, rel, units, fp32 = cpu_baseline(at1, xs, ys, sums, n_cpu, threads)
        note(f"CPU baseline: {v:.1f} probe-slices/s on {used} threads")
        full_s = args.steps * npos * S / v
        cpu = {"value": v, "unit": "probe-slices/s", "cores": used, "kind": "port",
               "sample": f"{n_cpu} probes x {S} slices of phonon configuration 1 ({units} probe-slices, FP64 "
                         f"host oracle, one OpenMP thread per probe on {threads} threads; nproc {nproc}, the "
                         f"one-GPU box's share 16)",
               "extrapolated_timed_steps_s": full_s,
               "extrapolation": f"{args.steps} steps x {npos} probes x {S} slices at this rate: {full_s:.0f} s "
                                f"against the GPU's {dt:.2f} s",
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
        # the dominant kernel is the slice step (row + column pass of one probe
        # group): algorithmic bytes 16 N + 12 N / G per probe-slice over the
        # event-timed launch pair
        "roofline": {"bound": "hbm", "kernel": "slice step (row + column pass)", "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": (pmc["row pass"] + pmc["column pass"]) if pmc else None,
                     "traffic_source": ("rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE of both passes (copy-kernel "
                                        "calibrated), per probe at the profiled launch shapes, scaled to "
                                        "probes_per_launch") if pmc else
                                       f"null: launch shapes {shapes} differ from the profiled {PMC_SHAPES}",
                     "launch_ms": step_ms, "launches_timed": min(n_row, n_col),
                     "algorithmic_bytes_per_launch": slice_bytes, "probes_per_launch": probes_per_launch,
                     "probe_slices_per_s_from_launches": probes_per_launch / (step_ms * 1e-3),
                     "copy_kernel_GBps": copy_bw / 1e9, "frac_of_copy": achieved / copy_bw,
                     "in_place_kernel_GBps": inplace_bw / 1e9, "frac_of_in_place": achieved / inplace_bw,
                     "passes": {
                         "row": {"launch_ms": row_ms, "launches_timed": n_row, "algorithmic_bytes": row_bytes,
                                 "frac": row_bytes / (row_ms * 1e-3) / HBM_PEAK,
                                 "traffic": pmc["row pass"] if pmc else None},
                         "column": {"launch_ms": col_ms, "launches_timed": n_col, "algorithmic_bytes": col_bytes,
                                    "frac": col_bytes / (col_ms * 1e-3) / HBM_PEAK,
                                    "traffic": pmc["column pass"] if pmc else None}}},
        # the whole timed job (probe builds, detector passes, batch boundaries
        # and host work included) against the same bytes per probe-slice
        "slice_roofline_frac": ps * (16 * N + 12 * N / G) / HBM_PEAK,
        # arithmetic, reported only (HBM bounds the path): two 2D FFT pairs'
        # worth of radix-2 flops per probe-slice, 2 * 5 N log2 N
        "flops": {"per_probe_slice": 10.0 * N * math.log2(N),
                  "achieved_TFLOPs": ps * 10.0 * N * math.log2(N) / 1e12,
                  "fp32_peak_TFLOPs": 157.3},  # published dense FP32 rate
        "band_skip": {"threshold": 0.0 if args.no_band_skip else 1e-9, "skipped_columns": band[1] - band[0] + 1 if band[1] >= band[0] else 0,
                      "of": N_GRID},
        "batched_vs_plain": {
            "ratio": ps / plain_ps,
            "1024": {"ratio": ps / plain_ps, "batched_probe_slices_per_s": ps, "plain_probe_slices_per_s": plain_ps,
                     "plain_probes": args.plain_probes},
            "512": {"ratio": batched512 / plain512, "batched_probe_slices_per_s": batched512,
                    "plain_probe_slices_per_s": plain512, "plain_probes": args.plain_probes,
                    "batched": "32x32 raster (1024 probes) x 80 slices, 2 phonon steps, default launch shapes"},
            "plain_path": "one probe per launch, atoms re-cut and potentials rebuilt per probe (general scanner), "
                          "seeded raster sample"},
        "ragged_mix": {"probe_slices_per_s": rag_ps, "probe_slices": rag_work,
                       "first_call_s": rag_first,
                       "workload": "8 seeded requests on 512^2/768^2/1024^2/2048^2 grids, Si 50-400 A, "
                                   "200/300 kV, 10-30 mrad, detector sums every 10 slices, one graph launch; "
                                   "median of 5 timed calls re-launching the graph the first call captured "
                                   "(first_call_s: that call, capture and instantiation included)"},
        "adjoint_vs_forward": {"ratio": t_fwd_adj / t_fwd, "forward_s": t_fwd, "forward_plus_adjoint_s": t_fwd_adj,
                               "workload": "512^2, 40 slices, 16x16 probes, 300 kV / 20 mrad, "
                                           "checkpoint every slice, default launch shapes"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
