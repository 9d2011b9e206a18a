"""The five benchmark workloads, each on its own configuration, against the FP64
oracle on the same seeded inputs (relative L2: 1e-4 on exit waves, 1e-3 on
intensities, detector sums and gradients).

  plain STEM     512^2, SrTiO3 10x10x40, 80 slices, 300 kV / 25 mrad, ADF
                 60-200 mrad, one probe per launch: the C++ general scanner
                 (run_STEM_general_simu, scan_simulator.cuh:114-183) and the
                 engine's per-probe path, 4 positions of the 32x32 raster
  batched STEM   1024^2, the same specimen with u_rms 0.08 A frozen phonons,
                 512 positions of the 64x64 raster in batches of 256 (one
                 em_stem call per phonon configuration, the second reusing
                 the first's graphs), 24 probes checked
  ragged         512^2 / 768^2 / 1024^2 / 2048^2 requests in one launch, Si
                 50-400 A, 200 / 300 kV, 10-30 mrad, detector sums every 10
                 slices, every probe checked
  inversion      512^2, 40 slices of smooth random transmission, 16x16 probes
                 at 300 kV / 20 mrad, far-field loss: torch.autograd gradient
                 w.r.t. the transmission against the oracle adjoint
The oracle's work is spread over host threads (its entry points release the
GIL)."""
import concurrent.futures as cf
import os
import subprocess
import sys

import numpy as np
import pytest

import pyoracle as O
import cudaem_amd.specimen as sp
from helpers import OracleSpecimen, probe_lens, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_simulators")
TOL_WAVE, TOL_INT = 1e-4, 1e-3
THREADS = max(1, min(16, os.cpu_count() or 1))


def _np(t):
    return t.detach().cpu().numpy()


def _pmap(f, items):
    with cf.ThreadPoolExecutor(max_workers=THREADS) as ex:
        return list(ex.map(f, items))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def em(cuda):
    import cudaem_amd
    return cudaem_amd


# ------------------------------------------------------------------ plain STEM
A_STO, DZ_STO = 3.905, 3.905 / 2
RASTER2 = [(5, 3), (16, 3), (5, 16), (16, 16)]  # (ix, iy) of the 32 x 32 raster


def _srtio3_f32(cells, nz):
    """SrTiO3 with the float32 coordinates the C++ program builds."""
    basis = [(38, 0, 0, 0), (22, .5, .5, .5), (8, .5, .5, 0), (8, .5, 0, .5), (8, 0, .5, .5)]
    xs, ys, zs, Zs = [], [], [], []
    for (Z, fx, fy, fz) in basis:
        for i in range(cells):
            for j in range(cells):
                for k in range(nz):
                    xs.append(np.float32((i + fx) * A_STO))
                    ys.append(np.float32((j + fy) * A_STO))
                    zs.append(np.float32((k + fz) * A_STO))
                    Zs.append(Z)
    f = lambda v: np.array(v, dtype=np.float32).astype(np.float64)
    x = f(xs)
    return {"x": x, "y": f(ys), "z": f(zs), "Z": np.array(Zs, np.int32), "occ": np.ones_like(x)}


def _config2_positions():
    """Positions the C++ program's 2 x 2 scan over [5/32, 3/32, 27/32, 29/32]
    of the field produces (float range times float field, then double)."""
    lx = np.float32(10) * np.float32(A_STO)
    f = np.float32
    x0, y0 = float(f(5 / 32) * lx), float(f(3 / 32) * lx)
    wx, wy = float(f(f(27 / 32) - f(5 / 32)) * lx), float(f(f(29 / 32) - f(3 / 32)) * lx)
    xs = [x0 + ix * wx / 2 for iy in range(2) for ix in range(2)]
    ys = [y0 + iy * wy / 2 for iy in range(2) for ix in range(2)]
    return float(lx), np.array(xs), np.array(ys)


def _config2_reference(at, L, xs, ys, n=512):
    """Per position: the atoms re-cut around the probe (2 A margin), sliced,
    the probe built at the window centre with the defocus at slice 0's middle,
    80 slices, detector 60-200 mrad."""
    lam = O.constants(300.0)[0]
    g2min, g2max = (np.sin(60e-3) / lam) ** 2, (np.sin(200e-3) / lam) ** 2
    lens = probe_lens(25.0)

    def one(i):
        c = sp.cut_window(at, xs[i] - L / 2, xs[i] + L / 2, ys[i] - L / 2, ys[i] + L / 2, margin=2.0)
        c = {k: (np.asarray(v, dtype=np.float32).astype(np.float64) if k in ("x", "y") else v) for k, v in c.items()}
        o = OracleSpecimen(c, n, n, L, L, DZ_STO, 300.0, O.BORDER_MIN)
        assert o.S == 80
        psi = O.probe(n, n, L, L, o.lam, lens, 0.0, 0.0, o.zmid0())
        Psi = O.multislice(o.propagator(), psi, V=o.potentials())
        return O.detector_sum(Psi, L, L, g2min, g2max)

    return np.array(_pmap(one, range(len(xs))))


@pytest.fixture(scope="module")
def config2_reference():
    at = _srtio3_f32(10, 40)
    L, xs, ys = _config2_positions()
    _say("config 2: oracle, 4 positions x 80 slices at 512^2")
    return at, L, xs, ys, _config2_reference(at, L, xs, ys)


@pytest.mark.timeout(900)
def test_config2_cpp_general_scanner(cuda, tmp_path, config2_reference):
    """scan_simulator's general scanning through the C++ surface."""
    at, L, xs, ys, ref = config2_reference
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_simulators is not built (run __graft_entry__.build())")
    _say("config 2: C++ general scanner")
    r = subprocess.run([EXE, "config2", str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "config2_stem.f32", dtype=np.float32).astype(np.float64)
    e = rel_l2(got, ref)
    _say("config 2 C++ general scanner", got, "oracle", ref, "relL2", e)
    assert e < TOL_INT


@pytest.mark.timeout(900)
def test_config2_engine_one_probe_per_launch(em, config2_reference):
    """The engine's per-probe path as the bench's plain side runs it: per
    position re-cut, slice, build the potentials, build one probe, run the
    slices, sum the detector."""
    at, L, xs, ys, ref = config2_reference
    got = []
    with em.Plan(512, 512, L, L, DZ_STO, 300.0, max_batch=1) as p:
        for i in range(len(xs)):
            c = sp.cut_window(at, xs[i] - L / 2, xs[i] + L / 2, ys[i] - L / 2, ys[i] + L / 2, margin=2.0)
            c = {k: (np.asarray(v, dtype=np.float32).astype(np.float64) if k in ("x", "y") else v)
                 for k, v in c.items()}
            S = p.set_specimen(c["x"], c["y"], c["z"], c["Z"], c["occ"], border=em.BORDER_MIN)
            assert S == 80
            zs, ze, _, _ = p.slice_info(0)
            V = p.potential_slices()
            psi = p.build_probes(probe_lens(25.0), [0.0], [0.0], defocus_shift=0.5 * (zs + ze))
            p.multislice_forward(psi, V=V)
            got.append(float(_np(p.detector_sums(psi, 60.0, 200.0))[0, 0]))
    e = rel_l2(np.array(got), ref)
    _say("config 2 engine per-probe", got, "oracle", ref, "relL2", e)
    assert e < TOL_INT


# ---------------------------------------------------------------- batched STEM
@pytest.mark.timeout(900)
def test_config3_batched_stem_frozen_phonons(em):
    """1024^2 em_stem over 512 raster positions in batches of 256, two
    frozen-phonon configurations (u_rms 0.08 A, seeded); the second call
    reuses the graphs the first captured.  16 probes of configuration 1 and 8
    of configuration 2 against the oracle."""
    import torch
    n = 1024
    at0 = sp.srtio3(10, 10, 40)
    L = at0["lx"]
    rx, ry = sp.raster(64, 64, L, L)
    rng = np.random.default_rng(33)
    pick = rng.choice(64 * 64, 512, replace=False)
    xs, ys = rx[pick], ry[pick]
    lens = probe_lens(25.0)
    got, specs = [], []
    with em.Plan(n, n, L, L, DZ_STO, 300.0, max_batch=256) as p:
        for iconf in (1, 2):
            at = sp.vibrate(at0, 0.08, 1234, iconf)
            S = p.set_specimen(at["x"], at["y"], at["z"], at["Z"], at["occ"], border=em.BORDER_MIN)
            assert S == 80
            got.append(_np(p.stem(lens, xs, ys, [60.0], [200.0]))[0].astype(np.float64))
            specs.append(at)
        torch.cuda.synchronize()
    for iconf, (at, k) in enumerate(zip(specs, (16, 8)), start=1):
        sub = np.random.default_rng(40 + iconf).choice(512, k, replace=False)
        o = OracleSpecimen(at, n, n, L, L, DZ_STO, 300.0, O.BORDER_MIN)
        _say(f"config 3: oracle potentials, configuration {iconf}")
        V = np.stack(_pmap(o.potential, range(o.S)))
        g2 = [np.array([(np.sin(a * 1e-3) / o.lam) ** 2]) for a in (60.0, 200.0)]
        _say(f"config 3: oracle STEM, {k} probes x 80 slices")
        ref, _ = O.stem(n, n, L, L, o.lam, lens, o.zmid0(), o.propagator(), V, xs[sub], ys[sub], g2[0], g2[1],
                        nthreads=THREADS)
        e = rel_l2(got[iconf - 1][sub], ref[0])
        _say(f"config 3 configuration {iconf} relL2", e)
        assert e < TOL_INT
    # the two configurations differ (the cached graph read the new specimen)
    assert rel_l2(got[1], got[0]) > 1e-3


# ---------------------------------------------------------------------- ragged
# (grid, kV, convergence mrad, probes, Si cells thick): thicknesses 54-402 A
RAGGED = [(512, 300.0, 30.0, 2, 74), (768, 200.0, 10.0, 1, 37), (1024, 300.0, 20.0, 2, 20),
          (2048, 200.0, 25.0, 1, 10)]
RAGGED_DET = ([10.0, 40.0], [40.0, 150.0])


def _oracle_frames(P, V, psi0, lx, ly, g2min, g2max, stride):
    """Detector sums of the exit wave after every stride-th slice and the
    last one: the slice sequence of wave_function.cuh:60-87 stepped one slice
    at a time (each step's exit wave FFT[t_k psi] is the wave after slice k
    with the final-slice rule; psi_{k+1} = IFFT[P FFT[t_k psi]])."""
    N = psi0.size
    psi = psi0.astype(np.complex128)
    S = V.shape[0]
    out = []
    for k in range(S):
        Psi = O.multislice(P, psi, V=V[k:k + 1])
        if (k + 1) % stride == 0 or k == S - 1:
            out.append([O.detector_sum(Psi, lx, ly, a, b) for a, b in zip(g2min, g2max)])
        if k < S - 1:
            psi = np.fft.ifft2(P * Psi) * N
    return np.array(out)


@pytest.mark.timeout(900)
def test_config4_ragged_mix_against_oracle(em):
    import torch
    rng = np.random.default_rng(17)
    plans, Vs, psis, jobs = [], [], [], []
    for n, kv, alpha, B, nz in RAGGED:
        at = sp.si_diamond(5, 5, nz)
        L = at["lx"]
        p = em.Plan(n, n, L, L, 5.431 / 4, kv, max_batch=B)
        S = p.set_specimen(at["x"], at["y"], at["z"], at["Z"], at["occ"], border=em.BORDER_MIN)
        assert S == 4 * nz
        zs, ze, _, _ = p.slice_info(0)
        xs, ys = rng.uniform(0, L, B), rng.uniform(0, L, B)
        lens = probe_lens(alpha)
        Vs.append(p.potential_slices())
        psis.append(p.build_probes(lens, xs, ys, defocus_shift=0.5 * (zs + ze)))
        plans.append(p)
        jobs.append((n, kv, lens, at, xs, ys))
    got = [_np(g) for g in em.ragged(plans, Vs, psis, *RAGGED_DET, stride=10)]
    torch.cuda.synchronize()
    for p in plans:
        p.close()
    del Vs, psis

    # oracle: specimens first (potentials over threads), then one task per probe
    tasks = []
    for r, (n, kv, lens, at, xs, ys) in enumerate(jobs):
        L = at["lx"]
        o = OracleSpecimen(at, n, n, L, L, 5.431 / 4, kv, O.BORDER_MIN)
        _say(f"config 4: oracle potentials, request {r} ({n}^2, {o.S} slices)")
        V = np.stack(_pmap(o.potential, range(o.S)))
        P = o.propagator()
        g2 = [[(np.sin(a * 1e-3) / o.lam) ** 2 for a in d] for d in RAGGED_DET]
        for b in range(len(xs)):
            tasks.append((r, b, P, V, O.probe(n, n, L, L, o.lam, lens, xs[b], ys[b], o.zmid0()), L, g2))
    _say(f"config 4: oracle slice loops, {len(tasks)} probes")
    res = _pmap(lambda t: _oracle_frames(t[2], t[3], t[4], t[5], t[5], t[6][0], t[6][1], 10), tasks)
    for (r, b, *_), frames in zip(tasks, res):
        g = got[r][:, :, b].astype(np.float64)
        assert g.shape == frames.shape, (r, g.shape, frames.shape)
        e = rel_l2(g, frames)
        _say(f"config 4 request {r} ({RAGGED[r][0]}^2) probe {b}: {g.shape[0]} frames, relL2 {e:.2e}")
        assert e < TOL_INT


# ------------------------------------------------------------------- inversion
@pytest.mark.timeout(900)
def test_config5_inverse_multislice_gradient(em):
    """The differentiable-inversion workload at full size through
    cudaem_amd.autograd: exit waves against the oracle forward, and the
    transmission gradient against the oracle adjoint of the same loss."""
    import torch
    from cudaem_amd.autograd import multislice
    n, S, scan, L = 512, 40, 16, 30.0
    g = np.random.default_rng(5)
    k = np.fft.fftfreq(n)
    lp = np.exp(-(k[None, :] ** 2 + k[:, None] ** 2) / 2e-3)

    def smooth(lo, hi):
        f = np.real(np.fft.ifft2(np.fft.fft2(g.standard_normal((S, n, n)), axes=(1, 2)) * lp, axes=(1, 2)))
        f = (f - f.min()) / (f.max() - f.min())
        return lo + (hi - lo) * f

    t = smooth(0.95, 1.0) * np.exp(1j * smooth(0.0, 0.5))
    t64 = t.astype(np.complex64)
    xs, ys = sp.raster(scan, scan, L, L)
    B = len(xs)
    with em.Plan(n, n, L, L, 2.0, 300.0, max_batch=B) as p:
        psi0 = p.build_probes(probe_lens(20.0), xs, ys)
        tt = torch.from_numpy(t64).cuda().requires_grad_(True)
        Psi = multislice(p, tt, psi0, every=1)
        I = Psi.real ** 2 + Psi.imag ** 2
        scale = 1.0 / float(I.detach().max())
        # seeded far-field targets at the intensities' mean level (a target
        # proportional to I would make I - target cancel where the factor is
        # near 1 and amplify FP32 differences in the residual)
        target = torch.from_numpy(g.uniform(0.5, 1.5, (B, n, n)).astype(np.float32)).cuda() * float(I.detach().mean())
        loss = (((I - target) * scale) ** 2).sum()
        loss.backward()
        grad = _np(tt.grad)
        Psi_g = _np(Psi)
        P = _np(p.propagator()).astype(np.complex128)
        p0 = _np(psi0).astype(np.complex128)
        tgt = _np(target).astype(np.float64)
        torch.cuda.synchronize()
    t_in = t64.astype(np.complex128)
    # dL/dPsi of the oracle's own forward; the adjoint over probe chunks, summed
    chunks = np.array_split(np.arange(B), THREADS * 2)

    def fwd(c):
        return np.stack([O.multislice(P, p0[b], t=t_in) for b in c])

    _say(f"config 5: oracle forward, {B} probes x {S} slices")
    Psi_o = np.concatenate(_pmap(fwd, chunks))
    e_wave = rel_l2(Psi_g, Psi_o)
    _say("config 5 exit-wave relL2", e_wave)
    assert e_wave < TOL_WAVE
    gPsi = 4 * scale ** 2 * (abs(Psi_o) ** 2 - tgt) * Psi_o
    _say("config 5: oracle adjoint")
    parts = _pmap(lambda c: O.adjoint(P, t_in, p0[c], gPsi[c])[0], chunks)
    gt_o = np.sum(parts, axis=0)
    e = rel_l2(grad, gt_o)
    _say("config 5 gradient relL2", e)
    assert e < TOL_INT
