"""Pins the FP64 oracle (CPU, no GPU): against golden vectors computed
independently (tests/golden/golden.npz: scipy constants, scipy Bessel
functions, numpy FFTs, quadrature of the 3D potential), brute-force DFTs, and
closed forms / invariants of the multislice (norm conservation, vacuum
propagation, detector pixel counts, weak-phase linearity, finite differences
of the adjoint)."""
import os

import numpy as np
import pytest

import pyoracle as O
from helpers import rel_l2

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


# ----------------------------------------------------------------- FFT
def brute_dft(x, sign):
    n = len(x)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [2, 3, 5, 8, 12, 16, 18, 25, 30, 7, 11])
def test_fft1d_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        np.testing.assert_allclose(O.fft1d(x, sign), brute_dft(x, sign), rtol=0, atol=1e-12 * n)


def test_fft_golden_vectors():
    for n in (24, 768):
        assert rel_l2(O.fft1d(GOLD[f"fft1_in_{n}"]), GOLD[f"fft1_out_{n}"]) < 1e-14
    assert rel_l2(O.fft2d(GOLD["fft2_in"]), GOLD["fft2_out"]) < 1e-14


@pytest.mark.parametrize("shape", [(8, 8), (16, 12), (32, 32), (24, 48)])
def test_fft2d_brute_force_parseval_shift(shape):
    rng = np.random.default_rng(1)
    ny, nx = shape
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ky, kx = np.arange(ny), np.arange(nx)
    Fy = np.exp(-2j * np.pi * np.outer(ky, ky) / ny)
    Fx = np.exp(-2j * np.pi * np.outer(kx, kx) / nx)
    X = O.fft2d(x)
    np.testing.assert_allclose(X, Fy @ x @ Fx.T, atol=1e-11 * nx * ny)
    # Parseval for the unnormalised transform
    assert abs(np.sum(abs(X) ** 2) - nx * ny * np.sum(abs(x) ** 2)) < 1e-9 * np.sum(abs(X) ** 2)
    # shift theorem: x[n - s] -> X[k] exp(-2 pi i k s / n)
    xs = np.roll(x, (3, 5), axis=(0, 1))
    ph = np.exp(-2j * np.pi * (3 * ky[:, None] / ny + 5 * kx[None, :] / nx))
    np.testing.assert_allclose(O.fft2d(xs), X * ph, atol=1e-10 * nx * ny)
    # inverse undoes forward up to N
    np.testing.assert_allclose(O.fft2d(X, +1), x * nx * ny, atol=1e-10 * nx * ny)


# ----------------------------------------------------------------- constants / potential
def test_constants_match_scipy_golden():
    for kv, lam, gam, sig in zip(GOLD["kv"], GOLD["lambda"], GOLD["gamma"], GOLD["sigma"]):
        l, g, s = O.constants(kv)
        assert abs(l / lam - 1) < 1e-9 and abs(g / gam - 1) < 1e-9 and abs(s / sig - 1) < 1e-9
    # the survey's table (rounded published values)
    l, g, s = O.constants(200.0)
    assert abs(l - 0.025079) < 1e-6 and abs(g - 1.391390) < 1e-6 and abs(s - 7.2884e-4) < 1e-8
    l, g, s = O.constants(300.0)
    assert abs(l - 0.019687) < 1e-6 and abs(g - 1.587085) < 1e-6 and abs(s - 6.5262e-4) < 1e-8


def test_bessel_matches_scipy_golden():
    for x, k0, k1 in zip(GOLD["bessel_x"], GOLD["k0"], GOLD["k1"]):
        a, b = O.bessel_k01(x)
        assert abs(a / k0 - 1) < 1e-12 and abs(b / k1 - 1) < 1e-12


@pytest.mark.parametrize("Z", [8, 22, 38])
def test_projected_potential_matches_quadrature(Z):
    a, b = O.moliere_coefs(Z)
    V, dV = O.vr(a, b, [], [], GOLD["vr_R"])
    # scipy carries CODATA 2022 constants, the oracle CODATA 2018: ~1e-9 apart
    np.testing.assert_allclose(V, GOLD[f"vr_{Z}"], rtol=1e-8)
    # Kirkland large-q limit: sum a_i = Z / (2 pi^2 a0)
    assert abs(a.sum() - Z / (2 * np.pi ** 2 * 0.529177210903)) < 1e-12 * Z
    # dV/dR against a central difference
    R = GOLD["vr_R"]
    h = 1e-6
    Vp, _ = O.vr(a, b, [], [], R + h)
    Vm, _ = O.vr(a, b, [], [], R - h)
    np.testing.assert_allclose(dV, (Vp - Vm) / (2 * h), rtol=1e-6)


def test_spline_interpolates_and_tapers():
    Z = 38
    R2, c0, c1, c2, c3 = O.build_spline(Z)
    a, b = O.moliere_coefs(Z)
    # nodes reproduce c0 exactly, and c0 is VR before the taper radius
    assert np.allclose(O.spline_eval(R2, (R2, c0, c1, c2, c3)), c0, rtol=1e-14, atol=1e-14)
    untapered = np.sqrt(R2) < O.SPLINE_TAP * O.SPLINE_RMAX
    Vn, _ = O.vr(a, b, [], [], np.sqrt(R2[untapered]))
    np.testing.assert_allclose(c0[untapered], Vn, rtol=1e-13)
    # midpoints against the analytic potential
    mid = 0.5 * (R2[:-1] + R2[1:])
    ok = np.sqrt(R2[1:]) < O.SPLINE_TAP * O.SPLINE_RMAX  # intervals wholly before the taper
    Vm, _ = O.vr(a, b, [], [], np.sqrt(mid[ok]))
    assert np.max(np.abs(O.spline_eval(mid[ok], (R2, c0, c1, c2, c3)) / Vm - 1)) < 1e-4
    # the taper brings the potential to zero at R_max
    assert abs(c0[-1]) < 1e-12 * abs(c0[0])
    # C1 continuity of the piecewise cubic at interior nodes
    d = R2[1:] - R2[:-1]
    left = ((c3[:-1] * d + c2[:-1]) * d + c1[:-1]) * d + c0[:-1]
    dleft = (3 * c3[:-1] * d + 2 * c2[:-1]) * d + c1[:-1]
    np.testing.assert_allclose(left, c0[1:], rtol=1e-9, atol=1e-9 * abs(c0).max())
    np.testing.assert_allclose(dleft, c1[1:], rtol=1e-8, atol=1e-8 * abs(c1).max())


# ----------------------------------------------------------------- slicing
def _sorted_z(at):
    return np.sort(at["z"], kind="stable")


def test_slice_counts_of_the_workloads():
    import cudaem_amd.specimen as sp
    cu = sp.fcc_cu(8, 8, 10)
    assert len(cu["x"]) == 2560
    b, i0, i1 = O.slicing(_sorted_z(cu), 1.8075, O.BORDER_MAX)
    assert len(i0) == 20
    st = sp.srtio3(10, 10, 40)
    assert len(st["x"]) == 20000
    b, i0, i1 = O.slicing(_sorted_z(st), 1.9525, O.BORDER_MIN)
    assert len(i0) == 80
    for nz in (10, 74):  # Si 54 A and 402 A thick
        si = sp.si_diamond(2, 2, nz)
        b, i0, i1 = O.slicing(_sorted_z(si), 5.431 / 4, O.BORDER_MIN)
        assert len(i0) == 4 * nz


@pytest.mark.parametrize("border", [O.BORDER_MIN, O.BORDER_MAX, O.BORDER_MIDDLE])
def test_slicing_partitions_atoms(border):
    rng = np.random.default_rng(3)
    z = np.sort(rng.uniform(-3, 17, 500))
    dz = 1.3
    b, i0, i1 = O.slicing(z, dz, border)
    assert np.allclose(np.diff(b), dz)
    owner = np.full(len(z), -1)
    for s in range(len(i0)):
        if i0[s] <= i1[s]:
            idx = np.arange(i0[s], i1[s] + 1)
            assert np.all(owner[idx] == -1)
            owner[idx] = s
            assert np.all((z[idx] >= b[s]) & (z[idx] < b[s + 1]))
    # brute force: atom j belongs to the slice whose half-open range holds it
    for j in range(len(z)):
        s = np.searchsorted(b, z[j], side="right") - 1
        if 0 <= s < len(i0):
            assert owner[j] == s
    if border != O.BORDER_MIDDLE:
        assert np.all(owner >= 0)


# ----------------------------------------------------------------- operators
def test_propagator_closed_form_and_composition():
    n, L, lam = 64, 12.0, O.constants(300.0)[0]
    g = np.fft.fftfreq(n, d=L / n)
    g2 = g[None, :] ** 2 + g[:, None] ** 2
    for dz in (1.0, 2.5):
        P = O.propagator(n, n, L, L, dz, False, lam)
        np.testing.assert_allclose(P, np.exp(-1j * np.pi * lam * dz * g2) / n ** 2, atol=1e-15)
    P1 = O.propagator(n, n, L, L, 1.0, False, lam)
    P2 = O.propagator(n, n, L, L, 2.5, False, lam)
    P3 = O.propagator(n, n, L, L, 3.5, False, lam)
    np.testing.assert_allclose(P1 * P2 * n * n, P3, atol=1e-14)
    # bwl: |P| N is the Fermi factor with 1% at 2/3 g_max + 0.25 (grid_dev.cuh:53-55)
    Pb = O.propagator(n, n, L, L, 1.0, True, lam)
    gl2, alpha = O.bwl_params(n, n, L, L)
    gl = (2.0 / 3.0) * (n // 2) / L
    assert abs(gl2 - gl * gl) < 1e-12
    np.testing.assert_allclose(abs(Pb) * n * n, 1 / (1 + np.exp(alpha * (g2 - gl2))), atol=1e-14)
    assert abs(1 / (1 + np.exp(alpha * ((gl + 0.25) ** 2 - gl2))) - 0.01) < 1e-12


def test_probe_is_inverse_fft_of_aperture_phase():
    n, L = 64, 10.0
    lam = O.constants(300.0)[0]
    lens = {"c_10": 50.0, "c_30": 2e4, "outer_aper_mrad": 30.0, "inner_aper_mrad": 5.0}
    xb, yb = 3.0, 7.0
    psi = O.probe(n, n, L, L, lam, lens, xb, yb, defocus_shift=0.0)
    g = np.fft.fftfreq(n, d=L / n)
    GX, GY = np.meshgrid(g, g)
    g2 = GX ** 2 + GY ** 2
    gmin, gmax = np.sin(5e-3) / lam, np.sin(30e-3) / lam
    ap = (g2 >= gmin ** 2) & (g2 < gmax ** 2)
    chi = -np.pi * 50.0 * lam * g2 - np.pi * 2e4 * lam ** 3 / 2 * g2 ** 2
    ph = chi - 2 * np.pi * (xb - L / 2) * GX - 2 * np.pi * (yb - L / 2) * GY
    ref = np.fft.ifft2(np.where(ap, np.exp(1j * ph), 0)) * n * n
    assert rel_l2(psi, ref) < 1e-12
    # the probe sits at (xb + L/2, yb + L/2) mod L
    iy, ix = np.unravel_index(np.argmax(abs(O.probe(n, n, L, L, lam, {"outer_aper_mrad": 30.0}, xb, yb))), psi.shape)
    assert abs(ix * L / n - (xb + L / 2) % L) < L / n and abs(iy * L / n - (yb + L / 2) % L) < L / n


def test_chi_azimuthal_terms():
    lam = O.constants(200.0)[0]
    lens = {"c_12": 10.0, "phi_12": 0.4, "c_23": 500.0, "phi_23": 0.1, "m": 1.0}
    gx, gy = np.array([0.3, -0.7, 0.0]), np.array([0.2, 0.5, -0.9])
    phi = np.arctan2(gy, gx)
    g = np.hypot(gx, gy)
    ref = (-np.pi * 10.0 * lam * g ** 2 * np.sin(2 * (phi - 0.4))
           - 2 * np.pi * 500.0 * lam ** 2 / 3 * g ** 3 * np.sin(3 * (phi - 0.1)) + phi)
    np.testing.assert_allclose(O.chi(lens, lam, gx, gy), ref, rtol=1e-13)


def test_potential_brute_force_small_grid():
    R2, c0, c1, c2, c3 = O.build_spline(14)
    tab = O.Tables([14])
    n, L = 32, 6.0
    x = np.array([1.0, 3.3, 5.9])
    y = np.array([2.0, 0.1, 4.4])
    occ = np.array([1.0, 0.5, 1.0])
    V = O.potential(n, n, L, L, 0, 2, x, y, occ, np.zeros(3, np.int32), tab, 0.7)
    ref = np.zeros((n, n))
    for a in range(3):
        for iy in range(n):
            for ix in range(n):
                r2 = (ix * L / n - x[a]) ** 2 + (iy * L / n - y[a]) ** 2  # non-periodic
                if r2 < R2[-1]:
                    ref[iy, ix] += 0.7 * occ[a] * O.spline_eval(np.array([r2]), (R2, c0, c1, c2, c3))[0]
    np.testing.assert_allclose(V, ref, rtol=1e-13, atol=1e-12)


# ----------------------------------------------------------------- multislice invariants
def _smooth_phase(rng, S, n, scale):
    f = rng.standard_normal((S, n, n))
    k = np.fft.fftfreq(n)
    lp = np.exp(-(k[None, :] ** 2 + k[:, None] ** 2) / 0.01)
    return scale * np.real(np.fft.ifft2(np.fft.fft2(f) * lp))


def test_multislice_norm_conservation_and_bwl_loss():
    rng = np.random.default_rng(11)
    n, L, S = 64, 12.0, 6
    lam = O.constants(200.0)[0]
    V = _smooth_phase(rng, S, n, 3.0)
    psi = O.probe(n, n, L, L, lam, {"outer_aper_mrad": 25.0}, 2.0, 3.0)
    P = O.propagator(n, n, L, L, 1.5, False, lam)
    Psi = O.multislice(P, psi, V=V)
    n_in = np.sum(abs(psi) ** 2)
    assert abs(np.sum(abs(Psi) ** 2) / (n * n * n_in) - 1) < 1e-12
    Pb = O.propagator(n, n, L, L, 1.5, True, lam)
    Psib = O.multislice(Pb, psi, V=V)
    assert np.sum(abs(Psib) ** 2) < n * n * n_in


def test_vacuum_probe_detector_counts():
    """Vacuum, bwl off: |Psi(g)|^2 = N^2 on the aperture, so a detector sums
    N^2 times the number of aperture pixels in its (closed) band."""
    n, L, S = 64, 20.0, 4
    lam = O.constants(300.0)[0]
    lens = {"outer_aper_mrad": 30.0}
    psi = O.probe(n, n, L, L, lam, lens, 5.0, 5.0)
    P = O.propagator(n, n, L, L, 2.0, False, lam)
    Psi = O.multislice(P, psi, V=np.zeros((S, n, n)))
    g = np.fft.fftfreq(n, d=L / n)
    g2 = g[None, :] ** 2 + g[:, None] ** 2
    ga = (np.sin(30e-3) / lam) ** 2
    for lo, hi in ((0.0, 15.0), (10.0, 25.0), (20.0, 40.0)):
        g2lo, g2hi = (np.sin(lo * 1e-3) / lam) ** 2, (np.sin(hi * 1e-3) / lam) ** 2
        count = np.sum((g2 < ga) & (g2 >= g2lo) & (g2 <= g2hi))
        s = O.detector_sum(Psi, L, L, g2lo, g2hi)
        assert abs(s - count * n ** 4) < 1e-8 * max(count * n ** 4, 1)


def test_weak_phase_object_is_linear():
    rng = np.random.default_rng(4)
    n = 32
    V = _smooth_phase(rng, 1, n, 1.0)
    eps = 1e-5
    P = np.ones((n, n)) / n ** 2
    Psi = O.multislice(P, np.ones((n, n), complex), V=eps * V)
    lin = np.fft.fft2(1 + 1j * eps * V[0])
    assert rel_l2(Psi, lin) < 1e-9


def test_thickness_frames_full_band_equal_wave_intensity():
    rng = np.random.default_rng(8)
    n, L, S = 32, 8.0, 5
    lam = O.constants(200.0)[0]
    V = _smooth_phase(rng, S, n, 2.0)
    P = O.propagator(n, n, L, L, 1.0, False, lam)
    psi = O.probe(n, n, L, L, lam, {"outer_aper_mrad": 40.0}, 1.0, 2.0)
    frames, w = O.thickness(P, V, psi, L, L, 0.0, 1e30, stride=1)
    ref = psi.copy()
    for k in range(S):
        ref = np.fft.ifft2(np.fft.fft2(ref * np.exp(1j * V[k])) * P) * n * n
        np.testing.assert_allclose(frames[k], abs(ref) ** 2, rtol=1e-9, atol=1e-9 * abs(ref).max() ** 2)
    f2, _ = O.thickness(P, V, psi, L, L, 0.0, 1e30, stride=2)
    assert f2.shape[0] == 3  # slices 2, 4 and the last


def test_ctf_image_without_aberrations_is_the_wave():
    rng = np.random.default_rng(9)
    n, L = 32, 8.0
    psi = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Psi = np.fft.fft2(psi)
    I, w = O.ctf_image(Psi, L, L, O.constants(200.0)[0], {})
    np.testing.assert_allclose(w, psi * n * n, atol=1e-9 * n * n)
    np.testing.assert_allclose(I, abs(psi * n * n) ** 2, rtol=1e-10)


# ----------------------------------------------------------------- adjoint
def _loss(P, t, psi0, target):
    L = 0.0
    for b in range(psi0.shape[0]):
        Psi = O.multislice(P, psi0[b], t=t)
        L += np.sum((abs(Psi) ** 2 - target[b]) ** 2)
    return L


def test_adjoint_finite_differences():
    rng = np.random.default_rng(21)
    n, S, B, L = 16, 3, 2, 5.0
    lam = O.constants(300.0)[0]
    P = O.propagator(n, n, L, L, 1.0, True, lam)
    t = rng.uniform(0.95, 1, (S, n, n)) * np.exp(1j * rng.uniform(0, 0.5, (S, n, n)))
    psi0 = np.stack([O.probe(n, n, L, L, lam, {"outer_aper_mrad": 60.0}, x, x) for x in (1.0, 3.0)]) / n
    target = rng.uniform(0.5, 1.5, (B, n, n)) * np.stack([abs(O.multislice(P, p, t=t)) ** 2 for p in psi0])
    Psi = np.stack([O.multislice(P, p, t=t) for p in psi0])
    gPsi = 4 * (abs(Psi) ** 2 - target) * Psi
    gt, Psi2 = O.adjoint(P, t, psi0, gPsi)
    assert rel_l2(Psi2, Psi) < 1e-14
    scale = np.abs(t).max()
    for _ in range(6):
        d = rng.standard_normal((S, n, n)) + 1j * rng.standard_normal((S, n, n))
        h = 1e-6 * scale
        fd = (_loss(P, t + h * d, psi0, target) - _loss(P, t - h * d, psi0, target)) / (2 * h)
        an = np.sum(np.real(np.conj(gt) * d))
        assert abs(fd - an) < 1e-6 * max(abs(an), 1e-30), (fd, an)


def test_adjoint_dot_product():
    """<G x, y> = <x, G^H y> for the linear map x -> exit wave at fixed t."""
    rng = np.random.default_rng(22)
    n, S = 16, 4
    P = O.propagator(n, n, 4.0, 4.0, 1.0, True, O.constants(200.0)[0])
    t = np.exp(1j * rng.uniform(0, 1, (S, n, n)))
    x = rng.standard_normal((1, n, n)) + 1j * rng.standard_normal((1, n, n))
    y = rng.standard_normal((1, n, n)) + 1j * rng.standard_normal((1, n, n))
    Gx = O.multislice(P, x[0], t=t)
    # the adjoint's gradient w.r.t. the incident wave is G^H y: recover it via
    # the t-gradient of slice 0, gt_0 = conj(psi_0) * lambda_0 with psi_0 = x
    gt, _ = O.adjoint(P, t, x, y)
    GHy_t0 = gt[0] / np.conj(x[0])  # = lambda_0 = conj(t_0)^-1 * (G^H y) ... lambda_0 before conj(t_0)
    GHy = np.conj(t[0]) * GHy_t0
    lhs = np.vdot(y[0], Gx)
    rhs = np.vdot(GHy, x[0])
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
