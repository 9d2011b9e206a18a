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
    # the closed forms' values at 200 and 300 kV, rounded to 5-6 significant digits
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
    z32 = z.astype(np.float32).astype(np.float64)  # the float atoms the slicer sees
    dz = 1.3
    b, i0, i1 = O.slicing(z, dz, border)
    assert np.allclose(np.diff(b), dz, rtol=1e-5)
    assert np.array_equal(b.astype(np.float32).astype(np.float64), b)  # float bounds
    owner = np.full(len(z), -1)
    for s in range(len(i0)):
        if i0[s] <= i1[s]:
            idx = np.arange(i0[s], i1[s] + 1)
            assert np.all(owner[idx] == -1)
            owner[idx] = s
            assert np.all((z32[idx] >= b[s]) & (z32[idx] < b[s + 1]))
    # brute force: atom j belongs to the slice whose half-open range holds it
    for j in range(len(z)):
        s = np.searchsorted(b, z32[j], side="right") - 1
        if 0 <= s < len(i0):
            assert owner[j] == s
    if border != O.BORDER_MIDDLE:
        assert np.all(owner >= 0)


@pytest.mark.parametrize("border", [O.BORDER_MIN, O.BORDER_MAX, O.BORDER_MIDDLE])
def test_slicing_is_float_arithmetic(border):
    """Bounds and memberships equal a brute-force float32 restatement bit for
    bit on phonon-displaced specimens, and an atom planted between a bound's
    float and double values is placed where the float arithmetic puts it."""
    import cudaem_amd.specimen as sp
    from slicing_ref import brute_force_slicing, planted_atom
    st = sp.vibrate(sp.srtio3(3, 3, 40), 0.08, 11, 2 + border)
    z = np.sort(st["z"], kind="stable")
    for zz in (z, np.sort(np.append(z, planted_atom(z, 1.9525, border)))):
        b, i0, i1 = O.slicing(zz, 1.9525, border)
        bf, i0f, i1f = brute_force_slicing(zz, 1.9525, border, O.Z_EPS)
        assert np.array_equal(b, bf) and np.array_equal(i0, i0f) and np.array_equal(i1, i1f)
        bd, i0d, i1d = brute_force_slicing(zz, 1.9525, border, O.Z_EPS, real=np.float64)
    # with the planted atom, double arithmetic would slice differently
    assert not (np.array_equal(i0, i0d) and np.array_equal(i1, i1d))


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


def test_fp32_stem_variant_tracks_the_fp64_oracle():
    """The FP32 timing variant of the STEM loop agrees with the FP64 loop to
    FP32 accuracy, and differs from it (so it really runs in FP32)."""
    rng = np.random.default_rng(11)
    n, L, S = 64, 16.0, 6
    lam = O.constants(300.0)[0]
    V = _smooth_phase(rng, S, n, 0.8)
    P = O.propagator(n, n, L, L, 2.0, True, lam)
    lens = {"outer_aper_mrad": 25.0}
    xb, yb = [3.0, 8.5, 12.0], [4.0, 9.5, 2.5]
    g2min = np.array([(np.sin(20e-3) / lam) ** 2, 0.0])
    g2max = np.array([(np.sin(60e-3) / lam) ** 2, (np.sin(20e-3) / lam) ** 2])
    r64, _ = O.stem(n, n, L, L, lam, lens, 0.0, P, V, xb, yb, g2min, g2max, nthreads=2)
    r32, _ = O.stem(n, n, L, L, lam, lens, 0.0, P, V, xb, yb, g2min, g2max, nthreads=2, fp32=True)
    e = np.linalg.norm(r32 - r64) / np.linalg.norm(r64)
    assert 1e-12 < e < 1e-4


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


# ----------------------------------------------------------------- fine slicing
FINE_TYPES = [1, 2, 3, 4, 5, 6]


def _fine_set(pt):
    return GOLD[f"fine_cl_{pt}"], GOLD[f"fine_cnl_{pt}"]


@pytest.mark.parametrize("n", [7, 64, 128])
def test_gauss_legendre_matches_numpy(n):
    x, w = O.gauss_legendre(n)
    np.testing.assert_allclose(x, GOLD[f"gl_x_{n}"], rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, GOLD[f"gl_w_{n}"], rtol=1e-10, atol=1e-15)
    # exact for polynomials of degree <= 2n - 1 on [-1/2, 1/2]
    for k in range(0, 2 * n, max(1, n // 4)):
        exact = 0.0 if k % 2 else 2 * 0.5 ** (k + 1) / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-14


@pytest.mark.parametrize("pt", FINE_TYPES)
def test_vr3d_derivative_by_finite_differences(pt):
    """dVrir = (1/r) dV/dr of every parameterisation."""
    cl, cnl = _fine_set(pt)
    r = np.linspace(0.15, 2.8, 23)
    h = 1e-6
    V, d = O.vr3d(pt, r, cl, cnl)
    Vp, _ = O.vr3d(pt, r + h, cl, cnl)
    Vm, _ = O.vr3d(pt, r - h, cl, cnl)
    fd = (Vp - Vm) / (2 * h) / r
    np.testing.assert_allclose(d, fd, rtol=2e-7, atol=1e-9 * np.abs(V).max())


def _projection_closed_form(pt, R):
    """int V(sqrt(z^2 + R^2)) dz over all z: Gaussians -> c sqrt(pi/b) e^{-b R^2};
    Yukawa c e^{-b r}/r -> 2 c K0(b R); Weickenmeier c erfc(b r)/r -> c E1(b^2 R^2);
    Lobato c (2/(b r) + 1) e^{-b r} -> c (4/b K0(b R) + 2 R K1(b R))."""
    from scipy.special import exp1, k0, k1
    cl, cnl = _fine_set(pt)
    if pt in (1, 2, 3):
        return sum(c * np.sqrt(np.pi / b) * np.exp(-b * R * R) for c, b in zip(cl, cnl))
    if pt == 4:
        return (sum(2 * c * k0(b * R) for c, b in zip(cl[:3], cnl[:3]))
                + sum(c * np.sqrt(np.pi / b) * np.exp(-b * R * R) for c, b in zip(cl[3:], cnl[3:])))
    if pt == 5:
        return sum(c * exp1(b * b * R * R) for c, b in zip(cl, cnl))
    return sum(c * (4 / b * k0(b * R) + 2 * R * k1(b * R)) for c, b in zip(cl, cnl))


@pytest.mark.parametrize("pt", FINE_TYPES)
def test_fine_spline_thick_slab_is_the_projection(pt):
    """A 12 A slab around the atom integrates the whole potential: the node
    values equal the closed-form projections; c1 = d c0 / d(R^2)."""
    cl, cnl = _fine_set(pt)
    R2, c0, c1, c2, c3 = O.fine_spline(pt, cl, cnl, 0.0, 12.0, nq=1024, tap=0.999)
    R = np.sqrt(R2)
    m = (R > 0.25) & (R < 2.9)
    np.testing.assert_allclose(c0[m], _projection_closed_form(pt, R[m]), rtol=1e-9, atol=1e-10 * abs(c0).max())
    h = 1e-5
    fd = (_projection_closed_form(pt, np.sqrt(R2[m] + h)) - _projection_closed_form(pt, np.sqrt(R2[m] - h))) / (2 * h)
    np.testing.assert_allclose(c1[m], fd, rtol=1e-6, atol=1e-9 * np.abs(c0).max())


@pytest.mark.parametrize("pt", FINE_TYPES)
def test_fine_spline_one_slice_matches_quadrature_golden(pt):
    """One 1.9525 A slice centred 0.7 A above the atom (the per-slice
    quantity of fine slicing): node values against adaptive quadrature of the
    published forms at the node radii."""
    cl, cnl = _fine_set(pt)
    R2, c0 = O.fine_spline(pt, cl, cnl, 0.7, 1.9525, tap=0.999)[:2]
    R = GOLD["fine_R"]
    idx = np.searchsorted(R2, R * R * (1 - 1e-12))
    np.testing.assert_allclose(R2[idx], R * R, rtol=1e-12)
    np.testing.assert_allclose(c0[idx], GOLD[f"fine_slab_{pt}"], rtol=1e-10, atol=1e-12 * abs(c0).max())


def test_fine_slices_sum_to_one_slab_and_moliere_3d_projects_to_moliere():
    """Splitting a slab into slices and summing is exact up to quadrature
    round-off; the Moliere 3D form over a slab holding the whole potential
    reproduces the K0 projected form the direct path uses."""
    cl, cnl = O.moliere_3d(38)
    dz = 1.9525
    L = 15 * dz
    parts = [O.fine_spline(O.PT_KIRKLAND_0_12, cl, cnl, -L / 2 + (k + 0.5) * dz, dz, tap=0.999)
             for k in range(15)]
    whole = O.fine_spline(O.PT_KIRKLAND_0_12, cl, cnl, 0.0, L, nq=2048, tap=0.999)
    R = np.sqrt(whole[0])
    m = (R > 0.2) & (R < 2.5)
    s = sum(p[1] for p in parts)
    np.testing.assert_allclose(s[m], whole[1][m], rtol=1e-9)
    proj = O.build_spline(38, tap=0.999)
    np.testing.assert_allclose(whole[1][m], proj[1][m], rtol=1e-6)


def test_fine_spline_taper_and_cubic():
    """Tapered nodes vanish at R_max; each cubic reproduces the node values
    and slopes at both ends of its interval."""
    cl, cnl = O.moliere_3d(22)
    R2, c0, c1, c2, c3 = O.fine_spline(O.PT_KIRKLAND_0_12, cl, cnl, 0.3, 1.9525)
    assert abs(c0[-1]) < 1e-12 * abs(c0).max()
    d = np.diff(R2)
    end = ((c3[:-1] * d + c2[:-1]) * d + c1[:-1]) * d + c0[:-1]
    slope = (3 * c3[:-1] * d + 2 * c2[:-1]) * d + c1[:-1]
    np.testing.assert_allclose(end, c0[1:], rtol=1e-10, atol=1e-12 * abs(c0).max())
    np.testing.assert_allclose(slope, c1[1:], rtol=1e-8, atol=1e-10 * abs(c1).max())


@pytest.mark.parametrize("border", [O.BORDER_MIN, O.BORDER_MAX, O.BORDER_MIDDLE])
def test_slicing_fine_membership_brute_force(border):
    from slicing_ref import brute_force_slicing
    rng = np.random.default_rng(7 + border)
    z = np.sort(rng.uniform(0, 20, 300))
    dz, r_int, eps = 1.35775, 3.0, O.Z_EPS
    b, i0, i1 = O.slicing_fine(z, dz, border, r_int)
    assert b[0] <= z[0] - r_int + 3 * eps and b[-1] >= z[-1] + r_int - 3 * eps
    np.testing.assert_allclose(np.diff(b), dz, rtol=1e-5)
    z32 = z.astype(np.float32)
    f = np.float32
    for s in range(len(i0)):
        zm = f(0.5) * (f(b[s]) + f(b[s + 1]))
        lo = f(min(f(zm - f(r_int)), f(b[s])) - f(eps))
        hi = f(max(f(zm + f(r_int)), f(b[s + 1])) + f(eps))
        idx = np.nonzero((z32 >= lo) & (z32 < hi))[0]
        if len(idx) == 0:
            assert i0[s] > i1[s]
        else:
            assert (i0[s], i1[s]) == (idx[0], idx[-1])
    bf, i0f, i1f = brute_force_slicing(z, dz, border, eps, fine=True, r_int=r_int)
    assert np.array_equal(b, bf) and np.array_equal(i0, i0f) and np.array_equal(i1, i1f)


def test_potential_fine_brute_force_and_slice_sum():
    """Each fine slice equals sigma-free sum over atoms of the potential
    integrated over the slice (adaptive quadrature at a few pixels); summed
    over the slices holding an atom it equals that atom's slab over their
    union."""
    from scipy.integrate import quad
    nx = ny = 24
    lx = ly = 6.0
    x = np.array([2.1, 3.9, 3.0])
    y = np.array([2.5, 3.2, 4.1])
    z = np.array([0.4, 0.9, 1.6])
    occ = np.array([1.0, 0.7, 1.0])
    types = np.array([0, 1, 0])
    pts = [O.PT_KIRKLAND_0_12, O.PT_LOBATO_0_12]
    cls = [O.moliere_3d(14)[0], GOLD["fine_cl_6"]]
    cnls = [O.moliere_3d(14)[1], GOLD["fine_cnl_6"]]
    b, i0, i1 = O.slicing_fine(z, 1.35775, O.BORDER_MIN, 3.0)
    S = len(i0)
    Vs = [O.potential_fine(nx, ny, lx, ly, b[s], b[s + 1], i0[s], i1[s], x, y, z, occ, types, pts, cls,
                           cnls, 1.0) for s in range(S)]
    dx = lx / nx
    checked = 0
    for s in (1, S // 2, S - 2):
        for iy in range(0, ny, 3):
            for ix in range(0, nx, 3):
                Rs = [np.hypot(ix * dx - x[a], iy * dx - y[a]) for a in range(i0[s], i1[s] + 1)]
                # pixels whose atoms are all off the core and off the taper zone,
                # including the spline interval that straddles its start
                if any(0.8 * O.SPLINE_RMAX <= R < O.SPLINE_RMAX or R < 0.2 for R in Rs):
                    continue
                tot = 0.0
                for a, R in zip(range(i0[s], i1[s] + 1), Rs):
                    if R >= O.SPLINE_RMAX:
                        continue
                    cl, cnl = O._coef6(cls[types[a]]), O._coef6(cnls[types[a]])
                    f = lambda zz: O.vr3d(pts[types[a]], np.array([np.hypot(zz, R)]), cl, cnl)[0][0]
                    tot += occ[a] * quad(f, b[s] - z[a], b[s + 1] - z[a], epsrel=1e-12)[0]
                # cubic interpolation between nodes: up to a few 1e-5 relative
                assert abs(Vs[s][iy, ix] - tot) <= 5e-5 * max(1.0, abs(tot))
                checked += 1
    assert checked > 20
    total = np.sum(Vs, axis=0)
    ref = np.zeros((ny, nx))
    for a in range(3):
        ks = [s for s in range(S) if i0[s] <= a <= i1[s]]
        assert ks == list(range(ks[0], ks[-1] + 1))
        zs, ze = b[ks[0]], b[ks[-1] + 1]
        ref += O.potential_fine(nx, ny, lx, ly, zs, ze, a, a, x, y, z, occ, types, pts, cls, cnls, 1.0,
                                nq=2048)
    np.testing.assert_allclose(total, ref, rtol=0, atol=1e-8 * abs(ref).max())


# ----------------------------------------------------------------- partial coherence
@pytest.mark.parametrize("n", [1, 2, 12, 24])
def test_gauss_hermite_matches_numpy(n):
    x, w = O.gauss_hermite(n)
    np.testing.assert_allclose(x, GOLD[f"gh_x_{n}"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(w, GOLD[f"gh_w_{n}"], rtol=1e-12, atol=1e-16)


def _two_beam(n=64, L=20.0, h=10, b=0.3):
    Psi = np.zeros((n, n), complex)
    Psi[0, 0] = n * n
    Psi[0, h] = b * n * n
    return Psi, h / L


def _cross(I, h):
    n = I.shape[0]
    return np.fft.fft2(I)[0, h] / (n * n)


LENS_PC = {"c_30": 1e7, "c_10": -500.0}


def test_partial_coherence_reduces_to_the_ctf():
    """No spread: the PCTF is the CTF image; one node at zero with weight 1
    makes the TCC the CTF image."""
    rng = np.random.default_rng(3)
    n, L = 32, 12.0
    lam = O.constants(300.0)[0]
    Psi = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    lens = dict(LENS_PC, outer_aper_mrad=30.0)
    I, _ = O.ctf_image(Psi, L, L, lam, lens)
    np.testing.assert_allclose(O.pctf_image(Psi, L, L, lam, lens, 0.0, 0.0), I, rtol=1e-13, atol=0)
    np.testing.assert_allclose(O.tcc_image(Psi, L, L, lam, lens, 25.0, 1, 0.4, 1), I, rtol=1e-12, atol=0)


def test_two_beam_temporal_damping_is_the_gaussian_characteristic_function():
    """A Gaussian defocus spread of standard deviation sigma damps the
    interference of beams 0 and g by exp(-(pi lambda sigma g^2)^2 / 2), in
    the PCTF envelope and in the TCC average alike."""
    lam = O.constants(300.0)[0]
    Psi, g = _two_beam()
    sigma = 40.0
    c = _cross(O.pctf_image(Psi, 20.0, 20.0, lam, LENS_PC, 0.0, 0.0), 10)
    p = _cross(O.pctf_image(Psi, 20.0, 20.0, lam, LENS_PC, sigma, 0.0, O.COH_TEMPORAL), 10)
    t = _cross(O.tcc_image(Psi, 20.0, 20.0, lam, LENS_PC, sigma, 30, 0.0, 1, O.COH_TEMPORAL), 10)
    exact = np.exp(-0.5 * (np.pi * lam * sigma * g * g) ** 2)
    assert abs(abs(p) / abs(c) - exact) < 1e-12
    assert abs(abs(t) / abs(c) - exact) < 1e-12
    assert abs(np.angle(t / c)) < 1e-12


def test_two_beam_spatial_damping_follows_the_chi_gradient():
    """Tilt spread: the exact TCC average agrees with the PCTF envelope
    exp(-(pi beta)^2 g^2 (c30 lambda^2 g^2 + c10)^2) to second order in the
    tilt, and rules out the opposite defocus sign."""
    lam = O.constants(300.0)[0]
    Psi, g = _two_beam()
    beta = 0.3
    c = abs(_cross(O.pctf_image(Psi, 20.0, 20.0, lam, LENS_PC, 0.0, 0.0), 10))
    p = abs(_cross(O.pctf_image(Psi, 20.0, 20.0, lam, LENS_PC, 0.0, beta, O.COH_SPATIAL), 10)) / c
    t = abs(_cross(O.tcc_image(Psi, 20.0, 20.0, lam, LENS_PC, 0.0, 1, beta, 20, O.COH_SPATIAL), 10)) / c
    b = beta * 1e-3
    plus = np.exp(-(np.pi * b) ** 2 * g * g * (1e7 * lam ** 2 * g * g - 500.0) ** 2)
    minus = np.exp(-(np.pi * b) ** 2 * g * g * (1e7 * lam ** 2 * g * g + 500.0) ** 2)
    assert abs(p - plus) < 1e-12
    assert abs(t - p) < 2e-3
    assert abs(t - minus) > 0.2
    # joint spreads: the coupled envelope (the 1/u terms) against the 2D + 1D average
    pj = abs(_cross(O.pctf_image(Psi, 20.0, 20.0, lam, LENS_PC, 40.0, beta), 10)) / c
    tj = abs(_cross(O.tcc_image(Psi, 20.0, 20.0, lam, LENS_PC, 40.0, 20, beta, 16), 10)) / c
    assert abs(tj - pj) < 2e-3


def test_tcc_of_a_weak_phase_object_is_close_to_the_pctf():
    """For a weak-phase object the image is linear in the scattered wave, the
    regime the quasi-coherent PCTF describes: the two images agree to a
    small fraction of the contrast."""
    rng = np.random.default_rng(9)
    n, L = 64, 16.0
    lam = O.constants(300.0)[0]
    phase = 0.02 * np.real(np.fft.ifft2(np.fft.fft2(rng.standard_normal((n, n))) *
                                        (np.add.outer(np.fft.fftfreq(n) ** 2, np.fft.fftfreq(n) ** 2) < 0.05)))
    Psi = np.fft.fft2(np.exp(1j * phase))
    lens = dict(LENS_PC, outer_aper_mrad=25.0)
    p = O.pctf_image(Psi, L, L, lam, lens, 30.0, 0.2)
    t = O.tcc_image(Psi, L, L, lam, lens, 30.0, 12, 0.2, 10)
    contrast = np.std(t)
    assert contrast > 1e-3 * np.mean(t)
    assert np.std(p - t) < 0.02 * contrast
