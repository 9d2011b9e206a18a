"""CPU tests of the C ABI and the host logic: the library loads and exports
every entry point of include/cudaem_amd.h, argument validation happens before
any device work, and the host-side physics (slicing, splines, constants)
matches the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cudaem_amd.h")


def header_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:em_status|int|const char\*)\s+(em_\w+)\(", txt, re.M)))


@pytest.fixture(scope="module")
def C():
    from cudaem_amd import _capi
    _capi.lib()
    return _capi


def test_library_exports_every_declared_entry_point(C):
    names = header_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(C.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(C.EXPORTS) == names


def test_lens_layout_is_shared():
    from cudaem_amd import _capi
    txt = open(HEADER).read()
    body = re.search(r"typedef struct em_lens \{(.*?)\} em_lens;", txt, re.S).group(1)
    fields = re.findall(r"\b(c_\d\d|phi_\d\d|m|inner_aper_mrad|outer_aper_mrad)\b", body)
    assert fields == _capi.LENS_FIELDS == O.LENS_FIELDS


def test_fft_sizes(C):
    for n in (64, 128, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096):
        assert C.lib().em_fft_size_supported(n) == 1
    for n in (0, 100, 192, 1000, 1280, 6144, 8192):
        assert C.lib().em_fft_size_supported(n) == 0


def test_validation_before_device_work(C):
    L = C.lib()
    h = ctypes.c_void_p()
    assert L.em_plan_create(None, None, 1, ctypes.byref(h)) == C.EM_ERR_INVALID
    g = C.em_grid(100, 100, 10.0, 10.0, 1.0, 1)
    b = C.em_beam(200.0, 0.0, 0.0)
    assert L.em_plan_create(ctypes.byref(g), ctypes.byref(b), 1, ctypes.byref(h)) == C.EM_ERR_UNSUPPORTED
    g = C.em_grid(128, 128, -1.0, 10.0, 1.0, 1)
    assert L.em_plan_create(ctypes.byref(g), ctypes.byref(b), 1, ctypes.byref(h)) == C.EM_ERR_INVALID
    g = C.em_grid(128, 128, 10.0, 10.0, 1.0, 1)
    assert L.em_plan_create(ctypes.byref(g), ctypes.byref(b), 0, ctypes.byref(h)) == C.EM_ERR_INVALID
    assert L.em_plan_destroy(None) == C.EM_ERR_INVALID
    assert L.em_stem(None, None, 0, None, None, 0, None, None, 1.0, 0, None) == C.EM_ERR_INVALID
    assert L.em_multislice_forward(None, None, None, 0, None, 0, None, 1) == C.EM_ERR_INVALID
    assert L.em_multislice_adjoint(None, None, None, 0, None, 1, None, 0, None, None, None) == C.EM_ERR_INVALID
    assert L.em_multislice_ragged(0, None, None, None, None, None, 0, None, None, 1, None, None) == C.EM_ERR_INVALID


def test_every_plan_entry_point_rejects_a_null_plan(C):
    """Each exported function that takes a plan (or plans) returns
    EM_ERR_INVALID for null handles and zero sizes, before any device call."""
    L = C.lib()
    checked = 0
    for name in sorted(C.EXPORTS):
        if name.startswith("em_host") or name in ("em_fft_size_supported", "em_plan_last_error"):
            continue
        f = getattr(L, name)
        args = [0 if t in (ctypes.c_int, ctypes.c_longlong) else 0.0 if t in (ctypes.c_double, ctypes.c_float)
                else None for t in f.argtypes]
        assert f(*args) == C.EM_ERR_INVALID, name
        checked += 1
    assert checked >= 35
    assert L.em_plan_last_error(None) == b"null plan"


def test_host_constants_match_oracle(C):
    for kv in (80.0, 200.0, 300.0):
        l, g, s = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        assert C.lib().em_host_constants(kv, ctypes.byref(l), ctypes.byref(g), ctypes.byref(s)) == 0
        lo, go, so = O.constants(kv)
        assert (l.value, g.value, s.value) == (lo, go, so)


def _host_slicing(C, z, dz, border):
    z = np.ascontiguousarray(z, dtype=np.float64)
    cap = 4096
    ns = ctypes.c_int()
    b = np.zeros(cap + 1)
    i0 = np.zeros(cap, np.int64)
    i1 = np.zeros(cap, np.int64)
    st = C.lib().em_host_slicing(len(z), z.ctypes.data_as(C._PD), dz, border, cap, ctypes.byref(ns),
                                 b.ctypes.data_as(C._PD), i0.ctypes.data_as(C._PLL), i1.ctypes.data_as(C._PLL))
    assert st == 0
    S = ns.value
    return b[:S + 1], i0[:S], i1[:S]


@pytest.mark.parametrize("border", [0, 1, 2])
def test_host_slicing_matches_oracle(C, border):
    import cudaem_amd.specimen as sp
    rng = np.random.default_rng(border)
    cases = [np.sort(rng.uniform(-5, 30, 700)), np.sort(sp.fcc_cu(3, 3, 10)["z"]),
             np.sort(sp.vibrate(sp.srtio3(2, 2, 40), 0.08, 5, 1)["z"]), np.array([1.0])]
    for z in cases:
        for dz in (1.8075, 1.9525, 0.7):
            b, i0, i1 = _host_slicing(C, z, dz, border)
            bo, i0o, i1o = O.slicing(z, dz, border)
            assert np.array_equal(b, bo) and np.array_equal(i0, i0o) and np.array_equal(i1, i1o)


def _workload_specimens():
    """Seeded phonon-displaced specimens of the four forward workloads, with
    their slice thickness and border: Cu 8x8x10 (HRTEM, 1.8075 A, top
    aligned), SrTiO3 10x10x40 (STEM, 1.9525 A, u_rms 0.08 A, the 2-probe
    general scan and the batched phonon steps), Si diamond 50-400 A
    (dz = a/4)."""
    import cudaem_amd.specimen as sp
    yield "cu", sp.vibrate(sp.fcc_cu(8, 8, 10), 0.08, 21, 1), 1.8075, 1
    yield "cu-static", sp.fcc_cu(8, 8, 10), 1.8075, 1
    for conf in (1, 2, 3, 4):
        yield f"srtio3-{conf}", sp.vibrate(sp.srtio3(10, 10, 40), 0.08, 1234, conf), 1.9525, 0
    for nz in (9, 37, 74):
        yield f"si-{nz}", sp.vibrate(sp.si_diamond(5, 5, nz), 0.08, 7, nz), 5.431 / 4, 0


def test_host_slicing_is_float_on_workload_specimens(C):
    """The product's slicer against a brute-force float32 restatement, bit
    for bit in bounds and every slice's (i0, i1), on the workloads' specimens;
    and on each, with an atom planted where double arithmetic would put it in
    another slice."""
    from slicing_ref import brute_force_slicing, planted_atom
    for name, at, dz, border in _workload_specimens():
        z = np.sort(at["z"], kind="stable")
        for zz in (z, np.sort(np.append(z, planted_atom(z, dz, border)))):
            b, i0, i1 = _host_slicing(C, zz, dz, border)
            bf, i0f, i1f = brute_force_slicing(zz, dz, border, O.Z_EPS)
            assert np.array_equal(b, bf), name
            assert np.array_equal(i0, i0f) and np.array_equal(i1, i1f), name
        bd, i0d, i1d = brute_force_slicing(zz, dz, border, O.Z_EPS, real=np.float64)
        assert not (np.array_equal(i0, i0d) and np.array_equal(i1, i1d)), name


@pytest.mark.parametrize("border", [0, 1, 2])
def test_host_slicing_fine_matches_oracle(C, border):
    """Fine-sliced bounds and atom windows equal the oracle's bit for bit."""
    import cudaem_amd.specimen as sp
    rng = np.random.default_rng(10 + border)
    cases = [np.sort(rng.uniform(-5, 30, 700)), np.sort(sp.vibrate(sp.srtio3(2, 2, 10), 0.08, 5, 1)["z"]),
             np.array([1.0])]
    cap = 4096
    for z in cases:
        for dz, r_int in ((1.9525, 3.0), (1.35775, 2.5), (0.7, 0.0)):
            ns = ctypes.c_int()
            b = np.zeros(cap + 1)
            i0 = np.zeros(cap, np.int64)
            i1 = np.zeros(cap, np.int64)
            st = C.lib().em_host_slicing_fine(len(z), z.ctypes.data_as(C._PD), dz, border, r_int, cap,
                                              ctypes.byref(ns), b.ctypes.data_as(C._PD),
                                              i0.ctypes.data_as(C._PLL), i1.ctypes.data_as(C._PLL))
            assert st == 0
            S = ns.value
            bo, i0o, i1o = O.slicing_fine(z, dz, border, r_int)
            assert np.array_equal(b[:S + 1], bo) and np.array_equal(i0[:S], i0o) and np.array_equal(i1[:S], i1o)


@pytest.mark.parametrize("kind,n", [(0, 7), (0, 128), (1, 1), (1, 12), (1, 24)])
def test_host_quadrature_matches_oracle(C, kind, n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert C.lib().em_host_quadrature(kind, n, x.ctypes.data_as(C._PD), w.ctypes.data_as(C._PD)) == 0
    xo, wo = O.gauss_legendre(n) if kind == 0 else O.gauss_hermite(n)
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-14)
    np.testing.assert_allclose(w, wo, rtol=1e-13, atol=1e-16)


def test_host_slicing_rejects_unsorted(C):
    z = np.array([1.0, 0.5])
    ns = ctypes.c_int()
    assert C.lib().em_host_slicing(2, z.ctypes.data_as(C._PD), 1.0, 0, 10, ctypes.byref(ns), None, None,
                                   None) == C.EM_ERR_INVALID


@pytest.mark.parametrize("Z", [8, 14, 22, 29, 38, 79])
def test_host_spline_matches_oracle(C, Z):
    nR = 128
    arrs = [np.zeros(nR) for _ in range(5)]
    assert C.lib().em_host_spline(Z, nR, O.SPLINE_RMIN, O.SPLINE_RMAX, O.SPLINE_TAP,
                                  *[a.ctypes.data_as(C._PD) for a in arrs]) == 0
    ref = O.build_spline(Z)
    for got, exp in zip(arrs, ref):
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12 * np.abs(exp).max())


# ----------------------------------------------------------------- property-based host logic
from hypothesis import given, settings, strategies as st  # noqa: E402


@settings(max_examples=60, deadline=None)
@given(z=st.lists(st.floats(-50.0, 50.0, allow_nan=False), min_size=1, max_size=300),
       dz=st.floats(0.3, 5.0), border=st.integers(0, 2), r_int=st.floats(0.0, 4.0), fine=st.booleans())
def test_host_slicing_property_matches_oracle(C, z, dz, border, r_int, fine):
    """Random z sets, slice thicknesses, borders and integration radii: the
    product's slicer equals the oracle's bit for bit, and every atom of a
    direct slicing sits in exactly one slice."""
    z = np.ascontiguousarray(np.sort(np.asarray(z, dtype=np.float64)))
    cap = 4096
    ns = ctypes.c_int()
    b = np.zeros(cap + 1)
    i0 = np.zeros(cap, np.int64)
    i1 = np.zeros(cap, np.int64)
    args = (len(z), z.ctypes.data_as(C._PD), dz, border)
    if fine:
        st_ = C.lib().em_host_slicing_fine(*args, r_int, cap, ctypes.byref(ns), b.ctypes.data_as(C._PD),
                                           i0.ctypes.data_as(C._PLL), i1.ctypes.data_as(C._PLL))
        bo, i0o, i1o = O.slicing_fine(z, dz, border, r_int)
    else:
        st_ = C.lib().em_host_slicing(*args, cap, ctypes.byref(ns), b.ctypes.data_as(C._PD),
                                      i0.ctypes.data_as(C._PLL), i1.ctypes.data_as(C._PLL))
        bo, i0o, i1o = O.slicing(z, dz, border)
    assert st_ == 0
    S = ns.value
    assert np.array_equal(b[:S + 1], bo) and np.array_equal(i0[:S], i0o) and np.array_equal(i1[:S], i1o)
    if not fine:
        counts = np.zeros(len(z), int)
        for s in range(S):
            if i0[s] <= i1[s]:
                counts[i0[s]:i1[s] + 1] += 1
        assert (counts <= 1).all()


@settings(max_examples=40, deadline=None)
@given(n=st.integers(1, 40), coefs=st.lists(st.floats(-3.0, 3.0), min_size=1, max_size=80))
def test_host_quadratures_are_exact_on_polynomials(C, n, coefs):
    """Gauss-Legendre on [-1/2, 1/2] and Gauss-Hermite (weight e^{-t^2})
    integrate polynomials of degree < 2n exactly."""
    c = np.asarray(coefs[:2 * n])
    x = np.zeros(n)
    w = np.zeros(n)
    assert C.lib().em_host_quadrature(0, n, x.ctypes.data_as(C._PD), w.ctypes.data_as(C._PD)) == 0
    exact = sum(ck * (0.0 if k % 2 else 2 * 0.5 ** (k + 1) / (k + 1)) for k, ck in enumerate(c))
    assert abs(np.dot(w, np.polyval(c[::-1], x)) - exact) < 1e-12 * max(1.0, np.abs(c).sum())
    assert C.lib().em_host_quadrature(1, n, x.ctypes.data_as(C._PD), w.ctypes.data_as(C._PD)) == 0
    from math import gamma
    exact = sum(ck * (0.0 if k % 2 else gamma((k + 1) / 2)) for k, ck in enumerate(c))
    scale = max(1.0, sum(abs(ck) * gamma((k + 1) / 2) for k, ck in enumerate(c)))
    assert abs(np.dot(w, np.polyval(c[::-1], x)) - exact) < 1e-10 * scale
