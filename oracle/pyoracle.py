"""ctypes front end of the FP64 host oracle (oracle/ms_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Arrays are numpy; complex arrays are complex128 and are passed to the C side
as interleaved doubles.  Lens parameters are dicts with the keys of LENS_FIELDS
(the field order of em_lens in include/cudaem_amd.h).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_ms.so")

LENS_FIELDS = [
    "c_10", "c_12", "phi_12", "c_21", "phi_21", "c_23", "phi_23", "c_30", "c_32", "phi_32",
    "c_34", "phi_34", "c_41", "phi_41", "c_43", "phi_43", "c_45", "phi_45", "c_50", "c_52",
    "phi_52", "c_54", "phi_54", "c_56", "phi_56", "m", "inner_aper_mrad", "outer_aper_mrad",
]

_D = ctypes.c_double
_I = ctypes.c_int
_LL = ctypes.c_longlong
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_PLL = ctypes.POINTER(ctypes.c_longlong)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_constants.argtypes = [_D, _PD, _PD, _PD]
        L.orc_bessel_k01.argtypes = [_D, _PD, _PD]
        L.orc_moliere_coefs.argtypes = [_I, _PD, _PD]
        L.orc_vr.argtypes = [_I, _PD, _PD, _I, _PD, _PD, _I, _PD, _PD, _PD]
        L.orc_build_spline.argtypes = [_I, _PD, _PD, _I, _PD, _PD, _D, _D, _D, _I,
                                       _PD, _PD, _PD, _PD, _PD]
        L.orc_spline_eval.argtypes = [_I, _PD, _I, _PD, _PD, _PD, _PD, _PD, _PD]
        L.orc_fft1d.argtypes = [_PD, _I, _I]
        L.orc_fft2d.argtypes = [_PD, _I, _I, _I]
        L.orc_slicing.argtypes = [_I, _PD, _D, _I, _D, _I, _PD, _PLL, _PLL]
        L.orc_slicing.restype = _I
        L.orc_potential.argtypes = [_I, _I, _D, _D, _LL, _LL, _PD, _PD, _PD, _PI, _I,
                                    _PD, _PD, _PD, _PD, _PD, _PD, _D, _PD]
        L.orc_propagator.argtypes = [_I, _I, _D, _D, _D, _I, _D, _D, _D, _PD]
        L.orc_bwl_params.argtypes = [_I, _I, _D, _D, _PD, _PD]
        L.orc_lens_g2.argtypes = [_PD, _D, _PD, _PD]
        L.orc_chi.argtypes = [_PD, _D, _D, _I, _PD, _PD, _PD]
        L.orc_probe.argtypes = [_I, _I, _D, _D, _D, _PD, _D, _D, _D, _PD]
        L.orc_multislice.argtypes = [_I, _I, _PD, _I, _PD, _PD, _PD, _PD]
        L.orc_detector_sum.argtypes = [_I, _I, _D, _D, _PD, _D, _D]
        L.orc_detector_sum.restype = _D
        L.orc_ctf_image.argtypes = [_I, _I, _D, _D, _D, _PD, _PD, _PD, _PD, _PD]
        L.orc_thickness.argtypes = [_I, _I, _D, _D, _PD, _I, _PD, _PD, _D, _D, _I, _PD]
        L.orc_thickness.restype = _I
        L.orc_stem.argtypes = [_I, _I, _D, _D, _D, _PD, _D, _PD, _I, _PD, _I, _PD, _PD,
                               _I, _PD, _PD, _D, _PD, _I]
        L.orc_stem.restype = _I
        L.orc_stem_f32.argtypes = L.orc_stem.argtypes
        L.orc_stem_f32.restype = _I
        L.orc_adjoint.argtypes = [_I, _I, _PD, _I, _PD, _I, _PD, _PD, _PD, _PD]
        L.orc_gauss_hermite.argtypes = [_I, _PD, _PD]
        L.orc_pctf_image.argtypes = [_I, _I, _D, _D, _D, _PD, _D, _D, _I, _PD, _PD, _PD, _PD]
        L.orc_tcc_image.argtypes = [_I, _I, _D, _D, _D, _PD, _D, _I, _D, _I, _I, _PD, _PD, _PD, _PD]
        L.orc_slicing_fine.argtypes = [_I, _PD, _D, _I, _D, _D, _I, _PD, _PLL, _PLL]
        L.orc_slicing_fine.restype = _I
        L.orc_vr3d.argtypes = [_I, _I, _PD, _PD, _PD, _PD, _PD]
        L.orc_gauss_legendre.argtypes = [_I, _PD, _PD]
        L.orc_moliere_3d.argtypes = [_I, _PD, _PD]
        L.orc_fine_spline.argtypes = [_I, _PD, _PD, _D, _D, _I, _D, _D, _D, _I, _PD, _PD, _PD, _PD, _PD]
        L.orc_potential_fine.argtypes = [_I, _I, _D, _D, _D, _D, _LL, _LL, _PD, _PD, _PD, _PD, _PI,
                                         _PI, _PD, _PD, _I, _D, _D, _D, _I, _D, _PD]
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(_PD) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def lens_array(lens):
    return np.array([float(lens.get(k, 0.0)) for k in LENS_FIELDS], dtype=np.float64)


# --------------------------------------------------------------------------- physics
def constants(e_kv):
    lam, gam, sig = _D(), _D(), _D()
    lib().orc_constants(e_kv, ctypes.byref(lam), ctypes.byref(gam), ctypes.byref(sig))
    return lam.value, gam.value, sig.value


def bessel_k01(x):
    k0, k1 = _D(), _D()
    lib().orc_bessel_k01(x, ctypes.byref(k0), ctypes.byref(k1))
    return k0.value, k1.value


def moliere_coefs(Z):
    a = np.zeros(3)
    b = np.zeros(3)
    lib().orc_moliere_coefs(int(Z), _pd(a), _pd(b))
    return a, b


def vr(a, b, c, d, R):
    a, b, c, d, R = _f64(a), _f64(b), _f64(c), _f64(d), _f64(R)
    V = np.zeros_like(R)
    dV = np.zeros_like(R)
    lib().orc_vr(len(a), _pd(a), _pd(b), len(c), _pd(c), _pd(d), len(R), _pd(R), _pd(V), _pd(dV))
    return V, dV


# The spline conventions the product also uses (cudaem_amd/csrc/em_host.cpp).
SPLINE_NR = 128
SPLINE_RMIN = 0.005
SPLINE_RMAX = 3.0
SPLINE_TAP = 0.85


def build_spline(Z, r_min=SPLINE_RMIN, r_max=SPLINE_RMAX, tap=SPLINE_TAP, nR=SPLINE_NR,
                 coefs=None):
    if coefs is None:
        a, b = moliere_coefs(Z)
        c = np.zeros(0)
        d = np.zeros(0)
    else:
        a, b, c, d = (_f64(x) for x in coefs)
    out = [np.zeros(nR) for _ in range(5)]
    lib().orc_build_spline(len(a), _pd(a), _pd(b), len(c), _pd(c), _pd(d), r_min, r_max, tap,
                           nR, *[_pd(o) for o in out])
    return out  # R2, c0, c1, c2, c3


def spline_eval(x, tab):
    x = _f64(x)
    out = np.zeros_like(x)
    R2, c0, c1, c2, c3 = tab
    lib().orc_spline_eval(len(x), _pd(x), len(R2), _pd(R2), _pd(c0), _pd(c1), _pd(c2), _pd(c3),
                          _pd(out))
    return out


# --------------------------------------------------------------------------- FFT
def fft1d(x, sign=-1):
    a = _c128(x).copy()
    lib().orc_fft1d(a.view(np.float64).ctypes.data_as(_PD), a.size, sign)
    return a


def fft2d(x, sign=-1):
    a = _c128(x).copy()
    ny, nx = a.shape
    lib().orc_fft2d(a.view(np.float64).ctypes.data_as(_PD), nx, ny, sign)
    return a


# --------------------------------------------------------------------------- slicing / potential
Z_EPS = 1e-4
BORDER_MIN, BORDER_MAX, BORDER_MIDDLE = 0, 1, 2


def slicing(z_sorted, dz, border, eps=Z_EPS, max_slices=100000):
    z = _f64(z_sorted)
    bounds = np.zeros(max_slices + 1)
    i0 = np.zeros(max_slices, dtype=np.int64)
    i1 = np.zeros(max_slices, dtype=np.int64)
    S = lib().orc_slicing(len(z), _pd(z), dz, border, eps, max_slices, _pd(bounds),
                          i0.ctypes.data_as(_PLL), i1.ctypes.data_as(_PLL))
    if S < 0:
        raise ValueError("too many slices")
    return bounds[:S + 1].copy(), i0[:S].copy(), i1[:S].copy()


class Tables:
    """Per-element spline tables indexed by type (0..ntypes-1)."""

    def __init__(self, Zs, **kw):
        self.Zs = list(Zs)
        tabs = [build_spline(Z, **kw) for Z in self.Zs]
        self.nR = len(tabs[0][0])
        self.R2, self.c0, self.c1, self.c2, self.c3 = (
            np.ascontiguousarray(np.stack([t[i] for t in tabs])) for i in range(5))
        self.rmax2 = np.array([t[0][-1] for t in tabs])


def potential(nx, ny, lx, ly, i0, i1, x, y, occ, types, tables, coeff):
    x, y, occ = _f64(x), _f64(y), _f64(occ)
    types = np.ascontiguousarray(types, dtype=np.int32)
    V = np.zeros((ny, nx))
    t = tables
    lib().orc_potential(nx, ny, lx, ly, int(i0), int(i1), _pd(x), _pd(y), _pd(occ),
                        types.ctypes.data_as(_PI), t.nR, _pd(t.R2), _pd(t.c0), _pd(t.c1),
                        _pd(t.c2), _pd(t.c3), _pd(t.rmax2), coeff, _pd(V))
    return V


# --------------------------------------------------------------------------- fine slicing
# Parameterisation types of the 3D atomic potential (the build's numbering).
PT_DOYLE_0_4, PT_PENG_0_4, PT_PENG_0_12, PT_KIRKLAND_0_12, PT_WEICKENMEIER_0_12, PT_LOBATO_0_12 = 1, 2, 3, 4, 5, 6
# Gauss-Legendre points per slice of the z integration (the product uses the same).
FINE_NQ = 128


def _coef6(c):
    out = np.zeros(6)
    c = np.asarray(c, dtype=np.float64).ravel()
    out[:len(c)] = c
    return out


def vr3d(pt, r, cl, cnl):
    r = _f64(r)
    cl, cnl = _coef6(cl), _coef6(cnl)
    V = np.zeros_like(r)
    d = np.zeros_like(r)
    lib().orc_vr3d(pt, len(r), _pd(r), _pd(cl), _pd(cnl), _pd(V), _pd(d))
    return V, d


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _pd(x), _pd(w))
    return x, w


def moliere_3d(Z):
    cl = np.zeros(6)
    cnl = np.zeros(6)
    lib().orc_moliere_3d(int(Z), _pd(cl), _pd(cnl))
    return cl, cnl


def fine_spline(pt, cl, cnl, a, dz, nq=FINE_NQ, r_min=SPLINE_RMIN, r_max=SPLINE_RMAX, tap=SPLINE_TAP,
                nR=SPLINE_NR):
    cl, cnl = _coef6(cl), _coef6(cnl)
    out = [np.zeros(nR) for _ in range(5)]
    lib().orc_fine_spline(pt, _pd(cl), _pd(cnl), a, dz, nq, r_min, r_max, tap, nR, *[_pd(o) for o in out])
    return out  # R2, c0, c1, c2, c3


def slicing_fine(z_sorted, dz, border, r_int, eps=Z_EPS, max_slices=100000):
    z = _f64(z_sorted)
    bounds = np.zeros(max_slices + 1)
    i0 = np.zeros(max_slices, dtype=np.int64)
    i1 = np.zeros(max_slices, dtype=np.int64)
    S = lib().orc_slicing_fine(len(z), _pd(z), dz, border, eps, r_int, max_slices, _pd(bounds),
                               i0.ctypes.data_as(_PLL), i1.ctypes.data_as(_PLL))
    if S < 0:
        raise ValueError("too many slices")
    return bounds[:S + 1].copy(), i0[:S].copy(), i1[:S].copy()


def potential_fine(nx, ny, lx, ly, zs, ze, i0, i1, x, y, z, occ, types, pts, cls, cnls, coeff,
                   nq=FINE_NQ, r_min=SPLINE_RMIN, r_max=SPLINE_RMAX, tap=SPLINE_TAP, nR=SPLINE_NR):
    """One fine-sliced slice; element e uses pts[e], cls[e], cnls[e]."""
    x, y, z, occ = _f64(x), _f64(y), _f64(z), _f64(occ)
    types = np.ascontiguousarray(types, dtype=np.int32)
    pts = np.ascontiguousarray(pts, dtype=np.int32)
    cl = np.ascontiguousarray(np.stack([_coef6(c) for c in cls]))
    cnl = np.ascontiguousarray(np.stack([_coef6(c) for c in cnls]))
    V = np.zeros((ny, nx))
    lib().orc_potential_fine(nx, ny, lx, ly, zs, ze, int(i0), int(i1), _pd(x), _pd(y), _pd(z), _pd(occ),
                             types.ctypes.data_as(_PI), pts.ctypes.data_as(_PI), _pd(cl), _pd(cnl), nq,
                             r_min, r_max, tap, nR, coeff, _pd(V))
    return V


def propagator(nx, ny, lx, ly, dz, bwl, lam, gx0=0.0, gy0=0.0):
    P = np.zeros((ny, nx), dtype=np.complex128)
    lib().orc_propagator(nx, ny, lx, ly, dz, int(bwl), lam, gx0, gy0,
                         P.view(np.float64).ctypes.data_as(_PD))
    return P


def bwl_params(nx, ny, lx, ly):
    a, b = _D(), _D()
    lib().orc_bwl_params(nx, ny, lx, ly, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def lens_g2(lens, lam):
    la = lens_array(lens)
    a, b = _D(), _D()
    lib().orc_lens_g2(_pd(la), lam, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def chi(lens, lam, gx, gy, defocus_shift=0.0):
    la = lens_array(lens)
    gx, gy = _f64(gx), _f64(gy)
    out = np.zeros_like(gx)
    lib().orc_chi(_pd(la), lam, defocus_shift, gx.size, _pd(gx), _pd(gy), _pd(out))
    return out


def probe(nx, ny, lx, ly, lam, lens, xb, yb, defocus_shift=0.0):
    la = lens_array(lens)
    psi = np.zeros((ny, nx), dtype=np.complex128)
    lib().orc_probe(nx, ny, lx, ly, lam, _pd(la), defocus_shift, xb, yb,
                    psi.view(np.float64).ctypes.data_as(_PD))
    return psi


def multislice(P, psi, V=None, t=None, checkpoints=False):
    ny, nx = P.shape
    P = _c128(P)
    out = _c128(psi).copy()
    S = (V.shape[0] if V is not None else t.shape[0])
    Vd = _f64(V) if V is not None else None
    td = _c128(t) if t is not None else None
    ck = np.zeros((S, ny, nx), dtype=np.complex128) if checkpoints else None
    lib().orc_multislice(nx, ny, P.view(np.float64).ctypes.data_as(_PD), S, _pd(Vd),
                         td.view(np.float64).ctypes.data_as(_PD) if td is not None else None,
                         out.view(np.float64).ctypes.data_as(_PD),
                         ck.view(np.float64).ctypes.data_as(_PD) if ck is not None else None)
    return (out, ck) if checkpoints else out


def detector_sum(Psi, lx, ly, g2min, g2max):
    Psi = _c128(Psi)
    ny, nx = Psi.shape
    return lib().orc_detector_sum(nx, ny, lx, ly, Psi.view(np.float64).ctypes.data_as(_PD),
                                  g2min, g2max)


def ctf_image(Psi, lx, ly, lam, lens):
    Psi = _c128(Psi)
    ny, nx = Psi.shape
    la = lens_array(lens)
    I = np.zeros((ny, nx))
    re = np.zeros((ny, nx))
    im = np.zeros((ny, nx))
    lib().orc_ctf_image(nx, ny, lx, ly, lam, _pd(la), Psi.view(np.float64).ctypes.data_as(_PD),
                        _pd(I), _pd(re), _pd(im))
    return I, re + 1j * im


# --------------------------------------------------------------------------- partial coherence
COH_BOTH, COH_TEMPORAL, COH_SPATIAL = 1, 2, 3


def gauss_hermite(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_hermite(n, _pd(x), _pd(w))
    return x, w


def pctf_image(Psi, lx, ly, lam, lens, sigma, beta_mrad, mode=COH_BOTH, full=False):
    """Quasi-coherent image: returns |.|^2, or (re, im) with full=True."""
    a = _c128(Psi)
    ny, nx = a.shape
    I, re, im = np.zeros((ny, nx)), np.zeros((ny, nx)), np.zeros((ny, nx))
    lib().orc_pctf_image(nx, ny, lx, ly, lam, _pd(lens_array(lens)), sigma, beta_mrad, mode,
                         a.view(np.float64).ctypes.data_as(_PD), _pd(I), _pd(re), _pd(im))
    return (re, im) if full else I


def tcc_image(Psi, lx, ly, lam, lens, sigma, nt, beta_mrad, ns, mode=COH_BOTH, full=False):
    a = _c128(Psi)
    ny, nx = a.shape
    I, re, im = np.zeros((ny, nx)), np.zeros((ny, nx)), np.zeros((ny, nx))
    lib().orc_tcc_image(nx, ny, lx, ly, lam, _pd(lens_array(lens)), sigma, nt, beta_mrad, ns, mode,
                        a.view(np.float64).ctypes.data_as(_PD), _pd(I), _pd(re), _pd(im))
    return (re, im) if full else I


def thickness(P, V, psi, lx, ly, g2min, g2max, stride=1):
    S, ny, nx = V.shape
    w = _c128(psi).copy()
    frames = np.zeros((S, ny, nx))
    nf = lib().orc_thickness(nx, ny, lx, ly, _c128(P).view(np.float64).ctypes.data_as(_PD), S,
                             _pd(_f64(V)), w.view(np.float64).ctypes.data_as(_PD), g2min, g2max,
                             stride, _pd(frames))
    return frames[:nf].copy(), w


def stem(nx, ny, lx, ly, lam, lens, defocus_shift, P, V, xb, yb, g2min, g2max, w=1.0,
         nthreads=0, fp32=False):
    """Detector sums [ndet][B] of probes at (xb, yb), one OpenMP thread per
    probe.  fp32: the same loop with the wave, tables and transforms in FP32
    (a timing variant; the oracle is the FP64 loop)."""
    la = lens_array(lens)
    xb, yb = _f64(xb), _f64(yb)
    g2min, g2max = _f64(g2min), _f64(g2max)
    B = len(xb)
    out = np.zeros((len(g2min), B))
    fn = lib().orc_stem_f32 if fp32 else lib().orc_stem
    used = fn(nx, ny, lx, ly, lam, _pd(la), defocus_shift,
                          _c128(P).view(np.float64).ctypes.data_as(_PD), V.shape[0],
                          _pd(_f64(V)), B, _pd(xb), _pd(yb), len(g2min), _pd(g2min),
                          _pd(g2max), w, _pd(out), int(nthreads))
    return out, used


def adjoint(P, t, psi0, gPsi):
    """grad of L w.r.t. t (dL/dRe + i dL/dIm) given gPsi = dL/dRe(Psi) + i dL/dIm(Psi)."""
    S, ny, nx = t.shape
    B = psi0.shape[0]
    t = _c128(t)
    psi0 = _c128(psi0)
    gPsi = _c128(gPsi)
    gt = np.zeros((S, ny, nx), dtype=np.complex128)
    Psi = np.zeros((B, ny, nx), dtype=np.complex128)
    v = lambda a: a.view(np.float64).ctypes.data_as(_PD)
    lib().orc_adjoint(nx, ny, v(_c128(P)), S, v(t), B, v(psi0), v(gPsi), v(gt), v(Psi))
    return gt, Psi
