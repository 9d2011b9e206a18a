"""Python front end of the C ABI: a Plan owns one em_plan.

Device buffers are torch tensors on the GPU (complex64 waves / transmissions,
float32 potentials and intensities).  Work is enqueued on torch's current
stream.  Shapes follow the C ABI: batches outermost, then [ny][nx].
"""
import ctypes

import numpy as np

from . import _capi as C


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dptr(a):
    return a.ctypes.data_as(C._PD)


def _need(t, shape, dtype, name):
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be on the GPU")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def lens_struct(lens):
    s = C.em_lens()
    for k in C.LENS_FIELDS:
        setattr(s, k, float((lens or {}).get(k, 0.0)))
    return s


def fft_size_supported(n):
    return bool(C.lib().em_fft_size_supported(int(n)))


class Plan:
    def __init__(self, nx, ny, lx, ly, dz, e0_kv, bwl=True, theta=0.0, phi=0.0, max_batch=1):
        self.nx, self.ny, self.lx, self.ly, self.dz = int(nx), int(ny), float(lx), float(ly), float(dz)
        self.e0_kv = float(e0_kv)
        self.max_batch = int(max_batch)
        g = C.em_grid(self.nx, self.ny, self.lx, self.ly, self.dz, 1 if bwl else 0)
        b = C.em_beam(self.e0_kv, float(theta), float(phi))
        h = ctypes.c_void_p()
        C.check(C.lib().em_plan_create(ctypes.byref(g), ctypes.byref(b), self.max_batch, ctypes.byref(h)),
                None, "em_plan_create")
        self._h = h
        self.n_slices = 0
        lam, sig, gam = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        C.check(C.lib().em_plan_constants(h, ctypes.byref(lam), ctypes.byref(sig), ctypes.byref(gam)), h)
        self.lam, self.sigma, self.gamma = lam.value, sig.value, gam.value

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            C.lib().em_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def _sync_stream(self):
        torch = _torch()
        s = torch.cuda.current_stream().cuda_stream
        C.check(C.lib().em_plan_set_stream(self._h, ctypes.c_void_p(s)), self._h, "set_stream")

    def _call(self, name, *args):
        self._sync_stream()
        C.check(getattr(C.lib(), name)(self._h, *args), self._h, name)

    @property
    def N(self):
        return self.nx * self.ny

    # -- configuration ------------------------------------------------------
    def set_probe_group(self, g):
        C.check(C.lib().em_plan_set_probe_group(self._h, int(g)), self._h)

    def group_size(self, nprobe):
        """Probes per group the fused passes use for a batch of nprobe probes."""
        g = ctypes.c_int(0)
        C.check(C.lib().em_plan_group_size(self._h, int(nprobe), ctypes.byref(g)), self._h)
        return g.value

    def set_pass_timing(self, on):
        """Record the duration of every interior row / column pass launch of
        the following stem() calls (they then run without graph capture)."""
        C.check(C.lib().em_plan_set_pass_timing(self._h, 1 if on else 0), self._h)

    def pass_times(self):
        """(row ms, column ms, row launches, column launches): average launch
        durations recorded since set_pass_timing(True); clears the records."""
        r, c = ctypes.c_double(0), ctypes.c_double(0)
        nr, nc = ctypes.c_int(0), ctypes.c_int(0)
        C.check(C.lib().em_plan_pass_times(self._h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(nr),
                                           ctypes.byref(nc)), self._h)
        return r.value, c.value, nr.value, nc.value

    def tuning(self):
        t = C.em_tuning()
        C.check(C.lib().em_plan_get_tuning(self._h, ctypes.byref(t)), self._h)
        return {k: getattr(t, k) for k, _ in C.em_tuning._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = C.em_tuning(**cur)
        C.check(C.lib().em_plan_set_tuning(self._h, ctypes.byref(t)), self._h, "set_tuning")

    def set_band_skip(self, threshold=1e-9):
        """Skip propagator columns below threshold (relative); 0 disables.
        Returns the skipped column range (lo, hi), lo > hi when none."""
        lo, hi = ctypes.c_int(), ctypes.c_int()
        C.check(C.lib().em_plan_set_band_skip(self._h, float(threshold), ctypes.byref(lo), ctypes.byref(hi)),
                self._h, "set_band_skip")
        return lo.value, hi.value

    def autotune(self, nprobe, nslices=4):
        """Times candidate launch shapes on this grid and keeps the fastest;
        returns (tuning, ms per probe-slice)."""
        ms = ctypes.c_double()
        self._call("em_plan_autotune", int(nprobe), int(nslices), ctypes.byref(ms))
        return self.tuning(), ms.value

    def set_element_coefs(self, Z, a, b, c=(), d=()):
        a, b, c, d = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d))
        C.check(C.lib().em_set_element_coefs(self._h, int(Z), len(a), _dptr(a), _dptr(b), len(c),
                                             _dptr(c), _dptr(d)), self._h)

    def set_spline_params(self, nR=128, r_min=0.005, r_max=3.0, tap=0.85):
        C.check(C.lib().em_set_spline_params(self._h, int(nR), r_min, r_max, tap), self._h)

    def set_element_params(self, Z, ptype, cl, cnl):
        """3D potential of element Z for fine slicing: parameterisation type
        (C.PT_*) and up to 6 linear / non-linear coefficients."""
        cl6 = np.zeros(6)
        cnl6 = np.ones(6)
        cl, cnl = np.asarray(cl, dtype=np.float64).ravel(), np.asarray(cnl, dtype=np.float64).ravel()
        if len(cl) > 6 or len(cnl) != len(cl):
            raise ValueError("up to 6 coefficient pairs")
        cl6[:len(cl)] = cl
        cnl6[:len(cnl)] = cnl
        C.check(C.lib().em_set_element_params(self._h, int(Z), int(ptype), _dptr(cl6), _dptr(cnl6)), self._h)

    def set_fine_slicing(self, enable=True, nq=128):
        """Fine slicing from the next set_specimen: per (atom, slice) splines of
        the 3D potential integrated over the slice with nq Gauss-Legendre points."""
        C.check(C.lib().em_plan_set_fine_slicing(self._h, 1 if enable else 0, int(nq)), self._h)

    def set_specimen(self, x, y, z, Z, occ=None, border=C.BORDER_MIN):
        x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        occ = None if occ is None else np.ascontiguousarray(occ, dtype=np.float64)
        n = len(x)
        if not (len(y) == n and len(z) == n and len(Z) == n and (occ is None or len(occ) == n)):
            raise ValueError("atom arrays must have equal length")
        a = C.em_atoms(n, _dptr(x), _dptr(y), _dptr(z), Z.ctypes.data_as(C._PI),
                       _dptr(occ) if occ is not None else None)
        ns = ctypes.c_int()
        C.check(C.lib().em_plan_set_specimen(self._h, ctypes.byref(a), int(border), ctypes.byref(ns)),
                self._h, "set_specimen")
        self.n_slices = ns.value
        return ns.value

    def slice_info(self, s):
        zs, ze = ctypes.c_double(), ctypes.c_double()
        i0, i1 = ctypes.c_longlong(), ctypes.c_longlong()
        C.check(C.lib().em_slice_info(self._h, int(s), ctypes.byref(zs), ctypes.byref(ze),
                                      ctypes.byref(i0), ctypes.byref(i1)), self._h)
        return zs.value, ze.value, i0.value, i1.value

    def spline_tables(self):
        nt, nR = ctypes.c_int(), ctypes.c_int()
        C.check(C.lib().em_spline_tables(self._h, 0, None, ctypes.byref(nt), ctypes.byref(nR),
                                         None, None, None, None, None), self._h)
        Zs = np.zeros(nt.value, dtype=np.int32)
        arrs = [np.zeros((nt.value, nR.value)) for _ in range(5)]
        C.check(C.lib().em_spline_tables(self._h, nt.value, Zs.ctypes.data_as(C._PI), ctypes.byref(nt),
                                         ctypes.byref(nR), *[_dptr(a) for a in arrs]), self._h)
        return Zs, arrs

    # -- operators ------------------------------------------------------------
    def potential_slices(self, s0=0, ns=None, out=None):
        torch = _torch()
        ns = self.n_slices - s0 if ns is None else ns
        if out is None:
            out = torch.empty((ns, self.ny, self.nx), dtype=torch.float32, device="cuda")
        _need(out, (ns, self.ny, self.nx), torch.float32, "V")
        self._call("em_potential_slices", int(s0), int(ns), _ptr(out))
        return out

    def propagator(self):
        torch = _torch()
        P = torch.empty((self.ny, self.nx), dtype=torch.complex64, device="cuda")
        self._call("em_propagator", _ptr(P))
        return P

    def build_probes(self, lens, x, y, defocus_shift=0.0, out=None):
        torch = _torch()
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        B = len(x)
        if out is None:
            out = torch.empty((B, self.ny, self.nx), dtype=torch.complex64, device="cuda")
        _need(out, (B, self.ny, self.nx), torch.complex64, "psi")
        ls = lens_struct(lens)
        self._call("em_build_probes", ctypes.byref(ls), float(defocus_shift), _dptr(x), _dptr(y), B, _ptr(out))
        return out

    def plane_wave(self, B=1, out=None):
        torch = _torch()
        if out is None:
            out = torch.empty((B, self.ny, self.nx), dtype=torch.complex64, device="cuda")
        _need(out, (B, self.ny, self.nx), torch.complex64, "psi")
        self._call("em_plane_wave", B, _ptr(out))
        return out

    def fft2d(self, data, direction=-1):
        torch = _torch()
        B = data.shape[0]
        _need(data, (B, self.ny, self.nx), torch.complex64, "data")
        self._call("em_fft2d", _ptr(data), B, int(direction))
        return data

    @staticmethod
    def n_checkpoints(S, every):
        return (S + every - 1) // every

    def multislice_forward(self, psi, V=None, t=None, ckpt=None, every=1):
        """psi: incident waves [B][ny][nx] (real space) -> exit waves (Fourier,
        unnormalised), in place.  ckpt: optional [ceil(S/every)][B][ny][nx]."""
        torch = _torch()
        if (V is None) == (t is None):
            raise ValueError("give exactly one of V and t")
        if every < 1:
            raise ValueError("every >= 1")
        B = psi.shape[0]
        _need(psi, (B, self.ny, self.nx), torch.complex64, "psi")
        S = (V if V is not None else t).shape[0]
        if V is not None:
            _need(V, (S, self.ny, self.nx), torch.float32, "V")
        else:
            _need(t, (S, self.ny, self.nx), torch.complex64, "t")
        if ckpt is not None:
            _need(ckpt, (self.n_checkpoints(S, every), B, self.ny, self.nx), torch.complex64, "ckpt")
        self._call("em_multislice_forward", _ptr(V), _ptr(t), S, _ptr(psi), B, _ptr(ckpt), int(every))
        return psi

    def multislice_forward_multi(self, psi, Vs):
        """psi: incident waves [B][ny][nx] (real space) -> exit waves (Fourier,
        unnormalised), in place; probe b runs through its own potential stack
        Vs[b] ([S_b][ny][nx] float32, sigma-scaled).  One batch for, e.g., the
        frozen-phonon configurations of one probe position."""
        torch = _torch()
        B = psi.shape[0]
        _need(psi, (B, self.ny, self.nx), torch.complex64, "psi")
        if len(Vs) != B:
            raise ValueError("one potential stack per probe")
        for V in Vs:
            _need(V, (V.shape[0], self.ny, self.nx), torch.float32, "V")
        arrV = (ctypes.c_void_p * B)(*[V.data_ptr() for V in Vs])
        S = (ctypes.c_int * B)(*[V.shape[0] for V in Vs])
        self._call("em_multislice_forward_multi", arrV, S, _ptr(psi), B)
        return psi

    def multislice_adjoint(self, ckpt, gPsi, V=None, t=None, gt=None, gV=None, gpsi0=None, every=1):
        """Gradient w.r.t. the transmission of a loss whose exit-wave gradient is
        gPsi (dL/dRe + i dL/dIm; overwritten).  Returns gt [S][ny][nx]."""
        torch = _torch()
        if (V is None) == (t is None):
            raise ValueError("give exactly one of V and t")
        S = (V if V is not None else t).shape[0]
        B = gPsi.shape[0]
        _need(gPsi, (B, self.ny, self.nx), torch.complex64, "gPsi")
        _need(ckpt, (self.n_checkpoints(S, every), B, self.ny, self.nx), torch.complex64, "ckpt")
        if V is not None:
            _need(V, (S, self.ny, self.nx), torch.float32, "V")
        else:
            _need(t, (S, self.ny, self.nx), torch.complex64, "t")
        if gt is None:
            gt = torch.empty((S, self.ny, self.nx), dtype=torch.complex64, device="cuda")
        _need(gt, (S, self.ny, self.nx), torch.complex64, "gt")
        if gV is not None:
            _need(gV, (S, self.ny, self.nx), torch.float32, "gV")
        if gpsi0 is not None:
            _need(gpsi0, (B, self.ny, self.nx), torch.complex64, "gpsi0")
        self._call("em_multislice_adjoint", _ptr(V), _ptr(t), S, _ptr(ckpt), int(every), _ptr(gPsi), B, _ptr(gt),
                   _ptr(gV), _ptr(gpsi0))
        return gt

    @staticmethod
    def _angles(inner, outer):
        inner = np.ascontiguousarray(np.atleast_1d(inner), dtype=np.float64)
        outer = np.ascontiguousarray(np.atleast_1d(outer), dtype=np.float64)
        if inner.shape != outer.shape or not (1 <= len(inner) <= 8):
            raise ValueError("1..8 detectors with matching inner/outer angles")
        return inner, outer

    def detector_sums(self, Psi, inner_mrad, outer_mrad, w=1.0, out=None, accumulate=False):
        torch = _torch()
        B = Psi.shape[0]
        _need(Psi, (B, self.ny, self.nx), torch.complex64, "Psi")
        inner, outer = self._angles(inner_mrad, outer_mrad)
        if out is None:
            out = torch.zeros((len(inner), B), dtype=torch.float32, device="cuda")
        _need(out, (len(inner), B), torch.float32, "out")
        self._call("em_detector_sums", _ptr(Psi), B, len(inner), _dptr(inner), _dptr(outer), float(w),
                   1 if accumulate else 0, _ptr(out))
        return out

    def ctf_image(self, lens, Psi, full_wave=False, w=1.0, out=None, accumulate=False):
        torch = _torch()
        B = Psi.shape[0]
        _need(Psi, (B, self.ny, self.nx), torch.complex64, "Psi")
        ls = lens_struct(lens)
        if full_wave:
            re, im = out if out is not None else (
                torch.zeros((B, self.ny, self.nx), dtype=torch.float32, device="cuda"),
                torch.zeros((B, self.ny, self.nx), dtype=torch.float32, device="cuda"))
            _need(re, (B, self.ny, self.nx), torch.float32, "re")
            _need(im, (B, self.ny, self.nx), torch.float32, "im")
            self._call("em_ctf_image", ctypes.byref(ls), _ptr(Psi), B, None, _ptr(re), _ptr(im), float(w),
                       1 if accumulate else 0)
            return re, im
        I = out if out is not None else torch.zeros((B, self.ny, self.nx), dtype=torch.float32, device="cuda")
        _need(I, (B, self.ny, self.nx), torch.float32, "I")
        self._call("em_ctf_image", ctypes.byref(ls), _ptr(Psi), B, _ptr(I), None, None, float(w),
                   1 if accumulate else 0)
        return I

    def _coherent_image(self, fn, lens, coh, Psi, full_wave, w, out, accumulate):
        torch = _torch()
        B = Psi.shape[0]
        _need(Psi, (B, self.ny, self.nx), torch.complex64, "Psi")
        ls = lens_struct(lens)
        cs = C.em_coherence(float(coh.get("dsf_sigma", 0.0)), int(coh.get("dsf_npoints", 1)),
                            float(coh.get("ssf_beta", 0.0)), int(coh.get("ssf_npoints", 1)),
                            int(coh.get("mode", C.COH_TEMPORAL_SPATIAL)))
        if full_wave:
            re, im = out if out is not None else (
                torch.zeros((B, self.ny, self.nx), dtype=torch.float32, device="cuda"),
                torch.zeros((B, self.ny, self.nx), dtype=torch.float32, device="cuda"))
            _need(re, (B, self.ny, self.nx), torch.float32, "re")
            _need(im, (B, self.ny, self.nx), torch.float32, "im")
            self._call(fn, ctypes.byref(ls), ctypes.byref(cs), _ptr(Psi), B, None, _ptr(re), _ptr(im), float(w),
                       1 if accumulate else 0)
            return re, im
        I = out if out is not None else torch.zeros((B, self.ny, self.nx), dtype=torch.float32, device="cuda")
        _need(I, (B, self.ny, self.nx), torch.float32, "I")
        self._call(fn, ctypes.byref(ls), ctypes.byref(cs), _ptr(Psi), B, _ptr(I), None, None, float(w),
                   1 if accumulate else 0)
        return I

    def pctf_image(self, lens, coh, Psi, full_wave=False, w=1.0, out=None, accumulate=False):
        """Quasi-coherent image; coh: dict dsf_sigma (A), ssf_beta (mrad), mode."""
        return self._coherent_image("em_pctf_image", lens, coh, Psi, full_wave, w, out, accumulate)

    def tcc_image(self, lens, coh, Psi, full_wave=False, w=1.0, out=None, accumulate=False):
        """Image averaged over defocus / tilt Gauss-Hermite nodes; coh adds
        dsf_npoints and ssf_npoints."""
        return self._coherent_image("em_tcc_image", lens, coh, Psi, full_wave, w, out, accumulate)

    def diffraction(self, Psi, w=1.0, out=None, accumulate=False):
        torch = _torch()
        B = Psi.shape[0]
        _need(Psi, (B, self.ny, self.nx), torch.complex64, "Psi")
        I = out if out is not None else torch.zeros((B, self.ny, self.nx), dtype=torch.float32, device="cuda")
        _need(I, (B, self.ny, self.nx), torch.float32, "I")
        self._call("em_diffraction", _ptr(Psi), B, _ptr(I), float(w), 1 if accumulate else 0)
        return I

    def stem(self, lens, x, y, inner_mrad, outer_mrad, w=1.0, out=None, accumulate=False):
        torch = _torch()
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if len(x) != len(y) or len(x) == 0:
            raise ValueError("probe positions")
        inner, outer = self._angles(inner_mrad, outer_mrad)
        if out is None:
            out = torch.zeros((len(inner), len(x)), dtype=torch.float32, device="cuda")
        _need(out, (len(inner), len(x)), torch.float32, "out")
        ls = lens_struct(lens)
        self._call("em_stem", ctypes.byref(ls), len(x), _dptr(x), _dptr(y), len(inner), _dptr(inner),
                   _dptr(outer), float(w), 1 if accumulate else 0, _ptr(out))
        return out

    def time_passes(self, V, wave, iters=20):
        """Average device ms of one fused row pass and one column pass over
        wave.shape[0] probes (HIP events on torch's current stream)."""
        torch = _torch()
        B = wave.shape[0]
        _need(V, (self.ny, self.nx), torch.float32, "V")
        _need(wave, (B, self.ny, self.nx), torch.complex64, "wave")
        r, c = ctypes.c_float(), ctypes.c_float()
        self._call("em_time_passes", _ptr(V), _ptr(wave), B, int(iters), ctypes.byref(r), ctypes.byref(c))
        return r.value, c.value

    def thickness(self, V, psi, inner_mrad, outer_mrad, stride=1, w=1.0, frames=None, accumulate=False):
        torch = _torch()
        S = V.shape[0]
        B = psi.shape[0]
        _need(V, (S, self.ny, self.nx), torch.float32, "V")
        _need(psi, (B, self.ny, self.nx), torch.complex64, "psi")
        nf_max = S if stride <= 1 else (S + stride - 1) // stride + 1
        if frames is None:
            frames = torch.zeros((nf_max, B, self.ny, self.nx), dtype=torch.float32, device="cuda")
        if frames.shape[0] < nf_max:
            raise ValueError("frames too small")
        _need(frames, (frames.shape[0], B, self.ny, self.nx), torch.float32, "frames")
        nf = ctypes.c_int()
        self._call("em_thickness", _ptr(V), S, _ptr(psi), B, float(inner_mrad), float(outer_mrad), int(stride),
                   float(w), 1 if accumulate else 0, _ptr(frames), ctypes.byref(nf))
        return frames[:nf.value]

    def thickness_multi(self, Vs, psi, inner_mrad, outer_mrad, stride=1, w=1.0, frames=None, accumulate=False):
        """thickness() with a potential stack per wave: Vs[b] is [S][ny][nx]
        (one S for all), e.g. the frozen-phonon configurations of one probe."""
        torch = _torch()
        B = psi.shape[0]
        if len(Vs) != B or B == 0:
            raise ValueError("one potential stack per wave")
        S = Vs[0].shape[0]
        for V in Vs:
            _need(V, (S, self.ny, self.nx), torch.float32, "V")
        _need(psi, (B, self.ny, self.nx), torch.complex64, "psi")
        nf_max = S if stride <= 1 else (S + stride - 1) // stride + 1
        if frames is None:
            frames = torch.zeros((nf_max, B, self.ny, self.nx), dtype=torch.float32, device="cuda")
        if frames.shape[0] < nf_max:
            raise ValueError("frames too small")
        _need(frames, (frames.shape[0], B, self.ny, self.nx), torch.float32, "frames")
        arrV = (ctypes.c_void_p * B)(*[V.data_ptr() for V in Vs])
        nf = ctypes.c_int()
        self._call("em_thickness_multi", arrV, S, _ptr(psi), B, float(inner_mrad), float(outer_mrad), int(stride),
                   float(w), 1 if accumulate else 0, _ptr(frames), ctypes.byref(nf))
        return frames[:nf.value]


def ragged(plans, Vs, psis, inner_mrad, outer_mrad, stride, outs=None):
    """One graph launch over requests on different grids (em_multislice_ragged).

    Returns per request a tensor [n_out, ndet, B] of detector sums after every
    `stride` slices and after the last one: `outs` when given (reused, so a
    repeated request set re-launches the graph kept on plans[0]), else new
    tensors."""
    torch = _torch()
    n = len(plans)
    if not (len(Vs) == n == len(psis)) or n == 0:
        raise ValueError("one V and one psi per plan")
    if len({id(p) for p in plans}) != n:
        raise ValueError("requests need distinct plans")
    if outs is not None and len(outs) != n:
        raise ValueError("one output per plan")
    inner, outer = Plan._angles(inner_mrad, outer_mrad)
    new_outs = []
    for i, (p, V, psi) in enumerate(zip(plans, Vs, psis)):
        S, B = V.shape[0], psi.shape[0]
        _need(V, (S, p.ny, p.nx), torch.float32, "V")
        _need(psi, (B, p.ny, p.nx), torch.complex64, "psi")
        nout = (S + stride - 1) // stride if S % stride == 0 else S // stride + 1
        if outs is None:
            new_outs.append(torch.zeros((nout, len(inner), B), dtype=torch.float32, device="cuda"))
        else:
            _need(outs[i], (nout, len(inner), B), torch.float32, "outs[%d]" % i)
            new_outs.append(outs[i])
    outs = new_outs
    arrP = (ctypes.c_void_p * n)(*[p.handle for p in plans])
    arrV = (ctypes.c_void_p * n)(*[V.data_ptr() for V in Vs])
    arrPsi = (ctypes.c_void_p * n)(*[q.data_ptr() for q in psis])
    arrO = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
    S = (ctypes.c_int * n)(*[V.shape[0] for V in Vs])
    B = (ctypes.c_int * n)(*[q.shape[0] for q in psis])
    stream = torch.cuda.current_stream().cuda_stream
    C.check(C.lib().em_multislice_ragged(n, arrP, arrV, S, arrPsi, B, len(inner), _dptr(inner), _dptr(outer),
                                         int(stride), arrO, ctypes.c_void_p(stream)), plans[0].handle,
            "em_multislice_ragged")
    return outs
