"""em_detector_sums_noalloc against em_detector_sums through the C ABI: equal
results (the same kernels, bit for bit) on two calls in a row, the second
with fewer probes, detectors and a smaller grid's plan; with accumulation;
and a clear error, before any device work, for sizes beyond the fixed
scratch."""
import ctypes

import numpy as np
import pytest

from helpers import probe_lens

pytestmark = pytest.mark.gpu


def _sums(C, fn, plan, psi, inner, outer, w=1.0, accumulate=0, out=None):
    import torch
    B = psi.shape[0]
    inner = np.ascontiguousarray(inner, dtype=np.float64)
    outer = np.ascontiguousarray(outer, dtype=np.float64)
    if out is None:
        out = torch.zeros((len(inner), B), dtype=torch.float32, device="cuda")
    st = getattr(C.lib(), fn)(plan.handle, ctypes.c_void_p(psi.data_ptr()), B, len(inner),
                              inner.ctypes.data_as(C._PD), outer.ctypes.data_as(C._PD), float(w), int(accumulate),
                              ctypes.c_void_p(out.data_ptr()))
    return st, out


def test_noalloc_detector_sums_match(cuda):
    import torch
    import cudaem_amd as em
    from cudaem_amd import _capi as C
    rng = np.random.default_rng(21)
    cases = [(512, 24, [0.0, 20.0, 60.0], [20.0, 60.0, 200.0]), (256, 5, [10.0], [80.0])]
    for n, B, inner, outer in cases:
        with em.Plan(n, n, 20.0, 20.0, 2.0, 300.0, max_batch=B) as p:
            psi = p.build_probes(probe_lens(25.0), rng.uniform(0, 20, B), rng.uniform(0, 20, B))
            p.fft2d(psi, -1)
            st0, ref = _sums(C, "em_detector_sums", p, psi, inner, outer, w=0.5)
            st1, got = _sums(C, "em_detector_sums_noalloc", p, psi, inner, outer, w=0.5)
            torch.cuda.synchronize()
            assert st0 == 0 and st1 == 0
            assert float(ref.abs().max()) > 0
            assert torch.equal(got, ref), (n, B)
            # accumulation into an existing output
            st2, acc = _sums(C, "em_detector_sums_noalloc", p, psi, inner, outer, w=0.25, accumulate=1,
                             out=ref.clone())
            st3, acc_ref = _sums(C, "em_detector_sums", p, psi, inner, outer, w=0.25, accumulate=1,
                                 out=ref.clone())
            torch.cuda.synchronize()
            assert st2 == 0 and st3 == 0
            assert torch.equal(acc, acc_ref)


def test_noalloc_rejects_sizes_beyond_the_scratch(cuda):
    import torch
    import cudaem_amd as em
    from cudaem_amd import _capi as C
    n = 1024
    with em.Plan(n, n, 20.0, 20.0, 2.0, 300.0, max_batch=1) as p:
        big = (1 << 18) // (8 * (n // 8)) + 1  # probes x 8 detectors x 128 tiles > 262144
        # a buffer of that many probes, so even a missed check reads in bounds
        psi = torch.zeros((big, n, n), dtype=torch.complex64, device="cuda")
        inner, outer = np.zeros(8), np.full(8, 50.0)
        out = torch.zeros((8, big), dtype=torch.float32, device="cuda")
        st = C.lib().em_detector_sums_noalloc(p.handle, ctypes.c_void_p(psi.data_ptr()), big, 8,
                                             inner.ctypes.data_as(C._PD), outer.ctypes.data_as(C._PD), 1.0, 0,
                                             ctypes.c_void_p(out.data_ptr()))
        assert st == C.EM_ERR_INVALID
        assert b"scratch" in C.lib().em_plan_last_error(p.handle)
