"""Brute-force slicing in a chosen real type, for the slicing tests.

The bounds follow gene_bounds (slice_generator.hpp:130-166) one scalar
operation at a time in `real` (numpy float32 by default: the reference
instantiates Slice_Generator only for float, specimen.cuh:57), and membership
is decided for every atom by direct comparison, not by the binary search of
fd_by_z (:108-129): slice k holds the atoms with z_start <= z < z_end (fine
mode: within the eps-widened integration window), the last atom also when it
sits exactly on the top bound.  Run in float64 it gives what double arithmetic
would, which is how the tests show the float choice matters.
"""
import numpy as np

BORDER_MIN, BORDER_MAX, BORDER_MIDDLE = 0, 1, 2


def bounds(z, dz, border, eps, real=np.float32, fine=False, r_int=0.0):
    f = real
    z = np.asarray(z).astype(real)
    dz, eps, r = f(dz), f(eps), f(r_int)
    vmin = f(z[0] - r) if fine else z[0]
    vmax = f(z[-1] + r) if fine else z[-1]
    nv = int(np.floor(f(f(vmax - vmin) + eps) / dz)) + 2
    b = np.zeros(nv, dtype=real)
    if border == BORDER_MIN:
        zb = f(vmin - f(f(2) * eps))
        for i in range(nv):
            b[i] = f(zb + f(f(i) * dz))
    elif border == BORDER_MAX:
        zt = f(vmax + f(f(2) * eps))
        for i in range(nv):
            b[nv - 1 - i] = f(zt - f(f(i) * dz))
    else:
        zb = f(f(0.5) * f(f(vmin + vmax) - f(f(nv - 1) * dz)))
        for i in range(nv):
            b[i] = f(zb + f(f(i) * dz))
    return b


def brute_force_slicing(z, dz, border, eps, real=np.float32, fine=False, r_int=0.0):
    """(bounds as float64, i0, i1) with empty slices as (1, 0)."""
    f = real
    zr = np.asarray(z).astype(real)
    b = bounds(z, dz, border, eps, real, fine, r_int)
    S = len(b) - 1
    i0 = np.ones(S, dtype=np.int64)
    i1 = np.zeros(S, dtype=np.int64)
    for s in range(S):
        zs, ze = b[s], b[s + 1]
        if fine:
            zm = f(f(0.5) * f(b[s] + b[s + 1]))
            zs = f(min(f(zm - f(r_int)), b[s]) - f(eps))
            ze = f(max(f(zm + f(r_int)), b[s + 1]) + f(eps))
        if zs > ze:
            continue
        top = zr[-1] <= ze
        inside = (zr >= zs) & ((zr < ze) | top)
        idx = np.nonzero(inside)[0]
        if len(idx):
            i0[s], i1[s] = idx[0], idx[-1]
    return b.astype(np.float64), i0, i1


def planted_atom(z, dz, border, eps=1e-4):
    """A z (double) on which float and double slicing disagree: it lies on
    one side of some bound in double arithmetic and, rounded to float, on the
    other side of that bound's float value."""
    bf = bounds(z, dz, border, eps, np.float32).astype(np.float64)
    bd = bounds(z, dz, border, eps, np.float64)
    for k in range(1, len(bf) - 1):
        lo, hi = sorted((bf[k], bd[k]))
        for cand in np.linspace(lo, hi, 9)[1:-1]:
            in_f = np.float64(np.float32(cand)) >= bf[k]
            in_d = cand >= bd[k]
            if in_f != in_d:
                return cand
    raise AssertionError("no bound whose float and double values separate an atom")
