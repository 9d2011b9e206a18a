"""Specimens: crystal builders and frozen-phonon displacements (host side).

The reference takes atoms from its missing atom_coor_data.hpp and displaces
them with the missing random_lib.hpp (vibrate_atom_positions, called at
scan_simulator.cuh:41-50 and single_simulator.cuh:65-74).  The build fixes a
seeded Gaussian displacement generator: splitmix64 uniforms, Box-Muller pairs
consumed in order (x, y, z per atom), state seeded from (seed, configuration).
"""
import numpy as np

# Unit-cell bases in fractional coordinates: (Z, fx, fy, fz)
FCC = [(0, 0.0, 0.0, 0.0), (0, 0.5, 0.5, 0.0), (0, 0.5, 0.0, 0.5), (0, 0.0, 0.5, 0.5)]
PEROVSKITE_SRTIO3 = [(38, 0.0, 0.0, 0.0), (22, 0.5, 0.5, 0.5), (8, 0.5, 0.5, 0.0),
                     (8, 0.5, 0.0, 0.5), (8, 0.0, 0.5, 0.5)]
DIAMOND = FCC + [(0, x + 0.25, y + 0.25, z + 0.25) for (_, x, y, z) in FCC]


def crystal(basis, a, nx, ny, nz, Z=None, b=None, c=None):
    """Atoms of an orthorhombic crystal of nx*ny*nz cells (lattice a, b, c).

    Basis entries with Z == 0 take the element Z.  Returns dict x, y, z, Z, occ
    and the box lengths lx, ly, lz."""
    b = a if b is None else b
    c = a if c is None else c
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    xs, ys, zs, Zs = [], [], [], []
    for (zz, fx, fy, fz) in basis:
        xs.append((ix + fx) * a)
        ys.append((iy + fy) * b)
        zs.append((iz + fz) * c)
        Zs.append(np.full(ix.size, zz if zz else Z, dtype=np.int32))
    x, y, z, Zv = (np.concatenate(v) for v in (xs, ys, zs, Zs))
    return dict(x=x, y=y, z=z, Z=Zv, occ=np.ones_like(x), lx=nx * a, ly=ny * b, lz=nz * c)


def fcc_cu(nx=8, ny=8, nz=10, a=3.615):
    return crystal(FCC, a, nx, ny, nz, Z=29)


def srtio3(nx=10, ny=10, nz=40, a=3.905):
    return crystal(PEROVSKITE_SRTIO3, a, nx, ny, nz)


def si_diamond(nx, ny, nz, a=5.431):
    return crystal(DIAMOND, a, nx, ny, nz, Z=14)


_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(state0, n):
    """n outputs of splitmix64 from state0 (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(state0) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def gaussians(seed, iconf, n):
    """n standard normal values of configuration iconf (Box-Muller pairs)."""
    with np.errstate(over="ignore"):
        state0 = np.uint64(seed) * _GAMMA + np.uint64(iconf)
    npair = (n + 1) // 2
    u = (splitmix64(state0, 2 * npair) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.empty(2 * npair)
    g[0::2] = r * np.cos(2 * np.pi * u2)
    g[1::2] = r * np.sin(2 * np.pi * u2)
    return g[:n]


def vibrate(atoms, u_rms, seed, iconf, dims=(True, True, True)):
    """Frozen-phonon configuration iconf: Gaussian displacements of rms u_rms
    (Angstrom) along the enabled dimensions."""
    n = len(atoms["x"])
    g = gaussians(seed, iconf, 3 * n).reshape(n, 3)
    out = dict(atoms)
    for j, key in enumerate(("x", "y", "z")):
        out[key] = atoms[key] + (u_rms * g[:, j] if dims[j] else 0.0)
    return out


def cut_window(atoms, x0, x1, y0, y1, margin=2.0):
    """remove_atoms_outside_xy_range + shift_atoms as used by the scanners
    (scan_simulator.cuh:56-65, 148-157): keep atoms within the window plus a
    margin and shift the window origin to (0, 0)."""
    m = ((atoms["x"] >= x0 - margin) & (atoms["x"] <= x1 + margin) &
         (atoms["y"] >= y0 - margin) & (atoms["y"] <= y1 + margin))
    out = {k: (v[m] if isinstance(v, np.ndarray) and v.shape == atoms["x"].shape else v)
           for k, v in atoms.items()}
    out["x"] = out["x"] - x0
    out["y"] = out["y"] - y0
    return out


def raster(n_x, n_y, lx, ly, x0=0.0, y0=0.0):
    """Probe positions of an n_x x n_y raster covering [x0, x0+lx) x [y0, y0+ly),
    row-major (y outer), x fastest, as STEM_intensity[iDet][y][x]
    (scan_simulator.cuh:100-105)."""
    xs = x0 + np.arange(n_x) * (lx / n_x)
    ys = y0 + np.arange(n_y) * (ly / n_y)
    X, Y = np.meshgrid(xs, ys)
    return X.ravel(), Y.ravel()
