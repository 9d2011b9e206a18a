"""Shared set-up of the parity tests: the same physical inputs go to the HIP
engine (cudaem_amd) and to the FP64 oracle (oracle/pyoracle.py)."""
import numpy as np

import pyoracle as O


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def sort_atoms(atoms):
    order = np.argsort(atoms["z"], kind="stable")
    return {k: (v[order] if isinstance(v, np.ndarray) and v.shape == atoms["x"].shape else v)
            for k, v in atoms.items()}


class OracleSpecimen:
    """Slicing + sigma-scaled potentials of a specimen, FP64."""

    def __init__(self, atoms, nx, ny, lx, ly, dz, e0_kv, border, theta=0.0):
        self.lam, self.gam, self.sig = O.constants(e0_kv)
        self.vr = self.sig / np.cos(theta)
        a = sort_atoms(atoms)
        self.atoms = a
        self.nx, self.ny, self.lx, self.ly, self.dz = nx, ny, lx, ly, dz
        self._border = border
        self.bounds, self.i0, self.i1 = O.slicing(a["z"], dz, border)
        self.S = len(self.i0)
        self.types = sorted(set(int(z) for z in a["Z"]))
        self.tables = O.Tables(self.types)
        self.tidx = np.searchsorted(np.array(self.types), a["Z"]).astype(np.int32)

    def potential(self, s):
        a = self.atoms
        return O.potential(self.nx, self.ny, self.lx, self.ly, self.i0[s], self.i1[s], a["x"], a["y"],
                           a["occ"], self.tidx, self.tables, self.vr)

    def potentials(self):
        return np.stack([self.potential(s) for s in range(self.S)])

    def use_fine_slicing(self, params=None, border=None, r_int=None):
        """Switch to fine slicing: params maps Z -> (type, cl, cnl); the
        default is Moliere's 3D form."""
        a = self.atoms
        r_int = O.SPLINE_RMAX if r_int is None else r_int
        border = self._border if border is None else border
        self.bounds, self.i0, self.i1 = O.slicing_fine(a["z"], self.dz, border, r_int)
        self.S = len(self.i0)
        params = params or {}
        self.fine = []
        for Z in self.types:
            if Z in params:
                self.fine.append(params[Z])
            else:
                cl, cnl = O.moliere_3d(Z)
                self.fine.append((O.PT_KIRKLAND_0_12, cl, cnl))

    def potential_fine(self, s):
        a = self.atoms
        pts = [f[0] for f in self.fine]
        return O.potential_fine(self.nx, self.ny, self.lx, self.ly, self.bounds[s], self.bounds[s + 1],
                                self.i0[s], self.i1[s], a["x"], a["y"], a["z"], a["occ"], self.tidx, pts,
                                [f[1] for f in self.fine], [f[2] for f in self.fine], self.vr)

    def zmid0(self):
        return 0.5 * (self.bounds[0] + self.bounds[1])

    def propagator(self, bwl=True):
        return O.propagator(self.nx, self.ny, self.lx, self.ly, self.dz, bwl, self.lam)


def probe_lens(alpha_mrad, c10=0.0, c30=0.0):
    return {"c_10": c10, "c_30": c30, "outer_aper_mrad": alpha_mrad}
