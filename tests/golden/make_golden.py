"""Writes tests/golden/golden.npz: reference values computed independently of
the oracle (scipy / numpy / numerical quadrature), used to pin it.

  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.constants as sc
import scipy.integrate as si
import scipy.special as ss

HERE = os.path.dirname(os.path.abspath(__file__))


def relativistic(e_kv):
    E = e_kv * 1e3 * sc.e                      # J
    mc2 = sc.m_e * sc.c ** 2
    p = np.sqrt(E * (E + 2 * mc2)) / sc.c      # momentum
    lam = sc.h / p * 1e10                      # Angstrom
    gam = 1 + E / mc2
    # sigma = 2 pi m e lambda / h^2 with the relativistic mass gamma m0 (rad / (V A))
    sig = 2 * np.pi * gam * sc.m_e * sc.e * (lam * 1e-10) / sc.h ** 2 * 1e-10
    return lam, gam, sig


def moliere_projected(Z, R):
    """int phi(sqrt(R^2+z^2)) dz of the Moliere potential, by quadrature (V*A)."""
    ek = sc.e ** 2 / (4 * np.pi * sc.epsilon_0) / sc.e * 1e10   # eV*A
    a0 = sc.physical_constants["Bohr radius"][0] * 1e10
    atf = 0.88534 * a0 * Z ** (-1.0 / 3.0)
    al = (0.35, 0.55, 0.10)
    be = (0.3, 1.2, 6.0)

    def phi(z):
        r = np.sqrt(R * R + z * z)
        return ek * Z / r * sum(a * np.exp(-b * r / atf) for a, b in zip(al, be))

    v, _ = si.quad(phi, 0, np.inf, epsabs=0, epsrel=1e-13, limit=500)
    return 2 * v


def main():
    out = {}
    kv = np.array([80.0, 200.0, 300.0])
    rel = np.array([relativistic(e) for e in kv])
    out["kv"] = kv
    out["lambda"], out["gamma"], out["sigma"] = rel[:, 0], rel[:, 1], rel[:, 2]
    xs = np.geomspace(1e-3, 40.0, 41)
    out["bessel_x"] = xs
    out["k0"] = ss.k0(xs)
    out["k1"] = ss.k1(xs)
    rng = np.random.default_rng(2024)
    for n in (24, 768):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"fft1_in_{n}"] = x
        out[f"fft1_out_{n}"] = np.fft.fft(x)
    x2 = rng.standard_normal((12, 20)) + 1j * rng.standard_normal((12, 20))
    out["fft2_in"] = x2
    out["fft2_out"] = np.fft.fft2(x2)
    Rs = np.array([0.02, 0.1, 0.3, 0.7, 1.2, 2.0, 2.9])
    out["vr_R"] = Rs
    for Z in (8, 22, 38):
        out[f"vr_{Z}"] = np.array([moliere_projected(Z, R) for R in Rs])
    np.savez(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"))


if __name__ == "__main__":
    main()
