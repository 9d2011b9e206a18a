"""The C++ simulator surface (include/cudaEM/simulator.hpp) end to end: the
drop-in program tests/cpp/test_simulators drives mulSimulator /
TEM_simulator / scan_simulator / Thickness_simulator as reference user code
would; every output is checked against the FP64 oracle on the same inputs.
Tolerances: 1e-3 relative L2 on intensities / detector sums, 1e-4 on waves."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
import cudaem_amd.specimen as sp
from helpers import OracleSpecimen, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_simulators")
N, CELLS, NZ, A, E0 = 128, 3, 4, 3.905, 300.0
L = CELLS * A
DZ = A / 2
DET_IN, DET_OUT = [0.0, 40.0], [20.0, 120.0]
COND = {"outer_aper_mrad": 20.0}


def atoms_f32():
    """The specimen of the C++ program, with its float32 coordinates."""
    basis = [(38, 0, 0, 0), (22, .5, .5, .5), (8, .5, .5, 0), (8, .5, 0, .5), (8, 0, .5, .5)]
    xs, ys, zs, Zs = [], [], [], []
    for (Z, fx, fy, fz) in basis:
        for i in range(CELLS):
            for j in range(CELLS):
                for k in range(NZ):
                    xs.append(np.float32((i + fx) * A))
                    ys.append(np.float32((j + fy) * A))
                    zs.append(np.float32((k + fz) * A))
                    Zs.append(Z)
    f = lambda v: np.array(v, dtype=np.float32).astype(np.float64)
    x = f(xs)
    return {"x": x, "y": f(ys), "z": f(zs), "Z": np.array(Zs, np.int32), "occ": np.ones_like(x)}


@pytest.fixture(scope="module")
def outputs(cuda, tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_simulators is not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("cpp")
    r = subprocess.run([EXE, "run", str(d)], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    return lambda name: np.fromfile(d / name, dtype=np.float32)


def shift(a):
    return np.fft.fftshift(a)


def test_cpp_hrtem(outputs):
    at = atoms_f32()
    o = OracleSpecimen(at, N, N, L, L, DZ, E0, O.BORDER_MAX)
    Psi = O.multislice(o.propagator(), np.ones((N, N), complex), V=o.potentials())
    I, _ = O.ctf_image(Psi, L, L, o.lam, {"c_30": 1e7, "c_10": -500.0})
    assert rel_l2(outputs("hrtem.f32").reshape(N, N), I) < 1e-3


@pytest.mark.parametrize("model", ["pctf", "tcc"])
def test_cpp_hrtem_partial_coherence(outputs, model):
    """illumination_model eIM_Partial_Coherent / eIM_Trans_Cross_Coef with
    the objective lens's defocus and source spreads."""
    at = atoms_f32()
    o = OracleSpecimen(at, N, N, L, L, DZ, E0, O.BORDER_MAX)
    Psi = O.multislice(o.propagator(), np.ones((N, N), complex), V=o.potentials())
    lens = {"c_30": 1e7, "c_10": -500.0, "outer_aper_mrad": 30.0}
    sig, beta = float(np.float32(40.0)), float(np.float32(0.3))
    if model == "pctf":
        I = O.pctf_image(Psi, L, L, o.lam, lens, sig, beta)
    else:
        I = O.tcc_image(Psi, L, L, o.lam, lens, sig, 5, beta, 3)
    assert rel_l2(outputs(f"hrtem_{model}.f32").reshape(N, N), I) < 1e-3


def test_cpp_ed_full_wave(outputs):
    at = atoms_f32()
    o = OracleSpecimen(at, N, N, L, L, DZ, E0, O.BORDER_MIN)
    Psi = O.multislice(o.propagator(), np.ones((N, N), complex), V=o.potentials())
    got = outputs("ed_re.f32").reshape(N, N) + 1j * outputs("ed_im.f32").reshape(N, N)
    assert rel_l2(got, shift(Psi.real) + 1j * shift(Psi.imag)) < 1e-4


def test_cpp_ed_fine_slicing(outputs):
    """ED exit wave through the fine-sliced specimen, Sr given in the Lobato
    form through multislice_params.element_potentials."""
    at = atoms_f32()
    o = OracleSpecimen(at, N, N, L, L, DZ, E0, O.BORDER_MIN)
    o.use_fine_slicing({38: (O.PT_LOBATO_0_12, [1.0, 3.0, 8.0, 15.0, 4.0], [4.0, 6.0, 10.0, 20.0, 40.0])})
    a = o.atoms
    Vo = np.stack([O.potential_fine(N, N, L, L, o.bounds[s], o.bounds[s + 1], o.i0[s], o.i1[s], a["x"], a["y"],
                                    a["z"], a["occ"], o.tidx, [f[0] for f in o.fine], [f[1] for f in o.fine],
                                    [f[2] for f in o.fine], o.vr, nq=64) for s in range(o.S)])
    Psi = O.multislice(o.propagator(), np.ones((N, N), complex), V=Vo)
    got = outputs("ed_fine_re.f32").reshape(N, N) + 1j * outputs("ed_fine_im.f32").reshape(N, N)
    assert rel_l2(got, shift(Psi.real) + 1j * shift(Psi.imag)) < 1e-4


def test_cpp_cbed_frozen_phonons(outputs):
    at = atoms_f32()
    acc = np.zeros((N, N))
    for iconf in (1, 2):
        v = sp.vibrate(at, np.float32(0.08).item(), 77, iconf, dims=(True, True, False))
        v = {k: (np.asarray(val, dtype=np.float32).astype(np.float64) if k in ("x", "y", "z") else val)
             for k, val in v.items()}
        o = OracleSpecimen(v, N, N, L, L, DZ, E0, O.BORDER_MIN)
        psi = O.probe(N, N, L, L, o.lam, COND, 5.0, 6.0, o.zmid0())
        Psi = O.multislice(o.propagator(), psi, V=o.potentials())
        acc += 0.5 * abs(Psi) ** 2
    assert rel_l2(outputs("cbed.f32").reshape(N, N), shift(acc)) < 1e-3


def _scan():
    return sp.raster(4, 3, L, L)


def _blank_total(lam):
    psi = O.probe(N, N, L, L, lam, COND, 0.0, 0.0)
    return np.sum(abs(np.fft.fft2(psi)) ** 2)


def test_cpp_stem_tile(outputs):
    at = atoms_f32()
    o = OracleSpecimen(at, N, N, L, L, DZ, E0, O.BORDER_MIN)
    xs, ys = _scan()
    g2 = [(np.sin(np.array(a) * 1e-3) / o.lam) ** 2 for a in (DET_IN, DET_OUT)]
    ref, _ = O.stem(N, N, L, L, o.lam, COND, o.zmid0(), o.propagator(), o.potentials(), xs, ys, g2[0], g2[1])
    got = outputs("stem_tile.f32").reshape(2, 3 * 4)
    assert rel_l2(got, ref) < 1e-3
    assert abs(outputs("stem_tile_total.f32")[0] / _blank_total(o.lam) - 1) < 1e-4


def test_cpp_stem_general(outputs):
    at = atoms_f32()
    xs, ys = _scan()
    lam = O.constants(E0)[0]
    g2 = [(np.sin(np.array(a) * 1e-3) / lam) ** 2 for a in (DET_IN, DET_OUT)]
    ref = np.zeros((2, len(xs)))
    for ip, (xp, yp) in enumerate(zip(xs, ys)):
        c = sp.cut_window(at, xp - L / 2, xp + L / 2, yp - L / 2, yp + L / 2, margin=2.0)
        c = {k: (np.asarray(v, dtype=np.float32).astype(np.float64) if k in ("x", "y") else v) for k, v in c.items()}
        o = OracleSpecimen(c, N, N, L, L, DZ, E0, O.BORDER_MIN)
        psi = O.probe(N, N, L, L, o.lam, COND, 0.0, 0.0, o.zmid0())
        Psi = O.multislice(o.propagator(), psi, V=o.potentials())
        for d in range(2):
            ref[d, ip] = O.detector_sum(Psi, L, L, g2[0][d], g2[1][d])
    got = outputs("stem_general.f32").reshape(2, len(xs))
    assert rel_l2(got, ref) < 1e-3


def test_cpp_thickness_frozen_phonons(outputs):
    """Three frozen-phonon configurations (run as one batch of waves, each
    with its own potentials) against the oracle's weighted sum of series."""
    at = atoms_f32()
    acc = None
    for iconf in (1, 2, 3):
        v = sp.vibrate(at, np.float32(0.08).item(), 91, iconf, dims=(True, True, False))
        v = {k: (np.asarray(val, dtype=np.float32).astype(np.float64) if k in ("x", "y", "z") else val)
             for k, val in v.items()}
        o = OracleSpecimen(v, N, N, L, L, DZ, E0, O.BORDER_MIN)
        psi = O.probe(N, N, L, L, o.lam, COND, 2.0, 3.0, o.zmid0())
        g2min, g2max = (np.sin(40e-3) / o.lam) ** 2, (np.sin(120e-3) / o.lam) ** 2
        frames, _ = O.thickness(o.propagator(), o.potentials(), psi, L, L, g2min, g2max, stride=3)
        acc = frames / 3 if acc is None else acc + frames / 3
    got = outputs("thickness_fp.f32").reshape(-1, N, N)
    assert got.shape == acc.shape
    assert rel_l2(got, acc) < 1e-3


def test_cpp_thickness(outputs):
    at = atoms_f32()
    o = OracleSpecimen(at, N, N, L, L, DZ, E0, O.BORDER_MIN)
    psi = O.probe(N, N, L, L, o.lam, COND, 2.0, 3.0, o.zmid0())
    g2min, g2max = (np.sin(40e-3) / o.lam) ** 2, (np.sin(120e-3) / o.lam) ** 2
    frames, _ = O.thickness(o.propagator(), o.potentials(), psi, L, L, g2min, g2max, stride=2)
    got = outputs("thickness.f32").reshape(-1, N, N)
    assert got.shape == frames.shape
    assert rel_l2(got, frames) < 1e-3
