"""CPU tests of the host-side specimen layer: crystal builders, the seeded
frozen-phonon generator (splitmix64 known-answer vector, Gaussian moments,
C++ / Python agreement) and the scanner helpers."""
import os
import subprocess

import numpy as np
import pytest

import cudaem_amd.specimen as sp
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_splitmix64_known_answers():
    # reference outputs of splitmix64 from state 0 (Vigna's splitmix64.c)
    ref = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC,
           0x1B39896A51A8749B]
    assert [int(v) for v in sp.splitmix64(np.uint64(0), 5)] == ref


def test_gaussian_moments_and_determinism():
    g = sp.gaussians(7, 1, 200001)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01
    assert abs(np.mean(g ** 4) - 3) < 0.05
    assert np.array_equal(g, sp.gaussians(7, 1, 200001))
    assert not np.array_equal(g[:100], sp.gaussians(7, 2, 100))


def test_crystals():
    cu = sp.fcc_cu(8, 8, 10)
    assert len(cu["x"]) == 2560 and set(cu["Z"]) == {29} and cu["lx"] == pytest.approx(28.92)
    st = sp.srtio3(10, 10, 40)
    assert len(st["x"]) == 20000
    assert np.bincount(st["Z"])[[8, 22, 38]].tolist() == [12000, 4000, 4000]
    # SrO and TiO2 planes alternate every a/2
    zs = np.unique(np.round(st["z"], 6))
    assert len(zs) == 80 and np.allclose(np.diff(zs), 3.905 / 2)
    si = sp.si_diamond(2, 2, 2)
    assert len(si["x"]) == 64
    d = np.sqrt((si["x"][:, None] - si["x"][None]) ** 2 + (si["y"][:, None] - si["y"][None]) ** 2 +
                (si["z"][:, None] - si["z"][None]) ** 2)
    nn = np.min(d + np.eye(len(d)) * 99, axis=1)
    assert np.allclose(nn, 5.431 * np.sqrt(3) / 4)


def test_vibrate_rms():
    st = sp.srtio3(6, 6, 6)
    v = sp.vibrate(st, 0.08, 1234, 2)
    dx = np.concatenate([v[k] - st[k] for k in ("x", "y", "z")])
    assert abs(np.sqrt(np.mean(dx ** 2)) - 0.08) < 0.003
    v2 = sp.vibrate(st, 0.08, 1234, 2, dims=(True, True, False))
    assert np.array_equal(v2["z"], st["z"]) and np.array_equal(v2["x"], v["x"])


def test_cpp_vibrate_matches_python(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "test_simulators")
    if not os.path.exists(exe):
        pytest.skip("C++ test program not built")
    subprocess.run([exe, "vibrate", str(tmp_path)], check=True)
    r = np.fromfile(tmp_path / "vibrate.f32", dtype=np.float32).reshape(7, 3)
    at = {"x": np.arange(7) * 0.5, "y": np.arange(7) * 1.0, "z": np.arange(7) * 0.25}
    v = sp.vibrate(at, 0.08, 1234, 3)
    d = np.stack([v["x"] - at["x"], v["y"] - at["y"], v["z"] - at["z"]], 1)
    np.testing.assert_allclose(r, d, atol=1e-7)


def test_cpp_atom_text_and_raw_io(tmp_path):
    """read_atoms_text parses Z x y z [occ] lines with comments and blanks;
    write_raw_data / read_raw_data round-trip float32 arrays."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_simulators")
    if not os.path.exists(exe):
        pytest.skip("C++ test program not built")
    (tmp_path / "atoms.txt").write_text("# SrTiO3 fragment\n38 0 0 0\n\n22 1.9525 1.9525 1.9525 0.5  # half\n"
                                        "8 1.9525 1.9525 0\n")
    subprocess.run([exe, "io", str(tmp_path)], check=True)
    r = np.fromfile(tmp_path / "atoms.f32", dtype=np.float32).reshape(-1, 5)
    np.testing.assert_array_equal(r, np.array([[38, 0, 0, 0, 1], [22, 1.9525, 1.9525, 1.9525, 0.5],
                                               [8, 1.9525, 1.9525, 0, 1]], dtype=np.float32))
    (tmp_path / "atoms.txt").write_text("8 1.0 2.0\n")
    assert subprocess.run([exe, "io", str(tmp_path)], capture_output=True).returncode != 0


def test_cpp_parameter_names_and_accessors(tmp_path):
    """multislice_params carries the reference's host names
    (temporal_spatial_incoh, nrot, spec_rot_u0, spec_rot_center_p) and its
    accessors: output_lamda / output_gama / output_sigma equal the oracle's
    constants; scan intervals, counts, seed, wave type and batching."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_simulators")
    if not os.path.exists(exe):
        pytest.skip("C++ test program not built")
    subprocess.run([exe, "params", str(tmp_path)], check=True)
    r = np.fromfile(tmp_path / "params.f64", dtype=np.float64)
    lam, gam, sig = O.constants(300.0)
    np.testing.assert_allclose(r[:3], [lam, gam, sig], rtol=1e-12)
    np.testing.assert_allclose(r[3:], [8, 6, 0.5 * 40 / 8, 30 / 6, 42, 1, 1, 48])


def test_cpp_atom_cut_matches_python(tmp_path):
    """remove_atoms_outside_xy_range + shift_atoms of the C++ surface keep and
    move the same atoms as cut_window (closed box, 2 A margin)."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_simulators")
    if not os.path.exists(exe):
        pytest.skip("C++ test program not built")
    subprocess.run([exe, "cut", str(tmp_path)], check=True)
    r = np.fromfile(tmp_path / "cut.f64", dtype=np.float64).reshape(-1, 5)
    i = np.arange(40)
    at = {"Z": 6 + i % 3, "x": -3.0 + 0.5 * i, "y": 7.0 - 0.35 * i, "z": 0.1 * i, "occ": 0.5 + 0.01 * i}
    c = sp.cut_window(at, -1.0, 9.0, -2.0, 4.0)
    assert 0 < len(c["x"]) < 40
    np.testing.assert_array_equal(r[:, 0], c["Z"])
    np.testing.assert_allclose(r[:, 1:], np.stack([c["x"], c["y"], c["z"], c["occ"]], axis=1), rtol=0, atol=1e-12)


def test_raster_and_cut():
    x, y = sp.raster(4, 3, 8.0, 6.0)
    assert len(x) == 12 and x[1] == 2.0 and y[4] == 2.0
    at = {"x": np.array([-3.0, -1.0, 5.0, 13.0]), "y": np.zeros(4), "z": np.zeros(4), "Z": np.ones(4, int)}
    c = sp.cut_window(at, 0.0, 10.0, -1.0, 1.0)
    assert np.allclose(c["x"], [-1.0, 5.0])
