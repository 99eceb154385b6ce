"""Text format of felis_amd.diag (felField's power/crossP files), no device:
numbers as a C++ stream prints a double by default (6 significant digits,
%g rules) right-aligned in 12 columns, and felField's header lines."""
import types

import numpy as np

from felis_amd import diag


def test_number_format():
    assert diag._g(0.0) == "           0"
    assert diag._g(1234567.0) == " 1.23457e+06"
    assert diag._g(-1.5e-5) == "    -1.5e-05"
    assert diag._g(0.0001) == "      0.0001"
    assert diag._g(123456.0) == "      123456"
    assert len(diag._g(-1.23456789e-300)) == 13  # wider than the field: no truncation


def test_power_file(tmp_path):
    d = types.SimpleNamespace(n_harm=2, slice_n=3, num_slice=1, lambda0=1e-9, dz=0.026)
    p = np.arange(12, dtype=float).reshape(2, 2, 3)
    diag.write_power(tmp_path / "power7", 7, d, p)
    lines = (tmp_path / "power7").read_text().splitlines()
    assert lines[0] == "#z = 0.182[m]"
    assert lines[1] == "#s[m]         Px1[W]       Py1[W]       Px2[W]       Py2[W]       "
    # slice 1: s = 1e-9; Px1, Py1, Px2, Py2 = p[0,0,1], p[1,0,1], p[0,1,1], p[1,1,1]
    assert lines[3] == "       1e-09            1            7            4           10 "


def test_cross_file(tmp_path):
    d = types.SimpleNamespace(n_harm=1, nx=2, ny=2, dx=0.5, dy=0.25, min_x=-1.0, min_y=-2.0, dz=1.0)
    f = np.arange(8, dtype=float).reshape(2, 1, 4)
    diag.write_cross(tmp_path / "crossP0", 0, d, f)
    lines = (tmp_path / "crossP0").read_text().split("\n")
    assert lines[0] == "#z = 0[m]"
    assert lines[1] == "# x[m]       y[m]         Ix1[J/m^2]  Iy1[J/m^2]  "
    # i = 1, j = 0: x = -0.5, y = -2, value at flat index i*ny + j = 2
    assert lines[5] == "        -0.5           -2            2            6 "
    assert lines[4] == ""  # a blank line after each x block


def _desc2():
    return types.SimpleNamespace(n_harm=1, harmony=[1], nx=2, ny=2, dx=0.5, dy=0.25, min_x=-1.0, min_y=-2.0,
                                 dz=0.5, slice_n=2, num_slice=1, lambda0=1e-6, min_z=0.0)


def test_static_fields_file(tmp_path):
    f = np.arange(8, dtype=float).reshape(2, 1, 4) * 2.0
    diag.write_fields_txt(tmp_path / "Power3", _desc2(), f, 2.0)
    lines = (tmp_path / "Power3").read_text().split("\n")
    assert lines[0] == "# x            y            Px1            Py1"
    # i = 1, j = 0 is the field's flat cell nx*j + i = 1 (no transposition)
    assert lines[4] == "        -0.5           -2            1            5 "


def test_particle_files(tmp_path):
    d = _desc2()
    p = {k: np.arange(3, dtype=float) + o for o, k in enumerate(("x", "y", "px", "py", "theta", "gamma"))}
    so = np.array([0, 1, 3], np.int64)
    w = np.array([0.5, 0.25])
    diag.write_particles_txt(tmp_path / "0Particle0", d, p, so, w, 0)
    lines = (tmp_path / "0Particle0").read_text().splitlines()
    assert lines[0] == "# x            y            px           py           theta        gamma        weight"
    assert lines[2] == "           1            2            3            4            5            6         0.25"
    diag.write_felparticle(tmp_path / "f", 4, d, p, so, w, 0, skip=2)
    lines = (tmp_path / "f").read_text().splitlines()
    assert lines[0] == "# z = 2, "
    # rows 0 and 2 of the species; row 2 is in slice 1: s = (2 pi + theta) lambda0 / (2 pi)
    assert len(lines) == 3
    s2 = (2 * np.pi + 6.0) * 1e-6 / (2 * np.pi)
    assert lines[2] == "%.6g 2 3 7 4 5 0.25" % s2
