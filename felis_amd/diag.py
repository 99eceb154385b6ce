"""Text outputs of the reference, from the live state: the diag tools'
files (diag/felField.cpp, diag/felParticle.cpp) and the Static mode's
savers (saveFile.cpp saveFieldsToTxt, saveParticleToTxt.cpp).

felField reads a field dump and writes two text files per step:
  power<it>   the slice power of every harmonic and polarisation [W]
              against s = slice * bucketZ (felField.cpp:83-89, 170-179)
  crossP<it>  the transverse fluence [J/m^2] on the grid (felField.cpp:91-97,
              183-197)
Here the sums come from the device (felis_field_diag) and these functions
write the same files, number for number in the same format (C++ stream
defaults: 6 significant digits, right-aligned in 12 columns).

    p, f = ctx.field_diag(spot_sig_r)
    diag.write_power("power100", 100, desc, p)
    diag.write_cross("crossP100", 100, desc, f)

The fluence file labels row (i, j) with x = i dx + minX, y = j dy + minY and
the value at flat cell index i*ny + j, as felField does; the field's own
flat index is j*nx + i, so for a square grid the labels are those of the
transposed cell (felField.cpp:148-151, 185-192).
"""
from __future__ import annotations

import numpy as np


def _g(v) -> str:
    """a double as `std::ostream << std::setw(12) << v` prints it"""
    return "%12s" % ("%g" % float(v))


def _z(v) -> str:
    return "%g" % float(v)


def write_power(path, step: int, desc, power) -> None:
    """power[2][n_harm][slice_n] (W) in felField's power<it> format."""
    nh, sn = desc.n_harm, desc.slice_n
    power = np.asarray(power).reshape(2, nh, sn)
    bucket = desc.num_slice * desc.lambda0
    with open(path, "w") as f:
        f.write("#z = " + _z(step * desc.dz) + "[m]\n")
        f.write("#s[m]         ")
        for h in range(nh):
            f.write("Px" + str(h + 1) + "[W]" + "       " + "Py" + str(h + 1) + "[W]" + "       ")
        f.write("\n")
        for i in range(sn):
            f.write(_g(i * bucket) + " ")
            for h in range(nh):
                f.write(_g(power[0, h, i]) + " ")
                f.write(_g(power[1, h, i]) + " ")
            f.write("\n")


def write_cross(path, step: int, desc, fluence) -> None:
    """fluence[2][n_harm][nx*ny] (J/m^2) in felField's crossP<it> format."""
    nh, nx, ny = desc.n_harm, desc.nx, desc.ny
    fluence = np.asarray(fluence).reshape(2, nh, nx * ny)
    with open(path, "w") as f:
        f.write("#z = " + _z(step * desc.dz) + "[m]\n")
        f.write("# x[m]       y[m]         ")
        for h in range(nh):
            f.write("Ix" + str(h + 1) + "[J/m^2]" + "  " + "Iy" + str(h + 1) + "[J/m^2]" + "  ")
        f.write("\n")
        for i in range(nx):
            x = i * desc.dx + desc.min_x
            for j in range(ny):
                y = j * desc.dy + desc.min_y
                f.write(_g(x) + " " + _g(y) + " ")
                for h in range(nh):
                    f.write(_g(fluence[0, h, i * ny + j]) + " ")
                    f.write(_g(fluence[1, h, i * ny + j]) + " ")
                f.write("\n")
            f.write("\n")


def write_fields_txt(path, desc, fluence, intensity_norm) -> None:
    """saveFieldsToTxt (saveFile.cpp:8-60), the Static mode's Power<it>: per
    cell (x, y) the sum over slices of |U|^2 of every harmonic, unnormalised
    (fluence / intensity_norm), at the field's own flat index j*nx + i."""
    nh, nx, ny = desc.n_harm, desc.nx, desc.ny
    raw = np.asarray(fluence).reshape(2, nh, nx * ny) / intensity_norm
    with open(path, "w") as f:
        f.write("# x            y")
        for h in range(nh):
            H = desc.harmony[h]
            f.write("            Px" + str(H) + "            Py" + str(H))
        f.write("\n")
        for i in range(nx):
            x = i * desc.dx + desc.min_x
            for j in range(ny):
                y = j * desc.dy + desc.min_y
                f.write(_g(x) + " " + _g(y) + " ")
                for h in range(nh):
                    f.write(_g(raw[0, h, nx * j + i]) + " " + _g(raw[1, h, nx * j + i]) + " ")
                f.write("\n")
            f.write("\n")


def write_particles_txt(path, desc, particles, seg_off, seg_weight, species: int) -> None:
    """saveParticlesToTxt (saveParticleToTxt.cpp:6-45), the Static mode's
    <s>Particle<it>: every particle of one species, slice by slice, as
    x y px py theta gamma weight."""
    sn = desc.slice_n
    with open(path, "w") as f:
        f.write("# x            y            px           py")
        f.write("           theta        gamma        weight\n")
        for sl in range(sn):
            seg = species * sn + sl
            w = _g(seg_weight[seg])
            for n in range(int(seg_off[seg]), int(seg_off[seg + 1])):
                f.write(" ".join(_g(particles[k][n]) for k in ("x", "y", "px", "py", "theta", "gamma")) + " " + w + "\n")


def write_felparticle(path, step: int, desc, particles, seg_off, seg_weight, species: int, skip: int = 1) -> None:
    """felParticle (diag/felParticle.cpp:56-110) on the live particles: every
    skip-th particle of one species (counted over the species' rows in slice
    order) as s = (slice dPhi + theta) bucketZ / dPhi, x, y, gamma, px, py,
    weight."""
    sn = desc.slice_n
    d_phi = 2 * np.pi * desc.num_slice
    bucket = desc.lambda0 * desc.num_slice
    with open(path, "w") as f:
        f.write("# z = %g, \n" % (desc.min_z + step * desc.dz))
        idx = 0
        for sl in range(sn):
            seg = species * sn + sl
            for n in range(int(seg_off[seg]), int(seg_off[seg + 1])):
                if idx % skip == 0:
                    th = sl * d_phi + particles["theta"][n]
                    f.write("%.6g %g %g %g %g %g %g\n" % (th * bucket / d_phi, particles["x"][n], particles["y"][n],
                                                           particles["gamma"][n], particles["px"][n],
                                                           particles["py"][n], seg_weight[seg]))
                idx += 1
