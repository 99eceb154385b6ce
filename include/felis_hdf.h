/* HDF5 dump and restore of a felis context in the reference's file layout.
 *
 * libfelis_hdf.so (host code, serial HDF5) on top of libfelis_hip.so: it
 * moves the state between the device and the files through the C-ABI of
 * felis.h. Files:
 *
 *   field<it>.h5     datasets "Ux<H>", "Uy<H>" per harmonic H: double
 *                    [sliceN][2*nx*ny] (re, im of cell j*nx + i, slices
 *                    1..sliceN of the reference's slots); int datasets
 *                    sliceN, harmony[numHarmony], numHarmony, nx, ny;
 *                    double datasets minZ, dz, bucketZ, powerNorm,
 *                    intensityNorm, dx, dy, minX, minY
 *   particle<it>.h5  dataset "<s>" per species s: double [totalCnt][8] rows
 *                    (theta, x, y, gamma, px, py, slice, weight), slices in
 *                    order, with long attributes cntP[1] and totalCnt[1];
 *                    double datasets minZ, dz, bucketZ, dPhi, lambda0,
 *                    gamma0[nSpecies]; int datasets numData, nSpecies, sliceN
 *
 * The caller names the file (the reference uses "field%d.h5" and
 * "particle%d.h5" with the iteration, in the working directory).
 * Return codes are felis_status values; felis_hdf_last_error() explains the
 * last failure of this library on the calling thread. */
#ifndef FELIS_HDF_H
#define FELIS_HDF_H

#include "felis.h"

#ifdef __cplusplus
extern "C" {
#endif

/* saveFieldHDF (saveFieldHDF.cpp:34-99, saveFieldComp 101-157, saveMeta
 * saveParticleHDF.cpp:175-219). spot_sig_r is Domain::spotSigR, used by the
 * power normalisation of the 1D path only (saveFieldHDF.cpp:67-73). */
int felis_save_field_hdf(felis_ctx *ctx, const char *file, double spot_sig_r);

/* restoreFieldHDF (restoreFieldHDF.cpp:46-107, restoreFieldComp 109-162):
 * nx, ny, dx, dy, minX, minY, sliceN and numHarmony must equal the
 * context's (FELIS_ERR_ARG otherwise, where the reference exits). Slices
 * 1..sliceN of Ux and Uy are replaced; the ghost slots keep their values. */
int felis_restore_field_hdf(felis_ctx *ctx, const char *file);

/* saveParticleHDF (saveParticleHDF.cpp:27-172, save_attr_HDF 221-268).
 * gamma0[n_species] is the reference's loadList[s].energy / mc2 + 1
 * (saveParticleHDF.cpp:145-148). */
int felis_save_particles_hdf(felis_ctx *ctx, const char *file, const double *gamma0);

/* restoreParticleHDF (restoreParticleHDF.cpp:80-257): nSpecies must equal
 * the context's; rows must be in slice order. Each slice's particles become
 * its segment, with the weight of its last row (the reference keeps the last
 * row's weight per slice). Unlike the reference, a slice with no rows is
 * restored empty and the slices after it keep their own rows (the
 * reference's range search, restoreParticleHDF.cpp:166-176, gives the next
 * slice a count that includes every earlier row). */
int felis_restore_particles_hdf(felis_ctx *ctx, const char *file);

const char *felis_hdf_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FELIS_HDF_H */
