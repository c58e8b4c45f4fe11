/*
 * pmpic_status.h -- status codes returned by every pmpic_* C-ABI entry point.
 *
 * The reference has no status codes: each failure is an `error stop` (or print + MPI_Finalize).
 * Each code below names the reference condition it replaces; the Fortran shim turns a non-zero
 * status into `error stop` with the text pmpic_status_message() returns.
 */
#ifndef PMPIC_STATUS_H
#define PMPIC_STATUS_H

enum pmpic_status {
  PMPIC_OK = 0,
  /* model_core/src/parcels/parcel_interpolation.F90:222 "parcel out of box (x direction)" */
  PMPIC_ERR_OUT_OF_BOX_X = 1,
  /* parcel_interpolation.F90:223 "parcel out of box (y direction)" */
  PMPIC_ERR_OUT_OF_BOX_Y = 2,
  /* parcel_interpolation.F90:224-227 "parcel out of box (z direction)" */
  PMPIC_ERR_OUT_OF_BOX_Z = 3,
  /* parcel_interpolation.F90:235-237 "delx/dely/delz wrong size" */
  PMPIC_ERR_DELTA = 4,
  /* components/parcel_splitting/src/parcel_splitting.F90:90-93 and
   * components/parcel_mixing/src/parcel_mixing.F90:196-199 "maxparcels reached" */
  PMPIC_ERR_MAXPARCELS = 5,
  /* model_core/src/fft/fft_ops.F90:642-645 "Error in tridiagonal solver: divide by zero" */
  PMPIC_ERR_ZERO_PIVOT = 6,
  /* model_core/src/fft/pencilfft.F90:114-117 "NX and/or NY are the wrong sizes for FFTE" */
  PMPIC_ERR_FFT_SIZE = 7,
  /* model_core/src/parcels/haloswap.F90:278-282 "parcel number not conserved" */
  PMPIC_ERR_COUNT = 8,
  /* parcel_mixing.F90:357-360: the backfill search runs below slot 1 (reads keep(0)) */
  PMPIC_ERR_BACKFILL = 9,
  /* invalid argument (sizes, ids, null pointers) */
  PMPIC_ERR_ARG = 10,
  /* HIP runtime failure (allocation, launch, copy) */
  PMPIC_ERR_HIP = 11,
  /* host allocation failure */
  PMPIC_ERR_NOMEM = 12,
  /* fft_ops.F90:528-550 laplinv called without exactly one boundary condition */
  PMPIC_ERR_BC = 13
};

#endif
