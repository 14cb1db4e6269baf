/*
 * parties_vof_io.h -- the reference's HDF5 output and restart files for the
 * 2-D VOF_DIFFUSE path, written from and read into a device subdomain
 * (libparties_vof_io.so, linked against HDF5 1.10).
 *
 * File layouts are the reference's (one rank: the whole domain):
 *  - Data file (IO/Output.c:277-774, Output_h5_data): /time; /grid/NX, NY,
 *    NZ (int), xu, xc (NX), yv, yc (NY), zw, zc (NZ); /u, /v, /w, /p; /VOF/F,
 *    C_L, C_S, C_G, psi, psi_LG, lap_C, bulk_S, ch_rhs_n, ch_rhs_nm1,
 *    normal_x, normal_y, normal_z, rho, mu, f_sigma_old_x/y/z.
 *  - Resume file (IO/Output.c:44-276, Output_h5_resume): /time, /dt,
 *    /dt_old, /ntime, /noutput, /noutput_2d, /output_time, /output_time_2d,
 *    /timer (34 doubles); /Resume/F, C_L, C_S, ch_rhs_n, ch_rhs_nm1, mu, rho,
 *    flux_x, flux_y, flux_z, normal_x, normal_y, normal_z, psi_LG,
 *    f_sigma_old_x/y/z.
 * Every field is a 3-D double dataset of (NZ, NY, NX) = (2, NYM+1, NXM+1)
 * values: the G range of Output_h5_flow_variable / _noghost_variable
 * (:1141-1260), both k slabs.  Fields the 2-D planar path does not carry
 * (w, C_S, normal_*, flux_z, f_sigma_old_z) are written as zeros; F is C_L,
 * which the reference's phase cache mirrors into F (VOF_DIFFUSE.c:108-126).
 *
 * Restart (IO/Resume.c:31-338 Resume_h5_resume + :341-690 Resume_h5_data):
 * the scalars, then u, v, p from the Data file with their boundaries and
 * ghost exchange, and the VOF fields of the Resume file with their ghost
 * exchange; rho and mu are recomputed when absent, C_G = clamp01(1 - C_L -
 * C_S) over the G range.  Errors return PV_EINVAL (missing dataset, wrong
 * dimensions) or PV_EHIP, with pv_io_last_error().
 */
#ifndef PARTIES_VOF_IO_H
#define PARTIES_VOF_IO_H

#include "parties_vof.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PV_IO_NTIMER 34      /* Include/DataTypes.h:993 N_TIMER_ELEMENTS */

/* Output bookkeeping kept in Parameters / Timer by the reference */
typedef struct pv_io_counters {
    int noutput, noutput_2d;
    double output_time, output_time_2d;
    double timer[PV_IO_NTIMER];
} pv_io_counters;

const char *pv_io_last_error(void);
/* Output_h5_data (IO/Output.c:277-774) */
int pv_Output_h5_data(pv_bag *bag, const char *filename);
/* Output_h5_resume (IO/Output.c:44-276); counters may be NULL (zeros) */
int pv_Output_h5_resume(pv_bag *bag, const char *filename, const pv_io_counters *counters);
/* Resume_h5_resume + Resume_h5_data (IO/Resume.c:31-690); counters may be NULL */
int pv_Resume_h5(pv_bag *bag, const char *resume_file, const char *data_file, pv_io_counters *counters);

#ifdef __cplusplus
}
#endif
#endif /* PARTIES_VOF_IO_H */
