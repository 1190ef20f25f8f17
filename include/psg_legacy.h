/*
 * Legacy drop-in entry points: the exact symbols and argument lists the
 * reference's Fortran drivers bind (gfortran mangling: lower case + "_",
 * every argument by address, no hidden string lengths).  Each flavour is a
 * separate shared library because the symbol names collide:
 *
 *   libpsg_legacy_ns3d_dns.so    timestep_cuda_ (36 args), set_gpu_
 *   libpsg_legacy_ns3d_gmres.so  timestep_cuda_ (18 args), set_gpu_
 *   libpsg_legacy_euler.so       timestep_cuda_ (28 args), set_gpu_
 *   libpsg_legacy_ns2d_gmres.so  timestep_cuda_ (19 args), set_gpu_
 *   libpsg_legacy_ns2d_dns.so    timestep_cuda_ (23 args), set_gpu_
 *   libpsg_legacy_bous_dns.so    timestep_cuda_ (37 args), set_gpu_
 *   libpsg_legacy_bous_gmres.so  timestep_cuda_ (37 args), timederiv_cuda_ (37 args), set_gpu_
 *
 * Semantics kept from the reference: state arrays are overwritten in place
 * with the final state, *Time is updated, the caller-built CN tables and
 * mask are used, DNS variants write their JOB<rank>/ side-effect files.
 * Changed on purpose: the GPU context (allocations, twiddles, pass buffers)
 * persists across calls with the same sizes instead of being rebuilt on
 * every call, and errors are reported on stderr and via psg_last_error()
 * instead of being discarded.  The reference returns nothing; so do these.
 *
 * Sizes the 18-argument GMRES variant does not receive (the reference
 * compiles them in, 3D_NavierStokes/src_gmres/kernels.h:1-12) are read from
 * PSG_NX / PSG_NY / PSG_NZ / PSG_ALPHA (default 128,128,128,1.0).
 */
#ifndef PSG_LEGACY_H
#define PSG_LEGACY_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } psg_double2;

/* 3D_NavierStokes/src/functions.h:3-20 (queries the device; exits if none) */
void set_gpu_(int *go);

/* 3D_NavierStokes/src/cuda.cu:18-54 */
void timestep_cuda_ns3d_dns(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, double *xir, double *etr,
                            double *ztr, double *nuzn, double *nu1, double *kx, double *ky, double *kz,
                            double *Time, double *TSTART, double *AMPFOR, double *DELT, double *ResThresh,
                            double *KFY, double *V2, double *alpha, int *ikF, int *ikN, int *LC, int *IKTX,
                            int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NSTOP,
                            int *NOUT, int *NOUTV, int *statsFLAG, int *RecFLAG, int *RANK);

/* 3D_NavierStokes/src_gmres/cuda.cu:10-28 */
void timestep_cuda_ns3d_gmres(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, double *nuzn, double *nu1,
                              double *kx, double *ky, double *kz, double *Time, double *TSTART, double *AMPFOR,
                              double *DELT, int *KFY, double *V2, int *ikF, int *ikN, int *LC, int *NSTOP);

/* 3D_Euler/src/cuda.cu:17-44 */
void timestep_cuda_euler(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, double *xir, double *etr,
                         double *ztr, double *kx, double *ky, double *kz, double *Time, double *TSTART,
                         double *DELT, double *alpha, int *ikF, int *ikN, int *LC, int *IKTX, int *IKTY,
                         int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NSTOP, int *NOUT,
                         int *NOUTV, int *statsFLAG);

/* 2D_NavierStokes/src_RK4/src_gmres/cuda.cu:16-34 */
void timestep_cuda_ns2d_gmres(psg_double2 *Z, double *ZR, double *kx, double *ky, double *Time, double *TSTART,
                              double *TNEWT, double *DELT, double *alpha, double *V2, int *ikF, int *ikN,
                              int *IKTX, int *IKTY, int *KTY, int *NX, int *NY, int *NSTOP, int *LC);

/* 23-argument variant called by 2D_NavierStokes/src_RK4/main.F90:219-221, whose
 * implementation is not in the reference; defined from the call site's types:
 * NIT RK4 steps of DELT, forcing NZKX -= 2 at (0,+-4), *TSTEP and *TSTART
 * untouched, stats.dat every NOUT steps when STFLAG=1. */
void timestep_cuda_ns2d_dns(psg_double2 *ZO, double *ZR, double *KX, double *KY, double *TSTEP, double *TSTART,
                            double *DELT, double *alpha, double *v2, double *SHIFTXin, double *ResThresh,
                            int *IKF, int *IKN, int *IKTX, int *IKTY, int *KTY, int *NX, int *NY, int *NIT,
                            int *NOUT, int *L, int *RCFLAG, int *STFLAG);

/* 3D_Boussinesq/src/cuda.cu:21-58: steps while tme <= Tstep; adaptive dt
 * (adaptFLAG) and JOB<rank>/{U,V,W,rho}NNN.raw dumps as out_low.h:5-110 */
void timestep_cuda_bous_dns(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, psg_double2 *rho, double *kx,
                            double *ky, double *kz, double *Time, double *TSTART, double *Tstep, double *AMPFOR,
                            double *DELT, double *ResThresh, double *KFY, double *V2, double *Ri, double *Sc,
                            double *Theta, double *alpha, double *Dtarget, int *ikF, int *ikN, int *LC, int *IKTX,
                            int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NOUT,
                            int *NOUTV, int *statsFLAG, int *RecFLAG, int *adaptFLAG, int *RANK);

/* 3D_Boussinesq/src/src_gmres/cuda.cu:12-49: steps of DELT while Tstep-tme > DELT,
 * then one short step to reach Tstep */
void timestep_cuda_bous_gmres(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, psg_double2 *rho, double *kx,
                              double *ky, double *kz, double *Time, double *TSTART, double *Tstep, double *AMPFOR,
                              double *DELT, double *ResThresh, double *KFY, double *V2, double *Ri, double *Sc,
                              double *Theta, double *alpha, double *Dtarget, int *ikF, int *ikN, int *LC,
                              int *IKTX, int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ,
                              int *NOUT, int *NOUTV, int *statsFLAG, int *RecFLAG, int *adaptFLAG, int *RANK);

/* 3D_Boussinesq/src/src_gmres/cuda.cu:319-356: dz/dt written over the 4 state fields */
void timederiv_cuda_(psg_double2 *xii, psg_double2 *eta, psg_double2 *zet, psg_double2 *rho, double *kx,
                     double *ky, double *kz, double *Time, double *TSTART, double *Tstep, double *AMPFOR,
                     double *DELT, double *ResThresh, double *KFY, double *V2, double *Ri, double *Sc,
                     double *Theta, double *alpha, double *Dtarget, int *ikF, int *ikN, int *LC, int *IKTX,
                     int *IKTY, int *IKTZ, int *KTZ, int *NKT, int *NX, int *NY, int *NZ, int *NOUT, int *NOUTV,
                     int *statsFLAG, int *RecFLAG, int *adaptFLAG, int *RANK);

/* In each libpsg_legacy_<flavour>.so the function above is exported as
 * timestep_cuda_ (an alias), which is what the Fortran call sites bind. */

#ifdef __cplusplus
}
#endif
#endif
