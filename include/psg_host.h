/*
 * psgpu-mi355x host layer (plain C): sizes, wavenumber tables, masks,
 * Crank-Nicolson tables, index maps, random initial conditions, the
 * params.txt / paramsEXTRA.txt readers and the field-file formats.
 * These are the Fortran-side duties of the reference drivers, restated in C
 * so the drop-in drivers and the legacy entry points share them.
 */
#ifndef PSG_HOST_H
#define PSG_HOST_H
#include <stdint.h>
#include <stdio.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int dim;
  int nx, ny, nz;
  double alpha;
  int ktx, kty, ktz;
  int iktx, ikty, iktz;
  int nx2;
  long nkt, nr, n2;
} psgh_grid;

/* 3D: global_params.F90:13-26 (NY, NX=NZ/alpha given explicitly); 2D: NZ=1 */
int psgh_grid3(psgh_grid *g, int nx, int ny, int nz, double alpha);
int psgh_grid2(psgh_grid *g, int nx, int ny, double alpha);
/* kx,ky,kz (kz NULL in 2D) and mask L: 3D_NavierStokes/src/init.F90:25-64,
 * 2D_NavierStokes/src_RK4/init.F90:25-38 */
void psgh_tables(const psgh_grid *g, double *kx, double *ky, double *kz, int *L);
long psgh_count_active(const psgh_grid *g);
/* NU = 10**(log10 v2 + 2 log10|k|), NUZN = (1-NU dt/2)/dt, NU1 = dt/(1+NU dt/2), 0 off-mask
 * (3D_NavierStokes/src/init.F90:194-198) */
void psgh_nutab(const psgh_grid *g, double v2, double dt, double *nu, double *nuzn, double *nu1);
/* padded<->truncated index maps as 3D_NavierStokes/src/main.F90:100-199 */
void psgh_ikf_ikn(const psgh_grid *g, int *ikf, int *ikn);
/* e^{-2 pi i k/n}, k<n, interleaved re,im (FP64, glibc sincos) */
void psgh_twiddles(int n, double *w);

/* Numerical Recipes ran1 and the reference ranno (3D_NavierStokes/src/init.F90:237-294) */
void psgh_rng_seed(int iseed);
double psgh_ranno0(void);
/* reference random ICs: 3D_NavierStokes/src/init.F90:151-230, 2D_NavierStokes/src_RK4/init.F90:138-208 */
void psgh_ic_ns3d_ref(const psgh_grid *g, int iseed, double ampv, double *xii, double *eta, double *zet);
void psgh_ic_ns2d_ref(const psgh_grid *g, int iseed, double ampv, double *z);
/* 3D_Euler/src/init.F90:137-190 (Taylor-Green pair plus a random kz = 1 shell in zet) and
 * 3D_Boussinesq/src/init.F90:234-305 (solenoidal random shell, rho = 0) */
void psgh_ic_euler_ref(const psgh_grid *g, int iseed, double ampv, double *xii, double *eta, double *zet);
void psgh_ic_bous_ref(const psgh_grid *g, int iseed, double ampv, double *xii, double *eta, double *zet,
                      double *rho);
/* benchmark ICs (SplitMix64 + Box-Muller): solenoidal velocity with E(k) ~ k^4 exp(-2(k/4)^2),
 * sum_L|u|^2 = energy, w = i k x u; optional buoyancy with sum|rho|^2 = rho_scale^2 energy */
void psgh_ic_solenoidal(const psgh_grid *g, uint64_t seed, double energy, double rho_scale,
                        double *xii, double *eta, double *zet, double *rho);
void psgh_ic_ns2d_cfg(const psgh_grid *g, uint64_t seed, double ampv, double *z);

/* ---------------- parameter files ----------------
 * Order-fixed "LABEL = value" lines, each label matched character by character
 * and the value read with the Fortran edit descriptor of the reference reader
 * (read_label/read_int/read_float/read_logical/read_string,
 * 3D_NavierStokes/src/parameters.F90:47-124): I10, F12.4 (no decimal point =>
 * the last 4 digits are the fraction), L1, A256.  One label table per reference
 * reader; values are returned by label. */
enum {
  PSGH_PARAMS_NS3D_DNS = 0,   /* 3D_NavierStokes/src/parameters.F90:134-149 */
  PSGH_PARAMS_NS3D_GMRES = 1, /* 3D_NavierStokes/src_gmres/parameters.F90:145-165 */
  PSGH_PARAMS_EULER = 2,      /* 3D_Euler/src/parameters.F90:127-136 */
  PSGH_PARAMS_BOUS_DNS = 3,   /* 3D_Boussinesq/src/parameters.F90:120-140 */
  PSGH_PARAMS_BOUS_GMRES = 4, /* 3D_Boussinesq/src/src_gmres/parameters.F90:140-160 */
  PSGH_PARAMS_NS2D_DNS = 5,   /* 2D_NavierStokes/src_RK4/parameters.F90:145-171 */
  PSGH_PARAMS_NS2D_GMRES = 6, /* 2D_NavierStokes/src_RK4/src_gmres/parameters.F90:146-169 */
  PSGH_EXTRA_NS2D = 7,        /* 2D_NavierStokes/src_RK4/src_gmres/parameters.F90:182-203 */
  PSGH_EXTRA_NS3D = 8,        /* 3D_NavierStokes/src_gmres/parameters.F90:184-200 */
  PSGH_EXTRA_BOUS = 9         /* 3D_Boussinesq/src/src_gmres/parameters.F90:179-195 */
};
#define PSGH_MAXKEYS 48
typedef struct {
  int n;
  char key[PSGH_MAXKEYS][40];
  char kind[PSGH_MAXKEYS];            /* 'I','F','L','A' */
  long ival[PSGH_MAXKEYS];            /* I and L (1/0) */
  double fval[PSGH_MAXKEYS];          /* F */
  char sval[PSGH_MAXKEYS][257];       /* A (trailing blanks trimmed) */
} psgh_kv;
/* returns 0, or -1 with msg = "bad label <label>" / I/O error text */
int psgh_params_read(const char *path, int layout, psgh_kv *kv, char *msg, int msglen);
/* lookups by label without " = " (e.g. "Re", "infile type"); missing -> default */
long psgh_kv_int(const psgh_kv *kv, const char *key, long dflt);
double psgh_kv_float(const psgh_kv *kv, const char *key, double dflt);
const char *psgh_kv_str(const psgh_kv *kv, const char *key);
/* Fortran Fw.d input conversion of a w-character field (exposed for tests) */
double psgh_fortran_f(const char *field, int w, int d);

/* ---------------- field files ----------------
 * Fortran unformatted sequential records: 4-byte length marker, payload,
 * 4-byte length marker (gfortran layout; records < 2 GiB). */
int psgh_rec_write(FILE *f, const void *const *parts, const long *sizes, int nparts);
/* read the next record into buf (capacity cap); returns payload length or -1 */
long psgh_rec_read(FILE *f, void *buf, long cap);
/* Zk.out header record "1,IKTX,IKTY[,IKTZ]" (3D_NavierStokes/src/main.F90:53-54) */
int psgh_zk_header(FILE *f, int nrec, int iktx, int ikty, int iktz /* <=0 in 2D */);
/* state record: t, v2, T, sx, sy, [sz,] NER, fields (src_gmres/main.F90:645-646) */
int psgh_zk_record(FILE *f, const double *scalars, int nscalars, const double *fields, long ncomplex);
/* raw physical dump: NR doubles, x fastest (3D_NavierStokes/src/cuda.cu:72-93) */
int psgh_raw_write(const char *path, const double *r, long nr);

/* ---------------- restart records ("infile type") ----------------
 * The reference's restart readers take record IREST of `infile` in one of
 * several layouts chosen by the "infile type" letter:
 *   PSGH_REC_NS2D  2D_NavierStokes/src_RK4/src_gmres/init.F90:41-106 (types A-L;
 *                  Fortran sequential, header jw,IKTX,IKTY; B's first record
 *                  holds the time only)
 *   PSGH_REC_NS3D  3D_NavierStokes/src/init.F90:66-147 and src_gmres/init.F90:
 *                  75-130 (A, B: C stream; C: sequential, fields only;
 *                  D: sequential with header jw,IKTX,IKTY,IKTZ)
 *   PSGH_REC_BOUS  3D_Boussinesq/src/init.F90:46-182 (A, B stream; C fields
 *                  only; D, E, G, H with a header; F without; F-H carry Ri and
 *                  alpha, H ignores its alpha)
 *   PSGH_REC_EULER 3D_Euler/src/init.F90:73-118 (A stream with the time;
 *                  B fields only; C with a header and the time)
 * Values that a type does not carry are left as the caller set them (the
 * reference keeps the params.txt value, or the previous record's). */
enum { PSGH_REC_NS2D = 0, PSGH_REC_NS3D = 1, PSGH_REC_BOUS = 2, PSGH_REC_EULER = 3 };
typedef struct {
  double time, v2, ri, alpha, period, shiftx, shifty, shiftz, ner, ner2, tunder, rsym;
  int has_v2, has_ri, has_alpha;   /* set by psgh_restart_read when the record carried them */
} psgh_restart;
/* 1 if `type` is a layout of `family`, else 0 */
int psgh_restart_known(int family, const char *type);
/* read record irest (1-based) into state (nf*nkt complex, the grid's storage
 * order) and the scalars it carries into r; returns 0 or -1 with msg set */
int psgh_restart_read(const char *path, int family, const char *type, long irest, int iktx, int ikty, int iktz,
                      int nf, long nkt, double *state, psgh_restart *r, char *msg, int msglen);
/* the same one record after the other: open (checks the type and the header
 * against the grid; NULL with msg set on failure), next (0 or -1), close */
typedef struct psgh_restart_file psgh_restart_file;
psgh_restart_file *psgh_restart_open(const char *path, int family, const char *type, int iktx, int ikty, int iktz,
                                     int nf, long nkt, char *msg, int msglen);
int psgh_restart_next(psgh_restart_file *rf, double *state, psgh_restart *r, char *msg, int msglen);
void psgh_restart_close(psgh_restart_file *rf);

#ifdef __cplusplus
}
#endif
#endif
