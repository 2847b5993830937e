/*
 * charm_amd.h -- C ABI of libcharm_amd.so, an MI355X (gfx950) implementation
 * of CHarm's point spherical-harmonic synthesis and analysis in double
 * precision.
 *
 * The structures and the charm_* functions below are drop-in replacements for
 * the same names in CHarm 0.4.9 built in double precision without MPI: same
 * signatures, same struct layouts (x86-64 LP64), same error codes and the same
 * argument checks.  Each declaration cites the CHarm interface it replaces
 * (paths relative to the CHarm tree).  Functions prefixed charm_amd_ are new:
 * device-resident and batched entry points with no counterpart in CHarm.
 *
 * Floating point arrays passed to charm_* functions are host memory.  Arrays
 * passed to charm_amd_*_dev functions are device memory (hipMalloc'ed), and
 * the work is enqueued on the given HIP stream (0 = the null stream).
 */
#ifndef CHARM_AMD_H
#define CHARM_AMD_H

#include <limits.h>
#include <stddef.h>
#include <stdbool.h>

#ifdef __cplusplus
/* C's _Bool and C++'s bool have the same size and representation on the
 * platforms this library supports. */
#define CHARM_AMD_BOOL bool
extern "C" {
#else
#define CHARM_AMD_BOOL _Bool
#endif

/* ------------------------------------------------------------------------ */
/* Error handling: src/err/err.h.in:58-294                                  */
/* ------------------------------------------------------------------------ */
#define CHARM_ERR_MAX_FILE 4096U
#define CHARM_ERR_MAX_FUNC 256U
#define CHARM_ERR_MAX_MSG 4096U
#define CHARM_ERR_MAX_LEVEL 10U

/* src/err/err.h.in:89-124 */
enum {
    CHARM_FAILURE = -1,
    CHARM_SUCCESS,
    CHARM_EMEM,
    CHARM_EFUNCARG,
    CHARM_EFILEIO,
    CHARM_EFFTWINIT,
    CHARM_EMPIPROCESS
};

/* src/err/err.h.in:159-222 (non-MPI build: 56 bytes) */
typedef struct {
    size_t level;
    char **file;
    size_t *line;
    char **func;
    int code;
    char *msg;
    CHARM_AMD_BOOL saturated;
    CHARM_AMD_BOOL distributed;
} charm_err;

/* src/err/err.h.in:236 */
charm_err *charm_err_init(void);
/* src/err/err.h.in:250 */
void charm_err_free(charm_err *err);
/* src/err/err.h.in:262 */
void charm_err_reset(charm_err *err);
/* src/err/err.h.in:280 */
void charm_err_handler(charm_err *err, CHARM_AMD_BOOL terminate);
/* src/err/err.h.in:294 */
CHARM_AMD_BOOL charm_err_isempty(const charm_err *err);

/* ------------------------------------------------------------------------ */
/* Spherical harmonic coefficients: src/shc/shc.h.in:139-602                */
/* ------------------------------------------------------------------------ */
/* src/shc/shc.h.in:139-381 (non-MPI build: 64 bytes).  c[m] points at
 * C_{m,m}; the rows for m = 0..nmax are contiguous from c[0]
 * (src/shc/shc_init_chunk.c:139-150).  S_{n,0} are stored as zeros. */
typedef struct {
    unsigned long nmax;
    double mu;
    double r;
    size_t nc;
    size_t ns;
    double **c;
    double **s;
    CHARM_AMD_BOOL owner;
    CHARM_AMD_BOOL distributed;
} charm_shc;

/* src/shc/shc.h.in:422-426 */
charm_shc *charm_shc_malloc(unsigned long nmax, double mu, double r);
/* src/shc/shc.h.in:434-438 */
charm_shc *charm_shc_calloc(unsigned long nmax, double mu, double r);
/* src/shc/shc.h.in:545-550 */
charm_shc *charm_shc_init(unsigned long nmax, double mu, double r, double *c, double *s);
/* src/shc/shc.h.in:602 */
void charm_shc_free(charm_shc *shcs);

/* Coefficient utilities (host code).  The ranges are degrees nmin..nmax of
 * every order; mu and R must match where two structures meet (within
 * charm_glob_threshold); a bad argument sets CHARM_EFUNCARG in err. */
/* src/shc/shc.h.in:555-587: copy of degrees nmin..nmax into a new zeroed
 * structure of degree nmax_shcs_out; NULL on a bad range */
charm_shc *charm_shc_copy(const charm_shc *shcs, unsigned long nmin, unsigned long nmax,
                          unsigned long nmax_shcs_out);
/* src/shc/shc.h.in:1320-1363: rop = op1 + op2 on degrees nmin..nmax
 * (nmax <= rop->nmax), operands zero padded beyond their nmax; in place
 * allowed */
void charm_shc_add(charm_shc *rop, charm_shc *op1, charm_shc *op2, unsigned long nmin,
                   unsigned long nmax, charm_err *err);
/* src/shc/shc.h.in:1371-1380 */
void charm_shc_sub(charm_shc *rop, charm_shc *op1, charm_shc *op2, unsigned long nmin,
                   unsigned long nmax, charm_err *err);
/* src/shc/shc.h.in:1388-1401 */
void charm_shc_mul(charm_shc *rop, charm_shc *op1, charm_shc *op2, unsigned long nmin,
                   unsigned long nmax, charm_err *err);
/* src/shc/shc.h.in:1474-1491 (nmax <= op2->nmax; no zero divisor except
 * S_{n,0}) */
void charm_shc_div(charm_shc *rop, charm_shc *op1, charm_shc *op2, unsigned long nmin,
                   unsigned long nmax, charm_err *err);
/* src/shc/shc.h.in:1408-1434: degree n times a[n - nmin] */
void charm_shc_mul_degree_wise(charm_shc *shcs, const double *a, unsigned long nmin,
                               unsigned long nmax, charm_err *err);
/* src/shc/shc.h.in:1441-1466: order m times a[m] (nmax + 1 factors) */
void charm_shc_mul_order_wise(charm_shc *shcs, const double *a, unsigned long nmin,
                              unsigned long nmax, charm_err *err);
/* src/shc/shc.h.in:1499-1507 */
void charm_shc_div_degree_wise(charm_shc *shcs, const double *a, unsigned long nmin,
                               unsigned long nmax, charm_err *err);
/* src/shc/shc.h.in:1515-1523 */
void charm_shc_div_order_wise(charm_shc *shcs, const double *a, unsigned long nmin,
                              unsigned long nmax, charm_err *err);
/* src/shc/shc.h.in:1545-1571: dv[n] = sum_m C_nm^2 + S_nm^2, n = 0..nmax */
void charm_shc_dv(const charm_shc *shcs, unsigned long nmax, double *dv, charm_err *err);
/* src/shc/shc.h.in:1578-1603: sqrt of dv */
void charm_shc_da(const charm_shc *shcs, unsigned long nmax, double *da, charm_err *err);
/* src/shc/shc.h.in:1610-1639: degree variances of shcs1 - shcs2 */
void charm_shc_ddv(const charm_shc *shcs1, const charm_shc *shcs2, unsigned long nmax,
                   double *ddv, charm_err *err);
/* src/shc/shc.h.in:1646-1675 */
void charm_shc_dda(const charm_shc *shcs1, const charm_shc *shcs2, unsigned long nmax,
                   double *dda, charm_err *err);
/* src/shc/shc.h.in:1693-1722: to munew and rnew, degree n times
 * (mu / munew) (R / rnew)^n */
void charm_shc_rescale(charm_shc *shcs, double munew, double rnew, charm_err *err);

/* Coefficient files (src/shc/shc.h.in:60-108, 618-1330).  The readers return
 * the maximum degree stored in the file (CHARM_SHC_NMAX_ERROR on error); with
 * nmax = CHARM_SHC_NMAX_MODEL and shcs = NULL they only read that degree.
 * They set shcs->mu, shcs->r and every coefficient (zeros above nmax). */
#define CHARM_SHC_NMAX_MODEL (ULONG_MAX - 1UL)
#define CHARM_SHC_NMAX_ERROR (ULONG_MAX)
/* ordering of charm_shc_write_tbl/dov: degree (N) or order (M) fastest */
enum { CHARM_SHC_WRITE_N, CHARM_SHC_WRITE_M };
/* src/shc/shc.h.in:721-724 (binary: nmax, mu, r, C and S by order) */
unsigned long charm_shc_read_bin(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err);
/* src/shc/shc.h.in:789-793 (ICGEM gfc 1.0/2.0; epoch "yyyyMMdd[.hhmm]" for
 * time variable models) */
unsigned long charm_shc_read_gfc(const char *pathname, unsigned long nmax, const char *epoch,
                                 charm_shc *shcs, charm_err *err);
/* src/shc/shc.h.in:863-866 */
unsigned long charm_shc_read_tbl(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err);
/* src/shc/shc.h.in:939-942 */
unsigned long charm_shc_read_dov(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err);
/* src/shc/shc.h.in:1013-1016 */
unsigned long charm_shc_read_mtx(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err);
/* src/shc/shc.h.in:1055-1058 */
void charm_shc_write_bin(const charm_shc *shcs, unsigned long nmax, const char *pathname,
                         charm_err *err);
/* src/shc/shc.h.in:1140-1145 (formatting: a printf conversion for doubles) */
void charm_shc_write_tbl(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         int ordering, const char *pathname, charm_err *err);
/* src/shc/shc.h.in:1233-1238 */
void charm_shc_write_dov(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         int ordering, const char *pathname, charm_err *err);
/* src/shc/shc.h.in:1292-1296 */
void charm_shc_write_mtx(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         const char *pathname, charm_err *err);

/* ------------------------------------------------------------------------ */
/* Evaluation points: src/crd/crd.h.in:72-628                               */
/* ------------------------------------------------------------------------ */
/* src/crd/crd.h.in:72-102 */
enum {
    CHARM_CRD_CELL_GRID = -2,
    CHARM_CRD_CELL_SCATTERED,
    CHARM_CRD_POINT_SCATTERED = 1,
    CHARM_CRD_POINT_GRID,
    CHARM_CRD_POINT_GRID_GL,
    CHARM_CRD_POINT_GRID_DH1,
    CHARM_CRD_POINT_GRID_DH2
};

/* src/crd/crd.h.in:125-332 (non-MPI build: 72 bytes) */
typedef struct {
    int type;
    size_t nlat;
    size_t nlon;
    size_t npoint;
    double *lat;
    double *lon;
    double *r;
    double *w;
    CHARM_AMD_BOOL owner;
    CHARM_AMD_BOOL distributed;
} charm_point;

/* src/crd/crd.h.in:477-480 */
charm_point *charm_crd_point_malloc(int type, size_t nlat, size_t nlon);
/* src/crd/crd.h.in:489-492 */
charm_point *charm_crd_point_calloc(int type, size_t nlat, size_t nlon);
/* src/crd/crd.h.in:542-548 */
charm_point *charm_crd_point_init(int type, size_t nlat, size_t nlon, double *lat,
                                  double *lon, double *r);
/* src/crd/crd.h.in:564 */
void charm_crd_point_free(charm_point *pnt);
/* src/crd/crd.h.in:614-616 */
charm_point *charm_crd_point_gl(unsigned long nmax, double r);
/* src/crd/crd.h.in:626-628 */
void charm_crd_point_gl_shape(unsigned long nmax, size_t *nlat, size_t *nlon);
/* src/crd/crd.h.in:650-653 (non-equiangular Driscoll--Healy grid) */
charm_point *charm_crd_point_dh1(unsigned long nmax, double r);
/* src/crd/crd.h.in:662-664 */
void charm_crd_point_dh1_shape(unsigned long nmax, size_t *nlat, size_t *nlon);
/* src/crd/crd.h.in:674-677 (equiangular Driscoll--Healy grid) */
charm_point *charm_crd_point_dh2(unsigned long nmax, double r);
/* src/crd/crd.h.in:686-688 */
void charm_crd_point_dh2_shape(unsigned long nmax, size_t *nlat, size_t *nlon);

/* ------------------------------------------------------------------------ */
/* Global tunables: src/glob/glob.h.in:66-149, src/glob/glob.c:16-23        */
/* ------------------------------------------------------------------------ */
extern double charm_glob_threshold;
extern double charm_glob_threshold2;
extern unsigned long charm_glob_polar_optimization_a1;
extern double charm_glob_polar_optimization_a2;
/* The polar optimisation is tested jointly over blocks of consecutive rings
 * (points): simd * charm_glob_shs_block_lat_multiplier in synthesis,
 * simd * charm_glob_sha_block_lat_multiplier in analysis, with simd =
 * charm_amd_polar_simd_size (src/misc/misc_polar_optimization_apply.c:16-41,
 * src/shs/shs_point_grd.c:860-862, src/shs/shs_point_sctr.c:518-520,
 * src/sha/sha_point.c:716-718).  The multipliers are CHarm's globals
 * (src/glob/glob.h.in:231,241; defaults SIMD_BLOCK_A = 8, SIMD_BLOCK_S = 4,
 * src/simd/simd.h:303-304); the SIMD size names the CHarm build whose blocks
 * are reproduced: 8 for AVX-512 (the default), 4 for AVX/AVX2, 2 for
 * SSE4.1/NEON, 1 for builds without SIMD (whose multipliers are both 8,
 * src/simd/simd.h:609-611).  Read at every call. */
extern size_t charm_glob_sha_block_lat_multiplier;
extern size_t charm_glob_shs_block_lat_multiplier;
extern int charm_amd_polar_simd_size;

/* ------------------------------------------------------------------------ */
/* Synthesis and analysis: src/shs/shs.h.in:168-173, src/sha/sha.h.in:164-169 */
/* ------------------------------------------------------------------------ */
/* src/shs/shs.h.in:168-173.  Point synthesis of shcs up to degree nmax at
 * pnt (GL grids, custom grids and scattered points); f receives pnt->npoint
 * values (row-major [nlat][nlon] for grids). */
void charm_shs_point(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                     double *f, charm_err *err);
/* src/shs/shs.h.in:239-243.  First-order gradient in the local north-oriented
 * frame: f[0], f[1], f[2] each receive pnt->npoint values of
 * fx = (1/r) df/dlat, fy = -1/(r cos lat) df/dlon, fz = df/dr. */
void charm_shs_point_grad1(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                           double **f, charm_err *err);
/* src/shs/shs.h.in:344-348.  Second-order gradient (Reed 1973): f[0..5]
 * receive fxx, fxy, fxz, fyy, fyz, fzz. */
void charm_shs_point_grad2(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                           double **f, charm_err *err);
/* src/shs/shs.h.in:428-435.  1/(r^(dlat+dlon) cos^dlon lat) times the
 * derivative of order dr, dlat, dlon in r, lat, lon; dr + dlat + dlon <= 2. */
void charm_shs_point_guru(const charm_point *pnt, const charm_shc *shcs, unsigned long nmax,
                          unsigned dr, unsigned dlat, unsigned dlon, double *f, charm_err *err);
/* src/sha/sha.h.in:164-169.  Point analysis of f sampled on the quadrature
 * grid pnt up to degree nmax; overwrites all of shcs->c and shcs->s. */
void charm_sha_point(const charm_point *pnt, const double *f, unsigned long nmax,
                     charm_shc *shcs, charm_err *err);

/* ------------------------------------------------------------------------ */
/* Global spectral gravity forward modelling: src/gfm/gfm.h.in:222-320       */
/* ------------------------------------------------------------------------ */
/* src/gfm/gfm.h.in:223-239.  Potential coefficients (mu = G M, r = R) of the
 * masses between the sphere of radius shape_radius_ref and the shape
 * (shape_shcs, mu = r = 1), density sum_i rho_i(lat, lon) r^i (density_shcs[i],
 * mu = r = 1), topography powers shape_power_min..max; optionally exports the
 * coefficients of each power to <path>-p<p>-i<i>.{txt,shcs}. */
void charm_gfm_global_density_3d(const charm_shc *shape_shcs, unsigned long shape_nmax,
                                 double shape_radius_ref, charm_shc **density_shcs,
                                 unsigned long *density_nmax, unsigned density_order,
                                 double grav_const, double mass, unsigned shape_power_min,
                                 unsigned shape_power_max, unsigned long potential_shcs_nmax,
                                 const char *shape_density_shcs_path,
                                 const char *potential_shcs_path, const char *shcs_file_format,
                                 charm_shc *potential_shcs, charm_err *err);
/* src/gfm/gfm.h.in:255-270 (laterally varying density) */
void charm_gfm_global_density_lateral(const charm_shc *shape_shcs, unsigned long shape_nmax,
                                      double shape_radius_ref, charm_shc *density_shcs,
                                      unsigned long density_nmax, double grav_const, double mass,
                                      unsigned shape_power_min, unsigned shape_power_max,
                                      unsigned long potential_shcs_nmax,
                                      const char *shape_density_shcs_path,
                                      const char *potential_shcs_path,
                                      const char *shcs_file_format, charm_shc *potential_shcs,
                                      charm_err *err);
/* src/gfm/gfm.h.in:305-319 (constant density) */
void charm_gfm_global_density_const(const charm_shc *shape_shcs, unsigned long shape_nmax,
                                    double shape_radius_ref, double density, double grav_const,
                                    double mass, unsigned shape_power_min,
                                    unsigned shape_power_max, unsigned long potential_shcs_nmax,
                                    const char *shape_shcs_path, const char *potential_shcs_path,
                                    const char *shcs_file_format, charm_shc *potential_shcs,
                                    charm_err *err);

/* ------------------------------------------------------------------------ */
/* Device-resident entry points (no counterpart in CHarm)                    */
/* ------------------------------------------------------------------------ */
/* A Gauss-Legendre grid whose latitude tables live in device memory, for
 * repeated transforms without host round trips.  Created from a host
 * charm_point of type CHARM_CRD_POINT_GRID_GL; all radii must be equal. */
typedef struct charm_amd_grid charm_amd_grid;

charm_amd_grid *charm_amd_grid_create(const charm_point *pnt, charm_err *err);
void charm_amd_grid_free(charm_amd_grid *grd);

/* SHS on a device-resident GL grid.  cs_dev: device array of 2*P doubles,
 * P = (nmax+1)(nmax+2)/2, holding (C_nm, S_nm) pairs in the m-major order of
 * charm_shc truncated to degree nmax.  f_dev: device array of nlat*nlon
 * doubles.  mu, R as in charm_shc.  Work is enqueued on stream. */
void charm_amd_shs_grid_dev(const charm_amd_grid *grd, unsigned long nmax, double mu,
                            double R, const double *cs_dev, double *f_dev,
                            void *stream, charm_err *err);
/* SHA on a device-resident GL grid; writes cs_dev (2*P doubles, layout as
 * above) for a model with constants mu, R (coefficients are rescaled from the
 * grid radius to R). */
void charm_amd_sha_grid_dev(const charm_amd_grid *grd, unsigned long nmax, double mu,
                            double R, const double *f_dev, double *cs_dev,
                            void *stream, charm_err *err);
/* SHS at npoint scattered points given by device arrays lat, lon, r. */
void charm_amd_shs_sctr_dev(size_t npoint, const double *lat_dev, const double *lon_dev,
                            const double *r_dev, unsigned long nmax, double mu, double R,
                            const double *cs_dev, double *f_dev, void *stream,
                            charm_err *err);
/* No counterpart in CHarm.  charm_amd_shs_sctr_dev without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): every
 * buffer beyond the inputs and f_dev lives in a static device array of
 * 128 MiB per GPU, held until the process ends.  At most 1048576 points and
 * degree 2190 (larger requests fail with CHARM_EFUNCARG).  The first call on
 * a device builds the recurrence tables there and must not be under stream
 * capture.  Calls that overlap on one GPU share that array and race: the
 * caller serialises them.  Same results as charm_amd_shs_sctr_dev. */
void charm_amd_shs_sctr_dev_noalloc(size_t npoint, const double *lat_dev,
                                    const double *lon_dev, const double *r_dev,
                                    unsigned long nmax, double mu, double R,
                                    const double *cs_dev, double *f_dev, void *stream,
                                    charm_err *err);
/* Derivative requests of the device-resident entry points below:
 * CHARM_AMD_DERIV_GRAD1 (3 outputs, as charm_shs_point_grad1),
 * CHARM_AMD_DERIV_GRAD2 (6 outputs, as charm_shs_point_grad2), or
 * 100 dr + 10 dlat + dlon (1 output, as charm_shs_point_guru). */
#define CHARM_AMD_DERIV_GRAD1 (-1)
#define CHARM_AMD_DERIV_GRAD2 (-2)
/* charm_shs_point_grad1/grad2/guru (src/shs/shs.h.in:239-243, 344-348,
 * 428-435) on a device-resident GL grid; the outputs follow one another in
 * f_dev, nlat * nlon values each. */
void charm_amd_shs_grid_deriv_dev(const charm_amd_grid *grd, unsigned long nmax, double mu,
                                  double R, const double *cs_dev, int what, double *f_dev,
                                  void *stream, charm_err *err);
/* The same at npoint scattered points given by device arrays. */
void charm_amd_shs_sctr_deriv_dev(size_t npoint, const double *lat_dev, const double *lon_dev,
                                  const double *r_dev, unsigned long nmax, double mu, double R,
                                  const double *cs_dev, int what, double *f_dev, void *stream,
                                  charm_err *err);

/* Synthesis of nset coefficient sets on one device-resident GL grid (no
 * counterpart in CHarm, where each set is a separate charm_shs_point call,
 * src/shs/shs.h.in:168-173).  cs_dev: nset consecutive m-major triangles of
 * interleaved (C, S) pairs of degree nmax; f_dev: nset grids [set][row][lon].
 * The grid radius must equal R.  The sums over degrees run on the FP64
 * matrix cores. */
void charm_amd_shs_grid_multi_dev(const charm_amd_grid *grd, unsigned long nmax, size_t nset,
                                  double mu, double R, const double *cs_dev, double *f_dev,
                                  void *stream, charm_err *err);

/* Copy the (C, S) pairs of a device model of degree nmax_src, cut to degree
 * nmax <= nmax_src, into dst_dev (the layout the *_dev functions take). */
void charm_amd_cs_truncate(const double *src_dev, unsigned long nmax_src, double *dst_dev,
                           unsigned long nmax, void *stream, charm_err *err);

/* Batched execution: many independent requests in one persistent launch per
 * transform kind, optionally replayed as a HIP graph.  Synthesis requests
 * (kind CHARM_AMD_SHS_GRID) read cs and write f; analysis requests
 * (CHARM_AMD_SHA_GRID) read f and write cs; layouts as for
 * charm_amd_shs_grid_dev / charm_amd_sha_grid_dev.  Scattered synthesis
 * requests (CHARM_AMD_SHS_SCTR) read cs and write f[npoint] at the points
 * lat/lon/r, as charm_amd_shs_sctr_dev (grid unused); their coordinates are
 * read when the batch is created.  All pointers are device memory and must
 * stay valid for the life of the batch; the contents of cs and f may change
 * between runs. */
enum { CHARM_AMD_SHS_GRID = 0, CHARM_AMD_SHA_GRID = 1, CHARM_AMD_SHS_SCTR = 2 };

typedef struct {
    int kind;
    unsigned long nmax;
    double mu;
    double R;
    const charm_amd_grid *grid;
    double *cs;
    double *f;
    size_t npoint;            /* CHARM_AMD_SHS_SCTR only */
    const double *lat, *lon, *r;
} charm_amd_request;

typedef struct charm_amd_batch charm_amd_batch;

charm_amd_batch *charm_amd_batch_create(const charm_amd_request *reqs, size_t nreq, int use_graph,
                                        charm_err *err);
/* Enqueue one execution of every request of the batch on stream.  With
 * use_graph, the first run executes eagerly and records a HIP graph that
 * later runs replay. */
void charm_amd_batch_run(charm_amd_batch *b, void *stream, charm_err *err);
int charm_amd_batch_uses_graph(const charm_amd_batch *b);
void charm_amd_batch_free(charm_amd_batch *b);

/* Serving (no counterpart in CHarm): a worker thread admits requests under an
 * HBM byte budget and a limit of launches in flight -- in arrival order, a
 * request that does not fit yet letting at most 8 later ones pass -- and runs
 * each on its own stream, so that requests overlap on the device; small
 * scattered-point requests admitted together run as one persistent launch.
 * Models and Gauss-Legendre grids (with their seed tables) stay resident.
 * charm_amd_server_submit copies a request's input arrays (lat, lon, r,
 * f_in) before it returns and builds its grid if needed; the output array
 * must stay valid until charm_amd_server_wait, which copies the result into
 * it, returns. */
enum { CHARM_AMD_SRV_SHS_SCTR = 0, CHARM_AMD_SRV_SHS_GRID = 1, CHARM_AMD_SRV_SHA_GRID = 2 };

typedef struct {
    int kind;
    int model;                /* SHS kinds: id from charm_amd_server_add_model */
    unsigned long nmax;       /* degree of the transform */
    size_t npoint;            /* SHS_SCTR: points given by lat, lon, r */
    const double *lat, *lon, *r;
    unsigned long grid_nmax;  /* grid kinds: the GL grid of degree grid_nmax >= nmax */
    double r_grid;            /* grid kinds: radius of the grid */
    const double *f_in;       /* SHA_GRID: grid values, (grid_nmax+1) x 2(grid_nmax+1) */
    double mu, R;             /* SHA_GRID: constants of the output coefficients */
    double *out;              /* SHS: values (npoint or grid); SHA: (C, S) pairs */
} charm_amd_srv_request;

typedef struct charm_amd_server charm_amd_server;

charm_amd_server *charm_amd_server_create(size_t hbm_budget_bytes, int max_inflight, charm_err *err);
/* Copies shcs up to degree nmax into HBM; returns the model id (-1 on error). */
int charm_amd_server_add_model(charm_amd_server *s, const charm_shc *shcs, unsigned long nmax,
                               charm_err *err);
/* Queues a request; returns a ticket (> 0), or -1 on error. */
long charm_amd_server_submit(charm_amd_server *s, const charm_amd_srv_request *r, charm_err *err);
/* Waits for a ticket; *latency_ms = completion time minus submission time. */
int charm_amd_server_wait(charm_amd_server *s, long ticket, double *latency_ms, charm_err *err);
/* Cancels a ticket: a queued request is dropped, an in-flight one has its
 * result discarded; charm_amd_server_wait then reports CHARM_EFUNCARG
 * ("cancelled").  Returns 0, 1 if the request had already completed (its
 * result stays available), -1 for an unknown ticket. */
int charm_amd_server_cancel(charm_amd_server *s, long ticket, charm_err *err);
/* Fault injection from the host: the next `count` launches fail as a device
 * error would, with CHARM error code `code` (CHARM_EMEM for an allocation
 * failure, CHARM_EFUNCARG for a failed kernel); their requests report it in
 * charm_amd_server_wait and the server carries on. */
void charm_amd_server_inject_error(charm_amd_server *s, int count, int code);
typedef struct {
    unsigned long submitted, completed, failed, cancelled;
    unsigned long launches;   /* launch units started (a request or a group) */
    unsigned long grouped;    /* requests that ran in a group of several */
    unsigned long max_group;  /* largest group */
} charm_amd_srv_stats;
void charm_amd_server_get_stats(charm_amd_server *s, charm_amd_srv_stats *st);
/* HBM bytes currently accounted to models, grids and requests in flight. */
size_t charm_amd_server_in_use(charm_amd_server *s);
void charm_amd_server_free(charm_amd_server *s);

/* Library information: number of visible HIP devices (0 if none). */
int charm_amd_device_count(void);

/* Kernel timing: when enabled, HIP events on the launching stream bracket
 * each launch of the named kernels ("shs_grid_kernel", "fft_c2r",
 * "sha_kernel", "sha_reduce_kernel", "fft_r2c", "sctr_kernel"); on = 2
 * brackets only the recurrence kernels ("shs_grid_kernel", "sha_kernel",
 * "sctr_kernel"), each event pair adding a few microseconds between kernels.
 * charm_amd_prof_read waits for the recorded events and returns the summed
 * milliseconds and the launch count of one kernel. */
void charm_amd_prof_enable(int on);
void charm_amd_prof_reset(void);
int charm_amd_prof_read(const char *name, double *total_ms, long *count);
/* Ring pairs per lane of the last grid synthesis (kind 0) or analysis
 * (kind 1) launch in this process, as chosen by the ring-tile tuner (0 before
 * the first launch).  Diagnostic; no counterpart in CHarm. */
int charm_amd_last_ring_k(int kind);

#ifdef __cplusplus
}
#endif

#endif /* CHARM_AMD_H */
