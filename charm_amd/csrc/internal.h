// Internal declarations shared by the host code and the HIP kernels.
#ifndef CHARM_AMD_INTERNAL_H
#define CHARM_AMD_INTERNAL_H

#include <cstddef>
#include <cstdint>
#include <vector>

#include "charm_amd.h"

namespace charm_amd {

bool nearly_equal(double a, double b, double eps);
void err_set(charm_err *err, const char *file, size_t line, const char *func, int code,
             const char *msg);
void err_propagate(charm_err *err, const char *file, size_t line, const char *func);

#define CHARM_AMD_ERR(err, code, msg) \
    ::charm_amd::err_set((err), __FILE__, __LINE__, __func__, (code), (msg))

// Number of (n, m) pairs with 0 <= m <= n <= nmax.
inline size_t tri_len(unsigned long nmax)
{
    return ((size_t)(nmax + 1) * (size_t)(nmax + 2)) / 2;
}

// Offset of column m (entries n = m..nmax) in the m-major triangle of degree
// nmax (src/shc/shc_init_chunk.c:139-150).
inline size_t col_off(unsigned long nmax, unsigned long m)
{
    return (size_t)m * (nmax + 1) - ((size_t)m * (m - 1)) / 2;
}

// ---------------------------------------------------------------------------
// Kernel launch parameters
// ---------------------------------------------------------------------------
// Lumped-coefficient synthesis on rings of a grid.
struct ShsGridArgs {
    int nmax;          // degree of the synthesis
    int nring;         // rings evaluated directly ("north" half when folded)
    int nlat;          // rows of the output
    int mcount;        // columns written per row (>= nmax+1; extra columns = 0)
    int ntile;         // ring tiles
    int unit;          // all radii equal to R: no (R/r)^k factors
    const double *t;   // [nring] sin(lat)
    const double *sx;  // sectorial seeds, [(m-1)*sld + i]
    const int *six;
    size_t sld;
    const double *rp;  // unit == 0: (R/r)^k at [k*rld + row], k = 0..nmax+1; rld = 0: [k]
    const double *ratio;  // unit == 0: R/r per row
    size_t rld;
    const unsigned char *mirror;  // [nring] 1 if row nlat-1-i is the mirror ring
    const double *al;  // alpha_nm of the scaled recurrence, m-major, degree al_nt
    int al_nt;
    const void *cs;    // (C_nm, S_nm) times s_nm, m-major, degree nmax; see sn
    const double *sn;  // s_nm (layout of al): cs holds unscaled pairs, scaled as
                       // they are staged; nullptr: cs is already scaled
    void *out;         // [row * ldo + m] double2
    size_t ldo;
    double mur;        // mu / R
    int fft_layout;    // 1: out = Hermitian half-spectrum for c2r; 0: raw (a, b)
    int k;             // ring pairs per lane (1..3)
    const int *pm0;    // polar optimisation: [nring] first order skipped at the
                       // ring (its CHarm ring block); nullptr: off
    unsigned char *dead;  // [m * ntile + tile]: 1 if every term of the wave's
                          // column is exactly zero (see dead_mode)
    int dead_mode;     // 0: unused; 1: record the flags; 2: skip flagged waves
};

struct ShaArgs {
    int nmax;
    int nring;         // ring pairs ("north" rings, incl. the equator)
    int nlat;          // ring i mirrors to row nlat - 1 - i
    int nrow;          // rows of spec and cw
    int ntile;
    const double *t;   // [nring]
    const double *sx;
    const int *six;
    size_t sld;
    const unsigned char *mirror;
    const double *cw;  // [nlat] quadrature weights w; the rows are weighted by cws * w
    double cws;        // the global constant of the analysis (sha_point.c:313-322)
    const void *spec;  // r2c spectrum of f, [row * lds + k] double2
    size_t lds;
    const double *al;  // alpha_nm, m-major, degree al_nt
    int al_nt;
    void *part;        // [slot][tri_len(nmax)] double2 partial sums (before s_nm)
    int k;             // ring pairs per lane (2..4)
    int tpg;           // ring tiles per block, added in tile order into slot tile / tpg
    const int *pm0;    // as in ShsGridArgs (analysis ring blocks)
    unsigned char *dead;  // as in ShsGridArgs
    int dead_mode;
    // Fused finish (fc != nullptr, one partial slot): the block of each order
    // quad applies s_nm and scale[n] to its slot and writes the coefficients
    // of degree nmax (what sha_reduce_kernel does otherwise).
    const double *sn;     // s_nm, m-major, degree al_nt
    const double *scale;  // [nmax + 1] or nullptr
    double *fc, *fs;      // outputs at (dcol_off(nmax, m) + n - m) * fstride
    size_t fstride;
};

struct SctrArgs {
    int nmax;
    size_t npoint;
    int nchunk;
    const int *mchunk;  // [nchunk+1] order ranges
    int unit;
    double R;
    const double *lat_t;  // sin(lat) per point
    const double *lat_u;  // cos(lat) per point
    const double *lon;
    const double *r;
    const double *dm;     // [nmax+1]
    const double *al;     // alpha_nm, m-major, degree al_nt
    int al_nt;
    const void *cs;       // (C_nm, S_nm) times s_nm, m-major, degree nmax; see sn
    const double *sn;     // as in ShsGridArgs
    double *part;         // [chunk * npoint + i]
    int k;                // points per lane (1, 2, 4)
    const int *pm0;       // polar optimisation: first order skipped per point
                          // (its CHarm point block), in the order of lat_t; nullptr: off
    const int *unit_dev;  // non-null: the radii test was made on the device and
                          // unit is taken from *unit_dev (both variants are
                          // launched; the one that does not apply returns at once)
};

// Many coefficient sets on one grid (multi.hip): synthesis by FP64 MFMA.
struct MultiShsArgs {
    int nmax;
    int nring;            // ring pairs
    int nlat;
    int nset;
    int mcount;           // orders written per row (nfc; orders above nmax get zeros)
    int nrt;              // ring tiles of 64 ring pairs
    int nst;              // set tiles of 16 sets
    int nunit;            // (order group of 8, set tile) units
    const double *t;
    const double *sx;
    const int *six;
    size_t sld;
    const unsigned char *mirror;
    const double *al;     // alpha_nm, degree al_nt
    int al_nt;
    const void *cs;       // (C s, S s) pairs degree-major: entry (m, n) of the triangle holds
                          // nsetp consecutive sets (launch_scale_cs_multi's layout)
    int nsetp;            // sets padded to a multiple of 16 (the padding holds zeros)
    void *spec;           // [set][row][lds] double2
    size_t lds;
    double mur;
    const int *pm0;       // as in ShsGridArgs; nullptr: off
};
int launch_shs_multi(const MultiShsArgs &a, void *stream);
// (C_nm s_nm, S_nm s_nm) of nset models of degree nmax (set-major input, one
// m-major triangle per set) into the layout of MultiShsArgs::cs (nsetp
// sets per triangle entry, the padding zeroed)
int launch_scale_cs_multi(int nmax, size_t nset, const void *cs_dev, const double *sn_dev, int nt,
                          void *out_dev, int nsetp, void *stream);

// Derivatives of the synthesis (src/shs/shs.h.in:176-436).  Every derivative
// is a sum of ordinary syntheses F[T] over transformed coefficient tables T,
// each weighted per point by a monomial in t = sin(lat), 1/u = 1/cos(lat),
// 1/r and q = R/r.
//
// A table maps (C, S) of degree n, order m to degree n - shift:
//   (c0 + c1 n + c2 n^2) [e_nm if shift] (C_nm, S_nm)
// followed by lon = 1: (C, S) -> m (S, -C), lon = 2: (C, S) -> -m^2 (C, S);
// entries whose source degree exceeds nmax are zero.  e_nm is the factor of
// P_{n-1,m} in cos(lat) dP_nm/dlat = -n t P_nm + e_nm P_{n-1,m}
// (src/leg/leg_func_enm.c:40-41, src/leg/leg_func_xnum.h:35-37).
struct GradTable {
    double c0, c1, c2;
    int shift, lon;
};
// coef t^pt u^-pu r^-pr q^pq F[table tab]
struct GradTerm {
    double coef;
    int tab, pt, pu, pr, pq, pad;
};
constexpr int GRAD_MAXTERM = 24;
constexpr int GRAD_MAXOUT = 6;
struct GradCombine {
    int nout;
    int beg[GRAD_MAXOUT + 1];           // terms of output o: beg[o] .. beg[o+1]-1
    GradTerm term[GRAD_MAXTERM];
    double *out[GRAD_MAXOUT];           // [point] per output
};
int launch_grad_table(int nmax, const void *cs_dev, const GradTable &g, void *out_dev,
                      void *stream);
// On grids whose rows go through the FFT, the weights depend on the row only,
// so outputs are combined from a few basis syntheses at the spectrum level:
// term = coef t^pt u^-pu r^-pr q^pq (i m)^lon X_basis[row][m]
struct SpecTerm {
    double coef;
    int basis, lon, pt, pu, pr, pq;
};
constexpr int SPEC_MAXTERM = 64;
struct SpecCombine {
    int nout;
    int beg[GRAD_MAXOUT + 1];
    SpecTerm term[SPEC_MAXTERM];
    void *out[GRAD_MAXOUT];             // [row][m] double2 per output
};
// basis b of row r at basis + (b * nrow + r) * ld; outputs likewise
int launch_grad_spec_combine(size_t nrow, size_t ld, const double *t_dev, const double *u_dev,
                             const double *r_dev, int unit, double R, const void *basis_dev,
                             const SpecCombine &c, void *stream);
// Point p of nrow x ncol lies on row p / ncol; t, u, r per row.  With unit
// set every radius is taken as R.  Table k's values start at tabs + k * tld.
int launch_grad_combine(size_t nrow, size_t ncol, const double *t_dev, const double *u_dev,
                        const double *r_dev, int unit, double R, const double *tabs_dev,
                        size_t tld, const GradCombine &c, void *stream);

// Kernel launchers (kernels.hip).  All return hipError_t as int.
// alpha_nm and s_nm of degree nt; ab_tmp holds tri_len(nt) double2 of scratch
int launch_leg_tables(int nt, const double *r_dev, const double *ri_dev, void *ab_tmp,
                      double *al_dev, double *sn_dev, void *stream);
// out = (C_nm s_nm, S_nm s_nm) for interleaved pairs cs of degree nmax
int launch_scale_cs(int nmax, const void *cs_dev, const double *sn_dev, int nt, void *out_dev,
                    void *stream);
int launch_seeds(int nring, const double *u_dev, const double *dm_dev, int nmax, size_t sld,
                 double *sx_dev, int *six_dev, void *stream);
int launch_rpows(int nrow, const double *ratio_dev, int kmax, size_t rld, double *rp_dev,
                 void *stream);
int launch_shs_grid(const ShsGridArgs &a, void *stream);
int launch_lon_direct(int nlat, int nmax, size_t nlon, const double *lon_dev, const void *lc_dev,
                      size_t ldl, double mur, double *f_dev, void *stream);
int launch_sha(const ShaArgs &a, void *stream);
int launch_sha_reduce(int nmax, int ntile, const void *part_dev, const double *sn_dev, int nt,
                      int nmax_out, const double *scale_dev, double *c_dev, double *s_dev,
                      size_t stride, void *stream);
int launch_sctr(const SctrArgs &a, void *stream);
int launch_sctr_reduce(size_t npoint, int nchunk, const double *part_dev, const int *perm_dev,
                       double mur, double *f_dev, void *stream);
int launch_sort_points(size_t n, const double *lat_dev, const double *lon_dev, const double *r_dev,
                       int *perm_dev, double *t_dev, double *u_dev, double *lon_out_dev,
                       double *r_out_dev, void *stream);
// The same in a caller's workspace of ws_bytes >= sort_points_ws_bytes(n)
// (0: the size query failed); the permutation ends in perm_dev or in the
// workspace, and *perm_out names it.
size_t sort_points_ws_bytes(size_t n);
int launch_sort_points_ws(size_t n, const double *lat_dev, const double *lon_dev,
                          const double *r_dev, int *perm_dev, void *ws, size_t ws_bytes,
                          double *t_dev, double *u_dev, double *lon_out_dev, double *r_out_dev,
                          const int **perm_out, void *stream);
// *flag_dev = 1 if every x[i] is nearly equal to ref (misc_is_nearly_equal),
// else 0, decided on the device
int launch_all_near_dev(size_t n, const double *x_dev, double ref, double eps, int *flag_dev,
                        void *stream);
// p_dev[0..n) = v[0..n), n <= 72, from kernel arguments (no host copy)
struct IntVals {
    int n;
    int v[72];
};
int launch_set_ints(int *p_dev, const IntVals &v, void *stream);
int sctr_points_per_block(int k);
// Polar optimisation at scattered points: m0[i] = first order skipped at the
// i-th point in the order of perm (nullptr: input order), for CHarm's blocks
// of bs consecutive input points, threshold thr >= 0.
int launch_polar_points(size_t n, const double *lat_dev, const int *perm_dev, size_t bs, int nmax,
                        double thr, int *m0_dev, void *stream);
// Polar optimisation on grid rings given by cos(lat) (device), blocks of bs
// consecutive rings; (n + bs - 1) / bs doubles of workspace
int launch_polar_rings_ws(size_t n, const double *u_dev, size_t bs, int nmax, double thr,
                          double *umax_ws, int *m0_dev, void *stream);
// out[k] = q^k for k = 0..n by repeated products in order; out[i] = v for i < n
int launch_powers(int n, double q, double *out_dev, void *stream);
int launch_fill(size_t n, double v, double *out_dev, void *stream);
// The same with (n + bs - 1) / bs doubles of caller workspace
int launch_polar_points_ws(size_t n, const double *lat_dev, const int *perm_dev, size_t bs,
                           int nmax, double thr, double *umax_ws, int *m0_dev, void *stream);
int launch_pack_cs(const double *c_dev, const double *s_dev, size_t n, void *cs_dev,
                   void *stream);
int launch_sincos(size_t n, const double *lat_dev, double *t_dev, double *u_dev, void *stream);
int launch_zero_ints(int *p, int n, void *stream);
int launch_cs_truncate(const double *src_dev, unsigned long nmax_src, double *dst_dev,
                       unsigned long nmax, void *stream);
// Persistent batch kernels: kind 0 = grid synthesis, 1 = analysis,
// 2 = scattered synthesis.
int batch_ring_k(int kind);
int batch_grid_size(int kind);
int launch_shs_batch(const void *reqs_dev, const void *items_dev, int nitems, int *counter,
                     int grid, void *stream);
int launch_sha_batch(const void *reqs_dev, const void *items_dev, int nitems, int *counter,
                     int grid, void *stream);
// kind 2 = scattered synthesis; items (request, point block << 6 | chunk)
int launch_sctr_batch(const void *reqs_dev, const void *items_dev, int nitems, int *counter,
                      int grid, void *stream);
int launch_all_near(size_t n, const double *x_dev, double ref, double eps, int *result,
                    void *stream);

// Longitude transforms (fft.hip): rows of n reals <-> n/2+1 complex values
// (hipFFT conventions, unnormalised); n even and at most 8192.
bool row_fft_supported(size_t n);
// device scratch a transform of nrow rows needs (rows longer than one LDS
// transform are split in two passes through this buffer); 0 if none
size_t row_fft_scratch(size_t n, size_t nrow);
int launch_row_fft(int c2r, size_t n, size_t nrow, double *f, size_t ldf, void *spec, size_t lds,
                   void *scratch, void *stream);
int prepare_row_fft(size_t n, void *stream);
// Row transforms of many small grids in a few launches (one per transform
// mode): the rows of job j are those of one launch_row_fft call
struct RowJob {
    size_t nlon, nrow;
    double *f;
    size_t ldf;
    void *spec;
    size_t lds;
};
struct RowGroupPart {
    int mode = 0, njob = 0, threads = 64;
    unsigned nrow = 0;
    size_t lds = 0;
    const void *args = nullptr;        // device RowArgs[njob]
    const unsigned *row0 = nullptr;    // device first row of each job
};
// Per-request finishing steps of a batch, one launch for all requests
struct ScaleJob {
    int nmax;
    const void *cs;  // double2 coefficients, m-major triangle of degree nmax
    void *out;       // the same times s_nm
};
struct ShaReduceJob {
    int nmax, ntile;
    const void *part;     // [ntile][tri_len(nmax)] double2
    const double *scale;  // [nmax + 1] or nullptr
    double *cs;           // interleaved (C, S) triangle of degree nmax
};
struct SctrReduceJob {
    size_t npoint;
    int nchunk;
    const double *part;  // [nchunk][npoint]
    const int *perm;
    double mur;
    double *f;
};
int launch_scale_cs_group(const ScaleJob *jobs, int njob, int nmax_max, const double *sn, int nt,
                          void *stream);
int launch_sha_reduce_group(const ShaReduceJob *jobs, int njob, int nmax_max, const double *sn,
                            int nt, void *stream);
int launch_sctr_reduce_group(const SctrReduceJob *jobs, int njob, size_t npoint_max, void *stream);
bool row_group_supported(size_t nlon);
int row_group_build(const RowJob *jobs, int njob, void *stream, std::vector<RowGroupPart> *parts,
                    std::vector<void *> *owned);
int row_group_launch(int c2r, const std::vector<RowGroupPart> &parts, void *stream);

// Polar optimisation threshold for a transform of degree nmax from the
// globals: max(a1, a2 nmax), or -1 (off) when a2 < 0
// (src/misc/misc_polar_optimization_threshold.c:16-20).
double polar_threshold(unsigned long nmax);
// Rings (points) per polar-optimisation block: SIMD size times the block
// multiplier of synthesis (analysis = 0) or analysis (analysis = 1).
size_t polar_block(int analysis);
// First order the polar optimisation skips for a block with largest cos(lat)
// umax (nmax + 1: none), and the same per ring for consecutive blocks of
// `block` rings of u[0..nring).
int polar_first_skip(unsigned long nmax, double thr, double umax);
void polar_ring_m0(const double *u, size_t nring, unsigned long nmax, double thr, size_t block,
                   int *m0);

// Scattered synthesis plan: points per lane and the order chunks (chunk c
// holds orders mc[c] .. mc[c+1]-1), from the point count and the degree only.
// r[k] = sqrt(k), ri[k] = 1/sqrt(k) for k <= 2 nt + 3 and dm[n] for n <= nt,
// the host inputs of launch_leg_tables (src/leg/leg_func_r_ri.c:15-30,
// src/leg/leg_func_dm.c:14-23)
void leg_host_tables(int nt, std::vector<double> &r, std::vector<double> &ri,
                     std::vector<double> &dm);

constexpr int SCTR_MAX_CHUNKS = 64;
int sctr_points_per_lane(size_t npoint);
int sctr_order_chunks(size_t npoint, unsigned long nmax, int *mc);  // mc: nchunk + 1 entries

// Device bytes one scattered synthesis or one grid synthesis/analysis
// allocates internally (for the server's admission by footprint), the bytes
// the library keeps cached for reuse, and a trim of that cache.
size_t sctr_call_scratch(size_t npoint, unsigned long nmax);
size_t grid_call_scratch(int analysis, size_t nring, size_t nlat, size_t nlon, unsigned long nmax);
size_t scratch_cached_bytes();
void scratch_trim(size_t keep_bytes);

// Several scattered syntheses in one persistent launch (each as
// charm_amd_shs_sctr_dev computes it, so with the same sums); device pointers,
// enqueued on stream.  Returns a hipError_t.
struct SctrJob {
    size_t npoint;
    const double *lat, *lon, *r;
    unsigned long nmax;
    double mu, R;
    const double *cs;  // (C, S) pairs of degree nmax
    double *f;
};
int run_sctr_jobs(const SctrJob *jobs, int njob, void *stream);

// Run-time ring-tile tuning (CHARM_AMD_TUNE=1) on or off for the calling
// thread; the serving worker switches it off.
void set_thread_tuning(bool on);


}  // namespace charm_amd

#endif
