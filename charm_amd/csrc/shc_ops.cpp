// Spherical harmonic coefficient utilities of the C ABI: copies, degree
// (difference) variances and amplitudes, rescaling to another mu and R, and
// coefficient-wise arithmetic.  Host code: each routine is one pass over the
// m-major triangle (O(nmax^2) flops, well under a millisecond at degree 2160
// next to the milliseconds of a synthesis), so nothing here goes to the GPU.
// Argument checks, zero padding and error codes follow CHarm 0.4.9
// (src/shc/shc.h.in:555-1730).
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "charm_amd.h"
#include "internal.h"

namespace charm_amd {
namespace {

// Calls f(m, n) for every order m = 0..nmax and degree n = max(m, nmin)..nmax,
// the degree range [nmin, nmax] of the triangle in storage order.
template <class F>
void for_range(unsigned long nmin, unsigned long nmax, F &&f)
{
    for (unsigned long m = 0; m <= nmax; m++)
        for (unsigned long n = std::max(m, nmin); n <= nmax; n++) f(m, n - m, n);
}

bool fail(charm_err *err, const char *msg)
{
    CHARM_AMD_ERR(err, CHARM_EFUNCARG, msg);
    return false;
}

// Checks of the binary operations (src/shc/shc_arithmetics_checks.c:17-107):
// rop, op1 and op2 share mu and R, and nmin <= nmax <= rop->nmax.
bool binary_ok(const charm_shc *rop, const charm_shc *op1, const charm_shc *op2,
               unsigned long nmin, unsigned long nmax, charm_err *err)
{
    if (!rop || !op1 || !op2) return fail(err, "\"rop\", \"op1\" and \"op2\" cannot be NULL.");
    const double t = charm_glob_threshold;
    if (!nearly_equal(rop->mu, op1->mu, t)) return fail(err, "\"rop->mu\" must be equal to \"op1->mu\".");
    if (!nearly_equal(rop->mu, op2->mu, t)) return fail(err, "\"rop->mu\" must be equal to \"op2->mu\".");
    if (!nearly_equal(rop->r, op1->r, t)) return fail(err, "\"rop->r\" must be equal to \"op1->r\".");
    if (!nearly_equal(rop->r, op2->r, t)) return fail(err, "\"rop->r\" must be equal to \"op2->r\".");
    if (nmin > nmax) return fail(err, "\"nmin\" cannot be larger than \"nmax\".");
    if (nmin > rop->nmax) return fail(err, "\"nmin\" cannot be larger than \"rop->nmax\".");
    if (nmax > rop->nmax) return fail(err, "\"nmax\" cannot be larger than \"rop->nmax\".");
    return true;
}

// Checks of the degree- and order-wise operations
// (src/shc/shc_arithmetics_wise_checks.c:18-61).
bool wise_ok(const charm_shc *shcs, const double *a, unsigned long nmin, unsigned long nmax,
             charm_err *err)
{
    if (!shcs || !a) return fail(err, "\"shcs\" and \"a\" cannot be NULL.");
    if (nmin > nmax) return fail(err, "\"nmin\" cannot be larger than \"nmax\".");
    if (nmin > shcs->nmax) return fail(err, "\"nmin\" cannot be larger than \"shcs->nmax\".");
    if (nmax > shcs->nmax) return fail(err, "\"nmax\" cannot be larger than \"shcs->nmax\".");
    return true;
}

enum class Bin { add, sub, mul };

// rop = op1 (+, -, *) op2 over degrees nmin..nmax; a degree beyond an
// operand's nmax reads that operand as zero (src/shc/shc_add.c:30-65,
// shc_sub.c, shc_mul.c:30-52).  Where one operand is padded, the other is
// copied (negated for op2 in a subtraction) rather than combined with 0.0,
// so that the sign of a zero coefficient is kept.  Element-wise, so rop may
// alias op1 and/or op2.
void binary(Bin op, charm_shc *rop, const charm_shc *op1, const charm_shc *op2,
            unsigned long nmin, unsigned long nmax)
{
    for_range(nmin, nmax, [&](unsigned long m, unsigned long k, unsigned long n) {
        const bool in1 = n <= op1->nmax, in2 = n <= op2->nmax;
        double c = 0.0, s = 0.0;
        if (op == Bin::mul) {
            if (in1 && in2) {
                c = op1->c[m][k] * op2->c[m][k];
                s = op1->s[m][k] * op2->s[m][k];
            }
        } else if (in1 && in2) {
            c = op == Bin::add ? op1->c[m][k] + op2->c[m][k] : op1->c[m][k] - op2->c[m][k];
            s = op == Bin::add ? op1->s[m][k] + op2->s[m][k] : op1->s[m][k] - op2->s[m][k];
        } else if (in1) {
            c = op1->c[m][k];
            s = op1->s[m][k];
        } else if (in2) {
            c = op == Bin::add ? op2->c[m][k] : -op2->c[m][k];
            s = op == Bin::add ? op2->s[m][k] : -op2->s[m][k];
        }
        rop->c[m][k] = c;
        rop->s[m][k] = s;
    });
}

// Degree (difference) variances (src/shc/shc_dv.c:38-68,
// shc_ddv.c:81-105): for each degree, the sum over m = 0..n of the squared
// C and S terms, accumulated in increasing m.
void variances(const charm_shc *a, const charm_shc *b, unsigned long nmax, double *out)
{
    for (unsigned long n = 0; n <= nmax; n++) {
        double acc = 0.0;
        for (unsigned long m = 0; m <= n; m++) {
            const double dc = b ? a->c[m][n - m] - b->c[m][n - m] : a->c[m][n - m];
            const double ds = b ? a->s[m][n - m] - b->s[m][n - m] : a->s[m][n - m];
            acc += dc * dc + ds * ds;
        }
        out[n] = acc;
    }
}

bool dv_ok(const charm_shc *shcs, unsigned long nmax, const double *out, charm_err *err)
{
    if (!shcs || !out) return fail(err, "\"shcs\" and the output array cannot be NULL.");
    if (nmax > shcs->nmax) return fail(err, "\"nmax\" cannot be larger than \"shcs->nmax\".");
    return true;
}

// src/shc/shc_ddv.c:37-78
bool ddv_ok(const charm_shc *s1, const charm_shc *s2, unsigned long nmax, const double *out,
            charm_err *err)
{
    if (!s1 || !s2 || !out)
        return fail(err, "\"shcs1\", \"shcs2\" and the output array cannot be NULL.");
    if (nmax > s1->nmax) return fail(err, "\"nmax\" cannot be larger than \"shcs1->nmax\".");
    if (nmax > s2->nmax) return fail(err, "\"nmax\" cannot be larger than \"shcs2->nmax\".");
    const double t = charm_glob_threshold;
    if (!nearly_equal(s1->r, s2->r, t))
        return fail(err, "To get meaningful difference degree variances/amplitudes, "
                         "\"shcs1->r\" must be equal to \"shcs2->r\".");
    if (!nearly_equal(s1->mu, s2->mu, t))
        return fail(err, "To get meaningful difference degree variances/amplitudes, "
                         "\"shcs1->mu\" must be equal to \"shcs2->mu\".");
    return true;
}

}  // namespace
}  // namespace charm_amd

using namespace charm_amd;

extern "C" {

// src/shc/shc_copy.c:16-57: NULL for an invalid range or a distributed input;
// the copy is a fresh calloc'ed structure of degree nmax_shcs_out.
charm_shc *charm_shc_copy(const charm_shc *shcs, unsigned long nmin, unsigned long nmax,
                          unsigned long nmax_shcs_out)
{
    if (!shcs || nmin > shcs->nmax || nmax > shcs->nmax || nmin > nmax || nmax > nmax_shcs_out ||
        shcs->distributed)
        return nullptr;
    charm_shc *out = charm_shc_calloc(nmax_shcs_out, shcs->mu, shcs->r);
    if (!out) return nullptr;
    for_range(nmin, nmax, [&](unsigned long m, unsigned long k, unsigned long) {
        out->c[m][k] = shcs->c[m][k];
        out->s[m][k] = shcs->s[m][k];
    });
    return out;
}

void charm_shc_add(charm_shc *rop, charm_shc *op1, charm_shc *op2, unsigned long nmin,
                   unsigned long nmax, charm_err *err)
{
    if (binary_ok(rop, op1, op2, nmin, nmax, err)) binary(Bin::add, rop, op1, op2, nmin, nmax);
}

void charm_shc_sub(charm_shc *rop, charm_shc *op1, charm_shc *op2, unsigned long nmin,
                   unsigned long nmax, charm_err *err)
{
    if (binary_ok(rop, op1, op2, nmin, nmax, err)) binary(Bin::sub, rop, op1, op2, nmin, nmax);
}

void charm_shc_mul(charm_shc *rop, charm_shc *op1, charm_shc *op2, unsigned long nmin,
                   unsigned long nmax, charm_err *err)
{
    if (binary_ok(rop, op1, op2, nmin, nmax, err)) binary(Bin::mul, rop, op1, op2, nmin, nmax);
}

// src/shc/shc_div.c:17-97.  op2 may not be padded (nmax <= op2->nmax) and no
// divisor in the range may be (nearly) zero, apart from S_{n,0}, which do not
// exist: for m = 0 only C is divided and rop's S_{n,0} are left as they are.
// Degrees beyond op1->nmax give zeros.  The whole range is checked before
// anything is written.
void charm_shc_div(charm_shc *rop, charm_shc *op1, charm_shc *op2, unsigned long nmin,
                   unsigned long nmax, charm_err *err)
{
    if (!binary_ok(rop, op1, op2, nmin, nmax, err)) return;
    if (nmax > op2->nmax) {
        fail(err, "\"nmax\" cannot be larger than \"op2->nmax\": zero padding \"op2\" "
                  "beyond its degree would divide by zero.");
        return;
    }
    const double t = charm_glob_threshold;
    bool zero = false;
    for_range(nmin, nmax, [&](unsigned long m, unsigned long k, unsigned long) {
        zero = zero || nearly_equal(op2->c[m][k], 0.0, t) ||
               (m > 0 && nearly_equal(op2->s[m][k], 0.0, t));
    });
    if (zero) {
        fail(err, "The coefficients of \"op2\" must not be zero in order to avoid division "
                  "by zero.  The only exceptions are the non-existing \"s\" coefficients of "
                  "order \"0\".");
        return;
    }
    for_range(nmin, nmax, [&](unsigned long m, unsigned long k, unsigned long n) {
        if (n <= op1->nmax) {
            rop->c[m][k] = op1->c[m][k] / op2->c[m][k];
            if (m > 0) rop->s[m][k] = op1->s[m][k] / op2->s[m][k];
        } else {
            rop->c[m][k] = 0.0;
            rop->s[m][k] = 0.0;
        }
    });
}

// src/shc/shc_mul_degree_wise.c:17-46: degree n scaled by a[n - nmin]
void charm_shc_mul_degree_wise(charm_shc *shcs, const double *a, unsigned long nmin,
                               unsigned long nmax, charm_err *err)
{
    if (!wise_ok(shcs, a, nmin, nmax, err)) return;
    for_range(nmin, nmax, [&](unsigned long m, unsigned long k, unsigned long n) {
        shcs->c[m][k] *= a[n - nmin];
        shcs->s[m][k] *= a[n - nmin];
    });
}

// src/shc/shc_mul_order_wise.c:17-49: order m scaled by a[m] (a holds
// nmax + 1 factors whatever nmin is)
void charm_shc_mul_order_wise(charm_shc *shcs, const double *a, unsigned long nmin,
                              unsigned long nmax, charm_err *err)
{
    if (!wise_ok(shcs, a, nmin, nmax, err)) return;
    for_range(nmin, nmax, [&](unsigned long m, unsigned long k, unsigned long) {
        shcs->c[m][k] *= a[m];
        shcs->s[m][k] *= a[m];
    });
}

// src/shc/shc_div_degree_wise.c:18-60: every one of the nmax - nmin + 1
// divisors must be nonzero (checked before anything is written)
void charm_shc_div_degree_wise(charm_shc *shcs, const double *a, unsigned long nmin,
                               unsigned long nmax, charm_err *err)
{
    if (!wise_ok(shcs, a, nmin, nmax, err)) return;
    for (unsigned long i = 0; i <= nmax - nmin; i++)
        if (nearly_equal(a[i], 0.0, charm_glob_threshold)) {
            fail(err, "All \"nmax - nmin + 1\" elements of \"a\" must be non-zero in order "
                      "to avoid division by zero.");
            return;
        }
    for_range(nmin, nmax, [&](unsigned long m, unsigned long k, unsigned long n) {
        shcs->c[m][k] /= a[n - nmin];
        shcs->s[m][k] /= a[n - nmin];
    });
}

// src/shc/shc_div_order_wise.c:18-63: all nmax + 1 divisors must be nonzero
void charm_shc_div_order_wise(charm_shc *shcs, const double *a, unsigned long nmin,
                              unsigned long nmax, charm_err *err)
{
    if (!wise_ok(shcs, a, nmin, nmax, err)) return;
    for (unsigned long i = 0; i <= nmax; i++)
        if (nearly_equal(a[i], 0.0, charm_glob_threshold)) {
            fail(err, "All \"nmax + 1\" elements of \"a\" must be non-zero in order to avoid "
                      "division by zero.");
            return;
        }
    for_range(nmin, nmax, [&](unsigned long m, unsigned long k, unsigned long) {
        shcs->c[m][k] /= a[m];
        shcs->s[m][k] /= a[m];
    });
}

void charm_shc_dv(const charm_shc *shcs, unsigned long nmax, double *dv, charm_err *err)
{
    if (dv_ok(shcs, nmax, dv, err)) variances(shcs, nullptr, nmax, dv);
}

// src/shc/shc_da.c: square roots of the degree variances
void charm_shc_da(const charm_shc *shcs, unsigned long nmax, double *da, charm_err *err)
{
    if (!dv_ok(shcs, nmax, da, err)) return;
    variances(shcs, nullptr, nmax, da);
    for (unsigned long n = 0; n <= nmax; n++) da[n] = std::sqrt(da[n]);
}

void charm_shc_ddv(const charm_shc *shcs1, const charm_shc *shcs2, unsigned long nmax,
                   double *ddv, charm_err *err)
{
    if (ddv_ok(shcs1, shcs2, nmax, ddv, err)) variances(shcs1, shcs2, nmax, ddv);
}

// src/shc/shc_dda.c:16-36
void charm_shc_dda(const charm_shc *shcs1, const charm_shc *shcs2, unsigned long nmax,
                   double *dda, charm_err *err)
{
    if (!ddv_ok(shcs1, shcs2, nmax, dda, err)) return;
    variances(shcs1, shcs2, nmax, dda);
    for (unsigned long n = 0; n <= nmax; n++) dda[n] = std::sqrt(dda[n]);
}

// src/shc/shc_rescale.c:19-88: degree n scaled by (mu / munew) (R / rnew)^n,
// the powers formed by repeated products in increasing n; then mu and r take
// the new values.
void charm_shc_rescale(charm_shc *shcs, double munew, double rnew, charm_err *err)
{
    if (!shcs) {
        fail(err, "\"shcs\" cannot be NULL.");
        return;
    }
    std::vector<double> f;
    try {
        f.resize((size_t)shcs->nmax + 1);
    } catch (const std::bad_alloc &) {
        CHARM_AMD_ERR(err, CHARM_EMEM, "Memory allocation failure.");
        return;
    }
    const double qmu = shcs->mu / munew, qr = shcs->r / rnew;
    f[0] = qmu;
    double p = qmu * qr;
    for (unsigned long n = 1; n <= shcs->nmax; n++) {
        f[n] = p;
        p *= qr;
    }
    for_range(0, shcs->nmax, [&](unsigned long m, unsigned long k, unsigned long n) {
        shcs->c[m][k] *= f[n];
        shcs->s[m][k] *= f[n];
    });
    shcs->mu = munew;
    shcs->r = rnew;
}

}  // extern "C"
