// Spherical harmonic coefficient files: CHarm's bin, gfc (ICGEM 1.0 and 2.0,
// static and time variable), tbl, dov and mtx formats
// (src/shc/shc.h.in:618-1330).
//
// A file is read in one piece; its data lines are split into ranges that
// worker threads convert to records (numbers parsed with strtod/strtoul, so
// values are the correctly rounded ones CHarm's misc_str2real gives), and the
// records are then applied in file order, which keeps CHarm's semantics
// (later lines overwrite earlier ones in tbl/dov, gfc records accumulate,
// the first bad line in file order is the one reported).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "charm_amd.h"
#include "internal.h"

namespace charm_amd {
namespace {

constexpr double PI_D = 3.14159265358979323846264338327950288;

struct Tok {
    const char *p = nullptr;
    size_t n = 0;
    std::string str() const { return std::string(p, n); }
    bool is(const char *s) const { return n == std::strlen(s) && std::memcmp(p, s, n) == 0; }
};

// up to maxtok whitespace-separated tokens of [b, e); returns the count
int tokenize(const char *b, const char *e, Tok *t, int maxtok)
{
    int k = 0;
    while (b < e && k < maxtok) {
        while (b < e && std::isspace((unsigned char)*b)) b++;
        if (b >= e) break;
        const char *s = b;
        while (b < e && !std::isspace((unsigned char)*b)) b++;
        t[k].p = s;
        t[k].n = (size_t)(b - s);
        k++;
    }
    return k;
}

// src/misc/misc_str2ul.c:18-30: decimal, fully consumed, no minus sign
bool str2ul(const Tok &t, unsigned long &v)
{
    if (t.n == 0 || t.n > 40) return false;
    char buf[48];
    std::memcpy(buf, t.p, t.n);
    buf[t.n] = '\0';
    if (std::strchr(buf, '-')) return false;
    char *end;
    errno = 0;
    v = std::strtoul(buf, &end, 10);
    return *end == '\0' && !errno && end != buf;
}

// src/misc/misc_str2real.c:20-58: Fortran exponents D/d read as e, trailing
// characters must be white space
bool str2real(const Tok &t, double &v)
{
    if (t.n == 0 || t.n > 126) return false;
    char buf[128];
    std::memcpy(buf, t.p, t.n);
    buf[t.n] = '\0';
    for (char *c = buf; *c; c++)
        if (*c == 'D' || *c == 'd') *c = 'e';
    char *end;
    errno = 0;
    v = std::strtod(buf, &end);
    if (end == buf || errno) return false;
    while (*end)
        if (!std::isspace((unsigned char)*end++)) return false;
    return true;
}

struct Lines {
    std::vector<const char *> b, e;
    size_t size() const { return b.size(); }
};

void split_lines(const char *p, const char *end, Lines &L)
{
    while (p < end) {
        const char *q = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        const char *qe = q ? q : end;
        L.b.push_back(p);
        L.e.push_back(qe);
        p = q ? q + 1 : end;
    }
}

bool read_whole(const char *path, const char *mode, std::string &out)
{
    FILE *f = std::fopen(path, mode);
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? std::fread(&out[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    out.resize(got);
    return true;
}

// Run fn(i0, i1) over [0, n) in contiguous ranges on up to 16 threads, at
// least `grain` items per thread
template <class F>
void parallel_ranges(size_t n, F fn, size_t grain = 20000)
{
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({hw, 16, std::max<size_t>(1, n / std::max<size_t>(grain, 1))});
    if (nt <= 1) {
        fn((size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (size_t k = 0; k < nt; k++) {
        const size_t i0 = n * k / nt, i1 = n * (k + 1) / nt;
        th.emplace_back([=] { fn(i0, i1); });
    }
    for (auto &t : th) t.join();
}

void set_err(charm_err *err, int code, const std::string &msg)
{
    CHARM_AMD_ERR(err, code, msg.c_str());
}

bool nmax_only(unsigned long nmax, const charm_shc *shcs)
{
    return nmax == CHARM_SHC_NMAX_MODEL && shcs == nullptr;
}

void reset_coeffs(charm_shc *shcs)
{
    std::fill(shcs->c[0], shcs->c[0] + shcs->nc, 0.0);
    std::fill(shcs->s[0], shcs->s[0] + shcs->ns, 0.0);
}

// First line of tbl, dov and mtx files: nmax mu r (src/shc/shc_read_mtdt.c)
bool read_mtdt(const Lines &L, const char *path, unsigned long &nmax_file, double &mu, double &r,
               charm_err *err)
{
    Tok t[3];
    if (L.size() == 0) {
        set_err(err, CHARM_EFILEIO,
                std::string("Failed to read the first line of \"") + path +
                    "\", where the maximum harmonic degree, the scaling constant and the radius "
                    "of the reference sphere are supposed to be given.");
        return false;
    }
    if (tokenize(L.b[0], L.e[0], t, 3) != 3) {
        set_err(err, CHARM_EFILEIO,
                std::string("Failed to read the metadata from the first line of \"") + path +
                    "\".");
        return false;
    }
    if (!str2ul(t[0], nmax_file)) {
        set_err(err, CHARM_EFILEIO,
                "Failed to convert the maximum harmonic degree to the \"unsigned long int\" data "
                "format.");
        return false;
    }
    if (!str2real(t[1], mu)) {
        set_err(err, CHARM_EFILEIO,
                "Failed to convert the scaling parameter to the \"REAL\" data format.");
        return false;
    }
    if (!str2real(t[2], r)) {
        set_err(err, CHARM_EFILEIO,
                "Failed to convert the radius of the reference sphere to the \"REAL\" data "
                "format.");
        return false;
    }
    if (r <= 0.0) {
        set_err(err, CHARM_EFUNCARG, "Spherical radius must be positive.");
        return false;
    }
    return true;
}

// Common checks after the metadata (src/shc/shc_read_tbl.c:96-130)
bool check_degrees(const char *path, unsigned long nmax, unsigned long nmax_file,
                   const charm_shc *shcs, charm_err *err)
{
    if (shcs->nmax < nmax) {
        char m[CHARM_ERR_MAX_MSG];
        std::snprintf(m, sizeof m,
                      "\"shcs->nmax = %lu\" cannot be smaller than the input parameter \"nmax = "
                      "%lu\".",
                      shcs->nmax, nmax);
        set_err(err, CHARM_EFUNCARG, m);
        return false;
    }
    if (nmax_file < nmax) {
        char m[CHARM_ERR_MAX_MSG];
        std::snprintf(m, sizeof m,
                      "Couldn't read coefficients up to degree \"nmax = %lu\", because the "
                      "maximum degree in \"%s\" is \"%lu\" only.",
                      nmax, path, nmax_file);
        set_err(err, CHARM_EFUNCARG, m);
        return false;
    }
    return true;
}

// One data line of a tbl or dov file
struct Rec {
    unsigned long n = 0, m = 0;
    double c = 0.0, s = 0.0;
    int status = 0;  // 0 ok, 1 skip (degree above nmax), <0 error kind
    bool sine = false;
};

// ---------------------------------------------------------------------------
// Epochs of time variable gfc models: yyyyMMdd or yyyyMMdd.hhmm as a fraction
// of a year, year + (doy + (h + min/60)/24) / ndays
// (src/shc/shc_read_gfc.c:163-410)
// ---------------------------------------------------------------------------
bool epoch_fraction(const char *date, double &t, std::string &why)
{
    while (std::isspace((unsigned char)*date)) date++;
    if (date[0] == '-') {
        why = std::string("Negative epoch \"") + date + "\".";
        return false;
    }
    const size_t nd = std::strlen(date);
    const std::string fmt = std::string("Wrong format of the epoch string \"") + date +
                            "\".  Supported formats are \"yyyyMMdd\" and \"yyyyMMdd.hhmm\".";
    if (nd < 8) {
        why = fmt;
        return false;
    }
    static const unsigned long monthdays[13] = {0, 31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    static const double before[13] = {0, 0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334};
    auto field = [&](size_t off, size_t len, unsigned long &v) {
        Tok t;
        t.p = date + off;
        t.n = len;
        return str2ul(t, v);
    };
    unsigned long year, month, day, hour = 0, min = 0;
    if (!field(0, 4, year)) {
        why = std::string("Failed to convert the year of the epoch string \"") + date + "\".";
        return false;
    }
    const bool leap = !(year % 400) || (!(year % 4) && (year % 100));
    if (!field(4, 2, month) || month < 1 || month > 12) {
        why = std::string("Invalid month in the epoch string \"") + date + "\".";
        return false;
    }
    if (!field(6, 2, day) || day < 1 || day > monthdays[month] + ((leap && month == 2) ? 1 : 0)) {
        why = std::string("Invalid day in the epoch string \"") + date + "\".";
        return false;
    }
    if (nd > 8) {
        if (nd != 13 || date[8] != '.') {
            why = fmt;
            return false;
        }
        if (!field(9, 2, hour) || hour > 24) {
            why = std::string("Invalid hours in the epoch string \"") + date + "\".";
            return false;
        }
        if (!field(11, 2, min) || min > 60) {
            why = std::string("Invalid minutes in the epoch string \"") + date + "\".";
            return false;
        }
    }
    if (hour == 24 && min == 60) {
        why = std::string("Invalid combination of hours and minutes in the epoch string \"") +
              date + "\".";
        return false;
    }
    double doy = (double)day + before[month] - 1.0;
    if (leap && month > 2) doy += 1.0;
    const double ndays = 365.0 + (leap ? 1.0 : 0.0);
    t = (double)year + (doy + ((double)hour + (double)min / 60.0) / 24.0) / ndays;
    return true;
}

// Runs the body of a C ABI function; an exception (a failed allocation for a
// corrupt or huge file, say) becomes a CHARM error instead of leaving the
// extern "C" function
template <class R, class F>
R guarded(charm_err *err, R fail, F &&body)
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        set_err(err, CHARM_EMEM, "Memory allocation failure.");
    } catch (const std::exception &e) {
        set_err(err, CHARM_EFILEIO, e.what());
    } catch (...) {
        set_err(err, CHARM_EFILEIO, "Unexpected failure.");
    }
    return fail;
}

}  // namespace
}  // namespace charm_amd

using namespace charm_amd;

extern "C" {

// ---------------------------------------------------------------------------
// bin: nmax (unsigned long), mu, r, then C and S, each order m = 0..nmax with
// degrees m..nmax (src/shc/shc_write_bin.c:18-187, shc_read_bin.c:26-256)
// ---------------------------------------------------------------------------
static unsigned long shc_read_bin_impl(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err)
{
    if (!nmax_only(nmax, shcs) && !shcs) {
        set_err(err, CHARM_EFUNCARG, "\"shcs\" cannot be NULL.");
        return CHARM_SHC_NMAX_ERROR;
    }
    FILE *f = std::fopen(pathname, "rb");
    if (!f) {
        set_err(err, CHARM_EFILEIO, std::string("Couldn't open \"") + pathname + "\".");
        return CHARM_SHC_NMAX_ERROR;
    }
    unsigned long nmax_file = CHARM_SHC_NMAX_ERROR;
    double mu = 0.0, r = 0.0;
    if (std::fread(&nmax_file, sizeof nmax_file, 1, f) != 1) {
        set_err(err, CHARM_EFILEIO, "Failed to read the maximum harmonic degree.");
        std::fclose(f);
        return nmax_file;
    }
    if (nmax_only(nmax, shcs)) {
        std::fclose(f);
        return nmax_file;
    }
    if (std::fread(&mu, sizeof mu, 1, f) != 1) {
        set_err(err, CHARM_EFILEIO, "Failed to read the scaling parameter.");
        std::fclose(f);
        return nmax_file;
    }
    if (std::fread(&r, sizeof r, 1, f) != 1) {
        set_err(err, CHARM_EFILEIO, "Failed to read the radius of the reference sphere.");
        std::fclose(f);
        return nmax_file;
    }
    if (nmax > nmax_file) {
        set_err(err, CHARM_EFUNCARG,
                "Not enough coefficients in the input file for the maximum harmonic degree "
                "\"nmax\".");
        std::fclose(f);
        return nmax_file;
    }
    if (shcs->nmax < nmax) {
        set_err(err, CHARM_EFUNCARG,
                "Too low maximum degree \"shcs->nmax\" to read coefficients up to degree "
                "\"nmax\".");
        std::fclose(f);
        return nmax_file;
    }
    shcs->mu = mu;
    shcs->r = r;
    reset_coeffs(shcs);
    // column by column: order m holds degrees m..nmax_file; the degrees above
    // nmax and the orders above nmax are skipped
    for (int cs = 0; cs < 2; cs++) {
        double **dst = cs ? shcs->s : shcs->c;
        bool ok = true;
        for (unsigned long m = 0; m <= nmax_file && ok; m++) {
            if (m <= nmax) {
                const size_t nread = nmax + 1 - m;
                ok = std::fread(dst[m], sizeof(double), nread, f) == nread;
                if (ok && nmax < nmax_file)
                    ok = std::fseek(f, (long)((nmax_file - nmax) * sizeof(double)), SEEK_CUR) == 0;
            } else {
                ok = std::fseek(f, (long)((nmax_file + 1 - m) * sizeof(double)), SEEK_CUR) == 0;
            }
        }
        if (!ok) {
            set_err(err, CHARM_EFILEIO, cs ? "Failed to read the \"S\" coefficients."
                                           : "Failed to read the \"C\" coefficients.");
            break;
        }
    }
    std::fclose(f);
    return nmax_file;
}

static void shc_write_bin_impl(const charm_shc *shcs, unsigned long nmax, const char *pathname,
                         charm_err *err)
{
    if (nmax > shcs->nmax) {
        set_err(err, CHARM_EFUNCARG,
                "Not enough coefficients in \"shcs\" to write up to degree \"nmax\".");
        return;
    }
    FILE *f = std::fopen(pathname, "wb");
    if (!f) {
        set_err(err, CHARM_EFILEIO, std::string("Couldn't create \"") + pathname + "\".");
        return;
    }
    bool ok = std::fwrite(&nmax, sizeof nmax, 1, f) == 1 &&
              std::fwrite(&shcs->mu, sizeof(double), 1, f) == 1 &&
              std::fwrite(&shcs->r, sizeof(double), 1, f) == 1;
    for (int cs = 0; cs < 2 && ok; cs++)
        for (unsigned long m = 0; m <= nmax && ok; m++) {
            const double *col = cs ? shcs->s[m] : shcs->c[m];
            ok = std::fwrite(col, sizeof(double), nmax + 1 - m, f) == nmax + 1 - m;
        }
    if (std::fclose(f) != 0) ok = false;
    if (!ok) set_err(err, CHARM_EFILEIO, "Failed writing to the output file.");
}

// ---------------------------------------------------------------------------
// tbl: "n m C S" per line, S optional for m = 0 (src/shc/shc_read_tbl.c)
// dov: "n m C" or "n -m S" per line (src/shc/shc_read_dov.c)
// ---------------------------------------------------------------------------
static unsigned long read_lines_format(const char *pathname, unsigned long nmax,
                                       charm_shc *shcs, charm_err *err, bool dov)
{
    if (!nmax_only(nmax, shcs) && !shcs) {
        set_err(err, CHARM_EFUNCARG, "\"shcs\" cannot be NULL.");
        return CHARM_SHC_NMAX_ERROR;
    }
    std::string text;
    if (!read_whole(pathname, "r", text)) {
        set_err(err, CHARM_EFILEIO, std::string("Couldn't open \"") + pathname + "\".");
        return CHARM_SHC_NMAX_ERROR;
    }
    Lines L;
    split_lines(text.data(), text.data() + text.size(), L);
    unsigned long nmax_file = CHARM_SHC_NMAX_ERROR;
    double mu = 0.0, r = 0.0;
    if (!read_mtdt(L, pathname, nmax_file, mu, r, err)) return nmax_file;
    if (nmax_only(nmax, shcs)) return nmax_file;
    shcs->mu = mu;
    shcs->r = r;
    if (!check_degrees(pathname, nmax, nmax_file, shcs, err)) return nmax_file;
    reset_coeffs(shcs);
    const size_t nl = L.size() - 1;
    std::vector<Rec> rec(nl);
    parallel_ranges(nl, [&](size_t i0, size_t i1) {
        Tok t[5];
        for (size_t i = i0; i < i1; i++) {
            Rec &q = rec[i];
            const int k = tokenize(L.b[i + 1], L.e[i + 1], t, 5);
            if (dov ? (k != 3) : (k != 3 && k != 4)) {
                q.status = -1;
                continue;
            }
            if (!str2ul(t[0], q.n)) {
                q.status = -2;
                continue;
            }
            if (q.n > nmax) {
                q.status = 1;
                continue;
            }
            Tok mt = t[1];
            if (dov && mt.n > 0 && mt.p[0] == '-') {
                q.sine = true;
                mt.p++;
                mt.n--;
            }
            if (!str2ul(mt, q.m)) {
                q.status = -3;
                continue;
            }
            if (q.m > q.n) {
                q.status = -6;
                continue;
            }
            if (!str2real(t[2], q.c)) {
                q.status = -4;
                continue;
            }
            if (!dov) {
                if (k == 3) {
                    if (q.m != 0) {
                        q.status = -5;
                        continue;
                    }
                    q.s = 0.0;
                } else if (!str2real(t[3], q.s)) {
                    q.status = -4;
                    continue;
                }
            }
        }
    });
    for (size_t i = 0; i < nl; i++) {
        const Rec &q = rec[i];
        if (q.status == 1) continue;
        if (q.status < 0) {
            static const char *what[] = {"",
                                         "has a wrong number of entries",
                                         "has an invalid harmonic degree",
                                         "has an invalid harmonic order",
                                         "has a coefficient that cannot be converted to the "
                                         "\"REAL\" data format",
                                         "gives no sine coefficient for an order above zero",
                                         "has a harmonic order larger than the degree"};
            std::string line(L.b[i + 1], L.e[i + 1]);
            set_err(err, CHARM_EFILEIO,
                    std::string("Line ") + std::to_string(i + 2) + " of \"" + pathname + "\" " +
                        what[-q.status] + ": \"" + line + "\".");
            return nmax_file;
        }
        if (dov) {
            if (q.sine)
                shcs->s[q.m][q.n - q.m] = q.c;
            else
                shcs->c[q.m][q.n - q.m] = q.c;
        } else {
            shcs->c[q.m][q.n - q.m] = q.c;
            shcs->s[q.m][q.n - q.m] = q.s;
        }
    }
    return nmax_file;
}

static unsigned long shc_read_tbl_impl(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err)
{
    return read_lines_format(pathname, nmax, shcs, err, false);
}

static unsigned long shc_read_dov_impl(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err)
{
    return read_lines_format(pathname, nmax, shcs, err, true);
}

// ---------------------------------------------------------------------------
// mtx: row i holds C_{i,0..i}, then S_{j,i+1} for j = i+1..nmax_file
// (src/shc/shc_read_mtx.c, shc_write_mtx.c)
// ---------------------------------------------------------------------------
static unsigned long shc_read_mtx_impl(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err)
{
    if (!nmax_only(nmax, shcs) && !shcs) {
        set_err(err, CHARM_EFUNCARG, "\"shcs\" cannot be NULL.");
        return CHARM_SHC_NMAX_ERROR;
    }
    std::string text;
    if (!read_whole(pathname, "r", text)) {
        set_err(err, CHARM_EFILEIO, std::string("Couldn't open \"") + pathname + "\".");
        return CHARM_SHC_NMAX_ERROR;
    }
    Lines L;
    split_lines(text.data(), text.data() + text.size(), L);
    unsigned long nmax_file = CHARM_SHC_NMAX_ERROR;
    double mu = 0.0, r = 0.0;
    if (!read_mtdt(L, pathname, nmax_file, mu, r, err)) return nmax_file;
    if (nmax_only(nmax, shcs)) return nmax_file;
    shcs->mu = mu;
    shcs->r = r;
    if (!check_degrees(pathname, nmax, nmax_file, shcs, err)) return nmax_file;
    reset_coeffs(shcs);
    // rows: the non-blank lines after the first one
    std::vector<size_t> rows;
    for (size_t i = 1; i < L.size() && rows.size() <= nmax; i++) {
        const char *p = L.b[i];
        while (p < L.e[i] && std::isspace((unsigned char)*p)) p++;
        if (p < L.e[i]) rows.push_back(i);
    }
    if (rows.size() < nmax + 1) {
        set_err(err, CHARM_EFILEIO,
                std::string("\"") + pathname + "\" has fewer coefficient rows than required.");
        return nmax_file;
    }
    std::vector<int> bad(nmax + 1, 0);
    parallel_ranges(nmax + 1, [&](size_t r0, size_t r1) {
        std::vector<Tok> t(nmax + 1);
        for (size_t row = r0; row < r1; row++) {
            const size_t li = rows[row];
            if (tokenize(L.b[li], L.e[li], t.data(), (int)nmax + 1) < (int)nmax + 1) {
                bad[row] = 1;
                continue;
            }
            for (unsigned long col = 0; col <= nmax; col++) {
                double v;
                if (!str2real(t[col], v)) {
                    bad[row] = 2;
                    break;
                }
                if (row >= col)
                    shcs->c[col][row - col] = v;
                else
                    shcs->s[row + 1][col - row - 1] = v;
            }
        }
    }, std::max<size_t>(1, 20000 / (nmax + 1)));
    for (unsigned long row = 0; row <= nmax; row++)
        if (bad[row]) {
            set_err(err, CHARM_EFILEIO,
                    std::string("Row ") + std::to_string(row) + " of \"" + pathname + "\" " +
                        (bad[row] == 1 ? "has too few entries."
                                       : "has an entry that cannot be converted to the \"REAL\" "
                                         "data format."));
            return nmax_file;
        }
    return nmax_file;
}

// ---------------------------------------------------------------------------
// gfc (ICGEM format, src/shc/shc_read_gfc.c:455-1409)
// ---------------------------------------------------------------------------
static unsigned long shc_read_gfc_impl(const char *pathname, unsigned long nmax, const char *epoch,
                                 charm_shc *shcs, charm_err *err)
{
    if (!nmax_only(nmax, shcs) && !shcs) {
        set_err(err, CHARM_EFUNCARG, "\"shcs\" cannot be NULL.");
        return CHARM_SHC_NMAX_ERROR;
    }
    std::string text;
    if (!read_whole(pathname, "r", text)) {
        set_err(err, CHARM_EFILEIO, std::string("Couldn't open \"") + pathname + "\".");
        return CHARM_SHC_NMAX_ERROR;
    }
    Lines L;
    split_lines(text.data(), text.data() + text.size(), L);
    const std::string P(pathname);

    // header: the last value of each keyword before end_of_head counts; a
    // value that does not convert is an error once the header has ended
    unsigned long nmax_file = CHARM_SHC_NMAX_ERROR;
    double mu_file = 0.0, r_file = 0.0;
    bool nmax_found = false, mu_found = false, r_found = false, errors_found = false;
    bool nmax_bad = false, mu_bad = false, r_bad = false, errors_bad = false, norm_bad = false;
    bool format_bad = false;
    std::string nmax_s, mu_s, r_s, errors_s, norm_s, format_s;
    bool e_no = false, e_cal = false, e_form = false, e_calform = false;
    bool icgem1 = false, icgem2 = false;
    size_t li = 0;
    bool eoh = false;
    for (; li < L.size() && !eoh; li++) {
        Tok t[2];
        const int k = tokenize(L.b[li], L.e[li], t, 2);
        if (k == 0) continue;
        if (t[0].is("end_of_head")) {
            eoh = true;
            continue;
        }
        if (k != 2) continue;
        if (t[0].is("max_degree")) {
            nmax_found = true;
            nmax_bad = !str2ul(t[1], nmax_file);
            nmax_s = t[1].str();
        } else if (t[0].is("earth_gravity_constant") || t[0].is("gravity_constant")) {
            mu_found = true;
            mu_bad = !str2real(t[1], mu_file);
            mu_s = t[1].str();
        } else if (t[0].is("radius")) {
            r_found = true;
            r_bad = !str2real(t[1], r_file);
            r_s = t[1].str();
        } else if (t[0].is("errors")) {
            errors_found = true;
            errors_bad = false;
            if (t[1].is("no"))
                e_no = true;
            else if (t[1].is("calibrated"))
                e_cal = true;
            else if (t[1].is("formal"))
                e_form = true;
            else if (t[1].is("calibrated_and_formal"))
                e_calform = true;
            else {
                errors_bad = true;
                errors_s = t[1].str();
            }
        } else if (t[0].is("norm")) {
            norm_bad = !t[1].is("fully_normalized");
            norm_s = t[1].str();
        } else if (t[0].is("format")) {
            format_bad = false;
            if (t[1].is("icgem1.0"))
                icgem1 = true;
            else if (t[1].is("icgem2.0"))
                icgem2 = true;
            else {
                format_bad = true;
                format_s = t[1].str();
            }
        }
    }
    if (!eoh) {
        set_err(err, CHARM_EFILEIO,
                "Couldn't find the \"end_of_head\" keyword in \"" + P +
                    "\" or couldn't read the file.");
        return nmax_file;
    }
    const char *missing = !nmax_found ? "max_degree"
                          : !mu_found ? "earth_gravity_constant\" or \"gravity_constant"
                          : !r_found  ? "radius"
                          : !errors_found ? "errors" : nullptr;
    if (missing) {
        set_err(err, CHARM_EFILEIO,
                std::string("Couldn't find the \"") + missing +
                    "\" keyword before reaching the \"end_of_head\" keyword in \"" + P + "\".");
        return nmax_file;
    }
    auto bad_value = [&](const char *key, const std::string &v, const char *type) {
        set_err(err, CHARM_EFILEIO,
                std::string("Couldn't convert the value \"") + v + "\" of the keyword \"" + key +
                    "\" in \"" + P + "\" to " + type + ".");
    };
    auto unsupported = [&](const char *key, const std::string &v) {
        set_err(err, CHARM_EFILEIO,
                std::string("Unsupported value \"") + v + "\" of the keyword \"" + key +
                    "\" in \"" + P + "\".");
    };
    if (nmax_bad) {
        bad_value("max_degree", nmax_s, "\"unsigned long int\"");
        return nmax_file;
    }
    if (mu_bad) {
        bad_value("earth_gravity_constant\" or \"gravity_constant", mu_s, "\"REAL\"");
        return nmax_file;
    }
    if (r_bad) {
        bad_value("radius", r_s, "\"REAL\"");
        return nmax_file;
    }
    if (errors_bad) {
        unsupported("errors", errors_s);
        return nmax_file;
    }
    if (norm_bad) {
        unsupported("norm", norm_s);
        return nmax_file;
    }
    if (format_bad) {
        unsupported("format", format_s);
        return nmax_file;
    }
    if (nmax_only(nmax, shcs)) return nmax_file;
    shcs->mu = mu_file;
    shcs->r = r_file;
    if (!icgem1 && !icgem2) icgem1 = true;
    const bool fmt2 = icgem2 && !icgem1;  // record layout of ICGEM 2.0
    double t_user = 0.0;
    if (epoch) {
        std::string why;
        if (!epoch_fraction(epoch, t_user, why)) {
            set_err(err, CHARM_EFUNCARG, why);
            return nmax_file;
        }
    }
    if (!check_degrees(pathname, nmax, nmax_file, shcs, err)) return nmax_file;
    if (fmt2 && epoch == nullptr) {
        set_err(err, CHARM_EFUNCARG,
                "\"epoch\" cannot be \"NULL\" for \"gfc\" files of the \"icgem2.0\" format.");
        return nmax_file;
    }
    reset_coeffs(shcs);

    // data records: tokens and numbers converted in parallel
    enum Key { K_NONE, K_GFC, K_GFCT, K_TRND, K_DOT, K_ASIN, K_ACOS };
    struct G {
        int key = K_NONE;
        int ns = 0;
        int status = 0;       // 0 ok, 1 degree above nmax, <0 error
        std::string msg;
        unsigned long n = 0, m = 0;
        double c = 0.0, s = 0.0, t0 = 0.0, t1 = 0.0, p = 0.0;
        bool has_t0 = false, has_t1 = false, has_p = false;
    };
    const size_t d0 = li, nd = L.size() - d0;
    std::vector<G> g(nd);
    // token slots of the epochs and the period per errors type and format
    const int slot_t0 = e_no ? 5 : ((e_cal || e_form) ? 7 : 9);
    const int slot_p1 = slot_t0;          // ICGEM 1.0: period where 2.0 has t0
    const int slot_p2 = slot_t0 + 2;      // ICGEM 2.0: period after t0, t1
    parallel_ranges(nd, [&](size_t i0, size_t i1) {
        Tok t[12];
        for (size_t i = i0; i < i1; i++) {
            G &q = g[i];
            const int ns = tokenize(L.b[d0 + i], L.e[d0 + i], t, 12);
            q.ns = ns;
            if (ns == 0) continue;
            if (t[0].is("gfc"))
                q.key = K_GFC;
            else if (t[0].is("gfct"))
                q.key = K_GFCT;
            else if (t[0].is("trnd"))
                q.key = K_TRND;
            else if (t[0].is("dot"))
                q.key = K_DOT;
            else if (t[0].is("asin"))
                q.key = K_ASIN;
            else if (t[0].is("acos"))
                q.key = K_ACOS;
            else
                continue;
            // entries required per line (keyword included): the largest
            // count any "errors" value seen in the header asks for
            int need;
            const int base = e_calform ? 4 : ((e_cal || e_form) ? 2 : 0);
            switch (q.key) {
            case K_GFC: need = e_calform ? 9 : ((e_cal || e_form) ? 7 : 4); break;
            case K_GFCT: need = (fmt2 ? 7 : 6) + base; break;
            case K_TRND:
            case K_DOT: need = (fmt2 ? 7 : 5) + base; break;
            default: need = (fmt2 ? 8 : 6) + base; break;
            }
            if (ns < need) {
                q.status = -1;
                q.msg = "has an incorrect number of entries";
                continue;
            }
            if (!str2ul(t[1], q.n)) {
                q.status = -1;
                q.msg = "has a harmonic degree that cannot be converted";
                continue;
            }
            if (q.n > nmax) {
                q.status = 1;
                continue;
            }
            if (!str2ul(t[2], q.m) || q.m > q.n) {
                q.status = -1;
                q.msg = "has an invalid harmonic order";
                continue;
            }
            if (!str2real(t[3], q.c)) {
                q.status = -1;
                q.msg = "has a C coefficient that cannot be converted";
                continue;
            }
            if (q.key == K_GFC && q.m == 0) {
                q.s = 0.0;
            } else {
                if (ns < 5 || !str2real(t[4], q.s)) {
                    q.status = -1;
                    q.msg = ns < 5 ? "has an incorrect number of entries"
                                   : "has an S coefficient that cannot be converted";
                    continue;
                }
            }
            if (q.key == K_GFC) continue;
            std::string why;
            const bool want_t0 = (q.key == K_GFCT) || fmt2;
            if (want_t0) {
                const std::string s0 = t[slot_t0].str();
                if (!epoch_fraction(s0.c_str(), q.t0, why)) {
                    q.status = -2;
                    q.msg = why;
                    continue;
                }
                q.has_t0 = true;
            }
            if (fmt2) {
                const std::string s1 = t[slot_t0 + 1].str();
                if (!epoch_fraction(s1.c_str(), q.t1, why)) {
                    q.status = -2;
                    q.msg = why;
                    continue;
                }
                q.has_t1 = true;
            }
            if (q.key == K_ASIN || q.key == K_ACOS) {
                if (!str2real(t[fmt2 ? slot_p2 : slot_p1], q.p)) {
                    q.status = -1;
                    q.msg = "has a period that cannot be converted";
                    continue;
                }
                q.has_p = true;
            }
        }
    });

    // apply in file order (the state of icgem1.0 groups: t0 of the last gfct)
    const bool no_epoch = !fmt2 && epoch == nullptr;
    double t = t_user, t0 = 0.0;
    unsigned long n = 0, m = 0;
    bool f_trnd = false, f_dot = false, f_asin = false, f_acos = false;
    for (size_t i = 0; i < nd; i++) {
        const G &q = g[i];
        if (q.key == K_NONE || q.status == 1) continue;
        if (q.status < 0) {
            set_err(err, q.status == -2 ? CHARM_EFUNCARG : CHARM_EFILEIO,
                    "Line " + std::to_string(d0 + i + 1) + " of \"" + P + "\" " + q.msg + ".");
            return nmax_file;
        }
        double cnm = q.c, snm = q.s;
        if (q.key == K_GFCT) {
            n = q.n;
            m = q.m;
            t0 = q.t0;
            if (fmt2 && !(t_user >= q.t0 && t_user < q.t1)) continue;
            if (f_trnd && f_dot) {
                set_err(err, CHARM_EFILEIO,
                        "Both \"trnd\" and \"dot\" coefficients were found for one \"gfct\" "
                        "record in \"" + P + "\".");
                return nmax_file;
            } else if ((f_trnd || f_dot) && f_asin && f_acos) {
            } else if (f_trnd || f_dot || f_asin || f_acos) {
                set_err(err, CHARM_EFILEIO,
                        "Incomplete set of \"gfct\", \"trnd\" (or \"dot\"), \"asin\" and "
                        "\"acos\" records before line " + std::to_string(d0 + i + 1) + " of \"" +
                            P + "\".");
                return nmax_file;
            }
            f_trnd = f_dot = f_asin = f_acos = false;
        } else if (q.key != K_GFC) {
            if (q.n != n || q.m != m) {
                set_err(err, CHARM_EFILEIO,
                        "Line " + std::to_string(d0 + i + 1) + " of \"" + P +
                            "\" does not follow the degree and order of its \"gfct\" record.");
                return nmax_file;
            }
            if (fmt2) {
                t0 = q.t0;
                if (!(t_user >= q.t0 && t_user < q.t1)) continue;
            }
            if (no_epoch) t = t0;
            const double dt = t - t0;
            if (q.key == K_TRND || q.key == K_DOT) {
                cnm *= dt;
                snm *= dt;
                if (q.key == K_DOT)
                    f_dot = f_asin = f_acos = true;
                else
                    f_trnd = true;
            } else {
                const double arg = (2.0 * PI_D / q.p) * dt;
                const double w = (q.key == K_ASIN) ? std::sin(arg) : std::cos(arg);
                cnm *= w;
                snm *= w;
                (q.key == K_ASIN ? f_asin : f_acos) = true;
            }
        }
        shcs->c[q.m][q.n - q.m] += cnm;
        shcs->s[q.m][q.n - q.m] += snm;
    }
    return nmax_file;
}

// ---------------------------------------------------------------------------
// Text writers (src/shc/shc_write_mtdt.c, shc_write_tbl.c, shc_write_dov.c,
// shc_write_mtx.c): every real through the caller's printf format
// ---------------------------------------------------------------------------
static FILE *open_text_out(const charm_shc *shcs, unsigned long nmax, const char *pathname,
                           const char *format, charm_err *err)
{
    FILE *f = std::fopen(pathname, "w");
    if (!f) {
        set_err(err, CHARM_EFILEIO, std::string("Couldn't create \"") + pathname + "\".");
        return nullptr;
    }
    if (nmax > shcs->nmax) {
        set_err(err, CHARM_EFUNCARG,
                "Not enough coefficients in \"shcs\" to write up to degree \"nmax\".");
        std::fclose(f);
        return nullptr;
    }
    if (std::fprintf(f, "%lu ", nmax) < 1 || std::fprintf(f, format, shcs->mu) < 1 ||
        std::fprintf(f, " ") < 1 || std::fprintf(f, format, shcs->r) < 1 ||
        std::fprintf(f, "\n") < 1) {
        set_err(err, CHARM_EFILEIO, "Failed to write the metadata.");
        std::fclose(f);
        return nullptr;
    }
    return f;
}

static void close_text_out(FILE *f, bool ok, charm_err *err)
{
    if (std::fclose(f) != 0) ok = false;
    if (!ok) set_err(err, CHARM_EFILEIO, "Failed writing to the output file.");
}

static void write_pairs(const charm_shc *shcs, unsigned long nmax, const char *format,
                        int ordering, const char *pathname, charm_err *err, bool dov)
{
    if (ordering != CHARM_SHC_WRITE_N && ordering != CHARM_SHC_WRITE_M) {
        set_err(err, CHARM_EFUNCARG, "Unsupported value of \"ordering\".");
        return;
    }
    FILE *f = open_text_out(shcs, nmax, pathname, format, err);
    if (!f) return;
    bool ok = true;
    auto put = [&](unsigned long n, unsigned long m) {
        const double c = shcs->c[m][n - m], s = shcs->s[m][n - m];
        if (dov) {
            ok = ok && std::fprintf(f, "%lu  %lu ", n, m) > 0 && std::fprintf(f, format, c) > 0 &&
                 std::fprintf(f, "\n") > 0;
            if (m > 0)
                ok = ok && std::fprintf(f, "%lu -%lu ", n, m) > 0 &&
                     std::fprintf(f, format, s) > 0 && std::fprintf(f, "\n") > 0;
        } else {
            ok = ok && std::fprintf(f, "%lu %lu ", n, m) > 0 && std::fprintf(f, format, c) > 0 &&
                 std::fprintf(f, " ") > 0 && std::fprintf(f, format, s) > 0 &&
                 std::fprintf(f, "\n") > 0;
        }
    };
    if (ordering == CHARM_SHC_WRITE_N) {
        for (unsigned long m = 0; m <= nmax && ok; m++)
            for (unsigned long n = m; n <= nmax && ok; n++) put(n, m);
    } else {
        for (unsigned long n = 0; n <= nmax && ok; n++)
            for (unsigned long m = 0; m <= n && ok; m++) put(n, m);
    }
    close_text_out(f, ok, err);
}

static void shc_write_tbl_impl(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         int ordering, const char *pathname, charm_err *err)
{
    write_pairs(shcs, nmax, formatting, ordering, pathname, err, false);
}

static void shc_write_dov_impl(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         int ordering, const char *pathname, charm_err *err)
{
    write_pairs(shcs, nmax, formatting, ordering, pathname, err, true);
}

static void shc_write_mtx_impl(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         const char *pathname, charm_err *err)
{
    FILE *f = open_text_out(shcs, nmax, pathname, formatting, err);
    if (!f) return;
    bool ok = true;
    for (unsigned long row = 0; row <= nmax && ok; row++) {
        for (unsigned long col = 0; col <= nmax && ok; col++) {
            const double v = (row >= col) ? shcs->c[col][row - col] : shcs->s[row + 1][col - row - 1];
            ok = std::fprintf(f, formatting, v) > 0 && (col == nmax || std::fprintf(f, " ") > 0);
        }
        ok = ok && std::fprintf(f, "\n") > 0;
    }
    close_text_out(f, ok, err);
}

unsigned long charm_shc_read_bin(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err)
{
    return guarded(err, CHARM_SHC_NMAX_ERROR, [&] { return shc_read_bin_impl(pathname, nmax, shcs, err); });
}

unsigned long charm_shc_read_tbl(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err)
{
    return guarded(err, CHARM_SHC_NMAX_ERROR, [&] { return shc_read_tbl_impl(pathname, nmax, shcs, err); });
}

unsigned long charm_shc_read_dov(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err)
{
    return guarded(err, CHARM_SHC_NMAX_ERROR, [&] { return shc_read_dov_impl(pathname, nmax, shcs, err); });
}

unsigned long charm_shc_read_mtx(const char *pathname, unsigned long nmax, charm_shc *shcs,
                                 charm_err *err)
{
    return guarded(err, CHARM_SHC_NMAX_ERROR, [&] { return shc_read_mtx_impl(pathname, nmax, shcs, err); });
}

unsigned long charm_shc_read_gfc(const char *pathname, unsigned long nmax, const char *epoch,
                                 charm_shc *shcs, charm_err *err)
{
    return guarded(err, CHARM_SHC_NMAX_ERROR,
                   [&] { return shc_read_gfc_impl(pathname, nmax, epoch, shcs, err); });
}

void charm_shc_write_bin(const charm_shc *shcs, unsigned long nmax, const char *pathname,
                         charm_err *err)
{
    guarded(err, 0, [&] { shc_write_bin_impl(shcs, nmax, pathname, err); return 0; });
}

void charm_shc_write_tbl(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         int ordering, const char *pathname, charm_err *err)
{
    guarded(err, 0, [&] { shc_write_tbl_impl(shcs, nmax, formatting, ordering, pathname, err); return 0; });
}

void charm_shc_write_dov(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         int ordering, const char *pathname, charm_err *err)
{
    guarded(err, 0, [&] { shc_write_dov_impl(shcs, nmax, formatting, ordering, pathname, err); return 0; });
}

void charm_shc_write_mtx(const charm_shc *shcs, unsigned long nmax, const char *formatting,
                         const char *pathname, charm_err *err)
{
    guarded(err, 0, [&] { shc_write_mtx_impl(shcs, nmax, formatting, pathname, err); return 0; });
}

}  // extern "C"
