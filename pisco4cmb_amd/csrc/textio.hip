/*
 * textio.hip -- host-side readers and writers of the reference's text
 * formats: whitespace-separated columns, one row per HEALPix pixel, as
 * numpy.savetxt writes them (sky maps: I Q U V per pixel,
 * scripts/create_cmb.py:96-97; tilde beams: I Q U V per beam pixel,
 * scripts/fields2beams.py:175-176).  No device code.
 *
 * The file is read whole, split at line boundaries into one part per thread,
 * and each part is parsed with strtof (correctly rounded to float, as a C++
 * stream extracts a float).  Blank lines and lines starting with '#' are
 * skipped.
 */
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "pisco.h"
#include "kernels.hpp"

namespace {

int io_fail(int code, const std::string &msg) { return pisco::set_error(code, msg.c_str()); }

bool skip_line(const char *b, const char *e)
{
    while (b < e && (*b == ' ' || *b == '\t' || *b == '\r'))
        b++;
    return b == e || *b == '#';
}

/* Parses the rows of [b, e) (whole lines) into cols starting at row r0.
 * Returns the number of rows parsed, or -1 on a bad line (*why says which). */
long parse_part(const char *b, const char *e, int ncols, float *const *cols, long r0,
                long capacity, std::string *why)
{
    long r = r0;
    const char *p = b;
    std::string line;
    while (p < e) {
        const char *nl = (const char *)memchr(p, '\n', (size_t)(e - p));
        const char *le = nl ? nl : e;
        if (!skip_line(p, le)) {
            if (r >= capacity) {
                r++;
            } else {
                line.assign(p, le);
                const char *s = line.c_str();
                char *end = nullptr;
                for (int k = 0; k < ncols; k++) {
                    errno = 0;
                    float v = strtof(s, &end);
                    if (end == s) {
                        *why = "row " + std::to_string(r) + ": expected " +
                               std::to_string(ncols) + " numbers";
                        return -1;
                    }
                    if (cols[k])
                        cols[k][r] = v;
                    s = end;
                }
                while (*s == ' ' || *s == '\t' || *s == '\r')
                    s++;
                if (*s != '\0' && *s != '#') {
                    *why = "row " + std::to_string(r) + ": more than " + std::to_string(ncols) +
                           " columns";
                    return -1;
                }
                r++;
            }
        }
        p = nl ? nl + 1 : e;
    }
    return r - r0;
}

long count_rows(const char *b, const char *e)
{
    long n = 0;
    const char *p = b;
    while (p < e) {
        const char *nl = (const char *)memchr(p, '\n', (size_t)(e - p));
        const char *le = nl ? nl : e;
        if (!skip_line(p, le))
            n++;
        p = nl ? nl + 1 : e;
    }
    return n;
}

int read_file(const char *path, std::vector<char> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return io_fail(PISCO_ERR_INVALID, std::string("cannot open ") + path);
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return io_fail(PISCO_ERR_INVALID, std::string("cannot seek ") + path);
    }
    long sz = ftell(f);
    rewind(f);
    buf.resize(sz > 0 ? (size_t)sz : 0);
    size_t got = sz > 0 ? fread(buf.data(), 1, (size_t)sz, f) : 0;
    fclose(f);
    if ((long)got != sz)
        return io_fail(PISCO_ERR_INVALID, std::string("short read of ") + path);
    return PISCO_OK;
}

} // namespace

extern "C" {

int pisco_read_txt_columns(const char *path, int ncols, long capacity, float *const cols[],
                           long *nrows)
{
    if (!path || !nrows || ncols < 1 || ncols > 64 || capacity < 0 || (capacity > 0 && !cols))
        return io_fail(PISCO_ERR_INVALID, "bad argument");
    std::vector<char> buf;
    int rc = read_file(path, buf);
    if (rc != PISCO_OK)
        return rc;
    const char *b = buf.data(), *e = b + buf.size();
    /* split at line boundaries, one part per thread */
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (buf.size() < (1u << 20))
        nt = 1;
    std::vector<const char *> cut(nt + 1);
    cut[0] = b;
    cut[nt] = e;
    for (unsigned t = 1; t < nt; t++) {
        const char *p = b + (size_t)buf.size() * t / nt;
        p = std::max(p, cut[t - 1]);
        const char *nl = (const char *)memchr(p, '\n', (size_t)(e - p));
        cut[t] = nl ? nl + 1 : e;
    }
    std::vector<long> rows(nt);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++)
            th.emplace_back([&, t] { rows[t] = count_rows(cut[t], cut[t + 1]); });
        for (auto &x : th)
            x.join();
    }
    long total = 0;
    std::vector<long> r0(nt);
    for (unsigned t = 0; t < nt; t++) {
        r0[t] = total;
        total += rows[t];
    }
    *nrows = total;
    if (total > capacity)
        return capacity == 0 ? PISCO_OK
                             : io_fail(PISCO_ERR_INVALID,
                                       std::string(path) + " has " + std::to_string(total) +
                                           " rows, more than the " + std::to_string(capacity) +
                                           " the arrays hold");
    std::vector<long> got(nt);
    std::vector<std::string> why(nt);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++)
            th.emplace_back([&, t] {
                got[t] = parse_part(cut[t], cut[t + 1], ncols, cols, r0[t], capacity, &why[t]);
            });
        for (auto &x : th)
            x.join();
    }
    for (unsigned t = 0; t < nt; t++)
        if (got[t] < 0)
            return io_fail(PISCO_ERR_INVALID, std::string(path) + ": " + why[t]);
    return PISCO_OK;
}

int pisco_write_txt_columns(const char *path, int ncols, long nrows, const float *const cols[])
{
    if (!path || ncols < 1 || ncols > 64 || nrows < 0 || !cols)
        return io_fail(PISCO_ERR_INVALID, "bad argument");
    for (int k = 0; k < ncols; k++)
        if (!cols[k])
            return io_fail(PISCO_ERR_INVALID, "NULL column");
    FILE *f = fopen(path, "wb");
    if (!f)
        return io_fail(PISCO_ERR_INVALID, std::string("cannot create ") + path);
    char line[64 * 32];
    for (long r = 0; r < nrows; r++) {
        int n = 0;
        for (int k = 0; k < ncols; k++)
            n += snprintf(line + n, sizeof line - (size_t)n, k ? " %.18e" : "%.18e",
                          (double)cols[k][r]);
        line[n++] = '\n';
        if (fwrite(line, 1, (size_t)n, f) != (size_t)n) {
            fclose(f);
            return io_fail(PISCO_ERR_INVALID, std::string("write failed: ") + path);
        }
    }
    if (fclose(f) != 0)
        return io_fail(PISCO_ERR_INVALID, std::string("close failed: ") + path);
    return PISCO_OK;
}

} /* extern "C" */
