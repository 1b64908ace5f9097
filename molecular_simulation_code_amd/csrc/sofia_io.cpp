// Checkpoint / restart and coordinate files of the hot path's host side.
//
//   restart files   dump_global_velocity_id_mpi / read_dump_global_id_v
//                   (Source/mod_datadump.f90:113-151, 212-252)
//   GRO files       read_in_Grofile / read_in_Grofile_box (Source/mod_GroFile.f90:89-159)
//                   and the coordinate writer of template_generator (:7-68)
//
// Restart layout (text): box, temp, num, timeclock on one line each, then one record per
// atom "id x y z vx vy vz" with ids as in the reference's (i5,6f12.6) and six f12.6 fields.
// The reference's i5 prints ***** from id 100000 on; here the id field widens instead
// (ids stay separated from the x field, whose f12.6 keeps leading blanks below 1e4 A).
// The reader is list-directed like the reference's: records may come in any order and go
// to the slot of their global id.
//
// GRO files are parsed by column (resid 0-4, resname 5-9, atom name 10-14, atom number
// 15-19, x y z from column 20 in 8-wide fields, nm), which stays correct where the
// reference's list-directed read needs special cases: from atom 10000 on the number runs
// into the name field.  Coordinates and box are converted from nm to A.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sofia_hip.h"

namespace sofia {
int io_fail(int code, const std::string &msg);   // sets sofia_last_error (sofia_api.hip)
}

namespace {

struct File {
    FILE *f = nullptr;
    explicit File(FILE *p) : f(p) {}
    ~File() { if (f) std::fclose(f); }
};

bool read_line(FILE *f, std::string &out) {
    out.clear();
    int ch;
    while ((ch = std::fgetc(f)) != EOF) {
        if (ch == '\n') return true;
        out.push_back((char)ch);
    }
    return !out.empty();
}

// fixed-width field [a, a + w) of a line as a double (blank or missing -> false)
bool field_d(const std::string &l, size_t a, size_t w, double *v) {
    if (l.size() <= a) return false;
    std::string s = l.substr(a, w);
    char *end = nullptr;
    errno = 0;
    const double x = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || errno) return false;
    *v = x;
    return true;
}

}  // namespace

extern "C" {

int sofia_write_restart(const char *path, float box, float temp, int32_t num, float timeclock,
                        int32_t np, const int32_t *global_id, const float *pos, const float *v) {
    if (!path || np < 0 || (np > 0 && (!global_id || !pos || !v)))
        return sofia::io_fail(SOFIA_E_ARG, "write_restart: null argument");
    File f(std::fopen(path, "w"));
    if (!f.f) return sofia::io_fail(SOFIA_E_ARG, std::string("write_restart: cannot open ") + path);
    // header lines in the reference's list-directed style
    std::fprintf(f.f, "   %.8g\n   %.8g\n %11d\n   %.8g\n", (double)box, (double)temp, num,
                 (double)timeclock);
    for (int32_t i = 0; i < np; ++i) {
        const float *x = pos + 3 * (size_t)i, *u = v + 3 * (size_t)i;
        std::fprintf(f.f, "%5d%12.6f%12.6f%12.6f%12.6f%12.6f%12.6f\n", global_id[i], (double)x[0],
                     (double)x[1], (double)x[2], (double)u[0], (double)u[1], (double)u[2]);
    }
    if (std::ferror(f.f)) return sofia::io_fail(SOFIA_E_ARG, "write_restart: write error");
    return SOFIA_OK;
}

int sofia_read_restart(const char *path, int32_t np, float *box, float *temp, int32_t *num,
                       float *timeclock, float *pos, float *v) {
    if (!path || np < 0 || !box || !temp || !num || !timeclock || (np > 0 && (!pos || !v)))
        return sofia::io_fail(SOFIA_E_ARG, "read_restart: null argument");
    File f(std::fopen(path, "r"));
    if (!f.f) return sofia::io_fail(SOFIA_E_ARG, std::string("read_restart: cannot open ") + path);
    double b, t, tc;
    long n;
    if (std::fscanf(f.f, "%lf %lf %ld %lf", &b, &t, &n, &tc) != 4)
        return sofia::io_fail(SOFIA_E_ARG, "read_restart: bad header (box, temp, num, timeclock)");
    *box = (float)b; *temp = (float)t; *num = (int32_t)n; *timeclock = (float)tc;
    std::vector<char> seen((size_t)np, 0);
    for (int32_t i = 0; i < np; ++i) {
        long j;
        double r[6];
        if (std::fscanf(f.f, "%ld %lf %lf %lf %lf %lf %lf", &j, r, r + 1, r + 2, r + 3, r + 4, r + 5) != 7)
            return sofia::io_fail(SOFIA_E_ARG, "read_restart: fewer atom records than np");
        if (j < 1 || j > np) return sofia::io_fail(SOFIA_E_ARG, "read_restart: global id out of range");
        if (seen[j - 1]) return sofia::io_fail(SOFIA_E_ARG, "read_restart: global id repeated");
        seen[j - 1] = 1;
        for (int c = 0; c < 3; ++c) {
            pos[3 * (size_t)(j - 1) + c] = (float)r[c];
            v[3 * (size_t)(j - 1) + c] = (float)r[3 + c];
        }
    }
    return SOFIA_OK;
}

int sofia_read_gro(const char *path, int32_t np_max, int32_t *np, float *pos, float *box) {
    if (!path || !np || !box) return sofia::io_fail(SOFIA_E_ARG, "read_gro: null argument");
    File f(std::fopen(path, "r"));
    if (!f.f) return sofia::io_fail(SOFIA_E_ARG, std::string("read_gro: cannot open ") + path);
    std::string line;
    if (!read_line(f.f, line)) return sofia::io_fail(SOFIA_E_ARG, "read_gro: empty file");   // title
    if (!read_line(f.f, line)) return sofia::io_fail(SOFIA_E_ARG, "read_gro: no atom count");
    char *end = nullptr;
    const long n = std::strtol(line.c_str(), &end, 10);
    if (end == line.c_str() || n < 0) return sofia::io_fail(SOFIA_E_ARG, "read_gro: bad atom count");
    *np = (int32_t)n;
    if (!pos) {   // size query: atom count and box only
        for (long i = 0; i < n; ++i)
            if (!read_line(f.f, line)) return sofia::io_fail(SOFIA_E_ARG, "read_gro: fewer atom lines than the count");
    } else {
        if (n > np_max) return sofia::io_fail(SOFIA_E_ARG, "read_gro: more atoms than np_max");
        for (long i = 0; i < n; ++i) {
            if (!read_line(f.f, line)) return sofia::io_fail(SOFIA_E_ARG, "read_gro: fewer atom lines than the count");
            double x, y, z;
            if (!field_d(line, 20, 8, &x) || !field_d(line, 28, 8, &y) || !field_d(line, 36, 8, &z))
                return sofia::io_fail(SOFIA_E_ARG, "read_gro: bad coordinate columns at atom " + std::to_string(i + 1));
            pos[3 * i] = (float)(x * 10.0);
            pos[3 * i + 1] = (float)(y * 10.0);
            pos[3 * i + 2] = (float)(z * 10.0);
        }
    }
    if (!read_line(f.f, line)) return sofia::io_fail(SOFIA_E_ARG, "read_gro: no box line");
    double b;
    if (!field_d(line, 0, 10, &b)) return sofia::io_fail(SOFIA_E_ARG, "read_gro: bad box line");
    *box = (float)(b * 10.0);   // cubic box, the first of the (f10.5 x 3) values
    return SOFIA_OK;
}

int sofia_write_gro(const char *path, const char *title, int32_t np, const int32_t *resid,
                    const char *resname, const char *atomname, const float *pos, float box) {
    if (!path || np < 0 || (np > 0 && (!resid || !resname || !atomname || !pos)))
        return sofia::io_fail(SOFIA_E_ARG, "write_gro: null argument");
    File f(std::fopen(path, "w"));
    if (!f.f) return sofia::io_fail(SOFIA_E_ARG, std::string("write_gro: cannot open ") + path);
    std::fprintf(f.f, "%s\n%5d\n", title ? title : "", np);
    for (int32_t i = 0; i < np; ++i) {
        char rn[6] = {0}, an[6] = {0};
        std::memcpy(rn, resname + 5 * (size_t)i, 5);
        std::memcpy(an, atomname + 5 * (size_t)i, 5);
        // ( i5 , 2a5 , i5 , 3f8.3): numbers wrap at 100000 as GRO files do
        std::fprintf(f.f, "%5d%-5.5s%5.5s%5d%8.3f%8.3f%8.3f\n", resid[i] % 100000, rn, an,
                     (i + 1) % 100000, (double)pos[3 * (size_t)i] / 10.0,
                     (double)pos[3 * (size_t)i + 1] / 10.0, (double)pos[3 * (size_t)i + 2] / 10.0);
    }
    const double b = (double)box / 10.0;
    std::fprintf(f.f, "%10.5f%10.5f%10.5f\n", b, b, b);
    if (std::ferror(f.f)) return sofia::io_fail(SOFIA_E_ARG, "write_gro: write error");
    return SOFIA_OK;
}

}  // extern "C"
