// simp_io.cpp -- snapshots in the reference's CSV format (host code).
//
// The writer reproduces output.f90:47-76 byte for byte: three files of a
// run folder, appended once per snapshot, with the formats
//   1001 (1x, I7, 4(1x, A1, 1x, e15.8E3))    particle_momentum.csv
//   1002 (1x, I7, 11(1x, A1, 1x, e15.8E3))   particle_state.csv
//   1003 (1x, 2(I7, 1x, A1, 1x), 1(e15.8E3)) particle_other.csv
// (output.f90:61-63).  The reader is the spin-up reader of
// initialize.f90:234-330 made to read what the writer writes: the
// reference's reader expects 14 state columns where output.f90 writes 11
// and reads the first snapshot of the files; this one reads the last.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/simp.h"

namespace {

// Fortran Ew.dEe with w = 15, d = 8, e = 3: [-]0.dddddddd E+eee, the
// optional leading zero dropped when the sign needs its column; the eight
// digits are the value correctly rounded to nearest (ties to even, as the
// compiled reference prints them).  Non-finite values are right-justified
// words, as the Fortran runtime writes them.
void put_e15(std::string &o, double v) {
  char b[48];
  if (std::isnan(v)) {
    o.append(12, ' ');
    o.append("NaN");
    return;
  }
  if (std::isinf(v)) {
    const char *w = v > 0 ? "Inf" : "-Inf";
    o.append(15 - std::strlen(w), ' ');
    o.append(w);
    return;
  }
  const bool neg = std::signbit(v);
  int e10 = 0;
  char dig[9];
  if (v == 0.0) {
    std::memset(dig, '0', 8);
  } else {
    // d.ddddddde+xx: eight significant digits, correctly rounded
    std::snprintf(b, sizeof b, "%.7e", std::fabs(v));
    dig[0] = b[0];
    std::memcpy(dig + 1, b + 2, 7);
    e10 = std::atoi(b + 10) + 1;   // 0.dddddddd x 10^e10
  }
  dig[8] = 0;
  o.append(neg ? "-." : "0.");
  o.append(dig, 8);
  std::snprintf(b, sizeof b, "E%c%03d", e10 < 0 ? '-' : '+', e10 < 0 ? -e10 : e10);
  o.append(b);
}

// Iw: right-justified, asterisks when the value does not fit
void put_i(std::string &o, long v, int w) {
  char b[32];
  const int n = std::snprintf(b, sizeof b, "%ld", v);
  if (n > w) {
    o.append(w, '*');
    return;
  }
  o.append(w - n, ' ');
  o.append(b, n);
}

std::string path_of(const char *folder, const char *name) {
  std::string p(folder);
  if (!p.empty() && p.back() != '/') p += '/';
  return p + name;
}

const char *kFiles[3] = {"particle_momentum.csv", "particle_state.csv", "particle_other.csv"};

int append_file(const std::string &path, const std::string &text) {
  FILE *f = std::fopen(path.c_str(), "ab");
  if (!f) return SIMP_E_ARG;
  const size_t n = std::fwrite(text.data(), 1, text.size(), f);
  const int rc = std::fclose(f);
  return (n == text.size() && rc == 0) ? SIMP_OK : SIMP_E_ARG;
}

bool read_lines(const std::string &path, std::vector<std::string> &lines) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::string cur;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) {
    for (size_t i = 0; i < n; ++i) {
      if (buf[i] == '\n') {
        lines.push_back(cur);
        cur.clear();
      } else if (buf[i] != '\r') {
        cur += buf[i];
      }
    }
  }
  std::fclose(f);
  if (!cur.empty()) lines.push_back(cur);
  return true;
}

// list-directed read of comma/space separated numbers
int parse_fields(const std::string &s, double *out, int cap) {
  const char *p = s.c_str();
  int k = 0;
  while (*p && k < cap) {
    while (*p == ' ' || *p == ',' || *p == '\t') ++p;
    if (!*p) break;
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(p, &end);
    if (end == p) return -1;
    out[k++] = v;
    p = end;
  }
  return k;
}

}  // namespace

extern "C" {

int simp_output_create(const char *folder) {
  if (!folder) return SIMP_E_ARG;
  for (const char *name : kFiles) {
    FILE *f = std::fopen(path_of(folder, name).c_str(), "wb");
    if (!f) return SIMP_E_ARG;
    std::fclose(f);
  }
  return SIMP_OK;
}

int simp_output_append(const char *folder, int n, const double *x, const double *vx,
                       const double *mass, const double *p, const double *hsml, const double *A,
                       const double *hth, const double *eps3, const double *sig3, int nb,
                       double current_time) {
  if (!folder || n < 0 || nb < 0 || (n > 0 && (!x || !vx || !mass || !p || !hsml || !A ||
                                               !hth || !eps3 || !sig3)))
    return SIMP_E_ARG;
  std::string m, s, o;
  m.reserve((size_t)n * 88);
  s.reserve((size_t)n * 216);
  for (int i = 0; i < n; ++i) {
    // output.f90:67-69: position in km, velocity in cm/s
    m += ' ';
    put_i(m, i + 1, 7);
    for (int d = 0; d < 2; ++d) {
      m += " , ";
      put_e15(m, x[2 * i + d] / 1000);
    }
    for (int d = 0; d < 2; ++d) {
      m += " , ";
      put_e15(m, vx[2 * i + d] * 100);
    }
    m += '\n';
    // output.f90:71-75
    const double f[11] = {mass[i], p[i], hsml[i], A[i], hth[i], eps3[3 * i], eps3[3 * i + 1],
                          eps3[3 * i + 2], sig3[3 * i], sig3[3 * i + 1], sig3[3 * i + 2]};
    s += ' ';
    put_i(s, i + 1, 7);
    for (double v : f) {
      s += " , ";
      put_e15(s, v);
    }
    s += '\n';
  }
  // output.f90:78
  o += ' ';
  put_i(o, n, 7);
  o += " , ";
  put_i(o, nb, 7);
  o += " , ";
  put_e15(o, current_time / 3600);
  o += '\n';
  int rc = append_file(path_of(folder, kFiles[0]), m);
  if (!rc) rc = append_file(path_of(folder, kFiles[1]), s);
  if (!rc) rc = append_file(path_of(folder, kFiles[2]), o);
  return rc;
}

int simp_output_read_last(const char *folder, int cap, int *n, int *nb, double *current_time,
                          double *x, double *vx, double *mass, double *p, double *hsml,
                          double *A, double *hth, double *eps3, double *sig3) {
  if (!folder || !n || !nb || !current_time || cap < 0) return SIMP_E_ARG;
  std::vector<std::string> lo, lm, ls;
  if (!read_lines(path_of(folder, kFiles[2]), lo) || lo.empty()) return SIMP_E_ARG;
  double g[3];
  if (parse_fields(lo.back(), g, 3) != 3) return SIMP_E_ARG;
  const int nt = (int)g[0];
  *n = nt;
  *nb = (int)g[1];
  *current_time = g[2] * 3600;   // initialize.f90:282
  if (nt > cap) return SIMP_E_CAPACITY;
  if (nt > 0 && (!x || !vx || !mass || !p || !hsml || !A || !hth || !eps3 || !sig3))
    return SIMP_E_ARG;
  if (!read_lines(path_of(folder, kFiles[0]), lm) || !read_lines(path_of(folder, kFiles[1]), ls))
    return SIMP_E_ARG;
  // every snapshot holds n records; the last one is the tail of the files
  if ((long)lm.size() < nt || (long)ls.size() < nt) return SIMP_E_ARG;
  const size_t m0 = lm.size() - nt, s0 = ls.size() - nt;
  for (int r = 0; r < nt; ++r) {
    double a[5], b[12];
    if (parse_fields(lm[m0 + r], a, 5) != 5 || parse_fields(ls[s0 + r], b, 12) != 12)
      return SIMP_E_ARG;
    const int i = (int)a[0] - 1;
    if (i < 0 || i >= nt || (int)b[0] - 1 != i) return SIMP_E_ARG;
    // initialize.f90:300-301: back to m and m/s
    x[2 * i] = a[1] * 1000;
    x[2 * i + 1] = a[2] * 1000;
    vx[2 * i] = a[3] / 100;
    vx[2 * i + 1] = a[4] / 100;
    mass[i] = b[1];
    p[i] = b[2];
    hsml[i] = b[3];
    A[i] = b[4];
    hth[i] = b[5];
    for (int c = 0; c < 3; ++c) {
      eps3[3 * i + c] = b[6 + c];
      sig3[3 * i + c] = b[9 + c];
    }
  }
  return SIMP_OK;
}

}  // extern "C"
