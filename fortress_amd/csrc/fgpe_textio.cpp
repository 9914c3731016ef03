// fgpe_textio.hip -- the text files of cgpe.x at scale: the field dumps
// written with the reference's FORMATs, byte for byte as Fortran formatted
// output writes them, and the initial_sol reader, both on all host threads.
//
// At 256^3 spin-2 a 3D dump is 16.8 M records of 161 bytes (2.7 GB); the
// Fortran runtime formats it on one thread.  Here rows are formatted in
// blocks by a pool of threads and written in order.
//
// Fw.d editing (what the reference's F descriptors produce with the
// runtime's default rounding): the exact binary value rounded to d decimals,
// ties to even; a minus sign for negative values, -0.0 and values that round
// to zero included; the leading zero of |v| < 1 dropped only when the field
// is otherwise one character too narrow; w asterisks when it does not fit;
// Inf/NaN right-justified.
//
// Host code only, built with the host C++ compiler: the math library calls
// (hypot, atan2, sqrt, cexp) are the C library's, as the Fortran runtime's.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fgpe.h"

namespace fgpe_impl {
int fail(int code, const char* fmt, ...);  // fgpe.hip: sets the message of fgpe_last_error
}
using fgpe_impl::fail;

namespace {

constexpr uint64_t POW10[19] = {1ULL,
                                10ULL,
                                100ULL,
                                1000ULL,
                                10000ULL,
                                100000ULL,
                                1000000ULL,
                                10000000ULL,
                                100000000ULL,
                                1000000000ULL,
                                10000000000ULL,
                                100000000000ULL,
                                1000000000000ULL,
                                10000000000000ULL,
                                100000000000000ULL,
                                1000000000000000ULL,
                                10000000000000000ULL,
                                100000000000000000ULL,
                                1000000000000000000ULL};

// Appends the Fw.d field of v to out (w <= 40, d <= 18).
void fmt_f(std::string& out, double v, int w, int d) {
  char field[64];
  const bool neg = std::signbit(v);
  if (!std::isfinite(v)) {
    const char* t = std::isnan(v) ? "NaN" : (neg ? "-Inf" : "Inf");
    const int len = (int)std::strlen(t);
    if (len > w) {
      out.append(w, '*');
    } else {
      out.append(w - len, ' ');
      out.append(t);
    }
    return;
  }
  const double a = std::fabs(v);
  // a = mant * 2^e exactly, mant < 2^53
  int e = 0;
  const double m = std::frexp(a, &e);
  const uint64_t mant = (uint64_t)std::ldexp(m, 53);
  e -= 53;
  // q = round(a * 10^d), ties to even, in 128-bit integers
  const unsigned __int128 num = (unsigned __int128)mant * POW10[d];
  unsigned __int128 q = 0;
  bool overflow = false;
  if (mant == 0) {
    q = 0;
  } else if (e >= 0) {
    if (e > 36) overflow = true;
    else q = num << e;
  } else {
    const int s = -e;
    if (s >= 127) {
      q = 0;  // below 2^53 * 2^-127: far under half a unit of the last decimal
    } else {
      q = num >> s;
      const unsigned __int128 rem = num & ((((unsigned __int128)1) << s) - 1);
      const unsigned __int128 half = ((unsigned __int128)1) << (s - 1);
      if (rem > half || (rem == half && (q & 1))) q += 1;
    }
  }
  if (overflow || q >= (((unsigned __int128)1) << 100)) {
    out.append(w, '*');
    return;
  }
  char dig[48];
  int nd = 0;
  do {
    dig[nd++] = (char)('0' + (int)(q % 10));
    q /= 10;
  } while (q != 0);
  while (nd < d + 1) dig[nd++] = '0';  // at least one digit before the point
  const int intd = nd - d;
  int len = (neg ? 1 : 0) + intd + 1 + d;
  bool zero_lead = false;
  if (len > w) {
    // the optional leading zero goes first
    if (intd == 1 && dig[nd - 1] == '0' && len - 1 <= w) {
      zero_lead = true;
      --len;
    } else {
      out.append(w, '*');
      return;
    }
  }
  int p = 0;
  for (int i = 0; i < w - len; ++i) field[p++] = ' ';
  if (neg) field[p++] = '-';
  for (int i = nd - 1; i >= d; --i) {
    if (zero_lead && i == nd - 1) continue;
    field[p++] = dig[i];
  }
  field[p++] = '.';
  for (int i = d - 1; i >= 0; --i) field[p++] = dig[i];
  out.append(field, p);
}

// layout of one dump kind: leading coordinate fields and value fields
struct Kind {
  int ncoord;      // 0 (3D), 1 (x) or 2 (x, y / x, z)
  int cw, cd;      // coordinate field
  int vw, vd;      // value fields
  bool blank;      // empty record after each x
  int slice;       // 0 full grid, 1 z = NZ/2+1 plane, 2 y = NY/2+1 plane
};

bool kind_of(int k, Kind& o) {
  switch (k) {
    case FGPE_DUMP_1D_TMP: o = {1, 11, 5, 14, 6, false, 0}; return true;
    case FGPE_DUMP_1D_SOL: o = {1, 11, 5, 18, 10, false, 0}; return true;
    case FGPE_DUMP_2D: o = {2, 8, 2, 10, 5, true, 0}; return true;
    case FGPE_DUMP_3D: o = {0, 0, 0, 16, 10, false, 0}; return true;
    case FGPE_DUMP_3D_XY: o = {2, 10, 4, 14, 8, true, 1}; return true;
    case FGPE_DUMP_3D_XZ: o = {2, 10, 4, 14, 8, true, 2}; return true;
  }
  return false;
}

int pool_size(int threads) {
  if (threads > 0) return threads;
  const unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 64u) : 1;
}

// Runs f(t) for t in [0, nt) on nt threads.
template <typename F> void parallel(int nt, F f) {
  if (nt <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t) th.emplace_back(f, t);
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" int fgpe_format_f(double v, int32_t w, int32_t d, char* out, size_t cap) {
  if (!out || w < 1 || w > 40 || d < 0 || d > 18 || d >= w || cap < (size_t)w + 1)
    return fail(FGPE_EINVAL, "bad F edit descriptor or buffer");
  std::string s;
  fmt_f(s, v, w, d);
  std::memcpy(out, s.data(), s.size());
  out[s.size()] = '\0';
  return 0;
}

extern "C" int fgpe_write_dump(const char* path, int32_t kind, const int32_t* n, int32_t ms, const double* phi,
                               const double* x, const double* y, const double* z, int32_t threads) {
  Kind k;
  if (!path || !n || !phi || !kind_of(kind, k)) return fail(FGPE_EINVAL, "bad dump arguments");
  if (ms != 3 && ms != 5) return fail(FGPE_EINVAL, "MS must be 3 or 5");
  const long nx = n[0], ny = n[1], nz = n[2];
  if (nx < 1 || ny < 1 || nz < 1) return fail(FGPE_EINVAL, "bad grid");
  if (k.ncoord >= 1 && !x) return fail(FGPE_EINVAL, "the dump needs x");
  if (k.ncoord == 2 && ((k.slice == 2 && !z) || (k.slice != 2 && !y))) return fail(FGPE_EINVAL, "the dump needs y/z");
  const long npts = nx * ny * nz;
  // records: index r -> (i, j, k) in the reference's loop order (i outer,
  // then j, then k); slices fix one axis
  const long nj = k.slice == 2 ? nz : ny;           // second loop of a 2D-like dump
  const long nk = (k.slice == 0 && kind == FGPE_DUMP_3D) ? nz : 1;
  const long rows = k.slice ? nx * nj : (kind == FGPE_DUMP_3D ? npts : nx * ny);
  const long per_x = k.slice ? nj : (kind == FGPE_DUMP_3D ? ny * nz : ny);
  FILE* f = fopen(path, "wb");
  if (!f) return fail(FGPE_EINVAL, "cannot write '%s'", path);
  const int nt = pool_size(threads);
  const long block = 1L << 18;  // records per formatting block and thread
  std::vector<std::string> buf(nt);
  int rc = 0;
  for (long r0 = 0; r0 < rows && !rc; r0 += block * nt) {
    parallel(nt, [&](int t) {
      std::string& s = buf[t];
      s.clear();
      const long a = r0 + (long)t * block, b = std::min(rows, a + block);
      if (a >= b) return;
      s.reserve((size_t)(b - a) * (size_t)(k.ncoord * k.cw + 2 * ms * k.vw + 2));
      for (long r = a; r < b; ++r) {
        const long i = r / per_x, rest = r % per_x;
        long jj, kk;
        if (k.slice == 1) {
          jj = rest;
          kk = nz / 2;
        } else if (k.slice == 2) {
          jj = ny / 2;
          kk = rest;
        } else {
          jj = rest / nk;
          kk = rest % nk;
        }
        if (k.ncoord >= 1) fmt_f(s, x[i], k.cw, k.cd);
        if (k.ncoord == 2) fmt_f(s, k.slice == 2 ? z[kk] : y[jj], k.cw, k.cd);
        const long pt = i + nx * (jj + ny * kk);
        for (int m = 0; m < ms; ++m) {
          const double re = phi[2 * (pt + npts * m)], im = phi[2 * (pt + npts * m) + 1];
          const double ab = std::hypot(re, im);
          fmt_f(s, ab * ab, k.vw, k.vd);
          fmt_f(s, std::atan2(im, re), k.vw, k.vd);
        }
        s.push_back('\n');
        if (k.blank && rest == per_x - 1) s.push_back('\n');
      }
    });
    for (int t = 0; t < nt && !rc; ++t)
      if (!buf[t].empty() && fwrite(buf[t].data(), 1, buf[t].size(), f) != buf[t].size())
        rc = fail(FGPE_EINVAL, "short write to '%s'", path);
  }
  if (fclose(f) != 0 && !rc) rc = fail(FGPE_EINVAL, "cannot close '%s'", path);
  return rc;
}

// ------------------------------------------------------------ initial_sol
namespace {

inline bool is_sep(char c) { return c == ' ' || c == '\t' || c == ',' || c == '\r'; }

// Parses the values of one record [p, e) into v (at most want); returns the
// count, or -1 on a token that is not a plain number (repeat counts, '/',
// quoted text: forms this reader does not take).
int parse_record(const char* p, const char* e, double* v, int want) {
  int got = 0;
  char tok[64];
  while (p < e) {
    while (p < e && is_sep(*p)) ++p;
    if (p >= e) break;
    const char* s = p;
    while (p < e && !is_sep(*p)) ++p;
    const long len = p - s;
    if (len >= (long)sizeof tok) return -1;
    for (long i = 0; i < len; ++i) tok[i] = (s[i] == 'D' || s[i] == 'd') ? 'E' : s[i];
    tok[len] = '\0';
    char* end = nullptr;
    const double x = std::strtod(tok, &end);
    if (end != tok + len) return -1;
    if (got < want) v[got] = x;
    ++got;
  }
  return got;
}

}  // namespace

extern "C" int fgpe_read_initial_sol(const char* path, int32_t dim, const int32_t* n, int32_t ms, double* phi,
                                     double* x, double* y, int32_t threads) {
  if (!path || !n || !phi || dim < 1 || dim > 3 || (ms != 3 && ms != 5)) return fail(FGPE_EINVAL, "bad arguments");
  const long nx = n[0], ny = dim >= 2 ? n[1] : 1, nz = dim >= 3 ? n[2] : 1, npts = nx * ny * nz;
  const int lead = dim == 3 ? 0 : dim;  // x (1D), x and y (2D) before the values
  const int want = lead + 2 * ms;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(FGPE_EINVAL, "cannot open '%s'", path);
  std::string text;
  {
    std::vector<char> chunk(1 << 24);
    size_t got;
    while ((got = fread(chunk.data(), 1, chunk.size(), f)) > 0) text.append(chunk.data(), got);
    fclose(f);
  }
  // one READ per record (SRC/cgpe3d.f90:391-409): record r of the file, blank
  // records skipped, holds the values of row r; the rows go in the reference's
  // order (i outer, then j, then k)
  std::vector<size_t> starts;  // non-blank records
  {
    size_t p = 0;
    const size_t L = text.size();
    while (p < L) {
      size_t q = text.find('\n', p);
      if (q == std::string::npos) q = L;
      bool blank = true;
      for (size_t i = p; i < q; ++i)
        if (!is_sep(text[i])) {
          blank = false;
          break;
        }
      if (!blank) starts.push_back(p);
      p = q + 1;
    }
  }
  if ((long)starts.size() < npts)
    return fail(FGPE_EINVAL, "'%s' has %zu data records, %ld needed", path, starts.size(), npts);
  const int nt = pool_size(threads);
  std::vector<int> bad(nt, -1);
  parallel(nt, [&](int t) {
    const long a = npts * t / nt, b = npts * (t + 1) / nt;
    double v[16];
    for (long r = a; r < b; ++r) {
      const char* s = text.data() + starts[r];
      const char* e = (const char*)std::memchr(s, '\n', text.size() - starts[r]);
      if (!e) e = text.data() + text.size();
      if (parse_record(s, e, v, want) < want) {
        if (bad[t] < 0) bad[t] = (int)std::min<long>(r, 2000000000L);
        continue;
      }
      const long i = r / (ny * nz), j = (r / nz) % ny, k = r % nz;
      const long pt = i + nx * (j + ny * k);
      // READ (1,*) X(I), Y(J), ...: the last record of each I / J sets it
      if (lead >= 1 && x && j == ny - 1) x[i] = v[0];
      if (lead == 2 && y && i == nx - 1) y[j] = v[1];
      for (int m = 0; m < ms; ++m) {
        // SQRT(rho) * EXP(CMPLX(0, theta)) as Fortran evaluates it: the
        // complex exponential (cexp), then the complex product with
        // (SQRT(rho), 0), whose zero parts fix the signs of zero results
        const double s2 = std::sqrt(v[lead + 2 * m]);
        const std::complex<double> e = std::exp(std::complex<double>(0.0, v[lead + 2 * m + 1]));
        const double zero = 0.0;
        phi[2 * (pt + npts * m)] = s2 * e.real() - zero * e.imag();
        phi[2 * (pt + npts * m) + 1] = s2 * e.imag() + zero * e.real();
      }
    }
  });
  for (int t = 0; t < nt; ++t)
    if (bad[t] >= 0)
      return fail(FGPE_EINVAL, "'%s': data record %d does not hold %d plain numbers", path, bad[t] + 1, want);
  return 0;
}
