// fr_io.cpp -- the reference's file formats, host side: MRC/CCP4 images and
// maps (iof.f:177-780: IOPEN, IREAD, IWRITE, ICLOSE, SET_MHDR) and parameter
// (par) file records (card10.f:92-300 reads them, lmain.f:913-923 writes them).
//
// MRC: a 1024-byte header (SET_MHDR's fields), NSYMBT bytes of extended
// header, then NZ sections of NY rows of NX values; modes 0 (bytes, read as
// unsigned as IREAD does), 1 (INTEGER*2) and 2 (REAL*4) are read, written
// files are mode 2 with the statistics ICLOSE puts in the header (DMIN, DMAX,
// the mean and RMS deviation from FP64 sums of the FP32 values and their FP32
// squares).  Byte-swapped files are recognised by the machine stamp.
//
// Par records: the v9 line FORMAT(I7,3F8.2,2F10.2,I8,I6,2F9.1,2F8.2,I10,
// F11.4,2F8.2) -- C PSI THETA PHI SHX SHY MAG FILM DF1 DF2 ANGAST OCC LogP
// SIGMA SCORE CHANGE -- written as Fortran's F/I editing does (right-justified,
// rounded, a field of '*' when the value does not fit).  Reading follows
// card10.f: lines up to 95 characters are pre-v9 records (FORMAT(I7,5F8.2,
// F8.0,I6,2F9.1,F8.2,F7.2)), up to 128 the fixed v9 columns without SIGMA,
// longer ones list-directed.
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/frealign_hip.h"
#include "fr_common.h"

namespace {

constexpr int kHdr = 1024;

struct MrcHeader {
  int nx, ny, nz, mode;
  int mx, my, mz;
  float cella[3];
  float dmin, dmax, dmean, rms;
  int nsymbt;
  bool swapped;
};

uint32_t bswap32(uint32_t v) {
  return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

template <typename T>
T get(const unsigned char* h, int off, bool swap) {
  uint32_t v;
  memcpy(&v, h + off, 4);
  if (swap) v = bswap32(v);
  T out;
  memcpy(&out, &v, 4);
  return out;
}

int read_header(FILE* f, MrcHeader* m) {
  unsigned char h[kHdr];
  FR_CHECK(fread(h, 1, kHdr, f) == (size_t)kHdr, "MRC header shorter than 1024 bytes");
  // machine stamp (bytes 213-216): 0x44 0x41 little endian, 0x11 0x11 big;
  // without a stamp, a mode outside 0..6 read natively means swapped data
  bool swap = false;
  if (h[212] == 0x11 && h[213] == 0x11)
    swap = true;
  else if (!(h[212] == 0x44 && (h[213] == 0x41 || h[213] == 0x44))) {
    const int mode = get<int>(h, 12, false);
    swap = mode < 0 || mode > 6;
  }
  m->swapped = swap;
  m->nx = get<int>(h, 0, swap);
  m->ny = get<int>(h, 4, swap);
  m->nz = get<int>(h, 8, swap);
  m->mode = get<int>(h, 12, swap);
  m->mx = get<int>(h, 28, swap);
  m->my = get<int>(h, 32, swap);
  m->mz = get<int>(h, 36, swap);
  for (int i = 0; i < 3; i++) m->cella[i] = get<float>(h, 40 + 4 * i, swap);
  m->dmin = get<float>(h, 76, swap);
  m->dmax = get<float>(h, 80, swap);
  m->dmean = get<float>(h, 84, swap);
  m->nsymbt = get<int>(h, 92, swap);
  m->rms = get<float>(h, 216, swap);
  FR_CHECK(m->mode >= 0 && m->mode <= 2, "MRC data mode %d not supported (0, 1, 2)", m->mode);
  FR_CHECK(m->nx > 0 && m->ny > 0 && m->nz > 0, "MRC dimensions %d x %d x %d", m->nx, m->ny,
           m->nz);
  FR_CHECK(m->nsymbt >= 0, "negative extended header size");
  return 0;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

// ---------------------------------------------------------------- par lines
constexpr long long kNintOverflow = 1LL << 40;

// Fortran F editing of one value into a field of width w with d decimals
// (NaN and infinities right-justified as NaN, Inf, -Inf; a value too wide
// for the field as w asterisks)
void put_f(std::string& s, double v, int w, int d) {
  char buf[64];
  if (isnan(v))
    snprintf(buf, sizeof buf, "%*s", w, "NaN");
  else if (isinf(v))
    snprintf(buf, sizeof buf, "%*s", w, v > 0 ? "Inf" : "-Inf");
  else
    snprintf(buf, sizeof buf, "%*.*f", w, d, v);
  if ((int)strlen(buf) > w) {
    s.append((size_t)w, '*');
    return;
  }
  s += buf;
}
void put_i(std::string& s, long long v, int w) {
  char buf[64];
  snprintf(buf, sizeof buf, "%*lld", w, v);
  if ((int)strlen(buf) > w || v == kNintOverflow) {
    s.append((size_t)w, '*');
    return;
  }
  s += buf;
}

// NINT of a REAL into a default (32-bit) INTEGER, halves away from zero; a
// value out of its range prints as a field of '*' (as flang's I editing of
// the overflowed NINT does)
long long nint(float v) {
  if (!(fabsf(v) < 2147483647.0f)) return kNintOverflow;
  return (long long)llroundf(v);
}

// a fixed-column field of the line as Fortran's Iw / Fw.d input editing
// reads it: blanks are ignored (a blank field is 0) and an F field without a
// decimal point or exponent carries d implied decimals; d < 0 for Iw
bool field(const char* line, int len, int pos, int w, int d, double* out) {
  char buf[64];
  int n = 0;
  bool point = false;
  for (int k = pos; k < pos + w && k < len; k++) {
    const char ch = line[k];
    if (ch == ' ') continue;
    if (ch == '.' || ch == 'e' || ch == 'E' || ch == 'd' || ch == 'D') point = true;
    buf[n++] = (ch == 'd' || ch == 'D') ? 'e' : ch;
  }
  buf[n] = 0;
  if (n == 0) {
    *out = 0.0;
    return true;
  }
  char* end = nullptr;
  double v = d < 0 ? (double)strtoll(buf, &end, 10) : strtod(buf, &end);
  if (!end || *end != 0) return false;
  if (d > 0 && !point) v /= pow(10.0, d);
  *out = v;
  return true;
}

}  // namespace

extern "C" {

int fr_mrc_info(const char* path, int* nxyz, int* mode, float* psize, float* stats,
                int* nsymbt) {
  FR_CHECK(path && nxyz, "null argument");
  File f;
  f.f = fopen(path, "rb");
  FR_CHECK(f.f != nullptr, "cannot open %s", path);
  MrcHeader m;
  if (int r = read_header(f.f, &m)) return r;
  nxyz[0] = m.nx;
  nxyz[1] = m.ny;
  nxyz[2] = m.nz;
  if (mode) *mode = m.mode;
  // SET_MHDR: PSIZE = CELLA(1) / MX
  if (psize) *psize = m.mx != 0 ? m.cella[0] / (float)m.mx : 0.0f;
  if (stats) {
    stats[0] = m.dmin;
    stats[1] = m.dmax;
    stats[2] = m.dmean;
    stats[3] = m.rms;
  }
  if (nsymbt) *nsymbt = m.nsymbt;
  return 0;
}

int fr_mrc_read(const char* path, long long first, long long count, float* out) {
  FR_CHECK(path && out, "null argument");
  File f;
  f.f = fopen(path, "rb");
  FR_CHECK(f.f != nullptr, "cannot open %s", path);
  MrcHeader m;
  if (int r = read_header(f.f, &m)) return r;
  FR_CHECK(first >= 0 && count >= 0 && first + count <= m.nz,
           "sections %lld .. %lld outside 0 .. %d", first, first + count - 1, m.nz - 1);
  const int bytes = m.mode == 0 ? 1 : (m.mode == 1 ? 2 : 4);
  const size_t sec = (size_t)m.nx * m.ny;
  const long long off = kHdr + (long long)m.nsymbt + (long long)first * (long long)sec * bytes;
  FR_CHECK(fseeko(f.f, off, SEEK_SET) == 0, "seek failed in %s", path);
  std::vector<unsigned char> raw(sec * bytes);
  for (long long s = 0; s < count; s++) {
    FR_CHECK(fread(raw.data(), 1, raw.size(), f.f) == raw.size(), "%s is shorter than its header",
             path);
    float* o = out + (size_t)s * sec;
    if (m.mode == 0) {  // iof.f:529-537: a byte read as 0..255
      for (size_t k = 0; k < sec; k++) o[k] = (float)raw[k];
    } else if (m.mode == 1) {
      for (size_t k = 0; k < sec; k++) {
        uint16_t v;
        memcpy(&v, &raw[2 * k], 2);
        if (m.swapped) v = (uint16_t)((v >> 8) | (v << 8));
        int16_t sv;
        memcpy(&sv, &v, 2);
        o[k] = (float)sv;
      }
    } else {
      for (size_t k = 0; k < sec; k++) {
        uint32_t v;
        memcpy(&v, &raw[4 * k], 4);
        if (m.swapped) v = bswap32(v);
        memcpy(&o[k], &v, 4);
      }
    }
  }
  return 0;
}

int fr_mrc_write(const char* path, int nx, int ny, int nz, float psize, const float* data,
                 const char* label) {
  FR_CHECK(path && data, "null argument");
  FR_CHECK(nx > 0 && ny > 0 && nz > 0, "bad dimensions");
  // IWRITE's running statistics (iof.f:631-637) and ICLOSE's (iof.f:694-699)
  float dmin = 1.0e30f, dmax = -1.0e30f;
  double mean = 0.0, rms = 0.0;
  const size_t n = (size_t)nx * ny * nz;
  for (size_t k = 0; k < n; k++) {
    const float r = data[k];
    mean = mean + (double)r;
    rms = rms + (double)(r * r);
    if (r > dmax) dmax = r;
    if (r < dmin) dmin = r;
  }
  mean = mean / ny / nz;
  mean = mean / nx;
  rms = rms / ny / nz;
  rms = rms / nx;
  rms = sqrt(rms - mean * mean);
  unsigned char h[kHdr];
  memset(h, 0, sizeof h);
  auto puti = [&](int off, int v) { memcpy(h + off, &v, 4); };
  auto putf = [&](int off, float v) { memcpy(h + off, &v, 4); };
  puti(0, nx);
  puti(4, ny);
  puti(8, nz);
  puti(12, 2);                        // MODE: REAL
  puti(16, 0);                        // NXSTART..
  puti(28, nx);                       // MX, MY, MZ
  puti(32, ny);
  puti(36, nz);
  putf(40, psize * (float)nx);        // CELLA = PSIZE * N
  putf(44, psize * (float)ny);
  putf(48, psize * (float)nz);
  putf(52, 90.0f);                    // CELLB
  putf(56, 90.0f);
  putf(60, 90.0f);
  puti(64, 1);                        // MAPC, MAPR, MAPS
  puti(68, 2);
  puti(72, 3);
  putf(76, dmin);
  putf(80, dmax);
  putf(84, (float)mean);
  puti(88, 1);                        // ISPG
  puti(92, 0);                        // NSYMBT
  memcpy(h + 208, "MAP ", 4);
  h[212] = 0x44;                      // little-endian machine stamp
  h[213] = 0x41;
  putf(216, (float)rms);
  const int nlabl = label && label[0] ? 1 : 0;
  puti(220, nlabl);
  memset(h + 224, ' ', 800);
  if (nlabl) memcpy(h + 224, label, strnlen(label, 80));
  File f;
  f.f = fopen(path, "wb");
  FR_CHECK(f.f != nullptr, "cannot create %s", path);
  FR_CHECK(fwrite(h, 1, kHdr, f.f) == (size_t)kHdr, "write failed: %s", path);
  FR_CHECK(fwrite(data, sizeof(float), n, f.f) == n, "write failed: %s", path);
  return 0;
}

// lmain.f:913-923: one v9 record from the values it prints (angles in
// degrees, shifts in A, OCC and SCORE in %, as the Fortran has computed them)
int fr_par_format(int c, float psi, float theta, float phi, float shx, float shy, float mag,
                  int film, float df1, float df2, float angast, float occ, float logp,
                  float sigma, float score, float change, char* out, int cap) {
  FR_CHECK(out != nullptr && cap >= 137, "output buffer must hold 137 characters");
  std::string s;
  put_i(s, c, 7);
  put_f(s, psi, 8, 2);
  put_f(s, theta, 8, 2);
  put_f(s, phi, 8, 2);
  put_f(s, shx, 10, 2);
  put_f(s, shy, 10, 2);
  put_i(s, nint(mag), 8);
  put_i(s, film, 6);
  put_f(s, df1, 9, 1);
  put_f(s, df2, 9, 1);
  put_f(s, angast, 8, 2);
  put_f(s, occ, 8, 2);
  put_i(s, nint(logp), 10);
  put_f(s, sigma, 11, 4);
  put_f(s, score, 8, 2);
  put_f(s, change, 8, 2);
  memcpy(out, s.c_str(), s.size() + 1);
  return 0;
}

// card10.f:136-172: one record into (C, PSI, THETA, PHI, SHX, SHY, MAG, FILM,
// DF1, DF2, ANGAST, OCC, LogP, SIGMA, SCORE) as written in the file (degrees,
// A, %).  *format: 0 pre-v9 (OCC = 100, SIGMA = 0, SCORE = the phase residual
// column), 1 v9 fixed (no SIGMA), 2 v9 list-directed.  Comment lines (first
// character 'C') return 1 and leave rec alone.
int fr_par_parse(const char* line, double* rec, int* format) {
  FR_CHECK(line && rec && format, "null argument");
  int len = (int)strlen(line);
  while (len > 0 && (line[len - 1] == '\n' || line[len - 1] == '\r')) len--;
  if (len > 0 && line[0] == 'C') return 1;
  // SLEN2: the length up to the last non-blank character
  int slen = len;
  while (slen > 0 && line[slen - 1] == ' ') slen--;
  for (int k = 0; k < 15; k++) rec[k] = 0.0;
  if (slen <= 95) {
    // FORMAT(I7,5F8.2,F8.0,I6,2F9.1,F8.2,F7.2): C PSI THETA PHI SHX SHY MAG FILM DF1 DF2 ANGAST PRES
    static const int w[12] = {7, 8, 8, 8, 8, 8, 8, 6, 9, 9, 8, 7};
    static const int dd[12] = {-1, 2, 2, 2, 2, 2, 0, -1, 1, 1, 2, 2};
    static const int slot[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 14};
    int pos = 0;
    for (int k = 0; k < 12; k++) {
      double v;
      FR_CHECK(field(line, len, pos, w[k], dd[k], &v), "par record: bad field %d: %s", k + 1,
               line);
      rec[slot[k]] = v;
      pos += w[k];
    }
    rec[11] = 100.0;  // card10.f:145-146
    rec[13] = 0.0;
    *format = 0;
    return 0;
  }
  if (slen <= 128) {
    // FORMAT(I7,3F8.2,2F10.2,I8,I6,2F9.1,2F8.2,I13,F8.2)
    static const int w[14] = {7, 8, 8, 8, 10, 10, 8, 6, 9, 9, 8, 8, 13, 8};
    static const int dd[14] = {-1, 2, 2, 2, 2, 2, -1, -1, 1, 1, 2, 2, -1, 2};
    static const int slot[14] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14};
    int pos = 0;
    for (int k = 0; k < 14; k++) {
      double v;
      FR_CHECK(field(line, len, pos, w[k], dd[k], &v), "par record: bad field %d: %s", k + 1,
               line);
      rec[slot[k]] = v;
      pos += w[k];
    }
    *format = 1;
    return 0;
  }
  // list-directed: 15 values separated by blanks or commas
  std::string t(line, (size_t)len);
  const char* p = t.c_str();
  for (int k = 0; k < 15; k++) {
    while (*p == ' ' || *p == ',' || *p == '\t') p++;
    FR_CHECK(*p != 0, "par record: %d values, 15 expected: %s", k, line);
    char* end = nullptr;
    rec[k] = strtod(p, &end);
    FR_CHECK(end != p, "par record: bad value %d: %s", k + 1, line);
    p = end;
  }
  *format = 2;
  return 0;
}

// calc_occ.f:179-217 for np particles in nc classes (arrays [class][particle]):
// each particle's occupancies renormalised to 100 %, moved by the fraction f
// towards the LogP-weighted share ON = EXP(LogP - max) A / sum * 100 (classes
// more than 10 below the maximum get 0), and its SIGMA replaced in every
// class by the occupancy-weighted mean; avg[i] is class i's average
// occupancy A(I).  FP32 throughout, in the reference's order.
int fr_calc_occ(int nc, int np, float f, const float* avg, float* occ, const float* algp,
                float* sig) {
  FR_CHECK(nc >= 1 && np >= 0 && avg && occ && algp && sig, "bad arguments");
  for (int k = 0; k < np; k++) {
    auto O = [&](int i) -> float& { return occ[(size_t)i * np + k]; };
    auto L = [&](int i) { return algp[(size_t)i * np + k]; };
    float sump = 0.0f;
    for (int i = 0; i < nc; i++) sump = sump + O(i);
    if (sump == 0.0f) {
      sump = 100.0f;
      for (int i = 0; i < nc; i++) O(i) = 100.0f / (float)nc;
    }
    for (int i = 0; i < nc; i++) O(i) = O(i) / sump * 100.0f;
    float amax = -1.0e30f;
    for (int i = 0; i < nc; i++) amax = fmaxf(L(i), amax);
    sump = 0.0f;
    for (int i = 0; i < nc; i++)
      if (amax - L(i) < 10.0f) sump = sump + expf(L(i) - amax) * avg[i];
    for (int i = 0; i < nc; i++) {
      float on = 0.0f;
      if (amax - L(i) < 10.0f) on = expf(L(i) - amax) * avg[i] / sump * 100.0f;
      O(i) = f * (on - O(i)) + O(i);
    }
    sump = 0.0f;
    for (int i = 0; i < nc; i++) sump = sump + sig[(size_t)i * np + k] * O(i) / 100.0f;
    for (int i = 0; i < nc; i++) sig[(size_t)i * np + k] = sump;
  }
  return 0;
}

}  // extern "C"
