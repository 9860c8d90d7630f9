// fd2_io.cpp -- the reference's file formats, host only.
//
//   hybridMANTIS_input.in reader   main/hybridMANTIS_penEasy.f:190-229, 284-293
//   myimage_ / detected_ writers   main/hybridMANTIS_tallyEnergyDepositionEvents.f:
//                                  418-458
//   EDE event records               tally .f:173-174 (x y z in cm, de in eV, hist)
//   console statistics block        tally .f:461-496
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "fd2.h"

namespace {

bool parse_double(const char* s, double* v) {
  char* end = nullptr;
  errno = 0;
  const double d = std::strtod(s, &end);
  if (end == s || errno) return false;
  *v = d;
  return true;
}

bool parse_int(const char* s, int32_t* v) {
  char* end = nullptr;
  errno = 0;
  const long l = std::strtol(s, &end, 10);
  if (end == s || errno || l < INT32_MIN || l > INT32_MAX) return false;
  *v = (int32_t)l;
  return true;
}

}  // namespace

extern "C" {

// Tokens, in file order: mynumhist minphotons maxphotons mynumbins detx dety
// detheight detradius detnC detnIC dettop detbulk detbeta detdmin detdmax
// detlboundx detlboundy detuboundx detuboundy detyield detpixel
// detsensorRefl rungpu machinenum.  A token is the first blank-delimited
// word of a line whose first non-blank character is not '#'.
int fd2_read_input(const char* path, fd2_params* out, char tokens[10][16]) {
  if (!path || !out) return FD2_E_INVAL;
  FILE* f = std::fopen(path, "r");
  if (!f) return FD2_E_IO;
  char tok[24][151];
  int n = 0;
  char line[4096];
  while (n < 24 && std::fgets(line, sizeof line, f)) {
    line[150] = '\0';  // the reference reads records as (a150)
    const char* s = line;
    while (*s == ' ' || *s == '\t') ++s;
    if (*s == '#' || *s == '\0' || *s == '\n' || *s == '\r') continue;
    size_t len = std::strcspn(s, " \t\r\n");
    if (len > 150) len = 150;
    std::memcpy(tok[n], s, len);
    tok[n][len] = '\0';
    ++n;
  }
  std::fclose(f);
  if (n < 24) return FD2_E_IO;
  fd2_params p;
  std::memset(&p, 0, sizeof p);
  bool ok = parse_int(tok[0], &p.mynumhist) && parse_int(tok[1], &p.minphotons) &&
            parse_int(tok[2], &p.maxphotons) && parse_int(tok[3], &p.mynumbins) &&
            parse_double(tok[4], &p.detx) && parse_double(tok[5], &p.dety) &&
            parse_double(tok[6], &p.detheight) && parse_double(tok[7], &p.detradius) &&
            parse_double(tok[8], &p.detnC) && parse_double(tok[9], &p.detnIC) &&
            parse_double(tok[10], &p.dettop) && parse_double(tok[11], &p.detbulk) &&
            parse_double(tok[12], &p.detbeta) && parse_double(tok[13], &p.detdmin) &&
            parse_double(tok[14], &p.detdmax) && parse_double(tok[15], &p.detlboundx) &&
            parse_double(tok[16], &p.detlboundy) && parse_double(tok[17], &p.detuboundx) &&
            parse_double(tok[18], &p.detuboundy) && parse_double(tok[19], &p.detyield) &&
            parse_int(tok[20], &p.detpixel) && parse_double(tok[21], &p.detsensorRefl) &&
            parse_int(tok[22], &p.rungpu) && parse_int(tok[23], &p.machinenum);
  if (!ok) return FD2_E_IO;
  *out = p;
  if (tokens) {
    // cdetz .. cmachinenum = first 10 characters of tokens 7, 8, 11, 12, 13,
    // 15, 22, 20, 21, 24 (main/hybridMANTIS_penEasy.f:284-293)
    static const int which[10] = {6, 7, 10, 11, 12, 14, 21, 19, 20, 23};
    for (int i = 0; i < 10; ++i) {
      std::memset(tokens[i], 0, 16);
      std::strncpy(tokens[i], tok[which[i]], 10);
    }
  }
  return FD2_OK;
}

int fd2_output_names(const char tokens[10][16], char* prf_name, size_t prf_len,
                     char* phs_name, size_t phs_len) {
  if (!tokens || !prf_name || !phs_name) return FD2_E_INVAL;
  char body[256];
  int k = std::snprintf(body, sizeof body, "%.10s_%.10s_%.10s_%.10s_%.10s_%.10s_%.10s_%.10s_%.10s_%.10s",
                        tokens[0], tokens[1], tokens[2], tokens[3], tokens[4],
                        tokens[5], tokens[6], tokens[7], tokens[8], tokens[9]);
  if (k < 0 || (size_t)k >= sizeof body) return FD2_E_INVAL;
  if ((size_t)std::snprintf(prf_name, prf_len, "myimage_%s.dat", body) >= prf_len) return FD2_E_INVAL;
  if ((size_t)std::snprintf(phs_name, phs_len, "detected_%s.dat", body) >= phs_len) return FD2_E_INVAL;
  return FD2_OK;
}

// PRF file: for y-pixel J = 0..pixelX (outer) and x-pixel I = 0..pixelY
// (inner) one (1x,F12.4) record of prf[I][J] / n_hist, each J block ended by
// a blank record; pixelX = int((ub-lb)/pitch) - 1 (tally .f:418, 448-453).
// Pixels past the tallied image (the reference writes int() pixels while
// tallying ceil()) print as zero.
int fd2_write_prf(const char* path, const fd2_params* p, const uint64_t* prf,
                  int32_t xdim, int32_t ydim, double n_hist) {
  if (!path || !p || !prf || xdim <= 0 || ydim <= 0 || !(n_hist > 0)) return FD2_E_INVAL;
  FILE* t = std::fopen(path, "r");
  if (t) {  // STATUS='NEW'
    std::fclose(t);
    return FD2_E_IO;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) return FD2_E_IO;
  const int pix_x = (int)((p->detuboundx - p->detlboundx) / p->detpixel) - 1;
  const int pix_y = (int)((p->detuboundy - p->detlboundy) / p->detpixel) - 1;
  const double inv_n = 1.0 / n_hist;
  for (int j = 0; j <= pix_x; ++j) {
    for (int i = 0; i <= pix_y; ++i) {
      const uint64_t c = (i < xdim && j < ydim) ? prf[(size_t)i * ydim + j] : 0;
      const double v = (double)c * inv_n;
      if (v >= 1e7)
        std::fputs(" ************\n", f);
      else
        std::fprintf(f, " %12.4f\n", v);
    }
    std::fputs("\n", f);
  }
  return std::fclose(f) == 0 ? FD2_OK : FD2_E_IO;
}

int fd2_write_phs(const char* path, const int32_t* phs, int32_t num_bins) {
  if (!path || !phs || num_bins <= 0) return FD2_E_INVAL;
  FILE* t = std::fopen(path, "r");
  if (t) {
    std::fclose(t);
    return FD2_E_IO;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) return FD2_E_IO;
  for (int i = 0; i < num_bins; ++i) {
    if (phs[i] > 99999999 || phs[i] < -9999999)
      std::fputs(" ********\n", f);
    else
      std::fprintf(f, " %8d\n", phs[i]);
  }
  return std::fclose(f) == 0 ? FD2_OK : FD2_E_IO;
}

// Fortran (1x,F12.2): a blank, then the value rounded to 2 decimals right
// aligned in 12 columns, or 12 asterisks when it does not fit
static void put_f12_2(FILE* f, double v) {
  char buf[64];
  const int len = std::snprintf(buf, sizeof buf, "%12.2f", v);
  if (len < 0 || len > 12 || !std::isfinite(v))
    std::fputs(" ************\n", f);
  else
    std::fprintf(f, " %s\n", buf);
}

int fd2_write_stats(const char* path, const double* stats, int rungpu, double penoptical_s) {
  if (!stats) return FD2_E_INVAL;
  FILE* f = path ? std::fopen(path, "w") : stdout;
  if (!f) return FD2_E_IO;
  static const char* const kArrows = ">>>>>>>>>>>>>>>>>>>>>>>>>>>>>>>>>>>>>>";
  static const char* const kLabels[6] = {"Generated", "Detected",         "Absorbed at top",
                                         "Absorbed in bulk", "Lost", "Out of Column"};
  std::fprintf(f, " \n%s\n    OPTICAL TRANSPORT STATISTICS\n \n", kArrows);
  std::fputs(rungpu == 1 ? "   !! Simulation in the GPU !!\n" : "   !! Simulation in the CPU !!\n",
             f);
  std::fputs(" \nTotal # optical photons:\n", f);
  for (int k = 0; k < 6; ++k) {  // glgen .. gloutofcol (gltheta1 is not reported)
    std::fprintf(f, "%s\n", kLabels[k]);
    put_f12_2(f, stats[k]);
  }
  std::fputs("Total GPU or CPU time (s)\n", f);
  put_f12_2(f, stats[7] * 0.001);  // glgputime is in ms
  std::fputs("Total optical real time from PENELOPE (s)\n", f);
  put_f12_2(f, penoptical_s);
  std::fprintf(f, "%s\n \n", kArrows);
  if (!path) return std::fflush(f) == 0 ? FD2_OK : FD2_E_IO;
  return std::fclose(f) == 0 ? FD2_OK : FD2_E_IO;
}

int fd2_read_ede(const char* path, fd2_event* events, int64_t max, int64_t* n) {
  if (!path || !n || (max > 0 && !events) || max < 0) return FD2_E_INVAL;
  FILE* f = std::fopen(path, "r");
  if (!f) return FD2_E_IO;
  char line[512];
  int64_t k = 0;
  int rc = FD2_OK;
  while (std::fgets(line, sizeof line, f)) {
    const char* s = line;
    while (*s == ' ' || *s == '\t') ++s;
    if (*s == '#' || *s == '\0' || *s == '\n' || *s == '\r') continue;
    double x, y, z, de;
    long h;
    size_t len = std::strcspn(line, "\r\n");
    if (len == 72) {
      // the record's fixed columns: 3 x (1x,1pe16.9), (1x,1pe12.4), i8.  A
      // history number of 8 digits touches the energy field, so the fields
      // are cut at their columns rather than at blanks.
      static const int start[6] = {0, 17, 34, 51, 64, 72};
      double v[4];
      bool ok = true;
      char field[24];
      for (int f = 0; f < 5 && ok; ++f) {
        const int w = start[f + 1] - start[f];
        std::memcpy(field, line + start[f], (size_t)w);
        field[w] = '\0';
        char* end = nullptr;
        errno = 0;
        if (f < 4) {
          v[f] = std::strtod(field, &end);
        } else {
          h = std::strtol(field, &end, 10);
        }
        if (end == field || errno) ok = false;
        while (ok && end && *end == ' ') ++end;
        if (ok && end && *end) ok = false;
      }
      if (!ok) {
        rc = FD2_E_IO;
        break;
      }
      x = v[0];
      y = v[1];
      z = v[2];
      de = v[3];
    } else if (std::sscanf(s, "%lf %lf %lf %lf %ld", &x, &y, &z, &de, &h) != 5) {
      rc = FD2_E_IO;
      break;
    }
    if (k >= max) {
      rc = FD2_E_INVAL;
      break;
    }
    fd2_event& e = events[k++];
    // cm -> um as the drivers do (main/hybridMANTIS_c_ver1_0.c:212-214)
    e.x = (float)(x * 10000.0f);
    e.y = (float)(y * 10000.0f);
    e.z = (float)(z * 10000.0f);
    e.E = (float)de;
    e.hist = (int32_t)h;
    e.N = 0;
    e.geom = 0;
    e.col_dx = 0.0f;
  }
  std::fclose(f);
  *n = k;
  return rc;
}

}  // extern "C"
