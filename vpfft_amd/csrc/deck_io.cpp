// deck_io.cpp -- the input deck, microstructure files and the deck-driven run
// (host C++ only; the strain steps go through the C ABI of vpfft.h).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/vpfft.h"
#include "../../include/vpfft_io.h"
#include "capi_internal.h"

namespace vpfft {
namespace {

bool file_exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// Lines of whitespace-separated tokens (space, tab, CR), like
// Utilities::Tokenize with delimiters " \t\n" (Src/Parser.cpp:20-59); a last
// line without a newline is kept.
std::vector<std::vector<std::string>> tokenize(const char* text) {
  std::vector<std::vector<std::string>> lines;
  std::vector<std::string> cur;
  std::string tok;
  for (const char* p = text;; p++) {
    const char c = *p;
    if (c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\0') {
      if (!tok.empty()) cur.push_back(tok), tok.clear();
      if (c == '\n' || c == '\0') {
        lines.push_back(cur);
        cur.clear();
      }
      if (c == '\0') break;
    } else {
      tok.push_back(c);
    }
  }
  return lines;
}

bool is_number(const std::string& s) {
  char* end = nullptr;
  errno = 0;
  std::strtod(s.c_str(), &end);
  return end != s.c_str() && *end == '\0';
}

// Keyword table and the required set, Src/Config.cpp:40-98.
enum Key { kGrid, kRate, kMaxNR, kTimeStep, kNumTime, kTol, kTwoGrain, kFileInit, kCRSS, kGdot, kRateSens, kNumKeys };
const char* const kKeyNames[kNumKeys] = {"GridDimension", "InputStrainRate", "MaxNRIter",    "TimeStep",
                                         "NumTimeIter",   "NRTolerence",     "TwoGrainInit", "FileInit",
                                         "CRSS",          "GammaDotBase",    "RateSensitivity"};

int parse_deck(const char* text, const std::string& deck_dir, vpfft_deck* d) {
  std::memset(d, 0, sizeof *d);
  d->init_method = VPFFT_INIT_NONE;
  bool seen[kNumKeys] = {};
  const auto lines = tokenize(text);
  for (size_t i = 0; i < lines.size(); i++) {
    const auto& t = lines[i];
    if (t.empty() || t[0][0] == '#') continue;  // Src/Config.cpp:106-113
    int key = 0;
    while (key < kNumKeys && t[0] != kKeyNames[key]) key++;
    if (key == kNumKeys)  // Src/Config.cpp:225-230
      return set_error(VPFFT_E_INVALID, "deck line %zu: syntax not recognized: keyword %s", i + 1, t[0].c_str());
    seen[key] = true;
    auto need = [&](size_t n) {
      if (t.size() < n + 1)
        return set_error(VPFFT_E_INVALID, "deck line %zu: %s expects %zu parameter(s)", i + 1, kKeyNames[key], n);
      for (size_t k = 1; k <= n; k++)
        if (!is_number(t[k]))
          return set_error(VPFFT_E_INVALID, "deck line %zu: %s: '%s' is not a number", i + 1, kKeyNames[key],
                           t[k].c_str());
      return 0;
    };
    int rc = 0;
    switch (key) {
      case kGrid:  // Src/Config.cpp:128-135
        if ((rc = need(3))) return rc;
        d->num_x = std::atoi(t[1].c_str());
        d->num_y = std::atoi(t[2].c_str());
        d->num_z = std::atoi(t[3].c_str());
        break;
      case kRate:  // three continuation lines, Src/Config.cpp:137-151
        for (int r = 0; r < 3; r++) {
          const size_t j = i + 1 + r;
          if (j >= lines.size() || lines[j].size() < 3 || !is_number(lines[j][0]) || !is_number(lines[j][1]) ||
              !is_number(lines[j][2]))
            return set_error(VPFFT_E_INVALID,
                             "deck line %zu: InputStrainRate expects three following lines of three numbers",
                             i + 1);
          for (int c = 0; c < 3; c++) d->input_strain_rate[r][c] = std::atof(lines[j][c].c_str());
        }
        i += 3;
        break;
      case kMaxNR:
        if ((rc = need(1))) return rc;
        d->max_nr_iter = std::atoi(t[1].c_str());
        break;
      case kTimeStep:  // a float member, Src/Config.h:40
        if ((rc = need(1))) return rc;
        d->time_step = (double)(float)std::atof(t[1].c_str());
        break;
      case kNumTime:
        if ((rc = need(1))) return rc;
        d->num_time_iter = std::atoi(t[1].c_str());
        break;
      case kTol:  // a float member, Src/Config.h:39
        if ((rc = need(1))) return rc;
        d->nr_tolerance = (double)(float)std::atof(t[1].c_str());
        break;
      case kTwoGrain:
        d->init_method = VPFFT_INIT_TWO_GRAIN;
        break;
      case kFileInit: {  // Src/Config.cpp:175-189
        if (t.size() < 2) return set_error(VPFFT_E_INVALID, "deck line %zu: FileInit expects a file name", i + 1);
        std::string f = t[1];
        if (!file_exists(f) && f[0] != '/' && !deck_dir.empty() && file_exists(deck_dir + "/" + f))
          f = deck_dir + "/" + f;
        if (!file_exists(f)) return set_error(VPFFT_E_INVALID, "FileInit %s not found", t[1].c_str());
        if (f.size() >= sizeof d->microstructure_file)
          return set_error(VPFFT_E_INVALID, "FileInit path too long");
        std::memcpy(d->microstructure_file, f.c_str(), f.size() + 1);
        d->init_method = VPFFT_INIT_FILE;
        break;
      }
      case kCRSS:
      case kGdot:
      case kRateSens:  // Src/Config.cpp:190-215
        if ((rc = need(VPFFT_NSLIP))) return rc;
        for (int k = 0; k < VPFFT_NSLIP; k++) {
          const char* s = t[k + 1].c_str();
          if (key == kCRSS) d->crss[k] = std::atof(s);
          if (key == kGdot) d->gamma_dot_base[k] = std::atof(s);
          if (key == kRateSens) d->rate_sensitivity[k] = std::atoi(s);
        }
        break;
    }
  }
  // required keywords, Src/Config.cpp:236-249
  std::string missing;
  for (int k = 0; k < kNumKeys; k++)
    if (!seen[k] && k != kTwoGrain && k != kFileInit) missing += std::string(missing.empty() ? "" : ", ") + kKeyNames[k];
  if (!missing.empty()) return set_error(VPFFT_E_INVALID, "deck: missing parameter(s): %s", missing.c_str());
  if (d->init_method == VPFFT_INIT_NONE)
    return set_error(VPFFT_E_INVALID, "deck: an initialization scheme (FileInit or TwoGrainInit) is required");
  if (d->num_x < 1 || d->num_y < 1 || d->num_z < 1) return set_error(VPFFT_E_INVALID, "deck: bad GridDimension");
  return VPFFT_OK;
}

int read_text(const char* path, std::string* out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(VPFFT_E_INVALID, "cannot open %s", path);
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out->append(buf, n);
  std::fclose(f);
  return VPFFT_OK;
}

// Whitespace token stream over a file, for the large per-voxel files.
class TokenReader {
 public:
  explicit TokenReader(FILE* f) : f_(f) {}
  // false at end of file
  bool next(std::string* tok) {
    tok->clear();
    for (;;) {
      if (pos_ == len_) {
        len_ = std::fread(buf_, 1, sizeof buf_, f_);
        pos_ = 0;
        if (len_ == 0) return !tok->empty();
      }
      const char c = buf_[pos_++];
      if (c == ' ' || c == '\t' || c == '\r' || c == '\n') {
        if (!tok->empty()) return true;
      } else {
        tok->push_back(c);
      }
    }
  }
  bool number(double* v) {
    if (!next(&tok_)) return false;
    char* end = nullptr;
    *v = std::strtod(tok_.c_str(), &end);
    return end != tok_.c_str() && *end == '\0';
  }
  bool integer(long* v) {
    if (!next(&tok_)) return false;
    char* end = nullptr;
    *v = std::strtol(tok_.c_str(), &end, 10);
    return end != tok_.c_str() && *end == '\0';
  }

 private:
  FILE* f_;
  char buf_[1 << 16];
  size_t pos_ = 0, len_ = 0;
  std::string tok_;
};

// SMatrix3x3::BuildActiveEulerMatrix (Src/LinearAlgebra.cpp:161-181) on
// DEGREE_TO_RADIAN angles (Src/LinearAlgebra.h:21,28): float cosf / sinf of
// the double angle, products in double.
void euler_deg_matrix(double phi_deg, double theta_deg, double psi_deg, double m[9]) {
  const double kPi = 3.14159265358979323846;
  const double r = kPi / 180.0;
  const double phi = phi_deg * r, theta = theta_deg * r, psi = psi_deg * r;
  const double cp = cosf((float)phi), ct = cosf((float)theta), cs = cosf((float)psi);
  const double sp = sinf((float)phi), st = sinf((float)theta), ss = sinf((float)psi);
  m[0] = cp * cs - sp * ct * ss;
  m[3] = sp * cs + cp * ct * ss;
  m[6] = st * ss;
  m[1] = -cp * ss - sp * ct * cs;
  m[4] = -sp * ss + cp * ct * cs;
  m[7] = st * cs;
  m[2] = sp * st;
  m[5] = -cp * st;
  m[8] = ct;
}

// EigenRep(const SMatrix3x3&): v_a = sum_ij M_ij B_a,ij in row-major order
// (Src/LinearAlgebra.h:184-248, 307-312; InnerProduct Src/LinearAlgebra.cpp:812-820).
void eigenrep(const double M[3][3], double v[5]) {
  const double s2 = 1.0 / std::sqrt(2.0), s6 = 1.0 / std::sqrt(6.0);
  const double B[5][9] = {{-s2, 0, 0, 0, s2, 0, 0, 0, 0},
                          {-s6, 0, 0, 0, -s6, 0, 0, 0, 2 * s6},
                          {0, 0, 0, 0, 0, s2, 0, s2, 0},
                          {0, 0, s2, 0, 0, 0, s2, 0, 0},
                          {0, s2, 0, s2, 0, 0, 0, 0, 0}};
  for (int a = 0; a < 5; a++) {
    double r = 0;
    for (int e = 0; e < 9; e++) r += M[e / 3][e % 3] * B[a][e];
    v[a] = r;
  }
}

}  // namespace
}  // namespace vpfft

using namespace vpfft;

extern "C" {

int vpfft_deck_parse(const char* text, vpfft_deck* out) {
  if (!text || !out) return set_error(VPFFT_E_INVALID, "null argument");
  return parse_deck(text, "", out);
}

int vpfft_deck_read(const char* path, vpfft_deck* out) {
  if (!path || !out) return set_error(VPFFT_E_INVALID, "null argument");
  std::string text;
  int rc = read_text(path, &text);
  if (rc) return rc;
  if (text.empty()) return set_error(VPFFT_E_INVALID, "unable to read config file %s", path);
  std::string dir(path);
  const size_t slash = dir.rfind('/');
  dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
  return parse_deck(text.c_str(), dir, out);
}

int vpfft_deck_material(const vpfft_deck* deck, vpfft_material* out) {
  if (!deck || !out) return set_error(VPFFT_E_INVALID, "null argument");
  std::memset(out, 0, sizeof *out);
  for (int k = 0; k < VPFFT_NSLIP; k++) {  // Voce defaults, Src/MaterialGrid.cpp:329-336
    out->gamma_dot_base[k] = deck->gamma_dot_base[k];
    out->rate_sensitivity[k] = deck->rate_sensitivity[k];
    out->tau0[k] = 11;
    out->tau1[k] = 11;
    out->theta0[k] = 430;
    out->theta1[k] = 110;
    for (int j = 0; j < VPFFT_NSLIP; j++) out->hardening[k * VPFFT_NSLIP + j] = 1;
  }
  return VPFFT_OK;
}

int vpfft_deck_strain_rate(const vpfft_deck* deck, double edot_out[5]) {
  if (!deck || !edot_out) return set_error(VPFFT_E_INVALID, "null argument");
  eigenrep(deck->input_strain_rate, edot_out);
  return VPFFT_OK;
}

int vpfft_euler_deg_to_matrix(const double* euler_deg, long n, double* orientation_out) {
  if (n < 0 || (n > 0 && (!euler_deg || !orientation_out))) return set_error(VPFFT_E_INVALID, "null argument");
  for (long i = 0; i < n; i++)
    euler_deg_matrix(euler_deg[3 * i], euler_deg[3 * i + 1], euler_deg[3 * i + 2], orientation_out + 9 * i);
  return VPFFT_OK;
}

int vpfft_two_grain_init(int nx, int ny, int nz, double* orientation_out, double* stress_out) {
  if (nx < 1 || ny < 1 || nz < 1 || !orientation_out) return set_error(VPFFT_E_INVALID, "bad argument");
  double left[9], right[9];
  euler_deg_matrix(-40, 0, 0, left);  // Src/UnitTester.cpp:52-53
  euler_deg_matrix(0, 40, 0, right);
  const int xh = nx / 2, yh = ny / 2;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const long v = ((long)i * ny + j) * nz + k;
        const bool l = (i >= xh && j >= yh) || (i < xh && j < yh);
        std::memcpy(orientation_out + 9 * v, l ? left : right, sizeof left);
        if (stress_out)
          for (int c = 0; c < 5; c++) stress_out[5 * v + c] = l ? 1.0 : -1.0;
      }
  return VPFFT_OK;
}

int vpfft_read_euler_file(const char* path, int nx, int ny, int nz, double* orientation_out, double* euler_out,
                          int32_t* grain_out, int32_t* phase_out) {
  if (!path || nx < 1 || ny < 1 || nz < 1 || !orientation_out) return set_error(VPFFT_E_INVALID, "bad argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(VPFFT_E_INVALID, "cannot open %s", path);
  const long n = (long)nx * ny * nz;
  std::vector<unsigned char> got(n, 0);
  TokenReader rd(f);
  int rc = VPFFT_OK;
  for (long r = 0; r < n && rc == VPFFT_OK; r++) {  // Src/UnitTester.cpp:246-262
    double e[3];
    long x, y, z, grain, phase;
    if (!rd.number(&e[0]) || !rd.number(&e[1]) || !rd.number(&e[2]) || !rd.integer(&x) || !rd.integer(&y) ||
        !rd.integer(&z) || !rd.integer(&grain) || !rd.integer(&phase)) {
      rc = set_error(VPFFT_E_INVALID, "%s: record %ld: expected 'phi1 Phi phi2 x y z grain phase'", path, r + 1);
      break;
    }
    if (x < 1 || x > nx || y < 1 || y > ny || z < 1 || z > nz) {
      rc = set_error(VPFFT_E_INVALID, "%s: record %ld: voxel (%ld %ld %ld) outside the %dx%dx%d grid", path, r + 1,
                     x, y, z, nx, ny, nz);
      break;
    }
    const long v = ((x - 1) * ny + (y - 1)) * nz + (z - 1);
    if (got[v]) {
      rc = set_error(VPFFT_E_INVALID, "%s: record %ld: voxel (%ld %ld %ld) given twice", path, r + 1, x, y, z);
      break;
    }
    got[v] = 1;
    euler_deg_matrix(e[0], e[1], e[2], orientation_out + 9 * v);
    if (euler_out)
      for (int c = 0; c < 3; c++) euler_out[3 * v + c] = e[c];
    if (grain_out) grain_out[v] = (int32_t)grain;
    if (phase_out) phase_out[v] = (int32_t)phase;
  }
  std::fclose(f);
  return rc;
}

int vpfft_write_euler_file(const char* path, int nx, int ny, int nz, const double* euler_deg, const int32_t* grain,
                           const int32_t* phase) {
  if (!path || nx < 1 || ny < 1 || nz < 1 || !euler_deg) return set_error(VPFFT_E_INVALID, "bad argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(VPFFT_E_INVALID, "cannot open %s", path);
  for (int k = 1; k <= nz; k++)  // Src/UnitTester.cpp:24-38
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const long v = ((long)(i - 1) * ny + (j - 1)) * nz + (k - 1);
        std::fprintf(f, "%.17g %.17g %.17g %d %d %d %d %d\n", euler_deg[3 * v], euler_deg[3 * v + 1],
                     euler_deg[3 * v + 2], i, j, k, grain ? (int)grain[v] : 1, phase ? (int)phase[v] : 1);
      }
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) return set_error(VPFFT_E_INVALID, "write error on %s", path);
  return VPFFT_OK;
}

int vpfft_write_material_file(const char* path, int nx, int ny, int nz, const double* orientation,
                              const double* stress) {
  if (!path || nx < 1 || ny < 1 || nz < 1 || !orientation) return set_error(VPFFT_E_INVALID, "bad argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(VPFFT_E_INVALID, "cannot open %s", path);
  const long n = (long)nx * ny * nz;
  for (long v = 0; v < n; v++) {  // ToIndex order, Src/UnitTester.cpp:125-138
    for (int c = 0; c < 9; c++) std::fprintf(f, "%.17g ", orientation[9 * v + c]);
    for (int c = 0; c < 5; c++) std::fprintf(f, "%.17g ", stress ? stress[5 * v + c] : 0.0);
    std::fputc('\n', f);
  }
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) return set_error(VPFFT_E_INVALID, "write error on %s", path);
  return VPFFT_OK;
}

int vpfft_read_material_file(const char* path, int nx, int ny, int nz, double* orientation_out, double* stress_out) {
  if (!path || nx < 1 || ny < 1 || nz < 1 || !orientation_out) return set_error(VPFFT_E_INVALID, "bad argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(VPFFT_E_INVALID, "cannot open %s", path);
  const long n = (long)nx * ny * nz;
  TokenReader rd(f);
  int rc = VPFFT_OK;
  for (long v = 0; v < n && rc == VPFFT_OK; v++) {  // Src/UnitTester.cpp:152-166
    double s[14];
    for (int c = 0; c < 14; c++)
      if (!rd.number(&s[c])) {
        rc = set_error(VPFFT_E_INVALID, "%s: voxel %ld: expected 9 orientation and 5 stress values", path, v);
        break;
      }
    if (rc) break;
    std::memcpy(orientation_out + 9 * v, s, 9 * sizeof(double));
    if (stress_out) std::memcpy(stress_out + 5 * v, s + 9, 5 * sizeof(double));
  }
  std::fclose(f);
  return rc;
}

// Count the tokens of the first non-empty line: 8 = Euler file, 14 = material file.
static int sniff_columns(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return -1;
  int cols = 0, c;
  bool in_tok = false;
  while ((c = std::fgetc(f)) != EOF) {
    if (c == '\n') {
      if (cols > 0) break;
      continue;
    }
    const bool ws = c == ' ' || c == '\t' || c == '\r';
    if (!ws && !in_tok) cols++;
    in_tok = !ws;
  }
  std::fclose(f);
  return cols;
}

int vpfft_run_deck(const char* deck_path, int precision, const char* stress_strain_path,
                   const char* material_out_path, vpfft_step_info* infos, int infos_capacity) {
  vpfft_deck d;
  int rc = vpfft_deck_read(deck_path, &d);
  if (rc) return rc;
  const long n = (long)d.num_x * d.num_y * d.num_z;
  std::vector<double> R(9 * n), S(5 * n), crss(12 * n);
  if (d.init_method == VPFFT_INIT_TWO_GRAIN) {
    rc = vpfft_two_grain_init(d.num_x, d.num_y, d.num_z, R.data(), S.data());
  } else {
    const int cols = sniff_columns(d.microstructure_file);
    if (cols == 8)
      rc = vpfft_read_euler_file(d.microstructure_file, d.num_x, d.num_y, d.num_z, R.data(), nullptr, nullptr,
                                 nullptr);
    else if (cols == 14)
      rc = vpfft_read_material_file(d.microstructure_file, d.num_x, d.num_y, d.num_z, R.data(), S.data());
    else
      rc = set_error(VPFFT_E_INVALID, "%s: %d columns; expected 8 (Euler file) or 14 (orientation + stress)",
                     d.microstructure_file, cols);
  }
  if (rc) return rc;
  for (long v = 0; v < n; v++) std::memcpy(&crss[12 * v], d.crss, sizeof d.crss);
  vpfft_material mat;
  vpfft_deck_material(&d, &mat);
  vpfft_grid_desc desc = {d.num_x, d.num_y, d.num_z, precision};
  void* g = nullptr;
  if ((rc = vpfft_grid_create(&desc, &mat, &g))) return rc;
  double edot[5];
  eigenrep(d.input_strain_rate, edot);
  std::vector<vpfft_step_info> info(d.num_time_iter > 0 ? d.num_time_iter : 1);
  rc = vpfft_grid_upload(g, R.data(), crss.data(), 0.0, S.data());
  if (!rc) rc = vpfft_grid_set_voxel_length(g, 1, 1, 1);
  if (!rc)
    rc = vpfft_run_vpfft(g, edot, d.max_nr_iter, d.num_time_iter, d.time_step, d.nr_tolerance, nullptr, info.data(),
                         stress_strain_path);
  if (!rc && material_out_path) {
    rc = vpfft_grid_download(g, VPFFT_FIELD_ORIENTATION, R.data());
    if (!rc) rc = vpfft_grid_download(g, VPFFT_FIELD_STRESS, S.data());
    if (!rc) rc = vpfft_write_material_file(material_out_path, d.num_x, d.num_y, d.num_z, R.data(), S.data());
  }
  if (!rc && infos)
    for (int s = 0; s < d.num_time_iter && s < infos_capacity; s++) infos[s] = info[s];
  vpfft_grid_destroy(g);
  return rc;
}

}  // extern "C"
