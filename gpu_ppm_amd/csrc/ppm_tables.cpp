// Readers of the reference's table formats; see ppm_tables.hpp.
#include "ppm_tables.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>

namespace ppm {

namespace {

bool slurp(const std::string& path, std::string& buf) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(n > 0 ? (size_t)n : 0);
  size_t got = n > 0 ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
  std::fclose(f);
  buf.resize(got);
  return true;
}

// Sequential whitespace-separated number reader.
struct Scanner {
  const char* p;
  const char* end;
  explicit Scanner(const std::string& s) : p(s.data()), end(s.data() + s.size()) {}
  bool next(double& v) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    if (p >= end) return false;
    const char* q = p;
    if (*q == '+') ++q;  // from_chars rejects a leading '+'
    auto r = std::from_chars(q, end, v);
    if (r.ec != std::errc()) {
      // values beyond double range or odd spellings: fall back to strtod
      char* e = nullptr;
      v = std::strtod(p, &e);
      if (e == p) return false;
      p = e;
      return true;
    }
    p = r.ptr;
    return true;
  }
};

ppm_status read_values(const std::string& path, std::vector<double>& out, size_t count,
                       std::string& err) {
  std::string buf;
  if (!slurp(path, buf)) {
    err = "could not open " + path;
    return PPM_ERR_IO;
  }
  Scanner s(buf);
  out.resize(count);
  for (size_t k = 0; k < count; ++k) {
    if (!s.next(out[k])) {
      err = "short or malformed file " + path + " (value " + std::to_string(k) + ")";
      return PPM_ERR_FORMAT;
    }
  }
  return PPM_OK;
}

ppm_status read_axis(const std::string& path, double& lo, double& hi, int& n, std::string& err) {
  std::vector<double> v;  // "min max N" (cpu/hunt.hpp:141-153)
  ppm_status st = read_values(path, v, 3, err);
  if (st) return st;
  lo = v[0];
  hi = v[1];
  n = (int)v[2];
  if (n < 2 || !(hi > lo)) {
    err = path + ": axis needs n >= 2 and max > min";
    return PPM_ERR_FORMAT;
  }
  return PPM_OK;
}

bool is_close(double a, double b) {  // cpu/nsx.hpp:150-156
  const double atol = 10 * std::numeric_limits<double>::epsilon(), rtol = 1e-8;
  return std::fabs(a - b) <= atol + rtol * std::max(std::fabs(a), std::fabs(b));
}

double axis_node(double lo, double hi, int n, int i) { return lo + (hi - lo) * i / (n - 1); }

constexpr uint64_t kCacheMagic = 0x3158534e4d5050ull;  // "PPMNSX1"

struct CacheHeader {
  uint64_t magic;
  uint64_t src_size;
  int64_t src_mtime;
  ppm_nsx_config cfg;
};

bool same_cfg(const ppm_nsx_config& a, const ppm_nsx_config& b) {
  return std::memcmp(&a, &b, sizeof a) == 0;
}

bool load_cache(const std::string& cache, const struct stat& src, const ppm_nsx_config& cfg,
                NsxHost& out) {
  FILE* f = std::fopen(cache.c_str(), "rb");
  if (!f) return false;
  CacheHeader h{};
  bool ok = std::fread(&h, sizeof h, 1, f) == 1 && h.magic == kCacheMagic &&
            h.src_size == (uint64_t)src.st_size && h.src_mtime == (int64_t)src.st_mtime &&
            same_cfg(h.cfg, cfg);
  if (ok) {
    size_t n = (size_t)cfg.n_logT * cfg.n_logg * cfg.n_logEkT * cfg.n_mu;
    out.mu.resize(cfg.n_mu);
    out.data.resize(n);
    ok = std::fread(out.mu.data(), sizeof(double), cfg.n_mu, f) == (size_t)cfg.n_mu &&
         std::fread(out.data.data(), sizeof(double), n, f) == n;
  }
  std::fclose(f);
  return ok;
}

void save_cache(const std::string& cache, const struct stat& src, const NsxHost& t) {
  std::string tmp = cache + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;  // read-only location: no cache
  CacheHeader h{kCacheMagic, (uint64_t)src.st_size, (int64_t)src.st_mtime, t.cfg};
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 &&
            std::fwrite(t.mu.data(), sizeof(double), t.mu.size(), f) == t.mu.size() &&
            std::fwrite(t.data.data(), sizeof(double), t.data.size(), f) == t.data.size();
  ok = (std::fclose(f) == 0) && ok;
  if (ok)
    std::rename(tmp.c_str(), cache.c_str());
  else
    std::remove(tmp.c_str());
}

}  // namespace

ppm_nsx_config default_nsx_config() {  // nsx_H_v200804, cpu/nsx.hpp:25-37
  ppm_nsx_config c{};
  c.n_logT = 35;
  c.n_logg = 14;
  c.n_logEkT = 166;
  c.n_mu = 67;
  c.logT_min = 5.10;
  c.logT_max = 6.80;
  c.logg_min = 13.7;
  c.logg_max = 15.0;
  c.logEkT_min = -1.30;
  c.logEkT_max = 2.00;
  return c;
}

ppm_status read_lensing(const std::string& dir, LensingHost& t, std::string& err) {
  ppm_status st;
  if ((st = read_axis(dir + "/u.txt", t.u_min, t.u_max, t.n_u, err))) return st;
  if ((st = read_axis(dir + "/cos_psi.txt", t.cp_min, t.cp_max, t.n_cp, err))) return st;
  if ((st = read_axis(dir + "/cos_alpha.txt", t.ca_min, t.ca_max, t.n_ca, err))) return st;
  if ((st = read_values(dir + "/cos_alpha_of_u_cos_psi.txt", t.ca, (size_t)t.n_u * t.n_cp, err)))
    return st;
  if ((st = read_values(dir + "/lf_of_u_cos_psi.txt", t.lf, (size_t)t.n_u * t.n_cp, err))) return st;
  if ((st = read_values(dir + "/cdt_over_R_of_u_cos_alpha.txt", t.dt, (size_t)t.n_u * t.n_ca, err)))
    return st;
  return PPM_OK;
}

ppm_status read_nsx(const std::string& path, const ppm_nsx_config& cfg, NsxHost& t, std::string& err) {
  const ppm_nsx_config& c = cfg;
  if (c.n_logT < 2 || c.n_logg < 2 || c.n_logEkT < 4 || c.n_mu < 4 || c.n_mu > 80) {
    err = "NSX config: need n_logT, n_logg >= 2, n_logEkT >= 4 and 4 <= n_mu <= 80";
    return PPM_ERR_ARG;
  }
  struct stat sb {};
  if (stat(path.c_str(), &sb) != 0) {
    err = "could not open NSX file " + path;
    return PPM_ERR_IO;
  }
  t.cfg = cfg;
  const std::string cache = path + ".ppmcache";
  if (load_cache(cache, sb, cfg, t)) return PPM_OK;

  std::string buf;
  if (!slurp(path, buf)) {
    err = "could not open NSX file " + path;
    return PPM_ERR_IO;
  }
  const size_t n = (size_t)c.n_logT * c.n_logg * c.n_logEkT * c.n_mu;
  t.mu.assign(c.n_mu, 0.0);
  t.data.resize(n);
  Scanner s(buf);
  size_t idx = 0;
  char msg[256];
  for (int iT = 0; iT < c.n_logT; ++iT)
    for (int ig = 0; ig < c.n_logg; ++ig)
      for (int iE = 0; iE < c.n_logEkT; ++iE)
        for (int im = 0; im < c.n_mu; ++im, ++idx) {
          double v[5];
          for (double& x : v)
            if (!s.next(x)) {
              std::snprintf(msg, sizeof msg, "error reading NSX data at %d %d %d %d", iT, ig, iE, im);
              err = msg;
              return PPM_ERR_FORMAT;
            }
          const double logEkT = v[0], mu = v[1], logInuT3 = v[2], logT = v[3], logg = v[4];
          if (t.mu[im] == 0) {
            t.mu[im] = mu;
          } else if (!is_close(t.mu[im], mu)) {
            std::snprintf(msg, sizeof msg, "inconsistent mu at index %d: %g vs %g", im, t.mu[im], mu);
            err = msg;
            return PPM_ERR_FORMAT;
          }
          if (!is_close(axis_node(c.logEkT_min, c.logEkT_max, c.n_logEkT, iE), logEkT) ||
              !is_close(axis_node(c.logT_min, c.logT_max, c.n_logT, iT), logT) ||
              !is_close(axis_node(c.logg_min, c.logg_max, c.n_logg, ig), logg)) {
            std::snprintf(msg, sizeof msg, "NSX row %d %d %d %d is off the uniform axes", iT, ig, iE, im);
            err = msg;
            return PPM_ERR_FORMAT;
          }
          t.data[idx] = std::pow(10.0, logInuT3);
        }
  for (int i = 0; i + 1 < c.n_mu; ++i)
    if (t.mu[i] < t.mu[i + 1]) {
      std::snprintf(msg, sizeof msg, "NSX mu is not decreasing at index %d", i);
      err = msg;
      return PPM_ERR_FORMAT;
    }
  save_cache(cache, sb, t);
  return PPM_OK;
}

ppm_status read_response(const std::string& dir, ResponseHost& r, std::string& err) {
  std::vector<double> num;
  ppm_status st = read_values(dir + "/num.txt", num, 2, err);
  if (st) return st;
  r.n_ch = (int)num[0];
  r.n_e = (int)num[1];
  if (r.n_ch < 1 || r.n_e < 1) {
    err = dir + "/num.txt: bad sizes";
    return PPM_ERR_FORMAT;
  }
  if ((st = read_values(dir + "/E_obs.txt", r.e_obs, (size_t)r.n_e, err))) return st;
  if ((st = read_values(dir + "/rsp.txt", r.rsp, (size_t)r.n_ch * r.n_e, err))) return st;
  for (double e : r.e_obs)
    if (!(e > 0)) {
      err = dir + "/E_obs.txt: energies must be positive";
      return PPM_ERR_FORMAT;
    }
  return PPM_OK;
}

}  // namespace ppm
