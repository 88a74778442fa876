// ppm_response: builds tables/ism_inst/<setting>/{num,E_obs,rsp}.txt from an
// instrument's redistribution matrix (RMF), effective area (ARF) and an
// interstellar transmission curve, so that real instrument files drop in.
// Restates c_ism_inst.py:30-101:
//   - RMF text: 3 header lines, the last column holds the matrix in
//     energy-major order (c_ism_inst.py:30-37); ARF text: 3 header lines,
//     columns E_lo E_hi area (c_ism_inst.py:33);
//   - trim to channels [ch_lo, ch_hi) and energies [e_lo, e_hi)
//     (c_ism_inst.py:9-12, 43-46), response = area x RMF (c_ism_inst.py:58);
//   - E_obs log-spaced between the first lower and the last upper edge,
//     ends exact (np.geomspace, c_ism_inst.py:61);
//   - C[j][i] = integral over fine bin j of the not-a-knot cubic spline
//     through the unit vector e_i on E_obs (scipy CubicSpline default,
//     c_ism_inst.py:75-82);
//   - ISM: not-a-knot cubic spline of column 2 over column 0 at the fine bin
//     centres, to the 5th power (c_ism_inst.py:85-89);
//   - rsp = (response x frac) @ C, |x| < 1e-300 set to 0 (c_ism_inst.py:96-98).
//
// usage: ppm_response --rmf FILE --arf FILE --ism FILE --out DIR
//                     [--setting std|high|ultra | --n-eobs N]
//                     [--e-range LO HI] [--ch-range LO HI]
// Defaults as the reference: setting std (128 energies), energies [0, 1800),
// channels [30, 300).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

[[noreturn]] void die(const std::string& msg) {
  std::cerr << "ppm_response: " << msg << "\n";
  std::exit(EXIT_FAILURE);
}

// Rows of whitespace-separated numbers after `skip` header lines.
std::vector<std::vector<double>> load_rows(const std::string& path, int skip) {
  std::ifstream f(path);
  if (!f) die("cannot read " + path);
  std::vector<std::vector<double>> rows;
  std::string line;
  for (int k = 0; k < skip && std::getline(f, line); ++k) {
  }
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::vector<double> r;
    double v;
    while (is >> v) r.push_back(v);
    if (!is.eof()) die("non-numeric field in " + path);
    if (!r.empty()) rows.push_back(std::move(r));
  }
  return rows;
}

// Piecewise cubic on knots x: on [x_k, x_{k+1}], y = c0 + c1 t + c2 t^2 + c3 t^3, t = x - x_k.
struct Spline {
  std::vector<double> x, c0, c1, c2, c3;

  int interval(double v) const {  // the boundary pieces extrapolate
    int lo = 0, hi = (int)x.size() - 2;
    while (lo < hi) {
      const int m = (lo + hi + 1) / 2;
      if (v >= x[m]) lo = m; else hi = m - 1;
    }
    return lo;
  }
  double eval(double v) const {
    const int k = interval(v);
    const double t = v - x[k];
    return c0[k] + t * (c1[k] + t * (c2[k] + t * c3[k]));
  }
  double antideriv(int k, double t) const {  // integral of piece k from x_k to x_k + t
    return t * (c0[k] + t * (c1[k] / 2 + t * (c2[k] / 3 + t * c3[k] / 4)));
  }
  double integrate(double a, double b) const {  // a <= b
    const int ka = interval(a), kb = interval(b);
    if (ka == kb) return antideriv(ka, b - x[ka]) - antideriv(ka, a - x[ka]);
    double s = antideriv(ka, x[ka + 1] - x[ka]) - antideriv(ka, a - x[ka]);
    for (int k = ka + 1; k < kb; ++k) s += antideriv(k, x[k + 1] - x[k]);
    return s + antideriv(kb, b - x[kb]);
  }
};

// Not-a-knot cubic spline through (x, y), n >= 4: first derivatives s from
// the continuity of the second derivative at the interior knots and of the
// third derivative at x_1 and x_{n-2}, the system scipy's CubicSpline solves.
Spline not_a_knot(const std::vector<double>& x, const std::vector<double>& y) {
  const int n = (int)x.size();
  if (n < 4) die("a not-a-knot spline needs at least 4 knots");
  std::vector<double> dx(n - 1), sl(n - 1);
  for (int k = 0; k < n - 1; ++k) {
    dx[k] = x[k + 1] - x[k];
    if (!(dx[k] > 0)) die("spline knots must increase strictly");
    sl[k] = (y[k + 1] - y[k]) / dx[k];
  }
  // tridiagonal rows: a[k] s[k-1] + b[k] s[k] + c[k] s[k+1] = r[k]
  std::vector<double> a(n, 0.0), b(n, 0.0), c(n, 0.0), r(n, 0.0);
  for (int k = 1; k < n - 1; ++k) {
    a[k] = dx[k];
    b[k] = 2 * (dx[k - 1] + dx[k]);
    c[k] = dx[k - 1];
    r[k] = 3 * (dx[k] * sl[k - 1] + dx[k - 1] * sl[k]);
  }
  {
    const double d = x[2] - x[0];
    b[0] = dx[1];
    c[0] = d;
    r[0] = ((dx[0] + 2 * d) * dx[1] * sl[0] + dx[0] * dx[0] * sl[1]) / d;
    const double e = x[n - 1] - x[n - 3];
    a[n - 1] = e;
    b[n - 1] = dx[n - 3];
    r[n - 1] = (dx[n - 2] * dx[n - 2] * sl[n - 3] + (2 * e + dx[n - 2]) * dx[n - 3] * sl[n - 2]) / e;
  }
  // Thomas algorithm (the end rows keep the system tridiagonal)
  for (int k = 1; k < n; ++k) {
    const double m = a[k] / b[k - 1];
    b[k] -= m * c[k - 1];
    r[k] -= m * r[k - 1];
  }
  std::vector<double> s(n);
  s[n - 1] = r[n - 1] / b[n - 1];
  for (int k = n - 2; k >= 0; --k) s[k] = (r[k] - c[k] * s[k + 1]) / b[k];
  Spline sp;
  sp.x = x;
  sp.c0.resize(n - 1);
  sp.c1.resize(n - 1);
  sp.c2.resize(n - 1);
  sp.c3.resize(n - 1);
  for (int k = 0; k < n - 1; ++k) {
    const double t = (s[k] + s[k + 1] - 2 * sl[k]) / dx[k];
    sp.c0[k] = y[k];
    sp.c1[k] = s[k];
    sp.c2[k] = (sl[k] - s[k]) / dx[k] - t;
    sp.c3[k] = t / dx[k];
  }
  return sp;
}

void write_file(const std::string& path, const std::string& text) {
  std::ofstream f(path);
  if (!f || !(f << text)) die("cannot write " + path);
}

}  // namespace

int main(int argc, char** argv) {
  std::string rmf_path, arf_path, ism_path, out;
  int n_eobs = 128, e_lo = 0, e_hi = 1800, ch_lo = 30, ch_hi = 300;
  for (int k = 1; k < argc; ++k) {
    const std::string a = argv[k];
    auto need = [&](int m) {
      if (k + m >= argc) die("missing value after " + a);
    };
    if (a == "--rmf") { need(1); rmf_path = argv[++k]; }
    else if (a == "--arf") { need(1); arf_path = argv[++k]; }
    else if (a == "--ism") { need(1); ism_path = argv[++k]; }
    else if (a == "--out") { need(1); out = argv[++k]; }
    else if (a == "--n-eobs") { need(1); n_eobs = std::atoi(argv[++k]); }
    else if (a == "--setting") {  // c_ism_inst.py:18-28
      need(1);
      const std::string s = argv[++k];
      n_eobs = s == "ultra" ? 512 : s == "high" ? 256 : s == "std" ? 128 : -1;
      if (n_eobs < 0) die("setting must be std, high or ultra");
    } else if (a == "--e-range") { need(2); e_lo = std::atoi(argv[++k]); e_hi = std::atoi(argv[++k]); }
    else if (a == "--ch-range") { need(2); ch_lo = std::atoi(argv[++k]); ch_hi = std::atoi(argv[++k]); }
    else die("unknown option " + a);
  }
  if (rmf_path.empty() || arf_path.empty() || ism_path.empty() || out.empty())
    die("usage: ppm_response --rmf FILE --arf FILE --ism FILE --out DIR [--setting std|high|ultra | --n-eobs N] "
        "[--e-range LO HI] [--ch-range LO HI]");
  if (n_eobs < 4) die("at least 4 observed energies");

  // ---- inputs (c_ism_inst.py:30-46)
  const auto rmf_rows = load_rows(rmf_path, 3);
  const auto arf = load_rows(arf_path, 3);
  const auto ism = load_rows(ism_path, 0);
  const int n_e_all = (int)arf.size();
  if (n_e_all < 1 || rmf_rows.size() % n_e_all) die("RMF entries are not a multiple of the ARF rows");
  const int n_ch_all = (int)(rmf_rows.size() / n_e_all);
  for (const auto& r : arf)
    if (r.size() < 3) die("ARF rows need E_lo E_hi area");
  e_hi = std::min(e_hi, n_e_all);
  ch_hi = std::min(ch_hi, n_ch_all);
  if (!(0 <= e_lo && e_lo < e_hi && 0 <= ch_lo && ch_lo < ch_hi)) die("empty trim range");
  const int n_e = e_hi - e_lo, n_ch = ch_hi - ch_lo;
  // rmf.reshape((N_E, -1)).T: entry (e, ch) is row e * n_ch_all + ch
  std::vector<double> resp((size_t)n_ch * n_e);
  std::vector<double> E_lo(n_e), E_hi(n_e), E_mid(n_e);
  for (int j = 0; j < n_e; ++j) {
    const auto& ar = arf[e_lo + j];
    E_lo[j] = ar[0];
    E_hi[j] = ar[1];
    E_mid[j] = (ar[0] + ar[1]) * 0.5;
    for (int c = 0; c < n_ch; ++c)
      resp[(size_t)c * n_e + j] = ar[2] * rmf_rows[(size_t)(e_lo + j) * n_ch_all + ch_lo + c].back();
  }

  // ---- observed energies (np.geomspace, ends exact)
  std::vector<double> E_obs(n_eobs);
  const double l0 = std::log10(E_lo[0]), l1 = std::log10(E_hi[n_e - 1]), step = (l1 - l0) / (n_eobs - 1);
  for (int k = 0; k < n_eobs; ++k) E_obs[k] = std::pow(10.0, l0 + k * step);
  E_obs[0] = E_lo[0];
  E_obs[n_eobs - 1] = E_hi[n_e - 1];

  // ---- spline integration matrix C [n_e][n_eobs] (c_ism_inst.py:75-82)
  std::vector<double> Cm((size_t)n_e * n_eobs);
  std::vector<double> unit(n_eobs, 0.0);
  for (int i = 0; i < n_eobs; ++i) {
    unit[i] = 1.0;
    const Spline cs = not_a_knot(E_obs, unit);
    unit[i] = 0.0;
    for (int j = 0; j < n_e; ++j) Cm[(size_t)j * n_eobs + i] = cs.integrate(E_lo[j], E_hi[j]);
  }

  // ---- ISM transmission^5 at the fine bin centres (c_ism_inst.py:85-89)
  std::vector<double> ism_e, ism_f;
  for (const auto& r : ism) {
    if (r.size() < 3) die("ISM rows need at least 3 columns");
    ism_e.push_back(r[0]);
    ism_f.push_back(r[2]);
  }
  const Spline is = not_a_knot(ism_e, ism_f);
  std::vector<double> frac(n_e);
  for (int j = 0; j < n_e; ++j) frac[j] = std::pow(is.eval(E_mid[j]), 5);

  // ---- rsp = (response x frac) @ C (c_ism_inst.py:96-98)
  std::vector<double> rsp((size_t)n_ch * n_eobs, 0.0);
  for (int c = 0; c < n_ch; ++c)
    for (int j = 0; j < n_e; ++j) {
      const double v = resp[(size_t)c * n_e + j] * frac[j];
      if (v == 0.0) continue;
      const double* row = &Cm[(size_t)j * n_eobs];
      double* o = &rsp[(size_t)c * n_eobs];
      for (int i = 0; i < n_eobs; ++i) o[i] += v * row[i];
    }
  for (double& v : rsp)
    if (std::fabs(v) < 1e-300) v = 0.0;

  // ---- outputs (np.savetxt's %.18e)
  mkdir(out.c_str(), 0755);
  char buf[64];
  std::string txt;
  for (double e : E_obs) {
    std::snprintf(buf, sizeof buf, "%.18e\n", e);
    txt += buf;
  }
  write_file(out + "/E_obs.txt", txt);
  write_file(out + "/num.txt", std::to_string(n_ch) + " " + std::to_string(n_eobs) + "\n");
  txt.clear();
  txt.reserve((size_t)n_ch * n_eobs * 26);
  for (int c = 0; c < n_ch; ++c) {
    for (int i = 0; i < n_eobs; ++i) {
      std::snprintf(buf, sizeof buf, i ? " %.18e" : "%.18e", rsp[(size_t)c * n_eobs + i]);
      txt += buf;
    }
    txt += "\n";
  }
  write_file(out + "/rsp.txt", txt);
  std::cout << "ppm_response: " << n_ch << " channels x " << n_eobs << " energies written to " << out << "\n";
  return 0;
}
