// gd_host.cpp -- one-time host setup: geometry/coefficients/filter table and
// the initial-condition generators.  Nothing here runs inside the time step.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "gamil_dycore.h"
#include "gd_geom.hpp"

namespace gd {

std::string build_geometry(int N, int M, bool use_filter, const int32_t *cutoff,
                           int n_cutoff, Geometry &g) {
  if (N < 8 || N > 8192) return "num_lon must be in [8, 8192]";
  if (M < 5 || M > 8192) return "num_lat must be in [5, 8192]";
  if (n_cutoff < 0 || (n_cutoff > 0 && !cutoff)) return "bad cutoff table";
  g = Geometry();
  g.N = N;
  g.M = M;
  g.full_lat.assign(M, 0);  g.half_lat.assign(M - 1, 0);
  g.full_lon.assign(N, 0);  g.half_lon.assign(N, 0);
  g.full_cos.assign(M, 0);  g.full_sin.assign(M, 0);
  g.half_cos.assign(M - 1, 0); g.half_sin.assign(M - 1, 0);
  g.full_f.assign(M, 0); g.full_c.assign(M, 0); g.full_dlon.assign(M, 0); g.full_dlat.assign(M, 0);
  g.half_f.assign(M - 1, 0); g.half_c.assign(M - 1, 0);
  g.half_dlon.assign(M - 1, 0); g.half_dlat.assign(M - 1, 0);
  // mesh_mod.F90:70-110
  g.dlon = 2 * kPi / N;
  for (int i = 0; i < N; ++i) {
    g.full_lon[i] = i * g.dlon;
    g.half_lon[i] = g.full_lon[i] + 0.5 * g.dlon;
  }
  g.dlat = kPi / (M - 1);
  for (int j = 0; j < M - 1; ++j) {
    g.full_lat[j] = -0.5 * kPi + j * g.dlat;
    g.half_lat[j] = g.full_lat[j] + 0.5 * g.dlat;
  }
  g.full_lat[M - 1] = 0.5 * kPi;
  for (int j = 0; j < M - 1; ++j) {
    g.half_cos[j] = std::cos(g.half_lat[j]);
    g.half_sin[j] = std::sin(g.half_lat[j]);
  }
  for (int j = 0; j < M; ++j) {
    g.full_cos[j] = std::cos(g.full_lat[j]);
    g.full_sin[j] = std::sin(g.full_lat[j]);
  }
  g.full_cos[0] = 0.0;
  g.full_cos[M - 1] = 0.0;
  g.full_sin[0] = -1.0;
  g.full_sin[M - 1] = 1.0;
  // data_mod.F90:31-47 (curvature 0 at the poles, computed before the reset)
  for (int j = 0; j < M; ++j) {
    g.full_f[j] = 2.0 * kOmega * g.full_sin[j];
    g.full_c[j] = (j == 0 || j == M - 1) ? 0.0 : g.full_sin[j] / g.full_cos[j] / kRadius;
    g.full_dlon[j] = kRadius * g.dlon * g.full_cos[j];
    g.full_dlat[j] = kRadius * g.dlat * g.full_cos[j];
  }
  for (int j = 0; j < M - 1; ++j) {
    g.half_f[j] = 2.0 * kOmega * g.half_sin[j];
    g.half_c[j] = g.half_sin[j] / g.half_cos[j] / kRadius;
    g.half_dlon[j] = kRadius * g.dlon * g.half_cos[j];
    g.half_dlat[j] = kRadius * g.dlat * g.half_cos[j];
  }
  // reset_cos_lat_at_poles (dycore_mod.F90:159-173)
  g.full_cos[0] = g.half_cos[0] * 0.25;
  g.full_dlon[0] = kRadius * g.dlon * g.full_cos[0];
  g.full_dlat[0] = kRadius * g.dlat * g.full_cos[0];
  g.full_cos[M - 1] = g.half_cos[M - 2] * 0.25;
  g.full_dlon[M - 1] = kRadius * g.dlon * g.full_cos[M - 1];
  g.full_dlat[M - 1] = kRadius * g.dlat * g.full_cos[M - 1];

  // filter_mod.F90:35-103.  Entry k (1-based) of the cutoff table drives full
  // row 1+k / half row k from the south and full row M-k / half row M-k from
  // the north (reference 1-based rows).  Flags for the north half rows sit
  // at M-k+1, one row poleward of the mask row M-k (filter_mod.F90:56 vs
  // 94-95); the k=1 flag falls outside the half-lat array and is dropped.
  g.full_flag.assign(M, 0); g.half_flag.assign(M - 1, 0);
  g.full_cut.assign(M, -1); g.half_cut.assign(M - 1, -1);
  for (int k = 1; k <= n_cutoff; ++k) {
    int c = cutoff[k - 1];
    if (c == 0) continue;
    if (c < 0) return "zonal_tend_filter_cutoff_wavenumber entries must be >= 0";
    if (2 * c + 2 > N) return "cutoff wavenumber exceeds num_lon/2 - 1";
    // 1-based reference rows -> 0-based
    int fs = 1 + k, hs = k, fn = M - k, hn_mask = (M - 1) - k + 1, hn_flag = M - k + 1;
    if (fs > M || fn < 1 || hs > M - 1 || hn_mask < 1)
      return "cutoff table longer than the grid allows";
    auto set_max = [](int &dst, int v) { if (v > dst) dst = v; };  // union of 0/1 masks
    set_max(g.full_cut[fs - 1], c);
    set_max(g.half_cut[hs - 1], c);
    set_max(g.full_cut[fn - 1], c);
    set_max(g.half_cut[hn_mask - 1], c);
    if (use_filter) {
      g.full_flag[fs - 1] = 1;
      g.half_flag[hs - 1] = 1;
      g.full_flag[fn - 1] = 1;
      if (hn_flag <= M - 1) g.half_flag[hn_flag - 1] = 1;
    }
  }
  // A flagged row without a mask would be zeroed by the reference and then
  // divided by a zero inner product; refuse such tables.
  for (int j = 0; j < M; ++j)
    if (g.full_flag[j] && g.full_cut[j] < 0) return "flagged full row without filter mask";
  for (int j = 0; j < M - 1; ++j)
    if (g.half_flag[j] && g.half_cut[j] < 0)
      return "flagged half row without filter mask (a zero entry inside the cutoff table)";
  // pole rows are never filtered (the full-row loops skip them)
  g.full_flag[0] = 0;
  g.full_flag[M - 1] = 0;
  return "";
}

// ---------------------------------------------------------------------------
// Initial conditions
// ---------------------------------------------------------------------------

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): portable 64-bit generator.
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double normal() {  // Box-Muller, one value per call
    double u1 = 1.0 - uniform();  // (0,1]
    double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
  }
};

// Fully normalised associated Legendre functions (4-pi normalisation, no
// Condon-Shortley phase) P[n][m] at x = sin(lat), c = cos(lat), n <= nmax,
// and dP/dlat via (1-x^2) dP_nm/dx = -n x P_nm + sqrt((2n+1)(n^2-m^2)/(2n-1)) P_{n-1,m}.
struct Legendre {
  int nmax;
  std::vector<double> P, dP;  // [(n)*(nmax+1)+m]
  explicit Legendre(int nm) : nmax(nm), P((nm + 1) * (nm + 1)), dP((nm + 1) * (nm + 1)) {}
  double &p(int n, int m) { return P[n * (nmax + 1) + m]; }
  double &dp(int n, int m) { return dP[n * (nmax + 1) + m]; }
  void eval(double x, double c) {
    std::fill(P.begin(), P.end(), 0.0);
    std::fill(dP.begin(), dP.end(), 0.0);
    p(0, 0) = 1.0;
    for (int m = 1; m <= nmax; ++m)
      p(m, m) = (m == 1 ? std::sqrt(3.0) : std::sqrt((2.0 * m + 1) / (2.0 * m))) * c * p(m - 1, m - 1);
    for (int m = 0; m < nmax; ++m) p(m + 1, m) = std::sqrt(2.0 * m + 3) * x * p(m, m);
    for (int m = 0; m <= nmax; ++m)
      for (int n = m + 2; n <= nmax; ++n) {
        double a = std::sqrt((4.0 * n * n - 1) / ((double)n * n - (double)m * m));
        double b = std::sqrt(((n - 1.0) * (n - 1.0) - (double)m * m) / (4.0 * (n - 1.0) * (n - 1.0) - 1));
        p(n, m) = a * (x * p(n - 1, m) - b * p(n - 2, m));
      }
    if (c > 1e-12)
      for (int n = 1; n <= nmax; ++n)
        for (int m = 0; m <= n; ++m) {
          double t = -n * x * p(n, m);
          if (n - 1 >= m)
            t += std::sqrt((2.0 * n + 1) * ((double)n * n - (double)m * m) / (2.0 * n - 1)) * p(n - 1, m);
          dp(n, m) = t / c;  // dP/dlat = (1-x^2) dP/dx / cos(lat)
        }
  }
};

struct SHField {
  int nmax;
  std::vector<double> ac, as;  // [(n)*(nmax+1)+m]
};

SHField draw_coefficients(uint64_t seed, int nmax, double slope) {
  SplitMix64 rng(seed);
  SHField f;
  f.nmax = nmax;
  f.ac.assign((nmax + 1) * (nmax + 1), 0.0);
  f.as.assign((nmax + 1) * (nmax + 1), 0.0);
  for (int n = 1; n <= nmax; ++n) {
    double sd = std::pow((double)n, slope);
    for (int m = 0; m <= n; ++m) {
      f.ac[n * (nmax + 1) + m] = sd * rng.normal();
      f.as[n * (nmax + 1) + m] = m == 0 ? 0.0 : sd * rng.normal();
    }
  }
  return f;
}

// Evaluates the balanced field of one coefficient set, rescaled to rms:
// h' = sin^2(lat) H - mean, u = -(g/(2 Omega a)) (2 cos H + sin dH/dlat),
// v = (g/(2 Omega a)) sin H_lon / cos, all times the rms scale factor.
void balanced_field(const Geometry &g, const SHField &f, double rms, std::vector<double> &h,
                    std::vector<double> &u, std::vector<double> &v) {
  const int N = g.N, M = g.M, nmax = f.nmax;
  h.assign((size_t)M * N, 0.0);
  u.assign((size_t)M * N, 0.0);
  v.assign((size_t)(M - 1) * N, 0.0);
  Legendre L(nmax);
  std::vector<double> cm(nmax + 1), sm(nmax + 1);
  const double kw = kGravity / (2.0 * kOmega * kRadius);
  auto trig = [&](double lon) {
    for (int m = 0; m <= nmax; ++m) { cm[m] = std::cos(m * lon); sm[m] = std::sin(m * lon); }
  };
  // height at full points (true latitudes; pole rows use cos = 0)
  for (int j = 0; j < M; ++j) {
    double x = std::sin(g.full_lat[j]);
    double c = (j == 0 || j == M - 1) ? 0.0 : std::cos(g.full_lat[j]);
    L.eval(x, c);
    for (int i = 0; i < N; ++i) {
      trig(g.full_lon[i]);
      double H = 0.0;
      for (int n = 1; n <= nmax; ++n)
        for (int m = 0; m <= n; ++m)
          H += L.p(n, m) * (f.ac[n * (nmax + 1) + m] * cm[m] + f.as[n * (nmax + 1) + m] * sm[m]);
      h[(size_t)j * N + i] = x * x * H;
    }
  }
  // u at (half lon, full lat), pole rows stay 0
  for (int j = 1; j < M - 1; ++j) {
    double x = std::sin(g.full_lat[j]), c = std::cos(g.full_lat[j]);
    L.eval(x, c);
    for (int i = 0; i < N; ++i) {
      trig(g.half_lon[i]);
      double H = 0.0, Hp = 0.0;
      for (int n = 1; n <= nmax; ++n)
        for (int m = 0; m <= n; ++m) {
          double t = f.ac[n * (nmax + 1) + m] * cm[m] + f.as[n * (nmax + 1) + m] * sm[m];
          H += L.p(n, m) * t;
          Hp += L.dp(n, m) * t;
        }
      u[(size_t)j * N + i] = -kw * (2.0 * c * H + x * Hp);
    }
  }
  // v at (full lon, half lat)
  for (int j = 0; j < M - 1; ++j) {
    double x = std::sin(g.half_lat[j]), c = std::cos(g.half_lat[j]);
    L.eval(x, c);
    for (int i = 0; i < N; ++i) {
      trig(g.full_lon[i]);
      double Hl = 0.0;
      for (int n = 1; n <= nmax; ++n)
        for (int m = 1; m <= n; ++m)
          Hl += L.p(n, m) * m * (-f.ac[n * (nmax + 1) + m] * sm[m] + f.as[n * (nmax + 1) + m] * cm[m]);
      v[(size_t)j * N + i] = kw * x * Hl / c;
    }
  }
  // area-weighted mean removal and rms rescale (weights cos(lat), poles 0)
  double sw = 0, s1 = 0, s2 = 0;
  for (int j = 0; j < M; ++j) {
    double w = (j == 0 || j == M - 1) ? 0.0 : std::cos(g.full_lat[j]);
    for (int i = 0; i < N; ++i) { sw += w; s1 += w * h[(size_t)j * N + i]; }
  }
  double mean = s1 / sw;
  for (int j = 0; j < M; ++j) {
    double w = (j == 0 || j == M - 1) ? 0.0 : std::cos(g.full_lat[j]);
    for (int i = 0; i < N; ++i) { double d = h[(size_t)j * N + i] - mean; s2 += w * d * d; }
  }
  double cur = std::sqrt(s2 / sw);
  double scale = cur > 0 ? rms / cur : 0.0;
  for (auto &x : h) x = (x - mean) * scale;
  for (auto &x : u) x *= scale;
  for (auto &x : v) x *= scale;
}

void store(const std::vector<double> &src, int rows, int N, double *dst, int64_t ld) {
  if (!dst) return;
  for (int j = 0; j < rows; ++j)
    std::memcpy(dst + (size_t)j * ld, src.data() + (size_t)j * N, sizeof(double) * N);
}

}  // namespace
}  // namespace gd

extern "C" int gd_ic_balanced(int32_t N, int32_t M, const gd_ic_params *p, double *u, double *v,
                              double *gdo, double *ghs, int64_t ld) {
  using namespace gd;
  if (!p || N < 8 || M < 5 || ld < N || p->nmax < 1 || p->nmax > 512) return GD_ERR_INVALID_ARG;
  Geometry g;
  std::string err = build_geometry(N, M, false, nullptr, 0, g);
  if (!err.empty()) return GD_ERR_INVALID_ARG;
  std::vector<double> h, uu, vv;
  balanced_field(g, draw_coefficients(p->seed, p->nmax, p->slope), p->rms_height, h, uu, vv);
  if (p->pert_rms > 0 && p->pert_nmax > 0) {
    std::vector<double> h2, u2, v2;
    balanced_field(g, draw_coefficients(p->pert_seed, p->pert_nmax, p->slope), p->pert_rms, h2, u2, v2);
    for (size_t k = 0; k < h.size(); ++k) h[k] += h2[k];
    for (size_t k = 0; k < uu.size(); ++k) uu[k] += u2[k];
    for (size_t k = 0; k < vv.size(); ++k) vv[k] += v2[k];
  }
  std::vector<double> hs((size_t)M * N, 0.0);
  if (p->n_mountains > 0) {
    SplitMix64 rng(p->mountain_seed);
    for (int k = 0; k < p->n_mountains; ++k) {
      double lon0 = 2 * kPi * rng.uniform();
      double lat0 = (-60.0 + 120.0 * rng.uniform()) * kPi / 180.0;
      double h0 = 500.0 + 1500.0 * rng.uniform();
      double r0 = (10.0 + 10.0 * rng.uniform()) * kPi / 180.0;
      for (int j = 0; j < M; ++j)
        for (int i = 0; i < N; ++i) {
          double lat = g.full_lat[j], lon = g.full_lon[i];
          double cd = std::sin(lat) * std::sin(lat0) + std::cos(lat) * std::cos(lat0) * std::cos(lon - lon0);
          cd = cd > 1 ? 1 : (cd < -1 ? -1 : cd);
          double d = std::acos(cd);
          hs[(size_t)j * N + i] += h0 * std::exp(-std::log(2.0) * (d / r0) * (d / r0));
        }
    }
  }
  std::vector<double> gdv((size_t)M * N), ghsv((size_t)M * N);
  for (size_t k = 0; k < gdv.size(); ++k) {
    ghsv[k] = kGravity * hs[k];
    gdv[k] = kGravity * (p->mean_depth + h[k]) - ghsv[k];
  }
  store(uu, M, N, u, ld);
  store(vv, M - 1, N, v, ld);
  store(gdv, M, N, gdo, ld);
  store(ghsv, M, N, ghs, ld);
  return GD_OK;
}

// rossby_haurwitz_wave_test_mod.F90:33-85 (mesh cos/sin as set by mesh_init,
// i.e. cos = 0 exactly on the pole rows, before the pole reset).
extern "C" int gd_ic_rossby_haurwitz(int32_t N, int32_t M, double R, double omg, double gd0,
                                     double *u, double *v, double *gdo, double *ghs, int64_t ld) {
  using namespace gd;
  if (N < 8 || M < 5 || ld < N) return GD_ERR_INVALID_ARG;
  Geometry g;
  if (!build_geometry(N, M, false, nullptr, 0, g).empty()) return GD_ERR_INVALID_ARG;
  std::vector<double> fcos(M), fsin(g.full_sin);
  for (int j = 0; j < M; ++j) fcos[j] = (j == 0 || j == M - 1) ? 0.0 : std::cos(g.full_lat[j]);
  const double a = kRadius, Om = kOmega;
  for (int j = 0; j < M; ++j) {
    double c = fcos[j], s = fsin[j];
    for (int i = 0; i < N; ++i) {
      double lon = g.half_lon[i];
      double aa = c;
      double bb = R * std::pow(c, R - 1) * (s * s) * std::cos(R * lon);
      double cc = std::pow(c, R + 1) * std::cos(R * lon);
      if (u) u[(size_t)j * ld + i] = a * omg * (aa + bb - cc);
    }
  }
  for (int j = 0; j < M - 1; ++j) {
    double c = g.half_cos[j], s = g.half_sin[j];
    for (int i = 0; i < N; ++i) {
      double lon = g.full_lon[i];
      double aa = R * std::pow(c, R - 1) * s * std::sin(R * lon);
      if (v) v[(size_t)j * ld + i] = -a * omg * aa;
    }
  }
  for (int j = 0; j < M; ++j) {
    double c = fcos[j];
    double A = 0.5 * omg * (2 * Om + omg) * (c * c) +
               0.25 * (omg * omg) *
                   ((R + 1) * std::pow(c, 2 * R + 2) + (2 * (R * R) - R - 2) * std::pow(c, 2 * R) -
                    2 * (R * R) * std::pow(c, 2 * R - 2));
    double B = 2 * (Om + omg) * omg * std::pow(c, R) *
               ((R * R) + 2 * R + 2 - ((R + 1) * (R + 1)) * (c * c)) / (R + 1) / (R + 2);
    double C = 0.25 * (omg * omg) * std::pow(c, 2 * R) * ((R + 1) * (c * c) - R - 2);
    for (int i = 0; i < N; ++i) {
      double lon = g.full_lon[i];
      if (gdo) gdo[(size_t)j * ld + i] = gd0 + (a * a) * (A + B * std::cos(R * lon) + C * std::cos(2 * R * lon));
      if (ghs) ghs[(size_t)j * ld + i] = 0.0;
    }
  }
  return GD_OK;
}

// mountain_zonal_flow_test_mod.F90:28-97 with alpha = 0 and smooth_mountain
// = .false. (u on the pole rows and v are zero).
extern "C" int gd_ic_mountain_zonal_flow(int32_t N, int32_t M, double *u, double *v, double *gdo,
                                         double *ghs, int64_t ld) {
  using namespace gd;
  if (N < 8 || M < 5 || ld < N) return GD_ERR_INVALID_ARG;
  Geometry g;
  if (!build_geometry(N, M, false, nullptr, 0, g).empty()) return GD_ERR_INVALID_ARG;
  const double u0 = 20.0, gd0 = 5960.0 * kGravity, lon0 = kPi * 1.5, lat0 = kPi / 6.0;
  const double ghs0 = 2000.0 * kGravity, R = kPi / 9.0;
  std::vector<double> hsv((size_t)M * N);
  for (int j = 0; j < M; ++j)
    for (int i = 0; i < N; ++i) {
      double dl = std::fabs(g.full_lon[i] - lon0);
      dl = std::fmin(dl, 2 * kPi - dl);
      double d = std::fmin(R, std::sqrt(dl * dl + (g.full_lat[j] - lat0) * (g.full_lat[j] - lat0)));
      hsv[(size_t)j * N + i] = ghs0 * (1.0 - d / R);
    }
  for (int j = 0; j < M; ++j) {
    double c = (j == 0 || j == M - 1) ? 0.0 : std::cos(g.full_lat[j]);
    double s = g.full_sin[j];
    for (int i = 0; i < N; ++i) {
      if (u) u[(size_t)j * ld + i] = (j == 0 || j == M - 1) ? 0.0 : u0 * (c * 1.0 + 0.0);
      double t = s * 1.0 - 0.0;
      if (gdo) gdo[(size_t)j * ld + i] = gd0 - (kRadius * kOmega * u0 + (u0 * u0) * 0.5) * (t * t) - hsv[(size_t)j * N + i];
      if (ghs) ghs[(size_t)j * ld + i] = hsv[(size_t)j * N + i];
    }
  }
  if (v)
    for (int j = 0; j < M - 1; ++j)
      for (int i = 0; i < N; ++i) v[(size_t)j * ld + i] = 0.0;
  return GD_OK;
}

// steady_geostrophic_flow_test_mod.F90:20-62 (Williamson test 2, alpha = 0):
// u0 = 2 pi a / 12 days on the interior full rows (0 on the pole rows), v from
// the reference's expression -u0 cos(lon) sin(alpha) (it names cos(lon)
// sin_lon), i.e. a signed zero, gd from the balanced height; mesh cos/sin as
// set by mesh_init (cos = 0 on the pole rows).
extern "C" int gd_ic_steady_geostrophic(int32_t N, int32_t M, double *u, double *v, double *gdo,
                                        double *ghs, int64_t ld) {
  using namespace gd;
  if (N < 8 || M < 5 || ld < N) return GD_ERR_INVALID_ARG;
  Geometry g;
  if (!build_geometry(N, M, false, nullptr, 0, g).empty()) return GD_ERR_INVALID_ARG;
  const double alpha = 0.0, ca = std::cos(alpha), sa = std::sin(alpha);
  const double u0 = 2 * kPi * kRadius / (12 * 86400), gd0 = 2.94e4;
  for (int j = 0; j < M; ++j) {
    const bool pole = j == 0 || j == M - 1;
    const double c = pole ? 0.0 : std::cos(g.full_lat[j]), s = g.full_sin[j];
    for (int i = 0; i < N; ++i) {
      if (u) u[(size_t)j * ld + i] = pole ? 0.0 : u0 * (c * ca + std::cos(g.half_lon[i]) * s * sa);
      const double t = s * ca - std::cos(g.full_lon[i]) * c * sa;
      if (gdo) gdo[(size_t)j * ld + i] = gd0 - (kRadius * kOmega * u0 + (u0 * u0) * 0.5) * (t * t);
      if (ghs) ghs[(size_t)j * ld + i] = 0.0;
    }
  }
  if (v)
    for (int j = 0; j < M - 1; ++j)
      for (int i = 0; i < N; ++i) v[(size_t)j * ld + i] = -u0 * std::cos(g.full_lon[i]) * sa;
  return GD_OK;
}

namespace {

// Zonal wind of the barotropically unstable jet (jet_zonal_flow_test_mod.F90:
// 94-105): u_max / e_n exp(1 / ((lat - lat0)(lat - lat1))) between lat0 = pi/7
// and lat1 = pi/2 - lat0, zero elsewhere.
struct Jet {
  double u_max = 80.0, lat0 = gd::kPi / 7.0, lat1 = gd::kPi / 2.0 - gd::kPi / 7.0;
  double en = std::exp(-4.0 / ((lat1 - lat0) * (lat1 - lat0)));
  double u(double lat) const {
    if (lat <= lat0 || lat >= lat1) return 0.0;
    return u_max / en * std::exp(1 / (lat - lat0) / (lat - lat1));
  }
  // gradient-wind balance d(gh)/dlat = -a u (f + tan(lat) u / a)  (84-92)
  double integrand(double lat) const {
    const double uu = u(lat), f = 2 * gd::kOmega * std::sin(lat);
    return gd::kRadius * uu * (f + std::tan(lat) / gd::kRadius * uu);
  }
};

// n-point Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on the
// Legendre recurrence (the textbook construction).
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(gd::kPi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

}  // namespace

// jet_zonal_flow_test_mod.F90:29-82 (Galewsky et al. 2004): u from the jet
// profile on every full row, v = 0, gd = gh0 - int_{-pi/2}^{lat} (gradient
// wind) on each row (gh0 = g 1e4 on the south pole row), plus, when perturb is
// nonzero, the Gaussian bump ghd cos(lat) exp(-((lon - pi)/alpha)^2)
// exp(-((pi/4 - lat)/beta)^2) with ghd = 120 g, alpha = 1/3, beta = 1/15.
// The integral is evaluated to round-off with composite 16-point
// Gauss-Legendre over the jet (the integrand vanishes outside (lat0, lat1));
// the reference calls QUADPACK qags with epsrel = 1e-3, whose result agrees
// with it to about its own error estimate.
extern "C" int gd_ic_jet_zonal_flow(int32_t N, int32_t M, int32_t perturb, double *u, double *v,
                                    double *gdo, double *ghs, int64_t ld) {
  using namespace gd;
  if (N < 8 || M < 5 || ld < N) return GD_ERR_INVALID_ARG;
  Geometry g;
  if (!build_geometry(N, M, false, nullptr, 0, g).empty()) return GD_ERR_INVALID_ARG;
  const Jet jet;
  const double gh0 = kGravity * 1.0e4, ghd = kGravity * 120, lat2 = kPi / 4.0;
  const double alpha = 1.0 / 3.0, beta = 1.0 / 15.0;
  std::vector<double> gx, gw;
  gauss_legendre(16, gx, gw);
  const int panels = 256;
  std::vector<double> row(M);
  for (int j = 0; j < M; ++j) {
    const double lat = g.full_lat[j];
    double integral = 0.0;
    const double b = std::fmin(lat, jet.lat1);
    if (j > 0 && b > jet.lat0) {
      const double hw = 0.5 * (b - jet.lat0) / panels;
      for (int p = 0; p < panels; ++p) {
        const double mid = jet.lat0 + (2 * p + 1) * hw;
        double acc = 0.0;
        for (int q = 0; q < 16; ++q) acc += gw[q] * jet.integrand(mid + hw * gx[q]);
        integral += acc * hw;
      }
    }
    row[j] = j == 0 ? gh0 : gh0 - integral;
  }
  for (int j = 0; j < M; ++j) {
    const double lat = g.full_lat[j];
    const double uj = jet.u(lat);
    for (int i = 0; i < N; ++i) {
      if (u) u[(size_t)j * ld + i] = uj;
      double gdv = row[j];
      if (perturb) {
        const double x = (g.full_lon[i] - kPi) / alpha, y = (lat2 - lat) / beta;
        gdv = gdv + ghd * std::cos(lat) * std::exp(-(x * x)) * std::exp(-(y * y));
      }
      if (gdo) gdo[(size_t)j * ld + i] = gdv;
      if (ghs) ghs[(size_t)j * ld + i] = 0.0;
    }
  }
  if (v)
    for (int j = 0; j < M - 1; ++j)
      for (int i = 0; i < N; ++i) v[(size_t)j * ld + i] = 0.0;
  return GD_OK;
}

namespace {

// Analytic linear waves of Shamir & Paldor (2016, QJRMS), as coded in
// shallow_water_waves_test_mod.F90:140-305: (n, k) = (5, 10), gH = 5e4,
// their own Earth constants.
struct ShamirPaldor {
  double omega = 7.29212e-5, g = 9.80616, a = 6371220.0, H0 = 5.0e3;
  double pi = 3.14159265358979323;
  int n = 5, k = 10;
  double sigma = 0.5 + std::sqrt(0.25 + 100.0);

  // Phase speed: the roots of k^2 C^3 - E_n C - 2 g H0 Omega / a^2 = 0 by
  // Cardano's formula (140-178); Rossby (0) = -min |C_j|, EIG (1) = max,
  // WIG (-1) = min.
  double phase_speed(int flag) const {
    const double En = g * H0 / (a * a) * ((n + sigma) * (n + sigma));
    const double k2 = (double)(k * k), k4 = (double)(k * k * k * k);
    const double D0 = 3.0 * k2 * En, D4 = -54.0 * k4 * g * H0 * omega / (a * a);
    double Cj[3];
    for (int j = 1; j <= 3; ++j) {
      std::complex<double> d = std::sqrt(std::complex<double>(D4 * D4 - 4.0 * (D0 * D0 * D0), 0.0));
      d = std::pow(0.5 * (D4 + d), 1.0 / 3.0);
      d = d * std::exp(std::complex<double>(0.0, 2.0 * pi * j * (1.0 / 3.0)));
      Cj[j - 1] = (-(1.0 / 3.0) / k2 * (d + D0 / d)).real();
    }
    if (flag == 0) return -std::fmin(std::fabs(Cj[0]), std::fmin(std::fabs(Cj[1]), std::fabs(Cj[2])));
    if (flag == 1) return std::fmax(Cj[0], std::fmax(Cj[1], Cj[2]));
    return std::fmin(Cj[0], std::fmin(Cj[1], Cj[2]));
  }

  // Latitude amplitudes (u, v, h) at one latitude (getPsi 180-208,
  // getAmplitudes 213-258).
  void amplitudes(double lat, int flag, double &ut, double &vt, double &ht) const {
    const double C = phase_speed(flag), o2 = 2.0 * omega;
    const double a3 = sigma * (sigma + 1) * (sigma + 2), a4 = a3 * (sigma + 3), a5 = a4 * (sigma + 4);
    const double s = std::sin(lat), c = std::cos(lat), t = std::tan(lat);
    const double s2 = s * s, s4 = s2 * s2;
    const double C5 = (4.0 * a5 * s4 - 20.0 * a4 * s2 + 15.0 * a3) * s / 15.0;
    const double C5p = (4.0 * a5 * s4 - 12.0 * a4 * s2 + 3.0 * a3) * c / 3.0;
    const double amp = 1.0e-8, cs = std::pow(c, sigma);
    const double psi = amp * cs * C5, dpsi = amp * cs * (-sigma * t * C5 + C5p);
    if (flag == 0) {
      const double Kp = (g * H0 + (a * a) * (C * C) * (c * c)) / (C * c);
      const double Km = (g * H0 - (a * a) * (C * C) * (c * c)) / (C * c);
      vt = std::sqrt(o2 * std::fabs(Km) / (c * c)) * psi;
      ht = std::sqrt(o2 * std::fabs(Km) * (a * a) * (H0 * H0)) / Km * (dpsi + t * (0.5 * Kp / Km - o2 / C) * psi);
    } else {
      const double W1 = ((k * C) * (k * C) - (o2 * o2) * s2) / (C * c);
      const double W2 = ((k * C) * (k * C) - (o2 * o2) * s2) / (o2 * (c * c));
      vt = std::sqrt(o2 * std::fabs(W1) * g * H0 / (c * c)) / W1 * (dpsi + t * (0.5 - o2 / W2 + o2 / C) * psi);
      ht = std::sqrt(o2 * std::fabs(W1) * (a * a) * H0 / g) * psi;
    }
    ut = (o2 * s / C) * vt + (g / a / c / C) * ht;
    vt = k * vt;
  }
};

}  // namespace

// shallow_water_waves_test_mod.F90:96-136 at time 0: h = hTilde cos(k lon),
// u = uTilde cos(k lon_half), v = vTilde cos(k lon - pi/2), gd = g h + 5e4,
// ghs = 0.  As in the reference (getFields gets only the full latitudes for
// its amplitudes, 261-305), v on half row j uses the amplitude of full row j.
// u on the two pole rows, |u| < 1e-140 in the reference, is stored as 0.
// flag: 0 Rossby (the reference's choice), 1 eastward / -1 westward
// inertia-gravity wave.
extern "C" int gd_ic_shallow_water_waves(int32_t N, int32_t M, int32_t flag, double *u, double *v,
                                         double *gdo, double *ghs, int64_t ld, double *phase_speed) {
  using namespace gd;
  if (N < 8 || M < 5 || ld < N || flag < -1 || flag > 1) return GD_ERR_INVALID_ARG;
  Geometry g;
  if (!build_geometry(N, M, false, nullptr, 0, g).empty()) return GD_ERR_INVALID_ARG;
  const ShamirPaldor sp;
  std::vector<double> ut(M), vt(M), ht(M);
  for (int j = 0; j < M; ++j) sp.amplitudes(g.full_lat[j], flag, ut[j], vt[j], ht[j]);
  const double k = sp.k;
  for (int j = 0; j < M; ++j)
    for (int i = 0; i < N; ++i) {
      const double h = ht[j] * std::cos(k * g.full_lon[i]);
      if (gdo) gdo[(size_t)j * ld + i] = sp.g * h + 5.0e4;
      if (ghs) ghs[(size_t)j * ld + i] = 0.0;
      if (u) u[(size_t)j * ld + i] = (j == 0 || j == M - 1) ? 0.0 : ut[j] * std::cos(k * g.half_lon[i]);
    }
  if (v)
    for (int j = 0; j < M - 1; ++j)
      for (int i = 0; i < N; ++i) v[(size_t)j * ld + i] = vt[j] * std::cos(k * g.full_lon[i] - 0.5 * sp.pi);
  if (phase_speed) *phase_speed = sp.phase_speed(flag);
  return GD_OK;
}

extern "C" int gd_cutoff_from_lat0(int32_t N, int32_t M, double lat0_deg, int32_t *cutoff, int32_t cap) {
  using namespace gd;
  if (N < 8 || M < 5 || !(lat0_deg > 0.0) || !(lat0_deg < 90.0)) return -1;
  const double dlat = kPi / (M - 1);
  const double c0 = std::cos(lat0_deg * kPi / 180.0);
  int n = 0;
  for (int k = 1; k < M; ++k) {
    const double lat = -0.5 * kPi + k * dlat;  // full row 1+k from the south pole
    if (std::fabs(lat * 180.0 / kPi) <= lat0_deg + 1e-9) break;  // rows strictly poleward of lat0
    int c = (int)std::floor(N / 2.0 * std::cos(lat) / c0);
    c = std::max(1, std::min(c, N / 2 - 1));
    if (cutoff && n < cap) cutoff[n] = c;
    ++n;
  }
  return n;
}
