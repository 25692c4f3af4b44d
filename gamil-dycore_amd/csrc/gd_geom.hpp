// gd_geom.hpp -- host-side lat-lon C-grid geometry, coefficients and filter
// table of one configuration (FP64, computed once at gd_create).
//
// Index convention used throughout the library: 0-based.  Full latitude row
// j = 0..M-1 (reference row j+1; rows 0 and M-1 are the poles).  Half
// latitude row j = 0..M-2 lies between full rows j and j+1.  Longitude
// i = 0..N-1; half-longitude point i lies between full points i and i+1.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gd {

// params_mod.F90:5-11 (real(8) through -fdefault-real-8)
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kOmega = 2.0 * kPi / 86400.0;
constexpr double kRadius = 6.37122e6;
constexpr double kGravity = 9.80616;

struct Geometry {
  int N = 0, M = 0;
  double dlon = 0, dlat = 0;
  // mesh_mod.F90:44-114 (full_cos at the poles already reset by
  // reset_cos_lat_at_poles, dycore_mod.F90:159-173)
  std::vector<double> full_lat, half_lat, full_lon, half_lon;
  std::vector<double> full_cos, half_cos, full_sin, half_sin;
  // data_mod.F90:31-47
  std::vector<double> full_f, full_c, full_dlon, full_dlat;
  std::vector<double> half_f, half_c, half_dlon, half_dlat;
  // filter_mod.F90:35-103: flag per row, kept cutoff wavenumber c per row
  // (-1: no mask).  The mask keeps FFTPACK5 slots 0..2c+1.
  std::vector<int> full_flag, half_flag, full_cut, half_cut;
};

// Builds the mesh, coefficients (with the pole cos reset) and filter table.
// Returns an empty string on success, else the reason the config is rejected.
std::string build_geometry(int N, int M, bool use_filter, const int32_t *cutoff,
                           int n_cutoff, Geometry &g);

// Number of FFTPACK5 real slots kept by cutoff c: indices 1..2c+2 -> 2c+2,
// capped at N.
inline int kept_slots(int N, int c) {
  int k = 2 * c + 2;
  return k > N ? N : k;
}

}  // namespace gd
