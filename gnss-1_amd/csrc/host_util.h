// host_util.h — host-only helpers shared by the engine (compiled with g++, no device code).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gnss.h"

namespace gnss {

// exp(-2 pi i m / n), m = 0..n-1, computed in long double and rounded to double.
void twiddle_table(int n, std::vector<double>& re_im);

// One nonzero spectral coefficient of the initial condition.
struct HostIcEntry {
  int ix, iy, iz, field;
  double re, im;
};

// Tanh jet + seeded divergence-free noise in spectral space, FP64, masked by the 2/3 rule.
// Entries are in the order the values were formed; a mode may appear once per field.
void build_initial_condition(const gnss_params& p, std::vector<HostIcEntry>& out);

// 2/3-rule mask factors: keep iff 9 * ((mx ax)^2 + (my ay)^2 + (mz az)^2) <= g2.
void mask_factors(int nx, int ny, int nz, int64_t& ax, int64_t& ay, int64_t& az, int64_t& g2);

}  // namespace gnss
