// host_util.cpp — FP64 host code of the engine: twiddle tables and the seeded initial condition.
//
// The initial condition follows the reference's initializeJet (initialize.cu:614-641): a tanh
// jet U(y) = 0.5 - 0.5 tanh(H/(4 theta)(2|y|/H - 1)) on u (initializeVelocityKernel_mgpu,
// initialize.cu:114-142) and on every scalar (initializeScalarKernel_mgpu, initialize.cu:160-176),
// plus velocity noise.  The reference's noise is a glibc-rand spectrum (generateNoise,
// initialize.cu:286-612) that cannot be reproduced portably; this build draws complex Gaussian
// coefficients from a counter-based SplitMix64 hash of (seed, component, mode), keeps integer
// shells 1 <= |m| <= kmax of the half space (Hermitian partners on the kz = 0 plane are the
// conjugates), projects them divergence-free, scales them to the requested per-component rms and
// adds them in spectral space.  The jet's spectrum is its exact DFT along y.  Everything is FP64
// and built on the host, so every precision form starts from the same numbers.
#include <math.h>
#include <stdlib.h>

#include <complex>

#include "host_util.h"

namespace gnss {

void twiddle_table(int n, std::vector<double>& t) {
  const long double pi = 3.141592653589793238462643383279502884L;
  t.resize(2 * (size_t)n);
  for (int m = 0; m < n; ++m) {
    long double a = 2.0L * pi * (long double)m / (long double)n;
    t[2 * m] = (double)cosl(a);
    t[2 * m + 1] = -(double)sinl(a);
  }
}

static int64_t gcd64(int64_t a, int64_t b) {
  while (b) {
    int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

void mask_factors(int nx, int ny, int nz, int64_t& ax, int64_t& ay, int64_t& az, int64_t& g2) {
  int64_t g = nx;
  g = g / gcd64(g, ny) * ny;
  g = g / gcd64(g, nz) * nz;
  ax = g / nx;
  ay = g / ny;
  az = g / nz;
  g2 = g * g;
}

static uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static void gauss(uint64_t seed, int comp, int mx, int my, int mz, double& gr, double& gi) {
  const uint64_t key = (uint64_t)(unsigned)comp | ((uint64_t)(unsigned)(mx + 4096) << 8) |
                       ((uint64_t)(unsigned)(my + 4096) << 24) | ((uint64_t)(unsigned)(mz + 4096) << 40);
  const uint64_t h1 = splitmix64(seed ^ splitmix64(key));
  const uint64_t h2 = splitmix64(h1);
  const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double r = sqrt(-2.0 * log(u1));
  const double a = 6.283185307179586476925286766559 * u2;
  gr = r * cos(a);
  gi = r * sin(a);
}

static bool kept(const gnss_params& p, int mx, int my, int mz) {
  int64_t ax, ay, az, g2;
  mask_factors(p.nx, p.ny, p.nz, ax, ay, az, g2);
  const int64_t a = mx * ax, b = my * ay, c = mz * az;
  return 9 * (a * a + b * b + c * c) <= g2;
}

void build_initial_condition(const gnss_params& p, std::vector<HostIcEntry>& out) {
  out.clear();
  const int nx = p.nx, ny = p.ny, nz = p.nz, S = p.n_scalars;
  const double nn = (double)nx * (double)ny * (double)nz;
  const double h = p.jet_h > 0 ? p.jet_h : 1.0;
  // jet: U depends on y only -> modes (0, my, 0): Nx Nz sum_j U(y_j) exp(-2 pi i my j / Ny)
  std::vector<double> tw;
  twiddle_table(ny, tw);
  // the colloid (last scalar with GNSS_FLAG_COLLOID) has its own thickness theta_c
  // (initializeColloidKernel_mgpu, initialize.cu:178-194)
  const bool colloid = (p.flags & GNSS_FLAG_COLLOID) && S >= 2;
  const double theta_c = p.theta_c > 0 ? p.theta_c : p.theta;
  std::vector<double> prof(ny), prof_c(ny);
  for (int j = 0; j < ny; ++j) {
    const double y = (double)j * p.ly / (double)ny;
    prof[j] = 0.5 * (1.0 + tanh((0.5 * h - fabs(y - 0.5 * p.ly)) / (2.0 * p.theta)));
    prof_c[j] = 0.5 * (1.0 + tanh((0.5 * h - fabs(y - 0.5 * p.ly)) / (2.0 * theta_c)));
  }
  // accumulation box of the noise: mx, my in [-K, K], mz in [0, K]; the jet enters at (0, my, 0)
  struct Acc {
    double re, im;
    bool set;
  };
  const int K = (p.noise_amp != 0.0 && p.noise_kmax >= 1.0) ? (int)floor(p.noise_kmax) : 0;
  const int bx = 2 * K + 1, bz = K + 1;
  std::vector<Acc> vel((size_t)3 * bx * bx * bz, Acc{0.0, 0.0, false});
  auto vidx = [&](int comp, int mx, int my, int mz) {
    return (((size_t)comp * bx + (size_t)(mx + K)) * bx + (size_t)(my + K)) * bz + (size_t)mz;
  };
  std::vector<double> jr(ny), ji(ny), cr(ny), ci(ny);
  for (int iy = 0; iy < ny; ++iy) {
    double re = 0.0, im = 0.0, rc = 0.0, ic = 0.0;
    for (int j = 0; j < ny; ++j) {
      const int m = (int)(((int64_t)iy * j) % ny);
      re += prof[j] * tw[2 * m];
      im += prof[j] * tw[2 * m + 1];
      rc += prof_c[j] * tw[2 * m];
      ic += prof_c[j] * tw[2 * m + 1];
    }
    const double sx = (double)nx * (double)nz;
    jr[iy] = re * sx;
    ji[iy] = im * sx;
    cr[iy] = rc * sx;
    ci[iy] = ic * sx;
    const int my = iy < ny / 2 ? iy : iy - ny;
    if (my >= -K && my <= K) {
      Acc& a = vel[vidx(0, 0, my, 0)];
      a.re = jr[iy];
      a.im = ji[iy];
      a.set = true;
    }
  }
  if (K > 0) {
    const double k2max = p.noise_kmax * p.noise_kmax;
    const double pi = 3.14159265358979323846;
    double energy = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      if (pass && !(energy > 0)) break;
      const double alpha = pass ? p.noise_amp / sqrt(energy / 3.0) : 0.0;
      for (int mx = -K; mx <= K; ++mx)
        for (int my = -K; my <= K; ++my)
          for (int mz = 0; mz <= K; ++mz) {
            const int m2 = mx * mx + my * my + mz * mz;
            if (m2 < 1 || (double)m2 > k2max) continue;
            if (!(mz > 0 || my > 0 || (my == 0 && mx > 0))) continue;
            if (2 * abs(mx) >= nx || 2 * abs(my) >= ny || 2 * mz >= nz) continue;
            const double kv[3] = {(2.0 * pi / p.lx) * mx, (2.0 * pi / p.ly) * my, (2.0 * pi / p.lz) * mz};
            std::complex<double> a[3];
            for (int comp = 0; comp < 3; ++comp) {
              double gr, gi;
              gauss(p.seed, comp, mx, my, mz, gr, gi);
              a[comp] = std::complex<double>(gr, gi);
            }
            const double kk = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
            // kd = k . a, then a -= k (kd / kk); written out in the same operation order as the
            // oracle's C99 complex arithmetic
            const double kdr = kv[0] * a[0].real() + kv[1] * a[1].real() + kv[2] * a[2].real();
            const double kdi = kv[0] * a[0].imag() + kv[1] * a[1].imag() + kv[2] * a[2].imag();
            const double qr = kdr / kk, qi = kdi / kk;
            for (int comp = 0; comp < 3; ++comp)
              a[comp] = std::complex<double>(a[comp].real() - kv[comp] * qr, a[comp].imag() - kv[comp] * qi);
            if (!pass) {
              for (int comp = 0; comp < 3; ++comp)
                energy += 2.0 * (a[comp].real() * a[comp].real() + a[comp].imag() * a[comp].imag());
              continue;
            }
            const double sc = nn * alpha;
            for (int comp = 0; comp < 3; ++comp) {
              const double vr = sc * a[comp].real(), vi = sc * a[comp].imag();
              Acc& x = vel[vidx(comp, mx, my, mz)];
              x.re += vr;
              x.im += vi;
              x.set = true;
              if (mz == 0) {
                Acc& y = vel[vidx(comp, -mx, -my, 0)];
                y.re += vr;
                y.im += -vi;
                y.set = true;
              }
            }
          }
    }
  }
  for (int comp = 0; comp < 3; ++comp)
    for (int mx = -K; mx <= K; ++mx)
      for (int my = -K; my <= K; ++my)
        for (int mz = 0; mz <= K; ++mz) {
          const Acc& a = vel[vidx(comp, mx, my, mz)];
          if (!a.set || !kept(p, mx, my, mz)) continue;
          out.push_back(HostIcEntry{mx >= 0 ? mx : mx + nx, my >= 0 ? my : my + ny, mz, comp, a.re, a.im});
        }
  for (int iy = 0; iy < ny; ++iy) {
    const int my = iy < ny / 2 ? iy : iy - ny;
    if ((my >= -K && my <= K) || !kept(p, 0, my, 0)) continue;
    out.push_back(HostIcEntry{0, iy, 0, 0, jr[iy], ji[iy]});
  }
  for (int s = 0; s < S; ++s)
    for (int iy = 0; iy < ny; ++iy) {
      const int my = iy < ny / 2 ? iy : iy - ny;
      if (!kept(p, 0, my, 0)) continue;
      const bool c = colloid && s == S - 1;
      out.push_back(HostIcEntry{0, iy, 0, 3 + s, c ? cr[iy] : jr[iy], c ? ci[iy] : ji[iy]});
    }
}

}  // namespace gnss
