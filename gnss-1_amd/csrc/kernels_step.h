// kernels_step.h — pointwise step kernels and FP64 statistics reductions.
#pragma once
#include "fft_device.h"
#include "kernels_fft.h"

namespace gnss {

template <typename T> struct CPtrs { cx<T>* p[MAXW]; };
template <typename T> struct RPtrs { T* p[MAXW]; };

// Grid geometry and the kept box of the 2/3 rule.
struct Geom {
  int nx, ny, nz, nzh, kzs;
  int keep_x, keep_y, keep_z;   // floor(n_i / 3): the largest |m_i| the mask can keep
  long long ax, ay, az, g2;     // mask: keep iff 9 * ((mx ax)^2 + (my ay)^2 + (mz az)^2) <= g2
};

// The dealiasing rule in integers: 9 * sum_i (m_i G / n_i)^2 <= G^2, G = lcm(nx, ny, nz).
// Follows deAliasKernel_mgpu (solver.cu:17-42) with the 2/3 filter (dnsparams.h:65).
__device__ __forceinline__ bool kept(const Geom& g, int mx, int my, int mz) {
  const long long a = mx * g.ax, b = my * g.ay, c = mz * g.az;
  return 9 * (a * a + b * b + c * c) <= g.g2;
}

// A box of spectral modes: outer kx (kept-mode map when kx_keep >= 0), ky likewise, kz < nkz.
struct Box {
  int bx, by, bz;       // extents
  int kx_keep, ky_keep; // -1 = all
};
__device__ __forceinline__ void box_index(const Geom& g, const Box& b, long long i, int& ix,
                                          int& iy, int& iz) {
  iz = (int)(i % b.bz);
  long long r = i / b.bz;
  const int oy = (int)(r % b.by);
  const int ox = (int)(r / b.by);
  ix = b.kx_keep < 0 ? ox : (ox <= b.kx_keep ? ox : g.nx - (2 * b.kx_keep + 1) + ox);
  iy = b.ky_keep < 0 ? oy : (oy <= b.ky_keep ? oy : g.ny - (2 * b.ky_keep + 1) + oy);
}

// i * k * f for real k: (-k f.y, k f.x)
template <typename V, typename T> __device__ __forceinline__ V ik(T k, V f) {
  V r;
  r.x = -k * f.y;
  r.y = k * f.x;
  return r;
}

// Spectral prologue of a substage: u, v, w, omega = i k x u^ (calcOmega*Kernel_mgpu,
// solver.cu:68-114) and grad phi = i k phi^ (gradientKernel_mgpu, solver.cu:233-256); with a
// colloid (the last scalar) also c^ itself and lap Z^ = -k^2 Z^ of its carrier Z = scalar 0
// (divergenceKernel_mgpu of the gradient, solver.cu:276-293, scalarAdvection solver.cu:423).
template <typename T>
__global__ void k_prologue(Geom g, Box b, CPtrs<T> q, CPtrs<T> ws, const T* __restrict__ kxv,
                           const T* __restrict__ kyv, const T* __restrict__ kzv, int S, int colloid,
                           long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int ix, iy, iz;
  box_index(g, b, i, ix, iy, iz);
  const long long o = ((long long)ix * g.ny + iy) * g.kzs + iz;
  const T kx = kxv[ix], ky = kyv[iy], kz = kzv[iz];
  const cx<T> u = q.p[0][o], v = q.p[1][o], w = q.p[2][o];
  ws.p[0][o] = u;
  ws.p[1][o] = v;
  ws.p[2][o] = w;
  ws.p[3][o] = csub(ik(ky, w), ik(kz, v));
  ws.p[4][o] = csub(ik(kz, u), ik(kx, w));
  ws.p[5][o] = csub(ik(kx, v), ik(ky, u));
  for (int s = 0; s < S; ++s) {
    const cx<T> ph = q.p[3 + s][o];
    ws.p[6 + 3 * s][o] = ik(kx, ph);
    ws.p[7 + 3 * s][o] = ik(ky, ph);
    ws.p[8 + 3 * s][o] = ik(kz, ph);
  }
  if (colloid) {
    const T k2 = kx * kx + ky * ky + kz * kz;
    const cx<T> zh = q.p[3][o];
    ws.p[6 + 3 * S][o] = q.p[2 + S][o];
    ws.p[7 + 3 * S][o] = mk<T>(-k2 * zh.x, -k2 * zh.y);
  }
}

// Physical products: u x w (crossProductKernel_mgpu, solver.cu:139-164, with the sign of
// computeRHSKernel_mgpu folded in), u . grad phi (dotKernel_mgpu, solver.cu:342-356) and, for a
// colloid c (the last scalar), u . grad c + alpha (lap Z c + grad Z . grad c)
// (colloidAdvectionKernel_mgpu, solver.cu:374-391).
template <typename T>
__global__ void k_products(RPtrs<T> p, RPtrs<T> nl, int S, int colloid, T alpha, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T u = p.p[0][i], v = p.p[1][i], w = p.p[2][i];
  const T ox = p.p[3][i], oy = p.p[4][i], oz = p.p[5][i];
  nl.p[0][i] = v * oz - w * oy;
  nl.p[1][i] = w * ox - u * oz;
  nl.p[2][i] = u * oy - v * ox;
  for (int s = 0; s < S - colloid; ++s)
    nl.p[3 + s][i] = u * p.p[6 + 3 * s][i] + v * p.p[7 + 3 * s][i] + w * p.p[8 + 3 * s][i];
  if (colloid) {
    const int c0 = 6 + 3 * (S - 1);
    const T cx = p.p[c0][i], cy = p.p[c0 + 1][i], cz = p.p[c0 + 2][i];
    const T c = p.p[6 + 3 * S][i], lz = p.p[7 + 3 * S][i];
    nl.p[2 + S][i] = alpha * (lz * c + cx * p.p[6][i] + cy * p.p[7][i] + cz * p.p[8][i]) + (u * cx + v * cy + w * cz);
  }
}

// The integrating factor is applied as E x = x - eps x with eps = 1 - E = -expm1(-nu k^2 tau):
// for the slowly decaying modes E lies within 1e-4 of 1, and E rounded to the working precision
// would carry a relative error of up to half an ulp, the same in every substage, so that the
// state drifts by about (substages) x ulp/2 (~1e-5 after 100 FP32 steps).  eps is computed to
// its own relative precision, so 1 - eps is exact to far below an ulp.
template <typename T> __device__ __forceinline__ T dexpm1(T x);
template <> __device__ __forceinline__ double dexpm1<double>(double x) { return expm1(x); }
template <> __device__ __forceinline__ float dexpm1<float>(float x) { return expm1f(x); }
// x E for E = 1 - eps, one rounding: fma(-eps, x, x)
template <typename T> __device__ __forceinline__ T apply_if(T x, T eps) { return fma(-eps, x, x); }
// E (q + b d) = q + (b d - eps (q + b d)): the increment is formed at its own magnitude and added
// to the state once, so each substage rounds the state once.
template <typename T> __device__ __forceinline__ T if_update(T q, T b, T d, T eps) {
  const T s = fma(b, d, q);
  return q + fma(b, d, -eps * s);
}

struct RkCoef {
  double a, b, tau, dt;   // dq <- a dq + dt rhs;  q <- E (q + b dq);  dq <- E dq;  E = exp(-nu k^2 tau)
  int first, last;        // first: a = 0 (dq not read); last: dq not written
  double nu[MAXF];        // viscosity (velocity) and diffusivities (scalars)
  // batched contexts (k_rk): members nbox elements of the box apart in the launch, mstride
  // elements apart in memory, their viscosities/diffusivities in nu_m[member * MAXF + field]
  long long nbox, mstride;
  const double* nu_m;
};

// Projection + scalar sign (computeRHSKernel_mgpu, solver.cu:461-521), mask (solver.cu:17-42)
// and one low-storage RK3 substage with exact integrating factors (replaces solver.cu:547-657).
template <typename T>
__global__ void k_rk(Geom g, Box b, CPtrs<T> q, CPtrs<T> dq, CPtrs<T> nl, const T* __restrict__ kxv,
                     const T* __restrict__ kyv, const T* __restrict__ kzv, RkCoef c, int F,
                     long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int m = 0;
  if (c.nbox > 0) {
    m = (int)(i / c.nbox);
    i -= m * c.nbox;
  }
  int ix, iy, iz;
  box_index(g, b, i, ix, iy, iz);
  const long long o = m * c.mstride + ((long long)ix * g.ny + iy) * g.kzs + iz;
  const int mz = mode_of(iz, g.nz);
  if (iz >= g.nzh || !kept(g, mode_of(ix, g.nx), mode_of(iy, g.ny), mz)) {
    for (int f = 0; f < F; ++f) {
      st_s(&q.p[f][o], mk<T>(0, 0));
      if (!c.last) st_s(&dq.p[f][o], mk<T>(0, 0));
    }
    return;
  }
  const T kx = kxv[ix], ky = kyv[iy], kz = kzv[iz];
  const T k2 = kx * kx + ky * ky + kz * kz;
  cx<T> rhs[MAXF];
  if (k2 == (T)0) {
#pragma unroll
    for (int f = 0; f < MAXF; ++f) rhs[f] = mk<T>(0, 0);
  } else {
    const cx<T> n0 = ld_s(&nl.p[0][o]), n1 = ld_s(&nl.p[1][o]), n2 = ld_s(&nl.p[2][o]);
    cx<T> kn;
    kn.x = (kx * n0.x + ky * n1.x + kz * n2.x) / k2;
    kn.y = (kx * n0.y + ky * n1.y + kz * n2.y) / k2;
    rhs[0] = csub(n0, cscale(kn, kx));
    rhs[1] = csub(n1, cscale(kn, ky));
    rhs[2] = csub(n2, cscale(kn, kz));
#pragma unroll
    for (int f = 3; f < MAXF; ++f) {
      const cx<T> v = f < F ? ld_s(&nl.p[f][o]) : mk<T>(0, 0);
      rhs[f] = mk<T>(-v.x, -v.y);
    }
  }
  const T a = (T)c.a, bb = (T)c.b, dt = (T)c.dt;
  // every field's loads before the first store (which the compiler cannot move them past)
  cx<T> oq[MAXF], od[MAXF];
#pragma unroll
  for (int f = 0; f < MAXF; ++f) {
    od[f] = (f < F && !c.first) ? ld_s(&dq.p[f][o]) : mk<T>(0, 0);
    oq[f] = f < F ? ld_s(&q.p[f][o]) : mk<T>(0, 0);
  }
#pragma unroll
  for (int f = 0; f < MAXF; ++f) {
    if (f >= F) break;
    const T eps = -dexpm1<T>(-(T)(c.nu_m ? c.nu_m[m * MAXF + f] : c.nu[f]) * k2 * (T)c.tau);
    cx<T> d = cscale(rhs[f], dt);
    if (!c.first) {
      d.x = a * od[f].x + d.x;
      d.y = a * od[f].y + d.y;
    }
    cx<T> qq = oq[f];
    qq.x = if_update(qq.x, bb, d.x, eps);
    qq.y = if_update(qq.y, bb, d.y, eps);
    st_s(&q.p[f][o], qq);
    if (!c.last) st_s(&dq.p[f][o], mk<T>(apply_if(d.x, eps), apply_if(d.y, eps)));
  }
}

// Zero every mode the 2/3 rule removes (and the kz padding): deAlias (solver.cu:44-66).
template <typename T>
__global__ void k_dealias(Geom g, CPtrs<T> q, int F, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int iz = (int)(i % g.kzs);
  const long long r = i / g.kzs;
  const int iy = (int)(r % g.ny), ix = (int)(r / g.ny);
  if (iz < g.nzh && kept(g, mode_of(ix, g.nx), mode_of(iy, g.ny), mode_of(iz, g.nz))) return;
  for (int f = 0; f < F; ++f) q.p[f][i] = mk<T>(0, 0);
}

// Mask in the ABI order [ky][kx][kz < nzh].
static __global__ void k_mask_abi(Geom g, unsigned char* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int iz = (int)(i % g.nzh);
  const long long r = i / g.nzh;
  const int ix = (int)(r % g.nx), iy = (int)(r / g.nx);
  out[i] = kept(g, mode_of(ix, g.nx), mode_of(iy, g.ny), mode_of(iz, g.nz)) ? 1 : 0;
}

// k_i = (2 pi / L_i) * m_i in FP64 (waveNumber_kernel, initialize.cu:678-691); `scale` is the
// host-computed 2 pi / L_i so the product is the same IEEE operation as on the host.
static __global__ void k_wavenumbers(double* __restrict__ k, int n, int len, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) k[i] = scale * (double)mode_of(i, n);
}

template <typename T>
__global__ void k_cast_wavenumbers(const double* __restrict__ in, T* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)in[i];
}

struct IcEntry {
  long long index;  // device linear index [kx][ky][kz]
  int field;
  int pad;
  double re, im;
};

template <typename T>
__global__ void k_ic_scatter(const IcEntry* __restrict__ e, int n, CPtrs<T> q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const IcEntry x = e[i];
  q.p[x.field][x.index] = mk<T>((T)x.re, (T)x.im);
}

// ---------------------------------------------------------------------------------------------
// Statistics: deterministic FP64 reductions (fixed per-thread order + fixed tree), one block per
// kx plane; the host sums the per-plane partials in plane order.
// ---------------------------------------------------------------------------------------------
constexpr int STAT_THREADS = 256;

// Fixed-order tree sum of NQ per-thread values over the block; every thread receives all NQ
// totals in `out` (registers).  Reuses `sh`, so it starts with a barrier.
template <int NQ>
__device__ __forceinline__ void block_sum(const double (&acc)[NQ], double* sh, double (&out)[NQ]) {
  const int tid = threadIdx.x;
  __syncthreads();
  for (int qq = 0; qq < NQ; ++qq) sh[qq * STAT_THREADS + tid] = acc[qq];
  __syncthreads();
  for (int s = STAT_THREADS / 2; s > 0; s >>= 1) {
    if (tid < s)
      for (int qq = 0; qq < NQ; ++qq) sh[qq * STAT_THREADS + tid] += sh[qq * STAT_THREADS + tid + s];
    __syncthreads();
  }
  for (int qq = 0; qq < NQ; ++qq) out[qq] = sh[qq * STAT_THREADS];
}

constexpr int SPEC_NQ = 4 + 2 * 4;

// Sums per kx plane (calcVrms/Epsilon/IntegralLength/ScalarDissipation kernels,
// statistics.cu:1494-1845): [0] sum w|u|^2, [1] sum w k^2|u|^2, [2] sum w|u|^2/(2|k|),
// [3] sum w|k.u|^2, [4+2s] sum w k^2|phi|^2, [5+2s] sum_{k!=0} w|phi|^2 — all divided by NN^2.
template <typename T>
__global__ void __launch_bounds__(STAT_THREADS)
k_spec_sums(Geom g, CPtrs<T> q, const double* __restrict__ kxv, const double* __restrict__ kyv,
            const double* __restrict__ kzv, int S, double inn2, double* __restrict__ part) {
  __shared__ double sh[SPEC_NQ * STAT_THREADS];
  const int ix = blockIdx.x;
  double acc[SPEC_NQ];
  for (int qq = 0; qq < SPEC_NQ; ++qq) acc[qq] = 0.0;
  const double kx = kxv[ix];
  const long long n = (long long)g.ny * g.nzh;
  for (long long i = threadIdx.x; i < n; i += STAT_THREADS) {
    const int iy = (int)(i / g.nzh), iz = (int)(i % g.nzh);
    const long long o = ((long long)ix * g.ny + iy) * g.kzs + iz;
    const double ky = kyv[iy], kz = kzv[iz];
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double w = (iz == 0 || 2 * iz == g.nz) ? 1.0 : 2.0;
    const cx<T> u = q.p[0][o], v = q.p[1][o], ww = q.p[2][o];
    const double ux = u.x, uy = u.y, vx = v.x, vy = v.y, wx = ww.x, wy = ww.y;
    const double e = (ux * ux + uy * uy + vx * vx + vy * vy + wx * wx + wy * wy) * inn2;
    const double dr = kx * ux + ky * vx + kz * wx, di = kx * uy + ky * vy + kz * wy;
    acc[0] += w * e;
    acc[1] += w * k2 * e;
    if (k2 > 0) acc[2] += w * e / (2.0 * sqrt(k2));
    acc[3] += w * (dr * dr + di * di) * inn2;
    for (int s = 0; s < S; ++s) {
      const cx<T> ph = q.p[3 + s][o];
      const double pr = ph.x, pi = ph.y;
      const double ep = (pr * pr + pi * pi) * inn2;
      acc[4 + 2 * s] += w * k2 * ep;
      if (k2 > 0) acc[5 + 2 * s] += w * ep;
    }
  }
  double tot[SPEC_NQ];
  block_sum<SPEC_NQ>(acc, sh, tot);
  if (threadIdx.x < SPEC_NQ) part[(long long)ix * SPEC_NQ + threadIdx.x] = tot[threadIdx.x];
}

struct ShellArgs {
  int equal_box;        // 1: shell = nint(|m|) in integers
  double fx, fy, fz;    // else shell = nint(sqrt((mx fx)^2 + (my fy)^2 + (mz fz)^2))
};

__device__ __forceinline__ long long isqrt_ll(long long v) {
  long long r = (long long)sqrt((double)v);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}
__device__ __forceinline__ int shell_of(const ShellArgs& a, int mx, int my, int mz) {
  if (a.equal_box) {
    const long long m2 = (long long)mx * mx + (long long)my * my + (long long)mz * mz;
    return (int)((isqrt_ll(4 * m2) + 1) / 2);
  }
  const double x = mx * a.fx, y = my * a.fy, z = mz * a.fz;
  return (int)floor(sqrt(x * x + y * y + z * z) + 0.5);
}

// Shell spectrum of one kx plane (the reference declares it but never computes it,
// statistics.cu:1847-1934).  Thread b owns shell b: for each ky it walks the kz interval of
// that shell (the shell index is non-decreasing in kz), so every sum has a fixed order.
template <typename T>
__global__ void __launch_bounds__(STAT_THREADS)
k_spectrum(Geom g, CPtrs<T> q, int S, int nb, ShellArgs sa, double inn2, double* __restrict__ spec) {
  const int ix = blockIdx.x;
  const int mx = mode_of(ix, g.nx);
  for (int b = threadIdx.x; b < nb; b += STAT_THREADS) {
    double acc[1 + 4];
    for (int s = 0; s <= S; ++s) acc[s] = 0.0;
    for (int iy = 0; iy < g.ny; ++iy) {
      const int my = mode_of(iy, g.ny);
      if (shell_of(sa, mx, my, 0) > b) continue;
      if (shell_of(sa, mx, my, g.nz / 2) < b) continue;
      // first kz with shell >= b, first kz with shell >= b + 1
      int lo = 0, hi = g.nzh;
      while (lo < hi) { int mid = (lo + hi) >> 1; if (shell_of(sa, mx, my, mid) >= b) hi = mid; else lo = mid + 1; }
      const int z0 = lo;
      hi = g.nzh;
      while (lo < hi) { int mid = (lo + hi) >> 1; if (shell_of(sa, mx, my, mid) >= b + 1) hi = mid; else lo = mid + 1; }
      const int z1 = lo;
      const long long base = ((long long)ix * g.ny + iy) * g.kzs;
      for (int iz = z0; iz < z1; ++iz) {
        const double w = (iz == 0 || 2 * iz == g.nz) ? 0.5 : 1.0;
        const cx<T> u = q.p[0][base + iz], v = q.p[1][base + iz], ww = q.p[2][base + iz];
        const double ux = u.x, uy = u.y, vx = v.x, vy = v.y, wx = ww.x, wy = ww.y;
        acc[0] += w * (ux * ux + uy * uy + vx * vx + vy * vy + wx * wx + wy * wy) * inn2;
        for (int s = 0; s < S; ++s) {
          const cx<T> ph = q.p[3 + s][base + iz];
          const double pr = ph.x, pi = ph.y;
          acc[1 + s] += w * (pr * pr + pi * pi) * inn2;
        }
      }
    }
    for (int s = 0; s <= S; ++s) spec[((long long)ix * (1 + S) + s) * nb + b] = acc[s];
  }
}

// x-z plane averages at one y (calcYprof_kernel_2D, statistics.cu:1936-1977), two passes:
// means, then <u'u'>, <v'v'>, <w'w'>, <u'v'>, <phi'phi'>, <v'phi'>, plus the CFL and |u| maxima.
template <typename T>
__global__ void __launch_bounds__(STAT_THREADS)
k_plane_stats(Geom g, RPtrs<T> ph, int S, double idx_, double idy_, double idz_,
              double* __restrict__ prof, double* __restrict__ cmax) {
  constexpr int NQ = 4 + 2 * 4;
  __shared__ double sh[NQ * STAT_THREADS];
  __shared__ double mean[3 + 4];
  const int y = blockIdx.x;
  const int F = 3 + S;
  const long long n = (long long)g.nx * g.nz;
  const double inv = 1.0 / (double)n;
  double acc[NQ];
  for (int qq = 0; qq < NQ; ++qq) acc[qq] = 0.0;
  for (long long i = threadIdx.x; i < n; i += STAT_THREADS) {
    const int x = (int)(i / g.nz), z = (int)(i % g.nz);
    const long long o = ((long long)x * g.ny + y) * g.nz + z;
    for (int f = 0; f < F; ++f) acc[f] += (double)ph.p[f][o];
  }
  double tot[NQ];
  block_sum<NQ>(acc, sh, tot);
  if (threadIdx.x < F) mean[threadIdx.x] = tot[threadIdx.x] * inv;
  __syncthreads();
  for (int qq = 0; qq < NQ; ++qq) acc[qq] = 0.0;
  double cm = 0.0, um = 0.0;
  for (long long i = threadIdx.x; i < n; i += STAT_THREADS) {
    const int x = (int)(i / g.nz), z = (int)(i % g.nz);
    const long long o = ((long long)x * g.ny + y) * g.nz + z;
    const double u = ph.p[0][o], v = ph.p[1][o], w = ph.p[2][o];
    const double du = u - mean[0], dv = v - mean[1], dw = w - mean[2];
    acc[0] += du * du;
    acc[1] += dv * dv;
    acc[2] += dw * dw;
    acc[3] += du * dv;
    for (int s = 0; s < S; ++s) {
      const double dp = (double)ph.p[3 + s][o] - mean[3 + s];
      acc[4 + 2 * s] += dp * dp;
      acc[5 + 2 * s] += dv * dp;
    }
    const double cf = fabs(u) * idx_ + fabs(v) * idy_ + fabs(w) * idz_;
    cm = fmax(cm, cf);
    um = fmax(um, fmax(fabs(u), fmax(fabs(v), fabs(w))));
  }
  block_sum<NQ>(acc, sh, tot);
  // max reductions (order independent)
  __syncthreads();
  sh[threadIdx.x] = cm;
  sh[STAT_THREADS + threadIdx.x] = um;
  __syncthreads();
  for (int s = STAT_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
      sh[STAT_THREADS + threadIdx.x] = fmax(sh[STAT_THREADS + threadIdx.x], sh[STAT_THREADS + threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int ny = g.ny;
    prof[0 * ny + y] = mean[0];
    prof[1 * ny + y] = mean[1];
    prof[2 * ny + y] = mean[2];
    prof[3 * ny + y] = tot[0] * inv;
    prof[4 * ny + y] = tot[1] * inv;
    prof[5 * ny + y] = tot[2] * inv;
    prof[6 * ny + y] = tot[3] * inv;
    for (int s = 0; s < S; ++s) {
      prof[(7 + 3 * s) * ny + y] = mean[3 + s];
      prof[(8 + 3 * s) * ny + y] = tot[4 + 2 * s] * inv;
      prof[(9 + 3 * s) * ny + y] = tot[5 + 2 * s] * inv;
    }
    cmax[2 * y] = sh[0];
    cmax[2 * y + 1] = sh[STAT_THREADS];
  }
}

// Any non-finite value in the kz = 0 column of a field (cheap NaN/Inf watchdog after a step).
template <typename T>
__global__ void k_finite(Geom g, CPtrs<T> q, int F, int* __restrict__ flag) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)g.nx * g.ny) return;
  for (int f = 0; f < F; ++f) {
    const cx<T> v = q.p[f][i * g.kzs];
    if (!isfinite((double)v.x) || !isfinite((double)v.y)) flag[0] = 1;
  }
}


// ---------------------------------------------------------------------------------------------
// x passes fused with the spectral ends of a substage.
// ---------------------------------------------------------------------------------------------

// Inverse x pass of the work fields of one field group, fed straight from the state.  Per
// (ky kept, kz chunk) tile it loads the group's state columns once (kept kx rows only), forms
// each work field in registers and transforms it:
//   GRP 0 (velocity): u, v, w -> u, v, w, omega = i k x u  (calcOmega*Kernel_mgpu, solver.cu:68-114)
//   GRP 1 (scalar):   phi_s   -> i k phi_s                 (gradientKernel_mgpu, solver.cu:233-256)
//   GRP 2 (colloid):  c, Z    -> i k c, c, -k^2 Z          (scalarAdvection, solver.cu:419-437)
// Replaces k_prologue and the separate inverse x passes of those work fields.
template <typename T, int L, int C, int GRP>
__global__ void __launch_bounds__(lb_threads(C * (L / Plan<L>::E)))
k_xinv_prologue(Geom g, CPtrs<T> q, CPtrs<T> ws, const T* __restrict__ kxv, const T* __restrict__ kyv,
                const T* __restrict__ kzv, const cx<T>* __restrict__ tw, int scalar, int nchunks, int G) {
  constexpr bool VEL = GRP == 0;
  constexpr int NIN = VEL ? 3 : (GRP == 2 ? 2 : 1);
  constexpr int NOUT = VEL ? 6 : (GRP == 2 ? 5 : 3);
  constexpr int E = Plan<L>::E;
  constexpr int TPC = L / E;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = (tid / C) % TPC;
  const int gg = tid / (C * TPC);
  const int chunk = blockIdx.x % nchunks;
  const int o = (blockIdx.x / nchunks) * G + gg;
  const int nouter = 2 * g.keep_y + 1;
  const bool active = o < nouter;
  const int iy = outer_index(active ? o : 0, g.keep_y, g.ny);
  const int iz = chunk * C + c;
  const long long col = (long long)iy * g.kzs + iz;
  const long long estride = (long long)g.ny * g.kzs;
  const T ky = kyv[iy], kz = kzv[iz];
  const int in0 = VEL ? 0 : 3 + scalar;
  const int out0 = VEL ? 0 : 6 + 3 * scalar;
  cx<T> qv[NIN][E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + e * TPC;
    const bool ld = active && row_kept(p, L, g.keep_x);
#pragma unroll
    for (int f = 0; f < NIN; ++f) {
      const int fi = f == 0 ? in0 : (GRP == 2 ? 3 : in0 + f);  // the colloid group also reads Z
      qv[f][e] = ld ? ld_s(&q.p[fi][col + p * estride]) : mk<T>(0, 0);
    }
  }
  const LdsCol idx{gg * L * C + c, C};
#pragma unroll 1
  for (int w = 0; w < NOUT; ++w) {
    cx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const T kx = kxv[t + e * TPC];
      if (VEL) {
        const cx<T> u = qv[0][e], vv = qv[1][e], ww = qv[NIN - 1][e];
        if (w < 3) v[e] = w == 0 ? u : (w == 1 ? vv : ww);
        else if (w == 3) v[e] = csub(ik(ky, ww), ik(kz, vv));
        else if (w == 4) v[e] = csub(ik(kz, u), ik(kx, ww));
        else v[e] = csub(ik(kx, vv), ik(ky, u));
      } else if (GRP == 2 && w == 3) {
        v[e] = qv[0][e];
      } else if (GRP == 2 && w == 4) {
        const T k2 = kx * kx + ky * ky + kz * kz;
        v[e] = mk<T>(-k2 * qv[NIN - 1][e].x, -k2 * qv[NIN - 1][e].y);
      } else {
        v[e] = ik(w == 0 ? kx : (w == 1 ? ky : kz), qv[0][e]);
      }
    }
    __syncthreads();  // the previous field's last LDS reads are done
    fft<T, L, +1>(v, lds, t, idx, tw);
    if (active) {
#pragma unroll
      for (int e = 0; e < E; ++e) st_s(&ws.p[out0 + w][col + (long long)(t + e * TPC) * estride], v[e]);
    }
  }
}

// One work field of a substage formed from one or two state fields (the spectral prologue,
// calcOmega*Kernel_mgpu solver.cu:68-114, gradientKernel_mgpu solver.cu:233-256, the colloid's
// c and lap Z, solver.cu:419-437):  kind 0: A;  1: i k_da A;  2: i (k_da A - k_db B);  3: -k^2 A.
// Axis d picks kx (per row of the x column), ky (per tile) or kz (per lane).
struct XOp {
  int kind, da, db;
};
template <typename T> struct XOpList {
  cx<T>* out[MAXW];
  const cx<T>* A[MAXW];
  const cx<T>* B[MAXW];
  XOp op[MAXW];
};

// Inverse x passes of work fields with their prologue folded into the loads: persistent and
// software-pipelined like k_col_multi (the next tile's inputs are loaded before the current tile
// transforms); tile i is tile i % tpf of list entry i / tpf.  TWO: the two-input operation
// (kind 2); otherwise only A is read.  Inputs and output share the spectral layout, so a column of A (and B) sits at the
// output column's offset; only kept kx rows are read, every row is written.
template <typename T, int L, int C, bool TWO>
__global__ void __launch_bounds__(lb_threads(C * (L / Plan<L>::E)))
k_xinv_op_pipe(XOpList<T> li, const T* __restrict__ kxv, const T* __restrict__ kyv,
               const T* __restrict__ kzv, const cx<T>* __restrict__ tw, ColArgs a, int tpf, int ntiles) {
  constexpr int E = Plan<L>::E;
  constexpr int TPC = L / E;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = (tid / C) % TPC;
  const int g = tid / (C * TPC);
  auto offset_of = [&](int tile, bool& active, int& iy, int& iz) {
    const int lt = tile % tpf;
    const int chunk = lt % a.nchunks;
    const int o = (lt / a.nchunks) * a.G + g;
    active = o < a.nouter;
    iy = outer_index(active ? o : 0, a.outer_keep, a.outer_n);
    iz = chunk * C + c;
    return (long long)iy * a.ostride + iz;
  };
  // A is loaded one tile ahead; B (two-input operation) when the tile is formed, which keeps the
  // register footprint of a 16-column tile without spills
  auto load_col = [&](bool second, int tile, cx<T>* va) {
    const cx<T>* __restrict__ src = second ? li.B[tile / tpf] : li.A[tile / tpf];
    bool active;
    int iy, iz;
    const long long off = offset_of(tile, active, iy, iz);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int p = t + e * TPC;
      const bool ld = active && row_kept(p, L, a.in_keep);
      va[e] = ld ? ld_s(&src[off + (long long)p * a.estride]) : mk<T>(0, 0);
    }
  };
  auto form = [&](int tile, const cx<T>* va, cx<T>* v) {
    bool active;
    int iy, iz;
    (void)offset_of(tile, active, iy, iz);
    const T ky = kyv[iy], kz = kzv[iz];
    const XOp op = li.op[tile / tpf];
    cx<T> vb[TWO ? E : 1];
    if (TWO) load_col(true, tile, vb);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const T kx = kxv[t + e * TPC];
      const T ka = op.da == 0 ? kx : (op.da == 1 ? ky : kz);
      const T kb = op.db == 0 ? kx : (op.db == 1 ? ky : kz);
      if (TWO) v[e] = csub(ik(ka, va[e]), ik(kb, vb[TWO ? e : 0]));
      else if (op.kind == 0) v[e] = va[e];
      else if (op.kind == 1) v[e] = ik(ka, va[e]);
      else {
        const T k2 = kx * kx + ky * ky + kz * kz;
        v[e] = mk<T>(-k2 * va[e].x, -k2 * va[e].y);
      }
    }
  };
  int tile = blockIdx.x;
  cx<T> va[E], v[E];
  load_col(false, tile, va);
  const LdsCol idx{g * L * C + c, C};
  while (true) {
    form(tile, va, v);
    const int next = tile + gridDim.x;
    const bool more = next < ntiles;
    if (more) load_col(false, next, va);
    fft<T, L, +1>(v, lds, t, idx, tw);
    bool active;
    int iy, iz;
    const long long off = offset_of(tile, active, iy, iz);
    if (active) {
      cx<T>* __restrict__ out = li.out[tile / tpf];
#pragma unroll
      for (int e = 0; e < E; ++e) st_s(&out[off + (long long)(t + e * TPC) * a.estride], v[e]);
    }
    if (!more) break;
    tile = next;
  }
}

// Inverse strided passes with the derivatives along the transformed axis applied on the way in
// (the late-derivative x mode): list entry i reads one or two input fields A, B and writes one or
// two transformed outputs out_j = (a_j + i b_j k) A + (c_j + i d_j k) B with a_j .. d_j in
// {-1, 0, 1}, k the wavenumber of the element's mode along the axis.  The x pass reads the state
// once per field (u, v, w, phi) and writes it with its x-derivatives; the y pass adds the
// y-derivatives and omega_z = v_x - i k_y u; the fused z pass forms the remaining terms with
// i k_z (replaces the spectral prologue of the reference's computeRHS, solver.cu:419-437, and the
// curl/gradient kernels it calls).
struct PairOp {
  int nin, nout;
  float a[2], b[2], c[2], d[2];
};
template <typename T> struct PairList {
  const cx<T>* A[MAXW];
  const cx<T>* B[MAXW];
  cx<T>* out[2][MAXW];
  PairOp op[MAXW];
};

// Bit e set: every row p of slab e (p in [e TPC, (e + 1) TPC), the rows a thread's element e takes
// in a length-L transform of TPC threads per column) has |mode| > L / 3, a row the 2/3 rule removes.
constexpr unsigned removed_slabs(int L, int TPC, int E) {
  unsigned m = 0;
  for (int e = 0; e < E; ++e) {
    bool all = true;
    for (int p = e * TPC; p < (e + 1) * TPC && all; ++p) {
      const int md = p < (L >> 1) ? p : p - L;
      all = (md < 0 ? -md : md) > L / 3;
    }
    if (all) m |= 1u << e;
  }
  return m;
}

// Forms of the two-output kernels: both outputs through the radix stages together in two LDS
// regions, A loaded with its tile (PAIR_BATCHED); one after the other through one region, the
// next tile's A loaded once the last output is formed (PAIR_SEQ: half the LDS); together, with the
// next tile's A loaded before they transform (PAIR_BATCHED_PF).
constexpr int PAIR_BATCHED = 0, PAIR_SEQ = 1, PAIR_BATCHED_PF = 2;

// Persistent like k_xinv_op_pipe: tile i is tile i % tpf of list entry i / tpf; B is loaded with
// A (one tile ahead in the prefetching forms).  NIN, NOUT: inputs and outputs of every entry of the
// launch; FORM: how two outputs share the block (one output: always with the next tile loaded
// before the transform).
// The twiddles are read from a copy in LDS, after the data regions: read from memory, each radix
// stage would wait for every load in flight, the next tile's included.  An entry
// may write its outputs over its own inputs (each tile is read before it is written, by the one
// block that owns it, and both outputs are formed from registers); no entry may read a field
// another entry writes.
template <typename T, int L, int C, int NIN, int NOUT, int FORM>
__global__ void __launch_bounds__(lb_threads(C * (L / Plan<L>::E)))
k_inv_pair_pipe(PairList<T> li, const T* __restrict__ kv, const cx<T>* __restrict__ tw, ColArgs a, int tpf,
                int ntiles) {
  constexpr int E = Plan<L>::E;
  constexpr int TPC = L / E;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = (tid / C) % TPC;
  const int g = tid / (C * TPC);
  // Element (row p = t + e TPC, column c) of a tile sits at tile_off(tile) + e TPC estride (the
  // same for the whole group: with one group per block, wave-uniform) plus the lane's
  // c + t estride, which the host keeps below 2^32 bytes.
  auto tile_off = [&](int tile, bool& active) {
    const int lt = tile % tpf;
    const int chunk = lt % a.nchunks;
    const int o = (lt / a.nchunks) * a.G + g;
    active = o < a.nouter;
    int om = active ? o : 0;
    const long long mo = member_split(om, a.nouter_m, a.mstride);
    const int io = outer_index(om, a.outer_keep, a.outer_n);
    return mo + (long long)io * a.ostride + chunk * C;
  };
  const unsigned lane_b = (unsigned)(c + (long long)t * a.estride) * (unsigned)sizeof(cx<T>);
  // the inputs hold the rows the 2/3 rule keeps (in_keep = L / 3, checked by the host): the slabs
  // of removed rows are zeros known at compile time, neither loaded nor transformed
  constexpr unsigned ZS = removed_slabs(L, TPC, E);
  auto load_col = [&](const cx<T>* __restrict__ src, int tile, cx<T>* va) {
    bool active;
    const long long off = tile_off(tile, active);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int p = t + e * TPC;
      va[e] = (!((ZS >> e) & 1u) && active && row_kept(p, L, a.in_keep))
                  ? ld_at(src + off + (long long)e * TPC * a.estride, lane_b)
                  : mk<T>(0, 0);
    }
  };
  constexpr bool SEQ = NOUT == 2 && FORM == PAIR_SEQ;
  // G groups of C columns; each group owns NREG regions of L x C entries
  constexpr int NREG = SEQ ? 1 : NOUT;
  const LdsCol idx{g * NREG * L * C + c, C};
  constexpr bool PF = NOUT == 1 || FORM != PAIR_BATCHED;
  // output j of entry op at this thread's positions
  // the wavenumbers of this thread's rows, loaded once before the loop (see wait_vmem below)
  T kr[E];
#pragma unroll
  for (int e = 0; e < E; ++e) kr[e] = kv[t + e * TPC];
  auto form = [&](const PairOp& op, int j, const cx<T>* va, const cx<T>* vb, cx<T>* r) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const T k = kr[e];
      const T ka = (T)op.b[j] * k, ra = (T)op.a[j];
      cx<T> x = mk<T>(ra * va[e].x - ka * va[e].y, ra * va[e].y + ka * va[e].x);
      if (NIN == 2) {
        const T kb = (T)op.d[j] * k, rb = (T)op.c[j];
        const cx<T> b = vb[NIN == 2 ? e : 0];
        x.x += rb * b.x - kb * b.y;
        x.y += rb * b.y + kb * b.x;
      }
      r[e] = x;
    }
  };
  int tile = blockIdx.x;
  // the prefetching forms load the next tile's B with its A
  cx<T> va[E], vb[NIN == 2 ? E : 1];
  if (PF) load_col(li.A[tile / tpf], tile, va);
  if (PF && NIN == 2) load_col(li.B[tile / tpf], tile, vb);
  cx<T>* twl = lds + (long long)a.G * NREG * L * C;
  for (int i = tid; i < L; i += blockDim.x) twl[i] = tw[i];
  // Nothing in flight when the loop starts: loads issued on paths that may skip some of them (the
  // copy loop above, lane predicates) leave the compiler unable to count the loads in flight, and
  // its first wait inside the loop would then be for every load, the next tile's included.
  wait_vmem();
  __syncthreads();
  while (true) {
    const int ent = tile / tpf;
    const PairOp op = li.op[ent];
    if (!PF) load_col(li.A[ent], tile, va);
    if (!PF && NIN == 2) load_col(li.B[ent], tile, vb);
    const int next = tile + gridDim.x;
    const bool more = next < ntiles;
    bool active;
    const long long off = tile_off(tile, active);
    if constexpr (SEQ) {
#pragma unroll
      for (int j = 0; j < NOUT; ++j) {
        cx<T> v[1][E];
        form(op, j, va, vb, v[0]);
        if (j == NOUT - 1 && more) {
          load_col(li.A[next / tpf], next, va);
          if (NIN == 2) load_col(li.B[next / tpf], next, vb);
        }
        fft_batch<T, L, +1, LdsCol, Plan<L>, false, 1, L * C>(v, lds, t, idx, twl);
        // one wait per tile, after its last transform and before its last stores: the counter
        // covers loads and stores alike, so any wait with both in flight waits for all of them;
        // here the next tile's loads have had a whole transform to arrive, where the compiler's
        // own wait came after the stores and exposed their latency once per tile
        if (j == NOUT - 1) wait_vmem();
        if (active) {
          cx<T>* __restrict__ dst = li.out[j][ent];
#pragma unroll
          for (int e = 0; e < E; ++e) st_at(dst + off + (long long)e * TPC * a.estride, lane_b, v[0][e]);
        }
      }
    } else {
      cx<T> v[NOUT][E];
#pragma unroll
      for (int j = 0; j < NOUT; ++j) form(op, j, va, vb, v[j]);
      if (PF && more) {
        load_col(li.A[next / tpf], next, va);
        if (NIN == 2) load_col(li.B[next / tpf], next, vb);
      }
      fft_batch<T, L, +1, LdsCol, Plan<L>, false, NOUT, L * C>(v, lds, t, idx, twl);
      wait_vmem();  // as in the SEQ form
      if (active) {
#pragma unroll
        for (int j = 0; j < NOUT; ++j) {
          cx<T>* __restrict__ dst = li.out[j][ent];
#pragma unroll
          for (int e = 0; e < E; ++e) st_at(dst + off + (long long)e * TPC * a.estride, lane_b, v[j][e]);
        }
      }
    }
    if (!more) break;
    tile = next;
  }
}

// The velocity group of the fused forward x pass parks its projected right-hand sides in LDS
// (one slot per thread, field and element of a slab the 2/3 rule keeps) when they fit beside the
// twiddles and the x wavenumbers: its registers are then free to load the state of several
// elements before any store.
// Entries of that region for blocks of `threads` threads (0: the sides stay in registers).
template <typename T, int L, int C, bool VEL> __host__ __device__ constexpr long long xfwd_rhs_entries(int threads) {
  constexpr int E = Plan<L>::E, TPC = L / E;
  constexpr int NK = E - __builtin_popcount(removed_slabs(L, TPC, E));
  constexpr long long most = (long long)NK * 3 * lb_threads(C * TPC) * (long long)sizeof(cx<T>) +
                             L * (long long)(sizeof(cx<T>) + sizeof(T));
  return (VEL && most <= 160 * 1024) ? (long long)NK * 3 * threads : 0;
}

// Forward x pass of one field group's nonlinear terms fused with the projection (velocity),
// the sign of the scalar term (computeRHSKernel_mgpu solver.cu:461-521), the mask and the RK3-IF
// update of the state (replaces solver.cu:547-657): the transformed nonlinear terms never return
// to HBM.
template <typename T, int L, int C, bool VEL>
__global__ void __launch_bounds__(lb_threads(C * (L / Plan<L>::E)))
k_xfwd_rk(Geom g, CPtrs<T> q, CPtrs<T> dq, CPtrs<T> nl, const T* __restrict__ kxv,
          const T* __restrict__ kyv, const T* __restrict__ kzv, const cx<T>* __restrict__ tw, RkCoef rc,
          int scalar, int nchunks, int G) {
  constexpr int NF = VEL ? 3 : 1;
  constexpr int E = Plan<L>::E;
  constexpr int TPC = L / E;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = (tid / C) % TPC;
  const int gg = tid / (C * TPC);
  const int chunk = blockIdx.x % nchunks;
  const int o = (blockIdx.x / nchunks) * G + gg;
  const int nouter = 2 * g.keep_y + 1;
  const bool active = o < nouter;
  const int iy = outer_index(active ? o : 0, g.keep_y, g.ny);
  const int iz = chunk * C + c;
  const long long estride = (long long)g.ny * g.kzs;
  // element (row p = t + e TPC, column c) at at(e) (wave-uniform with one group per block) plus the
  // lane's c + t estride, which the host keeps below 2^32 bytes
  const long long base = (long long)iy * g.kzs + chunk * C;
  auto at = [&](int e) { return base + (long long)e * TPC * estride; };
  const unsigned lane_b = (unsigned)(c + (long long)t * estride) * (unsigned)sizeof(cx<T>);
  const LdsCol idx{gg * L * C + c, C};
  const int f0 = VEL ? 0 : 3 + scalar;
  // every field's loads are issued before the first transform (the registers hold all NF fields
  // after the transforms anyway), so the later fields arrive while the earlier ones transform; the
  // twiddles are read from a copy in LDS after the G L C data entries (read from memory, each radix
  // stage would wait for all of those loads)
  cx<T> nv[NF][E];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int e = 0; e < E; ++e) nv[f][e] = active ? ld_at(nl.p[f0 + f] + at(e), lane_b) : mk<T>(0, 0);
  // The x wavenumbers are read from a copy in LDS after the twiddles: loaded from memory per element
  // inside the projection and the update, each was one more wait, and after the first stores a wait
  // for every store in flight as well.
  const T ky = kyv[iy], kz = kzv[iz];
  const long long rhs_n = xfwd_rhs_entries<T, L, C, VEL>(blockDim.x);
  cx<T>* tws = lds + ((long long)G * L * C > rhs_n ? (long long)G * L * C : rhs_n);
  T* kxl = reinterpret_cast<T*>(tws + L);
  for (int i = tid; i < L; i += blockDim.x) {
    tws[i] = tw[i];
    kxl[i] = kxv[i];
  }
  __syncthreads();
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    if (f > 0) __syncthreads();
    fft<T, L, -1>(nv[f], lds, t, idx, (const cx<T>*)tws);
  }
  if (!active) return;
  const int my = mode_of(iy, g.ny), mz = mode_of(iz, g.nz);
  const T a = (T)rc.a, bb = (T)rc.b, dt = (T)rc.dt;
  const T nu = (T)rc.nu[f0];
  // right-hand side of element e: the projected nonlinear term (velocity) or its negative (scalar)
  auto rhs_of = [&](int e, cx<T>* rhs) {
    const T kx = kxl[t + e * TPC];
    const T k2 = kx * kx + ky * ky + kz * kz;
    if (k2 == (T)0) {
#pragma unroll
      for (int f = 0; f < NF; ++f) rhs[f] = mk<T>(0, 0);
    } else if (VEL) {
      const cx<T> n0 = nv[0][e], n1 = nv[NF > 1 ? 1 : 0][e], n2 = nv[NF - 1][e];
      cx<T> kn;
      kn.x = (kx * n0.x + ky * n1.x + kz * n2.x) / k2;
      kn.y = (kx * n0.y + ky * n1.y + kz * n2.y) / k2;
      rhs[0] = csub(n0, cscale(kn, kx));
      if (NF > 1) {
        rhs[NF > 1 ? 1 : 0] = csub(n1, cscale(kn, ky));
        rhs[NF - 1] = csub(n2, cscale(kn, kz));
      }
    } else {
      rhs[0] = mk<T>(-nv[0][e].x, -nv[0][e].y);
    }
  };
  // Slabs of rows the 2/3 rule removes are neither read nor written (the host checks keep_x = L / 3).
  constexpr unsigned ZS = removed_slabs(L, TPC, E);
  constexpr bool RL = xfwd_rhs_entries<T, L, C, VEL>(1) > 0;
  // slot of (kept slab e, field f) in the LDS region of the right-hand sides: thread-private, so no
  // barrier (every thread has read its last exchange of the transforms, which end in one)
  auto slot = [&](int e, int f) {
    const int ek = e - __builtin_popcount(ZS & ((1u << e) - 1u));
    return lds + ((long long)(ek * NF + f) * blockDim.x + tid);
  };
  if constexpr (RL) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if ((ZS >> e) & 1u) continue;
      cx<T> rhs[NF];
      rhs_of(e, rhs);
#pragma unroll
      for (int f = 0; f < NF; ++f) *slot(e, f) = rhs[f];
    }
  }
  // The state and RK-register values of EB elements are loaded before any of them is stored: the
  // stores may alias the loads as far as the compiler knows, so it keeps them in program order,
  // and loading per element and field would cost one HBM round trip each.  A velocity group whose
  // right-hand sides stay in registers has no room to load ahead (its loads go field by field).
  constexpr bool PRE = !VEL || RL;
  constexpr int EB = PRE ? (E / 2 < 4 ? E / 2 : 4) : 1;
  static_assert(E % EB == 0, "element batches must divide the elements of a thread");
#pragma unroll
  for (int e0 = 0; e0 < E; e0 += EB) {
    cx<T> oq[EB][NF], od[EB][NF];
#pragma unroll
    for (int j = 0; j < EB && PRE; ++j) {
      const int p = t + (e0 + j) * TPC;
      const long long o2 = at(e0 + j);
      const bool live = !((ZS >> (e0 + j)) & 1u) && row_kept(p, L, g.keep_x) && iz < g.nzh &&
                        kept(g, mode_of(p, g.nx), my, mz);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        od[j][f] = (live && !rc.first) ? ld_at(dq.p[f0 + f] + o2, lane_b) : mk<T>(0, 0);
        oq[j][f] = live ? ld_at(q.p[f0 + f] + o2, lane_b) : mk<T>(0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      const int e = e0 + j;
      const int p = t + e * TPC;
      if (((ZS >> e) & 1u) || !row_kept(p, L, g.keep_x)) continue;
      const long long o2 = at(e);
      if (iz >= g.nzh || !kept(g, mode_of(p, g.nx), my, mz)) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          st_at(q.p[f0 + f] + o2, lane_b, mk<T>(0, 0));
          if (!rc.last) st_at(dq.p[f0 + f] + o2, lane_b, mk<T>(0, 0));
        }
        continue;
      }
      const T kx = kxl[p];
      const T k2 = kx * kx + ky * ky + kz * kz;
      cx<T> rhs[NF];
      if constexpr (RL) {
#pragma unroll
        for (int f = 0; f < NF; ++f) rhs[f] = *slot(e, f);
      } else {
        rhs_of(e, rhs);
      }
      const T eps = -dexpm1<T>(-nu * k2 * (T)rc.tau);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        cx<T> d = cscale(rhs[f], dt);
        if (!rc.first) {
          const cx<T> old = PRE ? od[j][f] : ld_at(dq.p[f0 + f] + o2, lane_b);
          d.x = a * old.x + d.x;
          d.y = a * old.y + d.y;
        }
        cx<T> qq = PRE ? oq[j][f] : ld_at(q.p[f0 + f] + o2, lane_b);
        qq.x = if_update(qq.x, bb, d.x, eps);
        qq.y = if_update(qq.y, bb, d.y, eps);
        st_at(q.p[f0 + f] + o2, lane_b, qq);
        if (!rc.last) st_at(dq.p[f0 + f] + o2, lane_b, mk<T>(apply_if(d.x, eps), apply_if(d.y, eps)));
      }
    }
  }
}

// The reductions of the allocation-free statistics call, on the device and in the order the host
// sums the same partials for gnss_stats (so both give the same bits): fin[q] = sum over x of
// part[x][q] (q < SPEC_NQ); fin[SPEC_NQ], fin[SPEC_NQ + 1] = max over y of the CFL sum and of
// max |u_i|; fin[SPEC_NQ + 2 + j] = sum over x of spec[x][j] (j < nspec shell entries).
static __global__ void __launch_bounds__(STAT_THREADS)
k_stats_reduce(const double* __restrict__ part, const double* __restrict__ spec, const double* __restrict__ cmax,
               int nx, int ny, int nspec, double* __restrict__ fin) {
  const int t = threadIdx.x;
  if (t < SPEC_NQ) {
    double acc = 0.0;
    for (int x = 0; x < nx; ++x) acc += part[(long long)x * SPEC_NQ + t];
    fin[t] = acc;
  } else if (t < SPEC_NQ + 2) {
    double m = 0.0;
    for (int y = 0; y < ny; ++y) m = fmax(m, cmax[2 * y + (t - SPEC_NQ)]);
    fin[t] = m;
  }
  for (int j = t; j < nspec; j += STAT_THREADS) {
    double acc = 0.0;
    for (int x = 0; x < nx; ++x) acc += spec[(long long)x * nspec + j];
    fin[SPEC_NQ + 2 + j] = acc;
  }
}

// Plain copy (bandwidth reference): block b copies the contiguous chunk [b chunk, (b + 1) chunk)
// of 16-byte elements with four nontemporal loads in flight per thread before the stores.  Of the
// copy forms measured on MI355X (grid-stride or chunked, 2-8 loads in flight, cached or
// nontemporal) this one moves the most bytes per second.
typedef float v4f __attribute__((ext_vector_type(4)));
static __global__ void __launch_bounds__(256) k_copy16(const v4f* __restrict__ in, v4f* __restrict__ out,
                                                       long long n, long long chunk) {
  const long long beg = (long long)blockIdx.x * chunk;
  const long long end = beg + chunk < n ? beg + chunk : n;
  long long i = beg + threadIdx.x;
  for (; i + 3 * 256 < end; i += 4 * 256) {
    v4f a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = __builtin_nontemporal_load(&in[i + u * 256]);
#pragma unroll
    for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(a[u], &out[i + u * 256]);
  }
  for (; i < end; i += 256) __builtin_nontemporal_store(__builtin_nontemporal_load(&in[i]), &out[i]);
}

// ---------------------------------------------------------------------------------------------
// Diagnostics
// ---------------------------------------------------------------------------------------------

// omega^ = i k x u^ over the whole spectral array (calcOmega*Kernel_mgpu, solver.cu:68-114);
// the kz padding is written as zero.
template <typename T>
__global__ void k_vorticity(Geom g, CPtrs<T> q, CPtrs<T> w, const T* __restrict__ kxv,
                            const T* __restrict__ kyv, const T* __restrict__ kzv, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int iz = (int)(i % g.kzs);
  const long long r = i / g.kzs;
  const int iy = (int)(r % g.ny), ix = (int)(r / g.ny);
  if (iz >= g.nzh) {
    for (int c = 0; c < 3; ++c) w.p[c][i] = mk<T>(0, 0);
    return;
  }
  const T kx = kxv[ix], ky = kyv[iy], kz = kzv[iz];
  const cx<T> u = q.p[0][i], v = q.p[1][i], ww = q.p[2][i];
  w.p[0][i] = csub(ik(ky, ww), ik(kz, v));
  w.p[1][i] = csub(ik(kz, u), ik(kx, ww));
  w.p[2][i] = csub(ik(kx, v), ik(ky, u));
}

// |(a, b, c)| in place over a (VectorMagnitude_kernel, statistics.cu:2005-2018).
template <typename T>
__global__ void k_magnitude(T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T x = a[i], y = b[i], z = c[i];
  a[i] = sqrt(x * x + y * y + z * z);
}

// One physical plane [x][y][z] -> FP64: axis 2 (z = index) -> [x][y], axis 1 -> [x][z],
// axis 0 -> [y][z] (writexyfields / writexzfields / writeyzfields, iofuncs.cu:187-260).
template <typename T>
__global__ void k_plane(const T* __restrict__ a, double* __restrict__ out, int nx, int ny, int nz, int axis,
                        int index) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long src;
  if (axis == 2) {
    if (i >= (long long)nx * ny) return;
    src = i * nz + index;
  } else if (axis == 1) {
    if (i >= (long long)nx * nz) return;
    src = ((i / nz) * ny + index) * nz + i % nz;
  } else {
    if (i >= (long long)ny * nz) return;
    src = (long long)index * ny * nz + i;
  }
  out[i] = (double)a[src];
}

// Iso-surface area by the indicator gradient (surfaceArea_kernel_mgpu, statistics.cu:1310-1460):
// per grid point -(grad F . grad chi) / |grad F| with chi = (F <= iso), central differences on the
// periodic grid; one block per x plane, fixed per-thread order and tree -> part[x] (FP64).
template <typename T>
__global__ void __launch_bounds__(STAT_THREADS)
k_surface_area(Geom g, const T* __restrict__ F, double iso, double dx, double dy, double dz,
               double* __restrict__ part) {
  __shared__ double sh[STAT_THREADS];
  const int x = blockIdx.x;
  const int xm = x == 0 ? g.nx - 1 : x - 1, xp = x == g.nx - 1 ? 0 : x + 1;
  const long long plane = (long long)g.ny * g.nz;
  double acc[1] = {0.0};
  for (long long i = threadIdx.x; i < plane; i += STAT_THREADS) {
    const int y = (int)(i / g.nz), z = (int)(i % g.nz);
    const int ym = y == 0 ? g.ny - 1 : y - 1, yp = y == g.ny - 1 ? 0 : y + 1;
    const int zm = z == 0 ? g.nz - 1 : z - 1, zp = z == g.nz - 1 ? 0 : z + 1;
    auto at = [&](int a, int b, int c) { return (double)F[((long long)a * g.ny + b) * g.nz + c]; };
    const double fxm = at(xm, y, z), fxp = at(xp, y, z), fym = at(x, ym, z), fyp = at(x, yp, z);
    const double fzm = at(x, y, zm), fzp = at(x, y, zp);
    const double gx = (fxp - fxm) / (2.0 * dx), gy = (fyp - fym) / (2.0 * dy), gz = (fzp - fzm) / (2.0 * dz);
    const double cxv = ((double)(fxp <= iso) - (double)(fxm <= iso)) / (2.0 * dx);
    const double cyv = ((double)(fyp <= iso) - (double)(fym <= iso)) / (2.0 * dy);
    const double czv = ((double)(fzp <= iso) - (double)(fzm <= iso)) / (2.0 * dz);
    if ((gx == 0 && gy == 0 && gz == 0) || (cxv == 0 && cyv == 0 && czv == 0)) continue;
    acc[0] += -(gx * cxv + gy * cyv + gz * czv) / sqrt(gx * gx + gy * gy + gz * gz);
  }
  double tot[1];
  block_sum<1>(acc, sh, tot);
  if (threadIdx.x == 0) part[x] = tot[0];
}

}  // namespace gnss
