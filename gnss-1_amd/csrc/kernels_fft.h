// kernels_fft.h — 3D R2C / C2R passes of the engine.
//
// Device layouts (one complex field = nx * ny * kzs values, kzs = nz/2 + 1 rounded up to the
// column chunk C):
//   physical  real    [x][y][z], z fastest, unpadded
//   spectral  complex [kx][ky][kz], kz fastest (row stride kzs)
// A 3D forward transform is three passes: z (R2C, contiguous rows), y and x (C2C over strided
// columns).  The reference does a batched 2D cuFFT over (y,z), a per-plane gather + copy
// transpose and an NY-looped 1D cuFFT over x (fftfuncs.cu:288-320); on one MI355X the transpose
// disappears: the strided passes read tiles of C contiguous columns straight from HBM into
// registers (lane = column), so the x<->y reordering happens in the register/LDS tile of the
// first radix stage.  The inverse runs the passes in the opposite order and folds the
// 1/(nx ny nz) scaling of scaleKernel_mgpu (cudafuncs.cu:19-31) into the z pass.
//
// Pruning for the 2/3 rule: a pass can skip loading rows whose |mode| exceeds a bound (they are
// known zeros), skip storing rows nobody reads, restrict its outer loop to kept modes and its
// column chunks to kz below a limit.  Zero inputs give bitwise the same outputs as computing
// them, so pruned and unpruned transforms agree exactly on the kept modes.
#pragma once
#include "fft_device.h"

namespace gnss {

// Outer-index map: identity, or the kept-mode set {0..keep} U {n-keep..n-1} of an axis.
__device__ __forceinline__ int outer_index(int o, int keep, int n) {
  if (keep < 0) return o;
  return o <= keep ? o : n - (2 * keep + 1) + o;
}
__device__ __forceinline__ bool row_kept(int p, int n, int keep) {
  if (keep < 0) return true;
  const int m = mode_of(p, n);
  return (m < 0 ? -m : m) <= keep;
}

struct ColArgs {
  long long estride;   // distance between consecutive elements along the transformed axis
  long long ostride;   // distance between consecutive outer indices
  int nchunks;         // column chunks of C contiguous columns
  int nouter;          // outer indices processed (after the kept-mode map)
  int G;               // outer indices per block
  int outer_keep;      // -1 identity, else kept-mode map over an axis of length outer_n
  int outer_n;
  int in_keep;         // -1 all rows, else rows with |mode| <= in_keep are loaded (others are 0)
  int out_keep;        // -1 all rows, else only rows with |mode| <= out_keep are stored
};

// C2C pass along a strided axis of length L, in place.  Block = G x TPC x C threads.
// Largest block a pass launches: one transform group, or 256 threads of several groups.
constexpr int lb_threads(int per_group) { return per_group > 256 ? per_group : 256; }

template <typename T, int L, int SIGN, int C>
__global__ void __launch_bounds__(lb_threads(C * (L / Plan<L>::E)))
k_col(cx<T>* __restrict__ data, const cx<T>* __restrict__ tw, ColArgs a) {
  constexpr int E = Plan<L>::E;
  constexpr int TPC = L / E;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = (tid / C) % TPC;
  const int g = tid / (C * TPC);
  const int chunk = blockIdx.x % a.nchunks;
  const int o = (blockIdx.x / a.nchunks) * a.G + g;
  const bool active = o < a.nouter;
  const long long io = outer_index(active ? o : 0, a.outer_keep, a.outer_n);
  cx<T>* base = data + io * a.ostride + (long long)chunk * C + c;
  cx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + e * TPC;
    if (active && row_kept(p, L, a.in_keep)) v[e] = base[(long long)p * a.estride];
    else v[e] = mk<T>(0, 0);
  }
  fft<T, L, SIGN>(v, lds, t, LdsCol{g * L * C + c, C}, tw);
  if (!active) return;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + e * TPC;
    if (row_kept(p, L, a.out_keep)) base[(long long)p * a.estride] = v[e];
  }
}

// Row stride (complex entries) of the LDS tile used by the z passes: M+1 values plus padding.
__host__ __device__ constexpr int zrow_stride(int M) { return M + 2 + (M >> 4); }

// R2C along contiguous z rows: in = rows of NZ reals, out = rows of kzs complex (kz < out_lim
// stored; entries between nz/2 and kzs are zero padding).
template <typename T, int NZ>
__global__ void __launch_bounds__(lb_threads((NZ / 2) / Plan<NZ / 2>::E))
k_r2c_z(const T* __restrict__ in, cx<T>* __restrict__ out, const cx<T>* __restrict__ twM,
        const cx<T>* __restrict__ twN, long long nrows, int kzs, int out_lim, int G) {
  constexpr int M = NZ / 2;
  constexpr int E = Plan<M>::E;
  constexpr int TPC = M / E;
  constexpr int RS = zrow_stride(M);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int t = tid % TPC;
  const int g = tid / TPC;
  const long long row = (long long)blockIdx.x * G + g;
  const bool active = row < nrows;
  const cx<T>* src = reinterpret_cast<const cx<T>*>(in + (active ? row : 0) * NZ);
  cx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = active ? src[t + e * TPC] : mk<T>(0, 0);
  const LdsRow idx{g * RS};
  fft<T, M, -1>(v, lds, t, idx, twM);
#pragma unroll
  for (int e = 0; e < E; ++e) lds[idx(t + e * TPC)] = v[e];
  __syncthreads();
  if (!active) return;
  cx<T>* dst = out + row * kzs;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + e * TPC;
    if (k < out_lim) {
      const cx<T> zk = v[e];
      const cx<T> zc = cconj(lds[idx(k == 0 ? 0 : M - k)]);
      const cx<T> p = cmul(twN[k], csub(zk, zc));
      cx<T> x;
      x.x = (T)0.5 * (zk.x + zc.x + p.y);
      x.y = (T)0.5 * (zk.y + zc.y - p.x);
      dst[k] = x;
    }
  }
  if (t == 0 && M < out_lim) dst[M] = mk<T>(v[0].x - v[0].y, (T)0);
  for (int k = M + 1 + t; k < kzs && k < out_lim; k += TPC) dst[k] = mk<T>(0, 0);
}

// C2R along contiguous z rows: in = rows of kzs complex (entries kz >= in_lim are treated as 0;
// imaginary parts of kz = 0 and kz = nz/2 are ignored), out = rows of NZ reals times `scale`.
template <typename T, int NZ>
__global__ void __launch_bounds__(lb_threads((NZ / 2) / Plan<NZ / 2>::E))
k_c2r_z(const cx<T>* __restrict__ in, T* __restrict__ out, const cx<T>* __restrict__ twM,
        const cx<T>* __restrict__ twN, long long nrows, int kzs, int in_lim, T scale, int G) {
  constexpr int M = NZ / 2;
  constexpr int E = Plan<M>::E;
  constexpr int TPC = M / E;
  constexpr int RS = zrow_stride(M);
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int t = tid % TPC;
  const int g = tid / TPC;
  const long long row = (long long)blockIdx.x * G + g;
  const bool active = row < nrows;
  const cx<T>* src = in + (active ? row : 0) * kzs;
  const LdsRow idx{g * RS};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + e * TPC;
    cx<T> x = (active && k < in_lim) ? src[k] : mk<T>(0, 0);
    if (k == 0) x.y = 0;
    lds[idx(k)] = x;
  }
  if (t == 0) {
    cx<T> x = (active && M < in_lim) ? src[M] : mk<T>(0, 0);
    x.y = 0;
    lds[idx(M)] = x;
  }
  __syncthreads();
  cx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + e * TPC;
    const cx<T> xk = lds[idx(k)];
    const cx<T> xc = cconj(lds[idx(M - k)]);
    const cx<T> p = cmul(cconj(twN[k]), csub(xk, xc));
    cx<T> z;
    z.x = xk.x + xc.x - p.y;
    z.y = xk.y + xc.y + p.x;
    v[e] = z;
  }
  __syncthreads();
  fft<T, M, +1>(v, lds, t, idx, twM);
  if (!active) return;
  cx<T>* dst = reinterpret_cast<cx<T>*>(out + row * NZ);
#pragma unroll
  for (int e = 0; e < E; ++e) dst[t + e * TPC] = cscale(v[e], scale);
}

}  // namespace gnss
