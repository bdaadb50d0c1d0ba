// kernels_mfma.h — strided C2C pass whose first radix-16 stage runs on the MFMA matrix cores.
//
// A radix-16 DFT applied to many butterflies is a matrix product: with X the 16 complex inputs of
// each butterfly as a column, Y = W16 X.  In real block form
//     [Yr]   [Wr  -Wi] [Xr]
//     [Yi] = [Wi   Wr] [Xi],     W = exp(SIGN 2 pi i q r / 16),
// a 32 x 32 real matrix times 32 x 32 columns (32 butterflies) — one v_mfma_f32_32x32x8f16 tile
// per 8-wide K chunk.  Operands are split FP16: each butterfly is scaled by an exact power of two
// (its largest component maps into [0.5, 1)), then x = hi + lo with hi = fp16(x) and
// lo = fp16(x - hi); products hi*hi + hi*lo + lo*hi are accumulated in FP32 (the lo*lo term is
// below FP32 rounding).  W is split the same way.  The remaining radix stages run on VALU.
//
// Used for the FP32 form when GNSS_FLAG_MFMA_R16 is set, for strided passes of length 256, 512
// or 1024 (plans 16 x 16, 16 x 8 x 4, 16 x 8 x 8; 16 elements per thread).
#pragma once
#include "fft_device.h"
#include "kernels_fft.h"

namespace gnss {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

template <int L> struct MfmaTail;
template <> struct MfmaTail<256> { using R = Radices<16>; };
template <> struct MfmaTail<512> { using R = Radices<8, 4>; };
template <> struct MfmaTail<1024> { using R = Radices<8, 8>; };

template <int L, int SIGN, typename Idx, typename Rad> struct MfmaTailRun;
template <int L, int SIGN, typename Idx, int... Rs> struct MfmaTailRun<L, SIGN, Idx, Radices<Rs...>> {
  __device__ __forceinline__ static void run(cx<float> (*v)[16], cx<float>* lds, int t, const Idx& idx,
                                             const cx<float>* __restrict__ tw) {
    Stages<float, L, 16, SIGN, 16, Idx, false, 1, 0, Rs...>::run(v, lds, t, idx, tw);
  }
};

// cos and sin of 2 pi j / 16 (exact constants, j = 0..15)
__device__ __forceinline__ void cs16(int j, float& c, float& s) {
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
  const float cv[16] = {1.f, c1, h, s1, 0.f, -s1, -h, -c1, -1.f, -c1, -h, -s1, 0.f, s1, h, c1};
  const float sv[16] = {0.f, s1, h, c1, 1.f, c1, h, s1, 0.f, -s1, -h, -c1, -1.f, -c1, -h, -s1};
  c = cv[j & 15];
  s = sv[j & 15];
}

template <int L, int SIGN, int C>
__global__ void __launch_bounds__(C * (L / 16))
k_col_mfma(cx<float>* __restrict__ data, const cx<float>* __restrict__ tw, ColArgs a) {
  constexpr int E = 16;
  constexpr int TPC = L / E;       // threads per column; also the number of radix-16 butterflies
  constexpr int NT = C * TPC;      // threads per block
  constexpr int TILES = C * TPC / 32;  // 32-butterfly tiles per block
  static_assert(TILES % (NT / 64) == 0, "whole tiles per wave");
  constexpr int TPW = TILES / (NT / 64);
  extern __shared__ unsigned char smem_raw[];
  cx<float>* lds = reinterpret_cast<cx<float>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = tid / C;
  const int chunk = blockIdx.x % a.nchunks;
  const int o = blockIdx.x / a.nchunks;  // one outer index per block (G = 1)
  const long long io = outer_index(o, a.outer_keep, a.outer_n);
  cx<float>* base = data + io * a.ostride + (long long)chunk * C + c;
  // 1. coalesced load into the LDS tile [pos][C]
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + e * TPC;
    lds[p * C + c] = row_kept(p, L, a.in_keep) ? base[(long long)p * a.estride] : mk<float>(0, 0);
  }
  // 2. constant operand: lane row m of the 32 x 32 block matrix, its 16 K entries (hi, lo)
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int half = lane >> 5;
  const int m = lane & 31;
  half4_t ahi[4], alo[4];
#pragma unroll
  for (int kc = 0; kc < 4; ++kc) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = 8 * kc + 4 * half + i;
      const int q = m & 15, r = k & 15;
      float cw, sw;
      cs16(q * r, cw, sw);
      const float wi = (float)SIGN * sw;
      float val;
      if (m < 16) val = (k < 16) ? cw : -wi;
      else val = (k < 16) ? wi : cw;
      const _Float16 hi = (_Float16)val;
      ahi[kc][i] = hi;
      alo[kc][i] = (_Float16)(val - (float)hi);
    }
  }
  __syncthreads();
  // 3. gather each wave's butterflies (inputs r of butterfly (cb, j) sit at position j + 16... r)
  //    lane holds r in {4 half + 0..3, 8 + 4 half + 0..3}
  cx<float> xin[TPW][8];
  int bcol[TPW], bj[TPW];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int n = (wave * TPW + u) * 32 + (lane & 31);  // butterfly index in the block
    bcol[u] = n / TPC;
    bj[u] = n % TPC;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = (i < 4) ? (4 * half + i) : (8 + 4 * half + (i - 4));
      xin[u][i] = lds[(bj[u] + r * TPC) * C + bcol[u]];
    }
  }
  __syncthreads();
  // 4. block-scaled split-FP16 MFMA, outputs to the Stockham positions j * 16 + q
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    float mx = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) mx = fmaxf(mx, fmaxf(fabsf(xin[u][i].x), fabsf(xin[u][i].y)));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    int ex = 0;
    if (mx > 0.f) (void)frexpf(mx, &ex);
    const float sc = ldexpf(1.f, -ex), inv = ldexpf(1.f, ex);
    half4_t bhi[4], blo[4];
#pragma unroll
    for (int kc = 0; kc < 4; ++kc)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        // K row k = 8 kc + 4 half + i: real parts for kc < 2, imaginary parts for kc >= 2
        const int slot = (kc & 1) * 4 + i;  // r = 4 half + i (kc even) or 8 + 4 half + i (kc odd)
        const float x = ((kc < 2) ? xin[u][slot].x : xin[u][slot].y) * sc;
        const _Float16 hi = (_Float16)x;
        bhi[kc][i] = hi;
        blo[kc][i] = (_Float16)(x - (float)hi);
      }
    float16_t acc = {0};
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
      acc = __builtin_amdgcn_mfma_f32_32x32x8f16(ahi[kc], bhi[kc], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8f16(ahi[kc], blo[kc], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8f16(alo[kc], bhi[kc], acc, 0, 0, 0);
    }
    // lane holds rows m = 8 (i / 4) + 4 half + (i % 4) of column n: real parts for i < 8
    // (q = m), imaginary parts for i >= 8 (q = m - 16) — the same 8 outputs q
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int q = 8 * (i / 4) + 4 * half + (i % 4);
      cx<float> y;
      y.x = acc[i] * inv;
      y.y = acc[i + 8] * inv;
      lds[(bj[u] * 16 + q) * C + bcol[u]] = y;
    }
  }
  __syncthreads();
  // 5. remaining radix stages on VALU from the strided positions
  cx<float> v[1][E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[0][e] = lds[(t + e * TPC) * C + c];
  __syncthreads();
  MfmaTailRun<L, SIGN, LdsCol, typename MfmaTail<L>::R>::run(v, lds, t, LdsCol{c, C}, tw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + e * TPC;
    if (row_kept(p, L, a.out_keep)) base[(long long)p * a.estride] = v[0][e];
  }
}

}  // namespace gnss
