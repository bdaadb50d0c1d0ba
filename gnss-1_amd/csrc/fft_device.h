// fft_device.h — device building blocks of the VALU radix FFT passes (gfx950).
//
// Replaces the cuFFT plans of the reference (fftfuncs.cu:133-231) with hand-written Stockham
// autosort transforms.  A transform of length L is computed by TPC = L/E threads; thread t holds
// E complex values in "strided" order: register e holds position t + e*TPC.  Each radix stage
// works entirely in registers; between stages the values are exchanged through LDS.  The input
// of the first stage and the output of the last stage are both in strided order, so a pass loads
// from and stores to HBM directly from registers: the lanes of a wavefront cover consecutive
// columns (strided passes) or consecutive positions (contiguous passes), which keeps every HBM
// access in runs of at least 128 bytes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gnss {

template <typename T> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename T> using cx = typename cplx_t<T>::type;

template <typename T> __device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { a.x += b.x; a.y += b.y; return a; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { a.x -= b.x; a.y -= b.y; return a; }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename V> __device__ __forceinline__ V cconj(V a) { a.y = -a.y; return a; }
template <typename V, typename S> __device__ __forceinline__ V cscale(V a, S s) { a.x *= s; a.y *= s; return a; }
// HBM element I/O of the transform passes.  A pass streams a whole field through and the next
// pass reads it only after every block has finished, so nothing is gained by keeping the lines in
// L2 or the last-level cache: with GNSS_NT_STORE / GNSS_NT_LOAD the element goes through a
// nontemporal (streaming) vector store / load.
#ifndef GNSS_NT_STORE
#define GNSS_NT_STORE 1
#endif
#ifndef GNSS_NT_LOAD
#define GNSS_NT_LOAD 0
#endif
template <typename V> __device__ __forceinline__ void st_s(V* p, const V& v) {
#if GNSS_NT_STORE
  using E = decltype(v.x);
  typedef E v2 __attribute__((ext_vector_type(2)));
  v2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
#else
  *p = v;
#endif
}
template <typename V> __device__ __forceinline__ V ld_s(const V* p) {
#if GNSS_NT_LOAD
  using E = decltype(p->x);
  typedef E v2 __attribute__((ext_vector_type(2)));
  const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  V r;
  r.x = w.x;
  r.y = w.y;
  return r;
#else
  return *p;
#endif
}
// The element at a (wave-uniform) base plus a 32-bit per-lane byte offset: addressed with the base
// in scalar registers and one vector register per lane for any number of bases, where a 64-bit
// element index costs a vector register pair per access.
template <typename V> __device__ __forceinline__ V ld_at(const V* base, unsigned byte_off) {
  return ld_s(reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + byte_off));
}
template <typename V> __device__ __forceinline__ void st_at(V* base, unsigned byte_off, const V& v) {
  st_s(reinterpret_cast<V*>(reinterpret_cast<char*>(base) + byte_off), v);
}

// Waits until none of this wave's vector-memory operations is in flight: s_waitcnt vmcnt(0), with
// the expcnt (bits 6:4) and lgkmcnt (bits 11:8) fields at their maxima so that only vmcnt (bits
// 3:0 and 15:14) is waited for.  Before a pipelined loop: loads issued ahead of it on a path that may
// skip some of them (a loop of unknown trip count, a lane predicate) would otherwise leave the
// compiler unable to count the loads in flight inside the loop, and its first wait there would be
// for every load, the next tile's prefetch included.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// multiply by (SIGN * i)
template <int SIGN, typename V> __device__ __forceinline__ V cmul_si(V a) {
  V r;
  if (SIGN < 0) { r.x = a.y; r.y = -a.x; } else { r.x = -a.y; r.y = a.x; }
  return r;
}

// ---------------------------------------------------------------------------------------------
// Small DFTs, out[q] = sum_r a[r] exp(SIGN 2 pi i r q / R), in place.
// ---------------------------------------------------------------------------------------------
template <int R, int SIGN, typename V> struct Dft;

template <int SIGN, typename V> struct Dft<2, SIGN, V> {
  __device__ __forceinline__ static void run(V* a) {
    V t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <int SIGN, typename V> struct Dft<3, SIGN, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = decltype(a[0].x);
    const T s = (T)0.86602540378443864676372317075294;
    V t = cadd(a[1], a[2]);
    V d = csub(a[1], a[2]);
    V m = csub(a[0], cscale(t, (T)0.5));
    V jd = cscale(cmul_si<SIGN>(d), s);
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, jd);
    a[2] = csub(m, jd);
  }
};

template <int SIGN, typename V> struct Dft<4, SIGN, V> {
  __device__ __forceinline__ static void run(V* a) {
    V t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    V t2 = cadd(a[1], a[3]), t3 = cmul_si<SIGN>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <int SIGN, typename V> struct Dft<8, SIGN, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = decltype(a[0].x);
    const T h = (T)0.70710678118654752440084436210485;
    V e[4] = {a[0], a[2], a[4], a[6]};
    V o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, SIGN, V>::run(e);
    Dft<4, SIGN, V>::run(o);
    // o[q] *= w8^q, w8 = exp(SIGN 2 pi i / 8)
    V o1 = o[1];
    o[1].x = h * (o1.x - (T)SIGN * o1.y);
    o[1].y = h * (o1.y + (T)SIGN * o1.x);
    o[2] = cmul_si<SIGN>(o[2]);
    V o3 = o[3];
    o[3].x = h * (-o3.x - (T)SIGN * o3.y);
    o[3].y = h * (-o3.y + (T)SIGN * o3.x);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = cadd(e[q], o[q]);
      a[q + 4] = csub(e[q], o[q]);
    }
  }
};

template <int SIGN, typename V> struct Dft<16, SIGN, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = decltype(a[0].x);
    const T c1 = (T)0.92387953251128675612818318939679;  // cos(pi/8)
    const T s1 = (T)0.38268343236508977172845998403040;  // sin(pi/8)
    const T h = (T)0.70710678118654752440084436210485;
    // 16 = 4 x 4: columns r (input index 4m + r), DFT4 over m, twiddle w16^(r q), DFT4 over r
    V b[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int m = 0; m < 4; ++m) b[r][m] = a[4 * m + r];
      Dft<4, SIGN, V>::run(b[r]);
    }
    // twiddles w16^(r q) for r, q in 1..3; w16^j = cos(2 pi j/16) + SIGN i sin(2 pi j/16)
    const T cw[10] = {(T)1, c1, h, s1, (T)0, -s1, -h, -c1, (T)-1, -c1};
    const T sw[10] = {(T)0, s1, h, c1, (T)1, c1, h, s1, (T)0, -s1};
#pragma unroll
    for (int r = 1; r < 4; ++r)
#pragma unroll
      for (int q = 1; q < 4; ++q) {
        const int j = r * q;  // <= 9
        V w;
        w.x = cw[j];
        w.y = (T)SIGN * sw[j];
        b[r][q] = cmul(b[r][q], w);
      }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      V c[4] = {b[0][q], b[1][q], b[2][q], b[3][q]};
      Dft<4, SIGN, V>::run(c);
#pragma unroll
      for (int p = 0; p < 4; ++p) a[q + 4 * p] = c[p];
    }
  }
};

// ---------------------------------------------------------------------------------------------
// Compile-time plans: radix sequence and elements per thread for each supported length.
// ---------------------------------------------------------------------------------------------
template <int... Rs> struct Radices {};

template <int L> struct Plan;
template <> struct Plan<8> { static constexpr int E = 8; using R = Radices<8>; };
template <> struct Plan<12> { static constexpr int E = 12; using R = Radices<4, 3>; };
template <> struct Plan<16> { static constexpr int E = 4; using R = Radices<4, 4>; };
template <> struct Plan<24> { static constexpr int E = 12; using R = Radices<4, 2, 3>; };
template <> struct Plan<32> { static constexpr int E = 8; using R = Radices<8, 4>; };
template <> struct Plan<48> { static constexpr int E = 12; using R = Radices<4, 4, 3>; };
template <> struct Plan<64> { static constexpr int E = 8; using R = Radices<8, 8>; };
template <> struct Plan<96> { static constexpr int E = 12; using R = Radices<4, 4, 2, 3>; };
template <> struct Plan<128> { static constexpr int E = 8; using R = Radices<8, 4, 4>; };
template <> struct Plan<192> { static constexpr int E = 12; using R = Radices<4, 4, 4, 3>; };
template <> struct Plan<256> { static constexpr int E = 8; using R = Radices<8, 8, 4>; };
template <> struct Plan<384> { static constexpr int E = 12; using R = Radices<4, 4, 4, 2, 3>; };
template <> struct Plan<512> { static constexpr int E = 8; using R = Radices<8, 8, 8>; };
template <> struct Plan<1024> { static constexpr int E = 16; using R = Radices<16, 8, 8>; };

// Twiddle read: table tw[m] = exp(-2 pi i m / L); the inverse uses the conjugate.
template <int SIGN, typename V> __device__ __forceinline__ V twid(const V* __restrict__ tw, int m) {
  V w = tw[m];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// Number of radices of a plan and the radix of stage i.
template <typename R> struct RadInfo;
template <int... Rs> struct RadInfo<Radices<Rs...>> {
  static constexpr int n = sizeof...(Rs);
  static constexpr int r[sizeof...(Rs)] = {Rs...};
};

// The twiddles one thread of a length-L transform (plan P) multiplies by, over all radix stages,
// read from the table once and then kept in registers.  A kernel that runs many transforms of one
// length in the same thread positions (the fused z pass) so issues no twiddle load, and waits for
// none, inside its radix stages.  Stage with span NS, block bt, input r >= 1 sits at
// w[off(NS) + bt * (R - 1) + r - 1]; the values are exp(-2 pi i m / L) (conjugated at use).
template <typename T, int L, typename P> struct TwRegs {
  using RI = RadInfo<typename P::R>;
  static constexpr int E = P::E;
  static constexpr int off(int ns) {
    int o = 0, s = 1;
    for (int i = 0; i < RI::n; ++i) {
      if (s == ns) return o;
      if (s > 1) o += (E / RI::r[i]) * (RI::r[i] - 1);
      s *= RI::r[i];
    }
    return o;
  }
  static constexpr int total = off(L);
  cx<T> w[total > 0 ? total : 1];

  __device__ __forceinline__ void load(const cx<T>* __restrict__ tw, int t) {
    constexpr int TPC = L / E;
    int s = 1, o = 0;
#pragma unroll
    for (int i = 0; i < RI::n; ++i) {
      const int R = RI::r[i], NB = E / R;
      if (s > 1) {
#pragma unroll
        for (int bt = 0; bt < NB; ++bt) {
          const int k = (t + bt * TPC) % s;
#pragma unroll
          for (int r = 1; r < R; ++r) w[o + bt * (R - 1) + r - 1] = tw[r * k * (L / (s * R))];
        }
        o += NB * (R - 1);
      }
      s *= R;
    }
  }
};

// Twiddle of input r of block bt in the stage with span NS: from the table, or from registers.
template <int SIGN, int L, int NS, int R, typename V>
__device__ __forceinline__ V stage_tw(const V* __restrict__ tw, int k, int bt, int r) {
  return twid<SIGN>(tw, r * k * (L / (NS * R)));
}
template <int SIGN, int L, int NS, int R, typename V, typename T, typename P>
__device__ __forceinline__ V stage_tw(const TwRegs<T, L, P>& tw, int, int bt, int r) {
  constexpr int o = TwRegs<T, L, P>::off(NS);
  V w = tw.w[o + bt * (R - 1) + r - 1];
  if (SIGN > 0) w.y = -w.y;
  return w;
}

// LDS index maps: column tile [pos][C] or row tile [row][pos] with padding.
struct LdsCol {
  int base, C;
  __device__ __forceinline__ int operator()(int pos) const { return base + pos * C; }
};
// Row tiles are padded by A entries every 2^S positions: with S >= 4 the padding of a radix-4/8
// stage's accesses separates into a per-thread base plus compile-time offsets (no address VALU).
template <int S, int A> struct LdsRowP {
  int base;
  __device__ __forceinline__ int operator()(int pos) const { return base + pos + A * (pos >> S); }
  // entries one row of M positions (plus the Nyquist slot) occupies
  __host__ __device__ static constexpr int stride(int M) { return M + 2 + A * (M >> S); }
};
using LdsRow = LdsRowP<4, 1>;

// Synchronisation between radix stages: a whole block, or only the calling wavefront when every
// transform lives inside one wavefront (its LDS traffic is then ordered by the wave itself and
// only compiler reordering has to be prevented).
template <bool WAVE> __device__ __forceinline__ void lds_sync() {
  if (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham stage with radix R after a span NS, applied to B independent transforms at once
// (transform b uses LDS region idx(pos) + b * BSTRIDE); recursion over the remaining radices.
// Batching lets one LDS round trip serve B transforms, so the exchange latency is paid once.
template <typename T, int L, int E, int SIGN, int NS, typename Idx, bool WAVE, int B, int BSTRIDE, int R,
          int... Rest>
struct Stages {
  template <typename TW>
  __device__ __forceinline__ static void run(cx<T> (*v)[E], cx<T>* lds, int t, const Idx& idx, const TW& tw) {
    constexpr int TPC = L / E;
    constexpr int NB = E / R;
    constexpr bool LAST = sizeof...(Rest) == 0;
    static_assert(E % R == 0, "radix must divide the elements per thread");
#pragma unroll
    for (int bt = 0; bt < NB; ++bt) {
      const int j = t + bt * TPC;
      const int k = (NS == 1) ? 0 : (j % NS);
      cx<T> w[R];
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = stage_tw<SIGN, L, NS, R, cx<T>>(tw, k, bt, r);
      }
#pragma unroll
      for (int b = 0; b < B; ++b) {
        cx<T> a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[b][bt + r * NB];
        if (NS > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w[r]);
        }
        Dft<R, SIGN, cx<T>>::run(a);
        if (LAST) {
#pragma unroll
          for (int q = 0; q < R; ++q) v[b][bt + q * NB] = a[q];
        } else {
          const int base = (j / NS) * NS * R + k;
#pragma unroll
          for (int q = 0; q < R; ++q) lds[idx(base + q * NS) + b * BSTRIDE] = a[q];
        }
      }
    }
    if constexpr (!LAST) {
      lds_sync<WAVE>();
#pragma unroll
      for (int b = 0; b < B; ++b)
#pragma unroll
        for (int e = 0; e < E; ++e) v[b][e] = lds[idx(t + e * TPC) + b * BSTRIDE];
      lds_sync<WAVE>();
      Stages<T, L, E, SIGN, NS * R, Idx, WAVE, B, BSTRIDE, Rest...>::run(v, lds, t, idx, tw);
    }
  }
};

// Plans for kernels that keep several fields in registers at once: the fewest elements per
// thread that still fit one transform into a 64-lane wavefront.  Lengths with a factor 3 use the
// default plan (at most 32 threads per transform).
template <int L> struct PlanLow : Plan<L> {};
template <> struct PlanLow<8> { static constexpr int E = 4; using R = Radices<4, 2>; };
template <> struct PlanLow<16> { static constexpr int E = 4; using R = Radices<4, 4>; };
template <> struct PlanLow<32> { static constexpr int E = 4; using R = Radices<4, 4, 2>; };
template <> struct PlanLow<64> { static constexpr int E = 4; using R = Radices<4, 4, 4>; };
template <> struct PlanLow<128> { static constexpr int E = 4; using R = Radices<4, 4, 4, 2>; };
template <> struct PlanLow<256> { static constexpr int E = 4; using R = Radices<4, 4, 4, 4>; };
template <> struct PlanLow<512> { static constexpr int E = 8; using R = Radices<8, 8, 8>; };

template <typename T, int L, int SIGN, typename Idx, typename P, bool WAVE, int B, int BSTRIDE, typename Rad>
struct FftL;
template <typename T, int L, int SIGN, typename Idx, typename P, bool WAVE, int B, int BSTRIDE, int... Rs>
struct FftL<T, L, SIGN, Idx, P, WAVE, B, BSTRIDE, Radices<Rs...>> {
  template <typename TW>
  __device__ __forceinline__ static void run(cx<T> (*v)[P::E], cx<T>* lds, int t, const Idx& idx, const TW& tw) {
    Stages<T, L, P::E, SIGN, 1, Idx, WAVE, B, BSTRIDE, Rs...>::run(v, lds, t, idx, tw);
  }
};

// Transform of length L held in strided order by TPC = L / P::E threads.  With WAVE = false it
// contains __syncthreads() (every thread of the block must call it); with WAVE = true the TPC
// threads of one transform must belong to one wavefront.
template <typename T, int L, int SIGN, typename Idx, typename P = Plan<L>, bool WAVE = false>
__device__ __forceinline__ void fft(cx<T>* v, cx<T>* lds, int t, const Idx& idx,
                                    const cx<T>* __restrict__ tw) {
  static_assert(!WAVE || (L / P::E) <= 64, "a wave-synchronised transform must fit one wavefront");
  FftL<T, L, SIGN, Idx, P, WAVE, 1, 0, typename P::R>::run(reinterpret_cast<cx<T>(*)[P::E]>(v), lds, t, idx, tw);
}

// B transforms at once; transform b exchanges through LDS region idx(pos) + b * BSTRIDE.  The
// twiddles come from the table (a pointer) or from a TwRegs of the same length and plan.
template <typename T, int L, int SIGN, typename Idx, typename P, bool WAVE, int B, int BSTRIDE, typename TW>
__device__ __forceinline__ void fft_batch(cx<T> (*v)[P::E], cx<T>* lds, int t, const Idx& idx, const TW& tw) {
  static_assert(!WAVE || (L / P::E) <= 64, "a wave-synchronised transform must fit one wavefront");
  FftL<T, L, SIGN, Idx, P, WAVE, B, BSTRIDE, typename P::R>::run(v, lds, t, idx, tw);
}

// Signed mode index of storage index i along an axis of length n.
__device__ __forceinline__ int mode_of(int i, int n) { return i < (n >> 1) ? i : i - n; }

}  // namespace gnss
