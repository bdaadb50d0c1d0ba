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

// Work rows: u, v, w, omega_x, omega_y, omega_z, (dphi/dx, dphi/dy, dphi/dz) per scalar, and
// with a colloid (the last scalar) two more: c itself and lap Z (Z = scalar 0).
constexpr int MAXF = 3 + 4;           // velocity + up to 4 scalars
constexpr int MAXW = 6 + 3 * 4 + 2;   // work rows of the largest field set

// Outer-index map: identity, or the kept-mode set {0..keep} U {n-keep..n-1} of an axis.
__device__ __forceinline__ int outer_index(int o, int keep, int n) {
  if (keep < 0) return o;
  return o <= keep ? o : n - (2 * keep + 1) + o;
}
// Batched contexts hold several members of one grid back to back (mstride elements apart) and a
// pass runs over every member's outer indices (nouter = members x nouter_m).  Splits outer index o
// into the member offset (returned) and the index within the member (left in o).
__device__ __forceinline__ long long member_split(int& o, int nouter_m, long long mstride) {
  if (nouter_m <= 0) return 0;
  const int m = o / nouter_m;
  o -= m * nouter_m;
  return m * mstride;
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
  int nouter_m = 0;    // batched contexts: outer indices per member (0: one member)
  long long mstride = 0;  // batched contexts: elements between members
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
  int om = active ? o : 0;
  const long long mo = member_split(om, a.nouter_m, a.mstride);
  const long long io = outer_index(om, a.outer_keep, a.outer_n);
  cx<T>* base = data + mo + io * a.ostride + (long long)chunk * C + c;
  cx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + e * TPC;
    if (active && row_kept(p, L, a.in_keep)) v[e] = ld_s(&base[(long long)p * a.estride]);
    else v[e] = mk<T>(0, 0);
  }
  fft<T, L, SIGN>(v, lds, t, LdsCol{g * L * C + c, C}, tw);
  if (!active) return;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + e * TPC;
    if (row_kept(p, L, a.out_keep)) st_s(&base[(long long)p * a.estride], v[e]);
  }
}

// Persistent, software-pipelined column pass over one or several fields: each block walks tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... and issues the HBM loads of its next tile before
// transforming the current one (stage barriers wait only for LDS traffic, so the loads stay in
// flight).  Tile i of the launch is tile i % tpf of field i / tpf: the fields of a pass share one
// launch gap and one tail.
template <typename T> struct FieldList {
  cx<T>* p[MAXW];
};

template <typename T, int L, int SIGN, int C>
__global__ void __launch_bounds__(lb_threads(C * (L / Plan<L>::E)))
k_col_multi(FieldList<T> fl, const cx<T>* __restrict__ tw, ColArgs a, int tpf, int ntiles) {
  constexpr int E = Plan<L>::E;
  constexpr int TPC = L / E;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int t = (tid / C) % TPC;
  const int g = tid / (C * TPC);
  auto base_of = [&](int tile, bool& active) {
    const int f = tile / tpf, lt = tile % tpf;
    const int chunk = lt % a.nchunks;
    const int o = (lt / a.nchunks) * a.G + g;
    active = o < a.nouter;
    int om = active ? o : 0;
    const long long mo = member_split(om, a.nouter_m, a.mstride);
    const long long io = outer_index(om, a.outer_keep, a.outer_n);
    return fl.p[f] + mo + io * a.ostride + (long long)chunk * C + c;
  };
  auto load = [&](int tile, cx<T>* v) {
    bool active;
    const cx<T>* base = base_of(tile, active);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int p = t + e * TPC;
      v[e] = (active && row_kept(p, L, a.in_keep)) ? ld_s(&base[(long long)p * a.estride]) : mk<T>(0, 0);
    }
  };
  int tile = blockIdx.x;
  cx<T> v[E], nv[E];
  load(tile, v);
  // The twiddles are read from a copy in LDS after the G column regions.  Held in registers from
  // loads before the loop, their first use inside it had to wait for every later load, the next
  // tile's included (the prefetch may be skipped, so the compiler cannot count it); the first
  // tile is waited for here for the same reason.
  cx<T>* twl = lds + (long long)a.G * L * C;
  for (int i = tid; i < L; i += blockDim.x) twl[i] = tw[i];
  wait_vmem();
  __syncthreads();
  const LdsCol idx{g * L * C + c, C};
  while (true) {
    const int next = tile + gridDim.x;
    const bool more = next < ntiles;
    if (more) load(next, nv);
    fft<T, L, SIGN>(v, lds, t, idx, twl);
    // One wait per tile, here: the vector-memory counter covers loads and stores alike and, with
    // both in flight, every wait is a wait for all of them.  Here the next tile's loads have had
    // the whole transform to arrive and this tile's stores are not yet issued; left to the
    // compiler, the wait came after the stores (for the copy of the next tile), exposing their
    // full latency once per tile.
    wait_vmem();
    bool active;
    cx<T>* base = base_of(tile, active);
    if (active) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int p = t + e * TPC;
        if (row_kept(p, L, a.out_keep)) st_s(&base[(long long)p * a.estride], v[e]);
      }
    }
    if (!more) break;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nv[e];
    tile = next;
  }
}

// Row stride (complex entries) of the LDS tile used by the z passes: M+1 values plus padding.
__host__ __device__ constexpr int zrow_stride(int M) { return M + 2 + (M >> 4); }

// One contiguous z row held by TPC threads as E complex values per thread.  The real row of NZ
// values is packed as M = NZ/2 complex pairs (z = 2n, 2n+1); the half-length transform is
// unpacked with exp(-2 pi i k / NZ) (R2C) or packed the same way before it (C2R).  TPC divides
// 64, so a row never straddles wavefronts and all row-local LDS traffic is wave-synchronised.
template <typename T, int NZ, typename P, typename Pad = LdsRow>
struct ZRow {
  static constexpr int M = NZ / 2;
  static constexpr int E = P::E;
  static constexpr int TPC = M / E;
  static constexpr int RS = Pad::stride(M);
  static_assert(64 % TPC == 0, "a z row must lie inside one wavefront");

  // Registers of one spectral row as loaded from HBM (entries k >= lim read as 0); thread 0 also
  // holds Re X[M] in nyq (a register of its own: loaded into half of x[0], it would have to wait
  // for the load of x[0] before it could be issued).
  struct Raw {
    cx<T> x[E];
    T nyq;
  };

  __device__ __forceinline__ static void load(const cx<T>* __restrict__ src, bool active, int lim, int t,
                                              Raw& r) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + e * TPC;
      r.x[e] = (active && k < lim) ? ld_s(&src[(unsigned)k]) : mk<T>(0, 0);
    }
    r.nyq = (t == 0 && active && M < lim) ? src[M].x : (T)0;
  }

  // Row formed on loading: SA A + SB i kz B (SA in {-1, 0, 1}, SB in {-1, 1}), entries k >= lim 0;
  // kzr[e]: kz of this thread's position k = t + e TPC, kz: the table (for k = M).  Every load of
  // the row is issued before any value is formed: forming a value inside its load's predicate would
  // wait for that load alone, one HBM round trip per element.
  template <int SA, int SB>
  __device__ __forceinline__ static void load_ikz(const cx<T>* __restrict__ A, const cx<T>* __restrict__ B,
                                                  const T* kzr, const T* __restrict__ kz, bool active, int lim,
                                                  int t, Raw& r) {
    auto form = [&](const cx<T>& b, const cx<T>& x, T kzk) {
      const T kk = (T)SB * kzk;
      cx<T> v = mk<T>(-kk * b.y, kk * b.x);
      if (SA != 0) {
        v.x += (T)SA * x.x;
        v.y += (T)SA * x.y;
      }
      return v;
    };
    cx<T> bv[E], av[SA != 0 ? E : 1];
    bool on[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + e * TPC;
      on[e] = active && k < lim;
      bv[e] = on[e] ? ld_s(&B[(unsigned)k]) : mk<T>(0, 0);
      if (SA != 0) av[e] = on[e] ? ld_s(&A[(unsigned)k]) : mk<T>(0, 0);
    }
    const bool nq = t == 0 && active && M < lim;
    const cx<T> bm = nq ? ld_s(&B[M]) : mk<T>(0, 0);
    const cx<T> am = (nq && SA != 0) ? ld_s(&A[M]) : mk<T>(0, 0);
#pragma unroll
    for (int e = 0; e < E; ++e) r.x[e] = on[e] ? form(bv[e], av[SA != 0 ? e : 0], kzr[e]) : mk<T>(0, 0);
    r.nyq = nq ? form(bm, am, kz[M]).x : (T)0;
  }

  // Twiddles of one thread for every row it transforms, read once and kept in registers: the
  // radix stages' (TwRegs) and the packing factors twN[k] of its positions k = t + e TPC.  As
  // (E/2) TPC = NZ/4, twN[k] of register e >= E/2 is -i times that of register e - E/2.
  static_assert(E % 2 == 0, "register twiddles need an even number of elements per thread");
  struct Tw {
    TwRegs<T, M, P> m;
    cx<T> n[E / 2];
    __device__ __forceinline__ void load(const cx<T>* __restrict__ twM, const cx<T>* __restrict__ twN, int t) {
      m.load(twM, t);
#pragma unroll
      for (int e = 0; e < E / 2; ++e) n[e] = twN[t + e * TPC];
    }
  };
  __device__ __forceinline__ static cx<T> ntw(const cx<T>* __restrict__ twN, int k, int) { return twN[k]; }
  __device__ __forceinline__ static cx<T> ntw(const Tw& tw, int, int e) {
    return e < E / 2 ? tw.n[e] : mk<T>(tw.n[e - E / 2].y, -tw.n[e - E / 2].x);
  }
  __device__ __forceinline__ static const cx<T>* mtw(const cx<T>* twM, const cx<T>*) { return twM; }
  __device__ __forceinline__ static const TwRegs<T, M, P>& mtw(const Tw& tw, const Tw&) { return tw.m; }

  // Value of register `e` in lane TPC - t of this row (t = 0 reads its own lane): the partner
  // index M - k of k = t + e TPC lives there in register E - 1 - e.
  __device__ __forceinline__ static cx<T> mirror(const cx<T>& x, int t) {
    const int src = (TPC - t) & (TPC - 1);
    cx<T> r;
    r.x = __shfl(x.x, src, TPC);
    r.y = __shfl(x.y, src, TPC);
    return r;
  }

  // C2R of B loaded rows to packed real rows, scaled by `scale` unless SCALE is false (Im of k = 0
  // and k = M ignored).  The pairing of k with M - k is a lane shuffle; thread 0, whose partners
  // are its own registers, selects what it sends before the shuffle, so the shuffles of all B rows
  // go out together.  Only the radix stages use LDS, and the B rows go through each stage together
  // (LDS region of row b at idx + b * RS).  twM, twN: the tables (pointers), or both the same Tw.
  template <int B, bool SCALE = true, typename TM, typename TN>
  __device__ __forceinline__ static void c2r(const Raw* r, cx<T>* lds, const Pad& idx, int t, const TM& twM,
                                             const TN& twN, T scale, cx<T> (*v)[E]) {
    const bool t0 = t == 0;
#pragma unroll
    for (int b = 0; b < B; ++b) {
      cx<T> x[E];
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = r[b].x[e];
      const T xm = r[b].nyq;  // Re X[M] in thread 0 (a C2R ignores Im X[0] and Im X[M])
      x[0].y = t0 ? (T)0 : x[0].y;
      cx<T> xc[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        // thread 0 pairs k = e TPC with M - k = (E - e) TPC, its own register E - e (M itself for e = 0)
        const cx<T> own = e == 0 ? mk<T>(xm, (T)0) : x[E - e];
        cx<T> snd = x[E - 1 - e];
        snd.x = t0 ? own.x : snd.x;
        snd.y = t0 ? own.y : snd.y;
        xc[e] = mirror(snd, t);
      }
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = t + e * TPC;
        const cx<T> xk = x[e];
        const cx<T> c = cconj(xc[e]);
        const cx<T> p = cmul(cconj(ntw(twN, k, e)), csub(xk, c));
        cx<T> z;
        z.x = xk.x + c.x - p.y;
        z.y = xk.y + c.y + p.x;
        v[b][e] = z;
      }
    }
    lds_sync<true>();
    fft_batch<T, M, +1, Pad, P, true, B, RS>(v, lds, t, idx, mtw(twM, twN));
    if constexpr (SCALE) {
#pragma unroll
      for (int b = 0; b < B; ++b)
#pragma unroll
        for (int e = 0; e < E; ++e) v[b][e] = cscale(v[b][e], scale);
    }
  }

  // Transform B packed real rows to half spectra; entries k < lim of row b are stored to dst[b]
  // (the Nyquist value at k = M and zero padding up to kzs included), multiplied by 2 h (h = 1/2:
  // the plain transform; the fused z pass folds the inverse transforms' 1/N^2 into it).
  template <int B, typename TM, typename TN>
  __device__ __forceinline__ static void r2c(cx<T> (*v)[E], cx<T>* const* dst, bool active, int lim,
                                             int kzs, cx<T>* lds, const Pad& idx, int t, const TM& twM,
                                             const TN& twN, T h = (T)0.5, bool drain = false) {
    lds_sync<true>();
    fft_batch<T, M, -1, Pad, P, true, B, RS>(v, lds, t, idx, mtw(twM, twN));
    // drain: wait for every memory operation in flight before the stores (a pipelined caller's
    // next rows, loaded before the transform; see wait_vmem)
    if (drain) wait_vmem();
    const bool t0 = t == 0;
#pragma unroll
    for (int b = 0; b < B; ++b) {
      cx<T> zc[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const cx<T> own = v[b][(E - e) % E];  // thread 0: k = e TPC pairs with its register (E - e) mod E
        cx<T> snd = v[b][E - 1 - e];
        snd.x = t0 ? own.x : snd.x;
        snd.y = t0 ? own.y : snd.y;
        zc[e] = mirror(snd, t);
      }
      if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int k = t + e * TPC;
          if (k < lim) {
            const cx<T> zk = v[b][e];
            const cx<T> c = cconj(zc[e]);
            const cx<T> p = cmul(ntw(twN, k, e), csub(zk, c));
            cx<T> x;
            x.x = h * (zk.x + c.x + p.y);
            x.y = h * (zk.y + c.y - p.x);
            st_s(&dst[b][(unsigned)k], x);
          }
        }
        if (t0 && M < lim) st_s(&dst[b][M], mk<T>(((T)2 * h) * (v[b][0].x - v[b][0].y), (T)0));
        for (int k = M + 1 + t; k < kzs && k < lim; k += TPC) st_s(&dst[b][k], mk<T>(0, 0));
      }
    }
  }
};

// R2C along contiguous z rows: in = rows of NZ reals, out = rows of kzs complex (kz < out_lim
// stored; entries between nz/2 and kzs are zero padding).
template <typename T, int NZ>
__global__ void __launch_bounds__(lb_threads((NZ / 2) / Plan<NZ / 2>::E))
k_r2c_z(const T* __restrict__ in, cx<T>* __restrict__ out, const cx<T>* __restrict__ twM,
        const cx<T>* __restrict__ twN, long long nrows, int kzs, int out_lim, int G) {
  using Z = ZRow<T, NZ, Plan<NZ / 2>>;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int t = threadIdx.x % Z::TPC;
  const int g = threadIdx.x / Z::TPC;
  const long long row = (long long)blockIdx.x * G + g;
  const bool active = row < nrows;
  const cx<T>* src = reinterpret_cast<const cx<T>*>(in + (active ? row : 0) * NZ);
  cx<T> v[1][Z::E];
#pragma unroll
  for (int e = 0; e < Z::E; ++e) v[0][e] = active ? ld_s(&src[t + e * Z::TPC]) : mk<T>(0, 0);
  cx<T>* dst[1] = {out + (active ? row : 0) * kzs};
  Z::template r2c<1>(v, dst, active, out_lim, kzs, lds, LdsRow{g * Z::RS}, t, twM, twN);
}

// C2R along contiguous z rows: in = rows of kzs complex (entries kz >= in_lim are treated as 0;
// imaginary parts of kz = 0 and kz = nz/2 are ignored), out = rows of NZ reals times `scale`.
template <typename T, int NZ>
__global__ void __launch_bounds__(lb_threads((NZ / 2) / Plan<NZ / 2>::E))
k_c2r_z(const cx<T>* __restrict__ in, T* __restrict__ out, const cx<T>* __restrict__ twM,
        const cx<T>* __restrict__ twN, long long nrows, int kzs, int in_lim, T scale, int G) {
  using Z = ZRow<T, NZ, Plan<NZ / 2>>;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int t = threadIdx.x % Z::TPC;
  const int g = threadIdx.x / Z::TPC;
  const long long row = (long long)blockIdx.x * G + g;
  const bool active = row < nrows;
  typename Z::Raw r[1];
  Z::load(in + (active ? row : 0) * kzs, active, in_lim, t, r[0]);
  cx<T> v[1][Z::E];
  Z::template c2r<1>(r, lds, LdsRow{g * Z::RS}, t, twM, twN, scale, v);
  if (!active) return;
  cx<T>* dst = reinterpret_cast<cx<T>*>(out + row * NZ);
#pragma unroll
  for (int e = 0; e < Z::E; ++e) st_s(&dst[t + e * Z::TPC], v[0][e]);
}

// Persistent, software-pipelined forms of k_r2c_z / k_c2r_z: each block walks row groups
// blockIdx.x, blockIdx.x + gridDim.x, ... and issues the HBM loads of its next rows before
// transforming the current ones.  Every row lives in one wavefront (wave-level LDS sync only),
// so the waves of a block never wait for each other.  Both twiddle tables are copied to LDS after
// the G row regions (M entries each): read from memory inside the loop, every twiddle wait would
// also wait for the next rows' loads.
template <typename T, int M>
__device__ __forceinline__ cx<T>* zpipe_twiddles(cx<T>* lds, int G, int RS, const cx<T>* __restrict__ twM,
                                                 const cx<T>* __restrict__ twN) {
  cx<T>* tl = lds + (long long)G * RS;
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    tl[i] = twM[i];
    tl[M + i] = twN[i];
  }
  wait_vmem();  // the first rows too: the loop then starts with no load of unknown count in flight
  __syncthreads();
  return tl;
}
template <typename T, int NZ>
__global__ void __launch_bounds__(lb_threads((NZ / 2) / Plan<NZ / 2>::E))
k_r2c_z_pipe(const T* __restrict__ in, cx<T>* __restrict__ out, const cx<T>* __restrict__ twM,
             const cx<T>* __restrict__ twN, long long nrows, int kzs, int out_lim, int G, long long ngroups) {
  using Z = ZRow<T, NZ, Plan<NZ / 2>>;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int t = threadIdx.x % Z::TPC;
  const int g = threadIdx.x / Z::TPC;
  auto load = [&](long long grp, cx<T>* v) {
    const long long row = grp * G + g;
    const bool active = row < nrows;
    const cx<T>* src = reinterpret_cast<const cx<T>*>(in + (active ? row : 0) * NZ);
#pragma unroll
    for (int e = 0; e < Z::E; ++e) v[e] = active ? ld_s(&src[t + e * Z::TPC]) : mk<T>(0, 0);
  };
  long long grp = blockIdx.x;
  cx<T> v[1][Z::E], nv[Z::E];
  load(grp, v[0]);
  const cx<T>* tl = zpipe_twiddles<T, Z::M>(lds, G, Z::RS, twM, twN);
  while (true) {
    const long long next = grp + gridDim.x;
    const bool more = next < ngroups;
    if (more) load(next, nv);
    const long long row = grp * G + g;
    const bool active = row < nrows;
    cx<T>* dst[1] = {out + (active ? row : 0) * kzs};
    Z::template r2c<1>(v, dst, active, out_lim, kzs, lds, LdsRow{g * Z::RS}, t, tl, tl + Z::M, (T)0.5, true);
    if (!more) break;
#pragma unroll
    for (int e = 0; e < Z::E; ++e) v[0][e] = nv[e];
    grp = next;
  }
}

template <typename T, int NZ>
__global__ void __launch_bounds__(lb_threads((NZ / 2) / Plan<NZ / 2>::E))
k_c2r_z_pipe(const cx<T>* __restrict__ in, T* __restrict__ out, const cx<T>* __restrict__ twM,
             const cx<T>* __restrict__ twN, long long nrows, int kzs, int in_lim, T scale, int G,
             long long ngroups) {
  using Z = ZRow<T, NZ, Plan<NZ / 2>>;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int t = threadIdx.x % Z::TPC;
  const int g = threadIdx.x / Z::TPC;
  long long grp = blockIdx.x;
  typename Z::Raw r[1], rn;
  {
    const long long row = grp * G + g;
    Z::load(in + (row < nrows ? row : 0) * kzs, row < nrows, in_lim, t, r[0]);
  }
  const cx<T>* tl = zpipe_twiddles<T, Z::M>(lds, G, Z::RS, twM, twN);
  while (true) {
    const long long next = grp + gridDim.x;
    const bool more = next < ngroups;
    if (more) {
      const long long row = next * G + g;
      Z::load(in + (row < nrows ? row : 0) * kzs, row < nrows, in_lim, t, rn);
    }
    const long long row = grp * G + g;
    const bool active = row < nrows;
    cx<T> v[1][Z::E];
    Z::template c2r<1>(r, lds, LdsRow{g * Z::RS}, t, tl, tl + Z::M, scale, v);
    wait_vmem();  // the next rows, before this group's stores (see wait_vmem)
    if (active) {
      cx<T>* dst = reinterpret_cast<cx<T>*>(out + row * NZ);
#pragma unroll
      for (int e = 0; e < Z::E; ++e) st_s(&dst[t + e * Z::TPC], v[0][e]);
    }
    if (!more) break;
    r[0] = rn;
    grp = next;
  }
}

// Variants of the fused z kernel (chosen by the autotuner): V = 0 the fewest elements per thread
// with the next rows' loads issued early (FP32), V = 1 the default plan (fewer radix stages, more
// registers), V = 2 the fewest elements per thread without early loads (fewest registers),
// V = 3 as 2 with a wider LDS row padding (fewer bank conflicts in the radix-4 exchanges: 5 entries
// every 32 for 8-byte, 3 every 16 for 16-byte elements), V = 4 the default plan without early loads,
// V = 5 as 3 with the register budget of four waves per SIMD, V = 6 as 3 with each thread's
// twiddles read once and held in registers (no twiddle load inside a radix stage), V = 7 as 6 with
// the inverse transforms' 1/N moved to the forward ones, V = 8 as 7 with the 16-byte rows padded
// by 2 entries every 8 (a 32-bank model of the FP64 exchanges puts fewer conflicts there).
template <typename T, int M, int V> struct FusedPlan {
  using type = PlanLow<M>;
  static constexpr bool prefetch = sizeof(T) == 4;
  using pad = LdsRow;
  static constexpr int min_waves = 1;  // waves per SIMD the register budget must allow
  static constexpr bool regtw = false;  // twiddles read once per thread and kept in registers
  static constexpr bool defer = false;  // 1/N of the inverse transforms applied as 1/N^2 on the products
};
template <typename T, int M> struct FusedPlan<T, M, 1> {
  using type = Plan<M>;
  static constexpr bool prefetch = sizeof(T) == 4;
  using pad = LdsRow;
  static constexpr int min_waves = 1;
  static constexpr bool regtw = false;
  static constexpr bool defer = false;
};
template <typename T, int M> struct FusedPlan<T, M, 2> {
  using type = PlanLow<M>;
  static constexpr bool prefetch = false;
  using pad = LdsRow;
  static constexpr int min_waves = 1;
  static constexpr bool regtw = false;
  static constexpr bool defer = false;
};
template <typename T, int M> struct FusedPlan<T, M, 3> {
  using type = PlanLow<M>;
  static constexpr bool prefetch = false;
  using pad = typename std::conditional<sizeof(T) == 4, LdsRowP<5, 5>, LdsRowP<4, 3>>::type;
  static constexpr int min_waves = 1;
  static constexpr bool regtw = false;
  static constexpr bool defer = false;
};
template <typename T, int M> struct FusedPlan<T, M, 4> {
  using type = Plan<M>;
  static constexpr bool prefetch = false;
  using pad = LdsRow;
  static constexpr int min_waves = 1;
  static constexpr bool regtw = false;
  static constexpr bool defer = false;
};
template <typename T, int M> struct FusedPlan<T, M, 5> : FusedPlan<T, M, 3> {
  static constexpr int min_waves = 4;
};
template <typename T, int M> struct FusedPlan<T, M, 6> : FusedPlan<T, M, 3> {
  static constexpr bool regtw = true;
};
template <typename T, int M> struct FusedPlan<T, M, 7> : FusedPlan<T, M, 6> {
  static constexpr bool defer = true;
};
template <typename T, int M> struct FusedPlan<T, M, 8> : FusedPlan<T, M, 7> {
  using pad = typename std::conditional<sizeof(T) == 4, LdsRowP<5, 5>, LdsRowP<3, 2>>::type;
};
// Half spectra of the variant below as loaded: positions k = t + e TPC (e < E/2) of the full-length
// plan's threads, i.e. half its elements per thread with the same threads per row.
template <int NZ> struct HalfPlan {
  static constexpr int E = Plan<NZ>::E / 2;
  using R = typename Plan<NZ>::R;
};
// V = 9: two real rows per complex transform of the full row length (z = a + i b), eight elements
// per thread in three radix-8 stages at 512 (two LDS exchanges instead of three, and no pairing of
// k with M - k per row); inverse 1/N on the forward side as in 7; twiddles in registers.
template <typename T, int M> struct FusedPlan<T, M, 9> {
  using type = HalfPlan<2 * M>;
  static constexpr bool prefetch = false;
  using pad = typename std::conditional<sizeof(T) == 4, LdsRowP<5, 5>, LdsRowP<4, 3>>::type;
  static constexpr int min_waves = 1;
  static constexpr bool regtw = true;
  static constexpr bool defer = true;
};


// kz entries a substage's rows hold: kz <= nz/3 (the 2/3 rule), rounded up to the widest column tile
// of the strided passes (16 complex in FP32, 8 in FP64); the engine's pruned kz limit.
template <typename T> __host__ __device__ constexpr int zfused_lim(int nz) {
  return (nz / 3 + 1 + (sizeof(T) == 8 ? 7 : 15)) / (sizeof(T) == 8 ? 8 : 16) * (sizeof(T) == 8 ? 8 : 16);
}

template <typename T> struct ZFusedArgs {
  const cx<T>* ws[MAXW];
  cx<T>* nl[MAXF];       // (u x omega)_x,y,z, then the advection term of each scalar
  long long nrows;
  int kzs, S, G;
  int colloid;           // 1: scalar S - 1 is a colloid drifting with alpha grad(scalar 0)
  T alpha;
  // late-derivative mode: the rows hold u, v, w, w_y, w_x, omega_z and per scalar phi_x, phi_y, phi;
  // omega_x = w_y - i kz v, omega_y = i kz u - w_x and phi_z = i kz phi are formed on loading
  int late;
  const T* kz;           // kz wavenumber of each stored z mode
};

template <typename V> __device__ __forceinline__ void fma_pair(V& acc, const V& a, const V& b) {
  acc.x += a.x * b.x;
  acc.y += a.y * b.y;
}
template <typename V> __device__ __forceinline__ void fms_pair(V& acc, const V& a, const V& b) {
  acc.x -= a.x * b.x;
  acc.y -= a.y * b.y;
}
template <typename V> __device__ __forceinline__ V mul_pair(const V& a, const V& b) {
  V r;
  r.x = a.x * b.x;
  r.y = a.y * b.y;
  return r;
}

// Fused z pass of one substage: C2R of the work rows, the physical products u x omega and
// u . grad phi (crossProductKernel_mgpu / dotKernel_mgpu, solver.cu:139-164, 342-356; the colloid
// term of colloidAdvectionKernel_mgpu, solver.cu:374-391), and R2C of the 3 + S products — the
// physical fields never leave registers.  Rows hold kz < lim only.
// Fields move through the radix stages three at a time (u, v, w; omega; products; grad phi),
// and in FP32 the HBM loads of the next three rows are issued before the current ones transform.
// Each row owns FUSED_B LDS regions of RS entries.
constexpr int FUSED_B = 3;
template <typename T, int NZ, int V, typename TM, typename TN>
__device__ __forceinline__ void zfused_row(const ZFusedArgs<T>& a, const TM& twM, const TN& twN, T scale,
                                           cx<T>* lds, int t, int g) {
  using FP = FusedPlan<T, NZ / 2, V>;
  using Z = ZRow<T, NZ, typename FP::type, typename FP::pad>;
  constexpr int E = Z::E;
  constexpr bool PREFETCH = FP::prefetch;
  // Variants with `defer`: the inverse transforms skip their 1/N; every product below is quadratic
  // in the transformed fields, so the forward transforms take 1/N^2 instead (h = 1/(2 N^2) in place
  // of 1/2, no extra instruction).
  constexpr bool DEFER = FP::defer;
  const T h = DEFER ? (T)0.5 * scale * scale : (T)0.5;
  // the pruned row length is the same for every launch of a grid (kz below the 2/3 rule, rounded up
  // to the widest column tile): as a constant it lets the compiler drop the loads, shuffles and
  // packing of the entries known to be zero and the unpacking of the entries never stored
  constexpr int LIM = zfused_lim<T>(NZ);
  const long long row = (long long)blockIdx.x * a.G + g;
  const bool active = row < a.nrows;
  const long long off = (active ? row : 0) * a.kzs;
  const typename FP::pad idx{g * Z::RS * FUSED_B};
  typename Z::Raw r0[3], r1[3];
  cx<T> vel[3][E], om[3][E];
#pragma unroll
  for (int b = 0; b < 3; ++b) Z::load(a.ws[b] + off, active, LIM, t, r0[b]);
  // late mode: kz of this thread's positions, read once per row
  T kzr[E];
#pragma unroll
  for (int e = 0; e < E; ++e) kzr[e] = (a.late && t + e * Z::TPC < LIM) ? a.kz[t + e * Z::TPC] : (T)0;
  auto load_omega = [&]() {
    if (a.late) {
      Z::template load_ikz<1, -1>(a.ws[3] + off, a.ws[1] + off, kzr, a.kz, active, LIM, t, r1[0]);  // w_y - i kz v
      Z::template load_ikz<-1, 1>(a.ws[4] + off, a.ws[0] + off, kzr, a.kz, active, LIM, t, r1[1]);  // i kz u - w_x
      Z::load(a.ws[5] + off, active, LIM, t, r1[2]);
    } else {
#pragma unroll
      for (int b = 0; b < 3; ++b) Z::load(a.ws[3 + b] + off, active, LIM, t, r1[b]);
    }
  };
  auto load_grad = [&](int s) {
    const int w0 = 6 + 3 * s;
    if (a.late) {
      Z::load(a.ws[w0] + off, active, LIM, t, r0[0]);
      Z::load(a.ws[w0 + 1] + off, active, LIM, t, r0[1]);
      Z::template load_ikz<0, 1>(a.ws[w0 + 2] + off, a.ws[w0 + 2] + off, kzr, a.kz, active, LIM, t, r0[2]);
    } else {
#pragma unroll
      for (int b = 0; b < 3; ++b) Z::load(a.ws[w0 + b] + off, active, LIM, t, r0[b]);
    }
  };
  if (PREFETCH) load_omega();
  Z::template c2r<3, !DEFER>(r0, lds, idx, t, twM, twN, scale, vel);  // u, v, w
  if (!PREFETCH) load_omega();
  // Plain scalars; with a colloid, scalar 0 (its carrier Z) goes last so that the velocity
  // registers can then take the drift velocity u + alpha grad Z.
  const int P = a.S - a.colloid;
  const int s_first = (a.colloid && P > 1) ? 1 : 0;
  if (PREFETCH && P > 0) load_grad(s_first);
  Z::template c2r<3, !DEFER>(r1, lds, idx, t, twM, twN, scale, om);  // omega
  // u x omega, in place of omega
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const cx<T> u = vel[0][e], v = vel[1][e], w = vel[2][e];
    const cx<T> ox = om[0][e], oy = om[1][e], oz = om[2][e];
    om[0][e] = mk<T>(v.x * oz.x - w.x * oy.x, v.y * oz.y - w.y * oy.y);
    om[1][e] = mk<T>(w.x * ox.x - u.x * oz.x, w.y * ox.y - u.y * oz.y);
    om[2][e] = mk<T>(u.x * oy.x - v.x * ox.x, u.y * oy.y - v.y * ox.y);
  }
  {
    cx<T>* dst[3] = {a.nl[0] + off, a.nl[1] + off, a.nl[2] + off};
    Z::template r2c<3>(om, dst, active, LIM, a.kzs, lds, idx, t, twM, twN, h);
  }
  for (int j = 0; j < P; ++j) {
    const int s = a.colloid ? (j + 1 < P ? j + 1 : 0) : j;
    if (!PREFETCH || j > 0) load_grad(s);
    Z::template c2r<3, !DEFER>(r0, lds, idx, t, twM, twN, scale, om);  // grad phi
    cx<T> n[1][E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const cx<T> u = vel[0][e], v = vel[1][e], w = vel[2][e];
      n[0][e] = mk<T>(u.x * om[0][e].x + v.x * om[1][e].x + w.x * om[2][e].x,
                      u.y * om[0][e].y + v.y * om[1][e].y + w.y * om[2][e].y);
    }
    if (a.colloid && s == 0) {
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int e = 0; e < E; ++e) fma_pair(vel[b][e], mk<T>(a.alpha, a.alpha), om[b][e]);
    }
    cx<T>* dst[1] = {a.nl[3 + s] + off};
    Z::template r2c<1>(n, dst, active, LIM, a.kzs, lds, idx, t, twM, twN, h);
  }
  if (a.colloid) {
    // colloidAdvectionKernel_mgpu (solver.cu:374-391): u.grad c + alpha (lap Z c + grad Z . grad c)
    // = (u + alpha grad Z) . grad c + alpha lap Z c, with u + alpha grad Z already in vel.
    const int sc = a.S - 1;
#pragma unroll
    for (int b = 0; b < 3; ++b) Z::load(a.ws[6 + 3 * sc + b] + off, active, LIM, t, r0[b]);
    Z::template c2r<3, !DEFER>(r0, lds, idx, t, twM, twN, scale, om);  // grad c
    cx<T> n[1][E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const cx<T> u = vel[0][e], v = vel[1][e], w = vel[2][e];
      n[0][e] = mk<T>(u.x * om[0][e].x + v.x * om[1][e].x + w.x * om[2][e].x,
                      u.y * om[0][e].y + v.y * om[1][e].y + w.y * om[2][e].y);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) Z::load(a.ws[6 + 3 * a.S + b] + off, active, LIM, t, r0[b]);
    Z::template c2r<2, !DEFER>(r0, lds, idx, t, twM, twN, scale, om);  // c, lap Z
#pragma unroll
    for (int e = 0; e < E; ++e) fma_pair(n[0][e], mk<T>(a.alpha * om[1][e].x, a.alpha * om[1][e].y), om[0][e]);
    cx<T>* dst[1] = {a.nl[3 + sc] + off};
    Z::template r2c<1>(n, dst, active, LIM, a.kzs, lds, idx, t, twM, twN, h);
  }
}

// Two real z rows a, b of length N = NZ as one complex row z = a + i b through one N-point transform.
// The C2R side builds Z = A + i B at all N positions from the two loaded half spectra (A and B are
// Hermitian: Z[N - k] = conj(A[k]) + i conj(B[k]), the partner entries a lane shuffle away); the
// R2C side splits Z = FFT(p1 + i p2) into P1[k] = (Z[k] + conj Z[N - k]) / 2 and
// P2[k] = (Z[k] - conj Z[N - k]) / 2i.  Positions k = t + e TPC, e < E/2, are the half spectrum;
// the partner N - k of a position in the upper half lives in lane TPC - t, register E/2 - 1 - j
// (j = e - E/2), or, in lane 0, in its own register E/2 - j (the Nyquist entry for j = 0).
template <typename T, int NZ, typename Pad>
struct ZPair {
  using P = Plan<NZ>;
  using H = ZRow<T, NZ, HalfPlan<NZ>>;  // loaders of one half spectrum (k < N/2, Nyquist in nyq)
  using Raw = typename H::Raw;
  static constexpr int N = NZ;
  static constexpr int E = P::E;
  static constexpr int HE = E / 2;
  static constexpr int TPC = N / E;
  static constexpr int RS = Pad::stride(N);
  static_assert(H::TPC == TPC && H::E == HE, "half rows and full rows must share their threads");
  static_assert(64 % TPC == 0, "a z row must lie inside one wavefront");

  __device__ __forceinline__ static cx<T> mirror(const cx<T>& x, int t) {
    const int src = (TPC - t) & (TPC - 1);
    cx<T> r;
    r.x = __shfl(x.x, src, TPC);
    r.y = __shfl(x.y, src, TPC);
    return r;
  }
  // Z at all positions from the half spectra ra, rb (their Im X[0] and Im X[N/2] ignored)
  __device__ __forceinline__ static void pack(const Raw& ra, const Raw& rb, int t, cx<T>* z) {
    const bool t0 = t == 0;
    cx<T> xa[HE], xb[HE];
#pragma unroll
    for (int e = 0; e < HE; ++e) {
      xa[e] = ra.x[e];
      xb[e] = rb.x[e];
    }
    xa[0].y = t0 ? (T)0 : xa[0].y;
    xb[0].y = t0 ? (T)0 : xb[0].y;
    cx<T> pa[HE], pb[HE];
#pragma unroll
    for (int j = 0; j < HE; ++j) {
      const cx<T> oa = j == 0 ? mk<T>(ra.nyq, (T)0) : xa[HE - j];
      const cx<T> ob = j == 0 ? mk<T>(rb.nyq, (T)0) : xb[HE - j];
      cx<T> sa = xa[HE - 1 - j], sb = xb[HE - 1 - j];
      sa.x = t0 ? oa.x : sa.x;
      sa.y = t0 ? oa.y : sa.y;
      sb.x = t0 ? ob.x : sb.x;
      sb.y = t0 ? ob.y : sb.y;
      pa[j] = mirror(sa, t);
      pb[j] = mirror(sb, t);
    }
#pragma unroll
    for (int e = 0; e < HE; ++e) z[e] = mk<T>(xa[e].x - xb[e].y, xa[e].y + xb[e].x);
#pragma unroll
    for (int j = 0; j < HE; ++j) z[HE + j] = mk<T>(pa[j].x + pb[j].y, pb[j].x - pa[j].y);
  }
  // P1, P2 of Z (after the forward transform) times 2h; entries k < lim stored to d1 / d2 (d2 may be
  // null: a lone row), with the Nyquist value at k = N/2 and zero padding up to kzs
  __device__ __forceinline__ static void unpack_store(const cx<T>* z, cx<T>* d1, cx<T>* d2, bool active, int lim,
                                                      int kzs, int t, T h) {
    const bool t0 = t == 0;
    cx<T> zm[HE];
#pragma unroll
    for (int e = 0; e < HE; ++e) {
      const cx<T> own = z[(E - e) % E];
      cx<T> snd = z[E - 1 - e];
      snd.x = t0 ? own.x : snd.x;
      snd.y = t0 ? own.y : snd.y;
      zm[e] = mirror(snd, t);
    }
    if (!active) return;
#pragma unroll
    for (int e = 0; e < HE; ++e) {
      const int k = t + e * TPC;
      if (k < lim) {
        st_s(&d1[(unsigned)k], mk<T>(h * (z[e].x + zm[e].x), h * (z[e].y - zm[e].y)));
        if (d2) st_s(&d2[(unsigned)k], mk<T>(h * (z[e].y + zm[e].y), h * (zm[e].x - z[e].x)));
      }
    }
    if (t0 && N / 2 < lim) {
      st_s(&d1[N / 2], mk<T>(((T)2 * h) * z[HE].x, (T)0));
      if (d2) st_s(&d2[N / 2], mk<T>(((T)2 * h) * z[HE].y, (T)0));
    }
    for (int k = N / 2 + 1 + t; k < kzs && k < lim; k += TPC) {
      st_s(&d1[k], mk<T>(0, 0));
      if (d2) st_s(&d2[k], mk<T>(0, 0));
    }
  }
};

// Fused z pass in pairs of rows (variant 9, no colloid): C2R of (u, v), (w, omega_x),
// (omega_y, omega_z), then per scalar (phi_x, phi_y), (phi_z, 0); R2C of the products paired in
// order ((u x omega)_x, (u x omega)_y), ((u x omega)_z, u.grad phi_0), (u.grad phi_1, ...) ...
constexpr int PAIR_B = 2;  // packed rows per radix-stage batch (LDS regions per row of the block)
template <typename T, int NZ, typename Pad, typename TW>
__device__ __forceinline__ void zfused_pair_row(const ZFusedArgs<T>& a, const TW& tw, T scale, cx<T>* lds, int t,
                                                int g) {
  using Z = ZPair<T, NZ, Pad>;
  using H = typename Z::H;
  using P = typename Z::P;
  constexpr int E = Z::E, HE = Z::HE, N = Z::N, RS = Z::RS;
  constexpr int LIM = zfused_lim<T>(NZ);
  const T h = (T)0.5 * scale * scale;  // the inverse transforms' 1/N taken here as 1/N^2
  const long long row = (long long)blockIdx.x * a.G + g;
  const bool active = row < a.nrows;
  const long long off = (active ? row : 0) * a.kzs;
  const Pad idx{g * RS * PAIR_B};
  T kzr[HE];
#pragma unroll
  for (int e = 0; e < HE; ++e) kzr[e] = (a.late && t + e * Z::TPC < LIM) ? a.kz[t + e * Z::TPC] : (T)0;
  typename Z::Raw r[4];
  // velocity and vorticity: Z1 = u + i v, Z2 = w + i omega_x, then Z3 = omega_y + i omega_z
  H::load(a.ws[0] + off, active, LIM, t, r[0]);
  H::load(a.ws[1] + off, active, LIM, t, r[1]);
  H::load(a.ws[2] + off, active, LIM, t, r[2]);
  if (a.late) H::template load_ikz<1, -1>(a.ws[3] + off, a.ws[1] + off, kzr, a.kz, active, LIM, t, r[3]);
  else H::load(a.ws[3] + off, active, LIM, t, r[3]);
  cx<T> za[2][E];
  Z::pack(r[0], r[1], t, za[0]);
  Z::pack(r[2], r[3], t, za[1]);
  lds_sync<true>();
  fft_batch<T, N, +1, Pad, P, true, 2, RS>(za, lds, t, idx, tw);
  if (a.late) H::template load_ikz<-1, 1>(a.ws[4] + off, a.ws[0] + off, kzr, a.kz, active, LIM, t, r[0]);
  else H::load(a.ws[4] + off, active, LIM, t, r[0]);
  H::load(a.ws[5] + off, active, LIM, t, r[1]);
  cx<T> zb[1][E];
  Z::pack(r[0], r[1], t, zb[0]);
  lds_sync<true>();
  fft_batch<T, N, +1, Pad, P, true, 1, RS>(zb, lds, t, idx, tw);
  // u x omega; the velocity stays in za[0] (u + i v) and w
  T w[E], pend[E];
  cx<T> q[1][E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const T u = za[0][e].x, v = za[0][e].y, ww = za[1][e].x;
    const T ox = za[1][e].y, oy = zb[0][e].x, oz = zb[0][e].y;
    q[0][e] = mk<T>(v * oz - ww * oy, ww * ox - u * oz);
    pend[e] = u * oy - v * ox;
    w[e] = ww;
  }
  lds_sync<true>();
  fft_batch<T, N, -1, Pad, P, true, 1, RS>(q, lds, t, idx, tw);
  Z::unpack_store(q[0], a.nl[0] + off, a.nl[1] + off, active, LIM, a.kzs, t, h);
  int pend_f = 2;  // product row waiting for a partner
  const typename Z::Raw zero{};
  for (int s = 0; s < a.S; ++s) {
    const int w0 = 6 + 3 * s;
    H::load(a.ws[w0] + off, active, LIM, t, r[0]);
    H::load(a.ws[w0 + 1] + off, active, LIM, t, r[1]);
    if (a.late) H::template load_ikz<0, 1>(a.ws[w0 + 2] + off, a.ws[w0 + 2] + off, kzr, a.kz, active, LIM, t, r[2]);
    else H::load(a.ws[w0 + 2] + off, active, LIM, t, r[2]);
    cx<T> zg[2][E];
    Z::pack(r[0], r[1], t, zg[0]);
    Z::pack(r[2], zero, t, zg[1]);
    lds_sync<true>();
    fft_batch<T, N, +1, Pad, P, true, 2, RS>(zg, lds, t, idx, tw);
    T n[E];
#pragma unroll
    for (int e = 0; e < E; ++e) n[e] = za[0][e].x * zg[0][e].x + za[0][e].y * zg[0][e].y + w[e] * zg[1][e].x;
    if (pend_f >= 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) q[0][e] = mk<T>(pend[e], n[e]);
      lds_sync<true>();
      fft_batch<T, N, -1, Pad, P, true, 1, RS>(q, lds, t, idx, tw);
      Z::unpack_store(q[0], a.nl[pend_f] + off, a.nl[3 + s] + off, active, LIM, a.kzs, t, h);
      pend_f = -1;
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) pend[e] = n[e];
      pend_f = 3 + s;
    }
  }
  if (pend_f >= 0) {
#pragma unroll
    for (int e = 0; e < E; ++e) q[0][e] = mk<T>(pend[e], (T)0);
    lds_sync<true>();
    fft_batch<T, N, -1, Pad, P, true, 1, RS>(q, lds, t, idx, tw);
    Z::unpack_store(q[0], a.nl[pend_f] + off, (cx<T>*)nullptr, active, LIM, a.kzs, t, h);
  }
}

template <typename T, int NZ, int V>
__global__ void __launch_bounds__(lb_threads((NZ / 2) / FusedPlan<T, NZ / 2, V>::type::E),
                                  (FusedPlan<T, NZ / 2, V>::min_waves))
k_zfused(ZFusedArgs<T> a, const cx<T>* __restrict__ twM, const cx<T>* __restrict__ twN, T scale) {
  using FP = FusedPlan<T, NZ / 2, V>;
  using Z = ZRow<T, NZ, typename FP::type, typename FP::pad>;
  extern __shared__ unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int t = threadIdx.x % Z::TPC;
  const int g = Z::TPC == 64 ? __builtin_amdgcn_readfirstlane(threadIdx.x / 64) : threadIdx.x / Z::TPC;
  if constexpr (V == 9) {
    TwRegs<T, NZ, Plan<NZ>> tw;
    tw.load(twN, t);
    zfused_pair_row<T, NZ, typename FP::pad>(a, tw, scale, lds, t, g);
  } else if constexpr (FP::regtw) {
    typename Z::Tw tw;
    tw.load(twM, twN, t);
    zfused_row<T, NZ, V>(a, tw, tw, scale, lds, t, g);
  } else {
    zfused_row<T, NZ, V>(a, twM, twN, scale, lds, t, g);
  }
}

}  // namespace gnss
