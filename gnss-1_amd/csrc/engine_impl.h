// engine_impl.h — Engine<T>: allocation, transforms, the RK3 step and statistics on one stream.
//
// Call structure kept from the reference driver (main.cu:84-252): create (allocate_memory +
// plans) -> init_wavenumbers -> init_jet (IC + forward + dealias) -> step / stats -> destroy.
#pragma once
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <sstream>

#include "engine.h"
#include "host_util.h"
#include "kernels_fft.h"
#include "kernels_mfma.h"
#include "kernels_step.h"

namespace gnss {

static const double RK_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double RK_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
static const double RK_C[4] = {0.0, 1.0 / 3.0, 3.0 / 4.0, 1.0};

inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }
inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

template <typename T>
class Engine : public EngineBase {
  using C = cx<T>;
  static constexpr int LDS_MAX = 64 * 1024;

 public:
  // nb > 1: a batched context of the members mp[0..nb) of one grid (see make_engine_batch_*)
  explicit Engine(const gnss_params& p, int nb = 1, const gnss_params* mp = nullptr) : p_(p) {
    if (nb < 1 || (nb > 1 && !mp)) throw Error(GNSS_ERR_ARG, "a batched context needs its members");
    nb_ = nb;
    mp_.assign(1, p);
    if (nb > 1) {
      mp_.assign(mp, mp + nb);
      const unsigned bad = GNSS_FLAG_COLLOID | GNSS_FLAG_MFMA_R16 | GNSS_FLAG_UNFUSED;
      for (const gnss_params& q : mp_)
        if (q.nx != p.nx || q.ny != p.ny || q.nz != p.nz || q.n_scalars != p.n_scalars ||
            q.precision != p.precision || q.flags != p.flags || (q.flags & bad) || q.lx != p.lx ||
            q.ly != p.ly || q.lz != p.lz || q.dt != p.dt || !(q.re > 0))
          throw Error(GNSS_ERR_ARG, "batched members must share grid, scalars, precision, flags, box and dt "
                                    "(no colloid, MFMA or unfused step)");
    }
    nx_ = p.nx; ny_ = p.ny; nz_ = p.nz;
    nzh_ = nz_ / 2 + 1;
    S_ = p.n_scalars;
    colloid_ = (p.flags & GNSS_FLAG_COLLOID) ? 1 : 0;
    if (colloid_ && S_ < 2)
      throw Error(GNSS_ERR_ARG, "GNSS_FLAG_COLLOID needs n_scalars >= 2 (scalar 0 carries the colloid)");
    if (colloid_ && !std::isfinite(p.alpha)) throw Error(GNSS_ERR_ARG, "alpha must be finite");
    F_ = 3 + S_;
    W_ = 6 + 3 * S_ + 2 * colloid_;
    cx_ = col_chunk(nx_);
    cy_ = col_chunk(ny_);
    xc_ = cx_;
    zf_ = 0;
    load_tuning();  // may set kp, which fixes the row stride below
    // kz padding and the pruned kz limit use the widest column tile any variant may choose; kp
    // extra tiles of padding move the x pass's row stride (ny kzs elements) off large powers of two
    const int cmax = sizeof(T) == 8 ? 8 : 16;
    kzs_ = round_up(nzh_, cmax) + kp_ * cmax;
    g_.nx = nx_; g_.ny = ny_; g_.nz = nz_; g_.nzh = nzh_; g_.kzs = kzs_;
    g_.keep_x = nx_ / 3; g_.keep_y = ny_ / 3; g_.keep_z = nz_ / 3;
    int64_t ax, ay, az, g2;
    mask_factors(nx_, ny_, nz_, ax, ay, az, g2);
    g_.ax = ax; g_.ay = ay; g_.az = az; g_.g2 = g2;
    kz_lim_ = std::min(round_up(g_.keep_z + 1, cmax), kzs_);
    nsp_ = (long long)nx_ * ny_ * kzs_;
    nph_ = (long long)nx_ * ny_ * nz_;
    GNSS_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    {
      int dev = 0;
      GNSS_HIP(hipGetDevice(&dev));
      GNSS_HIP(hipDeviceGetAttribute(&n_cu_, hipDeviceAttributeMultiprocessorCount, dev));
    }
    try {
      allocate();
      qb_ = q_;
      dqb_ = dq_;
      if (nb_ > 1) {
        // viscosity and diffusivities of every member for the batched RK3 update
        std::vector<double> nu((size_t)nb_ * MAXF, 0.0);
        for (int m = 0; m < nb_; ++m) {
          for (int f = 0; f < 3; ++f) nu[(size_t)m * MAXF + f] = 1.0 / mp_[m].re;
          for (int q = 0; q < S_; ++q) nu[(size_t)m * MAXF + 3 + q] = 1.0 / (mp_[m].re * mp_[m].sc[q]);
        }
        dalloc(d_num_, sizeof(double) * nu.size());
        GNSS_HIP(hipMemcpy(d_num_, nu.data(), sizeof(double) * nu.size(), hipMemcpyHostToDevice));
      }
    } catch (...) {
      release();
      throw;
    }
  }

  int members() const override { return nb_; }
  // Per-member calls address member m: its state and RK register, and its parameters.
  void select_member(int m) override {
    if (m < 0 || m >= nb_) throw Error(GNSS_ERR_ARG, "member out of range");
    if (nb_ == 1) return;
    for (int f = 0; f < F_; ++f) {
      q_[f] = qb_[f] + (long long)m * nsp_;
      dq_[f] = dqb_[f] + (long long)m * nsp_;
    }
    p_ = mp_[m];
  }

  ~Engine() override { release(); }

  // ---------------------------------------------------------------------------------------
  void sizes(gnss_sizes* o) const override {
    o->n_fields = F_;
    o->nzh = nzh_;
    o->n_prof = 7 + 3 * S_;
    o->n_bins = n_bins();
    o->n_physical = nph_;
    o->n_spectral = (long long)nx_ * ny_ * nzh_;
  }
  void time(double* t, int64_t* st) const override {
    if (t) *t = t_;
    if (st) *st = step_;
  }
  void set_clock(int64_t st, double t) override {
    step_ = st;
    t_ = t;
  }

  void init_wavenumbers() override {
    const double pi = 3.14159265358979323846;
    launch_wn(dkx_, nx_, nx_, 2.0 * pi / p_.lx);
    launch_wn(dky_, ny_, ny_, 2.0 * pi / p_.ly);
    launch_wn(dkz_, nz_, nzh_, 2.0 * pi / p_.lz);
    cast_wn(dkx_, tkx_, nx_);
    cast_wn(dky_, tky_, ny_);
    cast_wn(dkz_, tkz_, nzh_);
    sync();
    have_k_ = true;
  }

  void get_wavenumbers(double* kx, double* ky, double* kz) override {
    need_k();
    if (kx) GNSS_HIP(hipMemcpy(kx, dkx_, sizeof(double) * nx_, hipMemcpyDeviceToHost));
    if (ky) GNSS_HIP(hipMemcpy(ky, dky_, sizeof(double) * ny_, hipMemcpyDeviceToHost));
    if (kz) GNSS_HIP(hipMemcpy(kz, dkz_, sizeof(double) * nzh_, hipMemcpyDeviceToHost));
  }

  void get_mask(uint8_t* mask) override {
    need_k();
    const long long n = (long long)nx_ * ny_ * nzh_;
    unsigned char* d = (unsigned char*)scratch_bytes(n);
    hipLaunchKernelGGL(k_mask_abi, dim3(cdiv(n, 256)), dim3(256), 0, s_, g_, d, n);
    check_launch("k_mask_abi");
    GNSS_HIP(hipMemcpyAsync(mask, d, n, hipMemcpyDeviceToHost, s_));
    sync();
  }

  void init_jet() override {
    for (int m = 0; m < nb_; ++m) {
      if (nb_ > 1) select_member(m);
      init_jet_member();
    }
    if (nb_ > 1) select_member(0);
    t_ = 0;
    step_ = 0;
    have_ic_ = true;
  }
  void init_jet_member() {
    need_k();
    if (!(p_.theta > 0)) throw Error(GNSS_ERR_ARG, "theta must be > 0");
    std::vector<HostIcEntry> h;
    build_initial_condition(p_, h);
    std::vector<IcEntry> e(h.size());
    for (size_t i = 0; i < h.size(); ++i) {
      e[i].index = ((long long)h[i].ix * ny_ + h[i].iy) * kzs_ + h[i].iz;
      e[i].field = h[i].field;
      e[i].pad = 0;
      e[i].re = h[i].re;
      e[i].im = h[i].im;
    }
    for (int f = 0; f < F_; ++f) {
      GNSS_HIP(hipMemsetAsync(q_[f], 0, sizeof(C) * nsp_, s_));
      GNSS_HIP(hipMemsetAsync(dq_[f], 0, sizeof(C) * nsp_, s_));
    }
    if (!e.empty()) {
      IcEntry* d = (IcEntry*)scratch_bytes(sizeof(IcEntry) * e.size());
      GNSS_HIP(hipMemcpyAsync(d, e.data(), sizeof(IcEntry) * e.size(), hipMemcpyHostToDevice, s_));
      hipLaunchKernelGGL(k_ic_scatter<T>, dim3(cdiv((long long)e.size(), 256)), dim3(256), 0, s_, d,
                         (int)e.size(), ptrs(q_));
      check_launch("k_ic_scatter");
    }
    sync();
  }

  void forward(int f) override {
    check_field(f);
    fwd3d(phys_[f], q_[f], false);
    sync();
  }

  void inverse(int f) override {
    check_field(f);
    GNSS_HIP(hipMemcpyAsync(ws_[0], q_[f], sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
    inv3d(ws_[0], phys_[f], false);
    sync();
  }

  void dealias() override {
    need_k();
    hipLaunchKernelGGL(k_dealias<T>, dim3(cdiv(nsp_, 256)), dim3(256), 0, s_, g_, ptrs(q_), F_, nsp_);
    check_launch("k_dealias");
    sync();
  }

  void step(int nsteps) override {
    need_k();
    if (nsteps < 0) throw Error(GNSS_ERR_ARG, "nsteps must be >= 0");
    if (nsteps == 0) return;
    // the step starts from a dealiased state (deAlias at the top of solver_ps, solver.cu:704);
    // every substage keeps it dealiased afterwards
    for (int m = 0; m < nb_; ++m) {
      if (nb_ > 1) select_member(m);
      timed("k_dealias", (double)nsp_ * F_ * sizeof(C), [&] {
        hipLaunchKernelGGL(k_dealias<T>, dim3(cdiv(nsp_, 256)), dim3(256), 0, s_, g_, ptrs(q_), F_, nsp_);
      });
    }
    for (int n = 0; n < nsteps; ++n) enqueue_rk3_step();
    advance_clock(nsteps);
    check_finite();
  }

  hipStream_t stream() const override { return s_; }

  void enqueue_rk3_step() override {
    need_k();
    // a batched context steps all its members with the same launches, from member 0's base
    if (nb_ > 1) select_member(0);
    bm_ = nb_;
    try {
      for (int s = 0; s < 3; ++s) substage(s);
    } catch (...) {
      bm_ = 1;
      throw;
    }
    bm_ = 1;
  }

  void advance_clock(int nsteps) override {
    step_ += nsteps;
    t_ = (double)step_ * p_.dt;
  }

  // NaN/Inf watchdog on the kz = 0 column of every field.
  void check_finite() override {
    GNSS_HIP(hipMemsetAsync(d_flag_, 0, sizeof(int), s_));
    for (int m = 0; m < nb_; ++m) {
      if (nb_ > 1) select_member(m);
      hipLaunchKernelGGL(k_finite<T>, dim3(cdiv((long long)nx_ * ny_, 256)), dim3(256), 0, s_, g_,
                         ptrs(q_), F_, d_flag_);
      check_launch("k_finite");
    }
    if (nb_ > 1) select_member(0);
    int flag = 0;
    GNSS_HIP(hipMemcpyAsync(&flag, d_flag_, sizeof(int), hipMemcpyDeviceToHost, s_));
    sync();
    if (flag) throw Error(GNSS_ERR_NUMERIC, "non-finite value in the state after step " + std::to_string(step_));
  }

  void stats(gnss_stats_out* o, double* yprof, double* spectrum) override {
    need_k();
    const int nb = n_bins();
    stats_kernels(d_part_, d_spec_, d_prof_, d_cmax_);
    std::vector<double> part((size_t)nx_ * SPEC_NQ), spec((size_t)nx_ * (1 + S_) * nb),
        prof((size_t)(7 + 3 * S_) * ny_), cm((size_t)2 * ny_);
    GNSS_HIP(hipMemcpyAsync(part.data(), d_part_, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s_));
    GNSS_HIP(hipMemcpyAsync(spec.data(), d_spec_, sizeof(double) * spec.size(), hipMemcpyDeviceToHost, s_));
    GNSS_HIP(hipMemcpyAsync(prof.data(), d_prof_, sizeof(double) * prof.size(), hipMemcpyDeviceToHost, s_));
    GNSS_HIP(hipMemcpyAsync(cm.data(), d_cmax_, sizeof(double) * cm.size(), hipMemcpyDeviceToHost, s_));
    sync();
    double sum[SPEC_NQ + 2] = {0};
    for (int x = 0; x < nx_; ++x)
      for (int qq = 0; qq < SPEC_NQ; ++qq) sum[qq] += part[(size_t)x * SPEC_NQ + qq];
    for (int y = 0; y < ny_; ++y) {
      sum[SPEC_NQ] = std::max(sum[SPEC_NQ], cm[2 * y]);
      sum[SPEC_NQ + 1] = std::max(sum[SPEC_NQ + 1], cm[2 * y + 1]);
    }
    stats_finish(sum, o);
    if (yprof) memcpy(yprof, prof.data(), sizeof(double) * prof.size());
    if (spectrum) {
      for (int s = 0; s <= S_; ++s)
        for (int b = 0; b < nb; ++b) {
          double acc = 0;
          for (int x = 0; x < nx_; ++x) acc += spec[((size_t)x * (1 + S_) + s) * nb + b];
          spectrum[(size_t)s * nb + b] = acc;
        }
    }
  }

  // Scratch layout of stats_noalloc, regions 256-byte aligned: per-plane spectral sums, per-plane
  // shell sums, profiles, per-y maxima, then the reduced results (SPEC_NQ sums, 2 maxima, the
  // spectrum).
  struct StatScratch {
    size_t part, spec, prof, cmax, fin, total;
  };
  StatScratch stat_scratch_layout() const {
    auto up = [](size_t n) { return (n + 31) / 32 * 32; };
    const size_t nb = n_bins();
    StatScratch l;
    l.part = 0;
    l.spec = l.part + up((size_t)nx_ * SPEC_NQ);
    l.prof = l.spec + up((size_t)nx_ * (1 + S_) * nb);
    l.cmax = l.prof + up((size_t)(7 + 3 * S_) * ny_);
    l.fin = l.cmax + up((size_t)2 * ny_);
    l.total = l.fin + up(SPEC_NQ + 2 + (size_t)(1 + S_) * nb);
    return l;
  }
  size_t stats_scratch_doubles() const override { return stat_scratch_layout().total; }

  // Every partial the kernels read is written earlier in the same call, so the scratch needs no
  // clearing; the sums over planes run on the device in the order stats() sums them on the host.
  void stats_noalloc(double* scratch, gnss_stats_out* o, double* yprof, double* spectrum) override {
    need_k();
    const int nb = n_bins();
    const StatScratch l = stat_scratch_layout();
    double* fin = scratch + l.fin;
    pause_timer_ = true;
    try {
      stats_kernels(scratch + l.part, scratch + l.spec, scratch + l.prof, scratch + l.cmax);
    } catch (...) {
      pause_timer_ = false;
      throw;
    }
    pause_timer_ = false;
    hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(STAT_THREADS), 0, s_, scratch + l.part, scratch + l.spec,
                       scratch + l.cmax, nx_, ny_, (1 + S_) * nb, fin);
    check_launch("k_stats_reduce");
    double sum[SPEC_NQ + 2];
    GNSS_HIP(hipMemcpyAsync(sum, fin, sizeof(sum), hipMemcpyDeviceToHost, s_));
    if (yprof)
      GNSS_HIP(hipMemcpyAsync(yprof, scratch + l.prof, sizeof(double) * (7 + 3 * S_) * ny_, hipMemcpyDeviceToHost, s_));
    if (spectrum)
      GNSS_HIP(hipMemcpyAsync(spectrum, fin + SPEC_NQ + 2, sizeof(double) * (1 + S_) * nb, hipMemcpyDeviceToHost, s_));
    sync();
    stats_finish(sum, o);
  }

  // Statistics kernels: spectral sums and shell sums per kx plane, the physical fields, then the
  // planar averages and maxima per y.
  void stats_kernels(double* part, double* spec, double* prof, double* cmax) {
    const double nn = (double)nx_ * ny_ * nz_;
    const double inn2 = 1.0 / (nn * nn);
    const int nb = n_bins();
    timed("k_spec_sums", (double)F_ * nx_ * ny_ * nzh_ * sizeof(C), [&] {
      hipLaunchKernelGGL(k_spec_sums<T>, dim3(nx_), dim3(STAT_THREADS), 0, s_, g_, ptrs(q_), dkx_,
                         dky_, dkz_, S_, inn2, part);
    });
    ShellArgs sa;
    sa.equal_box = (p_.lx == p_.ly && p_.ly == p_.lz) ? 1 : 0;
    const double lm = std::max(p_.lx, std::max(p_.ly, p_.lz));
    sa.fx = lm / p_.lx; sa.fy = lm / p_.ly; sa.fz = lm / p_.lz;
    timed("k_spectrum", (double)F_ * nx_ * ny_ * nzh_ * sizeof(C), [&] {
      hipLaunchKernelGGL(k_spectrum<T>, dim3(nx_), dim3(STAT_THREADS), 0, s_, g_, ptrs(q_), S_, nb, sa,
                         inn2, spec);
    });
    for (int f = 0; f < F_; ++f) {
      GNSS_HIP(hipMemcpyAsync(ws_[0], q_[f], sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
      inv3d(ws_[0], phys_[f], false);
    }
    RPtrs<T> pp{};
    for (int f = 0; f < F_; ++f) pp.p[f] = phys_[f];
    timed("k_plane_stats", (double)F_ * nph_ * sizeof(T) * 2, [&] {
      hipLaunchKernelGGL(k_plane_stats<T>, dim3(ny_), dim3(STAT_THREADS), 0, s_, g_, pp, S_,
                         nx_ / p_.lx, ny_ / p_.ly, nz_ / p_.lz, prof, cmax);
    });
  }

  // Integral statistics from the summed spectral partials sum[0..SPEC_NQ) and the maxima of the
  // CFL sum and of |u_i| (sum[SPEC_NQ], sum[SPEC_NQ + 1]).
  void stats_finish(const double* sum, gnss_stats_out* o) const {
    if (!o) return;
    const double nu = 1.0 / p_.re;
    memset(o, 0, sizeof(*o));
    o->t = t_;
    o->step = step_;
    o->ke = 0.5 * sum[0];
    o->vrms = sqrt(sum[0] / 3.0);
    o->epsilon = nu * sum[1];
    o->enstrophy = 0.5 * sum[1];
    o->lint = o->ke > 0 ? 3.0 * 3.14159265358979323846 / 4.0 * sum[2] / o->ke : 0.0;
    o->lambda = o->epsilon > 0 ? sqrt(15.0 * nu * o->vrms * o->vrms / o->epsilon) : 0.0;
    o->eta = o->epsilon > 0 ? sqrt(sqrt(nu * nu * nu / o->epsilon)) : 0.0;
    o->divergence = sum[1] > 0 ? sqrt(sum[3] / sum[1]) : 0.0;
    for (int s = 0; s < S_; ++s) {
      o->chi[s] = 2.0 * (nu / p_.sc[s]) * sum[4 + 2 * s];
      o->scalar_var[s] = sum[5 + 2 * s];
    }
    o->cfl = p_.dt * sum[SPEC_NQ];
    o->umax = sum[SPEC_NQ + 1];
  }

  void get_field(int f, int space, double* host) override {
    check_field(f);
    if (space == GNSS_PHYSICAL) {
      std::vector<T> h(nph_);
      GNSS_HIP(hipMemcpyAsync(h.data(), phys_[f], sizeof(T) * nph_, hipMemcpyDeviceToHost, s_));
      sync();
      for (long long i = 0; i < nph_; ++i) host[i] = (double)h[i];
      return;
    }
    if (space != GNSS_SPECTRAL) throw Error(GNSS_ERR_ARG, "bad space");
    std::vector<C> h(nsp_);
    GNSS_HIP(hipMemcpyAsync(h.data(), q_[f], sizeof(C) * nsp_, hipMemcpyDeviceToHost, s_));
    sync();
    for (int y = 0; y < ny_; ++y)
      for (int x = 0; x < nx_; ++x) {
        const C* src = h.data() + ((size_t)x * ny_ + y) * kzs_;
        double* dst = host + 2 * ((size_t)y * nx_ + x) * nzh_;
        for (int k = 0; k < nzh_; ++k) {
          dst[2 * k] = (double)src[k].x;
          dst[2 * k + 1] = (double)src[k].y;
        }
      }
  }

  void set_field(int f, int space, const double* host) override {
    check_field(f);
    if (space == GNSS_PHYSICAL) {
      std::vector<T> h(nph_);
      for (long long i = 0; i < nph_; ++i) h[i] = (T)host[i];
      GNSS_HIP(hipMemcpyAsync(phys_[f], h.data(), sizeof(T) * nph_, hipMemcpyHostToDevice, s_));
      sync();
      return;
    }
    if (space != GNSS_SPECTRAL) throw Error(GNSS_ERR_ARG, "bad space");
    std::vector<C> h(nsp_);
    memset((void*)h.data(), 0, sizeof(C) * nsp_);
    for (int y = 0; y < ny_; ++y)
      for (int x = 0; x < nx_; ++x) {
        C* dst = h.data() + ((size_t)x * ny_ + y) * kzs_;
        const double* src = host + 2 * ((size_t)y * nx_ + x) * nzh_;
        for (int k = 0; k < nzh_; ++k) {
          dst[k].x = (T)src[2 * k];
          dst[k].y = (T)src[2 * k + 1];
        }
      }
    GNSS_HIP(hipMemcpyAsync(q_[f], h.data(), sizeof(C) * nsp_, hipMemcpyHostToDevice, s_));
    GNSS_HIP(hipMemsetAsync(dq_[f], 0, sizeof(C) * nsp_, s_));
    sync();
  }

  void timing_enable(bool on) override {
    sync();
    timer_.enable(on, s_);
  }
  std::string timing_report() override {
    sync();
    return timer_.report();
  }

  void bench_copy(int reps, double* ms, double* bytes) override {
    if (reps <= 0) throw Error(GNSS_ERR_ARG, "reps must be > 0");
    const long long n16 = (long long)(sizeof(C) * nsp_ / 16);
    const long long chunk = 4096;  // 64 KiB per block
    const long long grid = cdiv(n16, chunk);
    auto launch = [&] {
      hipLaunchKernelGGL(k_copy16, dim3((unsigned)grid), dim3(256), 0, s_, reinterpret_cast<const v4f*>(ws_[0]),
                         reinterpret_cast<v4f*>(ws_[1]), n16, chunk);
    };
    hipEvent_t e0, e1;
    GNSS_HIP(hipEventCreate(&e0));
    GNSS_HIP(hipEventCreate(&e1));
    launch();  // warm-up
    GNSS_HIP(hipEventRecord(e0, s_));
    for (int r = 0; r < reps; ++r) launch();
    GNSS_HIP(hipEventRecord(e1, s_));
    GNSS_HIP(hipEventSynchronize(e1));
    GNSS_HIP(hipGetLastError());
    float t = 0;
    GNSS_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    *bytes = 2.0 * (double)n16 * 16.0;
  }

  void bench_fft(int reps, double* ms_fwd, double* ms_inv) override {
    if (reps <= 0) throw Error(GNSS_ERR_ARG, "reps must be > 0");
    hipEvent_t e0, e1, e2;
    GNSS_HIP(hipEventCreate(&e0));
    GNSS_HIP(hipEventCreate(&e1));
    GNSS_HIP(hipEventCreate(&e2));
    // warm-up pair
    inv3d_keep(q_[0], ws_[0], phys_[0]);
    fwd3d(phys_[0], ws_[1], false);
    float tf = 0, ti = 0;
    for (int r = 0; r < reps; ++r) {
      GNSS_HIP(hipEventRecord(e0, s_));
      GNSS_HIP(hipMemcpyAsync(ws_[0], q_[0], sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
      GNSS_HIP(hipEventRecord(e1, s_));
      inv3d(ws_[0], phys_[0], false);
      GNSS_HIP(hipEventRecord(e2, s_));
      GNSS_HIP(hipEventSynchronize(e2));
      float a = 0, b = 0;
      GNSS_HIP(hipEventElapsedTime(&a, e1, e2));
      ti += a;
      GNSS_HIP(hipEventRecord(e0, s_));
      fwd3d(phys_[0], ws_[1], false);
      GNSS_HIP(hipEventRecord(e1, s_));
      GNSS_HIP(hipEventSynchronize(e1));
      GNSS_HIP(hipEventElapsedTime(&b, e0, e1));
      tf += b;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (ms_fwd) *ms_fwd = tf / reps;
    if (ms_inv) *ms_inv = ti / reps;
  }

 public:
  // -------------------------------------------------------------------------------------------
  // Autotuner: times the compiled variants of the step's passes on this context's buffers and
  // keeps the fastest (column tile width per strided axis, plan of the fused z kernel).
  // ------------------------------------------------------------------------------------------
  std::string tuning_key() const {
    char b[96];
    snprintf(b, sizeof b, "%s %d %d %d %d%s", sizeof(T) == 8 ? "f64" : "f32", nx_, ny_, nz_, S_,
             colloid_ ? "c" : "");
    return b;
  }
  std::string tuning_string() const {
    char b[128];
    snprintf(b, sizeof b, "cx=%d cy=%d zf=%d xf=%d cp=%d xc=%d zp=%d kp=%d cb=%d lp=%d yc=%d", cx_, cy_, zf_, xf_, cp_,
             xc_, zp_, kp_, cb_, lp_, yc_);
    return b;
  }
  void set_tuning(const std::string& spec) override {
    int cx = cx_, cy = cy_, zf = zf_, xf = xf_, cp = cp_, xc = xc_, zp = zp_, kp = kp_, cb = cb_, lp = lp_, yc = yc_;
    const char* p = spec.c_str();
    const char* q;
    if ((q = strstr(p, "cp=")) || (q = strstr(p, "\"cp\": "))) cp = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "xf=")) || (q = strstr(p, "\"xf\": "))) xf = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "cx=")) || (q = strstr(p, "\"cx\": "))) cx = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "cy=")) || (q = strstr(p, "\"cy\": "))) cy = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "zf=")) || (q = strstr(p, "\"zf\": "))) zf = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "xc=")) || (q = strstr(p, "\"xc\": "))) xc = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "zp=")) || (q = strstr(p, "\"zp\": "))) zp = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "kp=")) || (q = strstr(p, "\"kp\": "))) kp = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "cb=")) || (q = strstr(p, "\"cb\": "))) cb = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "lp=")) || (q = strstr(p, "\"lp\": "))) lp = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if ((q = strstr(p, "yc=")) || (q = strstr(p, "\"yc\": "))) yc = atoi(strchr(q, q[2] == '=' ? '=' : ':') + 1);
    if (!chunk_ok(nx_, cx) || !chunk_ok(ny_, cy) || !zvariant_ok(zf) || xf < 0 || xf > max_xmode() ||
        cp < 0 || cp > 1 || !xchunk_ok(xc) || zp < 0 || zp > 1 || kp < 0 || kp > 4 || cb < 0 || cb > 1 || lp < 0 ||
        lp > 3 || (yc != 0 && !ychunk_ok(yc)))
      throw Error(GNSS_ERR_ARG, "invalid tuning '" + spec + "' for " + tuning_key());
    if (kp != kp_ && nsp_ > 0)
      throw Error(GNSS_ERR_ARG, "kp (row padding) is fixed when the context is created: set it in the tuning table");
    cx_ = cx;
    cy_ = cy;
    zf_ = zf;
    xf_ = xf;
    cp_ = cp;
    xc_ = xc;
    zp_ = zp;
    kp_ = kp;
    cb_ = cb;
    lp_ = lp;
    yc_ = yc;
  }

  // -------------------------------------------------------------------------------------------
  // Diagnostics. Work buffers ws_ hold the intermediate fields (a complex work buffer holds a
  // physical field: 2 kzs >= nz); the state, dq and the physical field buffers are untouched.
  // -------------------------------------------------------------------------------------------
  bool is_colloid(int field) const override { return colloid_ && field == F_ - 1; }

  // physical values of `field` (or |omega|) in one of ws_[1], ws_[3]
  T* diag_field(int field) {
    need_k();
    if (field < GNSS_FIELD_VORTICITY || field >= F_)
      throw Error(GNSS_ERR_ARG, "field " + std::to_string(field) + " out of range");
    T* out;
    if (field >= 0) {
      out = reinterpret_cast<T*>(ws_[1]);
      GNSS_HIP(hipMemcpyAsync(ws_[0], q_[field], sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
      inv3d(ws_[0], out, false);
    } else {
      CPtrs<T> w{};
      for (int c = 0; c < 3; ++c) w.p[c] = ws_[c];
      hipLaunchKernelGGL(k_vorticity<T>, dim3((unsigned)cdiv(nsp_, 256)), dim3(256), 0, s_, g_, ptrs(q_), w, tkx_,
                         tky_, tkz_, nsp_);
      for (int c = 0; c < 3; ++c) inv3d(ws_[c], reinterpret_cast<T*>(ws_[3 + c]), false);
      out = reinterpret_cast<T*>(ws_[3]);
      hipLaunchKernelGGL(k_magnitude<T>, dim3((unsigned)cdiv(nph_, 256)), dim3(256), 0, s_, out,
                         reinterpret_cast<const T*>(ws_[4]), reinterpret_cast<const T*>(ws_[5]), nph_);
    }
    GNSS_HIP(hipGetLastError());
    return out;
  }

  void get_slice(int field, int axis, int index, double* host) override {
    const int n[3] = {nx_, ny_, nz_};
    if (axis < 0 || axis > 2 || index < 0 || index >= n[axis])
      throw Error(GNSS_ERR_ARG, "plane " + std::to_string(index) + " of axis " + std::to_string(axis) + " out of range");
    const T* a = diag_field(field);
    const long long np = axis == 2 ? (long long)nx_ * ny_ : (axis == 1 ? (long long)nx_ * nz_ : (long long)ny_ * nz_);
    double* d = reinterpret_cast<double*>(ws_[2]);  // np doubles <= one complex work buffer
    hipLaunchKernelGGL(k_plane<T>, dim3((unsigned)cdiv(np, 256)), dim3(256), 0, s_, a, d, nx_, ny_, nz_, axis, index);
    GNSS_HIP(hipGetLastError());
    GNSS_HIP(hipMemcpyAsync(host, d, sizeof(double) * np, hipMemcpyDeviceToHost, s_));
    sync();
  }

  double surface_area(int field, double iso) override {
    const T* a = diag_field(field);
    const double dx = p_.lx / nx_, dy = p_.ly / ny_, dz = p_.lz / nz_;
    timed("k_surface_area", (double)nph_ * sizeof(T), [&] {
      hipLaunchKernelGGL(k_surface_area<T>, dim3(nx_), dim3(STAT_THREADS), 0, s_, g_, a, iso, dx, dy, dz, d_part_);
    });
    std::vector<double> part(nx_);
    GNSS_HIP(hipMemcpyAsync(part.data(), d_part_, sizeof(double) * nx_, hipMemcpyDeviceToHost, s_));
    sync();
    double sum = 0.0;
    for (int x = 0; x < nx_; ++x) sum += part[x];
    return sum * dx * dy * dz;
  }

  std::string autotune(int reps) override {
    need_k();
    if (nb_ > 1) return autotune_batched(reps);
    if (reps <= 0) return "{\"key\": \"" + tuning_key() + "\", \"choice\": \"" + tuning_string() + "\"}";
    const bool was_on = timer_.on();
    timer_.enable(false, s_);
    hipEvent_t e0, e1;
    GNSS_HIP(hipEventCreate(&e0));
    GNSS_HIP(hipEventCreate(&e1));
    auto time_it = [&](auto&& fn) {
      fn();  // warm-up
      GNSS_HIP(hipEventRecord(e0, s_));
      for (int r = 0; r < reps; ++r) fn();
      GNSS_HIP(hipEventRecord(e1, s_));
      GNSS_HIP(hipEventSynchronize(e1));
      float ms = 0;
      GNSS_HIP(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms / reps;
    };
    std::string rep = "{\"key\": \"" + tuning_key() + "\", \"candidates\": {";
    char buf[160];
    bool first = true;
    // x and y passes as the step runs them (pruned inverse then forward), per tile width
    for (int axis = 0; axis < 2; ++axis) {
      const int L = axis == 0 ? nx_ : ny_;
      int best = axis == 0 ? cx_ : cy_;
      double best_ms = 1e30;
      for (int c : {4, 8, 16}) {
        if (!chunk_ok(L, c)) continue;
        (axis == 0 ? cx_ : cy_) = c;
        const double ms = time_it([&] {
          if (axis == 0) {
            col_pass(ws_[0], ColSpec{0, +1, kz_lim_ / cx_, g_.keep_y, g_.keep_x, -1});
            col_pass(ws_[0], ColSpec{0, -1, kz_lim_ / cx_, g_.keep_y, -1, g_.keep_x});
          } else {
            col_pass(ws_[0], ColSpec{1, +1, kz_lim_ / cy_, -1, g_.keep_y, -1});
            col_pass(ws_[0], ColSpec{1, -1, kz_lim_ / cy_, -1, -1, g_.keep_y});
          }
        });
        snprintf(buf, sizeof buf, "%s\"c%c=%d\": %.5f", first ? "" : ", ", axis == 0 ? 'x' : 'y', c, ms);
        rep += buf;
        first = false;
        if (ms < best_ms) { best_ms = ms; best = c; }
      }
      (axis == 0 ? cx_ : cy_) = best;
    }
    {
      // persistent pipelined column passes or one block per tile
      int best = cp_;
      double best_ms = 1e30;
      for (int v = 0; v <= 1; ++v) {
        cp_ = v;
        const double ms = time_it([&] {
          col_pass(ws_[0], ColSpec{1, +1, kz_lim_ / cy_, -1, g_.keep_y, -1});
          col_pass(ws_[0], ColSpec{1, -1, kz_lim_ / cy_, -1, -1, g_.keep_y});
          col_pass(ws_[0], ColSpec{0, +1, kz_lim_ / cx_, g_.keep_y, g_.keep_x, -1});
          col_pass(ws_[0], ColSpec{0, -1, kz_lim_ / cx_, g_.keep_y, -1, g_.keep_x});
        });
        snprintf(buf, sizeof buf, ", \"cp=%d\": %.5f", v, ms);
        rep += buf;
        if (ms < best_ms) { best_ms = ms; best = v; }
      }
      cp_ = best;
    }
    {
      // standalone z passes (transforms outside the step: statistics, fields, the 3D FFT):
      // one block per row group or persistent and pipelined
      int best = zp_;
      double best_ms = 1e30;
      const T scale = (T)(1.0 / ((double)nx_ * ny_ * nz_));
      for (int v = 0; v <= 1; ++v) {
        zp_ = v;
        const double ms = time_it([&] {
          z_c2r(ws_[0], reinterpret_cast<T*>(ws_[1]), nzh_, scale);
          z_r2c(reinterpret_cast<const T*>(ws_[1]), ws_[0], kzs_);
        });
        snprintf(buf, sizeof buf, ", \"zp=%d\": %.5f", v, ms);
        rep += buf;
        if (ms < best_ms) { best_ms = ms; best = v; }
      }
      zp_ = best;
    }
    {
      int best = zf_;
      double best_ms = 1e30;
      for (int v = 0; v <= 9; ++v) {
        if (!zvariant_ok(v)) continue;
        zf_ = v;
        const double ms = time_it([&] { z_fused(); });
        snprintf(buf, sizeof buf, ", \"zf=%d\": %.5f", v, ms);
        rep += buf;
        if (ms < best_ms) { best_ms = ms; best = v; }
      }
      zf_ = best;
    }
    // fused x passes (prologue / update inside the x transforms) or separate kernels: timed over
    // whole middle substages on a snapshot of the state, restored afterwards
    {
      std::vector<void*> snap;
      auto grab = [&](size_t bytes) {
        void* v = nullptr;
        GNSS_HIP(hipMalloc(&v, bytes));
        snap.push_back(v);
        return v;
      };
      try {
        for (int f = 0; f < F_; ++f) {
          GNSS_HIP(hipMemcpyAsync(grab(sizeof(C) * nsp_), q_[f], sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
          GNSS_HIP(hipMemcpyAsync(grab(sizeof(C) * nsp_), dq_[f], sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
        }
        // x-pass mode and tile, jointly with one launch per pass (cb = 1) or per field (cb = 0)
        int best = xf_, best_xc = xc_, best_cb = cb_, best_lp = lp_;
        yc_ = 0;  // the x-mode search times the late y passes with the x tile
        double best_ms = 1e30;
        for (int b = 0; b <= 1; ++b) {
          for (int v = 0; v <= max_xmode(); ++v) {
            for (int c : {4, 8, 16}) {
              if (v == 0 ? c != 4 : !xchunk_ok(c)) continue;  // xf = 0 is timed once per cb
              for (int l = (v >= 4 ? 0 : 1); l <= (v >= 4 ? 3 : 1); ++l) {  // lp: late modes only
                cb_ = b;
                xf_ = v;
                lp_ = l;
                if (v) xc_ = c;
                const double ms = time_it([&] { substage(1); });
                if (v >= 4) snprintf(buf, sizeof buf, ", \"cb=%d xf=%d xc=%d lp=%d\": %.5f", b, v, c, l, ms);
                else if (v) snprintf(buf, sizeof buf, ", \"cb=%d xf=%d xc=%d\": %.5f", b, v, c, ms);
                else snprintf(buf, sizeof buf, ", \"cb=%d xf=0\": %.5f", b, ms);
                rep += buf;
                if (ms < best_ms) { best_ms = ms; best = v; best_xc = xc_; best_cb = b; best_lp = l; }
              }
            }
          }
        }
        xf_ = best;
        xc_ = best_xc;
        cb_ = best_cb;
        lp_ = best_lp;
        // late modes: the y inverse passes' own tile
        yc_ = 0;
        if (best >= 4) {
          for (int c : {4, 8, 16}) {
            if (c == xc_ || !ychunk_ok(c)) continue;
            yc_ = c;
            const double ms = time_it([&] { substage(1); });
            snprintf(buf, sizeof buf, ", \"yc=%d\": %.5f", c, ms);
            rep += buf;
            if (ms < best_ms) {
              best_ms = ms;
            } else {
              yc_ = 0;
            }
          }
        }
        for (int f = 0; f < F_; ++f) {
          GNSS_HIP(hipMemcpyAsync(q_[f], snap[2 * f], sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
          GNSS_HIP(hipMemcpyAsync(dq_[f], snap[2 * f + 1], sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
        }
        sync();
      } catch (...) {
        (void)hipStreamSynchronize(s_);
        for (void* v : snap) (void)hipFree(v);
        throw;
      }
      for (void* v : snap) (void)hipFree(v);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    timer_.enable(was_on, s_);
    return rep + "}, \"choice\": \"" + tuning_string() + "\"}";
  }

  // A batched context's own entry ("<key>b"): its launches cover every member, so other pass
  // variants can win than for one member of the grid.  Batched steps always run the late x mode;
  // the fused z variant is timed alone, then the inverse passes' tile jointly with persistent or
  // per-tile column passes (cp), one launch per pass or per field (cb) and the late passes' output
  // form (lp), then the forward passes' tiles (cx, cy) and the late y passes' own tile (yc) — all
  // over whole middle substages of every member, on a snapshot of the state restored afterwards.
  std::string autotune_batched(int reps) {
    const std::string key = tuning_key() + "b";
    if (reps <= 0) return "{\"key\": \"" + key + "\", \"choice\": \"" + tuning_string() + "\"}";
    const bool was_on = timer_.on();
    timer_.enable(false, s_);
    std::vector<void*> snap;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string rep = "{\"key\": \"" + key + "\", \"members\": " + std::to_string(nb_) + ", \"candidates\": {";
    char buf[160];
    select_member(0);
    const size_t fbytes = sizeof(C) * (size_t)nsp_ * nb_;
    try {
      GNSS_HIP(hipEventCreate(&e0));
      GNSS_HIP(hipEventCreate(&e1));
      for (int f = 0; f < F_; ++f)
        for (C* src : {qb_[f], dqb_[f]}) {
          void* v = nullptr;
          GNSS_HIP(hipMalloc(&v, fbytes));
          snap.push_back(v);
          GNSS_HIP(hipMemcpyAsync(v, src, fbytes, hipMemcpyDeviceToDevice, s_));
        }
      bm_ = nb_;
      auto time_it = [&](auto&& fn) {
        fn();  // warm-up
        GNSS_HIP(hipEventRecord(e0, s_));
        for (int r = 0; r < reps; ++r) fn();
        GNSS_HIP(hipEventRecord(e1, s_));
        GNSS_HIP(hipEventSynchronize(e1));
        float ms = 0;
        GNSS_HIP(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / reps;
      };
      bool first = true;
      auto note = [&](const char* what, double ms) {  // `what` may point into buf
        char line[200];
        snprintf(line, sizeof line, "%s\"%s\": %.5f", first ? "" : ", ", what, ms);
        rep += line;
        first = false;
      };
      {
        int best = zf_;
        double best_ms = 1e30;
        for (int v = 0; v <= 9; ++v) {
          if (!zvariant_ok(v)) continue;
          zf_ = v;
          const double ms = time_it([&] { z_fused(); });
          snprintf(buf, sizeof buf, "zf=%d", v);
          note(buf, ms);
          if (ms < best_ms) { best_ms = ms; best = v; }
        }
        zf_ = best;
      }
      double best_ms = 1e30;
      {
        int bxc = xc_, bcp = cp_, bcb = cb_, blp = lp_;
        yc_ = 0;
        for (int c : {4, 8, 16}) {
          if (!xchunk_ok(c)) continue;
          for (int cp = 0; cp <= 1; ++cp)
            for (int cb = 0; cb <= 1; ++cb)
              for (int l = 0; l <= 3; ++l) {
                xc_ = c; cp_ = cp; cb_ = cb; lp_ = l;
                const double ms = time_it([&] { substage(1); });
                snprintf(buf, sizeof buf, "xc=%d cp=%d cb=%d lp=%d", c, cp, cb, l);
                note(buf, ms);
                if (ms < best_ms) { best_ms = ms; bxc = c; bcp = cp; bcb = cb; blp = l; }
              }
        }
        xc_ = bxc; cp_ = bcp; cb_ = bcb; lp_ = blp;
      }
      for (int axis = 0; axis < 2; ++axis) {
        int& knob = axis == 0 ? cx_ : cy_;
        const int keep = knob;
        int best = knob;
        for (int c : {4, 8, 16}) {
          if (c == keep || !chunk_ok(axis == 0 ? nx_ : ny_, c)) continue;
          knob = c;
          const double ms = time_it([&] { substage(1); });
          snprintf(buf, sizeof buf, "c%c=%d", axis == 0 ? 'x' : 'y', c);
          note(buf, ms);
          if (ms < best_ms) { best_ms = ms; best = c; }
        }
        knob = best;
      }
      for (int c : {4, 8, 16}) {
        if (c == xc_ || !ychunk_ok(c)) continue;
        const int keep = yc_;
        yc_ = c;
        const double ms = time_it([&] { substage(1); });
        snprintf(buf, sizeof buf, "yc=%d", c);
        note(buf, ms);
        if (ms < best_ms) best_ms = ms;
        else yc_ = keep;
      }
      bm_ = 1;
      for (int f = 0; f < F_; ++f) {
        GNSS_HIP(hipMemcpyAsync(qb_[f], snap[2 * f], fbytes, hipMemcpyDeviceToDevice, s_));
        GNSS_HIP(hipMemcpyAsync(dqb_[f], snap[2 * f + 1], fbytes, hipMemcpyDeviceToDevice, s_));
      }
      sync();
    } catch (...) {
      bm_ = 1;
      (void)hipStreamSynchronize(s_);
      for (void* v : snap) (void)hipFree(v);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      throw;
    }
    for (void* v : snap) (void)hipFree(v);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    timer_.enable(was_on, s_);
    return rep + "}, \"choice\": \"" + tuning_string() + "\"}";
  }

 private:
  // fused z variants built: FP32 0-7 (8 is 7 in FP32), FP64 0, 3, 6, 7 and 8
  // fused z variants built for this context (9, two rows per transform: no colloid, nz <= 512)
  bool zvariant_ok(int v) const {
    if (v == 9) return !colloid_ && nz_ <= 512;
    return sizeof(T) == 4 ? (v >= 0 && v <= 7) : (v == 0 || v == 3 || (v >= 6 && v <= 8));
  }
  // column tiles the fused x kernels are built for: FP64 4 or 8, FP32 8 or 16
  bool xchunk_ok(int c) const {
    return (sizeof(T) == 8 ? (c == 4 || c == 8) : (c == 8 || c == 16)) && chunk_ok(nx_, c);
  }
  // column tile of the late y inverse passes (0: the x tile xc)
  bool ychunk_ok(int c) const {
    return (sizeof(T) == 8 ? (c == 4 || c == 8) : (c == 8 || c == 16)) && chunk_ok(ny_, c);
  }
  bool chunk_ok(int L, int c) const {
    if (sizeof(T) == 8 ? (c != 4 && c != 8) : (c != 8 && c != 16)) return false;
    return (long long)L * c * (long long)sizeof(C) <= LDS_MAX;
  }

  // Table of tuned choices: one line per key, "<key>": {"cx": C, "cy": C, "zf": V}.  Looked up in
  // $GNSS_TUNING or tuning/gnss_tuning.json next to the library; absent entries keep defaults.
  void load_tuning() {
    std::string path;
    if (const char* e = getenv("GNSS_TUNING")) path = e;
    else path = library_dir() + "/../tuning/gnss_tuning.json";
    // a batched context first looks for its own entry ("<key>b": its launches cover every member,
    // so other tiles can win), then for the grid's
    for (int pass = (nb_ > 1 ? 0 : 1); pass < 2; ++pass) {
      FILE* f = fopen(path.c_str(), "r");
      if (!f) return;
      const std::string key = "\"" + tuning_key() + (pass == 0 ? "b" : "") + "\"";
      char line[512];
      bool found = false;
      while (fgets(line, sizeof line, f)) {
        if (strstr(line, key.c_str())) {
          found = true;
          try {
            set_tuning(strchr(line, '{') ? strchr(line, '{') : line);
          } catch (const Error&) {
          }
          break;
        }
      }
      fclose(f);
      if (found) return;
    }
  }

  // -------------------------------------------------------------------------------------------
  int col_chunk(int L) const {
    int c = sizeof(T) == 8 ? 8 : 16;
    while (c > 4 && (long long)L * c * (long long)sizeof(C) > LDS_MAX) c /= 2;
    return c;
  }
  int n_bins() const {
    if (p_.lx == p_.ly && p_.ly == p_.lz) {
      long long m2 = (long long)(nx_ / 2) * (nx_ / 2) + (long long)(ny_ / 2) * (ny_ / 2) + (long long)(nz_ / 2) * (nz_ / 2);
      long long r = (long long)sqrt((double)(4 * m2));
      while (r * r > 4 * m2) --r;
      while ((r + 1) * (r + 1) <= 4 * m2) ++r;
      return (int)((r + 1) / 2) + 1;
    }
    const double lm = std::max(p_.lx, std::max(p_.ly, p_.lz));
    const double a = (nx_ / 2) * lm / p_.lx, b = (ny_ / 2) * lm / p_.ly, c = (nz_ / 2) * lm / p_.lz;
    return (int)floor(sqrt(a * a + b * b + c * c) + 0.5) + 1;
  }

  void need_k() const {
    if (!have_k_) throw Error(GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  }
  void check_field(int f) const {
    if (f < 0 || f >= F_) throw Error(GNSS_ERR_ARG, "field " + std::to_string(f) + " out of range");
  }
  void sync() { GNSS_HIP(hipStreamSynchronize(s_)); }
  void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(GNSS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  }
  template <typename Fn> void timed(const char* name, double bytes, Fn&& fn) {
    const bool on = timer_.on() && !pause_timer_;
    if (on) timer_.begin(name, bytes);
    fn();
    check_launch(name);
    if (on) timer_.end();
  }

  template <typename P> void* dalloc(P*& ptr, size_t bytes) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, bytes ? bytes : 1);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(GNSS_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    allocs_.push_back(v);
    ptr = (P*)v;
    return v;
  }

  void* scratch_bytes(size_t n) {
    if (n > scratch_n_) {
      if (scratch_) {
        GNSS_HIP(hipStreamSynchronize(s_));
        GNSS_HIP(hipFree(scratch_));
        scratch_ = nullptr;
      }
      void* v = nullptr;
      GNSS_HIP(hipMalloc(&v, n));
      scratch_ = v;
      scratch_n_ = n;
    }
    return scratch_;
  }

  void allocate() {
    q_.assign(F_, nullptr); dq_.assign(F_, nullptr); nl_.assign(F_, nullptr);
    phys_.assign(F_, nullptr); ws_.assign(W_, nullptr);
    // step arrays hold every member of a batched context back to back; physical buffers serve one
    // member at a time (statistics, transforms, I/O)
    const long long nst = nsp_ * nb_;
    for (int f = 0; f < F_; ++f) {
      dalloc(q_[f], sizeof(C) * nst);
      dalloc(dq_[f], sizeof(C) * nst);
      dalloc(nl_[f], sizeof(C) * nst);
      dalloc(phys_[f], sizeof(T) * nph_);
    }
    for (int w = 0; w < W_; ++w) dalloc(ws_[w], sizeof(C) * nst);
    for (void* v : allocs_) GNSS_HIP(hipMemsetAsync(v, 0, 1, s_));
    for (int f = 0; f < F_; ++f) {
      GNSS_HIP(hipMemsetAsync(q_[f], 0, sizeof(C) * nst, s_));
      GNSS_HIP(hipMemsetAsync(dq_[f], 0, sizeof(C) * nst, s_));
      GNSS_HIP(hipMemsetAsync(nl_[f], 0, sizeof(C) * nst, s_));
      GNSS_HIP(hipMemsetAsync(phys_[f], 0, sizeof(T) * nph_, s_));
    }
    for (int w = 0; w < W_; ++w) GNSS_HIP(hipMemsetAsync(ws_[w], 0, sizeof(C) * nst, s_));
    dalloc(dkx_, sizeof(double) * nx_);
    dalloc(dky_, sizeof(double) * ny_);
    dalloc(dkz_, sizeof(double) * nzh_);
    dalloc(tkx_, sizeof(T) * nx_);
    dalloc(tky_, sizeof(T) * ny_);
    // the kz table padded with zeros to the row stride: the column tiles of the strided passes reach
    // kz_lim, which on small grids lies a tile beyond nz/2 + 1 (those columns are masked)
    dalloc(tkz_, sizeof(T) * kzs_);
    GNSS_HIP(hipMemsetAsync(tkz_, 0, sizeof(T) * kzs_, s_));
    const int nb = n_bins();
    dalloc(d_part_, sizeof(double) * nx_ * SPEC_NQ);
    dalloc(d_spec_, sizeof(double) * (size_t)nx_ * (1 + S_) * nb);
    dalloc(d_prof_, sizeof(double) * (size_t)(7 + 3 * S_) * ny_);
    dalloc(d_cmax_, sizeof(double) * 2 * ny_);
    dalloc(d_flag_, sizeof(int));
    // twiddle tables for every length the passes use
    for (int L : {nx_, ny_, nz_, nz_ / 2}) {
      if (tw_.count(L)) continue;
      std::vector<double> t;
      twiddle_table(L, t);
      std::vector<C> h(L);
      for (int m = 0; m < L; ++m) {
        h[m].x = (T)t[2 * m];
        h[m].y = (T)t[2 * m + 1];
      }
      C* d = nullptr;
      dalloc(d, sizeof(C) * L);
      GNSS_HIP(hipMemcpyAsync(d, h.data(), sizeof(C) * L, hipMemcpyHostToDevice, s_));
      tw_[L] = d;
    }
    sync();
  }

  // Physical work buffers of the unfused step, allocated on its first substage.  They are
  // published only once every allocation succeeded (dalloc records each one for release()).
  void ensure_unfused_buffers() {
    if (!wp_.empty()) return;
    std::vector<T*> wp(W_, nullptr), nlp(F_, nullptr);
    for (int w = 0; w < W_; ++w) dalloc(wp[w], sizeof(T) * nph_);
    for (int f = 0; f < F_; ++f) dalloc(nlp[f], sizeof(T) * nph_);
    nlp_.swap(nlp);
    wp_.swap(wp);
  }

  void release() {
    if (s_) (void)hipStreamSynchronize(s_);
    for (void* v : allocs_) (void)hipFree(v);
    allocs_.clear();
    if (scratch_) (void)hipFree(scratch_);
    scratch_ = nullptr;
    if (s_) (void)hipStreamDestroy(s_);
    s_ = nullptr;
  }

  CPtrs<T> ptrs(const std::vector<C*>& v) const {
    CPtrs<T> r{};
    for (size_t i = 0; i < v.size() && i < (size_t)MAXW; ++i) r.p[i] = v[i];
    return r;
  }
  RPtrs<T> rptrs(const std::vector<T*>& v) const {
    RPtrs<T> r{};
    for (size_t i = 0; i < v.size() && i < (size_t)MAXW; ++i) r.p[i] = v[i];
    return r;
  }

  void launch_wn(double* d, int n, int len, double scale) {
    hipLaunchKernelGGL(k_wavenumbers, dim3(cdiv(len, 256)), dim3(256), 0, s_, d, n, len, scale);
    check_launch("k_wavenumbers");
  }
  void cast_wn(const double* in, T* out, int n) {
    hipLaunchKernelGGL(k_cast_wavenumbers<T>, dim3(cdiv(n, 256)), dim3(256), 0, s_, in, out, n);
    check_launch("k_cast_wavenumbers");
  }

  // -------------------------------------------------------------------------------------------
  // Transform passes
  // -------------------------------------------------------------------------------------------
  struct ColSpec {
    int axis;        // 0 = x, 1 = y
    int sign;        // -1 forward, +1 inverse
    int nchunks;     // column chunks processed
    int outer_keep;  // kept-mode outer map (x pass: over ky), -1 = all
    int in_keep, out_keep;
  };

  // the pass over nf fields: persistent blocks over all their tiles (cp), else one launch per field
  template <int L, int SIGN, int CC>
  void col_launch(C* const* fields, int nf, const ColArgs& a, int G) {
    constexpr int TPC = L / Plan<L>::E;
    const int threads = CC * TPC * G;
    const size_t lds = (size_t)G * L * CC * sizeof(C);
    const long long blocks = (long long)a.nchunks * cdiv(a.nouter, G);
    if (cp_) {
      // the persistent form also holds the twiddle table in LDS, after the data regions
      const size_t ldsp = lds + (size_t)L * sizeof(C);
      // persistent: as many blocks as stay resident (LDS and wave limits of a CU), all CUs
      const long long per_cu = std::max<long long>(1, std::min<long long>(160 * 1024 / (long long)ldsp,
                                                                           2048 / threads));
      const long long grid = std::min<long long>(blocks * nf, per_cu * n_cu_);
      FieldList<T> fl{};
      for (int f = 0; f < nf; ++f) fl.p[f] = fields[f];
      hipLaunchKernelGGL((k_col_multi<T, L, SIGN, CC>), dim3((unsigned)grid), dim3(threads), ldsp, s_, fl,
                         tw_.at(L), a, (int)blocks, (int)(blocks * nf));
      return;
    }
    for (int f = 0; f < nf; ++f)
      hipLaunchKernelGGL((k_col<T, L, SIGN, CC>), dim3((unsigned)blocks), dim3(threads), lds, s_, fields[f],
                         tw_.at(L), a);
  }

  template <int L, int SIGN>
  void col_dispatch_c(C* const* fields, int nf, const ColArgs& a, int CC, int G) {
    if constexpr (sizeof(T) == 8) {
      if (CC == 8) col_launch<L, SIGN, 8>(fields, nf, a, G);
      else col_launch<L, SIGN, 4>(fields, nf, a, G);
    } else {
      if (CC == 16) col_launch<L, SIGN, 16>(fields, nf, a, G);
      else col_launch<L, SIGN, 8>(fields, nf, a, G);
    }
  }

  // First radix-16 stage on MFMA (split FP16, FP32 accumulate); FP32 lengths 256, 512, 1024.
  bool use_mfma(int L) const {
    return sizeof(T) == 4 && (p_.flags & GNSS_FLAG_MFMA_R16) && (L == 256 || L == 512 || L == 1024);
  }
  template <int L, int SIGN, int CC>
  void mfma_launch(C* data, const ColArgs& a) {
    if constexpr (sizeof(T) == 4) {
      const size_t lds = (size_t)L * CC * sizeof(C);
      const long long blocks = (long long)a.nchunks * a.nouter;
      hipLaunchKernelGGL((k_col_mfma<L, SIGN, CC>), dim3((unsigned)blocks), dim3(CC * (L / 16)), lds, s_,
                         data, tw_.at(L), a);
    }
  }
  template <int SIGN>
  void mfma_dispatch(int L, C* data, const ColArgs& a, int CC) {
    if (L == 1024) return mfma_launch<1024, SIGN, 8>(data, a);
    if (L == 512) return CC == 16 ? mfma_launch<512, SIGN, 16>(data, a) : mfma_launch<512, SIGN, 8>(data, a);
    if (L == 256) return CC == 16 ? mfma_launch<256, SIGN, 16>(data, a) : mfma_launch<256, SIGN, 8>(data, a);
    throw Error(GNSS_ERR_UNSUPPORTED, "MFMA radix-16 stage needs length 256, 512 or 1024");
  }

  template <int SIGN>
  void col_dispatch_l(int L, C* const* data, int nf, const ColArgs& a, int CC, int G) {
    switch (L) {
      case 16: col_dispatch_c<16, SIGN>(data, nf, a, CC, G); break;
      case 24: col_dispatch_c<24, SIGN>(data, nf, a, CC, G); break;
      case 32: col_dispatch_c<32, SIGN>(data, nf, a, CC, G); break;
      case 48: col_dispatch_c<48, SIGN>(data, nf, a, CC, G); break;
      case 64: col_dispatch_c<64, SIGN>(data, nf, a, CC, G); break;
      case 96: col_dispatch_c<96, SIGN>(data, nf, a, CC, G); break;
      case 128: col_dispatch_c<128, SIGN>(data, nf, a, CC, G); break;
      case 192: col_dispatch_c<192, SIGN>(data, nf, a, CC, G); break;
      case 256: col_dispatch_c<256, SIGN>(data, nf, a, CC, G); break;
      case 384: col_dispatch_c<384, SIGN>(data, nf, a, CC, G); break;
      case 512: col_dispatch_c<512, SIGN>(data, nf, a, CC, G); break;
      case 1024: col_dispatch_c<1024, SIGN>(data, nf, a, CC, G); break;
      default: throw Error(GNSS_ERR_UNSUPPORTED, "unsupported transform length " + std::to_string(L));
    }
  }

  static int elems_per_thread(int L) {
    switch (L) {
      case 8: return Plan<8>::E;
      case 12: return Plan<12>::E;
      case 16: return Plan<16>::E;
      case 24: return Plan<24>::E;
      case 32: return Plan<32>::E;
      case 48: return Plan<48>::E;
      case 64: return Plan<64>::E;
      case 96: return Plan<96>::E;
      case 128: return Plan<128>::E;
      case 192: return Plan<192>::E;
      case 256: return Plan<256>::E;
      case 384: return Plan<384>::E;
      case 512: return Plan<512>::E;
      case 1024: return Plan<1024>::E;
      default: throw Error(GNSS_ERR_UNSUPPORTED, "unsupported transform length " + std::to_string(L));
    }
  }

  void col_pass(C* data, const ColSpec& sp) { col_passes(&data, 1, sp); }

  // The same strided pass over nf fields; with cb (and cp) one persistent launch covers them all.
  void col_passes(C* const* fields, int nf, const ColSpec& sp) {
    const int L = sp.axis == 0 ? nx_ : ny_;
    const int CC = sp.axis == 0 ? cx_ : cy_;
    const int TPC = L / elems_per_thread(L);
    ColArgs a;
    if (sp.axis == 1) {
      a.estride = kzs_;
      a.ostride = (long long)ny_ * kzs_;
      a.nouter = nx_;
      a.outer_keep = -1;
      a.outer_n = nx_;
    } else {
      a.estride = (long long)ny_ * kzs_;
      a.ostride = kzs_;
      a.outer_keep = sp.outer_keep;
      a.outer_n = ny_;
      a.nouter = sp.outer_keep < 0 ? ny_ : 2 * sp.outer_keep + 1;
    }
    a.nchunks = sp.nchunks;
    a.in_keep = sp.in_keep;
    a.out_keep = sp.out_keep;
    batch_outer(a);
    int G = std::max(1, 256 / (CC * TPC));
    while (G > 1 && (long long)G * L * CC * (long long)sizeof(C) > LDS_MAX) --G;
    a.G = G;
    // algorithmic bytes: rows loaded + rows stored
    auto rows = [&](int keep) { return keep < 0 ? L : std::min(L, 2 * keep + 1); };
    const double bytes = (double)a.nouter * a.nchunks * CC * (rows(sp.in_keep) + rows(sp.out_keep)) * sizeof(C);
    const char* name = sp.axis == 0 ? (sp.sign < 0 ? "fft_x_fwd" : "fft_x_inv") : (sp.sign < 0 ? "fft_y_fwd" : "fft_y_inv");
    if (use_mfma(L)) {
      a.G = 1;
      const char* mname = sp.axis == 0 ? (sp.sign < 0 ? "fft_x_fwd_mfma" : "fft_x_inv_mfma")
                                       : (sp.sign < 0 ? "fft_y_fwd_mfma" : "fft_y_inv_mfma");
      for (int f = 0; f < nf; ++f)
        timed(mname, bytes, [&] {
          if (sp.sign < 0) mfma_dispatch<-1>(L, fields[f], a, CC);
          else mfma_dispatch<+1>(L, fields[f], a, CC);
        });
      return;
    }
    const int groups = cb_ ? 1 : nf, per = cb_ ? nf : 1;
    for (int gi = 0; gi < groups; ++gi)
      timed(name, bytes * per, [&] {
        if (sp.sign < 0) col_dispatch_l<-1>(L, fields + gi * per, per, a, CC, G);
        else col_dispatch_l<+1>(L, fields + gi * per, per, a, CC, G);
      });
  }

  template <int NZ>
  void zr2c_launch(const T* in, C* out, int out_lim) {
    constexpr int M = NZ / 2;
    constexpr int TPC = M / Plan<M>::E;
    int G = std::max(1, 256 / TPC);
    while (G > 1 && (long long)G * zrow_stride(M) * (long long)sizeof(C) > LDS_MAX) --G;
    const long long rows = (long long)nx_ * ny_;
    const size_t lds = (size_t)G * zrow_stride(M) * sizeof(C);
    if (zp_) {
      const long long groups = cdiv(rows, G);
      const size_t ldsp = lds + 2 * (size_t)M * sizeof(C);  // and both twiddle tables
      hipLaunchKernelGGL((k_r2c_z_pipe<T, NZ>), dim3((unsigned)std::min(groups, zpipe_blocks(ldsp, TPC * G))),
                         dim3(TPC * G), ldsp, s_, in, out, tw_.at(M), tw_.at(NZ), rows, kzs_, out_lim, G, groups);
      return;
    }
    hipLaunchKernelGGL((k_r2c_z<T, NZ>), dim3((unsigned)cdiv(rows, G)), dim3(TPC * G), lds, s_, in, out,
                       tw_.at(M), tw_.at(NZ), rows, kzs_, out_lim, G);
  }
  template <int NZ>
  void zc2r_launch(const C* in, T* out, int in_lim, T scale) {
    constexpr int M = NZ / 2;
    constexpr int TPC = M / Plan<M>::E;
    int G = std::max(1, 256 / TPC);
    while (G > 1 && (long long)G * zrow_stride(M) * (long long)sizeof(C) > LDS_MAX) --G;
    const long long rows = (long long)nx_ * ny_;
    const size_t lds = (size_t)G * zrow_stride(M) * sizeof(C);
    if (zp_) {
      const long long groups = cdiv(rows, G);
      const size_t ldsp = lds + 2 * (size_t)M * sizeof(C);  // and both twiddle tables
      hipLaunchKernelGGL((k_c2r_z_pipe<T, NZ>), dim3((unsigned)std::min(groups, zpipe_blocks(ldsp, TPC * G))),
                         dim3(TPC * G), ldsp, s_, in, out, tw_.at(M), tw_.at(NZ), rows, kzs_, in_lim, scale, G,
                         groups);
      return;
    }
    hipLaunchKernelGGL((k_c2r_z<T, NZ>), dim3((unsigned)cdiv(rows, G)), dim3(TPC * G), lds, s_, in, out,
                       tw_.at(M), tw_.at(NZ), rows, kzs_, in_lim, scale, G);
  }

  // blocks of a persistent z pass: as many as stay resident on every CU (LDS and wave limits)
  long long zpipe_blocks(size_t lds, int threads) const {
    const long long per_cu = std::max<long long>(1, std::min<long long>(160 * 1024 / std::max<long long>(1, (long long)lds),
                                                                       2048 / threads));
    return per_cu * n_cu_;
  }

#define GNSS_Z_SWITCH(FN, ...)                                                                 \
  switch (nz_) {                                                                               \
    case 16: FN<16>(__VA_ARGS__); break;                                                       \
    case 24: FN<24>(__VA_ARGS__); break;                                                       \
    case 32: FN<32>(__VA_ARGS__); break;                                                       \
    case 48: FN<48>(__VA_ARGS__); break;                                                       \
    case 64: FN<64>(__VA_ARGS__); break;                                                       \
    case 96: FN<96>(__VA_ARGS__); break;                                                       \
    case 128: FN<128>(__VA_ARGS__); break;                                                     \
    case 192: FN<192>(__VA_ARGS__); break;                                                     \
    case 256: FN<256>(__VA_ARGS__); break;                                                     \
    case 384: FN<384>(__VA_ARGS__); break;                                                     \
    case 512: FN<512>(__VA_ARGS__); break;                                                     \
    case 1024: FN<1024>(__VA_ARGS__); break;                                                   \
    default: throw Error(GNSS_ERR_UNSUPPORTED, "unsupported nz " + std::to_string(nz_));      \
  }

  void z_r2c(const T* in, C* out, int out_lim) {
    const double bytes = (double)nx_ * ny_ * ((double)nz_ * sizeof(T) + (double)std::min(out_lim, kzs_) * sizeof(C));
    timed("fft_z_r2c", bytes, [&] { GNSS_Z_SWITCH(zr2c_launch, in, out, out_lim); });
  }
  void z_c2r(const C* in, T* out, int in_lim, T scale) {
    const double bytes = (double)nx_ * ny_ * ((double)nz_ * sizeof(T) + (double)std::min(in_lim, nzh_) * sizeof(C));
    timed("fft_z_c2r", bytes, [&] { GNSS_Z_SWITCH(zc2r_launch, in, out, in_lim, scale); });
  }

  // Forward 3D R2C: in (physical) -> out (spectral).  prune: only the kept box is produced.
  void fwd3d(const T* in, C* out, bool prune) {
    if (!prune) {
      z_r2c(in, out, kzs_);
      col_pass(out, ColSpec{1, -1, kzs_ / cy_, -1, -1, -1});
      col_pass(out, ColSpec{0, -1, kzs_ / cx_, -1, -1, -1});
      return;
    }
    z_r2c(in, out, kz_lim_);
    col_pass(out, ColSpec{1, -1, kz_lim_ / cy_, -1, -1, g_.keep_y});
    col_pass(out, ColSpec{0, -1, kz_lim_ / cx_, g_.keep_y, -1, g_.keep_x});
  }

  // Inverse 3D C2R with 1/(nx ny nz): spec is overwritten (used as the work buffer).
  void inv3d(C* spec, T* out, bool prune) {
    const T scale = (T)(1.0 / ((double)nx_ * ny_ * nz_));
    if (!prune) {
      col_pass(spec, ColSpec{0, +1, kzs_ / cx_, -1, -1, -1});
      col_pass(spec, ColSpec{1, +1, kzs_ / cy_, -1, -1, -1});
      z_c2r(spec, out, nzh_, scale);
      return;
    }
    col_pass(spec, ColSpec{0, +1, kz_lim_ / cx_, g_.keep_y, g_.keep_x, -1});
    col_pass(spec, ColSpec{1, +1, kz_lim_ / cy_, -1, g_.keep_y, -1});
    z_c2r(spec, out, kz_lim_, scale);
  }

  void inv3d_keep(const C* spec, C* work, T* out) {
    GNSS_HIP(hipMemcpyAsync(work, spec, sizeof(C) * nsp_, hipMemcpyDeviceToDevice, s_));
    inv3d(work, out, false);
  }

  // -------------------------------------------------------------------------------------------
  // One RK3 substage: nonlinear terms of the current state, then the IF-RK update.
  // -------------------------------------------------------------------------------------------
  Box kept_box() const {
    Box b;
    b.bx = 2 * g_.keep_x + 1;
    b.by = 2 * g_.keep_y + 1;
    b.bz = kz_lim_;
    b.kx_keep = g_.keep_x;
    b.ky_keep = g_.keep_y;
    return b;
  }

  template <int NZ>
  void zfused_launch(const ZFusedArgs<T>& a0, T scale) {
    if constexpr (sizeof(T) == 4) {
      if (zf_ == 1) return zfused_launch_v<NZ, 1>(a0, scale);
      if (zf_ == 2) return zfused_launch_v<NZ, 2>(a0, scale);
      if (zf_ == 4) return zfused_launch_v<NZ, 4>(a0, scale);
      if (zf_ == 5) return zfused_launch_v<NZ, 5>(a0, scale);
    }
    if (zf_ == 3) return zfused_launch_v<NZ, 3>(a0, scale);
    if (zf_ == 6) return zfused_launch_v<NZ, 6>(a0, scale);
    if (zf_ == 7) return zfused_launch_v<NZ, 7>(a0, scale);
    if constexpr (sizeof(T) == 8) {
      if (zf_ == 8) return zfused_launch_v<NZ, 8>(a0, scale);
    }
    if constexpr (NZ <= 512) {
      if (zf_ == 9) return zfused_launch_v<NZ, 9>(a0, scale);
    }
    zfused_launch_v<NZ, 0>(a0, scale);
  }

  template <int NZ, int V>
  void zfused_launch_v(const ZFusedArgs<T>& a0, T scale) {
    constexpr int M = NZ / 2;
    using P = typename FusedPlan<T, M, V>::type;
    constexpr int TPC = M / P::E;
    int G = std::max(1, 256 / TPC);
    // LDS regions of one row: three fields of length M per radix stage, or (V = 9) PAIR_B rows of NZ
    const long long row_lds = V == 9 ? (long long)PAIR_B * FusedPlan<T, M, V>::pad::stride(NZ) * (long long)sizeof(C)
                                     : (long long)FUSED_B * FusedPlan<T, M, V>::pad::stride(M) * (long long)sizeof(C);
    while (G > 1 && G * row_lds > LDS_MAX) --G;
    ZFusedArgs<T> a = a0;
    a.G = G;
    hipLaunchKernelGGL((k_zfused<T, NZ, V>), dim3((unsigned)cdiv(a.nrows, G)), dim3(TPC * G),
                       (size_t)(G * row_lds), s_, a, tw_.at(M), tw_.at(NZ), scale);
  }

  // x passes fused with the spectral ends of a substage, one launch per field group
  // (velocity, then each scalar).
  // x-pass mode of a substage: 0 prologue kernel + separate x passes; 1 x passes fused with the
  // prologue and the RK3 update per field group; 2 per-field x inverse passes that form their
  // work field on the fly (forward side as in mode 0); 3 the inverse side of mode 2 with the
  // forward side of mode 1.  The MFMA radix-16 path uses mode 0.
  int x_mode() const {
    if (p_.flags & GNSS_FLAG_MFMA_R16) return 0;
    if (nb_ > 1) return 4;
    return lane32() ? xf_ : (xf_ == 0 ? 0 : 2);
  }
  // The late passes and the fused forward x pass address a tile with a 32-bit byte offset per lane:
  // its column plus its row t (at most 64 rows per column), rows up to ny kzs elements apart.
  bool lane32() const { return (64LL * ny_ * kzs_ + 64) * (long long)sizeof(C) <= 0xFFFFFFFFLL; }
  // modes 4 and 5: late derivatives (the x pass reads each state field once, the y and z passes
  // apply their own derivatives; see late_inverse), with mode 2's and mode 3's forward side.
  // Not built for the colloid, whose work fields need second derivatives.
  int max_xmode() const { return colloid_ ? 3 : 5; }

  // Work field w of a substage as an operation on one or two state fields.
  XOp xop_of(int w, const C*& A, const C*& B) const {
    B = nullptr;
    if (w < 3) { A = q_[w]; return XOp{0, 0, 0}; }
    if (w == 3) { A = q_[2]; B = q_[1]; return XOp{2, 1, 2}; }  // omega_x = i (ky w - kz v)
    if (w == 4) { A = q_[0]; B = q_[2]; return XOp{2, 2, 0}; }  // omega_y = i (kz u - kx w)
    if (w == 5) { A = q_[1]; B = q_[0]; return XOp{2, 0, 1}; }  // omega_z = i (kx v - ky u)
    if (w < 6 + 3 * S_) {                                       // grad phi_s
      A = q_[3 + (w - 6) / 3];
      return XOp{1, (w - 6) % 3, 0};
    }
    if (w == 6 + 3 * S_) { A = q_[2 + S_]; return XOp{0, 0, 0}; }  // colloid c
    A = q_[3];
    return XOp{3, 0, 0};                                            // lap Z
  }

  template <int L, int CC>
  void xop_launch(const int* ws, int n, bool two, ColArgs a) {
    constexpr int TPC = L / Plan<L>::E;
    int G = std::max(1, 256 / (CC * TPC));
    while (G > 1 && (long long)G * L * CC * (long long)sizeof(C) > LDS_MAX) --G;
    a.G = G;
    const int threads = CC * TPC * G;
    const size_t lds = (size_t)G * L * CC * sizeof(C);
    const long long blocks = (long long)a.nchunks * cdiv(a.nouter, G);
    const long long per_cu = std::max<long long>(1, std::min<long long>(160 * 1024 / (long long)lds, 2048 / threads));
    XOpList<T> li{};
    for (int i = 0; i < n; ++i) {
      const C* A;
      const C* B;
      li.op[i] = xop_of(ws[i], A, B);
      li.out[i] = ws_[ws[i]];
      li.A[i] = A;
      li.B[i] = B ? B : A;
    }
    const long long grid_all = std::min<long long>(blocks * n, per_cu * n_cu_);
    if (two)
      hipLaunchKernelGGL((k_xinv_op_pipe<T, L, CC, true>), dim3((unsigned)grid_all), dim3(threads), lds, s_, li,
                         tkx_, tky_, tkz_, tw_.at(L), a, (int)blocks, (int)(blocks * n));
    else
      hipLaunchKernelGGL((k_xinv_op_pipe<T, L, CC, false>), dim3((unsigned)grid_all), dim3(threads), lds, s_, li,
                         tkx_, tky_, tkz_, tw_.at(L), a, (int)blocks, (int)(blocks * n));
  }
  template <int L>
  void xop_c(const int* ws, int n, bool two, const ColArgs& a) {
    if constexpr (sizeof(T) == 8) {
      if (xc_ == 8) xop_launch<L, 8>(ws, n, two, a);
      else xop_launch<L, 4>(ws, n, two, a);
    } else {
      if (xc_ == 16) xop_launch<L, 16>(ws, n, two, a);
      else xop_launch<L, 8>(ws, n, two, a);
    }
  }
  // Late-derivative inverse passes (x modes 4, 5).  Work buffers: ws0 u, ws1 v, ws2 w, ws3 w_y,
  // ws4 w_x, ws5 v_x (then omega_z), per scalar s ws[6+3s] phi_x, ws[7+3s] phi_y, ws[8+3s] phi.
  // The x pass reads u, v, w, phi once each (4 inputs, 4 + 2 + S outputs instead of 12 + 3S
  // inputs and 6 + 3S outputs); the y pass reads 5 + 2S fields and writes 6 + 3S, adding w_y,
  // phi_y and omega_z = v_x - i k_y u; the fused z pass forms omega_x = w_y - i k_z v,
  // omega_y = i k_z u - w_x and phi_z = i k_z phi as it loads.
  void late_inverse() {
    if (!lane32()) throw Error(GNSS_ERR_UNSUPPORTED, "late passes address a tile with 32-bit byte offsets");
    const PairOp copy{1, 1, {1, 0}, {0, 0}, {0, 0}, {0, 0}};        // out0 = A
    const PairOp copy_d{1, 2, {1, 0}, {0, 1}, {0, 0}, {0, 0}};      // out0 = A, out1 = i k A
    const PairOp curl_z{2, 2, {1, 0}, {0, -1}, {0, 1}, {0, 0}};     // out0 = A, out1 = B - i k A
    PairList<T> lx{}, ly{};
    int nxe = 0, nye = 0;
    auto put = [](PairList<T>& l, int& n, const C* A, const C* B, C* o0, C* o1, const PairOp& op) {
      l.A[n] = A;
      l.B[n] = B ? B : A;
      l.out[0][n] = o0;
      l.out[1][n] = o1 ? o1 : o0;
      l.op[n] = op;
      ++n;
    };
    put(lx, nxe, q_[0], nullptr, ws_[0], nullptr, copy);
    put(lx, nxe, q_[1], nullptr, ws_[1], ws_[5], copy_d);
    put(lx, nxe, q_[2], nullptr, ws_[2], ws_[4], copy_d);
    for (int q = 0; q < S_; ++q) put(lx, nxe, q_[3 + q], nullptr, ws_[8 + 3 * q], ws_[6 + 3 * q], copy_d);
    put(ly, nye, ws_[0], ws_[5], ws_[0], ws_[5], curl_z);
    put(ly, nye, ws_[1], nullptr, ws_[1], nullptr, copy);
    put(ly, nye, ws_[2], nullptr, ws_[2], ws_[3], copy_d);
    put(ly, nye, ws_[4], nullptr, ws_[4], nullptr, copy);
    for (int q = 0; q < S_; ++q) {
      put(ly, nye, ws_[8 + 3 * q], nullptr, ws_[8 + 3 * q], ws_[7 + 3 * q], copy_d);
      put(ly, nye, ws_[6 + 3 * q], nullptr, ws_[6 + 3 * q], nullptr, copy);
    }
    const int CC = pair_chunk(), CY = yc_ ? yc_ : CC;
    ColArgs ax;
    ax.estride = (long long)ny_ * kzs_;
    ax.ostride = kzs_;
    ax.outer_keep = g_.keep_y;
    ax.outer_n = ny_;
    ax.nouter = 2 * g_.keep_y + 1;
    ax.nchunks = kz_lim_ / CC;
    ax.in_keep = g_.keep_x;
    ax.out_keep = -1;
    ax.G = 1;
    ColArgs ay;
    ay.estride = kzs_;
    ay.ostride = (long long)ny_ * kzs_;
    ay.outer_keep = -1;
    ay.outer_n = nx_;
    ay.nouter = nx_;
    ay.nchunks = kz_lim_ / CY;
    ay.in_keep = g_.keep_y;
    ay.out_keep = -1;
    ay.G = 1;
    batch_outer(ax);
    batch_outer(ay);
    pair_pass(lx, nxe, ax, nx_, tkx_, "fft_x_inv_late", CC);
    pair_pass(ly, nye, ay, ny_, tky_, "fft_y_inv_late", CY);
  }
  int pair_chunk() const { return xc_; }
  // A pass of a batched step covers the outer indices of every member.
  void batch_outer(ColArgs& a) const {
    if (bm_ <= 1) return;
    a.nouter_m = a.nouter;
    a.nouter *= bm_;
    a.mstride = nsp_;
  }
  // With lp = 1 an entry's two outputs go through one launch (lp = 2: transformed one after the
  // other through one LDS region, which halves the kernel's LDS; lp = 3: together, with the next
  // tile's input loaded before they transform); with lp = 0 they are split into
  // single-output entries that read the inputs again, the second outputs first (they read fields
  // the first outputs overwrite in place).  Then one launch per (inputs, outputs) kind with cb,
  // else one per entry.
  void pair_pass(const PairList<T>& li, int n, ColArgs a, int L, const T* kv, const char* name, int CC) {
    if (lp_) return pair_kinds(li, n, a, L, kv, name, CC);
    PairList<T> first{}, second{};
    int n1 = 0, n2 = 0;
    for (int i = 0; i < n; ++i) {
      const PairOp& op = li.op[i];
      const bool first_uses_b = op.c[0] != 0 || op.d[0] != 0;
      const PairOp o0{first_uses_b ? 2 : 1, 1, {op.a[0], 0}, {op.b[0], 0}, {op.c[0], 0}, {op.d[0], 0}};
      first.A[n1] = li.A[i];
      first.B[n1] = li.B[i];
      first.out[0][n1] = first.out[1][n1] = li.out[0][i];
      first.op[n1++] = o0;
      if (op.nout == 2) {
        const bool uses_b = op.c[1] != 0 || op.d[1] != 0;
        PairOp o1{uses_b ? 2 : 1, 1, {op.a[1], 0}, {op.b[1], 0}, {op.c[1], 0}, {op.d[1], 0}};
        second.A[n2] = li.A[i];
        second.B[n2] = li.B[i];
        second.out[0][n2] = second.out[1][n2] = li.out[1][i];
        second.op[n2++] = o1;
      }
    }
    pair_kinds(second, n2, a, L, kv, name, CC);
    pair_kinds(first, n1, a, L, kv, name, CC);
  }
  void pair_kinds(const PairList<T>& li, int n, ColArgs a, int L, const T* kv, const char* name, int CC) {
    auto rows = [&](int keep) { return keep < 0 ? L : std::min(L, 2 * keep + 1); };
    const double col = (double)a.nouter * a.nchunks * CC * sizeof(C);
    static const int KIN[4] = {1, 1, 2, 2}, KOUT[4] = {1, 2, 2, 1};
    for (int kind = 0; kind < 4; ++kind) {
      const int nin = KIN[kind], nout = KOUT[kind];
      std::vector<int> ents;
      for (int i = 0; i < n; ++i)
        if (li.op[i].nin == nin && li.op[i].nout == nout) ents.push_back(i);
      const int groups = cb_ ? (ents.empty() ? 0 : 1) : (int)ents.size();
      for (int gi = 0; gi < groups; ++gi) {
        const int i0 = cb_ ? 0 : gi, i1 = cb_ ? (int)ents.size() : gi + 1;
        PairList<T> sub{};
        for (int i = i0; i < i1; ++i) {
          const int j = ents[i], k = i - i0;
          sub.A[k] = li.A[j];
          sub.B[k] = li.B[j];
          sub.out[0][k] = li.out[0][j];
          sub.out[1][k] = li.out[1][j];
          sub.op[k] = li.op[j];
        }
        const int m = i1 - i0;
        timed(name, col * m * (nin * rows(a.in_keep) + nout * rows(a.out_keep)), [&] {
          switch (L) {
            case 16: pair_c<16>(sub, m, a, kv, CC, kind); break;
            case 24: pair_c<24>(sub, m, a, kv, CC, kind); break;
            case 32: pair_c<32>(sub, m, a, kv, CC, kind); break;
            case 48: pair_c<48>(sub, m, a, kv, CC, kind); break;
            case 64: pair_c<64>(sub, m, a, kv, CC, kind); break;
            case 96: pair_c<96>(sub, m, a, kv, CC, kind); break;
            case 128: pair_c<128>(sub, m, a, kv, CC, kind); break;
            case 192: pair_c<192>(sub, m, a, kv, CC, kind); break;
            case 256: pair_c<256>(sub, m, a, kv, CC, kind); break;
            case 384: pair_c<384>(sub, m, a, kv, CC, kind); break;
            case 512: pair_c<512>(sub, m, a, kv, CC, kind); break;
            case 1024: pair_c<1024>(sub, m, a, kv, CC, kind); break;
            default: throw Error(GNSS_ERR_UNSUPPORTED, "late pass length " + std::to_string(L));
          }
        });
      }
    }
  }
  template <int L>
  void pair_c(const PairList<T>& li, int n, const ColArgs& a, const T* kv, int CC, int kind) {
    if constexpr (sizeof(T) == 8) {
      if (CC == 8) pair_k<L, 8>(li, n, a, kv, kind);
      else pair_k<L, 4>(li, n, a, kv, kind);
    } else {
      if (CC == 16) pair_k<L, 16>(li, n, a, kv, kind);
      else pair_k<L, 8>(li, n, a, kv, kind);
    }
  }
  // two-output form of lp = 1, 2, 3
  template <int L, int CC, int NIN>
  void pair_k2(const PairList<T>& li, int n, const ColArgs& a, const T* kv) {
    if (lp_ == 2) pair_launch<L, CC, NIN, 2, PAIR_SEQ>(li, n, a, kv);
    else if (lp_ == 3) pair_launch<L, CC, NIN, 2, PAIR_BATCHED_PF>(li, n, a, kv);
    else pair_launch<L, CC, NIN, 2, PAIR_BATCHED>(li, n, a, kv);
  }
  template <int L, int CC>
  void pair_k(const PairList<T>& li, int n, const ColArgs& a, const T* kv, int kind) {
    if (kind == 0) pair_launch<L, CC, 1, 1, PAIR_BATCHED>(li, n, a, kv);
    else if (kind == 1) pair_k2<L, CC, 1>(li, n, a, kv);
    else if (kind == 2) pair_k2<L, CC, 2>(li, n, a, kv);
    else pair_launch<L, CC, 2, 1, PAIR_BATCHED>(li, n, a, kv);
  }
  template <int L, int CC, int NIN, int NOUT, int FORM>
  void pair_launch(const PairList<T>& li, int n, ColArgs a, const T* kv) {
    constexpr int TPC = L / Plan<L>::E;
    // two batched outputs use two LDS regions, and the twiddles have a copy after them: up to the
    // 160 KiB a gfx950 workgroup can hold
    constexpr long long PAIR_LDS_MAX = 160 * 1024;
    const long long region = (long long)(NOUT == 2 && FORM == PAIR_SEQ ? 1 : NOUT) * L * CC * (long long)sizeof(C);
    const long long twb = (long long)L * (long long)sizeof(C);
    int G = std::max(1, 256 / (CC * TPC));
    while (G > 1 && G * region + twb > PAIR_LDS_MAX) --G;
    if (region + twb > PAIR_LDS_MAX) throw Error(GNSS_ERR_UNSUPPORTED, "late pass tile does not fit the LDS");
    // the kernel takes the rows beyond the 2/3 rule's L / 3 for zeros without reading them
    if (a.in_keep != L / 3) throw Error(GNSS_ERR_STATE, "late pass input rows must be the kept ones");
    a.G = G;
    const int threads = CC * TPC * G;
    const size_t lds = (size_t)(G * region + twb);
    const long long blocks = (long long)a.nchunks * cdiv(a.nouter, G);
    const long long per_cu = std::max<long long>(1, std::min<long long>(160 * 1024 / (long long)lds, 2048 / threads));
    const long long grid = std::min<long long>(blocks * n, per_cu * n_cu_);
    hipLaunchKernelGGL((k_inv_pair_pipe<T, L, CC, NIN, NOUT, FORM>), dim3((unsigned)grid), dim3(threads), lds, s_, li,
                       kv, tw_.at(L), a, (int)blocks, (int)(blocks * n));
  }

  // x inverse passes of all work fields formed from the state (modes 2 and 3): one-input and
  // two-input fields in one launch each with cb, else one launch per field
  void x_inv_ops() {
    std::vector<int> one, two;
    for (int w = 0; w < W_; ++w) {
      const C* A;
      const C* B;
      (xop_of(w, A, B).kind == 2 ? two : one).push_back(w);
    }
    for (int pass = 0; pass < 2; ++pass) {
      const std::vector<int>& v = pass ? two : one;
      if (v.empty()) continue;
      if (cb_) {
        x_inv_op(v.data(), (int)v.size(), pass == 1);
      } else {
        for (int w : v) x_inv_op(&w, 1, pass == 1);
      }
    }
  }
  void x_inv_op(const int* ws, int n, bool two) {
    ColArgs a;
    a.estride = (long long)ny_ * kzs_;
    a.ostride = kzs_;
    a.outer_keep = g_.keep_y;
    a.outer_n = ny_;
    a.nouter = 2 * g_.keep_y + 1;
    a.nchunks = kz_lim_ / xc_;
    a.in_keep = g_.keep_x;
    a.out_keep = -1;
    a.G = 1;
    const int nin = two ? 2 : 1;
    const double bytes = n * (double)a.nouter * a.nchunks * xc_ * (nin * (2.0 * g_.keep_x + 1) + nx_) * sizeof(C);
    timed("fft_x_inv_op", bytes, [&] {
      switch (nx_) {
        case 16: xop_c<16>(ws, n, two, a); break;
        case 24: xop_c<24>(ws, n, two, a); break;
        case 32: xop_c<32>(ws, n, two, a); break;
        case 48: xop_c<48>(ws, n, two, a); break;
        case 64: xop_c<64>(ws, n, two, a); break;
        case 96: xop_c<96>(ws, n, two, a); break;
        case 128: xop_c<128>(ws, n, two, a); break;
        case 192: xop_c<192>(ws, n, two, a); break;
        case 256: xop_c<256>(ws, n, two, a); break;
        case 384: xop_c<384>(ws, n, two, a); break;
        case 512: xop_c<512>(ws, n, two, a); break;
        case 1024: xop_c<1024>(ws, n, two, a); break;
        default: throw Error(GNSS_ERR_UNSUPPORTED, "x pass length " + std::to_string(nx_));
      }
    });
  }

  template <int L, int CC, int GRP>
  void xgrp_launch(bool pro, int G, int nchunks, int scalar, const RkCoef& rc) {
    constexpr int TPC = L / Plan<L>::E;
    const int nouter = 2 * g_.keep_y + 1;
    const long long blocks = (long long)nchunks * cdiv(nouter, G);
    // the forward kernel keeps a copy of the twiddles and of the x wavenumbers after its data
    // region, which the velocity group's right-hand sides may take over once the transforms are done
    const long long data = (long long)G * L * CC;
    const long long rhs = xfwd_rhs_entries<T, L, CC, GRP == 0>(CC * TPC * G);
    const size_t lds = pro ? (size_t)data * sizeof(C)
                           : (size_t)(std::max(data, rhs) + L) * sizeof(C) + (size_t)L * sizeof(T);
    if (lds > 160 * 1024) throw Error(GNSS_ERR_UNSUPPORTED, "fused x pass tile does not fit the LDS");
    if (!pro && g_.keep_x != L / 3) throw Error(GNSS_ERR_STATE, "fused x pass rows must be the kept ones");
    if (pro)
      hipLaunchKernelGGL((k_xinv_prologue<T, L, CC, GRP>), dim3((unsigned)blocks), dim3(CC * TPC * G), lds, s_,
                         g_, ptrs(q_), ptrs(ws_), tkx_, tky_, tkz_, tw_.at(L), scalar, nchunks, G);
    else
      hipLaunchKernelGGL((k_xfwd_rk<T, L, CC, GRP == 0>), dim3((unsigned)blocks), dim3(CC * TPC * G), lds, s_, g_,
                         ptrs(q_), ptrs(dq_), ptrs(nl_), tkx_, tky_, tkz_, tw_.at(L), rc, scalar, nchunks, G);
  }
  template <int L, int CC>
  void xgrp_k(bool pro, int G, int nchunks, int grp, const RkCoef& rc) {
    if (grp < 0) xgrp_launch<L, CC, 0>(pro, G, nchunks, 0, rc);
    else if (pro && colloid_ && grp == S_ - 1) xgrp_launch<L, CC, 2>(pro, G, nchunks, grp, rc);
    else xgrp_launch<L, CC, 1>(pro, G, nchunks, grp, rc);
  }
  template <int L>
  void xgrp_c(bool pro, int G, int nchunks, int grp, const RkCoef& rc) {
    if constexpr (sizeof(T) == 8) {
      if (xc_ == 8) xgrp_k<L, 8>(pro, G, nchunks, grp, rc);
      else xgrp_k<L, 4>(pro, G, nchunks, grp, rc);
    } else {
      if (xc_ == 16) xgrp_k<L, 16>(pro, G, nchunks, grp, rc);
      else xgrp_k<L, 8>(pro, G, nchunks, grp, rc);
    }
  }
  // grp < 0: velocity group; grp = s: scalar s (the colloid group when s is the colloid).
  void x_fused(bool pro, int grp, const RkCoef& rc) {
    const int L = nx_, CC = xc_;
    const int TPC = L / elems_per_thread(L);
    int G = std::max(1, 256 / (CC * TPC));
    const long long twb = pro ? 0 : (long long)L * (long long)(sizeof(C) + sizeof(T));
    while (G > 1 && (long long)G * L * CC * (long long)sizeof(C) + twb > LDS_MAX) --G;
    const int nchunks = kz_lim_ / CC;
    const int nouter = 2 * g_.keep_y + 1;
    const double kept_rows = 2 * g_.keep_x + 1;
    const bool cgrp = pro && colloid_ && grp == S_ - 1;
    const int nin = grp < 0 ? 3 : (cgrp ? 2 : 1), nout = grp < 0 ? 6 : (cgrp ? 5 : 3);
    const double per_col = pro ? (nin * kept_rows + nout * (double)L)
                               : (nin * (double)L + nin * kept_rows * (2 + (rc.first ? 0 : 1) + (rc.last ? 0 : 1)));
    const double bytes = (double)nouter * nchunks * CC * per_col * sizeof(C);
    timed(pro ? "fft_x_inv_prologue" : "fft_x_fwd_rk", bytes, [&] {
      switch (L) {
        case 16: xgrp_c<16>(pro, G, nchunks, grp, rc); break;
        case 24: xgrp_c<24>(pro, G, nchunks, grp, rc); break;
        case 32: xgrp_c<32>(pro, G, nchunks, grp, rc); break;
        case 48: xgrp_c<48>(pro, G, nchunks, grp, rc); break;
        case 64: xgrp_c<64>(pro, G, nchunks, grp, rc); break;
        case 96: xgrp_c<96>(pro, G, nchunks, grp, rc); break;
        case 128: xgrp_c<128>(pro, G, nchunks, grp, rc); break;
        case 192: xgrp_c<192>(pro, G, nchunks, grp, rc); break;
        case 256: xgrp_c<256>(pro, G, nchunks, grp, rc); break;
        case 384: xgrp_c<384>(pro, G, nchunks, grp, rc); break;
        case 512: xgrp_c<512>(pro, G, nchunks, grp, rc); break;
        case 1024: xgrp_c<1024>(pro, G, nchunks, grp, rc); break;
        default: throw Error(GNSS_ERR_UNSUPPORTED, "fused x pass length " + std::to_string(L));
      }
    });
  }

  // z pass of a substage fused with the physical products: ws (after the x, y inverse passes)
  // -> nl (before the y, x forward passes), kz < kz_lim_.
  void z_fused() {
    ZFusedArgs<T> a{};
    for (int w = 0; w < W_; ++w) a.ws[w] = ws_[w];
    for (int f = 0; f < F_; ++f) a.nl[f] = nl_[f];
    a.nrows = (long long)bm_ * nx_ * ny_;  // members lie back to back: their rows continue
    a.kzs = kzs_;
    a.S = S_;
    a.colloid = colloid_;
    a.alpha = (T)p_.alpha;
    a.late = ((p_.flags & GNSS_FLAG_UNFUSED) == 0 && x_mode() >= 4) ? 1 : 0;
    a.kz = tkz_;
    if (kz_lim_ != zfused_lim<T>(nz_))
      throw Error(GNSS_ERR_UNSUPPORTED, "fused z pass: pruned kz limit " + std::to_string(kz_lim_) + " differs from " +
                                            std::to_string(zfused_lim<T>(nz_)));
    const T scale = (T)(1.0 / ((double)nx_ * ny_ * nz_));
    const double bytes = (double)bm_ * nx_ * ny_ * (double)kz_lim_ * sizeof(C) * (W_ + F_);
    timed("fft_z_fused", bytes, [&] { GNSS_Z_SWITCH(zfused_launch, a, scale); });
  }

  RkCoef rk_coef(int s) const {
    RkCoef c;
    c.a = RK_A[s];
    c.b = RK_B[s];
    c.tau = (RK_C[s + 1] - RK_C[s]) * p_.dt;
    c.dt = p_.dt;
    c.first = s == 0;
    c.last = s == 2;
    const double nu = 1.0 / p_.re;
    for (int f = 0; f < MAXF; ++f) c.nu[f] = 0;
    for (int f = 0; f < 3; ++f) c.nu[f] = nu;
    for (int q = 0; q < S_; ++q) c.nu[3 + q] = nu / p_.sc[q];
    c.nbox = 0;
    c.mstride = 0;
    c.nu_m = nullptr;
    return c;
  }

  void substage(int s) {
    const bool unfused = (p_.flags & GNSS_FLAG_UNFUSED) != 0;
    const Box b = kept_box();
    const long long nb = (long long)b.bx * b.by * b.bz;
    const RkCoef c = rk_coef(s);
    const int xm = unfused ? 0 : x_mode();
    const bool xf = xm == 1;                  // group-fused inverse side
    const bool xfwd = xm == 1 || xm == 3 || xm == 5;  // group-fused forward side
    const bool late = xm >= 4;                // late-derivative inverse side
    if (xf) {
      x_fused(true, -1, c);
      for (int q = 0; q < S_; ++q) x_fused(true, q, c);
    } else if (xm == 0) {
      timed("k_prologue", (double)nb * (F_ + W_) * sizeof(C), [&] {
        hipLaunchKernelGGL(k_prologue<T>, dim3((unsigned)cdiv(nb, 256)), dim3(256), 0, s_, g_, b, ptrs(q_),
                           ptrs(ws_), tkx_, tky_, tkz_, S_, colloid_, nb);
      });
    }
    if (unfused) {
      ensure_unfused_buffers();
      for (int w = 0; w < W_; ++w) inv3d(ws_[w], wp_[w], true);
      timed("k_products", (double)nph_ * (W_ + F_) * sizeof(T), [&] {
        hipLaunchKernelGGL(k_products<T>, dim3((unsigned)cdiv(nph_, 256)), dim3(256), 0, s_, rptrs(wp_),
                           rptrs(nlp_), S_, colloid_, (T)p_.alpha, nph_);
      });
      for (int f = 0; f < F_; ++f) fwd3d(nlp_[f], nl_[f], true);
    } else {
      if (late) {
        late_inverse();
      } else {
        if (xm >= 2) x_inv_ops();
        else if (xm == 0) col_passes(ws_.data(), W_, ColSpec{0, +1, kz_lim_ / cx_, g_.keep_y, g_.keep_x, -1});
        col_passes(ws_.data(), W_, ColSpec{1, +1, kz_lim_ / cy_, -1, g_.keep_y, -1});
      }
      z_fused();
      col_passes(nl_.data(), F_, ColSpec{1, -1, kz_lim_ / cy_, -1, -1, g_.keep_y});
      if (!xfwd) col_passes(nl_.data(), F_, ColSpec{0, -1, kz_lim_ / cx_, g_.keep_y, -1, g_.keep_x});
    }
    if (xfwd) {
      x_fused(false, -1, c);
      for (int q = 0; q < S_; ++q) x_fused(false, q, c);
      return;
    }
    RkCoef cm = c;
    if (bm_ > 1) {
      cm.nbox = nb;
      cm.mstride = nsp_;
      cm.nu_m = d_num_;
    }
    const long long n = nb * bm_;
    const double rk_bytes = (double)n * F_ * sizeof(C) * (c.first ? 3 : (c.last ? 4 : 5));
    timed("k_rk", rk_bytes, [&] {
      hipLaunchKernelGGL(k_rk<T>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s_, g_, b, ptrs(q_), ptrs(dq_),
                         ptrs(nl_), tkx_, tky_, tkz_, cm, F_, n);
    });
  }

  // -------------------------------------------------------------------------------------------
  gnss_params p_;
  int colloid_ = 0;
  int nx_ = 0, ny_ = 0, nz_ = 0, nzh_ = 0, kzs_ = 0, S_ = 0, F_ = 0, W_ = 0, cx_ = 8, cy_ = 8, zf_ = 0, xf_ = 1;
  int cp_ = 0, xc_ = 8, zp_ = 0, kp_ = 0, cb_ = 0, lp_ = 1, n_cu_ = 256;
  // batched contexts: members held (nb_), members the launches in flight cover (bm_: nb_ inside a
  // step, 1 elsewhere), member parameters, base pointers of the state, per-member viscosities
  int nb_ = 1, bm_ = 1, yc_ = 0;
  std::vector<gnss_params> mp_;
  std::vector<C*> qb_, dqb_;
  double* d_num_ = nullptr;
  int kz_lim_ = 0;
  long long nsp_ = 0, nph_ = 0;
  Geom g_{};
  hipStream_t s_ = nullptr;
  std::vector<C*> q_, dq_, ws_, nl_;
  std::vector<T*> wp_, nlp_, phys_;
  std::map<int, C*> tw_;
  double *dkx_ = nullptr, *dky_ = nullptr, *dkz_ = nullptr;
  T *tkx_ = nullptr, *tky_ = nullptr, *tkz_ = nullptr;
  double *d_part_ = nullptr, *d_spec_ = nullptr, *d_prof_ = nullptr, *d_cmax_ = nullptr;
  int* d_flag_ = nullptr;
  void* scratch_ = nullptr;
  size_t scratch_n_ = 0;
  std::vector<void*> allocs_;
  bool have_k_ = false, have_ic_ = false, pause_timer_ = false;
  double t_ = 0;
  int64_t step_ = 0;
  KernelTimer timer_;
};

}  // namespace gnss
