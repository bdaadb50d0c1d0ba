// engine.h — host-side engine of the pseudospectral step (one context = one HIP stream).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gnss.h"

namespace gnss {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define GNSS_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::gnss::Error(GNSS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Per-kernel timing with HIP events on the engine's stream.
class KernelTimer {
 public:
  void enable(bool on, hipStream_t s);
  bool on() const { return on_; }
  void begin(const char* name, double bytes);
  void end();
  std::string report();
  ~KernelTimer();

 private:
  struct Rec {
    std::string name;
    double bytes;
    hipEvent_t a, b;
  };
  struct Agg {
    long long launches = 0;
    double ms = 0, bytes = 0;
  };
  void drain();
  hipEvent_t get_event();
  bool on_ = false;
  hipStream_t stream_ = nullptr;
  std::vector<Rec> pending_;
  std::vector<hipEvent_t> pool_;
  std::map<std::string, Agg> agg_;
};

// Precision-independent interface used by the C ABI.
class EngineBase {
 public:
  virtual ~EngineBase() {}
  virtual void init_wavenumbers() = 0;
  virtual void get_wavenumbers(double* kx, double* ky, double* kz) = 0;
  virtual void get_mask(uint8_t* mask) = 0;
  virtual void init_jet() = 0;
  virtual void forward(int field) = 0;
  virtual void inverse(int field) = 0;
  virtual void dealias() = 0;
  virtual void step(int nsteps) = 0;
  virtual void stats(gnss_stats_out* out, double* yprof, double* spectrum) = 0;
  // The same statistics with every partial sum in `scratch` (device memory the caller owns, at
  // least stats_scratch_doubles() doubles): no allocation on the device or the host, no timer
  // events; results bitwise equal to stats().
  virtual size_t stats_scratch_doubles() const = 0;
  virtual void stats_noalloc(double* scratch, gnss_stats_out* out, double* yprof, double* spectrum) = 0;
  virtual void get_field(int field, int space, double* host) = 0;
  virtual void set_field(int field, int space, const double* host) = 0;
  virtual void sizes(gnss_sizes* out) const = 0;
  virtual void time(double* t, int64_t* step) const = 0;
  virtual void set_clock(int64_t step, double t) = 0;
  virtual void timing_enable(bool on) = 0;
  virtual std::string timing_report() = 0;
  virtual void bench_fft(int reps, double* ms_fwd, double* ms_inv) = 0;
  virtual void bench_copy(int reps, double* ms, double* bytes) = 0;
  virtual std::string autotune(int reps) = 0;
  virtual void set_tuning(const std::string& spec) = 0;
  // Diagnostics: one physical plane of a field (or |omega|), and an iso-surface area.
  virtual void get_slice(int field, int axis, int index, double* out) = 0;
  virtual double surface_area(int field, double iso) = 0;
  virtual bool is_colloid(int field) const = 0;
  // Ensemble support: the stream the engine launches on, the kernels of one RK3 step without any
  // host synchronisation, memset or allocation (capturable into a hipGraph), the host-side clock
  // of a replayed step, and the NaN/Inf watchdog.
  virtual hipStream_t stream() const = 0;
  virtual void enqueue_rk3_step() = 0;
  virtual void advance_clock(int nsteps) = 0;
  virtual void check_finite() = 0;
  // Batched contexts (one engine holding several members of one grid, stepped by shared
  // launches): the number of members, and the member that the per-member calls (init_jet,
  // stats, fields, diagnostics) address.
  virtual int members() const { return 1; }
  virtual void select_member(int m) = 0;
  std::string err;
};

// Grid sizes the radix plans cover (products of 2s and 3s with a compiled plan).
inline bool supported_len(int n) {
  switch (n) {
    case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256:
    case 384: case 512: case 1024:
      return true;
    default:
      return false;
  }
}

// The parameter checks shared by gnss_create and gnss_ensemble_create: an empty string when the
// set is acceptable, else what is wrong with it.
inline std::string params_problem(const gnss_params& p) {
  if (!supported_len(p.nx) || !supported_len(p.ny) || !supported_len(p.nz))
    return "grid " + std::to_string(p.nx) + "x" + std::to_string(p.ny) + "x" + std::to_string(p.nz) +
           ": each size must be one of 16 24 32 48 64 96 128 192 256 384 512 1024";
  if (!(p.lx > 0 && p.ly > 0 && p.lz > 0)) return "box lengths must be > 0";
  if (!(p.re > 0)) return "re must be > 0";
  if (!(p.dt > 0)) return "dt must be > 0";
  if (p.n_scalars < 0 || p.n_scalars > GNSS_MAX_SCALARS) return "n_scalars must be in [0, 4]";
  for (int s = 0; s < p.n_scalars; ++s)
    if (!(p.sc[s] > 0)) return "sc must be > 0";
  if (p.precision != GNSS_FP64 && p.precision != GNSS_FP32) return "precision must be GNSS_FP64 or GNSS_FP32";
  return "";
}

// Directory of the loaded engine library (for its tuning table).
std::string library_dir();

EngineBase* make_engine_f64(const gnss_params& p);
EngineBase* make_engine_f32(const gnss_params& p);
// nb members of one grid (same sizes, scalars, precision, flags, box and dt; their own Re, Sc,
// theta and seed) in one batched engine.
EngineBase* make_engine_batch_f64(const gnss_params* p, int nb);
EngineBase* make_engine_batch_f32(const gnss_params* p, int nb);

}  // namespace gnss
