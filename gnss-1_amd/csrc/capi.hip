// capi.hip — the extern "C" boundary of include/gnss.h for the MI355X engine.
//
// Every entry point catches engine exceptions and turns them into a status code plus a message
// retrievable with gnss_last_error (the reference returns void and aborts or prints,
// fftfuncs.cu:297-300; allocate.cu:109-142).
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>

#include <dlfcn.h>

#include "engine.h"

#define GNSS_API extern "C" __attribute__((visibility("default")))

using gnss::EngineBase;
using gnss::Error;

struct gnss_ctx {
  EngineBase* eng = nullptr;
  gnss_params p{};
  int dev = 0;  // device current at gnss_create (the engine's stream lives there)
  std::string err;
};

static thread_local std::string g_create_err;

// ---------------------------------------------------------------------------------------------
// Fixed device scratch of gnss_stats_noalloc: one copy per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel or a copy (on the host it names a stand-in, not
// device memory): gnss_create resolves the device address once per device with
// hipGetSymbolAddress while that device is current, and every use goes through that address.
// ---------------------------------------------------------------------------------------------
#define GNSS_STATS_SCRATCH_DOUBLES (5u << 20)  // 40 MiB
static __device__ __align__(256) double scratch[GNSS_STATS_SCRATCH_DOUBLES];
static constexpr int GNSS_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[GNSS_MAX_DEVICES] = {};

// The device address of `scratch` on device `dev` (current), resolved on the first call there.
static void* stats_scratch(int dev) {
  if (dev < 0 || dev >= GNSS_MAX_DEVICES) throw Error(GNSS_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      throw Error(GNSS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

namespace gnss {

// ---------------------------------------------------------------------------------------------
// KernelTimer
// ---------------------------------------------------------------------------------------------
void KernelTimer::enable(bool on, hipStream_t s) {
  drain();
  on_ = on;
  stream_ = s;
  if (on) agg_.clear();
}

hipEvent_t KernelTimer::get_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  GNSS_HIP(hipEventCreate(&e));
  return e;
}

void KernelTimer::begin(const char* name, double bytes) {
  Rec r;
  r.name = name;
  r.bytes = bytes;
  r.a = get_event();
  r.b = get_event();
  GNSS_HIP(hipEventRecord(r.a, stream_));
  pending_.push_back(r);
  if (pending_.size() > 4096) drain();
}

void KernelTimer::end() { GNSS_HIP(hipEventRecord(pending_.back().b, stream_)); }

void KernelTimer::drain() {
  if (pending_.empty()) return;
  GNSS_HIP(hipEventSynchronize(pending_.back().b));
  for (Rec& r : pending_) {
    float ms = 0;
    GNSS_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    Agg& a = agg_[r.name];
    a.launches += 1;
    a.ms += ms;
    a.bytes += r.bytes;
    pool_.push_back(r.a);
    pool_.push_back(r.b);
  }
  pending_.clear();
}

std::string KernelTimer::report() {
  drain();
  std::string s = "{";
  bool first = true;
  char buf[256];
  for (auto& kv : agg_) {
    snprintf(buf, sizeof buf, "%s\"%s\": {\"launches\": %lld, \"ms\": %.6f, \"bytes\": %.0f}", first ? "" : ", ",
             kv.first.c_str(), kv.second.launches, kv.second.ms, kv.second.bytes);
    s += buf;
    first = false;
  }
  return s + "}";
}

KernelTimer::~KernelTimer() {
  for (Rec& r : pending_) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

std::string library_dir() {
  Dl_info info;
  if (dladdr((void*)&library_dir, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? "." : p.substr(0, k);
  }
  return ".";
}

}  // namespace gnss

template <typename Fn> static int guard(gnss_ctx* c, Fn&& fn) {
  if (!c || !c->eng) return GNSS_ERR_ARG;
  try {
    fn();
    return GNSS_OK;
  } catch (const Error& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    c->err = e.what();
    return GNSS_ERR_DEVICE;
  }
}

GNSS_API int gnss_create(const gnss_params* p, gnss_ctx** out) {
  if (!p || !out) {
    g_create_err = "null argument";
    return GNSS_ERR_ARG;
  }
  *out = nullptr;
  const std::string bad = gnss::params_problem(*p);
  if (!bad.empty()) {
    g_create_err = bad;
    return GNSS_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    g_create_err = "no HIP device available";
    return GNSS_ERR_DEVICE;
  }
  gnss_ctx* c = new gnss_ctx();
  c->p = *p;
  try {
    GNSS_HIP(hipGetDevice(&c->dev));
    (void)stats_scratch(c->dev);  // once per device, with the context's device current
    c->eng = p->precision == GNSS_FP64 ? gnss::make_engine_f64(*p) : gnss::make_engine_f32(*p);
  } catch (const Error& e) {
    g_create_err = e.what();
    delete c;
    return e.code;
  } catch (const std::exception& e) {
    g_create_err = e.what();
    delete c;
    return GNSS_ERR_DEVICE;
  }
  *out = c;
  return GNSS_OK;
}

GNSS_API int gnss_destroy(gnss_ctx* c) {
  if (!c) return GNSS_OK;
  delete c->eng;
  delete c;
  return GNSS_OK;
}

GNSS_API const char* gnss_last_error(const gnss_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

GNSS_API int gnss_get_sizes(const gnss_ctx* c, gnss_sizes* o) {
  if (!c || !c->eng || !o) return GNSS_ERR_ARG;
  c->eng->sizes(o);
  return GNSS_OK;
}

GNSS_API int gnss_init_wavenumbers(gnss_ctx* c) { return guard(c, [&] { c->eng->init_wavenumbers(); }); }
GNSS_API int gnss_get_wavenumbers(gnss_ctx* c, double* kx, double* ky, double* kz) {
  return guard(c, [&] { c->eng->get_wavenumbers(kx, ky, kz); });
}
GNSS_API int gnss_get_mask(gnss_ctx* c, uint8_t* mask) {
  if (!mask) return GNSS_ERR_ARG;
  return guard(c, [&] { c->eng->get_mask(mask); });
}
GNSS_API int gnss_init_jet(gnss_ctx* c) { return guard(c, [&] { c->eng->init_jet(); }); }
GNSS_API int gnss_forward(gnss_ctx* c, int f) { return guard(c, [&] { c->eng->forward(f); }); }
GNSS_API int gnss_inverse(gnss_ctx* c, int f) { return guard(c, [&] { c->eng->inverse(f); }); }
GNSS_API int gnss_dealias(gnss_ctx* c) { return guard(c, [&] { c->eng->dealias(); }); }
GNSS_API int gnss_step(gnss_ctx* c, int n) { return guard(c, [&] { c->eng->step(n); }); }
GNSS_API int gnss_stats(gnss_ctx* c, gnss_stats_out* o, double* yprof, double* spectrum) {
  return guard(c, [&] { c->eng->stats(o, yprof, spectrum); });
}
// gnss_stats without any allocation: the partial sums live in the fixed per-device scratch above
// (resolved by gnss_create), so a caller that cannot allocate can take statistics.  Every call on
// one GPU shares that scratch: calls that overlap on one GPU (several threads or contexts) race.
// Each call writes every partial before reading it, so nothing carries over between calls.
GNSS_API int gnss_stats_noalloc(gnss_ctx* c, gnss_stats_out* o, double* yprof, double* spectrum) {
  return guard(c, [&] {
    const size_t need = c->eng->stats_scratch_doubles();
    if (need > GNSS_STATS_SCRATCH_DOUBLES)
      throw Error(GNSS_ERR_ALLOC, "gnss_stats_noalloc: this grid's statistics need " + std::to_string(need * 8 >> 10) +
                                      " KiB of scratch, above the fixed " +
                                      std::to_string((size_t)GNSS_STATS_SCRATCH_DOUBLES * 8 >> 20) +
                                      " MiB; use gnss_stats");
    void* ptr = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      ptr = g_scratch_ptr[c->dev];
    }
    if (!ptr) throw Error(GNSS_ERR_STATE, "gnss_stats_noalloc: scratch of this device not resolved");
    c->eng->stats_noalloc(static_cast<double*>(ptr), o, yprof, spectrum);
  });
}
GNSS_API int gnss_get_field(gnss_ctx* c, int f, int space, double* host) {
  if (!host) return GNSS_ERR_ARG;
  return guard(c, [&] { c->eng->get_field(f, space, host); });
}
GNSS_API int gnss_set_field(gnss_ctx* c, int f, int space, const double* host) {
  if (!host) return GNSS_ERR_ARG;
  return guard(c, [&] { c->eng->set_field(f, space, host); });
}
GNSS_API int gnss_get_time(const gnss_ctx* c, double* t, int64_t* step) {
  if (!c || !c->eng) return GNSS_ERR_ARG;
  c->eng->time(t, step);
  return GNSS_OK;
}
GNSS_API int gnss_timing_enable(gnss_ctx* c, int on) { return guard(c, [&] { c->eng->timing_enable(on != 0); }); }
GNSS_API int gnss_timing_report(gnss_ctx* c, char* buf, int len) {
  if (!buf || len <= 0) return GNSS_ERR_ARG;
  return guard(c, [&] {
    std::string s = c->eng->timing_report();
    if ((int)s.size() + 1 > len) throw Error(GNSS_ERR_ARG, "report buffer too small");
    memcpy(buf, s.c_str(), s.size() + 1);
  });
}
GNSS_API int gnss_bench_copy(gnss_ctx* c, int reps, double* ms, double* bytes) {
  if (!ms || !bytes) return GNSS_ERR_ARG;
  return guard(c, [&] { c->eng->bench_copy(reps, ms, bytes); });
}
GNSS_API int gnss_bench_fft(gnss_ctx* c, int reps, double* ms_fwd, double* ms_inv) {
  return guard(c, [&] { c->eng->bench_fft(reps, ms_fwd, ms_inv); });
}

// ---------------------------------------------------------------------------------------------
// Ensemble entry points are implemented in ensemble.hip.
// ---------------------------------------------------------------------------------------------

GNSS_API int gnss_autotune(gnss_ctx* c, int reps, char* report, int len) {
  return guard(c, [&] {
    std::string s = c->eng->autotune(reps);
    if (report && len > 0) {
      if ((int)s.size() + 1 > len) throw Error(GNSS_ERR_ARG, "report buffer too small");
      memcpy(report, s.c_str(), s.size() + 1);
    }
  });
}
GNSS_API int gnss_set_tuning(gnss_ctx* c, const char* spec) {
  if (!spec) return GNSS_ERR_ARG;
  return guard(c, [&] { c->eng->set_tuning(spec); });
}

// ---------------------------------------------------------------------------------------------
// Raw field files in the reference's format (iofuncs.cu:329-358 writer, 462-497 reader).
// ---------------------------------------------------------------------------------------------
// u v w, z for the scalar, c for the colloid (iofuncs.cu:262-327), further scalars z1, z2, ...
static std::string field_name(const gnss_ctx* c, int f) {
  static const char* vel[3] = {"u", "v", "w"};
  if (f < 3) return vel[f];
  if (c->eng->is_colloid(f)) return "c";
  return f == 3 ? std::string("z") : "z" + std::to_string(f - 3);
}

static void write_raw(const std::string& path, const double* a, long long n) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) throw Error(GNSS_ERR_ARG, "cannot open " + path + " for writing");
  const double hdr = (double)n * (double)sizeof(double);
  const bool ok = fwrite(&hdr, sizeof(double), 1, f) == 1 && (long long)fwrite(a, sizeof(double), n, f) == n;
  fclose(f);
  if (!ok) throw Error(GNSS_ERR_ARG, "short write to " + path);
}

static void read_raw(const std::string& path, double* a, long long n) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) throw Error(GNSS_ERR_ARG, "cannot open " + path);
  double hdr = 0;
  const bool got = fread(&hdr, sizeof(double), 1, f) == 1;
  if (!got || hdr != (double)n * (double)sizeof(double)) {
    fclose(f);
    throw Error(GNSS_ERR_ARG, path + ": size header does not match nx*ny*nz doubles");
  }
  const bool ok = (long long)fread(a, sizeof(double), n, f) == n;
  fclose(f);
  if (!ok) throw Error(GNSS_ERR_ARG, "short read from " + path);
}

GNSS_API int gnss_save_fields(gnss_ctx* c, const char* dir, int iter) {
  if (!dir) return GNSS_ERR_ARG;
  return guard(c, [&] {
    gnss_sizes sz;
    c->eng->sizes(&sz);
    std::vector<double> buf((size_t)sz.n_physical);
    for (int f = 0; f < sz.n_fields; ++f) {
      c->eng->inverse(f);
      c->eng->get_field(f, GNSS_PHYSICAL, buf.data());
      write_raw(std::string(dir) + "/" + field_name(c, f) + "." + std::to_string(iter), buf.data(), sz.n_physical);
    }
    double t;
    int64_t st;
    c->eng->time(&t, &st);
    const std::string sp = std::string(dir) + "/gnss_state." + std::to_string(iter);
    FILE* fp = fopen(sp.c_str(), "w");
    if (!fp) throw Error(GNSS_ERR_ARG, "cannot open " + sp);
    fprintf(fp, "step %lld t %.17g\n", (long long)st, t);
    fclose(fp);
  });
}

GNSS_API int gnss_load_fields(gnss_ctx* c, const char* dir, int iter) {
  if (!dir) return GNSS_ERR_ARG;
  return guard(c, [&] {
    gnss_sizes sz;
    c->eng->sizes(&sz);
    std::vector<double> buf((size_t)sz.n_physical);
    for (int f = 0; f < sz.n_fields; ++f) {
      read_raw(std::string(dir) + "/" + field_name(c, f) + "." + std::to_string(iter), buf.data(), sz.n_physical);
      c->eng->set_field(f, GNSS_PHYSICAL, buf.data());
      c->eng->forward(f);
    }
    c->eng->dealias();
    long long st = iter;
    double t = (double)iter * c->p.dt;
    const std::string sp = std::string(dir) + "/gnss_state." + std::to_string(iter);
    if (FILE* fp = fopen(sp.c_str(), "r")) {
      long long a;
      double b;
      if (fscanf(fp, "step %lld t %lf", &a, &b) == 2) {
        st = a;
        t = b;
      }
      fclose(fp);
    }
    c->eng->set_clock(st, t);
  });
}

// ---------------------------------------------------------------------------------------------
// Diagnostics: plane slices (iofuncs.cu:187-327) and iso-surface areas (statistics.cu:1310-1492).
// ---------------------------------------------------------------------------------------------
GNSS_API int gnss_get_slice(gnss_ctx* c, int field, int axis, int index, double* out) {
  if (!out) return GNSS_ERR_ARG;
  return guard(c, [&] { c->eng->get_slice(field, axis, index, out); });
}

GNSS_API int gnss_save_slices(gnss_ctx* c, const char* dir, int iter) {
  if (!dir) return GNSS_ERR_ARG;
  return guard(c, [&] {
    gnss_sizes sz;
    c->eng->sizes(&sz);
    const long long np = sz.n_physical / (2 * (sz.nzh - 1));
    std::vector<double> buf((size_t)np);
    for (int f = 0; f < sz.n_fields; ++f) {
      c->eng->get_slice(f, 2, (int)(sz.nzh - 1), buf.data());
      write_raw(std::string(dir) + "/" + field_name(c, f) + "_xy." + std::to_string(iter), buf.data(), np);
    }
  });
}

GNSS_API int gnss_surface_area(gnss_ctx* c, int field, double iso, double* area) {
  if (!area) return GNSS_ERR_ARG;
  return guard(c, [&] { *area = c->eng->surface_area(field, iso); });
}
