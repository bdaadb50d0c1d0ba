// ensemble.hip — batched ensembles of independent realisations of uneven sizes.
//
// The reference runs one realisation per process (main.cu:84-252).  Here an ensemble owns one
// engine per member; one RK3 step of the whole ensemble is captured once into a hipGraph whose
// branches are the members' kernel sequences (fork/join with events on a capture stream), so the
// small members run concurrently and fill the device, and each replay costs one graph launch.
// Every member uses the tuned pass variants of its size (tuning table loaded at creation).
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "engine.h"

#define GNSS_API extern "C" __attribute__((visibility("default")))

using gnss::EngineBase;
using gnss::Error;

struct gnss_ens {
  std::vector<EngineBase*> m;
  std::vector<gnss_params> p;
  std::vector<hipEvent_t> done;
  hipStream_t cap = nullptr;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ready = false;
  std::string err;
};

static thread_local std::string g_ens_err;

static void ens_free(gnss_ens* e) {
  if (!e) return;
  if (e->cap) (void)hipStreamSynchronize(e->cap);
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  for (hipEvent_t ev : e->done) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->cap) (void)hipStreamDestroy(e->cap);
  for (EngineBase* x : e->m) delete x;
  delete e;
}

template <typename Fn> static int eguard(gnss_ens* e, Fn&& fn) {
  if (!e) return GNSS_ERR_ARG;
  try {
    fn();
    return GNSS_OK;
  } catch (const Error& x) {
    e->err = x.what();
    return x.code;
  } catch (const std::exception& x) {
    e->err = x.what();
    return GNSS_ERR_DEVICE;
  }
}

GNSS_API int gnss_ensemble_create(int n, const gnss_params* members, gnss_ens** out) {
  if (n <= 0 || !members || !out) {
    g_ens_err = "need n > 0 members";
    return GNSS_ERR_ARG;
  }
  *out = nullptr;
  gnss_ens* e = new gnss_ens();
  try {
    for (int i = 0; i < n; ++i) {
      const gnss_params& p = members[i];
      if (!gnss::supported_len(p.nx) || !gnss::supported_len(p.ny) || !gnss::supported_len(p.nz))
        throw Error(GNSS_ERR_ARG, "member " + std::to_string(i) + ": unsupported grid");
      if (!(p.re > 0 && p.dt > 0 && p.lx > 0 && p.ly > 0 && p.lz > 0) || p.n_scalars < 0 ||
          p.n_scalars > GNSS_MAX_SCALARS)
        throw Error(GNSS_ERR_ARG, "member " + std::to_string(i) + ": bad parameters");
      e->p.push_back(p);
      e->m.push_back(p.precision == GNSS_FP64 ? gnss::make_engine_f64(p) : gnss::make_engine_f32(p));
      hipEvent_t ev;
      GNSS_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      e->done.push_back(ev);
    }
    GNSS_HIP(hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking));
    GNSS_HIP(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
  } catch (const Error& x) {
    g_ens_err = x.what();
    ens_free(e);
    return x.code;
  } catch (const std::exception& x) {
    g_ens_err = x.what();
    ens_free(e);
    return GNSS_ERR_DEVICE;
  }
  *out = e;
  return GNSS_OK;
}

GNSS_API int gnss_ensemble_destroy(gnss_ens* e) {
  ens_free(e);
  return GNSS_OK;
}

GNSS_API const char* gnss_ensemble_last_error(const gnss_ens* e) { return e ? e->err.c_str() : g_ens_err.c_str(); }

// Wavenumbers + jet initial condition of every member (seeds, theta, Re from each member's
// parameters), then capture of the one-step graph.
GNSS_API int gnss_ensemble_init(gnss_ens* e) {
  return eguard(e, [&] {
    for (EngineBase* x : e->m) {
      x->init_wavenumbers();
      x->init_jet();
    }
    if (e->exec) {
      GNSS_HIP(hipGraphExecDestroy(e->exec));
      e->exec = nullptr;
    }
    if (e->graph) {
      GNSS_HIP(hipGraphDestroy(e->graph));
      e->graph = nullptr;
    }
    GNSS_HIP(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal));
    try {
      GNSS_HIP(hipEventRecord(e->fork, e->cap));
      for (size_t i = 0; i < e->m.size(); ++i) {
        hipStream_t s = e->m[i]->stream();
        GNSS_HIP(hipStreamWaitEvent(s, e->fork, 0));
        e->m[i]->enqueue_rk3_step();
        GNSS_HIP(hipEventRecord(e->done[i], s));
        GNSS_HIP(hipStreamWaitEvent(e->cap, e->done[i], 0));
      }
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(e->cap, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    GNSS_HIP(hipStreamEndCapture(e->cap, &e->graph));
    GNSS_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    e->ready = true;
  });
}

GNSS_API int gnss_ensemble_step(gnss_ens* e, int nsteps) {
  return eguard(e, [&] {
    if (!e->ready) throw Error(GNSS_ERR_STATE, "call gnss_ensemble_init first");
    if (nsteps < 0) throw Error(GNSS_ERR_ARG, "nsteps must be >= 0");
    for (int k = 0; k < nsteps; ++k) GNSS_HIP(hipGraphLaunch(e->exec, e->cap));
    GNSS_HIP(hipStreamSynchronize(e->cap));
    for (EngineBase* x : e->m) {
      x->advance_clock(nsteps);
      x->check_finite();
    }
  });
}

GNSS_API int gnss_ensemble_stats(gnss_ens* e, int i, gnss_stats_out* o, double* yprof, double* spectrum) {
  return eguard(e, [&] {
    if (i < 0 || i >= (int)e->m.size()) throw Error(GNSS_ERR_ARG, "member out of range");
    e->m[i]->stats(o, yprof, spectrum);
  });
}

GNSS_API int gnss_ensemble_get_field(gnss_ens* e, int i, int f, int space, double* host) {
  return eguard(e, [&] {
    if (i < 0 || i >= (int)e->m.size() || !host) throw Error(GNSS_ERR_ARG, "member out of range");
    e->m[i]->get_field(f, space, host);
  });
}
