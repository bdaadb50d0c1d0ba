// ensemble.hip — batched ensembles of independent realisations of uneven sizes.
//
// The reference runs one realisation per process (main.cu:84-252).  Here the members of one grid
// (same sizes, scalars, precision, flags, box and dt; each with its own Re, Sc, theta and seed)
// share one batched engine: their state arrays lie back to back and every pass of the step is one
// launch over all of them, so 32 members of 64^3 cost the launches of one.  One RK3 step of the
// whole ensemble is captured once into a hipGraph whose branches are the engines' kernel sequences
// (fork/join with events on a capture stream), and each replay costs one graph launch.  Every
// engine uses the tuned pass variants of its grid (tuning table loaded at creation).
// GNSS_ENSEMBLE_BATCH=0 gives every member an engine of its own.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "engine.h"

#define GNSS_API extern "C" __attribute__((visibility("default")))

using gnss::EngineBase;
using gnss::Error;

struct gnss_ens {
  std::vector<EngineBase*> m;        // engines (one per group of members)
  std::vector<int> eng_of, mem_of;   // member i: its engine and its index within that engine
  std::vector<gnss_params> p;        // member parameters
  std::vector<hipEvent_t> done;      // per engine
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
      const std::string bad = gnss::params_problem(p);
      if (!bad.empty()) throw Error(GNSS_ERR_ARG, "member " + std::to_string(i) + ": " + bad);
      // the unfused step allocates its physical buffers on first use, which a captured step
      // cannot do
      if (p.flags & GNSS_FLAG_UNFUSED)
        throw Error(GNSS_ERR_UNSUPPORTED, "member " + std::to_string(i) +
                                              ": GNSS_FLAG_UNFUSED is not available in an ensemble");
      e->p.push_back(p);
    }
    // groups of members that one batched engine can hold, in order of first appearance
    const char* env = getenv("GNSS_ENSEMBLE_BATCH");
    const bool batch = !(env && env[0] == '0');
    const unsigned solo = GNSS_FLAG_COLLOID | GNSS_FLAG_MFMA_R16;
    auto same_grid = [&](const gnss_params& a, const gnss_params& b) {
      return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.n_scalars == b.n_scalars &&
             a.precision == b.precision && a.flags == b.flags && a.lx == b.lx && a.ly == b.ly &&
             a.lz == b.lz && a.dt == b.dt && !(a.flags & solo);
    };
    // Only grids up to 256^3 are batched (larger members fill the device by themselves), and a
    // batched engine holds fewer than 2^31 elements per field (the late passes' 32-bit offsets;
    // the row padding of a field is at most 5 tiles of 16): larger groups are split.
    auto max_members = [](const gnss_params& a) -> size_t {
      if ((double)a.nx * a.ny * a.nz > 256.0 * 256.0 * 256.0) return 1;
      const double field = (double)a.nx * a.ny * (a.nz / 2 + 1 + 6 * 16);
      return (size_t)std::max(1.0, std::floor(2147483647.0 / field));
    };
    std::vector<std::vector<int>> groups;
    for (int i = 0; i < n; ++i) {
      size_t g = groups.size();
      if (batch)
        for (size_t k = 0; k < groups.size(); ++k)
          if (same_grid(e->p[groups[k][0]], e->p[i]) && groups[k].size() < max_members(e->p[i])) g = k;
      if (g == groups.size()) groups.emplace_back();
      groups[g].push_back(i);
    }
    e->eng_of.assign(n, 0);
    e->mem_of.assign(n, 0);
    for (const std::vector<int>& grp : groups) {
      std::vector<gnss_params> gp;
      for (int i : grp) gp.push_back(e->p[i]);
      const bool f64 = gp[0].precision == GNSS_FP64;
      EngineBase* x = gp.size() == 1 ? (f64 ? gnss::make_engine_f64(gp[0]) : gnss::make_engine_f32(gp[0]))
                                     : (f64 ? gnss::make_engine_batch_f64(gp.data(), (int)gp.size())
                                            : gnss::make_engine_batch_f32(gp.data(), (int)gp.size()));
      e->m.push_back(x);
      for (size_t k = 0; k < grp.size(); ++k) {
        e->eng_of[grp[k]] = (int)e->m.size() - 1;
        e->mem_of[grp[k]] = (int)k;
      }
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

// Captures one RK3 step of every engine into the ensemble's graph (replacing any earlier one).
static void ens_capture(gnss_ens* e) {
  e->ready = false;
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
}

// Wavenumbers + jet initial condition of every member (seeds, theta, Re from each member's
// parameters), then capture of the one-step graph.
GNSS_API int gnss_ensemble_init(gnss_ens* e) {
  return eguard(e, [&] {
    for (EngineBase* x : e->m) {
      x->init_wavenumbers();
      x->init_jet();
    }
    ens_capture(e);
  });
}

// Tunes every batched engine of the ensemble on its own buffers (the "<key>b" entries of the
// tuning table), keeps the choices and captures the step graph again with them.  The report is a
// JSON list with one object per batched engine.  Needs gnss_ensemble_init first; the members'
// states are restored after the timed substages.
GNSS_API int gnss_ensemble_autotune(gnss_ens* e, int reps, char* report, int len) {
  return eguard(e, [&] {
    if (!e->ready) throw Error(GNSS_ERR_STATE, "call gnss_ensemble_init first");
    std::string s = "[";
    for (EngineBase* x : e->m) {
      if (x->members() < 2) continue;
      s += (s.size() > 1 ? ", " : "") + x->autotune(reps);
    }
    s += "]";
    ens_capture(e);
    if (report && len > 0) {
      if ((int)s.size() + 1 > len) throw Error(GNSS_ERR_ARG, "report buffer too small");
      memcpy(report, s.c_str(), s.size() + 1);
    }
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
    if (i < 0 || i >= (int)e->p.size()) throw Error(GNSS_ERR_ARG, "member out of range");
    EngineBase* x = e->m[e->eng_of[i]];
    x->select_member(e->mem_of[i]);
    x->stats(o, yprof, spectrum);
  });
}

GNSS_API int gnss_ensemble_get_field(gnss_ens* e, int i, int f, int space, double* host) {
  return eguard(e, [&] {
    if (i < 0 || i >= (int)e->p.size() || !host) throw Error(GNSS_ERR_ARG, "member out of range");
    EngineBase* x = e->m[e->eng_of[i]];
    x->select_member(e->mem_of[i]);
    // the physical buffers of a batched engine serve one member at a time: transform this
    // member's state into them first
    if (space == GNSS_PHYSICAL) x->inverse(f);
    x->get_field(f, space, host);
  });
}
