// gnss_main.cpp — host driver with the reference's call structure (main.cu:84-252):
//   parameters -> allocate (gnss_create) -> wavenumbers -> jet IC + forward + dealias (or a
//   restart from raw field files) -> stats(0), slices(0) -> loop { step; stats every n_stats;
//   2D slices every n_vis; checkpoint every n_checkpoint } -> free.
// The reference hard-codes its configuration in dnsparams.h; here it comes from the command line.
// The driver only calls include/gnss.h, so it runs on any library implementing that ABI (built
// against the MI355X engine here).
//
// usage: gnss_main [--n N | --nx X --ny Y --nz Z] [--re R] [--sc S ...] [--theta T] [--dt DT]
//                  [--steps K] [--stats-every M] [--kmax K] [--amp A] [--seed S] [--fp32]
//                  [--l L] [--ly LY] [--config jet64|jet256|jet512|jet1024]
//                  [--colloid ALPHA [--sc-c SCC] [--theta-c TC]] [--area ISO]
//                  [--out DIR] [--vis-every V] [--checkpoint-every C] [--restart ITER]
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/gnss.h"

static void check(int rc, gnss_ctx* c, const char* what) {
  if (rc != GNSS_OK) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, gnss_last_error(c));
    exit(1);
  }
}

static void print_stats(gnss_ctx* ctx, const gnss_stats_out& s, int S, double ms, double iso) {
  // full precision, so that runs on different implementations of the ABI can be compared
  printf("step %6lld  t %9.5f  KE %.16e  eps %.16e  chi0 %.16e  CFL %.16e  div %.2e  (%.3f ms/step)",
         (long long)s.step, s.t, s.ke, s.epsilon, S > 0 ? s.chi[0] : 0.0, s.cfl, s.divergence, ms);
  if (S > 0 && !std::isnan(iso)) {
    // scalar interface area (the reference's area_scalar, statistics.cu:2132)
    double a = 0;
    check(gnss_surface_area(ctx, 3, iso, &a), ctx, "gnss_surface_area");
    printf("  area(z=%g) %.6e", iso, a);
  }
  printf("\n");
}

int main(int argc, char** argv) {
  gnss_params p;
  memset(&p, 0, sizeof p);
  p.nx = p.ny = p.nz = 64;
  p.lx = p.ly = p.lz = 2.0 * M_PI;
  p.re = 1000.0;
  p.n_scalars = 1;
  p.sc[0] = 0.7;
  p.theta = 0.05;
  p.jet_h = 1.0;
  p.noise_amp = 0.05;
  p.noise_kmax = 8.0;
  p.seed = 20261202ULL;
  p.dt = 1e-3;
  p.precision = GNSS_FP64;
  int steps = 100, stats_every = 20, vis_every = 0, ckpt_every = 0, restart = -1;
  double colloid_alpha = NAN, sc_c = 0.7, iso = NAN;
  std::string out_dir = ".";
  std::vector<double> sc;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&](void) -> const char* {
      if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(2); }
      return argv[++i];
    };
    if (a == "--n") p.nx = p.ny = p.nz = atoi(val());
    else if (a == "--nx") p.nx = atoi(val());
    else if (a == "--ny") p.ny = atoi(val());
    else if (a == "--nz") p.nz = atoi(val());
    else if (a == "--l") p.lx = p.ly = p.lz = atof(val());
    else if (a == "--ly") p.ly = atof(val());
    else if (a == "--re") p.re = atof(val());
    else if (a == "--sc") sc.push_back(atof(val()));
    else if (a == "--theta") p.theta = atof(val());
    else if (a == "--dt") p.dt = atof(val());
    else if (a == "--steps") steps = atoi(val());
    else if (a == "--stats-every") stats_every = atoi(val());
    else if (a == "--kmax") p.noise_kmax = atof(val());
    else if (a == "--amp") p.noise_amp = atof(val());
    else if (a == "--seed") p.seed = strtoull(val(), nullptr, 10);
    else if (a == "--fp32") p.precision = GNSS_FP32;
    else if (a == "--colloid") colloid_alpha = atof(val());
    else if (a == "--sc-c") sc_c = atof(val());
    else if (a == "--theta-c") p.theta_c = atof(val());
    else if (a == "--area") iso = atof(val());
    else if (a == "--out") out_dir = val();
    else if (a == "--vis-every") vis_every = atoi(val());
    else if (a == "--checkpoint-every") ckpt_every = atoi(val());
    else if (a == "--restart") restart = atoi(val());
    else if (a == "--config") {
      std::string c = val();
      if (c == "jet64") { p.nx = p.ny = p.nz = 64; p.theta = 0.05; p.re = 1000; sc = {0.7}; p.noise_kmax = 8; p.dt = 1e-3; steps = 100; }
      else if (c == "jet256") { p.nx = p.ny = p.nz = 256; p.theta = 0.03; p.re = 3000; sc = {0.7, 2.0}; p.noise_kmax = 16; p.dt = 5e-4; steps = 200; }
      else if (c == "jet512") { p.nx = p.ny = p.nz = 512; p.theta = 0.02; p.re = 5000; sc = {1.0}; p.noise_kmax = 32; p.dt = 2.5e-4; steps = 100; p.precision = GNSS_FP32; }
      else if (c == "jet1024") { p.nx = p.ny = p.nz = 1024; p.theta = 0.01; p.re = 10000; sc = {1.0}; p.noise_kmax = 64; p.dt = 1e-4; steps = 50; p.precision = GNSS_FP32; }
      else { fprintf(stderr, "unknown config %s\n", c.c_str()); return 2; }
    } else {
      fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  if (!sc.empty()) {
    p.n_scalars = (int)sc.size();
    for (size_t i = 0; i < sc.size() && i < GNSS_MAX_SCALARS; ++i) p.sc[i] = sc[i];
  }
  if (!std::isnan(colloid_alpha)) {  // the colloid becomes the last scalar (dnsparams.h:28-33)
    if (p.n_scalars >= GNSS_MAX_SCALARS) { fprintf(stderr, "no room for the colloid scalar\n"); return 2; }
    p.sc[p.n_scalars++] = sc_c;
    p.flags |= GNSS_FLAG_COLLOID;
    p.alpha = colloid_alpha;
  }
  printf("gnss: grid %dx%dx%d  Re %g  scalars %d  dt %g  steps %d  %s\n", p.nx, p.ny, p.nz, p.re, p.n_scalars,
         p.dt, steps, p.precision == GNSS_FP64 ? "FP64" : "FP32");

  gnss_ctx* ctx = nullptr;
  check(gnss_create(&p, &ctx), nullptr, "gnss_create");           // allocate_memory + plans
  check(gnss_init_wavenumbers(ctx), ctx, "gnss_init_wavenumbers"); // initializeWaveNumbers
  int c0 = 0;
  if (restart >= 0) {
    check(gnss_load_fields(ctx, out_dir.c_str(), restart), ctx, "gnss_load_fields");  // loadData
    c0 = restart;
  } else {
    check(gnss_init_jet(ctx), ctx, "gnss_init_jet");  // initializeJet + forward + deAlias
  }
  gnss_stats_out st;
  check(gnss_stats(ctx, &st, nullptr, nullptr), ctx, "gnss_stats");  // calcTurbStats_mgpu(0)
  print_stats(ctx, st, p.n_scalars, 0.0, iso);
  if (vis_every > 0) check(gnss_save_slices(ctx, out_dir.c_str(), c0), ctx, "gnss_save_slices");  // save2Dfields(0)
  double total_ms = 0;
  for (int c = c0 + 1; c <= c0 + steps; ++c) {
    auto t0 = std::chrono::steady_clock::now();
    check(gnss_step(ctx, 1), ctx, "gnss_step");  // solver_ps
    auto t1 = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    total_ms += ms;
    if (stats_every > 0 && c % stats_every == 0) {
      check(gnss_stats(ctx, &st, nullptr, nullptr), ctx, "gnss_stats");
      print_stats(ctx, st, p.n_scalars, ms, iso);
    }
    if (vis_every > 0 && c % vis_every == 0)
      check(gnss_save_slices(ctx, out_dir.c_str(), c), ctx, "gnss_save_slices");  // save2Dfields
    if (ckpt_every > 0 && c % ckpt_every == 0)
      check(gnss_save_fields(ctx, out_dir.c_str(), c), ctx, "gnss_save_fields");  // save3Dfields
  }
  printf("total step time %.3f s, %.3f steps/s, %.4e grid-point updates/s\n", total_ms / 1000.0,
         steps / (total_ms / 1000.0), (double)p.nx * p.ny * p.nz * steps / (total_ms / 1000.0));
  gnss_destroy(ctx);  // deallocate_memory
  return 0;
}
