/*
 * gnss_oracle.c — FP64 CPU restatement of the pseudospectral NS + passive-scalar RK3 step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X engine: the parity
 * tests and the CPU-baseline leg of bench.py load it; the engine never does.  It implements the
 * same C ABI (include/gnss.h) so one test script can drive either library.
 *
 * What it restates (reference file:line it follows):
 *   layout / normalisation   unnormalised R2C forward, 1/(nx ny nz) inverse
 *                            (fftfuncs.cu:288-359, cudafuncs.cu:19-31)
 *   wavenumbers              k = (2 pi/L) m, m = i < n/2 ? i : i-n  (initialize.cu:678-691)
 *   dealias mask             spherical truncation (solver.cu:17-42) with the 2/3 factor
 *                            (dnsparams.h:65): keep iff sum_i (m_i/n_i)^2 <= 1/9, in integers
 *   vorticity                w^ = i k x u^  (solver.cu:68-114)
 *   cross product            u x w in physical space  (solver.cu:139-164, sign fixed by
 *                            computeRHSKernel_mgpu, solver.cu:504-511)
 *   scalar advection         u . grad(phi), convective form  (solver.cu:233-256, 342-356, 407-432)
 *   colloid (optional)       last scalar c with drift alpha grad Z (Z = scalar 0):
 *                            u . grad c + alpha (lap Z c + grad Z . grad c)
 *                            (solver.cu:276-293, 374-459), own profile thickness theta_c
 *                            (initialize.cu:178-194)
 *   projection               rhs_i = (delta_ij - k_i k_j/k^2) (u x w)^_j, k = 0 zeroed;
 *                            scalar rhs = -(u.grad phi)^  (solver.cu:461-521)
 *   time integration         DEPARTURE: low-storage RK3 (Williamson 1980, A = 0,-5/9,-153/128,
 *                            B = 1/3,15/16,8/15) with exact integrating factors exp(-nu k^2 dt)
 *                            replaces Euler/AB2 + Crank-Nicolson (solver.cu:547-616)
 *   jet initial condition    U(y) = 0.5 - 0.5 tanh(H/(4 theta)(2|y|/H - 1)), same profile for the
 *                            scalar (initialize.cu:114-176); the noise is the project's seeded
 *                            generator (the reference's glibc-rand spectrum, initialize.cu:286-612,
 *                            is not reproducible and not a target)
 *   statistics               KE, Vrms, epsilon, integral length, chi, lambda, eta
 *                            (statistics.cu:1494-1845, 2147-2149); x-z plane averages
 *                            (statistics.cu:1936-2003) plus Reynolds stresses and the shell
 *                            spectrum the reference declares but never computes
 *                            (struct_def.h:84-87, statistics.cu:1847-1934)
 *
 * How each part is pinned (tests/test_oracle.py): brute-force DFT and numpy.fft golden vectors
 * (FFT); closed forms for wavenumbers/mask (enumeration), single-mode exact viscous decay (IF),
 * 2D Taylor-Green exact decay (whole step), projection invariants, third-order convergence
 * (RK3), inviscid energy conservation, single-mode statistics, Hermitian/divergence/amplitude
 * properties of the initial condition; the colloid term by the exact steady state
 * c = exp(alpha Z / D_c) of a frozen carrier, its uniform-flow translation, and the plain-scalar
 * limit alpha = 0.
 *
 * FFT backends: a self-contained Stockham mixed-radix FFT (default), or the FFTW3 API resolved
 * at run time from libfftw3 / libmkl_rt (used for the threaded CPU baseline).
 */
#define _GNU_SOURCE
#include <complex.h>
#include <dlfcn.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/gnss.h"

#define EXPORT __attribute__((visibility("default")))

typedef double complex cplx;

static char g_create_err[256];

/* ------------------------------------------------------------------------------------------ */
/* 1D complex FFT: Stockham autosort, radices 4/2/3 plus a table-driven generic radix.        */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
  int n, nst;
  int rad[48];
  cplx* tw; /* tw[m] = exp(-2 pi i m / n), computed in long double and rounded */
} fftplan;

static void twiddle_table(cplx* tw, int n) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; ++m) {
    long double a = 2.0L * pi * (long double)m / (long double)n;
    tw[m] = (double)cosl(a) - I * (double)sinl(a);
  }
}

static int plan_init(fftplan* p, int n) {
  memset(p, 0, sizeof(*p));
  p->n = n;
  int r = n, k = 0;
  while (r % 4 == 0) { p->rad[k++] = 4; r /= 4; }
  while (r % 2 == 0) { p->rad[k++] = 2; r /= 2; }
  while (r % 3 == 0) { p->rad[k++] = 3; r /= 3; }
  for (int f = 5; r > 1; f += 2)
    while (r % f == 0) { p->rad[k++] = f; r /= f; }
  p->nst = k;
  p->tw = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  if (!p->tw) return -1;
  twiddle_table(p->tw, n);
  return 0;
}

static void plan_free(fftplan* p) { free(p->tw); p->tw = NULL; }

static inline cplx tw_get(const fftplan* p, int m, int sign) {
  cplx w = p->tw[m];
  return sign < 0 ? w : conj(w);
}

/* out[q] = sum_r in[r] exp(sign 2 pi i r q / R) */
static void dft_small(const fftplan* p, int R, const cplx* in, cplx* out, int sign) {
  if (R == 2) {
    out[0] = in[0] + in[1];
    out[1] = in[0] - in[1];
  } else if (R == 4) {
    cplx t0 = in[0] + in[2], t1 = in[0] - in[2];
    cplx t2 = in[1] + in[3], t3 = in[1] - in[3];
    /* multiply t3 by sign*i */
    cplx t3r = sign < 0 ? (cimag(t3) - I * creal(t3)) : (-cimag(t3) + I * creal(t3));
    out[0] = t0 + t2;
    out[2] = t0 - t2;
    out[1] = t1 + t3r;
    out[3] = t1 - t3r;
  } else if (R == 3) {
    const double c = -0.5, s = 0.86602540378443864676372317075294;
    cplx a = in[1] + in[2], b = in[1] - in[2];
    cplx m = in[0] + c * a;
    cplx jb = (double)sign * s * (I * b);
    out[0] = in[0] + a;
    out[1] = m + jb;
    out[2] = m - jb;
  } else {
    int stride = p->n / R;
    for (int q = 0; q < R; ++q) {
      cplx acc = 0;
      for (int r = 0; r < R; ++r) acc += in[r] * tw_get(p, ((r * q) % R) * stride, sign);
      out[q] = acc;
    }
  }
}

/* In-place transform of x (length n, unit stride) using scratch y (length n). */
static void fft_exec(const fftplan* p, cplx* x, cplx* y, int sign) {
  const int n = p->n;
  cplx *a = x, *b = y;
  int Ns = 1;
  cplx in[64], out[64];
  for (int s = 0; s < p->nst; ++s) {
    const int R = p->rad[s];
    const int m = n / R;
    const int tstride = n / (Ns * R);
    for (int j = 0; j < m; ++j) {
      const int k = j % Ns;
      for (int r = 0; r < R; ++r) {
        cplx v = a[j + r * m];
        if (r && k) v *= tw_get(p, r * k * tstride, sign);
        in[r] = v;
      }
      dft_small(p, R, in, out, sign);
      const int base = (j / Ns) * Ns * R + k;
      for (int q = 0; q < R; ++q) b[base + q * Ns] = out[q];
    }
    cplx* t = a; a = b; b = t;
    Ns *= R;
  }
  if (a != x) memcpy(x, a, sizeof(cplx) * (size_t)n);
}

/* ------------------------------------------------------------------------------------------ */
/* FFTW3 API resolved at run time (libfftw3 or MKL's FFTW3 interface).                        */
/* ------------------------------------------------------------------------------------------ */

typedef void* fftw_plan_t;
typedef struct {
  void* lib;
  const char* name;
  int (*init_threads)(void);
  void (*plan_with_nthreads)(int);
  fftw_plan_t (*plan_r2c_3d)(int, int, int, double*, cplx*, unsigned);
  fftw_plan_t (*plan_c2r_3d)(int, int, int, cplx*, double*, unsigned);
  void (*execute_dft_r2c)(fftw_plan_t, double*, cplx*);
  void (*execute_dft_c2r)(fftw_plan_t, cplx*, double*);
  void (*destroy_plan)(fftw_plan_t);
} fftw_api;

static fftw_api g_fftw;
static int g_fftw_tried, g_fftw_ok;
#define FFTW_ESTIMATE_FLAG (1U << 6)

static int fftw_load(void) {
  if (g_fftw_tried) return g_fftw_ok;
  g_fftw_tried = 1;
  const char* names[] = {"libfftw3.so.3", "libfftw3.so", "libmkl_rt.so", "libmkl_rt.so.1",
                         "/opt/conda/lib/libmkl_rt.so", NULL};
  for (int i = 0; names[i]; ++i) {
    void* h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    fftw_api a;
    memset(&a, 0, sizeof(a));
    a.lib = h;
    a.name = names[i];
    a.init_threads = (int (*)(void))dlsym(h, "fftw_init_threads");
    a.plan_with_nthreads = (void (*)(int))dlsym(h, "fftw_plan_with_nthreads");
    a.plan_r2c_3d = (fftw_plan_t(*)(int, int, int, double*, cplx*, unsigned))dlsym(h, "fftw_plan_dft_r2c_3d");
    a.plan_c2r_3d = (fftw_plan_t(*)(int, int, int, cplx*, double*, unsigned))dlsym(h, "fftw_plan_dft_c2r_3d");
    a.execute_dft_r2c = (void (*)(fftw_plan_t, double*, cplx*))dlsym(h, "fftw_execute_dft_r2c");
    a.execute_dft_c2r = (void (*)(fftw_plan_t, cplx*, double*))dlsym(h, "fftw_execute_dft_c2r");
    a.destroy_plan = (void (*)(fftw_plan_t))dlsym(h, "fftw_destroy_plan");
    if (a.plan_r2c_3d && a.plan_c2r_3d && a.execute_dft_r2c && a.execute_dft_c2r && a.destroy_plan) {
      g_fftw = a;
      g_fftw_ok = 1;
      return 1;
    }
    dlclose(h);
  }
  return 0;
}

static int g_backend = 0;   /* 0 builtin, 1 fftw api */
static int g_threads = 0;   /* 0: OpenMP default */

EXPORT int gnss_oracle_set_backend(int backend) {
  if (backend == 1 && !fftw_load()) return GNSS_ERR_UNSUPPORTED;
  g_backend = backend ? 1 : 0;
  return GNSS_OK;
}
EXPORT const char* gnss_oracle_backend_name(void) {
  return g_backend ? g_fftw.name : "builtin-stockham";
}
EXPORT int gnss_oracle_set_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
  return GNSS_OK;
}
EXPORT int gnss_oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* Context                                                                                    */
/* ------------------------------------------------------------------------------------------ */

struct gnss_ctx {
  gnss_params p;
  int nx, ny, nz, nzh, S, F, W;
  int colloid;      /* 1: scalar S - 1 is a colloid carried by u + alpha grad(scalar 0) */
  size_t nsp, nph;
  double nu, diff[GNSS_MAX_SCALARS];
  fftplan px, py, pz;
  double *kx, *ky, *kz;
  int *mx, *my, *mz;
  uint8_t* mask;    /* internal [x][y][kz] */
  cplx** q;         /* F spectral state, internal [kx][ky][kz] */
  cplx** dq;        /* F RK register */
  double** phys;    /* F physical buffers [x][y][z] */
  cplx** wsp;       /* work spectral (W = 6 + 3S, + 2 with a colloid) */
  double** wph;     /* work physical (W) */
  cplx** nl;        /* 3 + S nonlinear terms (spectral) */
  double** nlph;    /* 3 + S nonlinear terms (physical) */
  cplx* tmp;        /* one spectral scratch */
  fftw_plan_t fw, bw;
  double t;
  int64_t step;
  int have_k, have_ic;
  char err[256];
};

static int set_err(gnss_ctx* c, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(c ? c->err : g_create_err, 256, fmt, ap);
  va_end(ap);
  return code;
}

static int good_len(int n) {
  if (n < 4 || n > 4096 || (n & 1)) return 0;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

static inline int mode_of(int i, int n) { return i < n / 2 ? i : i - n; }

static int64_t gcd64(int64_t a, int64_t b) {
  while (b) { int64_t t = a % b; a = b; b = t; }
  return a;
}

/* ------------------------------------------------------------------------------------------ */
/* 3D transforms, internal layout: physical [x][y][z], spectral [kx][ky][kz<nzh]               */
/* ------------------------------------------------------------------------------------------ */

static void r2c3d_builtin(gnss_ctx* c, const double* in, cplx* out) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzh = c->nzh;
  /* z rows */
#pragma omp parallel
  {
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    cplx* a = (cplx*)malloc(sizeof(cplx) * nmax);
    cplx* b = (cplx*)malloc(sizeof(cplx) * nmax);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)nx * ny; ++r) {
      const double* src = in + (size_t)r * nz;
      for (int z = 0; z < nz; ++z) a[z] = src[z];
      fft_exec(&c->pz, a, b, -1);
      cplx* dst = out + (size_t)r * nzh;
      for (int k = 0; k < nzh; ++k) dst[k] = a[k];
    }
    /* y columns */
#pragma omp for schedule(static)
    for (long r = 0; r < (long)nx * nzh; ++r) {
      const int x = (int)(r / nzh), k = (int)(r % nzh);
      cplx* base = out + (size_t)x * ny * nzh + k;
      for (int y = 0; y < ny; ++y) a[y] = base[(size_t)y * nzh];
      fft_exec(&c->py, a, b, -1);
      for (int y = 0; y < ny; ++y) base[(size_t)y * nzh] = a[y];
    }
    /* x columns */
#pragma omp for schedule(static)
    for (long r = 0; r < (long)ny * nzh; ++r) {
      cplx* base = out + r;
      for (int x = 0; x < nx; ++x) a[x] = base[(size_t)x * ny * nzh];
      fft_exec(&c->px, a, b, -1);
      for (int x = 0; x < nx; ++x) base[(size_t)x * ny * nzh] = a[x];
    }
    free(a);
    free(b);
  }
}

static void c2r3d_builtin(gnss_ctx* c, const cplx* in, double* out, cplx* tmp) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzh = c->nzh;
  const double scale = 1.0 / ((double)nx * ny * nz);
  memcpy(tmp, in, sizeof(cplx) * c->nsp);
#pragma omp parallel
  {
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    cplx* a = (cplx*)malloc(sizeof(cplx) * nmax);
    cplx* b = (cplx*)malloc(sizeof(cplx) * nmax);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)ny * nzh; ++r) {
      cplx* base = tmp + r;
      for (int x = 0; x < nx; ++x) a[x] = base[(size_t)x * ny * nzh];
      fft_exec(&c->px, a, b, +1);
      for (int x = 0; x < nx; ++x) base[(size_t)x * ny * nzh] = a[x];
    }
#pragma omp for schedule(static)
    for (long r = 0; r < (long)nx * nzh; ++r) {
      const int x = (int)(r / nzh), k = (int)(r % nzh);
      cplx* base = tmp + (size_t)x * ny * nzh + k;
      for (int y = 0; y < ny; ++y) a[y] = base[(size_t)y * nzh];
      fft_exec(&c->py, a, b, +1);
      for (int y = 0; y < ny; ++y) base[(size_t)y * nzh] = a[y];
    }
    /* z rows: Hermitian extension; imaginary parts of kz = 0 and kz = nz/2 are ignored */
#pragma omp for schedule(static)
    for (long r = 0; r < (long)nx * ny; ++r) {
      const cplx* src = tmp + (size_t)r * nzh;
      a[0] = creal(src[0]);
      for (int k = 1; k < nz / 2; ++k) {
        a[k] = src[k];
        a[nz - k] = conj(src[k]);
      }
      a[nz / 2] = creal(src[nz / 2]);
      fft_exec(&c->pz, a, b, +1);
      double* dst = out + (size_t)r * nz;
      for (int z = 0; z < nz; ++z) dst[z] = creal(a[z]) * scale;
    }
    free(a);
    free(b);
  }
}

static void r2c3d(gnss_ctx* c, const double* in, cplx* out) {
  if (g_backend == 1 && c->fw) {
    /* FFTW r2c may not modify the input for out-of-place transforms */
    g_fftw.execute_dft_r2c(c->fw, (double*)in, out);
    return;
  }
  r2c3d_builtin(c, in, out);
}

static void c2r3d(gnss_ctx* c, const cplx* in, double* out) {
  if (g_backend == 1 && c->bw) {
    const double scale = 1.0 / ((double)c->nx * c->ny * c->nz);
    memcpy(c->tmp, in, sizeof(cplx) * c->nsp); /* c2r destroys its input */
    g_fftw.execute_dft_c2r(c->bw, c->tmp, out);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)c->nph; ++i) out[i] *= scale;
    return;
  }
  c2r3d_builtin(c, in, out, c->tmp);
}

/* ------------------------------------------------------------------------------------------ */
/* Creation                                                                                   */
/* ------------------------------------------------------------------------------------------ */

static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

static void ctx_free(gnss_ctx* c) {
  if (!c) return;
  const int W = c->W;
  for (int i = 0; i < c->F; ++i) {
    if (c->q) free(c->q[i]);
    if (c->dq) free(c->dq[i]);
    if (c->phys) free(c->phys[i]);
    if (c->nl) free(c->nl[i]);
    if (c->nlph) free(c->nlph[i]);
  }
  for (int i = 0; i < W; ++i) {
    if (c->wsp) free(c->wsp[i]);
    if (c->wph) free(c->wph[i]);
  }
  free(c->q); free(c->dq); free(c->phys); free(c->nl); free(c->nlph); free(c->wsp); free(c->wph);
  free(c->tmp);
  free(c->kx); free(c->ky); free(c->kz); free(c->mx); free(c->my); free(c->mz); free(c->mask);
  if (c->fw) g_fftw.destroy_plan(c->fw);
  if (c->bw) g_fftw.destroy_plan(c->bw);
  plan_free(&c->px); plan_free(&c->py); plan_free(&c->pz);
  free(c);
}

EXPORT const char* gnss_last_error(const gnss_ctx* c) { return c ? c->err : g_create_err; }

EXPORT int gnss_create(const gnss_params* p, gnss_ctx** out) {
  if (!p || !out) return set_err(NULL, GNSS_ERR_ARG, "null argument");
  *out = NULL;
  if (!good_len(p->nx) || !good_len(p->ny) || !good_len(p->nz))
    return set_err(NULL, GNSS_ERR_ARG, "grid %dx%dx%d: each size must be even, >= 4 and a product of 2s and 3s",
                   p->nx, p->ny, p->nz);
  if (!(p->lx > 0 && p->ly > 0 && p->lz > 0)) return set_err(NULL, GNSS_ERR_ARG, "box lengths must be > 0");
  if (!(p->re > 0)) return set_err(NULL, GNSS_ERR_ARG, "re must be > 0");
  if (!(p->dt > 0)) return set_err(NULL, GNSS_ERR_ARG, "dt must be > 0");
  if (p->n_scalars < 0 || p->n_scalars > GNSS_MAX_SCALARS)
    return set_err(NULL, GNSS_ERR_ARG, "n_scalars must be in [0, %d]", GNSS_MAX_SCALARS);
  for (int s = 0; s < p->n_scalars; ++s)
    if (!(p->sc[s] > 0)) return set_err(NULL, GNSS_ERR_ARG, "sc[%d] must be > 0", s);
  if (p->precision != GNSS_FP64)
    return set_err(NULL, GNSS_ERR_UNSUPPORTED, "the oracle computes in FP64 only");
  if ((p->flags & GNSS_FLAG_COLLOID) && p->n_scalars < 2)
    return set_err(NULL, GNSS_ERR_ARG, "GNSS_FLAG_COLLOID needs n_scalars >= 2 (scalar 0 carries the colloid)");
  if ((p->flags & GNSS_FLAG_COLLOID) && !isfinite(p->alpha))
    return set_err(NULL, GNSS_ERR_ARG, "alpha must be finite");
  gnss_ctx* c = (gnss_ctx*)xcalloc(1, sizeof(gnss_ctx));
  if (!c) return set_err(NULL, GNSS_ERR_ALLOC, "out of host memory");
  c->p = *p;
  c->nx = p->nx; c->ny = p->ny; c->nz = p->nz; c->nzh = p->nz / 2 + 1;
  c->S = p->n_scalars; c->F = 3 + c->S;
  c->colloid = (p->flags & GNSS_FLAG_COLLOID) ? 1 : 0;
  c->W = 6 + 3 * c->S + 2 * c->colloid;
  c->nsp = (size_t)c->nx * c->ny * c->nzh;
  c->nph = (size_t)c->nx * c->ny * c->nz;
  c->nu = 1.0 / p->re;
  for (int s = 0; s < c->S; ++s) c->diff[s] = c->nu / p->sc[s];
  if (plan_init(&c->px, c->nx) || plan_init(&c->py, c->ny) || plan_init(&c->pz, c->nz)) {
    ctx_free(c);
    return set_err(NULL, GNSS_ERR_ALLOC, "out of host memory (plans)");
  }
  const int F = c->F, W = c->W;
  c->q = (cplx**)xcalloc(F, sizeof(cplx*));
  c->dq = (cplx**)xcalloc(F, sizeof(cplx*));
  c->phys = (double**)xcalloc(F, sizeof(double*));
  c->nl = (cplx**)xcalloc(F, sizeof(cplx*));
  c->nlph = (double**)xcalloc(F, sizeof(double*));
  c->wsp = (cplx**)xcalloc(W, sizeof(cplx*));
  c->wph = (double**)xcalloc(W, sizeof(double*));
  int bad = !c->q || !c->dq || !c->phys || !c->nl || !c->nlph || !c->wsp || !c->wph;
  for (int i = 0; !bad && i < F; ++i) {
    c->q[i] = (cplx*)xcalloc(c->nsp, sizeof(cplx));
    c->dq[i] = (cplx*)xcalloc(c->nsp, sizeof(cplx));
    c->phys[i] = (double*)xcalloc(c->nph, sizeof(double));
    c->nl[i] = (cplx*)xcalloc(c->nsp, sizeof(cplx));
    c->nlph[i] = (double*)xcalloc(c->nph, sizeof(double));
    bad = !c->q[i] || !c->dq[i] || !c->phys[i] || !c->nl[i] || !c->nlph[i];
  }
  for (int i = 0; !bad && i < W; ++i) {
    c->wsp[i] = (cplx*)xcalloc(c->nsp, sizeof(cplx));
    c->wph[i] = (double*)xcalloc(c->nph, sizeof(double));
    bad = !c->wsp[i] || !c->wph[i];
  }
  c->tmp = (cplx*)xcalloc(c->nsp, sizeof(cplx));
  c->kx = (double*)xcalloc(c->nx, sizeof(double));
  c->ky = (double*)xcalloc(c->ny, sizeof(double));
  c->kz = (double*)xcalloc(c->nzh, sizeof(double));
  c->mx = (int*)xcalloc(c->nx, sizeof(int));
  c->my = (int*)xcalloc(c->ny, sizeof(int));
  c->mz = (int*)xcalloc(c->nzh, sizeof(int));
  c->mask = (uint8_t*)xcalloc(c->nsp, 1);
  if (bad || !c->tmp || !c->kx || !c->ky || !c->kz || !c->mx || !c->my || !c->mz || !c->mask) {
    ctx_free(c);
    return set_err(NULL, GNSS_ERR_ALLOC, "out of host memory (fields)");
  }
  if (g_backend == 1) {
    if (g_fftw.init_threads && g_fftw.plan_with_nthreads) {
      g_fftw.init_threads();
      g_fftw.plan_with_nthreads(g_threads > 0 ? g_threads : gnss_oracle_max_threads());
    }
    c->fw = g_fftw.plan_r2c_3d(c->nx, c->ny, c->nz, c->wph[0], c->tmp, FFTW_ESTIMATE_FLAG);
    c->bw = g_fftw.plan_c2r_3d(c->nx, c->ny, c->nz, c->tmp, c->wph[0], FFTW_ESTIMATE_FLAG);
    if (!c->fw || !c->bw) {
      ctx_free(c);
      return set_err(NULL, GNSS_ERR_DEVICE, "FFTW planning failed");
    }
  }
  *out = c;
  return GNSS_OK;
}

EXPORT int gnss_destroy(gnss_ctx* c) {
  ctx_free(c);
  return GNSS_OK;
}

static int n_bins_of(const gnss_ctx* c);

EXPORT int gnss_get_sizes(const gnss_ctx* c, gnss_sizes* o) {
  if (!c || !o) return GNSS_ERR_ARG;
  o->n_fields = c->F;
  o->nzh = c->nzh;
  o->n_prof = 7 + 3 * c->S;
  o->n_bins = n_bins_of(c);
  o->n_physical = (int64_t)c->nph;
  o->n_spectral = (int64_t)c->nsp;
  return GNSS_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Wavenumbers and mask                                                                       */
/* ------------------------------------------------------------------------------------------ */

EXPORT int gnss_init_wavenumbers(gnss_ctx* c) {
  if (!c) return GNSS_ERR_ARG;
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < c->nx; ++i) { c->mx[i] = mode_of(i, c->nx); c->kx[i] = (2.0 * pi / c->p.lx) * (double)c->mx[i]; }
  for (int i = 0; i < c->ny; ++i) { c->my[i] = mode_of(i, c->ny); c->ky[i] = (2.0 * pi / c->p.ly) * (double)c->my[i]; }
  for (int i = 0; i < c->nzh; ++i) { c->mz[i] = mode_of(i, c->nz); c->kz[i] = (2.0 * pi / c->p.lz) * (double)c->mz[i]; }
  /* 2/3 rule: keep iff 9 * sum_i (m_i G/n_i)^2 <= G^2, G = lcm(nx, ny, nz) */
  int64_t g = c->nx;
  g = g / gcd64(g, c->ny) * c->ny;
  g = g / gcd64(g, c->nz) * c->nz;
  const int64_t ax = g / c->nx, ay = g / c->ny, az = g / c->nz;
  for (int x = 0; x < c->nx; ++x)
    for (int y = 0; y < c->ny; ++y)
      for (int k = 0; k < c->nzh; ++k) {
        int64_t a = (int64_t)c->mx[x] * ax, b = (int64_t)c->my[y] * ay, d = (int64_t)c->mz[k] * az;
        c->mask[((size_t)x * c->ny + y) * c->nzh + k] = (9 * (a * a + b * b + d * d) <= g * g) ? 1 : 0;
      }
  c->have_k = 1;
  return GNSS_OK;
}

EXPORT int gnss_get_wavenumbers(gnss_ctx* c, double* kx, double* ky, double* kz) {
  if (!c) return GNSS_ERR_ARG;
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  if (kx) memcpy(kx, c->kx, sizeof(double) * c->nx);
  if (ky) memcpy(ky, c->ky, sizeof(double) * c->ny);
  if (kz) memcpy(kz, c->kz, sizeof(double) * c->nzh);
  return GNSS_OK;
}

EXPORT int gnss_get_mask(gnss_ctx* c, uint8_t* mask) {
  if (!c || !mask) return GNSS_ERR_ARG;
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  for (int y = 0; y < c->ny; ++y)
    for (int x = 0; x < c->nx; ++x)
      memcpy(mask + ((size_t)y * c->nx + x) * c->nzh, c->mask + ((size_t)x * c->ny + y) * c->nzh, c->nzh);
  return GNSS_OK;
}

static void apply_mask(gnss_ctx* c, cplx* f) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nsp; ++i)
    if (!c->mask[i]) f[i] = 0;
}

EXPORT int gnss_dealias(gnss_ctx* c) {
  if (!c) return GNSS_ERR_ARG;
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  for (int f = 0; f < c->F; ++f) apply_mask(c, c->q[f]);
  return GNSS_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Initial condition: tanh jet + seeded divergence-free noise, FP64, spectral space           */
/* ------------------------------------------------------------------------------------------ */

static uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Complex standard Gaussian (unit variance per real part) keyed by (seed, component, mode). */
static cplx noise_gauss(uint64_t seed, int comp, int mx, int my, int mz) {
  uint64_t key = (uint64_t)(unsigned)comp | ((uint64_t)(unsigned)(mx + 4096) << 8) |
                 ((uint64_t)(unsigned)(my + 4096) << 24) | ((uint64_t)(unsigned)(mz + 4096) << 40);
  uint64_t h1 = splitmix64(seed ^ splitmix64(key));
  uint64_t h2 = splitmix64(h1);
  double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double r = sqrt(-2.0 * log(u1));
  double a = 6.283185307179586476925286766559 * u2;
  return r * cos(a) + I * (r * sin(a));
}

/* jet and scalar profile (initializeVelocityKernel_mgpu / initializeScalarKernel_mgpu,
 * initialize.cu:114-176); the colloid's uses its own thickness theta_c
 * (initializeColloidKernel_mgpu, initialize.cu:178-194) */
static double jet_profile(const gnss_params* p, double theta, double y) {
  const double h = p->jet_h > 0 ? p->jet_h : 1.0;
  return 0.5 * (1.0 + tanh((0.5 * h - fabs(y - 0.5 * p->ly)) / (2.0 * theta)));
}

EXPORT int gnss_init_jet(gnss_ctx* c) {
  if (!c) return GNSS_ERR_ARG;
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  if (!(c->p.theta > 0)) return set_err(c, GNSS_ERR_ARG, "theta must be > 0");
  const int nx = c->nx, ny = c->ny, nzh = c->nzh;
  const double nn = (double)nx * (double)ny * (double)c->nz;
  for (int f = 0; f < c->F; ++f) {
    memset(c->q[f], 0, sizeof(cplx) * c->nsp);
    memset(c->dq[f], 0, sizeof(cplx) * c->nsp);
  }
  /* jet: field depends on y only -> nonzero only at kx = kz = 0; direct DFT along y */
  cplx* tw = (cplx*)malloc(sizeof(cplx) * ny);
  double* prof = (double*)malloc(sizeof(double) * 2 * ny);
  if (!tw || !prof) { free(tw); free(prof); return set_err(c, GNSS_ERR_ALLOC, "out of host memory"); }
  double* prof_c = prof + ny;
  const double theta_c = c->p.theta_c > 0 ? c->p.theta_c : c->p.theta;
  twiddle_table(tw, ny);
  for (int j = 0; j < ny; ++j) {
    prof[j] = jet_profile(&c->p, c->p.theta, (double)j * c->p.ly / (double)ny);
    prof_c[j] = jet_profile(&c->p, theta_c, (double)j * c->p.ly / (double)ny);
  }
  for (int iy = 0; iy < ny; ++iy) {
    double re = 0.0, im = 0.0, rc = 0.0, ic = 0.0;
    for (int j = 0; j < ny; ++j) {
      int m = (int)(((int64_t)iy * j) % ny);
      re += prof[j] * creal(tw[m]);
      im += prof[j] * cimag(tw[m]);
      rc += prof_c[j] * creal(tw[m]);
      ic += prof_c[j] * cimag(tw[m]);
    }
    const double sx = (double)nx * (double)c->nz;
    cplx v = (re * sx) + I * (im * sx);
    c->q[0][(size_t)iy * nzh] = v;
    for (int s = 0; s < c->S; ++s) c->q[3 + s][(size_t)iy * nzh] = v;
    if (c->colloid) c->q[2 + c->S][(size_t)iy * nzh] = (rc * sx) + I * (ic * sx);
  }
  free(tw);
  free(prof);
  /* noise on 1 <= |m|^2 <= kmax^2, canonical half space, projected, normalised */
  const double kmax = c->p.noise_kmax;
  if (c->p.noise_amp != 0.0 && kmax >= 1.0) {
    const int K = (int)floor(kmax);
    const double k2max = kmax * kmax;
    const double pi = 3.14159265358979323846;
    /* first pass: energy of the projected raw coefficients */
    double energy = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      double alpha = pass ? c->p.noise_amp / sqrt(energy / 3.0) : 0.0;
      if (pass && !(energy > 0)) break;
      for (int mx = -K; mx <= K; ++mx)
        for (int my = -K; my <= K; ++my)
          for (int mz = 0; mz <= K; ++mz) {
            const int m2 = mx * mx + my * my + mz * mz;
            if (m2 < 1 || (double)m2 > k2max) continue;
            if (!(mz > 0 || my > 0 || (my == 0 && mx > 0))) continue;
            if (2 * abs(mx) >= nx || 2 * abs(my) >= ny || 2 * mz >= c->nz) continue;
            const double kv[3] = {(2.0 * pi / c->p.lx) * mx, (2.0 * pi / c->p.ly) * my, (2.0 * pi / c->p.lz) * mz};
            cplx a[3];
            for (int comp = 0; comp < 3; ++comp) a[comp] = noise_gauss(c->p.seed, comp, mx, my, mz);
            const double kk = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
            const cplx kd = kv[0] * a[0] + kv[1] * a[1] + kv[2] * a[2];
            for (int comp = 0; comp < 3; ++comp) a[comp] -= kv[comp] * (kd / kk);
            if (!pass) {
              for (int comp = 0; comp < 3; ++comp) energy += 2.0 * (creal(a[comp]) * creal(a[comp]) + cimag(a[comp]) * cimag(a[comp]));
              continue;
            }
            const int ix = mx >= 0 ? mx : mx + nx, iy = my >= 0 ? my : my + ny;
            const size_t idx = ((size_t)ix * ny + iy) * nzh + mz;
            for (int comp = 0; comp < 3; ++comp) {
              const cplx v = (nn * alpha) * a[comp];
              c->q[comp][idx] += v;
              if (mz == 0) {
                const int jx = mx > 0 ? nx - mx : -mx, jy = my > 0 ? ny - my : -my;
                const size_t jdx = ((size_t)jx * ny + jy) * nzh;
                c->q[comp][jdx] += conj(v);
              }
            }
          }
    }
  }
  for (int f = 0; f < c->F; ++f) apply_mask(c, c->q[f]);
  c->t = 0.0;
  c->step = 0;
  c->have_ic = 1;
  return GNSS_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Step                                                                                       */
/* ------------------------------------------------------------------------------------------ */

static const double RK_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double RK_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
static const double RK_C[4] = {0.0, 1.0 / 3.0, 3.0 / 4.0, 1.0};

/* Nonlinear terms of the state q (spectral, masked): nl[0..2] = (u x w)^, nl[3+s] = (u.grad phi_s)^
 * (for the colloid: the drift-advection term above) */
static void nonlinear(gnss_ctx* c) {
  const int nx = c->nx, ny = c->ny, nzh = c->nzh, S = c->S;
  cplx** w = c->wsp;
  /* spectral prologue: u, v, w, omega = i k x u, grad phi = i k phi */
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)nx * ny; ++r) {
    const int x = (int)(r / ny), y = (int)(r % ny);
    const double kx = c->kx[x], ky = c->ky[y];
    for (int k = 0; k < nzh; ++k) {
      const size_t i = (size_t)r * nzh + k;
      const double kz = c->kz[k];
      const cplx u = c->q[0][i], v = c->q[1][i], ww = c->q[2][i];
      w[0][i] = u;
      w[1][i] = v;
      w[2][i] = ww;
      w[3][i] = I * (ky * ww - kz * v);
      w[4][i] = I * (kz * u - kx * ww);
      w[5][i] = I * (kx * v - ky * u);
      for (int s = 0; s < S; ++s) {
        const cplx ph = c->q[3 + s][i];
        w[6 + 3 * s][i] = I * (kx * ph);
        w[7 + 3 * s][i] = I * (ky * ph);
        w[8 + 3 * s][i] = I * (kz * ph);
      }
      if (c->colloid) {
        /* c itself and lap Z = div(grad Z) = -k^2 Z (divergenceKernel_mgpu, solver.cu:276-293,
         * applied to grad Z in scalarAdvection, solver.cu:422-423) */
        w[6 + 3 * S][i] = c->q[2 + S][i];
        w[7 + 3 * S][i] = -(kx * kx + ky * ky + kz * kz) * c->q[3][i];
      }
    }
  }
  const int W = c->W, P = S - c->colloid;
  const double alpha = c->p.alpha;
  for (int f = 0; f < W; ++f) c2r3d(c, w[f], c->wph[f]);
  double** p = c->wph;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nph; ++i) {
    const double u = p[0][i], v = p[1][i], ww = p[2][i];
    const double ox = p[3][i], oy = p[4][i], oz = p[5][i];
    c->nlph[0][i] = v * oz - ww * oy;
    c->nlph[1][i] = ww * ox - u * oz;
    c->nlph[2][i] = u * oy - v * ox;
    for (int s = 0; s < P; ++s)
      c->nlph[3 + s][i] = u * p[6 + 3 * s][i] + v * p[7 + 3 * s][i] + ww * p[8 + 3 * s][i];
    if (c->colloid) {
      /* colloidAdvectionKernel_mgpu (solver.cu:374-391): with udp = grad Z and divVdp = lap Z,
       * alpha (divVdp c + grad c . udp) + u . grad c */
      const int g = 6 + 3 * P;
      const double cx = p[g][i], cy = p[g + 1][i], cz = p[g + 2][i];
      const double cc = p[6 + 3 * S][i], lz = p[7 + 3 * S][i];
      c->nlph[2 + S][i] = alpha * (lz * cc + cx * p[6][i] + cy * p[7][i] + cz * p[8][i]) + (u * cx + v * cy + ww * cz);
    }
  }
  for (int f = 0; f < 3 + S; ++f) r2c3d(c, c->nlph[f], c->nl[f]);
}

static void rk_substage(gnss_ctx* c, int s) {
  const int nx = c->nx, ny = c->ny, nzh = c->nzh, S = c->S;
  const double dt = c->p.dt;
  const double a = RK_A[s], b = RK_B[s], tau = (RK_C[s + 1] - RK_C[s]) * dt;
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)nx * ny; ++r) {
    const int x = (int)(r / ny), y = (int)(r % ny);
    const double kx = c->kx[x], ky = c->ky[y];
    for (int k = 0; k < nzh; ++k) {
      const size_t i = (size_t)r * nzh + k;
      if (!c->mask[i]) {
        for (int f = 0; f < c->F; ++f) { c->q[f][i] = 0; c->dq[f][i] = 0; }
        continue;
      }
      const double kz = c->kz[k];
      const double k2 = kx * kx + ky * ky + kz * kz;
      cplx rhs[3 + GNSS_MAX_SCALARS];
      if (k2 == 0.0) {
        for (int f = 0; f < c->F; ++f) rhs[f] = 0;
      } else {
        const cplx n0 = c->nl[0][i], n1 = c->nl[1][i], n2 = c->nl[2][i];
        const cplx kn = (kx * n0 + ky * n1 + kz * n2) / k2;
        rhs[0] = n0 - kx * kn;
        rhs[1] = n1 - ky * kn;
        rhs[2] = n2 - kz * kn;
        for (int q = 0; q < S; ++q) rhs[3 + q] = -c->nl[3 + q][i];
      }
      for (int f = 0; f < c->F; ++f) {
        const double nu = f < 3 ? c->nu : c->diff[f - 3];
        /* E = exp(-nu k^2 tau) applied as 1 - eps, eps = -expm1(-nu k^2 tau): the same factor,
           formed so that an E close to 1 carries no rounding of its own.
           E (q + b d) = q + (b d - eps (q + b d)): the state is rounded once per substage. */
        const double eps = -expm1(-nu * k2 * tau);
        const cplx d = (s == 0 ? 0 : a * c->dq[f][i]) + dt * rhs[f];
        const cplx qf = c->q[f][i];
        c->q[f][i] = qf + (b * d - eps * (qf + b * d));
        c->dq[f][i] = d - eps * d;
      }
    }
  }
}

EXPORT int gnss_step(gnss_ctx* c, int nsteps) {
  if (!c) return GNSS_ERR_ARG;
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  if (nsteps < 0) return set_err(c, GNSS_ERR_ARG, "nsteps must be >= 0");
  if (nsteps == 0) return GNSS_OK;
  /* the step starts from a dealiased state (deAlias at the top of solver_ps, solver.cu:704) */
  for (int f = 0; f < c->F; ++f) apply_mask(c, c->q[f]);
  for (int n = 0; n < nsteps; ++n) {
    for (int s = 0; s < 3; ++s) {
      nonlinear(c);
      rk_substage(c, s);
    }
    c->step += 1;
    c->t = (double)c->step * c->p.dt;
  }
  /* NaN/Inf watchdog on the mean mode */
  for (int f = 0; f < c->F; ++f) {
    double v = creal(c->q[f][0]) + cimag(c->q[f][0]);
    if (!isfinite(v)) return set_err(c, GNSS_ERR_NUMERIC, "non-finite value in field %d", f);
  }
  return GNSS_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Statistics                                                                                 */
/* ------------------------------------------------------------------------------------------ */

static int64_t isqrt64(int64_t v) {
  int64_t r = (int64_t)sqrt((double)v);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return r;
}

static int equal_box(const gnss_ctx* c) { return c->p.lx == c->p.ly && c->p.ly == c->p.lz; }

/* Shell index nint(|m|) (equal box lengths: exact integer arithmetic), else nint(|k|/k0),
 * k0 = 2 pi / max(L). */
static int shell_of(const gnss_ctx* c, int mx, int my, int mz) {
  if (equal_box(c)) {
    int64_t m2 = (int64_t)mx * mx + (int64_t)my * my + (int64_t)mz * mz;
    return (int)((isqrt64(4 * m2) + 1) / 2);
  }
  double lm = c->p.lx;
  if (c->p.ly > lm) lm = c->p.ly;
  if (c->p.lz > lm) lm = c->p.lz;
  double a = mx * lm / c->p.lx, b = my * lm / c->p.ly, d = mz * lm / c->p.lz;
  return (int)floor(sqrt(a * a + b * b + d * d) + 0.5);
}

static int n_bins_of(const gnss_ctx* c) {
  return shell_of(c, c->nx / 2, c->ny / 2, c->nz / 2) + 1;
}

EXPORT int gnss_stats(gnss_ctx* c, gnss_stats_out* o, double* yprof, double* spectrum) {
  if (!c) return GNSS_ERR_ARG;
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzh = c->nzh, S = c->S;
  const double nn = (double)nx * ny * nz, inn2 = 1.0 / (nn * nn);
  const int nb = n_bins_of(c);
  /* spectral sums, per x plane, then planes in order (deterministic) */
  const int NQ = 5 + 2 * GNSS_MAX_SCALARS;
  double* part = (double*)calloc((size_t)nx * NQ, sizeof(double));
  double* spec = (double*)calloc((size_t)(1 + S) * nb, sizeof(double));
  if (!part || !spec) { free(part); free(spec); return set_err(c, GNSS_ERR_ALLOC, "out of host memory"); }
#pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; ++x) {
    double* pp = part + (size_t)x * NQ;
    const double kx = c->kx[x];
    for (int y = 0; y < ny; ++y) {
      const double ky = c->ky[y];
      for (int k = 0; k < nzh; ++k) {
        const size_t i = ((size_t)x * ny + y) * nzh + k;
        const double wgt = (k == 0 || 2 * k == nz) ? 1.0 : 2.0;
        const double kz = c->kz[k];
        const double k2 = kx * kx + ky * ky + kz * kz;
        const cplx u = c->q[0][i], v = c->q[1][i], w = c->q[2][i];
        const double e = (creal(u) * creal(u) + cimag(u) * cimag(u) + creal(v) * creal(v) + cimag(v) * cimag(v) +
                          creal(w) * creal(w) + cimag(w) * cimag(w)) * inn2;
        const cplx dv = kx * u + ky * v + kz * w;
        pp[0] += wgt * e;
        pp[1] += wgt * k2 * e;
        if (k2 > 0) pp[2] += wgt * e / (2.0 * sqrt(k2));
        pp[3] += wgt * (creal(dv) * creal(dv) + cimag(dv) * cimag(dv)) * inn2;
        for (int s = 0; s < S; ++s) {
          const cplx ph = c->q[3 + s][i];
          const double ep = (creal(ph) * creal(ph) + cimag(ph) * cimag(ph)) * inn2;
          pp[5 + 2 * s] += wgt * k2 * ep;
          if (k2 > 0) pp[6 + 2 * s] += wgt * ep;
        }
      }
    }
  }
  double sum[5 + 2 * GNSS_MAX_SCALARS] = {0};
  for (int x = 0; x < nx; ++x)
    for (int qq = 0; qq < NQ; ++qq) sum[qq] += part[(size_t)x * NQ + qq];
  /* spectrum: sequential, fixed order */
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int k = 0; k < nzh; ++k) {
        const size_t i = ((size_t)x * ny + y) * nzh + k;
        const double wgt = (k == 0 || 2 * k == nz) ? 1.0 : 2.0;
        const int bn = shell_of(c, c->mx[x], c->my[y], c->mz[k]);
        const cplx u = c->q[0][i], v = c->q[1][i], w = c->q[2][i];
        const double e = (creal(u) * creal(u) + cimag(u) * cimag(u) + creal(v) * creal(v) + cimag(v) * cimag(v) +
                          creal(w) * creal(w) + cimag(w) * cimag(w)) * inn2;
        spec[bn] += 0.5 * wgt * e;
        for (int s = 0; s < S; ++s) {
          const cplx ph = c->q[3 + s][i];
          spec[(size_t)(1 + s) * nb + bn] += 0.5 * wgt * (creal(ph) * creal(ph) + cimag(ph) * cimag(ph)) * inn2;
        }
      }
  /* physical: inverse transforms of u, v, w and the scalars */
  for (int f = 0; f < c->F; ++f) c2r3d(c, c->q[f], c->phys[f]);
  const int NP = 7 + 3 * S;
  double* prof = (double*)calloc((size_t)NP * ny, sizeof(double));
  double* cflp = (double*)calloc((size_t)ny * 2, sizeof(double));
  if (!prof || !cflp) { free(part); free(spec); free(prof); free(cflp); return set_err(c, GNSS_ERR_ALLOC, "oom"); }
  const double dx = c->p.lx / nx, dy = c->p.ly / ny, dz = c->p.lz / nz;
  const double inv_xz = 1.0 / ((double)nx * nz);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < ny; ++y) {
    double mean[3 + GNSS_MAX_SCALARS] = {0};
    for (int f = 0; f < c->F; ++f) {
      double acc = 0.0;
      for (int x = 0; x < nx; ++x) {
        const double* row = c->phys[f] + ((size_t)x * ny + y) * nz;
        for (int z = 0; z < nz; ++z) acc += row[z];
      }
      mean[f] = acc * inv_xz;
    }
    double uu = 0, vv = 0, ww = 0, uv = 0, pp[GNSS_MAX_SCALARS] = {0}, vp[GNSS_MAX_SCALARS] = {0};
    double cmax = 0, umax = 0;
    for (int x = 0; x < nx; ++x) {
      const size_t o = ((size_t)x * ny + y) * nz;
      for (int z = 0; z < nz; ++z) {
        const double u = c->phys[0][o + z], v = c->phys[1][o + z], w = c->phys[2][o + z];
        const double du = u - mean[0], dv = v - mean[1], dw = w - mean[2];
        uu += du * du; vv += dv * dv; ww += dw * dw; uv += du * dv;
        for (int s = 0; s < S; ++s) {
          const double dp = c->phys[3 + s][o + z] - mean[3 + s];
          pp[s] += dp * dp;
          vp[s] += dv * dp;
        }
        const double cf = fabs(u) / dx + fabs(v) / dy + fabs(w) / dz;
        if (cf > cmax) cmax = cf;
        double am = fabs(u);
        if (fabs(v) > am) am = fabs(v);
        if (fabs(w) > am) am = fabs(w);
        if (am > umax) umax = am;
      }
    }
    prof[0 * ny + y] = mean[0];
    prof[1 * ny + y] = mean[1];
    prof[2 * ny + y] = mean[2];
    prof[3 * ny + y] = uu * inv_xz;
    prof[4 * ny + y] = vv * inv_xz;
    prof[5 * ny + y] = ww * inv_xz;
    prof[6 * ny + y] = uv * inv_xz;
    for (int s = 0; s < S; ++s) {
      prof[(7 + 3 * s) * ny + y] = mean[3 + s];
      prof[(8 + 3 * s) * ny + y] = pp[s] * inv_xz;
      prof[(9 + 3 * s) * ny + y] = vp[s] * inv_xz;
    }
    cflp[2 * y] = cmax;
    cflp[2 * y + 1] = umax;
  }
  if (o) {
    memset(o, 0, sizeof(*o));
    o->t = c->t;
    o->step = c->step;
    o->ke = 0.5 * sum[0];
    o->vrms = sqrt(sum[0] / 3.0);
    o->epsilon = c->nu * sum[1];
    o->enstrophy = 0.5 * sum[1];
    o->lint = o->ke > 0 ? 3.0 * 3.14159265358979323846 / 4.0 * sum[2] / o->ke : 0.0;
    o->lambda = o->epsilon > 0 ? sqrt(15.0 * c->nu * o->vrms * o->vrms / o->epsilon) : 0.0;
    o->eta = o->epsilon > 0 ? sqrt(sqrt(c->nu * c->nu * c->nu / o->epsilon)) : 0.0;
    o->divergence = sum[1] > 0 ? sqrt(sum[3] / sum[1]) : 0.0;
    for (int s = 0; s < S; ++s) {
      o->chi[s] = 2.0 * c->diff[s] * sum[5 + 2 * s];
      o->scalar_var[s] = sum[6 + 2 * s];
    }
    double cm = 0, um = 0;
    for (int y = 0; y < ny; ++y) {
      if (cflp[2 * y] > cm) cm = cflp[2 * y];
      if (cflp[2 * y + 1] > um) um = cflp[2 * y + 1];
    }
    o->cfl = c->p.dt * cm;
    o->umax = um;
  }
  if (yprof) memcpy(yprof, prof, sizeof(double) * (size_t)NP * ny);
  if (spectrum) memcpy(spectrum, spec, sizeof(double) * (size_t)(1 + S) * nb);
  free(part); free(spec); free(prof); free(cflp);
  return GNSS_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Host I/O                                                                                   */
/* ------------------------------------------------------------------------------------------ */

EXPORT int gnss_get_field(gnss_ctx* c, int field, int space, double* host) {
  if (!c || !host) return GNSS_ERR_ARG;
  if (field < 0 || field >= c->F) return set_err(c, GNSS_ERR_ARG, "field %d out of range", field);
  if (space == GNSS_PHYSICAL) {
    memcpy(host, c->phys[field], sizeof(double) * c->nph);
    return GNSS_OK;
  }
  if (space != GNSS_SPECTRAL) return set_err(c, GNSS_ERR_ARG, "bad space");
  cplx* dst = (cplx*)host;
  for (int y = 0; y < c->ny; ++y)
    for (int x = 0; x < c->nx; ++x)
      memcpy(dst + ((size_t)y * c->nx + x) * c->nzh, c->q[field] + ((size_t)x * c->ny + y) * c->nzh,
             sizeof(cplx) * c->nzh);
  return GNSS_OK;
}

EXPORT int gnss_set_field(gnss_ctx* c, int field, int space, const double* host) {
  if (!c || !host) return GNSS_ERR_ARG;
  if (field < 0 || field >= c->F) return set_err(c, GNSS_ERR_ARG, "field %d out of range", field);
  if (space == GNSS_PHYSICAL) {
    memcpy(c->phys[field], host, sizeof(double) * c->nph);
    return GNSS_OK;
  }
  if (space != GNSS_SPECTRAL) return set_err(c, GNSS_ERR_ARG, "bad space");
  const cplx* src = (const cplx*)host;
  for (int y = 0; y < c->ny; ++y)
    for (int x = 0; x < c->nx; ++x)
      memcpy(c->q[field] + ((size_t)x * c->ny + y) * c->nzh, src + ((size_t)y * c->nx + x) * c->nzh,
             sizeof(cplx) * c->nzh);
  memset(c->dq[field], 0, sizeof(cplx) * c->nsp);
  return GNSS_OK;
}

EXPORT int gnss_forward(gnss_ctx* c, int field) {
  if (!c) return GNSS_ERR_ARG;
  if (field < 0 || field >= c->F) return set_err(c, GNSS_ERR_ARG, "field %d out of range", field);
  r2c3d(c, c->phys[field], c->q[field]);
  return GNSS_OK;
}

EXPORT int gnss_inverse(gnss_ctx* c, int field) {
  if (!c) return GNSS_ERR_ARG;
  if (field < 0 || field >= c->F) return set_err(c, GNSS_ERR_ARG, "field %d out of range", field);
  c2r3d(c, c->q[field], c->phys[field]);
  return GNSS_OK;
}

EXPORT int gnss_get_time(const gnss_ctx* c, double* t, int64_t* step) {
  if (!c) return GNSS_ERR_ARG;
  if (t) *t = c->t;
  if (step) *step = c->step;
  return GNSS_OK;
}

/* Raw field files in the reference's format (iofuncs.cu:329-358 writer, 462-497 reader). */
/* file names of the reference's field dumps (iofuncs.cu:329-432, 262-327): u v w, z for the
 * scalar, c for the colloid; further scalars z1, z2, ... */
static void field_name(const gnss_ctx* c, int f, char* out, size_t n) {
  static const char* vel[3] = {"u", "v", "w"};
  if (f < 3) snprintf(out, n, "%s", vel[f]);
  else if (c->colloid && f == c->F - 1) snprintf(out, n, "c");
  else if (f == 3) snprintf(out, n, "z");
  else snprintf(out, n, "z%d", f - 3);
}

EXPORT int gnss_save_fields(gnss_ctx* c, const char* dir, int iter) {
  if (!c || !dir) return GNSS_ERR_ARG;
  char nm[16], path[4096];
  for (int f = 0; f < c->F; ++f) {
    c2r3d(c, c->q[f], c->phys[f]);
    field_name(c, f, nm, sizeof nm);
    snprintf(path, sizeof path, "%s/%s.%d", dir, nm, iter);
    FILE* fp = fopen(path, "wb");
    if (!fp) return set_err(c, GNSS_ERR_ARG, "cannot open %s", path);
    const double hdr = (double)c->nph * (double)sizeof(double);
    int ok = fwrite(&hdr, sizeof(double), 1, fp) == 1 && fwrite(c->phys[f], sizeof(double), c->nph, fp) == c->nph;
    fclose(fp);
    if (!ok) return set_err(c, GNSS_ERR_ARG, "short write to %s", path);
  }
  snprintf(path, sizeof path, "%s/gnss_state.%d", dir, iter);
  FILE* fp = fopen(path, "w");
  if (!fp) return set_err(c, GNSS_ERR_ARG, "cannot open %s", path);
  fprintf(fp, "step %lld t %.17g\n", (long long)c->step, c->t);
  fclose(fp);
  return GNSS_OK;
}

EXPORT int gnss_load_fields(gnss_ctx* c, const char* dir, int iter) {
  if (!c || !dir) return GNSS_ERR_ARG;
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  char nm[16], path[4096];
  for (int f = 0; f < c->F; ++f) {
    field_name(c, f, nm, sizeof nm);
    snprintf(path, sizeof path, "%s/%s.%d", dir, nm, iter);
    FILE* fp = fopen(path, "rb");
    if (!fp) return set_err(c, GNSS_ERR_ARG, "cannot open %s", path);
    double hdr = 0;
    if (fread(&hdr, sizeof(double), 1, fp) != 1 || hdr != (double)c->nph * (double)sizeof(double)) {
      fclose(fp);
      return set_err(c, GNSS_ERR_ARG, "%s: size header does not match nx*ny*nz doubles", path);
    }
    int ok = fread(c->phys[f], sizeof(double), c->nph, fp) == c->nph;
    fclose(fp);
    if (!ok) return set_err(c, GNSS_ERR_ARG, "short read from %s", path);
    r2c3d(c, c->phys[f], c->q[f]);
    apply_mask(c, c->q[f]);
    memset(c->dq[f], 0, sizeof(cplx) * c->nsp);
  }
  c->step = iter;
  c->t = (double)iter * c->p.dt;
  snprintf(path, sizeof path, "%s/gnss_state.%d", dir, iter);
  FILE* fp = fopen(path, "r");
  if (fp) {
    long long a;
    double b;
    if (fscanf(fp, "step %lld t %lf", &a, &b) == 2) { c->step = a; c->t = b; }
    fclose(fp);
  }
  return GNSS_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Diagnostics: plane slices and iso-surface areas                                            */
/* ------------------------------------------------------------------------------------------ */

/* Physical values of field f, or of |omega| = |i k x u^| for GNSS_FIELD_VORTICITY
 * (VectorMagnitude_kernel, statistics.cu:2005-2018), in a buffer of nph doubles. */
static const double* diag_field(gnss_ctx* c, int f) {
  if (f >= 0) {
    c2r3d(c, c->q[f], c->phys[f]);
    return c->phys[f];
  }
  const int nx = c->nx, ny = c->ny, nzh = c->nzh;
  cplx** w = c->wsp;
  for (long r = 0; r < (long)nx * ny; ++r) {
    const int x = (int)(r / ny), y = (int)(r % ny);
    const double kx = c->kx[x], ky = c->ky[y];
    for (int k = 0; k < nzh; ++k) {
      const size_t i = (size_t)r * nzh + k;
      const double kz = c->kz[k];
      const cplx u = c->q[0][i], v = c->q[1][i], ww = c->q[2][i];
      w[0][i] = I * (ky * ww - kz * v);
      w[1][i] = I * (kz * u - kx * ww);
      w[2][i] = I * (kx * v - ky * u);
    }
  }
  for (int f2 = 0; f2 < 3; ++f2) c2r3d(c, w[f2], c->wph[f2]);
  double* m = c->wph[3];
  for (size_t i = 0; i < c->nph; ++i) {
    const double a = c->wph[0][i], b = c->wph[1][i], d = c->wph[2][i];
    m[i] = sqrt(a * a + b * b + d * d);
  }
  return m;
}

static int diag_check(gnss_ctx* c, int field) {
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  if (field < GNSS_FIELD_VORTICITY || field >= c->F)
    return set_err(c, GNSS_ERR_ARG, "field %d out of range", field);
  return GNSS_OK;
}

static void plane_of(const gnss_ctx* c, const double* a, int axis, int index, double* out) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  if (axis == 2) {
    for (int x = 0; x < nx; ++x)
      for (int y = 0; y < ny; ++y) out[(size_t)x * ny + y] = a[((size_t)x * ny + y) * nz + index];
  } else if (axis == 1) {
    for (int x = 0; x < nx; ++x)
      for (int z = 0; z < nz; ++z) out[(size_t)x * nz + z] = a[((size_t)x * ny + index) * nz + z];
  } else {
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) out[(size_t)y * nz + z] = a[((size_t)index * ny + y) * nz + z];
  }
}

EXPORT int gnss_get_slice(gnss_ctx* c, int field, int axis, int index, double* out) {
  if (!c || !out) return GNSS_ERR_ARG;
  int rc = diag_check(c, field);
  if (rc) return rc;
  const int n[3] = {c->nx, c->ny, c->nz};
  if (axis < 0 || axis > 2 || index < 0 || index >= n[axis])
    return set_err(c, GNSS_ERR_ARG, "plane %d of axis %d out of range", index, axis);
  plane_of(c, diag_field(c, field), axis, index, out);
  return GNSS_OK;
}

EXPORT int gnss_save_slices(gnss_ctx* c, const char* dir, int iter) {
  if (!c || !dir) return GNSS_ERR_ARG;
  if (!c->have_k) return set_err(c, GNSS_ERR_STATE, "call gnss_init_wavenumbers first");
  const size_t np = (size_t)c->nx * c->ny;
  double* buf = (double*)malloc(sizeof(double) * np);
  if (!buf) return set_err(c, GNSS_ERR_ALLOC, "out of host memory");
  char nm[16], path[4096];
  for (int f = 0; f < c->F; ++f) {
    plane_of(c, diag_field(c, f), 2, c->nz / 2, buf);
    field_name(c, f, nm, sizeof nm);
    snprintf(path, sizeof path, "%s/%s_xy.%d", dir, nm, iter);
    FILE* fp = fopen(path, "wb");
    if (!fp) { free(buf); return set_err(c, GNSS_ERR_ARG, "cannot open %s", path); }
    const double hdr = (double)np * (double)sizeof(double);
    int ok = fwrite(&hdr, sizeof(double), 1, fp) == 1 && fwrite(buf, sizeof(double), np, fp) == np;
    fclose(fp);
    if (!ok) { free(buf); return set_err(c, GNSS_ERR_ARG, "short write to %s", path); }
  }
  free(buf);
  return GNSS_OK;
}

/* surfaceArea_kernel_mgpu (statistics.cu:1310-1460): -(grad F . grad chi) / |grad F| summed
 * over the grid times the cell volume; chi = (F <= iso); central differences, periodic wrap
 * (the reference's halo exchange, statistics.cu:1177-1209). */
EXPORT int gnss_surface_area(gnss_ctx* c, int field, double iso, double* area) {
  if (!c || !area) return GNSS_ERR_ARG;
  int rc = diag_check(c, field);
  if (rc) return rc;
  const double* F = diag_field(c, field);
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const double dx = c->p.lx / nx, dy = c->p.ly / ny, dz = c->p.lz / nz;
  double sum = 0.0;
  for (int x = 0; x < nx; ++x) {
    const int xm = (x + nx - 1) % nx, xp = (x + 1) % nx;
    double sx = 0.0;
    for (int y = 0; y < ny; ++y) {
      const int ym = (y + ny - 1) % ny, yp = (y + 1) % ny;
      for (int z = 0; z < nz; ++z) {
        const int zm = (z + nz - 1) % nz, zp = (z + 1) % nz;
#define AT(a, b, d) F[((size_t)(a) * ny + (b)) * nz + (d)]
        const double fxm = AT(xm, y, z), fxp = AT(xp, y, z), fym = AT(x, ym, z), fyp = AT(x, yp, z);
        const double fzm = AT(x, y, zm), fzp = AT(x, y, zp);
#undef AT
        const double gx = (fxp - fxm) / (2.0 * dx), gy = (fyp - fym) / (2.0 * dy), gz = (fzp - fzm) / (2.0 * dz);
        const double cx = ((fxp <= iso) - (fxm <= iso)) / (2.0 * dx);
        const double cy = ((fyp <= iso) - (fym <= iso)) / (2.0 * dy);
        const double cz = ((fzp <= iso) - (fzm <= iso)) / (2.0 * dz);
        if ((gx == 0 && gy == 0 && gz == 0) || (cx == 0 && cy == 0 && cz == 0)) continue;
        sx += -(gx * cx + gy * cy + gz * cz) / sqrt(gx * gx + gy * gy + gz * gz);
      }
    }
    sum += sx;
  }
  *area = sum * dx * dy * dz;
  return GNSS_OK;
}

EXPORT int gnss_timing_enable(gnss_ctx* c, int e) { (void)e; return c ? set_err(c, GNSS_ERR_UNSUPPORTED, "no kernels on the CPU") : GNSS_ERR_ARG; }
EXPORT int gnss_timing_report(gnss_ctx* c, char* b, int n) { (void)b; (void)n; return c ? set_err(c, GNSS_ERR_UNSUPPORTED, "no kernels on the CPU") : GNSS_ERR_ARG; }
EXPORT int gnss_bench_fft(gnss_ctx* c, int r, double* a, double* b) { (void)r; (void)a; (void)b; return c ? set_err(c, GNSS_ERR_UNSUPPORTED, "engine only") : GNSS_ERR_ARG; }
EXPORT int gnss_bench_copy(gnss_ctx* c, int r, double* a, double* b) { (void)r; (void)a; (void)b; return c ? set_err(c, GNSS_ERR_UNSUPPORTED, "engine only") : GNSS_ERR_ARG; }

EXPORT int gnss_autotune(gnss_ctx* c, int r, char* b, int n) { (void)r; (void)b; (void)n; return c ? set_err(c, GNSS_ERR_UNSUPPORTED, "engine only") : GNSS_ERR_ARG; }
EXPORT int gnss_set_tuning(gnss_ctx* c, const char* s) { (void)s; return c ? set_err(c, GNSS_ERR_UNSUPPORTED, "engine only") : GNSS_ERR_ARG; }

/* Ensembles on the CPU run the members one after the other (same semantics, no batching). */
struct gnss_ens {
  int n;
  gnss_ctx** m;
  char err[256];
};

EXPORT int gnss_ensemble_create(int n, const gnss_params* members, gnss_ens** out) {
  if (n <= 0 || !members || !out) return GNSS_ERR_ARG;
  gnss_ens* e = (gnss_ens*)calloc(1, sizeof(gnss_ens));
  if (!e) return GNSS_ERR_ALLOC;
  e->n = n;
  e->m = (gnss_ctx**)calloc((size_t)n, sizeof(gnss_ctx*));
  if (!e->m) { free(e); return GNSS_ERR_ALLOC; }
  for (int i = 0; i < n; ++i) {
    int rc = gnss_create(&members[i], &e->m[i]);
    if (rc) {
      for (int j = 0; j < i; ++j) gnss_destroy(e->m[j]);
      free(e->m);
      free(e);
      return rc;
    }
  }
  *out = e;
  return GNSS_OK;
}
EXPORT int gnss_ensemble_destroy(gnss_ens* e) {
  if (!e) return GNSS_OK;
  for (int i = 0; i < e->n; ++i) gnss_destroy(e->m[i]);
  free(e->m);
  free(e);
  return GNSS_OK;
}
EXPORT const char* gnss_ensemble_last_error(const gnss_ens* e) { return e ? e->err : g_create_err; }
EXPORT int gnss_ensemble_init(gnss_ens* e) {
  if (!e) return GNSS_ERR_ARG;
  for (int i = 0; i < e->n; ++i) {
    int rc = gnss_init_wavenumbers(e->m[i]);
    if (!rc) rc = gnss_init_jet(e->m[i]);
    if (rc) return rc;
  }
  return GNSS_OK;
}
EXPORT int gnss_ensemble_step(gnss_ens* e, int nsteps) {
  if (!e) return GNSS_ERR_ARG;
  for (int i = 0; i < e->n; ++i) {
    int rc = gnss_step(e->m[i], nsteps);
    if (rc) return rc;
  }
  return GNSS_OK;
}
EXPORT int gnss_ensemble_stats(gnss_ens* e, int i, gnss_stats_out* o, double* yp, double* sp) {
  if (!e || i < 0 || i >= e->n) return GNSS_ERR_ARG;
  return gnss_stats(e->m[i], o, yp, sp);
}
EXPORT int gnss_ensemble_get_field(gnss_ens* e, int i, int f, int s, double* h) {
  if (!e || i < 0 || i >= e->n) return GNSS_ERR_ARG;
  if (s == GNSS_PHYSICAL) { /* the member's current state, as the engine gives it */
    int rc = gnss_inverse(e->m[i], f);
    if (rc) return rc;
  }
  return gnss_get_field(e->m[i], f, s, h);
}
/* The CPU oracle has no device scratch: the allocation-free statistics call is gnss_stats. */
EXPORT int gnss_stats_noalloc(gnss_ctx* c, gnss_stats_out* o, double* yp, double* sp) {
  return gnss_stats(c, o, yp, sp);
}
EXPORT int gnss_ensemble_autotune(gnss_ens* e, int r, char* b, int n) {
  (void)r; (void)b; (void)n;
  if (!e) return GNSS_ERR_ARG;
  snprintf(e->err, sizeof e->err, "engine only");
  return GNSS_ERR_UNSUPPORTED;
}
