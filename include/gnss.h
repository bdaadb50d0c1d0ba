/*
 * gnss.h — C ABI of the pseudospectral Navier–Stokes + passive-scalar time step.
 *
 * One ABI, two implementations:
 *   libgnss_hip.so     the MI355X engine (hand-written gfx950 HIP kernels), the product;
 *   libgnss_oracle.so  the FP64 CPU restatement used only by tests and the CPU baseline.
 *
 * The reference (mfkiwl/gnss-1) has no plugin interface: its "operator API" is a set of
 * C++ free functions taking structs of per-GPU pointer arrays by value and returning void.
 * Each entry point below names the reference function it replaces.  Differences that are
 * deliberate:
 *   - every call returns an int status (GNSS_OK == 0) and gnss_last_error() explains a failure
 *     (the reference aborts on CUDA errors and ignores cuFFT errors, fftfuncs.cu:297-300);
 *   - configuration is a runtime struct instead of compile-time macros (dnsparams.h:9-41);
 *   - the time integrator is low-storage RK3 with exact integrating factors instead of
 *     Euler/AB2 + Crank–Nicolson (solver.cu:547-616), viscosity is nu = 1/Re everywhere,
 *     and the dealiasing mask is the 2/3 rule (dnsparams.h:65 variant of solver.cu:17-42).
 *
 * Layouts at the ABI (host buffers, all FP64, caller-owned, only touched during the call):
 *   physical  real   [x][y][z], z fastest, nx*ny*nz doubles (the reference's rows are padded to
 *                    2*(nz/2+1) for its in-place cuFFT, cudafuncs.cu:13-17; the ABI is unpadded)
 *   spectral  complex[ky][kx][kz], kz fastest, kz in [0, nz/2], interleaved (re, im) doubles —
 *                    the reference's transposed spectral order (solver.cu:25).  Spectral values
 *                    are unnormalised: forward is an unnormalised R2C, inverse carries 1/(nx ny nz)
 *                    (fftfuncs.cu:288-359, cudafuncs.cu:19-31).
 *   mask      uint8  [ky][kx][kz], 1 = kept by the dealiasing rule.
 *
 * Field ids: 0 = u, 1 = v, 2 = w, 3 + s = passive scalar s (s < n_scalars); with
 * GNSS_FLAG_COLLOID the last scalar is the colloid.
 */
#ifndef GNSS_H
#define GNSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GNSS_MAX_SCALARS 4

enum gnss_status {
  GNSS_OK = 0,
  GNSS_ERR_ARG = 1,         /* bad argument or parameter set */
  GNSS_ERR_DEVICE = 2,      /* HIP runtime / kernel launch failure */
  GNSS_ERR_ALLOC = 3,       /* allocation failure (or above the configured cap) */
  GNSS_ERR_STATE = 4,       /* call out of order (e.g. step before init) */
  GNSS_ERR_NUMERIC = 5,     /* NaN/Inf detected in the state */
  GNSS_ERR_UNSUPPORTED = 6  /* size/feature not supported by this implementation */
};

enum gnss_precision { GNSS_FP64 = 0, GNSS_FP32 = 1 };
enum gnss_space { GNSS_SPECTRAL = 0, GNSS_PHYSICAL = 1 };
enum gnss_flags {
  GNSS_FLAG_MFMA_R16 = 1,   /* FP32 form: radix-16 DFT stages on MFMA with split-FP16 operands */
  GNSS_FLAG_UNFUSED = 2,    /* engine: run the step as separate transform/pointwise kernels */
  GNSS_FLAG_COLLOID = 4     /* the last scalar is a colloid c carried by u + alpha grad(scalar 0):
                               its nonlinear term is u.grad c + alpha (lap(Z) c + grad Z . grad c),
                               Z = scalar 0 (colloidAdvectionKernel_mgpu, solver.cu:374-405, and
                               scalarAdvection, solver.cu:407-459); needs n_scalars >= 2 */
};

/* Runtime replacement of the reference's compile-time configuration (dnsparams.h:9-41). */
typedef struct gnss_params {
  int nx, ny, nz;            /* grid; each a product of 2s and 3s, even, 8..1024 */
  double lx, ly, lz;         /* box lengths (reference LX, LY, LZ) */
  double re;                 /* Reynolds number; nu = 1/re */
  int n_scalars;             /* passive scalars, 0..GNSS_MAX_SCALARS */
  double sc[GNSS_MAX_SCALARS]; /* Schmidt numbers; diffusivity of scalar s is nu/sc[s] */
  double theta;              /* jet momentum thickness (reference theta, dnsparams.h:26) */
  double jet_h;              /* jet width H (0 means 1.0); centred at y = ly/2 */
  double noise_amp;          /* rms (per component) of the divergence-free velocity noise */
  double noise_kmax;         /* noise lives on integer shells 1 <= |m| <= noise_kmax */
  uint64_t seed;             /* seed of the portable counter-based noise generator */
  double dt;                 /* fixed time step (reference dt, dnsparams.h:23) */
  int precision;             /* GNSS_FP64 or GNSS_FP32 */
  int flags;                 /* gnss_flags */
  double alpha;              /* colloid drift coefficient (reference alpha, dnsparams.h:33) */
  double theta_c;            /* colloid initial-profile thickness (reference theta_c,
                                dnsparams.h:28; 0 means theta) */
} gnss_params;

/* Integral statistics (reference statistics struct, struct_def.h:13-29; calcTurbStats_mgpu,
 * statistics.cu:2039-2156).  All sums are FP64, over the full Hermitian spectrum, divided by
 * (nx ny nz)^2 to undo the unnormalised forward transform. */
typedef struct gnss_stats_out {
  double t;                  /* simulation time */
  int64_t step;              /* steps taken */
  double ke;                 /* 1/2 <u_i u_i> */
  double vrms;               /* sqrt(<u_i u_i>/3) */
  double epsilon;            /* nu * sum k^2 |u^|^2 */
  double enstrophy;          /* 1/2 <w_i w_i> = 1/2 sum k^2 |u^|^2 */
  double lint;               /* integral length 3 pi/4 * (sum |u^|^2/(2|k|)) / ke */
  double lambda;             /* Taylor microscale sqrt(15 nu vrms^2 / epsilon) */
  double eta;                /* Kolmogorov scale (nu^3/epsilon)^(1/4) */
  double chi[GNSS_MAX_SCALARS];     /* 2 (nu/sc) sum k^2 |phi^|^2 */
  double scalar_var[GNSS_MAX_SCALARS]; /* <phi'^2> over the volume */
  double cfl;                /* dt * max(|u|/dx + |v|/dy + |w|/dz) over grid points */
  double umax;               /* max |u_i| over grid points and components */
  double divergence;         /* sqrt(sum |k.u^|^2) / sqrt(sum k^2 |u^|^2) */
} gnss_stats_out;

/* Sizes of the variable-length outputs. */
typedef struct gnss_sizes {
  int n_fields;              /* 3 + n_scalars */
  int nzh;                   /* nz/2 + 1 */
  int n_prof;                /* rows of yprof: 7 + 3 * n_scalars */
  int n_bins;                /* shells of the spectrum */
  int64_t n_physical;        /* nx*ny*nz */
  int64_t n_spectral;        /* nx*ny*(nz/2+1) complex values */
} gnss_sizes;

typedef struct gnss_ctx gnss_ctx;

/* ≈ splitData + allocate_memory + plan1dFFT/plan2dFFT (main.cu:105-113, allocate.cu:9-224,
 *   fftfuncs.cu:133-231).  Allocates the state and work buffers and builds the FFT plans. */
int gnss_create(const gnss_params* params, gnss_ctx** out);
/* ≈ deallocate_memory (allocate.cu:227-371). */
int gnss_destroy(gnss_ctx* ctx);
/* Message for the last failed call on ctx (ctx == NULL: last failure of gnss_create). */
const char* gnss_last_error(const gnss_ctx* ctx);
int gnss_get_sizes(const gnss_ctx* ctx, gnss_sizes* out);

/* ≈ initializeWaveNumbers (initialize.cu:678-707): k_i = (2 pi / L_i) * m_i with integer mode
 *   index m = i < n/2 ? i : i - n (the Nyquist index n/2 maps to -n/2), and the 2/3 mask. */
int gnss_init_wavenumbers(gnss_ctx* ctx);
/* Copies the wavenumber vectors (nx, ny, nz/2+1 doubles) and the mask ([ky][kx][kz] bytes). */
int gnss_get_wavenumbers(gnss_ctx* ctx, double* kx, double* ky, double* kz);
int gnss_get_mask(gnss_ctx* ctx, uint8_t* mask);

/* ≈ initializeJet + forwardTransform x5 + deAlias (initialize.cu:614-641, main.cu:139-156):
 *   tanh jet U(y) = 0.5 (1 + tanh((H/2 - |y - ly/2|)/(2 theta))) on u and on every scalar,
 *   plus seeded divergence-free Gaussian noise on u, v, w, built in FP64 in spectral space and
 *   dealiased.  Resets time and step counters. */
int gnss_init_jet(gnss_ctx* ctx);

/* ≈ forwardTransform (fftfuncs.cu:288-320): spectral state of `field` <- unnormalised R2C of
 *   its physical buffer.  No dealiasing. */
int gnss_forward(gnss_ctx* ctx, int field);
/* ≈ inverseTransform (fftfuncs.cu:322-359): physical buffer of `field` <- C2R of its spectral
 *   state, scaled by 1/(nx ny nz).  Out of place: the spectral state is not modified. */
int gnss_inverse(gnss_ctx* ctx, int field);
/* ≈ deAlias (solver.cu:44-66), applied to every field of the state. */
int gnss_dealias(gnss_ctx* ctx);

/* ≈ solver_ps (solver.cu:700-765), called nsteps times: one low-storage RK3 step with exact
 *   viscous/diffusive integrating factors per call. */
int gnss_step(gnss_ctx* ctx, int nsteps);

/* ≈ calcTurbStats_mgpu (statistics.cu:2039-2156).  yprof: n_prof x ny doubles (rows U, V, W,
 *   <u'u'>, <v'v'>, <w'w'>, <u'v'>, then per scalar <phi>, <phi'phi'>, <v'phi'>), x-z plane
 *   averages per y.  spectrum: (1 + n_scalars) x n_bins doubles, shell sums over
 *   bin = nint(|m|) of 1/2 |u^|^2 (then 1/2 |phi^|^2 per scalar).  Either pointer may be NULL. */
int gnss_stats(gnss_ctx* ctx, gnss_stats_out* out, double* yprof, double* spectrum);
/* No counterpart in the reference: gnss_stats for callers that cannot allocate (code under HIP
 *   graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same outputs, bitwise
 *   equal to gnss_stats'; the partial sums live in a fixed 40 MiB device scratch per GPU (held until
 *   the process ends, its address resolved by gnss_create) instead of the context's buffers, and the
 *   call allocates nothing on the device or the host.  A grid whose partials need more scratch
 *   (only very anisotropic boxes: the shell count grows with the longest box side over the shortest)
 *   gets GNSS_ERR_ALLOC.  Every call on one GPU shares its scratch: calls that overlap on one GPU
 *   (from several threads or contexts) race. */
int gnss_stats_noalloc(gnss_ctx* ctx, gnss_stats_out* out, double* yprof, double* spectrum);

/* ≈ save3Dfields / importVelocity host copies (iofuncs.cu:329-432, 499-535).
 *   space == GNSS_SPECTRAL: the spectral state; GNSS_PHYSICAL: the field's physical buffer. */
int gnss_get_field(gnss_ctx* ctx, int field, int space, double* host);
int gnss_set_field(gnss_ctx* ctx, int field, int space, const double* host);
int gnss_get_time(const gnss_ctx* ctx, double* t, int64_t* step);

/* ≈ save3Dfields / write3Dfields_mgpu (iofuncs.cu:329-432): writes every field in physical space
 *   to <dir>/<name>.<iter> in the reference's raw format — an 8-byte header holding the byte count
 *   as a double, then nx*ny*nz doubles in [x][y][z] order — with names u, v, w, z (scalar 0),
 *   z1, z2, ... for further scalars and c for the colloid (GNSS_FLAG_COLLOID), plus
 *   <dir>/gnss_state.<iter> ("step <n> t <time>").  The
 *   RK3 integrator is self-starting, so fields + clock are the whole checkpoint. */
int gnss_save_fields(gnss_ctx* ctx, const char* dir, int iter);
/* ≈ loadData / importVelocity / importScalar (iofuncs.cu:462-535): reads the same files (the size
 *   header must equal nx*ny*nz*8, as iofuncs.cu:475-479 checks), forward-transforms and dealiases
 *   them, and restores the clock from gnss_state.<iter> (else step = iter, t = iter * dt). */
int gnss_load_fields(gnss_ctx* ctx, const char* dir, int iter);

/* Diagnostics (physical space, FP64 results; the state is not changed). */
#define GNSS_FIELD_VORTICITY (-1)  /* |omega| = |i k x u^| in physical space (VectorMagnitude,
                                      statistics.cu:2005-2037) */
/* ≈ writexyfields / writexzfields / writeyzfields (iofuncs.cu:187-260): one physical plane of
 *   field f (0 <= f < 3 + n_scalars, or GNSS_FIELD_VORTICITY).  axis 2: z = index, out [x][y]
 *   (nx*ny doubles); axis 1: y = index, out [x][z]; axis 0: x = index, out [y][z]. */
int gnss_get_slice(gnss_ctx* ctx, int field, int axis, int index, double* out);
/* ≈ save2Dfields (iofuncs.cu:262-327): writes <dir>/<name>_xy.<iter> for every field — the
 *   z = nz/2 plane, an 8-byte header holding the byte count as a double, then [x][y] doubles —
 *   with the names of gnss_save_fields (the colloid is "c", as in the reference). */
int gnss_save_slices(gnss_ctx* ctx, const char* dir, int iter);
/* ≈ calcSurfaceArea_mgpu / surfaceArea_kernel_mgpu (statistics.cu:1310-1492; the reference's
 *   scalar and enstrophy interface areas, statistics.cu:2110, 2132): area of the iso-surface
 *   F = iso of field f (or GNSS_FIELD_VORTICITY) as  sum over grid points of
 *   -(grad F . grad chi) / |grad F| dx dy dz,  chi = (F <= iso), with second-order central
 *   differences on the periodic grid (a point where either gradient vanishes adds nothing).
 *   The reference's halo exchange between slabs is the periodic wrap here (one GPU). */
int gnss_surface_area(gnss_ctx* ctx, int field, double iso, double* area);

/* Per-kernel timing with HIP events on the context's stream (engine only; the oracle returns
 *   GNSS_ERR_UNSUPPORTED).  The report is a JSON object: kernel -> {launches, ms, bytes}. */
int gnss_timing_enable(gnss_ctx* ctx, int enable);
int gnss_timing_report(gnss_ctx* ctx, char* buf, int len);
/* Times `reps` standalone forward+inverse 3D transform pairs of field 0 on the context's stream
 *   (engine only).  ms_fwd/ms_inv: average milliseconds per 3D R2C / C2R. */
int gnss_bench_fft(gnss_ctx* ctx, int reps, double* ms_fwd, double* ms_inv);
/* Times a plain device-to-device copy of one spectral work buffer into another (a vectorised
 *   grid-stride kernel on the context's stream): the practical HBM bandwidth the passes are
 *   measured against.  *ms = mean time per copy, *bytes = bytes read + written per copy.  No
 *   reference counterpart (measurement); engine only. */
int gnss_bench_copy(gnss_ctx* ctx, int reps, double* ms, double* bytes);

/* ≈ the plan selection of plan1dFFT / plan2dFFT (fftfuncs.cu:133-231), which the reference leaves
 *   to cuFFT's planner.  The engine times its own compiled variants of the step's passes (column
 *   tile width of the strided x/y passes, radix plan of the fused z pass, x passes fused with the
 *   spectral prologue/update or not) `reps` times each on this context's buffers, keeps the
 *   fastest and writes a JSON report (candidates and choice) to `report` (may be NULL).  The
 *   state is snapshotted and restored around the timed substages.  reps == 0 only reports the
 *   current choice.  gnss_create applies the choices stored in $GNSS_TUNING or
 *   tuning/gnss_tuning.json beside the library. */
int gnss_autotune(gnss_ctx* ctx, int reps, char* report, int len);
/* Forces a variant, e.g. "cx=8 cy=16 zf=1 xf=0" (tests compare variants against each other). */
int gnss_set_tuning(gnss_ctx* ctx, const char* spec);

/* Batched ensemble of independent realisations of uneven sizes (no reference counterpart:
 *   the reference runs one realisation per process).  Members of one grid (same sizes, scalars,
 *   precision, flags, box and dt) share one batched engine whose passes cover all of them in one
 *   launch (GNSS_ENSEMBLE_BATCH=0 in the environment: one engine per member); one RK3 step of the
 *   whole ensemble is one hipGraph launch.  Member indices follow the order given here. */
typedef struct gnss_ens gnss_ens;
/* Members are checked as gnss_create checks them; GNSS_FLAG_UNFUSED members are refused
 *   (GNSS_ERR_UNSUPPORTED): the captured step cannot allocate their physical buffers. */
int gnss_ensemble_create(int n, const gnss_params* members, gnss_ens** out);
int gnss_ensemble_destroy(gnss_ens* ens);
const char* gnss_ensemble_last_error(const gnss_ens* ens);
int gnss_ensemble_init(gnss_ens* ens);
int gnss_ensemble_step(gnss_ens* ens, int nsteps);
int gnss_ensemble_stats(gnss_ens* ens, int member, gnss_stats_out* out, double* yprof,
                        double* spectrum);
/* Member `member`'s field as gnss_get_field gives it; GNSS_PHYSICAL inverse-transforms that
 *   member's current spectral state first (an ensemble has no gnss_inverse of its own). */
int gnss_ensemble_get_field(gnss_ens* ens, int member, int field, int space, double* host);
/* ≈ plan1dFFT / plan2dFFT's planner choice (fftfuncs.cu:133-231), for the batched engines: times
 *   each batched engine's pass variants over whole substages of all its members, keeps the
 *   fastest (the "<key>b" entries of the tuning table) and captures the step graph again.  Needs
 *   gnss_ensemble_init first; the members' states are restored.  report (may be NULL) receives a
 *   JSON list with one object per batched engine. */
int gnss_ensemble_autotune(gnss_ens* ens, int reps, char* report, int len);

#ifdef __cplusplus
}
#endif
#endif /* GNSS_H */
