/*
 * cbs.h -- C ABI of the MI355X convergent-Born-series (CBS) engine.
 *
 * The library replaces the device side of FrequencyMaxwell.jl's solve path:
 * the dyadic Green operator, the CBS linear operator and preconditioner, the
 * curl used for H, and the Krylov loop that LinearSolve.jl runs.  Two routes:
 * cbs_solve and its relatives take the potential and source term the caller
 * built (the reference's host stages stay with the caller); cbs_solve_scene
 * and its relatives take the permittivity and plane waves and build the
 * potential, incident field and source term, and add the incident field and
 * crop the result, on the GPU.
 *
 * Conventions
 *   - Complex numbers are interleaved (re, im) doubles: a Julia
 *     Ptr{ComplexF64} passes as `double *`.
 *   - Arrays are Julia column-major.  A scalar grid array (the potential) has
 *     element (i,j,k) at  (k*Ny + j)*Nx + i;  a field array (nx,ny,nz,3) has
 *     element (i,j,k,c) at ((c*Nz + k)*Ny + j)*Nx + i  (0-based indices).
 *   - Complex scalars pass as separate re / im doubles.
 *   - Every entry point returns a cbs_status (0 = success).  No C++ exception
 *     crosses the ABI.  cbs_last_error() returns a thread-local message for
 *     the last failing call on the calling thread.
 *   - Non-convergence is NOT an error: the call returns CBS_OK with
 *     stats->converged = 0 (the reference only warns,
 *     src/solvers/convergent_born.jl:753-755).
 *   - The caller owns every host buffer; the library copies in and out and
 *     keeps no host pointer after a call returns.  The library owns every
 *     device allocation; *_destroy frees it.
 *   - A cbs_problem is used by one host thread at a time.  A cbs_ctx may be
 *     shared by threads for the serving calls (cbs_submit / cbs_poll /
 *     cbs_wait).
 */
#ifndef CBS_MI355X_H
#define CBS_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t cbs_status;
typedef struct cbs_ctx cbs_ctx;
typedef struct cbs_problem cbs_problem;

enum {
  CBS_OK = 0,
  CBS_EINVAL = -1,    /* bad argument: maps to Julia ArgumentError */
  CBS_ENOMEM = -2,    /* device allocation failed / HBM budget refused */
  CBS_EDEVICE = -3,   /* HIP runtime error */
  CBS_ENOTFOUND = -4, /* unknown ticket */
  CBS_EBUSY = -5,     /* serving queue full */
  CBS_ECANCELED = -6, /* request was cancelled */
  CBS_ENUMERIC = -7   /* NaN / Inf in an input, or a residual that became NaN,
                         Inf or overflowed during the iteration */
};

/* Working precision of the iteration. */
enum {
  CBS_FP64 = 0,        /* complex FP64 everywhere ("plain FP64" path) */
  CBS_FP32 = 1,        /* complex FP32 bulk, FP64 twiddles/ramps/norms */
  CBS_FP32_REFINE = 2  /* FP32 bulk inside an FP64 iterative-refinement loop */
};

/* Iterative method. */
enum {
  CBS_BORN = 0,     /* CBS / preconditioned Richardson: x <- x + gamma (b - A x) */
  CBS_BICGSTAB = 1, /* left-preconditioned BiCGStab (KrylovJL_BICGSTAB) */
  CBS_GMRES = 2     /* left-preconditioned GMRES (KrylovJL_GMRES, the default) */
};

/* Stopping rule on the preconditioned residual r = gamma (b - A x). */
enum {
  CBS_STOP_RHS = 0,   /* ||r|| <= tol * ||gamma b||                       */
  CBS_STOP_KRYLOV = 1 /* ||r|| <= abstol + tol * ||r0|| (Krylov.jl rule)  */
};

/* Problem description: the padded FFT grid and the DyadicGreen parameters
 * (src/core/dyadic_green.jl:67-92, src/solvers/convergent_born.jl:619-675). */
typedef struct {
  int64_t n[3];        /* padded grid Nx, Ny, Nz                           */
  double res[3];       /* voxel size dx, dy, dz (m)                        */
  double shift[3];     /* subpixel shift per axis (0 periodic, 0.25 absorbing) */
  int32_t averaging;   /* 1: G_avg = (G(+s) + G(-s))/2 (any axis absorbing) */
  double k2_re;        /* k_square = k_bg^2 + i eps: real part             */
  double k2_im;        /*                           imaginary part (= eps) */
  double eps_imag;     /* eps of the preconditioner gamma = i V / eps      */
  int32_t precision;   /* CBS_FP64 | CBS_FP32 | CBS_FP32_REFINE            */
} cbs_problem_desc;

typedef struct {
  int32_t method;       /* CBS_BORN | CBS_BICGSTAB | CBS_GMRES              */
  int32_t precision;    /* CBS_FP64 | CBS_FP32 | CBS_FP32_REFINE            */
  int32_t stop_rule;    /* CBS_STOP_RHS | CBS_STOP_KRYLOV                   */
  int32_t maxiters;     /* bound on iterations as the method counts them:
                         * Born sweeps; BiCGStab iterations (two operator
                         * applications each) and GMRES Arnoldi steps, as
                         * Krylov.jl's itmax (LinearSolve maxiters);
                         * FP32_REFINE: inner iterations plus FP64 residual
                         * evaluations */
  double tol;           /* relative tolerance (reltol)                     */
  double abstol;        /* absolute tolerance (CBS_STOP_KRYLOV only)       */
  int32_t gmres_restart;/* 0 = no restart (Krylov.jl default).  GMRES keeps at
                         * most 400 basis vectors: restart 0 needs maxiters
                         * <= 400, restart must be <= 400 (CBS_EINVAL) */
  int32_t max_refine;   /* FP32_REFINE: bound on FP64 refinement steps     */
  double inner_tol;     /* FP32_REFINE: inner relative tolerance (0 = auto)*/
  int32_t check_every;  /* host convergence check cadence (>=1)            */
  int32_t use_graph;    /* capture check_every iterations as a hipGraph    */
  int32_t x0_zero;      /* 1: x0 = 0; 0: x0 = rhs (reference, :727-728)    */
  int32_t compute_h;    /* cbs_solve: also return H_scat                   */
  double h_scale_re;    /* H = h_scale * curl(E): -i lambda/(2 pi 377)     */
  double h_scale_im;
  int32_t fuse_sweeps;  /* Born: 1 (default) runs each sweep's x inverse and
                         * the next sweep's x forward as one pass (same
                         * results, fewer HBM bytes); 0 runs the five
                         * separate sweeps.  Served requests always fuse. */
} cbs_solve_opts;

typedef struct {
  int32_t iters;          /* iterations as the method counts them (the
                           * reference's solver.iteration_count = LinearSolve
                           * sol.iters, convergent_born.jl:761-763): Born
                           * sweeps; BiCGStab iterations; GMRES Arnoldi steps
                           * (the warm-start residual not counted);
                           * FP32_REFINE: inner iterations plus the FP64
                           * residual evaluations                            */
  int32_t iters_fp64;     /* operator applications in FP64                   */
  int32_t refine_steps;   /* FP64 refinement steps (FP32_REFINE)             */
  int32_t converged;      /* 1 when the stopping rule was met                */
  double resid;           /* ||gamma (b - A x)|| / ||gamma b|| (FP64) of the returned x */
  double resid_true_fp64; /* same, recomputed in FP64 after the solve (0 if not run) */
  double time_ms;         /* device time of the solve                        */
  double bytes_moved;     /* algorithmic HBM bytes of the solve              */
  int64_t voxel_iters;    /* applications * N_vox                            */
  int32_t applications;   /* operator applications (Born: one per sweep;
                           * BiCGStab: two per iteration plus the warm start;
                           * GMRES: one per step plus one per cycle start,
                           * plus one in a cycle that meets the tolerance:
                           * the delayed re-orthogonalisation learns it one
                           * step late)                                      */
} cbs_solve_stats;

/* ---- context ------------------------------------------------------------ */
cbs_status cbs_ctx_create(cbs_ctx **out, int32_t device, uint64_t hbm_budget_bytes);
cbs_status cbs_ctx_destroy(cbs_ctx *ctx);
const char *cbs_last_error(void);
const char *cbs_version(void);
void cbs_default_opts(cbs_solve_opts *opts);

/* ---- problem ------------------------------------------------------------ */
/* Upload the padded, attenuated potential V (complex, Nx*Ny*Nz) -- the
 * `potential_device` of src/solvers/convergent_born.jl:650 -- and build the
 * Green operator of src/core/dyadic_green.jl:67-92.  Limits (CBS_EINVAL):
 * every axis 1..4096, and fewer than 2^32 spectrum elements per shift
 * pipeline (3 components of the padded-stride grid, about 1.4e9 voxels). */
cbs_status cbs_problem_create(cbs_ctx *ctx, const cbs_problem_desc *desc, const double *V,
                              cbs_problem **out);
cbs_status cbs_problem_destroy(cbs_problem *p);

/* conv!(::DyadicGreen, src): field <- G_avg field, in place.
 * Replaces src/core/dyadic_green.jl:127-151. */
cbs_status cbs_green_apply(cbs_problem *p, double *field_inout);

/* mul!(y, ::CBSLinearOperator, x, alpha, beta): y <- alpha (x - G_avg(V x)) + beta y.
 * Replaces src/solvers/convergent_born.jl:482-512. */
cbs_status cbs_operator_apply(cbs_problem *p, double *y, const double *x, double a_re,
                              double a_im, double b_re, double b_im);

/* Allocation-free variant of cbs_operator_apply -- no counterpart in the
 * reference.  For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * same operator and kernels on DEVICE arrays x and y (3*Nx*Ny*Nz complex in
 * the problem's working precision: complex FP64 for CBS_FP64 problems,
 * complex FP32 otherwise); the spectrum workspace and the launch tables live
 * in a device array of CBS_NOALLOC_SCRATCH_BYTES that the library holds on
 * every GPU it runs on until the process ends, so the call allocates nothing.
 * A grid whose workspace does not fit (FP64 with averaging: about 2.8 million
 * voxels; FP32 or periodic: more) is refused with CBS_EINVAL.  Calls on one GPU
 * share its scratch array: calls that overlap on the same GPU race and must be
 * serialised by the caller.  Runs on the context's stream and returns when the
 * result is in y. */
#define CBS_NOALLOC_SCRATCH_BYTES 268435456ULL /* 256 MiB */
cbs_status cbs_operator_apply_noalloc(cbs_problem *p, void *y, const void *x, double a_re,
                                      double a_im, double b_re, double b_im);

/* ldiv!(y, ::CBSPreconditioner, x): y <- (i/eps) V x.
 * Replaces src/solvers/convergent_born.jl:398-410. */
cbs_status cbs_precond_apply(cbs_problem *p, double *y, const double *x);

/* conv(::Curl, E) scaled: field <- scale * curl(field).  With scale =
 * -i lambda/(2 pi 377) this is _compute_magnetic_field,
 * src/solvers/convergent_born.jl:784-800 / src/core/curl.jl:55-116. */
cbs_status cbs_curl_apply(cbs_problem *p, double *field_inout, double scale_re,
                          double scale_im);

/* _solve_cbs_scattering: rhs = G_avg S, x0 = rhs, iterate to tolerance, then
 * H = h_scale * curl(x).  Replaces src/solvers/convergent_born.jl:703-774
 * (the LinearSolve.solve call at :750 and everything it drives).
 * S, E_scat, H_scat: padded fields (Nx,Ny,Nz,3).  H_scat may be NULL. */
cbs_status cbs_solve(cbs_problem *p, const double *S, double *E_scat, double *H_scat,
                     const cbs_solve_opts *opts, cbs_solve_stats *stats);

/* Device-resident form of cbs_solve, for callers that solve repeatedly or
 * time the iteration alone: set the source (host -> device), run, fetch. */
cbs_status cbs_problem_set_source(cbs_problem *p, const double *S);
cbs_status cbs_problem_run(cbs_problem *p, const cbs_solve_opts *opts, cbs_solve_stats *stats);
cbs_status cbs_problem_get_fields(cbs_problem *p, double *E_scat, double *H_scat,
                                  double h_re, double h_im);

/* Many right-hand sides of one problem (illumination angles x polarisations,
 * one `solve` per source): every sweep of the iteration is one launch over all
 * of them.  S[i], E[i] are padded fields; H may be NULL. */
cbs_status cbs_solve_multi(cbs_problem *p, int32_t nrhs, const double *const *S, double **E,
                           double **H, const cbs_solve_opts *opts, cbs_solve_stats *stats);

/* Many problems of possibly different grids, each with one right-hand side:
 * every sweep is one launch over all of them (uneven-size work table). */
cbs_status cbs_solve_batch(cbs_ctx *ctx, int32_t n, cbs_problem *const *probs,
                           const double *const *S, double **E, double **H,
                           const cbs_solve_opts *opts, cbs_solve_stats *stats);

/* Measurement aid -- no counterpart in the reference; for benchmarks that
 * need per-kernel durations on the library's own stream.  Runs one Born
 * sweep of the resident source term (cbs_problem_set_source first) in the
 * given precision, then launches each of the six sweep kernels `reps` times
 * back to back between two HIP events (the problem's work buffers are left
 * overwritten; a later solve starts afresh).
 * ms[k] (6 entries): average duration of one launch of kernel k (0 x forward
 * + V x, 1 y forward, 2 z + Green, 3 y inverse, 4 x inverse + Born update,
 * 5 the fused x pass of a Born sweep: x inverse + Born update + the next
 * sweep's V x and x forward); a fused Born sweep is kernels 1, 2, 3, 5.
 * bytes[k] (6 entries): algorithmic HBM bytes of one launch of kernel k. */
cbs_status cbs_time_sweeps(cbs_problem *p, int32_t precision, int32_t reps, double *ms,
                           double *bytes);

/* Measurement aid -- no counterpart in the reference: a plain device copy of
 * `bytes` bytes (each thread loads four non-temporal 16-byte values, then
 * stores them), timed with HIP events over `reps` launches; *ms is the average
 * duration of one launch.  It gives the achievable HBM bandwidth of the box
 * and the calibration kernel for HBM byte counters. */
cbs_status cbs_stream_copy(cbs_ctx *ctx, uint64_t bytes, int32_t reps, double *ms);

/* ---- device-side setup and teardown (scene) -----------------------------
 * The reference's whole solve(solver, sources, medium)
 * (src/solvers/convergent_born.jl:565-600) with its host stages moved to the
 * GPU: _initialize_solver! (:619-675: potential, padding, eps, attenuation
 * :893-944), generate_incident_field (src/sources/plane_wave.jl:198-268,
 * coherent sum field_generation.jl:42-78), the source term (:585-586), the
 * magnetic field (:784-800) and E_inc/H_inc add + crop_to_ROI
 * (src/fields/electromagnetic_field.jl:252-303).  The caller keeps the
 * scalar boundary-condition and plane-wave algebra (padding_pixels, window
 * values, unit vectors), which is O(1) host work. */
typedef struct {
  int64_t grid[3];          /* ROI grid (gx, gy, gz): the permittivity array     */
  double res[3];            /* voxel size per axis                               */
  int32_t pad[3];           /* padding pixels on each side (padding_pixels)      */
  int32_t att_len[3];       /* attenuation window length L per axis (0: none)    */
  const double *window[3];  /* sharpness-blended window w*sharp + (1 - sharp),
                               att_len[d] values each (boundary_conditions.jl
                               :632-658 and convergent_born.jl:911-937)          */
  double shift[3];          /* Green subpixel shift per axis                     */
  int32_t averaging;        /* +/- shift averaging (any axis absorbing)          */
  double eps_bg;            /* background permittivity                           */
  double wavelength;        /* vacuum wavelength                                 */
  int32_t iterations_max;   /* solver iterations_max (< 0: the Bornmax rule)     */
  int32_t precision;        /* working precision of the Green operator          */
} cbs_scene_desc;

typedef struct {
  double k_vec[3];          /* k_bg * unit propagation direction                 */
  double e_perp[6];         /* transverse unit E polarisation (x, y, z complex)  */
  double h_perp[6];         /* k_hat x e_perp / (377 / sqrt(eps_bg))             */
  double amplitude;         /* A                                                 */
  double phase;             /* phi: E = A exp(i (k.r + phi)) e_perp              */
} cbs_plane_wave;

typedef struct {
  int64_t n[3];             /* padded grid                                       */
  double eps_imag;          /* max(2^-20, 1.01 max |V0|)                         */
  double k2_re, k2_im;      /* k_square = k_bg^2 + i eps_imag                    */
  int32_t bornmax;          /* Bornmax (convergent_born.jl:653-660)              */
  int32_t maxiters;         /* max(2 Bornmax, 100)                               */
} cbs_scene_info;

/* Setup only: the padded potential V (Nx*Ny*Nz complex) and source term S
 * (Nx*Ny*Nz*3 complex) as the GPU builds them; either may be NULL. */
cbs_status cbs_scene_build(cbs_ctx *ctx, const cbs_scene_desc *scene, const double *eps,
                           int32_t nsrc, const cbs_plane_wave *srcs, double *V, double *S,
                           cbs_scene_info *info);
/* The whole solve on the device.  eps: ROI permittivity (gx*gy*gz complex).
 * E_tot, H_tot: total fields cropped to the ROI (gx*gy*gz*3 complex); H_tot
 * may be NULL.  opts->maxiters <= 0 takes info.maxiters (the reference rule);
 * opts->h_scale is replaced by -i wavelength / (2 pi 377). */
cbs_status cbs_solve_scene(cbs_ctx *ctx, const cbs_scene_desc *scene, const double *eps,
                           int32_t nsrc, const cbs_plane_wave *srcs, const cbs_solve_opts *opts,
                           double *E_tot, double *H_tot, cbs_solve_stats *stats,
                           cbs_scene_info *info);

/* Many right-hand sides of one scene (illumination angles x polarisations):
 * RHS i is the coherent sum of nsrc[i] plane waves, taken in order from
 * srcs.  One potential, one source term per RHS built on the device, every
 * sweep one launch over all of them; each RHS stops on its own.  E_tot[i] and
 * H_tot[i] are ROI-cropped totals; H_tot (or any entry, or E_tot[i]) may be
 * NULL to skip that download.  stats may be NULL or hold nrhs entries. */
cbs_status cbs_solve_scene_multi(cbs_ctx *ctx, const cbs_scene_desc *scene, const double *eps,
                                 int32_t nrhs, const int32_t *nsrc, const cbs_plane_wave *srcs,
                                 const cbs_solve_opts *opts, double **E_tot, double **H_tot,
                                 cbs_solve_stats *stats, cbs_scene_info *info);

/* ---- fields kept on the device ---------------------------------------
 * The reference's field analysis (src/fields/electromagnetic_field.jl:97-237)
 * on fields resident in HBM, so a caller can take a scalar, an intensity map
 * or a plane instead of two full 3-D fields.  Arrays are (nx, ny, nz, 3)
 * complex column-major like every field here. */
typedef struct cbs_field cbs_field;
/* Upload E and H (H may be NULL: no H-based analysis then). */
cbs_status cbs_field_create(cbs_ctx *ctx, const int64_t n[3], const double res[3],
                            double wavelength, const double *E, const double *H,
                            cbs_field **out);
cbs_status cbs_field_destroy(cbs_field *f);
/* Copy the fields back (either pointer may be NULL). */
cbs_status cbs_field_get(cbs_field *f, double *E, double *H);
/* field_energy (:112-127): (eps0 sum|E|^2 + mu0 sum|H|^2) / 2 * voxel volume. */
cbs_status cbs_field_energy(cbs_field *f, double *energy);
/* field_intensity (:162-180): component 0 = total, 1/2/3 = x/y/z; out: nx*ny*nz doubles. */
cbs_status cbs_field_intensity(cbs_field *f, int32_t component, double *out);
/* poynting_vector (:135-154): S = E x conj(H); out: nx*ny*nz*3 complex. */
cbs_status cbs_field_poynting(cbs_field *f, double *S);
/* extract_plane (:192-235): plane_axis 1..3, plane_index 1-based; E_plane and
 * H_plane (either may be NULL) have the field's shape with that axis of length 1. */
cbs_status cbs_field_extract_plane(cbs_field *f, int32_t plane_axis, int64_t plane_index,
                                   double *E_plane, double *H_plane);
/* cbs_solve_scene leaving the cropped total fields on the device. */
cbs_status cbs_solve_scene_field(cbs_ctx *ctx, const cbs_scene_desc *scene, const double *eps,
                                 int32_t nsrc, const cbs_plane_wave *srcs,
                                 const cbs_solve_opts *opts, cbs_field **out,
                                 cbs_solve_stats *stats, cbs_scene_info *info);

/* ---- serving: continuous batching with HBM-footprint admission ----------
 * No counterpart in the reference (it solves one problem per call).  A
 * server thread owns a HIP stream; admitted requests share every sweep
 * launch (one work table over all of them, grids may differ), join at the
 * next chunk boundary when a slot and HBM are free, and leave as soon as
 * their own stopping rule is met.  Each request's arithmetic is independent
 * of the others': its result is the same alone and in any mix.
 * Supported per request: method CBS_BORN or CBS_BICGSTAB, precision CBS_FP64
 * or CBS_FP32_REFINE. */
typedef struct {
  cbs_problem_desc desc;  /* grid and Green parameters                       */
  const double *V;        /* potential (Nx*Ny*Nz complex), copied at submit  */
  const double *S;        /* source term (Nx*Ny*Nz*3 complex), copied at submit */
  cbs_solve_opts opts;    /* per-request tolerance / method / precision; H is
                             computed with opts.h_scale when opts.compute_h  */
} cbs_request;

/* Bytes of HBM a request of this description holds while resident. */
cbs_status cbs_request_footprint(const cbs_problem_desc *desc, const cbs_solve_opts *opts,
                                 uint64_t *bytes);
/* Start the server thread; at most max_active requests being loaded or
 * solved at once (a solved request's download does not hold a slot).
 * Admission also respects the context's HBM budget (cbs_ctx_create). */
cbs_status cbs_server_start(cbs_ctx *ctx, int32_t max_active);
/* Stop the thread; pending and running requests end with CBS_ECANCELED. */
cbs_status cbs_server_stop(cbs_ctx *ctx);
/* How the resident requests share the GPU, chunk by chunk (after
 * cbs_server_start; the default is CBS_SERVE_LEAD with kappa 0.2).
 * CBS_SERVE_SHARE: every resident request gets an equal share of each chunk's
 * GPU time, in shared launches (good median latency, but the largest requests
 * stretch with the load).  CBS_SERVE_LEAD: each chunk runs one request, the
 * oldest whose estimated work (padded voxels^(4/3) x log10(1/tol)) is at most
 * kappa times the mean of the requests submitted so far, else the oldest:
 * first come first served with small requests passing, which keeps the tail
 * latency near one-at-a-time service and the median below it.  Admission,
 * transfers, setup and the FP64 residual steps of every resident request
 * overlap in both. */
enum { CBS_SERVE_SHARE = 0, CBS_SERVE_LEAD = 1 };
cbs_status cbs_server_policy(cbs_ctx *ctx, int32_t policy, double kappa);
/* Queue a request (copies V and S); returns immediately with a ticket.
 * CBS_ENOMEM if the request can never fit the budget. */
cbs_status cbs_submit(cbs_ctx *ctx, const cbs_request *req, uint64_t *ticket);
/* Queue a scene (device-side setup and teardown, see cbs_solve_scene): eps is
 * the ROI permittivity, copied at submit; cbs_wait then returns the total
 * fields cropped to the ROI (gx*gy*gz*3 complex).  opts->maxiters <= 0 takes
 * the reference budget max(2 Bornmax, 100); H uses -i wavelength/(2 pi 377). */
cbs_status cbs_submit_scene(cbs_ctx *ctx, const cbs_scene_desc *scene, const double *eps,
                            int32_t nsrc, const cbs_plane_wave *srcs, const cbs_solve_opts *opts,
                            uint64_t *ticket);
/* done = 1 when the result is ready (stats filled), 0 while queued/running. */
cbs_status cbs_poll(cbs_ctx *ctx, uint64_t ticket, int32_t *done, cbs_solve_stats *stats);
/* Block until the request ends; copy E_scat (and H_scat if computed; either
 * may be NULL) and release the ticket. */
cbs_status cbs_wait(cbs_ctx *ctx, uint64_t ticket, double *E_scat, double *H_scat,
                    cbs_solve_stats *stats);
cbs_status cbs_cancel(cbs_ctx *ctx, uint64_t ticket);

#ifdef __cplusplus
}
#endif

#endif /* CBS_MI355X_H */
