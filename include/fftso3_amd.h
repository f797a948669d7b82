/*
 * fftso3_amd.h -- C ABI of the MI355X (gfx950) SO(3) / S^2 harmonic-transform library.
 *
 * The reference (fdcl-gwu/FFTSO3) is a C++ class library with no C ABI; the entry points below
 * are what its class methods bind to (the C++ class layer in fftso3_amd/cpp/ calls them):
 *
 *   fftso3_plan_create        <- fdcl::FFTSO3_complex(int l_max) / init(l_max)
 *                                (include/fdcl_FFTSO3.hpp:45-53, src/fdcl_FFTSO3.cpp:3-14);
 *                                fdcl::FFTSO3_real (include/fdcl_FFTSO3.hpp:193-202);
 *                                fdcl::FFTS2_complex / FFTS2_real (include/fdcl_FFTS2.hpp:37-45,117-124)
 *   fftso3_forward            <- forward_transform(std::function) (include/fdcl_FFTSO3.hpp:76,84,232,240;
 *                                include/fdcl_FFTS2.hpp:59,136; src/fdcl_FFTSO3.cpp:654-850;
 *                                src/fdcl_FFTS2.cpp:125-192,424-436), taking the sampled grid
 *   fftso3_inverse            <- inverse_transform(F, point) (include/fdcl_FFTSO3.hpp:90,96,246,252;
 *                                include/fdcl_FFTS2.hpp:65,142; src/fdcl_FFTSO3.cpp:498-531,1141-1157;
 *                                src/fdcl_FFTS2.cpp:194-230,500-512), evaluated on the whole grid
 *   fftso3_plan_destroy       <- ~FFTSO3_complex() (include/fdcl_FFTSO3.hpp:50)
 *
 * Conventions (identical to the reference):
 *   B = l_max + 1, N = 2B; alpha_j = gamma_j = j pi / B, beta_k = (2k+1) pi / (4B).
 *   SO(3) coefficients: F(l,m,n) at off(l) + (m+l) + (n+l)(2l+1), off(l) = l(4l^2-1)/3
 *     (the column-major per-l blocks of FFTSO3_matrix, src/fdcl_FFTSO3_matrix.cpp:43-47);
 *     B(4B^2-1)/3 coefficients per function.
 *   S^2 coefficients: F(l,m) at l^2 + m + l (src/fdcl_FFTS2_matrix.cpp:43-47); B^2 per function.
 *   SO(3) grid: [k][j1][j2] = f(alpha_j1, beta_k, gamma_j2), N^3 values per function.
 *   S^2 grid:   [k][j]      = f(theta_k, phi_j),            N^2 values per function.
 *   Complex values are interleaved (re, im).  Functions of a batch are packed back to back in
 *   batch order (uneven plans: each with its own N), both for grids and for coefficients.
 *   Element type: FFTSO3_FP64 plans use double / complex double; every other precision uses
 *   float / complex float in memory.
 *
 * Ownership and threading: the caller owns the device grids and coefficient arrays; the plan owns
 * its tables and workspaces.  A plan is stream-ordered and may be used by one host thread at a
 * time; distinct plans are independent.  Entry points never abort: they return a status code and
 * fftso3_last_error() describes the last failure of the calling thread.
 */
#ifndef FFTSO3_AMD_H
#define FFTSO3_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fftso3_plan_s *fftso3_plan_t;

enum fftso3_status {
    FFTSO3_OK = 0,
    FFTSO3_EINVAL = 1,        /* bad argument (bandlimit, count, null pointer, kind ...) */
    FFTSO3_ENOMEM = 2,        /* device or host allocation failed */
    FFTSO3_EHIP = 3,          /* a HIP runtime call or kernel launch failed */
    FFTSO3_EUNSUPPORTED = 4   /* valid request this build does not implement */
};

enum fftso3_domain { FFTSO3_DOMAIN_SO3 = 0, FFTSO3_DOMAIN_S2 = 1 };
enum fftso3_kind { FFTSO3_COMPLEX = 0, FFTSO3_REAL = 1 };

/* Arithmetic of the beta contraction (the alpha/gamma FFTs run in the storage type; the
 * Wigner-d / Legendre recursion, quadrature weights and normalisations are always FP64). */
enum fftso3_precision {
    FFTSO3_FP64 = 0,     /* double storage, FP64 everywhere: the exact drop-in */
    FFTSO3_FP32 = 1,     /* float storage, FP32 contraction */
    FFTSO3_BF16X3 = 2,   /* float storage, contraction as split-BF16 (2 pieces, 3 products) on MFMA */
    FFTSO3_BF16X6 = 3    /* float storage, contraction as split-BF16 (3 pieces, 6 products) on MFMA */
};

/* --- plans ----------------------------------------------------------------------------------- */

/* Plan flags.  By default a float-storage (FP32 or split-BF16) SO(3) plan of several bandlimits
 * (each a multiple of 8, up to l_max = 63) runs each pipeline stage as one persistent launch over
 * all its bandlimit classes: workgroups claim (class, unit) items, largest bandlimit first, from
 * per-XCD queues with atomic counters; FFTSO3_FLAG_PER_CLASS_LAUNCHES asks for one launch per
 * class and stage instead (the same results to 1e-6 relative); FFTSO3_FLAG_PER_CLASS_CONTRACTIONS
 * keeps the single launches of the FFT and layout stages but launches each class's contractions
 * on their own (the same results as per-class launches).  fftso3_plan_autotune chooses among the
 * three on its own. */
#define FFTSO3_FLAG_PER_CLASS_LAUNCHES 1u
#define FFTSO3_FLAG_PER_CLASS_CONTRACTIONS 2u

/* Uniform batch: n_fn functions of bandlimit B = l_max + 1.  flags: 0 or the flags above. */
int fftso3_plan_create(fftso3_plan_t *plan, int domain, int l_max, int n_fn, int kind, int precision,
                       unsigned flags);

/* Uneven batch: function i has bandlimit l_max[i] + 1 (the reference's objects each have one
 * bandlimit, src/fdcl_FFTSO3.cpp:8-14; a batch of them of several bandlimits is one plan here). */
int fftso3_plan_create_uneven(fftso3_plan_t *plan, int domain, const int *l_max, int n_fn, int kind,
                              int precision, unsigned flags);

int fftso3_plan_destroy(fftso3_plan_t plan);

/* Functions per pass through the pipeline.  By default FP32 plans with l_max <= 63 transform each
 * bandlimit class in one batch-wide pass (the Wigner-d block of every (m,n) group is generated
 * once and shared by all functions); set_chunk(1|2|4|8) selects the chunked kernels instead.
 * get_chunk reports the functions per pass of the first bandlimit class. */
int fftso3_plan_set_chunk(fftso3_plan_t plan, int chunk);
int fftso3_plan_get_chunk(fftso3_plan_t plan, int *chunk);

/* Graph mode (on = 1): each (direction, input, output, stream) combination is captured into a
 * hipGraph on its first call and replayed on later calls with the same arguments (at most 16 are
 * kept; changing the plan's schedule drops them).  Calls on the NULL stream, and calls while
 * per-stage timing is on, run eagerly.  No counterpart in the reference, which has no device
 * path; it serves callers that repeat the same transform on the same buffers. */
int fftso3_plan_set_graph(fftso3_plan_t plan, int on);

/* Schedule tuning: times `reps` inverse + forward pairs under each candidate schedule (chunk size
 * of the chunked kernels; row blocking of the spectrum workspace and group split of the
 * batch-wide kernels; for several bandlimits, one launch per stage or per class) and keeps the
 * fastest.  With graph mode on (and a non-NULL stream) the timed pairs are graph replays.  d_grid
 * and d_coef serve as scratch: their contents are undefined afterwards.  No counterpart in the
 * reference. */
int fftso3_plan_autotune(fftso3_plan_t plan, void *d_grid, void *d_coef, void *stream, int reps);

/* Schedule of the batch-wide kernels, the space autotune searches: workspace row blocking klo
 * (0..7: rows [k >> klo][g][k & (2^klo - 1)]), group interleave gil (0..2: the records of 2^gil
 * groups side by side; used at l_max = 63 with klo = 0) and group split (1, 2, 4: more
 * workgroups per group when groups are few).  The defaults are 0, 1, 1.  No counterpart in the
 * reference. */
int fftso3_plan_set_schedule(fftso3_plan_t plan, int klo, int gil, int split);

/* The plan's current schedule as text, one line per bandlimit class (l_max, functions, path,
 * functions per pass, workspace row blocking and group interleave, group split), written to buf
 * (at most len bytes, NUL-terminated).  No counterpart in the reference. */
int fftso3_plan_describe(fftso3_plan_t plan, char *buf, int len);

/* Element offsets of function i in the packed grid / coefficient arrays, and the totals. */
long fftso3_plan_grid_offset(fftso3_plan_t plan, int fn);
long fftso3_plan_coef_offset(fftso3_plan_t plan, int fn);
long fftso3_plan_grid_elems(fftso3_plan_t plan);
long fftso3_plan_coef_elems(fftso3_plan_t plan);
/* Device workspace held by the plan, in bytes. */
long fftso3_plan_workspace_bytes(fftso3_plan_t plan);

/* --- transforms (device pointers; stream is a hipStream_t, NULL for the default stream) ------- */

/* grid -> coefficients (forward_transform) */
int fftso3_forward(fftso3_plan_t plan, const void *d_grid, void *d_coef, void *stream);
/* coefficients -> grid (inverse_transform at every grid point) */
int fftso3_inverse(fftso3_plan_t plan, const void *d_coef, void *d_grid, void *stream);

/* fftso3_inverse that allocates nothing and uses none of the plan's workspaces: the spectrum
 * workspace and coefficient staging live in a static 128 MiB device array of the library (one per
 * GPU, held until the process ends), so it can be recorded under stream capture or called from
 * serving loops that must not allocate.  No counterpart in the reference; it is for callers that
 * cannot allocate.  Plans whose inverse needs more than 128 MiB of workspace (e.g. more than 6
 * complex functions at l_max = 63 in FP32) are rejected with FFTSO3_EINVAL.  Every call on a GPU
 * shares that GPU's array: calls that overlap in time on one GPU (other streams, other threads)
 * race, so the caller orders them.  It always runs eagerly (graph mode and per-stage timing of the
 * plan do not apply) on the device that was current when the plan was created. */
int fftso3_inverse_noalloc(fftso3_plan_t plan, const void *d_coef, void *d_grid, void *stream);

/* Pointwise inverse of function fn of the plan's batch at n_points arbitrary points
 * (inverse_transform(F, alpha, beta, gamma), src/fdcl_FFTSO3.cpp:498-531 and 1141-1167;
 * inverse_transform(F, theta, phi), src/fdcl_FFTS2.cpp:194-230 and 500-512), batched:
 *   d_coef    the plan's whole coefficient array (as for fftso3_inverse)
 *   d_angles  doubles: (alpha, beta, gamma) per point for SO(3), (theta, phi) for S^2
 *   d_values  one value per point: complex for FFTSO3_COMPLEX plans, real for FFTSO3_REAL plans,
 *             in the plan's storage type.
 * The recursion and the sums run in FP64 at every precision. */
int fftso3_evaluate(fftso3_plan_t plan, int fn, const void *d_coef, const double *d_angles, long n_points,
                    void *d_values, void *stream);

/* Rotation of spherical-harmonic coefficients by n_rot rotations, through an SO(3) plan of one
 * bandlimit (its Wigner-d tables; l_max is the plan's):
 *   FFTSO3_REAL plans:    G_i^l = U^l(R_i) F^l, U the real harmonics (real_harmonics,
 *                         src/fdcl_FFTSO3.cpp:1062-1105; the rotation of example/example3.cpp:50-53)
 *   FFTSO3_COMPLEX plans: G_i^l = D^l(R_i) F^l (wigner_D, src/fdcl_FFTSO3.cpp:201-224)
 *   d_F       (l_max+1)^2 coefficients in the S^2 packed layout l^2 + m + l, double (complex
 *             double interleaved for complex plans), whatever the plan's precision
 *   d_angles  n_rot x (alpha, beta, gamma), the 3-2-3 Euler angles of R_i
 *   d_G       n_rot x (l_max+1)^2 coefficients, same type as d_F
 * FP64 recursion and sums (accumulated with atomic adds, so the last bits may differ between
 * calls). */
int fftso3_rotate_s2(fftso3_plan_t plan, const double *d_F, const double *d_angles, long n_rot, double *d_G,
                     void *stream);

/* --- diagnostics ----------------------------------------------------------------------------- */

const char *fftso3_status_string(int status);
const char *fftso3_last_error(void);

/* Error-path testing: the after-th device allocation made by the calling thread from now on
 * fails with FFTSO3_ENOMEM (after = 0 switches this off).  A plan whose reconfiguration failed
 * (set_chunk, autotune) refuses to transform until a later reconfiguration succeeds.  No
 * counterpart in the reference. */
int fftso3_debug_fail_alloc(int after);

/* Per-stage kernel timing with HIP events on the launch stream.  Stages:
 * 0 fft(forward) 1 contraction(forward) 2 basis(forward) 3 basis(inverse) 4 contraction(inverse)
 * 5 fft(inverse).  read() synchronises, adds the elapsed milliseconds and launch counts since the
 * last reset into ms[6] / count[6], and resets. */
int fftso3_timing_enable(fftso3_plan_t plan, int on);
int fftso3_timing_read(fftso3_plan_t plan, double *ms, long *count);

/* Sizes (no plan needed) */
long fftso3_so3_ncoef(int l_max);   /* (l_max+1)(4(l_max+1)^2-1)/3 */
long fftso3_s2_ncoef(int l_max);    /* (l_max+1)^2 */

#ifdef __cplusplus
}
#endif
#endif
