/*
 * stab_hip.h -- C ABI of the MI355X (gfx950) implementation of stab's
 * Chebyshev-collocation stability solve (reference: sscollis/stab, Fortran).
 *
 * The reference's operator interface for this path is the pair of Fortran
 * subroutines temporal(name, ind) (temporal.f90:2) and spatial(name, ind)
 * (spatial.f90:2).  They take every input from module `stuff`
 * (stuff.f90:9-55) plus the mean-flow file profile.<ind> (getmean.f90:27-32),
 * build a dense complex pencil and call LAPACK on one core
 * (ZGESV+ZGEEV at temporal.f90:774,803; ZGETRF/ZGETRS+ZGEEV at
 * spatial.f90:978,996,1002,1043).  This library replaces that arithmetic:
 * the pencil is assembled on the GPU and the eigenpairs nearest a shift are
 * found by a blocked LU of the shifted pencil plus shift-invert Arnoldi.
 *
 * Conventions: plain C types only; complex numbers are interleaved
 * (re, im) doubles, Fortran column-major arrays, 0-based indices.  The caller
 * owns every host buffer; a context owns device memory and one HIP stream.
 * Every entry point returns a status (STAB_OK or a negative STAB_E*); per-point
 * results carry LAPACK-style info codes.  Nothing calls exit/stop.
 * One context is used by one host thread at a time.
 */
#ifndef STAB_HIP_H
#define STAB_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } stab_z;

/* Status codes returned by every entry point. */
#define STAB_OK            0
#define STAB_E_ARG        -1   /* bad argument (stab_last_error names it) */
#define STAB_E_HIP        -2   /* HIP runtime error */
#define STAB_E_NOMEM      -3   /* device allocation failed */
#define STAB_E_NOPROFILE  -4   /* point refers to a profile slot never set */

/* Per-point info codes (stab_solve info[]) */
#define STAB_INFO_OK         0
#define STAB_INFO_SINGULAR   1   /* exact zero pivot in the LU of the shifted pencil */
#define STAB_INFO_NOCONV     2   /* not every wanted pair met the residual tolerance */
#define STAB_INFO_FALLBACK   3   /* stab_solve_leading: found by the FP64 Krylov-Schur fallback */

/* Leading-mode methods (stab_solve_leading) */
#define STAB_LEAD_KRYLOV 0   /* Krylov-Schur for pt.nev pairs nearest pt.shift, then the leading one */
#define STAB_LEAD_INVIT  1   /* residual inverse iteration from pt.shift, FP64 Krylov-Schur fallback */
#define STAB_TRACK       2   /* as STAB_LEAD_INVIT, but the converged eigenvalue nearest pt.shift is
                                accepted whether discrete or not (mode tracking; the fallback returns
                                the converged pair nearest the fallback shift) */

/* A mode is discrete when its eigenvector, normalised to a largest entry of
 * modulus 1, stays below STAB_FARFIELD_MAX on the nodes of the outer eighth of
 * the domain (nodes 1..ceil(ny/8), all five variables), and its phase speed
 * omega_r / alpha_r is more than STAB_CONTINUUM_BAND away from the edge
 * velocity 1 (the vorticity/entropy continuous spectrum); a spatial mode must
 * also travel downstream (alpha_r > 0). */
#define STAB_FARFIELD_MAX   0.1
#define STAB_CONTINUUM_BAND 0.01

/* Solve modes */
#define STAB_MODE_PLAIN_F64  0   /* FP64 LU + FP64 shift-invert Arnoldi */
#define STAB_MODE_MIXED      1   /* equilibrated FP32 LU + FP64 iterative refinement; points whose
                                      FP32 factors contract poorly are re-solved in FP64 */

/* Matrices returned by stab_pencil (reference form, for parity checks) */
#define STAB_MAT_A0  0   /* temporal A0  (temporal.f90:664-752) */
#define STAB_MAT_B0  1   /* temporal B0  (temporal.f90:628-662) */
#define STAB_MAT_C0  2   /* spatial  C0  (spatial.f90:700-835)  */
#define STAB_MAT_C1  3   /* spatial  C1  (spatial.f90:837-938)  */
#define STAB_MAT_C2  4   /* spatial  C2  (spatial.f90:940-959)  */
#define STAB_MAT_SHIFTED 5 /* the matrix factored by stab_solve: A0 - sigma B0
                              (temporal) or C0 + sigma C1 + sigma^2 C2 (spatial),
                              with Dirichlet rows replaced by unit rows */

/*
 * One parameter point: the subset of module `stuff` (stuff.f90:9-55) that
 * temporal/spatial read, plus the shift.  Field meanings follow input.f90:
 *   mattyp  0 constant mu, 1 Sutherland (input.f90:14-32)
 *   Te      edge temperature as input.f90 computes it: T0 as entered when
 *           mattyp = 1 (input.f90:24-25 runs before Ma is read), 0 otherwise
 *   itype   1 Chebyshev temporal (temporal.f90), 2 Chebyshev spatial
 *           (spatial.f90), 3 / 4 finite-difference temporal / spatial
 *           (fd_temporal.f90, fd_spatial.f90, as stab.f90:56-67 runs them:
 *           wall at node 0, the explicit fourth-order rows of stencils.f90,
 *           ny >= 10, ymax > 2 yi; itype 3 reads neither wallt nor top, as
 *           fd_temporal does not)
 *   ny, yi, ymax   grid (sgengrid.f90:26-45; ogengrid.f90:17-48 for itypes 3, 4);
 *                  yi must be nonzero
 *   top, wallt     boundary-condition flags (input.f90:57-59)
 *   alpha, beta    temporal wavenumbers; omega, beta for spatial
 *   shift          sigma: a guess of omega (temporal) or alpha (spatial),
 *                  i.e. the value the reference's getevec -v reads
 *                  (getevec.f90:160-173, thesis/TStest/run.sh:13-17)
 *   profile        slot set by stab_profile_set (profile.<ind>)
 *   nev            eigenpairs wanted nearest the shift (<= STAB_MAX_NEV)
 */
#define STAB_MAX_NEV 16
typedef struct {
  int32_t itype;     /* 1 temporal, 2 spatial, 3 / 4 finite-difference temporal / spatial */
  int32_t ny;
  int32_t mattyp;
  int32_t wallt;     /* 0 or 2 */
  int32_t top;       /* 0 or 1 (spatial) */
  int32_t profile;
  int32_t nev;
  int32_t ider;      /* must be 1: derivatives by collocation (temporal.f90:99-103) */
  double Ma, Re, Pr, Te;
  double yi, ymax;
  stab_z alpha, beta, omega;
  stab_z shift;
} stab_point;

typedef struct stab_ctx stab_ctx;

/* Solver statistics of the last stab_solve / stab_solve_leading call
 * (device-timed, ms).  When stab_solve_leading's first pass runs as a HIP graph
 * (graph = 1 or 2) it is timed as a whole: ms_total holds it, and ms_prep,
 * ms_assemble, ms_lu and ms_lu_kernel are 0 (set STAB_GRAPH=0 for them). */
typedef struct {
  double ms_prep, ms_assemble, ms_lu, ms_eig, ms_total;
  int64_t lu_flops;       /* algorithmic (8/3) n^3 per point, summed */
  int64_t assemble_bytes; /* algorithmic bytes written/read by assembly */
  int32_t npts, nconv;    /* points, points whose nev pairs all converged */
  int32_t restarts_max;   /* Krylov-Schur cycles of the slowest point */
  int32_t nfallback;      /* STAB_MODE_MIXED: points re-solved by the FP64 path */
  double ms_arnoldi;      /* device time of the Arnoldi-expansion kernel launches */
  int64_t op_applies;     /* shift-invert operator applications, summed over points */
  double solve_bytes;     /* algorithmic factor bytes read by those applications */
  double ms_lu_kernel;    /* device time of the LU kernel alone (ms_lu adds the block inverses) */
  int32_t graph;          /* stab_solve_leading's first pass: 0 issued kernel by kernel, 1 captured
                             into a HIP graph and launched, 2 replayed from the graph of an earlier
                             call with the same batch layout (STAB_GRAPH=0 turns graphs off) */
  int32_t lu_launches;    /* LU segments: runs of points sharing a work-group shape, split by work up to one per LU stream; each runs its 64-column block steps (a panel and a trailing launch each) on its own stream */
} stab_stats;

/* Context on HIP device `device`.  New in this library (the reference keeps
 * its state in module globals, stuff.f90:2-61). */
int stab_ctx_create(stab_ctx **ctx, int device);
int stab_ctx_destroy(stab_ctx *ctx);
const char *stab_last_error(const stab_ctx *ctx);

/* Mean-flow profile slot `slot`: nym rows of y and vt (nym x 5, column-major:
 * rho, u, v, w, T).  v is forced to zero and the cubic-spline second
 * derivatives are formed on the device, as getmean.f90:56-86 does after
 * reading profile.<ind>. */
int stab_profile_set(stab_ctx *ctx, int slot, int nym, const double *y, const double *vt);

/* Grid, interpolated mean flow and derivative operators of one point, as the
 * device builds them (sgengrid.f90:15-45, getmean.f90:95-112, chebyd.f90:2-60,
 * temporal.f90:136; for itypes 3, 4 ogengrid.f90:17-48 and the explicit rows of
 * fd_temporal.f90:630-890 / fd_spatial.f90:669-1160).  y/eta/deta/d2eta/vm may be NULL; vm is ny x 5
 * column-major; d1/d2 are ny x ny column-major or NULL. */
int stab_grid(stab_ctx *ctx, const stab_point *pt, double *y, double *eta, double *deta,
              double *d2eta, double *vm, double *d1, double *d2);

/* Assemble matrix `which` (STAB_MAT_*) of one point on the device and copy
 * it to `out` (n x n column-major, n = 5 ny).  Replaces the fill loops of
 * temporal.f90:600-752 / spatial.f90:685-959. */
int stab_pencil(stab_ctx *ctx, const stab_point *pt, int which, stab_z *out);

/* Batched solve: for each point, the nev eigenvalues nearest pt.shift and
 * their eigenvectors.  Replaces ZGESV+ZGEEV (temporal.f90:771-810) and
 * ZGETRF/ZGETRS+ZGEEV (spatial.f90:972-1058) for the wanted part of the
 * spectrum, with the eigenvector scaling of temporal.f90:867-879 /
 * getevec.f90:179-189 applied to every returned vector.
 *   eval   npts*STAB_MAX_NEV values (point p, pair k at p*STAB_MAX_NEV+k),
 *          sorted by distance to the shift
 *   evec   NULL, or for point p a block of nvec_p*STAB_MAX_NEV values at offset
 *          sum_{q<p} nvec_q*STAB_MAX_NEV, nvec = 5 ny (temporal) or 10 ny
 *          (spatial: z = [phi; alpha phi], as the companion of
 *          spatial.f90:1007-1016 returns), column k = pair k; columns
 *          nev..STAB_MAX_NEV-1 (and eval/resid entries past nev) are zero
 *   resid  npts*STAB_MAX_NEV relative residuals
 *          ||(A - w B) x|| / ((||A|| + |w| ||B||) ||x||) in FP64 (Frobenius-type
 *          norms of the pencil's coefficient operators), or NULL
 *   info   npts per-point STAB_INFO_* codes
 */
int stab_solve(stab_ctx *ctx, int npts, const stab_point *pts, int mode, stab_z *eval,
               stab_z *evec, double *resid, int32_t *info, stab_stats *stats);

/* Factor the shifted matrix of one point (STAB_MAT_SHIFTED) and solve
 * M x = b (n = 5 ny entries).  Replaces the ZGETRF/ZGETRS pair of
 * spatial.f90:978,996 (one right-hand side).  mode STAB_MODE_MIXED factors the
 * row-equilibrated matrix in FP32 and applies ir_steps FP64
 * iterative-refinement steps (ir_steps < 0: adaptive, at most -ir_steps, until
 * the correction is below 1e-14 relative); STAB_MODE_PLAIN_F64 factors in FP64 (ir_steps
 * then refine against the same FP64 matrix).  info: STAB_INFO_SINGULAR on an
 * exact zero pivot. */
int stab_shifted_solve(stab_ctx *ctx, const stab_point *pt, int mode, int ir_steps, const stab_z *b, stab_z *x,
                       int32_t *info);

/* LU factors of the shifted matrix of one point as stab_solve computes them:
 * lu is npad x npad column-major (npad = 5 ny rounded up to a multiple of 32;
 * rows/columns beyond 5 ny hold the identity), unit-lower L below the
 * diagonal and U on and above it; perm[r] is the input row that ends in row r
 * (P A = L U).  STAB_MODE_MIXED factors Dr A Dc in FP32 and returns the row and
 * column scales in scales[0:npad] and scales[npad:2 npad] (scales may be NULL);
 * Dr holds powers of two and Dc is the identity.
 * Replaces ZGETRF (spatial.f90:978). */
int stab_lu_factors(stab_ctx *ctx, const stab_point *pt, int mode, stab_z *lu, int32_t *perm, double *scales,
                    int32_t *info);

/* Size in bytes of the evec buffer stab_solve needs for these points. */
size_t stab_evec_bytes(int npts, const stab_point *pts);

/* Leading eigenmode per point: the least stable discrete mode (largest
 * Im omega for temporal points, smallest Im alpha for spatial points) within
 * the window the method searches -- the pt.nev pairs nearest pt.shift
 * (STAB_LEAD_KRYLOV) or the pairs the iteration reaches from pt.shift
 * (STAB_LEAD_INVIT) -- not over the whole spectrum.  Columns and entries
 * past what a point returns are zero.  The
 * reference leaves this choice to the user, who picks a mode from the sorted
 * spectrum (temporal.f90:844-879) with getevec -v and a guess
 * (getevec.f90:160-173, thesis/TStest/run.sh:13-17).
 *   method STAB_LEAD_KRYLOV: the pt.nev pairs nearest pt.shift (as
 *     stab_solve), then the leading discrete pair with residual < 1e-10.
 *   method STAB_LEAD_INVIT: factor at pt.shift (a predicted eigenvalue) in
 *     `mode` (STAB_MODE_MIXED: equilibrated FP32 LU) and run at most maxit
 *     residual inverse-iteration corrections x <- x - L(shift)^-1 L(lambda) x
 *     with FP64 residuals (to a relative residual of 1e-15 or the rounding
 *     floor).  Temporal points whose residual stays above 1e-14, that land on
 *     a mode that is not discrete, or that needed 3/4 of maxit are iterated
 *     again on the same factors with two vectors and 2 x 2 Rayleigh-Ritz (at
 *     most 2 maxit steps); the wanted one of its two pairs is taken when its
 *     residual is below 1e-14 and the other pair cannot outrank it.  A point
 *     still undecided (or a spatial point the first pass missed) is solved
 *     again by STAB_LEAD_KRYLOV in FP64 from fallback_shift[p] (pt.shift if
 *     NULL) and pt.nev, and reports STAB_INFO_FALLBACK.
 *   method STAB_TRACK: as STAB_LEAD_INVIT with "nearest the shift" in place of
 *     "leading discrete": the eigenvalue selection of getab.f90:160-198 /
 *     getalpha.f90:139-145 (nearest to a guess or an extrapolation) without
 *     the full spectrum.
 *   eval/resid/farfield/iters/info: npts entries (iters = operator solves);
 *   evec: NULL or sum over points of nvec (5 ny temporal, 10 ny spatial with
 *     z = [phi; alpha phi]), normalised as getevec.f90:179-189.
 * info STAB_INFO_NOCONV: no discrete pair met the bound. */
int stab_solve_leading(stab_ctx *ctx, int npts, const stab_point *pts, int mode, int method, int maxit,
                       const stab_z *fallback_shift, stab_z *eval, stab_z *evec, double *resid,
                       double *farfield, int32_t *iters, int32_t *info, stab_stats *stats);

/* Allocation-free variant of stab_solve_leading.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and kernels as
 * stab_solve_leading; every device buffer of the call is carved from one
 * static device array of STAB_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends), so the call makes no hipMalloc/hipFree.  Points whose
 * worst-case need (first pass plus an FP64 Krylov-Schur fallback of every
 * point) exceeds that size are refused with STAB_E_ARG before any work --
 * roughly five N = 128 points or twenty N = 64 points per call.  All calls on
 * one GPU share the array and leave their data in it: calls that overlap in
 * time on the same GPU (two host threads, two contexts) race.  Profiles
 * must already be set (stab_profile_set allocates). */
#define STAB_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int stab_solve_leading_noalloc(stab_ctx *ctx, int npts, const stab_point *pts, int mode, int method, int maxit,
                               const stab_z *fallback_shift, stab_z *eval, stab_z *evec, double *resid,
                               double *farfield, int32_t *iters, int32_t *info, stab_stats *stats);

/* ---------------------------------------------------------------------------
 * Shift predictor.  No counterpart in the reference, whose user supplies the
 * eigenvalue guess by hand (getevec -v, thesis/TStest/run.sh:13-17): a tanh
 * MLP that maps stability parameters to the leading eigenvalue, trained and
 * evaluated on the GPU (FP32 master weights, BF16 matrix products on MFMA
 * with FP32 accumulation, Adam), whose predictions serve as pt.shift for
 * stab_solve_leading(STAB_LEAD_INVIT).
 *   nin inputs (<= 64) -> depth hidden tanh layers of width units (multiple
 *   of 16, <= 512) -> nout linear outputs (<= 16).  Inputs are mapped
 *   x' = (x - in_shift) * in_scale and outputs y = y' / out_scale + out_shift.
 * Flat FP32 parameter layout with nin_p = nin rounded up to 16 and nout_p =
 * 16 (padded rows and columns hold zeros):
 *   W_0 [width x nin_p], b_0 [width], then for l = 1..depth-1 W_l [width x
 *   width], b_l [width], then W_o [nout_p x width], b_o [nout_p];
 *   every W row-major by output unit (W[o * in + i]).
 * Row counts passed to stab_mlp_loss_grad and batch sizes are multiples of 16.
 * ------------------------------------------------------------------------- */
typedef struct stab_mlp stab_mlp;
int stab_mlp_create(stab_ctx *ctx, int nin, int width, int depth, int nout, uint64_t seed, stab_mlp **out);
int stab_mlp_destroy(stab_mlp *mlp);
int stab_mlp_nparams(const stab_mlp *mlp);
int stab_mlp_get_params(stab_mlp *mlp, float *params);
/* Also resets the Adam moments and step count. */
int stab_mlp_set_params(stab_mlp *mlp, const float *params);
int stab_mlp_set_scaling(stab_mlp *mlp, const float *in_shift, const float *in_scale, const float *out_shift,
                         const float *out_scale);
/* y (n x nout, row-major) = network(x (n x nin, row-major)). */
int stab_mlp_predict(stab_mlp *mlp, int n, const float *x, float *y);
/* Mean squared error over the n rows and nout outputs (in the scaled output
 * space) and its gradient with respect to the flat parameters. */
int stab_mlp_loss_grad(stab_mlp *mlp, int n, const float *x, const float *t, float *loss, float *grad);
/* `steps` Adam steps (beta1 0.9, beta2 0.999, eps 1e-8, rate lr) on
 * mini-batches of `batch` rows drawn with replacement from the n rows by a
 * hash of (seed, step, row) -- the rows in order when batch == n.  The data
 * set stays on the device for the call; loss_hist (NULL or steps values)
 * receives the batch loss of every step. */
int stab_mlp_train(stab_mlp *mlp, int n, const float *x, const float *t, int batch, int steps, float lr,
                   uint64_t seed, float *loss_hist);

/* Measurement utility with no counterpart in the reference: copies `bytes`
 * (a multiple of 16) between two device buffers `reps` times with a plain
 * 16-byte-per-thread kernel and returns the mean time of one copy.  Gives the
 * achievable HBM bandwidth (2 bytes moved per byte copied) and the reference
 * kernel that scales a profiler's byte counters. */
int stab_copy_probe(stab_ctx *ctx, int64_t bytes, int reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* STAB_HIP_H */
