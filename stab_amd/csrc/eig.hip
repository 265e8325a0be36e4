// Shift-invert Krylov-Schur eigensolver kernels (one work-group per point).
//
// Operator (temporal): T v = (A - sigma B)^-1 B v on the solver pencil, whose
// eigenvalues theta = 1/(omega - sigma) are largest for the omega nearest the
// shift.  Operator (spatial): the linearisation
//     [[C0,0],[0,I]] z = alpha [[-C1,-C2],[I,0]] z,  z = [phi; alpha phi]
// of the reference's companion (spatial.f90:982-1016); its shift-invert
// product reduces to one n x n solve with L(sigma) = C0 + sigma C1 + sigma^2 C2:
//     u = L(sigma)^-1 (-C1 z1 - C2 (z2 + sigma z1)),  T z = [u; z1 + sigma u].
// Solves use the LU factors (FP64, or equilibrated FP32 followed by FP64
// iterative-refinement steps whose residuals use the matrix-free pencil).
#include "pencil.h"

#define KS_MMAX 64
#define KS_NEVMAX 16

// Per-point Krylov-Schur state in global memory (small_off points here):
//   H  (KS_MMAX+1) x KS_MMAX, column-major, ld = KS_MMAX+1
//   Q  KS_MMAX x KS_MMAX (ld = KS_MMAX), Y KS_MMAX x KS_NEVMAX, theta[KS_MMAX]
#define SM_H 0
#define SM_Q ((KS_MMAX + 1) * KS_MMAX)
#define SM_Y (SM_Q + KS_MMAX * KS_MMAX)
#define SM_TH (SM_Y + KS_MMAX * KS_NEVMAX)
#define SM_SIZE (SM_TH + KS_MMAX)

struct KSState {
  int j0;        // first column to expand
  int k;         // kept columns after the last restart
  int done;      // 1: every wanted pair converged; 2: stopped on stagnation
  int cycles;
  double resmax; // max relative Ritz residual estimate of the wanted pairs
  double best;   // smallest resmax seen
  int stall;     // cycles without a 10x improvement of best
  int irbad;     // 1 if a refined solve contracted poorly (FP32 factors inadequate)
  int steps;     // operator applications so far
  int pad;
};

// ---------------------------------------------------------------------------
// y <- M^-1 y with the factors of k_lu and the 64 x 64 diagonal-block
// inverses of k_blkinv (scales applied by the caller).  The solve reads the
// whole factor once and is bound by how many bytes one work-group keeps in
// flight, so every load is 16 bytes wide (two FP32 or one FP64 complex matrix
// entries of consecutive rows) and each lane issues 16 of them before using
// any (64 KB in flight per work-group).  The factor and block inverses are
// read with streaming (non-temporal) loads: a point's next solve comes after
// the other work-groups have streamed theirs, so they would only push the
// operators D1, D2, which every pencil product rereads, out of L2.  Per block
// of 64 rows:
//   1. y_top <- inv(D) y_top: lane (i, q) sums columns 16q..16q+15 of row i
//      of the column-major inverse (a wave reads 64 consecutive entries per
//      load); the four partial sums meet in tmp (NT = 256 entries);
//   2. the rows below (forward) or above (backward): lane (g, q) of a row
//      group g (E rows, E = 16 bytes / entry size) loads 16 columns of
//      quarter q and the quarters are summed by shuffles.
// Columns of a partial last block (bs < 64) beyond bs are never loaded.
// y, work in LDS (npad); tmp in LDS (NT entries).
// ---------------------------------------------------------------------------
#define SOLVE_BS 64
template <typename T> struct Row16;
template <> struct Row16<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  static constexpr int E = 2;
  __device__ static zd get(const V &v, int e) { return e == 0 ? mk((double)v.x, (double)v.y) : mk((double)v.z, (double)v.w); }
};
template <> struct Row16<double> {
  typedef double V __attribute__((ext_vector_type(2)));
  static constexpr int E = 1;
  __device__ static zd get(const V &v, int) { return mk(v.x, v.y); }
};

// For each right-hand side r < NR: tmp[r NT + 0..bs) <- inv(D) y_r[k0..k0+bs)
// with D the diagonal block at inv (column major, entries beyond bs zero),
// tmp[r NT + bs..64) <- 0, and y_r[k0..k0+bs) <- the same.  The inverse is
// loaded once for all right-hand sides.  Branch-free: indexes past bs are
// clamped to bs - 1 and meet zero entries.  Two barriers.
template <typename T, int NT, int NR>
__device__ inline void solve_diag(const cx<T> *__restrict__ inv, int k0, int bs, zd *const (&y)[NR], zd *tmp) {
  constexpr int BS = SOLVE_BS;
  const int t = threadIdx.x, i = t & (BS - 1), q = t >> 6;
  const cx<T> *col = inv + (size_t)(q * 16) * BS + i;
  cx<T> v[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = load_stream(col + j * BS);
  zd s[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) s[r] = czero<double>();
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const zd vj = cvt<double>(v[j]);
    const int kk = k0 + min(q * 16 + j, bs - 1);
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] = cfma(s[r], vj, y[r][kk]);
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) tmp[r * NT + t] = s[r];
  __syncthreads();
  if (t < BS) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      zd *tr = tmp + r * NT;
      const zd v = tr[t] + tr[BS + t] + tr[2 * BS + t] + tr[3 * BS + t];
      tr[t] = (t < bs) ? v : czero<double>();
      if (t < bs) y[r][k0 + t] = v;
    }
  }
  __syncthreads();
}

// y_r[r0..r1) -= A[r0..r1, k0..k0+bs) tmp[r NT + 0..bs) for r < NR (the matrix
// is loaded once for all right-hand sides); r0, r1 multiples of E.  PIPE
// (the inverse-iteration kernels) issues each entry's load for the next pass
// as soon as the entry is used, so loads stay in flight through the arithmetic
// and the reductions; the other kernels, tighter on registers, load pass by
// pass.  The arithmetic is the same either way.
// Branch-free: column indexes past bs are clamped to bs - 1 (tmp is zero
// there) and rows past r1 to the last row group (not written back).
template <typename T, int NT, int NR, bool PIPE = false>
__device__ inline void solve_update(const cx<T> *__restrict__ A, int lda, int k0, int bs, int r0, int r1,
                                    zd *const (&y)[NR], const zd *tmp) {
  typedef Row16<T> R;
  typedef typename R::V V;
  constexpr int E = R::E, G = NT / 4;        // G row groups of E rows per pass
  const int t = threadIdx.x, q = t & 3, g = t >> 2;
  const int ldv = lda / E;
  const V *base = (const V *)A;
  int coff[16];                              // this lane's 16 columns, in V units
#pragma unroll
  for (int j = 0; j < 16; ++j) coff[j] = (k0 + min(q * 16 + j, bs - 1)) * ldv;
  V v[16];
  auto load = [&](int rb) {
    const int rl = min(rb + g * E, r1 - E);
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = __builtin_nontemporal_load(base + coff[j] + rl / E);
  };
  // one pass over rows rb .. rb + G E; with NEXT, each entry's register is
  // reloaded with the next pass's entry (rows rn ..) as soon as it is used
  auto pass = [&](int rb, int rn, auto next) {
    constexpr bool NEXT = decltype(next)::value;
    const int row = rb + g * E, rln = min(rn + g * E, r1 - E);
    zd a[NR][E];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int e = 0; e < E; ++e) a[r][e] = czero<double>();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const zd mj = tmp[r * NT + q * 16 + j];
#pragma unroll
        for (int e = 0; e < E; ++e) a[r][e] = cfma(a[r][e], R::get(v[j], e), mj);
      }
      if (NEXT) v[j] = __builtin_nontemporal_load(base + coff[j] + rln / E);
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        a[r][e].x += quad_xor1(a[r][e].x); a[r][e].y += quad_xor1(a[r][e].y);
        a[r][e].x += quad_xor2(a[r][e].x); a[r][e].y += quad_xor2(a[r][e].y);
      }
    if (row < r1 && q == 0) {
#pragma unroll
      for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int e = 0; e < E; ++e) y[r][row + e] = y[r][row + e] - a[r][e];
    }
  };
  if (!PIPE) {
#pragma unroll 1
    for (int rb = r0; rb < r1; rb += G * E) {
      load(rb);
      pass(rb, rb, std::false_type());
    }
    return;
  }
  // the next pass's loads go out while this pass computes (in the registers
  // this pass has just used, so no second buffer)
  if (r0 >= r1) return;
  load(r0);
  int rb = r0;
#pragma unroll 1
  for (; rb + G * E < r1; rb += G * E) pass(rb, rb + G * E, std::true_type());
  pass(rb, rb, std::false_type());
}

// NR right-hand sides y[r] (LDS, npad) solved together; tmp holds NR NT
// entries, work npad.
template <typename T, int NT, int NR, bool PIPE = false>
__device__ __forceinline__ void lu_solve_n(const PointDev &P, const cx<T> *__restrict__ A, const int *__restrict__ ipiv,
                                           const cx<T> *__restrict__ blk, zd *const (&y)[NR], zd *tmp, zd *work) {
  static_assert(NT == 4 * SOLVE_BS, "four lanes per row of a 64-row block");
  static_assert(LU_PREC == 1 + 4 * SOLVE_BS, "one interchange record per 64-row block");
  constexpr int BS = SOLVE_BS;
  const int n = P.npad, lda = P.npad, t = threadIdx.x;
  const int *perm = ipiv + n;
  const int nblk = (n + BS - 1) / BS;
  const cx<T> *Li = blk, *Ui = blk + (size_t)nblk * BS * BS;   // layout of k_blkinv
  const int *rec0 = ipiv + 2 * n;                                // pivot mode, block records
  const bool progressive = rec0[0] != 0;
  if (!progressive) {
    // y <- P y (one gather through `work` per right-hand side)
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      for (int i = t; i < n; i += NT) work[i] = y[r][perm[i]];
      __syncthreads();
      for (int i = t; i < n; i += NT) y[r][i] = work[i];
      __syncthreads();
    }
  }
  for (int b = 0; b < nblk; ++b) {
    const int k0 = b * BS, bs = min(BS, n - k0);
    if (progressive) {
      // the block's net interchanges (rows >= k0, at most 2 BS of them and
      // never more than n - k0, so they fit in work)
      const int *rec = rec0 + 1 + (size_t)b * LU_PREC;
      const int ns = rec[0];
      if (ns > 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          for (int e = t; e < ns; e += NT) work[e] = y[r][rec[1 + 2 * BS + e]];
          __syncthreads();
          for (int e = t; e < ns; e += NT) y[r][rec[1 + e]] = work[e];
          __syncthreads();
        }
      }
    }
    solve_diag<T, NT, NR>(Li + (size_t)b * BS * BS, k0, bs, y, tmp);
    if (k0 + bs < n) {
      solve_update<T, NT, NR, PIPE>(A, lda, k0, bs, k0 + bs, n, y, tmp);
      __syncthreads();
    }
  }
  for (int b = nblk - 1; b >= 0; --b) {
    const int k0 = b * BS, bs = min(BS, n - k0);
    solve_diag<T, NT, NR>(Ui + (size_t)b * BS * BS, k0, bs, y, tmp);
    if (k0 > 0) {
      solve_update<T, NT, NR, PIPE>(A, lda, k0, bs, 0, k0, y, tmp);
      __syncthreads();
    }
  }
}

template <typename T, int NT, bool PIPE = false>
__device__ __forceinline__ void lu_solve(const PointDev &P, const cx<T> *__restrict__ A, const int *__restrict__ ipiv,
                                         const cx<T> *__restrict__ blk, zd *y, zd *tmp, zd *work) {
  zd *const ys[1] = {y};
  lu_solve_n<T, NT, 1, PIPE>(P, A, ipiv, blk, ys, tmp, work);
}

// Block-wide reduction of NV complex partial sums; result broadcast to all.
template <int NT, int NV>
__device__ void block_reduce(zd (&v)[NV], zd *red) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    for (int off = 32; off > 0; off >>= 1) {
      v[k].x += __shfl_xor(v[k].x, off);
      v[k].y += __shfl_xor(v[k].y, off);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w * NV + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    zd s = czero<double>();
    for (int ww = 0; ww < NT / 64; ++ww) s = s + red[ww * NV + k];
    v[k] = s;
  }
  __syncthreads();
}

template <int NT>
__device__ double block_norm2(const zd *x, int len, zd *red) {
  zd v[1] = {czero<double>()};
  for (int r = threadIdx.x; r < len; r += NT) v[0].x += abs2(x[r]);
  block_reduce<NT, 1>(v, red);
  return v[0].x;
}

// Shared-memory layout of the operator kernels (dynamic LDS, zd units).
// The full layout serves refinement and the spatial operator; the compact one
// (plain FP64, temporal only) keeps just the solve vector and the Krylov
// vector, so more work-groups fit on a CU.
struct OpLds {
  zd *y, *rhs, *res, *X1, *X2, *z, *red, *tmp, *w2;
  // nvecmax: longest Krylov vector (2 n for spatial points).  w2, the gather
  // buffer of the solves, aliases X1/X2 (10 ny >= npad), which are idle then.
  __device__ OpLds(zd *base, const PointDev &P, int npadmax, int nymax, int full, int nvecmax) {
    red = base;
    tmp = red + 64;
    y = tmp + 256;
    z = y + npadmax;
    if (full) {
      rhs = z + nvecmax;
      res = rhs + npadmax;
      X1 = res + npadmax;
      X2 = X1 + 5 * nymax;
      w2 = X1;
    } else {
      rhs = res = X1 = X2 = w2 = nullptr;
    }
  }
};

__host__ __device__ inline size_t op_lds_bytes(int npadmax, int nymax, int full, int nvecmax) {
  return sizeof(zd) * (320 + (size_t)npadmax + (size_t)nvecmax +
                       (full ? 2 * (size_t)npadmax + 10 * (size_t)nymax : 0));
}

// y = M^-1 rhs in FP64 quality: scaled solve with the factors, then
// iterative-refinement steps with the FP64 matrix-free shifted pencil.
// ir > 0: exactly ir steps.  ir < 0: adaptive, at most -ir steps, stopping once
// the correction is below 1e-14 of the solution or stops shrinking by half
// (the FP64 rounding floor of the residual has been reached).  The contraction
// per step is about kappa(M) u32.
template <typename T, int NB, int NT, bool PIPE = false>
__device__ bool shifted_solve(const PointDev &P, const cx<T> *A, const int *ipiv, const cx<T> *Lb, const cx<T> *Ub,
                              const double *dr, const double *dc, const zd *coef, const double *ops, int ir,
                              OpLds &L) {
  const int n = P.npad, t = threadIdx.x;
  for (int r = t; r < n; r += NT) L.y[r] = (r < P.n) ? L.rhs[r] * dr[r] : czero<double>();
  __syncthreads();
  lu_solve<T, NT, PIPE>(P, A, ipiv, Lb, L.y, L.tmp, L.w2);
  for (int r = t; r < n; r += NT) L.y[r] = L.y[r] * dc[r];
  __syncthreads();
  const bool adaptive = ir < 0;
  const int steps = adaptive ? -ir : ir;
  double dprev = 0.0;
  bool ok = !adaptive;
  for (int s = 0; s < steps; ++s) {
    pencil_apply<NT>(P, coef, ops, FORM_SHIFTED, P.shift, L.y, L.res, L.X1, L.X2, L.red);
    for (int r = t; r < n; r += NT) L.res[r] = (r < P.n) ? (L.rhs[r] - L.res[r]) * dr[r] : czero<double>();
    __syncthreads();
    lu_solve<T, NT, PIPE>(P, A, ipiv, Lb, L.res, L.tmp, L.w2);
    zd nrm[2] = {czero<double>(), czero<double>()};
    for (int r = t; r < n; r += NT) {
      const zd d = L.res[r] * dc[r];
      L.y[r] = L.y[r] + d;
      nrm[0].x += abs2(d);
      nrm[1].x += abs2(L.y[r]);
    }
    if (adaptive) {
      block_reduce<NT, 2>(nrm, L.red);
      // the FP32 factors are adequate for this shift when the first correction
      // is below 1e-2 of the solution (refinement then contracts fast)
      if (s == 0) ok = nrm[0].x <= 1e-4 * nrm[1].x;
      if (nrm[0].x <= 1e-28 * nrm[1].x || (s > 0 && nrm[0].x > 0.25 * dprev)) break;
      dprev = nrm[0].x;
    } else {
      __syncthreads();
    }
  }
  return ok;
}

// w = T v for point P; v read from global (length nvec), w left in L.z.
// FULL = false compiles only the plain FP64 temporal operator.
template <typename T, int NB, int NT, bool FULL = true>
__device__ bool apply_op(const PointDev &P, const cx<T> *A, const int *ipiv, const cx<T> *Lb, const cx<T> *Ub,
                         const double *dr, const double *dc, const zd *coef, const double *ops, int ir,
                         const zd *v, OpLds &L) {
  const int n = P.n, t = threadIdx.x;
  if (!FULL || (P.itype == 1 && L.rhs == nullptr)) {
    // plain FP64 temporal: y = B v in place, one solve, no scaling
    for (int r = t; r < n; r += NT) L.z[r] = v[r];
    __syncthreads();
    pencil_apply<NT>(P, coef, ops, FORM_BM, P.shift, L.z, L.y, L.X1, L.X2, L.red);
    for (int r = n + t; r < P.npad; r += NT) L.y[r] = czero<double>();
    __syncthreads();
    lu_solve<T, NT>(P, A, ipiv, Lb, L.y, L.tmp, L.z);
    for (int r = t; r < n; r += NT) L.z[r] = L.y[r];
    __syncthreads();
    return true;
  }
  if constexpr (!FULL) return true;
  bool ok;
  if (P.itype == 1) {
    for (int r = t; r < n; r += NT) L.z[r] = v[r];
    __syncthreads();
    pencil_apply<NT>(P, coef, ops, FORM_BM, P.shift, L.z, L.rhs, L.X1, L.X2, L.red);
    ok = shifted_solve<T, NB, NT>(P, A, ipiv, Lb, Ub, dr, dc, coef, ops, ir, L);
    for (int r = t; r < n; r += NT) L.z[r] = L.y[r];
    __syncthreads();
  } else {
    const zd sg = P.shift;
    // z holds [z1; z2 + sigma z1]
    for (int r = t; r < n; r += NT) { L.z[r] = v[r]; L.z[n + r] = v[n + r] + sg * v[r]; }
    __syncthreads();
    pencil_apply<NT>(P, coef, ops, FORM_C1, sg, L.z, L.rhs, L.X1, L.X2, L.red);
    pencil_apply<NT>(P, coef, ops, FORM_C2, sg, L.z + n, L.res, L.X1, L.X2, L.red);
    for (int r = t; r < n; r += NT) L.rhs[r] = czero<double>() - L.rhs[r] - L.res[r];
    __syncthreads();
    ok = shifted_solve<T, NB, NT>(P, A, ipiv, Lb, Ub, dr, dc, coef, ops, ir, L);
    for (int r = t; r < n; r += NT) { L.z[n + r] = L.z[r] + sg * L.y[r]; L.z[r] = L.y[r]; }
    __syncthreads();
  }
  return ok;
}

// Deterministic start vector: T applied to a hashed pseudo-random vector.
__device__ inline double hash01(unsigned a, unsigned b) {
  unsigned h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
  return (double)h * (1.0 / 4294967296.0) - 0.5;
}

// Seed of a point's start vector from the point's own parameters, so that its
// iterates do not depend on where it sits in the batch.
__device__ inline unsigned point_seed(const PointDev &P) {
  const double f[] = {P.alpha.x, P.alpha.y, P.beta.x, P.beta.y, P.omega.x, P.omega.y, P.shift.x, P.shift.y,
                      P.Re, P.Ma, P.Pr, P.Te, P.yi, P.ymax};
  unsigned long long h = 0xCBF29CE484222325ull ^ (unsigned long long)(P.ny * 4 + P.itype);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(f) / sizeof(f[0])); ++k) {
    h ^= (unsigned long long)__double_as_longlong(f[k]);
    h *= 0x100000001B3ull;
    h ^= h >> 29;
  }
  return (unsigned)(h ^ (h >> 32));
}

// ---------------------------------------------------------------------------
// Arnoldi expansion from column j0 to m (CGS2).  V column-major, ld = ldv.
// ---------------------------------------------------------------------------
template <typename T, int NB, int NT, bool FULL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(FULL ? 2 : 3))) void k_arnoldi(const PointDev *__restrict__ pts, const cx<T> *__restrict__ mats,
                                                const int *__restrict__ pivs, const cx<T> *__restrict__ blks,
                                                const double *__restrict__ scales, const zd *__restrict__ coef,
                                                const double *__restrict__ ops, zd *__restrict__ Vall,
                                                zd *__restrict__ small, KSState *__restrict__ st, int m, int ir,
                                                int npadmax, int nymax, int nvecmax, int start) {
  extern __shared__ zd lds[];
  const PointDev &P = pts[blockIdx.x];
  if (st[blockIdx.x].done) return;
  int j0 = st[blockIdx.x].j0;
  OpLds L(lds, P, npadmax, nymax, FULL ? 1 : 0, nvecmax);
  const int t = threadIdx.x, nv = P.nvec;
  const size_t ldv = (size_t)nv;
  zd *V = Vall + P.vec_off;
  zd *H = small + P.small_off + SM_H;
  const int ldh = KS_MMAX + 1;
  const cx<T> *A = mats + P.mat_off;
  const int *ipiv = pivs + P.piv_off;
  const cx<T> *Lb = blks + P.blk_off, *Ub = nullptr;
  const double *dr = scales + P.scale_off, *dc = dr + P.npad;

  bool allok = true;
  if (start) {
    // v0 = T r / ||T r||, r pseudo-random
    const unsigned seed = point_seed(P);
    for (int r = t; r < nv; r += NT) V[r] = mk(hash01(seed, 2 * r), hash01(seed, 2 * r + 1));
    __threadfence_block();
    __syncthreads();
    allok &= apply_op<T, NB, NT, FULL>(P, A, ipiv, Lb, Ub, dr, dc, coef, ops, ir, V, L);
    const double nrm = sqrt(block_norm2<NT>(L.z, nv, L.red));
    for (int r = t; r < nv; r += NT) V[r] = L.z[r] * (1.0 / nrm);
    for (int e = t; e < ldh * KS_MMAX; e += NT) H[e] = czero<double>();
    __syncthreads();
    j0 = 0;
  }
  for (int j = j0; j < m; ++j) {
    allok &= apply_op<T, NB, NT, FULL>(P, A, ipiv, Lb, Ub, dr, dc, coef, ops, ir, V + (size_t)j * ldv, L);
    // two passes of classical Gram-Schmidt against V[:, 0..j]
    for (int pass = 0; pass < 2; ++pass) {
      for (int l0 = 0; l0 <= j; l0 += 4) {
        zd h[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) h[k] = czero<double>();
        for (int r = t; r < nv; r += NT) {
          const zd w = L.z[r];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (l0 + k <= j) h[k] = cfmac(h[k], V[(size_t)(l0 + k) * ldv + r], w);
        }
        block_reduce<NT, 4>(h, L.red);
        for (int r = t; r < nv; r += NT) {
          zd w = L.z[r];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (l0 + k <= j) w = cfms(w, V[(size_t)(l0 + k) * ldv + r], h[k]);
          L.z[r] = w;
        }
        if (t < 4 && l0 + t <= j) {
          zd hv = h[0];
#pragma unroll
          for (int k = 0; k < 4; ++k) if (k == t) hv = h[k];
          H[(size_t)j * ldh + l0 + t] = H[(size_t)j * ldh + l0 + t] + hv;
        }
        __syncthreads();
      }
    }
    double nrm = sqrt(block_norm2<NT>(L.z, nv, L.red));
    if (!(nrm > 1e-300)) nrm = 1e-300;
    for (int r = t; r < nv; r += NT) V[(size_t)(j + 1) * ldv + r] = L.z[r] * (1.0 / nrm);
    if (t == 0) H[(size_t)j * ldh + j + 1] = mk(nrm, 0.0);
    __syncthreads();
  }
  if (t == 0) {
    st[blockIdx.x].steps += (m - j0) + (start ? 1 : 0);
    st[blockIdx.x].j0 = m;
    if (start) st[blockIdx.x].k = 0;
    if (!allok) st[blockIdx.x].irbad = 1;
  }
}

// ---------------------------------------------------------------------------
// Small dense eigenproblem of H_m (one wave): Householder reduction to
// Hessenberg form, complex single-shift QR to Schur form T = Q^H H Q, then the
// `keep` eigenvalues of largest modulus are moved to the leading positions in
// decreasing modulus.  Ritz vectors of the first nev, residual estimates and
// convergence are written back.
// ---------------------------------------------------------------------------
__device__ inline void givens(zd x, zd y, double &c, zd &s) {
  const double ax = cabs(x), ay = cabs(y);
  if (ay == 0.0) { c = 1.0; s = czero<double>(); return; }
  if (ax == 0.0) { c = 0.0; s = conj(y) * (1.0 / ay); return; }
  const double r = hypot(ax, ay);
  c = ax / r;
  s = (x * (1.0 / ax)) * conj(y) * (1.0 / r);
}

__global__ __launch_bounds__(64) void k_schur(const PointDev *__restrict__ pts, zd *__restrict__ small,
                                              KSState *__restrict__ st, int m, int keep, double tol) {
  extern __shared__ zd sl[];
  const PointDev &P = pts[blockIdx.x];
  KSState S = st[blockIdx.x];
  if (S.done) return;
  zd *Hs = sl, *Q = Hs + m * m, *v = Q + m * m, *zz = v + m;   // zz: KS_NEVMAX * m
  zd *Hg = small + P.small_off + SM_H;
  zd *Qg = small + P.small_off + SM_Q;
  zd *Yg = small + P.small_off + SM_Y;
  zd *THg = small + P.small_off + SM_TH;
  const int ldh = KS_MMAX + 1, t = threadIdx.x;
  const int nev = P.nev;
  for (int e = t; e < m * m; e += 64) {
    const int i = e % m, j = e / m;
    Hs[i * m + j] = Hg[(size_t)j * ldh + i];   // row-major in LDS
    Q[i * m + j] = (i == j) ? mk(1.0, 0.0) : czero<double>();
  }
  const zd hlast = Hg[(size_t)(m - 1) * ldh + m];
  __syncthreads();
  // --- Householder reduction to Hessenberg form
  for (int k = 0; k < m - 2; ++k) {
    const int len = m - k - 1;
    double nx = 0.0;
    for (int i = 0; i < len; ++i) nx += abs2(Hs[(k + 1 + i) * m + k]);
    double below = nx - abs2(Hs[(k + 1) * m + k]);
    if (below <= 1e-300 * (nx + 1e-300)) continue;     // already reduced
    nx = sqrt(nx);
    const zd x0 = Hs[(k + 1) * m + k];
    const double ax0 = cabs(x0);
    const zd ph = (ax0 > 0.0) ? x0 * (1.0 / ax0) : mk(1.0, 0.0);
    const zd alpha = ph * (-nx);
    // v = x - alpha e1, normalised
    for (int i = t; i < len; i += 64) v[i] = (i == 0) ? x0 - alpha : Hs[(k + 1 + i) * m + k];
    __syncthreads();
    double vn = 0.0;
    for (int i = 0; i < len; ++i) vn += abs2(v[i]);
    vn = sqrt(vn);
    __syncthreads();
    for (int i = t; i < len; i += 64) v[i] = v[i] * (1.0 / vn);
    __syncthreads();
    // H <- (I - 2 v v^H) H  on rows k+1.., columns k..m-1
    for (int c = k + t; c < m; c += 64) {
      zd s = czero<double>();
      for (int i = 0; i < len; ++i) s = cfmac(s, v[i], Hs[(k + 1 + i) * m + c]);
      for (int i = 0; i < len; ++i) Hs[(k + 1 + i) * m + c] = cfms(Hs[(k + 1 + i) * m + c], v[i], s * 2.0);
    }
    __syncthreads();
    // H <- H (I - 2 v v^H) on columns k+1.., all rows; Q likewise
    for (int rr = t; rr < 2 * m; rr += 64) {
      zd *row = (rr < m) ? &Hs[rr * m] : &Q[(rr - m) * m];
      zd s = czero<double>();
      for (int i = 0; i < len; ++i) s = cfma(s, row[k + 1 + i], v[i]);
      for (int i = 0; i < len; ++i) row[k + 1 + i] = cfms(row[k + 1 + i], s * 2.0, conj(v[i]));
    }
    __syncthreads();
    if (t == 0)
      for (int i = 2; i < len + 1; ++i) Hs[(k + i) * m + k] = czero<double>();
    __syncthreads();
  }
  // --- complex QR iterations to Schur form (full T and Q)
  const double eps = 2.220446049250313e-16;
  int hi = m - 1, iter = 0, total = 0;
  while (hi > 0 && total < 60 * m) {
    int lo = hi;
    while (lo > 0) {
      const double sub = cabs1(Hs[lo * m + lo - 1]);
      const double dia = cabs1(Hs[lo * m + lo]) + cabs1(Hs[(lo - 1) * m + lo - 1]);
      if (sub <= eps * dia || sub < 1e-300) break;
      --lo;
    }
    if (lo > 0) {
      __syncthreads();
      if (t == 0) Hs[lo * m + lo - 1] = czero<double>();
      __syncthreads();
    }
    if (lo == hi) { --hi; iter = 0; continue; }
    ++iter; ++total;
    zd mu;
    const zd a = Hs[(hi - 1) * m + hi - 1], b = Hs[(hi - 1) * m + hi], c = Hs[hi * m + hi - 1], d = Hs[hi * m + hi];
    if (iter % 11 == 10) {
      mu = d + mk(0.75 * cabs(c), 0.0);
    } else {
      const zd half = (a - d) * 0.5;
      zd disc = half * half + b * c;
      // complex sqrt
      const double r = cabs(disc);
      double sr = sqrt(0.5 * (r + fabs(disc.x)));
      zd sq;
      if (sr == 0.0) sq = czero<double>();
      else if (disc.x >= 0) sq = mk(sr, disc.y / (2.0 * sr));
      else sq = mk(fabs(disc.y) / (2.0 * sr), disc.y >= 0 ? sr : -sr);
      const zd m1 = (a + d) * 0.5 + sq, m2 = (a + d) * 0.5 - sq;
      mu = (cabs(m1 - d) < cabs(m2 - d)) ? m1 : m2;
    }
    for (int k = lo; k < hi; ++k) {
      zd x, y;
      if (k == lo) { x = Hs[lo * m + lo] - mu; y = Hs[(lo + 1) * m + lo]; }
      else { x = Hs[k * m + k - 1]; y = Hs[(k + 1) * m + k - 1]; }
      double cg; zd sg;
      givens(x, y, cg, sg);
      __syncthreads();
      const int c0 = (k > lo) ? k - 1 : lo;
      for (int jj = c0 + t; jj < m; jj += 64) {
        const zd p = Hs[k * m + jj], q = Hs[(k + 1) * m + jj];
        Hs[k * m + jj] = p * cg + sg * q;
        Hs[(k + 1) * m + jj] = q * cg - conj(sg) * p;
      }
      __syncthreads();
      const int rmax = min(k + 2, hi);
      for (int ii = t; ii <= rmax; ii += 64) {
        const zd p = Hs[ii * m + k], q = Hs[ii * m + k + 1];
        Hs[ii * m + k] = p * cg + conj(sg) * q;
        Hs[ii * m + k + 1] = q * cg - sg * p;
      }
      for (int ii = t; ii < m; ii += 64) {
        const zd p = Q[ii * m + k], q = Q[ii * m + k + 1];
        Q[ii * m + k] = p * cg + conj(sg) * q;
        Q[ii * m + k + 1] = q * cg - sg * p;
      }
      __syncthreads();
      if (k > lo && t == 0) Hs[(k + 1) * m + k - 1] = czero<double>();
      __syncthreads();
    }
  }
  __syncthreads();
  // --- reorder: bring the `keep` eigenvalues of largest modulus to the front
  for (int pos = 0; pos < keep; ++pos) {
    int best = pos;
    double bv = cabs(Hs[pos * m + pos]);
    for (int i = pos + 1; i < m; ++i) {
      const double vv = cabs(Hs[i * m + i]);
      if (vv > bv) { bv = vv; best = i; }
    }
    for (int k = best - 1; k >= pos; --k) {
      const zd t11 = Hs[k * m + k], t22 = Hs[(k + 1) * m + k + 1];
      double cg; zd sg;
      givens(Hs[k * m + k + 1], t22 - t11, cg, sg);
      __syncthreads();
      for (int jj = k + t; jj < m; jj += 64) {
        const zd p = Hs[k * m + jj], q = Hs[(k + 1) * m + jj];
        Hs[k * m + jj] = p * cg + sg * q;
        Hs[(k + 1) * m + jj] = q * cg - conj(sg) * p;
      }
      __syncthreads();
      for (int ii = t; ii <= k + 1; ii += 64) {
        const zd p = Hs[ii * m + k], q = Hs[ii * m + k + 1];
        Hs[ii * m + k] = p * cg + conj(sg) * q;
        Hs[ii * m + k + 1] = q * cg - sg * p;
      }
      for (int ii = t; ii < m; ii += 64) {
        const zd p = Q[ii * m + k], q = Q[ii * m + k + 1];
        Q[ii * m + k] = p * cg + conj(sg) * q;
        Q[ii * m + k + 1] = q * cg - sg * p;
      }
      __syncthreads();
      if (t == 0) {
        Hs[k * m + k] = t22;
        Hs[(k + 1) * m + k + 1] = t11;
        Hs[(k + 1) * m + k] = czero<double>();
      }
      __syncthreads();
    }
  }
  // --- eigenvectors of T for the first nev (lane i handles pair i)
  if (t < nev) {
    const int i = t;
    zd *z = zz + i * m;
    const zd ti = Hs[i * m + i];
    const double smin = fmax(eps * cabs(ti), 1e-300);
    for (int l = 0; l < m; ++l) z[l] = czero<double>();
    z[i] = mk(1.0, 0.0);
    for (int l = i - 1; l >= 0; --l) {
      zd s = czero<double>();
      for (int c = l + 1; c <= i; ++c) s = cfma(s, Hs[l * m + c], z[c]);
      zd d = Hs[l * m + l] - ti;
      if (cabs(d) < smin) d = mk(smin, 0.0);
      z[l] = czero<double>() - cdiv(s, d);
    }
  }
  __syncthreads();
  // y_i = Q z_i, normalised; residual estimate |h_{m,m-1}| |y_i[m-1]|
  __shared__ double s_est[KS_NEVMAX];
  if (t < nev) {
    const int i = t;
    const zd *z = zz + i * m;
    double nn = 0.0;
    zd ylast = czero<double>();
    for (int r = 0; r < m; ++r) {
      zd s = czero<double>();
      for (int c = 0; c <= i; ++c) s = cfma(s, Q[r * m + c], z[c]);
      nn += abs2(s);
      if (r == m - 1) ylast = s;
    }
    const double inv = 1.0 / sqrt(nn);
    for (int r = 0; r < m; ++r) {
      zd s = czero<double>();
      for (int c = 0; c <= i; ++c) s = cfma(s, Q[r * m + c], z[c]);
      Yg[(size_t)i * KS_MMAX + r] = s * inv;
    }
    const zd th = Hs[i * m + i];
    s_est[i] = cabs(hlast) * cabs(ylast) * inv / fmax(cabs(th), 1e-300);
  }
  __syncthreads();
  // write T (row-major -> column-major global H) and Q, theta
  for (int e = t; e < m * m; e += 64) {
    const int i = e % m, j = e / m;
    Qg[(size_t)j * KS_MMAX + i] = Q[i * m + j];
  }
  for (int i = t; i < m; i += 64) THg[i] = Hs[i * m + i];
  for (int e = t; e < keep * keep; e += 64) {
    const int i = e % keep, j = e / keep;
    Hg[(size_t)j * ldh + i] = (i <= j) ? Hs[i * m + j] : czero<double>();
  }
  __syncthreads();
  if (t == 0) {
    double mx = 0.0;
    for (int i = 0; i < nev; ++i) mx = fmax(mx, s_est[i]);
    S.resmax = mx;
    S.cycles += 1;
    if (S.cycles == 1 || mx < 0.1 * S.best) { S.best = fmin(mx, S.cycles == 1 ? mx : S.best); S.stall = 0; }
    else S.stall += 1;
    S.best = fmin(S.best, mx);
    // converged, or no tenfold progress in three cycles (an ill-conditioned
    // cluster among the wanted pairs): stop this point
    S.done = (mx <= tol) ? 1 : ((S.stall >= 3) ? 2 : 0);
    st[blockIdx.x] = S;
  }
}

// Krylov-Schur truncation to `keep` columns: V <- V Q(:, 0:keep), v_keep <- v_m,
// H <- [T_keep; b^H], b^H = h_{m,m-1} Q(m-1, 0:keep).
template <int NT, int KMAX>
__global__ __launch_bounds__(NT) void k_restart(const PointDev *__restrict__ pts, zd *__restrict__ Vall,
                                                zd *__restrict__ small, KSState *__restrict__ st, int m, int keep) {
  const PointDev &P = pts[blockIdx.x];
  KSState S = st[blockIdx.x];
  if (S.done) return;
  __shared__ zd s_q[KS_MMAX * KMAX];
  const int t = threadIdx.x, nv = P.nvec;
  const size_t ldv = (size_t)nv;
  zd *V = Vall + P.vec_off;
  zd *H = small + P.small_off + SM_H;
  const zd *Qg = small + P.small_off + SM_Q;
  const int ldh = KS_MMAX + 1;
  for (int e = t; e < m * keep; e += NT) {
    const int r = e % m, c = e / m;
    s_q[r * KMAX + c] = Qg[(size_t)c * KS_MMAX + r];
  }
  const zd hlast = H[(size_t)(m - 1) * ldh + m];
  __syncthreads();
  for (int r = t; r < nv; r += NT) {
    zd acc[KMAX];
#pragma unroll
    for (int c = 0; c < KMAX; ++c) acc[c] = czero<double>();
    for (int l = 0; l < m; ++l) {
      const zd v = V[(size_t)l * ldv + r];
#pragma unroll
      for (int c = 0; c < KMAX; ++c)
        if (c < keep) acc[c] = cfma(acc[c], v, s_q[l * KMAX + c]);
    }
    const zd vm = V[(size_t)m * ldv + r];
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
      if (c < keep) V[(size_t)c * ldv + r] = acc[c];
    V[(size_t)keep * ldv + r] = vm;
  }
  __syncthreads();
  // rebuild H: T block already written by k_schur in rows/cols < keep
  for (int e = t; e < ldh * KS_MMAX; e += NT) {
    const int i = e % ldh, j = e / ldh;
    if (j >= keep || i > keep) H[e] = czero<double>();
  }
  __syncthreads();
  // b^T = h_{m,m-1} e_{m-1}^T Q(:, 0:keep)
  for (int j = t; j < keep; j += NT) H[(size_t)j * ldh + keep] = hlast * s_q[(m - 1) * KMAX + j];
  if (t == 0) { S.j0 = keep; S.k = keep; st[blockIdx.x] = S; }
}

// Final Ritz pairs: x_i = V y_i, omega_i = sigma + 1/theta_i, getevec-style
// normalisation (first entry of largest modulus among the first 5 ny
// components becomes 1, getevec.f90:179-189), FP64 relative residual.
// Whether node i lies in the farfield: the outer eighth of the Chebyshev
// nodes (node 0 .. ceil(ny/8) - 1, eta near 1); for a finite-difference
// point, whose node 0 is the wall, the nodes of the outer half of the domain,
// y >= ymax/2, with y formed as k_nodes forms it (for yi = 2/46, ymax = 50
// the Chebyshev window starts at y = 25.7 too).
__device__ inline bool farfield_node(const PointDev &P, int i) {
  if (!P.fd) return i < (P.ny + 7) / 8;
  const double dy = 1.0 / (double)(P.ny - 1);
  const double aa = P.ymax * P.yi / (P.ymax - 2.0 * P.yi), bb = 1.0 + aa / P.ymax;
  const double eta = (double)i * dy;
  return aa * eta / (bb - eta) >= 0.5 * P.ymax;
}

// Largest modulus over the farfield nodes (all five variables) of a vector
// normalised to a largest entry of modulus 1.  Discrete modes decay there,
// the continuous spectrum does not.
template <int NT>
__device__ double farfield_ratio(const PointDev &P, const zd *x, zd *red) {
  double a = 0.0, b = 0.0;
  for (int r = threadIdx.x; r < P.n; r += NT) {
    const double v = abs2(x[r]);
    b = fmax(b, v);
    if (farfield_node(P, r / 5)) a = fmax(a, v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a = fmax(a, __shfl_xor(a, off));
    b = fmax(b, __shfl_xor(b, off));
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6] = mk(a, b); }
  __syncthreads();
  for (int w = 0; w < NT / 64; ++w) { a = fmax(a, red[w].x); b = fmax(b, red[w].y); }
  __syncthreads();
  return b > 0.0 ? sqrt(a / b) : 1.0;
}

template <int NT>
__global__ __launch_bounds__(NT) void k_final(const PointDev *__restrict__ pts, const zd *__restrict__ Vall,
                                              const zd *__restrict__ small, const zd *__restrict__ coef,
                                              const double *__restrict__ ops, const double *__restrict__ norms,
                                              int m, zd *__restrict__ eval, zd *__restrict__ evec,
                                              double *__restrict__ resid, double *__restrict__ ffo, int npadmax,
                                              int nymax, int nvecmax) {
  extern __shared__ zd lds[];
  const PointDev &P = pts[blockIdx.x];
  OpLds L(lds, P, npadmax, nymax, 1, nvecmax);
  const int t = threadIdx.x, nv = P.nvec, n = P.n;
  const size_t ldv = (size_t)nv;
  const zd *V = Vall + P.vec_off;
  const zd *Yg = small + P.small_off + SM_Y;
  const zd *THg = small + P.small_off + SM_TH;
  zd *out = evec + P.out_off;
  __shared__ double s_mx[NT];
  __shared__ int s_ix[NT];
  for (int i = 0; i < P.nev; ++i) {
    const zd th = THg[i];
    const zd om = P.shift + cdiv(mk(1.0, 0.0), th);
    for (int r = t; r < nv; r += NT) {
      zd s = czero<double>();
      for (int l = 0; l < m; ++l) s = cfma(s, V[(size_t)l * ldv + r], Yg[(size_t)i * KS_MMAX + l]);
      L.z[r] = s;
    }
    __syncthreads();
    // first entry of largest modulus among the first n components
    double bm = -1.0; int bi = 0x7fffffff;
    for (int r = t; r < n; r += NT) { const double a = cabs(L.z[r]); if (a > bm) { bm = a; bi = r; } }
    s_mx[t] = bm; s_ix[t] = bi;
    __syncthreads();
    if (t == 0) {
      for (int k = 1; k < NT; ++k)
        if (s_mx[k] > bm || (s_mx[k] == bm && s_ix[k] < bi)) { bm = s_mx[k]; bi = s_ix[k]; }
      s_ix[0] = clamp_index(bi, n);   // all-NaN vector: no entry compares greater
    }
    __syncthreads();
    const zd scale = L.z[s_ix[0]];
    __syncthreads();
    if (scale_ok(scale))
      for (int r = t; r < nv; r += NT) L.z[r] = cdiv(L.z[r], scale);
    __syncthreads();
    {
      const double ff = farfield_ratio<NT>(P, L.z, L.red);
      if (t == 0 && ffo) ffo[(size_t)blockIdx.x * KS_NEVMAX + i] = ff;
    }
    // FP64 residual at the Ritz value and at the Rayleigh-quotient value of the
    // same vector; the smaller one is kept.  An operator error along the wanted
    // eigenvector (the near-null direction of A - sigma B when the shift is
    // close) moves only the Ritz value, not the vector, so the quotient with
    // the exact pencil recovers the eigenvalue.
    double rn, xn, den;
    zd om_out = om;
    if (P.itype == 1) {
      pencil_apply<NT>(P, coef, ops, FORM_AM, P.shift, L.z, L.y, L.X1, L.X2, L.red);
      pencil_apply<NT>(P, coef, ops, FORM_BM, P.shift, L.z, L.rhs, L.X1, L.X2, L.red);
      zd d3[3] = {czero<double>(), czero<double>(), czero<double>()};
      for (int r = t; r < n; r += NT) {
        d3[0] = cfmac(d3[0], L.z[r], L.y[r]);
        d3[1] = cfmac(d3[1], L.z[r], L.rhs[r]);
      }
      block_reduce<NT, 3>(d3, L.red);
      const zd om_rq = (cabs(d3[1]) > 0.0) ? cdiv(d3[0], d3[1]) : om;
      zd part[3] = {czero<double>(), czero<double>(), czero<double>()};
      for (int r = t; r < n; r += NT) {
        part[0].x += abs2(L.y[r] - om * L.rhs[r]);
        part[1].x += abs2(L.y[r] - om_rq * L.rhs[r]);
        part[2].x += abs2(L.z[r]);
      }
      block_reduce<NT, 3>(part, L.red);
      const bool use_rq = part[1].x < part[0].x;
      om_out = use_rq ? om_rq : om;
      rn = sqrt(use_rq ? part[1].x : part[0].x);
      xn = sqrt(part[2].x);
      den = sqrt(norms[blockIdx.x * 4 + 0]) + cabs(om_out) * sqrt(norms[blockIdx.x * 4 + 1]);
    } else {
      // C0 phi -> y, C1 phi -> rhs, C2 phi -> res
      pencil_apply<NT>(P, coef, ops, FORM_C0, P.shift, L.z, L.y, L.X1, L.X2, L.red);
      pencil_apply<NT>(P, coef, ops, FORM_C1, P.shift, L.z, L.rhs, L.X1, L.X2, L.red);
      pencil_apply<NT>(P, coef, ops, FORM_C2, P.shift, L.z, L.res, L.X1, L.X2, L.red);
      zd d3[3] = {czero<double>(), czero<double>(), czero<double>()};
      for (int r = t; r < n; r += NT) {
        d3[0] = cfmac(d3[0], L.z[r], L.y[r]);
        d3[1] = cfmac(d3[1], L.z[r], L.rhs[r]);
        d3[2] = cfmac(d3[2], L.z[r], L.res[r]);
      }
      block_reduce<NT, 3>(d3, L.red);
      // roots of c2 a^2 + c1 a + c0 = 0 nearest the Ritz value
      zd om_rq = om;
      if (cabs(d3[2]) > 0.0) {
        const zd bb = cdiv(d3[1], d3[2] * 2.0), cc = cdiv(d3[0], d3[2]);
        zd disc = bb * bb - cc;
        const double rr = cabs(disc);
        const double sr = sqrt(0.5 * (rr + fabs(disc.x)));
        zd sq;
        if (sr == 0.0) sq = czero<double>();
        else if (disc.x >= 0) sq = mk(sr, disc.y / (2.0 * sr));
        else sq = mk(fabs(disc.y) / (2.0 * sr), disc.y >= 0 ? sr : -sr);
        const zd r1 = czero<double>() - bb + sq, r2 = czero<double>() - bb - sq;
        om_rq = (cabs(r1 - om) < cabs(r2 - om)) ? r1 : r2;
      } else if (cabs(d3[1]) > 0.0) {
        om_rq = czero<double>() - cdiv(d3[0], d3[1]);
      }
      zd part[3] = {czero<double>(), czero<double>(), czero<double>()};
      for (int r = t; r < n; r += NT) {
        part[0].x += abs2(L.y[r] + om * L.rhs[r] + (om * om) * L.res[r]);
        part[1].x += abs2(L.y[r] + om_rq * L.rhs[r] + (om_rq * om_rq) * L.res[r]);
        part[2].x += abs2(L.z[r]);
      }
      block_reduce<NT, 3>(part, L.red);
      const bool use_rq = part[1].x < part[0].x;
      om_out = use_rq ? om_rq : om;
      rn = sqrt(use_rq ? part[1].x : part[0].x);
      xn = sqrt(part[2].x);
      const double a = cabs(om_out);
      den = sqrt(norms[blockIdx.x * 4 + 0]) + a * sqrt(norms[blockIdx.x * 4 + 1]) + a * a * sqrt(norms[blockIdx.x * 4 + 2]);
    }
    if (t == 0) {
      eval[(size_t)blockIdx.x * KS_NEVMAX + i] = om_out;
      resid[(size_t)blockIdx.x * KS_NEVMAX + i] = rn / (den * xn);
    }
    for (int r = t; r < nv; r += NT) out[(size_t)i * nv + r] = L.z[r];
    __syncthreads();
  }
}

// One shifted solve M x = b (factors from k_lu, `ir` refinement steps); used by
// the stab_shifted_solve entry point.
template <typename T, int NB, int NT>
__global__ __launch_bounds__(NT) void k_solve_one(const PointDev *__restrict__ pts, const cx<T> *__restrict__ mats,
                                                  const int *__restrict__ pivs, const cx<T> *__restrict__ blks,
                                                  const double *__restrict__ scales, const zd *__restrict__ coef,
                                                  const double *__restrict__ ops, const zd *__restrict__ b,
                                                  zd *__restrict__ x, int ir, int npadmax, int nymax) {
  extern __shared__ zd lds[];
  const PointDev &P = pts[blockIdx.x];
  OpLds L(lds, P, npadmax, nymax, 1, npadmax);
  const cx<T> *A = mats + P.mat_off;
  const int *ipiv = pivs + P.piv_off;
  const cx<T> *Lb = blks + P.blk_off, *Ub = nullptr;
  const double *dr = scales + P.scale_off, *dc = dr + P.npad;
  for (int r = threadIdx.x; r < P.n; r += NT) L.rhs[r] = b[r];
  __syncthreads();
  shifted_solve<T, NB, NT>(P, A, ipiv, Lb, Ub, dr, dc, coef, ops, ir, L);
  for (int r = threadIdx.x; r < P.n; r += NT) x[r] = L.y[r];
}

// Principal square root of a complex number.
__device__ inline zd zsqrt(zd z) {
  const double r = cabs(z);
  const double sr = sqrt(0.5 * (r + fabs(z.x)));
  if (sr == 0.0) return czero<double>();
  if (z.x >= 0) return mk(sr, z.y / (2.0 * sr));
  return mk(fabs(z.y) / (2.0 * sr), z.y >= 0 ? sr : -sr);
}

// ---------------------------------------------------------------------------
// Two-vector residual inverse iteration with a 2 x 2 Rayleigh-Ritz step, for
// temporal points whose single-vector iteration (k_invit) missed.  There the
// two eigenvalues nearest the shift are nearly equidistant from it, so one
// vector contracts towards neither (or settles on the nearer one when the
// other is wanted); a basis X = [x1, x2] converges to both at the rate
// max |lambda_{1,2} - sigma| / |lambda_3 - sigma| per step:
//   X orthonormal (FP64); projected pencil (X^H A X, X^H B X); its two
//   eigenpairs (theta_k, y_k) from det(A~ - theta B~) = 0; Ritz vectors
//   u_k = X y_k with FP64 residuals r_k = A u_k - theta_k B u_k; new basis
//   x_k = u_k - L(sigma)^-1 r_k.
// Starts from the vector k_invit left in column 0 of the point's block (x1)
// and x2 = L(sigma)^-1 B x1, on the factors k_invit used.  A pair has
// converged when its relative residual is below tol or has not halved for
// three steps while below 1e-12.  The wanted pair (want 0: the least stable
// of the pairs that pass lead2_discrete; want 1: the one nearest the shift)
// is iterated until it has converged and the other pair has converged too, or
// has a residual below 1e-10 and can no longer outrank it (lead2_competes,
// the test the host applies to the results); at most maxit steps.
// Outputs, k = 0, 1: eval2[2 p + k], resid2[2 p + k], ff2[2 p + k] and the
// Ritz vector (getevec normalisation) in column k of p's block;
// iters[p] += solves.  It runs inside k_invit after a point's first pass
// (fused) or as k_invit2 after the whole first pass.
// ---------------------------------------------------------------------------
// The discrete-mode test of the leading-mode selection for temporal points:
// farfield ratio below STAB_FARFIELD_MAX and phase speed off the continuum.
__host__ __device__ inline bool lead2_discrete(double alpha_r, zd ev, double ff) {
  if (!(ff < STAB_FARFIELD_MAX)) return false;
  if (alpha_r == 0.0) return true;
  return fabs(ev.x / alpha_r - 1.0) > STAB_CONTINUUM_BAND;
}

// Whether pair o (residual rel_o, farfield ratio ff_o), not converged, could
// still outrank the converged pair w.  Its eigenvalue error is taken as
// 1e5 rel_o max(|theta_o|, 1e-2) (a relative condition number of 1e5; the
// discrete modes of these pencils stay below it); its farfield ratio is
// trusted once rel_o < 1e-8.
__host__ __device__ inline bool lead2_competes(int want, double alpha_r, zd sigma, zd th_w, zd th_o, double rel_o,
                                               double ff_o) {
  const double err = 1e5 * rel_o * fmax(cabs(th_o), 1e-2);
  if (want == 1) return cabs(th_o - sigma) < cabs(th_w - sigma) + err;
  const bool maybe_discrete = !(rel_o < 1e-8) || lead2_discrete(alpha_r, th_o, ff_o);
  return maybe_discrete && th_o.y > th_w.y - err;
}

// The second pass for point p (P = its descriptor), x1 holding the first
// pass's vector (getevec normalisation) on entry.  x1, x2, a1, a2 and tmp2
// (2 NT entries) are LDS; b1 and b2, read only by the lane that wrote them,
// may be global.  L supplies the solve's scratch (y, rhs, tmp, w2), the
// products' (X1, X2, red) and nothing else, so x1 .. a2 may alias L.z, L.res,
// L.y and L.rhs as long as a1, a2 are not L.rhs, L.y before the first step's
// solve is done (they are written only inside the loop).
// Whether the second pass takes a point after the first (maxit1 steps):
// a temporal point whose LU succeeded and which the first pass left
// unconverged, not discrete (unless want == 1), or converged only after 3/4
// of its steps.  The host applies the same rule to the same data.
__host__ __device__ inline bool lead2_needed(int info, int itype, int flag, int want, double alpha_r, zd ev,
                                             double ff, int iters, int maxit1) {
  if (info != 0 || itype != 1) return false;
  const bool ok1 = flag == 0 && (want == 1 || lead2_discrete(alpha_r, ev, ff));
  return !(ok1 && 4 * iters < 3 * (maxit1 + 1));
}

template <typename T, int NB, int NT>
__device__ __forceinline__ void invit2_point(const PointDev &P, int p, const cx<T> *__restrict__ mats,
                                             const int *__restrict__ pivs, const cx<T> *__restrict__ blks,
                                             const double *__restrict__ scales, const zd *__restrict__ coef,
                                             const double *__restrict__ ops, const double *__restrict__ norms,
                                             int maxit, double tol, int want, zd *__restrict__ eval2,
                                             zd *__restrict__ evec, double *__restrict__ resid2,
                                             double *__restrict__ ff2, int *__restrict__ iters, OpLds &L, zd *x1,
                                             zd *x2, zd *a1, zd *b1, zd *a2, zd *b2, zd *tmp2, double *s_mx,
                                             int *s_ix) {
  const int t = threadIdx.x, n = P.n, nv = P.nvec;
  const cx<T> *A = mats + P.mat_off;
  const int *ipiv = pivs + P.piv_off;
  const cx<T> *Lb = blks + P.blk_off, *Ub = nullptr;
  const double *dr = scales + P.scale_off, *dc = dr + P.npad;
  const double nA = sqrt(norms[p * 4 + 0]), nB = sqrt(norms[p * 4 + 1]);
  const zd sg = P.shift;
  zd *out = evec + P.out_off;
  pencil_apply<NT>(P, coef, ops, FORM_BM, sg, x1, L.rhs, L.X1, L.X2, L.red);
  shifted_solve<T, NB, NT, true>(P, A, ipiv, Lb, Ub, dr, dc, coef, ops, 0, L);
  for (int r = t; r < n; r += NT) x2[r] = L.y[r];
  __syncthreads();
  int nsolve = 1;
  zd th[2] = {sg, sg};
  double rel[2] = {1e300, 1e300}, best[2] = {1e300, 1e300};
  int stall[2] = {0, 0};
  for (int it = 0;; ++it) {
    // orthonormal basis
    {
      const double n1 = sqrt(block_norm2<NT>(x1, n, L.red));
      const double i1 = (n1 > 0.0) ? 1.0 / n1 : 1.0;
      zd d[1] = {czero<double>()};
      for (int r = t; r < n; r += NT) {
        x1[r] = x1[r] * i1;
        d[0] = cfmac(d[0], x1[r], x2[r]);
      }
      block_reduce<NT, 1>(d, L.red);
      for (int r = t; r < n; r += NT) x2[r] = x2[r] - d[0] * x1[r];
      __syncthreads();
      const double n2 = sqrt(block_norm2<NT>(x2, n, L.red));
      const double i2 = (n2 > 0.0) ? 1.0 / n2 : 1.0;
      for (int r = t; r < n; r += NT) x2[r] = x2[r] * i2;
      __syncthreads();
    }
    pencil_apply<NT>(P, coef, ops, FORM_AM, sg, x1, a1, L.X1, L.X2, L.red);
    pencil_apply<NT>(P, coef, ops, FORM_BM, sg, x1, b1, L.X1, L.X2, L.red);
    pencil_apply<NT>(P, coef, ops, FORM_AM, sg, x2, a2, L.X1, L.X2, L.red);
    pencil_apply<NT>(P, coef, ops, FORM_BM, sg, x2, b2, L.X1, L.X2, L.red);
    // projected pencil: pa[i][j] = x_i^H a_j, pb[i][j] = x_i^H b_j
    zd pr[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) pr[k] = czero<double>();
    for (int r = t; r < n; r += NT) {
      const zd u = x1[r], v = x2[r];
      pr[0] = cfmac(pr[0], u, a1[r]); pr[1] = cfmac(pr[1], u, a2[r]);
      pr[2] = cfmac(pr[2], v, a1[r]); pr[3] = cfmac(pr[3], v, a2[r]);
      pr[4] = cfmac(pr[4], u, b1[r]); pr[5] = cfmac(pr[5], u, b2[r]);
      pr[6] = cfmac(pr[6], v, b1[r]); pr[7] = cfmac(pr[7], v, b2[r]);
    }
    block_reduce<NT, 8>(pr, L.red);
    const zd A00 = pr[0], A01 = pr[1], A10 = pr[2], A11 = pr[3];
    const zd B00 = pr[4], B01 = pr[5], B10 = pr[6], B11 = pr[7];
    // det(A~ - theta B~) = qa theta^2 + qb theta + qc
    const zd qa = B00 * B11 - B01 * B10;
    const zd qb = czero<double>() - (A00 * B11 + A11 * B00 - A01 * B10 - A10 * B01);
    const zd qc = A00 * A11 - A01 * A10;
    if (!(cabs(qa) > 0.0)) break;            // B~ singular: leave rel at its last value
    {
      const zd sq = zsqrt(qb * qb - qa * qc * 4.0);
      // q = -(qb + s sq) / 2 with the sign that avoids cancellation
      const zd s1 = qb + sq, s2 = qb - sq;
      const zd q = (cabs(s1) >= cabs(s2) ? s1 : s2) * -0.5;
      const zd t0 = cdiv(q, qa), t1 = (cabs(q) > 0.0) ? cdiv(qc, q) : t0;
      // keep each pair's index from the previous step (its convergence state
      // is tracked per index)
      const bool swap = it > 0 && cabs(t0 - th[1]) + cabs(t1 - th[0]) < cabs(t0 - th[0]) + cabs(t1 - th[1]);
      th[0] = swap ? t1 : t0;
      th[1] = swap ? t0 : t1;
    }
    // 2-vectors y_k with (A~ - theta_k B~) y_k = 0, unit norm
    zd y[2][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const zd m00 = A00 - th[k] * B00, m01 = A01 - th[k] * B01, m10 = A10 - th[k] * B10, m11 = A11 - th[k] * B11;
      zd y0 = czero<double>() - m01, y1 = m00;
      if (abs2(m11) + abs2(m10) > abs2(m00) + abs2(m01)) { y0 = m11; y1 = czero<double>() - m10; }
      double ny = sqrt(abs2(y0) + abs2(y1));
      if (!(ny > 0.0)) { y0 = mk(k == 0 ? 1.0 : 0.0, 0.0); y1 = mk(k == 0 ? 0.0 : 1.0, 0.0); ny = 1.0; }
      y[k][0] = y0 * (1.0 / ny);
      y[k][1] = y1 * (1.0 / ny);
    }
    // Ritz vectors into x1, x2 and residuals into a1, a2 (element-wise in place)
    zd nr[4] = {czero<double>(), czero<double>(), czero<double>(), czero<double>()};
    for (int r = t; r < n; r += NT) {
      const zd u = x1[r], v = x2[r];
      const zd u0 = y[0][0] * u + y[0][1] * v, u1 = y[1][0] * u + y[1][1] * v;
      const zd r0 = (y[0][0] * a1[r] + y[0][1] * a2[r]) - th[0] * (y[0][0] * b1[r] + y[0][1] * b2[r]);
      const zd r1 = (y[1][0] * a1[r] + y[1][1] * a2[r]) - th[1] * (y[1][0] * b1[r] + y[1][1] * b2[r]);
      x1[r] = u0; x2[r] = u1; a1[r] = r0; a2[r] = r1;
      nr[0].x += abs2(u0); nr[1].x += abs2(u1); nr[2].x += abs2(r0); nr[3].x += abs2(r1);
    }
    block_reduce<NT, 4>(nr, L.red);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double un = sqrt(nr[k].x), den = (nA + cabs(th[k]) * nB) * un;
      rel[k] = (den > 0.0) ? sqrt(nr[2 + k].x) / den : 1e300;
    }
    bool conv[2];
    double ffk[2] = {1.0, 1.0};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (rel[k] < 0.5 * best[k]) { best[k] = rel[k]; stall[k] = 0; } else if (rel[k] < 1e-12) { ++stall[k]; }
      conv[k] = rel[k] < tol || stall[k] >= 3;
    }
    // the farfield ratios enter only the choice of the wanted pair and the
    // early stop, which need a converged pair (conv comes from block-wide
    // sums, so the branch is the same for the whole work-group)
    if (conv[0] || conv[1]) {
#pragma unroll
      for (int k = 0; k < 2; ++k) ffk[k] = farfield_ratio<NT>(P, k == 0 ? x1 : x2, L.red);
    }
    int w;
    if (want == 1) {
      w = (cabs(th[0] - sg) <= cabs(th[1] - sg)) ? 0 : 1;
    } else {
      const bool d0 = lead2_discrete(P.alpha.x, th[0], ffk[0]), d1 = lead2_discrete(P.alpha.x, th[1], ffk[1]);
      w = (d0 && (!d1 || th[0].y >= th[1].y)) ? 0 : (d1 ? 1 : -1);
    }
    // early stop: the other pair is already an eigenpair to 1e-10 (so the
    // basis holds the two eigenvalues nearest the shift, not a passing mix)
    // and cannot outrank the converged wanted one
    bool done = conv[0] && conv[1];
    if (!done && w >= 0 && conv[w] && rel[w] < 1e-14 && rel[1 - w] < 1e-10) {
      const int o = 1 - w;
      done = !lead2_competes(want, P.alpha.x, sg, th[w], th[o], rel[o], ffk[o]);
    }
    if (done || it >= maxit) break;
    // corrections x_k = u_k - L(sigma)^-1 r_k: both solves in one pass over
    // the factors, in place in a1, a2 (scaled rows, zero padding)
    for (int r = t; r < P.npad; r += NT) {
      a1[r] = (r < n) ? a1[r] * dr[r] : czero<double>();
      a2[r] = (r < n) ? a2[r] * dr[r] : czero<double>();
    }
    __syncthreads();
    {
      zd *const ys[2] = {a1, a2};
      lu_solve_n<T, NT, 2, true>(P, A, ipiv, Lb, ys, tmp2, L.w2);
    }
    for (int r = t; r < n; r += NT) {
      x1[r] = x1[r] - a1[r] * dc[r];
      x2[r] = x2[r] - a2[r] * dc[r];
    }
    __syncthreads();
    nsolve += 2;
  }
  // ---- both Ritz vectors, getevec normalisation (first entry of largest
  // modulus set to 1), farfield ratio, out to columns 0 and 1
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    zd *x = (k == 0) ? x1 : x2;
    double bm = -1.0;
    int bi = 0x7fffffff;
    for (int r = t; r < n; r += NT) { const double a = abs2(x[r]); if (a > bm) { bm = a; bi = r; } }
    s_mx[t] = bm; s_ix[t] = bi;
    __syncthreads();
    if (t == 0) {
      for (int j = 1; j < NT; ++j)
        if (s_mx[j] > bm || (s_mx[j] == bm && s_ix[j] < bi)) { bm = s_mx[j]; bi = s_ix[j]; }
      s_ix[0] = clamp_index(bi, n);   // all-NaN vector: no entry compares greater
    }
    __syncthreads();
    const zd scale = x[s_ix[0]];
    __syncthreads();
    if (scale_ok(scale))
      for (int r = t; r < n; r += NT) x[r] = cdiv(x[r], scale);
    __syncthreads();
    const double ff = farfield_ratio<NT>(P, x, L.red);
    for (int r = t; r < n; r += NT) out[(size_t)k * nv + r] = x[r];
    if (t == 0) {
      eval2[2 * p + k] = th[k];
      resid2[2 * p + k] = rel[k];
      ff2[2 * p + k] = ff;
    }
  }
  if (t == 0) iters[p] += nsolve;
}


// ---------------------------------------------------------------------------
// Residual inverse iteration (A. Neumaier, "Residual inverse iteration for the
// nonlinear eigenvalue problem", SIAM J. Numer. Anal. 22 (1985) 914-923) for
// the eigenpair nearest the factored shift sigma:
//     x <- x - L(sigma)^-1 L(lambda) x,   lambda: root of x^H L(lambda) x = 0
// nearest the previous value, with L(l) = A - l B (temporal solver pencil) or
// C0 + l C1 + l^2 C2 (spatial, on phi directly).  The correction
// L(sigma)^-1 r shrinks with the residual r, so FP32 factors limit only the
// contraction per step (|lambda - sigma| / |lambda_2 - sigma| plus about
// kappa u32), not the attainable accuracy, which the FP64 residual sets.
// The first step is one plain inverse-iteration step x = L(sigma)^-1 L'(sigma) r
// from a hashed start vector.  One work-group per point.  Stops when the FP64
// relative residual is below tol, when it has not halved for three steps while
// below 1e-12 (the rounding floor), after maxit corrections, or (temporal
// points with a second pass to follow, from the sixth correction on, above
// 1e-12) when its rate over the last three steps predicts more than
// 3/4 (maxit + 1) solves in all.  Outputs per point: lambda, the
// vector normalised by its first entry of largest modulus (z = [phi;
// lambda phi] for spatial points, at out_off), the residual, the farfield
// ratio, the number of solves and a convergence flag (0, or 2 when the
// residual stayed above 1e-14).  pass2: 0 when no second pass follows, 1
// when k_invit2 follows, 2 when the work-group itself then runs the second
// pass (invit2_point) for a point that needs it (the launch then holds 2 NT
// more LDS entries).
// ---------------------------------------------------------------------------
template <typename T, int NB, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void k_invit(const PointDev *__restrict__ pts, const cx<T> *__restrict__ mats,
                                              const int *__restrict__ pivs, const cx<T> *__restrict__ blks,
                                              const double *__restrict__ scales, const zd *__restrict__ coef,
                                              const double *__restrict__ ops, const double *__restrict__ norms,
                                              int maxit, double tol, zd *__restrict__ eval, zd *__restrict__ evec,
                                              double *__restrict__ resid, double *__restrict__ ffo,
                                              int *__restrict__ iters, int *__restrict__ flag, int npadmax,
                                              int nymax, const int *__restrict__ order, int pass2, int want,
                                              const int *__restrict__ info, zd *__restrict__ eval2,
                                              double *__restrict__ resid2, double *__restrict__ ff2) {
  extern __shared__ zd lds[];
  const int pi = order ? order[blockIdx.x] : blockIdx.x;   // largest points first
  const PointDev &P = pts[pi];
  OpLds L(lds, P, npadmax, nymax, 1, npadmax);
  const int t = threadIdx.x, n = P.n;
  const cx<T> *A = mats + P.mat_off;
  const int *ipiv = pivs + P.piv_off;
  const cx<T> *Lb = blks + P.blk_off, *Ub = nullptr;
  const double *dr = scales + P.scale_off, *dc = dr + P.npad;
  const bool temporal = (P.itype == 1);
  const double nA = sqrt(norms[pi * 4 + 0]), nB = sqrt(norms[pi * 4 + 1]),
               nC = sqrt(norms[pi * 4 + 2]);
  const zd sg = P.shift;
  zd *x = L.z;
  __shared__ double s_mx[NT];
  __shared__ int s_ix[NT];

  // ---- start: x = L(sigma)^-1 L'(sigma) r
  const unsigned seed = point_seed(P);
  for (int r = t; r < n; r += NT) x[r] = mk(hash01(seed, 2 * r), hash01(seed, 2 * r + 1));
  __syncthreads();
  if (temporal) {
    pencil_apply<NT>(P, coef, ops, FORM_BM, sg, x, L.rhs, L.X1, L.X2, L.red);
  } else {
    pencil_apply<NT>(P, coef, ops, FORM_C1, sg, x, L.rhs, L.X1, L.X2, L.red);
    pencil_apply<NT>(P, coef, ops, FORM_C2, sg, x, L.res, L.X1, L.X2, L.red);
    for (int r = t; r < n; r += NT) L.rhs[r] = L.rhs[r] + (sg * 2.0) * L.res[r];
    __syncthreads();
  }
  shifted_solve<T, NB, NT, true>(P, A, ipiv, Lb, Ub, dr, dc, coef, ops, 0, L);
  for (int r = t; r < n; r += NT) x[r] = L.y[r];
  __syncthreads();
  int nsolve = 1;

  zd lam = sg;
  double rel = 1e300, best = 1e300;
  double back[3] = {1e300, 1e300, 1e300};   // rel one, two and three steps back
  int stall = 0;
  for (int it = 0;; ++it) {
    // unit 2-norm
    {
      const double nx = sqrt(block_norm2<NT>(x, n, L.red));
      const double inv = (nx > 0.0) ? 1.0 / nx : 1.0;
      for (int r = t; r < n; r += NT) x[r] = x[r] * inv;
      __syncthreads();
    }
    // pencil products and the new lambda
    zd d3[3] = {czero<double>(), czero<double>(), czero<double>()};
    if (temporal) {
      pencil_apply<NT>(P, coef, ops, FORM_AM, sg, x, L.y, L.X1, L.X2, L.red);
      pencil_apply<NT>(P, coef, ops, FORM_BM, sg, x, L.rhs, L.X1, L.X2, L.red);
      for (int r = t; r < n; r += NT) {
        d3[0] = cfmac(d3[0], x[r], L.y[r]);
        d3[1] = cfmac(d3[1], x[r], L.rhs[r]);
      }
      block_reduce<NT, 3>(d3, L.red);
      if (cabs(d3[1]) > 0.0) lam = cdiv(d3[0], d3[1]);
    } else {
      pencil_apply<NT>(P, coef, ops, FORM_C0, sg, x, L.y, L.X1, L.X2, L.red);
      pencil_apply<NT>(P, coef, ops, FORM_C1, sg, x, L.rhs, L.X1, L.X2, L.red);
      pencil_apply<NT>(P, coef, ops, FORM_C2, sg, x, L.res, L.X1, L.X2, L.red);
      for (int r = t; r < n; r += NT) {
        d3[0] = cfmac(d3[0], x[r], L.y[r]);
        d3[1] = cfmac(d3[1], x[r], L.rhs[r]);
        d3[2] = cfmac(d3[2], x[r], L.res[r]);
      }
      block_reduce<NT, 3>(d3, L.red);
      if (cabs(d3[2]) > 0.0) {
        const zd bb = cdiv(d3[1], d3[2] * 2.0), cc = cdiv(d3[0], d3[2]);
        const zd disc = bb * bb - cc;
        const double rr = cabs(disc);
        const double sr = sqrt(0.5 * (rr + fabs(disc.x)));
        zd sq;
        if (sr == 0.0) sq = czero<double>();
        else if (disc.x >= 0) sq = mk(sr, disc.y / (2.0 * sr));
        else sq = mk(fabs(disc.y) / (2.0 * sr), disc.y >= 0 ? sr : -sr);
        const zd r1 = czero<double>() - bb + sq, r2 = czero<double>() - bb - sq;
        lam = (cabs(r1 - lam) < cabs(r2 - lam)) ? r1 : r2;
      } else if (cabs(d3[1]) > 0.0) {
        lam = czero<double>() - cdiv(d3[0], d3[1]);
      }
    }
    // residual r = L(lambda) x into rhs
    zd part[1] = {czero<double>()};
    if (temporal) {
      for (int r = t; r < n; r += NT) {
        const zd v = L.y[r] - lam * L.rhs[r];
        L.rhs[r] = v;
        part[0].x += abs2(v);
      }
    } else {
      const zd l2 = lam * lam;
      for (int r = t; r < n; r += NT) {
        const zd v = L.y[r] + lam * L.rhs[r] + l2 * L.res[r];
        L.rhs[r] = v;
        part[0].x += abs2(v);
      }
    }
    block_reduce<NT, 1>(part, L.red);
    const double al = cabs(lam);
    const double den = temporal ? (nA + al * nB) : (nA + al * nB + al * al * nC);
    rel = sqrt(part[0].x) / den;             // ||x|| = 1
    // the stall count only runs near the rounding floor: far from it the
    // residual can grow for a step or two while the vector turns
    if (rel < 0.5 * best) { best = rel; stall = 0; } else if (rel < 1e-12) { ++stall; }
    // a temporal point that, at its rate over the last three steps, would
    // reach tol only after 3/4 of maxit + 1 solves (or not at all) leaves now
    // when a second pass follows: that pass takes such points whatever this
    // one gives them (a second eigenvalue nearly as close to the shift), so
    // the steps would be wasted
    bool slow = false;
    if (pass2 && temporal && it >= 6 && rel >= 1e-12 && rel >= tol) {
      const double rate = (back[2] > 0.0 && back[2] < 1e300) ? cbrt(rel / back[2]) : 1.0;
      slow = !(rate < 1.0) || nsolve + log(tol / rel) / log(rate) > 0.75 * (maxit + 1);
    }
    back[2] = back[1]; back[1] = back[0]; back[0] = rel;
    if (rel < tol || it >= maxit || stall >= 3 || slow) break;
    // correction
    shifted_solve<T, NB, NT, true>(P, A, ipiv, Lb, Ub, dr, dc, coef, ops, 0, L);
    ++nsolve;
    for (int r = t; r < n; r += NT) x[r] = x[r] - L.y[r];
    __syncthreads();
  }
  // ---- getevec normalisation: first entry of largest modulus set to 1
  double bm = -1.0;
  int bi = 0x7fffffff;
  for (int r = t; r < n; r += NT) { const double a = abs2(x[r]); if (a > bm) { bm = a; bi = r; } }
  s_mx[t] = bm; s_ix[t] = bi;
  __syncthreads();
  if (t == 0) {
    for (int k = 1; k < NT; ++k)
      if (s_mx[k] > bm || (s_mx[k] == bm && s_ix[k] < bi)) { bm = s_mx[k]; bi = s_ix[k]; }
    s_ix[0] = clamp_index(bi, n);     // all-NaN vector: no entry compares greater
  }
  __syncthreads();
  const zd scale = x[s_ix[0]];
  __syncthreads();
  if (scale_ok(scale))
    for (int r = t; r < n; r += NT) x[r] = cdiv(x[r], scale);
  __syncthreads();
  const double ff = farfield_ratio<NT>(P, x, L.red);
  zd *out = evec + P.out_off;
  for (int r = t; r < n; r += NT) {
    out[r] = x[r];
    if (!temporal) out[n + r] = lam * x[r];
  }
  if (t == 0) {
    eval[pi] = lam;
    resid[pi] = rel;
    ffo[pi] = ff;
    iters[pi] = nsolve;
    flag[pi] = (rel < 1e-14) ? 0 : 2;
  }
  // ---- the second pass in the same work-group (pass2 == 2), for the points the
  // rule of k_invit2 takes, with 2 maxit steps: it starts as soon as this
  // point's first pass ends, beside the other points' first passes.  x1 is
  // x (L.z), x2 L.res, a1 L.y, a2 L.rhs, b1 and b2 columns 2 and 3 of the
  // point's vector block, tmp2 the 2 NT entries after the operator layout.
  if (pass2 == 2 && lead2_needed(info[pi], P.itype, (rel < 1e-14) ? 0 : 2, want, P.alpha.x, lam, ff, nsolve, maxit)) {
    zd *tmp2 = lds + op_lds_bytes(npadmax, nymax, 1, npadmax) / sizeof(zd);
    const size_t nv = P.nvec;
    invit2_point<T, NB, NT>(P, pi, mats, pivs, blks, scales, coef, ops, norms, 2 * maxit, tol, want, eval2, evec,
                            resid2, ff2, iters, L, x, L.res, L.y, out + 2 * nv, L.rhs, out + 3 * nv, tmp2, s_mx,
                            s_ix);
  }
}

// The second pass as a kernel of its own, after k_invit: one work-group per
// listed point p = sub[blockIdx.x], or per point order[blockIdx.x] that the
// rule takes.  LDS: the full operator layout followed by six npadmax vectors
// and 2 NT entries.
template <typename T, int NB, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void k_invit2(const PointDev *__restrict__ pts, const int *__restrict__ sub,
                                               const cx<T> *__restrict__ mats, const int *__restrict__ pivs,
                                               const cx<T> *__restrict__ blks, const double *__restrict__ scales,
                                               const zd *__restrict__ coef, const double *__restrict__ ops,
                                               const double *__restrict__ norms, int maxit, double tol, int want,
                                               zd *__restrict__ eval2, zd *__restrict__ evec,
                                               double *__restrict__ resid2, double *__restrict__ ff2,
                                               int *__restrict__ iters, int npadmax, int nymax,
                                               const int *__restrict__ order, const int *__restrict__ info,
                                               const int *__restrict__ flag, const zd *__restrict__ eval1,
                                               const double *__restrict__ ff1, int maxit1) {
  extern __shared__ zd lds[];
  // sub lists the points; without it work-group g takes point order[g] and
  // runs only when the first pass (maxit1 steps) left that temporal point
  // unconverged, not discrete (unless want == 1), or converged only after
  // 3/4 of its steps -- the host applies the same rule to the same data
  const int p = sub ? sub[blockIdx.x] : order[blockIdx.x];
  if (!sub) {
    const PointDev &Q = pts[p];
    if (!lead2_needed(info[p], Q.itype, flag[p], want, Q.alpha.x, eval1[p], ff1[p], iters[p], maxit1)) return;
  }
  const PointDev &P = pts[p];
  OpLds L(lds, P, npadmax, nymax, 1, npadmax);
  zd *x1 = L.X2 + 5 * nymax;
  zd *x2 = x1 + npadmax, *a1 = x2 + npadmax, *b1 = a1 + npadmax, *a2 = b1 + npadmax, *b2 = a2 + npadmax;
  zd *tmp2 = b2 + npadmax;                     // 2 NT entries for the two-column solve
  __shared__ double s_mx[NT];
  __shared__ int s_ix[NT];
  zd *out = evec + P.out_off;
  for (int r = threadIdx.x; r < P.n; r += NT) x1[r] = out[r];
  __syncthreads();
  invit2_point<T, NB, NT>(P, p, mats, pivs, blks, scales, coef, ops, norms, maxit, tol, want, eval2, evec, resid2,
                          ff2, iters, L, x1, x2, a1, b1, a2, b2, tmp2, s_mx, s_ix);
}

// Copy column sel[p] of each point's eigenvector block (P.out_off, nvec rows,
// ld nvec) to dst + doff[p].  grid: points; sel[p] < 0 skips the point.
__global__ __launch_bounds__(256) void k_gather_cols(const PointDev *__restrict__ pts, const zd *__restrict__ src,
                                                     const int *__restrict__ sel, const long long *__restrict__ doff,
                                                     zd *__restrict__ dst) {
  const PointDev &P = pts[blockIdx.x];
  const int k = sel[blockIdx.x];
  if (k < 0) return;
  const zd *s = src + P.out_off + (size_t)k * P.nvec;
  zd *d = dst + doff[blockIdx.x];
  for (int r = threadIdx.x; r < P.nvec; r += 256) d[r] = s[r];
}
