// Batched blocked complex LU with partial pivoting, one work-group per matrix.
//
// Right-looking, panel width NB.  The panel rows live in registers of the
// threads that own them (row r belongs to thread r % NT), so the pivot search,
// the row interchanges inside the panel and the rank-1 updates never touch
// memory.  The trailing matrix is then processed in strips of SW columns:
// the panel's row interchanges are applied to the strip through LDS, U12 is
// formed as inv(L11) * A12 (inv(L11) computed once per panel), and the strip
// is updated with A22 -= L21 U12, L21 read from registers and U12 broadcast
// from LDS.  The interchanges are applied to the columns left of the panel
// too, so P A = L U with P given by perm[] as well as by ipiv[].
// Outputs: L and U in place, ipiv (absolute row indices, as ZGETRF) and perm
// (ipiv + n).  k_blkinv then inverts the 64 x 64 diagonal blocks used by the
// blocked triangular solves.
#include "common.h"

// Elements of the diagonal-block inverse storage of one point (see k_blkinv).
__host__ __device__ inline long long blk_elems(int npad, int bs) {
  const long long nblk = (npad + bs - 1) / bs;
  return 2LL * nblk * bs * bs;
}

template <typename T, int NB, int NT, int R, int SW>
__global__ __launch_bounds__(NT) void k_lu(const PointDev *__restrict__ pts, cx<T> *__restrict__ mats,
                                           int *__restrict__ pivs, cx<T> *__restrict__ blks,
                                           int *__restrict__ info, const int *__restrict__ active) {
  typedef cx<T> C;
  if (active && !active[blockIdx.x]) return;
  const PointDev &P = pts[blockIdx.x];
  const int n = P.npad, lda = P.npad;
  C *A = mats + P.mat_off;
  int *ipiv = pivs + P.piv_off;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  constexpr int NW = NT / 64;

  __shared__ C s_prow[NB], s_crow[NB];
  __shared__ T s_rv[NW];
  __shared__ int s_ri[NW];
  __shared__ C s_blk[NB * NB];
  __shared__ C s_linv[NB * NB];
  __shared__ int s_piv[NB];
  __shared__ int s_srow[2 * NB], s_sorig[2 * NB];
  __shared__ int s_topsrc[NB], s_outrow[NB], s_outsrc[NB];
  __shared__ int s_nout, s_ns;
  __shared__ C s_stage[2 * NB * SW];
  __shared__ C s_u12[NB * SW];

  C Pn[R][NB];
  int myinfo = 0;

  for (int k0 = 0; k0 < n; k0 += NB) {
    // ---- load the panel rows owned by this thread
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int r = t + j * NT;
#pragma unroll
      for (int c = 0; c < NB; ++c)
        Pn[j][c] = (r >= k0 && r < n) ? A[(size_t)(k0 + c) * lda + r] : czero<T>();
    }
    // ---- factor the panel.  The column loop is not unrolled; every register
    // access below uses compile-time indices with predicates, so Pn stays in
    // VGPRs.
#pragma unroll 1
    for (int c = 0; c < NB; ++c) {
      const int col = k0 + c;
      T best = (T)-1;
      int bi = 0x7fffffff;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = t + j * NT;
        C pc = czero<T>();
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) if (cc == c) pc = Pn[j][cc];
        if (r >= col && r < n) {
          const T v = cabs1(pc);
          if (v > best) { best = v; bi = r; }
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const T ov = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) { s_rv[wv] = best; s_ri[wv] = bi; }
      __syncthreads();
      best = s_rv[0]; bi = s_ri[0];
#pragma unroll
      for (int w = 1; w < NW; ++w) {
        if (s_rv[w] > best || (s_rv[w] == best && s_ri[w] < bi)) { best = s_rv[w]; bi = s_ri[w]; }
      }
      // no finite candidate (NaN column): keep the diagonal so every index stays in range
      const int p = (bi >= col && bi < n) ? bi : col;
      const int pj = p / NT, cj = col / NT;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if (j == pj && (p % NT) == t) {
#pragma unroll
          for (int cc = 0; cc < NB; ++cc) s_prow[cc] = Pn[j][cc];
        }
        if (j == cj && (col % NT) == t) {
#pragma unroll
          for (int cc = 0; cc < NB; ++cc) s_crow[cc] = Pn[j][cc];
        }
      }
      __syncthreads();
      if (p != col) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          if (j == pj && (p % NT) == t) {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) Pn[j][cc] = s_crow[cc];
          }
          if (j == cj && (col % NT) == t) {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) Pn[j][cc] = s_prow[cc];
          }
        }
      }
      const C piv = s_prow[c];
      if (t == 0) s_piv[c] = p;
      if (piv.x == (T)0 && piv.y == (T)0) {
        if (myinfo == 0) myinfo = col + 1;
      } else {
        const C inv = cdiv(mk<T>(1, 0), piv);
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const int r = t + j * NT;
          if (r > col && r < n) {
            C l = czero<T>();
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) if (cc == c) l = Pn[j][cc];
            l = l * inv;
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) {
              if (cc == c) Pn[j][cc] = l;
              else if (cc > c) Pn[j][cc] = cfms(Pn[j][cc], l, s_prow[cc]);
            }
          }
        }
      }
      __syncthreads();
    }
    // ---- write the factored panel back; stage the diagonal block
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int r = t + j * NT;
      if (r >= k0 && r < n) {
#pragma unroll
        for (int c = 0; c < NB; ++c) A[(size_t)(k0 + c) * lda + r] = Pn[j][c];
        if (r < k0 + NB) {
#pragma unroll
          for (int c = 0; c < NB; ++c) s_blk[(r - k0) * NB + c] = Pn[j][c];
        }
      }
    }
    __syncthreads();
    // ---- inv(L11) (unit lower), one column per thread, for U12 = inv(L11) A12
    if (t < NB) {
      const int c = t;
      for (int i = 0; i < NB; ++i) {
        C s = (i == c) ? mk<T>(1, 0) : czero<T>();
        if (i > c)
          for (int k = c; k < i; ++k) s = cfms(s, s_blk[i * NB + k], s_linv[k * NB + c]);
        s_linv[i * NB + c] = (i < c) ? czero<T>() : s;
      }
    }
    __syncthreads();
    if (t < NB) ipiv[k0 + t] = s_piv[t];
    // ---- net permutation of the panel's interchanges (rows touched, origins)
    if (t == 0) {
      int ns = 0;
      for (int c = 0; c < NB; ++c) {
        const int a = k0 + c, b = s_piv[c];
        if (a == b) continue;
        int ia = -1, ib = -1;
        for (int e = 0; e < ns; ++e) { if (s_srow[e] == a) ia = e; if (s_srow[e] == b) ib = e; }
        if (ia < 0) { ia = ns; s_srow[ns] = a; s_sorig[ns] = a; ++ns; }
        if (ib < 0) { ib = ns; s_srow[ns] = b; s_sorig[ns] = b; ++ns; }
        const int tmp = s_sorig[ia]; s_sorig[ia] = s_sorig[ib]; s_sorig[ib] = tmp;
      }
      for (int c = 0; c < NB; ++c) s_topsrc[c] = k0 + c;
      int no = 0;
      for (int e = 0; e < ns; ++e) {
        const int row = s_srow[e];
        if (row < k0 + NB) s_topsrc[row - k0] = s_sorig[e];
        else if (s_sorig[e] != row) { s_outrow[no] = row; s_outsrc[no] = s_sorig[e]; ++no; }
      }
      s_nout = no;
      s_ns = ns;
    }
    __syncthreads();
    // ---- trailing strips
    const int nout = s_nout, nslot = NB + nout;
    for (int c0 = k0 + NB; c0 < n; c0 += SW) {
      for (int e = t; e < nslot * SW; e += NT) {
        const int slot = e % nslot, cc = e / nslot;
        const int src = (slot < NB) ? s_topsrc[slot] : s_outsrc[slot - NB];
        s_stage[slot * SW + cc] = A[(size_t)(c0 + cc) * lda + src];
      }
      __syncthreads();
      for (int e = t; e < nout * SW; e += NT) {
        const int o = e % nout, cc = e / nout;
        A[(size_t)(c0 + cc) * lda + s_outrow[o]] = s_stage[(NB + o) * SW + cc];
      }
      for (int e = t; e < NB * SW; e += NT) {
        const int i = e / SW, cc = e - i * SW;
        C u = czero<T>();
        for (int k = 0; k <= i; ++k) u = cfma(u, s_linv[i * NB + k], s_stage[k * SW + cc]);
        s_u12[i * SW + cc] = u;
        A[(size_t)(c0 + cc) * lda + k0 + i] = u;
      }
      __syncthreads();
      // A22 strip -= L21 U12: R rows x G columns of independent accumulators
      constexpr int G = (R >= 2) ? 2 : 4;
#pragma unroll 1
      for (int cg = 0; cg < SW; cg += G) {
        C acc[R][G];
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const int r = t + j * NT;
          const bool live = (r >= k0 + NB && r < n);
#pragma unroll
          for (int q = 0; q < G; ++q) acc[j][q] = live ? A[(size_t)(c0 + cg + q) * lda + r] : czero<T>();
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          C u[G];
#pragma unroll
          for (int q = 0; q < G; ++q) u[q] = s_u12[k * SW + cg + q];
#pragma unroll
          for (int j = 0; j < R; ++j)
#pragma unroll
            for (int q = 0; q < G; ++q) acc[j][q] = cfms(acc[j][q], Pn[j][k], u[q]);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const int r = t + j * NT;
          if (r >= k0 + NB && r < n) {
#pragma unroll
            for (int q = 0; q < G; ++q) A[(size_t)(c0 + cg + q) * lda + r] = acc[j][q];
          }
        }
      }
      __syncthreads();
    }
    // ---- the same interchanges on the columns left of the panel, so that L
    // ends in the final row order (solves then start with one gather)
    const int ns = s_ns;
    if (ns > 0) {
      for (int c0 = 0; c0 < k0; c0 += SW) {
        for (int e = t; e < ns * SW; e += NT) {
          const int slot = e % ns, cc = e / ns;
          s_stage[slot * SW + cc] = A[(size_t)(c0 + cc) * lda + s_sorig[slot]];
        }
        __syncthreads();
        for (int e = t; e < ns * SW; e += NT) {
          const int slot = e % ns, cc = e / ns;
          A[(size_t)(c0 + cc) * lda + s_srow[slot]] = s_stage[slot * SW + cc];
        }
        __syncthreads();
      }
    }
  }
  // ---- final row permutation: row r of L U is row perm[r] of the input
  int *perm = ipiv + n;
  for (int r = t; r < n; r += NT) perm[r] = r;
  __syncthreads();
  if (t == 0) {
    for (int c = 0; c < n; ++c) {
      const int p = ipiv[c];
      if (p != c) { const int a = perm[c]; perm[c] = perm[p]; perm[p] = a; }
    }
  }
  if (t == 0) info[blockIdx.x] = myinfo;
}

// Inverses of the diagonal blocks of L (unit lower) and U (upper) of order
// BS (the last block may be shorter and is stored padded to BS x BS), in
// FP64, stored as T, row-major: with nblk = ceil(npad / BS), block b of
// inv(L) at blk_off + b BS^2 and of inv(U) at blk_off + (nblk + b) BS^2.  The
// per-point allocation is 2 nblk BS^2 elements (blk_elems below).
// grid: x = block index, y = point.  One column per thread.
template <typename T, int BS>
__global__ __launch_bounds__(BS) void k_blkinv(const PointDev *__restrict__ pts, const cx<T> *__restrict__ mats,
                                               cx<T> *__restrict__ blks) {
  const PointDev &P = pts[blockIdx.y];
  const int n = P.npad, k0 = blockIdx.x * BS;
  if (k0 >= n) return;
  const int bs = min(BS, n - k0), t = threadIdx.x;
  const int nblk = (n + BS - 1) / BS;
  const cx<T> *A = mats + P.mat_off;
  cx<T> *Li = blks + P.blk_off + (size_t)blockIdx.x * BS * BS;
  cx<T> *Ui = blks + P.blk_off + (size_t)(nblk + blockIdx.x) * BS * BS;
  __shared__ zd s_a[BS * BS];     // row-major copy of the block
  __shared__ zd s_x[BS * BS];     // column t of the inverse being built
  for (int e = t; e < BS * BS; e += BS) {
    const int i = e / BS, j = e % BS;
    s_a[e] = (i < bs && j < bs) ? cvt<double>(A[(size_t)(k0 + j) * n + k0 + i]) : czero<double>();
  }
  __syncthreads();
  const int c = t;
  // L^-1 column c: forward substitution, unit diagonal
  for (int i = 0; i < BS; ++i) {
    zd s = (i == c) ? mk(1.0, 0.0) : czero<double>();
    if (i > c && i < bs)
      for (int k = c; k < i; ++k) s = cfms(s, s_a[i * BS + k], s_x[k * BS + c]);
    s_x[i * BS + c] = (i < c || i >= bs) ? czero<double>() : s;
  }
  for (int i = 0; i < BS; ++i) Li[i * BS + c] = cvt<T>(s_x[i * BS + c]);
  __syncthreads();
  // U^-1 column c: back substitution
  for (int i = BS - 1; i >= 0; --i) {
    zd s = (i == c) ? mk(1.0, 0.0) : czero<double>();
    if (i < c && c < bs)
      for (int k = i + 1; k <= c; ++k) s = cfms(s, s_a[i * BS + k], s_x[k * BS + c]);
    s_x[i * BS + c] = (i > c || c >= bs) ? czero<double>() : cdiv(s, s_a[i * BS + i]);
  }
  for (int i = 0; i < BS; ++i) Ui[i * BS + c] = cvt<T>(s_x[i * BS + c]);
}
