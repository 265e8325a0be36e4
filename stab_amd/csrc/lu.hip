// Batched blocked complex LU with partial pivoting, one work-group per matrix
// (the getrf structure: factor a block column, swap, triangular solve, update).
//
// Columns are taken in blocks of KB = 64 (the last may be 32 wide).  Inside a
// block, panels of NB columns (16 in FP32, 8 in FP64) are factored with their rows in registers
// of the threads that own them (row r belongs to thread r % NT), so the pivot
// search, the row interchanges inside the panel and the rank-1 updates never
// touch memory; each panel then updates the rest of its block column (all
// rows) in strips of SW columns, applying its interchanges through LDS and
// forming its rows of U as inv(L11) * A12.  When the block column is done:
//   * its net row permutation is applied once to the trailing columns and,
//     for the LAPACK form only (left = 1), to the columns left of it;
//     otherwise it is stored as the block's interchange record for the solves,
//   * X = inv(L11) is formed once (four lanes per column, shuffles) and
//     stored as the block's inverse for the triangular solves, and U12 = X A12
//     runs on the matrix cores, and
//   * A22 -= L21 U12 with depth KB on the matrix cores, complex arithmetic as
//     2 x 2 real blocks, [Cr; Ci] -= [[Lr, -Li], [Li, Lr]] [Ur; Ui], i.e.
//     four real MFMA products per tile, computed transposed (the MFMA row index
//     runs over matrix columns) so that each accumulator register of 16 lanes
//     covers 16 consecutive matrix rows and loads/stores coalesce.
// The trailing matrix is read and written once per 64 columns.  P A = L U
// with P given by ipiv[].  Outputs: L and U in place, ipiv (absolute row
// indices, as ZGETRF), the pivot mode word and records (common.h), info and
// the inverses of the 64 x 64 diagonal blocks of L; k_blkinv then inverts
// those of U.  With left = 1 (pivot mode 0, the LAPACK form) L is in final
// row order and perm (ipiv + n, P as a row list) is stored too; with left = 0
// the columns left of each block keep their rows (mode 1).
#include "common.h"

// 16 x 16 x 4 MFMA: D[m][n] += sum_k A[m][k] B[k][n]; lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16].  Accumulator register i of lane l
// holds D[m][l % 16] with m = 4 (l / 16) + i (f32) or m = 4 i + l / 16 (f64).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row_of(int lane, int i) { return 4 * i + (lane >> 4); }
  static constexpr int row_step(int i) { return 4 * i; }   // row_of(lane, i) - row_of(lane, 0)
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row_of(int lane, int i) { return 4 * (lane >> 4) + i; }
  static constexpr int row_step(int i) { return i; }       // row_of(lane, i) - row_of(lane, 0)
};

// Columns of one LDS tile of U12 in the MFMA update (f64: 32, f32: 64).
template <typename T> __host__ __device__ constexpr int lu_gw() { return sizeof(T) == 8 ? 32 : 64; }

// Elements of the diagonal-block inverse storage of one point (see k_blkinv).
__host__ __device__ inline long long blk_elems(int npad, int bs) {
  const long long nblk = (npad + bs - 1) / bs;
  return 2LL * nblk * bs * bs;
}

// Net permutation of the interchanges ipiv[k0 .. k0+w) (absolute rows, each
// >= its step): touched rows srow[e] receive the input row sorig[e];
// returns the number of touched rows (<= 2 w).  One thread; slot[] (one
// entry per row, -1 when the row is not in the list) finds a row's entry in
// O(1) and is -1 again on return.
__device__ inline int net_perm(const int *ipiv, int k0, int w, int *srow, int *sorig, int *slot) {
  int ns = 0;
  for (int c = 0; c < w; ++c) {
    const int a = k0 + c, b = ipiv[c];
    if (a == b) continue;
    int ia = slot[a], ib = slot[b];
    if (ia < 0) { ia = ns; srow[ns] = a; sorig[ns] = a; slot[a] = ns; ++ns; }
    if (ib < 0) { ib = ns; srow[ns] = b; sorig[ns] = b; slot[b] = ns; ++ns; }
    const int tmp = sorig[ia]; sorig[ia] = sorig[ib]; sorig[ib] = tmp;
  }
  for (int e = 0; e < ns; ++e) slot[srow[e]] = -1;
  return ns;
}

// The batched LU in two launches per 64-column block (K0 = 0, 64, ...) over
// the whole batch:
//   k_lu_panel  one work-group per matrix whose order exceeds K0: factors the
//               block column [K0, K0 + kb) (panels in registers, the in-block
//               strips, on the matrix cores in FP32, the interchanges),
//               stores the block's records, and forms X = inv(L11); thread t
//               holds rows K0 + t + j NT, j < R, so R = ceil((n - K0) / NT)
//               suffices and later block steps launch with a smaller R;
//   k_lu_trail  one work-group per 64-column strip (32 in FP64) right of the
//               block, for every such matrix: the block's row permutation,
//               U12 = X A12 and A22 -= L21 U12 on the matrix cores.
// Split this way, the strips of one matrix run on several CUs at once and the
// latency-bound panel steps of some matrices overlap the matrix-core updates
// of others; the arithmetic is that of one work-group doing both in turn.
template <typename T, int NB, int NT, int R, int SW>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(R <= 2 && NT == 256 ? 3 : 2))) void k_lu_panel(const PointDev *__restrict__ pts, cx<T> *__restrict__ mats,
                                                 int *__restrict__ pivs, cx<T> *__restrict__ blks,
                                                 int *__restrict__ info, const int *__restrict__ order, int left,
                                                 int K0) {
  typedef cx<T> C;
  // work-group -> point: `order` lists the points largest first, so the
  // batch's last work-groups are its cheapest
  const int pi = order[blockIdx.x];
  const PointDev &P = pts[pi];
  const int n = P.npad, lda = P.npad;
  if (K0 >= n) return;
  C *A = mats + P.mat_off;
  int *ipiv = pivs + P.piv_off;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  constexpr int NW = NT / 64;
  constexpr int KB = 64, SU = 2 * KB * 16;
  static_assert(KB % NB == 0, "panel width must divide the block width");
  static_assert(LU_PREC == 1 + 4 * KB, "interchange record: count, 2 KB rows, 2 KB sources");
  static_assert(LU_TREC == 1 + 3 * KB, "trailing record: count, KB top sources, KB out rows, KB out sources");

  // double-buffered by column parity: a column's slots are rewritten two
  // columns later, after a barrier every wave passes only once it has read
  // them.  s_cand2: each wave's best candidate row, s_crow2: the row of the
  // diagonal; both are published before the column's one barrier.
  __shared__ C s_cand2[2][NW][NB], s_crow2[2][NB];
  __shared__ T s_rv2[2][NW];
  __shared__ int s_ri2[2][NW];
  __shared__ C s_blk[NB * NB];
  __shared__ C s_linv[NB * NB];           // inv(L11) of the current panel
  __shared__ int s_piv[KB];
  __shared__ int s_srow[2 * KB], s_sorig[2 * KB];
  __shared__ int s_slot[1280];            // net_perm's row map (npad <= 1280)
  __shared__ int s_topsrc[KB], s_outrow[KB], s_outsrc[KB];
  __shared__ int s_nout, s_ns;
  __shared__ C s_stage[2 * NB * SW];
  __shared__ C s_u12[NB * SW];
  // packed strictly lower L11 for X (KB (KB - 1) / 2 entries), staging of the
  // left-column interchanges (at most 2 KB rows x 16 columns), the final row list
  __shared__ C s_u[SU];

  C Pn[R][NB];
  int myinfo = (K0 == 0) ? 0 : info[pi];   // the first zero pivot of the earlier blocks
  for (int r = t; r < 1280; r += NT) s_slot[r] = -1;
  __syncthreads();

  {
    const int kb = min(KB, n - K0);       // 64, or 32 for the last block (n is a multiple of 32)
    const int kend = K0 + kb;
    for (int k0 = K0; k0 < kend; k0 += NB) {
      // ---- load the panel rows owned by this thread
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = K0 + t + j * NT;
#pragma unroll
        for (int c = 0; c < NB; ++c)
        {
          // unconditional load from a row in range, then select: a load under
          // a divergent branch waits for its data before the branch joins
          const bool ok = r >= k0 && r < n;
          const C v = A[(size_t)(k0 + c) * lda + (ok ? r : k0)];
          Pn[j][c] = ok ? v : czero<T>();
        }
      }
      // ---- factor the panel.  Every register access below has compile-time
      // column indices (with predicates when the loop is not unrolled), so Pn
      // stays in VGPRs.  FP32 unrolls the column loop, so the rank-1 updates
      // touch only the columns right of the pivot; FP64 (more registers per
      // entry) keeps it rolled.
      constexpr int PANEL_UNROLL = sizeof(T) == 4 ? NB : 1;
#pragma unroll PANEL_UNROLL
      for (int c = 0; c < NB; ++c) {
        const int col = k0 + c;
        C *s_crow = s_crow2[c & 1];
        T *s_rv = s_rv2[c & 1];
        int *s_ri = s_ri2[c & 1];
        T best = (T)-1;
        int bi = 0x7fffffff;
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const int r = K0 + t + j * NT;
          C pc = czero<T>();
#pragma unroll
          for (int cc = 0; cc < NB; ++cc) if (cc == c) pc = Pn[j][cc];
          if (r >= col && r < n) {
            const T v = cabs1(pc);
            if (v > best) { best = v; bi = r; }
          }
        }
        // best is -1 or a magnitude (a NaN one never replaces it), so the
        // wave's arg-max is well defined
        wave_argmax(best, bi);
        // publish the wave's candidate (value, row index, the row itself) and
        // the diagonal row; one barrier then serves the pivot choice and the
        // interchange.  The wave's candidate rows are its own rows, so the
        // owner of row bi is a lane of this wave.
        if (lane == 0) { s_rv[wv] = best; s_ri[wv] = bi; }
        const int cj = (col - K0) / NT, bj = (bi < n) ? (bi - K0) / NT : -1;
#pragma unroll
        for (int j = 0; j < R; ++j) {
          if (j == bj && ((bi - K0) % NT) == t) {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) s_cand2[c & 1][wv][cc] = Pn[j][cc];
          }
          if (j == cj && ((col - K0) % NT) == t) {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) s_crow[cc] = Pn[j][cc];
          }
        }
        __syncthreads();
        best = s_rv[0]; bi = s_ri[0];
        int wb = 0;
#pragma unroll
        for (int w = 1; w < NW; ++w) {
          if (s_rv[w] > best || (s_rv[w] == best && s_ri[w] < bi)) { best = s_rv[w]; bi = s_ri[w]; wb = w; }
        }
        // no finite candidate (NaN column): keep the diagonal so every index stays in range
        const bool found = (bi >= col && bi < n);
        const int p = found ? bi : col;
        const C *s_prow = found ? s_cand2[c & 1][wb] : s_crow;
        const int pj = (p - K0) / NT;
        if (p != col) {
#pragma unroll
          for (int j = 0; j < R; ++j) {
            if (j == pj && ((p - K0) % NT) == t) {
#pragma unroll
              for (int cc = 0; cc < NB; ++cc) Pn[j][cc] = s_crow[cc];
            }
            if (j == cj && ((col - K0) % NT) == t) {
#pragma unroll
              for (int cc = 0; cc < NB; ++cc) Pn[j][cc] = s_prow[cc];
            }
          }
        }
        const C piv = s_prow[c];
        if (t == 0) s_piv[col - K0] = p;
        if (piv.x == (T)0 && piv.y == (T)0) {
          if (myinfo == 0) myinfo = col + 1;
        } else {
          const C inv = cdiv(mk<T>(1, 0), piv);
#pragma unroll
          for (int j = 0; j < R; ++j) {
            const int r = K0 + t + j * NT;
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
      }
      // ---- write the factored panel back; stage the diagonal block
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = K0 + t + j * NT;
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
      // ---- inv(L11) (unit lower), one column per thread, the column in
      // registers (the terms k < c vanish: x[k] = 0 there)
      if (t < NB) {
        const int c = t;
        C x[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          C s = (i == c) ? mk<T>(1, 0) : czero<T>();
#pragma unroll
          for (int k = 0; k < i; ++k) s = cfms(s, (k >= c) ? s_blk[i * NB + k] : czero<T>(), x[k]);
          x[i] = (i < c) ? czero<T>() : s;
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) s_linv[i * NB + c] = x[i];
      }
      __syncthreads();
      if (t < NB) ipiv[k0 + t] = s_piv[k0 - K0 + t];
      // ---- net permutation of the panel's interchanges: rows of the panel
      // top and the rows below it that change
      if (t == 0) {
        const int ns = net_perm(s_piv + (k0 - K0), k0, NB, s_srow, s_sorig, s_slot);
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
      // ---- the rest of the block column, in strips of SW columns
      const int nout = s_nout, nslot = NB + nout;
      const C *Li = s_linv;
      for (int c0 = k0 + NB; c0 < kend; c0 += SW) {
        const int sw = min(SW, kend - c0);      // strips never cross the block end
        for (int e = t; e < nslot * sw; e += NT) {
          const int slot = e % nslot, cc = e / nslot;
          const int src = (slot < NB) ? s_topsrc[slot] : s_outsrc[slot - NB];
          s_stage[slot * SW + cc] = A[(size_t)(c0 + cc) * lda + src];
        }
        __syncthreads();
        for (int e = t; e < nout * sw; e += NT) {
          const int o = e % nout, cc = e / nout;
          A[(size_t)(c0 + cc) * lda + s_outrow[o]] = s_stage[(NB + o) * SW + cc];
        }
        for (int e = t; e < NB * sw; e += NT) {
          const int i = e / sw, cc = e - i * sw;
          C u = czero<T>();
          for (int k = 0; k <= i; ++k) u = cfma(u, Li[i * NB + k], s_stage[k * SW + cc]);
          s_u12[i * SW + cc] = u;
          A[(size_t)(c0 + cc) * lda + k0 + i] = u;
        }
        __syncthreads();
        // A22 strip -= L21 U12.  FP32: on the matrix cores, as the trailing
        // kernel does it (C tile = the update transposed: lane l holds strip
        // columns 4 (l / 16) + i of row r0 + l % 16), one 16-row tile at a
        // time; the B operand L21(r0 + l % 16, 4 s + l / 16) belongs to
        // another lane of the same wave, so each wave stages the L21 rows of
        // two of its tiles (32 rows x NB) in its quarter of s_u, which the
        // strips do not otherwise use, and reads them back in operand
        // order.  Tiles are whole: rows below the panel start at k0 + NB, a
        // multiple of 16, and n is a multiple of 32.
        const int wrow = t - lane;
        if constexpr (sizeof(T) == 4 && SW == 16 && NB == 16 && NT == 64 * NW && 32 * NB * NW <= SU) {
          typedef typename Mfma<T>::v4 v4;
          C *lst = s_u + (size_t)wv * 32 * NB;
#pragma unroll
          for (int j = 0; j < R; ++j) {
            const int w0 = K0 + wrow + j * NT;
            if (w0 >= n || w0 + 63 < k0 + NB) continue;           // wave-uniform
#pragma unroll 1
            for (int half = 0; half < 2; ++half) {
              const int rb = w0 + 32 * half;                       // first row of the tile pair
              if (rb >= n || rb + 31 < k0 + NB) continue;          // wave-uniform
              if ((lane >> 5) == half) {
#pragma unroll
                for (int k = 0; k < NB; ++k) lst[(lane & 31) * NB + k] = Pn[j][k];
              }
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 1
              for (int a = 0; a < 2; ++a) {
                const int r0 = rb + 16 * a;
                if (r0 >= n || r0 < k0 + NB) continue;               // wave-uniform
                C *Ab = A + (size_t)(c0 + 4 * (lane >> 4)) * lda + r0 + (lane & 15);
                v4 dr, di;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                  const C x = Ab[(size_t)i * lda];
                  dr[i] = x.x;
                  di[i] = x.y;
                }
#pragma unroll
                for (int st = 0; st < NB / 4; ++st) {
                  const C l = lst[(16 * a + (lane & 15)) * NB + 4 * st + (lane >> 4)];
                  const C u = s_u12[(4 * st + (lane >> 4)) * SW + (lane & 15)];
                  const T nur = -u.x, ui = u.y, nui = -u.y;
                  dr = Mfma<T>::mac(nur, l.x, dr);
                  dr = Mfma<T>::mac(ui, l.y, dr);
                  di = Mfma<T>::mac(nui, l.x, di);
                  di = Mfma<T>::mac(nur, l.y, di);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) Ab[(size_t)i * lda] = mk<T>(dr[i], di[i]);
              }
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
          }
          __syncthreads();
          continue;
        }
        // VALU (FP64): one row slot at a time with G columns of independent
        // accumulators; a wave skips the slots in which none of its 64 rows
        // is below the panel and inside the matrix (the third slot of
        // matrices of order <= 2 NT, for instance)
        constexpr int G = (R >= 2) ? 2 : 4;
#pragma unroll 1
        for (int cg = 0; cg < sw; cg += G) {
#pragma unroll
          for (int j = 0; j < R; ++j) {
            const int w0 = K0 + wrow + j * NT;
            if (w0 >= n || w0 + 63 < k0 + NB) continue;
            const int r = K0 + t + j * NT;
            const bool live = (r >= k0 + NB && r < n);
            C acc[G];
#pragma unroll
            for (int q = 0; q < G; ++q) {
              const C v = A[(size_t)(c0 + cg + q) * lda + (live ? r : k0)];
              acc[q] = live ? v : czero<T>();
            }
#pragma unroll
            for (int k = 0; k < NB; ++k) {
#pragma unroll
              for (int q = 0; q < G; ++q) acc[q] = cfms(acc[q], Pn[j][k], s_u12[k * SW + cg + q]);
            }
            if (live) {
#pragma unroll
              for (int q = 0; q < G; ++q) A[(size_t)(c0 + cg + q) * lda + r] = acc[q];
            }
          }
        }
        __syncthreads();
      }
      // ---- the panel's interchanges on the block's columns left of it, so the
      // block's L is in one row order when the block is done
      const int ns = s_ns;
      if (ns > 0 && k0 > K0) {
        for (int c0 = K0; c0 < k0; c0 += SW) {
          const int w = min(SW, k0 - c0);
          for (int e = t; e < ns * w; e += NT) {
            const int slot = e % ns, cc = e / ns;
            s_stage[slot * SW + cc] = A[(size_t)(c0 + cc) * lda + s_sorig[slot]];
          }
          __syncthreads();
          for (int e = t; e < ns * w; e += NT) {
            const int slot = e % ns, cc = e / ns;
            A[(size_t)(c0 + cc) * lda + s_srow[slot]] = s_stage[slot * SW + cc];
          }
          __syncthreads();
        }
      }
    }  // panels of the block

    // ---- net permutation of the whole block: top rows K0 .. kend take the
    // input rows s_topsrc[], rows below take s_outsrc[] (always input rows of
    // the block top); s_srow/s_sorig list every touched row
    if (t == 0) {
      const int ns = net_perm(s_piv, K0, kb, s_srow, s_sorig, s_slot);
      for (int c = 0; c < kb; ++c) s_topsrc[c] = K0 + c;
      int no = 0;
      for (int e = 0; e < ns; ++e) {
        const int row = s_srow[e];
        if (row < kend) s_topsrc[row - K0] = s_sorig[e];
        else if (s_sorig[e] != row) { s_outrow[no] = row; s_outsrc[no] = s_sorig[e]; ++no; }
      }
      s_nout = no;
      s_ns = ns;
    }
    __syncthreads();
    const int bns = s_ns, bnout = s_nout;
    {
      // ---- the block's net interchanges: the record of the forward solves
      // (read in pivot mode 1) and the trailing record of the update kernel
      int *rec = ipiv + 2 * n + 1 + (size_t)(K0 / KB) * LU_PREC;
      for (int e = t; e < bns; e += NT) { rec[1 + e] = s_srow[e]; rec[1 + 2 * KB + e] = s_sorig[e]; }
      if (t == 0) rec[0] = bns;
      int *trec = ipiv + lu_trec_off(n, K0 / KB);
      for (int e = t; e < kb; e += NT) trec[1 + e] = s_topsrc[e];
      for (int e = t; e < bnout; e += NT) { trec[1 + KB + e] = s_outrow[e]; trec[1 + 2 * KB + e] = s_outsrc[e]; }
      if (t == 0) trec[0] = bnout;
    }
    // ---- left columns [0, K0) (pivot mode 0 only): move the touched rows,
    // 16 columns at a time, through LDS (bns * 16 <= 2 KB * 16 = SU entries)
    if (left && bns > 0 && K0 > 0) {
      for (int c0 = 0; c0 < K0; c0 += 16) {
        for (int e = t; e < bns * 16; e += NT) s_u[e] = A[(size_t)(c0 + e / bns) * lda + s_sorig[e % bns]];
        __syncthreads();
        for (int e = t; e < bns * 16; e += NT) A[(size_t)(c0 + e / bns) * lda + s_srow[e % bns]] = s_u[e];
        __syncthreads();
      }
    }
    // ---- X = inv(L11) (unit lower, zero outside kb x kb) in T arithmetic,
    // column-major into this block's slot of the diagonal-block inverses
    // (read by the triangular solves, and below by U12 = X A12).  Four lanes
    // per column: lane q of column c keeps x[m] = X(4 m + q, c) in registers
    // and sums the terms k = q (mod 4) of each substitution step; the sums meet
    // by shuffles.  The strictly lower triangle of L11 is staged in s_u, packed
    // by rows (row i at i (i - 1) / 2; 2016 entries), with zeros in rows and
    // columns past kb.
    C *X = blks + P.blk_off + (size_t)(K0 / KB) * KB * KB;
    static_assert(KB * (KB - 1) / 2 <= SU, "packed L11 fits in s_u");
    for (int e = t; e < KB * KB; e += NT) {
      const int i = e % KB, k = e / KB;
      if (k < i) {
        const C v = A[(size_t)(K0 + k) * lda + K0 + (i < kb ? i : 0)];
        s_u[i * (i - 1) / 2 + k] = (i < kb) ? v : czero<T>();
      }
    }
    __syncthreads();
    if (t < 4 * KB) {
      const int c = t >> 2, q = t & 3;
      C x[KB / 4];
#pragma unroll
      for (int m = 0; m < KB / 4; ++m) x[m] = czero<T>();
#pragma unroll
      for (int mi = 0; mi < KB / 4; ++mi) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 4 * mi + r;
          C s = czero<T>();
#pragma unroll
          for (int m = 0; m <= mi; ++m)     // entry (i, 4 m + q) of L11 when 4 m + q < i
            s = cfma(s, (m < mi || q < r) ? s_u[i * (i - 1) / 2 + 4 * m + q] : czero<T>(), x[m]);
          s.x += quad_xor1(s.x); s.y += quad_xor1(s.y);
          s.x += quad_xor2(s.x); s.y += quad_xor2(s.y);
          const C v = (i < c || i >= kb || c >= kb) ? czero<T>() : ((i == c) ? mk<T>(1, 0) : czero<T>() - s);
          if (r == q) x[mi] = v;
        }
      }
#pragma unroll
      for (int m = 0; m < KB / 4; ++m) X[c * KB + 4 * m + q] = x[m];
    }
    __syncthreads();
  // ---- final row permutation: row r of L U is row perm[r] of the input,
  // composed in LDS (s_slot, idle now) by one thread, then stored; only the
  // LAPACK form (pivot mode 0) has it: the mode-1 solves use the block records
  static_assert(sizeof(s_u) >= 1280 * sizeof(int), "the pivots fit in s_u");
  if (left && kend == n) {
    int *perm = ipiv + n;
    int *s_ip = reinterpret_cast<int *>(s_u);   // the pivots, read by one thread
    __syncthreads();
    for (int r = t; r < n; r += NT) { s_slot[r] = r; s_ip[r] = ipiv[r]; }
    __syncthreads();
    if (t == 0) {
      for (int c = 0; c < n; ++c) {
        const int p = s_ip[c];
        if (p != c) { const int a = s_slot[c]; s_slot[c] = s_slot[p]; s_slot[p] = a; }
      }
    }
    __syncthreads();
    for (int r = t; r < n; r += NT) perm[r] = s_slot[r];
  }
  if (t == 0) {
    info[pi] = myinfo;
    if (kend == n) ipiv[2 * n] = left ? 0 : 1;
  }
  }  // block K0
}


// Strip [c0, c0 + gw) of the columns right of block K0 of one matrix
// (tasks[2 b], tasks[2 b + 1] = point, c0; point -1 pads the grid): stage the
// block rows of the strip in final order (some come from below the block;
// the rows below that change then receive their input rows, which lie in
// the block top and are not written before U12 is stored), U12 = X S with
// X = inv(L11) of the block, and A22 -= L21 U12 over the rows below the
// block.  Complex products as 2 x 2 real blocks on the matrix cores,
// computed transposed (MFMA rows = matrix columns) so that each accumulator
// register of 16 lanes covers 16 consecutive matrix rows.
// Streaming load of one A22 entry: the trailing update reads each A22 tile
// once per block step and writes it back with store_stream (the next read is
// a block step later, after the batch's other tiles have passed through the
// caches), so L2 is left to the L21 panel that every strip of a point reads.
template <typename T> __device__ inline cx<T> load_stream(const cx<T> *p);
template <> __device__ inline zf load_stream<float>(const zf *p) {
  typedef float v2 __attribute__((ext_vector_type(2)));
  const v2 w = __builtin_nontemporal_load((const v2 *)p);
  return mk(w.x, w.y);
}
template <> __device__ inline zd load_stream<double>(const zd *p) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  const v2 w = __builtin_nontemporal_load((const v2 *)p);
  return mk(w.x, w.y);
}
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_lu_trail(const PointDev *__restrict__ pts, cx<T> *__restrict__ mats,
                                                  const int *__restrict__ pivs, const cx<T> *__restrict__ blks,
                                                  const int *__restrict__ tasks, int K0) {
  typedef cx<T> C;
  typedef typename Mfma<T>::v4 v4;
  constexpr int NT = 256, NW = NT / 64, KB = 64, GW = lu_gw<T>(), UP = GW + 1, CT = GW / 16;
  const int pi = tasks[2 * blockIdx.x], c0 = tasks[2 * blockIdx.x + 1];
  if (pi < 0) return;
  const PointDev &P = pts[pi];
  const int n = P.npad, lda = P.npad;
  if (K0 >= n) return;
  const int kb = min(KB, n - K0), kend = K0 + kb;
  if (c0 < kend || c0 >= n) return;
  C *A = mats + P.mat_off;
  const int *trec = pivs + P.piv_off + lu_trec_off(n, K0 / KB);
  const C *X = blks + P.blk_off + (size_t)(K0 / KB) * KB * KB;
  // the wave index as a scalar: the row offsets of the update below are then
  // wave-uniform and the addresses a per-lane base plus a scalar offset
  const int t = threadIdx.x, lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
  __shared__ int s_topsrc[KB], s_outrow[KB], s_outsrc[KB];
  __shared__ C s_u[KB * UP];
  const int nout = min(trec[0], KB);
  for (int e = t; e < kb; e += NT) s_topsrc[e] = trec[1 + e];
  for (int e = t; e < nout; e += NT) { s_outrow[e] = trec[1 + KB + e]; s_outsrc[e] = trec[1 + 2 * KB + e]; }
  __syncthreads();
  const int gw = min(GW, n - c0);
  // stage the block rows of the strip in final order (some come from
  // below the block); then the rows below that change receive their
  // input rows, which lie in the block top and are not written before
  // U12 is stored
  {
    // every load of the thread issued before the first store to LDS
    constexpr int NE = (KB * GW + NT - 1) / NT;
    C v[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const int e = t + q * NT, k = e % KB, j = e / KB;
      const bool ok = k < kb && j < gw;       // else a harmless in-range load, not stored
      v[q] = A[(size_t)(ok ? c0 + j : c0) * lda + (ok ? s_topsrc[k] : K0)];
    }
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const int e = t + q * NT, k = e % KB, j = e / KB;
      if (k < kb && j < gw) s_u[k * UP + j] = v[q];
    }
  }
  __syncthreads();
  for (int e = t; e < nout * gw; e += NT) {
    const int o = e % nout, j = e / nout;
    A[(size_t)(c0 + j) * lda + s_outrow[o]] = A[(size_t)(c0 + j) * lda + s_outsrc[o]];
  }
  // U12 = X S (S = the staged rows) on the matrix cores, transposed as
  // the update below: wave w takes column tile w % CTN and the row tiles
  // it owns (four per column tile, split over NW / CTN waves so that the
  // triangular work, 4 (rt + 1) depth-4 steps for row tile rt, balances)
  {
    constexpr int CTN = GW / 16, NWC = NW / CTN;
    static_assert(NWC == 1 || NWC == 2 || NWC == 4, "waves per U12 column tile");
    const C *Xs = X;
    const int ct = wv % CTN, h = wv / CTN;
    const bool colok = ct * 16 < gw;
    v4 ur[4], ui[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
      for (int i = 0; i < 4; ++i) { ur[rt][i] = (T)0; ui[rt][i] = (T)0; }
      const int own = (NWC == 1) ? 0 : (NWC == 2 ? ((rt == 0 || rt == 3) ? 0 : 1) : rt);
      if (own != h || !colok) continue;
      // X(i, k) = 0 for k > i: row tile rt needs k < 16 (rt + 1)
#pragma unroll
      for (int ks = 0; ks < 4 * (rt + 1); ++ks) {
        const int k = 4 * ks + (lane >> 4);
        const C sv = s_u[k * UP + ct * 16 + (lane & 15)];
        const C xv = Xs[(size_t)k * KB + rt * 16 + (lane & 15)];
        ur[rt] = Mfma<T>::mac(sv.x, xv.x, ur[rt]);
        ur[rt] = Mfma<T>::mac(-sv.y, xv.y, ur[rt]);
        ui[rt] = Mfma<T>::mac(sv.y, xv.x, ui[rt]);
        ui[rt] = Mfma<T>::mac(sv.x, xv.y, ui[rt]);
      }
    }
    __syncthreads();     // S fully read
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      const int own = (NWC == 1) ? 0 : (NWC == 2 ? ((rt == 0 || rt == 3) ? 0 : 1) : rt);
      if (own != h || !colok) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int j = ct * 16 + Mfma<T>::row_of(lane, i), row = rt * 16 + (lane & 15);
        if (j < gw) {
          const C u = mk<T>(ur[rt][i], ui[rt][i]);
          s_u[row * UP + j] = u;
          A[(size_t)(c0 + j) * lda + K0 + row] = u;
        }
      }
    }
    __syncthreads();
  }
  // A22 -= L21 U12: each wave takes chunks of 16 rows against the gw columns
  // of the strip (rows of all chunks are multiples of 16 below n, a multiple
  // of 32); the A22 tile of the wave's next chunk is loaded while the current
  // one is on the matrix cores
  // Addresses as one per-lane base plus wave-uniform offsets: tile entry
  // (ct, i) of rows r0 .. r0 + 15 is Ab[(ct 16 + row_step(i)) lda + r0], and
  // L21 entry of depth step s is Lb[(4 s) lda + r0].  Loads take clamped
  // in-range offsets and are selected afterwards, so none waits under a branch.
  const C *Lb = A + (size_t)(K0 + (lane >> 4)) * lda + (lane & 15);
  C *Ab = A + (size_t)(c0 + Mfma<T>::row_of(lane, 0)) * lda + (lane & 15);
  v4 dr[CT], di[CT], nr[CT], ni[CT];
  int r0 = kend + 16 * wv;
#pragma unroll
  for (int ct = 0; ct < CT; ++ct)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool ok = r0 < n && ct * 16 < gw;
      C x = load_stream(Ab + (ok ? (ct * 16 + Mfma<T>::row_step(i)) * lda + r0 : kend));
      if (!ok) x = czero<T>();
      dr[ct][i] = x.x;
      di[ct][i] = x.y;
    }
  for (; r0 < n; r0 += 16 * NW) {
    const int rn = r0 + 16 * NW;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool ok = rn < n && ct * 16 < gw;
        C x = load_stream(Ab + (ok ? (ct * 16 + Mfma<T>::row_step(i)) * lda + rn : r0));
        if (!ok) x = czero<T>();
        nr[ct][i] = x.x;
        ni[ct][i] = x.y;
      }
    // L21 of the chunk is loaded LH depth-4 steps at a time before the
    // MFMA chain that uses it (one memory latency per LH steps).  kb is 32 or
    // 64, so every step of a 4 LH = 32 deep pass is inside the block; the
    // column tiles past gw (the last strip of an order 32 mod 64) are
    // computed from in-range data and not stored: a guard here would split
    // the MFMA chain into branches that copy its accumulators
    constexpr int LH = 8;
    static_assert(4 * LH == 32, "kb is a multiple of 32");
#pragma unroll 1
    for (int kb0 = 0; kb0 < kb; kb0 += 4 * LH) {
      T lr[LH], li[LH];
#pragma unroll
      for (int st = 0; st < LH; ++st) {
        const C x = Lb[(kb0 + 4 * st) * lda + r0];
        lr[st] = x.x;
        li[st] = x.y;
      }
#pragma unroll
      for (int st = 0; st < LH; ++st) {
        const int k = kb0 + 4 * st + (lane >> 4);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const C uu = s_u[k * UP + ct * 16 + (lane & 15)];
          const T nur = -uu.x, ui = uu.y, nui = -uu.y;
          dr[ct] = Mfma<T>::mac(nur, lr[st], dr[ct]);
          dr[ct] = Mfma<T>::mac(ui, li[st], dr[ct]);
          di[ct] = Mfma<T>::mac(nui, lr[st], di[ct]);
          di[ct] = Mfma<T>::mac(nur, li[st], di[ct]);
        }
      }
    }
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (ct * 16 < gw) store_stream(Ab + (ct * 16 + Mfma<T>::row_step(i)) * lda + r0, mk<T>(dr[ct][i], di[ct][i]));
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) { dr[ct] = nr[ct]; di[ct] = ni[ct]; }
  }
}

// Inverses of the diagonal blocks of U (upper) of order BS = 64 (the last
// block may be shorter; entries outside it are stored as zeros), in FP64
// arithmetic, stored as T, column-major.  With nblk = ceil(npad / BS), block b
// of inv(L) is at blk_off + b BS^2 (written by k_lu_panel) and block b
// of inv(U) at blk_off + (nblk + b) BS^2; the per-point allocation is
// 2 nblk BS^2 elements (blk_elems above).
// grid: x = block index, y = point; 256 threads, four lanes per column: lane
// q of column c sums the terms k = q (mod 4) of each substitution step and
// keeps the entries x[k][c], k = q (mod 4), in registers; the step's sum is
// reduced over the four lanes by shuffles, so the 64-step recurrence needs
// no barrier.  The reciprocals of the 64 diagonal entries are formed once,
// in LDS, so each step multiplies instead of dividing.  Besides them only the
// block itself lives in LDS, as stored (the arithmetic converts to FP64).
template <typename T, int BS>
__global__ __launch_bounds__(4 * BS) void k_blkinv(const PointDev *__restrict__ pts, const cx<T> *__restrict__ mats,
                                                   cx<T> *__restrict__ blks) {
  static_assert(BS == 64, "four lanes per column of a 64-wide block");
  const PointDev &P = pts[blockIdx.y];
  const int n = P.npad, k0 = blockIdx.x * BS;
  if (k0 >= n) return;
  const int bs = min(BS, n - k0), t = threadIdx.x;
  const int nblk = (n + BS - 1) / BS;
  const cx<T> *A = mats + P.mat_off;
  cx<T> *Ui = blks + P.blk_off + (size_t)(nblk + blockIdx.x) * BS * BS;
  constexpr int LD = BS + 1;       // padded rows: the column-wise staging writes hit distinct banks
  __shared__ cx<T> s_a[BS * LD];  // row-major copy of the block, as stored (FP32: 33 KB, so more groups per CU)
  __shared__ zd s_rd[BS];          // 1 / U(i, i)
  for (int e = t; e < BS * BS; e += 4 * BS) {
    const int i = e % BS, j = e / BS;   // consecutive lanes read consecutive rows of column j
    const bool ok = i < bs && j < bs;
    const cx<T> v = A[(size_t)(k0 + (ok ? j : 0)) * n + k0 + (ok ? i : 0)];
    s_a[i * LD + j] = ok ? v : czero<T>();
  }
  if (t < BS) {
    const int i = min(t, bs - 1);
    s_rd[t] = cdiv(mk(1.0, 0.0), cvt<double>(A[(size_t)(k0 + i) * n + k0 + i]));
  }
  __syncthreads();
  const int c = t >> 2, q = t & 3;
  zd x[BS / 4];                   // x[m] = entry (4 m + q, c) of the inverse being built
  // inv(U) column c: x_i = (delta_ic - sum_{i < k <= c} U_ik x_k) / U_ii, i = c .. 0
#pragma unroll
  for (int m = 0; m < BS / 4; ++m) x[m] = czero<double>();
#pragma unroll
  for (int mi = BS / 4 - 1; mi >= 0; --mi) {
#pragma unroll
    for (int r = 3; r >= 0; --r) {
      const int i = 4 * mi + r;
      zd s = czero<double>();
#pragma unroll
      for (int m = mi; m < BS / 4; ++m) {
        const int k = 4 * m + q;
        if (k > i && k <= c) s = cfma(s, cvt<double>(s_a[i * LD + k]), x[m]);
      }
      s.x += quad_xor1(s.x); s.y += quad_xor1(s.y);
      s.x += quad_xor2(s.x); s.y += quad_xor2(s.y);
      const zd rhs = ((i == c) ? mk(1.0, 0.0) : czero<double>()) - s;
      const zd v = (i > c || c >= bs) ? czero<double>() : rhs * s_rd[i];
      if (r == q) x[mi] = v;
    }
  }
#pragma unroll
  for (int m = 0; m < BS / 4; ++m) Ui[c * BS + 4 * m + q] = cvt<T>(x[m]);
}
