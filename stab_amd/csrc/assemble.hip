// Dense assembly of the shifted pencil into the LU workspace, plus the norms
// of the pencil parts and the row scales used for FP32 equilibration.
#include "pencil.h"

// Row scales of the shifted matrix for the FP32 factorisation:
// dr = 2^-round(log2 max_c |M(r,c)|), a power of two, so the scaled FP64 entries
// are exact multiples of the unscaled ones.  Row scaling alone brings the
// condition number of these matrices from 1e8-1e10 down to 1e4-1e6 (column
// scaling after it changes it by less than 10%), and partial pivoting then
// acts as scaled partial pivoting.  One block per point, one thread per row;
// the five (BH, VYY, DH) column coefficients of a row are formed once and the
// loop runs over the nodes j of D1(i, j), D2(i, j).
__global__ __launch_bounds__(256) void k_rowscale(const PointDev *__restrict__ pts, const zd *__restrict__ coef,
                                                  const double *__restrict__ ops, double *__restrict__ scales) {
  const PointDev &P = pts[blockIdx.x];
  const int ny = P.ny;
  const double *D1 = ops + P.op_off, *D2 = D1 + (size_t)ny * ny, *dt2w = D2 + (size_t)ny * ny;
  for (int r = threadIdx.x; r < P.npad; r += 256) {
    const int kind = row_kind(P, r);
    double mx2 = 1.0;                       // padding and Dirichlet rows: |diag| = 1
    if (kind == ROW_GENERIC || kind == ROW_WALLE) {
      const int i = r / 5, p = r - 5 * i;
      const zd *node = coef + P.coef_off + (size_t)i * NODE_STRIDE;
      zd bh[5], dh[5];
      double vyy[5];
      for (int q = 0; q < 5; ++q) form_coef(P, node, FORM_SHIFTED, p, q, P.shift, bh[q], vyy[q], dh[q]);
      mx2 = 0.0;
      for (int j = 0; j < ny; ++j) {
        const double d1 = D1[(size_t)j * ny + i], d2 = D2[(size_t)j * ny + i];
        for (int q = 0; q < 5; ++q) {
          zd v;
          if (kind == ROW_GENERIC || (!walle_vyy_d1(P) && q < 4)) v = bh[q] * d1 - mk(vyy[q] * d2, 0.0);
          else if (q < 4) v = bh[q] * d1 - mk(vyy[q] * d1, 0.0);
          else v = mk(-vyy[q] * dt2w[j], 0.0);
          if (j == i) v = v + dh[q];
          mx2 = fmax(mx2, abs2(v));
        }
      }
    }
    const double mx = sqrt(mx2);
    scales[P.scale_off + r] = (mx > 0.0) ? exp2(-rint(log2(mx))) : 1.0;
  }
}

// Streaming (non-temporal) store of one matrix entry, for matrices written
// once and read back only after they have left the caches: the assembled
// pencil (read by the LU, after the whole batch is written; L2 is left to the
// operators every row reads) and the trailing update's A22 tiles.
template <typename T> __device__ inline void store_stream(cx<T> *p, cx<T> v);
template <> __device__ inline void store_stream<float>(zf *p, zf v) {
  typedef float v2 __attribute__((ext_vector_type(2)));
  v2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, (v2 *)p);
}
template <> __device__ inline void store_stream<double>(zd *p, zd v) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  v2 w = {v.x, v.y};
  __builtin_nontemporal_store(w, (v2 *)p);
}

// Write matrix `form` (scaled by dr, dc when `scaled`) column-major with
// ld = npad.  grid: x = chunks of CW columns, y = point.
template <typename T, int CW>
__global__ __launch_bounds__(256) void k_assemble(const PointDev *__restrict__ pts, const zd *__restrict__ coef,
                                                  const double *__restrict__ ops, const double *__restrict__ scales,
                                                  cx<T> *__restrict__ mats, int form, int scaled) {
  const PointDev P = pts[blockIdx.y];
  const int c0 = blockIdx.x * CW;
  if (c0 >= P.npad) return;
  const int ld = P.npad;
  cx<T> *M = mats + P.mat_off;
  const double *dr = scales + P.scale_off, *dc = dr + P.npad;
  for (int e = threadIdx.x; e < CW * ld; e += 256) {
    const int cc = e / ld, r = e - cc * ld, c = c0 + cc;
    if (c >= P.npad) break;
    zd v = pencil_entry(P, coef, ops, form, P.shift, r, c);
    if (scaled) v = v * (dr[r] * dc[c]);
    store_stream(M + (size_t)c * ld + r, cvt<T>(v));
  }
}

template __global__ void k_assemble<double, 8>(const PointDev *, const zd *, const double *, const double *, zd *, int, int);
template __global__ void k_assemble<float, 8>(const PointDev *, const zd *, const double *, const double *, zf *, int, int);

// The shifted matrix (scaled by dr, dc when `scaled`) for the solve paths,
// one thread per row: the row's five (BH, VYY, DH) column coefficients are
// formed once, then the columns are written node by node (consecutive lanes
// hold consecutive rows, so every store and every D1/D2 load coalesces).
// Same entries, in the same operation order, as pencil_entry(FORM_SHIFTED).
// grid: x = chunks of 256 rows, y = point.
template <typename T>
__global__ __launch_bounds__(256) void k_assemble_rows(const PointDev *__restrict__ pts, const zd *__restrict__ coef,
                                                       const double *__restrict__ ops, const double *__restrict__ scales,
                                                       cx<T> *__restrict__ mats, int scaled) {
  const PointDev &P = pts[blockIdx.y];
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= P.npad) return;
  const int ld = P.npad, ny = P.ny, n = P.n;
  cx<T> *M = mats + P.mat_off;
  const double *dr = scales + P.scale_off, *dc = dr + P.npad;
  const double rs = scaled ? dr[r] : 1.0;
  const int kind = row_kind(P, r);
  if (kind == ROW_PAD || kind == ROW_DIRICHLET) {
    const zd d = (kind == ROW_PAD) ? mk(1.0, 0.0) : dirichlet_diag(P, FORM_SHIFTED);
    for (int c = 0; c < P.npad; ++c) {
      zd v = (c == r) ? d : czero<double>();
      if (scaled) v = v * (rs * dc[c]);
      store_stream(M + (size_t)c * ld + r, cvt<T>(v));
    }
    return;
  }
  const int i = r / 5, p = r - 5 * i;
  const zd *node = coef + P.coef_off + (size_t)i * NODE_STRIDE;
  zd bh[5], dh[5];
  double vyy[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) form_coef(P, node, FORM_SHIFTED, p, q, P.shift, bh[q], vyy[q], dh[q]);
  const double *D1 = ops + P.op_off, *D2 = D1 + (size_t)ny * ny, *dt2w = D2 + (size_t)ny * ny;
  for (int j = 0; j < ny; ++j) {
    const double d1 = D1[(size_t)j * ny + i], d2 = D2[(size_t)j * ny + i];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      zd v;
      if (kind == ROW_GENERIC) {
        v = bh[q] * d1 - mk(vyy[q] * d2, 0.0);
      } else if (q < 4) {   // wall energy row (wallt = 2, or finite differences)
        v = walle_vyy_d1(P) ? bh[q] * d1 - mk(vyy[q] * d1, 0.0) : bh[q] * d1 - mk(vyy[q] * d2, 0.0);
      } else {
        v = bh[q] * 0.0 - mk(vyy[q] * dt2w[j], 0.0);
      }
      if (i == j) v = v + dh[q];
      const int c = 5 * j + q;
      if (scaled) v = v * (rs * dc[c]);
      store_stream(M + (size_t)c * ld + r, cvt<T>(v));
    }
  }
  for (int c = n; c < P.npad; ++c) store_stream(M + (size_t)c * ld + r, czero<T>());
}

template __global__ void k_assemble_rows<double>(const PointDev *, const zd *, const double *, const double *, zd *, int);
template __global__ void k_assemble_rows<float>(const PointDev *, const zd *, const double *, const double *, zf *, int);
