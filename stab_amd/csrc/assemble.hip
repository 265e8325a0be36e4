// Dense assembly of the shifted pencil into the LU workspace, plus the norms
// of the pencil parts and the row scales used for FP32 equilibration.
#include "pencil.h"

// Squared Frobenius norm of one row (i,p) of matrix `form` in closed form:
// sum_j sum_q |bh D1_ij - vyy D2_ij + delta_ij dh|^2 expands into the row sums
// S11, S22, S12 of the operators plus a correction at j = i.
__device__ inline double row_norm2(const PointDev &P, const zd *node, const double *ops, int form, int i, int p,
                                   int kind) {
  const int ny = P.ny;
  const double *D1 = ops + P.op_off, *D2 = D1 + (size_t)ny * ny, *dt2w = D2 + (size_t)ny * ny;
  const double *sums = dt2w + ny;
  const double s11 = sums[i], s22 = sums[ny + i], s12 = sums[2 * ny + i], sdt = sums[3 * ny];
  if (kind == ROW_DIRICHLET) return abs2(dirichlet_diag(P, form));
  const double d1 = D1[(size_t)i * ny + i], d2 = D2[(size_t)i * ny + i];
  double s = 0.0;
  for (int q = 0; q < 5; ++q) {
    zd bh, dh; double vyy;
    form_coef(P, node, form, p, q, P.shift, bh, vyy, dh);
    zd off;
    double body;
    if (kind == ROW_GENERIC || (!walle_vyy_d1(P) && q < 4)) {
      body = abs2(bh) * s11 + vyy * vyy * s22 - 2.0 * bh.x * vyy * s12;
      off = bh * d1 - mk(vyy * d2, 0.0);
    } else if (q < 4) {          // Chebyshev temporal wall energy row: (bh - vyy) D1
      const zd b2 = bh - mk(vyy, 0.0);
      body = abs2(b2) * s11;
      off = b2 * d1;
    } else {                     // column 4 of the wall energy row: -vyy Dt2(wall, :)
      body = vyy * vyy * sdt;
      off = mk(-vyy * dt2w[i], 0.0);
    }
    s += body - abs2(off) + abs2(off + dh);
  }
  return s;
}

// Per-point squared Frobenius norms of the pencil parts (the denominators of
// the relative residuals):  temporal [0] = A, [1] = B (solver form);
// spatial [0] = C0, [1] = C1, [2] = C2.  One block per point, thread per node.
__global__ __launch_bounds__(256) void k_norms(const PointDev *__restrict__ pts, const zd *__restrict__ coef,
                                               const double *__restrict__ ops, double *__restrict__ norms) {
  const PointDev P = pts[blockIdx.x];
  const int i = threadIdx.x;
  __shared__ double s_red[3][256];
  double a[3] = {0.0, 0.0, 0.0};
  if (i < P.ny) {
    const zd *node = coef + P.coef_off + (size_t)i * NODE_STRIDE;
    const int f0 = (P.itype == 1) ? FORM_AM : FORM_C0;
    const int f1 = (P.itype == 1) ? FORM_BM : FORM_C1;
    for (int p = 0; p < 5; ++p) {
      const int kind = row_kind(P, 5 * i + p);
      a[0] += row_norm2(P, node, ops, f0, i, p, kind);
      a[1] += row_norm2(P, node, ops, f1, i, p, kind);
      if (P.itype == 2) a[2] += row_norm2(P, node, ops, FORM_C2, i, p, kind);
    }
  }
  for (int k = 0; k < 3; ++k) s_red[k][i] = a[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (i < w)
      for (int k = 0; k < 3; ++k) s_red[k][i] += s_red[k][i + w];
    __syncthreads();
  }
  if (i < 3) norms[blockIdx.x * 4 + i] = s_red[i][0];
}

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
    M[(size_t)c * ld + r] = cvt<T>(v);
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
      M[(size_t)c * ld + r] = cvt<T>(v);
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
      M[(size_t)c * ld + r] = cvt<T>(v);
    }
  }
  for (int c = n; c < P.npad; ++c) M[(size_t)c * ld + r] = czero<T>();
}

template __global__ void k_assemble_rows<double>(const PointDev *, const zd *, const double *, const double *, zd *, int);
template __global__ void k_assemble_rows<float>(const PointDev *, const zd *, const double *, const double *, zf *, int);
