// Pencil entries and matrix-free products built from the per-node blocks.
//
// Every matrix of the path has the row-block form of temporal.f90:683-703 /
// spatial.f90:721-741:
//     M[(i,p),(j,q)] = BH_i[p,q] D1(i,j) - VYY_i[p,q] D2(i,j) + delta_ij DH_i[p,q]
// with the unknown (i,p) at index 5 i + p (temporal.f90:634-641), except
//   * Dirichlet rows (top node rows of the momentum/energy equations and wall
//     rows 1-3, plus the wall energy row when wallt = 0), and
//   * the wall energy row when wallt = 2, which uses D1 in columns 0-3 and the
//     operators Dt1 (zero wall row) and Dt2 = D1 Dt1 in column 4
//     (temporal.f90:724-752, spatial.f90:760-788).
// Finite-difference points (fd_temporal.f90, fd_spatial.f90) have the wall
// at node 0 and explicit rows in D1/D2; their wall momentum rows are
// Dirichlet, and their wall energy row (always for temporal, wallt = 2 for
// spatial) is the generic row in columns 0-3 and -VYY D2(wall, :) in column 4
// (fd_temporal.f90:658-700, fd_spatial.f90:699-748), the form of the wallt = 2
// row with D2's wall row in the place of Dt2's.  The top rows are Dirichlet
// except the spatial continuity row when top = 1 (fd_spatial.f90:881-903).
#pragma once
#include "common.h"

enum { ROW_GENERIC = 0, ROW_DIRICHLET = 1, ROW_WALLE = 2, ROW_PAD = 3 };

__device__ inline int row_kind(const PointDev &P, int r) {
  if (r >= P.n) return ROW_PAD;
  const int i = r / 5, p = r - 5 * (r / 5);
  if (P.fd) {
    if (i == P.ny - 1)                                            // fd_temporal.f90:626-629
      return (p == 0 && P.itype == 2 && P.top == 1) ? ROW_GENERIC : ROW_DIRICHLET;   // fd_spatial.f90:913-923
    if (i > 0 || p == 0) return ROW_GENERIC;
    if (p < 4) return ROW_DIRICHLET;                              // :606-615
    return (P.itype == 1 || P.wallt == 2) ? ROW_WALLE : ROW_DIRICHLET;   // :658-700, fd_spatial.f90:889-898
  }
  if (i == 0) {
    if (p != 0) return ROW_DIRICHLET;
    if (P.itype == 2 && P.top != 1) return ROW_DIRICHLET;    // spatial.f90:794-798
    return ROW_GENERIC;
  }
  if (i == P.ny - 1) {
    if (p == 0) return ROW_GENERIC;
    if (p < 4) return ROW_DIRICHLET;
    return (P.wallt == 2) ? ROW_WALLE : ROW_DIRICHLET;
  }
  return ROW_GENERIC;
}

// Whether the wall energy row of a point takes its VYY term in columns 0-3
// through D1 (the Chebyshev temporal form, temporal.f90:733-735) rather than
// through D2 (spatial.f90:767-769, fd_temporal.f90:666-690).
__device__ inline bool walle_vyy_d1(const PointDev &P) { return P.itype == 1 && !P.fd; }

// Effective (BH, VYY, DH) coefficients of row r = (i,p), column q, for a form.
struct RowCoef { zd bh, dh; double vyy; bool wall_d1_vyy; };

__device__ inline void form_coef(const PointDev &P, const zd *__restrict__ node, int form, int p, int q, zd sig,
                                 zd &bh, double &vyy, zd &dh) {
  const int e = p * 5 + q;
  const zd I = mk(0.0, 1.0);
  switch (form) {
    case FORM_A0:
    case FORM_AM:
    case FORM_C0:
      bh = node[SLOT_BH * 25 + e]; vyy = node[SLOT_VYY * 25 + e].x; dh = node[SLOT_DH * 25 + e];
      break;
    case FORM_B0:
    case FORM_BM:
      bh = czero<double>(); vyy = 0.0; dh = I * node[SLOT_G * 25 + e].x;
      break;
    case FORM_C1:
      bh = node[SLOT_BH1 * 25 + e]; vyy = 0.0; dh = node[SLOT_DH1 * 25 + e];
      break;
    case FORM_C2:
      bh = czero<double>(); vyy = 0.0; dh = node[SLOT_VXX * 25 + e];
      break;
    default:  // FORM_SHIFTED
      if (P.itype == 1) {
        bh = node[SLOT_BH * 25 + e]; vyy = node[SLOT_VYY * 25 + e].x;
        dh = node[SLOT_DH * 25 + e] - sig * (I * node[SLOT_G * 25 + e].x);
      } else {
        bh = node[SLOT_BH * 25 + e] + sig * node[SLOT_BH1 * 25 + e];
        vyy = node[SLOT_VYY * 25 + e].x;
        dh = node[SLOT_DH * 25 + e] + sig * node[SLOT_DH1 * 25 + e] + (sig * sig) * node[SLOT_VXX * 25 + e].x;
      }
  }
}

// Value of Dirichlet row r on its diagonal for a form.
__device__ inline zd dirichlet_diag(const PointDev &P, int form) {
  switch (form) {
    case FORM_B0: return mk(0.0, 1.0);           // temporal.f90:630-654: B0 = i
    case FORM_C0: return mk(-1.0, 0.0);          // spatial.f90:797-826: C0 = -1
    case FORM_SHIFTED: return (P.itype == 2) ? mk(-1.0, 0.0) : mk(1.0, 0.0);
    case FORM_AM: return mk(1.0, 0.0);
    default: return czero<double>();             // A0, C1, C2 rows are zero
  }
}

// Entry (r, c) of matrix `form` (padding rows/columns: identity for the
// shifted form, zero otherwise).
__device__ inline zd pencil_entry(const PointDev &P, const zd *__restrict__ coef, const double *__restrict__ ops,
                                  int form, zd sig, int r, int c) {
  const int kind = row_kind(P, r);
  if (kind == ROW_PAD || c >= P.n)
    return (r == c && form == FORM_SHIFTED) ? mk(1.0, 0.0) : czero<double>();
  if (kind == ROW_DIRICHLET) return (r == c) ? dirichlet_diag(P, form) : czero<double>();
  const int ny = P.ny;
  const int i = r / 5, p = r - 5 * i, j = c / 5, q = c - 5 * j;
  const zd *node = coef + P.coef_off + (size_t)i * NODE_STRIDE;
  zd bh, dh; double vyy;
  form_coef(P, node, form, p, q, sig, bh, vyy, dh);
  const double *D1 = ops + P.op_off, *D2 = D1 + (size_t)ny * ny, *dt2w = D2 + (size_t)ny * ny;
  const double d1 = D1[(size_t)j * ny + i];
  zd v;
  if (kind == ROW_GENERIC) {
    v = bh * d1 - mk(vyy * D2[(size_t)j * ny + i], 0.0);
  } else {   // ROW_WALLE (wallt = 2, or a finite-difference point)
    if (q < 4) {
      if (walle_vyy_d1(P)) v = bh * d1 - mk(vyy * d1, 0.0);                   // temporal.f90:733-735
      else v = bh * d1 - mk(vyy * D2[(size_t)j * ny + i], 0.0);                // spatial.f90:767-769
    } else {
      v = bh * 0.0 - mk(vyy * dt2w[j], 0.0);                                  // Dt1(ny,:) = 0
    }
  }
  if (i == j) v = v + dh;
  return v;
}

// ---------------------------------------------------------------------------
// Matrix-free product y = M x for one point inside a work-group.
// x, y: LDS vectors of P.n complex (node-major); X1, X2: LDS scratch of 5 ny.
// All threads of the block must call it.  Padding entries are not touched.
// ---------------------------------------------------------------------------
template <int NT>
__device__ void pencil_apply(const PointDev &P, const zd *__restrict__ coef, const double *__restrict__ ops, int form,
                             zd sig, const zd *x, zd *y, zd *X1, zd *X2, zd *red) {
  const int ny = P.ny, n = P.n, t = threadIdx.x;
  const double *D1 = ops + P.op_off, *D2 = D1 + (size_t)ny * ny, *dt2w = D2 + (size_t)ny * ny;
  const bool need_d2 = (form == FORM_A0 || form == FORM_AM || form == FORM_C0 || form == FORM_SHIFTED);
  const bool need_d1 = (form != FORM_B0 && form != FORM_BM && form != FORM_C2);
  if (need_d1 && need_d2 && 2 * ny <= NT) {
    // one half of the work-group forms X1 = D1 x, the other X2 = D2 x, one
    // node row per thread
    const int g = (t >= NT / 2) ? 1 : 0, i = t - g * (NT / 2);
    if (i < ny) {
      const double *D = g ? D2 : D1;
      zd a[5];
      for (int k = 0; k < 5; ++k) a[k] = czero<double>();
#pragma unroll 8
      for (int j = 0; j < ny; ++j) {
        const double d = D[(size_t)j * ny + i];
        for (int k = 0; k < 5; ++k) a[k] = a[k] + x[j * 5 + k] * d;
      }
      zd *X = g ? X2 : X1;
      for (int k = 0; k < 5; ++k) X[i * 5 + k] = a[k];
    }
  } else if (need_d1) {
    for (int i = t; i < ny; i += NT) {
      zd a[5], b[5];
      for (int k = 0; k < 5; ++k) { a[k] = czero<double>(); b[k] = czero<double>(); }
#pragma unroll 4
      for (int j = 0; j < ny; ++j) {
        const double d1 = D1[(size_t)j * ny + i];
        const double d2 = need_d2 ? D2[(size_t)j * ny + i] : 0.0;
        for (int k = 0; k < 5; ++k) {
          const zd xv = x[j * 5 + k];
          a[k] = a[k] + xv * d1;
          b[k] = b[k] + xv * d2;
        }
      }
      for (int k = 0; k < 5; ++k) { X1[i * 5 + k] = a[k]; X2[i * 5 + k] = b[k]; }
    }
  }
  if (need_d1) {
    if ((P.wallt == 2 || P.fd) && t < 64) {
      // Dt2(wall,:) . x_T  (temporal.f90:137-140; D2(wall,:) . x_T for a
      // finite-difference point): one wave reduces
      zd s = czero<double>();
      for (int j = t; j < ny; j += 64) s = s + x[j * 5 + 4] * dt2w[j];
      for (int off = 32; off > 0; off >>= 1) {
        s.x += __shfl_xor(s.x, off);
        s.y += __shfl_xor(s.y, off);
      }
      if (t == 0) red[0] = s;
    }
    __syncthreads();
  }
  for (int r = t; r < n; r += NT) {
    const int kind = row_kind(P, r);
    const int i = r / 5, p = r - 5 * i;
    zd v = czero<double>();
    if (kind == ROW_DIRICHLET) {
      v = dirichlet_diag(P, form) * x[r];
    } else {
      const zd *node = coef + P.coef_off + (size_t)i * NODE_STRIDE;
      for (int q = 0; q < 5; ++q) {
        zd bh, dh; double vyy;
        form_coef(P, node, form, p, q, sig, bh, vyy, dh);
        zd term;
        if (!need_d1) {
          term = czero<double>();
        } else if (kind == ROW_GENERIC) {
          term = bh * X1[i * 5 + q] - X2[i * 5 + q] * vyy;
        } else if (q < 4) {
          term = walle_vyy_d1(P) ? (bh * X1[i * 5 + q] - X1[i * 5 + q] * vyy) : (bh * X1[i * 5 + q] - X2[i * 5 + q] * vyy);
        } else {
          term = czero<double>() - red[0] * vyy;
        }
        v = v + term + dh * x[i * 5 + q];
      }
    }
    y[r] = v;
  }
  __syncthreads();
}

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
// spatial [0] = C0, [1] = C1, [2] = C2.  Called by every thread of a block
// of 256, thread i for node i (live when i < ny), once the node's blocks
// are in `coef`; thread 0 .. 2 write norms[0 .. 2].
__device__ inline void point_norms(const PointDev &P, const zd *__restrict__ coef, const double *__restrict__ ops,
                                   int i, double *__restrict__ norms) {
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
  if (i < 3) norms[i] = s_red[i][0];
}
