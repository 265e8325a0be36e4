// Dense assembly of the shifted pencil into the LU workspace, plus the row and
// column statistics used for norms and for FP32 equilibration.
#include "pencil.h"

// Per-point norm accumulators (squared Frobenius norms) of the pencil parts:
//   temporal: [0] = A (solver form), [1] = B (solver form)
//   spatial:  [0] = C0, [1] = C1, [2] = C2
// plus the row scales dr (FP32 equilibration; 1 for FP64).
// grid: x = row chunks of 256, y = point.
__global__ __launch_bounds__(256) void k_rowstats(const PointDev *__restrict__ pts, const zd *__restrict__ coef,
                                                  const double *__restrict__ ops, double *__restrict__ norms,
                                                  double *__restrict__ scales, int equilibrate) {
  const PointDev P = pts[blockIdx.y];
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= P.npad) return;
  const zd sig = P.shift;
  double mx = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const int f0 = (P.itype == 1) ? FORM_AM : FORM_C0;
  const int f1 = (P.itype == 1) ? FORM_BM : FORM_C1;
  for (int c = 0; c < P.npad; ++c) {
    const zd m = pencil_entry(P, coef, ops, FORM_SHIFTED, sig, r, c);
    mx = fmax(mx, cabs(m));
    if (r < P.n && c < P.n) {
      s0 += abs2(pencil_entry(P, coef, ops, f0, sig, r, c));
      s1 += abs2(pencil_entry(P, coef, ops, f1, sig, r, c));
      if (P.itype == 2) s2 += abs2(pencil_entry(P, coef, ops, FORM_C2, sig, r, c));
    }
  }
  double *sc = scales + P.scale_off;
  sc[r] = (equilibrate && mx > 0.0) ? 1.0 / mx : 1.0;
  atomicAdd(&norms[blockIdx.y * 4 + 0], s0);
  atomicAdd(&norms[blockIdx.y * 4 + 1], s1);
  if (P.itype == 2) atomicAdd(&norms[blockIdx.y * 4 + 2], s2);
}

// Column scales dc = 1 / max_r |dr_r M(r,c)|.  grid: x = column chunks, y = point.
__global__ __launch_bounds__(256) void k_colstats(const PointDev *__restrict__ pts, const zd *__restrict__ coef,
                                                  const double *__restrict__ ops, double *__restrict__ scales) {
  const PointDev P = pts[blockIdx.y];
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= P.npad) return;
  double *sc = scales + P.scale_off;
  double mx = 0.0;
  for (int r = 0; r < P.npad; ++r) mx = fmax(mx, sc[r] * cabs(pencil_entry(P, coef, ops, FORM_SHIFTED, P.shift, r, c)));
  sc[P.npad + c] = (mx > 0.0) ? 1.0 / mx : 1.0;
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
