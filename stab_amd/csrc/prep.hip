// Preparation kernels: spline of the mean-flow profile, Chebyshev operators,
// grid/mapping, interpolated mean flow and the per-node 5x5 coefficient blocks.
#include "pencil.h"
#include "fd_stencils.h"

// Second derivatives of the cubic spline through one profile column with the
// cantilever end condition, lambda = 1 (spline.f90:2-49).  One thread per
// variable: the recurrence is sequential.
__global__ void k_spline(const double *__restrict__ y, const double *__restrict__ vt,
                         double *__restrict__ fdp, double *__restrict__ work, int nym) {
  int k = threadIdx.x;
  if (k >= STAB_NDOF || nym < 4) return;
  const double *v = vt + (size_t)k * nym;
  double *f = fdp + (size_t)k * nym;
  double *b = work + (size_t)k * 3 * nym, *c = b + nym, *r = c + nym;
  const double lam = 1.0;
  c[0] = y[1] - y[0];
  for (int i = 1; i < nym - 1; ++i) {
    c[i] = y[i + 1] - y[i];
    double a = c[i - 1];
    b[i] = 2.0 * (a + c[i]);
    r[i] = 6.0 * ((v[i + 1] - v[i]) / c[i] - (v[i] - v[i - 1]) / c[i - 1]);
  }
  b[1] = b[1] + lam * c[0];
  b[nym - 2] = b[nym - 2] + lam * c[nym - 2];
  for (int i = 2; i < nym - 1; ++i) {
    double t = c[i - 1] / b[i - 1];     // A(I) = C(I-1)
    b[i] = b[i] - t * c[i - 1];
    r[i] = r[i] - t * r[i - 1];
  }
  f[nym - 2] = r[nym - 2] / b[nym - 2];
  for (int i = nym - 3; i >= 1; --i) f[i] = (r[i] - c[i] * f[i + 1]) / b[i];
  f[0] = lam * f[1];
  f[nym - 1] = lam * f[nym - 2];
}

// Chebyshev operators for one ny, in three launches: D1 by the brute-force
// Lagrange formula of chebyd.f90:25-57 (k_cheb_d1, one block), D2 = D1 D1
// (temporal.f90:136) and the last row of Dt2 = D1 Dt1 with Dt1 = D1 whose
// wall row is zeroed (temporal.f90:137-140) (k_cheb_d2, one thread per entry
// of D2, sums in ascending k), then the row sums (k_cheb_sums, one block).
// Column-major, ld = ny.
__global__ void k_cheb_d1(int ny, double *__restrict__ D1) {
  __shared__ double x[256], a[256];
  const int n = ny - 1;
  const double pi = acos(-1.0);
  for (int j = threadIdx.x; j < ny; j += blockDim.x) x[j] = cos(pi * (double)j / (double)n);
  __syncthreads();
  for (int j = threadIdx.x; j < ny; j += blockDim.x) {
    double p = 1.0;
    for (int k = 0; k < ny; ++k)
      if (k != j) p = p * (x[j] - x[k]);
    a[j] = p;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < ny; j += blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < ny; ++k)
      if (k != j) s = s + 1.0 / (x[j] - x[k]);
    D1[(size_t)j * ny + j] = s;
  }
  for (int e = threadIdx.x; e < ny * ny; e += blockDim.x) {
    const int j = e % ny, k = e / ny;     // entry (j, k)
    if (k != j) D1[(size_t)k * ny + j] = a[j] / (a[k] * (x[j] - x[k]));
  }
}

// grid: ceil(ny^2 / 256) blocks of 256 threads
__global__ void k_cheb_d2(int ny, const double *__restrict__ D1, double *__restrict__ D2, double *__restrict__ dt2w) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ny * ny) return;
  const int i = e % ny, j = e / ny;
  double s = 0.0;
  for (int k = 0; k < ny; ++k) s += D1[(size_t)k * ny + i] * D1[(size_t)j * ny + k];
  D2[(size_t)j * ny + i] = s;
  if (i == ny - 1) {
    double w = 0.0;
    for (int k = 0; k < ny - 1; ++k) w += D1[(size_t)k * ny + (ny - 1)] * D1[(size_t)j * ny + k];
    dt2w[j] = w;
  }
}

// row sums for the closed-form Frobenius norms of the pencil (point_norms):
// S11_i = sum_j D1_ij^2, S22_i = sum_j D2_ij^2, S12_i = sum_j D1_ij D2_ij,
// and Sdt = sum_j Dt2(wall, j)^2, after dt2w (ny entries)
__global__ void k_cheb_sums(int ny, const double *__restrict__ D1, const double *__restrict__ D2,
                            double *__restrict__ dt2w) {
  double *sums = dt2w + ny;
  for (int i = threadIdx.x; i < ny; i += blockDim.x) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = 0; j < ny; ++j) {
      const double d1 = D1[(size_t)j * ny + i], d2 = D2[(size_t)j * ny + i];
      a += d1 * d1; b += d2 * d2; c += d1 * d2;
    }
    sums[i] = a; sums[ny + i] = b; sums[2 * ny + i] = c;
  }
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int j = 0; j < ny; ++j) s += dt2w[j] * dt2w[j];
    sums[3 * ny] = s;
  }
}

// Finite-difference operator set for one ny (fd_temporal.f90, itype 3, or
// fd_spatial.f90, itype 4: `spatial`), one block, one thread per row i: the
// pencil's D1 and D2 rows (the wall row of D2 also in the slot of the wall
// energy row's second-derivative row), then, after the row sums, the
// mean-flow derivative d/deta of grad.f90:47-85.  Column-major, ld = ny; the
// row sums are k_cheb_sums' job.
__global__ void k_fd_ops(int ny, double *__restrict__ ops, int spatial) {
  double *D1 = ops, *D2 = D1 + (size_t)ny * ny, *d2w = D2 + (size_t)ny * ny;
  double *Dg = d2w + 4 * ny + 1;
  const double dy = 1.0 / (double)(ny - 1);       // ogengrid.f90:22
  const double h1 = 1.0 / dy, h2 = 1.0 / (dy * dy), w = 12.0 * dy;
  const int i = threadIdx.x;
  if (i >= ny) return;
  const bool grad_row = spatial && i >= ny - 2;    // fd_spatial.f90:854-903
  for (int j = 0; j < ny; ++j) {
    const double a = grad_row ? fd::g_coef(ny, i, j) / w : fd::d1_coef(ny, i, j, spatial) * h1;
    const double b = fd::d2_coef(ny, i, j, spatial) * h2;
    D1[(size_t)j * ny + i] = a;
    D2[(size_t)j * ny + i] = b;
    Dg[(size_t)j * ny + i] = fd::g_coef(ny, i, j) / w;
    if (i == 0) d2w[j] = b;
  }
}

// ---------------------------------------------------------------------------
// Per-node mean state (temporal.f90:126-307; spatial.f90:147-352, curve = 0)
// ---------------------------------------------------------------------------
struct NodeState {
  double rhom, u1m, u2m, u3m, tm;
  double gum[3][3], grhom[3], gtm[3], gpm[3];
  double divum, gdiv[3];
  double S[3][3], Sjj[3];
  double rmu, dmu, d2mu, con, dcon, d2con, rlm, dlm, d2lm;
  double gmu[3], gdmu[3], gcon[3], gdcon[3], glm[3], gdlm[3];
  double g11T, g22T, g33T, lapT;
};

// Material law getmat.f90:2-37 (array form, one element).
__device__ inline void getmat1(const PointDev &P, double t, double &rmu, double &rlm, double &con, double &dmu,
                               double &d2mu, double &dlm, double &d2lm, double &dcon, double &d2con) {
  const double cp = 1003.1, pt66 = 6.6666666666666666666e-1;
  if (P.mattyp == 0) {
    rmu = P.dm1; dmu = 0.0; d2mu = 0.0;
  } else {
    const double d1 = P.dm1, d2 = P.dm2, d3 = P.dm3;
    rmu = d1 * (t) / d2 * sqrt((t) / d2) * (d2 + d3) / (t + d3);
    dmu = (d1 * (3.0 * d3 + t) * (d3 + d2) * sqrt(t / d2)) / (2.0 * ((d3 + t) * (d3 + t)) * d2);
    d2mu = (d1 * (3.0 * d3 * d3 - 6.0 * d3 * t - t * t) * (d3 + d2)) /
           (4.0 * ((d3 + t) * (d3 + t) * (d3 + t)) * sqrt(t / d2) * (d2 * d2));
  }
  con = rmu * cp / P.Pr; dcon = dmu * cp / P.Pr; d2con = d2mu * cp / P.Pr;
  rlm = -pt66 * rmu; dlm = -pt66 * dmu; d2lm = -pt66 * d2mu;
}

__device__ void node_state(const PointDev &P, bool spatial, const double vm[5], const double g2[5],
                           const double g22[5], NodeState &s) {
  const double gamma = 1.4;
  s.rhom = vm[0]; s.u1m = vm[1]; s.u2m = spatial ? 0.0 : vm[2]; s.u3m = vm[3]; s.tm = vm[4];
  // only wall-normal (index 1) derivatives exist for a parallel flow
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) s.gum[a][b] = 0.0;
  s.gum[0][1] = g2[1]; s.gum[1][1] = g2[2]; s.gum[2][1] = g2[3];
  s.grhom[0] = 0.0; s.grhom[1] = g2[0]; s.grhom[2] = 0.0;
  s.gtm[0] = 0.0; s.gtm[1] = g2[4]; s.gtm[2] = 0.0;
  const double fact = 1.0 / (gamma * P.Ma * P.Ma);
  for (int a = 0; a < 3; ++a) s.gpm[a] = fact * (s.grhom[a] * s.tm + s.rhom * s.gtm[a]);
  if (!spatial) {
    s.divum = s.gum[0][0] + s.gum[1][1] + s.gum[2][2];
    s.gdiv[0] = 0.0 + 0.0;            // g11vm(:,2) + g12vm(:,3)
    s.gdiv[1] = 0.0 + g22[2];         // g12vm(:,2) + g22vm(:,3)
    s.gdiv[2] = 0.0;
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) s.S[a][b] = 0.5 * (s.gum[a][b] + s.gum[b][a]);
    s.Sjj[0] = 0.5 * (0.0 + 0.0 + g22[1] + 0.0 + 0.0 + 0.0);
    s.Sjj[1] = 0.5 * (0.0 + 0.0 + g22[2] + g22[2] + 0.0 + 0.0);
    s.Sjj[2] = 0.5 * (0.0 + 0.0 + g22[3] + 0.0 + 0.0 + 0.0);
  } else {
    // curve = 0: h = 1 and every metric derivative vanishes (spatial.f90:119-125)
    const double u2m = s.u2m;
    s.divum = (s.gum[0][0] + u2m * 0.0) / 1.0 + s.gum[1][1] + s.gum[2][2];
    s.gdiv[0] = 0.0;
    s.gdiv[1] = (0.0 + s.gum[1][1] * 0.0 + u2m * 0.0) + (g22[2] + 0.0);
    s.gdiv[2] = 0.0;
    s.S[0][0] = s.gum[0][0];
    s.S[0][1] = 0.5 * (s.gum[1][0] + s.gum[0][1]);
    s.S[0][2] = 0.5 * (s.gum[2][0] + s.gum[0][2]);
    s.S[1][0] = s.S[0][1];
    s.S[1][1] = s.gum[1][1];
    s.S[1][2] = 0.5 * (s.gum[2][1] + s.gum[1][2]);
    s.S[2][0] = s.S[0][2];
    s.S[2][1] = s.S[1][2];
    s.S[2][2] = s.gum[2][2];
    s.Sjj[0] = 0.5 * g22[1];
    s.Sjj[1] = g22[2];
    s.Sjj[2] = 0.5 * g22[3];
  }
  s.g11T = 0.0; s.g22T = g22[4]; s.g33T = 0.0;
  s.lapT = g22[4];                     // LapT with h = 1 (spatial.f90:302-303)
  double rmu, rlm, con, dmu, d2mu, dlm, d2lm, dcon, d2con;
  getmat1(P, s.tm * P.Te, rmu, rlm, con, dmu, d2mu, dlm, d2lm, dcon, d2con);
  const double Te = P.Te;
  s.rmu = rmu / P.rmue; s.dmu = dmu * Te / P.rmue; s.d2mu = d2mu * (Te * Te) / P.rmue;
  s.con = con / P.cone; s.dcon = dcon * Te / P.cone; s.d2con = d2con * (Te * Te) / P.cone;
  s.rlm = rlm / P.rlme; s.dlm = dlm * Te / P.rlme; s.d2lm = d2lm * (Te * Te) / P.rlme;
  for (int a = 0; a < 3; ++a) {
    s.gmu[a] = s.dmu * s.gtm[a]; s.gdmu[a] = s.d2mu * s.gtm[a];
    s.gcon[a] = s.dcon * s.gtm[a]; s.gdcon[a] = s.d2con * s.gtm[a];
    s.glm[a] = s.dlm * s.gtm[a]; s.gdlm[a] = s.d2lm * s.gtm[a];
  }
}

// One row (equation p) of the eleven 5x5 tensors, as columns q = 0..4.
struct Row5 { double G[5], A[5], B[5], C[5], D[5], Vxx[5], Vxy[5], Vyy[5], Vxz[5], Vyz[5], Vzz[5]; };

__device__ inline void zero_row(Row5 &r) {
  for (int q = 0; q < 5; ++q)
    r.G[q] = r.A[q] = r.B[q] = r.C[q] = r.D[q] = r.Vxx[q] = r.Vxy[q] = r.Vyy[q] = r.Vxz[q] = r.Vyz[q] = r.Vzz[q] = 0.0;
}

// Temporal formulation, equation p (temporal.f90:311-598).
__device__ void temporal_row(const PointDev &P, const NodeState &s, int p, Row5 &r) {
  const double gamma = 1.4, gamma1 = 0.4;
  const double Ma = P.Ma, Re = P.Re, Pr = P.Pr, gM2 = gamma * Ma * Ma;
  const bool nav = P.navier;
  zero_row(r);
  const double rhom = s.rhom, u1m = s.u1m, u2m = s.u2m, u3m = s.u3m, tm = s.tm;
  if (p == 0) {
    r.G[0] = 1.0;
    r.A[0] = u1m; r.A[1] = rhom;
    r.B[0] = u2m; r.B[2] = rhom;
    r.C[0] = u3m; r.C[3] = rhom;
    r.D[0] = s.divum; r.D[1] = s.grhom[0]; r.D[2] = s.grhom[1]; r.D[3] = s.grhom[2];
    return;
  }
  if (p >= 1 && p <= 3) {
    const int c = p - 1;             // momentum component
    r.G[p] = rhom;
    double *conv[3] = {r.A, r.B, r.C};
    conv[0][p] = rhom * u1m;
    conv[1][p] = rhom * u2m;
    conv[2][p] = rhom * u3m;
    conv[c][0] = tm / gM2;
    conv[c][4] = rhom / gM2;
    r.D[0] = u1m * s.gum[c][0] + u2m * s.gum[c][1] + u3m * s.gum[c][2] + s.gtm[c] / gM2;
    r.D[1] = rhom * s.gum[c][0];
    r.D[2] = rhom * s.gum[c][1];
    r.D[3] = rhom * s.gum[c][2];
    r.D[4] = s.grhom[c] / gM2;
    if (!nav) return;
    double f = P.rlme / (P.rmue * Re);
    if (c == 0) {        // temporal.f90:352-392
      r.A[1] = r.A[1] - f * s.glm[0];
      r.A[4] = r.A[4] - f * s.dlm * s.divum;
      r.B[2] = r.B[2] - f * s.glm[0];
      r.C[3] = r.C[3] - f * s.glm[0];
      r.D[4] = r.D[4] - f * (s.gdlm[0] * s.divum + s.dlm * s.gdiv[0]);
      r.Vxx[1] = f * s.rlm; r.Vxy[2] = f * s.rlm; r.Vxz[3] = f * s.rlm;
      f = 1.0 / Re;
      r.A[1] = r.A[1] - f * 2.0 * s.gmu[0];
      r.A[2] = r.A[2] - f * s.gmu[1];
      r.A[3] = r.A[3] - f * s.gmu[2];
      r.A[4] = r.A[4] - f * s.dmu * 2.0 * s.S[0][0];
      r.B[1] = r.B[1] - f * s.gmu[1];
      r.B[4] = r.B[4] - f * s.dmu * 2.0 * s.S[0][1];
      r.C[1] = r.C[1] - f * s.gmu[2];
      r.C[4] = r.C[4] - f * s.dmu * 2.0 * s.S[0][2];
      r.D[4] = r.D[4] - f * 2.0 * (s.gdmu[0] * s.S[0][0] + s.gdmu[1] * s.S[0][1] + s.gdmu[2] * s.S[0][2] + s.dmu * s.Sjj[0]);
      r.Vxx[1] = r.Vxx[1] + f * 2.0 * s.rmu;
      r.Vxy[2] = r.Vxy[2] + f * s.rmu;
      r.Vyy[1] = r.Vyy[1] + f * s.rmu;
      r.Vxz[3] = r.Vxz[3] + f * s.rmu;
      r.Vzz[1] = r.Vzz[1] + f * s.rmu;
    } else if (c == 1) { // temporal.f90:417-458
      r.A[1] = r.A[1] - f * s.glm[1];
      r.B[2] = r.B[2] - f * s.glm[1];
      r.B[4] = r.B[4] - f * s.dlm * s.divum;
      r.C[3] = r.C[3] - f * s.glm[1];
      r.D[4] = r.D[4] - f * (s.gdlm[1] * s.divum + s.dlm * s.gdiv[1]);
      r.Vxy[1] = f * s.rlm; r.Vyy[2] = f * s.rlm; r.Vyz[3] = f * s.rlm;
      f = 1.0 / Re;
      r.A[2] = r.A[2] - f * s.gmu[0];
      r.A[4] = r.A[4] - f * s.dmu * 2.0 * s.S[1][0];
      r.B[1] = r.B[1] - f * s.gmu[0];
      r.B[2] = r.B[2] - f * 2.0 * s.gmu[1];
      r.B[3] = r.B[3] - f * s.gmu[2];
      r.B[4] = r.B[4] - f * s.dmu * 2.0 * s.S[1][1];
      r.C[2] = r.C[2] - f * s.gmu[2];
      r.C[4] = r.C[4] - f * s.dmu * 2.0 * s.S[1][2];
      r.D[4] = r.D[4] - f * 2.0 * (s.gdmu[0] * s.S[1][0] + s.gdmu[1] * s.S[1][1] + s.gdmu[2] * s.S[1][2] + s.dmu * s.Sjj[1]);
      r.Vxx[2] = r.Vxx[2] + f * s.rmu;
      r.Vxy[1] = r.Vxy[1] + f * s.rmu;
      r.Vyy[2] = r.Vyy[2] + f * 2.0 * s.rmu;
      r.Vyz[3] = r.Vyz[3] + f * s.rmu;
      r.Vzz[2] = r.Vzz[2] + f * s.rmu;
    } else {             // temporal.f90:483-524
      r.A[1] = r.A[1] - f * s.glm[2];
      r.B[2] = r.B[2] - f * s.glm[2];
      r.C[3] = r.C[3] - f * s.glm[2];
      r.C[4] = r.C[4] - f * s.dlm * s.divum;
      r.D[4] = r.D[4] - f * (s.gdlm[2] * s.divum + s.dlm * s.gdiv[2]);
      r.Vxz[1] = f * s.rlm; r.Vyz[2] = f * s.rlm; r.Vzz[3] = f * s.rlm;
      f = 1.0 / Re;
      r.A[3] = r.A[3] - f * s.gmu[0];
      r.A[4] = r.A[4] - f * s.dmu * 2.0 * s.S[2][0];
      r.B[3] = r.B[3] - f * s.gmu[1];
      r.B[4] = r.B[4] - f * s.dmu * 2.0 * s.S[2][1];
      r.C[1] = r.C[1] - f * s.gmu[0];
      r.C[2] = r.C[2] - f * s.gmu[1];
      r.C[3] = r.C[3] - f * 2.0 * s.gmu[2];
      r.C[4] = r.C[4] - f * s.dmu * 2.0 * s.S[2][2];
      r.D[4] = r.D[4] - f * 2.0 * (s.gdmu[0] * s.S[2][0] + s.gdmu[1] * s.S[2][1] + s.gdmu[2] * s.S[2][2] + s.dmu * s.Sjj[2]);
      r.Vxx[3] = r.Vxx[3] + f * s.rmu;
      r.Vyy[3] = r.Vyy[3] + f * s.rmu;
      r.Vxz[1] = r.Vxz[1] + f * s.rmu;
      r.Vyz[2] = r.Vyz[2] + f * s.rmu;
      r.Vzz[3] = r.Vzz[3] + f * 2.0 * s.rmu;
    }
    return;
  }
  // energy (temporal.f90:530-598)
  r.G[0] = -gamma1 * tm / gamma;
  r.G[4] = rhom / gamma;
  r.A[0] = -gamma1 * u1m * tm / gamma; r.A[4] = rhom * u1m / gamma;
  r.B[0] = -gamma1 * u2m * tm / gamma; r.B[4] = rhom * u2m / gamma;
  r.C[0] = -gamma1 * u3m * tm / gamma; r.C[4] = rhom * u3m / gamma;
  r.D[0] = 1.0 / gamma * (u1m * s.gtm[0] + u2m * s.gtm[1] + u3m * s.gtm[2]);
  r.D[1] = rhom * s.gtm[0] - gamma1 * (Ma * Ma) * s.gpm[0];
  r.D[2] = rhom * s.gtm[1] - gamma1 * (Ma * Ma) * s.gpm[1];
  r.D[3] = rhom * s.gtm[2] - gamma1 * (Ma * Ma) * s.gpm[2];
  r.D[4] = -gamma1 / gamma * (u1m * s.grhom[0] + u2m * s.grhom[1] + u3m * s.grhom[2]);
  if (!nav) return;
  double f = 1.0 / (Pr * Re);
  r.A[4] = r.A[4] - f * (s.gcon[0] + s.dcon * s.gtm[0]);
  r.B[4] = r.B[4] - f * (s.gcon[1] + s.dcon * s.gtm[1]);
  r.C[4] = r.C[4] - f * (s.gcon[2] + s.dcon * s.gtm[2]);
  r.D[4] = r.D[4] - f * (s.gdcon[0] * s.gtm[0] + s.gdcon[1] * s.gtm[1] + s.gdcon[2] * s.gtm[2] +
                         s.dcon * (s.g11T + s.g22T + s.g33T));
  r.Vxx[4] = f * s.con; r.Vyy[4] = f * s.con; r.Vzz[4] = f * s.con;
  f = gamma1 * (Ma * Ma) * P.rlme / (Re * P.rmue);
  r.A[1] = r.A[1] - f * 2.0 * s.rlm * s.divum;
  r.B[2] = r.B[2] - f * 2.0 * s.rlm * s.divum;
  r.C[3] = r.C[3] - f * 2.0 * s.rlm * s.divum;
  r.D[4] = r.D[4] - f * s.dlm * s.divum * s.divum;
  f = 2.0 * gamma1 * (Ma * Ma) / Re;
  r.A[1] = r.A[1] - f * 2.0 * s.rmu * s.S[0][0];
  r.A[2] = r.A[2] - f * 2.0 * s.rmu * s.S[1][0];
  r.A[3] = r.A[3] - f * 2.0 * s.rmu * s.S[2][0];
  r.B[1] = r.B[1] - f * 2.0 * s.rmu * s.S[0][1];
  r.B[2] = r.B[2] - f * 2.0 * s.rmu * s.S[1][1];
  r.B[3] = r.B[3] - f * 2.0 * s.rmu * s.S[2][1];
  r.C[1] = r.C[1] - f * 2.0 * s.rmu * s.S[0][2];
  r.C[2] = r.C[2] - f * 2.0 * s.rmu * s.S[1][2];
  r.C[3] = r.C[3] - f * 2.0 * s.rmu * s.S[2][2];
  double ss = 0.0;
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) ss += s.S[a][b] * s.S[a][b];
  r.D[4] = r.D[4] - f * s.dmu * ss;
}

// Spatial formulation with curve = 0 (h = 1, metric derivatives zero), equation p
// (spatial.f90:356-673).  Terms multiplied by a vanishing metric derivative are
// omitted; with h = 1 every "/ h" is exact.
__device__ void spatial_row(const PointDev &P, const NodeState &s, int p, Row5 &r) {
  const double gamma = 1.4, gamma1 = 0.4;
  const double Ma = P.Ma, Re = P.Re, Pr = P.Pr, gM2 = gamma * Ma * Ma;
  const bool nav = P.navier;
  zero_row(r);
  const double rhom = s.rhom, u1m = s.u1m, u2m = s.u2m, u3m = s.u3m, tm = s.tm;
  if (p == 0) {
    r.G[0] = 1.0;
    r.A[0] = u1m; r.A[1] = rhom;
    r.B[0] = u2m; r.B[2] = rhom;
    r.C[0] = u3m; r.C[3] = rhom;
    r.D[0] = s.divum; r.D[1] = s.grhom[0]; r.D[2] = s.grhom[1]; r.D[3] = s.grhom[2];
    return;
  }
  if (p == 1) {
    r.G[1] = rhom;
    r.A[0] = tm / gM2; r.A[1] = rhom * u1m; r.A[4] = rhom / gM2;
    r.B[1] = rhom * u2m;
    r.C[1] = rhom * u3m;
    r.D[0] = u1m * s.gum[0][0] + u2m * s.gum[0][1] + u3m * s.gum[0][2] + s.gtm[0] / gM2;
    r.D[1] = rhom * s.gum[0][0];
    r.D[2] = rhom * s.gum[0][1];
    r.D[3] = rhom * s.gum[0][2];
    r.D[4] = s.grhom[0] / gM2;
    if (!nav) return;
    double f = P.rlme / (P.rmue * Re);
    r.A[1] = r.A[1] - f * s.glm[0];
    r.A[4] = r.A[4] - f * s.dlm * s.divum;
    r.B[2] = r.B[2] - f * s.glm[0];
    r.C[3] = r.C[3] - f * s.glm[0];
    r.D[4] = r.D[4] - f * (s.gdlm[0] * s.divum + s.dlm * s.gdiv[0]);
    r.Vxx[1] = f * s.rlm; r.Vxy[2] = f * s.rlm; r.Vxz[3] = f * s.rlm;
    f = 1.0 / Re;
    r.A[1] = r.A[1] - f * 2.0 * s.gmu[0];
    r.A[2] = r.A[2] - f * s.gmu[1];
    r.A[3] = r.A[3] - f * s.gmu[2];
    r.A[4] = r.A[4] - f * s.dmu * 2.0 * s.S[0][0];
    r.B[1] = r.B[1] - f * s.gmu[1];
    r.B[4] = r.B[4] - f * s.dmu * 2.0 * s.S[0][1];
    r.C[1] = r.C[1] - f * s.gmu[2];
    r.C[4] = r.C[4] - f * s.dmu * 2.0 * s.S[0][2];
    r.D[4] = r.D[4] - f * 2.0 * (s.gdmu[0] * s.S[0][0] + s.gdmu[1] * s.S[0][1] + s.gdmu[2] * s.S[0][2] + s.dmu * s.Sjj[0]);
    r.Vxx[1] = r.Vxx[1] + f * 2.0 * s.rmu;
    r.Vxy[2] = r.Vxy[2] + f * s.rmu;
    r.Vyy[1] = r.Vyy[1] + f * s.rmu;
    r.Vxz[3] = r.Vxz[3] + f * s.rmu;
    r.Vzz[1] = r.Vzz[1] + f * s.rmu;
    return;
  }
  if (p == 2) {
    r.G[2] = rhom;
    r.A[2] = rhom * u1m;
    r.B[0] = tm / gM2; r.B[2] = rhom * u2m; r.B[4] = rhom / gM2;
    r.C[2] = rhom * u3m;
    r.D[0] = u1m * s.gum[1][0] + u2m * s.gum[1][1] + u3m * s.gum[1][2] + s.gtm[1] / gM2;
    r.D[1] = rhom * s.gum[1][0];
    r.D[2] = rhom * s.gum[1][1];
    r.D[3] = rhom * s.gum[1][2];
    r.D[4] = s.grhom[1] / gM2;
    if (!nav) return;
    double f = P.rlme / (P.rmue * Re);
    r.A[1] = r.A[1] - f * s.glm[1];
    r.B[2] = r.B[2] - f * s.glm[1];
    r.B[4] = r.B[4] - f * s.dlm * s.divum;
    r.C[3] = r.C[3] - f * s.glm[1];
    r.D[4] = r.D[4] - f * (s.gdlm[1] * s.divum + s.dlm * s.gdiv[1]);
    r.Vxy[1] = f * s.rlm; r.Vyy[2] = f * s.rlm; r.Vyz[3] = f * s.rlm;
    f = 1.0 / Re;
    r.A[2] = r.A[2] - f * s.gmu[0];
    r.A[4] = r.A[4] - f * s.dmu * 2.0 * s.S[1][0];
    r.B[1] = r.B[1] - f * s.gmu[0];
    r.B[2] = r.B[2] - f * 2.0 * s.gmu[1];
    r.B[3] = r.B[3] - f * s.gmu[2];
    r.B[4] = r.B[4] - f * s.dmu * 2.0 * s.S[1][1];
    r.C[2] = r.C[2] - f * s.gmu[2];
    r.C[4] = r.C[4] - f * s.dmu * 2.0 * s.S[1][2];
    r.D[4] = r.D[4] - f * 2.0 * (s.gdmu[0] * s.S[1][0] + s.gdmu[1] * s.S[1][1] + s.gdmu[2] * s.S[1][2] + s.dmu * s.Sjj[1]);
    r.Vxx[2] = r.Vxx[2] + f * s.rmu;
    r.Vxy[1] = r.Vxy[1] + f * s.rmu;
    r.Vyy[2] = r.Vyy[2] + f * 2.0 * s.rmu;
    r.Vyz[3] = r.Vyz[3] + f * s.rmu;
    r.Vzz[2] = r.Vzz[2] + f * s.rmu;
    return;
  }
  if (p == 3) {
    r.G[3] = rhom;
    r.A[3] = rhom * u1m;
    r.B[3] = rhom * u2m;
    r.C[0] = tm / gM2; r.C[3] = rhom * u3m; r.C[4] = rhom / gM2;
    r.D[0] = u1m * s.gum[2][0] + u2m * s.gum[2][1] + u3m * s.gum[2][2] + s.gtm[2] / gM2;
    r.D[1] = rhom * s.gum[2][0];
    r.D[2] = rhom * s.gum[2][1];
    r.D[3] = rhom * s.gum[2][2];
    r.D[4] = s.grhom[2] / gM2;
    if (!nav) return;
    double f = P.rlme / (P.rmue * Re);
    r.A[1] = r.A[1] - f * s.glm[2];
    r.B[2] = r.B[2] - f * s.glm[2];
    r.C[3] = r.C[3] - f * s.glm[2];
    r.C[4] = r.C[4] - f * s.dlm * s.divum;
    r.D[4] = r.D[4] - f * (s.gdlm[2] * s.divum + s.dlm * s.gdiv[2]);
    r.Vxz[1] = f * s.rlm; r.Vyz[2] = f * s.rlm; r.Vzz[3] = f * s.rlm;
    f = 1.0 / Re;
    r.A[3] = r.A[3] - f * s.gmu[0];
    r.A[4] = r.A[4] - f * s.dmu * 2.0 * s.S[2][0];
    r.B[3] = r.B[3] - f * s.gmu[1];
    r.B[4] = r.B[4] - f * s.dmu * 2.0 * s.S[2][1];
    r.C[1] = r.C[1] - f * s.gmu[0];
    r.C[2] = r.C[2] - f * s.gmu[1];
    r.C[3] = r.C[3] - f * 2.0 * s.gmu[2];
    r.C[4] = r.C[4] - f * s.dmu * 2.0 * s.S[2][2];
    r.D[4] = r.D[4] - f * 2.0 * (s.gdmu[0] * s.S[2][0] + s.gdmu[1] * s.S[2][1] + s.gdmu[2] * s.S[2][2] + s.dmu * s.Sjj[2]);
    r.Vxx[3] = r.Vxx[3] + f * s.rmu;
    r.Vyy[3] = r.Vyy[3] + f * s.rmu;
    r.Vxz[1] = r.Vxz[1] + f * s.rmu;
    r.Vyz[2] = r.Vyz[2] + f * s.rmu;
    r.Vzz[3] = r.Vzz[3] + f * 2.0 * s.rmu;
    return;
  }
  // energy in the spatial form (spatial.f90:599-673)
  r.G[4] = rhom;
  r.A[1] = rhom * gamma1 * tm; r.A[4] = rhom * u1m;
  r.B[2] = rhom * gamma1 * tm; r.B[4] = rhom * u2m;
  r.C[3] = rhom * gamma1 * tm; r.C[4] = rhom * u3m;
  r.D[0] = u1m * s.gtm[0] + u2m * s.gtm[1] + u3m * s.gtm[2] + gamma1 * tm * s.divum;
  r.D[1] = rhom * s.gtm[0];
  r.D[2] = rhom * s.gtm[1];
  r.D[3] = rhom * s.gtm[2];
  r.D[4] = rhom * gamma1 * s.divum;
  if (!nav) return;
  double f = gamma / (Pr * Re);
  r.A[4] = r.A[4] - f * (s.gcon[0] + s.dcon * s.gtm[0]);
  r.B[4] = r.B[4] - f * (s.gcon[1] + s.dcon * s.gtm[1]);
  r.C[4] = r.C[4] - f * (s.gcon[2] + s.dcon * s.gtm[2]);
  r.D[4] = r.D[4] - f * (s.gdcon[0] * s.gtm[0] + s.gdcon[1] * s.gtm[1] + s.gdcon[2] * s.gtm[2] + s.dcon * s.lapT);
  r.Vxx[4] = f * s.con; r.Vyy[4] = f * s.con; r.Vzz[4] = f * s.con;
  f = gamma * gamma1 * (Ma * Ma) * P.rlme / (Re * P.rmue);
  r.A[1] = r.A[1] - f * 2.0 * s.rlm * s.divum;
  r.B[2] = r.B[2] - f * 2.0 * s.rlm * s.divum;
  r.C[3] = r.C[3] - f * 2.0 * s.rlm * s.divum;
  r.D[4] = r.D[4] - f * s.dlm * s.divum * s.divum;
  f = gamma * gamma1 * (Ma * Ma) / Re;
  r.A[1] = r.A[1] - f * 4.0 * s.rmu * s.S[0][0];
  r.A[2] = r.A[2] - f * 4.0 * s.rmu * s.S[1][0];
  r.A[3] = r.A[3] - f * 4.0 * s.rmu * s.S[2][0];
  r.B[1] = r.B[1] - f * 4.0 * s.rmu * s.S[0][1];
  r.B[2] = r.B[2] - f * 4.0 * s.rmu * s.S[1][1];
  r.B[3] = r.B[3] - f * 4.0 * s.rmu * s.S[2][1];
  r.C[1] = r.C[1] - f * 4.0 * s.rmu * s.S[0][2];
  r.C[2] = r.C[2] - f * 4.0 * s.rmu * s.S[1][2];
  r.C[3] = r.C[3] - f * 4.0 * s.rmu * s.S[2][2];
  double ss = 0.0;
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) ss += s.S[a][b] * s.S[a][b];
  r.D[4] = r.D[4] - f * 2.0 * s.dmu * ss;
}

// Cubic-spline value at xx (spline.f90:52-75): the interval is the first i
// with xx <= x[i+1]; x ascending, so a binary search finds the same i.
__device__ inline double speval(const double *x, const double *v, const double *f, int nym, double xx) {
  int lo = 0, hi = nym - 2;            // answer in [0, nym-2]
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (xx <= x[mid + 1]) hi = mid; else lo = mid + 1;
  }
  const int i = lo;
  const double dxm = xx - x[i], dxp = x[i + 1] - xx, del = x[i + 1] - x[i];
  return f[i] * dxp * (dxp * dxp / del - del) / 6.0 + f[i + 1] * dxm * (dxm * dxm / del - del) / 6.0 +
         v[i] * dxp / del + v[i + 1] * dxm / del;
}

// One block per point, one thread per node.  Writes the grid (4 ny), vm (5 ny),
// the node coefficient slots and the norms of the pencil parts (4 per point).
__global__ __launch_bounds__(256) void k_nodes(const PointDev *__restrict__ pts, const double *__restrict__ ops,
                                               const double *__restrict__ prof, zd *__restrict__ coef,
                                               double *__restrict__ grid, double *__restrict__ norms) {
  const PointDev P = pts[blockIdx.x];
  const int ny = P.ny, i = threadIdx.x;
  const bool spatial = (P.itype == 2);
  __shared__ double s_vm[5][256], s_g[5][256];
  const double *D1 = ops + P.op_off, *D2 = D1 + (size_t)ny * ny;
  const double *Dg = D2 + (size_t)ny * ny + 4 * ny + 1;     // finite-difference set only
  const double *py = prof + P.prof_off, *pv = py + P.nym, *pf = pv + 5 * (size_t)P.nym;
  double y = 0, eta = 0, deta = 0, d2eta = 0;
  double vm[5] = {0, 0, 0, 0, 0};
  if (i < ny && P.fd) {
    // ogengrid.f90:17-48 (yi != 0): uniform eta from the wall, y = aa eta / (bb - eta)
    const double dy = 1.0 / (double)(ny - 1);
    const double aa = P.ymax * P.yi / (P.ymax - 2.0 * P.yi), bb = 1.0 + aa / P.ymax;
    eta = (double)i * dy;
    y = aa * eta / (bb - eta);
    const double t = bb - eta, ab = aa * bb;
    deta = t * t / ab;
    d2eta = -2.0 * (t * t * t) / (ab * ab);
  } else if (i < ny) {
    // sgengrid.f90:15-45
    const double pi = 3.1415926535897932385;
    const double dth = pi / (double)(ny - 1);
    eta = cos((double)i * dth);
    if (P.ymax == 0.0) {
      const double L = P.yi;
      deta = (eta - 1.0) * (eta - 1.0) / (2.0 * L);
      d2eta = (eta - 1.0) * (eta - 1.0) * (eta - 1.0) / (2.0 * L * L);
      y = (eta != 1.0) ? L * (1.0 + eta) / (1.0 - eta) : 1.0e99;
    } else {
      const double yi = P.yi, ym = P.ymax;
      y = ym * yi * (1.0 + eta) / (1.0 + 2.0 * yi - eta);
      const double t = 2.0 * yi + 1.0 - eta;
      deta = t * t / (2.0 * ym * yi * (yi + 1.0));
      const double u = ym * yi * (yi + 1.0);
      d2eta = -0.5 * t * t * t / (u * u);
    }
  }
  if (i < ny) {
    // getmean.f90:95-112
    const double ymaxm = py[P.nym - 1];
    for (int k = 0; k < 5; ++k) {
      const double *vk = pv + (size_t)k * P.nym, *fk = pf + (size_t)k * P.nym;
      vm[k] = (y <= ymaxm) ? speval(py, vk, fk, P.nym, y) : vk[P.nym - 1];
    }
    vm[2] = (y <= ymaxm) ? vm[2] : 0.0;
    for (int k = 0; k < 5; ++k) s_vm[k][i] = vm[k];
    double *g = grid + (size_t)blockIdx.x * 9 * 256;
    g[i] = y; g[256 + i] = eta; g[512 + i] = deta; g[768 + i] = d2eta;
    for (int k = 0; k < 5; ++k) g[(4 + k) * 256 + i] = vm[k];
  }
  __syncthreads();
  double g2[5], g22[5];
  if (P.fd) {
    // grad applied to the mean flow and again to its derivative, both in
    // eta, then mapped to y (fd_temporal.f90:131-151)
    double a[5];
    for (int k = 0; k < 5; ++k) {
      a[k] = 0.0;
      if (i < ny)
        for (int j = 0; j < ny; ++j) a[k] += Dg[(size_t)j * ny + i] * s_vm[k][j];
      s_g[k][i] = a[k];
    }
    __syncthreads();
    for (int k = 0; k < 5 && i < ny; ++k) {
      double b = 0.0;
      for (int j = 0; j < ny; ++j) b += Dg[(size_t)j * ny + i] * s_g[k][j];
      g22[k] = b * (deta * deta) + a[k] * d2eta;
      g2[k] = a[k] * deta;
    }
  } else {
    // mean derivatives by collocation, mapped to y (temporal.f90:148-179)
    for (int k = 0; k < 5 && i < ny; ++k) {
      double a = 0.0, b = 0.0;
      for (int j = 0; j < ny; ++j) {
        a += D1[(size_t)j * ny + i] * s_vm[k][j];
        b += D2[(size_t)j * ny + i] * s_vm[k][j];
      }
      g22[k] = b * (deta * deta) + a * d2eta;
      g2[k] = a * deta;
    }
  }
  if (i < ny) {
    NodeState s;
    node_state(P, spatial, vm, g2, g22, s);
    // fd_spatial.f90:899-907 subtracts the time term of the wall energy row
    // (wallt = 2) with G of the node its assembly loops ended on, node ny-2 (or
    // ny-1 when top = 1), not the wall's: the energy row of the spatial G is
    // rho e_T (:574), so that term is -i omega rho(that node) in the T column.
    const bool wall_g_quirk = spatial && P.fd && P.wallt == 2 && i == 0;
    const double rho_q = s_vm[0][(P.top == 1) ? ny - 1 : ny - 2];
    zd *out = coef + P.coef_off + (size_t)i * NODE_STRIDE;
    const zd I = mk(0.0, 1.0);
    for (int p = 0; p < 5; ++p) {
      Row5 r;
      if (spatial) spatial_row(P, s, p, r); else temporal_row(P, s, p, r);
      for (int q = 0; q < 5; ++q) {
        const int e = p * 5 + q;
        zd bh, dh;
        const double vyy = r.Vyy[q];
        if (!spatial) {
          const zd al = P.alpha, be = P.beta;
          // Dh = D + im*alpha*A + im*beta*C + alpha**2*Vxx + alpha*beta*Vxz + beta**2*Vzz
          dh = mk(r.D[q], 0.0);
          dh = dh + (I * al) * r.A[q];
          dh = dh + (I * be) * r.C[q];
          dh = dh + (al * al) * r.Vxx[q];
          dh = dh + (al * be) * r.Vxz[q];
          dh = dh + (be * be) * r.Vzz[q];
          // Bh = B - im*alpha*Vxy - im*beta*Vyz, then Bh*deta - Vyy*d2eta
          bh = mk(r.B[q], 0.0);
          bh = bh - (I * al) * r.Vxy[q];
          bh = bh - (I * be) * r.Vyz[q];
          bh = bh * deta - mk(vyy * d2eta, 0.0);
        } else {
          const zd be = P.beta, om = P.omega;
          dh = mk(r.D[q], 0.0);
          dh = dh + (I * be) * r.C[q];
          dh = dh + (be * be) * r.Vzz[q];
          const double gq = (wall_g_quirk && p == 4) ? (q == 4 ? rho_q : 0.0) : r.G[q];
          dh = dh - (I * om) * gq;                         // time term (spatial.f90:808-809)
          bh = mk(r.B[q], 0.0);
          bh = bh - (I * be) * r.Vyz[q];
          bh = bh * deta - mk(vyy * d2eta, 0.0);
          out[SLOT_BH1 * 25 + e] = (mk(0.0, -1.0) * r.Vxy[q]) * deta;   // Bh = -im*Vxy, times deta
          zd d1 = I * r.A[q];
          d1 = d1 + be * r.Vxz[q];
          out[SLOT_DH1 * 25 + e] = d1;
          out[SLOT_VXX * 25 + e] = mk(r.Vxx[q], 0.0);
        }
        out[SLOT_BH * 25 + e] = bh;
        out[SLOT_DH * 25 + e] = dh;
        out[SLOT_VYY * 25 + e] = mk(vyy * (deta * deta), 0.0);
        out[SLOT_G * 25 + e] = mk(r.G[q], 0.0);
      }
    }
  }
  // the norms of the pencil parts, from the blocks this thread just wrote
  point_norms(P, coef, ops, i, norms + (size_t)blockIdx.x * 4);
}
