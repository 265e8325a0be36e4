// Device-side types shared by the kernels of libstab_hip (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#define STAB_NDOF 5

template <typename T> struct cx { T x, y; };
typedef cx<double> zd;
typedef cx<float> zf;

template <typename T> __host__ __device__ inline cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __host__ __device__ inline cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> __host__ __device__ inline cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> __host__ __device__ inline cx<T> operator-(cx<T> a) { return mk(-a.x, -a.y); }
template <typename T> __host__ __device__ inline cx<T> operator*(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// complex * real, written as the reference's Fortran promotion does it:
// (a.x + i a.y)(s + 0i) = (a.x s - a.y 0) + i(a.x 0 + a.y s), which is exactly (a.x s, a.y s).
template <typename T> __host__ __device__ inline cx<T> operator*(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T> __host__ __device__ inline cx<T> operator*(T s, cx<T> a) { return mk(a.x * s, a.y * s); }
template <typename T> __host__ __device__ inline cx<T> conj(cx<T> a) { return mk(a.x, -a.y); }
template <typename T> __host__ __device__ inline T abs2(cx<T> a) { return a.x * a.x + a.y * a.y; }
template <typename T> __host__ __device__ inline T cabs(cx<T> a) { return hypot(a.x, a.y); }
template <typename T> __host__ __device__ inline T cabs1(cx<T> a) { return fabs(a.x) + fabs(a.y); }
// Smith's algorithm for complex division
template <typename T> __host__ __device__ inline cx<T> cdiv(cx<T> a, cx<T> b) {
  if (fabs(b.x) >= fabs(b.y)) {
    T r = b.y / b.x, d = b.x + b.y * r;
    return mk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    T r = b.x / b.y, d = b.y + b.x * r;
    return mk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
// a - b*c and a + b*c in complex FMA form
template <typename T> __host__ __device__ inline cx<T> cfms(cx<T> a, cx<T> b, cx<T> c) {
  a.x = fma(-b.x, c.x, a.x); a.x = fma(b.y, c.y, a.x);
  a.y = fma(-b.x, c.y, a.y); a.y = fma(-b.y, c.x, a.y);
  return a;
}
template <typename T> __host__ __device__ inline cx<T> cfma(cx<T> a, cx<T> b, cx<T> c) {
  a.x = fma(b.x, c.x, a.x); a.x = fma(-b.y, c.y, a.x);
  a.y = fma(b.x, c.y, a.y); a.y = fma(b.y, c.x, a.y);
  return a;
}
// a + conj(b)*c
template <typename T> __host__ __device__ inline cx<T> cfmac(cx<T> a, cx<T> b, cx<T> c) {
  a.x = fma(b.x, c.x, a.x); a.x = fma(b.y, c.y, a.x);
  a.y = fma(b.x, c.y, a.y); a.y = fma(-b.y, c.x, a.y);
  return a;
}
template <typename T> __host__ __device__ inline cx<T> czero() { return mk<T>(0, 0); }
// A normalisation divisor is used only when it is finite and nonzero; a vector
// whose largest entry is 0, Inf or NaN is left as it is (its residual then
// reports the failure).
// Index found by an arg-max over n entries, or 0 when no entry compared
// greater than the start value (all NaN): always a valid index.
__host__ __device__ inline int clamp_index(int bi, int n) { return (bi >= 0 && bi < n) ? bi : 0; }
template <typename T> __host__ __device__ inline bool scale_ok(cx<T> a) {
  const T m = fmax(fabs(a.x), fabs(a.y));
  return m > (T)0 && (a.x - a.x) == (T)0 && (a.y - a.y) == (T)0;   // x - x is NaN for Inf and NaN
}

// Lane moves by DPP: one VALU operation each, no trip through the LDS unit
// that __shfl_xor's lane permute takes.  update_dpp(old, src, ctrl, row_mask,
// bank_mask, bound_ctrl): a lane whose source lane is outside its row, or
// whose row is masked off, keeps `old` (here its own value).
template <int CTRL, int RM = 0xF> __device__ inline int dpp_move(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM = 0xF> __device__ inline float dpp_move(float x) {
  return __int_as_float(dpp_move<CTRL, RM>(__float_as_int(x)));
}
template <int CTRL, int RM = 0xF> __device__ inline double dpp_move(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_move<CTRL, RM>((int)(unsigned int)(b & 0xffffffffLL));
  const int hi = dpp_move<CTRL, RM>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo));
}
// the value of lane 63 of the wave in every lane
__device__ inline int lane63(int x) { return __builtin_amdgcn_readlane(x, 63); }
__device__ inline float lane63(float x) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63)); }
__device__ inline double lane63(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned int)(b & 0xffffffffLL), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo));
}
// exchange with lane ^ 1 and lane ^ 2 (quad permutations [1 0 3 2], [2 3 0 1])
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E;
template <typename T> __device__ inline T quad_xor1(T x) { return dpp_move<DPP_XOR1>(x); }
template <typename T> __device__ inline T quad_xor2(T x) { return dpp_move<DPP_XOR2>(x); }

// Arg-max of (v, i) over the 64 lanes of a wave: larger v, and the lower i
// among equal v, a total order when no v is NaN, so the result does not
// depend on the order of the reduction; returned in every lane.  Row shifts
// by 1, 2, 4 and 8 leave each 16-lane row's maximum in its lane 15, the row
// broadcasts (15 to rows 1 and 3, 31 to rows 2 and 3) combine the rows into
// lane 63, which is then read in every lane.
template <typename T> __device__ inline void wave_argmax(T &v, int &i) {
#define STAB_AM_STEP(CTRL, RM)                                                 \
  {                                                                            \
    const T ov = dpp_move<CTRL, RM>(v);                                        \
    const int oi = dpp_move<CTRL, RM>(i);                                      \
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }                     \
  }
  STAB_AM_STEP(0x111, 0xF) STAB_AM_STEP(0x112, 0xF) STAB_AM_STEP(0x114, 0xF) STAB_AM_STEP(0x118, 0xF)
  STAB_AM_STEP(0x142, 0xA) STAB_AM_STEP(0x143, 0xC)
#undef STAB_AM_STEP
  v = lane63(v);
  i = lane63(i);
}

template <typename To, typename Ti> __host__ __device__ inline cx<To> cvt(cx<Ti> a) { return mk<To>((To)a.x, (To)a.y); }

// Per-node coefficient slots (each a 5x5 complex block, row-major p*5+q):
//   temporal: BH  = (B - i a Vxy - i b Vyz) eta' - Vyy eta''   (temporal.f90:609-617)
//             DH  = D + i a A + i b C + a^2 Vxx + a b Vxz + b^2 Vzz  (temporal.f90:604-605)
//             VYY = Vyy eta'^2 (real)                          (temporal.f90:616)
//             G   = G (real)                                    (temporal.f90:313-531)
//   spatial:  BH  = (B - i b Vyz) eta' - Vyy eta''              (spatial.f90:689-697)
//             DH  = D + i b C + b^2 Vzz - i w G                 (spatial.f90:687, 808-809)
//             VYY, G as above
//             BH1 = -i Vxy eta'                                 (spatial.f90:841-849)
//             DH1 = i A + b Vxz                                 (spatial.f90:839)
//             VXX = Vxx (real)                                  (spatial.f90:942-949)
enum { SLOT_BH = 0, SLOT_DH = 1, SLOT_VYY = 2, SLOT_G = 3, SLOT_BH1 = 4, SLOT_DH1 = 5, SLOT_VXX = 6, NSLOT = 7 };
#define NODE_STRIDE (NSLOT * 25)

// Device-side description of one point of a batch.
struct PointDev {
  int itype, ny, n, npad;        // n = 5 ny; npad = n rounded up to the LU block
  int mattyp, wallt, top, navier;
  int nev, nvec, prof, nym;      // nvec = n (temporal) or 2n (spatial)
  int op_off;                    // offset (doubles) of this point's operator set
  int fd;                        // 1: finite-difference temporal point (fd_temporal.f90; itype 3)
  double Ma, Re, Pr, Te, yi, ymax;
  double dm1, dm2, dm3;          // datmat (input.f90:28-35)
  double rmue, rlme, cone;       // edge properties (input.f90:43-44)
  zd alpha, beta, omega, shift;
  long long prof_off;            // offset (doubles) of the profile rows
  long long coef_off;            // offset (zd) of the node coefficients
  long long mat_off;             // offset (elements) of the LU matrix
  long long vec_off;             // offset (zd) of the Krylov basis
  long long small_off;           // offset (zd) of H / Q scratch
  long long piv_off;             // offset (int) of pivots
  long long blk_off;             // offset (elements) of the Linv/Uinv blocks
  long long scale_off;           // offset (double) of equilibration scales (2 npad)
  long long out_off;             // offset (zd) of the output eigenvectors
};

// Pivot storage of one point (ints at piv_off): ipiv[npad] (absolute rows,
// as ZGETRF), perm[npad] (mode 0 only: row r of L U is row perm[r] of the
// input), then a mode word and one record per 64-column block of the LU:
//   mode 0: every interchange has also been applied to the columns left of
//           its block (L in final row order, the LAPACK form);
//   mode 1: it has not: the L columns of block b keep the row order reached
//           after block b, and a forward solve applies each block's net
//           interchanges to the right-hand side before that block
//           (LINPACK's order of work).  Record b: ns, then ns touched rows
//           srow[] at +1 and their source rows sorig[] at +1+128:
//           y[srow[e]] <- y[sorig[e]].
// The records are written in both modes.  After them, one trailing record per
// block (LU_TREC ints), the block's net interchanges in the form the update of
// the columns right of it reads them: nout, then topsrc[64] (input row of each
// row of the block top), outrow[64] and outsrc[64] (rows below the block that
// receive an input row of the block top).  Only the LU reads these.
#define LU_PREC 257
#define LU_TREC 193
__host__ __device__ inline long long piv_elems(int npad) {
  return 2LL * npad + 1 + (long long)(LU_PREC + LU_TREC) * ((npad + 63) / 64);
}
__host__ __device__ inline long long lu_trec_off(int npad, int block) {
  return 2LL * npad + 1 + (long long)LU_PREC * ((npad + 63) / 64) + (long long)LU_TREC * block;
}

// Matrix forms produced by the assembly kernel
enum {
  FORM_A0 = 0, FORM_B0 = 1, FORM_C0 = 2, FORM_C1 = 3, FORM_C2 = 4, FORM_SHIFTED = 5,
  // temporal pencil as the solver uses it: Dirichlet rows of A become unit rows
  // and those of B zero rows, which maps the reference's spurious omega = 0
  // eigenvalues of those rows to infinity; every other eigenpair is unchanged.
  FORM_AM = 6, FORM_BM = 7,
};

__device__ inline int wave_lane() { return threadIdx.x & 63; }
