// Explicit finite-difference rows of the finite-difference solves
// (fd_temporal.f90, itype 3; fd_spatial.f90, itype 4) on the uniform eta
// grid, node 0 the wall and node ny-1 the top.  Host and device code.
//
// Stencils of module `stencils` (stencils.f90:1-128) that the two files use.
// The rational ones are written as fractions; the third-order boundary
// closures of Carpenter, Gottlieb & Abarbanel (1994) for the fourth-order
// interior scheme are fixed by a stability condition as well as by accuracy,
// so they are the decimals of stencils.f90:56-90.
#pragma once

namespace fd {

// first derivative, units of 1/deta
constexpr double GA[4] = {1.0 / 12.0, -2.0 / 3.0, 2.0 / 3.0, -1.0 / 12.0};   // nodes i-2, i-1, i+1, i+2
constexpr double GC[4][6] = {
    // row 0 (wall), nodes 0-5 (stencils.f90:56-61)
    {-1.87603205562073772290017534999, 3.18538322557789286703668141674, -1.81454567943752757247830550034,
     0.59165824105260274421658150053, -0.101052068000505624644095417028, 0.0145883364282753087693133500917},
    // row 1 (:65-70)
    {-0.384254232677925402043233692192, -0.290638947767348681070279077173, 0.67176478451531541138011989728,
     0.071081659837399872713651693126, -0.073630718761724245070378308432, 0.0056774548542830440901194873933},
    // row 2 (:74-79)
    {0.182885278686826206583549109115, -1.08001475417455516433881053795, 0.65787289649662525818641772732,
     0.1776170486891931456381189546, 0.076779836395827558602005515075, -0.0151403060939170046712807681567},
    // row 3 (:83-88)
    {-0.0341837103365291857798396404953, 0.224829025740103121729831613399, -0.89081233292845396245426338198,
     0.165299947710035014782196870489, 0.61343955208752529977826815383, -0.078572482272680288056193615249}};
// second derivative, units of 1/deta^2
constexpr double DA[5] = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};     // central (:97-101)
constexpr double DB[5] = {11.0 / 12.0, -5.0 / 3.0, 1.0 / 2.0, 1.0 / 3.0, -1.0 / 12.0};     // biased (:105-109)
constexpr double DD[5] = {35.0 / 12.0, -26.0 / 3.0, 19.0 / 2.0, -14.0 / 3.0, 11.0 / 12.0}; // one-sided (:120-124)
// mean-flow derivative of grad.f90:47-85, units of 1/(12 deta)
constexpr double G0[5] = {-25.0, 48.0, -36.0, 16.0, -3.0};
constexpr double G1[5] = {-3.0, -10.0, 18.0, -6.0, 1.0};

// Entry (i, j) of the pencil's first-derivative rows, times deta
// (fd_temporal.f90:633-890): closures on rows 0-3, the central row on rows
// 4 .. ny-5, rows ny-4 .. ny-2 the closures of rows 3 .. 1 mirrored and
// negated; the top row (Dirichlet) is zero.  fd_spatial.f90 (spatial =
// true) runs the central row to row ny-3 (:834-852); its rows ny-2 and ny-1
// are those of grad (g_coef, :854-903) and zero here.
__host__ __device__ inline double d1_coef(int ny, int i, int j, bool spatial = false) {
  if (i <= 3) return (j <= 5) ? GC[i][j] : 0.0;
  if (i >= ny - 1 || (spatial && i == ny - 2)) return 0.0;
  if (i <= ny - 5 || spatial) {
    const int k = j - i;
    return k == -2 ? GA[0] : k == -1 ? GA[1] : k == 1 ? GA[2] : k == 2 ? GA[3] : 0.0;
  }
  const int m = ny - 1 - i, c = ny - 1 - j;     // m = 3, 2, 1
  return (c >= 0 && c <= 5) ? -GC[m][c] : 0.0;
}

// Entry (i, j) of the pencil's second-derivative rows, times deta^2, as
// fd_temporal.f90:658-890 writes them: the central row at nodes i-2..i+2
// (nodes 1-5 on row 3, 0-4 on row 2), the biased row on rows 1 and ny-2
// (mirrored), the one-sided row on the wall row; on rows ny-4 and ny-3 the
// entries right of the diagonal enter with the opposite sign, except the
// last one of row ny-3 (:826-878).  The top row is zero.  fd_spatial.f90
// (spatial = true) keeps the central row on rows ny-4 and ny-3 (:834-852).
__host__ __device__ inline double d2_coef(int ny, int i, int j, bool spatial = false) {
  if (i == 0) return (j <= 4) ? DD[j] : 0.0;
  if (i == 1) return (j <= 4) ? DB[j] : 0.0;
  if (i == 2) return (j <= 4) ? DA[j] : 0.0;
  if (i >= ny - 1) return 0.0;
  if (i == ny - 2) {
    const int c = ny - 1 - j;
    return (c >= 0 && c <= 4) ? DB[c] : 0.0;
  }
  const int k = j - i + 2;                      // 0..4 over nodes i-2..i+2
  if (k < 0 || k > 4) return 0.0;
  if (!spatial && i == ny - 4 && k >= 3) return -DA[k];
  if (!spatial && i == ny - 3 && k == 3) return -DA[k];
  return DA[k];
}

// Entry (i, j) of the mean-flow derivative d/deta of grad.f90:47-85, times
// 12 deta: five-point rows at both ends (mirrored and negated at the top),
// the central fourth-order row between.
__host__ __device__ inline double g_coef(int ny, int i, int j) {
  if (i == 0) return (j <= 4) ? G0[j] : 0.0;
  if (i == 1) return (j <= 4) ? G1[j] : 0.0;
  const int c = ny - 1 - j;
  if (i == ny - 1) return (c >= 0 && c <= 4) ? -G0[c] : 0.0;
  if (i == ny - 2) return (c >= 0 && c <= 4) ? -G1[c] : 0.0;
  const int k = j - i;
  return k == -2 ? 1.0 : k == -1 ? -8.0 : k == 1 ? 8.0 : k == 2 ? -1.0 : 0.0;
}

}  // namespace fd
