// agx_mg.hip -- multigrid self-gravity on gfx950 (plain FP64).
//
// The reference keeps a multigrid hierarchy per MeshBlock (levels bnx, bnx/2, .., 1) plus
// a root grid with one cell per MeshBlock (multigrid/multigrid.cpp:32-144,
// multigrid_driver.cpp:343-349), and refreshes a one-cell ghost layer from the
// neighbours before every stencil (task_list/mg_task_list.cpp:287-510).  On a uniform
// periodic mesh that hierarchy is exactly one periodic grid per level: the block
// levels and the root grid tile it, every ghost equals the neighbour's interior cell,
// and the block-coarsest level coincides with the root grid's finest.  This file keeps
// one HBM array per level for the whole mesh and reads neighbours through periodic
// index wrap, so the ghost exchanges disappear while every stencil value is the one
// the reference computes.  The control flow (V-cycle, FMG, iteration, average removal)
// follows MultigridDriver step for step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cfloat>
#include <cmath>

#include "agx_internal.hpp"

using namespace agx;

namespace {

constexpr int BS = 256;

__device__ __forceinline__ int wrapm(int i, int n) { return (i == 0) ? n - 1 : i - 1; }
__device__ __forceinline__ int wrapp(int i, int n) { return (i == n - 1) ? 0 : i + 1; }
__device__ __forceinline__ int wrap(int i, int n) { return (i < 0) ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ int64_t LIX(int n1, int n2, int k, int j, int i) {
  return ((int64_t)k * n2 + j) * n1 + i;
}

// One red-black Gauss-Seidel update (MGGravity::Smooth, gravity/mg_gravity.cpp:259-267),
// the operand order of the reference.
template <typename T>
__device__ __forceinline__ T gs_value(T uc, T ukp, T ujp, T uip, T ukm, T ujm, T uim, T s, T dx2,
                                      T isix) {
  return uc - ((T(6.0) * uc - ukp - ujp - uip - ukm - ujm - uim) + s * dx2) * isix;
}

// MGGravity::Smooth (gravity/mg_gravity.cpp:220-273), omega = 1.15 (mg_gravity.hpp:43).
// The reference sweeps i = il + ((color + k + j) & 1) in steps of 2 in block-local
// indices with a one-cell ghost; with an even number of cells per block that is the set
// of global cells with (I + J + K + color) even.  One thread per updated cell.
template <typename T>
__global__ void __launch_bounds__(BS)
k_mg_smooth(const int *__restrict__ gate, T *__restrict__ u, const T *__restrict__ src, int n1, int n2, int n3,
            T dx2, T isix, int color) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  int nh = (n1 + 1) / 2;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  int64_t tot = (int64_t)nh * n2 * n3;
  if (t >= tot) return;
  int h = (int)(t % (unsigned)nh);
  unsigned r = t / (unsigned)nh;
  int j = (int)(r % (unsigned)n2);
  int k = (int)(r / (unsigned)n2);
  int i = 2 * h + ((color + k + j) & 1);
  if (i >= n1) return;
  int64_t id = LIX(n1, n2, k, j, i);
  T uc = u[id];
  T ukp = u[LIX(n1, n2, wrapp(k, n3), j, i)], ujp = u[LIX(n1, n2, k, wrapp(j, n2), i)];
  T uip = u[LIX(n1, n2, k, j, wrapp(i, n1))];
  T ukm = u[LIX(n1, n2, wrapm(k, n3), j, i)], ujm = u[LIX(n1, n2, k, wrapm(j, n2), i)];
  T uim = u[LIX(n1, n2, k, j, wrapm(i, n1))];
  u[id] = gs_value(uc, ukp, ujp, uip, ukm, ujm, uim, src[id], dx2, isix);
}

// The same red/black sweep with one thread per (i, i+1) pair of a row (n1 even): the
// pair is read and written as one 16-B (8-B in FP32) vector, so rows are read and
// written as whole cache lines; the cell of the pair that is not of this colour is
// written back unchanged (no other thread writes it, and its readers see the same
// value either way).  Arithmetic as k_mg_smooth.
template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

template <typename T>
__global__ void __launch_bounds__(BS)
k_mg_smooth_pairs(const int *__restrict__ gate, T *__restrict__ u, const T *__restrict__ src, int n1, int n2, int n3,
                  T dx2, T isix, int color) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  using V = typename Vec2<T>::type;
  const unsigned nh = (unsigned)n1 >> 1;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 pairs per level
  if (t >= nh * (unsigned)n2 * (unsigned)n3) return;
  const unsigned h = t % nh;
  const unsigned r = t / nh;
  const int j = (int)(r % (unsigned)n2);
  const int k = (int)(r / (unsigned)n2);
  const int odd = (color + k + j) & 1;   // 0: update i = 2h, 1: update i = 2h + 1
  const int i = 2 * (int)h + odd;
  V *u2 = reinterpret_cast<V *>(u);
  const int64_t row = ((int64_t)k * n2 + j) * nh;
  V c = u2[row + h];
  V kp = u2[((int64_t)wrapp(k, n3) * n2 + j) * nh + h];
  V jp = u2[((int64_t)k * n2 + wrapp(j, n2)) * nh + h];
  V km = u2[((int64_t)wrapm(k, n3) * n2 + j) * nh + h];
  V jm = u2[((int64_t)k * n2 + wrapm(j, n2)) * nh + h];
  T uc = odd ? c.y : c.x;
  T ukp = odd ? kp.y : kp.x, ujp = odd ? jp.y : jp.x;
  T ukm = odd ? km.y : km.x, ujm = odd ? jm.y : jm.x;
  T uip = odd ? u[LIX(n1, n2, k, j, wrapp(i, n1))] : c.y;
  T uim = odd ? c.x : u[LIX(n1, n2, k, j, wrapm(i, n1))];
  T v = gs_value(uc, ukp, ujp, uip, ukm, ujm, uim, src[LIX(n1, n2, k, j, i)], dx2, isix);
  if (odd) c.y = v;
  else c.x = v;
  u2[row + h] = c;
}

// Both sweeps of one red-black pass (red = colour c0 first, then black) in one launch:
// reads uin, writes uout, so every red update sees the old black values and every black
// update the new red ones, as two in-place sweeps do.  A workgroup owns a TI x TJ tile
// of (i,j) columns and a chunk of KC planes, walking the planes upwards: it keeps the
// old values of three planes of (tile + 2-cell ring) and the new red values of three
// planes of (tile + 1-cell ring) in LDS, and writes plane q-1 of its tile (red copied,
// black updated) after computing the red values of plane q.  Red cells of the ring and
// of the planes just outside the chunk are recomputed by every workgroup that needs
// them, from the same inputs.  Periodic wrap in all directions; every extent even.
// Arithmetic: gs_value, the operands of k_mg_smooth.
constexpr int RB_TI = 64, RB_TJ = 8, RB_KC = 16;
template <typename T>
__global__ void __launch_bounds__(RB_TI * RB_TJ)
k_mg_smooth_rb(const int *__restrict__ gate,
               const T *__restrict__ uin, T *__restrict__ uout, const T *__restrict__ src, int n1, int n2,
               int n3, T dx2, T isix, int c0) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  constexpr int TI = RB_TI, TJ = RB_TJ, NT = TI * TJ, OI = TI + 4, OJ = TJ + 4, RI = TI + 2, RJ = TJ + 2;
  __shared__ T s_old[3][OJ][OI];
  __shared__ T s_red[3][RJ][RI];
  const int tid = threadIdx.x;
  const int nti = n1 / TI, ntj = n2 / TJ;
  const unsigned wg = xcd_slot(blockIdx.x, gridDim.x);
  const int tile = wg % (unsigned)(nti * ntj), chunk = wg / (unsigned)(nti * ntj);
  const int i0 = (tile % nti) * TI, j0 = (tile / nti) * TJ;
  const int k0 = chunk * RB_KC, k1 = min(k0 + RB_KC, n3);   // black/written planes [k0, k1)
  auto slot = [](int q) { return ((q % 3) + 3) % 3; };
  auto wr = [](int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); };
  auto load_old = [&](int q) {
    const int kq = wr(q, n3), sl = slot(q);
    for (int t = tid; t < OI * OJ; t += NT) {
      const int jj = t / OI, ii = t - jj * OI;
      s_old[sl][jj][ii] = uin[LIX(n1, n2, kq, wr(j0 - 2 + jj, n2), wr(i0 - 2 + ii, n1))];
    }
  };
  auto red_plane = [&](int q) {
    const int kq = wr(q, n3), so = slot(q), sm = slot(q - 1), sp = slot(q + 1);
    for (int t = tid; t < RI * RJ; t += NT) {
      const int jj = t / RI, ii = t - jj * RI;
      const int gj = wr(j0 - 1 + jj, n2), gi = wr(i0 - 1 + ii, n1);
      if ((gi + gj + kq + c0) & 1) continue;   // black cell
      s_red[so][jj][ii] = gs_value(s_old[so][jj + 1][ii + 1], s_old[sp][jj + 1][ii + 1], s_old[so][jj + 2][ii + 1],
                                   s_old[so][jj + 1][ii + 2], s_old[sm][jj + 1][ii + 1], s_old[so][jj][ii + 1],
                                   s_old[so][jj + 1][ii], src[LIX(n1, n2, kq, gj, gi)], dx2, isix);
    }
  };
  auto black_plane = [&](int q) {   // q in [k0, k1): writes the tile of plane q
    const int so = slot(q), sm = slot(q - 1), sp = slot(q + 1);
    for (int t = tid; t < NT; t += NT) {
      const int jj = t / TI, ii = t - jj * TI;
      const int gj = j0 + jj, gi = i0 + ii;
      const int64_t id = LIX(n1, n2, q, gj, gi);
      T v;
      if (((gi + gj + q + c0) & 1) == 0) {
        v = s_red[so][jj + 1][ii + 1];
      } else {
        v = gs_value(s_old[so][jj + 2][ii + 2], s_red[sp][jj + 1][ii + 1], s_red[so][jj + 2][ii + 1],
                     s_red[so][jj + 1][ii + 2], s_red[sm][jj + 1][ii + 1], s_red[so][jj][ii + 1],
                     s_red[so][jj + 1][ii], src[id], dx2, isix);
      }
      uout[id] = v;
    }
  };
  // the old values of plane q+1 are fetched into registers one iteration ahead, so their
  // latency overlaps the red and black updates of the previous iteration
  constexpr int NPF = (OI * OJ + NT - 1) / NT;
  T pf[NPF];
  auto fetch = [&](int q) {
    const int kq = wr(q, n3);
#pragma unroll
    for (int r = 0; r < NPF; ++r) {
      const int t = tid + r * NT;
      if (t < OI * OJ) {
        const int jj = t / OI, ii = t - jj * OI;
        pf[r] = uin[LIX(n1, n2, kq, wr(j0 - 2 + jj, n2), wr(i0 - 2 + ii, n1))];
      }
    }
  };
  auto stash = [&](int q) {
    const int sl = slot(q);
#pragma unroll
    for (int r = 0; r < NPF; ++r) {
      const int t = tid + r * NT;
      if (t < OI * OJ) {
        const int jj = t / OI, ii = t - jj * OI;
        s_old[sl][jj][ii] = pf[r];
      }
    }
  };
  load_old(k0 - 2);
  load_old(k0 - 1);
  fetch(k0);
  for (int q = k0 - 1; q <= k1; ++q) {
    stash(q + 1);
    if (q + 2 <= k1 + 1) fetch(q + 2);
    __syncthreads();
    red_plane(q);
    __syncthreads();
    if (q - 1 >= k0) black_plane(q - 1);
    __syncthreads();
  }
}

template <typename T>
__device__ __forceinline__ T defect_at(const T *u, const T *src, int n1, int n2,
                                       int n3, int k, int j, int i, T idx2) {
  // MGGravity::CalculateDefect, gravity/mg_gravity.cpp:283-314
  int64_t id = LIX(n1, n2, k, j, i);
  T uc = u[id];
  return (T(6.0) * uc - u[LIX(n1, n2, wrapp(k, n3), j, i)] - u[LIX(n1, n2, k, wrapp(j, n2), i)] -
          u[LIX(n1, n2, k, j, wrapp(i, n1))] - u[LIX(n1, n2, wrapm(k, n3), j, i)] -
          u[LIX(n1, n2, k, wrapm(j, n2), i)] - u[LIX(n1, n2, k, j, wrapm(i, n1))]) * idx2 + src[id];
}

// 8-cell average of the fine defect over the children of coarse cell (k, j, i)
// (Restrict, multigrid.cpp:684-724).
template <typename T>
__device__ __forceinline__ T restricted_defect(const T *u, const T *src, int n1, int n2, int n3,
                                               int k, int j, int i, T idx2) {
  int fi = 2 * i, fj = 2 * j, fk = 2 * k;
  T d000 = defect_at(u, src, n1, n2, n3, fk, fj, fi, idx2);
  T d001 = defect_at(u, src, n1, n2, n3, fk, fj, fi + 1, idx2);
  T d010 = defect_at(u, src, n1, n2, n3, fk, fj + 1, fi, idx2);
  T d011 = defect_at(u, src, n1, n2, n3, fk, fj + 1, fi + 1, idx2);
  T d100 = defect_at(u, src, n1, n2, n3, fk + 1, fj, fi, idx2);
  T d101 = defect_at(u, src, n1, n2, n3, fk + 1, fj, fi + 1, idx2);
  T d110 = defect_at(u, src, n1, n2, n3, fk + 1, fj + 1, fi, idx2);
  T d111 = defect_at(u, src, n1, n2, n3, fk + 1, fj + 1, fi + 1, idx2);
  return T(0.125) * (d000 + d001 + d010 + d011 + d100 + d101 + d110 + d111);
}

template <typename T>
__device__ __forceinline__ typename Vec2<T>::type load2(const T *a, int64_t id) {   // id even
  return *reinterpret_cast<const typename Vec2<T>::type *>(a + id);
}

// restricted_defect with the fine values read as (i, i+1) pairs: a cross of 12 row pairs
// (k-1..k+2 x j-1..j+2 without the corners), 8 i-neighbours and 4 source pairs, 24 loads
// instead of 64.  Every defect and the average use the operands and order of
// restricted_defect / defect_at, so the result is the same bit for bit.
template <typename T>
__device__ __forceinline__ T restricted_defect_pairs(const T *u, const T *src, int n1, int n2, int n3,
                                                     int k, int j, int i, T idx2) {
  using V = typename Vec2<T>::type;
  const int fi = 2 * i, fj = 2 * j, fk = 2 * k;
  const int kk[4] = {wrapm(fk, n3), fk, fk + 1, wrapp(fk + 1, n3)};
  const int jj[4] = {wrapm(fj, n2), fj, fj + 1, wrapp(fj + 1, n2)};
  const int im = wrapm(fi, n1), ip = wrapp(fi + 1, n1);
  V P[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if ((a == 1 || a == 2) || (b == 1 || b == 2)) P[a][b] = load2(u, LIX(n1, n2, kk[a], jj[b], fi));
  T d[2][2][2];
#pragma unroll
  for (int dk = 0; dk < 2; ++dk)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const int a = 1 + dk, b = 1 + dj;
      const V sv = load2(src, LIX(n1, n2, kk[a], jj[b], fi));
      const T um = u[LIX(n1, n2, kk[a], jj[b], im)], up = u[LIX(n1, n2, kk[a], jj[b], ip)];
      // di = 0: cell fi (i+1 neighbour is the pair's second value)
      d[dk][dj][0] = (T(6.0) * P[a][b].x - P[a + 1][b].x - P[a][b + 1].x - P[a][b].y - P[a - 1][b].x -
                      P[a][b - 1].x - um) * idx2 + sv.x;
      // di = 1: cell fi + 1 (i-1 neighbour is the pair's first value)
      d[dk][dj][1] = (T(6.0) * P[a][b].y - P[a + 1][b].y - P[a][b + 1].y - up - P[a - 1][b].y -
                      P[a][b - 1].y - P[a][b].x) * idx2 + sv.y;
    }
  return T(0.125) * (d[0][0][0] + d[0][0][1] + d[0][1][0] + d[0][1][1] + d[1][0][0] + d[1][0][1] +
                     d[1][1][0] + d[1][1][1]);
}

// RestrictBlock (multigrid/multigrid.cpp:331-355): defect on the fine level, 8-cell
// average into the coarse source (Restrict, multigrid.cpp:684-724), coarse u zeroed
// (ZeroClearData, non-FAS).  One thread per coarse cell.
template <typename T>
__global__ void __launch_bounds__(BS)
k_mg_defect_restrict(const int *__restrict__ gate, const T *__restrict__ u, const T *__restrict__ src,
                     T *__restrict__ csrc, T *__restrict__ cu, int n1, int n2, int n3,
                     T idx2) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  int c1 = n1 >> 1, c2 = n2 >> 1, c3 = n3 >> 1;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t >= (int64_t)c1 * c2 * c3) return;
  int i = (int)(t % (unsigned)c1);
  unsigned r = t / (unsigned)c1;
  int j = (int)(r % (unsigned)c2);
  int k = (int)(r / (unsigned)c2);
  csrc[t] = restricted_defect_pairs(u, src, n1, n2, n3, k, j, i, idx2);
  cu[t] = T(0.0);
}

// Restrict (multigrid.cpp:684-724) of the source only, for RestrictFMGSource
// (multigrid.cpp:253-262).
template <typename T>
__global__ void __launch_bounds__(BS)
k_mg_restrict_src(const T *__restrict__ src, T *__restrict__ csrc, int n1, int n2,
                  int n3) {
  int c1 = n1 >> 1, c2 = n2 >> 1, c3 = n3 >> 1;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t >= (int64_t)c1 * c2 * c3) return;
  int i = (int)(t % (unsigned)c1);
  unsigned r = t / (unsigned)c1;
  int j = (int)(r % (unsigned)c2);
  int k = (int)(r / (unsigned)c2);
  int fi = 2 * i, fj = 2 * j, fk = 2 * k;
#define S(kk, jj, ii) src[LIX(n1, n2, kk, jj, ii)]
  csrc[t] = T(0.125) * (S(fk, fj, fi) + S(fk, fj, fi + 1) + S(fk, fj + 1, fi) + S(fk, fj + 1, fi + 1) +
                     S(fk + 1, fj, fi) + S(fk + 1, fj, fi + 1) + S(fk + 1, fj + 1, fi) +
                     S(fk + 1, fj + 1, fi + 1));
#undef S
}

// Trilinear correction of fine cell (fk, fj, fi) from the coarse level (c1, c2, c3)
// (ProlongateAndCorrect, multigrid.cpp:733-829).
template <typename T>
__device__ __forceinline__ T prolong_corr(const T *cu, int c1, int c2, int c3, int fk, int fj, int fi) {
  int i = fi >> 1, j = fj >> 1, k = fk >> 1;
  int ip = wrap(i + ((fi & 1) ? 1 : -1), c1);
  int jp = wrap(j + ((fj & 1) ? 1 : -1), c2);
  int kp = wrap(k + ((fk & 1) ? 1 : -1), c3);
#define C(kk, jj, ii) cu[LIX(c1, c2, kk, jj, ii)]
  return T(0.015625) * (T(27.0) * C(k, j, i) + C(kp, jp, ip) + T(9.0) * (C(k, j, ip) + C(k, jp, i) + C(kp, j, i)) +
                        T(3.0) * (C(kp, jp, i) + C(kp, j, ip) + C(k, jp, ip)));
#undef C
}

// ProlongateAndCorrect (multigrid/multigrid.cpp:733-829): trilinear (27,9,3,1)/64
// correction.  One thread per fine cell; (dk,dj,di) = -1 for the low child, +1 for the
// high child along each axis.
template <typename T>
__global__ void __launch_bounds__(BS)
k_mg_prolong(const int *__restrict__ gate, const T *__restrict__ cu, T *__restrict__ u, int n1, int n2, int n3) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t >= (int64_t)n1 * n2 * n3) return;
  int fi = (int)(t % (unsigned)n1);
  unsigned r = t / (unsigned)n1;
  int fj = (int)(r % (unsigned)n2);
  int fk = (int)(r / (unsigned)n2);
  u[t] += prolong_corr(cu, n1 >> 1, n2 >> 1, n3 >> 1, fk, fj, fi);
}

// k_mg_prolong with one thread per (i, i+1) pair of fine cells, the two children of coarse
// cell i along x1: they share the coarse values at i-1, i, i+1 of four coarse rows, and
// the fine pair is read and written as one vector.  Per cell the arithmetic of
// prolong_corr.
template <typename T>
__global__ void __launch_bounds__(BS)
k_mg_prolong_pairs(const int *__restrict__ gate, const T *__restrict__ cu, T *__restrict__ u, int n1, int n2, int n3) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  using V = typename Vec2<T>::type;
  const unsigned nh = (unsigned)n1 >> 1;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 pairs per level
  if (t >= nh * (unsigned)n2 * (unsigned)n3) return;
  const int i = (int)(t % nh);
  const unsigned r = t / nh;
  const int fj = (int)(r % (unsigned)n2), fk = (int)(r / (unsigned)n2);
  const int c1 = n1 >> 1, c2 = n2 >> 1, c3 = n3 >> 1;
  const int j = fj >> 1, k = fk >> 1;
  const int jp = wrap(j + ((fj & 1) ? 1 : -1), c2), kp = wrap(k + ((fk & 1) ? 1 : -1), c3);
  const int im = wrap(i - 1, c1), ipp = wrap(i + 1, c1);
#define C(kk, jj, ii) cu[LIX(c1, c2, kk, jj, ii)]
  const T a0 = C(k, j, i), am = C(k, j, im), ap = C(k, j, ipp);
  const T b0 = C(k, jp, i), bm = C(k, jp, im), bp = C(k, jp, ipp);
  const T c0 = C(kp, j, i), cm = C(kp, j, im), cp = C(kp, j, ipp);
  const T e0 = C(kp, jp, i), em = C(kp, jp, im), ep = C(kp, jp, ipp);
#undef C
  // prolong_corr with ip = i - 1 (low child) and i + 1 (high child)
  const T lo = T(0.015625) * (T(27.0) * a0 + em + T(9.0) * (am + b0 + c0) + T(3.0) * (e0 + cm + bm));
  const T hi = T(0.015625) * (T(27.0) * a0 + ep + T(9.0) * (ap + b0 + c0) + T(3.0) * (e0 + cp + bp));
  V *u2 = reinterpret_cast<V *>(u);
  const int64_t id = ((int64_t)fk * n2 + fj) * nh + i;
  V v = u2[id];
  v.x += lo;
  v.y += hi;
  u2[id] = v;
}

// FMGProlongate (multigrid/multigrid.cpp:838-1048): 27-point interpolation whose
// weights are products of the 1-D weights (5,30,-3) (low child) or (-3,30,5) (high
// child) over (c-1, c, c+1); summed in (k,j,i) row-major order, then / 32768.
template <typename T>
__global__ void __launch_bounds__(BS)
k_mg_fmg_prolong(const T *__restrict__ cu, T *__restrict__ u, int n1, int n2, int n3) {
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t >= (int64_t)n1 * n2 * n3) return;
  int fi = (int)(t % (unsigned)n1);
  unsigned r = t / (unsigned)n1;
  int fj = (int)(r % (unsigned)n2);
  int fk = (int)(r / (unsigned)n2);
  int c1 = n1 >> 1, c2 = n2 >> 1, c3 = n3 >> 1;
  int i = fi >> 1, j = fj >> 1, k = fk >> 1;
  const T wlo[3] = {T(5.0), T(30.0), T(-3.0)}, whi[3] = {T(-3.0), T(30.0), T(5.0)};
  const T *wk = (fk & 1) ? whi : wlo;
  const T *wj = (fj & 1) ? whi : wlo;
  const T *wi = (fi & 1) ? whi : wlo;
  T acc = T(0.0);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int kk = wrap(k - 1 + a, c3);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      int jj = wrap(j - 1 + b, c2);
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        int ii = wrap(i - 1 + e, c1);
        T term = (wk[a] * wj[b] * wi[e]) * cu[LIX(c1, c2, kk, jj, ii)];
        acc = (a == 0 && b == 0 && e == 0) ? term : acc + term;
      }
    }
  }
  u[t] = acc / T(32768.0);
}

// The coarse end of a V(1,1) cycle in one workgroup: from level `top` (at most
// COARSE_CELLS cells) down to the 1-cell level 0 (red sweep, black sweep, defect +
// restriction per level), level 0 zeroed (SolveCoarsestGrid of a 1-cell periodic root,
// multigrid_driver.cpp:781-787), and back up (prolongation + correction, red, black).
// Each step is the per-cell arithmetic of k_mg_smooth, k_mg_defect_restrict and
// k_mg_prolong, separated by workgroup barriers (which order global-memory accesses
// within the workgroup), so the result equals the per-level launches; it replaces about
// six launches per level.
constexpr int COARSE_CELLS = 4096, COARSE_BS = 1024, MAXLEV = 24;
template <typename T>
struct CoarseLevels {
  T *u[MAXLEV];
  T *src[MAXLEV];
  int n1[MAXLEV], n2[MAXLEV], n3[MAXLEV];
  T dx2[MAXLEV], idx2[MAXLEV];
};

template <typename T>
__device__ __forceinline__ void coarse_smooth(const CoarseLevels<T> &L, int l, int color, T isix) {
  const int n1 = L.n1[l], n2 = L.n2[l], n3 = L.n3[l];
  const int nh = (n1 + 1) / 2;
  T *u = L.u[l];
  const T *src = L.src[l];
  for (int t = threadIdx.x; t < nh * n2 * n3; t += COARSE_BS) {
    int h = t % nh, r = t / nh;
    int j = r % n2, k = r / n2;
    int i = 2 * h + ((color + k + j) & 1);
    if (i >= n1) continue;
    int64_t id = LIX(n1, n2, k, j, i);
    T uc = u[id];
    T ukp = u[LIX(n1, n2, wrapp(k, n3), j, i)], ujp = u[LIX(n1, n2, k, wrapp(j, n2), i)];
    T uip = u[LIX(n1, n2, k, j, wrapp(i, n1))];
    T ukm = u[LIX(n1, n2, wrapm(k, n3), j, i)], ujm = u[LIX(n1, n2, k, wrapm(j, n2), i)];
    T uim = u[LIX(n1, n2, k, j, wrapm(i, n1))];
    u[id] = gs_value(uc, ukp, ujp, uip, ukm, ujm, uim, src[id], L.dx2[l], isix);
  }
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(COARSE_BS)
k_mg_coarse(const int *__restrict__ gate, CoarseLevels<T> L, int top, int coffset, T isix, int first_sweep_done) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  for (int l = top; l > 0; --l) {
    if (!(first_sweep_done && l == top)) coarse_smooth(L, l, 0 ^ coffset, isix);
    coarse_smooth(L, l, 1 ^ coffset, isix);
    const int c1 = L.n1[l - 1], c2 = L.n2[l - 1], c3 = L.n3[l - 1];
    for (int t = threadIdx.x; t < c1 * c2 * c3; t += COARSE_BS) {
      int i = t % c1, r = t / c1;
      int j = r % c2, k = r / c2;
      L.src[l - 1][t] = restricted_defect<T>(L.u[l], L.src[l], L.n1[l], L.n2[l], L.n3[l], k, j, i, L.idx2[l]);
      L.u[l - 1][t] = T(0.0);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) L.u[0][0] = T(0.0);
  __syncthreads();
  for (int l = 1; l <= top; ++l) {
    const int n1 = L.n1[l], n2 = L.n2[l], n3 = L.n3[l];
    for (int t = threadIdx.x; t < n1 * n2 * n3; t += COARSE_BS) {
      int fi = t % n1, r = t / n1;
      int fj = r % n2, fk = r / n2;
      L.u[l][t] += prolong_corr<T>(L.u[l - 1], L.n1[l - 1], L.n2[l - 1], L.n3[l - 1], fk, fj, fi);
    }
    __syncthreads();
    coarse_smooth(L, l, 0 ^ coffset, isix);
    coarse_smooth(L, l, 1 ^ coffset, isix);
  }
}

// Reductions: a launch of at most NPART workgroups, workgroup w summing the contiguous
// range [t0, t1) of the level (XCD-aware order); the partials are summed by one
// workgroup (k_mg_final).  The order is fixed, so results are deterministic.
constexpr int NPART = 2048;
__device__ __forceinline__ void wg_range(unsigned n, unsigned &t0, unsigned &t1) {
  const unsigned w = xcd_slot(blockIdx.x, gridDim.x);
  const unsigned per = ((n + gridDim.x - 1) / gridDim.x + BS - 1) / BS * BS;
  t0 = w * per;
  t1 = (t0 + per < n) ? t0 + per : n;
  if (t0 > n) t0 = n;
}

// block-partial sums in FP64: mode 0 sum(a), mode 1 sum(def^2)
template <int MODE, typename T>
__global__ void __launch_bounds__(BS)
k_mg_partial(const int *__restrict__ gate, const T *__restrict__ a, const T *__restrict__ src, int n1, int n2, int n3,
             T idx2, double *__restrict__ part) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  __shared__ double sh[BS];
  const unsigned ncell = (unsigned)n1 * n2 * n3;   // < 2^32 cells per level
  unsigned t0, t1;
  wg_range(ncell, t0, t1);
  double v = 0.0;
  for (unsigned t = t0 + threadIdx.x; t < t1; t += BS) {
    if (MODE == 0) {
      v += a[t];
    } else {
      int i = (int)(t % (unsigned)n1);
      unsigned r = t / (unsigned)n1;
      int j = (int)(r % (unsigned)n2);
      int k = (int)(r / (unsigned)n2);
      double d = defect_at(a, src, n1, n2, n3, k, j, i, idx2);
      v += d * d;
    }
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(BS) k_mg_final(const int *__restrict__ gate,
                                  const double *__restrict__ part, int64_t n,
                                                 double *__restrict__ out) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  __shared__ double sh[BS];
  double v = 0.0;
#pragma unroll 8
  for (int64_t q = threadIdx.x; q < n; q += BS) v += part[q];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

template <typename T>
__global__ void __launch_bounds__(BS) k_mg_add(const int *__restrict__ gate, T *__restrict__ a, int64_t n,
                                               const double *__restrict__ val, double scale) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t < n) a[t] -= T((*val) * scale);
}

template <typename T>
__global__ void __launch_bounds__(BS) k_mg_zero(const int *__restrict__ gate, T *__restrict__ a, int64_t n) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t < n) a[t] = T(0.0);
}

// Mixed-precision defect correction, FP64 side.  The FP64 defect d = src - L(phi)
// (the form of MGGravity::CalculateDefect) becomes the FP32 source of the correction
// problem L(e) = d, and its squares are summed in FP64 for the defect norm.  With ein,
// phi is first corrected by the FP32 correction of the previous cycle (phi + e, into
// uout, read back by no other thread of this launch), so the correction and the next
// defect take one pass.  e0 (not ein) receives the start of the next correction V-cycle: its first
// sweep (colour cred) from e = 0 gives gs_value(0, ..., 0, d) at those cells and 0
// elsewhere, so that cycle starts with its second sweep.
__global__ void __launch_bounds__(BS)
k_mg_defect_to32(const int *__restrict__ gate,
                 const double *__restrict__ u, const float *__restrict__ ein, double *__restrict__ uout,
                 const double *__restrict__ src, float *__restrict__ d32, float *__restrict__ e0, int n1, int n2,
                 int n3, double idx2, float dx2f, float isixf, int cred, double *__restrict__ part) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  __shared__ double sh[BS];
  const unsigned ncell = (unsigned)n1 * n2 * n3;   // < 2^32 cells per level
  unsigned t0, t1;
  wg_range(ncell, t0, t1);
  double v = 0.0;
  auto phi = [&](int64_t q) { return ein ? u[q] + (double)ein[q] : u[q]; };
  for (unsigned t = t0 + threadIdx.x; t < t1; t += BS) {
    const int i = (int)(t % (unsigned)n1);
    const unsigned r = t / (unsigned)n1;
    const int j = (int)(r % (unsigned)n2);
    const int k = (int)(r / (unsigned)n2);
    const double uc = phi(t);
    const double d = (6.0 * uc - phi(LIX(n1, n2, wrapp(k, n3), j, i)) - phi(LIX(n1, n2, k, wrapp(j, n2), i)) -
                      phi(LIX(n1, n2, k, j, wrapp(i, n1))) - phi(LIX(n1, n2, wrapm(k, n3), j, i)) -
                      phi(LIX(n1, n2, k, wrapm(j, n2), i)) - phi(LIX(n1, n2, k, j, wrapm(i, n1)))) * idx2 + src[t];
    if (uout) uout[t] = uc;
    const float df = (float)d;
    d32[t] = df;
    e0[t] = (((i + j + k + cred) & 1) == 0)
                ? gs_value(0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, df, dx2f, isixf)
                : 0.0f;
    v += d * d;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// k_mg_defect_to32 for levels with n1 % DT_TI == 0 and n2 % DT_TJ == 0, tiled like
// k_mg_smooth_rb: a workgroup owns a DT_TI x DT_TJ tile of (i,j) columns and a chunk of
// DT_KC planes, and keeps phi (= u + e) of three planes of (tile + 1-cell ring) in LDS,
// the next plane being fetched into registers while the current one is processed, so
// u and e are read about once instead of seven times.  Per cell the arithmetic is that
// of k_mg_defect_to32 (bit for bit); the partial sums of d^2 are per workgroup.
constexpr int DT_TI = 64, DT_TJ = 8, DT_KC = 16;
template <bool CORR>
__global__ void __launch_bounds__(DT_TI * DT_TJ)
k_mg_defect_to32_t(const int *__restrict__ gate,
                   const double *__restrict__ u, const float *__restrict__ ein, double *__restrict__ uout,
                   const double *__restrict__ src, float *__restrict__ d32, float *__restrict__ e0, int n1,
                   int n2, int n3, double idx2, float dx2f, float isixf, int cred, double *__restrict__ part) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  constexpr int TI = DT_TI, TJ = DT_TJ, NT = TI * TJ, RI = TI + 2, RJ = TJ + 2;
  constexpr int NPF = (RI * RJ + NT - 1) / NT;
  __shared__ double s_phi[3][RJ][RI];
  __shared__ double sh[NT];
  const int tid = threadIdx.x;
  const int nti = n1 / TI, ntj = n2 / TJ;
  const unsigned wg = xcd_slot(blockIdx.x, gridDim.x);
  const int tile = wg % (unsigned)(nti * ntj), chunk = wg / (unsigned)(nti * ntj);
  const int i0 = (tile % nti) * TI, j0 = (tile / nti) * TJ;
  const int k0 = chunk * DT_KC, k1 = min(k0 + DT_KC, n3);   // planes [k0, k1) of this workgroup
  auto slot = [](int q) { return ((q % 3) + 3) % 3; };
  auto wr = [](int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); };
  double pf[NPF];
  auto fetch = [&](int q) {
    const int kq = wr(q, n3);
#pragma unroll
    for (int r = 0; r < NPF; ++r) {
      const int t = tid + r * NT;
      if (t < RI * RJ) {
        const int jj = t / RI, ii = t - jj * RI;
        const int64_t id = LIX(n1, n2, kq, wr(j0 - 1 + jj, n2), wr(i0 - 1 + ii, n1));
        pf[r] = CORR ? u[id] + (double)ein[id] : u[id];
      }
    }
  };
  auto stash = [&](int q) {
    const int sl = slot(q);
#pragma unroll
    for (int r = 0; r < NPF; ++r) {
      const int t = tid + r * NT;
      if (t < RI * RJ) {
        const int jj = t / RI, ii = t - jj * RI;
        s_phi[sl][jj][ii] = pf[r];
      }
    }
  };
  const int jj = tid / TI, ii = tid - (tid / TI) * TI;
  const int gj = j0 + jj, gi = i0 + ii;
  double v = 0.0;
  fetch(k0 - 1);
  stash(k0 - 1);
  fetch(k0);
  stash(k0);
  fetch(k0 + 1);
  for (int q = k0; q < k1; ++q) {
    stash(q + 1);   // the slot of plane q - 2, last read before the previous barrier
    if (q + 2 <= k1) fetch(q + 2);
    __syncthreads();
    const int so = slot(q), sm = slot(q - 1), sp = slot(q + 1);
    const double uc = s_phi[so][jj + 1][ii + 1];
    const double d = (6.0 * uc - s_phi[sp][jj + 1][ii + 1] - s_phi[so][jj + 2][ii + 1] - s_phi[so][jj + 1][ii + 2] -
                      s_phi[sm][jj + 1][ii + 1] - s_phi[so][jj][ii + 1] - s_phi[so][jj + 1][ii]) * idx2;
    const int64_t id = LIX(n1, n2, q, gj, gi);
    const double dd = d + src[id];
    if (CORR) uout[id] = uc;
    const float df = (float)dd;
    d32[id] = df;
    e0[id] = (((gi + gj + q + cred) & 1) == 0)
                 ? gs_value(0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, df, dx2f, isixf)
                 : 0.0f;
    v += dd * dd;
    __syncthreads();
  }
  sh[tid] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  if (tid == 0) part[blockIdx.x] = sh[0];
}

// phi += e (mode 0) or phi = e (mode 1), FP32 correction into the FP64 iterate; mode 2
// converts an FP64 array to FP32 (src32 = src).
__global__ void __launch_bounds__(BS)
k_mg_mix(double *__restrict__ u, float *__restrict__ e, int64_t n, int mode) {
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t >= n) return;
  if (mode == 0) u[t] += (double)e[t];
  else if (mode == 1) u[t] = (double)e[t];
  else e[t] = (float)u[t];
}

// Source mask of MultigridDriver::SetupMultigrid (multigrid_driver.cpp:430-436), in the
// form of the reference's collapse problem (pgen/collapse.cpp:66-82): the source is zero
// at cell centres with |x - c|^2 > r^2.  Centres as MGCoordinates (whole-mesh faces,
// end faces exact).
struct SrcMask {
  int on;
  int n[3];
  double xmin[3], xmax[3], c[3], r2;
};
__device__ __forceinline__ double mask_centre(const SrcMask &k, int d, int i) {   // cell i = 0 .. n-1
  const double h = (k.xmax[d] - k.xmin[d]) / k.n[d];
  const double a = (i == 0) ? k.xmin[d] : k.xmin[d] + i * h;
  const double b = (i + 1 == k.n[d]) ? k.xmax[d] : k.xmin[d] + (i + 1) * h;
  return 0.5 * (a + b);
}
__device__ __forceinline__ bool masked_out(const SrcMask &k, int I, int J, int K) {
  if (!k.on) return false;
  const double x = mask_centre(k, 0, I) - k.c[0], y = mask_centre(k, 1, J) - k.c[1],
               z = mask_centre(k, 2, K) - k.c[2];
  return x * x + y * y + z * z > k.r2;
}

// LoadSource (multigrid.cpp:194-231): src = fac * u(IDN) over every block interior,
// and in MGI mode LoadFinestData (multigrid.cpp:160-180): u = phi.
__global__ void __launch_bounds__(BS)
k_mg_load(Geo g, const double *__restrict__ u, const double *__restrict__ phi,
          double *__restrict__ src, double *__restrict__ mu, double fac, int load_phi, SrcMask mk) {
  int N1 = g.nb1 * g.bx1, N2 = g.nb2 * g.bx2, N3 = g.nb3 * g.bx3;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t >= (int64_t)N1 * N2 * N3) return;
  int I = (int)(t % (unsigned)N1);
  unsigned r = t / (unsigned)N1;
  int J = (int)(r % (unsigned)N2);
  int K = (int)(r / (unsigned)N2);
  int b = ((K / g.bx3) * g.nb2 + (J / g.bx2)) * g.nb1 + (I / g.bx1);
  int64_t id = IDX(g, K % g.bx3 + g.ng, J % g.bx2 + g.ng, I % g.bx1 + g.ng);
  double v = u[(int64_t)b * NHYDRO * g.S + id];
  src[t] = masked_out(mk, I, J, K) ? 0.0 : ((fac == 1.0) ? v : v * fac);
  if (load_phi) mu[t] = phi[(int64_t)b * g.S + id];
}

// RetrieveResult with one ghost layer (multigrid.cpp:269-288).
__global__ void __launch_bounds__(BS)
k_mg_retrieve(Geo g, const double *__restrict__ mu, double *__restrict__ phi) {
  int m1 = g.bx1 + 2, m2 = g.bx2 + 2, m3 = g.bx3 + 2;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;   // < 2^32 cells per level
  if (t >= (int64_t)m1 * m2 * m3 * g.nb) return;
  int li = (int)(t % (unsigned)m1);
  unsigned r = t / (unsigned)m1;
  int lj = (int)(r % (unsigned)m2);
  r /= (unsigned)m2;
  int lk = (int)(r % (unsigned)m3);
  int b = (int)(r / (unsigned)m3);
  int N1 = g.nb1 * g.bx1, N2 = g.nb2 * g.bx2, N3 = g.nb3 * g.bx3;
  int lx1 = b % g.nb1, lx2 = (b / g.nb1) % g.nb2, lx3 = b / (g.nb1 * g.nb2);
  int I = wrap(lx1 * g.bx1 + li - 1, N1), J = wrap(lx2 * g.bx2 + lj - 1, N2),
      K = wrap(lx3 * g.bx3 + lk - 1, N3);
  phi[(int64_t)b * g.S + IDX(g, lk - 1 + g.ng, lj - 1 + g.ng, li - 1 + g.ng)] = mu[LIX(N1, N2, K, J, I)];
}


// ================================================================ multipole boundaries
// Non-periodic multigrid with MGMultipole* boundaries (multigrid/mgbval_multipole.cpp)
// and the full approximation scheme the reference uses with them
// (multigrid_driver.cpp:292-295).  Each level is one grid over the whole mesh with a
// one-cell ghost frame (MGLevelB).  A ghost is 2 phi_s(p) - (the cell inside it), phi_s
// the multipole potential at the boundary point p; edge and corner ghosts take "the cell
// inside" along the last face that writes them (x3 after x2 after x1,
// MGBoundaryValues::ApplyPhysicalBoundaries, bvals_mg.cpp:321-368), resolved here as a
// chain so that one launch fills every ghost.  mp[0..2] is the expansion origin,
// mp[3..27] the scaled coefficients.
__device__ __forceinline__ unsigned BX(int n1, int n2, int k, int j, int i) {
  return ((unsigned)k * (unsigned)(n2 + 2) + (unsigned)j) * (unsigned)(n1 + 2) + (unsigned)i;
}

struct MPGeom {
  int n1, n2, n3;
  double xmin[3], xmax[3], h[3];
};

// MGCoordinates::CalculateMGCoordinates (multigrid.cpp:1321-1350) on the whole mesh:
// face i (1 .. n+1, end faces exact) and centre i (0 .. n+1) along direction d
__device__ __forceinline__ double mp_face(const MPGeom &G, int d, int i) {
  const int n = d == 0 ? G.n1 : (d == 1 ? G.n2 : G.n3);
  if (i == 1) return G.xmin[d];
  if (i == n + 1) return G.xmax[d];
  return G.xmin[d] + (i - 1) * G.h[d];
}
__device__ __forceinline__ double mp_centre(const MPGeom &G, int d, int i) {
  return 0.5 * (mp_face(G, d, i) + mp_face(G, d, i + 1));
}

// phi_s at (x, y, z) relative to the origin, mgbval_multipole.cpp:36-88 (mporder 2, 4)
__device__ double mp_phis(const double *__restrict__ mpc, int order, double x, double y, double z) {
  double x2 = x * x, y2 = y * y, z2 = z * z;
  double xy = x * y, yz = y * z, zx = z * x;
  double r2 = x2 + y2 + z2;
  double ir2 = 1.0 / r2, ir1 = sqrt(ir2);
  double ir3 = ir2 * ir1, ir5 = ir3 * ir2;
  if (order == 2)
    return ir1 * mpc[0] + ir3 * (mpc[1] * y + mpc[2] * z + mpc[3] * x) +
           ir5 * (mpc[4] * xy + mpc[5] * yz + (3.0 * z2 - r2) * mpc[6] + mpc[7] * zx + mpc[8] * 0.5 * (x2 - y2));
  double ir7 = ir5 * ir2, ir9 = ir7 * ir2;
  double hx2my2 = 0.5 * (x2 - y2);
  double x2mty2 = x2 - 3.0 * y2;
  double tx2my2 = 3.0 * x2 - y2;
  return ir1 * mpc[0] + ir3 * (mpc[1] * y + mpc[2] * z + mpc[3] * x) +
         ir5 * (mpc[4] * xy + mpc[5] * yz + (3.0 * z2 - r2) * mpc[6] + mpc[7] * zx + mpc[8] * 0.5 * (x2 - y2)) +
         ir7 * (y * tx2my2 * mpc[9] + x * x2mty2 * mpc[15] + xy * z * mpc[10] + z * hx2my2 * mpc[14] +
                (5.0 * z2 - r2) * (y * mpc[11] + x * mpc[13]) + z * (z2 - 3.0 * r2) * mpc[12]) +
         ir9 * (xy * hx2my2 * mpc[16] + 0.125 * (x2 * x2mty2 - y2 * tx2my2) * mpc[24] + yz * tx2my2 * mpc[17] +
                zx * x2mty2 * mpc[23] + (7.0 * z2 - r2) * (xy * mpc[18] + hx2my2 * mpc[22]) +
                (7.0 * z2 - 3.0 * r2) * (yz * mpc[19] + zx * mpc[21]) +
                (35.0 * z2 * z2 - 30.0 * z2 * r2 + 3.0 * r2 * r2) * mpc[20]);
}

// Ghost q of a level's frame and its chain of boundary points inwards.  q enumerates the
// shell: the two k planes, then the two j rows of each inner plane, then the two i ends
// of each inner row.  The reference fills x1 faces, then x2 faces over all i, then x3
// faces over all (i, j), so an edge or corner ghost is 2 phi_s(p) minus the value inside
// it along the last face that writes it (at most three points); the chain ends at the
// interior cell (kk, jj, ii).  False past the last ghost.
struct GhostChain {
  int k, j, i, kk, jj, ii, depth;
  double px[3], py[3], pz[3];
};
__device__ __forceinline__ bool ghost_chain(const MPGeom &G, unsigned q, GhostChain &c) {
  const int n1 = G.n1, n2 = G.n2, n3 = G.n3;
  const unsigned w1 = n1 + 2, w2 = n2 + 2;
  const unsigned np = 2 * w1 * w2, nr = 2 * (unsigned)n3 * w1, nc = 2 * (unsigned)n3 * n2;
  if (q < np) {
    const unsigned side = q / (w1 * w2), r = q - side * w1 * w2;
    c.k = side ? n3 + 1 : 0; c.j = r / w1; c.i = r - c.j * w1;
  } else if ((q -= np) < nr) {
    const unsigned row = q / w1;
    c.i = q - row * w1; c.k = 1 + row / 2; c.j = (row & 1) ? n2 + 1 : 0;
  } else if ((q -= nr) < nc) {
    const unsigned row = q / 2;
    c.i = (q & 1) ? n1 + 1 : 0; c.k = 1 + row / n2; c.j = 1 + row % n2;
  } else {
    return false;
  }
  c.depth = 0;
  c.kk = c.k; c.jj = c.j; c.ii = c.i;
  while (c.depth < 3) {
    const int d = c.depth;
    if (c.kk == 0 || c.kk == n3 + 1) {
      c.px[d] = mp_centre(G, 0, c.ii); c.py[d] = mp_centre(G, 1, c.jj);
      c.pz[d] = mp_face(G, 2, c.kk == 0 ? 1 : n3 + 1);
      c.kk = (c.kk == 0) ? 1 : n3;
    } else if (c.jj == 0 || c.jj == n2 + 1) {
      c.px[d] = mp_centre(G, 0, c.ii); c.pz[d] = mp_centre(G, 2, c.kk);
      c.py[d] = mp_face(G, 1, c.jj == 0 ? 1 : n2 + 1);
      c.jj = (c.jj == 0) ? 1 : n2;
    } else if (c.ii == 0 || c.ii == n1 + 1) {
      c.py[d] = mp_centre(G, 1, c.jj); c.pz[d] = mp_centre(G, 2, c.kk);
      c.px[d] = mp_face(G, 0, c.ii == 0 ? 1 : n1 + 1);
      c.ii = (c.ii == 0) ? 1 : n1;
    } else {
      break;
    }
    ++c.depth;
  }
  return true;
}

// The multipole expansion at every boundary point of every ghost's chain, once per solve
// and level (the coefficients and origin are fixed during a solve): bcv[3q + d].
__global__ void __launch_bounds__(BS)
k_mb_bc_fill(double *__restrict__ bcv, MPGeom G, const double *__restrict__ mp, int order) {
  const unsigned q = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  GhostChain c;
  if (!ghost_chain(G, q, c)) return;
  for (int d = 0; d < c.depth; ++d)
    bcv[3 * (int64_t)q + d] = mp_phis(mp + 3, order, c.px[d] - mp[0], c.py[d] - mp[1], c.pz[d] - mp[2]);
}

// MGMultipoleInnerX1..OuterX3 (mgbval_multipole.cpp:28-426) on every ghost of a level in
// one launch: ghost = 2 phi_s(p) - (the value inside it), along the chain, with phi_s from
// bcv (the same doubles k_mb_bc_fill computed).
__global__ void __launch_bounds__(BS)
k_mb_bc(const int *__restrict__ gate, double *__restrict__ u, MPGeom G, const double *__restrict__ bcv) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned q = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  GhostChain c;
  if (!ghost_chain(G, q, c)) return;
  const int n1 = G.n1, n2 = G.n2;
  double v = u[BX(n1, n2, c.kk, c.jj, c.ii)];
  for (int d = c.depth - 1; d >= 0; --d) v = 2.0 * bcv[3 * (int64_t)q + d] - v;
  u[BX(n1, n2, c.k, c.j, c.i)] = v;
}

// The same chain for a correction of the potential: the boundary values are those of
// the potential itself, so the correction's ghosts take 2 * 0 - (the value inside).
__global__ void __launch_bounds__(BS)
k_mb_bc_hom(const int *__restrict__ gate, float *__restrict__ u, MPGeom G) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned q = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  GhostChain c;
  if (!ghost_chain(G, q, c)) return;
  const int n1 = G.n1, n2 = G.n2;
  float v = u[BX(n1, n2, c.kk, c.jj, c.ii)];
  for (int d = c.depth - 1; d >= 0; --d) v = -v;
  u[BX(n1, n2, c.k, c.j, c.i)] = v;
}

// red-black update of the framed level (MGGravity::Smooth, mg_gravity.cpp:220-273)
template <typename T>
__global__ void __launch_bounds__(BS)
k_mb_smooth(const int *__restrict__ gate,
            T *__restrict__ u, const T *__restrict__ src, int n1, int n2, int n3, T dx2, T isix, int color) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const int nh = (n1 + 1) / 2;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (t >= (unsigned)nh * n2 * n3) return;
  const int h = t % nh, r = t / nh;
  const int j = 1 + r % n2, k = 1 + r / n2;
  const int i = 1 + 2 * h + ((color + k + j) & 1);   // (color + k + j) parity in framed indices
  if (i > n1) return;
  const unsigned id = BX(n1, n2, k, j, i), sj = n1 + 2, sk = (unsigned)(n1 + 2) * (n2 + 2);
  u[id] = gs_value(u[id], u[id + sk], u[id + sj], u[id + 1], u[id - sk], u[id - sj], u[id - 1], src[id], dx2,
                   isix);
}


template <typename T>
__device__ __forceinline__ T mb_defect(const T *u, const T *src, unsigned id, unsigned sj, unsigned sk, T idx2) {
  return (T(6.0) * u[id] - u[id + sk] - u[id + sj] - u[id + 1] - u[id - sk] - u[id - sj] - u[id - 1]) * idx2 +
         src[id];
}

// FAS: uold = u over the framed level (StoreOldData, multigrid.cpp:645-653), then
// src -= (6u - sum of neighbours) idx2 on the interior (CalculateFASRHS,
// mg_gravity.cpp:324-351)
template <typename T>
__global__ void __launch_bounds__(BS)
k_mb_fas(const int *__restrict__ gate,
         const T *__restrict__ u, T *__restrict__ src, T *__restrict__ uold, int n1, int n2, int n3, T idx2) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  const unsigned w1 = n1 + 2, w2 = n2 + 2;
  if (t >= w1 * w2 * (unsigned)(n3 + 2)) return;
  uold[t] = u[t];
  const int i = t % w1, r = t / w1, j = r % w2, k = r / w2;
  if (i < 1 || i > n1 || j < 1 || j > n2 || k < 1 || k > n3) return;
  const unsigned sj = w1, sk = w1 * w2;
  src[t] -= (T(6.0) * u[t] - u[t + sk] - u[t + sj] - u[t + 1] - u[t - sk] - u[t - sj] - u[t - 1]) * idx2;
}

// RestrictBlock (multigrid.cpp:331-355): defect average into the coarse source and, in
// FAS, the average of u into the coarse u.  One thread per coarse interior cell.
template <typename T>
__global__ void __launch_bounds__(BS)
k_mb_restrict(const int *__restrict__ gate,
              const T *__restrict__ u, const T *__restrict__ src, T *__restrict__ csrc, T *__restrict__ cu,
              int n1, int n2, int n3, T idx2) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const int c1 = n1 >> 1, c2 = n2 >> 1, c3 = n3 >> 1;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (t >= (unsigned)c1 * c2 * c3) return;
  const int i = 1 + t % c1, r = t / c1, j = 1 + r % c2, k = 1 + r / c2;
  const int fi = 2 * i - 1, fj = 2 * j - 1, fk = 2 * k - 1;
  const unsigned sj = n1 + 2, sk = (unsigned)(n1 + 2) * (n2 + 2);
#define F(kk, jj, ii) BX(n1, n2, kk, jj, ii)
  const unsigned f[8] = {F(fk, fj, fi), F(fk, fj, fi + 1), F(fk, fj + 1, fi), F(fk, fj + 1, fi + 1),
                         F(fk + 1, fj, fi), F(fk + 1, fj, fi + 1), F(fk + 1, fj + 1, fi), F(fk + 1, fj + 1, fi + 1)};
#undef F
  T d[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) d[q] = mb_defect(u, src, f[q], sj, sk, idx2);
  const unsigned c = BX(c1, c2, k, j, i);
  csrc[c] = T(0.125) * (d[0] + d[1] + d[2] + d[3] + d[4] + d[5] + d[6] + d[7]);
  cu[c] = T(0.125) * (u[f[0]] + u[f[1]] + u[f[2]] + u[f[3]] + u[f[4]] + u[f[5]] + u[f[6]] + u[f[7]]);
}

// ProlongateAndCorrect with FAS (multigrid.cpp:361-379, 733-829): the fine interior
// gains the trilinear interpolation of (u - uold) of the framed coarse level.
template <typename T>
__global__ void __launch_bounds__(BS)
k_mb_prolong(const int *__restrict__ gate,
             const T *__restrict__ cu, const T *__restrict__ cuold, T *__restrict__ fu, int n1, int n2, int n3) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (t >= (unsigned)n1 * n2 * n3) return;
  const int fi = 1 + t % n1, r = t / n1, fj = 1 + r % n2, fk = 1 + r / n2;
  const int c1 = n1 >> 1, c2 = n2 >> 1;
  const int i = (fi - 1) / 2 + 1, j = (fj - 1) / 2 + 1, k = (fk - 1) / 2 + 1;
  const int di = ((fi - 1) & 1) ? 1 : -1, dj = ((fj - 1) & 1) ? 1 : -1, dk = ((fk - 1) & 1) ? 1 : -1;
#define C(kk, jj, ii) (cu[BX(c1, c2, kk, jj, ii)] - cuold[BX(c1, c2, kk, jj, ii)])
  fu[BX(n1, n2, fk, fj, fi)] +=
      T(0.015625) * (T(27.0) * C(k, j, i) + C(k + dk, j + dj, i + di) +
                     T(9.0) * (C(k, j, i + di) + C(k, j + dj, i) + C(k + dk, j, i)) +
                     T(3.0) * (C(k + dk, j + dj, i) + C(k + dk, j, i + di) + C(k, j + dj, i + di)));
#undef C
}

// k_mb_prolong for the two x1 children (2i-1, 2i) of each coarse cell i, which share the
// coarse (u - uold) values at i-1, i, i+1; every fine cell's sum is k_mb_prolong's, term
// for term.  n1 even (every level above the root grid).
template <typename T>
__global__ void __launch_bounds__(BS)
k_mb_prolong_pairs(const int *__restrict__ gate, const T *__restrict__ cu, const T *__restrict__ cuold,
                   T *__restrict__ fu, int n1, int n2, int n3) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const int c1 = n1 >> 1, c2 = n2 >> 1;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (t >= (unsigned)c1 * n2 * n3) return;
  const int i = 1 + t % c1, r = t / c1, fj = 1 + r % n2, fk = 1 + r / n2;
  const int j = (fj - 1) / 2 + 1, k = (fk - 1) / 2 + 1;
  const int dj = ((fj - 1) & 1) ? 1 : -1, dk = ((fk - 1) & 1) ? 1 : -1;
#define C(kk, jj, ii) (cu[BX(c1, c2, kk, jj, ii)] - cuold[BX(c1, c2, kk, jj, ii)])
  const T c0 = C(k, j, i), cj = C(k, j + dj, i), ck = C(k + dk, j, i), cjk = C(k + dk, j + dj, i);
  const unsigned f = BX(n1, n2, fk, fj, 2 * i - 1);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int di = h ? 1 : -1;
    const T ci = C(k, j, i + di), cij = C(k, j + dj, i + di), cik = C(k + dk, j, i + di),
            cijk = C(k + dk, j + dj, i + di);
    fu[f + h] += T(0.015625) * (T(27.0) * c0 + cijk + T(9.0) * (ci + cj + ck) + T(3.0) * (cjk + cik + cij));
  }
#undef C
}

// Mixed precision with multipole boundaries: the FP64 defect of the framed finest level
// (ghosts filled) becomes the FP32 source of the correction cycle, whose correction
// starts at zero over the framed level.
__global__ void __launch_bounds__(BS)
k_mb_defect32(const int *__restrict__ gate,
              const double *__restrict__ u, const double *__restrict__ src, float *__restrict__ src32,
              float *__restrict__ e32, int n1, int n2, int n3, double idx2) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  const unsigned w1 = n1 + 2, w2 = n2 + 2;
  if (t >= w1 * w2 * (unsigned)(n3 + 2)) return;
  e32[t] = 0.f;
  const int i = t % w1, r = t / w1, j = r % w2, k = r / w2;
  if (i < 1 || i > n1 || j < 1 || j > n2 || k < 1 || k > n3) return;
  src32[t] = (float)mb_defect(u, src, t, w1, w1 * w2, idx2);
}

// phi += e on the interior of the framed finest level
__global__ void __launch_bounds__(BS)
k_mb_add32(const int *__restrict__ gate, double *__restrict__ u, const float *__restrict__ e, int n1, int n2, int n3) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  const unsigned w1 = n1 + 2, w2 = n2 + 2;
  if (t >= w1 * w2 * (unsigned)(n3 + 2)) return;
  const int i = t % w1, r = t / w1, j = r % w2, k = r / w2;
  if (i < 1 || i > n1 || j < 1 || j > n2 || k < 1 || k > n3) return;
  u[t] += (double)e[t];
}

// The coarse end of a multipole V(1,1) cycle (vcycle_b, and vcycle_b32 in FP32) in one
// workgroup: levels `top` (at most MB_COARSE_CELLS interior cells) down to 0 and back,
// each step the per-element arithmetic of k_mb_bc (k_mb_bc_hom for the FP32 correction,
// HOM), k_mb_fas, k_mb_smooth, k_mb_restrict and k_mb_prolong, separated by workgroup
// barriers, so the potential equals the per-level launches bit for bit.  Every barrier
// is reached by every thread: loops have workgroup-uniform bounds.
constexpr int MB_COARSE_CELLS = 512, MB_MAXC = 8;
template <typename T>
struct CoarseB {
  T *u[MB_MAXC], *src[MB_MAXC], *uold[MB_MAXC];
  const double *bcv[MB_MAXC];
  MPGeom G[MB_MAXC];
  double dx[MB_MAXC];
};
template <typename T, bool HOM>
__device__ __forceinline__ void mbc_bc(const CoarseB<T> &L, int l) {
  const MPGeom &G = L.G[l];
  const unsigned ng = 2u * (G.n1 + 2) * (G.n2 + 2) + 2u * G.n3 * (G.n1 + 2) + 2u * G.n3 * G.n2;
  T *u = L.u[l];
  for (unsigned q = threadIdx.x; q < ng; q += COARSE_BS) {
    GhostChain c;
    if (!ghost_chain(G, q, c)) continue;
    T v = u[BX(G.n1, G.n2, c.kk, c.jj, c.ii)];
    for (int d = c.depth - 1; d >= 0; --d) {
      if (HOM) v = -v;
      else v = 2.0 * L.bcv[l][3 * (int64_t)q + d] - v;
    }
    u[BX(G.n1, G.n2, c.k, c.j, c.i)] = v;
  }
  __syncthreads();
}
template <typename T>
__device__ __forceinline__ void mbc_smooth(const CoarseB<T> &L, int l, int color) {
  const int n1 = L.G[l].n1, n2 = L.G[l].n2, n3 = L.G[l].n3, nh = (n1 + 1) / 2;
  T *u = L.u[l];
  const T *src = L.src[l];
  const T dx2 = T(L.dx[l] * L.dx[l]), isix = T(1.15 / 6.0);
  for (unsigned t = threadIdx.x; t < (unsigned)nh * n2 * n3; t += COARSE_BS) {
    const int h = t % nh, r = t / nh;
    const int j = 1 + r % n2, k = 1 + r / n2;
    const int i = 1 + 2 * h + ((color + k + j) & 1);
    if (i > n1) continue;
    const unsigned id = BX(n1, n2, k, j, i), sj = n1 + 2, sk = (unsigned)(n1 + 2) * (n2 + 2);
    u[id] = gs_value(u[id], u[id + sk], u[id + sj], u[id + 1], u[id - sk], u[id - sj], u[id - 1], src[id], dx2,
                     isix);
  }
  __syncthreads();
}
template <typename T>
__device__ __forceinline__ void mbc_fas(const CoarseB<T> &L, int l) {
  const int n1 = L.G[l].n1, n2 = L.G[l].n2, n3 = L.G[l].n3;
  const unsigned w1 = n1 + 2, w2 = n2 + 2, sj = w1, sk = w1 * w2;
  const T idx2 = T(1.0 / (L.dx[l] * L.dx[l]));
  T *src = L.src[l], *uold = L.uold[l];
  const T *u = L.u[l];
  for (unsigned t = threadIdx.x; t < w1 * w2 * (unsigned)(n3 + 2); t += COARSE_BS) {
    uold[t] = u[t];
    const int i = t % w1, r = t / w1, j = r % w2, k = r / w2;
    if (i < 1 || i > n1 || j < 1 || j > n2 || k < 1 || k > n3) continue;
    src[t] -= (T(6.0) * u[t] - u[t + sk] - u[t + sj] - u[t + 1] - u[t - sk] - u[t - sj] - u[t - 1]) * idx2;
  }
  __syncthreads();
}
template <typename T>
__device__ __forceinline__ void mbc_restrict(const CoarseB<T> &L, int l) {   // level l -> l-1
  const int n1 = L.G[l].n1, n2 = L.G[l].n2, n3 = L.G[l].n3;
  const int c1 = n1 >> 1, c2 = n2 >> 1, c3 = n3 >> 1;
  const unsigned sj = n1 + 2, sk = (unsigned)(n1 + 2) * (n2 + 2);
  const T idx2 = T(1.0 / (L.dx[l] * L.dx[l]));
  const T *u = L.u[l], *src = L.src[l];
  for (unsigned t = threadIdx.x; t < (unsigned)c1 * c2 * c3; t += COARSE_BS) {
    const int i = 1 + t % c1, r = t / c1, j = 1 + r % c2, k = 1 + r / c2;
    const int fi = 2 * i - 1, fj = 2 * j - 1, fk = 2 * k - 1;
#define F(kk, jj, ii) BX(n1, n2, kk, jj, ii)
    const unsigned f[8] = {F(fk, fj, fi), F(fk, fj, fi + 1), F(fk, fj + 1, fi), F(fk, fj + 1, fi + 1),
                           F(fk + 1, fj, fi), F(fk + 1, fj, fi + 1), F(fk + 1, fj + 1, fi), F(fk + 1, fj + 1, fi + 1)};
#undef F
    T d[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) d[q] = mb_defect(u, src, f[q], sj, sk, idx2);
    const unsigned c = BX(c1, c2, k, j, i);
    L.src[l - 1][c] = T(0.125) * (d[0] + d[1] + d[2] + d[3] + d[4] + d[5] + d[6] + d[7]);
    L.u[l - 1][c] = T(0.125) * (u[f[0]] + u[f[1]] + u[f[2]] + u[f[3]] + u[f[4]] + u[f[5]] + u[f[6]] + u[f[7]]);
  }
  __syncthreads();
}
template <typename T>
__device__ __forceinline__ void mbc_prolong(const CoarseB<T> &L, int l) {   // level l-1 -> l
  const int n1 = L.G[l].n1, n2 = L.G[l].n2, n3 = L.G[l].n3;
  const int c1 = n1 >> 1, c2 = n2 >> 1;
  const T *cu = L.u[l - 1], *cuold = L.uold[l - 1];
  T *fu = L.u[l];
  for (unsigned t = threadIdx.x; t < (unsigned)n1 * n2 * n3; t += COARSE_BS) {
    const int fi = 1 + t % n1, r = t / n1, fj = 1 + r % n2, fk = 1 + r / n2;
    const int i = (fi - 1) / 2 + 1, j = (fj - 1) / 2 + 1, k = (fk - 1) / 2 + 1;
    const int di = ((fi - 1) & 1) ? 1 : -1, dj = ((fj - 1) & 1) ? 1 : -1, dk = ((fk - 1) & 1) ? 1 : -1;
#define C(kk, jj, ii) (cu[BX(c1, c2, kk, jj, ii)] - cuold[BX(c1, c2, kk, jj, ii)])
    fu[BX(n1, n2, fk, fj, fi)] +=
        T(0.015625) * (T(27.0) * C(k, j, i) + C(k + dk, j + dj, i + di) +
                       T(9.0) * (C(k, j, i + di) + C(k, j + dj, i) + C(k + dk, j, i)) +
                       T(3.0) * (C(k + dk, j + dj, i) + C(k + dk, j, i + di) + C(k, j + dj, i + di)));
#undef C
  }
  __syncthreads();
}

template <typename T, bool HOM>
__global__ void __launch_bounds__(COARSE_BS)
k_mb_coarse(const int *__restrict__ gate, CoarseB<T> L, int top, int fmglevel, int coffset, int ni) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  for (int l = top; l > 0; --l) {
    mbc_bc<T, HOM>(L, l);
    if (l < fmglevel) mbc_fas(L, l);
    mbc_smooth(L, l, 0 ^ coffset);
    mbc_bc<T, HOM>(L, l);
    mbc_smooth(L, l, 1 ^ coffset);
    mbc_bc<T, HOM>(L, l);
    mbc_restrict(L, l);
  }
  mbc_bc<T, HOM>(L, 0);
  mbc_fas(L, 0);
  for (int it = 0; it < ni; ++it) {
    mbc_smooth(L, 0, 0 ^ coffset);
    mbc_bc<T, HOM>(L, 0);
    mbc_smooth(L, 0, 1 ^ coffset);
    mbc_bc<T, HOM>(L, 0);
  }
  for (int l = 1; l <= top; ++l) {
    mbc_bc<T, HOM>(L, l - 1);
    mbc_prolong(L, l);
    mbc_bc<T, HOM>(L, l);
    mbc_smooth(L, l, 0 ^ coffset);
    mbc_bc<T, HOM>(L, l);
    mbc_smooth(L, l, 1 ^ coffset);
    mbc_bc<T, HOM>(L, l);
  }
}

// FMGProlongate (multigrid.cpp:838-1048) from the framed coarse level
__global__ void __launch_bounds__(BS)
k_mb_fmg_prolong(const double *__restrict__ cu, double *__restrict__ fu, int n1, int n2, int n3) {
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (t >= (unsigned)n1 * n2 * n3) return;
  const int fi = 1 + t % n1, r = t / n1, fj = 1 + r % n2, fk = 1 + r / n2;
  const int c1 = n1 >> 1, c2 = n2 >> 1;
  const int i = (fi - 1) / 2 + 1, j = (fj - 1) / 2 + 1, k = (fk - 1) / 2 + 1;
  const double wlo[3] = {5.0, 30.0, -3.0}, whi[3] = {-3.0, 30.0, 5.0};
  const double *wk = ((fk - 1) & 1) ? whi : wlo;
  const double *wj = ((fj - 1) & 1) ? whi : wlo;
  const double *wi = ((fi - 1) & 1) ? whi : wlo;
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const double term = (wk[a] * wj[b] * wi[e]) * cu[BX(c1, c2, k - 1 + a, j - 1 + b, i - 1 + e)];
        acc = (a == 0 && b == 0 && e == 0) ? term : acc + term;
      }
  fu[BX(n1, n2, fk, fj, fi)] = acc / 32768.0;
}

// RestrictFMGSource for a framed level (source only)
__global__ void __launch_bounds__(BS)
k_mb_restrict_src(const double *__restrict__ src, double *__restrict__ csrc, int n1, int n2, int n3) {
  const int c1 = n1 >> 1, c2 = n2 >> 1, c3 = n3 >> 1;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (t >= (unsigned)c1 * c2 * c3) return;
  const int i = 1 + t % c1, r = t / c1, j = 1 + r % c2, k = 1 + r / c2;
  const int fi = 2 * i - 1, fj = 2 * j - 1, fk = 2 * k - 1;
#define S(kk, jj, ii) src[BX(n1, n2, kk, jj, ii)]
  csrc[BX(c1, c2, k, j, i)] = 0.125 * (S(fk, fj, fi) + S(fk, fj, fi + 1) + S(fk, fj + 1, fi) + S(fk, fj + 1, fi + 1) +
                                       S(fk + 1, fj, fi) + S(fk + 1, fj, fi + 1) + S(fk + 1, fj + 1, fi) +
                                       S(fk + 1, fj + 1, fi + 1));
#undef S
}

// squared defect of the framed finest level, partial sums per workgroup
__global__ void __launch_bounds__(BS)
k_mb_defect_partial(const int *__restrict__ gate,
                    const double *__restrict__ u, const double *__restrict__ src, int n1, int n2, int n3,
                    double idx2, double *__restrict__ part) {
  if (gate && *gate) return;   // a closed gate: the recorded V-cycle loop has ended
  __shared__ double sh[BS];
  unsigned t0, t1;
  wg_range((unsigned)n1 * n2 * n3, t0, t1);
  const unsigned sj = n1 + 2, sk = (unsigned)(n1 + 2) * (n2 + 2);
  double v = 0.0;
  for (unsigned t = t0 + threadIdx.x; t < t1; t += BS) {
    const int i = 1 + t % n1, r = t / n1, j = 1 + r % n2, k = 1 + r / n2;
    const double d = mb_defect(u, src, BX(n1, n2, k, j, i), sj, sk, idx2);
    v += d * d;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Multipole moments of the finest source (Multigrid::CalculateMultipoleCoefficients,
// multigrid.cpp:1054-1264) or, with com, the centre-of-mass sums
// (Multigrid::CalculateCenterOfMass, multigrid.cpp:1270-1298); per-workgroup partials
// part[q * gridDim.x + w] for moment q.
constexpr int NMOM = 25;
__global__ void __launch_bounds__(BS)
k_mb_moments(const double *__restrict__ src, MPGeom G, const double *__restrict__ mp, int order, int com,
             double *__restrict__ part) {
  __shared__ double sh[BS];
  const int n1 = G.n1, n2 = G.n2, n3 = G.n3;
  unsigned t0, t1;
  wg_range((unsigned)n1 * n2 * n3, t0, t1);
  const double vol = (mp_face(G, 0, 2) - mp_face(G, 0, 1)) * (mp_face(G, 1, 2) - mp_face(G, 1, 1)) *
                     (mp_face(G, 2, 2) - mp_face(G, 2, 1));
  const double xo = com ? 0.0 : mp[0], yo = com ? 0.0 : mp[1], zo = com ? 0.0 : mp[2];
  const int nm = com ? 4 : (order == 4 ? 25 : 9);
  double m[NMOM];
#pragma unroll
  for (int q = 0; q < NMOM; ++q) m[q] = 0.0;
  for (unsigned t = t0 + threadIdx.x; t < t1; t += BS) {
    const int i = 1 + t % n1, r = t / n1, j = 1 + r % n2, k = 1 + r / n2;
    const double z = mp_centre(G, 2, k) - zo, y = mp_centre(G, 1, j) - yo, x = mp_centre(G, 0, i) - xo;
    const double s = src[BX(n1, n2, k, j, i)] * vol;
    m[0] += s;
    m[1] += s * y;
    m[2] += s * z;
    m[3] += s * x;
    if (com) continue;
    const double z2 = z * z, y2 = y * y, yz = y * z, x2 = x * x, xy = x * y, zx = z * x;
    const double r2 = x2 + y2 + z2;
    m[4] += s * xy;
    m[5] += s * yz;
    m[6] += s * (3.0 * z2 - r2);
    m[7] += s * zx;
    if (order == 2) {
      m[8] += s * 0.5 * (x2 - y2);
      continue;
    }
    const double hx2my2 = 0.5 * (x2 - y2), x2mty2 = x2 - 3.0 * y2, tx2my2 = 3.0 * x2 - y2;
    const double fz2mr2 = 5.0 * z2 - r2, sz2mr2 = 7.0 * z2 - r2, sz2mtr2 = 7.0 * z2 - 3.0 * r2;
    m[8] += s * hx2my2;
    m[9] += s * y * tx2my2;
    m[10] += s * xy * z;
    m[11] += s * y * fz2mr2;
    m[12] += s * z * (z2 - 3.0 * r2);
    m[13] += s * x * fz2mr2;
    m[14] += s * z * hx2my2;
    m[15] += s * x * x2mty2;
    m[16] += s * xy * hx2my2;
    m[17] += s * yz * tx2my2;
    m[18] += s * xy * sz2mr2;
    m[19] += s * yz * sz2mtr2;
    m[20] += s * (35.0 * z2 * z2 - 30.0 * z2 * r2 + 3.0 * r2 * r2);
    m[21] += s * zx * sz2mtr2;
    m[22] += s * hx2my2 * sz2mr2;
    m[23] += s * zx * x2mty2;
    m[24] += s * 0.125 * (x2 * x2mty2 - y2 * tx2my2);
  }
  for (int q = 0; q < nm; ++q) {
    sh[threadIdx.x] = m[q];
    __syncthreads();
    for (int st = BS / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(unsigned)q * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
}

// Sums the partials in a fixed order; com: origin = (m3, m1, m2) / m0
// (MultigridDriver::CalculateCenterOfMass, multigrid_driver.cpp:1984-2015); otherwise
// the coefficients scaled as ScaleMultipoleCoefficients (multigrid_driver.cpp:1930-1981).
__global__ void __launch_bounds__(BS)
k_mb_moments_final(const double *__restrict__ part, int nwg, int order, int com, int nodipole,
                   double *__restrict__ mp) {
  __shared__ double sh[BS];
  const int nm = com ? 4 : (order == 4 ? 25 : 9);
  double acc[NMOM];
  for (int q = 0; q < nm; ++q) {
    double v = 0.0;
    for (int w = threadIdx.x; w < nwg; w += BS) v += part[q * nwg + w];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int st = BS / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
      __syncthreads();
    }
    acc[q] = sh[0];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (com) {
    const double im = 1.0 / acc[0];
    mp[0] = im * acc[3];
    mp[1] = im * acc[1];
    mp[2] = im * acc[2];
    return;
  }
  const double PI_ = 3.14159265358979323846;
  const double c0 = -0.25 / PI_, c1 = -0.25 / PI_, c2 = -0.0625 / PI_, c2a = -0.75 / PI_;
  const double c30 = -0.0625 / PI_, c31 = -0.0625 * 1.5 / PI_, c32 = -0.25 * 15.0 / PI_, c33 = -0.0625 * 2.5 / PI_;
  const double c40 = -0.0625 * 0.0625 / PI_, c41 = -0.0625 * 2.5 / PI_, c42 = -0.0625 * 5.0 / PI_,
               c43 = -0.0625 * 17.5 / PI_, c44 = -0.25 * 35.0 / PI_;
  const double sc[NMOM] = {c0, c1, c1, c1, c2a, c2a, c2, c2a, c2a, c33, c32, c31, c30,
                           c31, c32, c33, c44, c43, c42, c41, c40, c41, c42, c43, c44};
  if (nodipole) acc[1] = acc[2] = acc[3] = 0.0;   // nodipole: those sums are never formed
  for (int q = 0; q < NMOM; ++q) mp[3 + q] = (q < nm) ? acc[q] * sc[q] : 0.0;
}

// LoadSource (multigrid.cpp:194-231) into the framed finest level; MGI mode also loads
// u = phi (LoadFinestData, multigrid.cpp:160-180)
__global__ void __launch_bounds__(BS)
k_mb_load(Geo g, const double *__restrict__ u, const double *__restrict__ phi, double *__restrict__ src,
          double *__restrict__ mu, double fac, int load_phi, SrcMask mk) {
  const int N1 = g.nb1 * g.bx1, N2 = g.nb2 * g.bx2, N3 = g.nb3 * g.bx3;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (t >= (unsigned)N1 * N2 * N3) return;
  const int I = t % N1, r = t / N1, J = r % N2, K = r / N2;
  const int b = ((K / g.bx3) * g.nb2 + (J / g.bx2)) * g.nb1 + (I / g.bx1);
  const int64_t id = IDX(g, K % g.bx3 + g.ng, J % g.bx2 + g.ng, I % g.bx1 + g.ng);
  const double v = u[(int64_t)b * NHYDRO * g.S + id];
  const unsigned f = BX(N1, N2, K + 1, J + 1, I + 1);
  src[f] = masked_out(mk, I, J, K) ? 0.0 : ((fac == 1.0) ? v : v * fac);
  if (load_phi) mu[f] = phi[(int64_t)b * g.S + id];
}

// RetrieveResult with one ghost layer (multigrid.cpp:269-288) from the framed finest level
__global__ void __launch_bounds__(BS)
k_mb_retrieve(Geo g, const double *__restrict__ mu, double *__restrict__ phi) {
  const int m1 = g.bx1 + 2, m2 = g.bx2 + 2, m3 = g.bx3 + 2;
  const unsigned t = xcd_slot(blockIdx.x, gridDim.x) * BS + threadIdx.x;
  if (t >= (unsigned)m1 * m2 * m3 * g.nb) return;
  const int li = t % m1;
  unsigned r = t / m1;
  const int lj = r % m2;
  r /= m2;
  const int lk = r % m3, b = r / m3;
  const int N1 = g.nb1 * g.bx1, N2 = g.nb2 * g.bx2;
  const int lx1 = b % g.nb1, lx2 = (b / g.nb1) % g.nb2, lx3 = b / (g.nb1 * g.nb2);
  const int I = lx1 * g.bx1 + li, J = lx2 * g.bx2 + lj, K = lx3 * g.bx3 + lk;   // framed: global + 1
  phi[(int64_t)b * g.S + IDX(g, lk - 1 + g.ng, lj - 1 + g.ng, li - 1 + g.ng)] = mu[BX(N1, N2, K, J, I)];
}

}  // namespace

namespace agx {

template <typename T>
static int64_t lcells(const MGLevelT<T> &L) { return (int64_t)L.n1 * L.n2 * L.n3; }
static int64_t pad32(int64_t n) { return (n + 63) / 64 * 64; }   // 64 elements: >= 256 B
static int64_t npart(int64_t n) { return std::min<int64_t>(grid1d(n, BS), NPART); }

static double finish_norm(agx_mesh *m);
static void first_defect_sum(agx_mesh *m);
static void loop_step(agx_mesh *m);

// ---------------------------------------------------------------- multipole hierarchy
static int64_t bcells(const MGLevelB &L) { return (int64_t)(L.n1 + 2) * (L.n2 + 2) * (L.n3 + 2); }
static int64_t bghosts(const MGLevelB &L) { return bcells(L) - (int64_t)L.n1 * L.n2 * L.n3; }

// Framed levels with the same dimensions and smoother cell sizes as the periodic ones.
static int mp_create(agx_mesh *m) {
  const int L = m->ntotal;
  m->levb.resize(L);
  int64_t tot = 0;
  for (int l = 0; l < L; ++l) {
    MGLevelB &B = m->levb[l];
    B.n1 = m->lev[l].n1; B.n2 = m->lev[l].n2; B.n3 = m->lev[l].n3;
    B.dx = m->lev[l].dx;
    B.h[0] = (m->d.x1max - m->d.x1min) / B.n1;
    B.h[1] = (m->d.x2max - m->d.x2min) / B.n2;
    B.h[2] = (m->d.x3max - m->d.x3min) / B.n3;
    tot += 3 * pad32(bcells(B)) + pad32(3 * bghosts(B));
  }
  AGX_HIP(hipMalloc(&m->mgpoolb, tot * (int64_t)sizeof(double)));
  AGX_HIP(hipMemsetAsync(m->mgpoolb, 0, tot * sizeof(double), m->stream));
  m->mg_bytes += tot * (int64_t)sizeof(double);
  double *p = m->mgpoolb;
  for (int l = 0; l < L; ++l) {
    MGLevelB &B = m->levb[l];
    const int64_t n = pad32(bcells(B));
    B.u = p; p += n;
    B.src = p; p += n;
    B.uold = p; p += n;
    B.bcv = p; p += pad32(3 * bghosts(B));
  }
  AGX_HIP(hipMalloc(&m->mpdev, sizeof(double) * (28 + NMOM * NPART)));
  AGX_HIP(hipMemsetAsync(m->mpdev, 0, sizeof(double) * (28 + NMOM * NPART), m->stream));
  m->mp_on = true;
  return 0;
}

static SrcMask src_mask(const agx_mesh *m) {
  SrcMask k;
  k.on = m->srcmask;
  const Geo &g = m->g;
  k.n[0] = g.nb1 * g.bx1; k.n[1] = g.nb2 * g.bx2; k.n[2] = g.nb3 * g.bx3;
  k.xmin[0] = m->d.x1min; k.xmin[1] = m->d.x2min; k.xmin[2] = m->d.x3min;
  k.xmax[0] = m->d.x1max; k.xmax[1] = m->d.x2max; k.xmax[2] = m->d.x3max;
  for (int d = 0; d < 3; ++d) k.c[d] = m->srcmask_c[d];
  k.r2 = m->srcmask_r2;
  return k;
}

static MPGeom mp_geom(const agx_mesh *m, const MGLevelB &B) {
  MPGeom G;
  G.n1 = B.n1; G.n2 = B.n2; G.n3 = B.n3;
  G.xmin[0] = m->d.x1min; G.xmin[1] = m->d.x2min; G.xmin[2] = m->d.x3min;
  G.xmax[0] = m->d.x1max; G.xmax[1] = m->d.x2max; G.xmax[2] = m->d.x3max;
  for (int d = 0; d < 3; ++d) G.h[d] = B.h[d];
  return G;
}

// ProlongateAndCorrect with FAS from the framed coarse level: the paired form where the
// fine rows have even length
template <typename T>
static void mb_prolong(agx_mesh *m, const T *cu, const T *cuold, T *fu, int n1, int n2, int n3) {
  if ((n1 & 1) == 0)
    hipLaunchKernelGGL(k_mb_prolong_pairs<T>, dim3(grid1d((int64_t)(n1 / 2) * n2 * n3, BS)), dim3(BS), 0,
                       m->stream, m->gate, cu, cuold, fu, n1, n2, n3);
  else
    hipLaunchKernelGGL(k_mb_prolong<T>, dim3(grid1d((int64_t)n1 * n2 * n3, BS)), dim3(BS), 0, m->stream,
                       m->gate, cu, cuold, fu, n1, n2, n3);
}

static void bc_b(agx_mesh *m, int l) {
  MGLevelB &B = m->levb[l];
  timer_begin(m, AGX_K_MG_OTHER);
  hipLaunchKernelGGL(k_mb_bc, dim3(grid1d(bghosts(B), BS)), dim3(BS), 0, m->stream, m->gate, B.u, mp_geom(m, B), B.bcv);
  timer_end(m);
}
static void smooth_b(agx_mesh *m, int l, int color) {
  MGLevelB &B = m->levb[l];
  const int64_t n = (int64_t)((B.n1 + 1) / 2) * B.n2 * B.n3;
  timer_begin(m, AGX_K_MG_SMOOTH);
  hipLaunchKernelGGL(k_mb_smooth<double>, dim3(grid1d(n, BS)), dim3(BS), 0, m->stream, m->gate,
                     B.u, B.src, B.n1, B.n2, B.n3,
                     B.dx * B.dx, 1.15 / 6.0, color ^ m->coffset);
  timer_end(m);
}
static void fas_b(agx_mesh *m, int l) {
  MGLevelB &B = m->levb[l];
  timer_begin(m, AGX_K_MG_OTHER);
  hipLaunchKernelGGL(k_mb_fas<double>, dim3(grid1d(bcells(B), BS)), dim3(BS), 0, m->stream, m->gate,
                     B.u, B.src, B.uold, B.n1,
                     B.n2, B.n3, 1.0 / (B.dx * B.dx));
  timer_end(m);
}

// SolveVCycle(1,1) (multigrid_driver.cpp:684-693) with the FAS steps of
// OneStepToCoarser / SolveCoarsestGrid / OneStepToFiner (:583-677, :778-819) and the
// MeshBlock task lists (mg_task_list.cpp:258-262, 287-510).  Boundaries are filled before
// every stencil step, as the reference's exchanges and ApplyPhysicalBoundaries are.
// Highest level k_mb_coarse takes (at most MB_COARSE_CELLS interior cells, at most
// MB_MAXC levels from 0), or -1 when switched off (agx_set_mg_coarse)
static int coarse_top_b(const agx_mesh *m) {
  if (!m->mg_coarse) return -1;
  int top = -1;
  for (int l = 0; l < m->ntotal && l < MB_MAXC; ++l) {
    const MGLevelB &B = m->levb[l];
    if ((int64_t)B.n1 * B.n2 * B.n3 <= MB_COARSE_CELLS) top = l;
  }
  return top;
}

static void coarse_b(agx_mesh *m, int top, int ni) {
  CoarseB<double> c;
  for (int l = 0; l <= top; ++l) {
    MGLevelB &B = m->levb[l];
    c.u[l] = B.u; c.src[l] = B.src; c.uold[l] = B.uold; c.bcv[l] = B.bcv;
    c.G[l] = mp_geom(m, B);
    c.dx[l] = B.dx;
  }
  timer_begin(m, AGX_K_MG_OTHER);
  hipLaunchKernelGGL((k_mb_coarse<double, false>), dim3(1), dim3(COARSE_BS), 0, m->stream, m->gate, c, top, m->fmglevel,
                     m->coffset, ni);
  timer_end(m);
}

// the same for the FP32 correction cycle (homogeneous boundary chain)
static void coarse_b32(agx_mesh *m, int top, int ni) {
  CoarseB<float> c;
  for (int l = 0; l <= top; ++l) {
    MGLevelB32 &B = m->levb32[l];
    c.u[l] = B.u; c.src[l] = B.src; c.uold[l] = B.uold; c.bcv[l] = nullptr;
    c.G[l] = mp_geom(m, m->levb[l]);
    c.dx[l] = B.dx;
  }
  timer_begin(m, AGX_K_MG_OTHER);
  hipLaunchKernelGGL((k_mb_coarse<float, true>), dim3(1), dim3(COARSE_BS), 0, m->stream, m->gate, c, top, m->fmglevel,
                     m->coffset, ni);
  timer_end(m);
}

static void vcycle_b(agx_mesh *m) {
  int start = m->current;
  m->coffset ^= 1;
  const Geo &g0 = m->g;
  const int ct = coarse_top_b(m);
  if (ct >= 0) {
    // the levels above the coarse kernel's range launch per step; from level
    // min(start, ct) down to 0 and back, one launch
    while (m->current > ct) {
      const int l = m->current;
      MGLevelB &F = m->levb[l], &C = m->levb[l - 1];
      bc_b(m, l);
      if (l < m->fmglevel) fas_b(m, l);
      smooth_b(m, l, 0);
      bc_b(m, l);
      smooth_b(m, l, 1);
      bc_b(m, l);
      timer_begin(m, AGX_K_MG_RESTRICT);
      hipLaunchKernelGGL(k_mb_restrict<double>, dim3(grid1d((int64_t)C.n1 * C.n2 * C.n3, BS)), dim3(BS), 0, m->stream,
                         m->gate, F.u, F.src, C.src, C.u, F.n1, F.n2, F.n3, 1.0 / (F.dx * F.dx));
      timer_end(m);
      m->current--;
    }
    const int ni0 = std::max(g0.nb1, std::max(g0.nb2, g0.nb3)) >> (m->nrootlevel - 1);
    coarse_b(m, m->current, ni0);
    while (m->current < start) {
      const int l = m->current + 1;
      MGLevelB &F = m->levb[l], &C = m->levb[l - 1];
      bc_b(m, l - 1);
      timer_begin(m, AGX_K_MG_PROLONG);
      mb_prolong(m, C.u, C.uold, F.u, F.n1, F.n2, F.n3);
      timer_end(m);
      bc_b(m, l);
      smooth_b(m, l, 0);
      bc_b(m, l);
      smooth_b(m, l, 1);
      bc_b(m, l);
      m->current++;
    }
    return;
  }
  while (m->current > 0) {
    const int l = m->current;
    MGLevelB &F = m->levb[l], &C = m->levb[l - 1];
    bc_b(m, l);
    if (l < m->fmglevel) fas_b(m, l);
    smooth_b(m, l, 0);
    bc_b(m, l);
    smooth_b(m, l, 1);
    bc_b(m, l);
    timer_begin(m, AGX_K_MG_RESTRICT);
    hipLaunchKernelGGL(k_mb_restrict<double>, dim3(grid1d((int64_t)C.n1 * C.n2 * C.n3, BS)), dim3(BS), 0, m->stream,
                       m->gate, F.u, F.src, C.src, C.u, F.n1, F.n2, F.n3, 1.0 / (F.dx * F.dx));
    timer_end(m);
    m->current--;
  }
  // SolveCoarsestGrid with FAS (no average subtraction): ni sweeps on the root level
  const Geo &g = m->g;
  int ni = std::max(g.nb1, std::max(g.nb2, g.nb3)) >> (m->nrootlevel - 1);
  bc_b(m, 0);
  fas_b(m, 0);
  for (int it = 0; it < ni; ++it) {
    smooth_b(m, 0, 0);
    bc_b(m, 0);
    smooth_b(m, 0, 1);
    bc_b(m, 0);
  }
  while (m->current < start) {
    const int l = m->current + 1;
    MGLevelB &F = m->levb[l], &C = m->levb[l - 1];
    bc_b(m, l - 1);
    timer_begin(m, AGX_K_MG_PROLONG);
    mb_prolong(m, C.u, C.uold, F.u, F.n1, F.n2, F.n3);
    timer_end(m);
    bc_b(m, l);
    smooth_b(m, l, 0);
    bc_b(m, l);
    smooth_b(m, l, 1);
    bc_b(m, l);
    m->current++;
  }
}

// SolveFMGCycle (multigrid_driver.cpp:700-714) after RestrictFMGSource
static void fmg_b(agx_mesh *m) {
  const int L = m->ntotal;
  for (int l = L - 1; l > 0; --l) {
    MGLevelB &F = m->levb[l], &C = m->levb[l - 1];
    hipLaunchKernelGGL(k_mb_restrict_src, dim3(grid1d((int64_t)C.n1 * C.n2 * C.n3, BS)), dim3(BS), 0, m->stream,
                       F.src, C.src, F.n1, F.n2, F.n3);
  }
  m->current = 0;
  for (m->fmglevel = 0; m->fmglevel < L; m->fmglevel++) {
    vcycle_b(m);
    if (m->fmglevel != L - 1) {   // FMGProlongate, multigrid_driver.cpp:553-575
      const int l = m->current + 1;
      MGLevelB &F = m->levb[l], &C = m->levb[l - 1];
      bc_b(m, l - 1);
      timer_begin(m, AGX_K_MG_PROLONG);
      hipLaunchKernelGGL(k_mb_fmg_prolong, dim3(grid1d((int64_t)F.n1 * F.n2 * F.n3, BS)), dim3(BS), 0, m->stream,
                         C.u, F.u, F.n1, F.n2, F.n3);
      timer_end(m);
      m->current++;
    }
  }
  m->fmglevel = L - 1;
}

// ---- mixed precision with multipole boundaries: FP32 V(1,1) cycles with FAS on the
// correction (ghosts by the homogeneous chain), inside the FP64 defect correction
static void bc_b32(agx_mesh *m, int l) {
  timer_begin(m, AGX_K_MG_OTHER);
  hipLaunchKernelGGL(k_mb_bc_hom, dim3(grid1d(bghosts(m->levb[l]), BS)), dim3(BS), 0, m->stream, m->gate,
                     m->levb32[l].u,
                     mp_geom(m, m->levb[l]));
  timer_end(m);
}
static void smooth_b32(agx_mesh *m, int l, int color) {
  MGLevelB32 &B = m->levb32[l];
  const int64_t n = (int64_t)((B.n1 + 1) / 2) * B.n2 * B.n3;
  timer_begin(m, AGX_K_MG_SMOOTH);
  hipLaunchKernelGGL(k_mb_smooth<float>, dim3(grid1d(n, BS)), dim3(BS), 0, m->stream, m->gate, B.u, B.src, B.n1, B.n2,
                     B.n3, float(B.dx * B.dx), float(1.15 / 6.0), color ^ m->coffset);
  timer_end(m);
}
static void fas_b32(agx_mesh *m, int l) {
  MGLevelB32 &B = m->levb32[l];
  timer_begin(m, AGX_K_MG_OTHER);
  hipLaunchKernelGGL(k_mb_fas<float>, dim3(grid1d(bcells(m->levb[l]), BS)), dim3(BS), 0, m->stream, m->gate, B.u, B.src,
                     B.uold, B.n1, B.n2, B.n3, float(1.0 / (B.dx * B.dx)));
  timer_end(m);
}
static void vcycle_b32(agx_mesh *m) {   // the per-level sequence of vcycle_b
  const int start = m->current;
  m->coffset ^= 1;
  const int ct = coarse_top_b(m);   // levels at or below it in one launch (coarse_b32)
  while (m->current > 0 && m->current > ct) {
    const int l = m->current;
    MGLevelB32 &F = m->levb32[l], &C = m->levb32[l - 1];
    bc_b32(m, l);
    if (l < m->fmglevel) fas_b32(m, l);
    smooth_b32(m, l, 0);
    bc_b32(m, l);
    smooth_b32(m, l, 1);
    bc_b32(m, l);
    timer_begin(m, AGX_K_MG_RESTRICT);
    hipLaunchKernelGGL(k_mb_restrict<float>, dim3(grid1d((int64_t)C.n1 * C.n2 * C.n3, BS)), dim3(BS), 0,
                       m->stream, m->gate, F.u, F.src, C.src, C.u, F.n1, F.n2, F.n3, float(1.0 / (F.dx * F.dx)));
    timer_end(m);
    m->current--;
  }
  const Geo &g = m->g;
  const int ni = std::max(g.nb1, std::max(g.nb2, g.nb3)) >> (m->nrootlevel - 1);
  if (ct >= 0) {
    coarse_b32(m, m->current, ni);
  } else {
    bc_b32(m, 0);
    fas_b32(m, 0);
    for (int it = 0; it < ni; ++it) {
      smooth_b32(m, 0, 0);
      bc_b32(m, 0);
      smooth_b32(m, 0, 1);
      bc_b32(m, 0);
    }
  }
  while (m->current < start) {
    const int l = m->current + 1;
    MGLevelB32 &F = m->levb32[l], &C = m->levb32[l - 1];
    bc_b32(m, l - 1);
    timer_begin(m, AGX_K_MG_PROLONG);
    mb_prolong(m, C.u, C.uold, F.u, F.n1, F.n2, F.n3);
    timer_end(m);
    bc_b32(m, l);
    smooth_b32(m, l, 0);
    bc_b32(m, l);
    smooth_b32(m, l, 1);
    bc_b32(m, l);
    m->current++;
  }
}

// One defect-correction step: ghosts of phi from the multipole expansion, d = defect in
// FP64, FP32 V(1,1) on the correction from zero, phi += e, ghosts again (the defect norm
// and the next step read them).
static void correction_step_b(agx_mesh *m) {
  const int L = m->ntotal;
  MGLevelB &F = m->levb[L - 1];
  MGLevelB32 &E = m->levb32[L - 1];
  bc_b(m, L - 1);
  timer_begin(m, AGX_K_MG_OTHER);
  hipLaunchKernelGGL(k_mb_defect32, dim3(grid1d(bcells(F), BS)), dim3(BS), 0, m->stream, m->gate, F.u, F.src, E.src,
                     E.u, F.n1, F.n2, F.n3, 1.0 / (F.dx * F.dx));
  timer_end(m);
  const int keep = m->fmglevel;
  m->current = L - 1;
  m->fmglevel = L - 1;
  vcycle_b32(m);
  m->fmglevel = keep;
  timer_begin(m, AGX_K_MG_OTHER);
  hipLaunchKernelGGL(k_mb_add32, dim3(grid1d(bcells(F), BS)), dim3(BS), 0, m->stream, m->gate, F.u, E.u, F.n1, F.n2,
                     F.n3);
  timer_end(m);
  bc_b(m, L - 1);
}

// squared sum of the framed finest level's defect into scal[0] (reads its ghost frame as
// it stands), and the defect norm from it
static void defect_sum_b(agx_mesh *m) {
  MGLevelB &F = m->levb[m->ntotal - 1];
  const int64_t nb = npart((int64_t)F.n1 * F.n2 * F.n3);
  hipLaunchKernelGGL(k_mb_defect_partial, dim3((unsigned)nb), dim3(BS), 0, m->stream, m->gate, F.u, F.src, F.n1,
                     F.n2, F.n3, 1.0 / (F.dx * F.dx), m->red);
  hipLaunchKernelGGL(k_mg_final, dim3(1), dim3(BS), 0, m->stream, m->gate, m->red, nb, m->scal);
}
static double defect_norm_b(agx_mesh *m) {
  defect_sum_b(m);
  return finish_norm(m);
}

// origin of the expansion given in the input (<gravity>/mporigin_x1..x3, mg_gravity.cpp:116-118)
__global__ void k_mb_set_origin(double *mp, double x1, double x2, double x3) {
  if (threadIdx.x) return;
  mp[0] = x1;
  mp[1] = x2;
  mp[2] = x3;
}

// MultipoleCoefficients of the finest source (SetupMultigrid, multigrid_driver.cpp:441-445)
static void mp_coefficients(agx_mesh *m) {
  MGLevelB &F = m->levb[m->ntotal - 1];
  const int64_t nb = npart((int64_t)F.n1 * F.n2 * F.n3);
  const MPGeom G = mp_geom(m, F);
  double *part = m->mpdev + 28;
  if (m->d.mp_auto_origin) {
    hipLaunchKernelGGL(k_mb_moments, dim3((unsigned)nb), dim3(BS), 0, m->stream, F.src, G, m->mpdev,
                       m->d.mporder, 1, part);
    hipLaunchKernelGGL(k_mb_moments_final, dim3(1), dim3(BS), 0, m->stream, part, (int)nb, m->d.mporder, 1, 0,
                       m->mpdev);
  } else {   // the fixed origin is part of the mesh description: written by a kernel, so
             // that a recorded solve holds it
    hipLaunchKernelGGL(k_mb_set_origin, dim3(1), dim3(64), 0, m->stream, m->mpdev, m->d.mp_origin[0],
                       m->d.mp_origin[1], m->d.mp_origin[2]);
  }
  hipLaunchKernelGGL(k_mb_moments, dim3((unsigned)nb), dim3(BS), 0, m->stream, F.src, G, m->mpdev, m->d.mporder,
                     0, part);
  hipLaunchKernelGGL(k_mb_moments_final, dim3(1), dim3(BS), 0, m->stream, part, (int)nb, m->d.mporder, 0,
                     m->d.mp_nodipole ? 1 : 0, m->mpdev);
  // the boundary values of every level for this solve (the ghost fills then only gather)
  for (int l = 0; l < m->ntotal; ++l) {
    MGLevelB &B = m->levb[l];
    hipLaunchKernelGGL(k_mb_bc_fill, dim3(grid1d(bghosts(B), BS)), dim3(BS), 0, m->stream, B.bcv, mp_geom(m, B),
                       m->mpdev, m->d.mporder);
  }
}

// MGGravityDriver::Solve (mg_gravity.cpp:166-211) with multipole boundaries, up to the
// iteration: load, multipole coefficients, FMG
static void solve_load_b(agx_mesh *m) {
  const Geo &g = m->g;
  const int L = m->ntotal;
  MGLevelB &F = m->levb[L - 1];
  double **A = m->arr.a;
  hipLaunchKernelGGL(k_mb_load, dim3(grid1d((int64_t)F.n1 * F.n2 * F.n3, BS)), dim3(BS), 0, m->stream, g,
                     A[AGX_ARR_U], A[AGX_ARR_PHI], F.src, F.u, m->d.four_pi_G, m->mg_mode == 1 ? 1 : 0, src_mask(m));
  m->current = L - 1;
  m->fmglevel = L - 1;
  mp_coefficients(m);
  m->mp_solved = true;
  if (m->mg_mode == 0) fmg_b(m);
  m->current = L - 1;
}

static int retrieve_b(agx_mesh *m) {
  const Geo &g = m->g;
  MGLevelB &F = m->levb[m->ntotal - 1];
  const int64_t nret = (int64_t)(g.bx1 + 2) * (g.bx2 + 2) * (g.bx3 + 2) * g.nb;
  hipLaunchKernelGGL(k_mb_retrieve, dim3(grid1d(nret, BS)), dim3(BS), 0, m->stream, g, F.u, m->arr.a[AGX_ARR_PHI]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "multipole mg_solve launches");
  return 0;
}

static int mg_solve_b(agx_mesh *m, int stage) {
  solve_load_b(m);
  int iters = 0;
  if (m->mg_eps >= 0.0) {   // SolveIterative, multigrid_driver.cpp:721-755
    first_defect_sum(m);
    double def = finish_norm(m);
    int n = 0;
    while (def > m->mg_eps) {
      loop_step(m);   // correction_step_b (mixed) or vcycle_b, then defect_sum_b
      iters++;
      double olddef = def;
      def = finish_norm(m);
      if (def / olddef > 0.8) {
        if (m->mg_eps == 0.0) break;
      }
      if (n > 100) break;
      n++;
    }
    m->last_norm = def;
    if (stage == 1 || stage == 2) m->mg_kpred[stage - 1] = iters;
  } else {   // SolveIterativeFixedTimes, multigrid_driver.cpp:762-771
    for (int n = 0; n < m->mg_niter; ++n) {
      if (m->mixed) correction_step_b(m);
      else vcycle_b(m);
      iters++;
    }
    m->last_norm = m->capturing ? NAN : defect_norm_b(m);   // diagnostic only, see mg_solve
  }
  m->last_iters = iters;
  return retrieve_b(m);
}

int mg_create(agx_mesh *m) {
  const agx_mesh_desc &d = m->d;
  const Geo &g = m->g;
  if (d.bnx1 != d.bnx2 || d.bnx1 != d.bnx3) {
    set_error("multigrid requires logically cubic MeshBlocks (multigrid.cpp:37-43)");
    return -20;
  }
  int nl = 0;
  for (int l = 0; l < 20; ++l)
    if ((1 << l) == d.bnx1) { nl = l + 1; break; }
  if (nl == 0) { set_error("MeshBlock size must be a power of two (multigrid.cpp:92-104)"); return -21; }
  int nrl = 0;
  for (int l = 0; l < 20; ++l)
    if (g.nb1 % (1 << l) == 0 && g.nb2 % (1 << l) == 0 && g.nb3 % (1 << l) == 0) nrl = l + 1;
  m->nmblevel = nl;
  m->nrootlevel = nrl;
  m->ntotal = nrl + nl - 1;
  m->mg_mode = d.mg_mode;
  m->mg_eps = d.mg_threshold;
  m->mg_niter = d.mg_niter;
  // cell sizes: block levels use the MeshBlock's rdx_/rdy_/rdz_, root levels the root
  // grid's (multigrid.cpp:61-63); taken from MeshBlock gid 0, whose upper corner is
  // UniformMeshGenerator(ComputeMeshGeneratorX(1, nb)) (mesh.cpp:1668-1760).
  auto blk0 = [](int nb, double xmin, double xmax) {
    if (nb == 1) return xmax;
    int64_t no = 1 - nb / 2, noc = 1 - (nb + 1) / 2;
    double rx = (double)(no + noc) / (2.0 * nb);
    return 0.5 * (xmin + xmax) + (rx * xmax - rx * xmin);
  };
  m->rdx_blk = (blk0(g.nb1, d.x1min, d.x1max) - d.x1min) / (double)d.bnx1;
  m->rdy_blk = (blk0(g.nb2, d.x2min, d.x2max) - d.x2min) / (double)d.bnx2;
  m->rdz_blk = (blk0(g.nb3, d.x3min, d.x3max) - d.x3min) / (double)d.bnx3;
  m->rdx_root = (d.x1max - d.x1min) / (double)g.nb1;
  m->rdy_root = (d.x2max - d.x2min) / (double)g.nb2;
  m->rdz_root = (d.x3max - d.x3min) / (double)g.nb3;
  int L = m->ntotal;
  m->lev.resize(L);
  int64_t tot = 0;
  for (int l = 0; l < L; ++l) {
    MGLevel &Lv = m->lev[l];
    int sh = L - 1 - l;
    Lv.n1 = d.nx1 >> sh;
    Lv.n2 = d.nx2 >> sh;
    Lv.n3 = d.nx3 >> sh;
    if (l >= nrl) Lv.dx = m->rdx_blk * (double)(1 << sh);
    else Lv.dx = m->rdx_root * (double)(1 << (nrl - 1 - l));
    tot += 3 * pad32(lcells(Lv));
  }
  m->mg_bytes = tot * (int64_t)sizeof(double);
  AGX_HIP(hipMalloc(&m->mgpool, m->mg_bytes));
  AGX_HIP(hipMemsetAsync(m->mgpool, 0, m->mg_bytes, m->stream));
  double *p = m->mgpool;
  for (int l = 0; l < L; ++l) {
    MGLevel &Lv = m->lev[l];
    int64_t n = pad32(lcells(Lv));   // every array 256-B aligned (vector loads of pairs)
    Lv.u = p; p += n;
    Lv.src = p; p += n;
    Lv.def = p; p += n;
  }
  if (m->mixed) {
    // FP32 hierarchy of the correction V-cycles: u (correction) and src per level
    m->lev32.resize(L);
    int64_t t32 = 0;
    for (int l = 0; l < L; ++l) t32 += 3 * pad32(lcells(m->lev[l]));
    AGX_HIP(hipMalloc(&m->mgpool32, t32 * (int64_t)sizeof(float)));
    AGX_HIP(hipMemsetAsync(m->mgpool32, 0, t32 * sizeof(float), m->stream));
    m->mg_bytes += t32 * (int64_t)sizeof(float);
    float *q = m->mgpool32;
    for (int l = 0; l < L; ++l) {
      MGLevel32 &Lv = m->lev32[l];
      Lv.n1 = m->lev[l].n1; Lv.n2 = m->lev[l].n2; Lv.n3 = m->lev[l].n3; Lv.dx = m->lev[l].dx;
      int64_t n = pad32(lcells(Lv));
      Lv.u = q; q += n;
      Lv.src = q; q += n;
      Lv.def = q; q += n;   // spare array of the fused red-black pass
    }
  }
  int64_t need = (lcells(m->lev[L - 1]) + BS - 1) / BS + 1;
  if (need > m->red_n) {
    if (m->red) AGX_HIP(hipFree(m->red));
    m->red_n = need;
    AGX_HIP(hipMalloc(&m->red, sizeof(double) * m->red_n));
  }
  if (d.mg_bc == 1) {
    int rc = mp_create(m);
    if (rc) return rc;
    if (m->mixed) {   // FP32 framed levels: correction, source, FAS copy
      m->levb32.resize(L);
      int64_t t32 = 0;
      for (int l = 0; l < L; ++l) t32 += 3 * pad32(bcells(m->levb[l]));
      AGX_HIP(hipMalloc(&m->mgpoolb32, t32 * (int64_t)sizeof(float)));
      AGX_HIP(hipMemsetAsync(m->mgpoolb32, 0, t32 * sizeof(float), m->stream));
      m->mg_bytes += t32 * (int64_t)sizeof(float);
      float *q = m->mgpoolb32;
      for (int l = 0; l < L; ++l) {
        MGLevelB32 &B = m->levb32[l];
        B.n1 = m->levb[l].n1; B.n2 = m->levb[l].n2; B.n3 = m->levb[l].n3; B.dx = m->levb[l].dx;
        const int64_t n = pad32(bcells(m->levb[l]));
        B.u = q; q += n;
        B.src = q; q += n;
        B.uold = q; q += n;
      }
    }
  }
  if (!m->gatedev) AGX_HIP(hipMalloc(&m->gatedev, sizeof(MGGate)));
  if (!m->gatehost) AGX_HIP(hipHostMalloc(&m->gatehost, sizeof(MGGate)));
  AGX_HIP(hipStreamSynchronize(m->stream));
  m->mg_on = 1;
  return 0;
}

void mg_destroy(agx_mesh *m) {
  if (m->mgpoolb) hipFree(m->mgpoolb);
  if (m->mpdev) hipFree(m->mpdev);
  if (m->mgpoolb32) hipFree(m->mgpoolb32);
  m->mgpoolb = nullptr;
  m->mgpoolb32 = nullptr;
  m->mpdev = nullptr;
  m->levb.clear();
  m->levb32.clear();
  m->mp_on = false;
  if (m->mgpool) hipFree(m->mgpool);
  if (m->mgpool32) hipFree(m->mgpool32);
  m->mgpool = nullptr;
  m->mgpool32 = nullptr;
  m->lev.clear();
  m->lev32.clear();
  if (m->gatedev) hipFree(m->gatedev);
  if (m->gatehost) hipHostFree(m->gatehost);
  m->gatedev = nullptr;
  m->gatehost = nullptr;
  m->mg_on = 0;
}

// ---------------------------------------------------------------- level operations
// Each works on one hierarchy H: the FP64 levels, or the FP32 correction levels of the
// mixed-precision form.
template <typename T>
static void smooth(agx_mesh *m, std::vector<MGLevelT<T>> &H, int l, int color) {
  MGLevelT<T> &L = H[l];
  double dx2 = L.dx * L.dx;
  double isix = 1.15 / 6.0;
  int64_t n = (int64_t)((L.n1 + 1) / 2) * L.n2 * L.n3;
  timer_begin(m, AGX_K_MG_SMOOTH);
  if ((L.n1 & 1) == 0)
    hipLaunchKernelGGL(k_mg_smooth_pairs<T>, dim3(grid1d(n, BS)), dim3(BS), 0, m->stream, m->gate, L.u, L.src, L.n1,
                       L.n2, L.n3, T(dx2), T(isix), color ^ m->coffset);
  else
    hipLaunchKernelGGL(k_mg_smooth<T>, dim3(grid1d(n, BS)), dim3(BS), 0, m->stream, m->gate, L.u, L.src, L.n1, L.n2,
                       L.n3, T(dx2), T(isix), color ^ m->coffset);
  timer_end(m);
}

// One red-black pass (colour 0 then colour 1) on level l: fused into one launch into the
// level's spare array (then swapped with u), or two sweeps.  By default the fused form
// runs on FP64 levels of at least 2^26 cells.  Measured on MI355X (V(1,1) cycles/s, two
// sweeps -> fused): 512^3 FP64 186 -> 231; at 256^3 (134 MB per level array, which the
// 256 MB last-level cache holds between the two sweeps) two sweeps are faster (1746 vs
// 1600), as they are for FP32 levels at 512^3 (231 vs 213).  agx_autotune times both
// (mg_fuse 0 or 1: fused on every level of at least 2^21 cells) and keeps the faster;
// the results are identical either way.
template <typename T>
static void smooth_pass(agx_mesh *m, std::vector<MGLevelT<T>> &H, int l) {
  MGLevelT<T> &L = H[l];
  const bool big = m->mg_fuse < 0 ? (sizeof(T) == 8 && lcells(L) >= ((int64_t)1 << 26))
                                  : (m->mg_fuse == 1 && lcells(L) >= ((int64_t)1 << 21));
  const bool fused = big && L.def && L.n1 % RB_TI == 0 && L.n2 % RB_TJ == 0 && (L.n3 & 1) == 0 && L.n3 >= 4;
  if (!fused) {
    smooth(m, H, l, 0);
    smooth(m, H, l, 1);
    return;
  }
  const int64_t nwg = (int64_t)(L.n1 / RB_TI) * (L.n2 / RB_TJ) * ((L.n3 + RB_KC - 1) / RB_KC);
  timer_begin(m, AGX_K_MG_SMOOTH);
  hipLaunchKernelGGL(k_mg_smooth_rb<T>, dim3((unsigned)nwg), dim3(RB_TI * RB_TJ), 0, m->stream, m->gate, L.u, L.def,
                     L.src, L.n1, L.n2, L.n3, T(L.dx * L.dx), T(1.15 / 6.0), m->coffset);
  timer_end(m);
  std::swap(L.u, L.def);
}

template <typename T>
static void defect_restrict(agx_mesh *m, std::vector<MGLevelT<T>> &H, int l) {
  MGLevelT<T> &F = H[l], &C = H[l - 1];
  double idx2 = 1.0 / (F.dx * F.dx);
  timer_begin(m, AGX_K_MG_RESTRICT);
  hipLaunchKernelGGL(k_mg_defect_restrict<T>, dim3(grid1d(lcells(C), BS)), dim3(BS), 0, m->stream, m->gate, F.u,
                     F.src, C.src, C.u, F.n1, F.n2, F.n3, T(idx2));
  timer_end(m);
}

template <typename T>
static void prolong(agx_mesh *m, std::vector<MGLevelT<T>> &H, int l) {   // l-1 -> l
  MGLevelT<T> &F = H[l], &C = H[l - 1];
  timer_begin(m, AGX_K_MG_PROLONG);
  if ((F.n1 & 1) == 0)   // every level above the root: n1 = 2 * (coarse n1)
    hipLaunchKernelGGL(k_mg_prolong_pairs<T>, dim3(grid1d(lcells(F) / 2, BS)), dim3(BS), 0, m->stream, m->gate,
                       C.u, F.u,
                       F.n1, F.n2, F.n3);
  else
    hipLaunchKernelGGL(k_mg_prolong<T>, dim3(grid1d(lcells(F), BS)), dim3(BS), 0, m->stream, m->gate, C.u, F.u, F.n1,
                       F.n2, F.n3);
  timer_end(m);
}

template <typename T>
static void fmg_prolong(agx_mesh *m, std::vector<MGLevelT<T>> &H, int l) {   // l-1 -> l
  MGLevelT<T> &F = H[l], &C = H[l - 1];
  timer_begin(m, AGX_K_MG_PROLONG);
  hipLaunchKernelGGL(k_mg_fmg_prolong<T>, dim3(grid1d(lcells(F), BS)), dim3(BS), 0, m->stream, C.u, F.u,
                     F.n1, F.n2, F.n3);
  timer_end(m);
}

// device sum of a level array (mode 0) or of its squared defect (mode 1) into scal[slot]
template <typename T>
static void level_sum(agx_mesh *m, std::vector<MGLevelT<T>> &H, int l, int mode, int slot) {
  MGLevelT<T> &L = H[l];
  int64_t nb = npart(lcells(L));
  double idx2 = 1.0 / (L.dx * L.dx);
  timer_begin(m, AGX_K_MG_OTHER);
  if (mode == 0)
    hipLaunchKernelGGL((k_mg_partial<0, T>), dim3((unsigned)nb), dim3(BS), 0, m->stream, m->gate, L.u, L.src, L.n1,
                       L.n2, L.n3, T(idx2), m->red);
  else
    hipLaunchKernelGGL((k_mg_partial<1, T>), dim3((unsigned)nb), dim3(BS), 0, m->stream, m->gate, L.u, L.src, L.n1,
                       L.n2, L.n3, T(idx2), m->red);
  hipLaunchKernelGGL(k_mg_final, dim3(1), dim3(BS), 0, m->stream, m->gate, m->red, nb, m->scal + slot);
  timer_end(m);
}

static void level_sum_src(agx_mesh *m, int l, int slot) {
  MGLevel &L = m->lev[l];
  int64_t nb = npart(lcells(L));
  hipLaunchKernelGGL((k_mg_partial<0, double>), dim3((unsigned)nb), dim3(BS), 0, m->stream, m->gate, L.src, L.src,
                     L.n1, L.n2, L.n3, 0.0, m->red);
  hipLaunchKernelGGL(k_mg_final, dim3(1), dim3(BS), 0, m->stream, m->gate, m->red, nb, m->scal + slot);
}

static double mesh_volume(const agx_mesh_desc &d) {
  return (d.x1max - d.x1min) * (d.x2max - d.x2min) * (d.x3max - d.x3min);
}

// MultigridDriver::SubtractAverage (multigrid_driver.cpp:305-330) on the finest FP64
// level: ave = sum * dx*dy*dz / volume, subtracted on the device without a host round
// trip.
static void subtract_average(agx_mesh *m, int type) {
  int l = m->ntotal - 1;
  MGLevel &L = m->lev[l];
  const agx_mesh_desc &d = m->d;
  double dv = m->rdx_blk * m->rdy_blk * m->rdz_blk;
  if (type == 0) level_sum_src(m, l, 1);
  else level_sum(m, m->lev, l, 0, 1);
  double scale = dv / mesh_volume(d);
  hipLaunchKernelGGL(k_mg_add<double>, dim3(grid1d(lcells(L), BS)), dim3(BS), 0, m->stream, m->gate,
                     type ? L.u : L.src, lcells(L), m->scal + 1, scale);
}

template <typename T>
static void zero_level(agx_mesh *m, std::vector<MGLevelT<T>> &H, int l) {
  MGLevelT<T> &L = H[l];
  hipLaunchKernelGGL(k_mg_zero<T>, dim3(grid1d(lcells(L), BS)), dim3(BS), 0, m->stream, m->gate, L.u, lcells(L));
}

// SolveCoarsestGrid, multigrid_driver.cpp:778-819 (periodic: fsubtract_average_)
template <typename T>
static void solve_coarsest(agx_mesh *m, std::vector<MGLevelT<T>> &H) {
  const Geo &g = m->g;
  int ni = g.nb1 > g.nb2 ? g.nb1 : g.nb2;
  ni = ni > g.nb3 ? ni : g.nb3;
  ni >>= (m->nrootlevel - 1);
  if (ni == 1) {
    zero_level(m, H, 0);
    return;
  }
  // the root grid's average of level 0 is subtracted from the root grid's finest level
  // (Multigrid::CalculateTotal at the current level, SubtractAverage on u_[nlevel_-1],
  // multigrid.cpp:602-641)
  MGLevelT<T> &LR = H[m->nrootlevel - 1];
  int ll = m->nrootlevel - 1;
  double dxl = m->rdx_root * (double)(1 << ll), dyl = m->rdy_root * (double)(1 << ll),
         dzl = m->rdz_root * (double)(1 << ll);
  double scale = dxl * dyl * dzl / mesh_volume(m->d);
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      for (int it = 0; it < ni; ++it) { smooth(m, H, 0, 0); smooth(m, H, 0, 1); }
    }
    level_sum(m, H, 0, 0, 2);
    hipLaunchKernelGGL(k_mg_add<T>, dim3(grid1d(lcells(LR), BS)), dim3(BS), 0, m->stream, m->gate, LR.u, lcells(LR),
                       m->scal + 2, scale);
  }
}

// SolveVCycle(1,1), multigrid_driver.cpp:684-693 with OneStepToCoarser (:629-677) and
// OneStepToFiner (:583-622).  On the uniform periodic hierarchy every level step is:
// coarser = red sweep, black sweep, defect+restrict; finer = prolongate+correct, red,
// black.
// Highest level handled by k_mg_coarse, or -1 if the root grid is not a single cell.
template <typename T>
static int coarse_top(agx_mesh *m, const std::vector<MGLevelT<T>> &H) {
  const MGLevelT<T> &L0 = H[0];
  if (L0.n1 != 1 || L0.n2 != 1 || L0.n3 != 1 || (int)H.size() > MAXLEV) return -1;
  int top = 0;
  while (top + 1 < (int)H.size() && lcells(H[top + 1]) <= COARSE_CELLS) ++top;
  return top;
}

template <typename T>
static void coarse_cycle(agx_mesh *m, std::vector<MGLevelT<T>> &H, int top, bool first_sweep_done) {
  CoarseLevels<T> c;
  for (int l = 0; l <= top; ++l) {
    c.u[l] = H[l].u;
    c.src[l] = H[l].src;
    c.n1[l] = H[l].n1; c.n2[l] = H[l].n2; c.n3[l] = H[l].n3;
    c.dx2[l] = T(H[l].dx * H[l].dx);
    c.idx2[l] = T(1.0 / (H[l].dx * H[l].dx));
  }
  timer_begin(m, AGX_K_MG_SMOOTH);
  hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(COARSE_BS), 0, m->stream, m->gate, c, top, m->coffset,
                     T(1.15 / 6.0), first_sweep_done ? 1 : 0);
  timer_end(m);
}

template <typename T>
static void vcycle(agx_mesh *m, std::vector<MGLevelT<T>> &H, bool first_sweep_done = false) {
  int start = m->current;
  m->coffset ^= 1;
  const int ct = coarse_top(m, H);
  while (m->current > 0 && m->current > ct) {
    int l = m->current;
    if (first_sweep_done && l == start) smooth(m, H, l, 1);
    else smooth_pass(m, H, l);
    defect_restrict(m, H, l);
    m->current--;
  }
  if (ct >= 0) coarse_cycle(m, H, m->current, first_sweep_done && m->current == start);
  else solve_coarsest(m, H);
  while (m->current < start) {
    int l = m->current + 1;
    prolong(m, H, l);
    smooth_pass(m, H, l);
    m->current++;
  }
}

// SolveFMGCycle, multigrid_driver.cpp:700-714, on the hierarchy H whose finest source
// is set; sources are restricted first (RestrictFMGSource, multigrid_driver.cpp:447-456).
template <typename T>
static void fmg(agx_mesh *m, std::vector<MGLevelT<T>> &H) {
  int L = m->ntotal;
  for (int l = L - 1; l > 0; --l) {
    MGLevelT<T> &Fl = H[l], &Cl = H[l - 1];
    hipLaunchKernelGGL(k_mg_restrict_src<T>, dim3(grid1d(lcells(Cl), BS)), dim3(BS), 0, m->stream, Fl.src,
                       Cl.src, Fl.n1, Fl.n2, Fl.n3);
  }
  m->current = 0;
  for (m->fmglevel = 0; m->fmglevel < L; m->fmglevel++) {
    vcycle(m, H);
    if (m->fmglevel != L - 1) {
      fmg_prolong(m, H, m->current + 1);
      m->current++;
    }
  }
  m->fmglevel = L - 1;
}

static double finish_norm(agx_mesh *m) {
  hipMemcpyAsync(m->scal_host, m->scal, sizeof(double), hipMemcpyDeviceToHost, m->stream);
  hipStreamSynchronize(m->stream);
  const agx_mesh_desc &d = m->d;
  double dv = m->rdx_blk * m->rdy_blk * m->rdz_blk;
  double norm = m->scal_host[0] * dv / mesh_volume(d);
  return std::sqrt(norm);
}

static double defect_norm(agx_mesh *m) {
  // MultigridDriver::CalculateDefectNorm (L2), multigrid_driver.cpp:826-858
  level_sum(m, m->lev, m->ntotal - 1, 1, 0);
  return finish_norm(m);
}

// Mixed precision: FP64 defect of the finest iterate into the FP32 finest source, its
// squared sum into scal[0], and the first sweep of the next correction V-cycle into the
// FP32 correction (k_mg_defect_to32).  With correct, phi += e of the cycle just run is
// applied in the same pass (into the spare array, then swapped).
static void defect_to32(agx_mesh *m, bool correct) {
  int l = m->ntotal - 1;
  MGLevel &F = m->lev[l];
  MGLevel32 &F32 = m->lev32[l];
  const int cred = 0 ^ (m->coffset ^ 1);   // colour of the next V-cycle's first sweep
  const bool tiled = F.n1 % DT_TI == 0 && F.n2 % DT_TJ == 0;
  const int64_t nb = tiled ? (int64_t)(F.n1 / DT_TI) * (F.n2 / DT_TJ) * ((F.n3 + DT_KC - 1) / DT_KC)
                           : npart(lcells(F));   // <= red_n (cells / BS + 1) either way
  const double idx2 = 1.0 / (F.dx * F.dx);
  const float dx2f = float(F32.dx * F32.dx), isixf = float(1.15 / 6.0);
  timer_begin(m, AGX_K_MG_OTHER);
  // e of the last cycle is read at neighbour cells, so the next cycle's start goes to the
  // spare FP32 array (then swapped), as the corrected phi goes to the FP64 one
  if (tiled && correct)
    hipLaunchKernelGGL(k_mg_defect_to32_t<true>, dim3((unsigned)nb), dim3(DT_TI * DT_TJ), 0, m->stream, m->gate, F.u,
                       F32.u, F.def, F.src, F32.src, F32.def, F.n1, F.n2, F.n3, idx2, dx2f, isixf, cred, m->red);
  else if (tiled)
    hipLaunchKernelGGL(k_mg_defect_to32_t<false>, dim3((unsigned)nb), dim3(DT_TI * DT_TJ), 0, m->stream, m->gate, F.u,
                       (const float *)nullptr, (double *)nullptr, F.src, F32.src, F32.def, F.n1, F.n2, F.n3, idx2,
                       dx2f, isixf, cred, m->red);
  else
    hipLaunchKernelGGL(k_mg_defect_to32, dim3((unsigned)nb), dim3(BS), 0, m->stream, m->gate, F.u,
                       correct ? F32.u : (const float *)nullptr, correct ? F.def : (double *)nullptr, F.src,
                       F32.src, F32.def, F.n1, F.n2, F.n3, idx2, dx2f, isixf, cred, m->red);
  hipLaunchKernelGGL(k_mg_final, dim3(1), dim3(BS), 0, m->stream, m->gate, m->red, nb, m->scal);
  timer_end(m);
  if (correct) std::swap(F.u, F.def);
  std::swap(F32.u, F32.def);
}

// One defect-correction step: FP32 V(1,1) on L(e) = d from e = 0 (its first sweep
// already applied by defect_to32), then phi += e together with the next defect.  The
// FP32 source and correction must come from defect_to32.
static void correction_step(agx_mesh *m) {
  m->current = m->ntotal - 1;
  vcycle(m, m->lev32, true);
  defect_to32(m, true);
}

// MGGravityDriver::Solve (mg_gravity.cpp:166-211) on the periodic hierarchy, up to the
// iteration: load, mean subtraction, FMG
static void solve_load(agx_mesh *m) {
  const Geo &g = m->g;
  int L = m->ntotal;
  MGLevel &F = m->lev[L - 1];
  double **A = m->arr.a;
  hipLaunchKernelGGL(k_mg_load, dim3(grid1d(lcells(F), BS)), dim3(BS), 0, m->stream, g, A[AGX_ARR_U],
                     A[AGX_ARR_PHI], F.src, F.u, m->d.four_pi_G, m->mg_mode == 1 ? 1 : 0, src_mask(m));
  m->current = L - 1;
  m->fmglevel = L - 1;
  subtract_average(m, 0);   // SetupMultigrid, multigrid_driver.cpp:438-439
  if (m->mg_mode == 0) {
    if (!m->mixed) {
      fmg(m, m->lev);
    } else {
      // FMG in FP32 from the FP64 source; the result seeds the FP64 iterate
      MGLevel32 &F32 = m->lev32[L - 1];
      hipLaunchKernelGGL(k_mg_mix, dim3(grid1d(lcells(F), BS)), dim3(BS), 0, m->stream, F.src, F32.src,
                         lcells(F), 2);
      fmg(m, m->lev32);
      hipLaunchKernelGGL(k_mg_mix, dim3(grid1d(lcells(F), BS)), dim3(BS), 0, m->stream, F.u, F32.u,
                         lcells(F), 1);
    }
    m->current = L - 1;
    subtract_average(m, 1);
  }
}

static int retrieve(agx_mesh *m) {
  const Geo &g = m->g;
  int64_t nret = (int64_t)(g.bx1 + 2) * (g.bx2 + 2) * (g.bx3 + 2) * g.nb;
  hipLaunchKernelGGL(k_mg_retrieve, dim3(grid1d(nret, BS)), dim3(BS), 0, m->stream, g, m->lev[m->ntotal - 1].u,
                     m->arr.a[AGX_ARR_PHI]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "mg_solve launches");
  return 0;
}

// SolveIterative (multigrid_driver.cpp:721-755): the finest defect's squared sum into
// scal[0] before the loop (the mixed form also writes the first correction cycle's
// inputs), and one pass of the loop with the new defect's squared sum into scal[0]
static void first_defect_sum(agx_mesh *m) {
  if (m->mp_on) defect_sum_b(m);
  else if (m->mixed) defect_to32(m, false);
  else level_sum(m, m->lev, m->ntotal - 1, 1, 0);
}
static void loop_step(agx_mesh *m) {
  if (m->mp_on) {
    if (m->mixed) correction_step_b(m);
    else vcycle_b(m);
    defect_sum_b(m);
  } else if (m->mixed) {
    correction_step(m);   // the defect comes with the correction
  } else {
    vcycle(m, m->lev);
    level_sum(m, m->lev, m->ntotal - 1, 1, 0);
  }
}

int mg_solve(agx_mesh *m, int stage) {
  if (m->mp_on) return mg_solve_b(m, stage);
  const bool mixed = m->mixed;
  solve_load(m);
  int iters = 0;
  if (m->mg_eps >= 0.0) {   // SolveIterative, multigrid_driver.cpp:721-755
    first_defect_sum(m);
    double def = finish_norm(m);
    int n = 0;
    while (def > m->mg_eps) {
      loop_step(m);
      iters++;
      double olddef = def;
      def = finish_norm(m);
      if (def / olddef > 0.8) {
        if (m->mg_eps == 0.0) break;
      }
      if (n > 100) break;
      n++;
    }
    m->last_norm = def;
    if (stage == 1 || stage == 2) m->mg_kpred[stage - 1] = iters;
    subtract_average(m, 1);
  } else {   // SolveIterativeFixedTimes, multigrid_driver.cpp:762-771
    if (mixed && m->mg_niter > 0) defect_to32(m, false);
    for (int n = 0; n < m->mg_niter; ++n) {
      if (mixed) correction_step(m);
      else vcycle(m, m->lev);
      iters++;
    }
    subtract_average(m, 1);
    // the defect norm is only reported (agx_mg_last); a cycle being recorded into a graph
    // cannot read it back, so it is left out there
    m->last_norm = m->capturing ? NAN : defect_norm(m);
  }
  m->last_iters = iters;
  return retrieve(m);
}

// ---- a threshold solve recorded into a graph
//
// The loop's V-cycle count is only known on the device, so the recorded loop runs k
// passes whose kernels all take the gate: after each pass k_mg_gate_check applies the
// loop's stopping rule to the new norm, and once it closes the gate the remaining passes
// return at once.  The host state a pass changes (red-black offset, swapped arrays) is
// captured after every recorded pass; mg_solve_settle reads how many passes ran and
// restores the state after that many, then runs any further passes eagerly as mg_solve
// would.  Every pass that runs is the same kernel sequence as mg_solve's, so the result
// is the eager one bit for bit.

// finish_norm on the device: sqrt(sum * dv / volume)
__global__ void k_mg_gate_open(MGGate *g, const double *__restrict__ scal, double dv, double vol, double eps) {
  if (threadIdx.x) return;
  const double def = sqrt(scal[0] * dv / vol);
  g->iters = 0;
  g->def = def;
  g->closed = (def > eps) ? 0 : 1;
}

// SolveIterative's tests after a V-cycle (multigrid_driver.cpp:733-753); n is the
// loop's count before its increment
__global__ void k_mg_gate_check(MGGate *g, const double *__restrict__ scal, double dv, double vol, double eps) {
  if (threadIdx.x || g->closed) return;
  const double olddef = g->def;
  const double def = sqrt(scal[0] * dv / vol);
  const int n = g->iters;
  g->iters = n + 1;
  g->def = def;
  if ((def / olddef > 0.8 && eps == 0.0) || n > 100 || !(def > eps)) g->closed = 1;
}

MGSnap mg_snap(const agx_mesh *m) {
  MGSnap s;
  s.lev = m->lev;
  s.lev32 = m->lev32;
  s.coffset = m->coffset;
  s.current = m->current;
  s.fmglevel = m->fmglevel;
  return s;
}

void mg_restore(agx_mesh *m, const MGSnap &s) {
  m->lev = s.lev;
  m->lev32 = s.lev32;
  m->coffset = s.coffset;
  m->current = s.current;
  m->fmglevel = s.fmglevel;
}

std::vector<uintptr_t> mg_key(const agx_mesh *m) {
  std::vector<uintptr_t> k;
  for (const MGLevel &L : m->lev) { k.push_back((uintptr_t)L.u); k.push_back((uintptr_t)L.def); }
  for (const MGLevel32 &L : m->lev32) { k.push_back((uintptr_t)L.u); k.push_back((uintptr_t)L.def); }
  k.push_back((uintptr_t)m->coffset);
  k.push_back((uintptr_t)m->current);
  k.push_back((uintptr_t)m->fmglevel);
  return k;
}

static double norm_dv(const agx_mesh *m) { return m->rdx_blk * m->rdy_blk * m->rdz_blk; }

int mg_solve_head(agx_mesh *m) {
  if (m->mp_on) solve_load_b(m);
  else solve_load(m);
  first_defect_sum(m);
  hipLaunchKernelGGL(k_mg_gate_open, dim3(1), dim3(64), 0, m->stream, m->gatedev, m->scal, norm_dv(m),
                     mesh_volume(m->d), m->mg_eps);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "mg_solve_head launches");
}

void mg_solve_gated(agx_mesh *m, int k, std::vector<MGSnap> &snaps) {
  snaps.clear();
  snaps.push_back(mg_snap(m));
  m->gate = reinterpret_cast<const int *>(reinterpret_cast<const char *>(m->gatedev) + offsetof(MGGate, closed));
  for (int i = 0; i < k; ++i) {
    loop_step(m);
    hipLaunchKernelGGL(k_mg_gate_check, dim3(1), dim3(64), 0, m->stream, m->gatedev, m->scal, norm_dv(m),
                       mesh_volume(m->d), m->mg_eps);
    snaps.push_back(mg_snap(m));
  }
  m->gate = nullptr;
}

int mg_solve_settle(agx_mesh *m, int stage, const std::vector<MGSnap> &snaps) {
  hipError_t e = hipMemcpyAsync(m->gatehost, m->gatedev, sizeof(MGGate), hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) return hip_fail(e, "mg_solve_settle");
  const MGGate gs = *m->gatehost;
  if (gs.iters < 0 || gs.iters >= (int)snaps.size()) {
    set_error("recorded multigrid loop reports %d V-cycles of %d", gs.iters, (int)snaps.size() - 1);
    return -70;
  }
  mg_restore(m, snaps[gs.iters]);
  double def = gs.def;
  int iters = gs.iters;
  if (!gs.closed) {   // the recorded passes ended before the loop did: go on as mg_solve
    int n = iters;
    while (def > m->mg_eps) {
      loop_step(m);
      iters++;
      double olddef = def;
      def = finish_norm(m);
      if (def / olddef > 0.8) {
        if (m->mg_eps == 0.0) break;
      }
      if (n > 100) break;
      n++;
    }
  }
  m->last_norm = def;
  m->last_iters = iters;
  if (stage == 1 || stage == 2) m->mg_kpred[stage - 1] = iters;
  return 0;
}

int mg_solve_tail(agx_mesh *m) {
  if (m->mp_on) return retrieve_b(m);
  subtract_average(m, 1);
  return retrieve(m);
}

// Times V-cycles (on whatever the finest level holds; the next solve reloads it) with
// the two red-black pass forms and keeps the faster (smooth_pass).
int mg_autotune(agx_mesh *m) {
  if (!m->mg_on || m->mp_on) return 0;
  hipEvent_t e0, e1;
  AGX_HIP(hipEventCreate(&e0));
  AGX_HIP(hipEventCreate(&e1));
  const bool timing = m->timer.enabled;
  m->timer.enabled = false;
  float best = 0.f;
  int best_f = 0;
  int rc = 0;
  for (int f = 0; f <= 1 && rc == 0; ++f) {
    m->mg_fuse = f;
    rc = mg_vcycles(m, 1);   // warm-up
    if (rc) break;
    AGX_HIP(hipEventRecord(e0, m->stream));
    rc = mg_vcycles(m, 3);
    if (rc) break;
    AGX_HIP(hipEventRecord(e1, m->stream));
    AGX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    AGX_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (f == 0 || ms < best) { best = ms; best_f = f; }
  }
  m->timer.enabled = timing;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  m->mg_fuse = best_f;
  m->graph_gen++;
  return rc;
}

// n V(1,1) cycles on the current finest state; in the mixed form one cycle is one
// defect-correction step (FP64 defect, FP32 V-cycle, FP64 correction).
int mg_vcycles(agx_mesh *m, int n) {
  m->current = m->ntotal - 1;
  if (m->mp_on) {
    // the FAS levels, the boundary values and the source come from a solve
    if (!m->mp_solved) {
      set_error("agx_mg_vcycles: multipole boundaries need an agx_mg_solve first (it loads the source, "
                "the multipole boundary values and the FAS levels)");
      return -22;
    }
    for (int q = 0; q < n; ++q) {
      if (m->mixed) correction_step_b(m);
      else vcycle_b(m);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "mg_vcycles launches");
  }
  if (m->mixed && n > 0) defect_to32(m, false);
  for (int q = 0; q < n; ++q) {
    if (m->mixed) correction_step(m);
    else vcycle(m, m->lev);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "mg_vcycles launches");
  return 0;
}

// ---------------------------------------------------------------- per-level steps
// The multigrid at the granularity of the reference's MultigridTaskList
// (multigrid/mg_task_list.cpp:183-280, 287-530) and MGGravity / Multigrid methods, on
// the FP64 hierarchy (every level of the whole mesh, finest = nlevels - 1), for callers
// that drive their own cycles.  They use the same kernels as agx_mg_solve, so a V(1,1)
// cycle composed of them equals agx_mg_vcycles(1) bit for bit.
static int mg_level_check(agx_mesh *m, int level, int lo, const char *who) {
  if (!m) { set_error("%s: null mesh", who); return -1; }
  if (!m->mg_on) { set_error("%s: the mesh has no multigrid gravity", who); return -21; }
  if (level < lo || level >= m->ntotal) {
    set_error("%s: level %d not in [%d, %d]", who, level, lo, m->ntotal - 1);
    return -2;
  }
  return 0;
}

static int launches_ok(const char *who) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, who);
}

}  // namespace agx

using namespace agx;

// MGGravityDriver::Solve's set-up: Multigrid::LoadSource (src = 4 pi G rho, multigrid.cpp:
// 194-220), the source mask and MultigridDriver::SetupMultigrid (mean subtraction on
// periodic meshes; multipole coefficients and boundary values, multigrid_driver.cpp:
// 430-460), then RestrictFMGSource (:447-456) down to the root level.
extern "C" int agx_mg_setup(agx_mesh *m) {
  int rc = mg_level_check(m, 0, 0, "agx_mg_setup");
  if (rc) return rc;
  const Geo &g = m->g;
  const int L = m->ntotal;
  double **A = m->arr.a;
  if (m->mp_on) {
    MGLevelB &F = m->levb[L - 1];
    hipLaunchKernelGGL(k_mb_load, dim3(grid1d((int64_t)F.n1 * F.n2 * F.n3, BS)), dim3(BS), 0, m->stream, g,
                       A[AGX_ARR_U], A[AGX_ARR_PHI], F.src, F.u, m->d.four_pi_G, m->mg_mode == 1 ? 1 : 0,
                       src_mask(m));
    mp_coefficients(m);
    m->mp_solved = true;
    for (int l = L - 1; l > 0; --l) {
      MGLevelB &Fl = m->levb[l], &Cl = m->levb[l - 1];
      hipLaunchKernelGGL(k_mb_restrict_src, dim3(grid1d((int64_t)Cl.n1 * Cl.n2 * Cl.n3, BS)), dim3(BS), 0,
                         m->stream, Fl.src, Cl.src, Fl.n1, Fl.n2, Fl.n3);
    }
  } else {
    MGLevel &F = m->lev[L - 1];
    hipLaunchKernelGGL(k_mg_load, dim3(grid1d(lcells(F), BS)), dim3(BS), 0, m->stream, g, A[AGX_ARR_U],
                       A[AGX_ARR_PHI], F.src, F.u, m->d.four_pi_G, m->mg_mode == 1 ? 1 : 0, src_mask(m));
    subtract_average(m, 0);
    for (int l = L - 1; l > 0; --l) {
      MGLevel &Fl = m->lev[l], &Cl = m->lev[l - 1];
      hipLaunchKernelGGL(k_mg_restrict_src<double>, dim3(grid1d(lcells(Cl), BS)), dim3(BS), 0, m->stream, Fl.src,
                         Cl.src, Fl.n1, Fl.n2, Fl.n3);
    }
  }
  m->current = L - 1;
  m->fmglevel = L - 1;
  return launches_ok("agx_mg_setup");
}

// MGGravity::Smooth (mg_gravity.cpp:220-273): one colour of the red-black sweep on a level
// (multipole meshes: interior cells; fill the ghosts first, agx_mg_fill_ghost).  The colour
// is combined with the driver's colour offset (agx_mg_coffset) as mg_gravity.cpp:227,243 do.
extern "C" int agx_mg_smooth(agx_mesh *m, int32_t level, int32_t color) {
  int rc = mg_level_check(m, level, 0, "agx_mg_smooth");
  if (rc) return rc;
  if (color != 0 && color != 1) { set_error("agx_mg_smooth: colour must be 0 or 1"); return -3; }
  if (m->mp_on) smooth_b(m, level, color);
  else smooth(m, m->lev, level, color);
  return launches_ok("agx_mg_smooth");
}

// The ghost fill of a level before a stencil step (MGBoundaryValues, bvals_mg.cpp:321-370
// with MGMultipoleInnerX1..OuterX3, mgbval_multipole.cpp:28-426).  Periodic levels have no
// stored ghosts (neighbours are read through the periodic index wrap): nothing to do.
extern "C" int agx_mg_fill_ghost(agx_mesh *m, int32_t level) {
  int rc = mg_level_check(m, level, 0, "agx_mg_fill_ghost");
  if (rc) return rc;
  if (m->mp_on) bc_b(m, level);
  return launches_ok("agx_mg_fill_ghost");
}

// Multigrid::StoreOldData + MGGravity::CalculateFASRHS (multigrid.cpp:648-653,
// mg_gravity.cpp:324-351) on a level; multipole (FAS) meshes only.
extern "C" int agx_mg_fas(agx_mesh *m, int32_t level) {
  int rc = mg_level_check(m, level, 0, "agx_mg_fas");
  if (rc) return rc;
  if (!m->mp_on) { set_error("agx_mg_fas: FAS runs with multipole boundaries only (multigrid_driver.cpp:292-295)"); return -22; }
  fas_b(m, level);
  return launches_ok("agx_mg_fas");
}

// MGGravity::CalculateDefect + Multigrid::Restrict (mg_gravity.cpp:283-314,
// multigrid.cpp:684-724): the defect of `level` restricted into the source of level-1
// (with FAS also u), in one pass.
extern "C" int agx_mg_restrict(agx_mesh *m, int32_t level) {
  int rc = mg_level_check(m, level, 1, "agx_mg_restrict");
  if (rc) return rc;
  if (m->mp_on) {
    MGLevelB &F = m->levb[level], &C = m->levb[level - 1];
    hipLaunchKernelGGL(k_mb_restrict<double>, dim3(grid1d((int64_t)C.n1 * C.n2 * C.n3, BS)), dim3(BS), 0,
                       m->stream, m->gate, F.u, F.src, C.src, C.u, F.n1, F.n2, F.n3, 1.0 / (F.dx * F.dx));
  } else {
    defect_restrict(m, m->lev, level);
  }
  return launches_ok("agx_mg_restrict");
}

// Multigrid::ProlongateAndCorrect (multigrid.cpp:733-829): level-1 -> level.
extern "C" int agx_mg_prolong(agx_mesh *m, int32_t level) {
  int rc = mg_level_check(m, level, 1, "agx_mg_prolong");
  if (rc) return rc;
  if (m->mp_on) {
    MGLevelB &F = m->levb[level], &C = m->levb[level - 1];
    mb_prolong(m, C.u, C.uold, F.u, F.n1, F.n2, F.n3);
  } else {
    prolong(m, m->lev, level);
  }
  return launches_ok("agx_mg_prolong");
}

// Multigrid::FMGProlongate (multigrid.cpp:838-1048): level-1 -> level (multipole meshes:
// fill level-1's ghosts first, as SolveFMGCycle's exchange does).
extern "C" int agx_mg_fmg_prolong(agx_mesh *m, int32_t level) {
  int rc = mg_level_check(m, level, 1, "agx_mg_fmg_prolong");
  if (rc) return rc;
  if (m->mp_on) {
    MGLevelB &F = m->levb[level], &C = m->levb[level - 1];
    hipLaunchKernelGGL(k_mb_fmg_prolong, dim3(grid1d((int64_t)F.n1 * F.n2 * F.n3, BS)), dim3(BS), 0, m->stream, C.u,
                       F.u, F.n1, F.n2, F.n3);
  } else {
    fmg_prolong(m, m->lev, level);
  }
  return launches_ok("agx_mg_fmg_prolong");
}

// MultigridDriver::SolveCoarsestGrid (multigrid_driver.cpp:778-819) on level 0.
extern "C" int agx_mg_solve_coarsest(agx_mesh *m) {
  int rc = mg_level_check(m, 0, 0, "agx_mg_solve_coarsest");
  if (rc) return rc;
  if (m->mp_on) {
    const Geo &g = m->g;
    const int ni = std::max(g.nb1, std::max(g.nb2, g.nb3)) >> (m->nrootlevel - 1);
    bc_b(m, 0);
    fas_b(m, 0);
    for (int it = 0; it < ni; ++it) {
      smooth_b(m, 0, 0);
      bc_b(m, 0);
      smooth_b(m, 0, 1);
      bc_b(m, 0);
    }
  } else {
    solve_coarsest(m, m->lev);
  }
  return launches_ok("agx_mg_solve_coarsest");
}

// MultigridDriver::CalculateDefectNorm (L2, multigrid_driver.cpp:826-858) of the finest level.
extern "C" int agx_mg_defect_norm(agx_mesh *m, double *norm) {
  int rc = mg_level_check(m, 0, 0, "agx_mg_defect_norm");
  if (rc) return rc;
  if (!norm) { set_error("agx_mg_defect_norm: null output"); return -1; }
  *norm = m->mp_on ? defect_norm_b(m) : defect_norm(m);
  return launches_ok("agx_mg_defect_norm");
}

// Multigrid::RetrieveResult (multigrid.cpp:266-288): the finest u, with one ghost cell,
// into Gravity::phi; on periodic meshes after SubtractAverage of u, as
// MGGravityDriver::Solve does (multigrid_driver.cpp:755-771).
extern "C" int agx_mg_retrieve(agx_mesh *m) {
  int rc = mg_level_check(m, 0, 0, "agx_mg_retrieve");
  if (rc) return rc;
  const Geo &g = m->g;
  const int64_t nret = (int64_t)(g.bx1 + 2) * (g.bx2 + 2) * (g.bx3 + 2) * g.nb;
  if (m->mp_on) {
    hipLaunchKernelGGL(k_mb_retrieve, dim3(grid1d(nret, BS)), dim3(BS), 0, m->stream, g, m->levb[m->ntotal - 1].u,
                       m->arr.a[AGX_ARR_PHI]);
  } else {
    subtract_average(m, 1);
    hipLaunchKernelGGL(k_mg_retrieve, dim3(grid1d(nret, BS)), dim3(BS), 0, m->stream, g, m->lev[m->ntotal - 1].u,
                       m->arr.a[AGX_ARR_PHI]);
  }
  return launches_ok("agx_mg_retrieve");
}

// The red-black colour offset of the driver (MultigridDriver::coffset_, flipped at the
// start of every V-cycle, multigrid_driver.cpp:686); flip != 0 flips it first.
extern "C" int agx_mg_coffset(agx_mesh *m, int32_t flip, int32_t *value) {
  int rc = mg_level_check(m, 0, 0, "agx_mg_coffset");
  if (rc) return rc;
  if (flip) m->coffset ^= 1;
  if (value) *value = m->coffset;
  return 0;
}

// Levels of the hierarchy: n = nlevels; dims[3 l + d] = cells of level l along d.
extern "C" int agx_mg_levels(agx_mesh *m, int32_t *n, int32_t *dims, int32_t cap) {
  int rc = mg_level_check(m, 0, 0, "agx_mg_levels");
  if (rc) return rc;
  if (n) *n = m->ntotal;
  for (int l = 0; dims && l < m->ntotal && l < cap; ++l) {
    const int n1 = m->mp_on ? m->levb[l].n1 : m->lev[l].n1, n2 = m->mp_on ? m->levb[l].n2 : m->lev[l].n2,
              n3 = m->mp_on ? m->levb[l].n3 : m->lev[l].n3;
    dims[3 * l] = n1; dims[3 * l + 1] = n2; dims[3 * l + 2] = n3;
  }
  return 0;
}

extern "C" int agx_mg_set_source_mask_sphere(agx_mesh *m, double xc, double yc, double zc, double radius) {
  if (!m) { set_error("agx_mg_set_source_mask_sphere: null mesh"); return -1; }
  if (!m->mg_on) { set_error("agx_mg_set_source_mask_sphere: the mesh has no multigrid gravity"); return -21; }
  m->srcmask = radius > 0.0 ? 1 : 0;
  m->srcmask_c[0] = xc; m->srcmask_c[1] = yc; m->srcmask_c[2] = zc;
  m->srcmask_r2 = radius * radius;   // collapse.cpp:68: rc2 = rc*rc
  m->graph_gen++;
  return 0;
}

extern "C" int agx_set_mg_coarse(agx_mesh *m, int32_t on) {
  if (!m) { set_error("agx_set_mg_coarse: null mesh"); return -1; }
  m->mg_coarse = on ? 1 : 0;
  m->graph_gen++;
  return 0;
}
