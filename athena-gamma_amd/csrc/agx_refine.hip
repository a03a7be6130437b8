// agx_refine.hip -- static-mesh-refinement operators for gfx950 (3D Cartesian blocks).
//
// Restriction, prolongation and flux/EMF correction of the reference's MeshRefinement
// and flux-correction boundary variables, batched over device-resident work lists: one
// launch runs every item of a list, a (G x nitems) grid in which workgroup (x, m) walks
// item m's index range with stride G * 256.  A list is checked and uploaded once
// (agx_ref_list_create), so the kernels index without further bounds tests.  Arrays are
// the reference's dense AthenaArray layouts, block-major; every expression keeps the
// reference's operand order, so results equal the reference's FP64 arithmetic
// (-ffp-contract=off).  On a 3D Cartesian block the areas, volumes and edge lengths are
// products of dx?f (coordinates/coordinates.cpp:328-537) and the face-area positions
// x?s? are x?v (coordinates/cartesian.cpp:106-119).
//
// All of these operators are bound by HBM traffic and index arithmetic.  Each reads
// O(1) values per output and does no reuse worth an LDS stage, and the items are small
// (a block face or a block of coarse cells).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "agx_internal.hpp"

struct agx_ref_list {
  agx_ref_geom g;
  int32_t kind = 0;
  int32_t nitems = 0;
  int32_t G = 1;              // workgroups per item
  int64_t emf_max = 0;        // largest per-item buffer length of an EMF list
  int64_t max_count = 0;      // largest per-item index count (times nvar must stay < 2^31)
  int32_t nphase = 1;         // EMF_SET lists: launches in order, items sorted by phase
  int32_t phase_off[17] = {0};
  agx_ref_item *items = nullptr;   // device
};

namespace agx {
namespace {

constexpr int RBS = 256;

struct RG {            // derived geometry (see agx_ref_geom)
  int n1, n2, n3;      // block cells incl. ghosts
  int c1, c2, c3;      // coarse buffer cells incl. ghosts
  int is, js, ks, ie, je, ke;
  int cis, cjs, cks;
  int h1, h2, h3;      // nx / 2
  int64_t bS;          // block arrays in the mesh's padded layout: slab per variable (0: dense)
  int bJ, bP;          // rows per plane and row pitch of that layout
};

RG make_rg(const agx_ref_geom &g) {
  RG r;
  r.n1 = g.nx1 + 2 * g.ng; r.n2 = g.nx2 + 2 * g.ng; r.n3 = g.nx3 + 2 * g.ng;
  r.c1 = g.nx1 / 2 + 2 * g.cng; r.c2 = g.nx2 / 2 + 2 * g.cng; r.c3 = g.nx3 / 2 + 2 * g.cng;
  r.is = r.js = r.ks = g.ng;
  r.ie = g.ng + g.nx1 - 1; r.je = g.ng + g.nx2 - 1; r.ke = g.ng + g.nx3 - 1;
  r.cis = r.cjs = r.cks = g.cng;
  r.h1 = g.nx1 / 2; r.h2 = g.nx2 / 2; r.h3 = g.nx3 / 2;
  r.bS = 0;
  r.bJ = r.bP = 0;
  if (g.layout == AGX_REF_LAYOUT_MESH) {   // as agx_mesh_create lays out its arrays
    r.bP = ((r.n1 + 1) + 31) / 32 * 32;
    r.bJ = r.n2 + 1;
    r.bS = (int64_t)(r.n3 + 1) * r.bJ * r.bP;
  }
  return r;
}

// element (b, n, k, j, i) of a block-major array of nvar variables of d3 x d2 x d1
__device__ __forceinline__ int64_t ix(int d3, int d2, int d1, int nvar, int b, int n, int k, int j, int i) {
  return ((((int64_t)b * nvar + n) * d3 + k) * d2 + j) * (int64_t)d1 + i;
}
// slabs per block of a mesh array of nvar variables: the mesh's hydro arrays (u, w, the
// fluxes) hold 5 slabs for either EOS, the isothermal one using the first 4
__device__ __forceinline__ int mesh_slabs(int nvar) { return nvar == 4 ? 5 : nvar; }
// the same element of a block-side array (MeshBlock data, not a coarse buffer): the dense
// layout above, or the mesh's padded one, where every cell, face and edge array of a
// block has its slabs of bS elements, bJ rows of bP elements per plane
__device__ __forceinline__ int64_t bx(const RG &r, int d3, int d2, int d1, int nvar, int b, int n, int k, int j,
                                      int i) {
  if (r.bS) return ((int64_t)b * mesh_slabs(nvar) + n) * r.bS + ((int64_t)k * r.bJ + j) * r.bP + i;
  return ix(d3, d2, d1, nvar, b, n, k, j, i);
}
// an array of either side, for the operators that move data both ways
struct Lay {
  int d1, d2, d3;      // dense dims
  int64_t S;           // > 0: padded layout (slab S, J rows of P per plane)
  int J, P;
};
__device__ __forceinline__ int64_t lx(const Lay &L, int nvar, int b, int n, int k, int j, int i) {
  if (L.S) return ((int64_t)b * mesh_slabs(nvar) + n) * L.S + ((int64_t)k * L.J + j) * L.P + i;
  return ix(L.d3, L.d2, L.d1, nvar, b, n, k, j, i);
}

__device__ __forceinline__ double ref_sign(double x) { return (x < 0.0) ? -1.0 : 1.0; }   // SIGN
// 0.5*(SIGN(gm) + SIGN(gp))*min(|gm|, |gp|), mesh_refinement.cpp:427-438
__device__ __forceinline__ double minmod_slope(double gm, double gp) {
  const double am = fabs(gm), ap = fabs(gp);
  return 0.5 * (ref_sign(gm) + ref_sign(gp)) * ((ap < am) ? ap : am);
}

struct Range {
  int si, sj, sk, w1, w2, w3;
};
__device__ __forceinline__ Range item_range(const agx_ref_item &it) {
  return Range{it.si, it.sj, it.sk, it.ei - it.si + 1, it.ej - it.sj + 1, it.ek - it.sk + 1};
}
// Per-item counts are below 2^32 (checked on the host), so the walk and the index
// decomposition use 32-bit unsigned arithmetic (64-bit divisions cost several times more).
__device__ __forceinline__ unsigned first_t() { return blockIdx.x * RBS + threadIdx.x; }
__device__ __forceinline__ unsigned stride_t() { return gridDim.x * RBS; }

// ---------------------------------------------------------------- restriction
// RestrictCellCenteredValues, mesh_refinement.cpp:106-136
__device__ __forceinline__ void
b_ref_restrict_cc(RG r, int nvar, const double *__restrict__ fine, agx_ref_coords fc,
                  double *__restrict__ coarse, const agx_ref_item &it) {
  const Range q = item_range(it);
  const unsigned w1 = q.w1, w2 = q.w2, per = w1 * w2 * q.w3, cnt = per * nvar;
  const double *d1 = fc.dx1f + (int64_t)it.fb * r.n1, *d2 = fc.dx2f + (int64_t)it.fb * r.n2,
               *d3 = fc.dx3f + (int64_t)it.fb * r.n3;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int n = (int)(t / per);
    unsigned s = t - n * per;
    const int ci = q.si + (int)(s % w1);
    s /= w1;
    const int cj = q.sj + (int)(s % w2), ck = q.sk + (int)(s / w2);
    const int i = (ci - r.cis) * 2 + r.is, j = (cj - r.cjs) * 2 + r.js, k = (ck - r.cks) * 2 + r.ks;
    double v[2][2][2];   // CellVolume at (k+a, j+b, i+c) = dx1f*dx2f*dx3f
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) v[a][b][c] = d1[i + c] * d2[j + b] * d3[k + a];
    const double tvol = ((v[0][0][0] + v[0][1][0]) + (v[0][0][1] + v[0][1][1])) +
                        ((v[1][0][0] + v[1][1][0]) + (v[1][0][1] + v[1][1][1]));
    auto F = [&](int a, int b, int c) { return fine[bx(r, r.n3, r.n2, r.n1, nvar, it.fb, n, k + a, j + b, i + c)]; };
    const double sum = (((F(0, 0, 0) * v[0][0][0] + F(0, 1, 0) * v[0][1][0]) +
                         (F(0, 0, 1) * v[0][0][1] + F(0, 1, 1) * v[0][1][1])) +
                        ((F(1, 0, 0) * v[1][0][0] + F(1, 1, 0) * v[1][1][0]) +
                         (F(1, 0, 1) * v[1][0][1] + F(1, 1, 1) * v[1][1][1])));
    coarse[ix(r.c3, r.c2, r.c1, nvar, it.cb, n, ck, cj, ci)] = sum / tvol;
  }
}

__global__ void __launch_bounds__(RBS)
k_ref_restrict_cc(RG r, int nvar, const double *__restrict__ fine, agx_ref_coords fc,
                  double *__restrict__ coarse, const agx_ref_item *__restrict__ items) {
  b_ref_restrict_cc(r, nvar, fine, fc, coarse, items[blockIdx.y]);
}

// RestrictFieldX1/X2/X3, mesh_refinement.cpp:185-214, 244-279, 321-344
template <int DIR>
__device__ __forceinline__ void
b_ref_restrict_fc(RG r, const double *__restrict__ fine, agx_ref_coords fc, double *__restrict__ coarse,
                  const agx_ref_item &it) {
  const Range q = item_range(it);
  const unsigned cnt = (unsigned)q.w1 * q.w2 * q.w3;
  const int f3 = r.n3 + (DIR == 2), f2 = r.n2 + (DIR == 1), f1 = r.n1 + (DIR == 0);
  const int q3 = r.c3 + (DIR == 2), q2 = r.c2 + (DIR == 1), q1 = r.c1 + (DIR == 0);
  const double *d1 = fc.dx1f + (int64_t)it.fb * r.n1, *d2 = fc.dx2f + (int64_t)it.fb * r.n2,
               *d3 = fc.dx3f + (int64_t)it.fb * r.n3;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int ci = q.si + (int)(t % (unsigned)q.w1);
    const unsigned s = t / (unsigned)q.w1;
    const int cj = q.sj + (int)(s % (unsigned)q.w2), ck = q.sk + (int)(s / (unsigned)q.w2);
    const int i = (ci - r.cis) * 2 + r.is, j = (cj - r.cjs) * 2 + r.js, k = (ck - r.cks) * 2 + r.ks;
    auto F = [&](int a, int b, int c) { return fine[bx(r, f3, f2, f1, 1, it.fb, 0, k + a, j + b, i + c)]; };
    double val;
    if (DIR == 0) {   // Face1Area = dx2f(j)*dx3f(k)
      const double a00 = d2[j] * d3[k], a01 = d2[j + 1] * d3[k];
      const double a10 = d2[j] * d3[k + 1], a11 = d2[j + 1] * d3[k + 1];
      const double tarea = a00 + a01 + a10 + a11;
      val = (F(0, 0, 0) * a00 + F(0, 1, 0) * a01 + F(1, 0, 0) * a10 + F(1, 1, 0) * a11) / tarea;
    } else if (DIR == 1) {   // Face2Area = dx1f(i)*dx3f(k)
      const double s00 = d1[i] * d3[k], s01 = d1[i + 1] * d3[k];
      const double s10 = d1[i] * d3[k + 1], s11 = d1[i + 1] * d3[k + 1];
      const double tarea = s00 + s01 + s10 + s11;
      val = (F(0, 0, 0) * s00 + F(0, 0, 1) * s01 + F(1, 0, 0) * s10 + F(1, 0, 1) * s11) / tarea;
    } else {   // Face3Area = dx1f(i)*dx2f(j)
      const double s00 = d1[i] * d2[j], s01 = d1[i + 1] * d2[j];
      const double s10 = d1[i] * d2[j + 1], s11 = d1[i + 1] * d2[j + 1];
      const double tarea = s00 + s01 + s10 + s11;
      val = (F(0, 0, 0) * s00 + F(0, 0, 1) * s01 + F(0, 1, 0) * s10 + F(0, 1, 1) * s11) / tarea;
    }
    coarse[ix(q3, q2, q1, 1, it.cb, 0, ck, cj, ci)] = val;
  }
}

template <int DIR>
__global__ void __launch_bounds__(RBS)
k_ref_restrict_fc(RG r, const double *__restrict__ fine, agx_ref_coords fc, double *__restrict__ coarse,
                  const agx_ref_item *__restrict__ items) {
  b_ref_restrict_fc<DIR>(r, fine, fc, coarse, items[blockIdx.y]);
}

// ---------------------------------------------------------------- prolongation
// ProlongateCellCenteredValues (3D), mesh_refinement.cpp:386-458
__device__ __forceinline__ void
b_ref_prolong_cc(RG r, int nvar, const double *__restrict__ coarse, agx_ref_coords cc, double *__restrict__ fine,
                 agx_ref_coords fc, const agx_ref_item &it) {
  const Range q = item_range(it);
  const unsigned w1 = q.w1, w2 = q.w2, per = w1 * w2 * q.w3, cnt = per * nvar;
  const double *X1 = cc.x1v + (int64_t)it.cb * r.c1, *X2 = cc.x2v + (int64_t)it.cb * r.c2,
               *X3 = cc.x3v + (int64_t)it.cb * r.c3;
  const double *x1 = fc.x1v + (int64_t)it.fb * r.n1, *x2 = fc.x2v + (int64_t)it.fb * r.n2,
               *x3 = fc.x3v + (int64_t)it.fb * r.n3;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int n = (int)(t / per);
    unsigned s = t - n * per;
    const int i = q.si + (int)(s % w1);
    s /= w1;
    const int j = q.sj + (int)(s % w2), k = q.sk + (int)(s / w2);
    const int fi = (i - r.cis) * 2 + r.is, fj = (j - r.cjs) * 2 + r.js, fk = (k - r.cks) * 2 + r.ks;
    const double dx3m = X3[k] - X3[k - 1], dx3p = X3[k + 1] - X3[k];
    const double dx3fm = X3[k] - x3[fk], dx3fp = x3[fk + 1] - X3[k];
    const double dx2m = X2[j] - X2[j - 1], dx2p = X2[j + 1] - X2[j];
    const double dx2fm = X2[j] - x2[fj], dx2fp = x2[fj + 1] - X2[j];
    const double dx1m = X1[i] - X1[i - 1], dx1p = X1[i + 1] - X1[i];
    const double dx1fm = X1[i] - x1[fi], dx1fp = x1[fi + 1] - X1[i];
    auto C = [&](int a, int b, int c) { return coarse[ix(r.c3, r.c2, r.c1, nvar, it.cb, n, k + a, j + b, i + c)]; };
    const double cv = C(0, 0, 0);
    const double g1 = minmod_slope((cv - C(0, 0, -1)) / dx1m, (C(0, 0, 1) - cv) / dx1p);
    const double g2 = minmod_slope((cv - C(0, -1, 0)) / dx2m, (C(0, 1, 0) - cv) / dx2p);
    const double g3 = minmod_slope((cv - C(-1, 0, 0)) / dx3m, (C(1, 0, 0) - cv) / dx3p);
    auto FO = [&](int a, int b, int c) -> double & {
      return fine[bx(r, r.n3, r.n2, r.n1, nvar, it.fb, n, fk + a, fj + b, fi + c)];
    };
    FO(0, 0, 0) = cv - (g1 * dx1fm + g2 * dx2fm + g3 * dx3fm);
    FO(0, 0, 1) = cv + (g1 * dx1fp - g2 * dx2fm - g3 * dx3fm);
    FO(0, 1, 0) = cv - (g1 * dx1fm - g2 * dx2fp + g3 * dx3fm);
    FO(0, 1, 1) = cv + (g1 * dx1fp + g2 * dx2fp - g3 * dx3fm);
    FO(1, 0, 0) = cv - (g1 * dx1fm + g2 * dx2fm - g3 * dx3fp);
    FO(1, 0, 1) = cv + (g1 * dx1fp - g2 * dx2fm + g3 * dx3fp);
    FO(1, 1, 0) = cv - (g1 * dx1fm - g2 * dx2fp - g3 * dx3fp);
    FO(1, 1, 1) = cv + (g1 * dx1fp + g2 * dx2fp + g3 * dx3fp);
  }
}

__global__ void __launch_bounds__(RBS)
k_ref_prolong_cc(RG r, int nvar, const double *__restrict__ coarse, agx_ref_coords cc, double *__restrict__ fine,
                 agx_ref_coords fc, const agx_ref_item *__restrict__ items) {
  b_ref_prolong_cc(r, nvar, coarse, cc, fine, fc, items[blockIdx.y]);
}

// ProlongateSharedFieldX1/X2/X3 (3D), mesh_refinement.cpp:541-581, 626-674, 718-762
template <int DIR>
__device__ __forceinline__ void
b_ref_prolong_fc_shared(RG r, const double *__restrict__ coarse, agx_ref_coords cc, double *__restrict__ fine,
                        agx_ref_coords fc, const agx_ref_item &it) {
  const Range q = item_range(it);
  const unsigned cnt = (unsigned)q.w1 * q.w2 * q.w3;
  const int f3 = r.n3 + (DIR == 2), f2 = r.n2 + (DIR == 1), f1 = r.n1 + (DIR == 0);
  const int q3 = r.c3 + (DIR == 2), q2 = r.c2 + (DIR == 1), q1 = r.c1 + (DIR == 0);
  const double *X1 = cc.x1v + (int64_t)it.cb * r.c1, *X2 = cc.x2v + (int64_t)it.cb * r.c2,
               *X3 = cc.x3v + (int64_t)it.cb * r.c3;
  const double *x1 = fc.x1v + (int64_t)it.fb * r.n1, *x2 = fc.x2v + (int64_t)it.fb * r.n2,
               *x3 = fc.x3v + (int64_t)it.fb * r.n3;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int i = q.si + (int)(t % (unsigned)q.w1);
    const unsigned s = t / (unsigned)q.w1;
    const int j = q.sj + (int)(s % (unsigned)q.w2), k = q.sk + (int)(s / (unsigned)q.w2);
    const int fi = (i - r.cis) * 2 + r.is, fj = (j - r.cjs) * 2 + r.js, fk = (k - r.cks) * 2 + r.ks;
    auto C = [&](int a, int b, int c) { return coarse[ix(q3, q2, q1, 1, it.cb, 0, k + a, j + b, i + c)]; };
    auto FO = [&](int a, int b, int c) -> double & { return fine[bx(r, f3, f2, f1, 1, it.fb, 0, fk + a, fj + b, fi + c)]; };
    const double cv = C(0, 0, 0);
    if (DIR == 0) {
      const double g2 = minmod_slope((cv - C(0, -1, 0)) / (X2[j] - X2[j - 1]), (C(0, 1, 0) - cv) / (X2[j + 1] - X2[j]));
      const double g3 = minmod_slope((cv - C(-1, 0, 0)) / (X3[k] - X3[k - 1]), (C(1, 0, 0) - cv) / (X3[k + 1] - X3[k]));
      FO(0, 0, 0) = cv - g2 * (X2[j] - x2[fj]) - g3 * (X3[k] - x3[fk]);
      FO(0, 1, 0) = cv + g2 * (x2[fj + 1] - X2[j]) - g3 * (X3[k] - x3[fk]);
      FO(1, 0, 0) = cv - g2 * (X2[j] - x2[fj]) + g3 * (x3[fk + 1] - X3[k]);
      FO(1, 1, 0) = cv + g2 * (x2[fj + 1] - X2[j]) + g3 * (x3[fk + 1] - X3[k]);
    } else if (DIR == 1) {
      const double g1 = minmod_slope((cv - C(0, 0, -1)) / (X1[i] - X1[i - 1]), (C(0, 0, 1) - cv) / (X1[i + 1] - X1[i]));
      const double g3 = minmod_slope((cv - C(-1, 0, 0)) / (X3[k] - X3[k - 1]), (C(1, 0, 0) - cv) / (X3[k + 1] - X3[k]));
      FO(0, 0, 0) = cv - g1 * (X1[i] - x1[fi]) - g3 * (X3[k] - x3[fk]);
      FO(0, 0, 1) = cv + g1 * (x1[fi + 1] - X1[i]) - g3 * (X3[k] - x3[fk]);
      FO(1, 0, 0) = cv - g1 * (X1[i] - x1[fi]) + g3 * (x3[fk + 1] - X3[k]);
      FO(1, 0, 1) = cv + g1 * (x1[fi + 1] - X1[i]) + g3 * (x3[fk + 1] - X3[k]);
    } else {
      const double g1 = minmod_slope((cv - C(0, 0, -1)) / (X1[i] - X1[i - 1]), (C(0, 0, 1) - cv) / (X1[i + 1] - X1[i]));
      const double g2 = minmod_slope((cv - C(0, -1, 0)) / (X2[j] - X2[j - 1]), (C(0, 1, 0) - cv) / (X2[j + 1] - X2[j]));
      FO(0, 0, 0) = cv - g1 * (X1[i] - x1[fi]) - g2 * (X2[j] - x2[fj]);
      FO(0, 0, 1) = cv + g1 * (x1[fi + 1] - X1[i]) - g2 * (X2[j] - x2[fj]);
      FO(0, 1, 0) = cv - g1 * (X1[i] - x1[fi]) + g2 * (x2[fj + 1] - X2[j]);
      FO(0, 1, 1) = cv + g1 * (x1[fi + 1] - X1[i]) + g2 * (x2[fj + 1] - X2[j]);
    }
  }
}

template <int DIR>
__global__ void __launch_bounds__(RBS)
k_ref_prolong_fc_shared(RG r, const double *__restrict__ coarse, agx_ref_coords cc, double *__restrict__ fine,
                        agx_ref_coords fc, const agx_ref_item *__restrict__ items) {
  b_ref_prolong_fc_shared<DIR>(r, coarse, cc, fine, fc, items[blockIdx.y]);
}

// ProlongateInternalField (3D), mesh_refinement.cpp:831-946: the 12 interior faces of
// each 2x2x2 group of fine cells from its 24 outer faces; groups share no written face.
// ProlongateInternalField for the 2x2x2 fine group of coarse cell (k, j, i) of block fb:
// the 12 faces inside the group from its 24 outer faces (mesh_refinement.cpp:831-946).
__device__ __forceinline__ void
internal_group(const RG &r, double *__restrict__ b1, double *__restrict__ b2, double *__restrict__ b3,
               const double *__restrict__ d1, const double *__restrict__ d2, const double *__restrict__ d3, int fb,
               int i, int j, int k) {
    const int fi = (i - r.cis) * 2 + r.is, fj = (j - r.cjs) * 2 + r.js, fk = (k - r.cks) * 2 + r.ks;
    auto B1 = [&](int kk, int jj, int ii) -> double & { return b1[bx(r, r.n3, r.n2, r.n1 + 1, 1, fb, 0, kk, jj, ii)]; };
    auto B2 = [&](int kk, int jj, int ii) -> double & { return b2[bx(r, r.n3, r.n2 + 1, r.n1, 1, fb, 0, kk, jj, ii)]; };
    auto B3 = [&](int kk, int jj, int ii) -> double & { return b3[bx(r, r.n3 + 1, r.n2, r.n1, 1, fb, 0, kk, jj, ii)]; };
    // Face?Area(fk+a, fj+b, ii)
    auto A1 = [&](int a, int b) { return d2[fj + b] * d3[fk + a]; };
    auto A2 = [&](int a, int ii) { return d1[ii] * d3[fk + a]; };
    auto A3 = [&](int b, int ii) { return d1[ii] * d2[fj + b]; };
    double Uxx = 0.0, Vyy = 0.0, Wzz = 0.0, Uxyz = 0.0, Vxyz = 0.0, Wxyz = 0.0;
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int js = 2 * jj - 1, fjj = fj + jj, fjp = fj + 2 * jj;
#pragma unroll
      for (int ii = 0; ii < 2; ++ii) {
        const int is = 2 * ii - 1, fii = fi + ii, fip = fi + 2 * ii;
        const double x2lo = B2(fk, fjp, fii) * A2(0, fii), x2hi = B2(fk + 1, fjp, fii) * A2(1, fii);
        const double x3lo = B3(fk, fjj, fii) * A3(jj, fii), x3hi = B3(fk + 2, fjj, fii) * A3(jj, fii);
        const double x1lo = B1(fk, fjj, fip) * A1(0, jj), x1hi = B1(fk + 1, fjj, fip) * A1(1, jj);
        Uxx += is * (js * (x2lo + x2hi) + (x3hi - x3lo));
        Vyy += js * ((x3hi - x3lo) + is * (x1lo + x1hi));
        Wzz += is * (x1hi - x1lo) + js * (x2hi - x2lo);
        Uxyz += is * js * (x1hi - x1lo);
        Vxyz += is * js * (x2hi - x2lo);
        Wxyz += is * js * (x3hi - x3lo);
      }
    }
    const double s1 = d1[fi] + d1[fi + 1], s2 = d2[fj] + d2[fj + 1], s3 = d3[fk] + d3[fk + 1];
    const double Sdx1 = s1 * s1, Sdx2 = s2 * s2, Sdx3 = s3 * s3;
    Uxx *= 0.125; Vyy *= 0.125; Wzz *= 0.125;
    Uxyz *= 0.125 / (Sdx2 + Sdx3);
    Vxyz *= 0.125 / (Sdx1 + Sdx3);
    Wxyz *= 0.125 / (Sdx1 + Sdx2);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        double v = 0.5 * (B1(fk + a, fj + b, fi) * A1(a, b) + B1(fk + a, fj + b, fi + 2) * A1(a, b)) + Uxx;
        v = a ? v + Sdx3 * Vxyz : v - Sdx3 * Vxyz;
        v = b ? v + Sdx2 * Wxyz : v - Sdx2 * Wxyz;
        B1(fk + a, fj + b, fi + 1) = v / A1(a, b);
      }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        double v = 0.5 * (B2(fk + a, fj, fi + c) * A2(a, fi + c) + B2(fk + a, fj + 2, fi + c) * A2(a, fi + c)) + Vyy;
        v = a ? v + Sdx3 * Uxyz : v - Sdx3 * Uxyz;
        v = c ? v + Sdx1 * Wxyz : v - Sdx1 * Wxyz;
        B2(fk + a, fj + 1, fi + c) = v / A2(a, fi + c);
      }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        double v = 0.5 * (B3(fk + 2, fj + b, fi + c) * A3(b, fi + c) + B3(fk, fj + b, fi + c) * A3(b, fi + c)) + Wzz;
        v = b ? v + Sdx2 * Uxyz : v - Sdx2 * Uxyz;
        v = c ? v + Sdx1 * Vxyz : v - Sdx1 * Vxyz;
        B3(fk + 1, fj + b, fi + c) = v / A3(b, fi + c);
      }
}

__device__ __forceinline__ void
b_ref_prolong_fc_internal(RG r, double *__restrict__ b1, double *__restrict__ b2, double *__restrict__ b3,
                          agx_ref_coords fc, const agx_ref_item &it) {
  const Range q = item_range(it);
  const unsigned cnt = (unsigned)q.w1 * q.w2 * q.w3;
  const int fb = it.fb;
  const double *d1 = fc.dx1f + (int64_t)fb * r.n1, *d2 = fc.dx2f + (int64_t)fb * r.n2,
               *d3 = fc.dx3f + (int64_t)fb * r.n3;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int i = q.si + (int)(t % (unsigned)q.w1);
    const unsigned s = t / (unsigned)q.w1;
    const int j = q.sj + (int)(s % (unsigned)q.w2), k = q.sk + (int)(s / (unsigned)q.w2);
    internal_group(r, b1, b2, b3, d1, d2, d3, fb, i, j, k);
  }
}

__global__ void __launch_bounds__(RBS)
k_ref_prolong_fc_internal(RG r, double *__restrict__ b1, double *__restrict__ b2, double *__restrict__ b3,
                          agx_ref_coords fc, const agx_ref_item *__restrict__ items) {
  b_ref_prolong_fc_internal(r, b1, b2, b3, fc, items[blockIdx.y]);
}

// Field::CalculateCellCenteredField, uniform branch (field.cpp:112-176: lw = rw = 0.5), then
// EquationOfState::PrimitiveToConserved (adiabatic_mhd.cpp:99-136; isothermal_mhd.cpp:
// 82-108 without energy) on the fine cells of each coarse range.
// bcc and u of fine cell (k, j, i) of block fb from its faces and w (the loop body of
// CalculateCellCenteredField + PrimitiveToConserved)
template <bool ISO>
__device__ __forceinline__ void
p2c_cell(const RG &r, const double *__restrict__ w, const double *__restrict__ b1, const double *__restrict__ b2,
         const double *__restrict__ b3, double *__restrict__ bcc, double *__restrict__ u, double igm1, int fb, int i,
         int j, int k) {
  constexpr int NH = ISO ? 4 : 5;
    const double bc1 = 0.5 * b1[bx(r, r.n3, r.n2, r.n1 + 1, 1, fb, 0, k, j, i)] +
                       0.5 * b1[bx(r, r.n3, r.n2, r.n1 + 1, 1, fb, 0, k, j, i + 1)];
    const double bc2 = 0.5 * b2[bx(r, r.n3, r.n2 + 1, r.n1, 1, fb, 0, k, j, i)] +
                       0.5 * b2[bx(r, r.n3, r.n2 + 1, r.n1, 1, fb, 0, k, j + 1, i)];
    const double bc3 = 0.5 * b3[bx(r, r.n3 + 1, r.n2, r.n1, 1, fb, 0, k, j, i)] +
                       0.5 * b3[bx(r, r.n3 + 1, r.n2, r.n1, 1, fb, 0, k + 1, j, i)];
    bcc[bx(r, r.n3, r.n2, r.n1, 3, fb, 0, k, j, i)] = bc1;
    bcc[bx(r, r.n3, r.n2, r.n1, 3, fb, 1, k, j, i)] = bc2;
    bcc[bx(r, r.n3, r.n2, r.n1, 3, fb, 2, k, j, i)] = bc3;
    const double w_d = w[bx(r, r.n3, r.n2, r.n1, NH, fb, 0, k, j, i)];
    const double w_vx = w[bx(r, r.n3, r.n2, r.n1, NH, fb, 1, k, j, i)];
    const double w_vy = w[bx(r, r.n3, r.n2, r.n1, NH, fb, 2, k, j, i)];
    const double w_vz = w[bx(r, r.n3, r.n2, r.n1, NH, fb, 3, k, j, i)];
    u[bx(r, r.n3, r.n2, r.n1, NH, fb, 0, k, j, i)] = w_d;
    u[bx(r, r.n3, r.n2, r.n1, NH, fb, 1, k, j, i)] = w_vx * w_d;
    u[bx(r, r.n3, r.n2, r.n1, NH, fb, 2, k, j, i)] = w_vy * w_d;
    u[bx(r, r.n3, r.n2, r.n1, NH, fb, 3, k, j, i)] = w_vz * w_d;
    if (!ISO) {
      const double w_p = w[bx(r, r.n3, r.n2, r.n1, NH, fb, 4, k, j, i)];
      u[bx(r, r.n3, r.n2, r.n1, NH, fb, 4, k, j, i)] =
          w_p * igm1 + 0.5 * (w_d * (w_vx * w_vx + w_vy * w_vy + w_vz * w_vz) + (bc1 * bc1 + bc2 * bc2 + bc3 * bc3));
    }
}

template <bool ISO>
__device__ __forceinline__ void
b_ref_prim_to_cons(RG r, const double *__restrict__ w, const double *__restrict__ b1,
                   const double *__restrict__ b2, const double *__restrict__ b3, double *__restrict__ bcc,
                   double *__restrict__ u, double igm1, const agx_ref_item &it) {
  const Range q = item_range(it);
  const unsigned w1 = 2 * q.w1, w2 = 2 * q.w2, cnt = w1 * w2 * 2 * q.w3;   // fine cells
  const int fi0 = (q.si - r.cis) * 2 + r.is, fj0 = (q.sj - r.cjs) * 2 + r.js, fk0 = (q.sk - r.cks) * 2 + r.ks;
  const int fb = it.fb;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int i = fi0 + (int)(t % w1);
    const unsigned s = t / w1;
    const int j = fj0 + (int)(s % w2), k = fk0 + (int)(s / w2);
    p2c_cell<ISO>(r, w, b1, b2, b3, bcc, u, igm1, fb, i, j, k);
  }
}

template <bool ISO>
__global__ void __launch_bounds__(RBS)
k_ref_prim_to_cons(RG r, const double *__restrict__ w, const double *__restrict__ b1,
                   const double *__restrict__ b2, const double *__restrict__ b3, double *__restrict__ bcc,
                   double *__restrict__ u, double igm1, const agx_ref_item *__restrict__ items) {
  b_ref_prim_to_cons<ISO>(r, w, b1, b2, b3, bcc, u, igm1, items[blockIdx.y]);
}

// Field::CalculateCellCenteredField (uniform branch) and EquationOfState::
// ConservedToPrimitive (eos/adiabatic_mhd.cpp:41-90, isothermal_mhd.cpp:40-73) on the
// coarse-buffer cells of each item's range, as ApplyPhysicalBoundariesOnCoarseLevel
// calls them (bvals_refine.cpp:336-384) before the cell prolongation: coarse faces ->
// coarse bcc, coarse conserved -> coarse primitive, floors written back.
template <bool ISO>
__global__ void __launch_bounds__(RBS)
k_ref_cons_to_prim(RG r, double *__restrict__ u, const double *__restrict__ b1, const double *__restrict__ b2,
                   const double *__restrict__ b3, double *__restrict__ w, double *__restrict__ bcc, double gm1,
                   double dfl, double pfl, const agx_ref_item *__restrict__ items) {
  const agx_ref_item it = items[blockIdx.y];
  const Range q = item_range(it);
  const unsigned w1 = q.w1, w2 = q.w2, cnt = w1 * w2 * q.w3;
  constexpr int NH = ISO ? 4 : 5;
  const int cb = it.cb;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int i = q.si + (int)(t % w1);
    const unsigned s = t / w1;
    const int j = q.sj + (int)(s % w2), k = q.sk + (int)(s / w2);
    const double bc1 = 0.5 * b1[ix(r.c3, r.c2, r.c1 + 1, 1, cb, 0, k, j, i)] +
                       0.5 * b1[ix(r.c3, r.c2, r.c1 + 1, 1, cb, 0, k, j, i + 1)];
    const double bc2 = 0.5 * b2[ix(r.c3, r.c2 + 1, r.c1, 1, cb, 0, k, j, i)] +
                       0.5 * b2[ix(r.c3, r.c2 + 1, r.c1, 1, cb, 0, k, j + 1, i)];
    const double bc3 = 0.5 * b3[ix(r.c3 + 1, r.c2, r.c1, 1, cb, 0, k, j, i)] +
                       0.5 * b3[ix(r.c3 + 1, r.c2, r.c1, 1, cb, 0, k + 1, j, i)];
    bcc[ix(r.c3, r.c2, r.c1, 3, cb, 0, k, j, i)] = bc1;
    bcc[ix(r.c3, r.c2, r.c1, 3, cb, 1, k, j, i)] = bc2;
    bcc[ix(r.c3, r.c2, r.c1, 3, cb, 2, k, j, i)] = bc3;
    const int64_t o0 = ix(r.c3, r.c2, r.c1, NH, cb, 0, k, j, i);
    const int64_t S = (int64_t)r.c3 * r.c2 * r.c1;
    double u_d = u[o0], u_e = ISO ? 0.0 : u[o0 + 4 * S], wv[5];
    c2p_eos<ISO>(u_d, u[o0 + S], u[o0 + 2 * S], u[o0 + 3 * S], u_e, bc1, bc2, bc3, gm1, dfl, pfl, wv);
    u[o0] = u_d;
    if (!ISO) u[o0 + 4 * S] = u_e;
#pragma unroll
    for (int n = 0; n < NH; ++n) w[o0 + n * S] = wv[n];
  }
}

// ---------------------------------------------------------------- fused launches
// One launch for several lists of a step whose items touch disjoint data: workgroup row
// y runs item y of the concatenated lists.  The arithmetic is the single-list kernels'.
struct ItemSpan {
  const agx_ref_item *p[4];
  int n[4];
};
__device__ __forceinline__ int span_pick(const ItemSpan &sp, int &y) {
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    if (y < sp.n[l]) return l;
    y -= sp.n[l];
  }
  return -1;
}

// RestrictCellCenteredValues + RestrictFieldX1..3 (the same-level ghost restriction of
// one slot): lists RESTRICT_CC, FC1, FC2, FC3
__global__ void __launch_bounds__(RBS)
k_ref_restrict_all(RG r, int nvar, const double *__restrict__ u, const double *__restrict__ b1,
                   const double *__restrict__ b2, const double *__restrict__ b3, agx_ref_coords fc,
                   double *__restrict__ cu, double *__restrict__ cb1, double *__restrict__ cb2,
                   double *__restrict__ cb3, ItemSpan sp) {
  int y = blockIdx.y;
  const int l = span_pick(sp, y);
  if (l == 0) b_ref_restrict_cc(r, nvar, u, fc, cu, sp.p[0][y]);
  else if (l == 1) b_ref_restrict_fc<0>(r, b1, fc, cb1, sp.p[1][y]);
  else if (l == 2) b_ref_restrict_fc<1>(r, b2, fc, cb2, sp.p[2][y]);
  else if (l == 3) b_ref_restrict_fc<2>(r, b3, fc, cb3, sp.p[3][y]);
}

// ProlongateCellCenteredValues + ProlongateSharedFieldX1..3: lists PROLONG_CC, FC1..3
__global__ void __launch_bounds__(RBS)
k_ref_prolong_all(RG r, int nvar, const double *__restrict__ cw, const double *__restrict__ cb1,
                  const double *__restrict__ cb2, const double *__restrict__ cb3, agx_ref_coords cc,
                  double *__restrict__ w, double *__restrict__ b1, double *__restrict__ b2, double *__restrict__ b3,
                  agx_ref_coords fc, ItemSpan sp) {
  int y = blockIdx.y;
  const int l = span_pick(sp, y);
  if (l == 0) b_ref_prolong_cc(r, nvar, cw, cc, w, fc, sp.p[0][y]);
  else if (l == 1) b_ref_prolong_fc_shared<0>(r, cb1, cc, b1, fc, sp.p[1][y]);
  else if (l == 2) b_ref_prolong_fc_shared<1>(r, cb2, cc, b2, fc, sp.p[2][y]);
  else if (l == 3) b_ref_prolong_fc_shared<2>(r, cb3, cc, b3, fc, sp.p[3][y]);
}

// ProlongateInternalField, then CalculateCellCenteredField + PrimitiveToConserved, per
// 2x2x2 group: the thread that forms a group's 12 inner faces then converts its 8 cells,
// whose other faces are the group's outer (shared) faces, final before this launch.
template <bool ISO>
__global__ void __launch_bounds__(RBS)
k_ref_internal_p2c(RG r, double *__restrict__ b1, double *__restrict__ b2, double *__restrict__ b3,
                   agx_ref_coords fc, const double *__restrict__ w, double *__restrict__ bcc,
                   double *__restrict__ u, double igm1, const agx_ref_item *__restrict__ items) {
  const agx_ref_item it = items[blockIdx.y];
  const Range q = item_range(it);
  const unsigned cnt = (unsigned)q.w1 * q.w2 * q.w3;
  const int fb = it.fb;
  const double *d1 = fc.dx1f + (int64_t)fb * r.n1, *d2 = fc.dx2f + (int64_t)fb * r.n2,
               *d3 = fc.dx3f + (int64_t)fb * r.n3;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int i = q.si + (int)(t % (unsigned)q.w1);
    const unsigned s = t / (unsigned)q.w1;
    const int j = q.sj + (int)(s % (unsigned)q.w2), k = q.sk + (int)(s / (unsigned)q.w2);
    internal_group(r, b1, b2, b3, d1, d2, d3, fb, i, j, k);
    const int fi = (i - r.cis) * 2 + r.is, fj = (j - r.cjs) * 2 + r.js, fk = (k - r.cks) * 2 + r.ks;
#pragma unroll
    for (int c = 0; c < 8; ++c)
      p2c_cell<ISO>(r, w, b1, b2, b3, bcc, u, igm1, fb, fi + (c & 1), fj + ((c >> 1) & 1), fk + (c >> 2));
  }
}

// ---------------------------------------------------------------- flux correction
// LoadFluxBoundaryBufferToCoarser (flux_correction_cc.cpp:77-169) on the finer block and
// SetFluxBoundaryFromFiner (:240-283) on the coarser one, as one direct write: value
// (n, a, b) of the reference's buffer is the finer block's average over fine faces
// (2a, 2b) + {0,1}^2 of its face opposite `fid`, stored at coarse face (lo_a + a, lo_b + b).
__global__ void __launch_bounds__(RBS)
k_ref_flux_from_finer(RG r, int nvar, const double *__restrict__ ff1, const double *__restrict__ ff2,
                      const double *__restrict__ ff3, agx_ref_coords fc, double *__restrict__ cf1,
                      double *__restrict__ cf2, double *__restrict__ cf3, const agx_ref_item *__restrict__ items) {
  const agx_ref_item it = items[blockIdx.y];
  const int dir = it.fid >> 1, side = it.fid & 1;
  const int na = dir == 2 ? r.h2 : r.h3;   // slow index of the face: k (x1, x2 faces) or j (x3)
  const int nbb = dir == 0 ? r.h2 : r.h1;  // fast index: j (x1 face) or i
  const unsigned per = (unsigned)na * nbb, cnt = per * nvar;
  const double *d1 = fc.dx1f + (int64_t)it.fb * r.n1, *d2 = fc.dx2f + (int64_t)it.fb * r.n2,
               *d3 = fc.dx3f + (int64_t)it.fb * r.n3;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int n = (int)(t / per);
    const unsigned s = t - n * per;
    const int a = (int)(s / (unsigned)nbb), b = (int)(s - (unsigned)a * nbb);
    if (dir == 0) {
      const int ic = side ? r.ie + 1 : r.is, i = side ? r.is : r.ie + 1;
      const int k = r.ks + 2 * a, j = r.js + 2 * b;
      const double amm = d2[j] * d3[k], amp = d2[j + 1] * d3[k], apm = d2[j] * d3[k + 1], app = d2[j + 1] * d3[k + 1];
      const double tarea = amm + amp + apm + app;
      auto FX = [&](int p, int q) { return ff1[bx(r, r.n3, r.n2, r.n1 + 1, nvar, it.fb, n, k + p, j + q, i)]; };
      const double v = (FX(0, 0) * amm + FX(0, 1) * amp + FX(1, 0) * apm + FX(1, 1) * app) / tarea;
      cf1[bx(r, r.n3, r.n2, r.n1 + 1, nvar, it.cb, n, r.ks + (it.fi2 ? r.h3 : 0) + a, r.js + (it.fi1 ? r.h2 : 0) + b, ic)] = v;
    } else if (dir == 1) {
      const int jc = side ? r.je + 1 : r.js, j = side ? r.js : r.je + 1;
      const int k = r.ks + 2 * a, i = r.is + 2 * b;
      const double s00 = d1[i] * d3[k], s01 = d1[i + 1] * d3[k], s10 = d1[i] * d3[k + 1], s11 = d1[i + 1] * d3[k + 1];
      const double tarea = s00 + s01 + s10 + s11;
      auto FX = [&](int p, int q) { return ff2[bx(r, r.n3, r.n2 + 1, r.n1, nvar, it.fb, n, k + p, j, i + q)]; };
      const double v = (FX(0, 0) * s00 + FX(0, 1) * s01 + FX(1, 0) * s10 + FX(1, 1) * s11) / tarea;
      cf2[bx(r, r.n3, r.n2 + 1, r.n1, nvar, it.cb, n, r.ks + (it.fi2 ? r.h3 : 0) + a, jc, r.is + (it.fi1 ? r.h1 : 0) + b)] = v;
    } else {
      const int kc = side ? r.ke + 1 : r.ks, k = side ? r.ks : r.ke + 1;
      const int j = r.js + 2 * a, i = r.is + 2 * b;
      const double s00 = d1[i] * d2[j], s01 = d1[i + 1] * d2[j], s10 = d1[i] * d2[j + 1], s11 = d1[i + 1] * d2[j + 1];
      const double tarea = s00 + s01 + s10 + s11;
      auto FX = [&](int p, int q) { return ff3[bx(r, r.n3 + 1, r.n2, r.n1, nvar, it.fb, n, k, j + p, i + q)]; };
      const double v = (FX(0, 0) * s00 + FX(0, 1) * s01 + FX(1, 0) * s10 + FX(1, 1) * s11) / tarea;
      cf3[bx(r, r.n3 + 1, r.n2, r.n1, nvar, it.cb, n, kc, r.js + (it.fi2 ? r.h2 : 0) + a, r.is + (it.fi1 ? r.h1 : 0) + b)] = v;
    }
  }
}

// number of values of the reference's EMF buffer to a coarser face neighbour (3D):
// two edge components, each restricted over pairs of fine edges
__host__ __device__ inline int64_t emf_part(const RG &r, int dir, int part) {
  if (dir == 0) return part == 0 ? (int64_t)(r.h3 + 1) * r.h2 : (int64_t)r.h3 * (r.h2 + 1);
  if (dir == 1) return part == 0 ? (int64_t)(r.h3 + 1) * r.h1 : (int64_t)r.h3 * (r.h1 + 1);
  return part == 0 ? (int64_t)(r.h2 + 1) * r.h1 : (int64_t)r.h2 * (r.h1 + 1);
}
// values of a to-coarser buffer for face fid 0..5 or edge 6 + eid (one restricted line)
__host__ __device__ inline int64_t emf_coarser_count(const RG &r, int fid) {
  if (fid < 6) return emf_part(r, fid >> 1, 0) + emf_part(r, fid >> 1, 1);
  const int eid = fid - 6;
  return eid < 4 ? r.h3 : (eid < 8 ? r.h2 : r.h1);
}

// LoadFluxBoundaryBufferToCoarser for face neighbours of a 3D block,
// flux_correction_fc.cpp:315-418: the buffer order of the reference
__global__ void __launch_bounds__(RBS)
k_ref_emf_to_coarser(RG r, const double *__restrict__ e1, const double *__restrict__ e2, const double *__restrict__ e3,
                     agx_ref_coords fc, double *__restrict__ buf, int64_t stride,
                     const agx_ref_item *__restrict__ items) {
  const agx_ref_item it = items[blockIdx.y];
  const int dir = it.fid >> 1, side = it.fid & 1, fb = it.fb;
  const double *d1 = fc.dx1f + (int64_t)fb * r.n1, *d2 = fc.dx2f + (int64_t)fb * r.n2,
               *d3 = fc.dx3f + (int64_t)fb * r.n3;
  auto E1 = [&](int k, int j, int i) { return e1[bx(r, r.n3 + 1, r.n2 + 1, r.n1, 1, fb, 0, k, j, i)]; };
  auto E2 = [&](int k, int j, int i) { return e2[bx(r, r.n3 + 1, r.n2, r.n1 + 1, 1, fb, 0, k, j, i)]; };
  auto E3 = [&](int k, int j, int i) { return e3[bx(r, r.n3, r.n2 + 1, r.n1 + 1, 1, fb, 0, k, j, i)]; };
  double *out = buf + (int64_t)blockIdx.y * stride;
  if (it.fid >= 6) {   // edge neighbours, flux_correction_fc.cpp:481-546: one line, pairs along it
    const int eid = it.fid - 6, lo = eid & 1, hi = (eid >> 1) & 1;
    const unsigned cnt = (unsigned)emf_coarser_count(r, it.fid);
    for (unsigned t = first_t(); t < cnt; t += stride_t()) {
      const int s2 = 2 * (int)t;
      double v;
      if (eid < 4) {
        const int i = lo ? r.ie + 1 : r.is, j = hi ? r.je + 1 : r.js, k = r.ks + s2;
        v = (E3(k, j, i) * d3[k] + E3(k + 1, j, i) * d3[k + 1]) / (d3[k] + d3[k + 1]);
      } else if (eid < 8) {
        const int i = lo ? r.ie + 1 : r.is, k = hi ? r.ke + 1 : r.ks, j = r.js + s2;
        v = (E2(k, j, i) * d2[j] + E2(k, j + 1, i) * d2[j + 1]) / (d2[j] + d2[j + 1]);
      } else {
        const int j = lo ? r.je + 1 : r.js, k = hi ? r.ke + 1 : r.ks, i = r.is + s2;
        v = (E1(k, j, i) * d1[i] + E1(k, j, i + 1) * d1[i + 1]) / (d1[i] + d1[i + 1]);
      }
      out[t] = v;
    }
    return;
  }
  const unsigned pa = (unsigned)emf_part(r, dir, 0), cnt = pa + (unsigned)emf_part(r, dir, 1);
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const bool first = t < pa;
    const unsigned s = first ? t : t - pa;
    double v;
    if (dir == 0) {
      const int i = side ? r.ie + 1 : r.is;
      if (first) {   // e2, k = ks..ke+1 step 2, j = js..je step 2
        const int k = r.ks + 2 * (int)(s / (unsigned)r.h2), j = r.js + 2 * (int)(s % (unsigned)r.h2);
        v = (E2(k, j, i) * d2[j] + E2(k, j + 1, i) * d2[j + 1]) / (d2[j] + d2[j + 1]);
      } else {       // e3, k = ks..ke step 2, j = js..je+1 step 2
        const int k = r.ks + 2 * (int)(s / (unsigned)(r.h2 + 1)), j = r.js + 2 * (int)(s % (unsigned)(r.h2 + 1));
        v = (E3(k, j, i) * d3[k] + E3(k + 1, j, i) * d3[k + 1]) / (d3[k] + d3[k + 1]);
      }
    } else if (dir == 1) {
      const int j = side ? r.je + 1 : r.js;
      if (first) {   // e1, k = ks..ke+1 step 2, i = is..ie step 2
        const int k = r.ks + 2 * (int)(s / (unsigned)r.h1), i = r.is + 2 * (int)(s % (unsigned)r.h1);
        v = (E1(k, j, i) * d1[i] + E1(k, j, i + 1) * d1[i + 1]) / (d1[i] + d1[i + 1]);
      } else {       // e3, k = ks..ke step 2, i = is..ie+1 step 2
        const int k = r.ks + 2 * (int)(s / (unsigned)(r.h1 + 1)), i = r.is + 2 * (int)(s % (unsigned)(r.h1 + 1));
        v = (E3(k, j, i) * d3[k] + E3(k + 1, j, i) * d3[k + 1]) / (d3[k] + d3[k + 1]);
      }
    } else {
      const int k = side ? r.ke + 1 : r.ks;
      if (first) {   // e1, j = js..je+1 step 2, i = is..ie step 2
        const int j = r.js + 2 * (int)(s / (unsigned)r.h1), i = r.is + 2 * (int)(s % (unsigned)r.h1);
        v = (E1(k, j, i) * d1[i] + E1(k, j, i + 1) * d1[i + 1]) / (d1[i] + d1[i + 1]);
      } else {       // e2, j = js..je step 2, i = is..ie+1 step 2
        const int j = r.js + 2 * (int)(s / (unsigned)(r.h1 + 1)), i = r.is + 2 * (int)(s % (unsigned)(r.h1 + 1));
        v = (E2(k, j, i) * d2[j] + E2(k, j + 1, i) * d2[j + 1]) / (d2[j] + d2[j + 1]);
      }
    }
    out[t] = v;
  }
}

// ---------------------------------------------------------------- ghost exchange (cells)
// The buffer round trips of CellCenteredBoundaryVariable (bvals/cc/bvals_cc.cpp:201-480)
// as direct box copies: item m copies nvar x [si..ei] x [sj..ej] x [sk..ek] of block (or
// coarse buffer) fb to the box of the same shape starting at (fid, fi1, fi2) = (i, j, k)
// of block (or coarse buffer) cb.  S = source dims, D = destination dims.
__global__ void __launch_bounds__(RBS)
k_ref_copy_cc(Lay sl, Lay dl, int nvar, const double *__restrict__ src, double *__restrict__ dst,
              const agx_ref_item *__restrict__ items) {
  const agx_ref_item it = items[blockIdx.y];
  const Range q = item_range(it);
  const unsigned w1 = q.w1, w2 = q.w2, per = w1 * w2 * q.w3, cnt = per * nvar;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int n = (int)(t / per);
    unsigned s = t - n * per;
    const int a = (int)(s % w1);
    s /= w1;
    const int b = (int)(s % w2), c = (int)(s / w2);
    dst[lx(dl, nvar, it.cb, n, it.fi2 + c, it.fi1 + b, it.fid + a)] = src[lx(sl, nvar, it.fb, n, q.sk + c, q.sj + b, q.si + a)];
  }
}

// ---------------------------------------------------------------- ghost exchange (faces)
// FaceCenteredBoundaryVariable's buffer round trips (bvals/fc/bvals_fc.cpp:344-977): a
// pack item copies component aux & 3 (1..3 = x1f, x2f, x3f) over [si..ei] x [sj..ej] x
// [sk..ek] of block (or coarse buffer) fb into buf + item * stride, in PackData order;
// an unpack item writes buf + fb * stride into the same-shaped range of block (or coarse
// buffer) cb.  Shared faces are read by one neighbour's pack and written by another's
// unpack, so every pack runs before any unpack, as with the reference's buffers.
// UNPACK = false: pack.  COARSE: the face arrays are coarse buffers.
template <bool UNPACK>
__global__ void __launch_bounds__(RBS)
k_ref_fc_move(Lay L, double *__restrict__ f1, double *__restrict__ f2, double *__restrict__ f3,
              double *__restrict__ buf, int64_t stride, const agx_ref_item *__restrict__ items) {
  const agx_ref_item it = items[blockIdx.y];
  const Range q = item_range(it);
  const int c = (it.aux & 3) - 1;
  double *f = c == 0 ? f1 : (c == 1 ? f2 : f3);
  L.d1 += (c == 0);
  L.d2 += (c == 1);
  L.d3 += (c == 2);
  const int b = UNPACK ? it.cb : it.fb;
  double *row = buf + (UNPACK ? (int64_t)it.fb : (int64_t)blockIdx.y) * stride;
  const unsigned w1 = q.w1, w2 = q.w2, cnt = w1 * w2 * q.w3;
  for (unsigned t = first_t(); t < cnt; t += stride_t()) {
    const int i = q.si + (int)(t % w1);
    const unsigned s = t / w1;
    const int j = q.sj + (int)(s % w2), k = q.sk + (int)(s / w2);
    const int64_t o = lx(L, 1, b, 0, k, j, i);
    if (UNPACK) f[o] = row[t];
    else row[t] = f[o];
  }
}

// ---------------------------------------------------------------- EMF correction steps
// The elements an item touches form one or two boxes of one edge array each (e1, e2,
// e3), walked k outer, j, i inner as the reference's loops, so the packed buffers keep
// the reference's order.  Faces are fid 0..5, edges 6 + eid (0..3 x1x2, 4..7 x1x3,
// 8..11 x2x3; bit 0 = upper side along the first direction, bit 1 along the second).
struct EBox {
  int arr;              // 0, 1, 2 = e1, e2, e3
  int k0, j0, i0, nk, nj, ni;
};
struct EBoxes {
  EBox b[2];
  unsigned cnt0, cnt;   // elements of box 0, of both
  double scale;         // AVERAGE: the factor
};

__host__ __device__ inline EBox ebox(int arr, int k0, int k1, int j0, int j1, int i0, int i1) {
  return EBox{arr, k0, j0, i0, k1 - k0 + 1, j1 - j0 + 1, i1 - i0 + 1};
}
__host__ __device__ inline int face_at(const RG &r, int dir, int upper) {
  return dir == 0 ? (upper ? r.ie + 1 : r.is) : (dir == 1 ? (upper ? r.je + 1 : r.js) : (upper ? r.ke + 1 : r.ks));
}

// Boxes of item `it` for an EMF kind:
//  SAME_LEVEL, SET_SAME: LoadFluxBoundaryBufferSameLevel / SetFluxBoundarySameLevel
//    (flux_correction_fc.cpp:51-313, 689-940): the whole face or edge;
//  SET_FINER: SetFluxBoundaryFromFiner (:944-1185): the quadrant fi1/fi2 of a face,
//    the half fi1 of an edge;
//  CLEAR: ClearCoarseFluxBoundary (:1214-1348): inside the face (not on its own edges),
//    or the whole edge;
//  AVERAGE: AverageFluxBoundary (:1355-1543): inside the face (fi1 = 0, a neighbour on
//    the same level), the two lines where the finer blocks meet (fi1 = 1), or the whole
//    edge (fi1 = number of finest-level blocks around it).
__host__ __device__ inline EBoxes emf_boxes(const RG &r, int kind, const agx_ref_item &it) {
  EBoxes o;
  o.scale = 0.5;
  o.b[1] = EBox{0, 0, 0, 0, 0, 0, 0};
  const int fid = it.fid;
  int il = r.is, iu = r.ie, jl = r.js, ju = r.je, kl = r.ks, ku = r.ke;
  const bool finer = kind == AGX_REF_EMF_SET_FINER;
  if (fid < 6) {
    const int dir = fid >> 1, x = face_at(r, dir, fid & 1);
    if (kind == AGX_REF_EMF_CLEAR || (kind == AGX_REF_EMF_AVERAGE && it.fi1 == 0)) {
      if (dir == 0) {
        o.b[0] = ebox(1, r.ks + 1, r.ke, r.js, r.je, x, x);
        o.b[1] = ebox(2, r.ks, r.ke, r.js + 1, r.je, x, x);
      } else if (dir == 1) {
        o.b[0] = ebox(0, r.ks + 1, r.ke, x, x, r.is, r.ie);
        o.b[1] = ebox(2, r.ks, r.ke, x, x, r.is + 1, r.ie);
      } else {
        o.b[0] = ebox(0, x, x, r.js + 1, r.je, r.is, r.ie);
        o.b[1] = ebox(1, x, x, r.js, r.je, r.is + 1, r.ie);
      }
    } else if (kind == AGX_REF_EMF_AVERAGE) {
      if (dir == 0) {
        o.b[0] = ebox(1, r.ks + r.h3, r.ks + r.h3, r.js, r.je, x, x);
        o.b[1] = ebox(2, r.ks, r.ke, r.js + r.h2, r.js + r.h2, x, x);
      } else if (dir == 1) {
        o.b[0] = ebox(0, r.ks + r.h3, r.ks + r.h3, x, x, r.is, r.ie);
        o.b[1] = ebox(2, r.ks, r.ke, x, x, r.is + r.h1, r.is + r.h1);
      } else {
        o.b[0] = ebox(0, x, x, r.js + r.h2, r.js + r.h2, r.is, r.ie);
        o.b[1] = ebox(1, x, x, r.js, r.je, r.is + r.h1, r.is + r.h1);
      }
    } else {
      if (dir == 0) {
        if (finer) {
          if (it.fi1 == 0) ju = r.js + r.h2 - 1; else jl = r.js + r.h2;
          if (it.fi2 == 0) ku = r.ks + r.h3 - 1; else kl = r.ks + r.h3;
        }
        o.b[0] = ebox(1, kl, ku + 1, jl, ju, x, x);
        o.b[1] = ebox(2, kl, ku, jl, ju + 1, x, x);
      } else if (dir == 1) {
        if (finer) {
          if (it.fi1 == 0) iu = r.is + r.h1 - 1; else il = r.is + r.h1;
          if (it.fi2 == 0) ku = r.ks + r.h3 - 1; else kl = r.ks + r.h3;
        }
        o.b[0] = ebox(0, kl, ku + 1, x, x, il, iu);
        o.b[1] = ebox(2, kl, ku, x, x, il, iu + 1);
      } else {
        if (finer) {
          if (it.fi1 == 0) iu = r.is + r.h1 - 1; else il = r.is + r.h1;
          if (it.fi2 == 0) ju = r.js + r.h2 - 1; else jl = r.js + r.h2;
        }
        o.b[0] = ebox(0, x, x, jl, ju + 1, il, iu);
        o.b[1] = ebox(1, x, x, jl, ju, il, iu + 1);
      }
    }
  } else {
    const int eid = fid - 6, lo = eid & 1, hi = (eid >> 1) & 1;
    if (kind == AGX_REF_EMF_AVERAGE) o.scale = 1.0 / (double)it.fi1;
    if (eid < 4) {
      const int i = face_at(r, 0, lo), j = face_at(r, 1, hi);
      if (finer) { if (it.fi1 == 0) ku = r.ks + r.h3 - 1; else kl = r.ks + r.h3; }
      o.b[0] = ebox(2, kl, ku, j, j, i, i);
    } else if (eid < 8) {
      const int i = face_at(r, 0, lo), k = face_at(r, 2, hi);
      if (finer) { if (it.fi1 == 0) ju = r.js + r.h2 - 1; else jl = r.js + r.h2; }
      o.b[0] = ebox(1, k, k, jl, ju, i, i);
    } else {
      const int j = face_at(r, 1, lo), k = face_at(r, 2, hi);
      if (finer) { if (it.fi1 == 0) iu = r.is + r.h1 - 1; else il = r.is + r.h1; }
      o.b[0] = ebox(0, k, k, j, j, il, iu);
    }
  }
  o.cnt0 = (unsigned)(o.b[0].nk * o.b[0].nj * o.b[0].ni);
  o.cnt = o.cnt0 + (unsigned)(o.b[1].nk * o.b[1].nj * o.b[1].ni);
  return o;
}

enum { EOP_PACK = 0, EOP_ADD = 1, EOP_ZERO = 2, EOP_SCALE = 3 };

// One kernel for the four element operations: PACK buf[item] <- e (block fb), ADD
// e (block cb) += buf[fb], ZERO e (block cb) = 0, SCALE e (block cb) *= factor.
template <int OP>
__global__ void __launch_bounds__(RBS)
k_ref_emf_op(RG r, int kind, double *__restrict__ e1, double *__restrict__ e2, double *__restrict__ e3,
             double *__restrict__ buf, int64_t stride, const agx_ref_item *__restrict__ items) {
  const agx_ref_item it = items[blockIdx.y];
  const EBoxes q = emf_boxes(r, kind, it);
  const int b = OP == EOP_PACK ? it.fb : it.cb;
  double *bb = OP == EOP_PACK ? buf + (int64_t)blockIdx.y * stride
                              : (OP == EOP_ADD ? buf + (int64_t)it.fb * stride : nullptr);
  for (unsigned t = first_t(); t < q.cnt; t += stride_t()) {
    const bool second = t >= q.cnt0;
    const EBox &x = second ? q.b[1] : q.b[0];
    unsigned s = second ? t - q.cnt0 : t;
    const int i = x.i0 + (int)(s % (unsigned)x.ni);
    s /= (unsigned)x.ni;
    const int j = x.j0 + (int)(s % (unsigned)x.nj), k = x.k0 + (int)(s / (unsigned)x.nj);
    double *e;
    int64_t off;
    if (x.arr == 0) { e = e1; off = bx(r, r.n3 + 1, r.n2 + 1, r.n1, 1, b, 0, k, j, i); }
    else if (x.arr == 1) { e = e2; off = bx(r, r.n3 + 1, r.n2, r.n1 + 1, 1, b, 0, k, j, i); }
    else { e = e3; off = bx(r, r.n3, r.n2 + 1, r.n1 + 1, 1, b, 0, k, j, i); }
    if (OP == EOP_PACK) bb[t] = e[off];
    else if (OP == EOP_ADD) e[off] += bb[t];
    else if (OP == EOP_ZERO) e[off] = 0.0;
    else e[off] *= q.scale;
  }
}

// ---------------------------------------------------------------- host side
// layout of the coarse buffers (always dense) or of the blocks
Lay side_lay(const RG &r, bool coarse) {
  if (coarse) return Lay{r.c1, r.c2, r.c3, 0, 0, 0};
  return Lay{r.n1, r.n2, r.n3, r.bS, r.bJ, r.bP};
}

bool in(int lo, int hi, int mn, int mx) { return lo <= hi && lo >= mn && hi <= mx; }

// fine index range of coarse range [lo, hi] (f = (c - cs) * 2 + fs), plus `extra` cells
bool fine_in(int lo, int hi, int cs, int fs, int extra, int mx) {
  return in((lo - cs) * 2 + fs, (hi - cs) * 2 + fs + extra, 0, mx);
}

int check_item(const RG &r, int kind, const agx_ref_item &it, int nbf, int nbc, int m) {
  auto bad = [&](const char *what) {
    set_error("refinement work list, item %d: %s", m, what);
    return -40;
  };
  const bool uses_cb = kind != AGX_REF_PROLONG_FC_INT && kind != AGX_REF_EMF_TO_COARSER &&
                       kind != AGX_REF_PRIM_TO_CONS && kind != AGX_REF_EMF_SAME_LEVEL &&
                       kind != AGX_REF_PACK_FC_B && kind != AGX_REF_PACK_FC_C;
  const bool uses_fb = kind != AGX_REF_EMF_CLEAR && kind != AGX_REF_EMF_AVERAGE && kind != AGX_REF_CONS_TO_PRIM;
  if (uses_fb && (it.fb < 0 || it.fb >= nbf)) return bad("fine-side block (or buffer slot) index out of range");
  if (uses_cb && (it.cb < 0 || it.cb >= nbc)) return bad("coarse-side block index out of range");
  const int cs[3] = {r.cis, r.cjs, r.cks}, fs[3] = {r.is, r.js, r.ks};
  const int cn[3] = {r.c1, r.c2, r.c3}, fn[3] = {r.n1, r.n2, r.n3};
  const int lo[3] = {it.si, it.sj, it.sk}, hi[3] = {it.ei, it.ej, it.ek};
  switch (kind) {
    case AGX_REF_RESTRICT_CC:
      for (int d = 0; d < 3; ++d) {
        if (!in(lo[d], hi[d], 0, cn[d] - 1)) return bad("coarse range outside the coarse buffer");
        if (!fine_in(lo[d], hi[d], cs[d], fs[d], 1, fn[d] - 1)) return bad("fine cells outside the block");
      }
      return 0;
    case AGX_REF_RESTRICT_FC1: case AGX_REF_RESTRICT_FC2: case AGX_REF_RESTRICT_FC3: {
      const int dir = kind - AGX_REF_RESTRICT_FC1;
      for (int d = 0; d < 3; ++d) {
        const int e = d == dir;   // one more face than cells along the face normal
        if (!in(lo[d], hi[d], 0, cn[d] - 1 + e)) return bad("coarse range outside the coarse faces");
        if (!fine_in(lo[d], hi[d], cs[d], fs[d], e ? 0 : 1, fn[d] - 1 + e)) return bad("fine faces outside the block");
      }
      return 0;
    }
    case AGX_REF_PROLONG_CC:
      for (int d = 0; d < 3; ++d) {
        if (!in(lo[d] - 1, hi[d] + 1, 0, cn[d] - 1)) return bad("coarse stencil outside the coarse buffer");
        if (!fine_in(lo[d], hi[d], cs[d], fs[d], 1, fn[d] - 1)) return bad("fine cells outside the block");
      }
      return 0;
    case AGX_REF_PROLONG_FC1: case AGX_REF_PROLONG_FC2: case AGX_REF_PROLONG_FC3: {
      const int dir = kind - AGX_REF_PROLONG_FC1;
      for (int d = 0; d < 3; ++d) {
        if (d == dir) {
          if (!in(lo[d], hi[d], 0, cn[d])) return bad("coarse range outside the coarse faces");
          if (!fine_in(lo[d], hi[d], cs[d], fs[d], 0, fn[d])) return bad("fine faces outside the block");
        } else {
          if (!in(lo[d] - 1, hi[d] + 1, 0, cn[d] - 1)) return bad("coarse stencil outside the coarse faces");
          if (!fine_in(lo[d], hi[d], cs[d], fs[d], 1, fn[d] - 1)) return bad("fine faces outside the block");
        }
      }
      return 0;
    }
    case AGX_REF_PROLONG_FC_INT:
      for (int d = 0; d < 3; ++d) {
        // faces fi..fi+2 along d (face dims n+1), cells fi..fi+1 across
        if (!in(lo[d], hi[d], 0, cn[d] - 1)) return bad("coarse range outside the coarse buffer");
        if (!fine_in(lo[d], hi[d], cs[d], fs[d], 2, fn[d])) return bad("fine faces outside the block");
      }
      return 0;
    case AGX_REF_PRIM_TO_CONS:
      for (int d = 0; d < 3; ++d) {
        // fine cells (c - cs)*2 + fs + {0,1}; their upper faces at +2 (face dims n+1)
        if (!in(lo[d], hi[d], 0, cn[d] - 1)) return bad("coarse range outside the coarse buffer");
        if (!fine_in(lo[d], hi[d], cs[d], fs[d], 1, fn[d] - 1)) return bad("fine cells outside the block");
      }
      return 0;
    case AGX_REF_FLUX_FROM_FINER:
      if (it.fid < 0 || it.fid > 5) return bad("fid must be 0..5");
      if ((it.fi1 & ~1) || (it.fi2 & ~1)) return bad("fi1 and fi2 must be 0 or 1");
      return 0;
    case AGX_REF_COPY_CC_BB: case AGX_REF_COPY_CC_CB: case AGX_REF_COPY_CC_BC: {
      const bool src_coarse = kind == AGX_REF_COPY_CC_CB, dst_coarse = kind == AGX_REF_COPY_CC_BC;
      const int dst0[3] = {it.fid, it.fi1, it.fi2};
      for (int d = 0; d < 3; ++d) {
        const int sn = src_coarse ? cn[d] : fn[d], dn = dst_coarse ? cn[d] : fn[d];
        if (!in(lo[d], hi[d], 0, sn - 1)) return bad("source range outside the source array");
        if (!in(dst0[d], dst0[d] + hi[d] - lo[d], 0, dn - 1)) return bad("destination box outside the destination array");
      }
      return 0;
    }
    case AGX_REF_CONS_TO_PRIM:
      for (int d = 0; d < 3; ++d)   // cells and their upper faces (face dims nc + 1)
        if (!in(lo[d], hi[d], 0, cn[d] - 1)) return bad("range outside the coarse buffer");
      return 0;
    case AGX_REF_PACK_FC_B: case AGX_REF_PACK_FC_C: case AGX_REF_UNPACK_FC_B: case AGX_REF_UNPACK_FC_C: {
      const int c = (it.aux & 3) - 1;
      if (c < 0 || (it.aux >> 2) > 15) return bad("aux must be a component 1..3 plus 4 x a launch group 0..15");
      const bool coarse = kind == AGX_REF_PACK_FC_C || kind == AGX_REF_UNPACK_FC_C;
      for (int d = 0; d < 3; ++d) {
        const int nn = (coarse ? cn[d] : fn[d]) + (d == c);
        if (!in(lo[d], hi[d], 0, nn - 1)) return bad("range outside the face array");
      }
      return 0;
    }
    case AGX_REF_EMF_TO_COARSER: case AGX_REF_EMF_SAME_LEVEL: case AGX_REF_EMF_SET_SAME:
    case AGX_REF_EMF_CLEAR:
      if (it.fid < 0 || it.fid > 17) return bad("fid must be a face 0..5 or 6 + an edge 0..11");
      return 0;
    case AGX_REF_EMF_SET_FINER:
      if (it.fid < 0 || it.fid > 17) return bad("fid must be a face 0..5 or 6 + an edge 0..11");
      if ((it.fi1 & ~1) || (it.fi2 & ~1) || (it.fid >= 6 && it.fi2 != 0))
        return bad("fi1 and fi2 must be 0 or 1 (fi2 0 on an edge)");
      return 0;
    case AGX_REF_EMF_AVERAGE:
      if (it.fid < 0 || it.fid > 17) return bad("fid must be a face 0..5 or 6 + an edge 0..11");
      if (it.fid < 6 ? (it.fi1 & ~1) != 0 : (it.fi1 < 2 || it.fi1 > 4))
        return bad("fi1 must be 0 or 1 on a face and the block count 2..4 on an edge");
      return 0;
    default:
      return bad("unknown operator kind");
  }
}

bool is_emf_box_kind(int kind) {
  return kind == AGX_REF_EMF_SAME_LEVEL || kind == AGX_REF_EMF_SET_SAME || kind == AGX_REF_EMF_SET_FINER ||
         kind == AGX_REF_EMF_CLEAR || kind == AGX_REF_EMF_AVERAGE;
}

// Launch order of an EMF_SET item: the reference adds neighbours in the order x1 faces,
// x2 faces, x3 faces, edges (bvals_base.cpp:299-700), and an element on a face's edge
// takes one value from each of these classes, so the classes run one after another.
// Finer quadrants of one face meet on two lines, which two items of the class write;
// each quadrant (or edge half) gets its own launch, so no launch writes an element twice.
int emf_phase(int kind, const agx_ref_item &it) {
  const int cls = it.fid < 6 ? it.fid >> 1 : 3;
  if (kind != AGX_REF_EMF_SET_FINER) return cls * 4;
  return cls * 4 + (it.fid < 6 ? it.fi1 + 2 * it.fi2 : it.fi1);
}

int64_t item_count(const RG &r, int kind, const agx_ref_item &it) {
  if (kind == AGX_REF_FLUX_FROM_FINER) {
    const int dir = it.fid >> 1;
    return (int64_t)(dir == 2 ? r.h2 : r.h3) * (dir == 0 ? r.h2 : r.h1);
  }
  if (kind == AGX_REF_EMF_TO_COARSER) return emf_coarser_count(r, it.fid);
  if (is_emf_box_kind(kind)) return emf_boxes(r, kind, it).cnt;
  if (kind == AGX_REF_PRIM_TO_CONS) return 8 * (int64_t)(it.ei - it.si + 1) * (it.ej - it.sj + 1) * (it.ek - it.sk + 1);
  return (int64_t)(it.ei - it.si + 1) * (it.ej - it.sj + 1) * (it.ek - it.sk + 1);
}

int check_call(const agx_ref_list *l, int kind_lo, int kind_hi, const char *what, int nvar = 1) {
  if (!l) { set_error("%s: null work list", what); return -41; }
  if (l->kind < kind_lo || l->kind > kind_hi) {
    set_error("%s: work list was built for operator kind %d", what, l->kind);
    return -42;
  }
  if (nvar < 1 || nvar > 64 || l->max_count * nvar >= ((int64_t)1 << 31)) {
    set_error("%s: nvar must be 1..64 and items x nvar below 2^31 elements", what);
    return -47;
  }
  return 0;
}

int launch_status(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, what);
}

}  // namespace
}  // namespace agx

using namespace agx;

extern "C" {

int agx_ref_list_create(const agx_ref_geom *g, int32_t kind, const agx_ref_item *items, int32_t nitems,
                        int32_t nb_fine, int32_t nb_coarse, agx_ref_list **out) {
  if (!g || !out || (nitems > 0 && !items) || nitems < 0) { set_error("agx_ref_list_create: null argument"); return -1; }
  *out = nullptr;
  if (kind < 0 || kind >= AGX_REF_KINDS) { set_error("agx_ref_list_create: unknown kind %d", kind); return -43; }
  if (g->nx1 < 2 || g->nx2 < 2 || g->nx3 < 2 || (g->nx1 | g->nx2 | g->nx3) & 1) {
    set_error("refinement operators need 3D blocks with even nx1, nx2, nx3 >= 2"); return -44;
  }
  if (g->ng < 1 || g->cng < 1) { set_error("ng and cng must be >= 1"); return -45; }
  if (g->layout != AGX_REF_LAYOUT_DENSE && g->layout != AGX_REF_LAYOUT_MESH) {
    set_error("agx_ref_geom.layout must be AGX_REF_LAYOUT_DENSE or AGX_REF_LAYOUT_MESH"); return -51;
  }
  if (nitems > 65535) { set_error("at most 65535 items per work list"); return -46; }
  if ((int64_t)(g->nx1 + 2 * g->ng) * (g->nx2 + 2 * g->ng) * (g->nx3 + 2 * g->ng) > (1 << 24)) {
    set_error("refinement operators take blocks of at most 2^24 cells"); return -49;
  }
  const RG r = make_rg(*g);
  int64_t maxc = 0, emf_max = 0;
  for (int m = 0; m < nitems; ++m) {
    int rc = check_item(r, kind, items[m], nb_fine, nb_coarse, m);
    if (rc) return rc;
    const int64_t c = item_count(r, kind, items[m]);
    maxc = std::max(maxc, c);
    if (kind == AGX_REF_EMF_TO_COARSER || is_emf_box_kind(kind) || (kind >= AGX_REF_PACK_FC_B && kind <= AGX_REF_UNPACK_FC_C))
      emf_max = std::max(emf_max, c);
  }
  // EMF steps that write into blocks: one item per (block, face or edge[, quadrant]), so
  // that the items of one launch write disjoint elements
  std::vector<agx_ref_item> sorted(items, items + nitems);
  int32_t phase_off[17] = {0};
  int nphase = 1;
  if (kind == AGX_REF_EMF_SET_SAME || kind == AGX_REF_EMF_SET_FINER || kind == AGX_REF_EMF_CLEAR ||
      kind == AGX_REF_EMF_AVERAGE) {
    std::vector<int64_t> keys(nitems);
    for (int m = 0; m < nitems; ++m) {
      const agx_ref_item &it = items[m];
      keys[m] = (((int64_t)it.cb * 18 + it.fid) * 2 + (kind == AGX_REF_EMF_SET_FINER ? it.fi1 : 0)) * 2 +
                (kind == AGX_REF_EMF_SET_FINER ? it.fi2 : 0);
    }
    std::vector<int64_t> k2 = keys;
    std::sort(k2.begin(), k2.end());
    if (std::adjacent_find(k2.begin(), k2.end()) != k2.end()) {
      set_error("refinement work list: two items for the same block, face or edge (and quadrant)");
      return -40;
    }
    if (kind == AGX_REF_EMF_SET_SAME || kind == AGX_REF_EMF_SET_FINER) {
      nphase = 16;
      std::stable_sort(sorted.begin(), sorted.end(), [&](const agx_ref_item &a, const agx_ref_item &b) {
        return emf_phase(kind, a) < emf_phase(kind, b);
      });
      int m = 0;
      for (int p = 0; p < nphase; ++p) {
        phase_off[p] = m;
        while (m < nitems && emf_phase(kind, sorted[m]) == p) ++m;
      }
      phase_off[nphase] = nitems;
    }
  }
  if (kind == AGX_REF_UNPACK_FC_B || kind == AGX_REF_UNPACK_FC_C) {
    // launch groups in increasing aux >> 2: the caller's neighbour order (faces, edges,
    // corners; finer quadrants in turn), for elements two neighbours both write
    nphase = 16;
    std::stable_sort(sorted.begin(), sorted.end(),
                     [](const agx_ref_item &a, const agx_ref_item &b) { return (a.aux >> 2) < (b.aux >> 2); });
    int m = 0;
    for (int p = 0; p < nphase; ++p) {
      phase_off[p] = m;
      while (m < nitems && (sorted[m].aux >> 2) == p) ++m;
    }
    phase_off[nphase] = nitems;
  }
  if (nphase == 1) phase_off[1] = nitems;
  agx_ref_list *l = new agx_ref_list;
  l->g = *g;
  l->kind = kind;
  l->nitems = nitems;
  l->emf_max = emf_max;
  l->max_count = maxc;
  l->nphase = nphase;
  std::copy(phase_off, phase_off + 17, l->phase_off);
  const int nvar_hint = (kind == AGX_REF_RESTRICT_CC || kind == AGX_REF_PROLONG_CC ||
                         kind == AGX_REF_FLUX_FROM_FINER || kind == AGX_REF_COPY_CC_BB ||
                         kind == AGX_REF_COPY_CC_CB || kind == AGX_REF_COPY_CC_BC) ? 5 : 1;
  l->G = (int)std::min<int64_t>(64, std::max<int64_t>(1, (maxc * nvar_hint + RBS - 1) / RBS));
  if (nitems > 0) {
    hipError_t e = hipMalloc(&l->items, sizeof(agx_ref_item) * nitems);
    if (e == hipSuccess) e = hipMemcpy(l->items, sorted.data(), sizeof(agx_ref_item) * nitems, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (l->items) hipFree(l->items);
      delete l;
      return hip_fail(e, "agx_ref_list_create");
    }
  }
  *out = l;
  return 0;
}

int agx_ref_list_destroy(agx_ref_list *l) {
  if (!l) return 0;
  if (l->items) hipFree(l->items);
  delete l;
  return 0;
}

int64_t agx_ref_emf_count(const agx_ref_geom *g, int32_t fid) {
  if (!g || fid < 0 || fid > 17) return -1;
  return emf_coarser_count(make_rg(*g), fid);
}

int64_t agx_ref_emf_count_same_level(const agx_ref_geom *g, int32_t fid) {
  if (!g || fid < 0 || fid > 17) return -1;
  agx_ref_item it{};
  it.fid = fid;
  return emf_boxes(make_rg(*g), AGX_REF_EMF_SAME_LEVEL, it).cnt;
}

int agx_ref_count_fine_edges(const int32_t nblevel[27], int32_t edge_flag[12], int32_t nedge_fine[12]) {
  if (!nblevel || !edge_flag || !nedge_fine) { set_error("agx_ref_count_fine_edges: null argument"); return -1; }
  if (nblevel[13] != 0) { set_error("agx_ref_count_fine_edges: nblevel[13] (the block itself) must be 0"); return -51; }
  auto lv = [&](int o3, int o2, int o1) { return nblevel[((o3 + 1) * 3 + o2 + 1) * 3 + o1 + 1]; };
  int eid = 0;
  // x1x2 edges (ox2 outer, ox1 inner), x1x3 (ox3, ox1), x2x3 (ox3, ox2): the four
  // positions around each edge; a finer one restarts the count at its level
  for (int plane = 0; plane < 3; ++plane)
    for (int a = -1; a <= 1; a += 2)
      for (int c = -1; c <= 1; c += 2) {
        int nf = 0, fl = 0;
        for (int u = std::max(a - 1, -1); u <= std::min(a + 1, 1); ++u)
          for (int v = std::max(c - 1, -1); v <= std::min(c + 1, 1); ++v) {
            const int l = plane == 0 ? lv(0, u, v) : (plane == 1 ? lv(u, 0, v) : lv(u, v, 0));
            if (l > 1) { set_error("agx_ref_count_fine_edges: levels differ by more than one"); return -51; }
            if (l > fl) { ++fl; nf = 0; }
            if (l == fl) ++nf;
          }
        edge_flag[eid] = fl == 0;
        nedge_fine[eid++] = nf;
      }
  return 0;
}

int agx_ref_restrict_cc(const agx_ref_list *l, int32_t nvar, const double *fine, agx_ref_coords fc,
                        double *coarse, void *stream) {
  if (int rc = check_call(l, AGX_REF_RESTRICT_CC, AGX_REF_RESTRICT_CC, "agx_ref_restrict_cc", nvar)) return rc;
  if (l->nitems == 0) return 0;
  hipLaunchKernelGGL(k_ref_restrict_cc, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     make_rg(l->g), nvar, fine, fc, coarse, l->items);
  return launch_status("agx_ref_restrict_cc");
}

int agx_ref_restrict_fc(const agx_ref_list *l, const double *fine, agx_ref_coords fc, double *coarse, void *stream) {
  if (int rc = check_call(l, AGX_REF_RESTRICT_FC1, AGX_REF_RESTRICT_FC3, "agx_ref_restrict_fc")) return rc;
  if (l->nitems == 0) return 0;
  const dim3 grid(l->G, l->nitems);
  const RG r = make_rg(l->g);
  const hipStream_t s = (hipStream_t)stream;
  if (l->kind == AGX_REF_RESTRICT_FC1)
    hipLaunchKernelGGL(k_ref_restrict_fc<0>, grid, dim3(RBS), 0, s, r, fine, fc, coarse, l->items);
  else if (l->kind == AGX_REF_RESTRICT_FC2)
    hipLaunchKernelGGL(k_ref_restrict_fc<1>, grid, dim3(RBS), 0, s, r, fine, fc, coarse, l->items);
  else
    hipLaunchKernelGGL(k_ref_restrict_fc<2>, grid, dim3(RBS), 0, s, r, fine, fc, coarse, l->items);
  return launch_status("agx_ref_restrict_fc");
}

int agx_ref_prolong_cc(const agx_ref_list *l, int32_t nvar, const double *coarse, agx_ref_coords cc,
                       double *fine, agx_ref_coords fc, void *stream) {
  if (int rc = check_call(l, AGX_REF_PROLONG_CC, AGX_REF_PROLONG_CC, "agx_ref_prolong_cc", nvar)) return rc;
  if (l->nitems == 0) return 0;
  hipLaunchKernelGGL(k_ref_prolong_cc, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     make_rg(l->g), nvar, coarse, cc, fine, fc, l->items);
  return launch_status("agx_ref_prolong_cc");
}

int agx_ref_prolong_fc_shared(const agx_ref_list *l, const double *coarse, agx_ref_coords cc, double *fine,
                              agx_ref_coords fc, void *stream) {
  if (int rc = check_call(l, AGX_REF_PROLONG_FC1, AGX_REF_PROLONG_FC3, "agx_ref_prolong_fc_shared")) return rc;
  if (l->nitems == 0) return 0;
  const dim3 grid(l->G, l->nitems);
  const RG r = make_rg(l->g);
  const hipStream_t s = (hipStream_t)stream;
  if (l->kind == AGX_REF_PROLONG_FC1)
    hipLaunchKernelGGL(k_ref_prolong_fc_shared<0>, grid, dim3(RBS), 0, s, r, coarse, cc, fine, fc, l->items);
  else if (l->kind == AGX_REF_PROLONG_FC2)
    hipLaunchKernelGGL(k_ref_prolong_fc_shared<1>, grid, dim3(RBS), 0, s, r, coarse, cc, fine, fc, l->items);
  else
    hipLaunchKernelGGL(k_ref_prolong_fc_shared<2>, grid, dim3(RBS), 0, s, r, coarse, cc, fine, fc, l->items);
  return launch_status("agx_ref_prolong_fc_shared");
}

int agx_ref_prolong_fc_internal(const agx_ref_list *l, double *b1f, double *b2f, double *b3f, agx_ref_coords fc,
                                void *stream) {
  if (int rc = check_call(l, AGX_REF_PROLONG_FC_INT, AGX_REF_PROLONG_FC_INT, "agx_ref_prolong_fc_internal"))
    return rc;
  if (l->nitems == 0) return 0;
  hipLaunchKernelGGL(k_ref_prolong_fc_internal, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     make_rg(l->g), b1f, b2f, b3f, fc, l->items);
  return launch_status("agx_ref_prolong_fc_internal");
}

int agx_ref_flux_from_finer(const agx_ref_list *l, int32_t nvar, const double *ff1, const double *ff2,
                            const double *ff3, agx_ref_coords fc, double *cf1, double *cf2, double *cf3,
                            void *stream) {
  if (int rc = check_call(l, AGX_REF_FLUX_FROM_FINER, AGX_REF_FLUX_FROM_FINER, "agx_ref_flux_from_finer", nvar))
    return rc;
  if (l->nitems == 0) return 0;
  hipLaunchKernelGGL(k_ref_flux_from_finer, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     make_rg(l->g), nvar, ff1, ff2, ff3, fc, cf1, cf2, cf3, l->items);
  return launch_status("agx_ref_flux_from_finer");
}

int agx_ref_emf_to_coarser(const agx_ref_list *l, const double *e1, const double *e2, const double *e3,
                           agx_ref_coords fc, double *buf, int64_t stride, void *stream) {
  if (int rc = check_call(l, AGX_REF_EMF_TO_COARSER, AGX_REF_EMF_TO_COARSER, "agx_ref_emf_to_coarser")) return rc;
  if (stride < l->emf_max) {
    set_error("agx_ref_emf_to_coarser: stride %lld below the largest item (%lld values)", (long long)stride,
              (long long)l->emf_max);
    return -48;
  }
  if (l->nitems == 0) return 0;
  hipLaunchKernelGGL(k_ref_emf_to_coarser, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     make_rg(l->g), e1, e2, e3, fc, buf, stride, l->items);
  return launch_status("agx_ref_emf_to_coarser");
}

int agx_ref_prim_to_cons(const agx_ref_list *l, const double *w, const double *b1f, const double *b2f,
                         const double *b3f, double *bcc, double *u, double gamma, int32_t iso, void *stream) {
  if (int rc = check_call(l, AGX_REF_PRIM_TO_CONS, AGX_REF_PRIM_TO_CONS, "agx_ref_prim_to_cons", iso ? 4 : 5))
    return rc;
  if (!iso && !(gamma > 1.0)) { set_error("agx_ref_prim_to_cons: gamma must be > 1"); return -50; }
  if (l->nitems == 0) return 0;
  const double igm1 = iso ? 0.0 : 1.0 / (gamma - 1.0);
  if (iso)
    hipLaunchKernelGGL(k_ref_prim_to_cons<true>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                       make_rg(l->g), w, b1f, b2f, b3f, bcc, u, igm1, l->items);
  else
    hipLaunchKernelGGL(k_ref_prim_to_cons<false>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                       make_rg(l->g), w, b1f, b2f, b3f, bcc, u, igm1, l->items);
  return launch_status("agx_ref_prim_to_cons");
}

// the lists of a fused launch: kinds in order, one geometry, at most 65535 rows
static int fuse_lists(const agx_ref_list *const *ls, const int *kinds, int nvar, const char *what, ItemSpan &sp,
                      int &G, const agx_ref_list *&first) {
  int rows = 0;
  G = 1;
  first = nullptr;
  for (int l = 0; l < 4; ++l) {
    sp.p[l] = nullptr;
    sp.n[l] = 0;
    if (!ls[l]) continue;
    if (int rc = check_call(ls[l], kinds[l], kinds[l], what, nvar)) return rc;
    if (first && (first->g.nx1 != ls[l]->g.nx1 || first->g.nx2 != ls[l]->g.nx2 || first->g.nx3 != ls[l]->g.nx3 ||
                  first->g.ng != ls[l]->g.ng || first->g.cng != ls[l]->g.cng)) {
      set_error("%s: the lists have different geometries", what);
      return -42;
    }
    if (!first) first = ls[l];
    sp.p[l] = ls[l]->items;
    sp.n[l] = ls[l]->nitems;
    rows += ls[l]->nitems;
    G = std::max(G, ls[l]->G);
  }
  if (rows > 65535) { set_error("%s: more than 65535 items in one launch", what); return -46; }
  return rows;
}

int agx_ref_restrict_all(const agx_ref_list *lcc, const agx_ref_list *lf1, const agx_ref_list *lf2,
                         const agx_ref_list *lf3, int32_t nvar, const double *u, const double *b1f, const double *b2f,
                         const double *b3f, agx_ref_coords fc, double *cu, double *cb1f, double *cb2f, double *cb3f,
                         void *stream) {
  const agx_ref_list *ls[4] = {lcc, lf1, lf2, lf3};
  const int kinds[4] = {AGX_REF_RESTRICT_CC, AGX_REF_RESTRICT_FC1, AGX_REF_RESTRICT_FC2, AGX_REF_RESTRICT_FC3};
  ItemSpan sp;
  int G;
  const agx_ref_list *f;
  const int rows = fuse_lists(ls, kinds, nvar, "agx_ref_restrict_all", sp, G, f);
  if (rows <= 0) return rows;
  hipLaunchKernelGGL(k_ref_restrict_all, dim3(G, rows), dim3(RBS), 0, (hipStream_t)stream, make_rg(f->g), nvar, u,
                     b1f, b2f, b3f, fc, cu, cb1f, cb2f, cb3f, sp);
  return launch_status("agx_ref_restrict_all");
}

int agx_ref_prolong_all(const agx_ref_list *lcc, const agx_ref_list *lf1, const agx_ref_list *lf2,
                        const agx_ref_list *lf3, int32_t nvar, const double *cw, const double *cb1f, const double *cb2f,
                        const double *cb3f, agx_ref_coords cc, double *w, double *b1f, double *b2f, double *b3f,
                        agx_ref_coords fc, void *stream) {
  const agx_ref_list *ls[4] = {lcc, lf1, lf2, lf3};
  const int kinds[4] = {AGX_REF_PROLONG_CC, AGX_REF_PROLONG_FC1, AGX_REF_PROLONG_FC2, AGX_REF_PROLONG_FC3};
  ItemSpan sp;
  int G;
  const agx_ref_list *f;
  const int rows = fuse_lists(ls, kinds, nvar, "agx_ref_prolong_all", sp, G, f);
  if (rows <= 0) return rows;
  hipLaunchKernelGGL(k_ref_prolong_all, dim3(G, rows), dim3(RBS), 0, (hipStream_t)stream, make_rg(f->g), nvar, cw,
                     cb1f, cb2f, cb3f, cc, w, b1f, b2f, b3f, fc, sp);
  return launch_status("agx_ref_prolong_all");
}

int agx_ref_internal_to_cons(const agx_ref_list *l, double *b1f, double *b2f, double *b3f, agx_ref_coords fc,
                             const double *w, double *bcc, double *u, double gamma, int32_t iso, void *stream) {
  if (int rc = check_call(l, AGX_REF_PROLONG_FC_INT, AGX_REF_PROLONG_FC_INT, "agx_ref_internal_to_cons", iso ? 4 : 5))
    return rc;
  if (!iso && !(gamma > 1.0)) { set_error("agx_ref_internal_to_cons: gamma must be > 1"); return -50; }
  if (l->nitems == 0) return 0;
  const double igm1 = iso ? 0.0 : 1.0 / (gamma - 1.0);
  if (iso)
    hipLaunchKernelGGL(k_ref_internal_p2c<true>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                       make_rg(l->g), b1f, b2f, b3f, fc, w, bcc, u, igm1, l->items);
  else
    hipLaunchKernelGGL(k_ref_internal_p2c<false>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                       make_rg(l->g), b1f, b2f, b3f, fc, w, bcc, u, igm1, l->items);
  return launch_status("agx_ref_internal_to_cons");
}

int agx_ref_cons_to_prim(const agx_ref_list *l, double *u, const double *b1f, const double *b2f, const double *b3f,
                         double *w, double *bcc, double gamma, double dfloor, double pfloor, int32_t iso,
                         void *stream) {
  if (int rc = check_call(l, AGX_REF_CONS_TO_PRIM, AGX_REF_CONS_TO_PRIM, "agx_ref_cons_to_prim", iso ? 4 : 5))
    return rc;
  if (!iso && !(gamma > 1.0)) { set_error("agx_ref_cons_to_prim: gamma must be > 1"); return -50; }
  if (l->nitems == 0) return 0;
  const RG r = make_rg(l->g);
  if (iso)
    hipLaunchKernelGGL(k_ref_cons_to_prim<true>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream, r, u, b1f,
                       b2f, b3f, w, bcc, 0.0, dfloor, pfloor, l->items);
  else
    hipLaunchKernelGGL(k_ref_cons_to_prim<false>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream, r, u,
                       b1f, b2f, b3f, w, bcc, gamma - 1.0, dfloor, pfloor, l->items);
  return launch_status("agx_ref_cons_to_prim");
}

// BoundaryValues::ProlongateBoundaries loop limits (bvals_refine.cpp:172-199) and the face
// ranges of ProlongateGhostCells: a face shared with a neighbour at the same or a finer
// level is not prolongated (that neighbour supplies it).
int agx_ref_prolong_ranges(const agx_ref_geom *g, const int32_t ox[3], const int32_t lx_odd[3],
                           const int32_t nblevel[27], int32_t out[12]) {
  if (!g || !ox || !lx_odd || !nblevel || !out) { set_error("agx_ref_prolong_ranges: null argument"); return -1; }
  for (int d = 0; d < 3; ++d)
    if (ox[d] < -1 || ox[d] > 1 || (lx_odd[d] & ~1)) {
      set_error("agx_ref_prolong_ranges: ox must be -1..1 and lx_odd 0 or 1"); return -51;
    }
  if (nblevel[((ox[2] + 1) * 3 + ox[1] + 1) * 3 + ox[0] + 1] >= 0) {
    set_error("agx_ref_prolong_ranges: the neighbour at ox is not coarser"); return -52;
  }
  const RG r = make_rg(*g);
  const int cn = (g->ng + 1) / 2 + 1 - 1;   // cnghost - 1, meshblock.cpp:83
  const int cs[3] = {r.cis, r.cjs, r.cks};
  const int ce[3] = {r.cis + r.h1 - 1, r.cjs + r.h2 - 1, r.cks + r.h3 - 1};
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    if (ox[d] == 0) {
      lo[d] = cs[d]; hi[d] = ce[d];
      if (lx_odd[d] == 0) hi[d] += cn;
      else lo[d] -= cn;
    } else if (ox[d] > 0) {
      lo[d] = ce[d] + 1; hi[d] = ce[d] + cn;
    } else {
      lo[d] = cs[d] - cn; hi[d] = cs[d] - 1;
    }
  }
  auto lev = [&](int a, int b, int c) { return nblevel[(c * 3 + b) * 3 + a]; };   // (i, j, k) in 0..2
  int fl[3], fh[3];
  for (int d = 0; d < 3; ++d) {
    fl[d] = lo[d];
    fh[d] = hi[d] + 1;
    int a[3] = {ox[0] + 1, ox[1] + 1, ox[2] + 1}, b[3] = {ox[0] + 1, ox[1] + 1, ox[2] + 1};
    a[d] = ox[d];       // the neighbour across the lower face of the range along d
    b[d] = ox[d] + 2;   // and across the upper face
    if (ox[d] >= 0 && lev(a[0], a[1], a[2]) >= 0) fl[d]++;
    if (ox[d] <= 0 && lev(b[0], b[1], b[2]) >= 0) fh[d]--;
  }
  out[0] = lo[0]; out[1] = hi[0]; out[2] = lo[1]; out[3] = hi[1]; out[4] = lo[2]; out[5] = hi[2];
  out[6] = fl[0]; out[7] = fh[0]; out[8] = fl[1]; out[9] = fh[1]; out[10] = fl[2]; out[11] = fh[2];
  return 0;
}

int agx_ref_copy_cc(const agx_ref_list *l, int32_t nvar, const double *src, double *dst, void *stream) {
  if (int rc = check_call(l, AGX_REF_COPY_CC_BB, AGX_REF_COPY_CC_BC, "agx_ref_copy_cc", nvar)) return rc;
  if (l->nitems == 0) return 0;
  const RG r = make_rg(l->g);
  const bool sc = l->kind == AGX_REF_COPY_CC_CB, dc = l->kind == AGX_REF_COPY_CC_BC;
  hipLaunchKernelGGL(k_ref_copy_cc, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream, side_lay(r, sc),
                     side_lay(r, dc), nvar, src, dst, l->items);
  return launch_status("agx_ref_copy_cc");
}

int agx_ref_pack_fc(const agx_ref_list *l, const double *x1f, const double *x2f, const double *x3f, double *buf,
                    int64_t stride, void *stream) {
  if (int rc = check_call(l, AGX_REF_PACK_FC_B, AGX_REF_PACK_FC_C, "agx_ref_pack_fc")) return rc;
  if (stride < l->emf_max) {
    set_error("agx_ref_pack_fc: stride %lld below the largest item (%lld values)", (long long)stride,
              (long long)l->emf_max);
    return -48;
  }
  if (l->nitems == 0) return 0;
  const RG r = make_rg(l->g);
  const bool c = l->kind == AGX_REF_PACK_FC_C;
  hipLaunchKernelGGL(k_ref_fc_move<false>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     side_lay(r, c), const_cast<double *>(x1f),
                     const_cast<double *>(x2f), const_cast<double *>(x3f), buf, stride, l->items);
  return launch_status("agx_ref_pack_fc");
}

int agx_ref_unpack_fc(const agx_ref_list *l, double *x1f, double *x2f, double *x3f, const double *buf, int64_t stride,
                      void *stream) {
  if (int rc = check_call(l, AGX_REF_UNPACK_FC_B, AGX_REF_UNPACK_FC_C, "agx_ref_unpack_fc")) return rc;
  if (stride < l->emf_max) {
    set_error("agx_ref_unpack_fc: stride %lld below the largest item (%lld values)", (long long)stride,
              (long long)l->emf_max);
    return -48;
  }
  const RG r = make_rg(l->g);
  const bool c = l->kind == AGX_REF_UNPACK_FC_C;
  for (int p = 0; p < l->nphase; ++p) {
    const int a = l->phase_off[p], n = l->phase_off[p + 1] - a;
    if (n == 0) continue;
    hipLaunchKernelGGL(k_ref_fc_move<true>, dim3(l->G, n), dim3(RBS), 0, (hipStream_t)stream, side_lay(r, c),
                       x1f, x2f, x3f, const_cast<double *>(buf), stride,
                       l->items + a);
    if (int rc = launch_status("agx_ref_unpack_fc")) return rc;
  }
  return 0;
}

int agx_ref_emf_same_level(const agx_ref_list *l, const double *e1, const double *e2, const double *e3, double *buf,
                           int64_t stride, void *stream) {
  if (int rc = check_call(l, AGX_REF_EMF_SAME_LEVEL, AGX_REF_EMF_SAME_LEVEL, "agx_ref_emf_same_level")) return rc;
  if (stride < l->emf_max) {
    set_error("agx_ref_emf_same_level: stride %lld below the largest item (%lld values)", (long long)stride,
              (long long)l->emf_max);
    return -48;
  }
  if (l->nitems == 0) return 0;
  hipLaunchKernelGGL(k_ref_emf_op<EOP_PACK>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     make_rg(l->g), l->kind, const_cast<double *>(e1), const_cast<double *>(e2),
                     const_cast<double *>(e3), buf, stride, l->items);
  return launch_status("agx_ref_emf_same_level");
}

int agx_ref_emf_set(const agx_ref_list *l, double *e1, double *e2, double *e3, const double *buf, int64_t stride,
                    void *stream) {
  if (int rc = check_call(l, AGX_REF_EMF_SET_SAME, AGX_REF_EMF_SET_FINER, "agx_ref_emf_set")) return rc;
  if (stride < l->emf_max) {
    set_error("agx_ref_emf_set: stride %lld below the largest item (%lld values)", (long long)stride,
              (long long)l->emf_max);
    return -48;
  }
  const RG r = make_rg(l->g);
  for (int p = 0; p < l->nphase; ++p) {
    const int a = l->phase_off[p], n = l->phase_off[p + 1] - a;
    if (n == 0) continue;
    hipLaunchKernelGGL(k_ref_emf_op<EOP_ADD>, dim3(l->G, n), dim3(RBS), 0, (hipStream_t)stream, r, l->kind, e1, e2,
                       e3, const_cast<double *>(buf), stride, l->items + a);
    if (int rc = launch_status("agx_ref_emf_set")) return rc;
  }
  return 0;
}

int agx_ref_emf_clear(const agx_ref_list *l, double *e1, double *e2, double *e3, void *stream) {
  if (int rc = check_call(l, AGX_REF_EMF_CLEAR, AGX_REF_EMF_CLEAR, "agx_ref_emf_clear")) return rc;
  if (l->nitems == 0) return 0;
  hipLaunchKernelGGL(k_ref_emf_op<EOP_ZERO>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     make_rg(l->g), l->kind, e1, e2, e3, nullptr, (int64_t)0, l->items);
  return launch_status("agx_ref_emf_clear");
}

int agx_ref_emf_average(const agx_ref_list *l, double *e1, double *e2, double *e3, void *stream) {
  if (int rc = check_call(l, AGX_REF_EMF_AVERAGE, AGX_REF_EMF_AVERAGE, "agx_ref_emf_average")) return rc;
  if (l->nitems == 0) return 0;
  hipLaunchKernelGGL(k_ref_emf_op<EOP_SCALE>, dim3(l->G, l->nitems), dim3(RBS), 0, (hipStream_t)stream,
                     make_rg(l->g), l->kind, e1, e2, e3, nullptr, (int64_t)0, l->items);
  return launch_status("agx_ref_emf_average");
}

int agx_stream_create(void **out) {
  if (!out) { set_error("agx_stream_create: null argument"); return -1; }
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);   // default priority
  if (e != hipSuccess) return hip_fail(e, "agx_stream_create");
  *out = st;
  return 0;
}

int agx_stream_destroy(void *stream) {
  if (!stream) return 0;
  hipError_t e = hipStreamDestroy((hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "agx_stream_destroy");
}

int agx_graph_begin(void *stream) {
  if (!stream) { set_error("agx_graph_begin: capture needs a stream from agx_stream_create"); return -1; }
  hipError_t e = hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
  return e == hipSuccess ? 0 : hip_fail(e, "agx_graph_begin");
}

int agx_graph_end(void *stream, void **exec_out) {
  if (!stream || !exec_out) { set_error("agx_graph_end: null argument"); return -1; }
  *exec_out = nullptr;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &graph);
  if (e != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    return hip_fail(e, "agx_graph_end (capture)");
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(e, "agx_graph_end (instantiate)");
  *exec_out = exec;
  return 0;
}

int agx_graph_launch(void *exec, void *stream) {
  if (!exec) { set_error("agx_graph_launch: null graph"); return -1; }
  hipError_t e = hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "agx_graph_launch");
}

int agx_graph_destroy(void *exec) {
  if (!exec) return 0;
  hipError_t e = hipGraphExecDestroy((hipGraphExec_t)exec);
  return e == hipSuccess ? 0 : hip_fail(e, "agx_graph_destroy");
}

int agx_ref_sync(void *stream) {
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "agx_ref_sync");
}

int agx_device_alloc(int64_t bytes, void **out) {
  if (!out || bytes < 0) { set_error("agx_device_alloc: bad argument"); return -1; }
  *out = nullptr;
  if (bytes == 0) return 0;
  hipError_t e = hipMalloc(out, (size_t)bytes);
  if (e != hipSuccess) { *out = nullptr; return hip_fail(e, "agx_device_alloc"); }
  return 0;
}

int agx_device_free(void *p) {
  if (!p) return 0;
  hipError_t e = hipFree(p);
  return e == hipSuccess ? 0 : hip_fail(e, "agx_device_free");
}

int agx_device_copy(void *dst, const void *src, int64_t bytes, int32_t kind) {
  if (bytes < 0 || kind < 0 || kind > 2 || (bytes > 0 && (!dst || !src))) {
    set_error("agx_device_copy: bad argument"); return -1;
  }
  if (bytes == 0) return 0;
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  hipError_t e = hipMemcpy(dst, src, (size_t)bytes, k);
  return e == hipSuccess ? 0 : hip_fail(e, "agx_device_copy");
}

}  // extern "C"
