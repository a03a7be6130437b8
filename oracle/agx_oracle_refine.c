/* agx_oracle_refine.c -- CPU restatement of the static-mesh-refinement operators
 * (test infrastructure: the checker of the product's refinement kernels, never shipped
 * or called by the product).
 *
 * Each function restates one reference routine for 3D Cartesian blocks, in plain
 * FP64 C with the reference's operand order, over a list of work items:
 *   orc_ref_restrict_cc        MeshRefinement::RestrictCellCenteredValues
 *                              (mesh/mesh_refinement.cpp:106-136)
 *   orc_ref_restrict_fc        RestrictFieldX1/X2/X3 (mesh_refinement.cpp:185-214,
 *                              244-279, 321-344)
 *   orc_ref_prolong_cc         ProlongateCellCenteredValues (mesh_refinement.cpp:386-458)
 *   orc_ref_prolong_fc_shared  ProlongateSharedFieldX1/X2/X3 (mesh_refinement.cpp:541-581,
 *                              626-674, 718-762)
 *   orc_ref_prolong_fc_internal ProlongateInternalField (mesh_refinement.cpp:831-946)
 *   orc_ref_flux_from_finer    CellCenteredBoundaryVariable::LoadFluxBoundaryBufferToCoarser
 *                              + SetFluxBoundaryFromFiner (bvals/cc/flux_correction_cc.cpp:
 *                              77-169, 230-283), the buffer replaced by a direct write
 *   orc_ref_emf_to_coarser     FaceCenteredBoundaryVariable::LoadFluxBoundaryBufferToCoarser,
 *                              face and edge neighbours of a 3D block
 *                              (bvals/fc/flux_correction_fc.cpp:315-546), packed in the
 *                              reference's buffer order
 *   orc_ref_copy_cc            the pack + unpack of the cell-centred ghost exchanges
 *                              (bvals/cc/bvals_cc.cpp:201-479)
 *   orc_ref_cons_to_prim       the coarse-buffer ConservedToPrimitive of
 *                              ApplyPhysicalBoundariesOnCoarseLevel (bvals/bvals_refine.cpp:336-384)
 *   orc_ref_pack_fc, _unpack_fc the pack and unpack of the face-field ghost exchanges
 *                              (bvals/fc/bvals_fc.cpp:344-977)
 *   orc_ref_emf_same_level     LoadFluxBoundaryBufferSameLevel (flux_correction_fc.cpp:51-313)
 *   orc_ref_emf_set            SetFluxBoundarySameLevel / SetFluxBoundaryFromFiner
 *                              (flux_correction_fc.cpp:689-1185)
 *   orc_ref_emf_clear          ClearCoarseFluxBoundary (flux_correction_fc.cpp:1214-1348)
 *   orc_ref_emf_average        AverageFluxBoundary (flux_correction_fc.cpp:1355-1543)
 *   orc_ref_count_fine_edges   CountFineEdges (bvals/fc/bvals_fc.cpp:1128-1191)
 *   orc_ref_prim_to_cons       Field::CalculateCellCenteredField (uniform branch, field/field.cpp:
 *                              112-176) + EquationOfState::PrimitiveToConserved
 *                              (eos/adiabatic_mhd.cpp:99-136, isothermal_mhd.cpp:82-108)
 *   orc_ref_prolong_ranges     the loop limits of BoundaryValues::ProlongateBoundaries and
 *                              ProlongateGhostCells (bvals/bvals_refine.cpp:172-199, face
 *                              ranges of ProlongateGhostCells)
 *
 * Layouts (dense, AthenaArray order, block-major): cells (nb, nvar, n3, n2, n1); faces
 * x1f (nb, n3, n2, n1+1), x2f (nb, n3, n2+1, n1), x3f (nb, n3+1, n2, n1); hydro fluxes
 * as faces with nvar; edges x1e (nb, n3+1, n2+1, n1), x2e (nb, n3+1, n2, n1+1),
 * x3e (nb, n3, n2+1, n1+1).  Fine blocks have n = nx + 2 ng cells per direction, the
 * coarse buffers nc = nx/2 + 2 cng with their interior from cng (the reference: cng =
 * NGHOST, mesh/meshblock.cpp:59-85).  Coordinates are per block:
 * x?v and dx?f of length n (fine) or nc (coarse); on Cartesian blocks the face-area
 * positions x?s? equal x?v (coordinates/cartesian.cpp:106-119), areas and volumes are
 * products of dx?f (coordinates/coordinates.cpp:436-537) and edge lengths are dx?f
 * (coordinates.cpp:367-373).
 */
#include <math.h>
#include <stdint.h>

#include "agx_oracle.h"

#define REF_SIGN(x) (((x) < 0.0) ? -1.0 : 1.0)   /* SIGN, defs.hpp.in:117 */

static inline double minmod_slope(double gm, double gp) {
  return 0.5 * (REF_SIGN(gm) + REF_SIGN(gp)) * fmin(fabs(gm), fabs(gp));
}

/* element (b, n, k, j, i) of a block-major array of nvar variables of d3 x d2 x d1 */
static inline int64_t ix(int64_t d3, int64_t d2, int64_t d1, int64_t nvar, int64_t b, int64_t n,
                         int64_t k, int64_t j, int64_t i) {
  return ((((b * nvar + n) * d3 + k) * d2 + j) * d1) + i;
}

typedef struct {
  int n1, n2, n3;     /* fine block cells incl. ghosts */
  int c1, c2, c3;     /* coarse buffer cells incl. ghosts */
  int is, js, ks;     /* fine interior start (= ng) */
  int ie, je, ke;
  int cis, cjs, cks;  /* coarse interior start (= cng) */
} rgeo;

static rgeo mkgeo(const orc_ref_geom *g) {
  rgeo r;
  r.n1 = g->nx1 + 2 * g->ng; r.n2 = g->nx2 + 2 * g->ng; r.n3 = g->nx3 + 2 * g->ng;
  r.c1 = g->nx1 / 2 + 2 * g->cng; r.c2 = g->nx2 / 2 + 2 * g->cng; r.c3 = g->nx3 / 2 + 2 * g->cng;
  r.is = r.js = r.ks = g->ng;
  r.ie = g->ng + g->nx1 - 1; r.je = g->ng + g->nx2 - 1; r.ke = g->ng + g->nx3 - 1;
  r.cis = r.cjs = r.cks = g->cng;
  return r;
}

/* ---------------------------------------------------------------- restriction */
void orc_ref_restrict_cc(const orc_ref_geom *g, int nvar, const double *fine, orc_ref_coords fc,
                         double *coarse, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  for (int m = 0; m < nit; ++m) {
    const int fb = it[m].fb, cb = it[m].cb;
    const double *d1 = fc.dx1f + (int64_t)fb * r.n1, *d2 = fc.dx2f + (int64_t)fb * r.n2,
                 *d3 = fc.dx3f + (int64_t)fb * r.n3;
    for (int n = 0; n < nvar; ++n)
      for (int ck = it[m].sk; ck <= it[m].ek; ++ck) {
        int k = (ck - r.cks) * 2 + r.ks;
        for (int cj = it[m].sj; cj <= it[m].ej; ++cj) {
          int j = (cj - r.cjs) * 2 + r.js;
          for (int ci = it[m].si; ci <= it[m].ei; ++ci) {
            int i = (ci - r.cis) * 2 + r.is;
            /* CellVolume: dx1f(i)*dx2f(j)*dx3f(k); v[a][b][c] at (k+a, j+b, i+c) */
            double v[2][2][2];
            for (int a = 0; a < 2; ++a)
              for (int bb = 0; bb < 2; ++bb)
                for (int c = 0; c < 2; ++c) v[a][bb][c] = d1[i + c] * d2[j + bb] * d3[k + a];
            double tvol = ((v[0][0][0] + v[0][1][0]) + (v[0][0][1] + v[0][1][1])) +
                          ((v[1][0][0] + v[1][1][0]) + (v[1][0][1] + v[1][1][1]));
#define F(a, bb, c) fine[ix(r.n3, r.n2, r.n1, nvar, fb, n, k + (a), j + (bb), i + (c))]
            double s = (((F(0, 0, 0) * v[0][0][0] + F(0, 1, 0) * v[0][1][0]) +
                         (F(0, 0, 1) * v[0][0][1] + F(0, 1, 1) * v[0][1][1])) +
                        ((F(1, 0, 0) * v[1][0][0] + F(1, 1, 0) * v[1][1][0]) +
                         (F(1, 0, 1) * v[1][0][1] + F(1, 1, 1) * v[1][1][1])));
#undef F
            coarse[ix(r.c3, r.c2, r.c1, nvar, cb, n, ck, cj, ci)] = s / tvol;
          }
        }
      }
  }
}

/* dir 0: x1f (fine (n3, n2, n1+1) -> coarse (c3, c2, c1+1)); 1: x2f; 2: x3f */
void orc_ref_restrict_fc(const orc_ref_geom *g, int dir, const double *fine, orc_ref_coords fc,
                         double *coarse, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int f3 = r.n3 + (dir == 2), f2 = r.n2 + (dir == 1), f1 = r.n1 + (dir == 0);
  const int q3 = r.c3 + (dir == 2), q2 = r.c2 + (dir == 1), q1 = r.c1 + (dir == 0);
  for (int m = 0; m < nit; ++m) {
    const int fb = it[m].fb, cb = it[m].cb;
    const double *d1 = fc.dx1f + (int64_t)fb * r.n1, *d2 = fc.dx2f + (int64_t)fb * r.n2,
                 *d3 = fc.dx3f + (int64_t)fb * r.n3;
    for (int ck = it[m].sk; ck <= it[m].ek; ++ck) {
      int k = (ck - r.cks) * 2 + r.ks;
      for (int cj = it[m].sj; cj <= it[m].ej; ++cj) {
        int j = (cj - r.cjs) * 2 + r.js;
        for (int ci = it[m].si; ci <= it[m].ei; ++ci) {
          int i = (ci - r.cis) * 2 + r.is;
#define F(a, bb, c) fine[ix(f3, f2, f1, 1, fb, 0, k + (a), j + (bb), i + (c))]
          double val;
          if (dir == 0) {   /* Face1Area(k, j, i) = dx2f(j)*dx3f(k) */
            double a00 = d2[j] * d3[k], a01 = d2[j + 1] * d3[k];
            double a10 = d2[j] * d3[k + 1], a11 = d2[j + 1] * d3[k + 1];
            double tarea = a00 + a01 + a10 + a11;
            val = (F(0, 0, 0) * a00 + F(0, 1, 0) * a01 + F(1, 0, 0) * a10 + F(1, 1, 0) * a11) / tarea;
          } else if (dir == 1) {   /* Face2Area(k, j, i) = dx1f(i)*dx3f(k) */
            double s00 = d1[i] * d3[k], s01 = d1[i + 1] * d3[k];
            double s10 = d1[i] * d3[k + 1], s11 = d1[i + 1] * d3[k + 1];
            double tarea = s00 + s01 + s10 + s11;
            val = (F(0, 0, 0) * s00 + F(0, 0, 1) * s01 + F(1, 0, 0) * s10 + F(1, 0, 1) * s11) / tarea;
          } else {   /* Face3Area(k, j, i) = dx1f(i)*dx2f(j) */
            double s00 = d1[i] * d2[j], s01 = d1[i + 1] * d2[j];
            double s10 = d1[i] * d2[j + 1], s11 = d1[i + 1] * d2[j + 1];
            double tarea = s00 + s01 + s10 + s11;
            val = (F(0, 0, 0) * s00 + F(0, 0, 1) * s01 + F(0, 1, 0) * s10 + F(0, 1, 1) * s11) / tarea;
          }
#undef F
          coarse[ix(q3, q2, q1, 1, cb, 0, ck, cj, ci)] = val;
        }
      }
    }
  }
}

/* ---------------------------------------------------------------- prolongation */
void orc_ref_prolong_cc(const orc_ref_geom *g, int nvar, const double *coarse, orc_ref_coords cc,
                        double *fine, orc_ref_coords fc, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  for (int m = 0; m < nit; ++m) {
    const int fb = it[m].fb, cb = it[m].cb;
    const double *X1 = cc.x1v + (int64_t)cb * r.c1, *X2 = cc.x2v + (int64_t)cb * r.c2,
                 *X3 = cc.x3v + (int64_t)cb * r.c3;
    const double *x1 = fc.x1v + (int64_t)fb * r.n1, *x2 = fc.x2v + (int64_t)fb * r.n2,
                 *x3 = fc.x3v + (int64_t)fb * r.n3;
    for (int n = 0; n < nvar; ++n)
      for (int k = it[m].sk; k <= it[m].ek; ++k) {
        int fk = (k - r.cks) * 2 + r.ks;
        double dx3m = X3[k] - X3[k - 1], dx3p = X3[k + 1] - X3[k];
        double dx3fm = X3[k] - x3[fk], dx3fp = x3[fk + 1] - X3[k];
        for (int j = it[m].sj; j <= it[m].ej; ++j) {
          int fj = (j - r.cjs) * 2 + r.js;
          double dx2m = X2[j] - X2[j - 1], dx2p = X2[j + 1] - X2[j];
          double dx2fm = X2[j] - x2[fj], dx2fp = x2[fj + 1] - X2[j];
          for (int i = it[m].si; i <= it[m].ei; ++i) {
            int fi = (i - r.cis) * 2 + r.is;
            double dx1m = X1[i] - X1[i - 1], dx1p = X1[i + 1] - X1[i];
            double dx1fm = X1[i] - x1[fi], dx1fp = x1[fi + 1] - X1[i];
#define C(a, bb, c) coarse[ix(r.c3, r.c2, r.c1, nvar, cb, n, k + (a), j + (bb), i + (c))]
            double cv = C(0, 0, 0);
            double g1 = minmod_slope((cv - C(0, 0, -1)) / dx1m, (C(0, 0, 1) - cv) / dx1p);
            double g2 = minmod_slope((cv - C(0, -1, 0)) / dx2m, (C(0, 1, 0) - cv) / dx2p);
            double g3 = minmod_slope((cv - C(-1, 0, 0)) / dx3m, (C(1, 0, 0) - cv) / dx3p);
#undef C
#define FO(a, bb, c) fine[ix(r.n3, r.n2, r.n1, nvar, fb, n, fk + (a), fj + (bb), fi + (c))]
            FO(0, 0, 0) = cv - (g1 * dx1fm + g2 * dx2fm + g3 * dx3fm);
            FO(0, 0, 1) = cv + (g1 * dx1fp - g2 * dx2fm - g3 * dx3fm);
            FO(0, 1, 0) = cv - (g1 * dx1fm - g2 * dx2fp + g3 * dx3fm);
            FO(0, 1, 1) = cv + (g1 * dx1fp + g2 * dx2fp - g3 * dx3fm);
            FO(1, 0, 0) = cv - (g1 * dx1fm + g2 * dx2fm - g3 * dx3fp);
            FO(1, 0, 1) = cv + (g1 * dx1fp - g2 * dx2fm + g3 * dx3fp);
            FO(1, 1, 0) = cv - (g1 * dx1fm - g2 * dx2fp - g3 * dx3fp);
            FO(1, 1, 1) = cv + (g1 * dx1fp + g2 * dx2fp + g3 * dx3fp);
#undef FO
          }
        }
      }
  }
}

/* Shared faces: dir 0 prolongates x1f from the coarse x1f buffer, slopes along x2, x3;
 * dir 1 x2f (slopes x1, x3); dir 2 x3f (slopes x1, x2).  The item range is in coarse
 * face indices. */
void orc_ref_prolong_fc_shared(const orc_ref_geom *g, int dir, const double *coarse,
                               orc_ref_coords cc, double *fine, orc_ref_coords fc,
                               const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int f3 = r.n3 + (dir == 2), f2 = r.n2 + (dir == 1), f1 = r.n1 + (dir == 0);
  const int q3 = r.c3 + (dir == 2), q2 = r.c2 + (dir == 1), q1 = r.c1 + (dir == 0);
  for (int m = 0; m < nit; ++m) {
    const int fb = it[m].fb, cb = it[m].cb;
    const double *X1 = cc.x1v + (int64_t)cb * r.c1, *X2 = cc.x2v + (int64_t)cb * r.c2,
                 *X3 = cc.x3v + (int64_t)cb * r.c3;
    const double *x1 = fc.x1v + (int64_t)fb * r.n1, *x2 = fc.x2v + (int64_t)fb * r.n2,
                 *x3 = fc.x3v + (int64_t)fb * r.n3;
    for (int k = it[m].sk; k <= it[m].ek; ++k) {
      int fk = (k - r.cks) * 2 + r.ks;
      for (int j = it[m].sj; j <= it[m].ej; ++j) {
        int fj = (j - r.cjs) * 2 + r.js;
        for (int i = it[m].si; i <= it[m].ei; ++i) {
          int fi = (i - r.cis) * 2 + r.is;
#define C(a, bb, c) coarse[ix(q3, q2, q1, 1, cb, 0, k + (a), j + (bb), i + (c))]
#define FO(a, bb, c) fine[ix(f3, f2, f1, 1, fb, 0, fk + (a), fj + (bb), fi + (c))]
          double cv = C(0, 0, 0);
          if (dir == 0) {
            double g2 = minmod_slope((cv - C(0, -1, 0)) / (X2[j] - X2[j - 1]),
                                     (C(0, 1, 0) - cv) / (X2[j + 1] - X2[j]));
            double g3 = minmod_slope((cv - C(-1, 0, 0)) / (X3[k] - X3[k - 1]),
                                     (C(1, 0, 0) - cv) / (X3[k + 1] - X3[k]));
            FO(0, 0, 0) = cv - g2 * (X2[j] - x2[fj]) - g3 * (X3[k] - x3[fk]);
            FO(0, 1, 0) = cv + g2 * (x2[fj + 1] - X2[j]) - g3 * (X3[k] - x3[fk]);
            FO(1, 0, 0) = cv - g2 * (X2[j] - x2[fj]) + g3 * (x3[fk + 1] - X3[k]);
            FO(1, 1, 0) = cv + g2 * (x2[fj + 1] - X2[j]) + g3 * (x3[fk + 1] - X3[k]);
          } else if (dir == 1) {
            double g1 = minmod_slope((cv - C(0, 0, -1)) / (X1[i] - X1[i - 1]),
                                     (C(0, 0, 1) - cv) / (X1[i + 1] - X1[i]));
            double g3 = minmod_slope((cv - C(-1, 0, 0)) / (X3[k] - X3[k - 1]),
                                     (C(1, 0, 0) - cv) / (X3[k + 1] - X3[k]));
            FO(0, 0, 0) = cv - g1 * (X1[i] - x1[fi]) - g3 * (X3[k] - x3[fk]);
            FO(0, 0, 1) = cv + g1 * (x1[fi + 1] - X1[i]) - g3 * (X3[k] - x3[fk]);
            FO(1, 0, 0) = cv - g1 * (X1[i] - x1[fi]) + g3 * (x3[fk + 1] - X3[k]);
            FO(1, 0, 1) = cv + g1 * (x1[fi + 1] - X1[i]) + g3 * (x3[fk + 1] - X3[k]);
          } else {
            double g1 = minmod_slope((cv - C(0, 0, -1)) / (X1[i] - X1[i - 1]),
                                     (C(0, 0, 1) - cv) / (X1[i + 1] - X1[i]));
            double g2 = minmod_slope((cv - C(0, -1, 0)) / (X2[j] - X2[j - 1]),
                                     (C(0, 1, 0) - cv) / (X2[j + 1] - X2[j]));
            FO(0, 0, 0) = cv - g1 * (X1[i] - x1[fi]) - g2 * (X2[j] - x2[fj]);
            FO(0, 0, 1) = cv + g1 * (x1[fi + 1] - X1[i]) - g2 * (X2[j] - x2[fj]);
            FO(0, 1, 0) = cv - g1 * (X1[i] - x1[fi]) + g2 * (x2[fj + 1] - X2[j]);
            FO(0, 1, 1) = cv + g1 * (x1[fi + 1] - X1[i]) + g2 * (x2[fj + 1] - X2[j]);
          }
#undef C
#undef FO
        }
      }
    }
  }
}

/* Interior faces of each 2x2x2 group of fine cells from its 24 outer faces; the item
 * range is in coarse cell indices. */
void orc_ref_prolong_fc_internal(const orc_ref_geom *g, double *b1, double *b2, double *b3,
                                 orc_ref_coords fc, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  for (int m = 0; m < nit; ++m) {
    const int fb = it[m].fb;
    const double *d1 = fc.dx1f + (int64_t)fb * r.n1, *d2 = fc.dx2f + (int64_t)fb * r.n2,
                 *d3 = fc.dx3f + (int64_t)fb * r.n3;
    for (int k = it[m].sk; k <= it[m].ek; ++k) {
      int fk = (k - r.cks) * 2 + r.ks;
      for (int j = it[m].sj; j <= it[m].ej; ++j) {
        int fj = (j - r.cjs) * 2 + r.js;
        for (int i = it[m].si; i <= it[m].ei; ++i) {
          int fi = (i - r.cis) * 2 + r.is;
#define B1(a, bb, c) b1[ix(r.n3, r.n2, r.n1 + 1, 1, fb, 0, (a), (bb), (c))]
#define B2(a, bb, c) b2[ix(r.n3, r.n2 + 1, r.n1, 1, fb, 0, (a), (bb), (c))]
#define B3(a, bb, c) b3[ix(r.n3 + 1, r.n2, r.n1, 1, fb, 0, (a), (bb), (c))]
/* areas: A1(a,b,i) = Face1Area(fk+a, fj+b, i) etc. */
#define A1(a, bb, c) (d2[fj + (bb)] * d3[fk + (a)])
#define A2(a, bb, c) (d1[(c)] * d3[fk + (a)])
#define A3(a, bb, c) (d1[(c)] * d2[fj + (bb)])
          double Uxx = 0.0, Vyy = 0.0, Wzz = 0.0, Uxyz = 0.0, Vxyz = 0.0, Wxyz = 0.0;
          for (int jj = 0; jj < 2; ++jj) {
            int js = 2 * jj - 1, fjj = fj + jj, fjp = fj + 2 * jj;
            for (int ii = 0; ii < 2; ++ii) {
              int is = 2 * ii - 1, fii = fi + ii, fip = fi + 2 * ii;
              double x2lo = B2(fk, fjp, fii) * A2(0, 2 * jj, fii);
              double x2hi = B2(fk + 1, fjp, fii) * A2(1, 2 * jj, fii);
              double x3lo = B3(fk, fjj, fii) * A3(0, jj, fii);
              double x3hi = B3(fk + 2, fjj, fii) * A3(2, jj, fii);
              double x1lo = B1(fk, fjj, fip) * A1(0, jj, fip);
              double x1hi = B1(fk + 1, fjj, fip) * A1(1, jj, fip);
              Uxx += is * (js * (x2lo + x2hi) + (x3hi - x3lo));
              Vyy += js * ((x3hi - x3lo) + is * (x1lo + x1hi));
              Wzz += is * (x1hi - x1lo) + js * (x2hi - x2lo);
              Uxyz += is * js * (x1hi - x1lo);
              Vxyz += is * js * (x2hi - x2lo);
              Wxyz += is * js * (x3hi - x3lo);
            }
          }
          double Sdx1 = (d1[fi] + d1[fi + 1]) * (d1[fi] + d1[fi + 1]);
          double Sdx2 = (d2[fj] + d2[fj + 1]) * (d2[fj] + d2[fj + 1]);
          double Sdx3 = (d3[fk] + d3[fk + 1]) * (d3[fk] + d3[fk + 1]);
          Uxx *= 0.125; Vyy *= 0.125; Wzz *= 0.125;
          Uxyz *= 0.125 / (Sdx2 + Sdx3);
          Vxyz *= 0.125 / (Sdx1 + Sdx3);
          Wxyz *= 0.125 / (Sdx1 + Sdx2);
          for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb) {
              double s3 = a ? 1.0 : -1.0, s2 = bb ? 1.0 : -1.0;
              /* x1f(fk+a, fj+bb, fi+1): + Uxx, sign(a) Sdx3*Vxyz, sign(bb) Sdx2*Wxyz */
              double t = 0.5 * (B1(fk + a, fj + bb, fi) * A1(a, bb, fi) +
                                B1(fk + a, fj + bb, fi + 2) * A1(a, bb, fi + 2)) + Uxx;
              t = (s3 > 0) ? t + Sdx3 * Vxyz : t - Sdx3 * Vxyz;
              t = (s2 > 0) ? t + Sdx2 * Wxyz : t - Sdx2 * Wxyz;
              B1(fk + a, fj + bb, fi + 1) = t / A1(a, bb, fi + 1);
            }
          for (int a = 0; a < 2; ++a)
            for (int c = 0; c < 2; ++c) {
              /* x2f(fk+a, fj+1, fi+c): + Vyy, sign(a) Sdx3*Uxyz, sign(c) Sdx1*Wxyz */
              double t = 0.5 * (B2(fk + a, fj, fi + c) * A2(a, 0, fi + c) +
                                B2(fk + a, fj + 2, fi + c) * A2(a, 2, fi + c)) + Vyy;
              t = a ? t + Sdx3 * Uxyz : t - Sdx3 * Uxyz;
              t = c ? t + Sdx1 * Wxyz : t - Sdx1 * Wxyz;
              B2(fk + a, fj + 1, fi + c) = t / A2(a, 1, fi + c);
            }
          for (int bb = 0; bb < 2; ++bb)
            for (int c = 0; c < 2; ++c) {
              /* x3f(fk+1, fj+bb, fi+c): + Wzz, sign(bb) Sdx2*Uxyz, sign(c) Sdx1*Vxyz */
              double t = 0.5 * (B3(fk + 2, fj + bb, fi + c) * A3(2, bb, fi + c) +
                                B3(fk, fj + bb, fi + c) * A3(0, bb, fi + c)) + Wzz;
              t = bb ? t + Sdx2 * Uxyz : t - Sdx2 * Uxyz;
              t = c ? t + Sdx1 * Vxyz : t - Sdx1 * Vxyz;
              B3(fk + 1, fj + bb, fi + c) = t / A3(1, bb, fi + c);
            }
#undef B1
#undef B2
#undef B3
#undef A1
#undef A2
#undef A3
        }
      }
    }
  }
}

/* ---------------------------------------------------------------- flux correction */
/* Item: fb = the finer block, cb = the coarser block, fid = the face of the coarser block
 * that the finer block touches (0 inner_x1 .. 5 outer_x3), fi1/fi2 = its quadrant on that
 * face (NeighborIndexes::fi1/fi2).  The coarser block's face fluxes of that quadrant are
 * replaced by the area-weighted averages of the finer block's fluxes on its opposite
 * face. */
void orc_ref_flux_from_finer(const orc_ref_geom *g, int nvar, const double *ff1, const double *ff2,
                             const double *ff3, orc_ref_coords fc, double *cf1, double *cf2,
                             double *cf3, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int h1 = g->nx1 / 2, h2 = g->nx2 / 2, h3 = g->nx3 / 2;
  for (int m = 0; m < nit; ++m) {
    const int fb = it[m].fb, cb = it[m].cb, fid = it[m].fid, dir = fid >> 1, side = fid & 1;
    const double *d1 = fc.dx1f + (int64_t)fb * r.n1, *d2 = fc.dx2f + (int64_t)fb * r.n2,
                 *d3 = fc.dx3f + (int64_t)fb * r.n3;
    for (int n = 0; n < nvar; ++n) {
      if (dir == 0) {
        int ic = side ? r.ie + 1 : r.is, i = side ? r.is : r.ie + 1;
        int jl = r.js + (it[m].fi1 ? h2 : 0), kl = r.ks + (it[m].fi2 ? h3 : 0);
        for (int kc = 0; kc < h3; ++kc)
          for (int jc = 0; jc < h2; ++jc) {
            int k = r.ks + 2 * kc, j = r.js + 2 * jc;
            double amm = d2[j] * d3[k], amp = d2[j + 1] * d3[k];
            double apm = d2[j] * d3[k + 1], app = d2[j + 1] * d3[k + 1];
            double tarea = amm + amp + apm + app;
#define FX(a, bb) ff1[ix(r.n3, r.n2, r.n1 + 1, nvar, fb, n, k + (a), j + (bb), i)]
            double v = (FX(0, 0) * amm + FX(0, 1) * amp + FX(1, 0) * apm + FX(1, 1) * app) / tarea;
#undef FX
            cf1[ix(r.n3, r.n2, r.n1 + 1, nvar, cb, n, kl + kc, jl + jc, ic)] = v;
          }
      } else if (dir == 1) {
        int jc0 = side ? r.je + 1 : r.js, j = side ? r.js : r.je + 1;
        int il = r.is + (it[m].fi1 ? h1 : 0), kl = r.ks + (it[m].fi2 ? h3 : 0);
        for (int kc = 0; kc < h3; ++kc)
          for (int ic = 0; ic < h1; ++ic) {
            int k = r.ks + 2 * kc, i = r.is + 2 * ic;
            double s00 = d1[i] * d3[k], s01 = d1[i + 1] * d3[k];
            double s10 = d1[i] * d3[k + 1], s11 = d1[i + 1] * d3[k + 1];
            double tarea = s00 + s01 + s10 + s11;
#define FX(a, c) ff2[ix(r.n3, r.n2 + 1, r.n1, nvar, fb, n, k + (a), j, i + (c))]
            double v = (FX(0, 0) * s00 + FX(0, 1) * s01 + FX(1, 0) * s10 + FX(1, 1) * s11) / tarea;
#undef FX
            cf2[ix(r.n3, r.n2 + 1, r.n1, nvar, cb, n, kl + kc, jc0, il + ic)] = v;
          }
      } else {
        int kc0 = side ? r.ke + 1 : r.ks, k = side ? r.ks : r.ke + 1;
        int il = r.is + (it[m].fi1 ? h1 : 0), jl = r.js + (it[m].fi2 ? h2 : 0);
        for (int jc = 0; jc < h2; ++jc)
          for (int ic = 0; ic < h1; ++ic) {
            int j = r.js + 2 * jc, i = r.is + 2 * ic;
            double s00 = d1[i] * d2[j], s01 = d1[i + 1] * d2[j];
            double s10 = d1[i] * d2[j + 1], s11 = d1[i + 1] * d2[j + 1];
            double tarea = s00 + s01 + s10 + s11;
#define FX(bb, c) ff3[ix(r.n3 + 1, r.n2, r.n1, nvar, fb, n, k, j + (bb), i + (c))]
            double v = (FX(0, 0) * s00 + FX(0, 1) * s01 + FX(1, 0) * s10 + FX(1, 1) * s11) / tarea;
#undef FX
            cf3[ix(r.n3 + 1, r.n2, r.n1, nvar, cb, n, kc0, jl + jc, il + ic)] = v;
          }
      }
    }
  }
}

/* Item: fb = the finer block, fid = its face (0..5) or 6 + its edge (NeighborBlock::eid)
 * towards the coarser neighbour.  buf + m * stride receives the restricted edge EMFs of
 * that face or edge in the reference's order. */
void orc_ref_emf_to_coarser(const orc_ref_geom *g, const double *e1, const double *e2,
                            const double *e3, orc_ref_coords fc, double *buf, int64_t stride,
                            const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  for (int m = 0; m < nit; ++m) {
    const int fb = it[m].fb, fid = it[m].fid, dir = fid >> 1, side = fid & 1;
    const double *d1 = fc.dx1f + (int64_t)fb * r.n1, *d2 = fc.dx2f + (int64_t)fb * r.n2,
                 *d3 = fc.dx3f + (int64_t)fb * r.n3;
    double *out = buf + (int64_t)m * stride;
    int64_t p = 0;
#define E1(k, j, i) e1[ix(r.n3 + 1, r.n2 + 1, r.n1, 1, fb, 0, (k), (j), (i))]
#define E2(k, j, i) e2[ix(r.n3 + 1, r.n2, r.n1 + 1, 1, fb, 0, (k), (j), (i))]
#define E3(k, j, i) e3[ix(r.n3, r.n2 + 1, r.n1 + 1, 1, fb, 0, (k), (j), (i))]
    if (dir == 0) {
      int i = side ? r.ie + 1 : r.is;
      for (int k = r.ks; k <= r.ke + 1; k += 2)
        for (int j = r.js; j <= r.je; j += 2) {
          double l1 = d2[j], l2 = d2[j + 1];   /* GetEdge2Length = dx2f(j) */
          out[p++] = (E2(k, j, i) * l1 + E2(k, j + 1, i) * l2) / (l1 + l2);
        }
      for (int k = r.ks; k <= r.ke; k += 2)
        for (int j = r.js; j <= r.je + 1; j += 2) {
          double l1 = d3[k], l2 = d3[k + 1];   /* GetEdge3Length = dx3f(k) */
          out[p++] = (E3(k, j, i) * l1 + E3(k + 1, j, i) * l2) / (l1 + l2);
        }
    } else if (dir == 1) {
      int j = side ? r.je + 1 : r.js;
      for (int k = r.ks; k <= r.ke + 1; k += 2)
        for (int i = r.is; i <= r.ie; i += 2)   /* Edge1Length = dx1f(i) */
          out[p++] = (E1(k, j, i) * d1[i] + E1(k, j, i + 1) * d1[i + 1]) / (d1[i] + d1[i + 1]);
      for (int k = r.ks; k <= r.ke; k += 2)
        for (int i = r.is; i <= r.ie + 1; i += 2)
          out[p++] = (E3(k, j, i) * d3[k] + E3(k + 1, j, i) * d3[k + 1]) / (d3[k] + d3[k + 1]);
    } else if (dir == 2) {
      int k = side ? r.ke + 1 : r.ks;
      for (int j = r.js; j <= r.je + 1; j += 2)
        for (int i = r.is; i <= r.ie; i += 2)
          out[p++] = (E1(k, j, i) * d1[i] + E1(k, j, i + 1) * d1[i + 1]) / (d1[i] + d1[i + 1]);
      for (int j = r.js; j <= r.je; j += 2)
        for (int i = r.is; i <= r.ie + 1; i += 2)
          out[p++] = (E2(k, j, i) * d2[j] + E2(k, j + 1, i) * d2[j + 1]) / (d2[j] + d2[j + 1]);
    } else {   /* edge neighbours, flux_correction_fc.cpp:481-546 */
      const int eid = fid - 6, lo = eid & 1, hi = (eid >> 1) & 1;
      if (eid < 4) {          /* x1x2 edge: e3 along k */
        int i = lo ? r.ie + 1 : r.is, j = hi ? r.je + 1 : r.js;
        for (int k = r.ks; k <= r.ke; k += 2)
          out[p++] = (E3(k, j, i) * d3[k] + E3(k + 1, j, i) * d3[k + 1]) / (d3[k] + d3[k + 1]);
      } else if (eid < 8) {   /* x1x3 edge: e2 along j */
        int i = lo ? r.ie + 1 : r.is, k = hi ? r.ke + 1 : r.ks;
        for (int j = r.js; j <= r.je; j += 2)
          out[p++] = (E2(k, j, i) * d2[j] + E2(k, j + 1, i) * d2[j + 1]) / (d2[j] + d2[j + 1]);
      } else {                /* x2x3 edge: e1 along i */
        int j = lo ? r.je + 1 : r.js, k = hi ? r.ke + 1 : r.ks;
        for (int i = r.is; i <= r.ie; i += 2)
          out[p++] = (E1(k, j, i) * d1[i] + E1(k, j, i + 1) * d1[i + 1]) / (d1[i] + d1[i + 1]);
      }
    }
#undef E1
#undef E2
#undef E3
  }
}

/* ---------------------------------------------------------------- ghost exchange (cells)
 * The pack + unpack of CellCenteredBoundaryVariable's same-level, to-coarser and
 * to-finer exchanges (bvals/cc/bvals_cc.cpp:201-479) as one copy per item: nvar x the
 * source range of block or coarse buffer fb (src_coarse) to the box from (fid, fi1, fi2)
 * = (i, j, k) of block or coarse buffer cb (dst_coarse), in PackData / UnpackData
 * order (n, k, j, i). */
void orc_ref_copy_cc(const orc_ref_geom *g, int nvar, const double *src, int src_coarse, double *dst,
                     int dst_coarse, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int s1 = src_coarse ? r.c1 : r.n1, s2 = src_coarse ? r.c2 : r.n2, s3 = src_coarse ? r.c3 : r.n3;
  const int d1 = dst_coarse ? r.c1 : r.n1, d2 = dst_coarse ? r.c2 : r.n2, d3 = dst_coarse ? r.c3 : r.n3;
  for (int m = 0; m < nit; ++m) {
    for (int n = 0; n < nvar; ++n)
      for (int k = it[m].sk; k <= it[m].ek; ++k)
        for (int j = it[m].sj; j <= it[m].ej; ++j)
          for (int i = it[m].si; i <= it[m].ei; ++i)
            dst[ix(d3, d2, d1, nvar, it[m].cb, n, it[m].fi2 + k - it[m].sk, it[m].fi1 + j - it[m].sj,
                   it[m].fid + i - it[m].si)] = src[ix(s3, s2, s1, nvar, it[m].fb, n, k, j, i)];
  }
}

/* Field::CalculateCellCenteredField (uniform branch, field/field.cpp:112-176) over the
 * item's coarse cells, then EquationOfState::ConservedToPrimitive (eos/adiabatic_mhd.cpp:
 * 41-90; isothermal_mhd.cpp:40-73 without energy) on them, floors written back to u:
 * the call in BoundaryValues::ApplyPhysicalBoundariesOnCoarseLevel (bvals/
 * bvals_refine.cpp:336-384).  Arrays are coarse buffers; items in list order. */
void orc_ref_cons_to_prim(const orc_ref_geom *g, double *u, const double *b1, const double *b2,
                          const double *b3, double *w, double *bcc, double gamma, double dfl,
                          double pfl, int iso, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int nh = iso ? 4 : 5;
  const double gm1 = gamma - 1.0;
  const int64_t S = (int64_t)r.c3 * r.c2 * r.c1;
  for (int m = 0; m < nit; ++m) {
    const int cb = it[m].cb;
    for (int k = it[m].sk; k <= it[m].ek; ++k)
      for (int j = it[m].sj; j <= it[m].ej; ++j)
        for (int i = it[m].si; i <= it[m].ei; ++i) {
          bcc[ix(r.c3, r.c2, r.c1, 3, cb, 0, k, j, i)] = 0.5 * b1[ix(r.c3, r.c2, r.c1 + 1, 1, cb, 0, k, j, i)] +
                                                         0.5 * b1[ix(r.c3, r.c2, r.c1 + 1, 1, cb, 0, k, j, i + 1)];
          bcc[ix(r.c3, r.c2, r.c1, 3, cb, 1, k, j, i)] = 0.5 * b2[ix(r.c3, r.c2 + 1, r.c1, 1, cb, 0, k, j, i)] +
                                                         0.5 * b2[ix(r.c3, r.c2 + 1, r.c1, 1, cb, 0, k, j + 1, i)];
          bcc[ix(r.c3, r.c2, r.c1, 3, cb, 2, k, j, i)] = 0.5 * b3[ix(r.c3 + 1, r.c2, r.c1, 1, cb, 0, k, j, i)] +
                                                         0.5 * b3[ix(r.c3 + 1, r.c2, r.c1, 1, cb, 0, k + 1, j, i)];
        }
    for (int k = it[m].sk; k <= it[m].ek; ++k)
      for (int j = it[m].sj; j <= it[m].ej; ++j)
        for (int i = it[m].si; i <= it[m].ei; ++i) {
          const int64_t o = ix(r.c3, r.c2, r.c1, nh, cb, 0, k, j, i);
          double u_d = u[o];
          const double u_m1 = u[o + S], u_m2 = u[o + 2 * S], u_m3 = u[o + 3 * S];
          u_d = (u_d > dfl) ? u_d : dfl;
          u[o] = u_d;
          w[o] = u_d;
          const double di = 1.0 / u_d;
          w[o + S] = u_m1 * di;
          w[o + 2 * S] = u_m2 * di;
          w[o + 3 * S] = u_m3 * di;
          if (iso) continue;
          const double c1 = bcc[ix(r.c3, r.c2, r.c1, 3, cb, 0, k, j, i)];
          const double c2 = bcc[ix(r.c3, r.c2, r.c1, 3, cb, 1, k, j, i)];
          const double c3 = bcc[ix(r.c3, r.c2, r.c1, 3, cb, 2, k, j, i)];
          const double pb = 0.5 * (c1 * c1 + c2 * c2 + c3 * c3);
          const double e_k = 0.5 * di * (u_m1 * u_m1 + u_m2 * u_m2 + u_m3 * u_m3);
          double u_e = u[o + 4 * S];
          double w_p = gm1 * (u_e - e_k - pb);
          u_e = (w_p > pfl) ? u_e : ((pfl / gm1) + e_k + pb);
          w_p = (w_p > pfl) ? w_p : pfl;
          u[o + 4 * S] = u_e;
          w[o + 4 * S] = w_p;
        }
  }
}

/* ---------------------------------------------------------------- ghost exchange (faces)
 * PackData / UnpackData of FaceCenteredBoundaryVariable (bvals/fc/bvals_fc.cpp:344-977)
 * with the ranges given per item: component aux & 3 (1..3 = x1f, x2f, x3f) of block or
 * coarse buffer (coarse) fb into buf + m * stride (pack), or buf + fb * stride into
 * block or coarse buffer cb (unpack), items in list order. */
static void fc_move(const orc_ref_geom *g, double *f1, double *f2, double *f3, int coarse, double *buf,
                    int64_t stride, int unpack, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int n1 = coarse ? r.c1 : r.n1, n2 = coarse ? r.c2 : r.n2, n3 = coarse ? r.c3 : r.n3;
  for (int m = 0; m < nit; ++m) {
    const int c = (it[m].aux & 3) - 1;
    double *f = c == 0 ? f1 : (c == 1 ? f2 : f3);
    const int d1 = n1 + (c == 0), d2 = n2 + (c == 1), d3 = n3 + (c == 2);
    const int b = unpack ? it[m].cb : it[m].fb;
    double *row = buf + (unpack ? (int64_t)it[m].fb : (int64_t)m) * stride;
    int64_t p = 0;
    for (int k = it[m].sk; k <= it[m].ek; ++k)
      for (int j = it[m].sj; j <= it[m].ej; ++j)
        for (int i = it[m].si; i <= it[m].ei; ++i) {
          double *e = &f[ix(d3, d2, d1, 1, b, 0, k, j, i)];
          if (unpack) *e = row[p++];
          else row[p++] = *e;
        }
  }
}

void orc_ref_pack_fc(const orc_ref_geom *g, double *f1, double *f2, double *f3, int coarse, double *buf,
                     int64_t stride, const orc_ref_item *it, int nit) {
  fc_move(g, f1, f2, f3, coarse, buf, stride, 0, it, nit);
}

void orc_ref_unpack_fc(const orc_ref_geom *g, double *f1, double *f2, double *f3, int coarse, double *buf,
                       int64_t stride, const orc_ref_item *it, int nit) {
  fc_move(g, f1, f2, f3, coarse, buf, stride, 1, it, nit);
}

/* ---------------------------------------------------------------- EMF correction
 * The steps of FaceCenteredBoundaryVariable::SendFluxCorrection / ReceiveFluxCorrection
 * (bvals/fc/flux_correction_fc.cpp) on 3D Cartesian blocks without shearing or polar
 * boundaries.  A face is fid 0..5 (BoundaryFace), an edge 6 + eid (NeighborBlock::eid:
 * 0..3 x1x2, 4..7 x1x3, 8..11 x2x3; bit 0 picks the upper side along the first
 * direction, bit 1 along the second). */
#define EE1(b, k, j, i) e1[ix(r.n3 + 1, r.n2 + 1, r.n1, 1, (b), 0, (k), (j), (i))]
#define EE2(b, k, j, i) e2[ix(r.n3 + 1, r.n2, r.n1 + 1, 1, (b), 0, (k), (j), (i))]
#define EE3(b, k, j, i) e3[ix(r.n3, r.n2 + 1, r.n1 + 1, 1, (b), 0, (k), (j), (i))]

/* position of the face / edge of a block: fixed index along each normal direction */
static inline int face_at(const rgeo *r, int dir, int upper) {
  if (dir == 0) return upper ? r->ie + 1 : r->is;
  if (dir == 1) return upper ? r->je + 1 : r->js;
  return upper ? r->ke + 1 : r->ks;
}

/* LoadFluxBoundaryBufferSameLevel (flux_correction_fc.cpp:51-313), 3D, no shearing box.
 * Item: fb = the block, fid = its face or edge towards the neighbour; buf + m * stride
 * receives the block's own edge EMFs there in the reference's order. */
void orc_ref_emf_same_level(const orc_ref_geom *g, const double *e1, const double *e2,
                            const double *e3, double *buf, int64_t stride, const orc_ref_item *it,
                            int nit) {
  rgeo r = mkgeo(g);
  for (int m = 0; m < nit; ++m) {
    const int b = it[m].fb, fid = it[m].fid;
    double *out = buf + (int64_t)m * stride;
    int64_t p = 0;
    if (fid < 6) {
      const int dir = fid >> 1, x = face_at(&r, dir, fid & 1);
      if (dir == 0) {
        for (int k = r.ks; k <= r.ke + 1; k++)
          for (int j = r.js; j <= r.je; j++) out[p++] = EE2(b, k, j, x);
        for (int k = r.ks; k <= r.ke; k++)
          for (int j = r.js; j <= r.je + 1; j++) out[p++] = EE3(b, k, j, x);
      } else if (dir == 1) {
        for (int k = r.ks; k <= r.ke + 1; k++)
          for (int i = r.is; i <= r.ie; i++) out[p++] = EE1(b, k, x, i);
        for (int k = r.ks; k <= r.ke; k++)
          for (int i = r.is; i <= r.ie + 1; i++) out[p++] = EE3(b, k, x, i);
      } else {
        for (int j = r.js; j <= r.je + 1; j++)
          for (int i = r.is; i <= r.ie; i++) out[p++] = EE1(b, x, j, i);
        for (int j = r.js; j <= r.je; j++)
          for (int i = r.is; i <= r.ie + 1; i++) out[p++] = EE2(b, x, j, i);
      }
    } else {
      const int eid = fid - 6, lo = eid & 1, hi = (eid >> 1) & 1;
      if (eid < 4) {
        const int i = face_at(&r, 0, lo), j = face_at(&r, 1, hi);
        for (int k = r.ks; k <= r.ke; k++) out[p++] = EE3(b, k, j, i);
      } else if (eid < 8) {
        const int i = face_at(&r, 0, lo), k = face_at(&r, 2, hi);
        for (int j = r.js; j <= r.je; j++) out[p++] = EE2(b, k, j, i);
      } else {
        const int j = face_at(&r, 1, lo), k = face_at(&r, 2, hi);
        for (int i = r.is; i <= r.ie; i++) out[p++] = EE1(b, k, j, i);
      }
    }
  }
}

/* SetFluxBoundarySameLevel (flux_correction_fc.cpp:689-940, finer = 0) and
 * SetFluxBoundaryFromFiner (:944-1185, finer = 1), 3D, no shearing or polar boundary.
 * Item: cb = the receiving block, fb = the buffer slot (buf + fb * stride), fid = the
 * receiving block's face or edge, fi1 / fi2 = the finer neighbour's half along the
 * first / second tangential direction (faces) or along the edge (fi1).  Items run in
 * list order, which is the reference's neighbour order. */
void orc_ref_emf_set(const orc_ref_geom *g, double *e1, double *e2, double *e3, const double *buf,
                     int64_t stride, int finer, const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int nx1 = g->nx1, nx2 = g->nx2, nx3 = g->nx3;
  for (int m = 0; m < nit; ++m) {
    const int b = it[m].cb, fid = it[m].fid;
    const double *in = buf + (int64_t)it[m].fb * stride;
    int64_t p = 0;
    int il = r.is, iu = r.ie, jl = r.js, ju = r.je, kl = r.ks, ku = r.ke;
    if (fid < 6) {
      const int dir = fid >> 1, x = face_at(&r, dir, fid & 1);
      if (dir == 0) {
        if (finer) {
          if (it[m].fi1 == 0) ju = r.js + nx2 / 2 - 1; else jl = r.js + nx2 / 2;
          if (it[m].fi2 == 0) ku = r.ks + nx3 / 2 - 1; else kl = r.ks + nx3 / 2;
        }
        for (int k = kl; k <= ku + 1; k++)
          for (int j = jl; j <= ju; j++) EE2(b, k, j, x) += in[p++];
        for (int k = kl; k <= ku; k++)
          for (int j = jl; j <= ju + 1; j++) EE3(b, k, j, x) += in[p++];
      } else if (dir == 1) {
        if (finer) {
          if (it[m].fi1 == 0) iu = r.is + nx1 / 2 - 1; else il = r.is + nx1 / 2;
          if (it[m].fi2 == 0) ku = r.ks + nx3 / 2 - 1; else kl = r.ks + nx3 / 2;
        }
        for (int k = kl; k <= ku + 1; k++)
          for (int i = il; i <= iu; i++) EE1(b, k, x, i) += in[p++];
        for (int k = kl; k <= ku; k++)
          for (int i = il; i <= iu + 1; i++) EE3(b, k, x, i) += in[p++];
      } else {
        if (finer) {
          if (it[m].fi1 == 0) iu = r.is + nx1 / 2 - 1; else il = r.is + nx1 / 2;
          if (it[m].fi2 == 0) ju = r.js + nx2 / 2 - 1; else jl = r.js + nx2 / 2;
        }
        for (int j = jl; j <= ju + 1; j++)
          for (int i = il; i <= iu; i++) EE1(b, x, j, i) += in[p++];
        for (int j = jl; j <= ju; j++)
          for (int i = il; i <= iu + 1; i++) EE2(b, x, j, i) += in[p++];
      }
    } else {
      const int eid = fid - 6, lo = eid & 1, hi = (eid >> 1) & 1;
      if (eid < 4) {
        const int i = face_at(&r, 0, lo), j = face_at(&r, 1, hi);
        if (finer) { if (it[m].fi1 == 0) ku = r.ks + nx3 / 2 - 1; else kl = r.ks + nx3 / 2; }
        for (int k = kl; k <= ku; k++) EE3(b, k, j, i) += in[p++];
      } else if (eid < 8) {
        const int i = face_at(&r, 0, lo), k = face_at(&r, 2, hi);
        if (finer) { if (it[m].fi1 == 0) ju = r.js + nx2 / 2 - 1; else jl = r.js + nx2 / 2; }
        for (int j = jl; j <= ju; j++) EE2(b, k, j, i) += in[p++];
      } else {
        const int j = face_at(&r, 1, lo), k = face_at(&r, 2, hi);
        if (finer) { if (it[m].fi1 == 0) iu = r.is + nx1 / 2 - 1; else il = r.is + nx1 / 2; }
        for (int i = il; i <= iu; i++) EE1(b, k, j, i) += in[p++];
      }
    }
  }
}

/* ClearCoarseFluxBoundary (flux_correction_fc.cpp:1214-1348), 3D.  Item: cb = the block,
 * fid = a face whose neighbour is finer (its EMFs inside the face, not on the face's
 * own edges, are zeroed) or an edge with a finer block around it (edge_flag_ false). */
void orc_ref_emf_clear(const orc_ref_geom *g, double *e1, double *e2, double *e3,
                       const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  for (int m = 0; m < nit; ++m) {
    const int b = it[m].cb, fid = it[m].fid;
    if (fid < 6) {
      const int dir = fid >> 1, x = face_at(&r, dir, fid & 1);
      if (dir == 0) {
        for (int k = r.ks + 1; k <= r.ke; k++)
          for (int j = r.js; j <= r.je; j++) EE2(b, k, j, x) = 0.0;
        for (int k = r.ks; k <= r.ke; k++)
          for (int j = r.js + 1; j <= r.je; j++) EE3(b, k, j, x) = 0.0;
      } else if (dir == 1) {
        for (int k = r.ks + 1; k <= r.ke; k++)
          for (int i = r.is; i <= r.ie; i++) EE1(b, k, x, i) = 0.0;
        for (int k = r.ks; k <= r.ke; k++)
          for (int i = r.is + 1; i <= r.ie; i++) EE3(b, k, x, i) = 0.0;
      } else {
        for (int j = r.js + 1; j <= r.je; j++)
          for (int i = r.is; i <= r.ie; i++) EE1(b, x, j, i) = 0.0;
        for (int j = r.js; j <= r.je; j++)
          for (int i = r.is + 1; i <= r.ie; i++) EE2(b, x, j, i) = 0.0;
      }
    } else {
      const int eid = fid - 6, lo = eid & 1, hi = (eid >> 1) & 1;
      if (eid < 4) {
        const int i = face_at(&r, 0, lo), j = face_at(&r, 1, hi);
        for (int k = r.ks; k <= r.ke; k++) EE3(b, k, j, i) = 0.0;
      } else if (eid < 8) {
        const int i = face_at(&r, 0, lo), k = face_at(&r, 2, hi);
        for (int j = r.js; j <= r.je; j++) EE2(b, k, j, i) = 0.0;
      } else {
        const int j = face_at(&r, 1, lo), k = face_at(&r, 2, hi);
        for (int i = r.is; i <= r.ie; i++) EE1(b, k, j, i) = 0.0;
      }
    }
  }
}

/* AverageFluxBoundary (flux_correction_fc.cpp:1355-1543), 3D, no shearing box.  Item: cb =
 * the block; fid = a face with a block or periodic boundary and fi1 = 0 for a neighbour
 * on the same level (the EMFs inside the face are halved) or 1 for finer neighbours (the
 * two lines where the finer blocks meet are halved); or fid = an edge and fi1 = the
 * number of finest-level blocks around it (nedge_fine_ >= 2; the edge is scaled by
 * 1/fi1). */
void orc_ref_emf_average(const orc_ref_geom *g, double *e1, double *e2, double *e3,
                         const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int nx1 = g->nx1, nx2 = g->nx2, nx3 = g->nx3;
  for (int m = 0; m < nit; ++m) {
    const int b = it[m].cb, fid = it[m].fid;
    if (fid < 6) {
      const int dir = fid >> 1, x = face_at(&r, dir, fid & 1);
      if (it[m].fi1 == 0) {
        if (dir == 0) {
          for (int k = r.ks + 1; k <= r.ke; k++)
            for (int j = r.js; j <= r.je; j++) EE2(b, k, j, x) *= 0.5;
          for (int k = r.ks; k <= r.ke; k++)
            for (int j = r.js + 1; j <= r.je; j++) EE3(b, k, j, x) *= 0.5;
        } else if (dir == 1) {
          for (int k = r.ks + 1; k <= r.ke; k++)
            for (int i = r.is; i <= r.ie; i++) EE1(b, k, x, i) *= 0.5;
          for (int k = r.ks; k <= r.ke; k++)
            for (int i = r.is + 1; i <= r.ie; i++) EE3(b, k, x, i) *= 0.5;
        } else {
          for (int j = r.js + 1; j <= r.je; j++)
            for (int i = r.is; i <= r.ie; i++) EE1(b, x, j, i) *= 0.5;
          for (int j = r.js; j <= r.je; j++)
            for (int i = r.is + 1; i <= r.ie; i++) EE2(b, x, j, i) *= 0.5;
        }
      } else {
        if (dir == 0) {
          const int k = r.ks + nx3 / 2, j = r.js + nx2 / 2;
          for (int jj = r.js; jj <= r.je; jj++) EE2(b, k, jj, x) *= 0.5;
          for (int kk = r.ks; kk <= r.ke; kk++) EE3(b, kk, j, x) *= 0.5;
        } else if (dir == 1) {
          const int k = r.ks + nx3 / 2, i = r.is + nx1 / 2;
          for (int ii = r.is; ii <= r.ie; ii++) EE1(b, k, x, ii) *= 0.5;
          for (int kk = r.ks; kk <= r.ke; kk++) EE3(b, kk, x, i) *= 0.5;
        } else {
          const int j = r.js + nx2 / 2, i = r.is + nx1 / 2;
          for (int ii = r.is; ii <= r.ie; ii++) EE1(b, x, j, ii) *= 0.5;
          for (int jj = r.js; jj <= r.je; jj++) EE2(b, x, jj, i) *= 0.5;
        }
      }
    } else {
      const int eid = fid - 6, lo = eid & 1, hi = (eid >> 1) & 1;
      const double div = 1.0 / (double)it[m].fi1;
      if (eid < 4) {
        const int i = face_at(&r, 0, lo), j = face_at(&r, 1, hi);
        for (int k = r.ks; k <= r.ke; k++) EE3(b, k, j, i) *= div;
      } else if (eid < 8) {
        const int i = face_at(&r, 0, lo), k = face_at(&r, 2, hi);
        for (int j = r.js; j <= r.je; j++) EE2(b, k, j, i) *= div;
      } else {
        const int j = face_at(&r, 1, lo), k = face_at(&r, 2, hi);
        for (int i = r.is; i <= r.ie; i++) EE1(b, k, j, i) *= div;
      }
    }
  }
}
#undef EE1
#undef EE2
#undef EE3

/* CountFineEdges (bvals/fc/bvals_fc.cpp:1128-1191) for a 3D block.  nblevel[27] holds the
 * neighbourhood's levels relative to the block at ((ox3+1)*3 + ox2+1)*3 + ox1+1 (any
 * negative value for a coarser or missing neighbour); edge_flag[e] = no finer block
 * touches edge e, nedge_fine[e] = the number of blocks at the finest level around it. */
void orc_ref_count_fine_edges(const int *nblevel, int *edge_flag, int *nedge_fine) {
#define LV(o3, o2, o1) nblevel[((o3) + 1) * 9 + ((o2) + 1) * 3 + (o1) + 1]
  int eid = 0;
  for (int pass = 0; pass < 3; ++pass)   /* x1x2 (ox2, ox1), x1x3 (ox3, ox1), x2x3 (ox3, ox2) */
    for (int os = -1; os <= 1; os += 2)
      for (int of = -1; of <= 1; of += 2) {
        int nf = 0, fl = 0;
        const int as = os - 1 < -1 ? -1 : os - 1, ae = os + 1 > 1 ? 1 : os + 1;
        const int bs = of - 1 < -1 ? -1 : of - 1, be = of + 1 > 1 ? 1 : of + 1;
        for (int a = as; a <= ae; a++)
          for (int c = bs; c <= be; c++) {
            const int l = pass == 0 ? LV(0, a, c) : (pass == 1 ? LV(a, 0, c) : LV(a, c, 0));
            if (l > fl) { fl++; nf = 0; }
            if (l == fl) nf++;
          }
        edge_flag[eid] = (fl == 0);
        nedge_fine[eid++] = nf;
      }
#undef LV
}

/* ---------------------------------------------------------------- ghost prolongation */
void orc_ref_prim_to_cons(const orc_ref_geom *g, const double *w, const double *b1, const double *b2,
                          const double *b3, double *bcc, double *u, double gamma, int iso,
                          const orc_ref_item *it, int nit) {
  rgeo r = mkgeo(g);
  const int nh = iso ? 4 : 5;
  const double igm1 = iso ? 0.0 : 1.0 / (gamma - 1.0);
  for (int m = 0; m < nit; ++m) {
    const int fb = it[m].fb;
    const int fil = (it[m].si - r.cis) * 2 + r.is, fiu = (it[m].ei - r.cis) * 2 + r.is + 1;
    const int fjl = (it[m].sj - r.cjs) * 2 + r.js, fju = (it[m].ej - r.cjs) * 2 + r.js + 1;
    const int fkl = (it[m].sk - r.cks) * 2 + r.ks, fku = (it[m].ek - r.cks) * 2 + r.ks + 1;
    for (int k = fkl; k <= fku; ++k)
      for (int j = fjl; j <= fju; ++j)
        for (int i = fil; i <= fiu; ++i) {
          /* uniform branch: lw = rw = 0.5 */
          double c1 = 0.5 * b1[ix(r.n3, r.n2, r.n1 + 1, 1, fb, 0, k, j, i)] +
                      0.5 * b1[ix(r.n3, r.n2, r.n1 + 1, 1, fb, 0, k, j, i + 1)];
          double c2 = 0.5 * b2[ix(r.n3, r.n2 + 1, r.n1, 1, fb, 0, k, j, i)] +
                      0.5 * b2[ix(r.n3, r.n2 + 1, r.n1, 1, fb, 0, k, j + 1, i)];
          double c3 = 0.5 * b3[ix(r.n3 + 1, r.n2, r.n1, 1, fb, 0, k, j, i)] +
                      0.5 * b3[ix(r.n3 + 1, r.n2, r.n1, 1, fb, 0, k + 1, j, i)];
          bcc[ix(r.n3, r.n2, r.n1, 3, fb, 0, k, j, i)] = c1;
          bcc[ix(r.n3, r.n2, r.n1, 3, fb, 1, k, j, i)] = c2;
          bcc[ix(r.n3, r.n2, r.n1, 3, fb, 2, k, j, i)] = c3;
#define W(n) w[ix(r.n3, r.n2, r.n1, nh, fb, (n), k, j, i)]
#define U(n) u[ix(r.n3, r.n2, r.n1, nh, fb, (n), k, j, i)]
          double w_d = W(0), w_vx = W(1), w_vy = W(2), w_vz = W(3);
          U(0) = w_d;
          U(1) = w_vx * w_d;
          U(2) = w_vy * w_d;
          U(3) = w_vz * w_d;
          if (!iso)
            U(4) = W(4) * igm1 + 0.5 * (w_d * (w_vx * w_vx + w_vy * w_vy + w_vz * w_vz) + (c1 * c1 + c2 * c2 + c3 * c3));
#undef W
#undef U
        }
  }
}

int orc_ref_prolong_ranges(const orc_ref_geom *g, const int *ox, const int *lx_odd, const int *nblevel,
                           int *out) {
  rgeo r = mkgeo(g);
  int cnghost = (g->ng + 1) / 2 + 1, cn = cnghost - 1;   /* meshblock.cpp:83 */
  if (nblevel[((ox[2] + 1) * 3 + ox[1] + 1) * 3 + ox[0] + 1] >= 0) return -1;
  int cs[3] = {r.cis, r.cjs, r.cks};
  int ce[3] = {r.cis + g->nx1 / 2 - 1, r.cjs + g->nx2 / 2 - 1, r.cks + g->nx3 / 2 - 1};
  int s[3], e[3];
  for (int d = 0; d < 3; ++d) {   /* bvals_refine.cpp:172-199 */
    if (ox[d] == 0) {
      s[d] = cs[d]; e[d] = ce[d];
      if (lx_odd[d] == 0) e[d] += cn; else s[d] -= cn;
    } else if (ox[d] > 0) {
      s[d] = ce[d] + 1; e[d] = ce[d] + cn;
    } else {
      s[d] = cs[d] - cn; e[d] = cs[d] - 1;
    }
  }
  /* face ranges of ProlongateGhostCells: nblevel[ox3+1][ox2+1][ox1] etc. */
#define LEV(i, j, k) nblevel[((k) * 3 + (j)) * 3 + (i)]
  int il = s[0], iu = e[0] + 1, jl = s[1], ju = e[1] + 1, kl = s[2], ku = e[2] + 1;
  if (ox[0] >= 0 && LEV(ox[0], ox[1] + 1, ox[2] + 1) >= 0) il++;
  if (ox[0] <= 0 && LEV(ox[0] + 2, ox[1] + 1, ox[2] + 1) >= 0) iu--;
  if (ox[1] >= 0 && LEV(ox[0] + 1, ox[1], ox[2] + 1) >= 0) jl++;
  if (ox[1] <= 0 && LEV(ox[0] + 1, ox[1] + 2, ox[2] + 1) >= 0) ju--;
  if (ox[2] >= 0 && LEV(ox[0] + 1, ox[1] + 1, ox[2]) >= 0) kl++;
  if (ox[2] <= 0 && LEV(ox[0] + 1, ox[1] + 1, ox[2] + 2) >= 0) ku--;
#undef LEV
  out[0] = s[0]; out[1] = e[0]; out[2] = s[1]; out[3] = e[1]; out[4] = s[2]; out[5] = e[2];
  out[6] = il; out[7] = iu; out[8] = jl; out[9] = ju; out[10] = kl; out[11] = ku;
  return 0;
}
