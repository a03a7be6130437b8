/*
 * agx_oracle.c -- CPU FP64 restatement of the reference algorithm for the MHD +
 * multigrid self-gravity hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, smoke() and bench.py's
 * cpu_baseline leg; never by the product library.  Built with -ffp-contract=off so
 * that every expression rounds exactly as written; operand groupings follow the
 * reference so the result is the reference's IEEE sequence.
 *
 * Citations are file:line in the reference's src/ tree.
 *
 * Parity pinning (how this restatement is checked; see tests/test_oracle_*.py):
 *   - Riemann solvers: consistency F(U,U) = F_phys(U) for HLLD/HLLE, HLLD exact
 *     resolution of an isolated contact, mirror symmetry of the flux.
 *   - PLM: exact on linear data, zero slope at extrema.
 *   - CT: div B of the face field is preserved to round-off.
 *   - Full VL2+PLM+HLLD+CT: the reference's own regression thresholds for 3D linear
 *     waves (tst/regression/scripts/tests/mhd/mhd_linwave.py:97-162), second-order
 *     convergence, and bit-identical L- and R-going fast-wave errors (:159-162).
 *   - Uniform state: zero update for any solver; conservation to round-off.
 *   - Multigrid: closed-form Poisson solution (pgen/poisson.cpp:85-93), brute-force
 *     FFT solve of the discrete 7-point Laplacian, linearity of every operator.
 * No golden vectors exist in the reference for this path (tst/regression/data holds
 * only SR-shock and MRI files), and the reference cannot be built here without its
 * generated src/defs.hpp; the checks above are what pin this file.
 */
#include "agx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NHYDRO 5
#define NWAVE 7
#define IDN 0
#define IVX 1
#define IVY 2
#define IVZ 3
#define IEN 4
#define IPR 4
#define IBY 5
#define IBZ 6
#define IB1 0
#define IB2 1
#define IB3 2
#define SQR(x) ((x) * (x))
#define SMALL_NUMBER 1.0e-8
#define TINY_NUMBER 1.0e-20

enum {
  ARR_U = 0, ARR_U1, ARR_W, ARR_BCC, ARR_B1F, ARR_B2F, ARR_B3F, ARR_R1F, ARR_R2F,
  ARR_R3F, ARR_FX1, ARR_FX2, ARR_FX3, ARR_E3X1, ARR_E2X1, ARR_E1X2, ARR_E3X2,
  ARR_E1X3, ARR_E2X3, ARR_WX1, ARR_WX2, ARR_WX3, ARR_E1, ARR_E2, ARR_E3, ARR_PHI,
  ARR_COUNT
};

/* ---------------------------------------------------------------- multigrid data */
typedef struct {
  int nlev;          /* levels, 0 = coarsest */
  int n1[24], n2[24], n3[24];   /* active cells per level */
  double *u[24], *src[24], *def[24];
  double *uold[24];  /* FAS: u at StoreOldData (multigrid.cpp:645-653) */
  double rdx;        /* cell size of the finest level */
  double xmin[3], xmax[3];   /* RegionSize of the block (or of the root grid) */
  int bnd[6];        /* physical multipole boundary on face ix1, ox1, ix2, ox2, ix3, ox3 */
} mglevels;

struct orc_ctx {
  orc_desc d;
  int nb1, nb2, nb3, nblocks;
  int n1, n2, n3, ng;
  int is, ie, js, je, ks, ke;
  int64_t ncell, nf1, nf2, nf3, ne1, ne2, ne3;
  double *arr[ARR_COUNT];
  int64_t arrsize[ARR_COUNT];
  double *x1f, *x1v, *dx1f, *x2f, *x2v, *dx2f, *x3f, *x3v, *dx3f;  /* per block */
  double *cce;       /* cell-centred E scratch, [3][ncell] per block */
  double time, dt, tlim_hint;
  int ncycle;
  /* multigrid */
  mglevels *mgb;     /* one per block */
  mglevels mgr;      /* root grid */
  int nmblevel, nrootlevel, ntotal, current, fmglevel, coffset;
  double last_ave, last_norm;
  int last_iters;
  int mg_mode, mg_niter;
  double mg_eps;
  int srcmask;       /* source mask: src = 0 where |x - c|^2 > r^2 (pgen/collapse.cpp:66-82) */
  double srcmask_c[3], srcmask_r2;
  int iso;           /* isothermal EOS (eos = 1): NHYDRO = 4, the energy slot stays 0 */
  double noise_eps;  /* test-only: primitive perturbation after each C2P (orc_set_noise) */
  uint64_t noise_state;
  int nhydro;
  int fas;           /* ffas_: multipole boundaries (multigrid_driver.cpp:292-295) */
  double mpo[3];     /* expansion origin (mpo_) */
  double mpc[25];    /* scaled multipole coefficients (mpcoeff_) */
};

/* ---------------------------------------------------------------- helpers */
#define CIDX(c, k, j, i) ((((int64_t)(k)) * (c)->n2 + (j)) * (c)->n1 + (i))
#define F1IDX(c, k, j, i) ((((int64_t)(k)) * (c)->n2 + (j)) * ((c)->n1 + 1) + (i))
#define F2IDX(c, k, j, i) ((((int64_t)(k)) * ((c)->n2 + 1) + (j)) * (c)->n1 + (i))
#define F3IDX(c, k, j, i) ((((int64_t)(k)) * (c)->n2 + (j)) * (c)->n1 + (i))
#define E1IDX(c, k, j, i) ((((int64_t)(k)) * ((c)->n2 + 1) + (j)) * (c)->n1 + (i))
#define E2IDX(c, k, j, i) ((((int64_t)(k)) * (c)->n2 + (j)) * ((c)->n1 + 1) + (i))
#define E3IDX(c, k, j, i) ((((int64_t)(k)) * ((c)->n2 + 1) + (j)) * ((c)->n1 + 1) + (i))

static inline double dmin(double a, double b) { return (a < b) ? a : b; }   /* std::min */
static inline double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

/* ComputeMeshGeneratorX with sym_interval=true, mesh/mesh.hpp:389-405 */
static double meshgen_x(int64_t index, int64_t nrange) {
  int64_t noffset = index - (nrange) / 2;
  int64_t noffset_ceil = index - (nrange + 1) / 2;
  return (double)(noffset + noffset_ceil) / (2.0 * nrange);
}
/* UniformMeshGeneratorX1..3, mesh/mesh.hpp:465-488 */
static double uniform_gen(double x, double xmin, double xmax) {
  return 0.5 * (xmin + xmax) + (x * xmax - x * xmin);
}

/* Uniform-mesh face/centre coordinates of one block along one direction:
 * coordinates.cpp:127-145 (faces), cartesian.cpp:26-36 (centres),
 * mesh.cpp:1668-1690 (block extents). */
static void block_coords(int nmesh, int nblk, int nbx, int ng, int lx, double xmin,
                         double xmax, double *xf, double *xv, double *dxf) {
  double bmin = (lx == 0) ? xmin : uniform_gen(meshgen_x(lx, nblk), xmin, xmax);
  double bmax = (lx == nblk - 1) ? xmax : uniform_gen(meshgen_x(lx + 1, nblk), xmin, xmax);
  double dx = (bmax - bmin) / nbx;
  int il = ng, iu = ng + nbx - 1;
  for (int i = il - ng; i <= iu + ng + 1; ++i) {
    int64_t noffset = (int64_t)(i - il) + (int64_t)lx * nbx;
    xf[i] = uniform_gen(meshgen_x(noffset, nmesh), xmin, xmax);
  }
  xf[il] = bmin;
  xf[iu + 1] = bmax;
  for (int i = il - ng; i <= iu + ng; ++i) dxf[i] = dx;
  for (int i = il - ng; i <= iu + ng; ++i) xv[i] = 0.5 * (xf[i + 1] + xf[i]);
}

static void *xcalloc(size_t n) {
  void *p = calloc(n ? n : 1, 1);
  if (!p) abort();
  return p;
}

/* ---------------------------------------------------------------- create */
static void mg_alloc(mglevels *m, int nx1, int nx2, int nx3, int nlev, double rdx) {
  m->nlev = nlev;
  m->rdx = rdx;
  for (int l = 0; l < nlev; ++l) {
    int ll = nlev - 1 - l;
    m->n1[l] = nx1 >> ll;
    m->n2[l] = nx2 >> ll;
    m->n3[l] = nx3 >> ll;
    size_t sz = (size_t)(m->n1[l] + 2) * (m->n2[l] + 2) * (m->n3[l] + 2);
    m->u[l] = (double *)xcalloc(sz * sizeof(double));
    m->src[l] = (double *)xcalloc(sz * sizeof(double));
    m->def[l] = (double *)xcalloc(sz * sizeof(double));
    m->uold[l] = (double *)xcalloc(sz * sizeof(double));
  }
}

orc_ctx *orc_create(const orc_desc *d) {
  /* the multigrid hierarchy is built from logically cubic MeshBlocks of 2^n cells
   * (multigrid.cpp:35-42 refuses others); anything else is refused here too */
  if (d->self_gravity && (d->bnx1 != d->bnx2 || d->bnx1 != d->bnx3 || (d->bnx1 & (d->bnx1 - 1)))) return NULL;
  orc_ctx *c = (orc_ctx *)xcalloc(sizeof(orc_ctx));
  c->d = *d;
  c->nb1 = d->nx1 / d->bnx1;
  c->nb2 = d->nx2 / d->bnx2;
  c->nb3 = d->nx3 / d->bnx3;
  c->nblocks = c->nb1 * c->nb2 * c->nb3;
  c->ng = d->nghost;
  c->n1 = d->bnx1 + 2 * c->ng;
  c->n2 = d->bnx2 + 2 * c->ng;
  c->n3 = d->bnx3 + 2 * c->ng;
  c->is = c->ng; c->ie = c->ng + d->bnx1 - 1;
  c->js = c->ng; c->je = c->ng + d->bnx2 - 1;
  c->ks = c->ng; c->ke = c->ng + d->bnx3 - 1;
  c->ncell = (int64_t)c->n1 * c->n2 * c->n3;
  c->nf1 = (int64_t)c->n3 * c->n2 * (c->n1 + 1);
  c->nf2 = (int64_t)c->n3 * (c->n2 + 1) * c->n1;
  c->nf3 = (int64_t)(c->n3 + 1) * c->n2 * c->n1;
  c->ne1 = (int64_t)(c->n3 + 1) * (c->n2 + 1) * c->n1;
  c->ne2 = (int64_t)(c->n3 + 1) * c->n2 * (c->n1 + 1);
  c->ne3 = (int64_t)c->n3 * (c->n2 + 1) * (c->n1 + 1);
  int64_t per[ARR_COUNT] = {
      NHYDRO * c->ncell, NHYDRO * c->ncell, NHYDRO * c->ncell, 3 * c->ncell,
      c->nf1, c->nf2, c->nf3, c->nf1, c->nf2, c->nf3,
      NHYDRO * c->nf1, NHYDRO * c->nf2, NHYDRO * c->nf3,
      c->nf1, c->nf1, c->nf2, c->nf2, c->nf3, c->nf3,
      c->nf1, c->nf2, c->nf3, c->ne1, c->ne2, c->ne3, c->ncell};
  for (int a = 0; a < ARR_COUNT; ++a) {
    c->arrsize[a] = per[a];
    c->arr[a] = (double *)xcalloc((size_t)per[a] * c->nblocks * sizeof(double));
  }
  c->cce = (double *)xcalloc((size_t)3 * c->ncell * c->nblocks * sizeof(double));
  int nb = c->nblocks;
  c->x1f = (double *)xcalloc(sizeof(double) * nb * (c->n1 + 1));
  c->x1v = (double *)xcalloc(sizeof(double) * nb * (c->n1 + 1));
  c->dx1f = (double *)xcalloc(sizeof(double) * nb * (c->n1 + 1));
  c->x2f = (double *)xcalloc(sizeof(double) * nb * (c->n2 + 1));
  c->x2v = (double *)xcalloc(sizeof(double) * nb * (c->n2 + 1));
  c->dx2f = (double *)xcalloc(sizeof(double) * nb * (c->n2 + 1));
  c->x3f = (double *)xcalloc(sizeof(double) * nb * (c->n3 + 1));
  c->x3v = (double *)xcalloc(sizeof(double) * nb * (c->n3 + 1));
  c->dx3f = (double *)xcalloc(sizeof(double) * nb * (c->n3 + 1));
  for (int g = 0; g < nb; ++g) {
    int lx1 = g % c->nb1, lx2 = (g / c->nb1) % c->nb2, lx3 = g / (c->nb1 * c->nb2);
    block_coords(d->nx1, c->nb1, d->bnx1, c->ng, lx1, d->x1min, d->x1max,
                 c->x1f + g * (c->n1 + 1), c->x1v + g * (c->n1 + 1), c->dx1f + g * (c->n1 + 1));
    block_coords(d->nx2, c->nb2, d->bnx2, c->ng, lx2, d->x2min, d->x2max,
                 c->x2f + g * (c->n2 + 1), c->x2v + g * (c->n2 + 1), c->dx2f + g * (c->n2 + 1));
    block_coords(d->nx3, c->nb3, d->bnx3, c->ng, lx3, d->x3min, d->x3max,
                 c->x3f + g * (c->n3 + 1), c->x3v + g * (c->n3 + 1), c->dx3f + g * (c->n3 + 1));
  }
  c->dt = DBL_MAX;
  c->iso = (d->eos == 1);
  c->nhydro = c->iso ? 4 : NHYDRO;
  /* multigrid hierarchy: multigrid.cpp:32-144, multigrid_driver.cpp:343-349 */
  if (d->self_gravity) {
    int nl = 0;
    for (int l = 0; l < 20; ++l)
      if ((1 << l) == d->bnx1) { nl = l + 1; break; }
    c->nmblevel = nl;
    c->mgb = (mglevels *)xcalloc(sizeof(mglevels) * nb);
    for (int g = 0; g < nb; ++g) {
      int lx1 = g % c->nb1, lx2 = (g / c->nb1) % c->nb2, lx3 = g / (c->nb1 * c->nb2);
      double bmin = (lx1 == 0) ? d->x1min : uniform_gen(meshgen_x(lx1, c->nb1), d->x1min, d->x1max);
      double bmax = (lx1 == c->nb1 - 1) ? d->x1max
                                         : uniform_gen(meshgen_x(lx1 + 1, c->nb1), d->x1min, d->x1max);
      (void)lx2; (void)lx3;
      mg_alloc(&c->mgb[g], d->bnx1, d->bnx2, d->bnx3, nl, (bmax - bmin) / (double)d->bnx1);
      int lx[3] = {lx1, lx2, lx3}, nbd[3] = {c->nb1, c->nb2, c->nb3};
      double mn[3] = {d->x1min, d->x2min, d->x3min}, mx[3] = {d->x1max, d->x2max, d->x3max};
      for (int dd = 0; dd < 3; ++dd) {
        mglevels *m = &c->mgb[g];
        m->xmin[dd] = (lx[dd] == 0) ? mn[dd] : uniform_gen(meshgen_x(lx[dd], nbd[dd]), mn[dd], mx[dd]);
        m->xmax[dd] = (lx[dd] == nbd[dd] - 1) ? mx[dd]
                                              : uniform_gen(meshgen_x(lx[dd] + 1, nbd[dd]), mn[dd], mx[dd]);
        m->bnd[2 * dd] = (d->mg_bc == 1) && lx[dd] == 0;
        m->bnd[2 * dd + 1] = (d->mg_bc == 1) && lx[dd] == nbd[dd] - 1;
      }
    }
    int nrl = 0;
    for (int l = 0; l < 20; ++l)
      if (c->nb1 % (1 << l) == 0 && c->nb2 % (1 << l) == 0 && c->nb3 % (1 << l) == 0) nrl = l + 1;
    c->nrootlevel = nrl;
    mg_alloc(&c->mgr, c->nb1, c->nb2, c->nb3, nrl, (d->x1max - d->x1min) / (double)c->nb1);
    c->mgr.xmin[0] = d->x1min; c->mgr.xmin[1] = d->x2min; c->mgr.xmin[2] = d->x3min;
    c->mgr.xmax[0] = d->x1max; c->mgr.xmax[1] = d->x2max; c->mgr.xmax[2] = d->x3max;
    for (int f = 0; f < 6; ++f) c->mgr.bnd[f] = (d->mg_bc == 1);
    c->fas = (d->mg_bc == 1);
    c->mpo[0] = d->mp_origin[0]; c->mpo[1] = d->mp_origin[1]; c->mpo[2] = d->mp_origin[2];
    c->ntotal = nrl + nl - 1;
    c->mg_mode = d->mg_mode;
    c->mg_niter = d->mg_niter;
    c->mg_eps = d->mg_threshold;
  }
  return c;
}

static void mg_free(mglevels *m) {
  for (int l = 0; l < m->nlev; ++l) {
    free(m->u[l]); free(m->src[l]); free(m->def[l]); free(m->uold[l]);
  }
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  for (int a = 0; a < ARR_COUNT; ++a) free(c->arr[a]);
  free(c->cce);
  free(c->x1f); free(c->x1v); free(c->dx1f);
  free(c->x2f); free(c->x2v); free(c->dx2f);
  free(c->x3f); free(c->x3v); free(c->dx3f);
  if (c->mgb) {
    for (int g = 0; g < c->nblocks; ++g) mg_free(&c->mgb[g]);
    free(c->mgb);
    mg_free(&c->mgr);
  }
  free(c);
}

int64_t orc_size(orc_ctx *c, int what) {
  switch (what) {
    case 0: return c->ncell;
    case 1: return c->nf1;
    case 2: return c->nf2;
    case 3: return c->nf3;
    case 4: return c->ne1;
    case 5: return c->ne2;
    case 6: return c->ne3;
    case 7: return c->nblocks;
    default: return -1;
  }
}

double *orc_array(orc_ctx *c, int which) {
  if (which < 0 || which >= ARR_COUNT) return NULL;
  return c->arr[which];
}

void orc_set_time(orc_ctx *c, double time, double dt, int ncycle) {
  c->time = time; c->dt = dt; c->ncycle = ncycle;
}
void orc_get_time(orc_ctx *c, double *out3) {
  out3[0] = c->time; out3[1] = c->dt; out3[2] = c->ncycle;
}

/* Static mesh refinement: block g becomes the MeshBlock at LogicalLocation (level[g],
 * lx[3g..3g+2]) of a root grid of nrb blocks over box (x1min, x1max, ...): its
 * coordinates as Mesh::SetBlockSizeAndBoundaries (mesh.cpp:1668-1730) and the
 * Coordinates constructor (coordinates.cpp:86-145) set them on the uniform mesh of the
 * block's level. */
void orc_set_levels(orc_ctx *c, const int *nrb, const double *box, const int *level, const int64_t *lx) {
  const orc_desc *d = &c->d;
  for (int g = 0; g < c->nblocks; ++g) {
    const int L = level[g];
    block_coords((nrb[0] * d->bnx1) << L, nrb[0] << L, d->bnx1, c->ng, (int)lx[3 * g], box[0], box[1],
                 c->x1f + g * (c->n1 + 1), c->x1v + g * (c->n1 + 1), c->dx1f + g * (c->n1 + 1));
    block_coords((nrb[1] * d->bnx2) << L, nrb[1] << L, d->bnx2, c->ng, (int)lx[3 * g + 1], box[2], box[3],
                 c->x2f + g * (c->n2 + 1), c->x2v + g * (c->n2 + 1), c->dx2f + g * (c->n2 + 1));
    block_coords((nrb[2] * d->bnx3) << L, nrb[2] << L, d->bnx3, c->ng, (int)lx[3 * g + 2], box[4], box[5],
                 c->x3f + g * (c->n3 + 1), c->x3v + g * (c->n3 + 1), c->dx3f + g * (c->n3 + 1));
  }
}

void orc_coords(orc_ctx *c, int gid, int dir, double *xf, double *xv, double *dxf) {
  int n = (dir == 0) ? c->n1 : (dir == 1 ? c->n2 : c->n3);
  double *sf = (dir == 0) ? c->x1f : (dir == 1 ? c->x2f : c->x3f);
  double *sv = (dir == 0) ? c->x1v : (dir == 1 ? c->x2v : c->x3v);
  double *sd = (dir == 0) ? c->dx1f : (dir == 1 ? c->dx2f : c->dx3f);
  memcpy(xf, sf + (int64_t)gid * (n + 1), sizeof(double) * (n + 1));
  memcpy(xv, sv + (int64_t)gid * (n + 1), sizeof(double) * n);
  memcpy(dxf, sd + (int64_t)gid * (n + 1), sizeof(double) * n);
}

/* neighbour block id in direction (o1,o2,o3), periodic lattice */
static int nbr_gid(const orc_ctx *c, int g, int o1, int o2, int o3) {
  int lx1 = g % c->nb1, lx2 = (g / c->nb1) % c->nb2, lx3 = g / (c->nb1 * c->nb2);
  lx1 = (lx1 + o1 + c->nb1) % c->nb1;
  lx2 = (lx2 + o2 + c->nb2) % c->nb2;
  lx3 = (lx3 + o3 + c->nb3) % c->nb3;
  return (lx3 * c->nb2 + lx2) * c->nb1 + lx1;
}

/* ---------------------------------------------------------------- boundaries */
/* Same-level cell-centred ghost exchange: every ghost cell receives the neighbour's
 * interior cell at the same physical location (BufferUtility::PackData/UnpackData,
 * utils/buffer_utils.cpp:24-89; ranges bvals/cc/bvals_cc.cpp:201-336).  All senders
 * pack before any receiver unpacks, so reading interiors in place is equivalent. */
void orc_bvals_cc(orc_ctx *c, int which) {
  int nvar = (which == ARR_PHI) ? 1 : NHYDRO;
  double *a = c->arr[which];
  int64_t bstride = c->arrsize[which];
  int bx1 = c->d.bnx1, bx2 = c->d.bnx2, bx3 = c->d.bnx3;
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < c->nblocks; ++g) {
    for (int k = 0; k < c->n3; ++k) {
      int o3 = (k < c->ks) ? -1 : (k > c->ke ? 1 : 0);
      for (int j = 0; j < c->n2; ++j) {
        int o2 = (j < c->js) ? -1 : (j > c->je ? 1 : 0);
        for (int i = 0; i < c->n1; ++i) {
          int o1 = (i < c->is) ? -1 : (i > c->ie ? 1 : 0);
          if (!o1 && !o2 && !o3) continue;
          int s = nbr_gid(c, g, o1, o2, o3);
          int si = i - o1 * bx1, sj = j - o2 * bx2, sk = k - o3 * bx3;
          for (int n = 0; n < nvar; ++n)
            a[g * bstride + n * c->ncell + CIDX(c, k, j, i)] =
                a[s * bstride + n * c->ncell + CIDX(c, sk, sj, si)];
        }
      }
    }
  }
}

/* Same-level face-field ghost exchange, bvals/fc/bvals_fc.cpp:344-402 (pack) and
 * :583-700 (unpack).  The face shared by two blocks (e.g. x1f(ie+1) / x1f(is) of
 * the +x1 neighbour) is not exchanged: each block keeps its own copy. */
void orc_bvals_fc(orc_ctx *c, int reg) {
  int a1 = reg ? ARR_R1F : ARR_B1F, a2 = reg ? ARR_R2F : ARR_B2F, a3 = reg ? ARR_R3F : ARR_B3F;
  int bx1 = c->d.bnx1, bx2 = c->d.bnx2, bx3 = c->d.bnx3;
  double *b1 = c->arr[a1], *b2 = c->arr[a2], *b3 = c->arr[a3];
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < c->nblocks; ++g) {
    /* x1f: face index range is..ie+1 is "interior" along x1 */
    for (int k = 0; k < c->n3; ++k) {
      int o3 = (k < c->ks) ? -1 : (k > c->ke ? 1 : 0);
      for (int j = 0; j < c->n2; ++j) {
        int o2 = (j < c->js) ? -1 : (j > c->je ? 1 : 0);
        for (int i = 0; i <= c->n1; ++i) {
          int o1 = (i < c->is) ? -1 : (i > c->ie + 1 ? 1 : 0);
          if (!o1 && !o2 && !o3) continue;
          int s = nbr_gid(c, g, o1, o2, o3);
          b1[g * c->nf1 + F1IDX(c, k, j, i)] =
              b1[s * c->nf1 + F1IDX(c, k - o3 * bx3, j - o2 * bx2, i - o1 * bx1)];
        }
      }
    }
    for (int k = 0; k < c->n3; ++k) {
      int o3 = (k < c->ks) ? -1 : (k > c->ke ? 1 : 0);
      for (int j = 0; j <= c->n2; ++j) {
        int o2 = (j < c->js) ? -1 : (j > c->je + 1 ? 1 : 0);
        for (int i = 0; i < c->n1; ++i) {
          int o1 = (i < c->is) ? -1 : (i > c->ie ? 1 : 0);
          if (!o1 && !o2 && !o3) continue;
          int s = nbr_gid(c, g, o1, o2, o3);
          b2[g * c->nf2 + F2IDX(c, k, j, i)] =
              b2[s * c->nf2 + F2IDX(c, k - o3 * bx3, j - o2 * bx2, i - o1 * bx1)];
        }
      }
    }
    for (int k = 0; k <= c->n3; ++k) {
      int o3 = (k < c->ks) ? -1 : (k > c->ke + 1 ? 1 : 0);
      for (int j = 0; j < c->n2; ++j) {
        int o2 = (j < c->js) ? -1 : (j > c->je ? 1 : 0);
        for (int i = 0; i < c->n1; ++i) {
          int o1 = (i < c->is) ? -1 : (i > c->ie ? 1 : 0);
          if (!o1 && !o2 && !o3) continue;
          int s = nbr_gid(c, g, o1, o2, o3);
          b3[g * c->nf3 + F3IDX(c, k, j, i)] =
              b3[s * c->nf3 + F3IDX(c, k - o3 * bx3, j - o2 * bx2, i - o1 * bx1)];
        }
      }
    }
  }
}

/* ---------------------------------------------------------------- EOS */
/* EquationOfState::FastMagnetosonicSpeed, eos/adiabatic_mhd.cpp:149-156 */
static inline double fast_speed(double gamma, const double *prim, double bx) {
  double asq = gamma * prim[IPR];
  double vaxsq = bx * bx;
  double ct2 = (prim[IBY] * prim[IBY] + prim[IBZ] * prim[IBZ]);
  double qsq = vaxsq + ct2 + asq;
  double tmp = vaxsq + ct2 - asq;
  return sqrt(0.5 * (qsq + sqrt(tmp * tmp + 4.0 * asq * ct2)) / prim[IDN]);
}

/* EquationOfState::FastMagnetosonicSpeed, eos/isothermal_mhd.cpp:122-129 */
static inline double fast_speed_iso(double cs, const double *prim, double bx) {
  double asq = (cs * cs) * prim[IDN];
  double vaxsq = bx * bx;
  double ct2 = prim[IBY] * prim[IBY] + prim[IBZ] * prim[IBZ];
  double qsq = vaxsq + ct2 + asq;
  double tmp = vaxsq + ct2 - asq;
  return sqrt(0.5 * (qsq + sqrt(tmp * tmp + 4.0 * asq * ct2)) / prim[IDN]);
}

static inline double fast_speed_eos(const orc_ctx *c, const double *prim, double bx) {
  return c->iso ? fast_speed_iso(c->d.iso_cs, prim, bx) : fast_speed(c->d.gamma, prim, bx);
}

/* Field::CalculateCellCenteredField (uniform branch, field/field.cpp:112-180) then
 * EquationOfState::ConservedToPrimitive (eos/adiabatic_mhd.cpp:41-90), over the
 * interior plus all ghost zones (time_integrator.cpp:1971-1977: every face has a
 * neighbour on a periodic lattice). */
void orc_cons2prim(orc_ctx *c) {
  double gm1 = c->d.gamma - 1.0;
  double dfl = c->d.dfloor, pfl = c->d.pfloor;
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < c->nblocks; ++g) {
    double *u = c->arr[ARR_U] + g * c->arrsize[ARR_U];
    double *w = c->arr[ARR_W] + g * c->arrsize[ARR_W];
    double *bcc = c->arr[ARR_BCC] + g * c->arrsize[ARR_BCC];
    const double *b1 = c->arr[ARR_B1F] + g * c->nf1;
    const double *b2 = c->arr[ARR_B2F] + g * c->nf2;
    const double *b3 = c->arr[ARR_B3F] + g * c->nf3;
    int64_t nc = c->ncell;
    for (int k = 0; k < c->n3; ++k)
      for (int j = 0; j < c->n2; ++j)
        for (int i = 0; i < c->n1; ++i) {
          int64_t id = CIDX(c, k, j, i);
          bcc[IB1 * nc + id] = 0.5 * b1[F1IDX(c, k, j, i)] + 0.5 * b1[F1IDX(c, k, j, i + 1)];
          bcc[IB2 * nc + id] = 0.5 * b2[F2IDX(c, k, j, i)] + 0.5 * b2[F2IDX(c, k, j + 1, i)];
          bcc[IB3 * nc + id] = 0.5 * b3[F3IDX(c, k, j, i)] + 0.5 * b3[F3IDX(c, k + 1, j, i)];
        }
    for (int k = 0; k < c->n3; ++k)
      for (int j = 0; j < c->n2; ++j)
        for (int i = 0; i < c->n1; ++i) {
          int64_t id = CIDX(c, k, j, i);
          double u_d = u[IDN * nc + id];
          double u_m1 = u[IVX * nc + id], u_m2 = u[IVY * nc + id], u_m3 = u[IVZ * nc + id];
          double u_e = u[IEN * nc + id];
          u_d = (u_d > dfl) ? u_d : dfl;
          u[IDN * nc + id] = u_d;
          w[IDN * nc + id] = u_d;
          double di = 1.0 / u_d;
          w[IVX * nc + id] = u_m1 * di;
          w[IVY * nc + id] = u_m2 * di;
          w[IVZ * nc + id] = u_m3 * di;
          if (c->iso) continue;   /* eos/isothermal_mhd.cpp:40-73: no energy */
          double bcc1 = bcc[IB1 * nc + id], bcc2 = bcc[IB2 * nc + id], bcc3 = bcc[IB3 * nc + id];
          double pb = 0.5 * (SQR(bcc1) + SQR(bcc2) + SQR(bcc3));
          double e_k = 0.5 * di * (SQR(u_m1) + SQR(u_m2) + SQR(u_m3));
          double w_p = gm1 * (u_e - e_k - pb);
          u_e = (w_p > pfl) ? u_e : ((pfl / gm1) + e_k + pb);
          w_p = (w_p > pfl) ? w_p : pfl;
          u[IEN * nc + id] = u_e;
          w[IPR * nc + id] = w_p;
        }
  }
}

/* ---------------------------------------------------------------- Riemann solvers */
/* Hydro::GetWeightForCT, hydro/hydro.cpp:154-158 */
static inline double ct_weight(double dflx, double rhol, double rhor, double dx, double dt) {
  double v_over_c = (1024.0) * dt * dflx / (dx * (rhol + rhor));
  double tmp_min = dmin(0.5, v_over_c);
  return 0.5 + dmax(-0.5, tmp_min);
}

typedef struct { double d, mx, my, mz, e, by, bz; } cons1d;

/* HLLD for adiabatic MHD, hydro/rsolvers/mhd/hlld.cpp:38-382 (one interface). */
static void hlld(double gamma, const double *wli, const double *wri, double bxi,
                 double *flxi) {
  double igm1 = 1.0 / (gamma - 1.0);
  cons1d ul, ur, ulst, uldst, urdst, urst, fl, fr;
  double spd[5];
  double bxsq = bxi * bxi;
  double pbl = 0.5 * (bxsq + (SQR(wli[IBY]) + SQR(wli[IBZ])));
  double pbr = 0.5 * (bxsq + (SQR(wri[IBY]) + SQR(wri[IBZ])));
  double kel = 0.5 * wli[IDN] * (SQR(wli[IVX]) + (SQR(wli[IVY]) + SQR(wli[IVZ])));
  double ker = 0.5 * wri[IDN] * (SQR(wri[IVX]) + (SQR(wri[IVY]) + SQR(wri[IVZ])));
  ul.d = wli[IDN];
  ul.mx = wli[IVX] * ul.d;
  ul.my = wli[IVY] * ul.d;
  ul.mz = wli[IVZ] * ul.d;
  ul.e = wli[IPR] * igm1 + kel + pbl;
  ul.by = wli[IBY];
  ul.bz = wli[IBZ];
  ur.d = wri[IDN];
  ur.mx = wri[IVX] * ur.d;
  ur.my = wri[IVY] * ur.d;
  ur.mz = wri[IVZ] * ur.d;
  ur.e = wri[IPR] * igm1 + ker + pbr;
  ur.by = wri[IBY];
  ur.bz = wri[IBZ];

  double cfl = fast_speed(gamma, wli, bxi);
  double cfr = fast_speed(gamma, wri, bxi);
  spd[0] = dmin(wli[IVX] - cfl, wri[IVX] - cfr);
  spd[4] = dmax(wli[IVX] + cfl, wri[IVX] + cfr);

  double ptl = wli[IPR] + pbl;
  double ptr = wri[IPR] + pbr;
  fl.d = ul.mx;
  fl.mx = ul.mx * wli[IVX] + ptl - bxsq;
  fl.my = ul.my * wli[IVX] - bxi * ul.by;
  fl.mz = ul.mz * wli[IVX] - bxi * ul.bz;
  fl.e = wli[IVX] * (ul.e + ptl - bxsq) - bxi * (wli[IVY] * ul.by + wli[IVZ] * ul.bz);
  fl.by = ul.by * wli[IVX] - bxi * wli[IVY];
  fl.bz = ul.bz * wli[IVX] - bxi * wli[IVZ];
  fr.d = ur.mx;
  fr.mx = ur.mx * wri[IVX] + ptr - bxsq;
  fr.my = ur.my * wri[IVX] - bxi * ur.by;
  fr.mz = ur.mz * wri[IVX] - bxi * ur.bz;
  fr.e = wri[IVX] * (ur.e + ptr - bxsq) - bxi * (wri[IVY] * ur.by + wri[IVZ] * ur.bz);
  fr.by = ur.by * wri[IVX] - bxi * wri[IVY];
  fr.bz = ur.bz * wri[IVX] - bxi * wri[IVZ];

  double sdl = spd[0] - wli[IVX];
  double sdr = spd[4] - wri[IVX];
  spd[2] = (sdr * ur.mx - sdl * ul.mx + (ptl - ptr)) / (sdr * ur.d - sdl * ul.d);
  double sdml = spd[0] - spd[2];
  double sdmr = spd[4] - spd[2];
  double sdml_inv = 1.0 / sdml;
  double sdmr_inv = 1.0 / sdmr;
  ulst.d = ul.d * sdl * sdml_inv;
  urst.d = ur.d * sdr * sdmr_inv;
  double ulst_d_inv = 1.0 / ulst.d;
  double urst_d_inv = 1.0 / urst.d;
  double sqrtdl = sqrt(ulst.d);
  double sqrtdr = sqrt(urst.d);
  spd[1] = spd[2] - fabs(bxi) / sqrtdl;
  spd[3] = spd[2] + fabs(bxi) / sqrtdr;

  double ptstl = ptl + ul.d * sdl * (spd[2] - wli[IVX]);
  double ptstr = ptr + ur.d * sdr * (spd[2] - wri[IVX]);
  double ptst = 0.5 * (ptstr + ptstl);

  ulst.mx = ulst.d * spd[2];
  if (fabs(ul.d * sdl * sdml - bxsq) < (SMALL_NUMBER) * ptst) {
    ulst.my = ulst.d * wli[IVY];
    ulst.mz = ulst.d * wli[IVZ];
    ulst.by = ul.by;
    ulst.bz = ul.bz;
  } else {
    double tmp = bxi * (sdl - sdml) / (ul.d * sdl * sdml - bxsq);
    ulst.my = ulst.d * (wli[IVY] - ul.by * tmp);
    ulst.mz = ulst.d * (wli[IVZ] - ul.bz * tmp);
    tmp = (ul.d * SQR(sdl) - bxsq) / (ul.d * sdl * sdml - bxsq);
    ulst.by = ul.by * tmp;
    ulst.bz = ul.bz * tmp;
  }
  double vbstl = (ulst.mx * bxi + (ulst.my * ulst.by + ulst.mz * ulst.bz)) * ulst_d_inv;
  ulst.e = (sdl * ul.e - ptl * wli[IVX] + ptst * spd[2] +
            bxi * (wli[IVX] * bxi + (wli[IVY] * ul.by + wli[IVZ] * ul.bz) - vbstl)) * sdml_inv;

  urst.mx = urst.d * spd[2];
  if (fabs(ur.d * sdr * sdmr - bxsq) < (SMALL_NUMBER) * ptst) {
    urst.my = urst.d * wri[IVY];
    urst.mz = urst.d * wri[IVZ];
    urst.by = ur.by;
    urst.bz = ur.bz;
  } else {
    double tmp = bxi * (sdr - sdmr) / (ur.d * sdr * sdmr - bxsq);
    urst.my = urst.d * (wri[IVY] - ur.by * tmp);
    urst.mz = urst.d * (wri[IVZ] - ur.bz * tmp);
    tmp = (ur.d * SQR(sdr) - bxsq) / (ur.d * sdr * sdmr - bxsq);
    urst.by = ur.by * tmp;
    urst.bz = ur.bz * tmp;
  }
  double vbstr = (urst.mx * bxi + (urst.my * urst.by + urst.mz * urst.bz)) * urst_d_inv;
  urst.e = (sdr * ur.e - ptr * wri[IVX] + ptst * spd[2] +
            bxi * (wri[IVX] * bxi + (wri[IVY] * ur.by + wri[IVZ] * ur.bz) - vbstr)) * sdmr_inv;

  if (0.5 * bxsq < (SMALL_NUMBER) * ptst) {
    uldst = ulst;
    urdst = urst;
  } else {
    double invsumd = 1.0 / (sqrtdl + sqrtdr);
    double bxsig = (bxi > 0.0 ? 1.0 : -1.0);
    uldst.d = ulst.d;
    urdst.d = urst.d;
    uldst.mx = ulst.mx;
    urdst.mx = urst.mx;
    double tmp = invsumd * (sqrtdl * (ulst.my * ulst_d_inv) + sqrtdr * (urst.my * urst_d_inv) +
                            bxsig * (urst.by - ulst.by));
    uldst.my = uldst.d * tmp;
    urdst.my = urdst.d * tmp;
    tmp = invsumd * (sqrtdl * (ulst.mz * ulst_d_inv) + sqrtdr * (urst.mz * urst_d_inv) +
                     bxsig * (urst.bz - ulst.bz));
    uldst.mz = uldst.d * tmp;
    urdst.mz = urdst.d * tmp;
    tmp = invsumd * (sqrtdl * urst.by + sqrtdr * ulst.by +
                     bxsig * sqrtdl * sqrtdr * ((urst.my * urst_d_inv) - (ulst.my * ulst_d_inv)));
    uldst.by = urdst.by = tmp;
    tmp = invsumd * (sqrtdl * urst.bz + sqrtdr * ulst.bz +
                     bxsig * sqrtdl * sqrtdr * ((urst.mz * urst_d_inv) - (ulst.mz * ulst_d_inv)));
    uldst.bz = urdst.bz = tmp;
    tmp = spd[2] * bxi + (uldst.my * uldst.by + uldst.mz * uldst.bz) / uldst.d;
    uldst.e = ulst.e - sqrtdl * bxsig * (vbstl - tmp);
    urdst.e = urst.e + sqrtdr * bxsig * (vbstr - tmp);
  }

  uldst.d = spd[1] * (uldst.d - ulst.d);
  uldst.mx = spd[1] * (uldst.mx - ulst.mx);
  uldst.my = spd[1] * (uldst.my - ulst.my);
  uldst.mz = spd[1] * (uldst.mz - ulst.mz);
  uldst.e = spd[1] * (uldst.e - ulst.e);
  uldst.by = spd[1] * (uldst.by - ulst.by);
  uldst.bz = spd[1] * (uldst.bz - ulst.bz);
  ulst.d = spd[0] * (ulst.d - ul.d);
  ulst.mx = spd[0] * (ulst.mx - ul.mx);
  ulst.my = spd[0] * (ulst.my - ul.my);
  ulst.mz = spd[0] * (ulst.mz - ul.mz);
  ulst.e = spd[0] * (ulst.e - ul.e);
  ulst.by = spd[0] * (ulst.by - ul.by);
  ulst.bz = spd[0] * (ulst.bz - ul.bz);
  urdst.d = spd[3] * (urdst.d - urst.d);
  urdst.mx = spd[3] * (urdst.mx - urst.mx);
  urdst.my = spd[3] * (urdst.my - urst.my);
  urdst.mz = spd[3] * (urdst.mz - urst.mz);
  urdst.e = spd[3] * (urdst.e - urst.e);
  urdst.by = spd[3] * (urdst.by - urst.by);
  urdst.bz = spd[3] * (urdst.bz - urst.bz);
  urst.d = spd[4] * (urst.d - ur.d);
  urst.mx = spd[4] * (urst.mx - ur.mx);
  urst.my = spd[4] * (urst.my - ur.my);
  urst.mz = spd[4] * (urst.mz - ur.mz);
  urst.e = spd[4] * (urst.e - ur.e);
  urst.by = spd[4] * (urst.by - ur.by);
  urst.bz = spd[4] * (urst.bz - ur.bz);

  const cons1d *f;
  if (spd[0] >= 0.0) {
    f = &fl;
    flxi[IDN] = f->d; flxi[IVX] = f->mx; flxi[IVY] = f->my; flxi[IVZ] = f->mz;
    flxi[IEN] = f->e; flxi[IBY] = f->by; flxi[IBZ] = f->bz;
  } else if (spd[4] <= 0.0) {
    f = &fr;
    flxi[IDN] = f->d; flxi[IVX] = f->mx; flxi[IVY] = f->my; flxi[IVZ] = f->mz;
    flxi[IEN] = f->e; flxi[IBY] = f->by; flxi[IBZ] = f->bz;
  } else if (spd[1] >= 0.0) {
    flxi[IDN] = fl.d + ulst.d;
    flxi[IVX] = fl.mx + ulst.mx;
    flxi[IVY] = fl.my + ulst.my;
    flxi[IVZ] = fl.mz + ulst.mz;
    flxi[IEN] = fl.e + ulst.e;
    flxi[IBY] = fl.by + ulst.by;
    flxi[IBZ] = fl.bz + ulst.bz;
  } else if (spd[2] >= 0.0) {
    flxi[IDN] = fl.d + ulst.d + uldst.d;
    flxi[IVX] = fl.mx + ulst.mx + uldst.mx;
    flxi[IVY] = fl.my + ulst.my + uldst.my;
    flxi[IVZ] = fl.mz + ulst.mz + uldst.mz;
    flxi[IEN] = fl.e + ulst.e + uldst.e;
    flxi[IBY] = fl.by + ulst.by + uldst.by;
    flxi[IBZ] = fl.bz + ulst.bz + uldst.bz;
  } else if (spd[3] > 0.0) {
    flxi[IDN] = fr.d + urst.d + urdst.d;
    flxi[IVX] = fr.mx + urst.mx + urdst.mx;
    flxi[IVY] = fr.my + urst.my + urdst.my;
    flxi[IVZ] = fr.mz + urst.mz + urdst.mz;
    flxi[IEN] = fr.e + urst.e + urdst.e;
    flxi[IBY] = fr.by + urst.by + urdst.by;
    flxi[IBZ] = fr.bz + urst.bz + urdst.bz;
  } else {
    flxi[IDN] = fr.d + urst.d;
    flxi[IVX] = fr.mx + urst.mx;
    flxi[IVY] = fr.my + urst.my;
    flxi[IVZ] = fr.mz + urst.mz;
    flxi[IEN] = fr.e + urst.e;
    flxi[IBY] = fr.by + urst.by;
    flxi[IBZ] = fr.bz + urst.bz;
  }
}

/* Low-dissipation HLLD (Minoshima et al. 2021) for adiabatic MHD,
 * hydro/rsolvers/mhd/lhlld.cpp:38-389 (one interface).  dvn, dvt are the normal and
 * transverse velocity differences of Hydro::CalculateVelocityDifferences at this face
 * (the shock detector, lhlld.cpp:159-161).  Restated literally: every star and
 * double-star difference is formed before the flux is selected (lhlld.cpp:287-378). */
static void lhlld(double gamma, const double *wli, const double *wri, double bxi, double dvn,
                  double dvt, double *flxi) {
  double igm1 = 1.0 / (gamma - 1.0);
  cons1d ul, ur, ulst, uldst, urdst, urst, fl, fr;
  double spd[5];
  double bxsq = bxi * bxi;
  double pbl = 0.5 * (bxsq + (SQR(wli[IBY]) + SQR(wli[IBZ])));
  double pbr = 0.5 * (bxsq + (SQR(wri[IBY]) + SQR(wri[IBZ])));
  double kel = 0.5 * wli[IDN] * (SQR(wli[IVX]) + (SQR(wli[IVY]) + SQR(wli[IVZ])));
  double ker = 0.5 * wri[IDN] * (SQR(wri[IVX]) + (SQR(wri[IVY]) + SQR(wri[IVZ])));
  ul.d = wli[IDN];
  ul.mx = wli[IVX] * ul.d;
  ul.my = wli[IVY] * ul.d;
  ul.mz = wli[IVZ] * ul.d;
  ul.e = wli[IPR] * igm1 + kel + pbl;
  ul.by = wli[IBY];
  ul.bz = wli[IBZ];
  ur.d = wri[IDN];
  ur.mx = wri[IVX] * ur.d;
  ur.my = wri[IVY] * ur.d;
  ur.mz = wri[IVZ] * ur.d;
  ur.e = wri[IPR] * igm1 + ker + pbr;
  ur.by = wri[IBY];
  ur.bz = wri[IBZ];
  /* Step 2: speeds (lhlld.cpp:115-120) */
  double cfl = fast_speed(gamma, wli, bxi);
  double cfr = fast_speed(gamma, wri, bxi);
  spd[0] = dmin(wli[IVX] - cfl, wri[IVX] - cfr);
  spd[4] = dmax(wli[IVX] + cfl, wri[IVX] + cfr);
  double cfmax = dmax(cfl, cfr);
  /* Step 3: L/R fluxes (lhlld.cpp:130-148) */
  double ptl = wli[IPR] + pbl;
  double ptr = wri[IPR] + pbr;
  fl.d = ul.mx;
  fl.mx = ul.mx * wli[IVX] + ptl - bxsq;
  fl.my = ul.my * wli[IVX] - bxi * ul.by;
  fl.mz = ul.mz * wli[IVX] - bxi * ul.bz;
  fl.e = wli[IVX] * (ul.e + ptl - bxsq) - bxi * (wli[IVY] * ul.by + wli[IVZ] * ul.bz);
  fl.by = ul.by * wli[IVX] - bxi * wli[IVY];
  fl.bz = ul.bz * wli[IVX] - bxi * wli[IVZ];
  fr.d = ur.mx;
  fr.mx = ur.mx * wri[IVX] + ptr - bxsq;
  fr.my = ur.my * wri[IVX] - bxi * ur.by;
  fr.mz = ur.mz * wri[IVX] - bxi * ur.bz;
  fr.e = wri[IVX] * (ur.e + ptr - bxsq) - bxi * (wri[IVY] * ur.by + wri[IVZ] * ur.bz);
  fr.by = ur.by * wri[IVX] - bxi * wri[IVY];
  fr.bz = ur.bz * wri[IVX] - bxi * wri[IVZ];
  /* Step 4: middle and Alfven speeds with the shock detector (lhlld.cpp:150-178) */
  double sdl = spd[0] - wli[IVX];
  double sdr = spd[4] - wri[IVX];
  double sdld = sdl * ul.d;
  double sdrd = sdr * ur.d;
  double th1 = dmin(1.0, (cfmax - dmin(dvn, 0.0)) / (cfmax - dmin(dvt, 0.0)));
  double th = th1 * th1 * th1 * th1;
  spd[2] = (sdr * ur.mx - sdl * ul.mx + th * (ptl - ptr)) / (sdrd - sdld);
  double sdml = spd[0] - spd[2];
  double sdmr = spd[4] - spd[2];
  double sdml_inv = 1.0 / sdml;
  double sdmr_inv = 1.0 / sdmr;
  ulst.d = sdld * sdml_inv;
  urst.d = sdrd * sdmr_inv;
  double ulst_d_inv = 1.0 / ulst.d;
  double urst_d_inv = 1.0 / urst.d;
  double sqrtdl = sqrt(ulst.d);
  double sqrtdr = sqrt(urst.d);
  spd[1] = spd[2] - fabs(bxi) / sqrtdl;
  spd[3] = spd[2] + fabs(bxi) / sqrtdr;
  /* Step 5: intermediate states; the total pressure with the low-Mach factor phi
   * (lhlld.cpp:182-187) */
  double clsq = ((pbl + kel) + sqrt(SQR(pbl + kel) - 2.0 * kel * bxsq)) / ul.d;
  double crsq = ((pbr + ker) + sqrt(SQR(pbr + ker) - 2.0 * ker * bxsq)) / ur.d;
  double chi = dmin(1.0, sqrt(dmax(clsq, crsq)) / cfmax);
  double phi = chi * (2.0 - chi);
  double ptst = (sdrd * ptl - sdld * ptr + phi * sdrd * sdld * (wri[IVX] - wli[IVX])) / (sdrd - sdld);
  ulst.mx = ulst.d * spd[2];
  if (fabs(sdld * sdml - bxsq) < SMALL_NUMBER * ptst) {
    ulst.my = ulst.d * wli[IVY];
    ulst.mz = ulst.d * wli[IVZ];
    ulst.by = ul.by;
    ulst.bz = ul.bz;
  } else {
    double tmp = bxi * (sdl - sdml) / (sdld * sdml - bxsq);
    ulst.my = ulst.d * (wli[IVY] - ul.by * tmp);
    ulst.mz = ulst.d * (wli[IVZ] - ul.bz * tmp);
    tmp = (sdld * sdl - bxsq) / (sdld * sdml - bxsq);
    ulst.by = ul.by * tmp;
    ulst.bz = ul.bz * tmp;
  }
  double vbstl = (ulst.mx * bxi + (ulst.my * ulst.by + ulst.mz * ulst.bz)) * ulst_d_inv;
  ulst.e = (sdl * ul.e - ptl * wli[IVX] + ptst * spd[2] +
            bxi * (wli[IVX] * bxi + (wli[IVY] * ul.by + wli[IVZ] * ul.bz) - vbstl)) * sdml_inv;
  urst.mx = urst.d * spd[2];
  if (fabs(sdrd * sdmr - bxsq) < SMALL_NUMBER * ptst) {
    urst.my = urst.d * wri[IVY];
    urst.mz = urst.d * wri[IVZ];
    urst.by = ur.by;
    urst.bz = ur.bz;
  } else {
    double tmp = bxi * (sdr - sdmr) / (sdrd * sdmr - bxsq);
    urst.my = urst.d * (wri[IVY] - ur.by * tmp);
    urst.mz = urst.d * (wri[IVZ] - ur.bz * tmp);
    tmp = (sdrd * sdr - bxsq) / (sdrd * sdmr - bxsq);
    urst.by = ur.by * tmp;
    urst.bz = ur.bz * tmp;
  }
  double vbstr = (urst.mx * bxi + (urst.my * urst.by + urst.mz * urst.bz)) * urst_d_inv;
  urst.e = (sdr * ur.e - ptr * wri[IVX] + ptst * spd[2] +
            bxi * (wri[IVX] * bxi + (wri[IVY] * ur.by + wri[IVZ] * ur.bz) - vbstr)) * sdmr_inv;
  if (0.5 * bxsq < SMALL_NUMBER * ptst) {
    uldst = ulst;
    urdst = urst;
  } else {
    double invsumd = 1.0 / (sqrtdl + sqrtdr);
    double bxsig = (bxi > 0.0 ? 1.0 : -1.0);
    uldst.d = ulst.d;
    urdst.d = urst.d;
    uldst.mx = ulst.mx;
    urdst.mx = urst.mx;
    double tmp = invsumd * (sqrtdl * (ulst.my * ulst_d_inv) + sqrtdr * (urst.my * urst_d_inv) +
                            bxsig * (urst.by - ulst.by));
    uldst.my = uldst.d * tmp;
    urdst.my = urdst.d * tmp;
    tmp = invsumd * (sqrtdl * (ulst.mz * ulst_d_inv) + sqrtdr * (urst.mz * urst_d_inv) +
                     bxsig * (urst.bz - ulst.bz));
    uldst.mz = uldst.d * tmp;
    urdst.mz = urdst.d * tmp;
    tmp = invsumd * (sqrtdl * urst.by + sqrtdr * ulst.by +
                     bxsig * sqrtdl * sqrtdr * ((urst.my * urst_d_inv) - (ulst.my * ulst_d_inv)));
    uldst.by = urdst.by = tmp;
    tmp = invsumd * (sqrtdl * urst.bz + sqrtdr * ulst.bz +
                     bxsig * sqrtdl * sqrtdr * ((urst.mz * urst_d_inv) - (ulst.mz * ulst_d_inv)));
    uldst.bz = urdst.bz = tmp;
    tmp = spd[2] * bxi + (uldst.my * uldst.by + uldst.mz * uldst.bz) / uldst.d;
    uldst.e = ulst.e - sqrtdl * bxsig * (vbstl - tmp);
    urdst.e = urst.e + sqrtdr * bxsig * (vbstr - tmp);
  }
  /* Step 6: flux differences of every state, then the selection (lhlld.cpp:287-378) */
  cons1d dld, dl, drd, dr;
  dld.d = spd[1] * (uldst.d - ulst.d); dld.mx = spd[1] * (uldst.mx - ulst.mx);
  dld.my = spd[1] * (uldst.my - ulst.my); dld.mz = spd[1] * (uldst.mz - ulst.mz);
  dld.e = spd[1] * (uldst.e - ulst.e); dld.by = spd[1] * (uldst.by - ulst.by);
  dld.bz = spd[1] * (uldst.bz - ulst.bz);
  dl.d = spd[0] * (ulst.d - ul.d); dl.mx = spd[0] * (ulst.mx - ul.mx);
  dl.my = spd[0] * (ulst.my - ul.my); dl.mz = spd[0] * (ulst.mz - ul.mz);
  dl.e = spd[0] * (ulst.e - ul.e); dl.by = spd[0] * (ulst.by - ul.by);
  dl.bz = spd[0] * (ulst.bz - ul.bz);
  drd.d = spd[3] * (urdst.d - urst.d); drd.mx = spd[3] * (urdst.mx - urst.mx);
  drd.my = spd[3] * (urdst.my - urst.my); drd.mz = spd[3] * (urdst.mz - urst.mz);
  drd.e = spd[3] * (urdst.e - urst.e); drd.by = spd[3] * (urdst.by - urst.by);
  drd.bz = spd[3] * (urdst.bz - urst.bz);
  dr.d = spd[4] * (urst.d - ur.d); dr.mx = spd[4] * (urst.mx - ur.mx);
  dr.my = spd[4] * (urst.my - ur.my); dr.mz = spd[4] * (urst.mz - ur.mz);
  dr.e = spd[4] * (urst.e - ur.e); dr.by = spd[4] * (urst.by - ur.by);
  dr.bz = spd[4] * (urst.bz - ur.bz);
  const cons1d *f;
  if (spd[0] >= 0.0 || spd[4] <= 0.0) {
    f = (spd[0] >= 0.0) ? &fl : &fr;
    flxi[IDN] = f->d; flxi[IVX] = f->mx; flxi[IVY] = f->my; flxi[IVZ] = f->mz;
    flxi[IEN] = f->e; flxi[IBY] = f->by; flxi[IBZ] = f->bz;
  } else if (spd[1] >= 0.0) {
    flxi[IDN] = fl.d + dl.d; flxi[IVX] = fl.mx + dl.mx; flxi[IVY] = fl.my + dl.my;
    flxi[IVZ] = fl.mz + dl.mz; flxi[IEN] = fl.e + dl.e; flxi[IBY] = fl.by + dl.by;
    flxi[IBZ] = fl.bz + dl.bz;
  } else if (spd[2] >= 0.0) {
    flxi[IDN] = fl.d + dl.d + dld.d; flxi[IVX] = fl.mx + dl.mx + dld.mx;
    flxi[IVY] = fl.my + dl.my + dld.my; flxi[IVZ] = fl.mz + dl.mz + dld.mz;
    flxi[IEN] = fl.e + dl.e + dld.e; flxi[IBY] = fl.by + dl.by + dld.by;
    flxi[IBZ] = fl.bz + dl.bz + dld.bz;
  } else if (spd[3] > 0.0) {
    flxi[IDN] = fr.d + dr.d + drd.d; flxi[IVX] = fr.mx + dr.mx + drd.mx;
    flxi[IVY] = fr.my + dr.my + drd.my; flxi[IVZ] = fr.mz + dr.mz + drd.mz;
    flxi[IEN] = fr.e + dr.e + drd.e; flxi[IBY] = fr.by + dr.by + drd.by;
    flxi[IBZ] = fr.bz + dr.bz + drd.bz;
  } else {
    flxi[IDN] = fr.d + dr.d; flxi[IVX] = fr.mx + dr.mx; flxi[IVY] = fr.my + dr.my;
    flxi[IVZ] = fr.mz + dr.mz; flxi[IEN] = fr.e + dr.e; flxi[IBY] = fr.by + dr.by;
    flxi[IBZ] = fr.bz + dr.bz;
  }
}

/* Hydro::CalculateVelocityDifferences, 3D branch (calculate_velocity_differences.cpp:
 * 23-68), at the face below cell (k,j,i) along dir: the normal velocity jump dvn and the
 * smallest one-sided transverse velocity difference dvt of the two cells beside the face. */
static void velocity_differences(const orc_ctx *c, const double *w, int dir, int k, int j, int i,
                                 double *dvn, double *dvt) {
  const int64_t nc = c->ncell;
#define WV(n, kk, jj, ii) w[(n) * nc + CIDX(c, kk, jj, ii)]
  if (dir == 0) {
    *dvn = WV(IVX, k, j, i) - WV(IVX, k, j, i - 1);
    double dvl = dmin(WV(IVY, k, j + 1, i - 1) - WV(IVY, k, j, i - 1), WV(IVY, k, j, i - 1) - WV(IVY, k, j - 1, i - 1));
    double dvr = dmin(WV(IVY, k, j + 1, i) - WV(IVY, k, j, i), WV(IVY, k, j, i) - WV(IVY, k, j - 1, i));
    double dwl = dmin(WV(IVZ, k + 1, j, i - 1) - WV(IVZ, k, j, i - 1), WV(IVZ, k, j, i - 1) - WV(IVZ, k - 1, j, i - 1));
    double dwr = dmin(WV(IVZ, k + 1, j, i) - WV(IVZ, k, j, i), WV(IVZ, k, j, i) - WV(IVZ, k - 1, j, i));
    *dvt = dmin(dmin(dvl, dvr), dmin(dwl, dwr));
  } else if (dir == 1) {
    *dvn = WV(IVY, k, j, i) - WV(IVY, k, j - 1, i);
    double dvl = dmin(WV(IVZ, k + 1, j - 1, i) - WV(IVZ, k, j - 1, i), WV(IVZ, k, j - 1, i) - WV(IVZ, k - 1, j - 1, i));
    double dvr = dmin(WV(IVZ, k + 1, j, i) - WV(IVZ, k, j, i), WV(IVZ, k, j, i) - WV(IVZ, k - 1, j, i));
    double dwl = dmin(WV(IVX, k, j - 1, i + 1) - WV(IVX, k, j - 1, i), WV(IVX, k, j - 1, i) - WV(IVX, k, j - 1, i - 1));
    double dwr = dmin(WV(IVX, k, j, i + 1) - WV(IVX, k, j, i), WV(IVX, k, j, i) - WV(IVX, k, j, i - 1));
    *dvt = dmin(dmin(dvl, dvr), dmin(dwl, dwr));
  } else {
    *dvn = WV(IVZ, k, j, i) - WV(IVZ, k - 1, j, i);
    double dvl = dmin(WV(IVX, k - 1, j, i + 1) - WV(IVX, k - 1, j, i), WV(IVX, k - 1, j, i) - WV(IVX, k - 1, j, i - 1));
    double dvr = dmin(WV(IVX, k, j, i + 1) - WV(IVX, k, j, i), WV(IVX, k, j, i) - WV(IVX, k, j, i - 1));
    double dwl = dmin(WV(IVY, k - 1, j + 1, i) - WV(IVY, k - 1, j, i), WV(IVY, k - 1, j, i) - WV(IVY, k - 1, j - 1, i));
    double dwr = dmin(WV(IVY, k, j + 1, i) - WV(IVY, k, j, i), WV(IVY, k, j, i) - WV(IVY, k, j - 1, i));
    *dvt = dmin(dmin(dvl, dvr), dmin(dwl, dwr));
  }
#undef WV
}

/* HLLE for adiabatic MHD, hydro/rsolvers/mhd/hlle_mhd.cpp:25-182 (one interface). */
static void hlle(double gamma, int iso, double iso_cs, const double *wli, const double *wri,
                 double bxi, double *flxi) {
  double gm1 = gamma - 1.0;
  double wroe[NWAVE], fl[NWAVE], fr[NWAVE];
  double sqrtdl = sqrt(wli[IDN]);
  double sqrtdr = sqrt(wri[IDN]);
  double isdlpdr = 1.0 / (sqrtdl + sqrtdr);
  wroe[IDN] = sqrtdl * sqrtdr;
  wroe[IVX] = (sqrtdl * wli[IVX] + sqrtdr * wri[IVX]) * isdlpdr;
  wroe[IVY] = (sqrtdl * wli[IVY] + sqrtdr * wri[IVY]) * isdlpdr;
  wroe[IVZ] = (sqrtdl * wli[IVZ] + sqrtdr * wri[IVZ]) * isdlpdr;
  wroe[IBY] = (sqrtdr * wli[IBY] + sqrtdl * wri[IBY]) * isdlpdr;
  wroe[IBZ] = (sqrtdr * wli[IBZ] + sqrtdl * wri[IBZ]) * isdlpdr;
  double x = 0.5 * (SQR(wli[IBY] - wri[IBY]) + SQR(wli[IBZ] - wri[IBZ])) / (SQR(sqrtdl + sqrtdr));
  double y = 0.5 * (wli[IDN] + wri[IDN]) / wroe[IDN];
  double pbl = 0.5 * (bxi * bxi + SQR(wli[IBY]) + SQR(wli[IBZ]));
  double pbr = 0.5 * (bxi * bxi + SQR(wri[IBY]) + SQR(wri[IBZ]));
  double el = 0.0, er = 0.0, hroe = 0.0;
  if (!iso) {   /* hlle_mhd.cpp:83-87 */
    el = wli[IPR] / gm1 + 0.5 * wli[IDN] * (SQR(wli[IVX]) + SQR(wli[IVY]) + SQR(wli[IVZ])) + pbl;
    er = wri[IPR] / gm1 + 0.5 * wri[IDN] * (SQR(wri[IVX]) + SQR(wri[IVY]) + SQR(wri[IVZ])) + pbr;
    hroe = ((el + wli[IPR] + pbl) / sqrtdl + (er + wri[IPR] + pbr) / sqrtdr) * isdlpdr;
  }
  double cl = iso ? fast_speed_iso(iso_cs, wli, bxi) : fast_speed(gamma, wli, bxi);
  double cr = iso ? fast_speed_iso(iso_cs, wri, bxi) : fast_speed(gamma, wri, bxi);
  double btsq = SQR(wroe[IBY]) + SQR(wroe[IBZ]);
  double vaxsq = bxi * bxi / wroe[IDN];
  double bt_starsq, twid_asq;
  if (!iso) {   /* hlle_mhd.cpp:98-106: eq. B18 (adiabatic) or B39 (isothermal) */
    bt_starsq = (gm1 - (gm1 - 1.0) * y) * btsq;
    double hp = hroe - (vaxsq + btsq / wroe[IDN]);
    double vsq = SQR(wroe[IVX]) + SQR(wroe[IVY]) + SQR(wroe[IVZ]);
    twid_asq = dmax((gm1 * (hp - 0.5 * vsq) - (gm1 - 1.0) * x), 0.0);
  } else {
    bt_starsq = btsq * y;
    twid_asq = iso_cs * iso_cs + x;
  }
  double ct2 = bt_starsq / wroe[IDN];
  double tsum = vaxsq + ct2 + twid_asq;
  double tdif = vaxsq + ct2 - twid_asq;
  double cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
  double cfsq = 0.5 * (tsum + cf2_cs2);
  double a = sqrt(cfsq);
  double al = dmin((wroe[IVX] - a), (wli[IVX] - cl));
  double ar = dmax((wroe[IVX] + a), (wri[IVX] + cr));
  double bp = ar > 0.0 ? ar : 0.0;
  double bm = al < 0.0 ? al : 0.0;
  double vxl = wli[IVX] - bm;
  double vxr = wri[IVX] - bp;
  fl[IDN] = wli[IDN] * vxl;
  fr[IDN] = wri[IDN] * vxr;
  fl[IVX] = wli[IDN] * wli[IVX] * vxl + pbl - SQR(bxi);
  fr[IVX] = wri[IDN] * wri[IVX] * vxr + pbr - SQR(bxi);
  fl[IVY] = wli[IDN] * wli[IVY] * vxl - bxi * wli[IBY];
  fr[IVY] = wri[IDN] * wri[IVY] * vxr - bxi * wri[IBY];
  fl[IVZ] = wli[IDN] * wli[IVZ] * vxl - bxi * wli[IBZ];
  fr[IVZ] = wri[IDN] * wri[IVZ] * vxr - bxi * wri[IBZ];
  if (!iso) {   /* hlle_mhd.cpp:140-150 */
    fl[IVX] += wli[IPR];
    fr[IVX] += wri[IPR];
    fl[IEN] = el * vxl + wli[IVX] * (wli[IPR] + pbl - bxi * bxi);
    fr[IEN] = er * vxr + wri[IVX] * (wri[IPR] + pbr - bxi * bxi);
    fl[IEN] -= bxi * (wli[IBY] * wli[IVY] + wli[IBZ] * wli[IVZ]);
    fr[IEN] -= bxi * (wri[IBY] * wri[IVY] + wri[IBZ] * wri[IVZ]);
  } else {
    fl[IVX] += (iso_cs * iso_cs) * wli[IDN];
    fr[IVX] += (iso_cs * iso_cs) * wri[IDN];
    fl[IEN] = 0.0;
    fr[IEN] = 0.0;
  }
  fl[IBY] = wli[IBY] * vxl - bxi * wli[IVY];
  fr[IBY] = wri[IBY] * vxr - bxi * wri[IVY];
  fl[IBZ] = wli[IBZ] * vxl - bxi * wli[IVZ];
  fr[IBZ] = wri[IBZ] * vxr - bxi * wri[IVZ];
  double tmp = 0.0;
  if (bp != bm) tmp = 0.5 * (bp + bm) / (bp - bm);
  for (int n = 0; n < NWAVE; ++n)
    flxi[n] = 0.5 * (fl[n] + fr[n]) + (fl[n] - fr[n]) * tmp;
}

/* HLLD for isothermal MHD (Mignone 2007) at one interface,
 * hydro/rsolvers/mhd/hlld_iso.cpp:37-254; SMALL_NUMBER = 1e-8 (hlld_iso.cpp:31).  The
 * energy slot of flxi is set to 0 (NHYDRO = 4 in the reference). */
typedef struct { double d, mx, my, mz, by, bz; } cons_iso;

static void hlld_iso(double cs, double dfloor, const double *wli, const double *wri,
                     double bxi, double *flxi) {
  const double SMALL = 1.0e-8;
  cons_iso ul, ur, ulst, urst, ucst, fl, fr;
  double spd[5];
  ul.d = wli[IDN];
  ul.mx = wli[IVX] * ul.d;
  ul.my = wli[IVY] * ul.d;
  ul.mz = wli[IVZ] * ul.d;
  ul.by = wli[IBY];
  ul.bz = wli[IBZ];
  ur.d = wri[IDN];
  ur.mx = wri[IVX] * ur.d;
  ur.my = wri[IVY] * ur.d;
  ur.mz = wri[IVZ] * ur.d;
  ur.by = wri[IBY];
  ur.bz = wri[IBZ];
  /* speeds, Miyoshi & Kusano eqn. (67) (hlld_iso.cpp:97-101) */
  double cfl = fast_speed_iso(cs, wli, bxi);
  double cfr = fast_speed_iso(cs, wri, bxi);
  spd[0] = dmin(wli[IVX] - cfl, wri[IVX] - cfr);
  spd[4] = dmax(wli[IVX] + cfl, wri[IVX] + cfr);
  /* L/R fluxes (hlld_iso.cpp:103-122) */
  double bxsq = bxi * bxi;
  double ptl = SQR(cs) * wli[IDN] + 0.5 * (bxsq + SQR(wli[IBY]) + SQR(wli[IBZ]));
  double ptr = SQR(cs) * wri[IDN] + 0.5 * (bxsq + SQR(wri[IBY]) + SQR(wri[IBZ]));
  fl.d = ul.mx;
  fl.mx = ul.mx * wli[IVX] + ptl - bxsq;
  fl.my = ul.my * wli[IVX] - bxi * ul.by;
  fl.mz = ul.mz * wli[IVX] - bxi * ul.bz;
  fl.by = ul.by * wli[IVX] - bxi * wli[IVY];
  fl.bz = ul.bz * wli[IVX] - bxi * wli[IVZ];
  fr.d = ur.mx;
  fr.mx = ur.mx * wri[IVX] + ptr - bxsq;
  fr.my = ur.my * wri[IVX] - bxi * ur.by;
  fr.mz = ur.mz * wri[IVX] - bxi * ur.bz;
  fr.by = ur.by * wri[IVX] - bxi * wri[IVY];
  fr.bz = ur.bz * wri[IVX] - bxi * wri[IVZ];
  /* HLL averages and Alfven speeds, Mignone eqns. (15), (17), (29) (hlld_iso.cpp:124-150) */
  double idspd = 1.0 / (spd[4] - spd[0]);
  double dhll = (spd[4] * ur.d - spd[0] * ul.d - fr.d + fl.d) * idspd;
  dhll = dmax(dhll, dfloor);
  double sqrtdhll = sqrt(dhll);
  double fdhll = (spd[4] * fl.d - spd[0] * fr.d + spd[4] * spd[0] * (ur.d - ul.d)) * idspd;
  double fmxhll = (spd[4] * fl.mx - spd[0] * fr.mx + spd[4] * spd[0] * (ur.mx - ul.mx)) * idspd;
  double ustar = fdhll / dhll;
  double mxhll = (spd[4] * ur.mx - spd[0] * ul.mx - fr.mx + fl.mx) * idspd;
  spd[1] = ustar - fabs(bxi) / sqrtdhll;
  spd[3] = ustar + fabs(bxi) / sqrtdhll;
  /* Ul*, eqns. (20), (24), (30)-(33) (hlld_iso.cpp:152-173) */
  ulst.d = dhll;
  ulst.mx = mxhll;
  double tmp = (spd[0] - spd[1]) * (spd[0] - spd[3]);
  if (fabs(spd[0] - spd[1]) < (SMALL) * cs) {
    ulst.my = ul.my; ulst.mz = ul.mz; ulst.by = ul.by; ulst.bz = ul.bz;
  } else {
    double mfact = bxi * (ustar - wli[IVX]) / tmp;
    double bfact = (ul.d * SQR(spd[0] - wli[IVX]) - bxsq) / (dhll * tmp);
    ulst.my = dhll * wli[IVY] - ul.by * mfact;
    ulst.mz = dhll * wli[IVZ] - ul.bz * mfact;
    ulst.by = ul.by * bfact;
    ulst.bz = ul.bz * bfact;
  }
  /* Ur* (hlld_iso.cpp:175-195) */
  urst.d = dhll;
  urst.mx = mxhll;
  tmp = (spd[4] - spd[1]) * (spd[4] - spd[3]);
  if (fabs(spd[4] - spd[3]) < (SMALL) * cs) {
    urst.my = ur.my; urst.mz = ur.mz; urst.by = ur.by; urst.bz = ur.bz;
  } else {
    double mfact = bxi * (ustar - wri[IVX]) / tmp;
    double bfact = (ur.d * SQR(spd[4] - wri[IVX]) - bxsq) / (dhll * tmp);
    urst.my = dhll * wri[IVY] - ur.by * mfact;
    urst.mz = dhll * wri[IVZ] - ur.bz * mfact;
    urst.by = ur.by * bfact;
    urst.bz = ur.bz * bfact;
  }
  /* Uc*, eqns. (34)-(37) (hlld_iso.cpp:197-204) */
  double x = sqrtdhll * (bxi > 0.0 ? 1.0 : -1.0);
  ucst.d = dhll;
  ucst.mx = mxhll;
  ucst.my = 0.5 * (ulst.my + urst.my + (urst.by - ulst.by) * x);
  ucst.mz = 0.5 * (ulst.mz + urst.mz + (urst.bz - ulst.bz) * x);
  ucst.by = 0.5 * (ulst.by + urst.by + (urst.my - ulst.my) / x);
  ucst.bz = 0.5 * (ulst.bz + urst.bz + (urst.mz - ulst.mz) / x);
  /* flux selection, eqns. (38a)-(38e) (hlld_iso.cpp:206-245) */
  if (spd[0] >= 0.0) {
    flxi[IDN] = fl.d; flxi[IVX] = fl.mx; flxi[IVY] = fl.my; flxi[IVZ] = fl.mz;
    flxi[IBY] = fl.by; flxi[IBZ] = fl.bz;
  } else if (spd[4] <= 0.0) {
    flxi[IDN] = fr.d; flxi[IVX] = fr.mx; flxi[IVY] = fr.my; flxi[IVZ] = fr.mz;
    flxi[IBY] = fr.by; flxi[IBZ] = fr.bz;
  } else if (spd[1] >= 0.0) {
    flxi[IDN] = fl.d + spd[0] * (ulst.d - ul.d);
    flxi[IVX] = fl.mx + spd[0] * (ulst.mx - ul.mx);
    flxi[IVY] = fl.my + spd[0] * (ulst.my - ul.my);
    flxi[IVZ] = fl.mz + spd[0] * (ulst.mz - ul.mz);
    flxi[IBY] = fl.by + spd[0] * (ulst.by - ul.by);
    flxi[IBZ] = fl.bz + spd[0] * (ulst.bz - ul.bz);
  } else if (spd[3] <= 0.0) {
    flxi[IDN] = fr.d + spd[4] * (urst.d - ur.d);
    flxi[IVX] = fr.mx + spd[4] * (urst.mx - ur.mx);
    flxi[IVY] = fr.my + spd[4] * (urst.my - ur.my);
    flxi[IVZ] = fr.mz + spd[4] * (urst.mz - ur.mz);
    flxi[IBY] = fr.by + spd[4] * (urst.by - ur.by);
    flxi[IBZ] = fr.bz + spd[4] * (urst.bz - ur.bz);
  } else {
    flxi[IDN] = dhll * ustar;
    flxi[IVX] = fmxhll;
    flxi[IVY] = ucst.my * ustar - bxi * ucst.by;
    flxi[IVZ] = ucst.mz * ustar - bxi * ucst.bz;
    flxi[IBY] = ucst.by * ustar - bxi * ucst.my / ucst.d;
    flxi[IBZ] = ucst.bz * ustar - bxi * ucst.mz / ucst.d;
  }
  flxi[IEN] = 0.0;
}

/* Roe solver for adiabatic MHD, hydro/rsolvers/mhd/roe_mhd.cpp:43-214 with RoeFlux
 * (:245-498); falls back to the LLF flux for the hydro variables when an intermediate
 * density is negative (:190-201). */
static void roe(double gamma, const double *wli, const double *wri, double bxi, double *flxi) {
  double gm1 = gamma - 1.0;
  double wroe[NWAVE], fl[NWAVE], fr[NWAVE], du[NWAVE], ev[NWAVE];
  double sqrtdl = sqrt(wli[IDN]);
  double sqrtdr = sqrt(wri[IDN]);
  double isdlpdr = 1.0 / (sqrtdl + sqrtdr);
  wroe[IDN] = sqrtdl * sqrtdr;
  wroe[IVX] = (sqrtdl * wli[IVX] + sqrtdr * wri[IVX]) * isdlpdr;
  wroe[IVY] = (sqrtdl * wli[IVY] + sqrtdr * wri[IVY]) * isdlpdr;
  wroe[IVZ] = (sqrtdl * wli[IVZ] + sqrtdr * wri[IVZ]) * isdlpdr;
  wroe[IBY] = (sqrtdr * wli[IBY] + sqrtdl * wri[IBY]) * isdlpdr;
  wroe[IBZ] = (sqrtdr * wli[IBZ] + sqrtdl * wri[IBZ]) * isdlpdr;
  double x = 0.5 * (SQR(wli[IBY] - wri[IBY]) + SQR(wli[IBZ] - wri[IBZ])) / (SQR(sqrtdl + sqrtdr));
  double y = 0.5 * (wli[IDN] + wri[IDN]) / wroe[IDN];
  double pbl = 0.5 * (bxi * bxi + SQR(wli[IBY]) + SQR(wli[IBZ]));
  double pbr = 0.5 * (bxi * bxi + SQR(wri[IBY]) + SQR(wri[IBZ]));
  double el = wli[IPR] / gm1 + 0.5 * wli[IDN] * (SQR(wli[IVX]) + SQR(wli[IVY]) + SQR(wli[IVZ])) + pbl;
  double er = wri[IPR] / gm1 + 0.5 * wri[IDN] * (SQR(wri[IVX]) + SQR(wri[IVY]) + SQR(wri[IVZ])) + pbr;
  wroe[IPR] = ((el + wli[IPR] + pbl) / sqrtdl + (er + wri[IPR] + pbr) / sqrtdr) * isdlpdr;
  double mxl = wli[IDN] * wli[IVX];
  double mxr = wri[IDN] * wri[IVX];
  fl[IDN] = mxl;
  fr[IDN] = mxr;
  fl[IVX] = mxl * wli[IVX] + pbl - SQR(bxi);
  fr[IVX] = mxr * wri[IVX] + pbr - SQR(bxi);
  fl[IVY] = mxl * wli[IVY] - bxi * wli[IBY];
  fr[IVY] = mxr * wri[IVY] - bxi * wri[IBY];
  fl[IVZ] = mxl * wli[IVZ] - bxi * wli[IBZ];
  fr[IVZ] = mxr * wri[IVZ] - bxi * wri[IBZ];
  fl[IVX] += wli[IPR];
  fr[IVX] += wri[IPR];
  fl[IEN] = (el + wli[IPR] + pbl - bxi * bxi) * wli[IVX];
  fr[IEN] = (er + wri[IPR] + pbr - bxi * bxi) * wri[IVX];
  fl[IEN] -= bxi * (wli[IBY] * wli[IVY] + wli[IBZ] * wli[IVZ]);
  fr[IEN] -= bxi * (wri[IBY] * wri[IVY] + wri[IBZ] * wri[IVZ]);
  fl[IBY] = wli[IBY] * wli[IVX] - bxi * wli[IVY];
  fr[IBY] = wri[IBY] * wri[IVX] - bxi * wri[IVY];
  fl[IBZ] = wli[IBZ] * wli[IVX] - bxi * wli[IVZ];
  fr[IBZ] = wri[IBZ] * wri[IVX] - bxi * wri[IVZ];
  du[IDN] = wri[IDN] - wli[IDN];
  du[IVX] = wri[IDN] * wri[IVX] - wli[IDN] * wli[IVX];
  du[IVY] = wri[IDN] * wri[IVY] - wli[IDN] * wli[IVY];
  du[IVZ] = wri[IDN] * wri[IVZ] - wli[IDN] * wli[IVZ];
  du[IEN] = er - el;
  du[IBY] = wri[IBY] - wli[IBY];
  du[IBZ] = wri[IBZ] - wli[IBZ];
  for (int n = 0; n < NWAVE; ++n) flxi[n] = 0.5 * (fl[n] + fr[n]);
  int llf_flag = 0;
  {
    /* RoeFlux, adiabatic branch (roe_mhd.cpp:245-498) */
    const double b1 = bxi;
    double d = wroe[IDN], v1 = wroe[IVX], v2 = wroe[IVY], v3 = wroe[IVZ];
    double b2 = wroe[IBY], b3 = wroe[IBZ];
    double di = 1.0 / d;
    double btsq = b2 * b2 + b3 * b3;
    double vaxsq = b1 * b1 * di;
    double vsq = v1 * v1 + v2 * v2 + v3 * v3;
    double hp = wroe[IPR] - (vaxsq + btsq * di);
    double bt_starsq = (gm1 - (gm1 - 1.0) * y) * btsq;
    double twid_csq = dmax((gm1 * (hp - 0.5 * vsq) - (gm1 - 1.0) * x), 1.0e-20);
    double ct2 = bt_starsq * di;
    double tsum = vaxsq + ct2 + twid_csq;
    double tdif = vaxsq + ct2 - twid_csq;
    double cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_csq * ct2);
    double cfsq = 0.5 * (tsum + cf2_cs2);
    double cf = sqrt(cfsq);
    double cssq = twid_csq * vaxsq / cfsq;
    double cs = sqrt(cssq);
    double bt = sqrt(btsq);
    double bt_star = sqrt(bt_starsq);
    double bet2 = 0.0, bet3 = 0.0;
    if (bt != 0.0) { bet2 = b2 / bt; bet3 = b3 / bt; }
    double bet2_star = bet2 / sqrt(gm1 - (gm1 - 1.0) * y);
    double bet3_star = bet3 / sqrt(gm1 - (gm1 - 1.0) * y);
    double bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
    double vbet = v2 * bet2_star + v3 * bet3_star;
    double q2_star = 0.0, q3_star = 0.0;
    if (bet_starsq != 0.0) { q2_star = bet2_star / bet_starsq; q3_star = bet3_star / bet_starsq; }
    double alpha_f, alpha_s;
    if ((cfsq - cssq) <= 0.0) { alpha_f = 1.0; alpha_s = 0.0; }
    else if ((twid_csq - cssq) <= 0.0) { alpha_f = 0.0; alpha_s = 1.0; }
    else if ((cfsq - twid_csq) <= 0.0) { alpha_f = 1.0; alpha_s = 0.0; }
    else {
      alpha_f = sqrt((twid_csq - cssq) / (cfsq - cssq));
      alpha_s = sqrt((cfsq - twid_csq) / (cfsq - cssq));
    }
    double sqrtd = sqrt(d);
    double isqrtd = 1.0 / sqrtd;
    double s = (b1 < 0.0) ? -1.0 : 1.0;
    double twid_c = sqrt(twid_csq);
    double qf = cf * alpha_f * s;
    double qs = cs * alpha_s * s;
    double af_prime = twid_c * alpha_f * isqrtd;
    double as_prime = twid_c * alpha_s * isqrtd;
    double afpbb = af_prime * bt_star * bet_starsq;
    double aspbb = as_prime * bt_star * bet_starsq;
    double vqstr = (v2 * q2_star + v3 * q3_star);
    double vax = sqrt(vaxsq);
    double norm = 0.5 / twid_csq;
    double cff = norm * alpha_f * cf;
    double css = norm * alpha_s * cs;
    double qf_hat = qf * norm;
    double qs_hat = qs * norm;
    double af = norm * af_prime * d;
    double as = norm * as_prime * d;
    double afpb = norm * af_prime * bt_star;
    double aspb = norm * as_prime * bt_star;
    ev[0] = v1 - cf; ev[1] = v1 - vax; ev[2] = v1 - cs; ev[3] = v1;
    ev[4] = v1 + cs; ev[5] = v1 + vax; ev[6] = v1 + cf;
    double a[NWAVE];
    double alpha_f_bar = alpha_f * gm1 * norm;
    double alpha_s_bar = alpha_s * gm1 * norm;
    double gm1a = gm1 / twid_csq;
    a[0] = du[0] * (alpha_f_bar * (vsq - hp) + cff * (cf + v1) - qs_hat * vqstr - aspb);
    a[0] -= du[1] * (alpha_f_bar * v1 + cff);
    a[0] -= du[2] * (alpha_f_bar * v2 - qs_hat * q2_star);
    a[0] -= du[3] * (alpha_f_bar * v3 - qs_hat * q3_star);
    a[0] += du[4] * alpha_f_bar;
    a[0] += du[5] * (as * q2_star - alpha_f_bar * b2);
    a[0] += du[6] * (as * q3_star - alpha_f_bar * b3);
    a[1] = du[0] * (v2 * bet3 - v3 * bet2);
    a[1] -= du[2] * bet3;
    a[1] += du[3] * bet2;
    a[1] -= du[5] * sqrtd * bet3 * s;
    a[1] += du[6] * sqrtd * bet2 * s;
    a[1] *= 0.5;
    a[2] = du[0] * (alpha_s_bar * (vsq - hp) + css * (cs + v1) + qf_hat * vqstr + afpb);
    a[2] -= du[1] * (alpha_s_bar * v1 + css);
    a[2] -= du[2] * (alpha_s_bar * v2 + qf_hat * q2_star);
    a[2] -= du[3] * (alpha_s_bar * v3 + qf_hat * q3_star);
    a[2] += du[4] * alpha_s_bar;
    a[2] -= du[5] * (af * q2_star + alpha_s_bar * b2);
    a[2] -= du[6] * (af * q3_star + alpha_s_bar * b3);
    a[3] = du[0] * (1.0 - gm1a * (0.5 * vsq - (gm1 - 1.0) * x / gm1));
    a[3] += du[1] * gm1a * v1;
    a[3] += du[2] * gm1a * v2;
    a[3] += du[3] * gm1a * v3;
    a[3] -= du[4] * gm1a;
    a[3] += du[5] * gm1a * b2;
    a[3] += du[6] * gm1a * b3;
    a[4] = du[0] * (alpha_s_bar * (vsq - hp) + css * (cs - v1) - qf_hat * vqstr + afpb);
    a[4] -= du[1] * (alpha_s_bar * v1 - css);
    a[4] -= du[2] * (alpha_s_bar * v2 - qf_hat * q2_star);
    a[4] -= du[3] * (alpha_s_bar * v3 - qf_hat * q3_star);
    a[4] += du[4] * alpha_s_bar;
    a[4] -= du[5] * (af * q2_star + alpha_s_bar * b2);
    a[4] -= du[6] * (af * q3_star + alpha_s_bar * b3);
    a[5] = du[0] * (v3 * bet2 - v2 * bet3);
    a[5] += du[2] * bet3;
    a[5] -= du[3] * bet2;
    a[5] -= du[5] * sqrtd * bet3 * s;
    a[5] += du[6] * sqrtd * bet2 * s;
    a[5] *= 0.5;
    a[6] = du[0] * (alpha_f_bar * (vsq - hp) + cff * (cf - v1) + qs_hat * vqstr - aspb);
    a[6] -= du[1] * (alpha_f_bar * v1 - cff);
    a[6] -= du[2] * (alpha_f_bar * v2 + qs_hat * q2_star);
    a[6] -= du[3] * (alpha_f_bar * v3 + qs_hat * q3_star);
    a[6] += du[4] * alpha_f_bar;
    a[6] += du[5] * (as * q2_star - alpha_f_bar * b2);
    a[6] += du[6] * (as * q3_star - alpha_f_bar * b3);
    double coeff[NWAVE];
    for (int n = 0; n < NWAVE; ++n) coeff[n] = -0.5 * fabs(ev[n]) * a[n];
    double dens = wli[IDN] + a[0] * alpha_f;
    if (dens < 0.0) llf_flag = 1;
    dens += a[2] * alpha_s;
    if (dens < 0.0) llf_flag = 1;
    dens += a[3];
    if (dens < 0.0) llf_flag = 1;
    dens += a[4] * alpha_s;
    if (dens < 0.0) llf_flag = 1;
    flxi[0] += coeff[0] * alpha_f;
    flxi[0] += coeff[2] * alpha_s;
    flxi[0] += coeff[3];
    flxi[0] += coeff[4] * alpha_s;
    flxi[0] += coeff[6] * alpha_f;
    flxi[1] += coeff[0] * (alpha_f * (v1 - cf));
    flxi[1] += coeff[2] * (alpha_s * (v1 - cs));
    flxi[1] += coeff[3] * v1;
    flxi[1] += coeff[4] * (alpha_s * (v1 + cs));
    flxi[1] += coeff[6] * (alpha_f * (v1 + cf));
    flxi[2] += coeff[0] * (alpha_f * v2 + qs * bet2_star);
    flxi[2] -= coeff[1] * bet3;
    flxi[2] += coeff[2] * (alpha_s * v2 - qf * bet2_star);
    flxi[2] += coeff[3] * v2;
    flxi[2] += coeff[4] * (alpha_s * v2 + qf * bet2_star);
    flxi[2] += coeff[5] * bet3;
    flxi[2] += coeff[6] * (alpha_f * v2 - qs * bet2_star);
    flxi[3] += coeff[0] * (alpha_f * v3 + qs * bet3_star);
    flxi[3] += coeff[1] * bet2;
    flxi[3] += coeff[2] * (alpha_s * v3 - qf * bet3_star);
    flxi[3] += coeff[3] * v3;
    flxi[3] += coeff[4] * (alpha_s * v3 + qf * bet3_star);
    flxi[3] -= coeff[5] * bet2;
    flxi[3] += coeff[6] * (alpha_f * v3 - qs * bet3_star);
    flxi[4] += coeff[0] * (alpha_f * (hp - v1 * cf) + qs * vbet + aspbb);
    flxi[4] -= coeff[1] * (v2 * bet3 - v3 * bet2);
    flxi[4] += coeff[2] * (alpha_s * (hp - v1 * cs) - qf * vbet - afpbb);
    flxi[4] += coeff[3] * (0.5 * vsq + (gm1 - 1.0) * x / gm1);
    flxi[4] += coeff[4] * (alpha_s * (hp + v1 * cs) + qf * vbet - afpbb);
    flxi[4] += coeff[5] * (v1 * bet3 - v3 * bet2);
    flxi[4] += coeff[6] * (alpha_f * (hp + v1 * cf) - qs * vbet + aspbb);
    flxi[5] += coeff[0] * as_prime * bet2_star;
    flxi[5] -= coeff[1] * bet3 * s * isqrtd;
    flxi[5] -= coeff[2] * af_prime * bet2_star;
    flxi[5] -= coeff[4] * af_prime * bet2_star;
    flxi[5] -= coeff[5] * bet3 * s * isqrtd;
    flxi[5] += coeff[6] * as_prime * bet2_star;
    flxi[6] += coeff[0] * as_prime * bet3_star;
    flxi[6] += coeff[1] * bet2 * s * isqrtd;
    flxi[6] -= coeff[2] * af_prime * bet3_star;
    flxi[6] -= coeff[4] * af_prime * bet3_star;
    flxi[6] += coeff[5] * bet2 * s * isqrtd;
    flxi[6] += coeff[6] * as_prime * bet3_star;
  }
  if (ev[0] >= 0.0) {
    for (int n = 0; n < NWAVE; ++n) flxi[n] = fl[n];
  }
  if (ev[NWAVE - 1] <= 0.0) {
    for (int n = 0; n < NWAVE; ++n) flxi[n] = fr[n];
  }
  if (llf_flag != 0) {
    double cfl = fast_speed(gamma, wli, bxi);
    double cfr = fast_speed(gamma, wri, bxi);
    double a = 0.5 * dmax((fabs(wli[IVX]) + cfl), (fabs(wri[IVX]) + cfr));
    flxi[IDN] = 0.5 * (fl[IDN] + fr[IDN]) - a * du[IDN];
    flxi[IVX] = 0.5 * (fl[IVX] + fr[IVX]) - a * du[IVX];
    flxi[IVY] = 0.5 * (fl[IVY] + fr[IVY]) - a * du[IVY];
    flxi[IVZ] = 0.5 * (fl[IVZ] + fr[IVZ]) - a * du[IVZ];
    flxi[IEN] = 0.5 * (fl[IEN] + fr[IEN]) - a * du[IEN];
  }
}

/* ---------------------------------------------------------------- fluxes */
/* Hydro::CalculateFluxes, hydro/calculate_fluxes.cpp:36-378, orders 1 (dc.cpp:24-104)
 * and 2 (plm.cpp:27-351, uniform Cartesian van Leer branch).  Interface states are
 * evaluated per face directly; the pencil scratch arrays of the reference hold the
 * same values. */
/* SIGN, defs.hpp.in:117 */
static inline double sgn(double x) { return (x < 0.0) ? -1.0 : 1.0; }

/* PiecewiseParabolicX1..3 for one variable of one cell, uniform Cartesian mesh
 * (reconstruct/ppm.cpp:58-338; coefficients c1..c4 = 1/2, c5 = 1/6, c6 = -1/6 from
 * reconstruct/reconstruction.cpp:429-438, hplus/hminus ratios = 2).  q[0..4] are the
 * values at i-2..i+2; returns qminus (state at i-1/2) and qplus (state at i+1/2). */
static void ppm_cell(const double *q, double *qminus_out, double *qplus_out) {
  const double C2 = 1.25;
  const double c1 = 0.5, c2 = 0.5, c3 = 0.5, c4 = 0.5, c5 = 1.0 / 6.0, c6 = -1.0 / 6.0;
  double q_im2 = q[0], q_im1 = q[1], qc = q[2], q_ip1 = q[3], q_ip2 = q[4];
  double qa = (qc - q_im1);
  double qb = (q_ip1 - qc);
  double dd_im1 = c1 * qa + c2 * (q_im1 - q_im2);
  double dd = c1 * qb + c2 * qa;
  double dd_ip1 = c1 * (q_ip2 - q_ip1) + c2 * qb;
  double dph = (c3 * q_im1 + c4 * qc) + (c5 * dd_im1 + c6 * dd);
  double dph_ip1 = (c3 * qc + c4 * q_ip1) + (c5 * dd + c6 * dd_ip1);
  double d2qc_im1 = q_im2 + qc - 2.0 * q_im1;
  double d2qc = q_im1 + q_ip1 - 2.0 * qc;
  double d2qc_ip1 = qc + q_ip2 - 2.0 * q_ip1;
  {
    double qa_tmp = dph - q_im1;
    double qb_tmp = qc - dph;
    double a = 3.0 * (q_im1 + qc - 2.0 * dph);
    double b = d2qc_im1, cc = d2qc;
    double qd = 0.0;
    if (sgn(a) == sgn(b) && sgn(a) == sgn(cc))
      qd = sgn(a) * dmin(C2 * fabs(b), dmin(C2 * fabs(cc), fabs(a)));
    double dph_tmp = 0.5 * (q_im1 + qc) - qd / 6.0;
    if (qa_tmp * qb_tmp < 0.0) dph = dph_tmp;
  }
  {
    double qa_tmp = dph_ip1 - qc;
    double qb_tmp = q_ip1 - dph_ip1;
    double a = 3.0 * (qc + q_ip1 - 2.0 * dph_ip1);
    double b = d2qc, cc = d2qc_ip1;
    double qd = 0.0;
    if (sgn(a) == sgn(b) && sgn(a) == sgn(cc))
      qd = sgn(a) * dmin(C2 * fabs(b), dmin(C2 * fabs(cc), fabs(a)));
    double dphip1_tmp = 0.5 * (qc + q_ip1) - qd / 6.0;
    if (qa_tmp * qb_tmp < 0.0) dph_ip1 = dphip1_tmp;
  }
  double d2qf = 6.0 * (dph + dph_ip1 - 2.0 * qc);
  double qminus = dph, qplus = dph_ip1;
  double dqf_minus = qc - qminus;
  double dqf_plus = qplus - qc;
  {
    double qa_tmp = dqf_minus * dqf_plus;
    double qb_tmp = (q_ip1 - qc) * (qc - q_im1);
    double a = d2qc_im1, b = d2qc, cc = d2qc_ip1, qd = d2qf;
    double qe = 0.0;
    if (sgn(a) == sgn(b) && sgn(a) == sgn(cc) && sgn(a) == sgn(qd))
      qe = sgn(qd) * dmin(dmin(C2 * fabs(a), C2 * fabs(b)), dmin(C2 * fabs(cc), fabs(qd)));
    a = dmax(fabs(q_im1), fabs(q_im2));
    b = dmax(dmax(fabs(qc), fabs(q_ip1)), fabs(q_ip2));
    double rho = 0.0;
    if (fabs(qd) > (1.0e-12) * dmax(a, b)) rho = qe / qd;
    double tmp_m = qc - rho * dqf_minus;
    double tmp_p = qc + rho * dqf_plus;
    double tmp2_m = qc - 2.0 * dqf_plus;
    double tmp2_p = qc + 2.0 * dqf_minus;
    if ((qa_tmp <= 0.0 || qb_tmp <= 0.0)) {
      if (rho <= (1.0 - (1.0e-12))) {
        qminus = tmp_m;
        qplus = tmp_p;
      }
    } else {
      if (fabs(dqf_minus) >= 2.0 * fabs(dqf_plus)) qminus = tmp2_m;
      if (fabs(dqf_plus) >= 2.0 * fabs(dqf_minus)) qplus = tmp2_p;
    }
  }
  *qminus_out = qminus;
  *qplus_out = qplus;
}

/* The eigensystem of the primitive-variable MHD Jacobian at one cell state, shared by
 * Reconstruction::LeftEigenmatrixDotVector and RightEigenmatrixDotVector (characteristic.cpp:
 * 36-262, 280-500; Stone et al. 2008, ApJS 178, 137, eqs. A10, A16, A17).  w is the cell's
 * primitive state in the 7-slot order with (IBY, IBZ) its transverse field, b1 its
 * cell-centred normal field; asq = gamma P / rho (adiabatic) or iso_cs^2 (isothermal). */
typedef struct {
  double d, id, sqrtd, isqrtd, cf, cs, a, asq, bet2, bet3, alpha_f, alpha_s, s;
} orc_eigen;

static void char_eigen(const orc_ctx *c, double b1, const double *w, orc_eigen *e) {
  double iso_cs2 = SQR(c->d.iso_cs);
  e->d = w[IDN];
  e->id = 1.0 / w[IDN];
  e->sqrtd = sqrt(w[IDN]);
  e->isqrtd = 1.0 / e->sqrtd;
  double btsq = SQR(w[IBY]) + SQR(w[IBZ]);
  double bxsq = b1 * b1;
  double gamp = c->iso ? iso_cs2 * w[IDN] : c->d.gamma * w[IPR];
  double tdif = bxsq + btsq - gamp;                        /* eq. A10 */
  double cf2_cs2 = sqrt(tdif * tdif + 4.0 * gamp * btsq);
  double cfsq = 0.5 * (bxsq + btsq + gamp + cf2_cs2);
  double cssq = gamp * bxsq / cfsq;
  cfsq *= e->id;
  e->cf = sqrt(cfsq);
  cssq *= e->id;
  e->cs = sqrt(cssq);
  if (c->iso) { e->asq = iso_cs2; e->a = c->d.iso_cs; }
  else { e->asq = gamp * e->id; e->a = sqrt(e->asq); }
  double bt = sqrt(btsq);                                  /* eq. A17 */
  e->bet2 = 0.0;
  e->bet3 = 0.0;
  if (bt != 0.0) { e->bet2 = w[IBY] / bt; e->bet3 = w[IBZ] / bt; }
  double asq = e->asq;                                     /* eq. A16 */
  if ((cfsq - cssq) <= 0.0) { e->alpha_f = 1.0; e->alpha_s = 0.0; }
  else if ((asq - cssq) <= 0.0) { e->alpha_f = 0.0; e->alpha_s = 1.0; }
  else if ((cfsq - asq) <= 0.0) { e->alpha_f = 1.0; e->alpha_s = 0.0; }
  else {
    e->alpha_f = sqrt((asq - cssq) / (cfsq - cssq));
    e->alpha_s = sqrt((cfsq - asq) / (cfsq - cssq));
  }
  e->s = sgn(b1);
}

/* Number of characteristic fields: 7 adiabatic, 6 isothermal (no entropy wave). */
static inline int char_nwave(const orc_ctx *c) { return c->iso ? 6 : 7; }

/* LeftEigenmatrixDotVector: the primitive vector v (7 slots, mesh-frame velocities, ivx the
 * slot of the sweep's velocity) to its characteristic fields ch: eq. A18 (adiabatic,
 * characteristic.cpp:85-124: fast-, Alfven-, slow-, entropy, slow+, Alfven+, fast+) or
 * eq. A22 (isothermal, :166-197: the same without the entropy wave). */
static void char_left(const orc_ctx *c, const orc_eigen *e, int ivx, const double *v, double *ch) {
  int ivy = IVX + ((ivx - IVX) + 1) % 3, ivz = IVX + ((ivx - IVX) + 2) % 3;
  double vt = e->bet2 * v[ivy] + e->bet3 * v[ivz];
  double bt = e->bet2 * v[IBY] + e->bet3 * v[IBZ];
  double alf_m = 0.5 * (e->bet2 * (v[IBZ] * e->s * e->isqrtd + v[ivz]) -
                        e->bet3 * (v[IBY] * e->s * e->isqrtd + v[ivy]));
  double alf_p = 0.5 * (e->bet2 * (v[IBZ] * e->s * e->isqrtd - v[ivz]) -
                        e->bet3 * (v[IBY] * e->s * e->isqrtd - v[ivy]));
  double af_prime = 0.5 * e->alpha_f / (e->a * e->sqrtd);
  double as_prime = 0.5 * e->alpha_s / (e->a * e->sqrtd);
  if (!c->iso) {
    double nf = 0.5 / e->asq;
    double qf = nf * e->cf * e->alpha_f * e->s;
    double qs = nf * e->cs * e->alpha_s * e->s;
    double pr = v[IPR] * e->id;
    ch[0] = nf * e->alpha_f * (pr - e->cf * v[ivx]) + qs * vt + as_prime * bt;
    ch[1] = alf_m;
    ch[2] = nf * e->alpha_s * (pr - e->cs * v[ivx]) - qf * vt - af_prime * bt;
    ch[3] = v[IDN] - v[IPR] / e->asq;
    ch[4] = nf * e->alpha_s * (pr + e->cs * v[ivx]) + qf * vt - af_prime * bt;
    ch[5] = alf_p;
    ch[6] = nf * e->alpha_f * (pr + e->cf * v[ivx]) - qs * vt + as_prime * bt;
  } else {
    double c2 = e->asq;
    double qf = 0.5 * (e->cf * e->alpha_f * e->s) / c2;
    double qs = 0.5 * (e->cs * e->alpha_s * e->s) / c2;
    double dn = v[IDN] * e->id;
    ch[0] = 0.5 * e->alpha_f * (dn - e->cf * v[ivx] / c2) + qs * vt + as_prime * bt;
    ch[1] = alf_m;
    ch[2] = 0.5 * e->alpha_s * (dn - e->cs * v[ivx] / c2) - qf * vt - af_prime * bt;
    ch[3] = 0.5 * e->alpha_s * (dn + e->cs * v[ivx] / c2) + qf * vt - af_prime * bt;
    ch[4] = alf_p;
    ch[5] = 0.5 * e->alpha_f * (dn + e->cf * v[ivx] / c2) - qs * vt + as_prime * bt;
  }
}

/* RightEigenmatrixDotVector: characteristic fields ch back to the primitive vector v
 * (eq. A12; characteristic.cpp:330-352 adiabatic, :400-419 isothermal, whose state has no
 * pressure: that slot is set to 0 here). */
static void char_right(const orc_ctx *c, const orc_eigen *e, int ivx, const double *ch, double *v) {
  int ivy = IVX + ((ivx - IVX) + 1) % 3, ivz = IVX + ((ivx - IVX) + 2) % 3;
  /* slots of the wave pairs: fast (fm, fp), Alfven (am, ap), slow (sm, sp) */
  const int fm = 0, am = 1, sm = 2;
  const int sp = c->iso ? 3 : 4, ap = c->iso ? 4 : 5, fp = c->iso ? 5 : 6;
  double qf = e->cf * e->alpha_f * e->s;
  double qs = e->cs * e->alpha_s * e->s;
  double af = e->a * e->alpha_f * e->sqrtd;
  double as = e->a * e->alpha_s * e->sqrtd;
  double compr = e->alpha_f * (ch[fm] + ch[fp]) + e->alpha_s * (ch[sm] + ch[sp]);
  double tr = qs * (ch[fm] - ch[fp]) + qf * (ch[sp] - ch[sm]);
  double bc = as * (ch[fm] + ch[fp]) - af * (ch[sm] + ch[sp]);
  v[IDN] = e->d * compr;
  if (!c->iso) v[IDN] += ch[3];
  v[ivx] = e->cf * e->alpha_f * (ch[fp] - ch[fm]) + e->cs * e->alpha_s * (ch[sp] - ch[sm]);
  v[ivy] = e->bet2 * tr + e->bet3 * (ch[ap] - ch[am]);
  v[ivz] = e->bet3 * tr + e->bet2 * (ch[am] - ch[ap]);
  v[IPR] = c->iso ? 0.0 : e->d * e->asq * compr;
  v[IBY] = e->bet2 * bc - e->bet3 * e->s * e->sqrtd * (ch[ap] + ch[am]);
  v[IBZ] = e->bet3 * bc + e->bet2 * e->s * e->sqrtd * (ch[ap] + ch[am]);
}

static inline void recon_cell(const orc_ctx *c, const double *w, const double *bcc,
                              int dir, int order, int g, int k, int j, int i,
                              double *wm, double *wp) {
  /* returns reconstructed state at the lower (wm: "wr" of face at this cell's lower
   * side) and upper (wp: "wl" of face at upper side) faces of cell (k,j,i) */
  int64_t nc = c->ncell;
  int64_t id = CIDX(c, k, j, i);
  int64_t st = (dir == 0) ? 1 : (dir == 1 ? c->n1 : (int64_t)c->n1 * c->n2);
  /* variable order in the 7-slot wave array: IDN, IVX, IVY, IVZ, IPR, IBY, IBZ with
   * (IBY,IBZ) = the two transverse bcc components (plm.cpp:46-56,160-170,273-283) */
  int tb1 = (dir == 0) ? IB2 : (dir == 1 ? IB3 : IB1);
  int tb2 = (dir == 0) ? IB3 : (dir == 1 ? IB1 : IB2);
  const int ch = c->d.char_proj && order > 1;
  orc_eigen eg;
  if (ch) {
    double wc[NWAVE];
    for (int n = 0; n < NHYDRO; ++n) wc[n] = w[n * nc + id];
    wc[IBY] = bcc[tb1 * nc + id];
    wc[IBZ] = bcc[tb2 * nc + id];
    char_eigen(c, bcc[dir * nc + id], wc, &eg);
  }
  if (order == 3) {
    if (ch) {
      /* ppm.cpp:99-107,313-316: the five cells' primitive vectors projected with the
       * centre cell's left eigenvectors, each characteristic field limited on its own,
       * the two interface values projected back with its right eigenvectors */
      double qc[5][NWAVE], v[NWAVE], cm[NWAVE], cp[NWAVE];
      for (int t = 0; t < 5; ++t) {
        int64_t o = id + (t - 2) * st;
        for (int n = 0; n < NHYDRO; ++n) v[n] = w[n * nc + o];
        v[IBY] = bcc[tb1 * nc + o];
        v[IBZ] = bcc[tb2 * nc + o];
        char_left(c, &eg, IVX + dir, v, qc[t]);
      }
      for (int n = 0; n < char_nwave(c); ++n) {
        double q5[5];
        for (int t = 0; t < 5; ++t) q5[t] = qc[t][n];
        ppm_cell(q5, &cm[n], &cp[n]);
      }
      char_right(c, &eg, IVX + dir, cm, wm);
      char_right(c, &eg, IVX + dir, cp, wp);
    } else {
      for (int n = 0; n < NWAVE; ++n) {
        const double *a = (n < NHYDRO) ? w + n * nc : bcc + (n == IBY ? tb1 : tb2) * nc;
        double q5[5];
        for (int t = 0; t < 5; ++t) q5[t] = a[id + (t - 2) * st];
        ppm_cell(q5, &wm[n], &wp[n]);
      }
    }
    /* ApplyPrimitiveFloors on both reconstructed states (ppm.cpp:333-337,
     * eos/adiabatic_mhd.cpp:162-172) */
    double dfl = c->d.dfloor, pfl = c->d.pfloor;
    wm[IDN] = (wm[IDN] > dfl) ? wm[IDN] : dfl;
    wp[IDN] = (wp[IDN] > dfl) ? wp[IDN] : dfl;
    if (!c->iso) {   /* eos/isothermal_mhd.cpp:135-140 floors the density only */
      wm[IPR] = (wm[IPR] > pfl) ? wm[IPR] : pfl;
      wp[IPR] = (wp[IPR] > pfl) ? wp[IPR] : pfl;
    }
    return;
  }
  double q[NWAVE][3];
  for (int n = 0; n < NHYDRO; ++n) {
    q[n][0] = w[n * nc + id - st]; q[n][1] = w[n * nc + id]; q[n][2] = w[n * nc + id + st];
  }
  q[IBY][0] = bcc[tb1 * nc + id - st]; q[IBY][1] = bcc[tb1 * nc + id]; q[IBY][2] = bcc[tb1 * nc + id + st];
  q[IBZ][0] = bcc[tb2 * nc + id - st]; q[IBZ][1] = bcc[tb2 * nc + id]; q[IBZ][2] = bcc[tb2 * nc + id + st];
  if (order == 1) {
    for (int n = 0; n < NWAVE; ++n) { wm[n] = q[n][1]; wp[n] = q[n][1]; }
    return;
  }
  double dxp, dxm;
  if (dir == 0) {
    const double *xf = c->x1f + g * (c->n1 + 1), *xv = c->x1v + g * (c->n1 + 1),
                 *dx = c->dx1f + g * (c->n1 + 1);
    dxp = (xf[i + 1] - xv[i]) / dx[i];
    dxm = (xv[i] - xf[i]) / dx[i];
  } else if (dir == 1) {
    const double *xf = c->x2f + g * (c->n2 + 1), *xv = c->x2v + g * (c->n2 + 1),
                 *dx = c->dx2f + g * (c->n2 + 1);
    dxp = (xf[j + 1] - xv[j]) / dx[j];
    dxm = (xv[j] - xf[j]) / dx[j];
  } else {
    const double *xf = c->x3f + g * (c->n3 + 1), *xv = c->x3v + g * (c->n3 + 1),
                 *dx = c->dx3f + g * (c->n3 + 1);
    dxp = (xf[k + 1] - xv[k]) / dx[k];
    dxm = (xv[k] - xf[k]) / dx[k];
  }
  if (ch) {
    /* plm.cpp:59-63,96-99,109-136: the two one-sided differences projected with the
     * cell's left eigenvectors, the van Leer slope of each characteristic field, the
     * slope projected back, then the EOS floors on both states */
    double dl[NWAVE], dr[NWAVE], cl[NWAVE], cr[NWAVE], cm[NWAVE], dm[NWAVE];
    for (int n = 0; n < NWAVE; ++n) { dl[n] = q[n][1] - q[n][0]; dr[n] = q[n][2] - q[n][1]; }
    char_left(c, &eg, IVX + dir, dl, cl);
    char_left(c, &eg, IVX + dir, dr, cr);
    for (int n = 0; n < char_nwave(c); ++n) {
      double dw2 = cl[n] * cr[n];
      cm[n] = 2.0 * dw2 / (cl[n] + cr[n]);
      if (dw2 <= 0.0) cm[n] = 0.0;
    }
    char_right(c, &eg, IVX + dir, cm, dm);
    for (int n = 0; n < NWAVE; ++n) {
      wp[n] = q[n][1] + dxp * dm[n];
      wm[n] = q[n][1] - dxm * dm[n];
    }
    double dfl = c->d.dfloor, pfl = c->d.pfloor;
    wm[IDN] = (wm[IDN] > dfl) ? wm[IDN] : dfl;
    wp[IDN] = (wp[IDN] > dfl) ? wp[IDN] : dfl;
    if (!c->iso) {
      wm[IPR] = (wm[IPR] > pfl) ? wm[IPR] : pfl;
      wp[IPR] = (wp[IPR] > pfl) ? wp[IPR] : pfl;
    }
    return;
  }
  for (int n = 0; n < NWAVE; ++n) {
    double dwl = (q[n][1] - q[n][0]);
    double dwr = (q[n][2] - q[n][1]);
    double dw2 = dwl * dwr;
    double dwm = 2.0 * dw2 / (dwl + dwr);
    if (dw2 <= 0.0) dwm = 0.0;
    wp[n] = q[n][1] + dxp * dwm;
    wm[n] = q[n][1] - dxm * dwm;
  }
}

void orc_calculate_fluxes(orc_ctx *c, int order) {
  double gamma = c->d.gamma, dt = c->dt;
  int rs = c->d.rsolver;
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < c->nblocks; ++g) {
    const double *w = c->arr[ARR_W] + g * c->arrsize[ARR_W];
    const double *bcc = c->arr[ARR_BCC] + g * c->arrsize[ARR_BCC];
    for (int dir = 0; dir < 3; ++dir) {
      const double *bn; double *flx, *ey, *ez, *wct;
      int il, iu, jl, ju, kl, ku;
      int64_t nf;
      if (dir == 0) {
        bn = c->arr[ARR_B1F] + g * c->nf1; nf = c->nf1;
        flx = c->arr[ARR_FX1] + g * c->arrsize[ARR_FX1];
        ey = c->arr[ARR_E3X1] + g * c->nf1; ez = c->arr[ARR_E2X1] + g * c->nf1;
        wct = c->arr[ARR_WX1] + g * c->nf1;
        il = c->is; iu = c->ie + 1; jl = c->js - 1; ju = c->je + 1; kl = c->ks - 1; ku = c->ke + 1;
      } else if (dir == 1) {
        bn = c->arr[ARR_B2F] + g * c->nf2; nf = c->nf2;
        flx = c->arr[ARR_FX2] + g * c->arrsize[ARR_FX2];
        ey = c->arr[ARR_E1X2] + g * c->nf2; ez = c->arr[ARR_E3X2] + g * c->nf2;
        wct = c->arr[ARR_WX2] + g * c->nf2;
        il = c->is - 1; iu = c->ie + 1; jl = c->js; ju = c->je + 1; kl = c->ks - 1; ku = c->ke + 1;
      } else {
        bn = c->arr[ARR_B3F] + g * c->nf3; nf = c->nf3;
        flx = c->arr[ARR_FX3] + g * c->arrsize[ARR_FX3];
        ey = c->arr[ARR_E2X3] + g * c->nf3; ez = c->arr[ARR_E1X3] + g * c->nf3;
        wct = c->arr[ARR_WX3] + g * c->nf3;
        il = c->is - 1; iu = c->ie + 1; jl = c->js - 1; ju = c->je + 1; kl = c->ks; ku = c->ke + 1;
      }
      /* ivx/ivy/ivz cyclic permutation, hlld.cpp:43-44 */
      int ivx = IVX + dir, ivy = IVX + (dir + 1) % 3, ivz = IVX + (dir + 2) % 3;
      for (int k = kl; k <= ku; ++k)
        for (int j = jl; j <= ju; ++j)
          for (int i = il; i <= iu; ++i) {
            double lm[NWAVE], lp[NWAVE], rm[NWAVE], rp[NWAVE];
            /* left state: upper face of the cell below; right: lower face of cell */
            int ki = k - (dir == 2), ji = j - (dir == 1), ii = i - (dir == 0);
            recon_cell(c, w, bcc, dir, order, g, ki, ji, ii, lm, lp);
            recon_cell(c, w, bcc, dir, order, g, k, j, i, rm, rp);
            double wli[NWAVE], wri[NWAVE], flxi[NWAVE];
            wli[IDN] = lp[IDN]; wli[IVX] = lp[ivx]; wli[IVY] = lp[ivy]; wli[IVZ] = lp[ivz];
            wli[IPR] = lp[IPR]; wli[IBY] = lp[IBY]; wli[IBZ] = lp[IBZ];
            wri[IDN] = rm[IDN]; wri[IVX] = rm[ivx]; wri[IVY] = rm[ivy]; wri[IVZ] = rm[ivz];
            wri[IPR] = rm[IPR]; wri[IBY] = rm[IBY]; wri[IBZ] = rm[IBZ];
            int64_t fid = (dir == 0) ? F1IDX(c, k, j, i) : (dir == 1 ? F2IDX(c, k, j, i) : F3IDX(c, k, j, i));
            double bxi = bn[fid];
            if (rs == 1) hlle(gamma, c->iso, c->d.iso_cs, wli, wri, bxi, flxi);
            else if (c->iso) hlld_iso(c->d.iso_cs, c->d.dfloor, wli, wri, bxi, flxi);
            else if (rs == 2) roe(gamma, wli, wri, bxi, flxi);
            else if (rs == 3) {
              double dvn, dvt;
              velocity_differences(c, w, dir, k, j, i, &dvn, &dvt);
              lhlld(gamma, wli, wri, bxi, dvn, dvt, flxi);
            } else hlld(gamma, wli, wri, bxi, flxi);
            flx[IDN * nf + fid] = flxi[IDN];
            flx[ivx * nf + fid] = flxi[IVX];
            flx[ivy * nf + fid] = flxi[IVY];
            flx[ivz * nf + fid] = flxi[IVZ];
            flx[IEN * nf + fid] = flxi[IEN];
            ey[fid] = -flxi[IBY];
            ez[fid] = flxi[IBZ];
            double dxw = (dir == 0) ? c->dx1f[g * (c->n1 + 1) + i]
                                    : (dir == 1 ? c->dx2f[g * (c->n2 + 1) + j] : c->dx3f[g * (c->n3 + 1) + k]);
            wct[fid] = ct_weight(flxi[IDN], wli[IDN], wri[IDN], dxw, dt);
          }
    }
  }
}

/* ---------------------------------------------------------------- corner EMF */
/* Field::ComputeCornerE, 3D Newtonian branch, field/calculate_corner_e.cpp:130-229 */
void orc_corner_e(orc_ctx *c) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < c->nblocks; ++g) {
    int64_t nc = c->ncell;
    const double *w = c->arr[ARR_W] + g * c->arrsize[ARR_W];
    const double *bcc = c->arr[ARR_BCC] + g * c->arrsize[ARR_BCC];
    double *cce = c->cce + (int64_t)g * 3 * nc;
    for (int k = c->ks - 1; k <= c->ke + 1; ++k)
      for (int j = c->js - 1; j <= c->je + 1; ++j)
        for (int i = c->is - 1; i <= c->ie + 1; ++i) {
          int64_t id = CIDX(c, k, j, i);
          cce[IB1 * nc + id] = w[IVZ * nc + id] * bcc[IB2 * nc + id] - w[IVY * nc + id] * bcc[IB3 * nc + id];
          cce[IB2 * nc + id] = w[IVX * nc + id] * bcc[IB3 * nc + id] - w[IVZ * nc + id] * bcc[IB1 * nc + id];
          cce[IB3 * nc + id] = w[IVY * nc + id] * bcc[IB1 * nc + id] - w[IVX * nc + id] * bcc[IB2 * nc + id];
        }
    const double *e3x1 = c->arr[ARR_E3X1] + g * c->nf1, *e2x1 = c->arr[ARR_E2X1] + g * c->nf1;
    const double *e1x2 = c->arr[ARR_E1X2] + g * c->nf2, *e3x2 = c->arr[ARR_E3X2] + g * c->nf2;
    const double *e1x3 = c->arr[ARR_E1X3] + g * c->nf3, *e2x3 = c->arr[ARR_E2X3] + g * c->nf3;
    const double *wx1 = c->arr[ARR_WX1] + g * c->nf1, *wx2 = c->arr[ARR_WX2] + g * c->nf2,
                 *wx3 = c->arr[ARR_WX3] + g * c->nf3;
    double *e1 = c->arr[ARR_E1] + g * c->ne1, *e2 = c->arr[ARR_E2] + g * c->ne2,
           *e3 = c->arr[ARR_E3] + g * c->ne3;
    const double *c1 = cce + IB1 * nc, *c2 = cce + IB2 * nc, *c3 = cce + IB3 * nc;
#define CC(a, k, j, i) a[CIDX(c, k, j, i)]
#define A1(a, k, j, i) a[F1IDX(c, k, j, i)]
#define A2(a, k, j, i) a[F2IDX(c, k, j, i)]
#define A3(a, k, j, i) a[F3IDX(c, k, j, i)]
    for (int k = c->ks; k <= c->ke + 1; ++k)
      for (int j = c->js; j <= c->je + 1; ++j)
        for (int i = c->is; i <= c->ie + 1; ++i) {
          double de1_l3 = (1.0 - A2(wx2, k - 1, j, i)) * (A3(e1x3, k, j, i) - CC(c1, k - 1, j, i)) +
                          (A2(wx2, k - 1, j, i)) * (A3(e1x3, k, j - 1, i) - CC(c1, k - 1, j - 1, i));
          double de1_r3 = (1.0 - A2(wx2, k, j, i)) * (A3(e1x3, k, j, i) - CC(c1, k, j, i)) +
                          (A2(wx2, k, j, i)) * (A3(e1x3, k, j - 1, i) - CC(c1, k, j - 1, i));
          double de1_l2 = (1.0 - A3(wx3, k, j - 1, i)) * (A2(e1x2, k, j, i) - CC(c1, k, j - 1, i)) +
                          (A3(wx3, k, j - 1, i)) * (A2(e1x2, k - 1, j, i) - CC(c1, k - 1, j - 1, i));
          double de1_r2 = (1.0 - A3(wx3, k, j, i)) * (A2(e1x2, k, j, i) - CC(c1, k, j, i)) +
                          (A3(wx3, k, j, i)) * (A2(e1x2, k - 1, j, i) - CC(c1, k - 1, j, i));
          e1[E1IDX(c, k, j, i)] = 0.25 * (de1_l3 + de1_r3 + de1_l2 + de1_r2 + A2(e1x2, k - 1, j, i) +
                                          A2(e1x2, k, j, i) + A3(e1x3, k, j - 1, i) + A3(e1x3, k, j, i));

          double de2_l3 = (1.0 - A1(wx1, k - 1, j, i)) * (A3(e2x3, k, j, i) - CC(c2, k - 1, j, i)) +
                          (A1(wx1, k - 1, j, i)) * (A3(e2x3, k, j, i - 1) - CC(c2, k - 1, j, i - 1));
          double de2_r3 = (1.0 - A1(wx1, k, j, i)) * (A3(e2x3, k, j, i) - CC(c2, k, j, i)) +
                          (A1(wx1, k, j, i)) * (A3(e2x3, k, j, i - 1) - CC(c2, k, j, i - 1));
          double de2_l1 = (1.0 - A3(wx3, k, j, i - 1)) * (A1(e2x1, k, j, i) - CC(c2, k, j, i - 1)) +
                          (A3(wx3, k, j, i - 1)) * (A1(e2x1, k - 1, j, i) - CC(c2, k - 1, j, i - 1));
          double de2_r1 = (1.0 - A3(wx3, k, j, i)) * (A1(e2x1, k, j, i) - CC(c2, k, j, i)) +
                          (A3(wx3, k, j, i)) * (A1(e2x1, k - 1, j, i) - CC(c2, k - 1, j, i));
          e2[E2IDX(c, k, j, i)] = 0.25 * (de2_l3 + de2_r3 + de2_l1 + de2_r1 + A3(e2x3, k, j, i - 1) +
                                          A3(e2x3, k, j, i) + A1(e2x1, k - 1, j, i) + A1(e2x1, k, j, i));

          double de3_l2 = (1.0 - A1(wx1, k, j - 1, i)) * (A2(e3x2, k, j, i) - CC(c3, k, j - 1, i)) +
                          (A1(wx1, k, j - 1, i)) * (A2(e3x2, k, j, i - 1) - CC(c3, k, j - 1, i - 1));
          double de3_r2 = (1.0 - A1(wx1, k, j, i)) * (A2(e3x2, k, j, i) - CC(c3, k, j, i)) +
                          (A1(wx1, k, j, i)) * (A2(e3x2, k, j, i - 1) - CC(c3, k, j, i - 1));
          double de3_l1 = (1.0 - A2(wx2, k, j, i - 1)) * (A1(e3x1, k, j, i) - CC(c3, k, j, i - 1)) +
                          (A2(wx2, k, j, i - 1)) * (A1(e3x1, k, j - 1, i) - CC(c3, k, j - 1, i - 1));
          double de3_r1 = (1.0 - A2(wx2, k, j, i)) * (A1(e3x1, k, j, i) - CC(c3, k, j, i)) +
                          (A2(wx2, k, j, i)) * (A1(e3x1, k, j - 1, i) - CC(c3, k, j - 1, i));
          e3[E3IDX(c, k, j, i)] = 0.25 * (de3_l1 + de3_r1 + de3_l2 + de3_r2 + A2(e3x2, k, j, i - 1) +
                                          A2(e3x2, k, j, i) + A1(e3x1, k, j - 1, i) + A1(e3x1, k, j, i));
        }
#undef CC
#undef A1
#undef A2
#undef A3
  }
}

/* ---------------------------------------------------------------- integrate */
/* One VL2 stage of TimeIntegratorTaskList (time_integrator.cpp:1563-1660):
 * WeightedAve + register swap (weighted_ave.cpp:33-236, :1583-1584, :1637-1640),
 * AddFluxDivergence (add_flux_divergence.cpp:39-96), SelfGravity source
 * (hydro/srcterms/self_gravity.cpp:35-125), Field::CT (field/ct.cpp:31-116).
 * VL2 weights: stage 1 delta=1, beta=0.5; stage 2 delta=0, beta=1
 * (time_integrator.cpp:199-234). */
static void swap_ptr(double **a, double **b) { double *t = *a; *a = *b; *b = t; }

void orc_integrate(orc_ctx *c, int stage) {
  double beta = (stage == 1) ? 0.5 : 1.0;
  double delta = (stage == 1) ? 1.0 : 0.0;
  int nb = c->nblocks;
  if (stage == 1) {  /* StartupTaskList, time_integrator.cpp:1389-1397 */
    memset(c->arr[ARR_U1], 0, sizeof(double) * c->arrsize[ARR_U1] * nb);
    memset(c->arr[ARR_R1F], 0, sizeof(double) * c->nf1 * nb);
    memset(c->arr[ARR_R2F], 0, sizeof(double) * c->nf2 * nb);
    memset(c->arr[ARR_R3F], 0, sizeof(double) * c->nf3 * nb);
  }
  /* u1 += delta*u over the interior (weighted_ave.cpp:38-97 with wght[0]=1) */
  if (delta != 0.0) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < nb; ++g) {
      double *u = c->arr[ARR_U] + g * c->arrsize[ARR_U];
      double *u1 = c->arr[ARR_U1] + g * c->arrsize[ARR_U1];
      for (int n = 0; n < NHYDRO; ++n)
        for (int k = c->ks; k <= c->ke; ++k)
          for (int j = c->js; j <= c->je; ++j)
            for (int i = c->is; i <= c->ie; ++i) {
              int64_t id = n * c->ncell + CIDX(c, k, j, i);
              u1[id] += delta * u[id];
            }
      double *b1 = c->arr[ARR_B1F] + g * c->nf1, *r1 = c->arr[ARR_R1F] + g * c->nf1;
      double *b2 = c->arr[ARR_B2F] + g * c->nf2, *r2 = c->arr[ARR_R2F] + g * c->nf2;
      double *b3 = c->arr[ARR_B3F] + g * c->nf3, *r3 = c->arr[ARR_R3F] + g * c->nf3;
      for (int k = c->ks; k <= c->ke; ++k)
        for (int j = c->js; j <= c->je; ++j)
          for (int i = c->is; i <= c->ie + 1; ++i) r1[F1IDX(c, k, j, i)] += delta * b1[F1IDX(c, k, j, i)];
      for (int k = c->ks; k <= c->ke; ++k)
        for (int j = c->js; j <= c->je + 1; ++j)
          for (int i = c->is; i <= c->ie; ++i) r2[F2IDX(c, k, j, i)] += delta * b2[F2IDX(c, k, j, i)];
      for (int k = c->ks; k <= c->ke + 1; ++k)
        for (int j = c->js; j <= c->je; ++j)
          for (int i = c->is; i <= c->ie; ++i) r3[F3IDX(c, k, j, i)] += delta * b3[F3IDX(c, k, j, i)];
    }
  }
  /* gamma = (0,1,0): swap u <-> u1, b <-> b1 */
  swap_ptr(&c->arr[ARR_U], &c->arr[ARR_U1]);
  swap_ptr(&c->arr[ARR_B1F], &c->arr[ARR_R1F]);
  swap_ptr(&c->arr[ARR_B2F], &c->arr[ARR_R2F]);
  swap_ptr(&c->arr[ARR_B3F], &c->arr[ARR_R3F]);

  double wght = beta * c->dt;
  int grav = c->d.self_gravity;
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < nb; ++g) {
    int64_t nc = c->ncell;
    double *u = c->arr[ARR_U] + g * c->arrsize[ARR_U];
    const double *f1 = c->arr[ARR_FX1] + g * c->arrsize[ARR_FX1];
    const double *f2 = c->arr[ARR_FX2] + g * c->arrsize[ARR_FX2];
    const double *f3 = c->arr[ARR_FX3] + g * c->arrsize[ARR_FX3];
    const double *dx1 = c->dx1f + g * (c->n1 + 1), *dx2 = c->dx2f + g * (c->n2 + 1),
                 *dx3 = c->dx3f + g * (c->n3 + 1);
    for (int k = c->ks; k <= c->ke; ++k)
      for (int j = c->js; j <= c->je; ++j)
        for (int i = c->is; i <= c->ie; ++i) {
          /* Face?Area / CellVolume, coordinates.cpp:436-537 */
          double a1 = dx2[j] * dx3[k], a1p = dx2[j] * dx3[k];
          double a2 = dx1[i] * dx3[k], a2p = dx1[i] * dx3[k];
          double a3 = dx1[i] * dx2[j], a3p = dx1[i] * dx2[j];
          double vol = dx1[i] * dx2[j] * dx3[k];
          for (int n = 0; n < c->nhydro; ++n) {
            double dflx = (a1p * f1[n * c->nf1 + F1IDX(c, k, j, i + 1)] - a1 * f1[n * c->nf1 + F1IDX(c, k, j, i)]);
            dflx += (a2p * f2[n * c->nf2 + F2IDX(c, k, j + 1, i)] - a2 * f2[n * c->nf2 + F2IDX(c, k, j, i)]);
            dflx += (a3p * f3[n * c->nf3 + F3IDX(c, k + 1, j, i)] - a3 * f3[n * c->nf3 + F3IDX(c, k, j, i)]);
            u[n * nc + CIDX(c, k, j, i)] -= wght * dflx / vol;
          }
        }
    if (grav) {
      /* SelfGravity source with dt = beta*dt (time_integrator.cpp:1650-1655) */
      const double *w = c->arr[ARR_W] + g * c->arrsize[ARR_W];
      const double *phi = c->arr[ARR_PHI] + g * c->arrsize[ARR_PHI];
      double dtg = wght;
      for (int k = c->ks; k <= c->ke; ++k)
        for (int j = c->js; j <= c->je; ++j)
          for (int i = c->is; i <= c->ie; ++i) {
            int64_t id = CIDX(c, k, j, i);
            double dtodx1 = dtg / dx1[i];   /* dx1v == dx1f on a uniform mesh */
            double phic = phi[id];
            double phil = 0.5 * (phi[id - 1] + phi[id]);
            double phir = 0.5 * (phi[id] + phi[id + 1]);
            u[IVX * nc + id] -= dtodx1 * w[IDN * nc + id] * (phir - phil);
            double fl = f1[IDN * c->nf1 + F1IDX(c, k, j, i)], fr = f1[IDN * c->nf1 + F1IDX(c, k, j, i + 1)];
            if (!c->iso) u[IEN * nc + id] -= dtodx1 * (fl * (phic - phil) + fr * (phir - phic));
          }
      for (int k = c->ks; k <= c->ke; ++k)
        for (int j = c->js; j <= c->je; ++j)
          for (int i = c->is; i <= c->ie; ++i) {
            int64_t id = CIDX(c, k, j, i);
            double dtodx2 = dtg / dx2[j];
            double phic = phi[id];
            double phil = 0.5 * (phi[id - c->n1] + phi[id]);
            double phir = 0.5 * (phi[id] + phi[id + c->n1]);
            u[IVY * nc + id] -= dtodx2 * w[IDN * nc + id] * (phir - phil);
            if (!c->iso) u[IEN * nc + id] -= dtodx2 * (f2[IDN * c->nf2 + F2IDX(c, k, j, i)] * (phic - phil) +
                                          f2[IDN * c->nf2 + F2IDX(c, k, j + 1, i)] * (phir - phic));
          }
      int64_t sk = (int64_t)c->n1 * c->n2;
      for (int k = c->ks; k <= c->ke; ++k)
        for (int j = c->js; j <= c->je; ++j)
          for (int i = c->is; i <= c->ie; ++i) {
            int64_t id = CIDX(c, k, j, i);
            double dtodx3 = dtg / dx3[k];
            double phic = phi[id];
            double phil = 0.5 * (phi[id - sk] + phi[id]);
            double phir = 0.5 * (phi[id] + phi[id + sk]);
            u[IVZ * nc + id] -= dtodx3 * w[IDN * nc + id] * (phir - phil);
            if (!c->iso) u[IEN * nc + id] -= dtodx3 * (f3[IDN * c->nf3 + F3IDX(c, k, j, i)] * (phic - phil) +
                                          f3[IDN * c->nf3 + F3IDX(c, k + 1, j, i)] * (phir - phic));
          }
    }
    /* Field::CT */
    double *b1 = c->arr[ARR_B1F] + g * c->nf1, *b2 = c->arr[ARR_B2F] + g * c->nf2,
           *b3 = c->arr[ARR_B3F] + g * c->nf3;
    const double *e1 = c->arr[ARR_E1] + g * c->ne1, *e2 = c->arr[ARR_E2] + g * c->ne2,
                 *e3 = c->arr[ARR_E3] + g * c->ne3;
    for (int k = c->ks; k <= c->ke; ++k)
      for (int j = c->js; j <= c->je; ++j)
        for (int i = c->is; i <= c->ie + 1; ++i) {
          double area = dx2[j] * dx3[k];
          double len = dx3[k], len_p1 = dx3[k];
          b1[F1IDX(c, k, j, i)] -= (wght / area) * (len_p1 * e3[E3IDX(c, k, j + 1, i)] - len * e3[E3IDX(c, k, j, i)]);
          len = dx2[j]; len_p1 = dx2[j];
          b1[F1IDX(c, k, j, i)] += (wght / area) * (len_p1 * e2[E2IDX(c, k + 1, j, i)] - len * e2[E2IDX(c, k, j, i)]);
        }
    for (int k = c->ks; k <= c->ke; ++k)
      for (int j = c->js; j <= c->je + 1; ++j)
        for (int i = c->is; i <= c->ie; ++i) {
          double area = dx1[i] * dx3[k];
          double len = dx3[k], len_ip1 = dx3[k];
          b2[F2IDX(c, k, j, i)] += (wght / area) * (len_ip1 * e3[E3IDX(c, k, j, i + 1)] - len * e3[E3IDX(c, k, j, i)]);
          len = dx1[i]; double len_p1 = dx1[i];
          b2[F2IDX(c, k, j, i)] -= (wght / area) * (len_p1 * e1[E1IDX(c, k + 1, j, i)] - len * e1[E1IDX(c, k, j, i)]);
        }
    for (int k = c->ks; k <= c->ke + 1; ++k)
      for (int j = c->js; j <= c->je; ++j)
        for (int i = c->is; i <= c->ie; ++i) {
          double area = dx1[i] * dx2[j];
          double len = dx2[j], len_ip1 = dx2[j];
          b3[F3IDX(c, k, j, i)] -= (wght / area) * (len_ip1 * e2[E2IDX(c, k, j, i + 1)] - len * e2[E2IDX(c, k, j, i)]);
          len = dx1[i]; double len_p1 = dx1[i];
          b3[F3IDX(c, k, j, i)] += (wght / area) * (len_p1 * e1[E1IDX(c, k, j + 1, i)] - len * e1[E1IDX(c, k, j, i)]);
        }
  }
}

/* ---------------------------------------------------------------- time step */
/* Hydro::NewBlockTimeStep (hydro/new_blockdt.cpp:42-190), min over all blocks. */
double orc_new_block_dt(orc_ctx *c) {
  double mind = DBL_MAX;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : mind)
  for (int g = 0; g < c->nblocks; ++g) {
    int64_t nc = c->ncell;
    const double *w = c->arr[ARR_W] + g * c->arrsize[ARR_W];
    const double *bcc = c->arr[ARR_BCC] + g * c->arrsize[ARR_BCC];
    const double *b1 = c->arr[ARR_B1F] + g * c->nf1, *b2 = c->arr[ARR_B2F] + g * c->nf2,
                 *b3 = c->arr[ARR_B3F] + g * c->nf3;
    const double *dx1 = c->dx1f + g * (c->n1 + 1), *dx2 = c->dx2f + g * (c->n2 + 1),
                 *dx3 = c->dx3f + g * (c->n3 + 1);
    double m = DBL_MAX;
    for (int k = c->ks; k <= c->ke; ++k)
      for (int j = c->js; j <= c->je; ++j)
        for (int i = c->is; i <= c->ie; ++i) {
          int64_t id = CIDX(c, k, j, i);
          double wi[NWAVE];
          wi[IDN] = w[IDN * nc + id]; wi[IVX] = w[IVX * nc + id]; wi[IVY] = w[IVY * nc + id];
          wi[IVZ] = w[IVZ * nc + id]; wi[IPR] = w[IPR * nc + id];
          double bx = bcc[IB1 * nc + id] + fabs(b1[F1IDX(c, k, j, i)] - bcc[IB1 * nc + id]);
          wi[IBY] = bcc[IB2 * nc + id]; wi[IBZ] = bcc[IB3 * nc + id];
          double cf = fast_speed_eos(c, wi, bx);
          double dt1 = dx1[i] / (fabs(wi[IVX]) + cf);
          wi[IBY] = bcc[IB3 * nc + id]; wi[IBZ] = bcc[IB1 * nc + id];
          bx = bcc[IB2 * nc + id] + fabs(b2[F2IDX(c, k, j, i)] - bcc[IB2 * nc + id]);
          cf = fast_speed_eos(c, wi, bx);
          double dt2 = dx2[j] / (fabs(wi[IVY]) + cf);
          wi[IBY] = bcc[IB1 * nc + id]; wi[IBZ] = bcc[IB2 * nc + id];
          bx = bcc[IB3 * nc + id] + fabs(b3[F3IDX(c, k, j, i)] - bcc[IB3 * nc + id]);
          cf = fast_speed_eos(c, wi, bx);
          double dt3 = dx3[k] / (fabs(wi[IVZ]) + cf);
          m = dmin(m, dt1); m = dmin(m, dt2); m = dmin(m, dt3);
        }
    m *= c->d.cfl;
    mind = dmin(mind, m);
  }
  return mind;
}

/* Mesh::NewTimeStep, mesh/mesh.cpp:1078-1119 */
void orc_new_time_step(orc_ctx *c, double tlim) {
  double bdt = orc_new_block_dt(c);
  double dt = 2.0 * c->dt;
  dt = dmin(dt, bdt);
  if (c->time < tlim && (tlim - c->time) < dt) dt = tlim - c->time;
  c->dt = dt;
}

/* The blocking part of Mesh::Initialize (mesh.cpp:1401-1649): first MG solve, ghost
 * exchange of u and b, ConservedToPrimitive on interior + ghosts, first dt. */
void orc_initialize(orc_ctx *c) {
  c->dt = DBL_MAX;
  orc_initialize_restarted(c);
}

/* Mesh::Initialize after the restart constructor (mesh.cpp:1401-1651 with the time, dt
 * and ncycle of the file, mesh.cpp:551-700): the first NewTimeStep is capped by twice
 * the restored dt (set with orc_set_time first). */
void orc_initialize_restarted(orc_ctx *c) {
  if (c->d.self_gravity) orc_mg_solve(c, 0);
  orc_bvals_cc(c, ARR_U);
  orc_bvals_fc(c, 0);
  orc_cons2prim(c);
  orc_new_time_step(c, c->tlim_hint > 0 ? c->tlim_hint : DBL_MAX);
}

/* One cycle of the main loop (main.cpp:458-489) with the VL2 task list: stage 1 uses
 * donor-cell reconstruction (time_integrator.cpp:1466-1467), stage 2 xorder; MG solve
 * after each stage (main.cpp:459-466); then time += dt and NewTimeStep. */
/* Test-only perturbation of the primitives the next reconstruction reads (not part of
 * the reference's algorithm): eps > 0 multiplies every w and bcc value by (1 + eps u),
 * u uniform in [-1, 1) from a xorshift64* stream seeded with seed; eps < 0 rounds them to
 * FP32; eps = 0 turns it off.  It emulates FP32 storage of the primitives, to measure how
 * far the FP64 discretization itself moves under perturbations of that size. */
void orc_set_noise(orc_ctx *c, double eps, uint64_t seed) {
  c->noise_eps = eps;
  c->noise_state = seed * 0x9E3779B97F4A7C15ull + 1;
}

static void perturb_primitives(orc_ctx *c) {
  double *arrs[2] = {c->arr[ARR_W], c->arr[ARR_BCC]};
  int64_t n[2] = {c->arrsize[ARR_W] * c->nblocks, c->arrsize[ARR_BCC] * c->nblocks};
  for (int a = 0; a < 2; ++a)
    for (int64_t i = 0; i < n[a]; ++i) {
      if (c->noise_eps < 0.0) {
        arrs[a][i] = (double)(float)arrs[a][i];
        continue;
      }
      uint64_t x = c->noise_state;
      x ^= x >> 12; x ^= x << 25; x ^= x >> 27;
      c->noise_state = x;
      double u = (double)((x * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
      arrs[a][i] *= 1.0 + c->noise_eps * (2.0 * u - 1.0);
    }
}

void orc_cycle(orc_ctx *c, double tlim) {
  for (int stage = 1; stage <= 2; ++stage) {
    orc_calculate_fluxes(c, stage == 1 ? 1 : c->d.xorder);
    orc_corner_e(c);
    orc_integrate(c, stage);
    orc_bvals_cc(c, ARR_U);
    orc_bvals_fc(c, 0);
    orc_cons2prim(c);
    if (c->noise_eps != 0.0) perturb_primitives(c);
    if (c->d.self_gravity) orc_mg_solve(c, stage);
  }
  c->ncycle++;
  c->time += c->dt;
  orc_new_time_step(c, tlim);
}

/* ================================================================ multigrid */
#define MIDX(m, l, k, j, i) ((((int64_t)(k)) * ((m)->n2[l] + 2) + (j)) * ((m)->n1[l] + 2) + (i))

/* MGGravity::Smooth, gravity/mg_gravity.cpp:220-273 (omega = 1.15, mg_gravity.hpp:43) */
static void mg_smooth(orc_ctx *c, mglevels *m, int l, int color) {
  int ll = m->nlev - 1 - l;
  double dx = m->rdx * (double)(1 << ll);
  double dx2 = SQR(dx);
  double isix = 1.15 / 6.0;
  color ^= c->coffset;
  int nx = m->n1[l] + 2, ny = m->n2[l] + 2;
  double *u = m->u[l];
  const double *src = m->src[l];
  int64_t sj = nx, sk = (int64_t)nx * ny;
  for (int k = 1; k <= m->n3[l]; ++k)
    for (int j = 1; j <= m->n2[l]; ++j) {
      int cc = (color + k + j) & 1;
      for (int i = 1 + cc; i <= m->n1[l]; i += 2) {
        int64_t id = MIDX(m, l, k, j, i);
        u[id] -= ((6.0 * u[id] - u[id + sk] - u[id + sj] - u[id + 1] - u[id - sk] - u[id - sj] - u[id - 1]) +
                  src[id] * dx2) * isix;
      }
    }
}

/* MGGravity::CalculateDefect, gravity/mg_gravity.cpp:283-314 */
static void mg_defect(mglevels *m, int l) {
  int ll = m->nlev - 1 - l;
  double dx = m->rdx * (double)(1 << ll);
  double idx2 = 1.0 / SQR(dx);
  int nx = m->n1[l] + 2, ny = m->n2[l] + 2;
  int64_t sj = nx, sk = (int64_t)nx * ny;
  const double *u = m->u[l], *src = m->src[l];
  double *def = m->def[l];
  for (int k = 1; k <= m->n3[l]; ++k)
    for (int j = 1; j <= m->n2[l]; ++j)
      for (int i = 1; i <= m->n1[l]; ++i) {
        int64_t id = MIDX(m, l, k, j, i);
        def[id] = (6.0 * u[id] - u[id + sk] - u[id + sj] - u[id + 1] - u[id - sk] - u[id - sj] - u[id - 1]) * idx2 +
                  src[id];
      }
}

/* Multigrid::Restrict, multigrid/multigrid.cpp:684-724 (fine level l -> l-1) */
static void mg_restrict(mglevels *m, double *dst, const double *src, int l) {
  int lc = l - 1;
  for (int k = 1; k <= m->n3[lc]; ++k) {
    int fk = 2 * k - 1;
    for (int j = 1; j <= m->n2[lc]; ++j) {
      int fj = 2 * j - 1;
      for (int i = 1; i <= m->n1[lc]; ++i) {
        int fi = 2 * i - 1;
#define S(kk, jj, ii) src[MIDX(m, l, kk, jj, ii)]
        dst[MIDX(m, lc, k, j, i)] = 0.125 * (S(fk, fj, fi) + S(fk, fj, fi + 1) + S(fk, fj + 1, fi) + S(fk, fj + 1, fi + 1) +
                                             S(fk + 1, fj, fi) + S(fk + 1, fj, fi + 1) + S(fk + 1, fj + 1, fi) +
                                             S(fk + 1, fj + 1, fi + 1));
#undef S
      }
    }
  }
}

/* Multigrid::ProlongateAndCorrect, multigrid/multigrid.cpp:733-829 (l-1 -> l) */
static void mg_prolongate(mglevels *m, int l, int fas) {
  int lc = l - 1;
  if (fas) {   /* ProlongateAndCorrectBlock, FAS: u -= uold on the coarse level first */
    size_t sz = (size_t)(m->n1[lc] + 2) * (m->n2[lc] + 2) * (m->n3[lc] + 2);
    for (size_t q = 0; q < sz; ++q) m->u[lc][q] -= m->uold[lc][q];
  }
  const double *s = m->u[lc];
  double *d = m->u[l];
  for (int k = 1; k <= m->n3[lc]; ++k) {
    int fk = 2 * (k - 1) + 1;
    for (int j = 1; j <= m->n2[lc]; ++j) {
      int fj = 2 * (j - 1) + 1;
      for (int i = 1; i <= m->n1[lc]; ++i) {
        int fi = 2 * (i - 1) + 1;
#define S(kk, jj, ii) s[MIDX(m, lc, kk, jj, ii)]
#define D(kk, jj, ii) d[MIDX(m, l, kk, jj, ii)]
        for (int ok = 0; ok < 2; ++ok)
          for (int oj = 0; oj < 2; ++oj)
            for (int oi = 0; oi < 2; ++oi) {
              int dk = ok ? 1 : -1, dj = oj ? 1 : -1, di = oi ? 1 : -1;
              D(fk + ok, fj + oj, fi + oi) +=
                  0.015625 * (27.0 * S(k, j, i) + S(k + dk, j + dj, i + di) +
                              9.0 * (S(k, j, i + di) + S(k, j + dj, i) + S(k + dk, j, i)) +
                              3.0 * (S(k + dk, j + dj, i) + S(k + dk, j, i + di) + S(k, j + dj, i + di)));
            }
#undef S
#undef D
      }
    }
  }
}

/* Multigrid::FMGProlongate, multigrid/multigrid.cpp:838-1048: tricubic-type
 * interpolation, the tensor product of 1-D weights (5,30,-3)/32 toward the fine
 * cell's side, summed term by term in (k,j,i) row-major order and divided by 32768. */
static void mg_fmg_prolongate(mglevels *m, int l) {
  int lc = l - 1;
  const double *s = m->u[lc];
  double *d = m->u[l];
  static const double wlo[3] = {5.0, 30.0, -3.0}, whi[3] = {-3.0, 30.0, 5.0};
  for (int k = 1; k <= m->n3[lc]; ++k) {
    int fk = 2 * (k - 1) + 1;
    for (int j = 1; j <= m->n2[lc]; ++j) {
      int fj = 2 * (j - 1) + 1;
      for (int i = 1; i <= m->n1[lc]; ++i) {
        int fi = 2 * (i - 1) + 1;
        for (int ok = 0; ok < 2; ++ok)
          for (int oj = 0; oj < 2; ++oj)
            for (int oi = 0; oi < 2; ++oi) {
              const double *wk = ok ? whi : wlo, *wj = oj ? whi : wlo, *wi = oi ? whi : wlo;
              double acc = 0.0;
              int first = 1;
              for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                  for (int e = 0; e < 3; ++e) {
                    double coef = wk[a] * wj[b] * wi[e];
                    double t = coef * s[MIDX(m, lc, k - 1 + a, j - 1 + b, i - 1 + e)];
                    if (first) { acc = t; first = 0; }
                    else acc = acc + t;
                  }
              d[MIDX(m, l, fk + ok, fj + oj, fi + oi)] = acc / 32768.0;
            }
      }
    }
  }
}

/* Same-level ghost fill of every block's level l (bvals_mg.cpp:447-520,634-747,
 * 866-915): each ghost cell copies the neighbour block's interior cell. */
static void mg_phys_bc(const orc_ctx *c, mglevels *m, int l);

static void mg_fill_blocks(orc_ctx *c, int l, int which) {
  int nb = c->nblocks;
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < nb; ++g) {
    mglevels *m = &c->mgb[g];
    int n = m->n1[l];
    double *a = (which == 0) ? m->u[l] : m->src[l];
    const int lxo[3] = {g % c->nb1, (g / c->nb1) % c->nb2, g / (c->nb1 * c->nb2)};
    for (int k = 0; k <= n + 1; ++k) {
      int o3 = (k == 0) ? -1 : (k == n + 1 ? 1 : 0);
      for (int j = 0; j <= n + 1; ++j) {
        int o2 = (j == 0) ? -1 : (j == n + 1 ? 1 : 0);
        for (int i = 0; i <= n + 1; ++i) {
          int o1 = (i == 0) ? -1 : (i == n + 1 ? 1 : 0);
          if (!o1 && !o2 && !o3) continue;
          if (c->d.mg_bc == 1 && (lxo[0] + o1 < 0 || lxo[0] + o1 >= c->nb1 || lxo[1] + o2 < 0 ||
                                  lxo[1] + o2 >= c->nb2 || lxo[2] + o3 < 0 || lxo[2] + o3 >= c->nb3))
            continue;   /* no neighbour beyond a physical boundary */
          mglevels *s = &c->mgb[nbr_gid(c, g, o1, o2, o3)];
          double *sa = (which == 0) ? s->u[l] : s->src[l];
          a[MIDX(m, l, k, j, i)] = sa[MIDX(m, l, k - o3 * n, j - o2 * n, i - o1 * n)];
        }
      }
    }
  }
  if (c->d.mg_bc == 1 && which == 0)
    for (int g = 0; g < nb; ++g) mg_phys_bc(c, &c->mgb[g], l);
}

/* Root grid periodic physical boundaries, MGBoundaryValues::ApplyPhysicalBoundaries
 * (bvals_mg.cpp:321-370) with MGPeriodic* (multigrid/mgbval_periodic.cpp:26-144):
 * x1 faces over interior (j,k), then x2 over all i, then x3 over all (i,j). */
static void mg_root_bc(orc_ctx *c, mglevels *m, int l) {
  if (c->d.mg_bc == 1) {
    mg_phys_bc(c, m, l);
    return;
  }
  int n1 = m->n1[l], n2 = m->n2[l], n3 = m->n3[l];
  double *u = m->u[l];
  for (int k = 1; k <= n3; ++k)
    for (int j = 1; j <= n2; ++j) {
      u[MIDX(m, l, k, j, 0)] = u[MIDX(m, l, k, j, n1)];
      u[MIDX(m, l, k, j, n1 + 1)] = u[MIDX(m, l, k, j, 1)];
    }
  for (int k = 1; k <= n3; ++k)
    for (int i = 0; i <= n1 + 1; ++i) {
      u[MIDX(m, l, k, 0, i)] = u[MIDX(m, l, k, n2, i)];
      u[MIDX(m, l, k, n2 + 1, i)] = u[MIDX(m, l, k, 1, i)];
    }
  for (int j = 0; j <= n2 + 1; ++j)
    for (int i = 0; i <= n1 + 1; ++i) {
      u[MIDX(m, l, 0, j, i)] = u[MIDX(m, l, n3, j, i)];
      u[MIDX(m, l, n3 + 1, j, i)] = u[MIDX(m, l, 1, j, i)];
    }
}

static void mg_zero(mglevels *m, int l) {
  size_t sz = (size_t)(m->n1[l] + 2) * (m->n2[l] + 2) * (m->n3[l] + 2);
  memset(m->u[l], 0, sz * sizeof(double));
}

/* MGCoordinates::CalculateMGCoordinates (multigrid.cpp:1321-1350), ngh = 1: face
 * coordinate x?f(i) and centre x?v(i) of level l, i = 0..n+1 (0 and n+1 ghosts). */
static double mgc_f(const mglevels *m, int l, int dir, int i) {
  int nc = dir == 0 ? m->n1[l] : (dir == 1 ? m->n2[l] : m->n3[l]);
  if (i == 1) return m->xmin[dir];
  if (i == nc + 1) return m->xmax[dir];
  double dx = (m->xmax[dir] - m->xmin[dir]) / nc;
  return m->xmin[dir] + (i - 1) * dx;
}
static double mgc_v(const mglevels *m, int l, int dir, int i) {
  return 0.5 * (mgc_f(m, l, dir, i) + mgc_f(m, l, dir, i + 1));
}

/* Multipole potential at (x, y, z) relative to the origin, MGMultipoleInnerX1..
 * OuterX3 (multigrid/mgbval_multipole.cpp:28-426), mporder 2 or 4. */
static double mp_phis(const orc_ctx *c, double x, double y, double z) {
  const double *mp = c->mpc;
  double x2 = x * x, y2 = y * y, z2 = z * z;
  double xy = x * y, yz = y * z, zx = z * x;
  double r2 = x2 + y2 + z2;
  double ir2 = 1.0 / r2, ir1 = sqrt(ir2);
  double ir3 = ir2 * ir1, ir5 = ir3 * ir2;
  if (c->d.mporder == 2)
    return ir1 * mp[0] + ir3 * (mp[1] * y + mp[2] * z + mp[3] * x) +
           ir5 * (mp[4] * xy + mp[5] * yz + (3.0 * z2 - r2) * mp[6] + mp[7] * zx + mp[8] * 0.5 * (x2 - y2));
  double ir7 = ir5 * ir2, ir9 = ir7 * ir2;
  double hx2my2 = 0.5 * (x2 - y2);
  double x2mty2 = x2 - 3.0 * y2;
  double tx2my2 = 3.0 * x2 - y2;
  return ir1 * mp[0] + ir3 * (mp[1] * y + mp[2] * z + mp[3] * x) +
         ir5 * (mp[4] * xy + mp[5] * yz + (3.0 * z2 - r2) * mp[6] + mp[7] * zx + mp[8] * 0.5 * (x2 - y2)) +
         ir7 * (y * tx2my2 * mp[9] + x * x2mty2 * mp[15] + xy * z * mp[10] + z * hx2my2 * mp[14] +
                (5.0 * z2 - r2) * (y * mp[11] + x * mp[13]) + z * (z2 - 3.0 * r2) * mp[12]) +
         ir9 * (xy * hx2my2 * mp[16] + 0.125 * (x2 * x2mty2 - y2 * tx2my2) * mp[24] + yz * tx2my2 * mp[17] +
                zx * x2mty2 * mp[23] + (7.0 * z2 - r2) * (xy * mp[18] + hx2my2 * mp[22]) +
                (7.0 * z2 - 3.0 * r2) * (yz * mp[19] + zx * mp[21]) +
                (35.0 * z2 * z2 - 30.0 * z2 * r2 + 3.0 * r2 * r2) * mp[20]);
}

/* MGBoundaryValues::ApplyPhysicalBoundaries (bvals_mg.cpp:321-368) with the multipole
 * functions: ghost = 2 phi_s(face point) - adjacent cell.  x1 faces over the interior
 * (j, k) (plus the j/k ghosts where that face is not physical), then x2 faces over all
 * i, then x3 faces over all (i, j). */
static void mg_phys_bc(const orc_ctx *c, mglevels *m, int l) {
  int n1 = m->n1[l], n2 = m->n2[l], n3 = m->n3[l];
  double *u = m->u[l];
  const double xo = c->mpo[0], yo = c->mpo[1], zo = c->mpo[2];
  int bjs = m->bnd[2] ? 1 : 0, bje = m->bnd[3] ? n2 : n2 + 1;
  int bks = m->bnd[4] ? 1 : 0, bke = m->bnd[5] ? n3 : n3 + 1;
  for (int side = 0; side < 2; ++side) {
    if (!m->bnd[side]) continue;
    int ia = side ? n1 : 1, ig = side ? n1 + 1 : 0;
    double x = mgc_f(m, l, 0, side ? n1 + 1 : 1) - xo;
    for (int k = bks; k <= bke; ++k) {
      double z = mgc_v(m, l, 2, k) - zo;
      for (int j = bjs; j <= bje; ++j) {
        double y = mgc_v(m, l, 1, j) - yo;
        u[MIDX(m, l, k, j, ig)] = 2.0 * mp_phis(c, x, y, z) - u[MIDX(m, l, k, j, ia)];
      }
    }
  }
  for (int side = 0; side < 2; ++side) {
    if (!m->bnd[2 + side]) continue;
    int ja = side ? n2 : 1, jg = side ? n2 + 1 : 0;
    double y = mgc_f(m, l, 1, side ? n2 + 1 : 1) - yo;
    for (int k = bks; k <= bke; ++k) {
      double z = mgc_v(m, l, 2, k) - zo;
      for (int i = 0; i <= n1 + 1; ++i) {
        double x = mgc_v(m, l, 0, i) - xo;
        u[MIDX(m, l, k, jg, i)] = 2.0 * mp_phis(c, x, y, z) - u[MIDX(m, l, k, ja, i)];
      }
    }
  }
  for (int side = 0; side < 2; ++side) {
    if (!m->bnd[4 + side]) continue;
    int ka = side ? n3 : 1, kg = side ? n3 + 1 : 0;
    double z = mgc_f(m, l, 2, side ? n3 + 1 : 1) - zo;
    for (int j = 0; j <= n2 + 1; ++j) {
      double y = mgc_v(m, l, 1, j) - yo;
      for (int i = 0; i <= n1 + 1; ++i) {
        double x = mgc_v(m, l, 0, i) - xo;
        u[MIDX(m, l, kg, j, i)] = 2.0 * mp_phis(c, x, y, z) - u[MIDX(m, l, ka, j, i)];
      }
    }
  }
}

/* FAS: StoreOldData (multigrid.cpp:645-653) and MGGravity::CalculateFASRHS
 * (gravity/mg_gravity.cpp:324-351) on level l. */
static void mg_fas_rhs(mglevels *m, int l) {
  int ll = m->nlev - 1 - l;
  double dx = m->rdx * (double)(1 << ll);
  double idx2 = 1.0 / SQR(dx);
  size_t sz = (size_t)(m->n1[l] + 2) * (m->n2[l] + 2) * (m->n3[l] + 2);
  memcpy(m->uold[l], m->u[l], sz * sizeof(double));
  int nx = m->n1[l] + 2, ny = m->n2[l] + 2;
  int64_t sj = nx, sk = (int64_t)nx * ny;
  const double *u = m->u[l];
  double *src = m->src[l];
  for (int k = 1; k <= m->n3[l]; ++k)
    for (int j = 1; j <= m->n2[l]; ++j)
      for (int i = 1; i <= m->n1[l]; ++i) {
        int64_t id = MIDX(m, l, k, j, i);
        src[id] -= (6.0 * u[id] - u[id + sk] - u[id + sj] - u[id + 1] - u[id - sk] - u[id - sj] - u[id - 1]) * idx2;
      }
}

/* MultigridDriver::CalculateCenterOfMass (multigrid_driver.cpp:1984-2015) with
 * Multigrid::CalculateCenterOfMass (multigrid.cpp:1270-1298), blocks in gid order. */
static void mp_center_of_mass(orc_ctx *c) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int g = 0; g < c->nblocks; ++g) {
    mglevels *m = &c->mgb[g];
    int l = m->nlev - 1;
    double vol = (mgc_f(m, l, 0, 2) - mgc_f(m, l, 0, 1)) * (mgc_f(m, l, 1, 2) - mgc_f(m, l, 1, 1)) *
                 (mgc_f(m, l, 2, 2) - mgc_f(m, l, 2, 1));
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    for (int k = 1; k <= m->n3[l]; ++k) {
      double z = mgc_v(m, l, 2, k);
      for (int j = 1; j <= m->n2[l]; ++j) {
        double y = mgc_v(m, l, 1, j);
        for (int i = 1; i <= m->n1[l]; ++i) {
          double x = mgc_v(m, l, 0, i);
          double s = m->src[l][MIDX(m, l, k, j, i)] * vol;
          m0 += s;
          m1 += s * y;
          m2 += s * z;
          m3 += s * x;
        }
      }
    }
    acc[0] += m0; acc[1] += m1; acc[2] += m2; acc[3] += m3;
  }
  double im = 1.0 / acc[0];
  c->mpo[0] = im * acc[3];
  c->mpo[1] = im * acc[1];
  c->mpo[2] = im * acc[2];
}

/* MultigridDriver::CalculateMultipoleCoefficients + ScaleMultipoleCoefficients
 * (multigrid_driver.cpp:1903-1981) with Multigrid::CalculateMultipoleCoefficients
 * (multigrid.cpp:1054-1264); with "nodipole" the dipole sums are not formed
 * (multigrid.cpp:1072-1130, 1200-1226) and stay zero. */
static void mp_coefficients(orc_ctx *c) {
  const int np = (c->d.mporder == 4) ? 25 : 9;
  double acc[25] = {0.0};
  const double xo = c->mpo[0], yo = c->mpo[1], zo = c->mpo[2];
  for (int g = 0; g < c->nblocks; ++g) {
    mglevels *m = &c->mgb[g];
    int l = m->nlev - 1;
    double vol = (mgc_f(m, l, 0, 2) - mgc_f(m, l, 0, 1)) * (mgc_f(m, l, 1, 2) - mgc_f(m, l, 1, 1)) *
                 (mgc_f(m, l, 2, 2) - mgc_f(m, l, 2, 1));
    double mm[25] = {0.0};
    for (int k = 1; k <= m->n3[l]; ++k) {
      double z = mgc_v(m, l, 2, k) - zo;
      double z2 = z * z;
      for (int j = 1; j <= m->n2[l]; ++j) {
        double y = mgc_v(m, l, 1, j) - yo;
        double y2 = y * y, yz = y * z;
        for (int i = 1; i <= m->n1[l]; ++i) {
          double x = mgc_v(m, l, 0, i) - xo;
          double x2 = x * x, xy = x * y, zx = z * x;
          double r2 = x2 + y2 + z2;
          double s = m->src[l][MIDX(m, l, k, j, i)] * vol;
          mm[0] += s;
          if (!c->d.mp_nodipole) {
            mm[1] += s * y;
            mm[2] += s * z;
            mm[3] += s * x;
          }
          mm[4] += s * xy;
          mm[5] += s * yz;
          mm[6] += s * (3.0 * z2 - r2);
          mm[7] += s * zx;
          if (np == 9) {
            mm[8] += s * 0.5 * (x2 - y2);
            continue;
          }
          double hx2my2 = 0.5 * (x2 - y2);
          double x2mty2 = x2 - 3.0 * y2;
          double tx2my2 = 3.0 * x2 - y2;
          double fz2mr2 = 5.0 * z2 - r2;
          double sz2mr2 = 7.0 * z2 - r2;
          double sz2mtr2 = 7.0 * z2 - 3.0 * r2;
          mm[8] += s * hx2my2;
          mm[9] += s * y * tx2my2;
          mm[10] += s * xy * z;
          mm[11] += s * y * fz2mr2;
          mm[12] += s * z * (z2 - 3.0 * r2);
          mm[13] += s * x * fz2mr2;
          mm[14] += s * z * hx2my2;
          mm[15] += s * x * x2mty2;
          mm[16] += s * xy * hx2my2;
          mm[17] += s * yz * tx2my2;
          mm[18] += s * xy * sz2mr2;
          mm[19] += s * yz * sz2mtr2;
          mm[20] += s * (35.0 * z2 * z2 - 30.0 * z2 * r2 + 3.0 * r2 * r2);
          mm[21] += s * zx * sz2mtr2;
          mm[22] += s * hx2my2 * sz2mr2;
          mm[23] += s * zx * x2mty2;
          mm[24] += s * 0.125 * (x2 * x2mty2 - y2 * tx2my2);
        }
      }
    }
    for (int q = 0; q < np; ++q) acc[q] += mm[q];
  }
  const double PI_ = 3.14159265358979323846;
  const double c0 = -0.25 / PI_, c1 = -0.25 / PI_, c2 = -0.0625 / PI_, c2a = -0.75 / PI_;
  const double c30 = -0.0625 / PI_, c31 = -0.0625 * 1.5 / PI_, c32 = -0.25 * 15.0 / PI_,
               c33 = -0.0625 * 2.5 / PI_;
  const double c40 = -0.0625 * 0.0625 / PI_, c41 = -0.0625 * 2.5 / PI_, c42 = -0.0625 * 5.0 / PI_,
               c43 = -0.0625 * 17.5 / PI_, c44 = -0.25 * 35.0 / PI_;
  const double sc[25] = {c0, c1, c1, c1, c2a, c2a, c2, c2a, c2a, c33, c32, c31, c30,
                         c31, c32, c33, c44, c43, c42, c41, c40, c41, c42, c43, c44};
  for (int q = 0; q < 25; ++q) c->mpc[q] = (q < np) ? acc[q] * sc[q] : 0.0;
}

/* blocks: smooth red/black with ghost fills, then RestrictBlock
 * (mg_task_list.cpp:382-... nsmooth=1; multigrid.cpp:331-355) */
static void mg_blocks_to_coarser(orc_ctx *c, int bl) {
  mg_fill_blocks(c, bl, 0);
  if (c->fas && c->current < c->fmglevel)   /* MultigridTaskList::CalculateFASRHS */
    for (int g = 0; g < c->nblocks; ++g) mg_fas_rhs(&c->mgb[g], bl);
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < c->nblocks; ++g) mg_smooth(c, &c->mgb[g], bl, 0);
  mg_fill_blocks(c, bl, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < c->nblocks; ++g) mg_smooth(c, &c->mgb[g], bl, 1);
  mg_fill_blocks(c, bl, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < c->nblocks; ++g) {
    mglevels *m = &c->mgb[g];
    mg_defect(m, bl);
    mg_restrict(m, m->src[bl - 1], m->def[bl], bl);
    if (c->fas) mg_restrict(m, m->u[bl - 1], m->u[bl], bl);   /* FAS restricts u too */
    else mg_zero(m, bl - 1);
  }
}

static void mg_root_to_coarser(orc_ctx *c, int rl) {
  mglevels *m = &c->mgr;
  mg_root_bc(c, m, rl);
  if (c->fas && rl < c->fmglevel) mg_fas_rhs(m, rl);
  mg_smooth(c, m, rl, 0);
  mg_root_bc(c, m, rl);
  mg_smooth(c, m, rl, 1);
  mg_root_bc(c, m, rl);
  mg_defect(m, rl);
  mg_restrict(m, m->src[rl - 1], m->def[rl], rl);
  if (c->fas) mg_restrict(m, m->u[rl - 1], m->u[rl], rl);
  else mg_zero(m, rl - 1);
}

/* TransferFromBlocksToRoot (multigrid_driver.cpp:467-527), non-FAS */
static void mg_blocks_to_root(orc_ctx *c, int initflag) {
  mglevels *r = &c->mgr;
  int rl = c->nrootlevel - 1;
  for (int g = 0; g < c->nblocks; ++g) {
    int lx1 = g % c->nb1, lx2 = (g / c->nb1) % c->nb2, lx3 = g / (c->nb1 * c->nb2);
    r->src[rl][MIDX(r, rl, lx3 + 1, lx2 + 1, lx1 + 1)] = c->mgb[g].src[0][MIDX(&c->mgb[g], 0, 1, 1, 1)];
    if (c->fas && !initflag)
      r->u[rl][MIDX(r, rl, lx3 + 1, lx2 + 1, lx1 + 1)] = c->mgb[g].u[0][MIDX(&c->mgb[g], 0, 1, 1, 1)];
  }
}

/* TransferFromRootToBlocks / SetFromRootGrid (multigrid.cpp:480-550) */
static void mg_root_to_blocks(orc_ctx *c, int folddata) {
  mglevels *r = &c->mgr;
  int rl = c->nrootlevel - 1;
  for (int g = 0; g < c->nblocks; ++g) {
    int lx1 = g % c->nb1, lx2 = (g / c->nb1) % c->nb2, lx3 = g / (c->nb1 * c->nb2);
    mglevels *m = &c->mgb[g];
    for (int k = 0; k <= 2; ++k)
      for (int j = 0; j <= 2; ++j)
        for (int i = 0; i <= 2; ++i)
        {
          m->u[0][MIDX(m, 0, k, j, i)] = r->u[rl][MIDX(r, rl, lx3 + k, lx2 + j, lx1 + i)];
          if (folddata) m->uold[0][MIDX(m, 0, k, j, i)] = r->uold[rl][MIDX(r, rl, lx3 + k, lx2 + j, lx1 + i)];
        }
  }
}

static void mg_one_step_to_coarser(orc_ctx *c) {
  int cur = c->current;
  if (cur >= c->nrootlevel) {
    int bl = cur - c->nrootlevel + 1;
    mg_blocks_to_coarser(c, bl);
    if (cur == c->nrootlevel) {
      mg_blocks_to_root(c, 0);
      if (!c->fas) mg_zero(&c->mgr, c->nrootlevel - 1);
    }
  } else {
    mg_root_to_coarser(c, cur);
  }
  c->current--;
}

/* SolveCoarsestGrid, multigrid_driver.cpp:778-819 (periodic) */
static void mg_root_total_sub(orc_ctx *c, int l) {
  mglevels *m = &c->mgr;
  int ll = m->nlev - 1 - l;
  double dx = m->rdx * (double)(1 << ll);
  double rdy = (c->d.x2max - c->d.x2min) / (double)c->nb2;
  double rdz = (c->d.x3max - c->d.x3min) / (double)c->nb3;
  double dy = rdy * (double)(1 << ll), dz = rdz * (double)(1 << ll);
  double s = 0.0;
  for (int k = 1; k <= m->n3[l]; ++k)
    for (int j = 1; j <= m->n2[l]; ++j)
      for (int i = 1; i <= m->n1[l]; ++i) s += m->u[l][MIDX(m, l, k, j, i)];
  double vol = (c->d.x1max - c->d.x1min) * (c->d.x2max - c->d.x2min) * (c->d.x3max - c->d.x3min);
  double ave = s * dx * dy * dz / vol;
  /* Multigrid::SubtractAverage acts on the finest level of the root object
   * (multigrid.cpp:627-641) */
  mglevels *mm = m;
  int lf = m->nlev - 1;
  size_t sz = (size_t)(mm->n1[lf] + 2) * (mm->n2[lf] + 2) * (mm->n3[lf] + 2);
  for (size_t q = 0; q < sz; ++q) mm->u[lf][q] -= ave;
}

static void mg_solve_coarsest(orc_ctx *c) {
  int ni = c->nb1 > c->nb2 ? c->nb1 : c->nb2;
  ni = ni > c->nb3 ? ni : c->nb3;
  ni >>= (c->nrootlevel - 1);
  /* periodic: fsubtract_average_; multipole: FAS without average subtraction */
  if (!c->fas && ni == 1) {
    mg_zero(&c->mgr, 0);
  } else {
    if (!c->fas) mg_root_total_sub(c, 0);
    mg_root_bc(c, &c->mgr, 0);
    if (c->fas) mg_fas_rhs(&c->mgr, 0);
    for (int it = 0; it < ni; ++it) {
      mg_smooth(c, &c->mgr, 0, 0);
      mg_root_bc(c, &c->mgr, 0);
      mg_smooth(c, &c->mgr, 0, 1);
      mg_root_bc(c, &c->mgr, 0);
    }
    if (!c->fas) mg_root_total_sub(c, 0);
  }
}

static void mg_one_step_to_finer(orc_ctx *c) {
  int cur = c->current;
  if (cur == c->nrootlevel - 1) {          /* multigrid_driver.cpp:586-590 */
    mg_root_bc(c, &c->mgr, cur);
    mg_root_to_blocks(c, c->fas);
  }
  if (cur >= c->nrootlevel - 1) {          /* block levels, SetMGTaskListToFiner */
    int bl = cur - c->nrootlevel + 1;      /* coarse block level */
    /* flag 1 (first step on block levels): the coarse ghosts came from the root grid;
     * otherwise they are exchanged first (mg_task_list.cpp:287-300) */
    if (cur != c->nrootlevel - 1) mg_fill_blocks(c, bl, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < c->nblocks; ++g) mg_prolongate(&c->mgb[g], bl + 1, c->fas);
    mg_fill_blocks(c, bl + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < c->nblocks; ++g) mg_smooth(c, &c->mgb[g], bl + 1, 0);
    mg_fill_blocks(c, bl + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < c->nblocks; ++g) mg_smooth(c, &c->mgb[g], bl + 1, 1);
    if (cur == c->ntotal - 2) mg_fill_blocks(c, bl + 1, 0);   /* flag 2: last */
  } else {                                  /* root grid */
    mglevels *m = &c->mgr;
    mg_root_bc(c, m, cur);
    mg_prolongate(m, cur + 1, c->fas);
    mg_root_bc(c, m, cur + 1);
    mg_smooth(c, m, cur + 1, 0);
    mg_root_bc(c, m, cur + 1);
    mg_smooth(c, m, cur + 1, 1);
  }
  c->current++;
}

static void mg_vcycle(orc_ctx *c) {   /* SolveVCycle(1,1), multigrid_driver.cpp:684-693 */
  int start = c->current;
  c->coffset ^= 1;
  while (c->current > 0) mg_one_step_to_coarser(c);
  mg_solve_coarsest(c);
  while (c->current < start) mg_one_step_to_finer(c);
}

static void mg_fmg_prolongate_step(orc_ctx *c) {   /* multigrid_driver.cpp:553-575 */
  int cur = c->current;
  if (cur == c->nrootlevel - 1) {
    mg_root_bc(c, &c->mgr, cur);
    mg_root_to_blocks(c, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < c->nblocks; ++g) mg_fmg_prolongate(&c->mgb[g], 1);
  } else if (cur >= c->nrootlevel - 1) {
    int bl = cur - c->nrootlevel + 1;
    mg_fill_blocks(c, bl, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int g = 0; g < c->nblocks; ++g) mg_fmg_prolongate(&c->mgb[g], bl + 1);
  } else {
    mg_root_bc(c, &c->mgr, cur);
    mg_fmg_prolongate(&c->mgr, cur + 1);
  }
  c->current++;
}

/* SubtractAverage (multigrid_driver.cpp:305-330) for src (type 0) or u (type 1) of
 * the blocks' finest level; the per-block totals are CalculateTotal
 * (multigrid.cpp:602-625). */
static double mg_subtract_average(orc_ctx *c, int type) {
  int l = c->nmblevel - 1;
  double total = 0.0;
  for (int g = 0; g < c->nblocks; ++g) {
    mglevels *m = &c->mgb[g];
    double *a = type ? m->u[l] : m->src[l];
    double s = 0.0;
    for (int k = 1; k <= m->n3[l]; ++k)
      for (int j = 1; j <= m->n2[l]; ++j)
        for (int i = 1; i <= m->n1[l]; ++i) s += a[MIDX(m, l, k, j, i)];
    double dx = m->rdx;
    int lx2 = (g / c->nb1) % c->nb2, lx3 = g / (c->nb1 * c->nb2);
    (void)lx2; (void)lx3;
    double dy = (c->d.x2max - c->d.x2min) / (double)c->d.nx2;   /* rdy_ of the block */
    double dz = (c->d.x3max - c->d.x3min) / (double)c->d.nx3;
    {
      /* rdy_/rdz_ are block extents / block cells (multigrid.cpp:61-63) */
      double b2min = (lx2 == 0) ? c->d.x2min : uniform_gen(meshgen_x(lx2, c->nb2), c->d.x2min, c->d.x2max);
      double b2max = (lx2 == c->nb2 - 1) ? c->d.x2max : uniform_gen(meshgen_x(lx2 + 1, c->nb2), c->d.x2min, c->d.x2max);
      double b3min = (lx3 == 0) ? c->d.x3min : uniform_gen(meshgen_x(lx3, c->nb3), c->d.x3min, c->d.x3max);
      double b3max = (lx3 == c->nb3 - 1) ? c->d.x3max : uniform_gen(meshgen_x(lx3 + 1, c->nb3), c->d.x3min, c->d.x3max);
      dy = (b2max - b2min) / (double)c->d.bnx2;
      dz = (b3max - b3min) / (double)c->d.bnx3;
    }
    total += s * dx * dy * dz;
  }
  double vol = (c->d.x1max - c->d.x1min) * (c->d.x2max - c->d.x2min) * (c->d.x3max - c->d.x3min);
  double ave = total / vol;
  for (int g = 0; g < c->nblocks; ++g) {
    mglevels *m = &c->mgb[g];
    double *a = type ? m->u[l] : m->src[l];
    size_t sz = (size_t)(m->n1[l] + 2) * (m->n2[l] + 2) * (m->n3[l] + 2);
    for (size_t q = 0; q < sz; ++q) a[q] -= ave;
  }
  return ave;
}

/* MultigridDriver::CalculateDefectNorm (L2), multigrid_driver.cpp:826-858 with
 * Multigrid::CalculateDefectNorm, multigrid.cpp:555-595 (current = finest) */
static double mg_defect_norm(orc_ctx *c) {
  int l = c->nmblevel - 1;
  double norm = 0.0;
  for (int g = 0; g < c->nblocks; ++g) {
    mglevels *m = &c->mgb[g];
    mg_defect(m, l);
    double s = 0.0;
    for (int k = 1; k <= m->n3[l]; ++k)
      for (int j = 1; j <= m->n2[l]; ++j)
        for (int i = 1; i <= m->n1[l]; ++i) s += SQR(m->def[l][MIDX(m, l, k, j, i)]);
    int lx2 = (g / c->nb1) % c->nb2, lx3 = g / (c->nb1 * c->nb2);
    double b2min = (lx2 == 0) ? c->d.x2min : uniform_gen(meshgen_x(lx2, c->nb2), c->d.x2min, c->d.x2max);
    double b2max = (lx2 == c->nb2 - 1) ? c->d.x2max : uniform_gen(meshgen_x(lx2 + 1, c->nb2), c->d.x2min, c->d.x2max);
    double b3min = (lx3 == 0) ? c->d.x3min : uniform_gen(meshgen_x(lx3, c->nb3), c->d.x3min, c->d.x3max);
    double b3max = (lx3 == c->nb3 - 1) ? c->d.x3max : uniform_gen(meshgen_x(lx3 + 1, c->nb3), c->d.x3min, c->d.x3max);
    double dy = (b2max - b2min) / (double)c->d.bnx2, dz = (b3max - b3min) / (double)c->d.bnx3;
    norm += s * m->rdx * dy * dz;
  }
  double vol = (c->d.x1max - c->d.x1min) * (c->d.x2max - c->d.x2min) * (c->d.x3max - c->d.x3min);
  norm /= vol;
  return sqrt(norm);
}

void orc_mg_set_mode(orc_ctx *c, int mode, double threshold, int niter) {
  c->mg_mode = mode; c->mg_eps = threshold; c->mg_niter = niter;
}
/* MultigridDriver's srcmask_ (multigrid_driver.cpp:430-436) with the collapse problem's
 * mask (pgen/collapse.cpp:66-82), centred at c; radius <= 0 switches it off. */
void orc_mg_set_source_mask_sphere(orc_ctx *c, double xc, double yc, double zc, double radius) {
  c->srcmask = radius > 0.0;
  c->srcmask_c[0] = xc; c->srcmask_c[1] = yc; c->srcmask_c[2] = zc;
  c->srcmask_r2 = radius * radius;
}
double orc_mg_last_norm(orc_ctx *c) { return c->last_norm; }
int orc_mg_last_iters(orc_ctx *c) { return c->last_iters; }
void orc_mg_multipole(orc_ctx *c, double *out28) {
  for (int q = 0; q < 3; ++q) out28[q] = c->mpo[q];
  for (int q = 0; q < 25; ++q) out28[3 + q] = c->mpc[q];
}
double orc_mp_phis(orc_ctx *c, double x, double y, double z) { return mp_phis(c, x, y, z); }

/* MGGravityDriver::Solve, gravity/mg_gravity.cpp:166-211 (periodic boundaries:
 * fsubtract_average_ = true, no FAS; multipole boundaries with FAS; optional source
 * mask). */
void orc_mg_solve(orc_ctx *c, int stage) {
  (void)stage;
  int lf = c->nmblevel - 1;
  double fpg = c->d.four_pi_G;
  /* LoadSource (multigrid.cpp:194-231) and, in MGI mode, LoadFinestData (:160-180) */
  for (int g = 0; g < c->nblocks; ++g) {
    mglevels *m = &c->mgb[g];
    const double *u = c->arr[ARR_U] + g * c->arrsize[ARR_U];
    const double *phi = c->arr[ARR_PHI] + g * c->arrsize[ARR_PHI];
    for (int k = 1; k <= m->n3[lf]; ++k)
      for (int j = 1; j <= m->n2[lf]; ++j)
        for (int i = 1; i <= m->n1[lf]; ++i) {
          int64_t hid = CIDX(c, k - 1 + c->ng, j - 1 + c->ng, i - 1 + c->ng);
          double v = u[IDN * c->ncell + hid];
          m->src[lf][MIDX(m, lf, k, j, i)] = (fpg == 1.0) ? v : v * fpg;
          if (c->mg_mode == 1) m->u[lf][MIDX(m, lf, k, j, i)] = phi[hid];
        }
    if (c->srcmask) {   /* Multigrid::ApplySourceMask (multigrid.cpp:239-246) */
      for (int k = 1; k <= m->n3[lf]; ++k) {
        double z = mgc_v(m, lf, 2, k) - c->srcmask_c[2];
        for (int j = 1; j <= m->n2[lf]; ++j) {
          double y = mgc_v(m, lf, 1, j) - c->srcmask_c[1];
          for (int i = 1; i <= m->n1[lf]; ++i) {
            double x = mgc_v(m, lf, 0, i) - c->srcmask_c[0];
            double r2 = x * x + y * y + z * z;
            if (r2 > c->srcmask_r2) m->src[lf][MIDX(m, lf, k, j, i)] = 0.0;
          }
        }
      }
    }
  }
  c->current = c->ntotal - 1;
  c->fmglevel = c->ntotal - 1;
  /* SetupMultigrid, multigrid_driver.cpp:343-460: periodic boundaries subtract the mean
   * source; multipole boundaries compute the expansion (about the centre of mass with
   * auto_mporigin) instead */
  if (!c->fas) {
    c->last_ave = mg_subtract_average(c, 0);
  } else {
    if (c->d.mp_auto_origin) mp_center_of_mass(c);
    mp_coefficients(c);
  }
  if (c->mg_mode == 0) {
    for (int g = 0; g < c->nblocks; ++g) {   /* RestrictFMGSource, multigrid.cpp:253-262 */
      mglevels *m = &c->mgb[g];
      for (int l = m->nlev - 1; l > 0; --l) mg_restrict(m, m->src[l - 1], m->src[l], l);
    }
    mg_blocks_to_root(c, 1);
    for (int l = c->mgr.nlev - 1; l > 0; --l) mg_restrict(&c->mgr, c->mgr.src[l - 1], c->mgr.src[l], l);
    c->current = 0;
    /* SolveFMGCycle, multigrid_driver.cpp:700-714 */
    for (c->fmglevel = 0; c->fmglevel < c->ntotal; c->fmglevel++) {
      mg_vcycle(c);
      if (c->fmglevel != c->ntotal - 1) mg_fmg_prolongate_step(c);
    }
    c->fmglevel = c->ntotal - 1;
    if (!c->fas) mg_subtract_average(c, 1);
  }
  int iters = 0;
  if (c->mg_eps >= 0.0) {       /* SolveIterative, multigrid_driver.cpp:721-755 */
    double def = mg_defect_norm(c);
    int n = 0;
    while (def > c->mg_eps) {
      mg_vcycle(c);
      iters++;
      double olddef = def;
      def = mg_defect_norm(c);
      if (def / olddef > 0.8) {
        if (c->mg_eps == 0.0) break;
      }
      if (n > 100) break;
      n++;
    }
    c->last_norm = def;
    if (!c->fas) mg_subtract_average(c, 1);
  } else {                       /* SolveIterativeFixedTimes, :762-771 */
    for (int n = 0; n < c->mg_niter; ++n) { mg_vcycle(c); iters++; }
    if (!c->fas) mg_subtract_average(c, 1);
    c->last_norm = mg_defect_norm(c);
  }
  c->last_iters = iters;
  /* RetrieveResult with one ghost layer (multigrid.cpp:269-288) */
  for (int g = 0; g < c->nblocks; ++g) {
    mglevels *m = &c->mgb[g];
    double *phi = c->arr[ARR_PHI] + g * c->arrsize[ARR_PHI];
    int n = m->n1[lf];
    for (int k = 0; k <= n + 1; ++k)
      for (int j = 0; j <= n + 1; ++j)
        for (int i = 0; i <= n + 1; ++i)
          phi[CIDX(c, k - 1 + c->ng, j - 1 + c->ng, i - 1 + c->ng)] = m->u[lf][MIDX(m, lf, k, j, i)];
  }
}

/* ================================================================ problem setup */
/* Adiabatic-MHD right eigenvectors in conserved variables at a uniform state, the
 * Roe-matrix eigensystem used by pgen/linear_wave.cpp:626-800 (x=0, y=1). */
static void mhd_right_eigenvectors(double gamma, double d, double v1, double v2, double v3,
                                   double h, double b1, double b2, double b3,
                                   double ev[7], double rem[7][7]) {
  double gm1 = gamma - 1.0, x = 0.0, y = 1.0;
  double vsq = v1 * v1 + v2 * v2 + v3 * v3;
  double btsq = b2 * b2 + b3 * b3;
  double bt_starsq = (gm1 - (gm1 - 1.0) * y) * btsq;
  double vaxsq = b1 * b1 / d;
  double hp = h - (vaxsq + btsq / d);
  double twid_asq = dmax((gm1 * (hp - 0.5 * vsq) - (gm1 - 1.0) * x), TINY_NUMBER);
  double ct2 = bt_starsq / d;
  double tsum = vaxsq + ct2 + twid_asq;
  double tdif = vaxsq + ct2 - twid_asq;
  double cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
  double cfsq = 0.5 * (tsum + cf2_cs2);
  double cf = sqrt(cfsq);
  double cssq = twid_asq * vaxsq / cfsq;
  double cs = sqrt(cssq);
  double bt = sqrt(btsq);
  double bet2, bet3;
  if (bt == 0.0) { bet2 = 1.0; bet3 = 0.0; }
  else { bet2 = b2 / bt; bet3 = b3 / bt; }
  double bt_star = sqrt(bt_starsq);
  double bet2_star = bet2 / sqrt(gm1 - (gm1 - 1.0) * y);
  double bet3_star = bet3 / sqrt(gm1 - (gm1 - 1.0) * y);
  double bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
  double vbet = v2 * bet2_star + v3 * bet3_star;
  double alpha_f, alpha_s;
  if ((cfsq - cssq) == 0.0) { alpha_f = 1.0; alpha_s = 0.0; }
  else if ((twid_asq - cssq) <= 0.0) { alpha_f = 0.0; alpha_s = 1.0; }
  else if ((cfsq - twid_asq) <= 0.0) { alpha_f = 1.0; alpha_s = 0.0; }
  else {
    alpha_f = sqrt((twid_asq - cssq) / (cfsq - cssq));
    alpha_s = sqrt((cfsq - twid_asq) / (cfsq - cssq));
  }
  double sqrtd = sqrt(d), isqrtd = 1.0 / sqrtd;
  double s = (b1 < 0.0) ? -1.0 : 1.0;
  double twid_a = sqrt(twid_asq);
  double qf = cf * alpha_f * s, qs = cs * alpha_s * s;
  double af_prime = twid_a * alpha_f * isqrtd, as_prime = twid_a * alpha_s * isqrtd;
  double afpbb = af_prime * bt_star * bet_starsq, aspbb = as_prime * bt_star * bet_starsq;
  double vax = sqrt(vaxsq);
  ev[0] = v1 - cf; ev[1] = v1 - vax; ev[2] = v1 - cs; ev[3] = v1;
  ev[4] = v1 + cs; ev[5] = v1 + vax; ev[6] = v1 + cf;
  rem[0][0] = alpha_f; rem[0][1] = 0.0; rem[0][2] = alpha_s; rem[0][3] = 1.0;
  rem[0][4] = alpha_s; rem[0][5] = 0.0; rem[0][6] = alpha_f;
  rem[1][0] = alpha_f * ev[0]; rem[1][1] = 0.0; rem[1][2] = alpha_s * ev[2]; rem[1][3] = v1;
  rem[1][4] = alpha_s * ev[4]; rem[1][5] = 0.0; rem[1][6] = alpha_f * ev[6];
  double qa = alpha_f * v2, qb = alpha_s * v2, qc = qs * bet2_star, qd = qf * bet2_star;
  rem[2][0] = qa + qc; rem[2][1] = -bet3; rem[2][2] = qb - qd; rem[2][3] = v2;
  rem[2][4] = qb + qd; rem[2][5] = bet3; rem[2][6] = qa - qc;
  qa = alpha_f * v3; qb = alpha_s * v3; qc = qs * bet3_star; qd = qf * bet3_star;
  rem[3][0] = qa + qc; rem[3][1] = bet2; rem[3][2] = qb - qd; rem[3][3] = v3;
  rem[3][4] = qb + qd; rem[3][5] = -bet2; rem[3][6] = qa - qc;
  rem[4][0] = alpha_f * (hp - v1 * cf) + qs * vbet + aspbb;
  rem[4][1] = -(v2 * bet3 - v3 * bet2);
  rem[4][2] = alpha_s * (hp - v1 * cs) - qf * vbet - afpbb;
  rem[4][3] = 0.5 * vsq + (gm1 - 1.0) * x / gm1;
  rem[4][4] = alpha_s * (hp + v1 * cs) + qf * vbet - afpbb;
  rem[4][5] = -rem[4][1];
  rem[4][6] = alpha_f * (hp + v1 * cf) - qs * vbet + aspbb;
  rem[5][0] = as_prime * bet2_star; rem[5][1] = -bet3 * s * isqrtd;
  rem[5][2] = -af_prime * bet2_star; rem[5][3] = 0.0; rem[5][4] = rem[5][2];
  rem[5][5] = rem[5][1]; rem[5][6] = rem[5][0];
  rem[6][0] = as_prime * bet3_star; rem[6][1] = bet2 * s * isqrtd;
  rem[6][2] = -af_prime * bet3_star; rem[6][3] = 0.0; rem[6][4] = rem[6][2];
  rem[6][5] = rem[6][1]; rem[6][6] = rem[6][0];
}

typedef struct {
  double sin_a2, cos_a2, sin_a3, cos_a3, k_par, d0, p0, u0, bx0, by0, bz0, dby, dbz;
  double rem[7][7], ev[7];
} lw_params;

/* Mesh::InitUserMeshData of pgen/linear_wave.cpp:78-190 (defaults: angles from the
 * box, ang_2/3_vert false). */
static void lw_setup(orc_ctx *c, int wave_flag, double amp, double vflow, lw_params *p) {
  double gam = c->d.gamma, gm1 = gam - 1.0;
  double x1size = c->d.x1max - c->d.x1min, x2size = c->d.x2max - c->d.x2min,
         x3size = c->d.x3max - c->d.x3min;
  double ang_3 = atan(x1size / x2size);
  p->sin_a3 = sin(ang_3); p->cos_a3 = cos(ang_3);
  double ang_2 = atan(0.5 * (x1size * p->cos_a3 + x2size * p->sin_a3) / x3size);
  p->sin_a2 = sin(ang_2); p->cos_a2 = cos(ang_2);
  double x1 = x1size * p->cos_a2 * p->cos_a3;
  double x2 = x2size * p->cos_a2 * p->sin_a3;
  double x3 = x3size * p->sin_a2;
  double lambda = x1;
  if (ang_3 != 0.0) lambda = dmin(lambda, x2);
  if (ang_2 != 0.0) lambda = dmin(lambda, x3);
  p->k_par = 2.0 * (M_PI) / lambda;
  p->d0 = 1.0; p->u0 = vflow;
  p->bx0 = 1.0; p->by0 = sqrt(2.0); p->bz0 = 0.5;
  p->p0 = 1.0 / gam;
  double h0 = ((p->p0 / gm1 + 0.5 * p->d0 * (p->u0 * p->u0)) + p->p0) / p->d0;
  h0 += (p->bx0 * p->bx0 + p->by0 * p->by0 + p->bz0 * p->bz0) / p->d0;
  mhd_right_eigenvectors(gam, p->d0, p->u0, 0.0, 0.0, h0, p->bx0, p->by0, p->bz0, p->ev, p->rem);
  p->dby = amp * p->rem[NWAVE - 2][wave_flag];
  p->dbz = amp * p->rem[NWAVE - 1][wave_flag];
}

static double lwA1(const lw_params *p, double x1, double x2, double x3) {
  double x = x1 * p->cos_a2 * p->cos_a3 + x2 * p->cos_a2 * p->sin_a3 + x3 * p->sin_a2;
  double y = -x1 * p->sin_a3 + x2 * p->cos_a3;
  double Ay = p->bz0 * x - (p->dbz / p->k_par) * cos(p->k_par * (x));
  double Az = -p->by0 * x + (p->dby / p->k_par) * cos(p->k_par * (x)) + p->bx0 * y;
  return -Ay * p->sin_a3 - Az * p->sin_a2 * p->cos_a3;
}
static double lwA2(const lw_params *p, double x1, double x2, double x3) {
  double x = x1 * p->cos_a2 * p->cos_a3 + x2 * p->cos_a2 * p->sin_a3 + x3 * p->sin_a2;
  double y = -x1 * p->sin_a3 + x2 * p->cos_a3;
  double Ay = p->bz0 * x - (p->dbz / p->k_par) * cos(p->k_par * (x));
  double Az = -p->by0 * x + (p->dby / p->k_par) * cos(p->k_par * (x)) + p->bx0 * y;
  return Ay * p->cos_a3 - Az * p->sin_a2 * p->sin_a3;
}
static double lwA3(const lw_params *p, double x1, double x2, double x3) {
  double x = x1 * p->cos_a2 * p->cos_a3 + x2 * p->cos_a2 * p->sin_a3 + x3 * p->sin_a2;
  double y = -x1 * p->sin_a3 + x2 * p->cos_a3;
  double Az = -p->by0 * x + (p->dby / p->k_par) * cos(p->k_par * (x)) + p->bx0 * y;
  return Az * p->cos_a2;
}

/* MeshBlock::ProblemGenerator, pgen/linear_wave.cpp:431-597 (uniform mesh: the
 * vector potential is sampled at edge centres). */
void orc_pgen_linear_wave(orc_ctx *c, int wave_flag, double amp, double vflow) {
  lw_params p;
  lw_setup(c, wave_flag, amp, vflow, &p);
  double gm1 = c->d.gamma - 1.0;
  for (int g = 0; g < c->nblocks; ++g) {
    const double *x1f = c->x1f + g * (c->n1 + 1), *x1v = c->x1v + g * (c->n1 + 1), *dx1 = c->dx1f + g * (c->n1 + 1);
    const double *x2f = c->x2f + g * (c->n2 + 1), *x2v = c->x2v + g * (c->n2 + 1), *dx2 = c->dx2f + g * (c->n2 + 1);
    const double *x3f = c->x3f + g * (c->n3 + 1), *x3v = c->x3v + g * (c->n3 + 1), *dx3 = c->dx3f + g * (c->n3 + 1);
    int64_t na = (int64_t)(c->n1 + 1) * (c->n2 + 1) * (c->n3 + 1);
    double *a1 = (double *)xcalloc(na * sizeof(double)), *a2 = (double *)xcalloc(na * sizeof(double)),
           *a3 = (double *)xcalloc(na * sizeof(double));
#define AI(k, j, i) ((((int64_t)(k)) * (c->n2 + 1) + (j)) * (c->n1 + 1) + (i))
    for (int k = c->ks; k <= c->ke + 1; k++)
      for (int j = c->js; j <= c->je + 1; j++)
        for (int i = c->is; i <= c->ie + 1; i++) {
          if (i != c->ie + 1) a1[AI(k, j, i)] = lwA1(&p, x1v[i], x2f[j], x3f[k]);
          if (j != c->je + 1) a2[AI(k, j, i)] = lwA2(&p, x1f[i], x2v[j], x3f[k]);
          if (k != c->ke + 1) a3[AI(k, j, i)] = lwA3(&p, x1f[i], x2f[j], x3v[k]);
        }
    double *b1 = c->arr[ARR_B1F] + g * c->nf1, *b2 = c->arr[ARR_B2F] + g * c->nf2,
           *b3 = c->arr[ARR_B3F] + g * c->nf3;
    for (int k = c->ks; k <= c->ke; k++)
      for (int j = c->js; j <= c->je; j++)
        for (int i = c->is; i <= c->ie + 1; i++)
          b1[F1IDX(c, k, j, i)] = (a3[AI(k, j + 1, i)] - a3[AI(k, j, i)]) / dx2[j] -
                                  (a2[AI(k + 1, j, i)] - a2[AI(k, j, i)]) / dx3[k];
    for (int k = c->ks; k <= c->ke; k++)
      for (int j = c->js; j <= c->je + 1; j++)
        for (int i = c->is; i <= c->ie; i++)
          b2[F2IDX(c, k, j, i)] = (a1[AI(k + 1, j, i)] - a1[AI(k, j, i)]) / dx3[k] -
                                  (a3[AI(k, j, i + 1)] - a3[AI(k, j, i)]) / dx1[i];
    for (int k = c->ks; k <= c->ke + 1; k++)
      for (int j = c->js; j <= c->je; j++)
        for (int i = c->is; i <= c->ie; i++)
          b3[F3IDX(c, k, j, i)] = (a2[AI(k, j, i + 1)] - a2[AI(k, j, i)]) / dx1[i] -
                                  (a1[AI(k, j + 1, i)] - a1[AI(k, j, i)]) / dx2[j];
#undef AI
    free(a1); free(a2); free(a3);
    double *u = c->arr[ARR_U] + g * c->arrsize[ARR_U];
    int64_t nc = c->ncell;
    for (int k = c->ks; k <= c->ke; k++)
      for (int j = c->js; j <= c->je; j++)
        for (int i = c->is; i <= c->ie; i++) {
          int64_t id = CIDX(c, k, j, i);
          double x = p.cos_a2 * (x1v[i] * p.cos_a3 + x2v[j] * p.sin_a3) + x3v[k] * p.sin_a2;
          double sn = sin(p.k_par * x);
          u[IDN * nc + id] = p.d0 + amp * sn * p.rem[0][wave_flag];
          double mx = p.d0 * vflow + amp * sn * p.rem[1][wave_flag];
          double my = amp * sn * p.rem[2][wave_flag];
          double mz = amp * sn * p.rem[3][wave_flag];
          u[IVX * nc + id] = mx * p.cos_a2 * p.cos_a3 - my * p.sin_a3 - mz * p.sin_a2 * p.cos_a3;
          u[IVY * nc + id] = mx * p.cos_a2 * p.sin_a3 + my * p.cos_a3 - mz * p.sin_a2 * p.sin_a3;
          u[IVZ * nc + id] = mx * p.sin_a2 + mz * p.cos_a2;
          u[IEN * nc + id] = p.p0 / gm1 + 0.5 * p.d0 * p.u0 * p.u0 + amp * sn * p.rem[4][wave_flag];
          u[IEN * nc + id] += 0.5 * (p.bx0 * p.bx0 + p.by0 * p.by0 + p.bz0 * p.bz0);
        }
  }
}

/* Mesh::UserWorkAfterLoop of pgen/linear_wave.cpp:199-425 (correct_err = false).
 * out[0..7] = L1 errors (d, M1, M2, M3, E, B1c, B2c, B3c), out[8] = RMS-L1,
 * out[9..16] = max errors, out[17] = largest max/L1. */
void orc_linear_wave_errors(orc_ctx *c, int wave_flag, double amp, double vflow, double *out) {
  lw_params p;
  lw_setup(c, wave_flag, amp, vflow, &p);
  double gm1 = c->d.gamma - 1.0;
  double l1[8] = {0}, mx_[8] = {0};
  for (int g = 0; g < c->nblocks; ++g) {
    const double *x1v = c->x1v + g * (c->n1 + 1), *dx1 = c->dx1f + g * (c->n1 + 1);
    const double *x2v = c->x2v + g * (c->n2 + 1), *dx2 = c->dx2f + g * (c->n2 + 1);
    const double *x3v = c->x3v + g * (c->n3 + 1), *dx3 = c->dx3f + g * (c->n3 + 1);
    const double *u = c->arr[ARR_U] + g * c->arrsize[ARR_U];
    const double *bcc = c->arr[ARR_BCC] + g * c->arrsize[ARR_BCC];
    int64_t nc = c->ncell;
    for (int k = c->ks; k <= c->ke; k++)
      for (int j = c->js; j <= c->je; j++)
        for (int i = c->is; i <= c->ie; i++) {
          int64_t id = CIDX(c, k, j, i);
          double x = p.cos_a2 * (x1v[i] * p.cos_a3 + x2v[j] * p.sin_a3) + x3v[k] * p.sin_a2;
          double sn = sin(p.k_par * x);
          double ref[8];
          ref[0] = p.d0 + amp * sn * p.rem[0][wave_flag];
          double mx = p.d0 * vflow + amp * sn * p.rem[1][wave_flag];
          double my = amp * sn * p.rem[2][wave_flag];
          double mz = amp * sn * p.rem[3][wave_flag];
          ref[1] = mx * p.cos_a2 * p.cos_a3 - my * p.sin_a3 - mz * p.sin_a2 * p.cos_a3;
          ref[2] = mx * p.cos_a2 * p.sin_a3 + my * p.cos_a3 - mz * p.sin_a2 * p.sin_a3;
          ref[3] = mx * p.sin_a2 + mz * p.cos_a2;
          double e0 = p.p0 / gm1 + 0.5 * p.d0 * p.u0 * p.u0 + amp * sn * p.rem[4][wave_flag];
          e0 += 0.5 * (p.bx0 * p.bx0 + p.by0 * p.by0 + p.bz0 * p.bz0);
          double bx = p.bx0, by = p.by0 + amp * sn * p.rem[5][wave_flag], bz = p.bz0 + amp * sn * p.rem[6][wave_flag];
          ref[5] = bx * p.cos_a2 * p.cos_a3 - by * p.sin_a3 - bz * p.sin_a2 * p.cos_a3;
          ref[6] = bx * p.cos_a2 * p.sin_a3 + by * p.cos_a3 - bz * p.sin_a2 * p.sin_a3;
          ref[7] = bx * p.sin_a2 + bz * p.cos_a2;
          ref[4] = e0;
          double vol = dx1[i] * dx2[j] * dx3[k];
          for (int n = 0; n < 5; ++n) {
            double e = fabs(ref[n] - u[n * nc + id]);
            l1[n] += e * vol;
            mx_[n] = dmax(e, mx_[n]);
          }
          for (int n = 0; n < 3; ++n) {
            double e = fabs(ref[5 + n] - bcc[n * nc + id]);
            l1[5 + n] += e * vol;
            mx_[5 + n] = dmax(e, mx_[5 + n]);
          }
        }
  }
  double vol = (c->d.x1max - c->d.x1min) * (c->d.x2max - c->d.x2min) * (c->d.x3max - c->d.x3min);
  double rms = 0.0, mm = 0.0;
  for (int n = 0; n < 8; ++n) l1[n] /= vol;
  for (int n = 0; n < 8; ++n) {
    rms += SQR(l1[n]);
    mm = dmax(mm, mx_[n] / l1[n]);
  }
  for (int n = 0; n < 8; ++n) { out[n] = l1[n]; out[9 + n] = mx_[n]; }
  out[8] = sqrt(rms);
  out[17] = mm;
}

/* max |div B| * dx over interior cells (the reference has no such diagnostic; this
 * is the discrete divergence of the face field, normalised by the cell width). */
double orc_max_divb(orc_ctx *c) {
  double m = 0.0;
  for (int g = 0; g < c->nblocks; ++g) {
    const double *b1 = c->arr[ARR_B1F] + g * c->nf1, *b2 = c->arr[ARR_B2F] + g * c->nf2,
                 *b3 = c->arr[ARR_B3F] + g * c->nf3;
    const double *dx1 = c->dx1f + g * (c->n1 + 1), *dx2 = c->dx2f + g * (c->n2 + 1),
                 *dx3 = c->dx3f + g * (c->n3 + 1);
    for (int k = c->ks; k <= c->ke; k++)
      for (int j = c->js; j <= c->je; j++)
        for (int i = c->is; i <= c->ie; i++) {
          double dv = (b1[F1IDX(c, k, j, i + 1)] - b1[F1IDX(c, k, j, i)]) / dx1[i] +
                      (b2[F2IDX(c, k, j + 1, i)] - b2[F2IDX(c, k, j, i)]) / dx2[j] +
                      (b3[F3IDX(c, k + 1, j, i)] - b3[F3IDX(c, k, j, i)]) / dx3[k];
          m = dmax(m, fabs(dv));
        }
  }
  return m;
}

/* volume sums of the 5 conserved variables and of the 3 face-field components */
void orc_conserved_totals(orc_ctx *c, double *out8) {
  for (int n = 0; n < 8; ++n) out8[n] = 0.0;
  for (int g = 0; g < c->nblocks; ++g) {
    const double *u = c->arr[ARR_U] + g * c->arrsize[ARR_U];
    const double *b1 = c->arr[ARR_B1F] + g * c->nf1, *b2 = c->arr[ARR_B2F] + g * c->nf2,
                 *b3 = c->arr[ARR_B3F] + g * c->nf3;
    for (int k = c->ks; k <= c->ke; k++)
      for (int j = c->js; j <= c->je; j++)
        for (int i = c->is; i <= c->ie; i++) {
          for (int n = 0; n < 5; ++n) out8[n] += u[n * c->ncell + CIDX(c, k, j, i)];
          out8[5] += b1[F1IDX(c, k, j, i)];
          out8[6] += b2[F2IDX(c, k, j, i)];
          out8[7] += b3[F3IDX(c, k, j, i)];
        }
  }
}

/* HistoryOutput::WriteOutputFile, the built-in sums of the non-orbital branch
 * (outputs/history.cpp:37-39,135-200): volume sums over the interior cells of every
 * MeshBlock, block by block in k, j, i order, of mass, the three momenta, the three
 * partitioned kinetic energies, the total energy (adiabatic only), the gravitational
 * energy 0.5*rho*phi (with self-gravity) and the three cell-centred magnetic energies.
 * out gets NHYDRO + 3 + (self-gravity) + 3 values in that order; returns the count. */
int orc_history(orc_ctx *c, double *out) {
  const int nh = c->nhydro, sg = c->d.self_gravity ? 1 : 0;
  const int nout = nh + 3 + sg + 3;
  for (int n = 0; n < nout; ++n) out[n] = 0.0;
  for (int g = 0; g < c->nblocks; ++g) {
    const double *u = c->arr[ARR_U] + g * c->arrsize[ARR_U];
    const double *bcc = c->arr[ARR_BCC] + g * c->arrsize[ARR_BCC];
    const double *phi = c->arr[ARR_PHI] + g * c->arrsize[ARR_PHI];
    for (int k = c->ks; k <= c->ke; k++)
      for (int j = c->js; j <= c->je; j++)
        for (int i = c->is; i <= c->ie; i++) {
          /* Coordinates::CellVolume (coordinates.cpp:521-531): dx1f*dx2f*dx3f */
          const double vol = c->dx1f[g * (c->n1 + 1) + i] * c->dx2f[g * (c->n2 + 1) + j] *
                             c->dx3f[g * (c->n3 + 1) + k];
          const int64_t id = CIDX(c, k, j, i);
          const double u_d = u[IDN * c->ncell + id], u_mx = u[IVX * c->ncell + id];
          const double u_my = u[IVY * c->ncell + id], u_mz = u[IVZ * c->ncell + id];
          out[0] += vol * u_d;
          out[1] += vol * u_mx;
          out[2] += vol * u_my;
          out[3] += vol * u_mz;
          out[4] += vol * 0.5 * (u_mx * u_mx) / u_d;
          out[5] += vol * 0.5 * (u_my * u_my) / u_d;
          out[6] += vol * 0.5 * (u_mz * u_mz) / u_d;
          if (!c->iso) out[7] += vol * u[IEN * c->ncell + id];
          if (sg) out[nh + 3] += vol * 0.5 * u_d * phi[id];
          const double b1 = bcc[IB1 * c->ncell + id], b2 = bcc[IB2 * c->ncell + id];
          const double b3 = bcc[IB3 * c->ncell + id];
          out[nh + 3 + sg] += vol * 0.5 * b1 * b1;
          out[nh + 3 + sg + 1] += vol * 0.5 * b2 * b2;
          out[nh + 3 + sg + 2] += vol * 0.5 * b3 * b3;
        }
  }
  return nout;
}

/* ================================================================ unit-level probes */
/* One Riemann-solver call in the solver's local frame (wave arrays ordered IDN, IVX
 * (normal), IVY, IVZ, IPR, IBY, IBZ): rs 0 = HLLD, 1 = HLLE.  flx[7] as flxi. */
void orc_riemann(int rs, double gamma, const double *wl, const double *wr, double bx,
                 double *flx) {
  if (rs == 1) hlle(gamma, 0, 0.0, wl, wr, bx, flx);
  else if (rs == 2) roe(gamma, wl, wr, bx, flx);
  else if (rs == 3) lhlld(gamma, wl, wr, bx, 0.0, 0.0, flx);
  else hlld(gamma, wl, wr, bx, flx);
}

/* LHLLD at one interface with the shock detector's velocity differences given */
void orc_riemann_lhlld(double gamma, const double *wl, const double *wr, double bx, double dvn,
                       double dvt, double *flx) {
  lhlld(gamma, wl, wr, bx, dvn, dvt, flx);
}

/* isothermal solvers at one interface: rs 0 HLLD-iso, 1 HLLE (isothermal branch) */
void orc_riemann_iso(int rs, double cs, double dfloor, const double *wl, const double *wr,
                     double bx, double *flx) {
  if (rs == 1) hlle(0.0, 1, cs, wl, wr, bx, flx);
  else hlld_iso(cs, dfloor, wl, wr, bx, flx);
}

/* PLM on a 1-D pencil q[0..n-1] with face-offset factors dxp/dxm (plm.cpp:69-77,
 * 113-119): wl[i+1] (upper face of cell i) and wr[i] (lower face), i = 1..n-2. */
void orc_plm_pencil(int n, const double *q, const double *dxp, const double *dxm,
                    double *wl, double *wr) {
  for (int i = 1; i < n - 1; ++i) {
    double dwl = (q[i] - q[i - 1]);
    double dwr = (q[i + 1] - q[i]);
    double dw2 = dwl * dwr;
    double dwm = 2.0 * dw2 / (dwl + dwr);
    if (dw2 <= 0.0) dwm = 0.0;
    wl[i + 1] = q[i] + dxp[i] * dwm;
    wr[i] = q[i] - dxm[i] * dwm;
  }
}

/* The eigenvector projections at one state (w7: IDN, IVX, IVY, IVZ, IPR, IBY, IBZ in the
 * mesh frame with (IBY, IBZ) the field transverse to dir; b1 the normal field): L[r][n] =
 * characteristic field r of the unit primitive vector n, R[n][r] = primitive slot n of
 * the unit characteristic vector r (7 x 7, row-major; isothermal fills 6 fields). */
void orc_char_matrices(int iso, double gamma, double iso_cs, int dir, const double *w7, double b1,
                       double *L, double *R) {
  orc_ctx c;
  memset(&c, 0, sizeof(c));
  c.iso = iso;
  c.d.gamma = gamma;
  c.d.iso_cs = iso_cs;
  orc_eigen e;
  char_eigen(&c, b1, w7, &e);
  for (int n = 0; n < NWAVE * NWAVE; ++n) L[n] = R[n] = 0.0;
  for (int n = 0; n < NWAVE; ++n) {
    double v[NWAVE] = {0}, ch[NWAVE] = {0};
    v[n] = 1.0;
    char_left(&c, &e, IVX + dir, v, ch);
    for (int r = 0; r < NWAVE; ++r) L[r * NWAVE + n] = ch[r];
  }
  for (int r = 0; r < char_nwave(&c); ++r) {
    double ch[NWAVE] = {0}, v[NWAVE];
    ch[r] = 1.0;
    char_right(&c, &e, IVX + dir, ch, v);
    for (int n = 0; n < NWAVE; ++n) R[n * NWAVE + r] = v[n];
  }
}

/* The reconstructed states of one cell of block g along dir (the 7-slot mesh-frame order):
 * wm at its lower face, wp at its upper face, as the flux calculation uses them. */
void orc_recon_cell(const orc_ctx *c, int dir, int order, int g, int k, int j, int i, double *wm,
                    double *wp) {
  const double *w = c->arr[ARR_W] + g * c->arrsize[ARR_W];
  const double *bcc = c->arr[ARR_BCC] + g * c->arrsize[ARR_BCC];
  recon_cell(c, w, bcc, dir, order, g, k, j, i, wm, wp);
}

/* PPM on a 1-D pencil: wl[i+1] (upper face of cell i), wr[i], i = 2..n-3. */
void orc_ppm_pencil(int n, const double *q, double *wl, double *wr) {
  for (int i = 2; i < n - 2; ++i) {
    double m, p;
    ppm_cell(q + i - 2, &m, &p);
    wl[i + 1] = p;
    wr[i] = m;
  }
}
