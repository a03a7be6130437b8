/*
 * agx_oracle.h -- CPU FP64 restatement of the reference's per-MeshBlock MHD update
 * and multigrid self-gravity solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
 * (athena-gamma_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * reference's src/).  Array layouts are identical to the device layout of the product
 * library so that host copies compare element by element:
 *   cell arrays   [block][var][k][j][i]        n_d = bnx_d + 2*nghost
 *   x1 faces      [block][k][j][i]  i in 0..n1 (n1+1 entries)
 *   x2 faces      [block][k][j][i]  j in 0..n2
 *   x3 faces      [block][k][j][i]  k in 0..n3
 *   edges x1e     [block][k 0..n3][j 0..n2][i 0..n1-1]
 *   edges x2e     [block][k 0..n3][j 0..n2-1][i 0..n1]
 *   edges x3e     [block][k 0..n3-1][j 0..n2][i 0..n1]
 * Blocks are numbered gid = (lx3*nb2 + lx2)*nb1 + lx1 (uniform root-level lattice).
 */
#ifndef AGX_ORACLE_H_
#define AGX_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and types as agx_mesh_desc in include/athena_gamma_amd.h. */
typedef struct orc_desc {
  int32_t nx1, nx2, nx3;        /* mesh cells */
  int32_t bnx1, bnx2, bnx3;     /* MeshBlock cells */
  int32_t nghost;
  int32_t xorder;               /* 1 donor cell, 2 PLM, 3 PPM */
  int32_t rsolver;              /* 0 HLLD, 1 HLLE, 2 Roe, 3 LHLLD */
  int32_t eos;                  /* 0 adiabatic */
  int32_t self_gravity;         /* 0 off, 1 multigrid */
  int32_t mg_mode;              /* 0 FMG (+iterate), 1 MGI */
  int32_t mg_niter;             /* fixed V-cycles when mg_threshold < 0 */
  int32_t precision;            /* 0 FP64 (oracle ignores) */
  double x1min, x1max, x2min, x2max, x3min, x3max;
  double gamma, iso_cs, cfl, dfloor, pfloor;
  double four_pi_G, mg_threshold;
  int32_t mg_bc;                /* 0 periodic, 1 multipole on all six faces */
  int32_t mporder;              /* 2 or 4 (multipole) */
  int32_t mp_auto_origin;       /* 1: expansion about the centre of mass */
  double mp_origin[3];          /* expansion origin when mp_auto_origin = 0 */
  int32_t mp_nodipole;          /* nodipole: dipole moments left out */
  int32_t char_proj;            /* xorder "2c"/"3c": characteristic projection */
} orc_desc;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_desc *d);
void orc_destroy(orc_ctx *c);

/* sizes (per block, in doubles) */
int64_t orc_size(orc_ctx *c, int what);   /* 0 cell, 1 x1f, 2 x2f, 3 x3f, 4 x1e, 5 x2e, 6 x3e, 7 nblocks */

/* raw array access: which = see AGX_ARR_* numbering in the product header */
double *orc_array(orc_ctx *c, int which);
void orc_set_time(orc_ctx *c, double time, double dt, int ncycle);
void orc_get_time(orc_ctx *c, double *out3);

/* coordinates of block gid: fills x1f (n1+1), x1v (n1), etc. dir 0..2 */
void orc_coords(orc_ctx *c, int gid, int dir, double *xf, double *xv, double *dxf);
/* per-block coordinates of a static-refinement tree (mesh.cpp:1668-1730) */
void orc_set_levels(orc_ctx *c, const int *nrb, const double *box, const int *level, const int64_t *lx);

/* reference steps, all blocks */
void orc_bvals_cc(orc_ctx *c, int which_arr);           /* bvals_cc.cpp:201-336 */
void orc_bvals_fc(orc_ctx *c, int b_register);           /* bvals_fc.cpp:344-402,583-... */
void orc_cons2prim(orc_ctx *c);                          /* adiabatic_mhd.cpp:41-90 */
void orc_calculate_fluxes(orc_ctx *c, int order);        /* calculate_fluxes.cpp:36-378 */
void orc_corner_e(orc_ctx *c);                           /* calculate_corner_e.cpp:28-236 */
void orc_integrate(orc_ctx *c, int stage);               /* time_integrator.cpp:1563-1660 */
double orc_new_block_dt(orc_ctx *c);                     /* new_blockdt.cpp:42-190 */
void orc_new_time_step(orc_ctx *c, double tlim);         /* mesh.cpp:1078-1119 */
void orc_initialize(orc_ctx *c);                         /* mesh.cpp:1367-1660 (subset) */
void orc_initialize_restarted(orc_ctx *c);               /* the same after a restart file */
void orc_cycle(orc_ctx *c, double tlim);                 /* main.cpp:430-489 */

/* multigrid gravity */
void orc_mg_solve(orc_ctx *c, int stage);                /* mg_gravity.cpp:166-211 */
double orc_mg_last_norm(orc_ctx *c);
int orc_mg_last_iters(orc_ctx *c);
/* multipole boundaries: out[0..2] = expansion origin, out[3..27] = scaled coefficients of
 * the last solve; the boundary potential phi_s at (x, y, z) relative to that origin */
void orc_mg_multipole(orc_ctx *c, double *out28);
double orc_mp_phis(orc_ctx *c, double x, double y, double z);
void orc_mg_set_mode(orc_ctx *c, int mode, double threshold, int niter);
void orc_mg_set_source_mask_sphere(orc_ctx *c, double xc, double yc, double zc, double radius);

/* problem generators / diagnostics (pgen restatements) */
void orc_pgen_linear_wave(orc_ctx *c, int wave_flag, double amp, double vflow);
void orc_linear_wave_errors(orc_ctx *c, int wave_flag, double amp, double vflow,
                            double *out);                /* linear_wave.cpp:199-425 */
double orc_max_divb(orc_ctx *c);

/* unit-level probes */
void orc_riemann(int rs, double gamma, const double *wl, const double *wr, double bx,
                 double *flx);
void orc_riemann_lhlld(double gamma, const double *wl, const double *wr, double bx, double dvn,
                       double dvt, double *flx);      /* lhlld.cpp:38-389 */
void orc_riemann_iso(int rs, double cs, double dfloor, const double *wl, const double *wr,
                     double bx, double *flx);
void orc_plm_pencil(int n, const double *q, const double *dxp, const double *dxm,
                    double *wl, double *wr);
void orc_ppm_pencil(int n, const double *q, double *wl, double *wr);
void orc_set_noise(orc_ctx *c, double eps, uint64_t seed);
void orc_char_matrices(int iso, double gamma, double iso_cs, int dir, const double *w7, double b1,
                       double *L, double *R);
void orc_recon_cell(const orc_ctx *c, int dir, int order, int g, int k, int j, int i, double *wm,
                    double *wp);
void orc_conserved_totals(orc_ctx *c, double *out8);
/* history.cpp:135-200: built-in volume sums; returns the count written to out */
int orc_history(orc_ctx *c, double *out);

/* static-mesh-refinement operators (agx_oracle_refine.c), 3D Cartesian blocks in the
 * dense layouts described there; items are processed in list order */
typedef struct {
  int32_t fb, cb;                   /* fine-side and coarse-side block */
  int32_t si, ei, sj, ej, sk, ek;   /* coarse index range (inclusive) */
  int32_t fid, fi1, fi2, aux;       /* face and quadrant (flux/EMF correction) */
} orc_ref_item;
typedef struct {
  int32_t nx1, nx2, nx3;            /* interior cells of a block */
  int32_t ng, cng;                  /* ghost width of the blocks and of the coarse buffers */
  int32_t pad;
} orc_ref_geom;
typedef struct {
  const double *x1v, *x2v, *x3v, *dx1f, *dx2f, *dx3f;   /* per block, length n */
} orc_ref_coords;
void orc_ref_restrict_cc(const orc_ref_geom *g, int nvar, const double *fine, orc_ref_coords fc,
                         double *coarse, const orc_ref_item *it, int nit);
void orc_ref_restrict_fc(const orc_ref_geom *g, int dir, const double *fine, orc_ref_coords fc,
                         double *coarse, const orc_ref_item *it, int nit);
void orc_ref_prolong_cc(const orc_ref_geom *g, int nvar, const double *coarse, orc_ref_coords cc,
                        double *fine, orc_ref_coords fc, const orc_ref_item *it, int nit);
void orc_ref_prolong_fc_shared(const orc_ref_geom *g, int dir, const double *coarse,
                               orc_ref_coords cc, double *fine, orc_ref_coords fc,
                               const orc_ref_item *it, int nit);
void orc_ref_prolong_fc_internal(const orc_ref_geom *g, double *b1, double *b2, double *b3,
                                 orc_ref_coords fc, const orc_ref_item *it, int nit);
void orc_ref_flux_from_finer(const orc_ref_geom *g, int nvar, const double *ff1, const double *ff2,
                             const double *ff3, orc_ref_coords fc, double *cf1, double *cf2,
                             double *cf3, const orc_ref_item *it, int nit);
void orc_ref_emf_to_coarser(const orc_ref_geom *g, const double *e1, const double *e2,
                            const double *e3, orc_ref_coords fc, double *buf, int64_t stride,
                            const orc_ref_item *it, int nit);
void orc_ref_copy_cc(const orc_ref_geom *g, int nvar, const double *src, int src_coarse, double *dst,
                     int dst_coarse, const orc_ref_item *it, int nit);
void orc_ref_cons_to_prim(const orc_ref_geom *g, double *u, const double *b1, const double *b2,
                          const double *b3, double *w, double *bcc, double gamma, double dfl,
                          double pfl, int iso, const orc_ref_item *it, int nit);
void orc_ref_pack_fc(const orc_ref_geom *g, double *f1, double *f2, double *f3, int coarse, double *buf,
                     int64_t stride, const orc_ref_item *it, int nit);
void orc_ref_unpack_fc(const orc_ref_geom *g, double *f1, double *f2, double *f3, int coarse, double *buf,
                       int64_t stride, const orc_ref_item *it, int nit);
void orc_ref_emf_same_level(const orc_ref_geom *g, const double *e1, const double *e2,
                            const double *e3, double *buf, int64_t stride, const orc_ref_item *it,
                            int nit);
void orc_ref_emf_set(const orc_ref_geom *g, double *e1, double *e2, double *e3, const double *buf,
                     int64_t stride, int finer, const orc_ref_item *it, int nit);
void orc_ref_emf_clear(const orc_ref_geom *g, double *e1, double *e2, double *e3,
                       const orc_ref_item *it, int nit);
void orc_ref_emf_average(const orc_ref_geom *g, double *e1, double *e2, double *e3,
                         const orc_ref_item *it, int nit);
void orc_ref_count_fine_edges(const int *nblevel, int *edge_flag, int *nedge_fine);
void orc_ref_prim_to_cons(const orc_ref_geom *g, const double *w, const double *b1, const double *b2,
                          const double *b3, double *bcc, double *u, double gamma, int iso,
                          const orc_ref_item *it, int nit);
int orc_ref_prolong_ranges(const orc_ref_geom *g, const int *ox, const int *lx_odd, const int *nblevel,
                           int *out);

#ifdef __cplusplus
}
#endif
#endif
