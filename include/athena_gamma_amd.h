/*
 * athena_gamma_amd.h -- C ABI of the MI355X (gfx950) MHD + multigrid-gravity library.
 *
 * One agx_mesh is one device-resident Mesh: every MeshBlock of a uniform periodic
 * root-level lattice, with all registered arrays (conserved/primitive variables, face
 * fields, fluxes, EMFs, potential, multigrid hierarchy) in HBM.  Each entry point runs
 * one task of the reference's task lists for ALL MeshBlocks in one batch of kernel
 * launches on the mesh's HIP stream, and replaces the reference method cited beside
 * it (paths relative to the reference's src/).  Device memory is allocated only by
 * agx_mesh_create, agx_ref_list_create and agx_device_alloc (and the HIP runtime's own
 * objects by agx_stream_create and agx_graph_end).
 *
 * Conventions
 *   - Return value: 0 on success; > 0 a HIP error code; < 0 a contract violation
 *     (bad argument, unsupported option).  agx_last_error() returns the message of the
 *     last failure on the calling thread.  This mirrors ATHENA_ERROR (defs.hpp.in:
 *     119-123): a host shim turns a non-zero return into std::runtime_error.
 *   - Host arrays are FP64, block-major, in the reference's AthenaArray order
 *     (n,k,j,i), i fastest, including ghost zones:
 *        cell  [nblocks][nvar][n3][n2][n1]           n_d = bnx_d + 2*nghost
 *        x1f   [nblocks][n3][n2][n1+1]   x2f [nblocks][n3][n2+1][n1]
 *        x3f   [nblocks][n3+1][n2][n1]
 *        x1e   [nblocks][n3+1][n2+1][n1] x2e [nblocks][n3+1][n2][n1+1]
 *        x3e   [nblocks][n3][n2+1][n1+1]
 *     Block index = (lx3*nb2 + lx2)*nb1 + lx1 (lattice order; the reference numbers
 *     MeshBlocks (gid) in the Z order of its block tree, and agx_block_gids gives the
 *     lattice index of each gid).  Device storage is padded (row pitch a multiple of
 *     16 doubles); agx_upload/agx_download convert.
 *   - Not re-entrant per mesh: call a given agx_mesh from one host thread.  Different
 *     meshes are independent.
 */
#ifndef ATHENA_GAMMA_AMD_H_
#define ATHENA_GAMMA_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime parameters of one Mesh: the <mesh>, <meshblock>, <time>, <hydro> and
 * <gravity> input blocks the path reads (mesh.cpp:63-300, eos/adiabatic_mhd.cpp:26-31,
 * reconstruct/reconstruction.cpp:39-100, gravity/mg_gravity.cpp:41-120). */
typedef struct agx_mesh_desc {
  int32_t nx1, nx2, nx3;        /* <mesh>/nx1..3 */
  int32_t bnx1, bnx2, bnx3;     /* <meshblock>/nx1..3 */
  int32_t nghost;               /* NGHOST (configure --nghost): >= 2, >= 3 for PPM */
  int32_t xorder;               /* <time>/xorder: 1 donor cell, 2 PLM, 3 PPM */
  int32_t rsolver;              /* configure --flux: AGX_RSOLVER_* */
  int32_t eos;                  /* configure --eos: 0 adiabatic, 1 isothermal (NHYDRO = 4:
                                   u, u1, w and the fluxes carry 4 variables on the host
                                   side; HLLD means hlld_iso.cpp; Roe is rejected) */
  int32_t self_gravity;         /* configure --grav: 0 none, 1 multigrid */
  int32_t mg_mode;              /* <gravity>/mgmode: 0 FMG, 1 MGI */
  int32_t mg_niter;             /* <gravity>/niteration (used when threshold < 0) */
  int32_t precision;            /* AGX_PREC_FP64 or AGX_PREC_MIXED (see below) */
  double x1min, x1max, x2min, x2max, x3min, x3max;   /* <mesh>/x?min, x?max */
  double gamma;                 /* <hydro>/gamma */
  double iso_cs;                /* <hydro>/iso_sound_speed (isothermal EOS) */
  double cfl;                   /* <time>/cfl_number (clamped to 0.5 in 3D) */
  double dfloor, pfloor;        /* <hydro>/dfloor, pfloor */
  double four_pi_G;             /* Mesh::SetFourPiG */
  double mg_threshold;          /* <gravity>/threshold (absolute L2 defect norm) */
  int32_t mg_bc;                /* <gravity>/ix1_bc..ox3_bc: 0 periodic (all faces), 1 multipole
                                   (all faces; mgbval_multipole.cpp, with FAS as
                                   multigrid_driver.cpp:292-295 requires) */
  int32_t mporder;              /* <gravity>/mporder: 2 or 4 (multipole only) */
  int32_t mp_auto_origin;       /* <gravity>/auto_mporigin: 1 = expand about the centre of mass */
  double mp_origin[3];          /* <gravity>/mporigin_x1..x3 (auto_mporigin = 0) */
  int32_t mp_nodipole;          /* <gravity>/nodipole: 1 = leave out the dipole moments
                                   (multigrid.cpp:1072, 1200); needs mp_auto_origin = 0
                                   (mg_gravity.cpp:106-113) */
  int32_t char_proj;            /* <time>/xorder = "2c" or "3c": 1 = limit PLM slopes or PPM
                                   values in the characteristic variables of the cell
                                   (reconstruction.cpp:53-66; characteristic.cpp); needs
                                   xorder 2 or 3 */
} agx_mesh_desc;

/* Riemann solvers (configure --flux): hlld.cpp, hlle_mhd.cpp, roe_mhd.cpp, and lhlld.cpp
 * (low-dissipation HLLD with the shock detector of calculate_velocity_differences.cpp;
 * adiabatic only, configure.py:321-322) */
enum { AGX_RSOLVER_HLLD = 0, AGX_RSOLVER_HLLE = 1, AGX_RSOLVER_ROE = 2, AGX_RSOLVER_LHLLD = 3 };
/* Precision forms.  FP64: every array and every operation in FP64, results identical to
 * the reference's FP64 arithmetic.  MIXED: reconstruction and the Riemann solver run in
 * FP32 from FP32 copies of w and bcc; fluxes, face EMFs and CT weights are stored in FP32;
 * the conserved-variable update, self-gravity source, C2P with floors, corner EMFs, CT
 * and the face field stay FP64 (div B at FP64 round-off).  The FP32 arrays (AGX_ARR_W,
 * AGX_ARR_BCC, AGX_ARR_FX?, AGX_ARR_E?X?, AGX_ARR_WX?) are converted by agx_upload /
 * agx_download, so the host side stays FP64.  In MIXED, agx_compute_corner_e and
 * agx_integrate (task-level FP64 steps) return an error; agx_cycle runs the fused form. */
enum { AGX_PREC_FP64 = 0, AGX_PREC_MIXED = 1 };

/* Registered arrays (MeshBlock members in the reference). */
enum {
  AGX_ARR_U = 0,      /* Hydro::u (NHYDRO=5)            hydro/hydro.cpp:32 */
  AGX_ARR_U1,         /* Hydro::u1 register             hydro/hydro.cpp:34 */
  AGX_ARR_W,          /* Hydro::w (primitive)           hydro/hydro.cpp:33 */
  AGX_ARR_BCC,        /* Field::bcc (3)                 field/field.cpp:30 */
  AGX_ARR_B1F, AGX_ARR_B2F, AGX_ARR_B3F,   /* Field::b.x?f  field/field.cpp:28 */
  AGX_ARR_R1F, AGX_ARR_R2F, AGX_ARR_R3F,   /* Field::b1.x?f field/field.cpp:29 */
  AGX_ARR_FX1, AGX_ARR_FX2, AGX_ARR_FX3,   /* Hydro::flux[3] hydro/hydro.cpp:40-42 */
  AGX_ARR_E3X1, AGX_ARR_E2X1, AGX_ARR_E1X2, AGX_ARR_E3X2, AGX_ARR_E1X3, AGX_ARR_E2X3,
                      /* face EMFs e?_x?f         field/field.cpp:33-38 */
  AGX_ARR_WX1, AGX_ARR_WX2, AGX_ARR_WX3,   /* Field::wght.x?f field/field.cpp:32 */
  AGX_ARR_E1, AGX_ARR_E2, AGX_ARR_E3,      /* Field::e.x?e   field/field.cpp:31 */
  AGX_ARR_PHI,        /* Gravity::phi                   gravity/gravity.cpp:37 */
  AGX_ARR_COUNT
};

/* Kernel classes for the built-in HIP-event timers. */
enum {
  AGX_K_FLUX_X1 = 0, AGX_K_FLUX_X2, AGX_K_FLUX_X3, AGX_K_CORNER_E, AGX_K_INTEGRATE,
  AGX_K_BVALS, AGX_K_C2P, AGX_K_NEWDT, AGX_K_MG_SMOOTH, AGX_K_MG_RESTRICT,
  AGX_K_MG_PROLONG, AGX_K_MG_OTHER, AGX_K_COUNT
};

typedef struct agx_mesh agx_mesh;

/* ---- lifecycle (Mesh::Mesh mesh.cpp:63, MeshBlock::MeshBlock meshblock.cpp:47-185,
 *      Coordinates coordinates.cpp:26-145, MGGravityDriver ctor mg_gravity.cpp:41-138) */
int agx_mesh_create(const agx_mesh_desc *desc, agx_mesh **out);
int agx_mesh_destroy(agx_mesh *m);
const char *agx_last_error(void);
/* number of MeshBlocks and per-block array dims: dims[0..3] = nvar, n3', n2', n1' */
int agx_array_dims(agx_mesh *m, int arr, int64_t *nblocks, int64_t dims[4]);
/* device bytes held by the mesh */
int64_t agx_mesh_bytes(agx_mesh *m);
/* Device address of registered array `arr` (no counterpart in the reference, where each
 * MeshBlock owns its AthenaArrays): the refinement operators below run on a mesh's own
 * arrays through it, with agx_ref_geom.layout = AGX_REF_LAYOUT_MESH.  agx_integrate and
 * agx_cycle rotate the u/u1 and b/b1 registers, so query the address again after them. */
int agx_array_ptr(agx_mesh *m, int arr, void **ptr);
/* The HIP stream every launch on the mesh runs on: agx_ref_* calls on the mesh's arrays
 * pass it so that they are ordered with the task-level entry points. */
int agx_mesh_stream(agx_mesh *m, void **stream);
/* Static mesh refinement: block b of the mesh becomes the MeshBlock at LogicalLocation
 * (level[b], lx[3b..3b+2]) of a tree whose root grid has nrb[0..2] blocks of desc.bnx
 * cells over box = {x1min, x1max, x2min, x2max, x3min, x3max}; its cell widths and PLM
 * factors follow Mesh::SetBlockSizeAndBoundaries on the uniform mesh generator
 * (mesh/mesh.cpp:1668-1730, coordinates/coordinates.cpp:127-145).  The mesh is then
 * advanced with the task-level FP64 steps (agx_calculate_fluxes, agx_compute_corner_e,
 * agx_integrate, agx_cons2prim, agx_new_time_step) and the refinement operators for flux
 * and EMF correction and for the ghost zones across levels, in the order of
 * TimeIntegratorTaskList (time_integrator.cpp:1460-1558); the entry points built on the
 * single-level lattice (agx_cycle, agx_initialize, agx_bvals_exchange, agx_autotune,
 * restart and VTK output) return -60.  Refused for the mixed form, with self-gravity
 * (multigrid on refined meshes is not built) and for an odd nghost (as the
 * MeshRefinement constructor refuses it, mesh_refinement.cpp:56-62; PPM takes 4). */
int agx_mesh_set_levels(agx_mesh *m, const int32_t nrb[3], const double box[6], const int32_t *level,
                        const int64_t *lx);

/* ---- host <-> device (AthenaArray copies at init/output only) */
int agx_upload(agx_mesh *m, int arr, const double *host);
int agx_download(agx_mesh *m, int arr, double *host);
/* MeshBlocks of uneven sizes on one level (no counterpart in the reference's interface:
 * there each MeshBlock, whatever its size, owns its AthenaArrays).  The device holds the
 * mesh as the uniform lattice of desc.bnx blocks, so one launch updates every block and
 * the results equal those of the uniform decomposition.  Block b covers mesh cells
 * origin[3b+d] .. origin[3b+d] + size[3b+d] - 1 along direction d; both are multiples of
 * bnx_d.  host[b] is its array in the layout above with n_d = size_d + 2*nghost (+1 along
 * a face array's normal).  Upload copies the block's interior cells (faces: both boundary
 * faces included).  Download fills the whole array, ghosts included, from the lattice
 * (periodic mesh).  Cell arrays (U, U1, W, BCC, PHI) and face fields (B?F, R?F) only;
 * for initialisation and output (one lattice round trip per call). */
int agx_upload_blocks(agx_mesh *m, int arr, int32_t nblocks, const int32_t *origin, const int32_t *size,
                      const double *const *host);
int agx_download_blocks(agx_mesh *m, int arr, int32_t nblocks, const int32_t *origin, const int32_t *size,
                        double *const *host);
int agx_set_time(agx_mesh *m, double time, double dt, int32_t ncycle);
int agx_get_time(agx_mesh *m, double *time, double *dt, int32_t *ncycle);
int agx_synchronize(agx_mesh *m);

/* ---- VL2 task-list steps, all MeshBlocks per call */
/* Hydro::CalculateFluxes(w, b, bcc, order), hydro/calculate_fluxes.cpp:36-378, with
 * Reconstruction::{DonorCell,PiecewiseLinear,PiecewiseParabolic}X1..3
 * (reconstruct/dc.cpp:24-104, plm.cpp:27-351, ppm.cpp:44-923), Hydro::RiemannSolver
 * (hydro/rsolvers/mhd/hlld.cpp:38-382, hlld_iso.cpp:37-254, hlle_mhd.cpp:25-182,
 * roe_mhd.cpp:43-498) and
 * Hydro::GetWeightForCT
 * (hydro/hydro.cpp:154-158). */
int agx_calculate_fluxes(agx_mesh *m, int order);
/* Field::ComputeCornerE(w, bcc), field/calculate_corner_e.cpp:28-236 (3D). */
int agx_compute_corner_e(agx_mesh *m);
/* TimeIntegratorTaskList::IntegrateHydro + AddSourceTerms + IntegrateField for VL2
 * stage `stage` (time_integrator.cpp:1563-1660): MeshBlock::WeightedAve and register
 * swap (mesh/weighted_ave.cpp:33-236), Hydro::AddFluxDivergence
 * (hydro/add_flux_divergence.cpp:39-96), HydroSourceTerms::SelfGravity
 * (hydro/srcterms/self_gravity.cpp:35-125) and Field::CT (field/ct.cpp:31-116). */
int agx_integrate(agx_mesh *m, int stage);
/* SendHydro/ReceiveHydro/SetBoundaries for Hydro::u and Field::b, same-level
 * neighbours (bvals/bvals_var.cpp:212-296, bvals/cc/bvals_cc.cpp:201-336,
 * bvals/fc/bvals_fc.cpp:344-402,583-700, utils/buffer_utils.cpp:24-89). */
int agx_bvals_exchange(agx_mesh *m, int which);   /* bit 1: u, bit 2: b, bit 4: phi */
/* EquationOfState::ConservedToPrimitive (eos/adiabatic_mhd.cpp:41-90,
 * eos/isothermal_mhd.cpp:40-73) with
 * Field::CalculateCellCenteredField (field/field.cpp:112-180), interior + ghosts. */
int agx_cons2prim(agx_mesh *m);
/* Hydro::NewBlockTimeStep (hydro/new_blockdt.cpp:42-190), min over MeshBlocks. */
int agx_new_block_dt(agx_mesh *m, double *min_dt);
/* Mesh::NewTimeStep (mesh/mesh.cpp:1078-1119). */
int agx_new_time_step(agx_mesh *m, double tlim);
/* Mesh::Initialize, the part after the problem generator (mesh/mesh.cpp:1401-1649):
 * first gravity solve, ghost exchange, ConservedToPrimitive, first time step. */
int agx_initialize(agx_mesh *m, double tlim);
/* ncycles iterations of the main loop (main.cpp:430-489): both VL2 stages with a
 * gravity solve after each (main.cpp:459-466), time += dt, Mesh::NewTimeStep.
 * Stops early once time >= tlim.  Returns the number of cycles run in *done.  Each stage
 * runs as fluxes + one fused kernel (corner EMFs, CT, update, C2P, and in the second
 * stage the time-step bound) + ghost exchange: the same values as the task-level
 * entry points above, in fewer passes over HBM. */
int agx_cycle(agx_mesh *m, double tlim, int32_t ncycles, int32_t *done);

/* ---- multigrid self-gravity */
/* MGGravityDriver::Solve(stage), gravity/mg_gravity.cpp:166-211, with
 * MultigridDriver::SetupMultigrid/SolveFMGCycle/SolveIterative(FixedTimes)
 * (multigrid/multigrid_driver.cpp:343-460,700-771) on periodic boundaries. */
int agx_mg_solve(agx_mesh *m, int stage);
/* <gravity>/mgmode, threshold, niteration at run time (mg_gravity.cpp:44-53). */
int agx_mg_set_mode(agx_mesh *m, int mode, double threshold, int32_t niter);
/* defect norm and V-cycle count of the last solve */
int agx_mg_last(agx_mesh *m, double *norm, int32_t *iters);
/* n V(1,1) cycles on the current finest-level state without setup or norms
 * (MultigridDriver::SolveVCycle, multigrid_driver.cpp:684-693): the unit of the
 * V-cycles/s benchmark.  With multipole boundaries (mg_bc = 1) the cycles continue the
 * last agx_mg_solve (its source, boundary values and FAS levels); before any solve the
 * call returns -22. */
int agx_mg_vcycles(agx_mesh *m, int32_t n);

/* ---- history diagnostics: HistoryOutput::WriteOutputFile, the built-in volume sums
 * (outputs/history.cpp:37-39,135-200) over the interior cells of every MeshBlock:
 *   mass, 1-mom, 2-mom, 3-mom, 1-KE, 2-KE, 3-KE, [tot-E (adiabatic)], [grav-E =
 *   sum 0.5 vol rho phi (self-gravity)], 1-ME, 2-ME, 3-ME (cell-centred 0.5 vol bcc^2),
 * NHYDRO + 3 + (self-gravity) + 3 values, in that order.  The sums are reduced on the
 * device in a fixed order (reproducible bit for bit; the reference's sequential sum
 * differs only in rounding).  agx_history writes them to host memory hst (nmax >= the
 * count) from the mesh's current u, bcc and phi. */
int agx_history(agx_mesh *m, double *hst, int32_t nmax);
/* The same sums over dense device arrays in the reference's AthenaArray order:
 * u (nblocks, nhydro, n3, n2, n1), bcc (nblocks, 3, n3, n2, n1), phi (nblocks, n3, n2, n1)
 * or null (no grav-E), n_d = nx_d + 2 ng; vol[b] = the cell volume dx1 dx2 dx3 of block b
 * (device array); hst = device array of the result.  Runs on `stream` (null: the default
 * stream) and returns after the result is written.  Partial sums go to device memory this
 * call allocates and frees. */
int agx_history_sum(const double *u, const double *bcc, const double *phi, int32_t nblocks, int32_t nhydro,
                    int32_t nx1, int32_t nx2, int32_t nx3, int32_t ng, const double *vol, double *hst,
                    void *stream);
/* agx_history_sum without any device allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under stream capture, tight loops, harnesses that
 * forbid hipMalloc).  Same arguments and results (bit for bit); the partial sums go to a
 * static device array of 2048 x 12 doubles (192 KiB) that each GPU holds until the process
 * ends, and the call returns once its launches are enqueued on `stream` (no
 * synchronisation).  Every size fits (the work split caps the partials at 2048 rows);
 * nblocks * nx3 * nx2 > 2^31 - 1 is rejected.  All calls on one GPU share that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race. */
int agx_history_sum_noalloc(const double *u, const double *bcc, const double *phi, int32_t nblocks,
                            int32_t nhydro, int32_t nx1, int32_t nx2, int32_t nx3, int32_t ng, const double *vol,
                            double *hst, void *stream);

/* ---- restart files (RestartOutput::WriteOutputFile, outputs/restart.cpp:37-207, and
 * the restart constructors Mesh::Mesh(pin, resfile), mesh/mesh.cpp:551-700).  The file
 * layout is the reference's: the parameter text (ParameterInput::ParameterDump) ending in
 * "<par_end>", the header (nbtotal, root_level, RegionSize mesh_size, time, dt, ncycle,
 * datasize), the ID list (LogicalLocation + cost per MeshBlock, Z order of
 * MeshBlockTree::GetMeshBlockList) and per MeshBlock u then b.x1f, b.x2f, b.x3f with
 * ghosts.  Uniform root-level meshes.  params: the caller's parameter text without
 * "<par_end>", or null for the inputs the descriptor carries.  The gravitational
 * potential is not stored; agx_initialize after agx_restart_read solves for it, exchanges
 * ghosts, converts to primitives and takes the first step with the file's dt as the cap
 * (mesh.cpp:1646-1649 with the restored dt). */
int agx_restart_write(agx_mesh *m, const char *path, const char *params);
int agx_restart_read(agx_mesh *m, const char *path);
/* The same on host arrays in the layout of agx_upload (lattice block order, NHYDRO = 4
 * for the isothermal EOS); no device work. */
int agx_restart_write_host(const char *path, const agx_mesh_desc *d, const char *params, double time, double dt,
                           int32_t ncycle, const double *u, const double *b1f, const double *b2f,
                           const double *b3f);
int agx_restart_read_host(const char *path, const agx_mesh_desc *d, double *time, double *dt, int32_t *ncycle,
                          double *u, double *b1f, double *b2f, double *b3f);

/* ---- VTK legacy output (VTKOutput::WriteOutputFile, outputs/vtk.cpp:54-220): one file
 * per MeshBlock, "<basename>.block<gid>.<file_id>.<file_number as %05d>.vtk", binary
 * big-endian float32 on a rectilinear grid of the block's face coordinates, the cell
 * data of OutputType::LoadOutputData (outputs/outputs.cpp:355-660) for the comma-
 * separated tokens of <output>/variable: prim (rho, press, vel, phi, Bcc), cons (dens,
 * Etot, mom, phi, Bcc), D, d, E, p, m, m1-m3, v, v1-v3 (vx-vz), phi, bcc, bcc1-bcc3;
 * other tokens are refused (-3).  ghost != 0: include_ghost_zones.  Slices and sums are
 * not supported. */
int agx_vtk_write(agx_mesh *m, const char *basename, const char *file_id, int32_t file_number,
                  const char *variable, int32_t ghost);
/* The same from host arrays in the layout of agx_download (phi may be null without
 * self-gravity); no device work. */
int agx_vtk_write_host(const char *basename, const char *file_id, int32_t file_number, const char *variable,
                       int32_t ghost, const agx_mesh_desc *d, double time, int32_t ncycle, const double *u,
                       const double *w, const double *bcc, const double *phi);
/* The reference's MeshBlock numbering (MeshBlockTree::GetMeshBlockList, Z order,
 * mesh/meshblock_tree.cpp:336-352): lattice_of_gid[gid] = (lx3*nb2 + lx2)*nb1 + lx1 for
 * gid = 0 .. nblocks-1; no device work. */
int agx_block_gids(const agx_mesh_desc *d, int64_t *lattice_of_gid);

/* ---- multigrid at task granularity (MultigridTaskList, multigrid/mg_task_list.cpp:
 * 183-280, 287-530; MGGravity::Smooth/CalculateDefect/CalculateFASRHS, mg_gravity.hpp:
 * 34-40; Multigrid::Restrict/ProlongateAndCorrect/FMGProlongate, multigrid.hpp:120-127)
 * for callers that drive their own cycles.  Each step runs on one level of the FP64
 * hierarchy of the whole mesh (level 0 = root, agx_mg_levels gives the count and sizes)
 * with the kernels of agx_mg_solve, so a V(1,1) composed of them equals agx_mg_vcycles(1)
 * bit for bit.  Order of a V(1,1) from the finest level F, after agx_mg_coffset(m, 1, 0):
 *   periodic:  l = F..1: smooth(l,0), smooth(l,1), restrict(l); solve_coarsest;
 *              l = 1..F: prolong(l), smooth(l,0), smooth(l,1)
 *   multipole: l = F..1: fill_ghost(l), [fas(l) if l < F], smooth(l,0), fill_ghost(l),
 *              smooth(l,1), fill_ghost(l), restrict(l); solve_coarsest;
 *              l = 1..F: fill_ghost(l-1), prolong(l), fill_ghost(l), smooth(l,0),
 *              fill_ghost(l), smooth(l,1), fill_ghost(l)
 * agx_mg_setup loads the source (LoadSource, the source mask, SetupMultigrid: mean
 * subtraction or multipole coefficients) and restricts it to every level (RestrictFMGSource);
 * agx_mg_retrieve writes phi (RetrieveResult).  Periodic levels keep no ghosts (periodic
 * index wrap): agx_mg_fill_ghost does nothing there.  agx_mg_fas: multipole (FAS) only. */
int agx_mg_setup(agx_mesh *m);
int agx_mg_smooth(agx_mesh *m, int32_t level, int32_t color);
int agx_mg_fill_ghost(agx_mesh *m, int32_t level);
int agx_mg_fas(agx_mesh *m, int32_t level);
int agx_mg_restrict(agx_mesh *m, int32_t level);          /* level -> level - 1 */
int agx_mg_prolong(agx_mesh *m, int32_t level);           /* level - 1 -> level */
int agx_mg_fmg_prolong(agx_mesh *m, int32_t level);       /* level - 1 -> level */
int agx_mg_solve_coarsest(agx_mesh *m);
int agx_mg_defect_norm(agx_mesh *m, double *norm);       /* finest level, L2 */
int agx_mg_retrieve(agx_mesh *m);
int agx_mg_coffset(agx_mesh *m, int32_t flip, int32_t *value);
int agx_mg_levels(agx_mesh *m, int32_t *nlevels, int32_t *dims, int32_t cap);

/* ---- tuning (no counterpart in the reference: launch shapes of this library).
 * agx_autotune times each form of a VL2 stage on the current state, for the first
 * (donor-cell) stage and for the second (the mesh's order) separately, and keeps the
 * fastest for each; *choice gets the second stage's form.  Forms 0..3: three flux launches, then one stage
 * kernel over 16x16, 32x8, 64x4 (i,j) columns per 256-thread workgroup or 64x8 per
 * 512-thread workgroup; forms 4..5: the whole stage (fluxes in LDS) in one launch over
 * 16x16 or 32x8 columns (form 5 is the default).  With multigrid
 * gravity (periodic boundaries) it also times V-cycles with the red-black pass fused into
 * one launch and as two sweeps, and keeps the faster.  It writes only scratch registers
 * (u1, b1, fluxes, the multigrid levels, which every solve reloads), so call it between
 * cycles, e.g. after agx_initialize.  agx_set_tile forces a shape.  Results do not depend
 * on either choice. */
int agx_autotune(agx_mesh *m, int32_t *choice);
int agx_set_tile(agx_mesh *m, int32_t tile);   /* the same form for both stages */
/* multigrid red-black pass form: -1 by level size (fused on FP64 levels of >= 2^26
 * cells, the default), 0 two sweeps, 1 fused on every level of >= 2^21 cells */
int agx_set_mg_fuse(agx_mesh *m, int32_t mode);
/* Multipole V-cycles: run the coarse levels (at most 512 cells each) as one
 * single-workgroup launch (on = 1, the default) or one launch per step (0).  The
 * potentials are identical bit for bit; no counterpart in the reference.  Applies to
 * the FP64 V-cycles and to the FP32 correction cycles of the mixed form. */
int agx_set_mg_coarse(agx_mesh *m, int32_t on);
/* Source mask of MultigridDriver::SetupMultigrid (multigrid_driver.cpp:430-436): after
 * LoadSource, the source is zero at every finest-level cell centre with |x - c|^2 > r^2,
 * the mask the reference's collapse problem enrolls (pgen/collapse.cpp:66-82, there with
 * c = 0).  Applied before the mean subtraction (periodic) or the multipole expansion.
 * radius <= 0 switches it off.  The reference takes an arbitrary user function here;
 * this entry point offers its spherical form. */
int agx_mg_set_source_mask_sphere(agx_mesh *m, double xc, double yc, double zc, double radius);

/* ---- hipGraph replay of agx_cycle (no counterpart in the reference: launch mechanism).
 * With enable != 0, agx_cycle captures one whole VL2 cycle (fluxes, stage kernels,
 * ghost C2P, the dt reduction and the device-side time advance) into a hipGraph on first
 * use and replays it, one graph launch per cycle; dt and time stay on the device and are
 * read back once per cycle for the tlim test.  Results are identical to the eager path.
 * With self-gravity the graph also holds both multigrid solves of the cycle when they run
 * a fixed number of V-cycles (threshold < 0, niteration; SolveIterativeFixedTimes,
 * multigrid_driver.cpp:762-771), periodic or multipole;
 * agx_mg_last then reports the V-cycle count and a NaN norm (the norm is only diagnostic
 * and is not read back).  Solves to a threshold (SolveIterative, multigrid_driver.cpp:
 * 721-755) replay as three graphs per cycle, split where the host reads how many V-cycles
 * each solve ran: each solve's loop is recorded with the count the same stage's previous
 * solve needed, a device-side check applies the loop's stopping rule after every V-cycle
 * and the recorded V-cycles after the stop return at once, and a solve that needs more
 * goes on eagerly; agx_mg_last reports the norm and count as the eager path does.
 * Cycles with the built-in timers on run eagerly.  Every setter that changes the
 * recorded work (multigrid mode, pass form, coarse kernel, source mask, stage form)
 * discards the recorded graphs.  On by default; enable = 0 selects the eager path. */
int agx_set_graph(agx_mesh *m, int32_t enable);

/* ---- static-mesh-refinement operators.
 * Batched forms of the per-MeshBlock refinement routines for 3D Cartesian blocks.  They
 * work on device arrays in the reference's dense AthenaArray order, block-major:
 *   cells      (nb, nvar, n3, n2, n1)
 *   faces      x1f (nb, n3, n2, n1+1), x2f (nb, n3, n2+1, n1), x3f (nb, n3+1, n2, n1)
 *   fluxes     as faces, with nvar
 *   edges      x1e (nb, n3+1, n2+1, n1), x2e (nb, n3+1, n2, n1+1), x3e (nb, n3, n2+1, n1+1)
 * Blocks have n = nx + 2 ng cells per direction.  Their coarse buffers (coarse_cons_,
 * coarse_b_) have nc = nx/2 + 2 cng, coarse interior from index cng; the reference uses
 * cng = NGHOST (mesh/meshblock.cpp:59-85: ncc1 = nx1/2 + 2*NGHOST, cis = NGHOST).  Coordinates are
 * per-block device tables x?v and dx?f of length n (blocks) or nc (coarse buffers).  On
 * Cartesian blocks the face positions x?s? are x?v (coordinates/cartesian.cpp:106-119),
 * areas and volumes are products of dx?f, and edge lengths are dx?f
 * (coordinates/coordinates.cpp:328-537).
 * A work list (agx_ref_list) holds the items of one operator in device memory.  It is
 * built and checked once, since refinement is static; every item of a list runs in one
 * launch.  Items: fb = block on the fine side, cb = block or coarse buffer on the coarse
 * side, [si..ei] x [sj..ej] x [sk..ek] = coarse index range (the reference's
 * csi..cek / si..ek arguments), fid / fi1 / fi2 = face and quadrant (NeighborBlock::fid,
 * NeighborIndexes::fi1/fi2).  All calls are asynchronous on `stream` (a hipStream_t, or
 * NULL for the default stream).  Items of one list must not write the same element. */
typedef struct {
  int32_t fb, cb;
  int32_t si, ei, sj, ej, sk, ek;
  int32_t fid, fi1, fi2, aux;   /* aux: face component and launch group of the face
                                   pack/unpack kinds, 0 otherwise */
} agx_ref_item;
typedef struct {
  int32_t nx1, nx2, nx3;   /* interior cells of a block (even) */
  int32_t ng, cng;         /* ghost width of the blocks and of the coarse buffers (= ng in
                              the reference) */
  int32_t layout;          /* layout of the block-side arrays (MeshBlock data; coarse
                              buffers are always dense): AGX_REF_LAYOUT_DENSE, or
                              AGX_REF_LAYOUT_MESH for the arrays of an agx_mesh with the
                              same block size and ghost width (agx_array_dims), so that the
                              operators work on a mesh's own arrays (whose hydro arrays
                              hold 5 slabs per block for either EOS) */
} agx_ref_geom;
enum { AGX_REF_LAYOUT_DENSE = 0, AGX_REF_LAYOUT_MESH = 1 };
typedef struct {
  const double *x1v, *x2v, *x3v, *dx1f, *dx2f, *dx3f;
} agx_ref_coords;
enum {
  AGX_REF_RESTRICT_CC = 0,     /* MeshRefinement::RestrictCellCenteredValues, mesh_refinement.cpp:106-136 */
  AGX_REF_RESTRICT_FC1 = 1,    /* RestrictFieldX1, mesh_refinement.cpp:185-214 */
  AGX_REF_RESTRICT_FC2 = 2,    /* RestrictFieldX2, mesh_refinement.cpp:244-279 */
  AGX_REF_RESTRICT_FC3 = 3,    /* RestrictFieldX3, mesh_refinement.cpp:321-344 */
  AGX_REF_PROLONG_CC = 4,      /* ProlongateCellCenteredValues, mesh_refinement.cpp:386-458 */
  AGX_REF_PROLONG_FC1 = 5,     /* ProlongateSharedFieldX1, mesh_refinement.cpp:541-581 */
  AGX_REF_PROLONG_FC2 = 6,     /* ProlongateSharedFieldX2, mesh_refinement.cpp:626-674 */
  AGX_REF_PROLONG_FC3 = 7,     /* ProlongateSharedFieldX3, mesh_refinement.cpp:718-762 */
  AGX_REF_PROLONG_FC_INT = 8,  /* ProlongateInternalField, mesh_refinement.cpp:831-946 */
  AGX_REF_FLUX_FROM_FINER = 9, /* flux_correction_cc.cpp:77-169 (LoadFluxBoundaryBufferToCoarser)
                                  + 240-283 (SetFluxBoundaryFromFiner): fb = finer block,
                                  cb = coarser block, fid = face of cb, fi1/fi2 = quadrant */
  AGX_REF_EMF_TO_COARSER = 10, /* flux_correction_fc.cpp:315-546 (LoadFluxBoundaryBufferToCoarser,
                                  3D face and edge neighbours): fb = finer block, fid = its face
                                  0..5 or 6 + its edge (NeighborBlock::eid) */
  AGX_REF_PRIM_TO_CONS = 11,   /* Field::CalculateCellCenteredField (uniform branch, field.cpp:112-176)
                                  + EquationOfState::PrimitiveToConserved (adiabatic_mhd.cpp:99-136,
                                  isothermal_mhd.cpp:82-108) on the fine cells of each coarse range,
                                  the last step of ProlongateGhostCells (bvals_refine.cpp) */
  AGX_REF_EMF_SAME_LEVEL = 12, /* LoadFluxBoundaryBufferSameLevel, flux_correction_fc.cpp:51-313:
                                  fb = block, fid = its face or 6 + edge */
  AGX_REF_EMF_SET_SAME = 13,   /* SetFluxBoundarySameLevel, flux_correction_fc.cpp:689-940 */
  AGX_REF_EMF_SET_FINER = 14,  /* SetFluxBoundaryFromFiner, flux_correction_fc.cpp:944-1185:
                                  cb = receiving block, fb = buffer slot, fid = cb's face or
                                  6 + edge, fi1/fi2 = the finer block's quadrant (fi1 = half
                                  on an edge) */
  AGX_REF_EMF_CLEAR = 15,      /* ClearCoarseFluxBoundary, flux_correction_fc.cpp:1214-1348: cb,
                                  fid = a face with finer neighbours or an edge with a finer
                                  block around it */
  AGX_REF_EMF_AVERAGE = 16,    /* AverageFluxBoundary, flux_correction_fc.cpp:1355-1543: cb, fid =
                                  a face with a block or periodic boundary and fi1 = 0 (same
                                  level) or 1 (finer), or an edge and fi1 = nedge_fine (2..4) */
  AGX_REF_COPY_CC_BB = 17,     /* cell-centred ghost exchange between blocks of one level:
                                  LoadBoundaryBufferSameLevel + SetBoundarySameLevel
                                  (bvals/cc/bvals_cc.cpp:201-221, 300-341) */
  AGX_REF_COPY_CC_CB = 18,     /* coarse buffer of a finer block -> ghosts of a coarser one:
                                  the pack of LoadBoundaryBufferToCoarser (bvals_cc.cpp:223-248,
                                  after AGX_REF_RESTRICT_CC) + SetBoundaryFromFiner (:407-479) */
  AGX_REF_COPY_CC_BC = 19,     /* coarser block -> coarse buffer of a finer one:
                                  LoadBoundaryBufferToFiner + SetBoundaryFromCoarser
                                  (bvals_cc.cpp:250-298, 343-405).  Copy items: fb = source,
                                  [si..ek] = source range, cb = destination, (fid, fi1, fi2) =
                                  the destination's first (i, j, k) */
  AGX_REF_PACK_FC_B = 20,      /* face-field ghost exchange (bvals/fc/bvals_fc.cpp:344-977), pack
                                  from block faces: fb = block, [si..ek] = range of component
                                  aux & 3 (1..3 = x1f, x2f, x3f), into buf + item * stride */
  AGX_REF_PACK_FC_C = 21,      /* the same from coarse-buffer faces (after AGX_REF_RESTRICT_FC*) */
  AGX_REF_UNPACK_FC_B = 22,    /* unpack into block faces: cb = block, [si..ek] = range of
                                  component aux & 3, fb = buffer slot; items run in launch groups
                                  of increasing aux >> 2 (0..15) */
  AGX_REF_UNPACK_FC_C = 23,    /* the same into coarse-buffer faces */
  AGX_REF_CONS_TO_PRIM = 24,   /* CalculateCellCenteredField + ConservedToPrimitive on coarse-buffer
                                  cells, as ApplyPhysicalBoundariesOnCoarseLevel calls them
                                  (bvals_refine.cpp:336-384): cb = block, [si..ek] = coarse cells */
  AGX_REF_KINDS = 25
};
typedef struct agx_ref_list agx_ref_list;
/* Checks every item against the geometry and the pool sizes (nb_fine blocks on the fine
 * side, nb_coarse on the coarse side) for operator `kind`, and copies the list to the
 * device. */
int agx_ref_list_create(const agx_ref_geom *g, int32_t kind, const agx_ref_item *items, int32_t nitems,
                        int32_t nb_fine, int32_t nb_coarse, agx_ref_list **out);
int agx_ref_list_destroy(agx_ref_list *l);
/* number of values one item of an AGX_REF_EMF_TO_COARSER list writes (the buffer size
 * of the reference) for face fid 0..5 or edge 6 + eid */
int64_t agx_ref_emf_count(const agx_ref_geom *g, int32_t fid);
/* the same for an AGX_REF_EMF_SAME_LEVEL item (BoundaryBase's same-level EMF buffer) */
int64_t agx_ref_emf_count_same_level(const agx_ref_geom *g, int32_t fid);
/* fine cells -> coarse buffer (kind RESTRICT_CC) */
int agx_ref_restrict_cc(const agx_ref_list *l, int32_t nvar, const double *fine, agx_ref_coords fine_coords,
                        double *coarse, void *stream);
/* fine faces -> coarse faces of the direction of the list (RESTRICT_FC1..3) */
int agx_ref_restrict_fc(const agx_ref_list *l, const double *fine, agx_ref_coords fine_coords, double *coarse,
                        void *stream);
/* coarse buffer -> fine cells, minmod-limited (PROLONG_CC) */
int agx_ref_prolong_cc(const agx_ref_list *l, int32_t nvar, const double *coarse, agx_ref_coords coarse_coords,
                       double *fine, agx_ref_coords fine_coords, void *stream);
/* coarse faces -> fine faces shared with the coarse grid (PROLONG_FC1..3) */
int agx_ref_prolong_fc_shared(const agx_ref_list *l, const double *coarse, agx_ref_coords coarse_coords,
                              double *fine, agx_ref_coords fine_coords, void *stream);
/* interior fine faces from the shared ones, divergence-preserving (PROLONG_FC_INT) */
int agx_ref_prolong_fc_internal(const agx_ref_list *l, double *b1f, double *b2f, double *b3f,
                                agx_ref_coords fine_coords, void *stream);
/* coarser-block face fluxes of each quadrant <- area-weighted finer fluxes (FLUX_FROM_FINER) */
int agx_ref_flux_from_finer(const agx_ref_list *l, int32_t nvar, const double *fine_x1flux,
                            const double *fine_x2flux, const double *fine_x3flux, agx_ref_coords fine_coords,
                            double *coarse_x1flux, double *coarse_x2flux, double *coarse_x3flux, void *stream);
/* restricted edge EMFs of each item's face or edge into buf + item * stride (EMF_TO_COARSER) */
int agx_ref_emf_to_coarser(const agx_ref_list *l, const double *e1, const double *e2, const double *e3,
                           agx_ref_coords fine_coords, double *buf, int64_t stride, void *stream);
/* EMF correction across blocks, one cycle (FaceCenteredBoundaryVariable::SendFluxCorrection
 * and ReceiveFluxCorrection, flux_correction_fc.cpp:623-688, 1610-1660).  Edge arrays are
 * x1e (nb, n3+1, n2+1, n1), x2e (nb, n3+1, n2, n1+1), x3e (nb, n3, n2+1, n1+1).  Order of
 * one correction: pack every buffer (agx_ref_emf_same_level, agx_ref_emf_to_coarser) from
 * the blocks' own EMFs; then agx_ref_emf_set with the SET_SAME list, agx_ref_emf_clear,
 * agx_ref_emf_set with the SET_FINER list, agx_ref_emf_average.  A block's items follow
 * its neighbour order; the library launches the x1-face, x2-face, x3-face and edge items
 * one class after another (and finer quadrants one after another), which gives every
 * element the reference's summation order.  One item per (cb, fid[, fi1, fi2]) in a
 * list (checked). */
/* box copies of nvar cell-centred variables (COPY_CC_BB: blocks to blocks, COPY_CC_CB:
 * coarse buffers to blocks, COPY_CC_BC: blocks to coarse buffers); items of one list
 * must not write the same element */
int agx_ref_copy_cc(const agx_ref_list *l, int32_t nvar, const double *src, double *dst, void *stream);
/* face-field ghost exchange: packs (PACK_FC_B / _C) read x?f of blocks or coarse
 * buffers into buf + item * stride; unpacks (UNPACK_FC_B / _C) write buf + fb * stride
 * into x?f, one launch per group (aux >> 2) in increasing order.  Run every pack of an
 * exchange before its unpacks. */
int agx_ref_pack_fc(const agx_ref_list *l, const double *x1f, const double *x2f, const double *x3f, double *buf,
                    int64_t stride, void *stream);
int agx_ref_unpack_fc(const agx_ref_list *l, double *x1f, double *x2f, double *x3f, const double *buf,
                      int64_t stride, void *stream);
/* each item's own EMFs on its face or edge into buf + item * stride (EMF_SAME_LEVEL) */
int agx_ref_emf_same_level(const agx_ref_list *l, const double *e1, const double *e2, const double *e3,
                           double *buf, int64_t stride, void *stream);
/* e += buf + item.fb * stride on each item's face, quadrant or edge (EMF_SET_SAME, EMF_SET_FINER) */
int agx_ref_emf_set(const agx_ref_list *l, double *e1, double *e2, double *e3, const double *buf,
                    int64_t stride, void *stream);
/* zero the EMFs that finer neighbours will supply (EMF_CLEAR) */
int agx_ref_emf_clear(const agx_ref_list *l, double *e1, double *e2, double *e3, void *stream);
/* divide the summed EMFs by the number of contributions (EMF_AVERAGE) */
int agx_ref_emf_average(const agx_ref_list *l, double *e1, double *e2, double *e3, void *stream);
/* FaceCenteredBoundaryVariable::CountFineEdges (bvals/fc/bvals_fc.cpp:1128-1191): from the
 * 27 neighbour levels relative to the block (index as for agx_ref_prolong_ranges,
 * nblevel[13] = 0, negative = coarser or missing), edge_flag[e] = 1 when no finer block
 * touches edge e and nedge_fine[e] = the number of finest-level blocks around it.  Host
 * only. */
int agx_ref_count_fine_edges(const int32_t nblevel[27], int32_t edge_flag[12], int32_t nedge_fine[12]);
/* fine cells of the item ranges: bcc from the faces, then the conserved variables from
 * w and bcc (nvar 5 with the adiabatic EOS, gamma; 4 with iso != 0) (PRIM_TO_CONS) */
int agx_ref_prim_to_cons(const agx_ref_list *l, const double *w, const double *b1f, const double *b2f,
                         const double *b3f, double *bcc, double *u, double gamma, int32_t iso, void *stream);
/* Fused launches for one slot of the ghost prolongation (no counterpart in the
 * reference; the lists and results are those of the single-list calls).  Any list may
 * be NULL.  restrict_all: RESTRICT_CC (u -> cu, nvar) and RESTRICT_FC1..3 (b -> cb) in
 * one launch; prolong_all: PROLONG_CC (cw -> w) and PROLONG_FC1..3 (shared faces);
 * internal_to_cons: a PROLONG_FC_INT list, each 2x2x2 group's inner faces then bcc and u
 * of its 8 cells (agx_ref_prolong_fc_internal followed by agx_ref_prim_to_cons on the
 * same ranges). */
int agx_ref_restrict_all(const agx_ref_list *lcc, const agx_ref_list *lf1, const agx_ref_list *lf2,
                         const agx_ref_list *lf3, int32_t nvar, const double *u, const double *b1f, const double *b2f,
                         const double *b3f, agx_ref_coords fine_coords, double *cu, double *cb1f, double *cb2f,
                         double *cb3f, void *stream);
int agx_ref_prolong_all(const agx_ref_list *lcc, const agx_ref_list *lf1, const agx_ref_list *lf2,
                        const agx_ref_list *lf3, int32_t nvar, const double *cw, const double *cb1f, const double *cb2f,
                        const double *cb3f, agx_ref_coords coarse_coords, double *w, double *b1f, double *b2f,
                        double *b3f, agx_ref_coords fine_coords, void *stream);
int agx_ref_internal_to_cons(const agx_ref_list *l, double *b1f, double *b2f, double *b3f, agx_ref_coords fine_coords,
                             const double *w, double *bcc, double *u, double gamma, int32_t iso, void *stream);
/* coarse conserved variables and faces -> coarse bcc and primitives on the item ranges,
 * floors written back to u (CONS_TO_PRIM; nvar 5 with the adiabatic EOS, 4 with iso) */
int agx_ref_cons_to_prim(const agx_ref_list *l, double *u, const double *b1f, const double *b2f, const double *b3f,
                         double *w, double *bcc, double gamma, double dfloor, double pfloor, int32_t iso,
                         void *stream);
/* Loop limits of BoundaryValues::ProlongateBoundaries for one coarser neighbour
 * (bvals_refine.cpp:172-199) and of the face prolongation in ProlongateGhostCells.  ox[3]
 * is the neighbour's offset (NeighborIndexes::ox1..3); lx_odd[3] = loc.lx1..3 & 1;
 * nblevel[27] holds the levels of the 3x3x3 neighbourhood minus this block's level, at
 * index ((ox3+1)*3 + ox2+1)*3 + ox1+1 (-1 where there is no neighbour).  out[0..5] =
 * si, ei, sj, ej, sk, ek (coarse cells: PROLONG_CC, PROLONG_FC_INT and PRIM_TO_CONS
 * items), out[6..7] = il, iu (PROLONG_FC1
 * faces, with sj..ek), out[8..9] = jl, ju (PROLONG_FC2), out[10..11] = kl, ku
 * (PROLONG_FC3).  The reference's cnghost = (NGHOST + 1)/2 + 1 with NGHOST = g->ng.  Host
 * only; no device work. */
int agx_ref_prolong_ranges(const agx_ref_geom *g, const int32_t ox[3], const int32_t lx_odd[3],
                           const int32_t nblevel[27], int32_t out[12]);
/* waits for `stream` */
int agx_ref_sync(void *stream);
/* Streams and graphs for callers without a runtime of their own (no counterpart in the
 * reference).  agx_stream_create makes a non-blocking stream of default priority.
 * agx_graph_begin / agx_graph_end capture the launches issued on that stream in
 * between (the refinement entry points with that stream) into an executable graph;
 * agx_graph_launch replays it on a stream.  A graph keeps the pointers and work lists
 * it was captured with: replay only while they are alive and unchanged. */
int agx_stream_create(void **out);
int agx_stream_destroy(void *stream);
int agx_graph_begin(void *stream);
int agx_graph_end(void *stream, void **exec_out);
int agx_graph_launch(void *exec, void *stream);
int agx_graph_destroy(void *exec);

/* ---- device memory for callers without an allocator of their own (no counterpart in the
 * reference; the refinement operators take device pointers).  Plain device allocations
 * on the current device; copies are synchronous.  kind: 0 host to device, 1 device to
 * host, 2 device to device. */
int agx_device_alloc(int64_t bytes, void **out);
int agx_device_free(void *p);
int agx_device_copy(void *dst, const void *src, int64_t bytes, int32_t kind);

/* ---- built-in timers (HIP events on the mesh stream around each launch) */
int agx_timing_enable(agx_mesh *m, int enable);
int agx_timing_reset(agx_mesh *m);
int agx_timing_get(agx_mesh *m, int kclass, double *ms_total, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* ATHENA_GAMMA_AMD_H_ */
