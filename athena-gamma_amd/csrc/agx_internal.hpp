// agx_internal.hpp -- shared definitions of the gfx950 MHD + multigrid library.
//
// Device layout: every registered array of every MeshBlock lives in one HBM buffer,
// block-major.  A per-variable slab is a padded box of (n3+1) planes x (n2+1) rows x
// P elements (P = pitch, a multiple of 32 elements, P >= n1+1), so cell, face and edge
// arrays share one index function IDX(k,j,i) = (k*J + j)*P + i with J = n2+1.  Each
// array's base is offset so that element i = is of every row opens a 128-B line.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/athena_gamma_amd.h"

namespace agx {

constexpr int NHYDRO = 5;
constexpr int NWAVE = 7;
constexpr int IDN = 0, IVX = 1, IVY = 2, IVZ = 3, IEN = 4, IPR = 4, IBY = 5, IBZ = 6;
constexpr int IB1 = 0, IB2 = 1, IB3 = 2;

// Geometry of the block-batched arrays, passed by value to every kernel.
struct Geo {
  int nb;                  // MeshBlocks
  int n1, n2, n3;          // cells per block including ghosts
  int ng;                  // ghost width
  int is, ie, js, je, ks, ke;
  int bx1, bx2, bx3;       // interior cells per block
  int nb1, nb2, nb3;       // block lattice
  int P;                   // row pitch (doubles)
  int PL;                  // stride of the per-direction PLM coefficient rows (>= max n_d)
  int J;                   // rows per plane = n2 + 1
  int64_t S;               // per-variable slab = (n3+1)*J*P
};

__host__ __device__ inline int64_t IDX(const Geo &g, int k, int j, int i) {
  return ((int64_t)k * g.J + j) * g.P + i;
}

// XCD-aware workgroup order.  The dispatcher hands workgroup n of a launch to XCD
// n mod 8, and each XCD has its own L2, so with the plain order a workgroup's stencil
// neighbours (the next rows, planes or tiles) run on other XCDs and their shared lines
// are fetched once per XCD.  xcd_slot renumbers the launch so that each XCD runs one
// contiguous range of the index space, in order; it is a bijection on [0, total).
__host__ __device__ inline unsigned xcd_slot(unsigned lin, unsigned total) {
  constexpr unsigned NXCD = 8;
  const unsigned q = total / NXCD, r = total % NXCD;
  const unsigned x = lin % NXCD, s = lin / NXCD;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + s;
}

// (x, y, z) workgroup index of this workgroup in the XCD-aware order
__device__ __forceinline__ uint3 xcd_block() {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned n = xcd_slot(lin, gx * gy * gridDim.z);
  const unsigned yz = n / gx;
  return make_uint3(n - yz * gx, yz % gy, yz / gy);
}

// Device scalar slots (agx_mesh::scal, 64 doubles): 0-3 reduction results; the time
// step, time and tlim of the fused cycle, which the flux and stage kernels read from
// device memory so that a captured cycle replays with the current step.
constexpr int SC_DT = 40, SC_TIME = 41, SC_TLIM = 42;

// Per-block constant geometry on the device (uniform Cartesian mesh).
struct BlockCoef {
  double dx[3];            // dx?f (== dx?v) of the block, coordinates.cpp:127-145
};

// Internal registers after the public AGX_ARR_* set (not reachable through the C ABI):
// the second primitive/bcc buffer and the third face-field register used by the fused
// stage kernel, which reads one buffer while it writes the other.
constexpr int ARR_W2 = AGX_ARR_COUNT, ARR_BCC2 = AGX_ARR_COUNT + 1;
constexpr int ARR_Q1F = AGX_ARR_COUNT + 2, ARR_Q2F = AGX_ARR_COUNT + 3, ARR_Q3F = AGX_ARR_COUNT + 4;
constexpr int ARR_TOTAL = AGX_ARR_COUNT + 5;

// Registered device arrays (AGX_ARR_* numbering, then the internal registers).  In the
// mixed-precision form the primitive, bcc, flux, face-EMF and CT-weight arrays hold FP32
// values (esz = 4) behind the same pointer; everything else is FP64.
struct Arrays {
  double *a[ARR_TOTAL];
  int nvar[ARR_TOTAL];
  int esz[ARR_TOTAL];
};

// Multigrid level data: one periodic grid per level covering the whole mesh.  The
// mixed-precision form keeps a second, FP32 hierarchy for the correction V-cycles.
template <typename T>
struct MGLevelT {
  int n1, n2, n3;
  double dx;               // cell size used by the smoother/defect at this level
  T *u, *src, *def;
};
using MGLevel = MGLevelT<double>;
using MGLevel32 = MGLevelT<float>;

// Multigrid level with a one-cell ghost frame, for non-periodic (multipole) boundaries:
// (n1+2) x (n2+2) x (n3+2) cells, index ((k*(n2+2) + j)*(n1+2) + i), k,j,i in [0, n+1].
// uold is the FAS copy of u (Multigrid::StoreOldData).
struct MGLevelB {
  int n1, n2, n3;
  double dx;               // cell size of the smoother/defect (as MGLevel::dx)
  double h[3];             // cell size per direction for coordinates: (xmax - xmin) / n
  double *u, *src, *uold;
  double *bcv;             // multipole boundary values of each ghost's chain (3 per ghost)
};

// FP32 framed level of the mixed-precision multipole form: the correction, its source
// and the FAS copy; the ghosts hold the homogeneous multipole condition of the correction
struct MGLevelB32 {
  int n1, n2, n3;
  double dx;
  float *u, *src, *uold;
};

// Device state of a threshold solve whose V-cycle loop is recorded for a fixed number of
// passes (the graph replay of agx_cycle).  While closed != 0 every multigrid kernel of the
// loop returns at once; k_mg_gate_check closes it where SolveIterative's loop stops.
struct MGGate {
  int closed;
  int iters;     // V-cycles the loop has run
  double def;    // defect norm after the last of them (before the loop: the first norm)
};

// Host-side multigrid state that a V-cycle changes: the red-black offset, the level
// cursor and the arrays the fused pass and the mixed correction swap.
struct MGSnap {
  std::vector<MGLevel> lev;
  std::vector<MGLevel32> lev32;
  int coffset = 0, current = 0, fmglevel = 0;
};

// History sums (agx_diag.hip): at most HIST_NW_MAX workgroups, each writing HIST_NV
// partial sums (5 hydro + 3 kinetic + 1 gravitational + 3 magnetic at most).
constexpr int HIST_NW_MAX = 2048;
constexpr int HIST_NV = 12;
constexpr int64_t HIST_PARTIAL_DOUBLES = (int64_t)HIST_NW_MAX * HIST_NV;

struct Timer {
  bool enabled = false;
  std::vector<hipEvent_t> ev;        // pairs
  std::vector<int> cls;              // class of each pair
  size_t used = 0;
  double ms[AGX_K_COUNT] = {0};
  int64_t launches[AGX_K_COUNT] = {0};
};

}  // namespace agx

struct agx_mesh {
  agx_mesh_desc d;
  agx::Geo g;
  agx::Arrays arr;
  bool mixed = false;                // AGX_PREC_MIXED
  bool iso = false;                  // isothermal EOS (NHYDRO = 4)
  int tile = 5;                      // stage form of the last VL2 stage (agx_autotune; 5 = k_vl2, 32 x 8)
  int tile1 = 5;                     // stage form of the first (donor-cell) stage
  hipStream_t stream = nullptr;
  double *pool = nullptr;            // all MHD arrays
  int64_t pool_bytes = 0;
  agx::BlockCoef *coef = nullptr;    // [nb]
  double *plm = nullptr;             // PLM coefficients [nb][3 dirs][2][P]
  double *scal = nullptr;            // device scalars (dt reduction etc.), 64 doubles
  double *scal_host = nullptr;       // pinned host mirror
  double *red = nullptr;             // reduction partials
  int64_t red_n = 0;
  double *hist = nullptr;            // history partial sums + result (agx_history)
  double *cellvol = nullptr;         // [nb] cell volume dx1*dx2*dx3 of each MeshBlock
  uint8_t *nearf = nullptr;          // mixed form: per-face near-tie flags of the FP32 PPM fluxes
  int64_t nearf_n = 0;               // bytes of nearf, nb*S rounded up to the fix-pass span
  double *dtpart = nullptr;          // per-workgroup time-step partials of the fused stage
  int64_t dtpart_n = 0;
  std::vector<agx::BlockCoef> coef_h;
  std::vector<double> plm_h;
  double time = 0.0, dt = 0.0;
  int32_t ncycle = 0;
  bool restarted = false;            // dt came from a restart file: agx_initialize keeps it
  bool tree = false;                 // blocks on refinement levels (agx_mesh_set_levels)
  int copies_state = -1;             // the two copies of every shared face agree: 1 yes, 0 no,
                                     // -1 not checked since the last upload of a face array
  // multigrid
  int mg_on = 0;
  int nrootlevel = 0, nmblevel = 0, ntotal = 0, current = 0, fmglevel = 0, coffset = 0;
  int mg_mode = 0, mg_niter = 0;
  int mg_fuse = -1;
  int mg_coarse = 1;   // multipole V-cycles: the coarse levels in one workgroup (k_mb_coarse)
  // source mask (agx_mg_set_source_mask_sphere): src = 0 where |x - c|^2 > r^2
  int srcmask = 0;
  double srcmask_c[3] = {0.0, 0.0, 0.0};
  double srcmask_r2 = 0.0;                  // red-black pass form: -1 by level size, 0 two sweeps, 1 fused
  double mg_eps = -1.0, last_norm = 0.0, last_ave = 0.0;
  int32_t last_iters = 0;
  agx::MGGate *gatedev = nullptr;    // gate of a recorded threshold loop (device)
  agx::MGGate *gatehost = nullptr;   // its pinned host mirror
  const int *gate = nullptr;         // &gatedev->closed while a gated loop is enqueued, else null
  int mg_kpred[2] = {-1, -1};        // V-cycles of the last threshold solve of each stage
  std::vector<agx::MGLevel> lev;
  std::vector<agx::MGLevel32> lev32;   // mixed precision only
  double *mgpool = nullptr;
  float *mgpool32 = nullptr;
  int64_t mg_bytes = 0;
  // multipole boundaries (mg_bc = 1): framed levels, coefficients and origin on the device
  bool mp_on = false;
  bool mp_solved = false;            // a multipole solve has loaded the levels (mg_vcycles needs it)
  std::vector<agx::MGLevelB> levb;
  double *mgpoolb = nullptr;
  std::vector<agx::MGLevelB32> levb32;   // mixed precision with multipole boundaries
  float *mgpoolb32 = nullptr;
  double *mpdev = nullptr;           // [0..2] origin, [3..27] coefficients, [28..] partials
  double rdx_root = 0.0, rdy_root = 0.0, rdz_root = 0.0;
  double rdx_blk = 0.0, rdy_blk = 0.0, rdz_blk = 0.0;
  agx::Timer timer;
  // hipGraph replay of agx_cycle: one executable graph per register assignment (the face
  // registers rotate with period two cycles), keyed by the pointer table it starts from
  bool graph_on = true;
  struct CycleGraph {
    hipGraphExec_t exec = nullptr;
    double *before[agx::ARR_TOTAL];
    double *after[agx::ARR_TOTAL];
    int coffset_before = 0, coffset_after = 0;   // multigrid red-black offset (self-gravity)
    int gen = 0;                                  // graph_gen when recorded
  };
  // threshold solves: the cycle is recorded in three segments, each ending where the
  // host reads how many of the recorded V-cycles ran (mg_solve_settle)
  struct SegGraph {
    hipGraphExec_t exec = nullptr;
    int seg = 0, k = 0, gen = 0;
    double *before[agx::ARR_TOTAL];
    double *after[agx::ARR_TOTAL];
    std::vector<uintptr_t> mgkey;          // multigrid state at the start (mg_key)
    std::vector<agx::MGSnap> snaps;        // multigrid state after 0..k recorded passes
  };
  std::vector<SegGraph> seggraphs;
  bool capturing = false;            // a cycle is being recorded into a graph
  int graph_gen = 0;                 // bumped by every setter that changes the recorded work
  std::vector<CycleGraph> graphs;
};

namespace agx {

// Face coordinates xf[0..n] and centres xv[0..n-1] (n = nbx + 2 ng) of the block at
// logical position lx along one direction of a uniform mesh, and its cell width
// (coordinates/coordinates.cpp:127-145, cartesian.cpp:26-36, mesh/mesh.cpp:1668-1690).
void block_coords(int nmesh, int nblk, int nbx, int ng, int lx, double xmin, double xmax,
                  std::vector<double> &xf, std::vector<double> &xv, double &dx);

// Lattice index (lx3*nb2 + lx2)*nb1 + lx1 of each MeshBlock in the reference's gid order:
// the Z order in which MeshBlockTree::GetMeshBlockList lists the root grid
// (mesh/meshblock_tree.cpp:87-110,336-352).
std::vector<int64_t> gid_lattice_order(int nb1, int nb2, int nb3);

void set_error(const char *fmt, ...);
// entry points built on the lattice of one level (its ghost exchange, file layouts) refuse
// a mesh with refinement levels
int refuse_tree(const agx_mesh *m, const char *what);
int hip_fail(hipError_t e, const char *what);

// history sums over dense AthenaArray inputs (agx_diag.hip)
int history_check_dense(const double *u, const double *bcc, int32_t nblocks, int32_t nhydro, int32_t nx1,
                        int32_t nx2, int32_t nx3, int32_t ng, const double *vol, const double *hst,
                        const char *who);
int history_dense(const double *u, const double *bcc, const double *phi, int32_t nblocks, int32_t nhydro,
                  int32_t nx1, int32_t nx2, int32_t nx3, int32_t ng, const double *vol, double *partial,
                  double *hst, hipStream_t st);

#define AGX_HIP(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return agx::hip_fail(e_, #call); \
  } while (0)

// launch bracketing for the built-in timers
void timer_begin(agx_mesh *m, int cls);
void timer_end(agx_mesh *m);
int timer_flush(agx_mesh *m);

inline unsigned grid1d(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// multigrid entry points implemented in agx_mg.hip
int mg_create(agx_mesh *m);
void mg_destroy(agx_mesh *m);
int mg_solve(agx_mesh *m, int stage);
// A threshold solve in parts that enqueue no host read-back, for graph recording:
// mg_solve_head (load, FMG, the first defect norm, the gate), mg_solve_gated (k V-cycles
// gated on the device by the stopping rule; the host state after 0..k of them into snaps),
// then, after the stream has run them, mg_solve_settle (the host state of the V-cycles
// that ran; the loop continued eagerly if it had not stopped) and mg_solve_tail (mean
// subtraction, retrieve).
int mg_solve_head(agx_mesh *m);
void mg_solve_gated(agx_mesh *m, int k, std::vector<agx::MGSnap> &snaps);
int mg_solve_settle(agx_mesh *m, int stage, const std::vector<agx::MGSnap> &snaps);
int mg_solve_tail(agx_mesh *m);
agx::MGSnap mg_snap(const agx_mesh *m);
void mg_restore(agx_mesh *m, const agx::MGSnap &s);
std::vector<uintptr_t> mg_key(const agx_mesh *m);
int mg_vcycles(agx_mesh *m, int n);
int mg_autotune(agx_mesh *m);

// C2P of one cell in registers (eos/adiabatic_mhd.cpp:52-86).  Returns true when a
// floor changed the conserved state.
__device__ __forceinline__ bool c2p_cell(double &u_d, double u_m1, double u_m2, double u_m3,
                                         double &u_e, double bcc1, double bcc2, double bcc3,
                                         double gm1, double dfl, double pfl, double *wv) {
  bool fl = !(u_d > dfl);
  u_d = (u_d > dfl) ? u_d : dfl;
  double di = 1.0 / u_d;
  double pb = 0.5 * (bcc1 * bcc1 + bcc2 * bcc2 + bcc3 * bcc3);
  double e_k = 0.5 * di * (u_m1 * u_m1 + u_m2 * u_m2 + u_m3 * u_m3);
  double w_p = gm1 * (u_e - e_k - pb);
  fl = fl || !(w_p > pfl);
  u_e = (w_p > pfl) ? u_e : ((pfl / gm1) + e_k + pb);
  w_p = (w_p > pfl) ? w_p : pfl;
  wv[0] = u_d;
  wv[1] = u_m1 * di;
  wv[2] = u_m2 * di;
  wv[3] = u_m3 * di;
  wv[4] = w_p;
  return fl;
}

// Isothermal C2P of one cell (eos/isothermal_mhd.cpp:40-73): density floor, velocities.
__device__ __forceinline__ void c2p_cell_iso(double &u_d, double u_m1, double u_m2, double u_m3,
                                             double dfl, double *wv) {
  u_d = (u_d > dfl) ? u_d : dfl;
  double di = 1.0 / u_d;
  wv[0] = u_d;
  wv[1] = u_m1 * di;
  wv[2] = u_m2 * di;
  wv[3] = u_m3 * di;
  wv[4] = 0.0;
}

// C2P of one cell for either EOS; u_e is untouched in the isothermal case.
template <bool ISO>
__device__ __forceinline__ void c2p_eos(double &u_d, double u_m1, double u_m2, double u_m3,
                                        double &u_e, double bcc1, double bcc2, double bcc3,
                                        double gm1, double dfl, double pfl, double *wv) {
  if (ISO) c2p_cell_iso(u_d, u_m1, u_m2, u_m3, dfl, wv);
  else c2p_cell(u_d, u_m1, u_m2, u_m3, u_e, bcc1, bcc2, bcc3, gm1, dfl, pfl, wv);
}

}  // namespace agx
