// agx_diag.hip -- history diagnostics (volume sums) of the gfx950 MHD library.
//
// HistoryOutput::WriteOutputFile (outputs/history.cpp:37-39,135-200) sums, over the
// interior cells of every MeshBlock, vol * (rho, M1, M2, M3, 0.5 M_d^2 / rho, E) plus the
// gravitational energy 0.5 vol rho phi and the cell-centred magnetic energies
// 0.5 vol bcc_d^2.  Here one launch sums a contiguous range of cell rows per workgroup
// into per-workgroup partials, and one single-workgroup launch adds the partials in a
// fixed order, so the result is reproducible bit for bit run to run (the reference adds
// the same per-cell terms sequentially, so the two agree to rounding of the sum order).
//
// Entry points sharing the kernels:
//   agx_history       all MeshBlocks of an agx_mesh (device-resident state);
//   agx_history_sum   dense AthenaArray inputs; partials in memory it allocates.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "agx_internal.hpp"

namespace agx {

constexpr int HIST_BS = 256;                 // threads per workgroup
constexpr int64_t HIST_CELLS_PER_WG = 4096;  // target work per workgroup

// Strided view of the arrays the sums read.  Element (b, n, k, j, i) of u is
// u[b*u_bs + n*u_vs + k*sk + j*sj + i]; bcc and phi use their own block/variable strides
// and the same k/j strides.  Interior cells are k in [ks, ks+nx3), j in [js, js+nx2),
// i in [is, is+nx1).
template <typename TB>
struct HistView {
  const double *u;
  const TB *bcc;
  const double *phi;                 // null: no gravitational energy
  const double *vol;                 // per-block cell volume (dx1 dx2 dx3), device
  int64_t u_bs, u_vs, b_bs, b_vs, p_bs, sk, sj;
  int nb, nh, nx1, nx2, nx3, is, js, ks;
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Per-workgroup partial sums of the history terms over rows [r0, r1) of the
// (block, k, j) row space; partial[w * HIST_NV + v].
template <typename TB>
__global__ void __launch_bounds__(HIST_BS)
k_hist_partial(HistView<TB> h, int64_t nrows, int nwg, double *__restrict__ partial) {
  __shared__ double s_part[HIST_BS / 64][HIST_NV];
  const int w = blockIdx.x;
  const int64_t r0 = nrows * w / nwg, r1 = nrows * (w + 1) / nwg;
  const int64_t ncell = (r1 - r0) * h.nx1;
  const int nout = h.nh + 3 + (h.phi ? 1 : 0) + 3;
  const int ime = h.nh + 3 + (h.phi ? 1 : 0);
  double acc[HIST_NV];
#pragma unroll
  for (int v = 0; v < HIST_NV; ++v) acc[v] = 0.0;
  for (int64_t l = threadIdx.x; l < ncell; l += HIST_BS) {
    const int64_t row = r0 + l / h.nx1;
    const int i = h.is + (int)(l % h.nx1);
    const int64_t bk = row / h.nx2;
    const int j = h.js + (int)(row - bk * h.nx2);
    const int b = (int)(bk / h.nx3);
    const int k = h.ks + (int)(bk - (int64_t)b * h.nx3);
    const int64_t id = k * h.sk + j * h.sj + i;
    const double vol = h.vol[b];
    const double *u = h.u + b * h.u_bs + id;
    const double u_d = u[0], u_mx = u[h.u_vs], u_my = u[2 * h.u_vs], u_mz = u[3 * h.u_vs];
    // history.cpp:156-186, per-cell terms formed as written there
    acc[0] += vol * u_d;
    acc[1] += vol * u_mx;
    acc[2] += vol * u_my;
    acc[3] += vol * u_mz;
    acc[4] += vol * 0.5 * (u_mx * u_mx) / u_d;
    acc[5] += vol * 0.5 * (u_my * u_my) / u_d;
    acc[6] += vol * 0.5 * (u_mz * u_mz) / u_d;
    if (h.nh == 5) acc[7] += vol * u[4 * h.u_vs];
    if (h.phi) acc[h.nh + 3] += vol * 0.5 * u_d * h.phi[b * h.p_bs + id];
    const TB *bc = h.bcc + b * h.b_bs + id;
    const double b1 = (double)bc[0], b2 = (double)bc[h.b_vs], b3 = (double)bc[2 * h.b_vs];
    acc[ime] += vol * 0.5 * b1 * b1;
    acc[ime + 1] += vol * 0.5 * b2 * b2;
    acc[ime + 2] += vol * 0.5 * b3 * b3;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < HIST_NV; ++v) {
    const double s = wave_sum(acc[v]);
    if (lane == 0) s_part[wv][v] = s;
  }
  __syncthreads();
  if (threadIdx.x < nout) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < HIST_BS / 64; ++q) s += s_part[q][threadIdx.x];
    partial[(int64_t)w * HIST_NV + threadIdx.x] = s;
  }
}

// Sum of the nwg partial rows, value by value, in a fixed order (thread t adds rows
// t, t + 256, ...; then a fixed tree): out[v], v < nout.
__global__ void __launch_bounds__(HIST_BS)
k_hist_final(const double *__restrict__ partial, int nwg, int nout, double *__restrict__ out) {
  __shared__ double s_part[HIST_BS / 64];
  for (int v = 0; v < nout; ++v) {
    double s = 0.0;
    for (int w = threadIdx.x; w < nwg; w += HIST_BS) s += partial[(int64_t)w * HIST_NV + v];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < HIST_BS / 64; ++q) t += s_part[q];
      out[v] = t;
    }
    __syncthreads();
  }
}

static int hist_nwg(int64_t nrows, int nx1) {
  const int64_t cells = nrows * nx1;
  int64_t nw = (cells + HIST_CELLS_PER_WG - 1) / HIST_CELLS_PER_WG;
  if (nw < 1) nw = 1;
  if (nw > HIST_NW_MAX) nw = HIST_NW_MAX;
  if (nw > nrows) nw = nrows;
  return (int)nw;
}

template <typename TB>
static int launch_history(const HistView<TB> &h, double *partial, double *out, hipStream_t st) {
  const int64_t nrows = (int64_t)h.nb * h.nx3 * h.nx2;
  const int nwg = hist_nwg(nrows, h.nx1);
  const int nout = h.nh + 3 + (h.phi ? 1 : 0) + 3;
  hipLaunchKernelGGL(k_hist_partial<TB>, dim3(nwg), dim3(HIST_BS), 0, st, h, nrows, nwg, partial);
  hipLaunchKernelGGL(k_hist_final, dim3(1), dim3(HIST_BS), 0, st, partial, nwg, nout, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "history launches");
}

int history_check_dense(const double *u, const double *bcc, int32_t nblocks, int32_t nhydro, int32_t nx1,
                       int32_t nx2, int32_t nx3, int32_t ng, const double *vol, const double *hst,
                       const char *who) {
  if (!u || !bcc || !vol || !hst) { set_error("%s: null u, bcc, vol or output pointer", who); return -1; }
  if (nblocks <= 0 || nx1 <= 0 || nx2 <= 0 || nx3 <= 0 || ng < 0) {
    set_error("%s: nblocks and nx1..3 must be positive, ng >= 0", who);
    return -2;
  }
  if (nhydro != 4 && nhydro != 5) { set_error("%s: nhydro must be 4 (isothermal) or 5 (adiabatic)", who); return -3; }
  if ((int64_t)nblocks * nx3 * nx2 > INT32_MAX) {
    set_error("%s: more than 2^31 - 1 cell rows (nblocks * nx3 * nx2)", who);
    return -4;
  }
  return 0;
}

static HistView<double> dense_view(const double *u, const double *bcc, const double *phi, int32_t nblocks,
                                   int32_t nhydro, int32_t nx1, int32_t nx2, int32_t nx3, int32_t ng,
                                   const double *vol) {
  HistView<double> h;
  const int64_t n1 = nx1 + 2 * ng, n2 = nx2 + 2 * ng, n3 = nx3 + 2 * ng;
  const int64_t cs = n1 * n2 * n3;
  h.u = u; h.bcc = bcc; h.phi = phi; h.vol = vol;
  h.u_vs = cs; h.u_bs = nhydro * cs;
  h.b_vs = cs; h.b_bs = 3 * cs;
  h.p_bs = cs;
  h.sk = n1 * n2; h.sj = n1;
  h.nb = nblocks; h.nh = nhydro;
  h.nx1 = nx1; h.nx2 = nx2; h.nx3 = nx3;
  h.is = h.js = h.ks = ng;
  return h;
}

// Sums over dense AthenaArray inputs with the partials in `partial` (at least
// HIST_PARTIAL_DOUBLES doubles); the result goes to the device array hst.
int history_dense(const double *u, const double *bcc, const double *phi, int32_t nblocks, int32_t nhydro,
                  int32_t nx1, int32_t nx2, int32_t nx3, int32_t ng, const double *vol, double *partial,
                  double *hst, hipStream_t st) {
  const HistView<double> h = dense_view(u, bcc, phi, nblocks, nhydro, nx1, nx2, nx3, ng, vol);
  return launch_history(h, partial, hst, st);
}

}  // namespace agx

using namespace agx;

extern "C" {

int agx_history_sum(const double *u, const double *bcc, const double *phi, int32_t nblocks, int32_t nhydro,
                    int32_t nx1, int32_t nx2, int32_t nx3, int32_t ng, const double *vol, double *hst,
                    void *stream) {
  int rc = history_check_dense(u, bcc, nblocks, nhydro, nx1, nx2, nx3, ng, vol, hst, "agx_history_sum");
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double *partial = nullptr;
  hipError_t e = hipMalloc(&partial, sizeof(double) * HIST_PARTIAL_DOUBLES);
  if (e != hipSuccess) return hip_fail(e, "agx_history_sum: hipMalloc(partials)");
  rc = history_dense(u, bcc, phi, nblocks, nhydro, nx1, nx2, nx3, ng, vol, partial, hst, st);
  // the partials are freed only after the launches that use them have finished
  e = hipStreamSynchronize(st);
  hipFree(partial);
  if (rc) return rc;
  return e == hipSuccess ? 0 : hip_fail(e, "agx_history_sum");
}

int agx_history(agx_mesh *m, double *hst, int32_t nmax) {
  if (!m || !hst) { set_error("agx_history: null mesh or output"); return -1; }
  const Geo &g = m->g;
  const int nh = m->iso ? 4 : 5;
  const int nout = nh + 3 + (m->mg_on ? 1 : 0) + 3;
  if (nmax < nout) { set_error("agx_history: output holds %d values, %d needed", nmax, nout); return -2; }
  int rc = 0;
  double *part = m->hist;
  double *out = m->hist + HIST_PARTIAL_DOUBLES;
  const double *phi = m->mg_on ? m->arr.a[AGX_ARR_PHI] : nullptr;
  auto fill = [&](auto *bcc) {
    using TB = std::remove_const_t<std::remove_pointer_t<decltype(bcc)>>;
    HistView<TB> h;
    h.u = m->arr.a[AGX_ARR_U]; h.bcc = bcc; h.phi = phi; h.vol = m->cellvol;
    h.u_vs = g.S; h.u_bs = (int64_t)NHYDRO * g.S;
    h.b_vs = g.S; h.b_bs = 3 * g.S;
    h.p_bs = g.S;
    h.sk = (int64_t)g.J * g.P; h.sj = g.P;
    h.nb = g.nb; h.nh = nh;
    h.nx1 = g.bx1; h.nx2 = g.bx2; h.nx3 = g.bx3;
    h.is = g.is; h.js = g.js; h.ks = g.ks;
    return launch_history(h, part, out, m->stream);
  };
  if (m->arr.esz[AGX_ARR_BCC] == 4) rc = fill(reinterpret_cast<const float *>(m->arr.a[AGX_ARR_BCC]));
  else rc = fill(static_cast<const double *>(m->arr.a[AGX_ARR_BCC]));
  if (rc) return rc;
  AGX_HIP(hipMemcpyAsync(hst, out, sizeof(double) * nout, hipMemcpyDeviceToHost, m->stream));
  AGX_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

}  // extern "C"
