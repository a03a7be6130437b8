// agx_mhd.hip -- gfx950 kernels and C-ABI entry points of the per-MeshBlock MHD update.
//
// Plain FP64 path.  Built with -ffp-contract=off: every expression below rounds in
// the operand order of the reference method it restates (cited per kernel, paths
// relative to the reference's src/), so results match the CPU FP64 restatement
// bit for bit.
//
// Launch shape: one thread per face / edge / cell; grid.z = MeshBlock, grid.y = the
// k-plane, grid.x = 256-thread tiles over the (j,i) plane, i fastest, so a wavefront
// reads 64 consecutive doubles of a 128-B aligned row.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <algorithm>
#include <cstring>
#include <new>
#include <type_traits>

#include "agx_internal.hpp"

using namespace agx;

namespace agx {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int refuse_tree(const agx_mesh *m, const char *what) {
  if (!m->tree) return 0;
  set_error("%s: the mesh has refinement levels; it runs through the task-level entry points and the "
            "refinement operators", what);
  return -60;
}

int hip_fail(hipError_t e, const char *what) {
  set_error("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return (int)e > 0 ? (int)e : 1;
}

// ---------------------------------------------------------------- timers
void timer_begin(agx_mesh *m, int cls) {
  Timer &t = m->timer;
  if (!t.enabled) return;
  if (t.used * 2 + 2 > t.ev.size()) timer_flush(m);
  hipEventRecord(t.ev[2 * t.used], m->stream);
  t.cls[t.used] = cls;
}

void timer_end(agx_mesh *m) {
  Timer &t = m->timer;
  if (!t.enabled) return;
  hipEventRecord(t.ev[2 * t.used + 1], m->stream);
  t.used++;
}

int timer_flush(agx_mesh *m) {
  Timer &t = m->timer;
  if (t.used == 0) return 0;
  AGX_HIP(hipEventSynchronize(t.ev[2 * t.used - 1]));
  for (size_t q = 0; q < t.used; ++q) {
    float ms = 0.f;
    AGX_HIP(hipEventElapsedTime(&ms, t.ev[2 * q], t.ev[2 * q + 1]));
    t.ms[t.cls[q]] += ms;
    t.launches[t.cls[q]]++;
  }
  t.used = 0;
  return 0;
}

}  // namespace agx

// ================================================================ device helpers
template <typename T>
__device__ __forceinline__ T dmin(T a, T b) { return (b < a) ? b : a; }  // std::min
template <typename T>
__device__ __forceinline__ T dmax(T a, T b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double xsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float xsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double xabs(double x) { return fabs(x); }
__device__ __forceinline__ float xabs(float x) { return fabsf(x); }
#define SQR(x) ((x) * (x))

struct PlaneIdx {
  int b, k, j, i;
};

// (b,k,j,i) of this thread for the box [kl..ku]x[jl..ju]x[il..iu]; false if outside.
// With xcd, workgroups are taken in the XCD-aware order (xcd_block): consecutive planes
// of a block then run on one XCD, which pays off for kernels with a k stencil.
__device__ __forceinline__ bool plane_index(int il, int iu, int jl, int ju, int kl, PlaneIdx &p,
                                            bool xcd = false) {
  const unsigned ni = iu - il + 1, nj = ju - jl + 1;
  const uint3 wg = xcd ? xcd_block() : make_uint3(blockIdx.x, blockIdx.y, blockIdx.z);
  const unsigned q = wg.x * blockDim.x + threadIdx.x;
  if (q >= ni * nj) return false;
  const unsigned qj = q / ni;
  p.j = jl + (int)qj;
  p.i = il + (int)(q - qj * ni);
  p.k = kl + (int)wg.y;
  p.b = wg.z;
  return true;
}

static dim3 plane_grid(int il, int iu, int jl, int ju, int kl, int ku, int nb) {
  int64_t n = (int64_t)(iu - il + 1) * (ju - jl + 1);
  return dim3(grid1d(n, 256), (unsigned)(ku - kl + 1), (unsigned)nb);
}

// Stage forms and tiles: TI x TJ (i,j) columns per workgroup of TI*TJ threads (256, or
// 512 for the 64 x 8 tile); grid.x = tiles of one MeshBlock, grid.y = MeshBlock.  Forms
// 0-3 run the fluxes as three k_flux launches and the rest of the stage as k_stage with
// the tile shown; forms 4-5 run the whole stage as one k_vl2 launch.  The form is picked
// per mesh by agx_autotune (default: 5).
constexpr int NTILE = 6, TILE_VL2 = 4, TILE_DEFAULT = 5;
constexpr int TILE_I[NTILE] = {16, 32, 64, 64, 16, 32}, TILE_J[NTILE] = {16, 8, 4, 8, 16, 8};
static dim3 stage_grid(const Geo &g, int tile) {
  int nti = (g.bx1 + TILE_I[tile] - 1) / TILE_I[tile], ntj = (g.bx2 + TILE_J[tile] - 1) / TILE_J[tile];
  return dim3((unsigned)(nti * ntj), (unsigned)g.nb);
}

template <typename X>
static inline X *as(double *p) { return reinterpret_cast<X *>(p); }

// EquationOfState::FastMagnetosonicSpeed, eos/adiabatic_mhd.cpp:149-156
template <typename T>
__device__ __forceinline__ T fast_speed(T gamma, const T *prim, T bx) {
  T asq = gamma * prim[IPR];
  T vaxsq = bx * bx;
  T ct2 = (prim[IBY] * prim[IBY] + prim[IBZ] * prim[IBZ]);
  T qsq = vaxsq + ct2 + asq;
  T tmp = vaxsq + ct2 - asq;
  return xsqrt(T(0.5) * (qsq + xsqrt(tmp * tmp + T(4.0) * asq * ct2)) / prim[IDN]);
}

// EquationOfState::FastMagnetosonicSpeed, eos/isothermal_mhd.cpp:122-129
template <typename T>
__device__ __forceinline__ T fast_speed_iso(T cs, const T *prim, T bx) {
  T asq = (cs * cs) * prim[IDN];
  T vaxsq = bx * bx;
  T ct2 = prim[IBY] * prim[IBY] + prim[IBZ] * prim[IBZ];
  T qsq = vaxsq + ct2 + asq;
  T tmp = vaxsq + ct2 - asq;
  return xsqrt(T(0.5) * (qsq + xsqrt(tmp * tmp + T(4.0) * asq * ct2)) / prim[IDN]);
}

template <bool ISO, typename T>
__device__ __forceinline__ T fast_speed_eos(T gamma, T cs, const T *prim, T bx) {
  return ISO ? fast_speed_iso(cs, prim, bx) : fast_speed(gamma, prim, bx);
}

// Hydro::GetWeightForCT, hydro/hydro.cpp:154-158
template <typename T>
__device__ __forceinline__ T ct_weight(T dflx, T rhol, T rhor, T dx,
                                            T dt) {
  T v_over_c = (T(1024.0)) * dt * dflx / (dx * (rhol + rhor));
  T tmp_min = dmin(T(0.5), v_over_c);
  return T(0.5) + dmax(-T(0.5), tmp_min);
}

template <typename T>
struct Cons1D {
  T d, mx, my, mz, e, by, bz;
};

// HLLD (Miyoshi & Kusano 2005) for adiabatic MHD at one interface,
// hydro/rsolvers/mhd/hlld.cpp:56-369, SMALL_NUMBER = 1e-8 (hlld.cpp:32); with LOW, the
// low-dissipation variant LHLLD (Minoshima, Kitamura & Miyoshi 2021),
// hydro/rsolvers/mhd/lhlld.cpp:56-389: the shock detector theta = th1^4 (from the
// normal and transverse velocity differences dvn, dvt of the face) scales the pressure
// jump in S_M, and phi = chi (2 - chi) the velocity jump in the star total pressure
// (lhlld.cpp:158-187); the rest is HLLD, with lhlld.cpp's grouping in the one place it
// differs (the transverse-field factor, :212).
template <typename T, bool LOW>
__device__ __forceinline__ void hlld_core(T gamma, const T *wli, const T *wri, T bxi, T dvn, T dvt,
                                          T *flxi) {
  const T SMALL = T(1.0e-8);
  T igm1 = T(1.0) / (gamma - T(1.0));
  Cons1D<T> ul, ur, ulst, uldst, urdst, urst, fl, fr;
  T spd[5];
  T bxsq = bxi * bxi;
  T pbl = T(0.5) * (bxsq + (SQR(wli[IBY]) + SQR(wli[IBZ])));
  T pbr = T(0.5) * (bxsq + (SQR(wri[IBY]) + SQR(wri[IBZ])));
  T kel = T(0.5) * wli[IDN] * (SQR(wli[IVX]) + (SQR(wli[IVY]) + SQR(wli[IVZ])));
  T ker = T(0.5) * wri[IDN] * (SQR(wri[IVX]) + (SQR(wri[IVY]) + SQR(wri[IVZ])));
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

  T cfl = fast_speed(gamma, wli, bxi);
  T cfr = fast_speed(gamma, wri, bxi);
  spd[0] = dmin(wli[IVX] - cfl, wri[IVX] - cfr);
  spd[4] = dmax(wli[IVX] + cfl, wri[IVX] + cfr);

  T ptl = wli[IPR] + pbl;
  T ptr = wri[IPR] + pbr;
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

  T sdl = spd[0] - wli[IVX];
  T sdr = spd[4] - wri[IVX];
  const T cfmax = dmax(cfl, cfr);
  if (LOW) {   // shock detector, lhlld.cpp:159-162
    T th1 = dmin(T(1.0), (cfmax - dmin(dvn, T(0.0))) / (cfmax - dmin(dvt, T(0.0))));
    T th = th1 * th1 * th1 * th1;
    spd[2] = (sdr * ur.mx - sdl * ul.mx + th * (ptl - ptr)) / (sdr * ur.d - sdl * ul.d);
  } else {
    spd[2] = (sdr * ur.mx - sdl * ul.mx + (ptl - ptr)) / (sdr * ur.d - sdl * ul.d);
  }
  T sdml = spd[0] - spd[2];
  T sdmr = spd[4] - spd[2];
  T sdml_inv = T(1.0) / sdml;
  T sdmr_inv = T(1.0) / sdmr;
  ulst.d = ul.d * sdl * sdml_inv;
  urst.d = ur.d * sdr * sdmr_inv;
  T ulst_d_inv = T(1.0) / ulst.d;
  T urst_d_inv = T(1.0) / urst.d;
  T sqrtdl = xsqrt(ulst.d);
  T sqrtdr = xsqrt(urst.d);
  spd[1] = spd[2] - xabs(bxi) / sqrtdl;
  spd[3] = spd[2] + xabs(bxi) / sqrtdr;

  T ptst;
  if (LOW) {   // low-Mach star pressure, lhlld.cpp:182-187
    T sdld = sdl * ul.d, sdrd = sdr * ur.d;
    T clsq = ((pbl + kel) + xsqrt(SQR(pbl + kel) - T(2.0) * kel * bxsq)) / ul.d;
    T crsq = ((pbr + ker) + xsqrt(SQR(pbr + ker) - T(2.0) * ker * bxsq)) / ur.d;
    T chi = dmin(T(1.0), xsqrt(dmax(clsq, crsq)) / cfmax);
    T phi = chi * (T(2.0) - chi);
    ptst = (sdrd * ptl - sdld * ptr + phi * sdrd * sdld * (wri[IVX] - wli[IVX])) / (sdrd - sdld);
  } else {
    T ptstl = ptl + ul.d * sdl * (spd[2] - wli[IVX]);
    T ptstr = ptr + ur.d * sdr * (spd[2] - wri[IVX]);
    ptst = T(0.5) * (ptstr + ptstl);
  }

  ulst.mx = ulst.d * spd[2];
  if (xabs(ul.d * sdl * sdml - bxsq) < SMALL * ptst) {
    ulst.my = ulst.d * wli[IVY];
    ulst.mz = ulst.d * wli[IVZ];
    ulst.by = ul.by;
    ulst.bz = ul.bz;
  } else {
    T tmp = bxi * (sdl - sdml) / (ul.d * sdl * sdml - bxsq);
    ulst.my = ulst.d * (wli[IVY] - ul.by * tmp);
    ulst.mz = ulst.d * (wli[IVZ] - ul.bz * tmp);
    tmp = (LOW ? (sdl * ul.d) * sdl - bxsq : ul.d * SQR(sdl) - bxsq) / (ul.d * sdl * sdml - bxsq);
    ulst.by = ul.by * tmp;
    ulst.bz = ul.bz * tmp;
  }
  T vbstl = (ulst.mx * bxi + (ulst.my * ulst.by + ulst.mz * ulst.bz)) * ulst_d_inv;
  ulst.e = (sdl * ul.e - ptl * wli[IVX] + ptst * spd[2] +
            bxi * (wli[IVX] * bxi + (wli[IVY] * ul.by + wli[IVZ] * ul.bz) - vbstl)) * sdml_inv;

  urst.mx = urst.d * spd[2];
  if (xabs(ur.d * sdr * sdmr - bxsq) < SMALL * ptst) {
    urst.my = urst.d * wri[IVY];
    urst.mz = urst.d * wri[IVZ];
    urst.by = ur.by;
    urst.bz = ur.bz;
  } else {
    T tmp = bxi * (sdr - sdmr) / (ur.d * sdr * sdmr - bxsq);
    urst.my = urst.d * (wri[IVY] - ur.by * tmp);
    urst.mz = urst.d * (wri[IVZ] - ur.bz * tmp);
    tmp = (LOW ? (sdr * ur.d) * sdr - bxsq : ur.d * SQR(sdr) - bxsq) / (ur.d * sdr * sdmr - bxsq);
    urst.by = ur.by * tmp;
    urst.bz = ur.bz * tmp;
  }
  T vbstr = (urst.mx * bxi + (urst.my * urst.by + urst.mz * urst.bz)) * urst_d_inv;
  urst.e = (sdr * ur.e - ptr * wri[IVX] + ptst * spd[2] +
            bxi * (wri[IVX] * bxi + (wri[IVY] * ur.by + wri[IVZ] * ur.bz) - vbstr)) * sdmr_inv;

  if (T(0.5) * bxsq < SMALL * ptst) {
    uldst = ulst;
    urdst = urst;
  } else {
    T invsumd = T(1.0) / (sqrtdl + sqrtdr);
    T bxsig = (bxi > T(0.0) ? T(1.0) : -T(1.0));
    uldst.d = ulst.d;
    urdst.d = urst.d;
    uldst.mx = ulst.mx;
    urdst.mx = urst.mx;
    T tmp = invsumd * (sqrtdl * (ulst.my * ulst_d_inv) + sqrtdr * (urst.my * urst_d_inv) +
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

  // flux selection (hlld.cpp:315-369); the differences are formed only for the branch
  // taken, which rounds identically to forming all of them first.
  if (spd[0] >= T(0.0)) {
    flxi[IDN] = fl.d; flxi[IVX] = fl.mx; flxi[IVY] = fl.my; flxi[IVZ] = fl.mz;
    flxi[IEN] = fl.e; flxi[IBY] = fl.by; flxi[IBZ] = fl.bz;
  } else if (spd[4] <= T(0.0)) {
    flxi[IDN] = fr.d; flxi[IVX] = fr.mx; flxi[IVY] = fr.my; flxi[IVZ] = fr.mz;
    flxi[IEN] = fr.e; flxi[IBY] = fr.by; flxi[IBZ] = fr.bz;
  } else if (spd[1] >= T(0.0)) {
    flxi[IDN] = fl.d + spd[0] * (ulst.d - ul.d);
    flxi[IVX] = fl.mx + spd[0] * (ulst.mx - ul.mx);
    flxi[IVY] = fl.my + spd[0] * (ulst.my - ul.my);
    flxi[IVZ] = fl.mz + spd[0] * (ulst.mz - ul.mz);
    flxi[IEN] = fl.e + spd[0] * (ulst.e - ul.e);
    flxi[IBY] = fl.by + spd[0] * (ulst.by - ul.by);
    flxi[IBZ] = fl.bz + spd[0] * (ulst.bz - ul.bz);
  } else if (spd[2] >= T(0.0)) {
    flxi[IDN] = fl.d + spd[0] * (ulst.d - ul.d) + spd[1] * (uldst.d - ulst.d);
    flxi[IVX] = fl.mx + spd[0] * (ulst.mx - ul.mx) + spd[1] * (uldst.mx - ulst.mx);
    flxi[IVY] = fl.my + spd[0] * (ulst.my - ul.my) + spd[1] * (uldst.my - ulst.my);
    flxi[IVZ] = fl.mz + spd[0] * (ulst.mz - ul.mz) + spd[1] * (uldst.mz - ulst.mz);
    flxi[IEN] = fl.e + spd[0] * (ulst.e - ul.e) + spd[1] * (uldst.e - ulst.e);
    flxi[IBY] = fl.by + spd[0] * (ulst.by - ul.by) + spd[1] * (uldst.by - ulst.by);
    flxi[IBZ] = fl.bz + spd[0] * (ulst.bz - ul.bz) + spd[1] * (uldst.bz - ulst.bz);
  } else if (spd[3] > T(0.0)) {
    flxi[IDN] = fr.d + spd[4] * (urst.d - ur.d) + spd[3] * (urdst.d - urst.d);
    flxi[IVX] = fr.mx + spd[4] * (urst.mx - ur.mx) + spd[3] * (urdst.mx - urst.mx);
    flxi[IVY] = fr.my + spd[4] * (urst.my - ur.my) + spd[3] * (urdst.my - urst.my);
    flxi[IVZ] = fr.mz + spd[4] * (urst.mz - ur.mz) + spd[3] * (urdst.mz - urst.mz);
    flxi[IEN] = fr.e + spd[4] * (urst.e - ur.e) + spd[3] * (urdst.e - urst.e);
    flxi[IBY] = fr.by + spd[4] * (urst.by - ur.by) + spd[3] * (urdst.by - urst.by);
    flxi[IBZ] = fr.bz + spd[4] * (urst.bz - ur.bz) + spd[3] * (urdst.bz - urst.bz);
  } else {
    flxi[IDN] = fr.d + spd[4] * (urst.d - ur.d);
    flxi[IVX] = fr.mx + spd[4] * (urst.mx - ur.mx);
    flxi[IVY] = fr.my + spd[4] * (urst.my - ur.my);
    flxi[IVZ] = fr.mz + spd[4] * (urst.mz - ur.mz);
    flxi[IEN] = fr.e + spd[4] * (urst.e - ur.e);
    flxi[IBY] = fr.by + spd[4] * (urst.by - ur.by);
    flxi[IBZ] = fr.bz + spd[4] * (urst.bz - ur.bz);
  }
}

template <typename T>
__device__ __forceinline__ void hlld(T gamma, const T *wli, const T *wri, T bxi, T *flxi) {
  hlld_core<T, false>(gamma, wli, wri, bxi, T(0.0), T(0.0), flxi);
}

// HLLE at one interface, hydro/rsolvers/mhd/hlle_mhd.cpp:56-171: adiabatic, or the
// isothermal branches (ISO; :83-87, :98-106, :140-150, energy flux 0).
template <bool ISO, typename T>
__device__ __forceinline__ void hlle(T gamma, T iso_cs, const T *wli, const T *wri,
                                     T bxi, T *flxi) {
  T gm1 = gamma - T(1.0);
  T wroe[NWAVE], fl[NWAVE], fr[NWAVE];
  T sqrtdl = xsqrt(wli[IDN]);
  T sqrtdr = xsqrt(wri[IDN]);
  T isdlpdr = T(1.0) / (sqrtdl + sqrtdr);
  wroe[IDN] = sqrtdl * sqrtdr;
  wroe[IVX] = (sqrtdl * wli[IVX] + sqrtdr * wri[IVX]) * isdlpdr;
  wroe[IVY] = (sqrtdl * wli[IVY] + sqrtdr * wri[IVY]) * isdlpdr;
  wroe[IVZ] = (sqrtdl * wli[IVZ] + sqrtdr * wri[IVZ]) * isdlpdr;
  wroe[IBY] = (sqrtdr * wli[IBY] + sqrtdl * wri[IBY]) * isdlpdr;
  wroe[IBZ] = (sqrtdr * wli[IBZ] + sqrtdl * wri[IBZ]) * isdlpdr;
  T x = T(0.5) * (SQR(wli[IBY] - wri[IBY]) + SQR(wli[IBZ] - wri[IBZ])) / (SQR(sqrtdl + sqrtdr));
  T y = T(0.5) * (wli[IDN] + wri[IDN]) / wroe[IDN];
  T pbl = T(0.5) * (bxi * bxi + SQR(wli[IBY]) + SQR(wli[IBZ]));
  T pbr = T(0.5) * (bxi * bxi + SQR(wri[IBY]) + SQR(wri[IBZ]));
  T el = T(0.0), er = T(0.0), hroe = T(0.0);
  if (!ISO) {
    el = wli[IPR] / gm1 + T(0.5) * wli[IDN] * (SQR(wli[IVX]) + SQR(wli[IVY]) + SQR(wli[IVZ])) + pbl;
    er = wri[IPR] / gm1 + T(0.5) * wri[IDN] * (SQR(wri[IVX]) + SQR(wri[IVY]) + SQR(wri[IVZ])) + pbr;
    hroe = ((el + wli[IPR] + pbl) / sqrtdl + (er + wri[IPR] + pbr) / sqrtdr) * isdlpdr;
  }
  T cl = fast_speed_eos<ISO>(gamma, iso_cs, wli, bxi);
  T cr = fast_speed_eos<ISO>(gamma, iso_cs, wri, bxi);
  T btsq = SQR(wroe[IBY]) + SQR(wroe[IBZ]);
  T vaxsq = bxi * bxi / wroe[IDN];
  T bt_starsq, twid_asq;
  if (!ISO) {
    bt_starsq = (gm1 - (gm1 - T(1.0)) * y) * btsq;
    T hp = hroe - (vaxsq + btsq / wroe[IDN]);
    T vsq = SQR(wroe[IVX]) + SQR(wroe[IVY]) + SQR(wroe[IVZ]);
    twid_asq = dmax((gm1 * (hp - T(0.5) * vsq) - (gm1 - T(1.0)) * x), T(0.0));
  } else {
    bt_starsq = btsq * y;
    twid_asq = iso_cs * iso_cs + x;
  }
  T ct2 = bt_starsq / wroe[IDN];
  T tsum = vaxsq + ct2 + twid_asq;
  T tdif = vaxsq + ct2 - twid_asq;
  T cf2_cs2 = xsqrt(tdif * tdif + T(4.0) * twid_asq * ct2);
  T cfsq = T(0.5) * (tsum + cf2_cs2);
  T a = xsqrt(cfsq);
  T al = dmin((wroe[IVX] - a), (wli[IVX] - cl));
  T ar = dmax((wroe[IVX] + a), (wri[IVX] + cr));
  T bp = ar > T(0.0) ? ar : T(0.0);
  T bm = al < T(0.0) ? al : T(0.0);
  T vxl = wli[IVX] - bm;
  T vxr = wri[IVX] - bp;
  fl[IDN] = wli[IDN] * vxl;
  fr[IDN] = wri[IDN] * vxr;
  fl[IVX] = wli[IDN] * wli[IVX] * vxl + pbl - SQR(bxi);
  fr[IVX] = wri[IDN] * wri[IVX] * vxr + pbr - SQR(bxi);
  fl[IVY] = wli[IDN] * wli[IVY] * vxl - bxi * wli[IBY];
  fr[IVY] = wri[IDN] * wri[IVY] * vxr - bxi * wri[IBY];
  fl[IVZ] = wli[IDN] * wli[IVZ] * vxl - bxi * wli[IBZ];
  fr[IVZ] = wri[IDN] * wri[IVZ] * vxr - bxi * wri[IBZ];
  if (!ISO) {
    fl[IVX] += wli[IPR];
    fr[IVX] += wri[IPR];
    fl[IEN] = el * vxl + wli[IVX] * (wli[IPR] + pbl - bxi * bxi);
    fr[IEN] = er * vxr + wri[IVX] * (wri[IPR] + pbr - bxi * bxi);
    fl[IEN] -= bxi * (wli[IBY] * wli[IVY] + wli[IBZ] * wli[IVZ]);
    fr[IEN] -= bxi * (wri[IBY] * wri[IVY] + wri[IBZ] * wri[IVZ]);
  } else {
    fl[IVX] += (iso_cs * iso_cs) * wli[IDN];
    fr[IVX] += (iso_cs * iso_cs) * wri[IDN];
    fl[IEN] = T(0.0);
    fr[IEN] = T(0.0);
  }
  fl[IBY] = wli[IBY] * vxl - bxi * wli[IVY];
  fr[IBY] = wri[IBY] * vxr - bxi * wri[IVY];
  fl[IBZ] = wli[IBZ] * vxl - bxi * wli[IVZ];
  fr[IBZ] = wri[IBZ] * vxr - bxi * wri[IVZ];
  T tmp = T(0.0);
  if (bp != bm) tmp = T(0.5) * (bp + bm) / (bp - bm);
#pragma unroll
  for (int n = 0; n < NWAVE; ++n) flxi[n] = T(0.5) * (fl[n] + fr[n]) + (fl[n] - fr[n]) * tmp;
}

// HLLD for isothermal MHD (Mignone 2007) at one interface,
// hydro/rsolvers/mhd/hlld_iso.cpp:37-254; SMALL_NUMBER = 1e-8 (hlld_iso.cpp:31).
template <typename T>
struct ConsIso {
  T d, mx, my, mz, by, bz;
};

template <typename T>
__device__ __forceinline__ void hlld_iso(T cs, T dfloor, const T *wli, const T *wri, T bxi,
                                         T *flxi) {
  const T SMALL = T(1.0e-8);
  ConsIso<T> ul, ur, ulst, urst, ucst, fl, fr;
  T spd[5];
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
  T cfl = fast_speed_iso(cs, wli, bxi);
  T cfr = fast_speed_iso(cs, wri, bxi);
  spd[0] = dmin(wli[IVX] - cfl, wri[IVX] - cfr);
  spd[4] = dmax(wli[IVX] + cfl, wri[IVX] + cfr);
  T bxsq = bxi * bxi;
  T ptl = SQR(cs) * wli[IDN] + T(0.5) * (bxsq + SQR(wli[IBY]) + SQR(wli[IBZ]));
  T ptr = SQR(cs) * wri[IDN] + T(0.5) * (bxsq + SQR(wri[IBY]) + SQR(wri[IBZ]));
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
  // HLL averages, Mignone eqns. (15), (17), (29)
  T idspd = T(1.0) / (spd[4] - spd[0]);
  T dhll = (spd[4] * ur.d - spd[0] * ul.d - fr.d + fl.d) * idspd;
  dhll = dmax(dhll, dfloor);
  T sqrtdhll = xsqrt(dhll);
  T fdhll = (spd[4] * fl.d - spd[0] * fr.d + spd[4] * spd[0] * (ur.d - ul.d)) * idspd;
  T fmxhll = (spd[4] * fl.mx - spd[0] * fr.mx + spd[4] * spd[0] * (ur.mx - ul.mx)) * idspd;
  T ustar = fdhll / dhll;
  T mxhll = (spd[4] * ur.mx - spd[0] * ul.mx - fr.mx + fl.mx) * idspd;
  spd[1] = ustar - xabs(bxi) / sqrtdhll;
  spd[3] = ustar + xabs(bxi) / sqrtdhll;
  // Ul*, Ur*: eqns. (20), (24), (30)-(33)
  ulst.d = dhll;
  ulst.mx = mxhll;
  T tmp = (spd[0] - spd[1]) * (spd[0] - spd[3]);
  if (xabs(spd[0] - spd[1]) < (SMALL) * cs) {
    ulst.my = ul.my; ulst.mz = ul.mz; ulst.by = ul.by; ulst.bz = ul.bz;
  } else {
    T mfact = bxi * (ustar - wli[IVX]) / tmp;
    T bfact = (ul.d * SQR(spd[0] - wli[IVX]) - bxsq) / (dhll * tmp);
    ulst.my = dhll * wli[IVY] - ul.by * mfact;
    ulst.mz = dhll * wli[IVZ] - ul.bz * mfact;
    ulst.by = ul.by * bfact;
    ulst.bz = ul.bz * bfact;
  }
  urst.d = dhll;
  urst.mx = mxhll;
  tmp = (spd[4] - spd[1]) * (spd[4] - spd[3]);
  if (xabs(spd[4] - spd[3]) < (SMALL) * cs) {
    urst.my = ur.my; urst.mz = ur.mz; urst.by = ur.by; urst.bz = ur.bz;
  } else {
    T mfact = bxi * (ustar - wri[IVX]) / tmp;
    T bfact = (ur.d * SQR(spd[4] - wri[IVX]) - bxsq) / (dhll * tmp);
    urst.my = dhll * wri[IVY] - ur.by * mfact;
    urst.mz = dhll * wri[IVZ] - ur.bz * mfact;
    urst.by = ur.by * bfact;
    urst.bz = ur.bz * bfact;
  }
  // Uc*, eqns. (34)-(37)
  T x = sqrtdhll * (bxi > T(0.0) ? T(1.0) : T(-1.0));
  ucst.d = dhll;
  ucst.mx = mxhll;
  ucst.my = T(0.5) * (ulst.my + urst.my + (urst.by - ulst.by) * x);
  ucst.mz = T(0.5) * (ulst.mz + urst.mz + (urst.bz - ulst.bz) * x);
  ucst.by = T(0.5) * (ulst.by + urst.by + (urst.my - ulst.my) / x);
  ucst.bz = T(0.5) * (ulst.bz + urst.bz + (urst.mz - ulst.mz) / x);
  // flux selection, eqns. (38a)-(38e)
  if (spd[0] >= T(0.0)) {
    flxi[IDN] = fl.d; flxi[IVX] = fl.mx; flxi[IVY] = fl.my; flxi[IVZ] = fl.mz;
    flxi[IBY] = fl.by; flxi[IBZ] = fl.bz;
  } else if (spd[4] <= T(0.0)) {
    flxi[IDN] = fr.d; flxi[IVX] = fr.mx; flxi[IVY] = fr.my; flxi[IVZ] = fr.mz;
    flxi[IBY] = fr.by; flxi[IBZ] = fr.bz;
  } else if (spd[1] >= T(0.0)) {
    flxi[IDN] = fl.d + spd[0] * (ulst.d - ul.d);
    flxi[IVX] = fl.mx + spd[0] * (ulst.mx - ul.mx);
    flxi[IVY] = fl.my + spd[0] * (ulst.my - ul.my);
    flxi[IVZ] = fl.mz + spd[0] * (ulst.mz - ul.mz);
    flxi[IBY] = fl.by + spd[0] * (ulst.by - ul.by);
    flxi[IBZ] = fl.bz + spd[0] * (ulst.bz - ul.bz);
  } else if (spd[3] <= T(0.0)) {
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
  flxi[IEN] = T(0.0);
}

// SIGN, defs.hpp.in:117
template <typename T>
__device__ __forceinline__ T sgn(T x) { return (x < T(0.0)) ? -T(1.0) : T(1.0); }

// PiecewiseParabolicX1..3 for one variable of one cell on a uniform Cartesian mesh
// (reconstruct/ppm.cpp:58-338): CW84 fourth-order interface values with c1..c4 = 1/2,
// c5 = 1/6, c6 = -1/6 (reconstruction.cpp:429-438), the CS08 extremum-preserving
// limiter with C2 = 1.25 and the 1e-12 round-off guards.  q[0..4] = values at i-2..i+2.
// GUARD (FP32 only): the limiter runs on the stencil minus its centre value (exact where
// neighbours are within a factor 2, as in smooth or nearly uniform fields), so FP32
// rounding scales with the stencil's variation rather than its magnitude; the centre
// value is added back at the end.  The limiter differs from itself under translation only
// in the 1e-12 guard on d2qf, which keeps the magnitudes of the values themselves.  *near
// is set when one of the discontinuous tests -- the sign of (dph - q_im1)(qc - dph) at
// either interface, of dqf_minus dqf_plus, and the d2qf tests (its sign and the 1e-12
// guard) where they are used -- has an operand within NEAR_TIE x FP32 rounding of the
// largest |offset value|, so FP32 arithmetic cannot be trusted to decide it as FP64 does
// on the same inputs.  A constant stencil is exact (every operand is 0 in both).
constexpr float NEAR_TIE = 64.0f * 5.9604645e-8f;   // 64 x 2^-24
template <typename T, bool GUARD = false>
__device__ __forceinline__ void ppm_cell(const T *q, T &qminus, T &qplus, bool *near = nullptr) {
  const T C2 = T(1.25);
  const T c1 = T(0.5), c2 = T(0.5), c3 = T(0.5), c4 = T(0.5), c5 = T(1.0) / T(6.0), c6 = -T(1.0) / T(6.0);
  T q_im2 = q[0], q_im1 = q[1], qc = q[2], q_ip1 = q[3], q_ip2 = q[4];
  const T ma = dmax(xabs(q_im1), xabs(q_im2));
  const T mb = dmax(dmax(xabs(qc), xabs(q_ip1)), xabs(q_ip2));
  const T t12 = T(1.0e-12) * dmax(ma, mb);   // round-off guard on d2qf
  T ref = T(0.0);
  if constexpr (GUARD) {
    ref = qc;
    q_im2 -= ref;
    q_im1 -= ref;
    q_ip1 -= ref;
    q_ip2 -= ref;
    qc = T(0.0);
  }
  T qa = (qc - q_im1);
  T qb = (q_ip1 - qc);
  T dd_im1 = c1 * qa + c2 * (q_im1 - q_im2);
  T dd = c1 * qb + c2 * qa;
  T dd_ip1 = c1 * (q_ip2 - q_ip1) + c2 * qb;
  T dph = (c3 * q_im1 + c4 * qc) + (c5 * dd_im1 + c6 * dd);
  T dph_ip1 = (c3 * qc + c4 * q_ip1) + (c5 * dd + c6 * dd_ip1);
  T d2qc_im1 = q_im2 + qc - T(2.0) * q_im1;
  T d2qc = q_im1 + q_ip1 - T(2.0) * qc;
  T tie = T(0.0);   // GUARD: operands below this decide no discontinuous test
  bool vary = false;   // GUARD: the stencil is not constant
  if constexpr (GUARD) {
    const T mq = dmax(dmax(xabs(q_im2), xabs(q_im1)), dmax(xabs(q_ip1), xabs(q_ip2)));
    vary = mq > T(0.0);
    tie = T(NEAR_TIE) * mq;
    // below 1e-12 the products of the sign tests can underflow in FP32: decide in FP64
    *near = vary && (mq < T(1.0e-12) || xabs(dph - q_im1) <= tie || xabs(qc - dph) <= tie ||
                     xabs(dph_ip1 - qc) <= tie || xabs(q_ip1 - dph_ip1) <= tie);
  }
  T d2qc_ip1 = qc + q_ip2 - T(2.0) * q_ip1;
  {
    T qa_tmp = dph - q_im1;
    T qb_tmp = qc - dph;
    T a = T(3.0) * (q_im1 + qc - T(2.0) * dph);
    T qd = T(0.0);
    if (sgn(a) == sgn(d2qc_im1) && sgn(a) == sgn(d2qc))
      qd = sgn(a) * dmin(C2 * xabs(d2qc_im1), dmin(C2 * xabs(d2qc), xabs(a)));
    T dph_tmp = T(0.5) * (q_im1 + qc) - qd / T(6.0);
    if (qa_tmp * qb_tmp < T(0.0)) dph = dph_tmp;
  }
  {
    T qa_tmp = dph_ip1 - qc;
    T qb_tmp = q_ip1 - dph_ip1;
    T a = T(3.0) * (qc + q_ip1 - T(2.0) * dph_ip1);
    T qd = T(0.0);
    if (sgn(a) == sgn(d2qc) && sgn(a) == sgn(d2qc_ip1))
      qd = sgn(a) * dmin(C2 * xabs(d2qc), dmin(C2 * xabs(d2qc_ip1), xabs(a)));
    T dphip1_tmp = T(0.5) * (qc + q_ip1) - qd / T(6.0);
    if (qa_tmp * qb_tmp < T(0.0)) dph_ip1 = dphip1_tmp;
  }
  T d2qf = T(6.0) * (dph + dph_ip1 - T(2.0) * qc);
  T qm = dph, qp = dph_ip1;
  T dqf_minus = qc - qm;
  T dqf_plus = qp - qc;
  T qa_tmp = dqf_minus * dqf_plus;
  T qb_tmp = (q_ip1 - qc) * (qc - q_im1);
  if constexpr (GUARD) {   // (q_ip1 - qc), (qc - q_im1): exact signs from the inputs themselves
    const bool ext = qa_tmp <= T(0.0) || qb_tmp <= T(0.0);
    *near = *near || (vary && (xabs(dqf_minus) <= tie || xabs(dqf_plus) <= tie ||
                               (qb_tmp == T(0.0) && q_ip1 != qc && qc != q_im1) ||
                               (ext && (xabs(d2qf) <= tie ||
                                        xabs(xabs(d2qf) - t12) <= tie + T(9.5367432e-7) * t12))));
  }
  T qe = T(0.0);
  if (sgn(d2qc_im1) == sgn(d2qc) && sgn(d2qc_im1) == sgn(d2qc_ip1) && sgn(d2qc_im1) == sgn(d2qf))
    qe = sgn(d2qf) * dmin(dmin(C2 * xabs(d2qc_im1), C2 * xabs(d2qc)), dmin(C2 * xabs(d2qc_ip1), xabs(d2qf)));
  T rho = T(0.0);
  if (xabs(d2qf) > t12) rho = qe / d2qf;
  T tmp_m = qc - rho * dqf_minus;
  T tmp_p = qc + rho * dqf_plus;
  T tmp2_m = qc - T(2.0) * dqf_plus;
  T tmp2_p = qc + T(2.0) * dqf_minus;
  if ((qa_tmp <= T(0.0) || qb_tmp <= T(0.0))) {
    if (rho <= (T(1.0) - (T(1.0e-12)))) {
      qm = tmp_m;
      qp = tmp_p;
    }
  } else {
    if (xabs(dqf_minus) >= T(2.0) * xabs(dqf_plus)) qm = tmp2_m;
    if (xabs(dqf_plus) >= T(2.0) * xabs(dqf_minus)) qp = tmp2_p;
  }
  if constexpr (GUARD) {
    qm += ref;
    qp += ref;
  }
  qminus = qm;
  qplus = qp;
}

// Eigensystem of the primitive-variable MHD Jacobian at one cell state, shared by
// Reconstruction::LeftEigenmatrixDotVector and RightEigenmatrixDotVector
// (reconstruct/characteristic.cpp:36-262, 280-500; Stone et al. 2008, ApJS 178, 137,
// eqs. A10, A16, A17).  w: the cell's state in the solver's local frame (IVX along the
// sweep, IBY/IBZ the transverse field); b1: its cell-centred normal field.  asq is
// gamma P / rho, or iso_cs^2 for the isothermal EOS.
template <typename T>
struct CharEigen {
  T d, id, sqrtd, isqrtd, cf, cs, a, asq, bet2, bet3, alpha_f, alpha_s, s;
};

template <typename T, bool ISO>
__device__ __forceinline__ void char_eigen(T gamma, T iso_cs, T b1, const T *w, CharEigen<T> &e) {
  const T iso_cs2 = iso_cs * iso_cs;
  e.d = w[IDN];
  e.id = T(1.0) / w[IDN];
  e.sqrtd = xsqrt(w[IDN]);
  e.isqrtd = T(1.0) / e.sqrtd;
  T btsq = SQR(w[IBY]) + SQR(w[IBZ]);
  T bxsq = b1 * b1;
  T gamp = ISO ? iso_cs2 * w[IDN] : gamma * w[IPR];
  T tdif = bxsq + btsq - gamp;   // fast and slow speeds, eq. A10
  T cf2_cs2 = xsqrt(tdif * tdif + T(4.0) * gamp * btsq);
  T cfsq = T(0.5) * (bxsq + btsq + gamp + cf2_cs2);
  T cssq = gamp * bxsq / cfsq;
  cfsq *= e.id;
  e.cf = xsqrt(cfsq);
  cssq *= e.id;
  e.cs = xsqrt(cssq);
  if (ISO) {
    e.asq = iso_cs2;
    e.a = iso_cs;
  } else {
    e.asq = gamp * e.id;
    e.a = xsqrt(e.asq);
  }
  T bt = xsqrt(btsq);   // transverse unit vector, eq. A17 (zero when B_t = 0)
  e.bet2 = T(0.0);
  e.bet3 = T(0.0);
  if (bt != T(0.0)) {
    e.bet2 = w[IBY] / bt;
    e.bet3 = w[IBZ] / bt;
  }
  const T asq = e.asq;   // normalisations, eq. A16
  if ((cfsq - cssq) <= T(0.0)) { e.alpha_f = T(1.0); e.alpha_s = T(0.0); }
  else if ((asq - cssq) <= T(0.0)) { e.alpha_f = T(0.0); e.alpha_s = T(1.0); }
  else if ((cfsq - asq) <= T(0.0)) { e.alpha_f = T(1.0); e.alpha_s = T(0.0); }
  else {
    e.alpha_f = xsqrt((asq - cssq) / (cfsq - cssq));
    e.alpha_s = xsqrt((cfsq - asq) / (cfsq - cssq));
  }
  e.s = sgn(b1);
}

// LeftEigenmatrixDotVector in the local frame: primitive vector v to its characteristic
// fields ch (eq. A18, characteristic.cpp:85-124: fast-, Alfven-, slow-, entropy, slow+,
// Alfven+, fast+; isothermal eq. A22, :166-197: the same without the entropy field).
template <typename T, bool ISO>
__device__ __forceinline__ void char_left(const CharEigen<T> &e, const T *v, T *ch) {
  T vt = e.bet2 * v[IVY] + e.bet3 * v[IVZ];
  T bt = e.bet2 * v[IBY] + e.bet3 * v[IBZ];
  T alf_m = T(0.5) * (e.bet2 * (v[IBZ] * e.s * e.isqrtd + v[IVZ]) - e.bet3 * (v[IBY] * e.s * e.isqrtd + v[IVY]));
  T alf_p = T(0.5) * (e.bet2 * (v[IBZ] * e.s * e.isqrtd - v[IVZ]) - e.bet3 * (v[IBY] * e.s * e.isqrtd - v[IVY]));
  T af_prime = T(0.5) * e.alpha_f / (e.a * e.sqrtd);
  T as_prime = T(0.5) * e.alpha_s / (e.a * e.sqrtd);
  if (!ISO) {
    T nf = T(0.5) / e.asq;
    T qf = nf * e.cf * e.alpha_f * e.s;
    T qs = nf * e.cs * e.alpha_s * e.s;
    T pr = v[IPR] * e.id;
    ch[0] = nf * e.alpha_f * (pr - e.cf * v[IVX]) + qs * vt + as_prime * bt;
    ch[1] = alf_m;
    ch[2] = nf * e.alpha_s * (pr - e.cs * v[IVX]) - qf * vt - af_prime * bt;
    ch[3] = v[IDN] - v[IPR] / e.asq;
    ch[4] = nf * e.alpha_s * (pr + e.cs * v[IVX]) + qf * vt - af_prime * bt;
    ch[5] = alf_p;
    ch[6] = nf * e.alpha_f * (pr + e.cf * v[IVX]) - qs * vt + as_prime * bt;
  } else {
    const T c2 = e.asq;
    T qf = T(0.5) * (e.cf * e.alpha_f * e.s) / c2;
    T qs = T(0.5) * (e.cs * e.alpha_s * e.s) / c2;
    T dn = v[IDN] * e.id;
    ch[0] = T(0.5) * e.alpha_f * (dn - e.cf * v[IVX] / c2) + qs * vt + as_prime * bt;
    ch[1] = alf_m;
    ch[2] = T(0.5) * e.alpha_s * (dn - e.cs * v[IVX] / c2) - qf * vt - af_prime * bt;
    ch[3] = T(0.5) * e.alpha_s * (dn + e.cs * v[IVX] / c2) + qf * vt - af_prime * bt;
    ch[4] = alf_p;
    ch[5] = T(0.5) * e.alpha_f * (dn + e.cf * v[IVX] / c2) - qs * vt + as_prime * bt;
    ch[6] = T(0.0);
  }
}

// RightEigenmatrixDotVector in the local frame: characteristic fields ch back to the
// primitive vector v (eq. A12, characteristic.cpp:330-352; isothermal :400-419, whose
// state has no pressure: that slot is set to 0).
template <typename T, bool ISO>
__device__ __forceinline__ void char_right(const CharEigen<T> &e, const T *ch, T *v) {
  constexpr int fm = 0, am = 1, sm = 2;   // wave pairs: fast, Alfven, slow
  constexpr int sp = ISO ? 3 : 4, ap = ISO ? 4 : 5, fp = ISO ? 5 : 6;
  T qf = e.cf * e.alpha_f * e.s;
  T qs = e.cs * e.alpha_s * e.s;
  T af = e.a * e.alpha_f * e.sqrtd;
  T as = e.a * e.alpha_s * e.sqrtd;
  T compr = e.alpha_f * (ch[fm] + ch[fp]) + e.alpha_s * (ch[sm] + ch[sp]);
  T tr = qs * (ch[fm] - ch[fp]) + qf * (ch[sp] - ch[sm]);
  T bc = as * (ch[fm] + ch[fp]) - af * (ch[sm] + ch[sp]);
  v[IDN] = e.d * compr;
  if (!ISO) v[IDN] += ch[3];
  v[IVX] = e.cf * e.alpha_f * (ch[fp] - ch[fm]) + e.cs * e.alpha_s * (ch[sp] - ch[sm]);
  v[IVY] = e.bet2 * tr + e.bet3 * (ch[ap] - ch[am]);
  v[IVZ] = e.bet3 * tr + e.bet2 * (ch[am] - ch[ap]);
  v[IPR] = ISO ? T(0.0) : e.d * e.asq * compr;
  v[IBY] = e.bet2 * bc - e.bet3 * e.s * e.sqrtd * (ch[ap] + ch[am]);
  v[IBZ] = e.bet3 * bc + e.bet2 * e.s * e.sqrtd * (ch[ap] + ch[am]);
}

// Reconstruction orders of face_flux beyond the reference's xorder 1..3: PLM and PPM
// limited in characteristic variables (xorder "2c" and "3c", reconstruction.cpp:53-66).
constexpr int ORDER_PLM_CHAR = 4, ORDER_PPM_CHAR = 5;

// PPM of one cell as the flux calculation uses it.  FP64: ppm_cell.  FP32 (mixed form):
// ppm_cell in FP32, re-run in FP64 on the same (FP32) inputs wherever a discontinuous
// limiter test is a near-tie, so every branch is taken as FP64 arithmetic takes it on
// those inputs and the FP32 rounding only moves the values continuously.
__device__ __forceinline__ void ppm_limited(const double *q, double &qminus, double &qplus) {
  ppm_cell(q, qminus, qplus);
}
__device__ __forceinline__ void ppm_limited(const float *q, float &qminus, float &qplus) {
  bool near;
  ppm_cell<float, true>(q, qminus, qplus, &near);
  if (near) {
    const double qd[5] = {q[0], q[1], q[2], q[3], q[4]};
    double m, p;
    ppm_cell(qd, m, p);
    qminus = float(m);
    qplus = float(p);
  }
}

// The PPM limiter of face_flux.  FLAG (FP32 only): the guarded FP32 limiter alone, its
// near-tie flag OR-ed into near, and no FP64 re-run in line -- the caller recomputes a
// flagged face with ppm_limited in a second pass (k_flux_fix), which gives the same
// values as ppm_limited in line while keeping the FP64 code out of the main flux kernel.
template <bool FLAG, typename T>
__device__ __forceinline__ void ppm_face(const T *q, T &qminus, T &qplus, bool &near) {
  if constexpr (FLAG) {
    bool n;
    ppm_cell<T, true>(q, qminus, qplus, &n);
    near = near || n;
  } else {
    ppm_limited(q, qminus, qplus);
  }
}

// Roe solver for adiabatic MHD at one interface, hydro/rsolvers/mhd/roe_mhd.cpp:56-212
// with RoeFlux (:245-498): Roe-averaged eigensystem, wave-strength dissipation, LLF
// fallback for the hydro variables when an intermediate density is negative.
template <typename T>
__device__ __forceinline__ void roe(T gamma, const T *wli, const T *wri,
                                    T bxi, T *flxi) {
  T gm1 = gamma - T(1.0);
  T wroe[NWAVE], fl[NWAVE], fr[NWAVE], du[NWAVE], ev[NWAVE];
  T sqrtdl = xsqrt(wli[IDN]);
  T sqrtdr = xsqrt(wri[IDN]);
  T isdlpdr = T(1.0) / (sqrtdl + sqrtdr);
  wroe[IDN] = sqrtdl * sqrtdr;
  wroe[IVX] = (sqrtdl * wli[IVX] + sqrtdr * wri[IVX]) * isdlpdr;
  wroe[IVY] = (sqrtdl * wli[IVY] + sqrtdr * wri[IVY]) * isdlpdr;
  wroe[IVZ] = (sqrtdl * wli[IVZ] + sqrtdr * wri[IVZ]) * isdlpdr;
  wroe[IBY] = (sqrtdr * wli[IBY] + sqrtdl * wri[IBY]) * isdlpdr;
  wroe[IBZ] = (sqrtdr * wli[IBZ] + sqrtdl * wri[IBZ]) * isdlpdr;
  T x = T(0.5) * (SQR(wli[IBY] - wri[IBY]) + SQR(wli[IBZ] - wri[IBZ])) / (SQR(sqrtdl + sqrtdr));
  T y = T(0.5) * (wli[IDN] + wri[IDN]) / wroe[IDN];
  T pbl = T(0.5) * (bxi * bxi + SQR(wli[IBY]) + SQR(wli[IBZ]));
  T pbr = T(0.5) * (bxi * bxi + SQR(wri[IBY]) + SQR(wri[IBZ]));
  T el = wli[IPR] / gm1 + T(0.5) * wli[IDN] * (SQR(wli[IVX]) + SQR(wli[IVY]) + SQR(wli[IVZ])) + pbl;
  T er = wri[IPR] / gm1 + T(0.5) * wri[IDN] * (SQR(wri[IVX]) + SQR(wri[IVY]) + SQR(wri[IVZ])) + pbr;
  wroe[IPR] = ((el + wli[IPR] + pbl) / sqrtdl + (er + wri[IPR] + pbr) / sqrtdr) * isdlpdr;
  T mxl = wli[IDN] * wli[IVX];
  T mxr = wri[IDN] * wri[IVX];
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
#pragma unroll
  for (int n = 0; n < NWAVE; ++n) flxi[n] = T(0.5) * (fl[n] + fr[n]);
  int llf_flag = 0;
  {
    /* RoeFlux, adiabatic branch (roe_mhd.cpp:245-498) */
    const T b1 = bxi;
    T d = wroe[IDN], v1 = wroe[IVX], v2 = wroe[IVY], v3 = wroe[IVZ];
    T b2 = wroe[IBY], b3 = wroe[IBZ];
    T di = T(1.0) / d;
    T btsq = b2 * b2 + b3 * b3;
    T vaxsq = b1 * b1 * di;
    T vsq = v1 * v1 + v2 * v2 + v3 * v3;
    T hp = wroe[IPR] - (vaxsq + btsq * di);
    T bt_starsq = (gm1 - (gm1 - T(1.0)) * y) * btsq;
    T twid_csq = dmax((gm1 * (hp - T(0.5) * vsq) - (gm1 - T(1.0)) * x), T(1.0e-20));
    T ct2 = bt_starsq * di;
    T tsum = vaxsq + ct2 + twid_csq;
    T tdif = vaxsq + ct2 - twid_csq;
    T cf2_cs2 = xsqrt(tdif * tdif + T(4.0) * twid_csq * ct2);
    T cfsq = T(0.5) * (tsum + cf2_cs2);
    T cf = xsqrt(cfsq);
    T cssq = twid_csq * vaxsq / cfsq;
    T cs = xsqrt(cssq);
    T bt = xsqrt(btsq);
    T bt_star = xsqrt(bt_starsq);
    T bet2 = T(0.0), bet3 = T(0.0);
    if (bt != T(0.0)) { bet2 = b2 / bt; bet3 = b3 / bt; }
    T bet2_star = bet2 / xsqrt(gm1 - (gm1 - T(1.0)) * y);
    T bet3_star = bet3 / xsqrt(gm1 - (gm1 - T(1.0)) * y);
    T bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
    T vbet = v2 * bet2_star + v3 * bet3_star;
    T q2_star = T(0.0), q3_star = T(0.0);
    if (bet_starsq != T(0.0)) { q2_star = bet2_star / bet_starsq; q3_star = bet3_star / bet_starsq; }
    T alpha_f, alpha_s;
    if ((cfsq - cssq) <= T(0.0)) { alpha_f = T(1.0); alpha_s = T(0.0); }
    else if ((twid_csq - cssq) <= T(0.0)) { alpha_f = T(0.0); alpha_s = T(1.0); }
    else if ((cfsq - twid_csq) <= T(0.0)) { alpha_f = T(1.0); alpha_s = T(0.0); }
    else {
      alpha_f = xsqrt((twid_csq - cssq) / (cfsq - cssq));
      alpha_s = xsqrt((cfsq - twid_csq) / (cfsq - cssq));
    }
    T sqrtd = xsqrt(d);
    T isqrtd = T(1.0) / sqrtd;
    T s = (b1 < T(0.0)) ? -T(1.0) : T(1.0);
    T twid_c = xsqrt(twid_csq);
    T qf = cf * alpha_f * s;
    T qs = cs * alpha_s * s;
    T af_prime = twid_c * alpha_f * isqrtd;
    T as_prime = twid_c * alpha_s * isqrtd;
    T afpbb = af_prime * bt_star * bet_starsq;
    T aspbb = as_prime * bt_star * bet_starsq;
    T vqstr = (v2 * q2_star + v3 * q3_star);
    T vax = xsqrt(vaxsq);
    T norm = T(0.5) / twid_csq;
    T cff = norm * alpha_f * cf;
    T css = norm * alpha_s * cs;
    T qf_hat = qf * norm;
    T qs_hat = qs * norm;
    T af = norm * af_prime * d;
    T as = norm * as_prime * d;
    T afpb = norm * af_prime * bt_star;
    T aspb = norm * as_prime * bt_star;
    ev[0] = v1 - cf; ev[1] = v1 - vax; ev[2] = v1 - cs; ev[3] = v1;
    ev[4] = v1 + cs; ev[5] = v1 + vax; ev[6] = v1 + cf;
    T a[NWAVE];
    T alpha_f_bar = alpha_f * gm1 * norm;
    T alpha_s_bar = alpha_s * gm1 * norm;
    T gm1a = gm1 / twid_csq;
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
    a[1] *= T(0.5);
    a[2] = du[0] * (alpha_s_bar * (vsq - hp) + css * (cs + v1) + qf_hat * vqstr + afpb);
    a[2] -= du[1] * (alpha_s_bar * v1 + css);
    a[2] -= du[2] * (alpha_s_bar * v2 + qf_hat * q2_star);
    a[2] -= du[3] * (alpha_s_bar * v3 + qf_hat * q3_star);
    a[2] += du[4] * alpha_s_bar;
    a[2] -= du[5] * (af * q2_star + alpha_s_bar * b2);
    a[2] -= du[6] * (af * q3_star + alpha_s_bar * b3);
    a[3] = du[0] * (T(1.0) - gm1a * (T(0.5) * vsq - (gm1 - T(1.0)) * x / gm1));
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
    a[5] *= T(0.5);
    a[6] = du[0] * (alpha_f_bar * (vsq - hp) + cff * (cf - v1) + qs_hat * vqstr - aspb);
    a[6] -= du[1] * (alpha_f_bar * v1 - cff);
    a[6] -= du[2] * (alpha_f_bar * v2 + qs_hat * q2_star);
    a[6] -= du[3] * (alpha_f_bar * v3 + qs_hat * q3_star);
    a[6] += du[4] * alpha_f_bar;
    a[6] += du[5] * (as * q2_star - alpha_f_bar * b2);
    a[6] += du[6] * (as * q3_star - alpha_f_bar * b3);
    T coeff[NWAVE];
    for (int n = 0; n < NWAVE; ++n) coeff[n] = -T(0.5) * xabs(ev[n]) * a[n];
    T dens = wli[IDN] + a[0] * alpha_f;
    if (dens < T(0.0)) llf_flag = 1;
    dens += a[2] * alpha_s;
    if (dens < T(0.0)) llf_flag = 1;
    dens += a[3];
    if (dens < T(0.0)) llf_flag = 1;
    dens += a[4] * alpha_s;
    if (dens < T(0.0)) llf_flag = 1;
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
    flxi[4] += coeff[3] * (T(0.5) * vsq + (gm1 - T(1.0)) * x / gm1);
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
  if (ev[0] >= T(0.0)) {
    for (int n = 0; n < NWAVE; ++n) flxi[n] = fl[n];
  }
  if (ev[NWAVE - 1] <= T(0.0)) {
    for (int n = 0; n < NWAVE; ++n) flxi[n] = fr[n];
  }
  if (llf_flag != 0) {
    T cfl = fast_speed(gamma, wli, bxi);
    T cfr = fast_speed(gamma, wri, bxi);
    T a = T(0.5) * dmax((xabs(wli[IVX]) + cfl), (xabs(wri[IVX]) + cfr));
    flxi[IDN] = T(0.5) * (fl[IDN] + fr[IDN]) - a * du[IDN];
    flxi[IVX] = T(0.5) * (fl[IVX] + fr[IVX]) - a * du[IVX];
    flxi[IVY] = T(0.5) * (fl[IVY] + fr[IVY]) - a * du[IVY];
    flxi[IVZ] = T(0.5) * (fl[IVZ] + fr[IVZ]) - a * du[IVZ];
    flxi[IEN] = T(0.5) * (fl[IEN] + fr[IEN]) - a * du[IEN];
  }
}

// ================================================================ kernels
// One face of Hydro::CalculateFluxes along direction dir (hydro/calculate_fluxes.cpp:
// 56-321): reconstruction from the cells on both sides (ORDER 1: reconstruct/dc.cpp:
// 24-104; ORDER 2: the uniform-mesh van Leer branch of reconstruct/plm.cpp:27-351;
// ORDER 3: reconstruct/ppm.cpp:58-338), the Riemann solver, the face EMFs and the CT
// upwind weight.  The states are assembled directly in the solver's local frame (IVX the
// velocity along dir, IBY/IBZ the transverse bcc components of plm.cpp:46-56,160-170,
// 273-283); each variable is reconstructed on its own, so this equals reconstructing in
// the mesh frame and permuting.  dir may be a compile-time constant (k_flux) or a
// run-time value (the fused stage kernel, one code path for every direction).
// wb, bb: w and bcc of the block; id: index of the cell above the face; st: stride
// along dir; pc: the block's PLM factor rows for dir; cidx: the cell index along dir.
// Outputs: f[0..4] = mass, momentum along dir, dir+1, dir+2, energy flux; ey = -F(By),
// ez = F(Bz) (hlld.cpp:371-377); wct = GetWeightForCT.
template <int ORDER, int RS, typename T, bool ISO, bool FLAG = false>
__device__ __forceinline__ void face_flux(int dir, const T *__restrict__ wb, const T *__restrict__ bb,
                                          int64_t S, int64_t id, int64_t st, int64_t st1, int64_t st2,
                                          const double *__restrict__ pc, int PL, int cidx, T bxi, T gamma, T cs,
                                          T dfl, T pfl, T dx, T dt, T *f, T &ey, T &ez, T &wct,
                                          bool *nearp = nullptr) {
  const int d1 = (dir == 2) ? 0 : dir + 1, d2 = (dir == 0) ? 2 : dir - 1;   // (dir+1)%3, (dir+2)%3
  bool near = false;   // FLAG: a PPM limiter test of this face was a near-tie (ppm_face)
  T wli[NWAVE], wri[NWAVE];
  if constexpr (ORDER >= ORDER_PLM_CHAR) {
    // limiting in characteristic variables (plm.cpp:59-63,96-99,109-136; ppm.cpp:
    // 99-107,313-316): each cell's stencil is projected with the left eigenvectors of that
    // cell's own state, each field limited on its own, and the result projected back
    // with its right eigenvectors.  NP cells c-NP/2 .. c+NP/2-1 around the face.
    constexpr int NP = (ORDER == ORDER_PPM_CHAR) ? 6 : 4, H = NP / 2;
    constexpr int NWC = ISO ? 6 : 7;   // characteristic fields
    T q[NP][NWAVE];
    const T *src[NWAVE] = {wb, wb + (int64_t)(IVX + dir) * S, wb + (int64_t)(IVX + d1) * S,
                           wb + (int64_t)(IVX + d2) * S, wb + (int64_t)IPR * S, bb + (int64_t)d1 * S,
                           bb + (int64_t)d2 * S};
#pragma unroll
    for (int t = 0; t < NP; ++t)
#pragma unroll
      for (int n = 0; n < NWAVE; ++n) q[t][n] = (ISO && n == IPR) ? T(0.0) : src[n][id + (t - H) * st];
    const T *bn = bb + (int64_t)dir * S + id;
#pragma unroll
    for (int side = 0; side < 2; ++side) {   // 0: cell c-1 (left state), 1: cell c (right)
      const int cc = H - 1 + side;             // the cell's index in q
      CharEigen<T> e;
      char_eigen<T, ISO>(gamma, cs, bn[(side - 1) * st], q[cc], e);
      T res[NWAVE], lim[NWAVE];
      if constexpr (ORDER == ORDER_PPM_CHAR) {
        T ch[5][NWAVE];
#pragma unroll
        for (int t = 0; t < 5; ++t) char_left<T, ISO>(e, q[cc - 2 + t], ch[t]);
#pragma unroll
        for (int n = 0; n < NWC; ++n) {
          T c5[5] = {ch[0][n], ch[1][n], ch[2][n], ch[3][n], ch[4][n]}, qm, qp;
          ppm_face<FLAG>(c5, qm, qp, near);
          lim[n] = side ? qm : qp;
        }
        if (NWC < NWAVE) lim[NWAVE - 1] = T(0.0);
        char_right<T, ISO>(e, lim, side ? wri : wli);
      } else {
        T dl[NWAVE], dr[NWAVE], cl[NWAVE], cr[NWAVE];
#pragma unroll
        for (int n = 0; n < NWAVE; ++n) {
          dl[n] = q[cc][n] - q[cc - 1][n];
          dr[n] = q[cc + 1][n] - q[cc][n];
        }
        char_left<T, ISO>(e, dl, cl);
        char_left<T, ISO>(e, dr, cr);
#pragma unroll
        for (int n = 0; n < NWC; ++n) {
          T dw2 = cl[n] * cr[n];
          lim[n] = T(2.0) * dw2 / (cl[n] + cr[n]);
          if (dw2 <= T(0.0)) lim[n] = T(0.0);
        }
        if (NWC < NWAVE) lim[NWAVE - 1] = T(0.0);
        char_right<T, ISO>(e, lim, res);
        if (side == 0) {
          const T f = T(pc[cidx - 1]);
#pragma unroll
          for (int n = 0; n < NWAVE; ++n) wli[n] = q[cc][n] + f * res[n];
        } else {
          const T f = T(pc[PL + cidx]);
#pragma unroll
          for (int n = 0; n < NWAVE; ++n) wri[n] = q[cc][n] - f * res[n];
        }
      }
    }
  } else {
#pragma unroll
  for (int n = 0; n < NWAVE; ++n) {
    if (ISO && n == IPR) {   // no pressure in the isothermal state vector
      wli[n] = wri[n] = T(0.0);
      continue;
    }
    const T *q;
    if (n == IDN) q = wb + id;
    else if (n == IVX) q = wb + (int64_t)(IVX + dir) * S + id;
    else if (n == IVY) q = wb + (int64_t)(IVX + d1) * S + id;
    else if (n == IVZ) q = wb + (int64_t)(IVX + d2) * S + id;
    else if (n == IPR) q = wb + (int64_t)IPR * S + id;
    else if (n == IBY) q = bb + (int64_t)d1 * S + id;
    else q = bb + (int64_t)d2 * S + id;
    if (ORDER == 1) {
      wli[n] = q[-st];
      wri[n] = q[0];
    } else if (ORDER == 3) {
      T q6[6];
#pragma unroll
      for (int t = 0; t < 6; ++t) q6[t] = q[(t - 3) * st];
      T m0, p0, m1, p1;
      ppm_face<FLAG>(q6, m0, p0, near);       // cell c-1: values at c-3..c+1
      ppm_face<FLAG>(q6 + 1, m1, p1, near);   // cell c:   values at c-2..c+2
      wli[n] = p0;
      wri[n] = m1;
    } else {
      T qm2 = q[-2 * st], qm1 = q[-st], q0 = q[0], qp1 = q[st];
      // left state: upper face of cell c-1
      T dwl = (qm1 - qm2), dwr = (q0 - qm1);
      T dw2 = dwl * dwr;
      T dwm = T(2.0) * dw2 / (dwl + dwr);
      if (dw2 <= T(0.0)) dwm = T(0.0);
      wli[n] = qm1 + T(pc[cidx - 1]) * dwm;
      // right state: lower face of cell c
      dwl = (q0 - qm1);
      dwr = (qp1 - q0);
      dw2 = dwl * dwr;
      dwm = T(2.0) * dw2 / (dwl + dwr);
      if (dw2 <= T(0.0)) dwm = T(0.0);
      wri[n] = q0 - T(pc[PL + cidx]) * dwm;
    }
  }
  }
  if (ORDER >= 3) {
    // ApplyPrimitiveFloors on both reconstructed states (ppm.cpp:333-337, and after a
    // characteristic PLM, plm.cpp:128-136); the isothermal EOS floors the density only
    // (eos/isothermal_mhd.cpp:135-140)
    wli[IDN] = (wli[IDN] > dfl) ? wli[IDN] : dfl;
    wri[IDN] = (wri[IDN] > dfl) ? wri[IDN] : dfl;
    if (!ISO) {
      wli[IPR] = (wli[IPR] > pfl) ? wli[IPR] : pfl;
      wri[IPR] = (wri[IPR] > pfl) ? wri[IPR] : pfl;
    }
  }
  T flxi[NWAVE];
  if (RS == AGX_RSOLVER_HLLE) hlle<ISO>(gamma, cs, wli, wri, bxi, flxi);
  else if (ISO) hlld_iso(cs, dfl, wli, wri, bxi, flxi);
  else if (RS == AGX_RSOLVER_ROE) roe(gamma, wli, wri, bxi, flxi);
  else if (RS == AGX_RSOLVER_LHLLD) {
    // Hydro::CalculateVelocityDifferences at this face (calculate_velocity_differences.cpp:
    // 23-68): the normal velocity jump, and the smallest one-sided difference of each
    // transverse velocity along its own direction in the two cells beside the face
    const T *vn = wb + (int64_t)(IVX + dir) * S, *v1 = wb + (int64_t)(IVX + d1) * S;
    const T *v2 = wb + (int64_t)(IVX + d2) * S;
    const int64_t l = id - st;
    T dvn = vn[id] - vn[l];
    T dvl = dmin(v1[l + st1] - v1[l], v1[l] - v1[l - st1]);
    T dvr = dmin(v1[id + st1] - v1[id], v1[id] - v1[id - st1]);
    T dwl = dmin(v2[l + st2] - v2[l], v2[l] - v2[l - st2]);
    T dwr = dmin(v2[id + st2] - v2[id], v2[id] - v2[id - st2]);
    T dvt = dmin(dmin(dvl, dvr), dmin(dwl, dwr));
    hlld_core<T, true>(gamma, wli, wri, bxi, dvn, dvt, flxi);
  } else hlld(gamma, wli, wri, bxi, flxi);
  f[0] = flxi[IDN];
  f[1] = flxi[IVX];
  f[2] = flxi[IVY];
  f[3] = flxi[IVZ];
  f[4] = ISO ? T(0.0) : flxi[IEN];
  ey = -flxi[IBY];
  ez = flxi[IBZ];
  wct = ct_weight(flxi[IDN], wli[IDN], wri[IDN], dx, dt);
  if constexpr (FLAG) *nearp = near;
}

// FP32 PPM faces (the mixed form) are computed in two passes: k_flux with the guarded FP32
// limiter only, flagging each face where a limiter test was a near-tie, then k_flux_fix
// recomputing the flagged faces with the FP64 re-run (ppm_limited).
template <int ORDER, typename T>
constexpr bool flux_flagged() {
  return std::is_same<T, float>::value && (ORDER == 3 || ORDER == ORDER_PPM_CHAR);
}

// One face of Hydro::CalculateFluxes in direction DIR: face_flux at (b, id), written to
// the face arrays; FLAG also writes the face's near-tie flag (0 or 1) to nearf.
template <int DIR, int ORDER, int RS, typename T, bool ISO, bool FLAG>
__device__ __forceinline__ void flux_face(const Geo &g, int b, int64_t id, int cidx, const T *__restrict__ w,
                                          const T *__restrict__ bcc, const double *__restrict__ bn,
                                          T *__restrict__ flx, T *__restrict__ ey, T *__restrict__ ez,
                                          T *__restrict__ wct, const BlockCoef *__restrict__ coef,
                                          const double *__restrict__ plm, double gamma_, double cs_,
                                          const double *__restrict__ dtp, double dfl_, double pfl_,
                                          uint8_t *__restrict__ nearf) {
  const int64_t sd[3] = {1, (int64_t)g.P, (int64_t)g.J * g.P};   // strides along i, j, k
  const double *pc = plm + ((int64_t)b * 3 + DIR) * 2 * g.PL;    // [dxp][dxm] rows
  T f[5], fey, fez, fw;
  bool near;
  face_flux<ORDER, RS, T, ISO, FLAG>(DIR, w + (int64_t)b * NHYDRO * g.S, bcc + (int64_t)b * 3 * g.S, g.S, id,
                                     sd[DIR], sd[(DIR + 1) % 3], sd[(DIR + 2) % 3], pc, g.PL, cidx,
                                     T(bn[(int64_t)b * g.S + id]), T(gamma_), T(cs_), T(dfl_), T(pfl_),
                                     T(coef[b].dx[DIR]), T(*dtp), f, fey, fez, fw, &near);
  constexpr int ivx = IVX + DIR, ivy = IVX + (DIR + 1) % 3, ivz = IVX + (DIR + 2) % 3;
  T *fb = flx + (int64_t)b * NHYDRO * g.S + id;
  fb[IDN * g.S] = f[0];
  fb[ivx * g.S] = f[1];
  fb[ivy * g.S] = f[2];
  fb[ivz * g.S] = f[3];
  if (!ISO) fb[IEN * g.S] = f[4];
  int64_t fo = (int64_t)b * g.S + id;
  ey[fo] = fey;
  ez[fo] = fez;
  wct[fo] = fw;
  if constexpr (FLAG) nearf[fo] = near ? 1 : 0;
}

// Hydro::CalculateFluxes for direction DIR, one thread per face (face_flux).  For FP32 PPM
// nearf (one byte per cell of the block arrays) receives each face's near-tie flag.
template <int DIR, int ORDER, int RS, typename T, bool ISO>
__global__ void __launch_bounds__(256)
k_flux(Geo g, const T *__restrict__ w, const T *__restrict__ bcc,
       const double *__restrict__ bn, T *__restrict__ flx, T *__restrict__ ey,
       T *__restrict__ ez, T *__restrict__ wct, const BlockCoef *__restrict__ coef,
       const double *__restrict__ plm, double gamma_, double cs_, const double *__restrict__ dtp, double dfl_,
       double pfl_, uint8_t *__restrict__ nearf) {
  int il = g.is - (DIR != 0), iu = g.ie + 1;
  int jl = g.js - (DIR != 1), ju = g.je + 1;
  int kl = g.ks - (DIR != 2);
  PlaneIdx p;
  if (!plane_index(il, iu, jl, ju, kl, p, DIR == 2)) return;   // x3: 4-plane stencil
  const int cidx = (DIR == 0) ? p.i : (DIR == 1 ? p.j : p.k);   // index of the cell above the face
  flux_face<DIR, ORDER, RS, T, ISO, flux_flagged<ORDER, T>()>(g, p.b, IDX(g, p.k, p.j, p.i), cidx, w, bcc, bn,
                                                                flx, ey, ez, wct, coef, plm, gamma_, cs_, dtp,
                                                                dfl_, pfl_, nearf);
}

// Second pass of FP32 PPM fluxes: each workgroup scans FIX_SPAN bytes of nearf (16 per
// thread, one 16-byte load), gathers the flagged faces into an LDS list, clears their
// flags (so nearf is all zero between passes) and recomputes those faces with the FP64
// re-run of the limiter, so the near-tie faces occupy whole waves instead of single
// lanes of every wave of k_flux.  nflag (the length of nearf) is a multiple of FIX_SPAN.
constexpr int FIX_PER = 16, FIX_SPAN = 256 * FIX_PER;
template <int DIR, int ORDER, int RS, bool ISO>
__global__ void __launch_bounds__(256)
k_flux_fix(Geo g, const float *__restrict__ w, const float *__restrict__ bcc,
           const double *__restrict__ bn, float *__restrict__ flx, float *__restrict__ ey,
           float *__restrict__ ez, float *__restrict__ wct, const BlockCoef *__restrict__ coef,
           const double *__restrict__ plm, double gamma_, double cs_, const double *__restrict__ dtp,
           double dfl_, double pfl_, uint8_t *__restrict__ nearf, int64_t nflag) {
  __shared__ int list[FIX_SPAN];
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * FIX_SPAN;
  const int64_t off = base + (int64_t)threadIdx.x * FIX_PER;
  if (off < nflag) {
    uint4 v = *reinterpret_cast<const uint4 *>(nearf + off);
    if (v.x | v.y | v.z | v.w) {
      const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int t = 0; t < FIX_PER; ++t)
        if ((wv[t >> 2] >> (8 * (t & 3))) & 0xffu) list[atomicAdd(&cnt, 1)] = threadIdx.x * FIX_PER + t;
      *reinterpret_cast<uint4 *>(nearf + off) = make_uint4(0u, 0u, 0u, 0u);
    }
  }
  __syncthreads();
  const int n = cnt;
  const int il = g.is - (DIR != 0), jl = g.js - (DIR != 1), kl = g.ks - (DIR != 2);
  for (int e = threadIdx.x; e < n; e += 256) {
    const int64_t fo = base + list[e];
    const int b = (int)(fo / g.S);
    const int64_t id = fo - (int64_t)b * g.S;
    const int k = (int)(id / ((int64_t)g.J * g.P));
    const int r = (int)(id - (int64_t)k * g.J * g.P);
    const int j = r / g.P, i = r - j * g.P;
    // only faces of this direction are ever flagged; the test keeps the stencil in range
    if (b >= g.nb || i < il || i > g.ie + 1 || j < jl || j > g.je + 1 || k < kl || k > g.ke + 1) continue;
    const int cidx = (DIR == 0) ? i : (DIR == 1 ? j : k);
    flux_face<DIR, ORDER, RS, float, ISO, false>(g, b, id, cidx, w, bcc, bn, flx, ey, ez, wct, coef, plm, gamma_,
                                                 cs_, dtp, dfl_, pfl_, nullptr);
  }
}

// Field::ComputeCornerE, 3D Newtonian branch (field/calculate_corner_e.cpp:130-229).
// The cell-centred E = -v x B (:176-183) is recomputed per edge instead of stored.
__device__ __forceinline__ double cce_comp(const double *wb, const double *bb, int64_t S,
                                           int64_t id, int comp) {
  if (comp == 0) return wb[IVZ * S + id] * bb[IB2 * S + id] - wb[IVY * S + id] * bb[IB3 * S + id];
  if (comp == 1) return wb[IVX * S + id] * bb[IB3 * S + id] - wb[IVZ * S + id] * bb[IB1 * S + id];
  return wb[IVY * S + id] * bb[IB1 * S + id] - wb[IVX * S + id] * bb[IB2 * S + id];
}

// GS05 upwinded corner EMFs (calculate_corner_e.cpp:187-229), one edge each, in the
// reference's operand order.  Arguments are named by position relative to the edge
// (k,j,i): c_* are cell-centred E of the four cells around the edge, e?x?_* the face
// EMFs and w?_* the CT weights of the four faces meeting at the edge.
__device__ __forceinline__ double edge_e1(double c_k, double c_kjm, double c_km, double c_kmjm,
                                          double e1x3, double e1x3_jm, double e1x2, double e1x2_km,
                                          double w2, double w2_km, double w3, double w3_jm) {
  double de1_l3 = (1.0 - w2_km) * (e1x3 - c_km) + (w2_km) * (e1x3_jm - c_kmjm);
  double de1_r3 = (1.0 - w2) * (e1x3 - c_k) + (w2) * (e1x3_jm - c_kjm);
  double de1_l2 = (1.0 - w3_jm) * (e1x2 - c_kjm) + (w3_jm) * (e1x2_km - c_kmjm);
  double de1_r2 = (1.0 - w3) * (e1x2 - c_k) + (w3) * (e1x2_km - c_km);
  return 0.25 * (de1_l3 + de1_r3 + de1_l2 + de1_r2 + e1x2_km + e1x2 + e1x3_jm + e1x3);
}

__device__ __forceinline__ double edge_e2(double c_k, double c_kim, double c_km, double c_kmim,
                                          double e2x3, double e2x3_im, double e2x1, double e2x1_km,
                                          double w1, double w1_km, double w3, double w3_im) {
  double de2_l3 = (1.0 - w1_km) * (e2x3 - c_km) + (w1_km) * (e2x3_im - c_kmim);
  double de2_r3 = (1.0 - w1) * (e2x3 - c_k) + (w1) * (e2x3_im - c_kim);
  double de2_l1 = (1.0 - w3_im) * (e2x1 - c_kim) + (w3_im) * (e2x1_km - c_kmim);
  double de2_r1 = (1.0 - w3) * (e2x1 - c_k) + (w3) * (e2x1_km - c_km);
  return 0.25 * (de2_l3 + de2_r3 + de2_l1 + de2_r1 + e2x3_im + e2x3 + e2x1_km + e2x1);
}

__device__ __forceinline__ double edge_e3(double c_k, double c_im, double c_jm, double c_jmim,
                                          double e3x2, double e3x2_im, double e3x1, double e3x1_jm,
                                          double w1, double w1_jm, double w2, double w2_im) {
  double de3_l2 = (1.0 - w1_jm) * (e3x2 - c_jm) + (w1_jm) * (e3x2_im - c_jmim);
  double de3_r2 = (1.0 - w1) * (e3x2 - c_k) + (w1) * (e3x2_im - c_im);
  double de3_l1 = (1.0 - w2_im) * (e3x1 - c_im) + (w2_im) * (e3x1_jm - c_jmim);
  double de3_r1 = (1.0 - w2) * (e3x1 - c_k) + (w2) * (e3x1_jm - c_jm);
  return 0.25 * (de3_l1 + de3_r1 + de3_l2 + de3_r2 + e3x2_im + e3x2 + e3x1_jm + e3x1);
}

__global__ void __launch_bounds__(256)
k_corner_e(Geo g, const double *__restrict__ w, const double *__restrict__ bcc,
           const double *__restrict__ e3x1, const double *__restrict__ e2x1,
           const double *__restrict__ e1x2, const double *__restrict__ e3x2,
           const double *__restrict__ e1x3, const double *__restrict__ e2x3,
           const double *__restrict__ wx1, const double *__restrict__ wx2,
           const double *__restrict__ wx3, double *__restrict__ e1, double *__restrict__ e2,
           double *__restrict__ e3) {
  PlaneIdx p;
  if (!plane_index(g.is, g.ie + 1, g.js, g.je + 1, g.ks, p)) return;
  const int64_t S = g.S;
  const int64_t sj = g.P, sk = (int64_t)g.J * g.P;
  const int64_t id = IDX(g, p.k, p.j, p.i);
  const int64_t fo = (int64_t)p.b * S + id;
  const double *wb = w + (int64_t)p.b * NHYDRO * S;
  const double *bb = bcc + (int64_t)p.b * 3 * S;
  // e1 at (k,j,i): cells (k-1|k, j-1|j, i)
  e1[fo] = edge_e1(cce_comp(wb, bb, S, id, 0), cce_comp(wb, bb, S, id - sj, 0),
                   cce_comp(wb, bb, S, id - sk, 0), cce_comp(wb, bb, S, id - sk - sj, 0),
                   e1x3[fo], e1x3[fo - sj], e1x2[fo], e1x2[fo - sk], wx2[fo], wx2[fo - sk],
                   wx3[fo], wx3[fo - sj]);
  // e2 at (k,j,i): cells (k-1|k, j, i-1|i)
  e2[fo] = edge_e2(cce_comp(wb, bb, S, id, 1), cce_comp(wb, bb, S, id - 1, 1),
                   cce_comp(wb, bb, S, id - sk, 1), cce_comp(wb, bb, S, id - sk - 1, 1),
                   e2x3[fo], e2x3[fo - 1], e2x1[fo], e2x1[fo - sk], wx1[fo], wx1[fo - sk],
                   wx3[fo], wx3[fo - 1]);
  // e3 at (k,j,i): cells (k, j-1|j, i-1|i)
  e3[fo] = edge_e3(cce_comp(wb, bb, S, id, 2), cce_comp(wb, bb, S, id - 1, 2),
                   cce_comp(wb, bb, S, id - sj, 2), cce_comp(wb, bb, S, id - sj - 1, 2),
                   e3x2[fo], e3x2[fo - 1], e3x1[fo], e3x1[fo - sj], wx1[fo], wx1[fo - sj],
                   wx2[fo], wx2[fo - 1]);
}

// Conserved-variable part of one VL2 stage for the interior cells
// (time_integrator.cpp:1563-1660): dst = src - wght*div(F)/V (AddFluxDivergence,
// hydro/add_flux_divergence.cpp:50-94, x1 then x2 then x3) followed by the self-gravity
// source with dt = wght (hydro/srcterms/self_gravity.cpp:35-125, x1 then x2 then x3).
// Stage 1: src = u, dst = u1 (the WeightedAve copy u1 = 0 + 1.0*u is exact);
// stage 2: src = dst = u1 (in place).  The host swaps u and u1 afterwards.
__global__ void __launch_bounds__(256)
k_integrate_u(Geo g, const double *src, double *dst, const double *__restrict__ f1,
              const double *__restrict__ f2, const double *__restrict__ f3,
              const double *__restrict__ w, const double *__restrict__ phi,
              const BlockCoef *__restrict__ coef, double wght, int grav, int iso) {
  PlaneIdx p;
  if (!plane_index(g.is, g.ie, g.js, g.je, g.ks, p)) return;
  const int64_t S = g.S;
  const int64_t sj = g.P, sk = (int64_t)g.J * g.P;
  const int64_t id = IDX(g, p.k, p.j, p.i);
  const double dx1 = coef[p.b].dx[0], dx2 = coef[p.b].dx[1], dx3 = coef[p.b].dx[2];
  const double a1 = dx2 * dx3, a2 = dx1 * dx3, a3 = dx1 * dx2;
  const double vol = dx1 * dx2 * dx3;
  const int64_t cb = (int64_t)p.b * NHYDRO * S + id;
  double un[NHYDRO];
#pragma unroll
  for (int n = 0; n < NHYDRO; ++n) {
    const int64_t o = cb + n * S;
    double dflx = (a1 * f1[o + 1] - a1 * f1[o]);
    dflx += (a2 * f2[o + sj] - a2 * f2[o]);
    dflx += (a3 * f3[o + sk] - a3 * f3[o]);
    un[n] = src[o] - wght * dflx / vol;
  }
  if (grav) {
    const int64_t pb = (int64_t)p.b * S + id;
    const double rho = w[cb];
    const double phic = phi[pb];
    {
      double dtodx1 = wght / dx1;
      double phil = 0.5 * (phi[pb - 1] + phic);
      double phir = 0.5 * (phic + phi[pb + 1]);
      un[IVX] -= dtodx1 * rho * (phir - phil);
      if (!iso) un[IEN] -= dtodx1 * (f1[cb] * (phic - phil) + f1[cb + 1] * (phir - phic));
    }
    {
      double dtodx2 = wght / dx2;
      double phil = 0.5 * (phi[pb - sj] + phic);
      double phir = 0.5 * (phic + phi[pb + sj]);
      un[IVY] -= dtodx2 * rho * (phir - phil);
      if (!iso) un[IEN] -= dtodx2 * (f2[cb] * (phic - phil) + f2[cb + sj] * (phir - phic));
    }
    {
      double dtodx3 = wght / dx3;
      double phil = 0.5 * (phi[pb - sk] + phic);
      double phir = 0.5 * (phic + phi[pb + sk]);
      un[IVZ] -= dtodx3 * rho * (phir - phil);
      if (!iso) un[IEN] -= dtodx3 * (f3[cb] * (phic - phil) + f3[cb + sk] * (phir - phic));
    }
  }
#pragma unroll
  for (int n = 0; n < NHYDRO; ++n) dst[cb + n * S] = un[n];
}

// Field::CT for one face component (field/ct.cpp:39-114), interior faces; src/dst as in
// k_integrate_u.  Uniform mesh: Face?Area and Edge?Length are the block's dx products.
template <int DIR>
__global__ void __launch_bounds__(256)
k_ct(Geo g, const double *src, double *dst, const double *__restrict__ e1,
     const double *__restrict__ e2, const double *__restrict__ e3,
     const BlockCoef *__restrict__ coef, double wght) {
  PlaneIdx p;
  if (!plane_index(g.is, g.ie + (DIR == 0), g.js, g.je + (DIR == 1), g.ks, p)) return;
  const int64_t sj = g.P, sk = (int64_t)g.J * g.P;
  const int64_t fo = (int64_t)p.b * g.S + IDX(g, p.k, p.j, p.i);
  const double dx1 = coef[p.b].dx[0], dx2 = coef[p.b].dx[1], dx3 = coef[p.b].dx[2];
  double b = src[fo];
  if (DIR == 0) {
    double area = dx2 * dx3;
    b -= (wght / area) * (dx3 * e3[fo + sj] - dx3 * e3[fo]);
    b += (wght / area) * (dx2 * e2[fo + sk] - dx2 * e2[fo]);
  } else if (DIR == 1) {
    double area = dx1 * dx3;
    b += (wght / area) * (dx3 * e3[fo + 1] - dx3 * e3[fo]);
    b -= (wght / area) * (dx1 * e1[fo + sk] - dx1 * e1[fo]);
  } else {
    double area = dx1 * dx2;
    b -= (wght / area) * (dx2 * e2[fo + 1] - dx2 * e2[fo]);
    b += (wght / area) * (dx1 * e1[fo + sj] - dx1 * e1[fo]);
  }
  dst[fo] = b;
}

// Same-level ghost exchange of a cell-centred array (bvals/cc/bvals_cc.cpp:201-336):
// every ghost cell copies the neighbour block's interior cell at the same location.
// One launch over all cells of all blocks; interior threads exit.
__device__ __forceinline__ int nbr_block(const Geo &g, int b, int o1, int o2, int o3) {
  int lx1 = b % g.nb1, lx2 = (b / g.nb1) % g.nb2, lx3 = b / (g.nb1 * g.nb2);
  lx1 += o1; lx1 += (lx1 < 0) ? g.nb1 : 0; lx1 -= (lx1 >= g.nb1) ? g.nb1 : 0;
  lx2 += o2; lx2 += (lx2 < 0) ? g.nb2 : 0; lx2 -= (lx2 >= g.nb2) ? g.nb2 : 0;
  lx3 += o3; lx3 += (lx3 < 0) ? g.nb3 : 0; lx3 -= (lx3 >= g.nb3) ? g.nb3 : 0;
  return (lx3 * g.nb2 + lx2) * g.nb1 + lx1;
}

__global__ void __launch_bounds__(256)
k_bvals_cc(Geo g, double *a, int nvar) {
  PlaneIdx p;
  if (!plane_index(0, g.n1 - 1, 0, g.n2 - 1, 0, p)) return;
  int o1 = (p.i < g.is) ? -1 : (p.i > g.ie ? 1 : 0);
  int o2 = (p.j < g.js) ? -1 : (p.j > g.je ? 1 : 0);
  int o3 = (p.k < g.ks) ? -1 : (p.k > g.ke ? 1 : 0);
  if (!(o1 | o2 | o3)) return;
  int s = nbr_block(g, p.b, o1, o2, o3);
  int64_t di = IDX(g, p.k, p.j, p.i);
  int64_t si = IDX(g, p.k - o3 * g.bx3, p.j - o2 * g.bx2, p.i - o1 * g.bx1);
  for (int n = 0; n < nvar; ++n)
    a[((int64_t)p.b * nvar + n) * g.S + di] = a[((int64_t)s * nvar + n) * g.S + si];
}

// Same-level face-field ghost exchange (bvals/fc/bvals_fc.cpp:344-402 pack,
// :583-700 unpack).  Along its own direction a face component's "interior" range is
// is..ie+1: the face shared with the neighbour is kept by both blocks, not exchanged.
template <int DIR>
__global__ void __launch_bounds__(256)
k_bvals_fc(Geo g, double *a) {
  PlaneIdx p;
  if (!plane_index(0, g.n1 - 1 + (DIR == 0), 0, g.n2 - 1 + (DIR == 1), 0, p)) return;
  int o1 = (p.i < g.is) ? -1 : (p.i > g.ie + (DIR == 0) ? 1 : 0);
  int o2 = (p.j < g.js) ? -1 : (p.j > g.je + (DIR == 1) ? 1 : 0);
  int o3 = (p.k < g.ks) ? -1 : (p.k > g.ke + (DIR == 2) ? 1 : 0);
  if (!(o1 | o2 | o3)) return;
  int s = nbr_block(g, p.b, o1, o2, o3);
  a[(int64_t)p.b * g.S + IDX(g, p.k, p.j, p.i)] =
      a[(int64_t)s * g.S + IDX(g, p.k - o3 * g.bx3, p.j - o2 * g.bx2, p.i - o1 * g.bx1)];
}

// c2p_cell, c2p_cell_iso, c2p_eos: agx_internal.hpp (shared with the refinement operators).

// Field::CalculateCellCenteredField (uniform branch, field/field.cpp:112-180) and
// EquationOfState::ConservedToPrimitive (eos/adiabatic_mhd.cpp:41-90, isothermal_mhd.cpp:
// 40-73) over interior + ghosts; the floors are written back to u.  TW is the storage
// type of w and bcc (double, or float in the mixed-precision form).
template <typename TW, bool ISO>
__global__ void __launch_bounds__(256)
k_cons2prim(Geo g, double *__restrict__ u, TW *__restrict__ w, TW *__restrict__ bcc,
            const double *__restrict__ b1, const double *__restrict__ b2,
            const double *__restrict__ b3, double gm1, double dfl, double pfl) {
  PlaneIdx p;
  if (!plane_index(0, g.n1 - 1, 0, g.n2 - 1, 0, p)) return;
  const int64_t S = g.S;
  const int64_t id = IDX(g, p.k, p.j, p.i);
  const int64_t fo = (int64_t)p.b * S + id;
  const double bcc1 = 0.5 * b1[fo] + 0.5 * b1[fo + 1];
  const double bcc2 = 0.5 * b2[fo] + 0.5 * b2[fo + g.P];
  const double bcc3 = 0.5 * b3[fo] + 0.5 * b3[fo + (int64_t)g.J * g.P];
  const int64_t bo = (int64_t)p.b * 3 * S + id;
  bcc[bo] = TW(bcc1);
  bcc[bo + S] = TW(bcc2);
  bcc[bo + 2 * S] = TW(bcc3);
  const int64_t c = (int64_t)p.b * NHYDRO * S + id;
  double u_d = u[c], u_e = ISO ? 0.0 : u[c + 4 * S], wv[5];
  c2p_eos<ISO>(u_d, u[c + S], u[c + 2 * S], u[c + 3 * S], u_e, bcc1, bcc2, bcc3, gm1, dfl, pfl, wv);
  u[c] = u_d;
  if (!ISO) u[c + 4 * S] = u_e;
  constexpr int NH = ISO ? 4 : NHYDRO;
#pragma unroll
  for (int n = 0; n < NH; ++n) w[c + n * S] = TW(wv[n]);
}

// Hydro::NewBlockTimeStep per cell (hydro/new_blockdt.cpp:78-97): min over the three
// directions of dx/(|v|+c_f) with bx = bcc + |b_face - bcc| (lower face).
template <bool ISO>
__device__ __forceinline__ double cell_dt(double gamma, double cs, const double *wv, double c1,
                                          double c2, double c3, double f1, double f2, double f3,
                                          const double *dx) {
  double wi[NWAVE];
  wi[IDN] = wv[0]; wi[IVX] = wv[1]; wi[IVY] = wv[2]; wi[IVZ] = wv[3];
  wi[IPR] = ISO ? 0.0 : wv[4];
  double bx = c1 + fabs(f1 - c1);
  wi[IBY] = c2; wi[IBZ] = c3;
  double cf = fast_speed_eos<ISO>(gamma, cs, wi, bx);
  double dt1 = dx[0] / (fabs(wi[IVX]) + cf);
  wi[IBY] = c3; wi[IBZ] = c1;
  bx = c2 + fabs(f2 - c2);
  cf = fast_speed_eos<ISO>(gamma, cs, wi, bx);
  double dt2 = dx[1] / (fabs(wi[IVY]) + cf);
  wi[IBY] = c1; wi[IBZ] = c2;
  bx = c3 + fabs(f3 - c3);
  cf = fast_speed_eos<ISO>(gamma, cs, wi, bx);
  double dt3 = dx[2] / (fabs(wi[IVZ]) + cf);
  double m = DBL_MAX;
  m = dmin(m, dt1);
  m = dmin(m, dt2);
  m = dmin(m, dt3);
  return m;
}

// Minimum over a workgroup of NW wavefronts (exact, so the order is free): wavefront
// shuffle, then LDS across the wavefronts; every thread returns the result.
template <int NW>
__device__ __forceinline__ double wg_min(double m, double *sh) {
  for (int off = 32; off > 0; off >>= 1) m = dmin(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) r = dmin(r, sh[w]);
  return r;
}
__device__ __forceinline__ double wg_min256(double m, double *sh) { return wg_min<4>(m, sh); }

// ================================================================ fused stage kernel
// One VL2 stage after the fluxes: corner EMFs (field/calculate_corner_e.cpp:130-229),
// CT (field/ct.cpp:39-114), flux divergence and self-gravity source
// (hydro/add_flux_divergence.cpp:50-94, hydro/srcterms/self_gravity.cpp:35-125), bcc
// (field/field.cpp:112-180) and C2P (eos/adiabatic_mhd.cpp:41-90) of the interior cells,
// and in the last stage the per-cell time-step bound (hydro/new_blockdt.cpp:57-160),
// in one pass over HBM.  Every value is formed with the same operations as the task-level
// kernels, so results are identical.
//
// A workgroup owns a TI x TJ tile of (i,j) columns of one MeshBlock (thread = column) and
// walks the cell planes upwards.  Iteration c computes the cell-centred E of plane c+1
// and the corner EMFs of edge plane c and plane c into LDS (three planes of cell-centred
// E, two of edges), then, after one barrier, each thread does the CT of the faces of its
// cell in plane c-1 and of its x3 face in face plane c, and updates that cell.  Faces
// shared by two cells are computed by both from identical inputs; fluxes, face EMFs and
// weights stream from HBM once.  Outputs go to registers other than the inputs (u may
// be in place: each cell reads only its own src), so tiles that share halo values never
// see each other's writes.
// Ghost copies written by their source (same-level periodic exchange, bvals_cc.cpp:201-336
// and bvals_fc.cpp:344-402,583-700, run "push" instead of "pull"): index x of an array
// whose own range along an axis is [lo, hi] (cells: [is, ie]; the normal axis of a face
// array: [is, ie+1], the shared face being kept by both blocks) is a ghost of the +1
// neighbour at x - bx when that lands in [lo - ng, lo - 1], and of the -1 neighbour at
// x + bx when that lands in [hi + 1, hi + ng].  Returns the offset range [omin, omax].
__device__ __forceinline__ void push_range(int x, int lo, int hi, int bx, int ng, int &omin,
                                           int &omax) {
  omax = (x - bx >= lo - ng && x - bx <= lo - 1) ? 1 : 0;
  omin = (x + bx >= hi + 1 && x + bx <= hi + ng) ? -1 : 0;
}

// Calls f(dest_block, dest_index) for every neighbour ghost that copies element (k,j,i)
// of block b; r[0..5] = (o1min, o1max, o2min, o2max, o3min, o3max).
template <typename F>
__device__ __forceinline__ void for_each_ghost(const Geo &g, int b, int k, int j, int i,
                                               const int *r, F f) {
  if (!(r[0] | r[1] | r[2] | r[3] | r[4] | r[5])) return;
  for (int o3 = r[4]; o3 <= r[5]; ++o3)
    for (int o2 = r[2]; o2 <= r[3]; ++o2)
      for (int o1 = r[0]; o1 <= r[1]; ++o1) {
        if (!(o1 | o2 | o3)) continue;
        f(nbr_block(g, b, o1, o2, o3), IDX(g, k - o3 * g.bx3, j - o2 * g.bx2, i - o1 * g.bx1));
      }
}

template <typename TW>
struct StageArgs {
  const double *src;                 // conserved state the stage starts from
  double *dst;                       // updated conserved state (interior)
  const double *rho;                 // conserved state whose density feeds self-gravity
  const double *phi;
  const TW *f1, *f2, *f3;            // fluxes
  const TW *e3x1, *e2x1, *e1x2, *e3x2, *e1x3, *e2x3;   // face EMFs
  const TW *wx1, *wx2, *wx3;         // CT upwind weights
  const TW *w, *bcc;                 // primitives / bcc the fluxes were computed from
  TW *wn, *bccn;                     // new primitives / bcc (interior)
  const double *s1, *s2, *s3;        // CT source faces
  double *d1, *d2, *d3;              // CT destination faces (interior)
  const BlockCoef *coef;
  const double *dtp;                 // device time step (scal[SC_DT])
  double wfac;                       // stage weight: wght = wfac * dt
  double gm1, dfl, pfl, gamma, cs;
  int grav;
  double *dtpart;                    // per-workgroup min of the dt bound, or null
};

// PUSH: the two copies of every shared face agree (agx_mesh::copies_state), so the
// ghosts receive this cell's floored u, w and bcc (the same bits a ghost C2P would give);
// otherwise they receive the un-floored u and are converted by k_ghost_c2p.
template <int TI, int TJ, typename TW, bool ISO, bool PUSH>
__global__ void __launch_bounds__(TI * TJ, 4)   // >= 4 waves per SIMD: <= 128 VGPRs
k_stage(Geo g, StageArgs<TW> a) {
  static_assert(TI * TJ == 256 || TI * TJ == 512, "k_stage runs 256- or 512-thread workgroups");
  constexpr int NH = ISO ? 4 : NHYDRO;   // conserved variables (no energy if isothermal)
  constexpr int NT = TI * TJ, CI = TI + 2, CJ = TJ + 2;
  __shared__ double s_cce[3][3][CJ][CI];      // planes c-1, c, c+1 (slot = plane mod 3)
  __shared__ double s_e1[2][TJ + 1][TI];      // edge planes c-1, c (slot = plane & 1)
  __shared__ double s_e2[2][TJ][TI + 1];
  __shared__ double s_e3[2][TJ + 1][TI + 1];
  __shared__ double s_red[TI * TJ / 64];
  const int tid = threadIdx.x;
  const int tx = tid % TI, ty = tid / TI;
  const int nti = (g.bx1 + TI - 1) / TI;
  const uint3 wg = xcd_block();   // whole MeshBlocks per XCD: tiles share halo lines in L2
  const int b = wg.y;
  const int i0 = g.is + (int)(wg.x % nti) * TI;
  const int j0 = g.js + (int)(wg.x / nti) * TJ;
  const int ni = min(TI, g.ie + 1 - i0), nj = min(TJ, g.je + 1 - j0);
  const bool own = (tx < ni) && (ty < nj);
  const int i = i0 + tx, j = j0 + ty;
  const int64_t S = g.S, sj = g.P, sk = (int64_t)g.J * g.P;
  const int64_t fb = (int64_t)b * S;
  const int64_t cbase = (int64_t)b * NHYDRO * S, bbase = (int64_t)b * 3 * S;
  const double dx[3] = {a.coef[b].dx[0], a.coef[b].dx[1], a.coef[b].dx[2]};
  const double dx1 = dx[0], dx2 = dx[1], dx3 = dx[2];
  const double wght = a.wfac * *a.dtp;
  const int wi = ni + 2, ncce = wi * (nj + 2);
  const int np = (nj + 1) * (ni + 1);
  // cell-centred E = -v x B at plane c, cells (j0-1..j0+nj) x (i0-1..i0+ni), into slot
  auto cce_plane = [&](int c) {
    const int sl = (c + 3) % 3;
    for (int q = tid; q < ncce; q += NT) {
      const int jj = q / wi, ii = q - jj * wi;
      const int64_t id = IDX(g, c, j0 - 1 + jj, i0 - 1 + ii);
      const TW *wb = a.w + cbase + id;
      const TW *bb = a.bcc + bbase + id;
      const double vx = wb[IVX * S], vy = wb[IVY * S], vz = wb[IVZ * S];
      const double bx = bb[0], by = bb[S], bz = bb[2 * S];
      s_cce[sl][0][jj][ii] = vz * by - vy * bz;
      s_cce[sl][1][jj][ii] = vx * bz - vz * bx;
      s_cce[sl][2][jj][ii] = vy * bx - vx * by;
    }
  };
  cce_plane(g.ks - 1);
  cce_plane(g.ks);
  __syncthreads();
  double b3_lo = 0.0;       // new b3 at the lower face of cell plane c-1
  double dtmin = DBL_MAX;
  for (int c = g.ks; c <= g.ke + 1; ++c) {
    const int cb = c & 1, pb = cb ^ 1;
    const int cc = c % 3, pc = (c + 2) % 3;
    // (A) cell-centred E of plane c+1 (used by the next iteration), then the edges at the
    // points (c, j0+jj, i0+ii), jj in [0,nj], ii in [0,ni]: e1 and e2 of edge plane c
    // (lower edges of cell plane c; e1 for ii < ni, e2 for jj < nj) and e3 of plane c.
    // One thread per point, so weights shared by the three edges load once.
    if (c + 1 <= g.ke + 1) cce_plane(c + 1);
    const bool do3 = (c <= g.ke);
    for (int q = tid; q < np; q += NT) {
      const int jj = q / (ni + 1), ii = q - jj * (ni + 1);
      const bool do1 = ii < ni, do2 = jj < nj;
      const int64_t fo = fb + IDX(g, c, j0 + jj, i0 + ii);
      const double w3 = (do1 || do2) ? (double)a.wx3[fo] : 0.0;
      if (do1) {
        s_e1[cb][jj][ii] = edge_e1(
            s_cce[cc][0][jj + 1][ii + 1], s_cce[cc][0][jj][ii + 1], s_cce[pc][0][jj + 1][ii + 1],
            s_cce[pc][0][jj][ii + 1], (double)a.e1x3[fo], (double)a.e1x3[fo - sj],
            (double)a.e1x2[fo], (double)a.e1x2[fo - sk], (double)a.wx2[fo],
            (double)a.wx2[fo - sk], w3, (double)a.wx3[fo - sj]);
      }
      if (do2) {
        s_e2[cb][jj][ii] = edge_e2(
            s_cce[cc][1][jj + 1][ii + 1], s_cce[cc][1][jj + 1][ii], s_cce[pc][1][jj + 1][ii + 1],
            s_cce[pc][1][jj + 1][ii], (double)a.e2x3[fo], (double)a.e2x3[fo - 1],
            (double)a.e2x1[fo], (double)a.e2x1[fo - sk], (double)a.wx1[fo],
            (double)a.wx1[fo - sk], w3, (double)a.wx3[fo - 1]);
      }
      if (do3) {
        s_e3[cb][jj][ii] = edge_e3(
            s_cce[cc][2][jj + 1][ii + 1], s_cce[cc][2][jj + 1][ii], s_cce[cc][2][jj][ii + 1],
            s_cce[cc][2][jj][ii], (double)a.e3x2[fo], (double)a.e3x2[fo - 1],
            (double)a.e3x1[fo], (double)a.e3x1[fo - sj], (double)a.wx1[fo],
            (double)a.wx1[fo - sj], (double)a.wx2[fo], (double)a.wx2[fo - 1]);
      }
    }
    __syncthreads();
    // (B) own column: CT of the x3 face of face plane c, then for cell (c-1, j, i) the
    // CT of its four x1/x2 faces (the upper ones are the neighbours' lower faces,
    // recomputed here from the same edges), flux divergence, self-gravity, bcc, C2P and
    // the dt bound.
    double b3_hi = 0.0;
    if (own) {
      {
        const int64_t fo = fb + IDX(g, c, j, i);
        double v = a.s3[fo];
        const double area = dx1 * dx2;
        v -= (wght / area) * (dx2 * s_e2[cb][ty][tx + 1] - dx2 * s_e2[cb][ty][tx]);
        v += (wght / area) * (dx1 * s_e1[cb][ty + 1][tx] - dx1 * s_e1[cb][ty][tx]);
        a.d3[fo] = v;
        b3_hi = v;
        int r[6];
        push_range(i, g.is, g.ie, g.bx1, g.ng, r[0], r[1]);
        push_range(j, g.js, g.je, g.bx2, g.ng, r[2], r[3]);
        push_range(c, g.ks, g.ke + 1, g.bx3, g.ng, r[4], r[5]);
        for_each_ghost(g, b, c, j, i, r, [&](int nb, int64_t id2) { a.d3[(int64_t)nb * S + id2] = v; });
      }
      if (c - 1 >= g.ks) {
        const int64_t id = IDX(g, c - 1, j, i);
        const int64_t fo = fb + id;
        const double area1 = dx2 * dx3, area2 = dx1 * dx3;
        double b1l = a.s1[fo], b1h = a.s1[fo + 1], b2l = a.s2[fo], b2h = a.s2[fo + sj];
        b1l -= (wght / area1) * (dx3 * s_e3[pb][ty + 1][tx] - dx3 * s_e3[pb][ty][tx]);
        b1l += (wght / area1) * (dx2 * s_e2[cb][ty][tx] - dx2 * s_e2[pb][ty][tx]);
        b1h -= (wght / area1) * (dx3 * s_e3[pb][ty + 1][tx + 1] - dx3 * s_e3[pb][ty][tx + 1]);
        b1h += (wght / area1) * (dx2 * s_e2[cb][ty][tx + 1] - dx2 * s_e2[pb][ty][tx + 1]);
        b2l += (wght / area2) * (dx3 * s_e3[pb][ty][tx + 1] - dx3 * s_e3[pb][ty][tx]);
        b2l -= (wght / area2) * (dx1 * s_e1[cb][ty][tx] - dx1 * s_e1[pb][ty][tx]);
        b2h += (wght / area2) * (dx3 * s_e3[pb][ty + 1][tx + 1] - dx3 * s_e3[pb][ty + 1][tx]);
        b2h -= (wght / area2) * (dx1 * s_e1[cb][ty + 1][tx] - dx1 * s_e1[pb][ty + 1][tx]);
        a.d1[fo] = b1l;
        if (tx == ni - 1) a.d1[fo + 1] = b1h;
        a.d2[fo] = b2l;
        if (ty == nj - 1) a.d2[fo + sj] = b2h;
        {
          int rc[6], r[6];
          push_range(i, g.is, g.ie, g.bx1, g.ng, rc[0], rc[1]);
          push_range(j, g.js, g.je, g.bx2, g.ng, rc[2], rc[3]);
          push_range(c - 1, g.ks, g.ke, g.bx3, g.ng, rc[4], rc[5]);
          // x1 faces i (and i+1 at the tile's upper edge)
          r[2] = rc[2]; r[3] = rc[3]; r[4] = rc[4]; r[5] = rc[5];
          push_range(i, g.is, g.ie + 1, g.bx1, g.ng, r[0], r[1]);
          for_each_ghost(g, b, c - 1, j, i, r, [&](int nb, int64_t id2) { a.d1[(int64_t)nb * S + id2] = b1l; });
          if (tx == ni - 1) {
            push_range(i + 1, g.is, g.ie + 1, g.bx1, g.ng, r[0], r[1]);
            for_each_ghost(g, b, c - 1, j, i + 1, r,
                           [&](int nb, int64_t id2) { a.d1[(int64_t)nb * S + id2] = b1h; });
          }
          // x2 faces j (and j+1 at the tile's upper edge)
          r[0] = rc[0]; r[1] = rc[1];
          push_range(j, g.js, g.je + 1, g.bx2, g.ng, r[2], r[3]);
          for_each_ghost(g, b, c - 1, j, i, r, [&](int nb, int64_t id2) { a.d2[(int64_t)nb * S + id2] = b2l; });
          if (ty == nj - 1) {
            push_range(j + 1, g.js, g.je + 1, g.bx2, g.ng, r[2], r[3]);
            for_each_ghost(g, b, c - 1, j + 1, i, r,
                           [&](int nb, int64_t id2) { a.d2[(int64_t)nb * S + id2] = b2h; });
          }
          // cell copies are written after the update below, with the same ranges rc
          r[0] = rc[0]; r[1] = rc[1]; r[2] = rc[2]; r[3] = rc[3];
        }
        const double a1 = dx2 * dx3, a2 = dx1 * dx3, a3 = dx1 * dx2;
        const double vol = dx1 * dx2 * dx3;
        const int64_t cb0 = cbase + id;
        double un[NHYDRO];
        un[IEN] = 0.0;
#pragma unroll
        for (int n = 0; n < NH; ++n) {
          const int64_t o = cb0 + n * S;
          double dflx = (a1 * (double)a.f1[o + 1] - a1 * (double)a.f1[o]);
          dflx += (a2 * (double)a.f2[o + sj] - a2 * (double)a.f2[o]);
          dflx += (a3 * (double)a.f3[o + sk] - a3 * (double)a.f3[o]);
          un[n] = a.src[o] - wght * dflx / vol;
        }
        if (a.grav) {
          const double rho = a.rho[cb0];
          const double phic = a.phi[fo];
          {
            double dtodx1 = wght / dx1;
            double phil = 0.5 * (a.phi[fo - 1] + phic);
            double phir = 0.5 * (phic + a.phi[fo + 1]);
            un[IVX] -= dtodx1 * rho * (phir - phil);
            if (!ISO)
              un[IEN] -= dtodx1 * ((double)a.f1[cb0] * (phic - phil) + (double)a.f1[cb0 + 1] * (phir - phic));
          }
          {
            double dtodx2 = wght / dx2;
            double phil = 0.5 * (a.phi[fo - sj] + phic);
            double phir = 0.5 * (phic + a.phi[fo + sj]);
            un[IVY] -= dtodx2 * rho * (phir - phil);
            if (!ISO)
              un[IEN] -= dtodx2 * ((double)a.f2[cb0] * (phic - phil) + (double)a.f2[cb0 + sj] * (phir - phic));
          }
          {
            double dtodx3 = wght / dx3;
            double phil = 0.5 * (a.phi[fo - sk] + phic);
            double phir = 0.5 * (phic + a.phi[fo + sk]);
            un[IVZ] -= dtodx3 * rho * (phir - phil);
            if (!ISO)
              un[IEN] -= dtodx3 * ((double)a.f3[cb0] * (phic - phil) + (double)a.f3[cb0 + sk] * (phir - phic));
          }
        }
        const double bcc1 = 0.5 * b1l + 0.5 * b1h;
        const double bcc2 = 0.5 * b2l + 0.5 * b2h;
        const double bcc3 = 0.5 * b3_lo + 0.5 * b3_hi;
        const int64_t bo = bbase + id;
        a.bccn[bo] = TW(bcc1);
        a.bccn[bo + S] = TW(bcc2);
        a.bccn[bo + 2 * S] = TW(bcc3);
        // ghost copies of the un-floored conserved state (bvals_cc.cpp:201-336 sends u
        // before ConservedToPrimitive; the receiver converts its ghosts itself)
        int rc[6];
        push_range(i, g.is, g.ie, g.bx1, g.ng, rc[0], rc[1]);
        push_range(j, g.js, g.je, g.bx2, g.ng, rc[2], rc[3]);
        push_range(c - 1, g.ks, g.ke, g.bx3, g.ng, rc[4], rc[5]);
        if constexpr (!PUSH)
          for_each_ghost(g, b, c - 1, j, i, rc, [&](int nb, int64_t id2) {
            const int64_t c2 = (int64_t)nb * NHYDRO * S + id2;
#pragma unroll
            for (int n = 0; n < NH; ++n) a.dst[c2 + n * S] = un[n];
          });
        // C2P writes its floors back into u (adiabatic_mhd.cpp:66-84)
        double wv[5];
        c2p_eos<ISO>(un[IDN], un[IVX], un[IVY], un[IVZ], un[IEN], bcc1, bcc2, bcc3, a.gm1, a.dfl, a.pfl,
                     wv);
#pragma unroll
        for (int n = 0; n < NH; ++n) a.dst[cb0 + n * S] = un[n];
#pragma unroll
        for (int n = 0; n < NH; ++n) a.wn[cb0 + n * S] = TW(wv[n]);
        // PUSH: a ghost's bcc and its C2P are this cell's, bit for bit
        if constexpr (PUSH)
          for_each_ghost(g, b, c - 1, j, i, rc, [&](int nb, int64_t id2) {
            const int64_t c2 = (int64_t)nb * NHYDRO * S + id2, b2 = (int64_t)nb * 3 * S + id2;
#pragma unroll
            for (int n = 0; n < NH; ++n) {
              a.dst[c2 + n * S] = un[n];
              a.wn[c2 + n * S] = TW(wv[n]);
            }
            a.bccn[b2] = TW(bcc1);
            a.bccn[b2 + S] = TW(bcc2);
            a.bccn[b2 + 2 * S] = TW(bcc3);
          });
        if (a.dtpart)
          dtmin = dmin(dtmin, cell_dt<ISO>(a.gamma, a.cs, wv, bcc1, bcc2, bcc3, b1l, b2l, b3_lo, dx));
      }
    }
    b3_lo = b3_hi;
    __syncthreads();
  }
  if (a.dtpart) {
    double m = wg_min<TI * TJ / 64>(dtmin, s_red);
    if (tid == 0) a.dtpart[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = m;
  }
}

// ================================================================ whole VL2 stage in one kernel
// k_vl2 does everything of one VL2 stage after the ghost exchange: the fluxes of all
// three directions (face_flux: reconstruction, Riemann solver, face EMFs, CT weights),
// the corner EMFs, CT, the flux divergence with the self-gravity source, bcc, C2P with
// floors, the ghost copies of u and the faces, and in the last stage the time-step bound.
// Fluxes, face EMFs and weights live only in LDS and registers: per interior cell it
// reads w, bcc, the faces and u (and phi) and writes u, w, bcc and the faces, so the HBM
// traffic of a stage is that of k_stage minus the 24 flux words per cell, and the flux
// kernels' own traffic is gone.  Every value is formed by the same operations, in the
// same order, as the split path (k_flux + k_stage), so results are identical bit for bit.
//
// A workgroup owns a TI x TJ tile of (i,j) columns of one MeshBlock and walks the cell
// planes c = ks-1 .. ke+1.  Per plane:
//   F  faces of plane c: x1 faces i0..i0+ni of rows j0-1..j0+nj, x2 faces j0..j0+nj of
//      columns i0-1..i0+ni, and x3 faces of face plane c (below cell plane c) of the tile
//      and its one-cell ring without corners; then the cell-centred E of plane c.  Thread
//      t solves the x1, x2 and x3 faces of its own column, and the rest (tile edges and
//      ring, 5 ni + 5 nj + 4 faces) go one per thread.  Face EMFs and weights go to LDS;
//      the hydro fluxes of a thread's own x1 and x3 faces stay in its registers, those of
//      the x2 faces and of the tile's right-edge x1 faces go to LDS.
//   E  corner EMFs: e1, e2 of edge plane c (from faces of planes c-1 and c) and e3 of
//      plane c, one point per thread, as k_stage.
//   U  CT of the own x3 face of face plane c; the update of the own cell of plane c-1
//      (its x1/x2 flux differences were summed in the previous iteration, its x3 fluxes
//      are the lower and upper faces solved in the previous and this iteration); then the
//      x1/x2 flux differences of the own cell of plane c, and the face EMFs, weights and
//      cell-centred E that the next plane's edges need are kept as "previous plane".
// Three barriers per plane.  Thread (tx, ty) reads the x1 flux of face i+1 from lane
// +1 (the same row of the same wavefront, since TI divides 64), the last column from LDS.
template <typename TW>
struct VL2Args {
  const TW *w, *bcc;                 // primitives / bcc of the stage's fluxes
  TW *wn, *bccn;                     // new primitives / bcc (interior)
  const double *bn1, *bn2, *bn3;     // face field whose normal components enter the Riemann problems
  const double *src;                 // conserved state the stage starts from
  double *dst;                       // updated conserved state (interior + pushed ghosts)
  const double *rho;                 // conserved state whose density feeds self-gravity
  const double *phi;
  const double *s1, *s2, *s3;        // CT source faces
  double *d1, *d2, *d3;              // CT destination faces
  const BlockCoef *coef;
  const double *plm;                 // PLM face-offset factors [nb][3][2][PL]
  const double *dtp;                 // device time step
  double wfac;                       // stage weight: wght = wfac * dt
  double gm1, dfl, pfl, gamma, cs;
  int grav;
  double *dtpart;                    // per-workgroup min of the dt bound, or null
};

// fluxes in the mesh frame from the solver's frame of direction dir
template <typename T>
__device__ __forceinline__ void flux_to_mesh(int dir, const T *f, T *out) {
  out[0] = f[0];
  out[4] = f[4];
  if (dir == 0) { out[1] = f[1]; out[2] = f[2]; out[3] = f[3]; }
  else if (dir == 1) { out[1] = f[3]; out[2] = f[1]; out[3] = f[2]; }
  else { out[1] = f[2]; out[2] = f[3]; out[3] = f[1]; }
}

template <int TI, int TJ, int ORDER, int RS, typename TW, bool ISO, bool PUSH>
__global__ void __launch_bounds__(TI * TJ, 2)   // >= 2 waves per SIMD
k_vl2(Geo g, VL2Args<TW> a) {
  static_assert(64 % TI == 0 && TI * TJ == 256, "k_vl2 tiles: 256 threads, rows within a wavefront");
  constexpr int NH = ISO ? 4 : NHYDRO;
  constexpr int NT = TI * TJ;
  // current plane: x1 faces [ey=e3x1, ez=e2x1, w1][row jj+1][face ii]; x2 faces
  // [ey=e1x2, ez=e3x2, w2][face jj][column ii+1]; x3 faces [ey=e2x3, ez=e1x3, w3][jj+1][ii+1]
  __shared__ TW s_f1[3][TJ + 2][TI + 1];
  __shared__ TW s_f2[3][TJ + 1][TI + 2];
  __shared__ TW s_f3[3][TJ + 2][TI + 2];
  // previous plane: x1 (e2x1, w1) of rows [0, nj); x2 (e1x2, w2) of columns [0, ni)
  __shared__ TW s_f1p[2][TJ][TI + 1];
  __shared__ TW s_f2p[2][TJ + 1][TI];
  // cell-centred E of plane c (3 components) and of plane c-1 (component 0 on columns
  // [0, ni), component 1 on rows [0, nj))
  __shared__ double s_cce[3][TJ + 2][TI + 2];
  __shared__ double s_ccp0[TJ + 2][TI];
  __shared__ double s_ccp1[TJ][TI + 2];
  __shared__ double s_e1[2][TJ + 1][TI];      // edge planes c-1, c (slot = plane & 1)
  __shared__ double s_e2[2][TJ][TI + 1];
  __shared__ double s_e3[2][TJ + 1][TI + 1];
  __shared__ TW s_h2[NHYDRO][TJ + 1][TI];     // x2 hydro fluxes (mesh frame) of plane c
  __shared__ TW s_h1e[NHYDRO][TJ];            // x1 hydro fluxes of the tile's right-edge faces
  __shared__ double s_red[NT / 64];
  const int tid = threadIdx.x;
  const int tx = tid % TI, ty = tid / TI;
  const int nti = (g.bx1 + TI - 1) / TI;
  const uint3 wg = xcd_block();   // whole MeshBlocks per XCD: tiles share halo lines in L2
  const int b = wg.y;
  const int i0 = g.is + (int)(wg.x % nti) * TI;
  const int j0 = g.js + (int)(wg.x / nti) * TJ;
  const int ni = min(TI, g.ie + 1 - i0), nj = min(TJ, g.je + 1 - j0);
  const bool own = (tx < ni) && (ty < nj);
  const int i = i0 + tx, j = j0 + ty;
  const int64_t S = g.S, sj = g.P, sk = (int64_t)g.J * g.P;
  const int64_t fb = (int64_t)b * S;
  const int64_t cbase = (int64_t)b * NHYDRO * S, bbase = (int64_t)b * 3 * S;
  const double dx[3] = {a.coef[b].dx[0], a.coef[b].dx[1], a.coef[b].dx[2]};
  const double dx1 = dx[0], dx2 = dx[1], dx3 = dx[2];
  const double dtv = *a.dtp;
  const double wght = a.wfac * dtv;
  const TW gT = TW(a.gamma), csT = TW(a.cs), dflT = TW(a.dfl), pflT = TW(a.pfl), dtT = TW(dtv);
  const TW dxT[3] = {TW(dx1), TW(dx2), TW(dx3)};
  const TW *wb = a.w + cbase;
  const TW *bb = a.bcc + bbase;
  const double *pcb = a.plm + (int64_t)b * 3 * 2 * g.PL;
  const int nx1e = nj + 2 * (ni + 1), nx2e = ni + 2 * (nj + 1), nx3e = 2 * ni + 2 * nj;
  const int ntask = 3 * NT + nx1e + nx2e + nx3e;
  const int wi = ni + 2, ncce = wi * (nj + 2);
  const int np = (nj + 1) * (ni + 1);
  const double a1 = dx2 * dx3, a2 = dx1 * dx3, a3 = dx1 * dx2;
  const double vol = dx1 * dx2 * dx3;

  TW f1own[5], f3hi[5];          // this plane's own x1 and x3 face fluxes (mesh frame)
  double f3lo[5];                // x3 fluxes of the own cell's lower face
  double dflx12[5];              // own cell's x1 + x2 flux differences (next cell plane)
  double g1l = 0.0, g1h = 0.0, g2l = 0.0, g2h = 0.0;   // their mass fluxes (self-gravity)
#pragma unroll
  for (int n = 0; n < 5; ++n) { f1own[n] = f3hi[n] = TW(0.0); f3lo[n] = dflx12[n] = 0.0; }
  double b3_lo = 0.0;            // new b3 at the lower face of cell plane c-1
  double dtmin = DBL_MAX;

  for (int c = g.ks - 1; c <= g.ke + 1; ++c) {
    const int cb = c & 1, pb = cb ^ 1;
    // ---- F: faces of plane c, cell-centred E of plane c
    const bool x3 = (c >= g.ks);
    for (int q = tid; q < ntask; q += NT) {
      int dir, ii, jj;
      if (q < 3 * NT) {
        dir = q / NT;
        ii = tx;
        jj = ty;
        if (!own) continue;
      } else {
        int e = q - 3 * NT;
        if (e < nx1e) {                        // x1: right edge, then rows -1 and nj
          dir = 0;
          if (e < nj) { ii = ni; jj = e; }
          else if (e - nj <= ni) { jj = -1; ii = e - nj; }
          else { jj = nj; ii = e - nj - (ni + 1); }
        } else if ((e -= nx1e) < nx2e) {       // x2: top edge, then columns -1 and ni
          dir = 1;
          if (e < ni) { jj = nj; ii = e; }
          else if (e - ni <= nj) { ii = -1; jj = e - ni; }
          else { ii = ni; jj = e - ni - (nj + 1); }
        } else {                               // x3: the ring without its corners
          e -= nx2e;
          dir = 2;
          if (e < ni) { jj = -1; ii = e; }
          else if (e < 2 * ni) { jj = nj; ii = e - ni; }
          else if (e < 2 * ni + nj) { ii = -1; jj = e - 2 * ni; }
          else { ii = ni; jj = e - 2 * ni - nj; }
        }
      }
      if (dir == 2 && !x3) continue;
      const int fi = i0 + ii, fj = j0 + jj;
      const int64_t id = IDX(g, c, fj, fi);
      // strides along dir and the two transverse directions (dir+1, dir+2) mod 3
      const int64_t st = (dir == 0) ? 1 : (dir == 1 ? sj : sk);
      const int64_t st1 = (dir == 0) ? sj : (dir == 1 ? sk : 1);
      const int64_t st2 = (dir == 0) ? sk : (dir == 1 ? 1 : sj);
      const int cidx = (dir == 0) ? fi : (dir == 1 ? fj : c);
      const double *bn = (dir == 0) ? a.bn1 : (dir == 1 ? a.bn2 : a.bn3);
      const TW dxd = (dir == 0) ? dxT[0] : (dir == 1 ? dxT[1] : dxT[2]);
      TW f[5], ey, ez, wc;
      face_flux<ORDER, RS, TW, ISO>(dir, wb, bb, S, id, st, st1, st2, pcb + (int64_t)dir * 2 * g.PL, g.PL, cidx,
                                    TW(bn[fb + id]), gT, csT, dflT, pflT, dxd, dtT, f, ey, ez, wc);
      if (dir == 0) {
        s_f1[0][jj + 1][ii] = ey; s_f1[1][jj + 1][ii] = ez; s_f1[2][jj + 1][ii] = wc;
      } else if (dir == 1) {
        s_f2[0][jj][ii + 1] = ey; s_f2[1][jj][ii + 1] = ez; s_f2[2][jj][ii + 1] = wc;
      } else {
        s_f3[0][jj + 1][ii + 1] = ey; s_f3[1][jj + 1][ii + 1] = ez; s_f3[2][jj + 1][ii + 1] = wc;
      }
      TW fm[5];
      flux_to_mesh(dir, f, fm);
      if (q < NT) {
#pragma unroll
        for (int n = 0; n < 5; ++n) f1own[n] = fm[n];
      } else if (q < 2 * NT) {
#pragma unroll
        for (int n = 0; n < NH; ++n) s_h2[n][jj][ii] = fm[n];
      } else if (q < 3 * NT) {
#pragma unroll
        for (int n = 0; n < 5; ++n) f3hi[n] = fm[n];
      } else if (dir == 0 && ii == ni && jj >= 0 && jj < nj) {
#pragma unroll
        for (int n = 0; n < NH; ++n) s_h1e[n][jj] = fm[n];
      } else if (dir == 1 && jj == nj && ii >= 0 && ii < ni) {
#pragma unroll
        for (int n = 0; n < NH; ++n) s_h2[n][nj][ii] = fm[n];
      }
    }
    for (int q = tid; q < ncce; q += NT) {
      const int jj = q / wi, ii = q - jj * wi;
      const int64_t id = IDX(g, c, j0 - 1 + jj, i0 - 1 + ii);
      const TW *wq = wb + id;
      const TW *bq = bb + id;
      const double vx = wq[IVX * S], vy = wq[IVY * S], vz = wq[IVZ * S];
      const double bx = bq[0], by = bq[S], bz = bq[2 * S];
      s_cce[0][jj][ii] = vz * by - vy * bz;
      s_cce[1][jj][ii] = vx * bz - vz * bx;
      s_cce[2][jj][ii] = vy * bx - vx * by;
    }
    __syncthreads();
    // ---- E: corner EMFs of edge plane c (e1, e2) and of plane c (e3)
    if (c >= g.ks) {
      const bool do3 = (c <= g.ke);
      for (int q = tid; q < np; q += NT) {
        const int jj = q / (ni + 1), ii = q - jj * (ni + 1);
        const bool do1 = ii < ni, do2 = jj < nj;
        if (do1) {
          s_e1[cb][jj][ii] = edge_e1(
              s_cce[0][jj + 1][ii + 1], s_cce[0][jj][ii + 1], s_ccp0[jj + 1][ii], s_ccp0[jj][ii],
              (double)s_f3[1][jj + 1][ii + 1], (double)s_f3[1][jj][ii + 1], (double)s_f2[0][jj][ii + 1],
              (double)s_f2p[0][jj][ii], (double)s_f2[2][jj][ii + 1], (double)s_f2p[1][jj][ii],
              (double)s_f3[2][jj + 1][ii + 1], (double)s_f3[2][jj][ii + 1]);
        }
        if (do2) {
          s_e2[cb][jj][ii] = edge_e2(
              s_cce[1][jj + 1][ii + 1], s_cce[1][jj + 1][ii], s_ccp1[jj][ii + 1], s_ccp1[jj][ii],
              (double)s_f3[0][jj + 1][ii + 1], (double)s_f3[0][jj + 1][ii], (double)s_f1[1][jj + 1][ii],
              (double)s_f1p[0][jj][ii], (double)s_f1[2][jj + 1][ii], (double)s_f1p[1][jj][ii],
              (double)s_f3[2][jj + 1][ii + 1], (double)s_f3[2][jj + 1][ii]);
        }
        if (do3) {
          s_e3[cb][jj][ii] = edge_e3(
              s_cce[2][jj + 1][ii + 1], s_cce[2][jj + 1][ii], s_cce[2][jj][ii + 1], s_cce[2][jj][ii],
              (double)s_f2[1][jj][ii + 1], (double)s_f2[1][jj][ii], (double)s_f1[0][jj + 1][ii],
              (double)s_f1[0][jj][ii], (double)s_f1[2][jj + 1][ii], (double)s_f1[2][jj][ii],
              (double)s_f2[2][jj][ii + 1], (double)s_f2[2][jj][ii]);
        }
      }
    }
    __syncthreads();
    // ---- U: CT of the own x3 face of face plane c, update of the own cell of plane c-1
    double b3_hi = 0.0;
    if (own && c >= g.ks) {
      {
        const int64_t fo = fb + IDX(g, c, j, i);
        double v = a.s3[fo];
        const double area = dx1 * dx2;
        v -= (wght / area) * (dx2 * s_e2[cb][ty][tx + 1] - dx2 * s_e2[cb][ty][tx]);
        v += (wght / area) * (dx1 * s_e1[cb][ty + 1][tx] - dx1 * s_e1[cb][ty][tx]);
        a.d3[fo] = v;
        b3_hi = v;
        int r[6];
        push_range(i, g.is, g.ie, g.bx1, g.ng, r[0], r[1]);
        push_range(j, g.js, g.je, g.bx2, g.ng, r[2], r[3]);
        push_range(c, g.ks, g.ke + 1, g.bx3, g.ng, r[4], r[5]);
        for_each_ghost(g, b, c, j, i, r, [&](int nb, int64_t id2) { a.d3[(int64_t)nb * S + id2] = v; });
      }
      if (c - 1 >= g.ks) {
        const int64_t id = IDX(g, c - 1, j, i);
        const int64_t fo = fb + id;
        const double area1 = dx2 * dx3, area2 = dx1 * dx3;
        double b1l = a.s1[fo], b1h = a.s1[fo + 1], b2l = a.s2[fo], b2h = a.s2[fo + sj];
        b1l -= (wght / area1) * (dx3 * s_e3[pb][ty + 1][tx] - dx3 * s_e3[pb][ty][tx]);
        b1l += (wght / area1) * (dx2 * s_e2[cb][ty][tx] - dx2 * s_e2[pb][ty][tx]);
        b1h -= (wght / area1) * (dx3 * s_e3[pb][ty + 1][tx + 1] - dx3 * s_e3[pb][ty][tx + 1]);
        b1h += (wght / area1) * (dx2 * s_e2[cb][ty][tx + 1] - dx2 * s_e2[pb][ty][tx + 1]);
        b2l += (wght / area2) * (dx3 * s_e3[pb][ty][tx + 1] - dx3 * s_e3[pb][ty][tx]);
        b2l -= (wght / area2) * (dx1 * s_e1[cb][ty][tx] - dx1 * s_e1[pb][ty][tx]);
        b2h += (wght / area2) * (dx3 * s_e3[pb][ty + 1][tx + 1] - dx3 * s_e3[pb][ty + 1][tx]);
        b2h -= (wght / area2) * (dx1 * s_e1[cb][ty + 1][tx] - dx1 * s_e1[pb][ty + 1][tx]);
        a.d1[fo] = b1l;
        if (tx == ni - 1) a.d1[fo + 1] = b1h;
        a.d2[fo] = b2l;
        if (ty == nj - 1) a.d2[fo + sj] = b2h;
        {
          int rc[6], r[6];
          push_range(i, g.is, g.ie, g.bx1, g.ng, rc[0], rc[1]);
          push_range(j, g.js, g.je, g.bx2, g.ng, rc[2], rc[3]);
          push_range(c - 1, g.ks, g.ke, g.bx3, g.ng, rc[4], rc[5]);
          r[2] = rc[2]; r[3] = rc[3]; r[4] = rc[4]; r[5] = rc[5];
          push_range(i, g.is, g.ie + 1, g.bx1, g.ng, r[0], r[1]);
          for_each_ghost(g, b, c - 1, j, i, r, [&](int nb, int64_t id2) { a.d1[(int64_t)nb * S + id2] = b1l; });
          if (tx == ni - 1) {
            push_range(i + 1, g.is, g.ie + 1, g.bx1, g.ng, r[0], r[1]);
            for_each_ghost(g, b, c - 1, j, i + 1, r,
                           [&](int nb, int64_t id2) { a.d1[(int64_t)nb * S + id2] = b1h; });
          }
          r[0] = rc[0]; r[1] = rc[1];
          push_range(j, g.js, g.je + 1, g.bx2, g.ng, r[2], r[3]);
          for_each_ghost(g, b, c - 1, j, i, r, [&](int nb, int64_t id2) { a.d2[(int64_t)nb * S + id2] = b2l; });
          if (ty == nj - 1) {
            push_range(j + 1, g.js, g.je + 1, g.bx2, g.ng, r[2], r[3]);
            for_each_ghost(g, b, c - 1, j + 1, i, r,
                           [&](int nb, int64_t id2) { a.d2[(int64_t)nb * S + id2] = b2h; });
          }
        }
        const int64_t cb0 = cbase + id;
        double un[NHYDRO];
        un[IEN] = 0.0;
#pragma unroll
        for (int n = 0; n < NH; ++n) {
          const int64_t o = cb0 + n * S;
          double dflx = dflx12[n];
          dflx += (a3 * (double)f3hi[n] - a3 * f3lo[n]);
          un[n] = a.src[o] - wght * dflx / vol;
        }
        if (a.grav) {
          const double rho = a.rho[cb0];
          const double phic = a.phi[fo];
          {
            double dtodx1 = wght / dx1;
            double phil = 0.5 * (a.phi[fo - 1] + phic);
            double phir = 0.5 * (phic + a.phi[fo + 1]);
            un[IVX] -= dtodx1 * rho * (phir - phil);
            if (!ISO) un[IEN] -= dtodx1 * (g1l * (phic - phil) + g1h * (phir - phic));
          }
          {
            double dtodx2 = wght / dx2;
            double phil = 0.5 * (a.phi[fo - sj] + phic);
            double phir = 0.5 * (phic + a.phi[fo + sj]);
            un[IVY] -= dtodx2 * rho * (phir - phil);
            if (!ISO) un[IEN] -= dtodx2 * (g2l * (phic - phil) + g2h * (phir - phic));
          }
          {
            double dtodx3 = wght / dx3;
            double phil = 0.5 * (a.phi[fo - sk] + phic);
            double phir = 0.5 * (phic + a.phi[fo + sk]);
            un[IVZ] -= dtodx3 * rho * (phir - phil);
            if (!ISO) un[IEN] -= dtodx3 * (f3lo[IDN] * (phic - phil) + (double)f3hi[IDN] * (phir - phic));
          }
        }
        const double bcc1 = 0.5 * b1l + 0.5 * b1h;
        const double bcc2 = 0.5 * b2l + 0.5 * b2h;
        const double bcc3 = 0.5 * b3_lo + 0.5 * b3_hi;
        const int64_t bo = bbase + id;
        a.bccn[bo] = TW(bcc1);
        a.bccn[bo + S] = TW(bcc2);
        a.bccn[bo + 2 * S] = TW(bcc3);
        int rc[6];
        push_range(i, g.is, g.ie, g.bx1, g.ng, rc[0], rc[1]);
        push_range(j, g.js, g.je, g.bx2, g.ng, rc[2], rc[3]);
        push_range(c - 1, g.ks, g.ke, g.bx3, g.ng, rc[4], rc[5]);
        if constexpr (!PUSH)
          for_each_ghost(g, b, c - 1, j, i, rc, [&](int nb, int64_t id2) {
            const int64_t c2 = (int64_t)nb * NHYDRO * S + id2;
#pragma unroll
            for (int n = 0; n < NH; ++n) a.dst[c2 + n * S] = un[n];
          });
        double wv[5];
        c2p_eos<ISO>(un[IDN], un[IVX], un[IVY], un[IVZ], un[IEN], bcc1, bcc2, bcc3, a.gm1, a.dfl, a.pfl, wv);
#pragma unroll
        for (int n = 0; n < NH; ++n) a.dst[cb0 + n * S] = un[n];
#pragma unroll
        for (int n = 0; n < NH; ++n) a.wn[cb0 + n * S] = TW(wv[n]);
        if constexpr (PUSH)   // as in k_stage
          for_each_ghost(g, b, c - 1, j, i, rc, [&](int nb, int64_t id2) {
            const int64_t c2 = (int64_t)nb * NHYDRO * S + id2, b2 = (int64_t)nb * 3 * S + id2;
#pragma unroll
            for (int n = 0; n < NH; ++n) {
              a.dst[c2 + n * S] = un[n];
              a.wn[c2 + n * S] = TW(wv[n]);
            }
            a.bccn[b2] = TW(bcc1);
            a.bccn[b2 + S] = TW(bcc2);
            a.bccn[b2 + 2 * S] = TW(bcc3);
          });
        if (a.dtpart)
          dtmin = dmin(dtmin, cell_dt<ISO>(a.gamma, a.cs, wv, bcc1, bcc2, bcc3, b1l, b2l, b3_lo, dx));
      }
    }
    b3_lo = b3_hi;
    // x1 + x2 flux differences of the own cell of plane c (summed as k_stage does)
    TW f1hi[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) f1hi[n] = __shfl_down(f1own[n], 1, 64);
    if (own && c >= g.ks && c <= g.ke) {
#pragma unroll
      for (int n = 0; n < NH; ++n) {
        const double fh = (tx == ni - 1) ? (double)s_h1e[n][ty] : (double)f1hi[n];
        double d = (a1 * fh - a1 * (double)f1own[n]);
        d += (a2 * (double)s_h2[n][ty + 1][tx] - a2 * (double)s_h2[n][ty][tx]);
        dflx12[n] = d;
      }
      g1l = (double)f1own[IDN];
      g1h = (tx == ni - 1) ? (double)s_h1e[IDN][ty] : (double)f1hi[IDN];
      g2l = (double)s_h2[IDN][ty][tx];
      g2h = (double)s_h2[IDN][ty + 1][tx];
    }
#pragma unroll
    for (int n = 0; n < 5; ++n) f3lo[n] = (double)f3hi[n];
    // the face data and cell-centred E of plane c become the previous plane
    for (int q = tid; q < nj * (ni + 1); q += NT) {
      const int jj = q / (ni + 1), ii = q - jj * (ni + 1);
      s_f1p[0][jj][ii] = s_f1[1][jj + 1][ii];
      s_f1p[1][jj][ii] = s_f1[2][jj + 1][ii];
    }
    for (int q = tid; q < (nj + 1) * ni; q += NT) {
      const int jj = q / ni, ii = q - jj * ni;
      s_f2p[0][jj][ii] = s_f2[0][jj][ii + 1];
      s_f2p[1][jj][ii] = s_f2[2][jj][ii + 1];
    }
    for (int q = tid; q < (nj + 2) * ni; q += NT) {
      const int jj = q / ni, ii = q - jj * ni;
      s_ccp0[jj][ii] = s_cce[0][jj][ii + 1];
    }
    for (int q = tid; q < nj * (ni + 2); q += NT) {
      const int jj = q / (ni + 2), ii = q - jj * (ni + 2);
      s_ccp1[jj][ii] = s_cce[1][jj + 1][ii];
    }
    __syncthreads();
  }
  if (a.dtpart) {
    double m = wg_min<NT / 64>(dtmin, s_red);
    if (tid == 0) a.dtpart[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = m;
  }
}

// The q-th element of the shell of the box [0,N1)x[0,N2)x[0,N3) that lies outside the
// inner box [a1,b1]x[a2,b2]x[a3,b3]: whole planes first, then whole rows of the inner
// planes, then the cells of the inner rows.
struct Shell {
  int N1, N2, N3, a1, b1, a2, b2, a3, b3;
  int64_t nk, nj, ni;
};

static Shell make_shell(int N1, int N2, int N3, int a1, int b1, int a2, int b2, int a3, int b3) {
  Shell s{N1, N2, N3, a1, b1, a2, b2, a3, b3, 0, 0, 0};
  s.nk = (int64_t)(a3 + (N3 - 1 - b3)) * N2 * N1;
  s.nj = (int64_t)(b3 - a3 + 1) * (a2 + (N2 - 1 - b2)) * N1;
  s.ni = (int64_t)(b3 - a3 + 1) * (b2 - a2 + 1) * (a1 + (N1 - 1 - b1));
  return s;
}

__device__ __forceinline__ bool shell_elem(const Shell &s, unsigned q, int &k, int &j, int &i) {
  // 32-bit index math: a shell has fewer than 2^31 elements per MeshBlock
  const unsigned N1 = s.N1, plane = (unsigned)s.N1 * (unsigned)s.N2;
  const unsigned nk = (unsigned)s.nk, nj = (unsigned)s.nj, ni = (unsigned)s.ni;
  if (q < nk) {
    int pidx = (int)(q / plane);
    unsigned r = q - (unsigned)pidx * plane;
    k = (pidx < s.a3) ? pidx : s.b3 + 1 + (pidx - s.a3);
    j = (int)(r / N1);
    i = (int)(r - (unsigned)j * N1);
    return true;
  }
  q -= nk;
  if (q < nj) {
    const unsigned rows_out = s.a2 + (s.N2 - 1 - s.b2);
    const unsigned per = rows_out * N1;
    int kk = (int)(q / per);
    unsigned r = q - (unsigned)kk * per;
    int rr = (int)(r / N1);
    i = (int)(r - (unsigned)rr * N1);
    k = s.a3 + kk;
    j = (rr < s.a2) ? rr : s.b2 + 1 + (rr - s.a2);
    return true;
  }
  q -= nj;
  if (q < ni) {
    const unsigned cols_out = s.a1 + (s.N1 - 1 - s.b1);
    const unsigned nrows = s.b2 - s.a2 + 1;
    const unsigned per = nrows * cols_out;
    int kk = (int)(q / per);
    unsigned r = q - (unsigned)kk * per;
    int jj = (int)(r / cols_out);
    int cc = (int)(r - (unsigned)jj * cols_out);
    k = s.a3 + kk;
    j = s.a2 + jj;
    i = (cc < s.a1) ? cc : s.b1 + 1 + (cc - s.a1);
    return true;
  }
  return false;
}

// ConservedToPrimitive over the ghost zones (time_integrator.cpp:1971-1977,
// eos/adiabatic_mhd.cpp:41-90 or isothermal_mhd.cpp:40-73) after the exchange: bcc
// from this block's own faces (its boundary faces and the received ghost faces), C2P of
// the received conserved state, floors written back.  One thread per ghost cell.  The
// two copies of a face shared by neighbouring blocks are computed separately and may
// differ by rounding, so a ghost's primitives are not copied from its source cell.
template <typename TW, bool ISO>
__global__ void __launch_bounds__(256)
k_ghost_c2p(Geo g, Shell s, double *__restrict__ u, TW *__restrict__ w, TW *__restrict__ bcc,
            const double *__restrict__ b1, const double *__restrict__ b2,
            const double *__restrict__ b3, double gm1, double dfl, double pfl) {
  const uint3 wg = xcd_block();
  const unsigned q = wg.x * 256 + threadIdx.x;
  if (q >= s.nk + s.nj + s.ni) return;
  int k, j, i;
  shell_elem(s, q, k, j, i);
  const int b = wg.z;
  const int64_t S = g.S;
  const int64_t id = IDX(g, k, j, i);
  const int64_t c = (int64_t)b * NHYDRO * S + id;
  const int64_t fo = (int64_t)b * S + id;
  const double bcc1 = 0.5 * b1[fo] + 0.5 * b1[fo + 1];
  const double bcc2 = 0.5 * b2[fo] + 0.5 * b2[fo + g.P];
  const double bcc3 = 0.5 * b3[fo] + 0.5 * b3[fo + (int64_t)g.J * g.P];
  const int64_t bo = (int64_t)b * 3 * S + id;
  bcc[bo] = TW(bcc1);
  bcc[bo + S] = TW(bcc2);
  bcc[bo + 2 * S] = TW(bcc3);
  double u_d = u[c], u_e = ISO ? 0.0 : u[c + 4 * S], wv[5];
  c2p_eos<ISO>(u_d, u[c + S], u[c + 2 * S], u[c + 3 * S], u_e, bcc1, bcc2, bcc3, gm1, dfl, pfl, wv);
  u[c] = u_d;
  if (!ISO) u[c + 4 * S] = u_e;
  constexpr int NH = ISO ? 4 : NHYDRO;
#pragma unroll
  for (int n = 0; n < NH; ++n) w[c + n * S] = TW(wv[n]);
}

// Hydro::NewBlockTimeStep (hydro/new_blockdt.cpp:57-160): per-cell bound, min over all
// interior cells of all blocks, one partial per workgroup; k_min_final reduces the
// partials (min is exact, so the reduction order does not matter).
template <typename TW, bool ISO>
__global__ void __launch_bounds__(256)
k_newdt(Geo g, const TW *__restrict__ w, const TW *__restrict__ bcc,
        const double *__restrict__ b1, const double *__restrict__ b2,
        const double *__restrict__ b3, const BlockCoef *__restrict__ coef, double gamma,
        double cs, double *__restrict__ part) {
  __shared__ double sh[4];
  PlaneIdx p;
  double m = DBL_MAX;
  if (plane_index(g.is, g.ie, g.js, g.je, g.ks, p)) {
    const int64_t S = g.S;
    const int64_t id = IDX(g, p.k, p.j, p.i);
    const int64_t c = (int64_t)p.b * NHYDRO * S + id;
    const int64_t bo = (int64_t)p.b * 3 * S + id;
    const int64_t fo = (int64_t)p.b * S + id;
    double wv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    constexpr int NH = ISO ? 4 : NHYDRO;
#pragma unroll
    for (int n = 0; n < NH; ++n) wv[n] = w[c + n * S];
    m = cell_dt<ISO>(gamma, cs, wv, bcc[bo], bcc[bo + S], bcc[bo + 2 * S], b1[fo], b2[fo], b3[fo],
                     coef[p.b].dx);
  }
  m = wg_min256(m, sh);
  if (threadIdx.x == 0) part[((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = m;
}

__global__ void __launch_bounds__(256) k_min_final(const double *__restrict__ part, int64_t n,
                                                   double *__restrict__ out) {
  __shared__ double sh[256];
  double m = DBL_MAX;
  for (int64_t q = threadIdx.x; q < n; q += 256) m = dmin(m, part[q]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = dmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// ================================================================ host side
namespace agx {

double meshgen_x(int64_t index, int64_t nrange) {   // mesh/mesh.hpp:389-405
  int64_t noffset = index - (nrange) / 2;
  int64_t noffset_ceil = index - (nrange + 1) / 2;
  return (double)(noffset + noffset_ceil) / (2.0 * nrange);
}
double uniform_gen(double x, double xmin, double xmax) {   // mesh/mesh.hpp:465-488
  return 0.5 * (xmin + xmax) + (x * xmax - x * xmin);
}

// Face/centre coordinates of one block along one direction on a uniform mesh
// (coordinates/coordinates.cpp:127-145, cartesian.cpp:26-36, mesh/mesh.cpp:1668-1690).
void block_coords(int nmesh, int nblk, int nbx, int ng, int lx, double xmin, double xmax,
                  std::vector<double> &xf, std::vector<double> &xv, double &dx) {
  double bmin = (lx == 0) ? xmin : uniform_gen(meshgen_x(lx, nblk), xmin, xmax);
  double bmax = (lx == nblk - 1) ? xmax : uniform_gen(meshgen_x(lx + 1, nblk), xmin, xmax);
  dx = (bmax - bmin) / nbx;
  int n = nbx + 2 * ng;
  xf.assign(n + 1, 0.0);
  xv.assign(n, 0.0);
  int il = ng, iu = ng + nbx - 1;
  for (int i = 0; i <= n; ++i) {
    int64_t noffset = (int64_t)(i - il) + (int64_t)lx * nbx;
    xf[i] = uniform_gen(meshgen_x(noffset, nmesh), xmin, xmax);
  }
  xf[il] = bmin;
  xf[iu + 1] = bmax;
  for (int i = 0; i < n; ++i) xv[i] = 0.5 * (xf[i + 1] + xf[i]);
}

}  // namespace agx

using agx::block_coords;

namespace {

int check_desc(const agx_mesh_desc *d) {
  if (!d) { set_error("null descriptor"); return -1; }
  if (d->nx1 <= 0 || d->nx2 <= 0 || d->nx3 <= 0 || d->bnx1 <= 0 || d->bnx2 <= 0 || d->bnx3 <= 0) {
    set_error("mesh and MeshBlock sizes must be positive"); return -2;
  }
  if (d->nx1 % d->bnx1 || d->nx2 % d->bnx2 || d->nx3 % d->bnx3) {
    set_error("mesh size must be a multiple of the MeshBlock size"); return -3;
  }
  if (d->bnx1 < 2 || d->bnx2 < 2 || d->bnx3 < 2) {
    set_error("only 3D meshes are supported (every MeshBlock dimension >= 2)"); return -4;
  }
  if (d->nghost < 2) { set_error("nghost must be >= 2"); return -5; }
  if (d->xorder < 1 || d->xorder > 3) {
    set_error("xorder=%d not supported (1: donor cell, 2: PLM, 3: PPM)", d->xorder); return -6;
  }
  if (d->xorder == 3 && d->nghost < 3) {
    set_error("xorder=3 (PPM) needs nghost >= 3 (reconstruction.cpp:90-100)"); return -13;
  }
  if (d->char_proj != 0 && d->char_proj != 1) { set_error("char_proj must be 0 or 1"); return -6; }
  if (d->char_proj && d->xorder == 1) {   // reconstruction.cpp:49-71 knows no "1c"
    set_error("characteristic projection needs xorder 2 or 3"); return -6;
  }
  if (d->nghost > d->bnx1 || d->nghost > d->bnx2 || d->nghost > d->bnx3) {
    set_error("nghost larger than a MeshBlock"); return -7;
  }
  if (d->rsolver != AGX_RSOLVER_HLLD && d->rsolver != AGX_RSOLVER_HLLE &&
      d->rsolver != AGX_RSOLVER_ROE && d->rsolver != AGX_RSOLVER_LHLLD) {
    set_error("unknown Riemann solver %d", d->rsolver); return -8;
  }
  if (d->eos == 1 && d->rsolver == AGX_RSOLVER_LHLLD) {   // configure.py:321-322
    set_error("LHLLD flux cannot be used with the isothermal EOS"); return -9;
  }
  if (d->eos != 0 && d->eos != 1) { set_error("eos must be 0 (adiabatic) or 1 (isothermal)"); return -9; }
  if (d->eos == 1 && d->rsolver == AGX_RSOLVER_ROE) {
    set_error("the isothermal EOS supports HLLD (hlld_iso.cpp) and HLLE only"); return -9;
  }
  if (d->eos == 1 && !(d->iso_cs > 0.0)) { set_error("iso_cs must be > 0"); return -9; }
  if (d->precision != AGX_PREC_FP64 && d->precision != AGX_PREC_MIXED) {
    set_error("precision %d not supported (0: FP64, 1: mixed)", d->precision); return -10;
  }
  if (!(d->gamma > 1.0)) { set_error("gamma must be > 1"); return -11; }
  if (d->mg_bc != 0 && d->mg_bc != 1) {
    set_error("mg_bc %d not supported (0: periodic, 1: multipole)", d->mg_bc); return -14;
  }
  if (d->mg_bc == 1 && d->mp_auto_origin && d->mp_nodipole) {
    set_error("auto_mporigin and nodipole cannot be used together (mg_gravity.cpp:106-113)"); return -14;
  }
  if (d->mg_bc == 1 && d->mporder != 2 && d->mporder != 4) {
    set_error("mporder must be 2 or 4 (mg_gravity.cpp:97-105)"); return -14;
  }
  return 0;
}

}  // namespace

extern "C" {

const char *agx_last_error(void) { return agx::g_err.c_str(); }

// bytes of one registered array: nv slabs per block, plus room for the alignment offset,
// rounded to 256 B
static int64_t arr_bytes(int nv, const Geo &g, int esz) {
  int64_t b = ((int64_t)nv * g.nb * g.S + 32) * esz;
  return (b + 255) / 256 * 256;
}

// Host tables of block b at logical location lx of a lattice of nblk blocks over nmesh
// cells between xmin and xmax per direction (Mesh::SetBlockSizeAndBoundaries on a uniform
// mesh, mesh.cpp:1668-1690): its cell width and the PLM face-offset factors
// (reconstruct/plm.cpp:116-117,224-225,330-331).
static void set_block_geometry(agx_mesh *m, int b, const int nmesh[3], const int nblk[3], const int lx[3],
                               const double xmin[3], const double xmax[3]) {
  const Geo &g = m->g;
  const int nbx[3] = {g.bx1, g.bx2, g.bx3};
  for (int dd = 0; dd < 3; ++dd) {
    std::vector<double> xf, xv;
    double dx;
    block_coords(nmesh[dd], nblk[dd], nbx[dd], g.ng, lx[dd], xmin[dd], xmax[dd], xf, xv, dx);
    m->coef_h[b].dx[dd] = dx;
    double *pp = &m->plm_h[((size_t)b * 3 + dd) * 2 * g.PL];
    const int n = nbx[dd] + 2 * g.ng;
    for (int i = 0; i < n; ++i) {
      pp[i] = (xf[i + 1] - xv[i]) / dx;
      pp[g.PL + i] = (xv[i] - xf[i]) / dx;
    }
  }
}

int agx_mesh_create(const agx_mesh_desc *desc, agx_mesh **out) {
  if (!out) { set_error("null output pointer"); return -1; }
  *out = nullptr;
  int rc = check_desc(desc);
  if (rc) return rc;
  agx_mesh *m = new (std::nothrow) agx_mesh();
  if (!m) { set_error("host allocation failed"); return -12; }
  m->d = *desc;
  // CFL clamp of the VL2 integrator in 3D, time_integrator.cpp:212-215,885-896
  if (m->d.cfl > 0.5) m->d.cfl = 0.5;
  // characteristic projection is built into the split forms only: 32 x 8 for its stage
  if (m->d.char_proj && m->d.xorder > 1) m->tile = 1;
  Geo &g = m->g;
  g.nb1 = desc->nx1 / desc->bnx1;
  g.nb2 = desc->nx2 / desc->bnx2;
  g.nb3 = desc->nx3 / desc->bnx3;
  g.nb = g.nb1 * g.nb2 * g.nb3;
  g.ng = desc->nghost;
  g.bx1 = desc->bnx1; g.bx2 = desc->bnx2; g.bx3 = desc->bnx3;
  g.n1 = g.bx1 + 2 * g.ng; g.n2 = g.bx2 + 2 * g.ng; g.n3 = g.bx3 + 2 * g.ng;
  g.is = g.ng; g.ie = g.ng + g.bx1 - 1;
  g.js = g.ng; g.je = g.ng + g.bx2 - 1;
  g.ks = g.ng; g.ke = g.ng + g.bx3 - 1;
  // row pitch: a multiple of 32 elements, so FP64 and FP32 rows both start on a 128-B line
  g.P = ((g.n1 + 1) + 31) / 32 * 32;
  g.PL = g.n1 > g.n2 ? g.n1 : g.n2;
  g.PL = g.PL > g.n3 ? g.PL : g.n3;
  g.J = g.n2 + 1;
  g.S = (int64_t)(g.n3 + 1) * g.J * g.P;
  // variables per array: the AGX_ARR_* set, then W2, BCC2, Q1F, Q2F, Q3F
  static const int nv[ARR_TOTAL] = {5, 5, 5, 3, 1, 1, 1, 1, 1, 1, 5, 5, 5,
                                    1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                    5, 3, 1, 1, 1};
  m->mixed = (desc->precision == AGX_PREC_MIXED);
  m->iso = (desc->eos == 1);
  int64_t total = 0;
  for (int a = 0; a < ARR_TOTAL; ++a) {
    bool f32 = m->mixed && (a == AGX_ARR_W || a == AGX_ARR_BCC || a == ARR_W2 || a == ARR_BCC2 ||
                            (a >= AGX_ARR_FX1 && a <= AGX_ARR_WX3));
    m->arr.nvar[a] = nv[a];
    m->arr.esz[a] = f32 ? 4 : 8;
    // every array starts on a 256-B boundary (S is a multiple of 16 elements)
    total += arr_bytes(nv[a], g, m->arr.esz[a]);
  }
  m->pool_bytes = total;
  hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { int r = hip_fail(e, "hipStreamCreate"); delete m; return r; }
  e = hipMalloc(&m->pool, m->pool_bytes);
  if (e != hipSuccess) { int r = hip_fail(e, "hipMalloc(pool)"); agx_mesh_destroy(m); return r; }
  e = hipMemsetAsync(m->pool, 0, m->pool_bytes, m->stream);
  if (e != hipSuccess) { int r = hip_fail(e, "hipMemset(pool)"); agx_mesh_destroy(m); return r; }
  // every array starts OFF elements past a 256-B boundary so that the first interior
  // element of each row (i = is) opens a 128-B line: wavefront-wide row segments of the
  // interior then map onto whole lines (no partially read or written lines)
  const int OFF = (32 - g.ng % 32) % 32;
  char *ptr = reinterpret_cast<char *>(m->pool);
  for (int a = 0; a < ARR_TOTAL; ++a) {
    m->arr.a[a] = reinterpret_cast<double *>(ptr + (int64_t)OFF * m->arr.esz[a]);
    ptr += arr_bytes(nv[a], g, m->arr.esz[a]);
  }
  // per-block geometry
  m->coef_h.resize(g.nb);
  m->plm_h.assign((size_t)g.nb * 3 * 2 * g.PL, 0.0);
  {
    const int nmesh[3] = {desc->nx1, desc->nx2, desc->nx3}, nblk[3] = {g.nb1, g.nb2, g.nb3};
    const double xmin[3] = {desc->x1min, desc->x2min, desc->x3min};
    const double xmax[3] = {desc->x1max, desc->x2max, desc->x3max};
    for (int b = 0; b < g.nb; ++b) {
      const int lx[3] = {b % g.nb1, (b / g.nb1) % g.nb2, b / (g.nb1 * g.nb2)};
      set_block_geometry(m, b, nmesh, nblk, lx, xmin, xmax);
    }
  }
  e = hipMalloc(&m->coef, sizeof(BlockCoef) * g.nb);
  if (e == hipSuccess) e = hipMalloc(&m->plm, sizeof(double) * m->plm_h.size());
  if (e == hipSuccess) e = hipMalloc(&m->scal, sizeof(double) * 64);
  {
    dim3 gr = plane_grid(g.is, g.ie, g.js, g.je, g.ks, g.ke, g.nb);
    m->red_n = (int64_t)gr.x * gr.y * gr.z;
    m->dtpart_n = 0;
    for (int t = 0; t < NTILE; ++t) {
      dim3 gs = stage_grid(g, t);
      m->dtpart_n = std::max(m->dtpart_n, (int64_t)gs.x * gs.y);
    }
  }
  if (e == hipSuccess) e = hipMalloc(&m->red, sizeof(double) * m->red_n);
  if (e == hipSuccess) e = hipMalloc(&m->dtpart, sizeof(double) * m->dtpart_n);
  if (e == hipSuccess) e = hipMalloc(&m->hist, sizeof(double) * (HIST_PARTIAL_DOUBLES + HIST_NV));
  if (e == hipSuccess) e = hipMalloc(&m->cellvol, sizeof(double) * g.nb);
  if (m->mixed && g.ng >= 3) {   // near-tie flags of the FP32 PPM flux pass, all zero between passes
    m->nearf_n = ((int64_t)g.nb * g.S + FIX_SPAN - 1) / FIX_SPAN * FIX_SPAN;
    if (e == hipSuccess) e = hipMalloc(&m->nearf, m->nearf_n);
    if (e == hipSuccess) e = hipMemsetAsync(m->nearf, 0, m->nearf_n, m->stream);
  }
  // Coordinates::CellVolume, Cartesian: dx1f*dx2f*dx3f (coordinates.cpp:521-531)
  std::vector<double> vol_h(g.nb);
  for (int b = 0; b < g.nb; ++b) vol_h[b] = m->coef_h[b].dx[0] * m->coef_h[b].dx[1] * m->coef_h[b].dx[2];
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->cellvol, vol_h.data(), sizeof(double) * g.nb, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipHostMalloc(&m->scal_host, sizeof(double) * 64, hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->coef, m->coef_h.data(), sizeof(BlockCoef) * g.nb, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->plm, m->plm_h.data(), sizeof(double) * m->plm_h.size(), hipMemcpyHostToDevice,
                       m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) { int r = hip_fail(e, "mesh geometry"); agx_mesh_destroy(m); return r; }
  m->dt = DBL_MAX;
  if (desc->self_gravity) {
    int r = mg_create(m);
    if (r) { agx_mesh_destroy(m); return r; }
  }
  *out = m;
  return 0;
}

static void graph_reset(agx_mesh *m);

int agx_mesh_destroy(agx_mesh *m) {
  if (!m) return 0;
  if (m->stream) hipStreamSynchronize(m->stream);
  graph_reset(m);
  for (auto &ev : m->timer.ev) hipEventDestroy(ev);
  mg_destroy(m);
  if (m->pool) hipFree(m->pool);
  if (m->coef) hipFree(m->coef);
  if (m->plm) hipFree(m->plm);
  if (m->scal) hipFree(m->scal);
  if (m->scal_host) hipHostFree(m->scal_host);
  if (m->red) hipFree(m->red);
  if (m->dtpart) hipFree(m->dtpart);
  if (m->hist) hipFree(m->hist);
  if (m->cellvol) hipFree(m->cellvol);
  if (m->nearf) hipFree(m->nearf);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

// host (tight) dims of each registered array: nvar, n3', n2', n1'
static void host_dims(const agx_mesh *m, int arr, int64_t dims[4]) {
  const Geo &g = m->g;
  int64_t n1 = g.n1, n2 = g.n2, n3 = g.n3;
  int64_t d1 = n1, d2 = n2, d3 = n3;
  switch (arr) {
    case AGX_ARR_B1F: case AGX_ARR_R1F: case AGX_ARR_FX1: case AGX_ARR_E3X1: case AGX_ARR_E2X1:
    case AGX_ARR_WX1: d1 = n1 + 1; break;
    case AGX_ARR_B2F: case AGX_ARR_R2F: case AGX_ARR_FX2: case AGX_ARR_E1X2: case AGX_ARR_E3X2:
    case AGX_ARR_WX2: d2 = n2 + 1; break;
    case AGX_ARR_B3F: case AGX_ARR_R3F: case AGX_ARR_FX3: case AGX_ARR_E1X3: case AGX_ARR_E2X3:
    case AGX_ARR_WX3: d3 = n3 + 1; break;
    case AGX_ARR_E1: d3 = n3 + 1; d2 = n2 + 1; break;
    case AGX_ARR_E2: d3 = n3 + 1; d1 = n1 + 1; break;
    case AGX_ARR_E3: d2 = n2 + 1; d1 = n1 + 1; break;
    default: break;
  }
  dims[0] = m->arr.nvar[arr];
  // isothermal: NHYDRO = 4 on the host side (the device keeps a fifth, unused slot)
  if (m->iso && (arr == AGX_ARR_U || arr == AGX_ARR_U1 || arr == AGX_ARR_W || arr == AGX_ARR_FX1 ||
                 arr == AGX_ARR_FX2 || arr == AGX_ARR_FX3))
    dims[0] = 4;
  dims[1] = d3;
  dims[2] = d2;
  dims[3] = d1;
}

int agx_array_dims(agx_mesh *m, int arr, int64_t *nblocks, int64_t dims[4]) {
  if (!m || arr < 0 || arr >= AGX_ARR_COUNT || !dims) { set_error("bad argument"); return -1; }
  if (nblocks) *nblocks = m->g.nb;
  host_dims(m, arr, dims);
  return 0;
}

int64_t agx_mesh_bytes(agx_mesh *m) { return m ? m->pool_bytes + m->mg_bytes : 0; }

int agx_array_ptr(agx_mesh *m, int arr, void **ptr) {
  if (!m || arr < 0 || arr >= AGX_ARR_COUNT || !ptr) { set_error("agx_array_ptr: bad argument"); return -1; }
  *ptr = m->arr.a[arr];
  return 0;
}

int agx_mesh_stream(agx_mesh *m, void **stream) {
  if (!m || !stream) { set_error("agx_mesh_stream: bad argument"); return -1; }
  *stream = m->stream;
  return 0;
}

int agx_mesh_set_levels(agx_mesh *m, const int32_t nrb[3], const double box[6], const int32_t *level,
                        const int64_t *lx) {
  if (!m || !nrb || !box || !level || !lx) { set_error("agx_mesh_set_levels: null argument"); return -1; }
  if (m->mixed) { set_error("agx_mesh_set_levels: the refined mesh runs the task-level FP64 steps only"); return -2; }
  if (m->d.self_gravity) { set_error("agx_mesh_set_levels: multigrid on refined meshes is not built"); return -3; }
  const Geo &g = m->g;
  if (g.ng % 2) {   // MeshRefinement constructor, mesh_refinement.cpp:56-62
    set_error("agx_mesh_set_levels: an odd ghost width (%d) is incompatible with mesh refinement", g.ng);
    return -8;
  }
  const int nbx[3] = {g.bx1, g.bx2, g.bx3};
  for (int d = 0; d < 3; ++d) {
    if (nrb[d] < 1 || nbx[d] % 2) { set_error("agx_mesh_set_levels: root blocks >= 1 and even block sizes"); return -4; }
    if (!(box[2 * d + 1] > box[2 * d])) { set_error("agx_mesh_set_levels: empty box"); return -5; }
  }
  for (int b = 0; b < g.nb; ++b) {
    if (level[b] < 0 || level[b] > 20) { set_error("agx_mesh_set_levels: block %d: level %d", b, level[b]); return -6; }
    for (int d = 0; d < 3; ++d)
      if (lx[3 * b + d] < 0 || lx[3 * b + d] >= ((int64_t)nrb[d] << level[b])) {
        set_error("agx_mesh_set_levels: block %d: location outside the level's lattice", b);
        return -7;
      }
  }
  const double xmin[3] = {box[0], box[2], box[4]}, xmax[3] = {box[1], box[3], box[5]};
  for (int b = 0; b < g.nb; ++b) {
    const int L = level[b];
    const int nmesh[3] = {(nrb[0] * nbx[0]) << L, (nrb[1] * nbx[1]) << L, (nrb[2] * nbx[2]) << L};
    const int nblk[3] = {nrb[0] << L, nrb[1] << L, nrb[2] << L};
    const int lxb[3] = {(int)lx[3 * b], (int)lx[3 * b + 1], (int)lx[3 * b + 2]};
    set_block_geometry(m, b, nmesh, nblk, lxb, xmin, xmax);
  }
  std::vector<double> vol_h(g.nb);
  for (int b = 0; b < g.nb; ++b) vol_h[b] = m->coef_h[b].dx[0] * m->coef_h[b].dx[1] * m->coef_h[b].dx[2];
  AGX_HIP(hipMemcpyAsync(m->coef, m->coef_h.data(), sizeof(BlockCoef) * g.nb, hipMemcpyHostToDevice, m->stream));
  AGX_HIP(hipMemcpyAsync(m->plm, m->plm_h.data(), sizeof(double) * m->plm_h.size(), hipMemcpyHostToDevice,
                         m->stream));
  AGX_HIP(hipMemcpyAsync(m->cellvol, vol_h.data(), sizeof(double) * g.nb, hipMemcpyHostToDevice, m->stream));
  AGX_HIP(hipStreamSynchronize(m->stream));
  m->tree = true;
  graph_reset(m);
  return 0;
}

// Strided 3D copies between the tight host layout and the padded device layout.  FP32
// arrays of the mixed-precision form are staged through a host buffer and converted.
// Host <-> device copy of one registered array.  Each device slab (one variable of one
// block, padded layout) is packed or unpacked on the host in a pinned staging buffer and
// moved whole with one contiguous copy; two staging halves alternate so that packing
// overlaps the transfer.  Padding elements are written as zeros (they hold zeros from
// agx_mesh_create on and no kernel writes them).
static int copy_array(agx_mesh *m, int arr, double *host, bool to_device) {
  if (!m || arr < 0 || arr >= AGX_ARR_COUNT || !host) { set_error("bad argument"); return -1; }
  const Geo &g = m->g;
  int64_t dims[4];
  host_dims(m, arr, dims);
  const int esz = m->arr.esz[arr];
  const int64_t nslab = (int64_t)g.nb * dims[0];
  const int64_t nslab_el = dims[1] * dims[2] * dims[3];
  const int64_t slab_bytes = g.S * esz;
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nslab, (int64_t(64) << 20) / slab_bytes));
  const int nv_dev = m->arr.nvar[arr];   // device slabs per block (may exceed the host's)
  char *dbase = reinterpret_cast<char *>(m->arr.a[arr]);
  char *stage = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  hipError_t e = hipHostMalloc((void **)&stage, (size_t)(2 * chunk * slab_bytes), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  auto pack = [&](int64_t s, char *buf) {   // host slab s -> padded slab in buf
    std::memset(buf, 0, (size_t)slab_bytes);
    const double *h = host + s * nslab_el;
    for (int64_t k = 0; k < dims[1]; ++k)
      for (int64_t j = 0; j < dims[2]; ++j) {
        const double *row = h + (k * dims[2] + j) * dims[3];
        const int64_t o = IDX(g, (int)k, (int)j, 0);
        if (esz == 4) {
          float *d = reinterpret_cast<float *>(buf) + o;
          for (int64_t i = 0; i < dims[3]; ++i) d[i] = (float)row[i];
        } else {
          std::memcpy(reinterpret_cast<double *>(buf) + o, row, (size_t)dims[3] * sizeof(double));
        }
      }
  };
  auto unpack = [&](int64_t s, const char *buf) {
    double *h = host + s * nslab_el;
    for (int64_t k = 0; k < dims[1]; ++k)
      for (int64_t j = 0; j < dims[2]; ++j) {
        double *row = h + (k * dims[2] + j) * dims[3];
        const int64_t o = IDX(g, (int)k, (int)j, 0);
        if (esz == 4) {
          const float *d = reinterpret_cast<const float *>(buf) + o;
          for (int64_t i = 0; i < dims[3]; ++i) row[i] = (double)d[i];
        } else {
          std::memcpy(row, reinterpret_cast<const double *>(buf) + o, (size_t)dims[3] * sizeof(double));
        }
      }
  };
  int h = 0;
  for (int64_t c0 = 0; e == hipSuccess && c0 < nslab; c0 += chunk, h ^= 1) {
    const int64_t n = std::min(chunk, nslab - c0);
    char *buf = stage + h * chunk * slab_bytes;
    if (used[h]) e = hipEventSynchronize(ev[h]);
    for (int64_t s = c0; e == hipSuccess && s < c0 + n; ++s) {
      const int64_t sd = (s / dims[0]) * nv_dev + (s % dims[0]);
      char *sb = buf + (s - c0) * slab_bytes;
      if (to_device) {
        pack(s, sb);
        e = hipMemcpyAsync(dbase + sd * slab_bytes, sb, (size_t)slab_bytes, hipMemcpyHostToDevice, m->stream);
      } else {
        e = hipMemcpyAsync(sb, dbase + sd * slab_bytes, (size_t)slab_bytes, hipMemcpyDeviceToHost, m->stream);
      }
    }
    if (e == hipSuccess && !to_device) {
      e = hipStreamSynchronize(m->stream);
      for (int64_t s = c0; e == hipSuccess && s < c0 + n; ++s) unpack(s, buf + (s - c0) * slab_bytes);
    }
    if (e == hipSuccess) e = hipEventRecord(ev[h], m->stream);
    used[h] = true;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (ev[0]) hipEventDestroy(ev[0]);
  if (ev[1]) hipEventDestroy(ev[1]);
  if (stage) hipHostFree(stage);
  if (e != hipSuccess) return hip_fail(e, to_device ? "agx_upload" : "agx_download");
  return 0;
}

int agx_upload(agx_mesh *m, int arr, const double *host) {
  if (m && (arr == AGX_ARR_B1F || arr == AGX_ARR_B2F || arr == AGX_ARR_B3F)) m->copies_state = -1;
  return copy_array(m, arr, const_cast<double *>(host), true);
}
int agx_download(agx_mesh *m, int arr, double *host) { return copy_array(m, arr, host, false); }

// MeshBlocks of uneven sizes on one level, mapped onto the device lattice of bnx blocks by
// global index.  Cell arrays (u, u1, w, bcc, phi) and face fields (b, b1).  One download
// of the lattice array, the per-block copies on the host, and (upload) one upload back.
static int copy_blocks(agx_mesh *m, int arr, int32_t nblocks, const int32_t *origin, const int32_t *size,
                       double *const *host, bool to_device) {
  if (!m || arr < 0 || arr >= AGX_ARR_COUNT || nblocks < 0 || (nblocks > 0 && (!origin || !size || !host))) {
    set_error("agx_%s_blocks: bad argument", to_device ? "upload" : "download");
    return -1;
  }
  int fdir = -1;   // face normal of a face array
  switch (arr) {
    case AGX_ARR_U: case AGX_ARR_U1: case AGX_ARR_W: case AGX_ARR_BCC: case AGX_ARR_PHI: break;
    case AGX_ARR_B1F: case AGX_ARR_R1F: fdir = 0; break;
    case AGX_ARR_B2F: case AGX_ARR_R2F: fdir = 1; break;
    case AGX_ARR_B3F: case AGX_ARR_R3F: fdir = 2; break;
    default: set_error("agx_%s_blocks: array %d is not a cell array or face field", to_device ? "upload" : "download", arr);
      return -53;
  }
  const Geo &g = m->g;
  const int bn[3] = {g.bx1, g.bx2, g.bx3}, nl[3] = {g.nb1, g.nb2, g.nb3}, ng = g.ng;
  const int N[3] = {g.bx1 * g.nb1, g.bx2 * g.nb2, g.bx3 * g.nb3};
  for (int b = 0; b < nblocks; ++b)
    for (int d = 0; d < 3; ++d) {
      const int o = origin[3 * b + d], z = size[3 * b + d];
      if (o < 0 || z < bn[d] || o % bn[d] || z % bn[d] || o + z > N[d]) {
        set_error("block %d: origin and size must be multiples of the MeshBlock size %d inside the mesh (dir %d)",
                  b, bn[d], d + 1);
        return -54;
      }
    }
  int64_t dims[4];
  host_dims(m, arr, dims);
  const int64_t nv = dims[0], L3 = dims[1], L2 = dims[2], L1 = dims[3];   // lattice block dims
  std::vector<double> lat((size_t)(g.nb * nv * L3 * L2 * L1));
  int rc = copy_array(m, arr, lat.data(), false);
  if (rc) return rc;
  // lattice element of global (cell or face) index I along d; second copy of a face shared
  // by two lattice blocks (the lower block's upper face), or -1
  auto lat_idx = [&](int d, int I, int &lb, int &loc, int &lb2, int &loc2) {
    I = ((I % N[d]) + N[d]) % N[d];
    lb = I / bn[d];
    loc = I % bn[d] + ng;
    lb2 = loc2 = -1;
    if (d == fdir && I % bn[d] == 0) {
      lb2 = (lb - 1 + nl[d]) % nl[d];
      loc2 = bn[d] + ng;
    }
  };
  for (int b = 0; b < nblocks; ++b) {
    int n[3];
    for (int d = 0; d < 3; ++d) n[d] = size[3 * b + d] + 2 * ng + (d == fdir ? 1 : 0);
    double *h = host[b];
    for (int v = 0; v < nv; ++v)
      for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
          for (int i = 0; i < n[0]; ++i) {
            const int loc3[3] = {i, j, k};
            bool interior = true;
            for (int d = 0; d < 3; ++d) {
              const int hi = ng + size[3 * b + d] - 1 + (d == fdir ? 1 : 0);
              if (loc3[d] < ng || loc3[d] > hi) interior = false;
            }
            if (to_device && !interior) continue;
            int lb[3], lc[3], lb2[3], lc2[3];
            for (int d = 0; d < 3; ++d) lat_idx(d, origin[3 * b + d] + loc3[d] - ng, lb[d], lc[d], lb2[d], lc2[d]);
            const int64_t blk = ((int64_t)lb[2] * nl[1] + lb[1]) * nl[0] + lb[0];
            const int64_t li = (((blk * nv + v) * L3 + lc[2]) * L2 + lc[1]) * L1 + lc[0];
            double &hv = h[(((int64_t)v * n[2] + k) * n[1] + j) * n[0] + i];
            if (!to_device) { hv = lat[li]; continue; }
            lat[li] = hv;
            if (fdir >= 0 && lb2[fdir] >= 0) {   // the same face in the lower lattice block
              int b2[3] = {lb[0], lb[1], lb[2]}, c2[3] = {lc[0], lc[1], lc[2]};
              b2[fdir] = lb2[fdir];
              c2[fdir] = lc2[fdir];
              const int64_t blk2 = ((int64_t)b2[2] * nl[1] + b2[1]) * nl[0] + b2[0];
              lat[(((blk2 * nv + v) * L3 + c2[2]) * L2 + c2[1]) * L1 + c2[0]] = hv;
            }
          }
  }
  return to_device ? copy_array(m, arr, lat.data(), true) : 0;
}

int agx_upload_blocks(agx_mesh *m, int arr, int32_t nblocks, const int32_t *origin, const int32_t *size,
                      const double *const *host) {
  if (m && (arr == AGX_ARR_B1F || arr == AGX_ARR_B2F || arr == AGX_ARR_B3F)) m->copies_state = -1;
  return copy_blocks(m, arr, nblocks, origin, size, const_cast<double *const *>(host), true);
}

int agx_download_blocks(agx_mesh *m, int arr, int32_t nblocks, const int32_t *origin, const int32_t *size,
                        double *const *host) {
  return copy_blocks(m, arr, nblocks, origin, size, host, false);
}

int agx_set_time(agx_mesh *m, double time, double dt, int32_t ncycle) {
  if (!m) { set_error("null mesh"); return -1; }
  m->time = time; m->dt = dt; m->ncycle = ncycle;
  return 0;
}
int agx_get_time(agx_mesh *m, double *time, double *dt, int32_t *ncycle) {
  if (!m) { set_error("null mesh"); return -1; }
  if (time) *time = m->time;
  if (dt) *dt = m->dt;
  if (ncycle) *ncycle = m->ncycle;
  return 0;
}
int agx_synchronize(agx_mesh *m) {
  if (!m) { set_error("null mesh"); return -1; }
  AGX_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

#define LAUNCH_CHECK(what)                                     \
  do {                                                         \
    hipError_t e_ = hipGetLastError();                         \
    if (e_ != hipSuccess) return agx::hip_fail(e_, what);      \
  } while (0)

// Writes the host's dt, time and tlim into the device scalars read by the flux and stage
// kernels and by k_advance.
__global__ void k_set_scalars(double *__restrict__ s, double dt, double time, double tlim) {
  if (threadIdx.x == 0) {
    s[SC_DT] = dt;
    s[SC_TIME] = time;
    s[SC_TLIM] = tlim;
  }
}

// End of a cycle on the device: time += dt, then Mesh::NewTimeStep (mesh.cpp:1078-1119)
// from the block minimum in s[0] (k_min_final): dt = min(cfl * min, 2 dt), clipped to
// tlim - time.  The same operations as agx_cycle's host path.
__global__ void k_advance(double *__restrict__ s, double cfl) {
  if (threadIdx.x == 0) {
    const double dt = s[SC_DT], tlim = s[SC_TLIM];
    const double time = s[SC_TIME] + dt;
    const double bdt = s[0] * cfl;
    double ndt = 2.0 * dt;
    ndt = (bdt < ndt) ? bdt : ndt;
    if (time < tlim && (tlim - time) < ndt) ndt = tlim - time;
    s[SC_TIME] = time;
    s[SC_DT] = ndt;
  }
}

// The two copies of every shared face (the upper boundary face of a block and the lower
// boundary face of its lattice neighbour, periodic), compared bit for bit over the
// face's interior; any difference clears *flag.  One thread per face.
extern "C++" {
template <int DIR>
__global__ void k_face_copies(Geo g, const double *__restrict__ f, double *__restrict__ flag) {
  PlaneIdx p;
  if (DIR == 0 && !plane_index(g.js, g.je, g.ks, g.ke, 0, p)) return;
  if (DIR == 1 && !plane_index(g.is, g.ie, g.ks, g.ke, 0, p)) return;
  if (DIR == 2 && !plane_index(g.is, g.ie, g.js, g.je, 0, p)) return;
  int64_t up, lo;
  int nb;
  if (DIR == 0) {
    up = IDX(g, p.j, p.i, g.ie + 1); lo = IDX(g, p.j, p.i, g.is); nb = nbr_block(g, p.b, 1, 0, 0);
  } else if (DIR == 1) {
    up = IDX(g, p.j, g.je + 1, p.i); lo = IDX(g, p.j, g.js, p.i); nb = nbr_block(g, p.b, 0, 1, 0);
  } else {
    up = IDX(g, g.ke + 1, p.j, p.i); lo = IDX(g, g.ks, p.j, p.i); nb = nbr_block(g, p.b, 0, 0, 1);
  }
  const long long x = __double_as_longlong(f[(int64_t)p.b * g.S + up]);
  const long long y = __double_as_longlong(f[(int64_t)nb * g.S + lo]);
  if (x != y) *flag = 0.0;
}
}  // extern "C++"

__global__ void k_set_value(double *__restrict__ p, double v) {
  if (threadIdx.x == 0) *p = v;
}

// Whether the two copies of every shared face agree (m->copies_state), checked once after
// each upload of a face array.  While they agree, every block computes the same value for
// a shared edge, CT changes both copies alike and they keep agreeing, so the stage kernels
// push each cell's floored u, w and bcc to the ghosts that copy it instead of converting
// the ghosts afterwards from the block's own faces (the same bits).
constexpr int SC_COPIES = 48;
static int ensure_copies_checked(agx_mesh *m) {
  if (m->copies_state >= 0) return 0;
  if (m->tree) { m->copies_state = 0; return 0; }
  const Geo &g = m->g;
  double **A = m->arr.a;
  double *flag = m->scal + SC_COPIES;
  hipLaunchKernelGGL(k_set_value, dim3(1), dim3(64), 0, m->stream, flag, 1.0);
  hipLaunchKernelGGL(k_face_copies<0>, plane_grid(g.js, g.je, g.ks, g.ke, 0, 0, g.nb), dim3(256), 0, m->stream, g,
                     A[AGX_ARR_B1F], flag);
  hipLaunchKernelGGL(k_face_copies<1>, plane_grid(g.is, g.ie, g.ks, g.ke, 0, 0, g.nb), dim3(256), 0, m->stream, g,
                     A[AGX_ARR_B2F], flag);
  hipLaunchKernelGGL(k_face_copies<2>, plane_grid(g.is, g.ie, g.js, g.je, 0, 0, g.nb), dim3(256), 0, m->stream, g,
                     A[AGX_ARR_B3F], flag);
  LAUNCH_CHECK("k_face_copies");
  AGX_HIP(hipMemcpyAsync(m->scal_host + SC_COPIES, flag, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  AGX_HIP(hipStreamSynchronize(m->stream));
  m->copies_state = m->scal_host[SC_COPIES] == 1.0 ? 1 : 0;
  m->graph_gen++;
  return 0;
}

static int push_scalars(agx_mesh *m, double tlim) {
  hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, m->stream, m->scal, m->dt, m->time, tlim);
  LAUNCH_CHECK("k_set_scalars");
  return 0;
}

static int calculate_fluxes(agx_mesh *m, int order);

int agx_calculate_fluxes(agx_mesh *m, int order) {
  if (!m) { set_error("null mesh"); return -1; }
  int rc = push_scalars(m, DBL_MAX);
  return rc ? rc : calculate_fluxes(m, order);
}

extern "C++" {
// k_flux in direction DIR from the face-centred field B to the face arrays F, EY, EZ, WX;
// FP32 PPM adds the k_flux_fix pass over the faces k_flux flagged.
template <int DIR, int ORD, int RS, typename TW, bool ISO>
static void launch_flux(agx_mesh *m, int B, int F, int EY, int EZ, int WX) {
  const Geo &g = m->g;
  double **A = m->arr.a;
  hipLaunchKernelGGL((k_flux<DIR, ORD, RS, TW, ISO>),
                     plane_grid(g.is - (DIR != 0), g.ie + 1, g.js - (DIR != 1), g.je + 1, g.ks - (DIR != 2),
                                g.ke + 1, g.nb),
                     dim3(256), 0, m->stream, g, as<TW>(A[AGX_ARR_W]), as<TW>(A[AGX_ARR_BCC]), A[B], as<TW>(A[F]),
                     as<TW>(A[EY]), as<TW>(A[EZ]), as<TW>(A[WX]), m->coef, m->plm, m->d.gamma, m->d.iso_cs,
                     m->scal + SC_DT, m->d.dfloor, m->d.pfloor, m->nearf);
  if constexpr (flux_flagged<ORD, TW>()) {
    hipLaunchKernelGGL((k_flux_fix<DIR, ORD, RS, ISO>), dim3((unsigned)(m->nearf_n / FIX_SPAN)), dim3(256), 0,
                       m->stream, g, as<float>(A[AGX_ARR_W]), as<float>(A[AGX_ARR_BCC]), A[B], as<float>(A[F]),
                       as<float>(A[EY]), as<float>(A[EZ]), as<float>(A[WX]), m->coef, m->plm, m->d.gamma,
                       m->d.iso_cs, m->scal + SC_DT, m->d.dfloor, m->d.pfloor, m->nearf, m->nearf_n);
  }
}
}  // extern "C++"

static int calculate_fluxes(agx_mesh *m, int order) {
  if (order < 1 || order > 3) { set_error("order %d not supported", order); return -2; }
  if (order == 3 && m->g.ng < 3) { set_error("PPM needs nghost >= 3 (reconstruction.cpp:90-100)"); return -3; }
  const int rs = m->d.rsolver;
#define FLUX_LAUNCH_T(DIR, ORD, RS, B, F, EY, EZ, WX, TW, ISO)                                 \
  launch_flux<DIR, ORD, RS, TW, ISO>(m, B, F, EY, EZ, WX)
#define FLUX_LAUNCH_E(DIR, ORD, RS, B, F, EY, EZ, WX, ISO)                                     \
  if (m->mixed) FLUX_LAUNCH_T(DIR, ORD, RS, B, F, EY, EZ, WX, float, ISO);                    \
  else FLUX_LAUNCH_T(DIR, ORD, RS, B, F, EY, EZ, WX, double, ISO)
#define FLUX_LAUNCH(DIR, ORD, RS, B, F, EY, EZ, WX)                                            \
  if (m->iso) { FLUX_LAUNCH_E(DIR, ORD, (RS == AGX_RSOLVER_HLLE ? AGX_RSOLVER_HLLE : AGX_RSOLVER_HLLD), \
                              B, F, EY, EZ, WX, true); }                                       \
  else { FLUX_LAUNCH_E(DIR, ORD, RS, B, F, EY, EZ, WX, false); }
#define FLUX_ORD(DIR, ORD, B, F, EY, EZ, WX)                                                    \
  if (rs == AGX_RSOLVER_HLLE) { FLUX_LAUNCH(DIR, ORD, AGX_RSOLVER_HLLE, B, F, EY, EZ, WX); }    \
  else if (rs == AGX_RSOLVER_ROE) { FLUX_LAUNCH(DIR, ORD, AGX_RSOLVER_ROE, B, F, EY, EZ, WX); } \
  else if (rs == AGX_RSOLVER_LHLLD) { FLUX_LAUNCH_E(DIR, ORD, AGX_RSOLVER_LHLLD, B, F, EY, EZ, WX, false); } \
  else { FLUX_LAUNCH(DIR, ORD, AGX_RSOLVER_HLLD, B, F, EY, EZ, WX); }
  // xorder "2c"/"3c": PLM/PPM limited in characteristic variables
  const bool chp = m->d.char_proj && order > 1;
#define FLUX_DIR(DIR, B, F, EY, EZ, WX)                                                         \
  timer_begin(m, AGX_K_FLUX_X1 + DIR);                                                        \
  if (order == 1) {                                                                           \
    FLUX_ORD(DIR, 1, B, F, EY, EZ, WX)                                                        \
  } else if (order == 2 && chp) {                                                             \
    FLUX_ORD(DIR, ORDER_PLM_CHAR, B, F, EY, EZ, WX)                                           \
  } else if (order == 2) {                                                                    \
    FLUX_ORD(DIR, 2, B, F, EY, EZ, WX)                                                        \
  } else if (chp) {                                                                           \
    FLUX_ORD(DIR, ORDER_PPM_CHAR, B, F, EY, EZ, WX)                                           \
  } else {                                                                                    \
    FLUX_ORD(DIR, 3, B, F, EY, EZ, WX)                                                        \
  }                                                                                           \
  timer_end(m);                                                                               \
  LAUNCH_CHECK("k_flux");
  // face EMF targets: x1 -> (e3_x1f, e2_x1f), x2 -> (e1_x2f, e3_x2f), x3 -> (e2_x3f, e1_x3f)
  // (calculate_fluxes.cpp:86,200,306)
  FLUX_DIR(0, AGX_ARR_B1F, AGX_ARR_FX1, AGX_ARR_E3X1, AGX_ARR_E2X1, AGX_ARR_WX1)
  FLUX_DIR(1, AGX_ARR_B2F, AGX_ARR_FX2, AGX_ARR_E1X2, AGX_ARR_E3X2, AGX_ARR_WX2)
  FLUX_DIR(2, AGX_ARR_B3F, AGX_ARR_FX3, AGX_ARR_E2X3, AGX_ARR_E1X3, AGX_ARR_WX3)
#undef FLUX_DIR
#undef FLUX_ORD
#undef FLUX_LAUNCH
#undef FLUX_LAUNCH_E
#undef FLUX_LAUNCH_T
  return 0;
}

int agx_compute_corner_e(agx_mesh *m) {
  if (!m) { set_error("null mesh"); return -1; }
  if (m->mixed) { set_error("the task-level corner-EMF entry point runs in FP64 only"); return -2; }
  const Geo &g = m->g;
  double **A = m->arr.a;
  timer_begin(m, AGX_K_CORNER_E);
  hipLaunchKernelGGL(k_corner_e, plane_grid(g.is, g.ie + 1, g.js, g.je + 1, g.ks, g.ke + 1, g.nb),
                     dim3(256), 0, m->stream, g, A[AGX_ARR_W], A[AGX_ARR_BCC], A[AGX_ARR_E3X1],
                     A[AGX_ARR_E2X1], A[AGX_ARR_E1X2], A[AGX_ARR_E3X2], A[AGX_ARR_E1X3],
                     A[AGX_ARR_E2X3], A[AGX_ARR_WX1], A[AGX_ARR_WX2], A[AGX_ARR_WX3], A[AGX_ARR_E1],
                     A[AGX_ARR_E2], A[AGX_ARR_E3]);
  timer_end(m);
  LAUNCH_CHECK("k_corner_e");
  return 0;
}

int agx_integrate(agx_mesh *m, int stage) {
  if (!m) { set_error("null mesh"); return -1; }
  if (stage != 1 && stage != 2) { set_error("VL2 has stages 1 and 2, got %d", stage); return -2; }
  if (m->mixed) { set_error("the task-level integrate entry point runs in FP64 only"); return -3; }
  const Geo &g = m->g;
  double **A = m->arr.a;
  // VL2 stage weights, time_integrator.cpp:199-234: beta = 0.5, 1.0
  const double wght = (stage == 1 ? 0.5 : 1.0) * m->dt;
  const int uin = (stage == 1) ? AGX_ARR_U : AGX_ARR_U1;
  timer_begin(m, AGX_K_INTEGRATE);
  hipLaunchKernelGGL(k_integrate_u, plane_grid(g.is, g.ie, g.js, g.je, g.ks, g.ke, g.nb), dim3(256), 0,
                     m->stream, g, A[uin], A[AGX_ARR_U1], A[AGX_ARR_FX1], A[AGX_ARR_FX2], A[AGX_ARR_FX3],
                     A[AGX_ARR_W], A[AGX_ARR_PHI], m->coef, wght, m->d.self_gravity ? 1 : 0, m->iso ? 1 : 0);
  timer_end(m);
  LAUNCH_CHECK("k_integrate_u");
  const int b1 = (stage == 1) ? AGX_ARR_B1F : AGX_ARR_R1F;
  const int b2 = (stage == 1) ? AGX_ARR_B2F : AGX_ARR_R2F;
  const int b3 = (stage == 1) ? AGX_ARR_B3F : AGX_ARR_R3F;
  timer_begin(m, AGX_K_INTEGRATE);
  hipLaunchKernelGGL(k_ct<0>, plane_grid(g.is, g.ie + 1, g.js, g.je, g.ks, g.ke, g.nb), dim3(256), 0,
                     m->stream, g, A[b1], A[AGX_ARR_R1F], A[AGX_ARR_E1], A[AGX_ARR_E2], A[AGX_ARR_E3],
                     m->coef, wght);
  hipLaunchKernelGGL(k_ct<1>, plane_grid(g.is, g.ie, g.js, g.je + 1, g.ks, g.ke, g.nb), dim3(256), 0,
                     m->stream, g, A[b2], A[AGX_ARR_R2F], A[AGX_ARR_E1], A[AGX_ARR_E2], A[AGX_ARR_E3],
                     m->coef, wght);
  hipLaunchKernelGGL(k_ct<2>, plane_grid(g.is, g.ie, g.js, g.je, g.ks, g.ke + 1, g.nb), dim3(256), 0,
                     m->stream, g, A[b3], A[AGX_ARR_R3F], A[AGX_ARR_E1], A[AGX_ARR_E2], A[AGX_ARR_E3],
                     m->coef, wght);
  timer_end(m);
  LAUNCH_CHECK("k_ct");
  // gamma = (0,1,0): swap registers (time_integrator.cpp:1583-1584,1637-1640)
  std::swap(A[AGX_ARR_U], A[AGX_ARR_U1]);
  std::swap(A[AGX_ARR_B1F], A[AGX_ARR_R1F]);
  std::swap(A[AGX_ARR_B2F], A[AGX_ARR_R2F]);
  std::swap(A[AGX_ARR_B3F], A[AGX_ARR_R3F]);
  return 0;
}

int agx_bvals_exchange(agx_mesh *m, int which) {
  if (!m) { set_error("null mesh"); return -1; }
  if (int rc = refuse_tree(m, "agx_bvals_exchange")) return rc;
  const Geo &g = m->g;
  double **A = m->arr.a;
  timer_begin(m, AGX_K_BVALS);
  if (which & 1)
    hipLaunchKernelGGL(k_bvals_cc, plane_grid(0, g.n1 - 1, 0, g.n2 - 1, 0, g.n3 - 1, g.nb), dim3(256), 0,
                       m->stream, g, A[AGX_ARR_U], NHYDRO);
  if (which & 2) {
    hipLaunchKernelGGL(k_bvals_fc<0>, plane_grid(0, g.n1, 0, g.n2 - 1, 0, g.n3 - 1, g.nb), dim3(256), 0,
                       m->stream, g, A[AGX_ARR_B1F]);
    hipLaunchKernelGGL(k_bvals_fc<1>, plane_grid(0, g.n1 - 1, 0, g.n2, 0, g.n3 - 1, g.nb), dim3(256), 0,
                       m->stream, g, A[AGX_ARR_B2F]);
    hipLaunchKernelGGL(k_bvals_fc<2>, plane_grid(0, g.n1 - 1, 0, g.n2 - 1, 0, g.n3, g.nb), dim3(256), 0,
                       m->stream, g, A[AGX_ARR_B3F]);
  }
  if (which & 4)
    hipLaunchKernelGGL(k_bvals_cc, plane_grid(0, g.n1 - 1, 0, g.n2 - 1, 0, g.n3 - 1, g.nb), dim3(256), 0,
                       m->stream, g, A[AGX_ARR_PHI], 1);
  timer_end(m);
  LAUNCH_CHECK("k_bvals");
  return 0;
}

extern "C++" {
template <typename TW>
static void launch_cons2prim(agx_mesh *m) {
  const Geo &g = m->g;
  double **A = m->arr.a;
  const dim3 grid = plane_grid(0, g.n1 - 1, 0, g.n2 - 1, 0, g.n3 - 1, g.nb);
  if (m->iso)
    hipLaunchKernelGGL((k_cons2prim<TW, true>), grid, dim3(256), 0, m->stream, g, A[AGX_ARR_U],
                       as<TW>(A[AGX_ARR_W]), as<TW>(A[AGX_ARR_BCC]), A[AGX_ARR_B1F], A[AGX_ARR_B2F],
                       A[AGX_ARR_B3F], m->d.gamma - 1.0, m->d.dfloor, m->d.pfloor);
  else
    hipLaunchKernelGGL((k_cons2prim<TW, false>), grid, dim3(256), 0, m->stream, g, A[AGX_ARR_U],
                       as<TW>(A[AGX_ARR_W]), as<TW>(A[AGX_ARR_BCC]), A[AGX_ARR_B1F], A[AGX_ARR_B2F],
                       A[AGX_ARR_B3F], m->d.gamma - 1.0, m->d.dfloor, m->d.pfloor);
}
}  // extern "C++"

int agx_cons2prim(agx_mesh *m) {
  if (!m) { set_error("null mesh"); return -1; }
  timer_begin(m, AGX_K_C2P);
  if (m->mixed) launch_cons2prim<float>(m);
  else launch_cons2prim<double>(m);
  timer_end(m);
  LAUNCH_CHECK("k_cons2prim");
  return 0;
}

// min over the nwg partials in part -> host; returns min_dt_hyperbolic * cfl
// (new_blockdt.cpp:158).
static int finish_block_dt(agx_mesh *m, const double *part, int64_t nwg, double *min_dt) {
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(256), 0, m->stream, part, nwg, m->scal);
  LAUNCH_CHECK("k_min_final");
  AGX_HIP(hipMemcpyAsync(m->scal_host, m->scal, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  AGX_HIP(hipStreamSynchronize(m->stream));
  double v = m->scal_host[0] * m->d.cfl;
  if (min_dt) *min_dt = v;
  return 0;
}

extern "C++" {
template <typename TW>
static void launch_newdt(agx_mesh *m, dim3 grid) {
  const Geo &g = m->g;
  double **A = m->arr.a;
  if (m->iso)
    hipLaunchKernelGGL((k_newdt<TW, true>), grid, dim3(256), 0, m->stream, g, as<TW>(A[AGX_ARR_W]),
                       as<TW>(A[AGX_ARR_BCC]), A[AGX_ARR_B1F], A[AGX_ARR_B2F], A[AGX_ARR_B3F], m->coef,
                       m->d.gamma, m->d.iso_cs, m->red);
  else
    hipLaunchKernelGGL((k_newdt<TW, false>), grid, dim3(256), 0, m->stream, g, as<TW>(A[AGX_ARR_W]),
                       as<TW>(A[AGX_ARR_BCC]), A[AGX_ARR_B1F], A[AGX_ARR_B2F], A[AGX_ARR_B3F], m->coef,
                       m->d.gamma, m->d.iso_cs, m->red);
}
}  // extern "C++"

int agx_new_block_dt(agx_mesh *m, double *min_dt) {
  if (!m) { set_error("null mesh"); return -1; }
  const Geo &g = m->g;
  dim3 grid = plane_grid(g.is, g.ie, g.js, g.je, g.ks, g.ke, g.nb);
  int64_t nwg = (int64_t)grid.x * grid.y * grid.z;
  timer_begin(m, AGX_K_NEWDT);
  if (m->mixed) launch_newdt<float>(m, grid);
  else launch_newdt<double>(m, grid);
  LAUNCH_CHECK("k_newdt");
  int rc = finish_block_dt(m, m->red, nwg, min_dt);
  timer_end(m);
  return rc;
}

// Mesh::NewTimeStep, mesh/mesh.cpp:1083-1107
static void set_new_dt(agx_mesh *m, double bdt, double tlim) {
  double dt = 2.0 * m->dt;
  dt = (bdt < dt) ? bdt : dt;
  if (m->time < tlim && (tlim - m->time) < dt) dt = tlim - m->time;
  m->dt = dt;
}

int agx_new_time_step(agx_mesh *m, double tlim) {
  double bdt = 0.0;
  int rc = agx_new_block_dt(m, &bdt);
  if (rc) return rc;
  set_new_dt(m, bdt, tlim);
  return 0;
}

int agx_initialize(agx_mesh *m, double tlim) {
  if (!m) { set_error("null mesh"); return -1; }
  if (int rc = refuse_tree(m, "agx_initialize")) return rc;
  int rc = 0;
  if (m->d.self_gravity) rc = mg_solve(m, 0);          // mesh.cpp:1401-1402
  if (!rc) rc = agx_bvals_exchange(m, 3);               // mesh.cpp:1404-1453
  if (!rc) rc = agx_cons2prim(m);                       // mesh.cpp:1543-1545
  if (rc) return rc;
  // a fresh Mesh starts from dt = DBL_MAX (mesh.cpp:91); after agx_restart_read the dt of
  // the file caps the first step, as the restart constructor's dt does (mesh.cpp:667)
  if (!m->restarted) m->dt = DBL_MAX;
  m->restarted = false;
  return agx_new_time_step(m, tlim);                    // mesh.cpp:1646-1649
}

// One VL2 stage with the fused kernels (time_integrator.cpp:1460-2034): fluxes, then
// k_stage (corner EMFs, CT, update, C2P with floors, the neighbours' ghost copies of the
// new faces and of the un-floored conserved state and, when dtpart is set, the time-step
// partials), the register rotation, then C2P over the ghost zones.  Same results as
// agx_calculate_fluxes + agx_compute_corner_e + agx_integrate + agx_bvals_exchange +
// agx_cons2prim.
//
// Registers: stage 1 reads u and b and writes u1 and b1; the swap makes (u, b) the
// half-step state and (u1, b1) the start state.  Stage 2 updates u1 in place from the
// fluxes of the half-step state and writes the new faces to the third register; the
// rotation leaves b = new faces and b1 = half-step faces, as the reference's swap does.
extern "C++" {
template <typename TW, bool PUSH>
static void launch_stage_t(agx_mesh *m, int tile, const StageArgs<TW> &a, dim3 grid) {
  const Geo &g = m->g;
  if (m->iso) {
    if (tile == 1) hipLaunchKernelGGL((k_stage<32, 8, TW, true, PUSH>), grid, dim3(256), 0, m->stream, g, a);
    else if (tile == 2) hipLaunchKernelGGL((k_stage<64, 4, TW, true, PUSH>), grid, dim3(256), 0, m->stream, g, a);
    else if (tile == 3) hipLaunchKernelGGL((k_stage<64, 8, TW, true, PUSH>), grid, dim3(512), 0, m->stream, g, a);
    else hipLaunchKernelGGL((k_stage<16, 16, TW, true, PUSH>), grid, dim3(256), 0, m->stream, g, a);
  } else {
    if (tile == 1) hipLaunchKernelGGL((k_stage<32, 8, TW, false, PUSH>), grid, dim3(256), 0, m->stream, g, a);
    else if (tile == 2) hipLaunchKernelGGL((k_stage<64, 4, TW, false, PUSH>), grid, dim3(256), 0, m->stream, g, a);
    else if (tile == 3) hipLaunchKernelGGL((k_stage<64, 8, TW, false, PUSH>), grid, dim3(512), 0, m->stream, g, a);
    else hipLaunchKernelGGL((k_stage<16, 16, TW, false, PUSH>), grid, dim3(256), 0, m->stream, g, a);
  }
}

template <typename TW>
static void launch_stage(agx_mesh *m, int tile, int stage, double wfac, double *dtpart) {
  const Geo &g = m->g;
  double **A = m->arr.a;
  StageArgs<TW> a;
  a.src = (stage == 1) ? A[AGX_ARR_U] : A[AGX_ARR_U1];
  a.dst = A[AGX_ARR_U1];
  a.rho = A[AGX_ARR_U];
  a.phi = A[AGX_ARR_PHI];
  a.f1 = as<TW>(A[AGX_ARR_FX1]); a.f2 = as<TW>(A[AGX_ARR_FX2]); a.f3 = as<TW>(A[AGX_ARR_FX3]);
  a.e3x1 = as<TW>(A[AGX_ARR_E3X1]); a.e2x1 = as<TW>(A[AGX_ARR_E2X1]);
  a.e1x2 = as<TW>(A[AGX_ARR_E1X2]); a.e3x2 = as<TW>(A[AGX_ARR_E3X2]);
  a.e1x3 = as<TW>(A[AGX_ARR_E1X3]); a.e2x3 = as<TW>(A[AGX_ARR_E2X3]);
  a.wx1 = as<TW>(A[AGX_ARR_WX1]); a.wx2 = as<TW>(A[AGX_ARR_WX2]); a.wx3 = as<TW>(A[AGX_ARR_WX3]);
  a.w = as<TW>(A[AGX_ARR_W]); a.bcc = as<TW>(A[AGX_ARR_BCC]);
  a.wn = as<TW>(A[ARR_W2]); a.bccn = as<TW>(A[ARR_BCC2]);
  const int src = (stage == 1) ? AGX_ARR_B1F : AGX_ARR_R1F;
  const int dst = (stage == 1) ? AGX_ARR_R1F : ARR_Q1F;
  a.s1 = A[src]; a.s2 = A[src + 1]; a.s3 = A[src + 2];
  a.d1 = A[dst]; a.d2 = A[dst + 1]; a.d3 = A[dst + 2];
  a.coef = m->coef;
  a.dtp = m->scal + SC_DT;
  a.wfac = wfac;
  a.gm1 = m->d.gamma - 1.0;
  a.dfl = m->d.dfloor;
  a.pfl = m->d.pfloor;
  a.gamma = m->d.gamma;
  a.cs = m->d.iso_cs;
  a.grav = m->d.self_gravity ? 1 : 0;
  a.dtpart = dtpart;
  const dim3 grid = stage_grid(g, tile);
  if (m->copies_state == 1) launch_stage_t<TW, true>(m, tile, a, grid);
  else launch_stage_t<TW, false>(m, tile, a, grid);
}
}  // extern "C++"


// The whole stage as one k_vl2 launch (fused form); tile 0 = 16 x 16, 1 = 32 x 8 columns.
extern "C++" {
template <int TI, int TJ, int ORDER, typename TW, bool ISO, bool PUSH>
static void launch_vl2_rs(agx_mesh *m, int rs, const VL2Args<TW> &a, dim3 grid) {
  if (ISO) {
    if (rs == AGX_RSOLVER_HLLE)
      hipLaunchKernelGGL((k_vl2<TI, TJ, ORDER, AGX_RSOLVER_HLLE, TW, ISO, PUSH>), grid, dim3(256), 0, m->stream, m->g, a);
    else
      hipLaunchKernelGGL((k_vl2<TI, TJ, ORDER, AGX_RSOLVER_HLLD, TW, ISO, PUSH>), grid, dim3(256), 0, m->stream, m->g, a);
  } else {
    if (rs == AGX_RSOLVER_HLLE)
      hipLaunchKernelGGL((k_vl2<TI, TJ, ORDER, AGX_RSOLVER_HLLE, TW, ISO, PUSH>), grid, dim3(256), 0, m->stream, m->g, a);
    else if (rs == AGX_RSOLVER_ROE)
      hipLaunchKernelGGL((k_vl2<TI, TJ, ORDER, AGX_RSOLVER_ROE, TW, ISO, PUSH>), grid, dim3(256), 0, m->stream, m->g, a);
    else if (rs == AGX_RSOLVER_LHLLD)
      hipLaunchKernelGGL((k_vl2<TI, TJ, ORDER, AGX_RSOLVER_LHLLD, TW, ISO, PUSH>), grid, dim3(256), 0, m->stream, m->g, a);
    else
      hipLaunchKernelGGL((k_vl2<TI, TJ, ORDER, AGX_RSOLVER_HLLD, TW, ISO, PUSH>), grid, dim3(256), 0, m->stream, m->g, a);
  }
}

template <int TI, int TJ, typename TW, bool ISO>
static void launch_vl2_order(agx_mesh *m, int order, const VL2Args<TW> &a, dim3 grid) {
  if (m->copies_state == 1) {
    if (order == 1) launch_vl2_rs<TI, TJ, 1, TW, ISO, true>(m, m->d.rsolver, a, grid);
    else if (order == 2) launch_vl2_rs<TI, TJ, 2, TW, ISO, true>(m, m->d.rsolver, a, grid);
    else launch_vl2_rs<TI, TJ, 3, TW, ISO, true>(m, m->d.rsolver, a, grid);
  } else {
    if (order == 1) launch_vl2_rs<TI, TJ, 1, TW, ISO, false>(m, m->d.rsolver, a, grid);
    else if (order == 2) launch_vl2_rs<TI, TJ, 2, TW, ISO, false>(m, m->d.rsolver, a, grid);
    else launch_vl2_rs<TI, TJ, 3, TW, ISO, false>(m, m->d.rsolver, a, grid);
  }
}

template <typename TW>
static void launch_vl2(agx_mesh *m, int tile, int stage, int order, double wfac, double *dtpart) {
  const Geo &g = m->g;
  double **A = m->arr.a;
  VL2Args<TW> a;
  a.w = as<TW>(A[AGX_ARR_W]); a.bcc = as<TW>(A[AGX_ARR_BCC]);
  a.wn = as<TW>(A[ARR_W2]); a.bccn = as<TW>(A[ARR_BCC2]);
  a.bn1 = A[AGX_ARR_B1F]; a.bn2 = A[AGX_ARR_B2F]; a.bn3 = A[AGX_ARR_B3F];
  a.src = (stage == 1) ? A[AGX_ARR_U] : A[AGX_ARR_U1];
  a.dst = A[AGX_ARR_U1];
  a.rho = A[AGX_ARR_U];
  a.phi = A[AGX_ARR_PHI];
  const int src = (stage == 1) ? AGX_ARR_B1F : AGX_ARR_R1F;
  const int dst = (stage == 1) ? AGX_ARR_R1F : ARR_Q1F;
  a.s1 = A[src]; a.s2 = A[src + 1]; a.s3 = A[src + 2];
  a.d1 = A[dst]; a.d2 = A[dst + 1]; a.d3 = A[dst + 2];
  a.coef = m->coef;
  a.plm = m->plm;
  a.dtp = m->scal + SC_DT;
  a.wfac = wfac;
  a.gm1 = m->d.gamma - 1.0;
  a.dfl = m->d.dfloor;
  a.pfl = m->d.pfloor;
  a.gamma = m->d.gamma;
  a.cs = m->d.iso_cs;
  a.grav = m->d.self_gravity ? 1 : 0;
  a.dtpart = dtpart;
  const dim3 grid = stage_grid(g, tile);
  if (m->iso) {
    if (tile == 5) launch_vl2_order<32, 8, TW, true>(m, order, a, grid);
    else launch_vl2_order<16, 16, TW, true>(m, order, a, grid);
  } else {
    if (tile == 5) launch_vl2_order<32, 8, TW, false>(m, order, a, grid);
    else launch_vl2_order<16, 16, TW, false>(m, order, a, grid);
  }
}
}  // extern "C++"

extern "C++" {
template <typename TW>
static void launch_ghost_c2p(agx_mesh *m) {
  const Geo &g = m->g;
  double **A = m->arr.a;
  const Shell cs = make_shell(g.n1, g.n2, g.n3, g.is, g.ie, g.js, g.je, g.ks, g.ke);
  const dim3 grid(grid1d(cs.nk + cs.nj + cs.ni, 256), 1, g.nb);
  if (m->iso)
    hipLaunchKernelGGL((k_ghost_c2p<TW, true>), grid, dim3(256), 0, m->stream, g, cs, A[AGX_ARR_U],
                       as<TW>(A[AGX_ARR_W]), as<TW>(A[AGX_ARR_BCC]), A[AGX_ARR_B1F], A[AGX_ARR_B2F],
                       A[AGX_ARR_B3F], m->d.gamma - 1.0, m->d.dfloor, m->d.pfloor);
  else
    hipLaunchKernelGGL((k_ghost_c2p<TW, false>), grid, dim3(256), 0, m->stream, g, cs, A[AGX_ARR_U],
                       as<TW>(A[AGX_ARR_W]), as<TW>(A[AGX_ARR_BCC]), A[AGX_ARR_B1F], A[AGX_ARR_B2F],
                       A[AGX_ARR_B3F], m->d.gamma - 1.0, m->d.dfloor, m->d.pfloor);
}
}  // extern "C++"

static int stage_fused(agx_mesh *m, int stage, bool want_dt) {
  double **A = m->arr.a;
  const int order = (stage == 1) ? 1 : m->d.xorder;
  const double wfac = (stage == 1 ? 0.5 : 1.0);   // wght = wfac * dt on the device
  // each stage runs the form tuned for its reconstruction order; the last stage's form
  // sets the layout of the time-step partials (m->tile)
  const int tile = (stage == 1) ? m->tile1 : m->tile;
  if (tile >= TILE_VL2) {
    if (order == 3 && m->g.ng < 3) { set_error("PPM needs nghost >= 3 (reconstruction.cpp:90-100)"); return -3; }
    if (order > 1 && m->d.char_proj) {
      set_error("characteristic projection runs in the split stage forms only"); return -4;
    }
    timer_begin(m, AGX_K_INTEGRATE);
    if (m->mixed) launch_vl2<float>(m, tile, stage, order, wfac, want_dt ? m->dtpart : nullptr);
    else launch_vl2<double>(m, tile, stage, order, wfac, want_dt ? m->dtpart : nullptr);
    timer_end(m);
    LAUNCH_CHECK("k_vl2");
  } else {
    int rc = calculate_fluxes(m, order);
    if (rc) return rc;
    timer_begin(m, AGX_K_INTEGRATE);
    if (m->mixed) launch_stage<float>(m, tile, stage, wfac, want_dt ? m->dtpart : nullptr);
    else launch_stage<double>(m, tile, stage, wfac, want_dt ? m->dtpart : nullptr);
    timer_end(m);
    LAUNCH_CHECK("k_stage");
  }
  std::swap(A[AGX_ARR_U], A[AGX_ARR_U1]);
  std::swap(A[AGX_ARR_W], A[ARR_W2]);
  std::swap(A[AGX_ARR_BCC], A[ARR_BCC2]);
  for (int d = 0; d < 3; ++d) {
    if (stage == 1) {
      std::swap(A[AGX_ARR_B1F + d], A[AGX_ARR_R1F + d]);
    } else {
      double *half = A[AGX_ARR_B1F + d], *start = A[AGX_ARR_R1F + d];
      A[AGX_ARR_B1F + d] = A[ARR_Q1F + d];
      A[AGX_ARR_R1F + d] = half;
      A[ARR_Q1F + d] = start;
    }
  }
  if (m->copies_state == 1) return 0;   // the stage kernel pushed the ghosts' u, w and bcc
  timer_begin(m, AGX_K_BVALS);
  if (m->mixed) launch_ghost_c2p<float>(m);
  else launch_ghost_c2p<double>(m);
  timer_end(m);
  LAUNCH_CHECK("k_ghost_c2p");
  return 0;
}

static void graph_reset(agx_mesh *m) {
  for (auto &cg : m->graphs)
    if (cg.exec) hipGraphExecDestroy(cg.exec);
  m->graphs.clear();
  for (auto &sg : m->seggraphs)
    if (sg.exec) hipGraphExecDestroy(sg.exec);
  m->seggraphs.clear();
}

// One fused cycle with everything on the device (flux and stage kernels read dt from
// m->scal, k_advance forms the next dt and time), for stream capture.
static int cycle_device(agx_mesh *m) {
  const dim3 sg = stage_grid(m->g, m->tile);
  int rc = stage_fused(m, 1, false);
  if (!rc && m->d.self_gravity) rc = mg_solve(m, 1);   // main.cpp:464-465
  if (!rc) rc = stage_fused(m, 2, true);
  if (!rc && m->d.self_gravity) rc = mg_solve(m, 2);
  if (rc) return rc;
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(256), 0, m->stream, m->dtpart, (int64_t)sg.x * sg.y, m->scal);
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, m->stream, m->scal, m->d.cfl);
  LAUNCH_CHECK("k_advance");
  return 0;
}

// The executable graph of the cycle that starts from the current register assignment,
// captured on first use; the host's register table is left as the cycle leaves it.
static int graph_for_cycle(agx_mesh *m, hipGraphExec_t *exec) {
  double **A = m->arr.a;
  if (!m->graphs.empty() && m->graphs[0].gen != m->graph_gen) graph_reset(m);   // settings changed
  for (auto &cg : m->graphs)
    if (std::memcmp(cg.before, A, sizeof(cg.before)) == 0 && cg.coffset_before == m->coffset) {
      std::memcpy(A, cg.after, sizeof(cg.after));
      m->coffset = cg.coffset_after;
      *exec = cg.exec;
      return 0;
    }
  if (m->graphs.size() >= 4) graph_reset(m);
  agx_mesh::CycleGraph cg;
  std::memcpy(cg.before, A, sizeof(cg.before));
  cg.coffset_before = m->coffset;
  cg.gen = m->graph_gen;
  hipGraph_t graph = nullptr;
  AGX_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  m->capturing = true;
  int rc = cycle_device(m);
  m->capturing = false;
  hipError_t e = hipStreamEndCapture(m->stream, &graph);
  if (rc || e != hipSuccess) {
    std::memcpy(A, cg.before, sizeof(cg.before));
    m->coffset = cg.coffset_before;
    if (graph) hipGraphDestroy(graph);
    return rc ? rc : hip_fail(e, "hipStreamEndCapture");
  }
  e = hipGraphInstantiate(&cg.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) {
    std::memcpy(A, cg.before, sizeof(cg.before));
    m->coffset = cg.coffset_before;
    return hip_fail(e, "hipGraphInstantiate");
  }
  std::memcpy(cg.after, A, sizeof(cg.after));
  cg.coffset_after = m->coffset;
  m->graphs.push_back(cg);
  *exec = cg.exec;
  return 0;
}

// agx_cycle with each cycle replayed as one hipGraph launch; the host reads back dt and
// time after each cycle (the tlim test), as the eager path does.
static int cycle_graphs(agx_mesh *m, double tlim, int32_t ncycles, int32_t *done) {
  int32_t n = 0;
  int rc = push_scalars(m, tlim);
  while (!rc && n < ncycles && m->time < tlim) {
    hipGraphExec_t exec = nullptr;
    rc = graph_for_cycle(m, &exec);
    if (rc) break;
    hipError_t e = hipGraphLaunch(exec, m->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(m->scal_host + SC_DT, m->scal + SC_DT, 2 * sizeof(double), hipMemcpyDeviceToHost,
                         m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "cycle graph"); break; }
    m->dt = m->scal_host[SC_DT];
    m->time = m->scal_host[SC_TIME];
    m->ncycle++;
    if (m->d.self_gravity) {   // the recorded solves: a fixed V-cycle count, no norm read back
      m->last_norm = NAN;
      m->last_iters = m->mg_niter;
    }
    ++n;
  }
  if (done) *done = n;
  return rc;
}

// A cycle with threshold solves, in three segments that each end where the host learns
// how many V-cycles a solve ran (mg_solve_settle): 0 = stage 1 and the first solve up to
// the end of its recorded loop; 1 = that solve's tail, stage 2 and the second solve's
// loop; 2 = the second solve's tail and the next time step.
static int cycle_segment(agx_mesh *m, int seg, int k, std::vector<MGSnap> &snaps) {
  int rc = 0;
  if (seg > 0) rc = mg_solve_tail(m);
  if (!rc && seg < 2) rc = stage_fused(m, seg + 1, seg == 1);
  if (!rc && seg < 2) rc = mg_solve_head(m);
  if (rc) return rc;
  if (seg < 2) {
    mg_solve_gated(m, k, snaps);
    return 0;
  }
  const dim3 sg = stage_grid(m->g, m->tile);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(256), 0, m->stream, m->dtpart, (int64_t)sg.x * sg.y, m->scal);
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, m->stream, m->scal, m->d.cfl);
  LAUNCH_CHECK("k_advance");
  snaps.assign(1, mg_snap(m));
  return 0;
}

// Index into m->seggraphs of the graph of segment seg (k recorded V-cycles) that starts
// from the current register table and multigrid state, recorded on first use; the host
// state is left as the segment leaves it when all k V-cycles run.
static int graph_for_segment(agx_mesh *m, int seg, int k, int *index) {
  double **A = m->arr.a;
  if (!m->seggraphs.empty() && m->seggraphs[0].gen != m->graph_gen) graph_reset(m);   // settings changed
  const std::vector<uintptr_t> key = mg_key(m);
  for (size_t i = 0; i < m->seggraphs.size(); ++i) {
    agx_mesh::SegGraph &sgr = m->seggraphs[i];
    if (sgr.seg == seg && sgr.k == k && sgr.mgkey == key && std::memcmp(sgr.before, A, sizeof(sgr.before)) == 0) {
      std::memcpy(A, sgr.after, sizeof(sgr.after));
      mg_restore(m, sgr.snaps.back());
      *index = (int)i;
      return 0;
    }
  }
  if (m->seggraphs.size() >= 32) graph_reset(m);
  agx_mesh::SegGraph sgr;
  sgr.seg = seg;
  sgr.k = k;
  sgr.gen = m->graph_gen;
  sgr.mgkey = key;
  std::memcpy(sgr.before, A, sizeof(sgr.before));
  const MGSnap start = mg_snap(m);
  hipGraph_t graph = nullptr;
  AGX_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  m->capturing = true;
  int rc = cycle_segment(m, seg, k, sgr.snaps);
  m->capturing = false;
  hipError_t e = hipStreamEndCapture(m->stream, &graph);
  if (!rc && e == hipSuccess) {
    e = hipGraphInstantiate(&sgr.exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) sgr.exec = nullptr;
  }
  if (graph) hipGraphDestroy(graph);
  if (rc || e != hipSuccess) {
    std::memcpy(A, sgr.before, sizeof(sgr.before));
    mg_restore(m, start);
    m->gate = nullptr;
    return rc ? rc : hip_fail(e, "segment graph");
  }
  std::memcpy(sgr.after, A, sizeof(sgr.after));
  m->seggraphs.push_back(std::move(sgr));
  *index = (int)m->seggraphs.size() - 1;
  return 0;
}

// agx_cycle with threshold solves: three graph launches per cycle.  The loop of each solve
// is recorded with as many V-cycles as the same stage's solve ran last time (none before
// the first cycle); passes beyond the stop return at once, and a solve that needs more
// continues eagerly, so every cycle equals the eager one bit for bit.
static int cycle_graphs_threshold(agx_mesh *m, double tlim, int32_t ncycles, int32_t *done) {
  int32_t n = 0;
  int rc = push_scalars(m, tlim);
  while (!rc && n < ncycles && m->time < tlim) {
    for (int seg = 0; seg < 3 && !rc; ++seg) {
      const int k = seg < 2 ? std::max(m->mg_kpred[seg], 0) : 0;
      int idx = -1;
      rc = graph_for_segment(m, seg, k, &idx);
      if (rc) break;
      hipError_t e = hipGraphLaunch(m->seggraphs[idx].exec, m->stream);
      if (e != hipSuccess) { rc = hip_fail(e, "cycle segment graph"); break; }
      if (seg < 2) rc = mg_solve_settle(m, seg + 1, m->seggraphs[idx].snaps);
    }
    if (rc) break;
    hipError_t e = hipMemcpyAsync(m->scal_host + SC_DT, m->scal + SC_DT, 2 * sizeof(double),
                                  hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) { rc = hip_fail(e, "cycle graph"); break; }
    m->dt = m->scal_host[SC_DT];
    m->time = m->scal_host[SC_TIME];
    m->ncycle++;
    ++n;
  }
  if (done) *done = n;
  return rc;
}

int agx_set_graph(agx_mesh *m, int32_t enable) {
  if (!m) { set_error("null mesh"); return -1; }
  m->graph_on = enable != 0;
  if (!m->graph_on) graph_reset(m);
  return 0;
}

int agx_cycle(agx_mesh *m, double tlim, int32_t ncycles, int32_t *done) {
  if (!m) { set_error("null mesh"); return -1; }
  if (int rc = refuse_tree(m, "agx_cycle")) return rc;
  if (int rc = ensure_copies_checked(m)) return rc;
  // with self-gravity the fixed-count solve (niteration, SolveIterativeFixedTimes) is
  // recorded whole; the threshold form (SolveIterative) in segments that end where the
  // host reads how many V-cycles ran
  const bool sg_rec = m->d.self_gravity && m->mg_on;
  if (m->graph_on && !m->timer.enabled && (!m->d.self_gravity || (sg_rec && m->mg_eps < 0.0)))
    return cycle_graphs(m, tlim, ncycles, done);
  if (m->graph_on && !m->timer.enabled && sg_rec) return cycle_graphs_threshold(m, tlim, ncycles, done);
  int32_t n = 0;
  int rc = 0;
  const dim3 sg = stage_grid(m->g, m->tile);
  while (n < ncycles && m->time < tlim) {
    rc = push_scalars(m, tlim);
    for (int stage = 1; stage <= 2 && !rc; ++stage) {
      // VL2 stage 1 is donor cell (time_integrator.cpp:1466-1467), inside stage_fused;
      // the last stage also leaves the NewBlockTimeStep partials in m->dtpart
      rc = stage_fused(m, stage, stage == 2);
      if (!rc && m->d.self_gravity) rc = mg_solve(m, stage);   // main.cpp:464-465
    }
    if (rc) break;
    m->ncycle++;
    m->time += m->dt;
    double bdt = 0.0;
    timer_begin(m, AGX_K_NEWDT);
    rc = finish_block_dt(m, m->dtpart, (int64_t)sg.x * sg.y, &bdt);
    timer_end(m);
    if (rc) break;
    set_new_dt(m, bdt, tlim);
    ++n;
  }
  if (done) *done = n;
  return rc;
}

int agx_mg_solve(agx_mesh *m, int stage) {
  if (!m) { set_error("null mesh"); return -1; }
  if (!m->mg_on) { set_error("mesh created without self-gravity"); return -2; }
  return mg_solve(m, stage);
}

int agx_mg_set_mode(agx_mesh *m, int mode, double threshold, int32_t niter) {
  if (!m) { set_error("null mesh"); return -1; }
  if (mode != 0 && mode != 1) { set_error("mgmode must be 0 (FMG) or 1 (MGI)"); return -2; }
  if (threshold < 0.0 && niter < 0) {
    set_error("either threshold or niteration must be set (mg_gravity.cpp:55-63)"); return -3;
  }
  m->mg_mode = mode; m->mg_eps = threshold; m->mg_niter = niter;
  m->graph_gen++;
  return 0;
}

int agx_mg_last(agx_mesh *m, double *norm, int32_t *iters) {
  if (!m) { set_error("null mesh"); return -1; }
  if (norm) *norm = m->last_norm;
  if (iters) *iters = m->last_iters;
  return 0;
}

int agx_mg_vcycles(agx_mesh *m, int32_t n) {
  if (!m) { set_error("null mesh"); return -1; }
  if (!m->mg_on) { set_error("mesh created without self-gravity"); return -2; }
  return mg_vcycles(m, n);
}

// Stage-form autotuner (no counterpart in the reference).  Times every form on the
// current state with stage-1 launches at the mesh's reconstruction order (the split
// forms with their k_flux launches), which read u, w, bcc and b and write only scratch
// registers (fluxes, u1, b1, the second w/bcc), so the state is unchanged; keeps the
// fastest.  Call between cycles.
int agx_autotune(agx_mesh *m, int32_t *choice) {
  if (!m) { set_error("null mesh"); return -1; }
  if (int rc = refuse_tree(m, "agx_autotune")) return rc;
  if (int rc = ensure_copies_checked(m)) return rc;
  hipEvent_t e0, e1;
  AGX_HIP(hipEventCreate(&e0));
  AGX_HIP(hipEventCreate(&e1));
  const double wfac = 0.5;
  graph_reset(m);
  const bool timing = m->timer.enabled;
  m->timer.enabled = false;
  int best_of[4] = {0, 0, 0, 0};   // per reconstruction order
  for (int order : {1, m->d.xorder}) {
    float best = 0.f;
    int best_t = 0;
    // characteristic projection has the split forms only
    const int ntile = (order > 1 && m->d.char_proj) ? TILE_VL2 : NTILE;
    for (int t = 0; t < ntile; ++t) {
      float ms = 0.f;
      for (int rep = 0; rep < 4; ++rep) {
        if (rep == 1) AGX_HIP(hipEventRecord(e0, m->stream));
        if (t >= TILE_VL2) {   // the whole stage in one launch
          if (m->mixed) launch_vl2<float>(m, t, 1, order, wfac, nullptr);
          else launch_vl2<double>(m, t, 1, order, wfac, nullptr);
        } else {               // fluxes, then the stage kernel
          int rc = calculate_fluxes(m, order);
          if (rc) return rc;
          if (m->mixed) launch_stage<float>(m, t, 1, wfac, nullptr);
          else launch_stage<double>(m, t, 1, wfac, nullptr);
        }
      }
      AGX_HIP(hipEventRecord(e1, m->stream));
      AGX_HIP(hipEventSynchronize(e1));
      AGX_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (t == 0 || ms < best) { best = ms; best_t = t; }
    }
    best_of[order] = best_t;
  }
  m->timer.enabled = timing;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  LAUNCH_CHECK("stage forms (autotune)");
  m->tile1 = best_of[1];
  m->tile = best_of[m->d.xorder];
  if (choice) *choice = m->tile;
  return mg_autotune(m);
}

int agx_set_mg_fuse(agx_mesh *m, int32_t mode) {
  if (!m) { set_error("null mesh"); return -1; }
  if (mode < -1 || mode > 1) { set_error("red-black pass form %d not in {-1, 0, 1}", mode); return -2; }
  m->mg_fuse = mode;
  m->graph_gen++;
  return 0;
}

int agx_set_tile(agx_mesh *m, int32_t tile) {
  if (!m) { set_error("null mesh"); return -1; }
  if (tile < 0 || tile >= NTILE) { set_error("stage form %d not in [0, %d)", tile, NTILE); return -2; }
  if (tile >= TILE_VL2 && m->d.char_proj && m->d.xorder > 1) {
    set_error("characteristic projection runs in the split stage forms (0..%d) only", TILE_VL2 - 1);
    return -3;
  }
  m->tile = m->tile1 = tile;
  graph_reset(m);
  return 0;
}

int agx_timing_enable(agx_mesh *m, int enable) {
  if (!m) { set_error("null mesh"); return -1; }
  Timer &t = m->timer;
  if (enable && t.ev.empty()) {
    t.ev.resize(2 * 4096);
    t.cls.resize(4096);
    for (auto &ev : t.ev) AGX_HIP(hipEventCreate(&ev));
  }
  if (!enable) { int r = timer_flush(m); if (r) return r; }
  t.enabled = enable != 0;
  return 0;
}

int agx_timing_reset(agx_mesh *m) {
  if (!m) { set_error("null mesh"); return -1; }
  int r = timer_flush(m);
  for (int c = 0; c < AGX_K_COUNT; ++c) { m->timer.ms[c] = 0.0; m->timer.launches[c] = 0; }
  return r;
}

int agx_timing_get(agx_mesh *m, int kclass, double *ms_total, int64_t *launches) {
  if (!m || kclass < 0 || kclass >= AGX_K_COUNT) { set_error("bad argument"); return -1; }
  int r = timer_flush(m);
  if (r) return r;
  if (ms_total) *ms_total = m->timer.ms[kclass];
  if (launches) *launches = m->timer.launches[kclass];
  return 0;
}

}  // extern "C"
