// agx_bridge.hpp -- the reference-independent half of the host shim that binds the C ABI
// (include/athena_gamma_amd.h) into Athena++ / Athena-gamma.
//
// DescFromInput reads the input deck through the reference's ParameterInput interface
// (GetOrAddString / GetOrAddInteger / GetOrAddReal / GetOrAddBoolean, parameter_input.hpp:
// 86-97) -- any class with those four methods works -- and fills an agx_mesh_desc with
// the same defaults and the same refusals the reference's constructors apply.  What the
// deck does not hold comes in as arguments: the mesh geometry and CFL number (public
// members of Mesh), the build's configure options (NGHOST, the EOS, --flux, --grav) and
// 4 pi G, which the problem generator passes to Mesh::SetFourPiG (mesh.hpp:376-377; the
// member itself is private).  Errors throw std::runtime_error, as ATHENA_ERROR does
// (defs.hpp.in:119-123).
#ifndef AGX_BRIDGE_HPP_
#define AGX_BRIDGE_HPP_

#include <cctype>
#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <string>

#include "athena_gamma_amd.h"

namespace agx_bridge {

// A non-zero return of a C-ABI call as an exception carrying agx_last_error().
inline void Check(int rc, const char *what) {
  if (rc != 0)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + agx_last_error());
}

// Mesh geometry: Mesh::mesh_size (nx1..3, x?min, x?max) and Mesh::cfl_number.
struct MeshGeom {
  int nx1, nx2, nx3;
  double x1min, x1max, x2min, x2max, x3min, x3max;
  double cfl_number;
};

// configure-time choices of the build (defs.hpp): NGHOST, NON_BAROTROPIC_EOS,
// SELF_GRAVITY_ENABLED == 2 (multigrid), and the --flux solver name.
struct BuildOptions {
  int nghost;
  bool adiabatic;
  bool multigrid_gravity;
  std::string flux;   // "hlld", "lhlld", "hlle" or "roe"
};

inline int RiemannSolver(const std::string &flux) {
  if (flux == "hlld") return AGX_RSOLVER_HLLD;
  if (flux == "lhlld") return AGX_RSOLVER_LHLLD;
  if (flux == "hlle") return AGX_RSOLVER_HLLE;
  if (flux == "roe") return AGX_RSOLVER_ROE;
  throw std::runtime_error("--flux=" + flux + " has no device form (hlld, lhlld, hlle, roe)");
}

// <gravity>/ix1_bc..ox3_bc (mg_gravity.cpp:79-90; names of GetMGBoundaryFlag,
// boundary_flag.cpp:149-170): the library solves with all six faces periodic or all six
// multipole.  Any other mix is refused rather than approximated.
template <class Pin>
int GravityBoundary(Pin *pin) {
  static const char *faces[6] = {"ix1_bc", "ox1_bc", "ix2_bc", "ox2_bc", "ix3_bc", "ox3_bc"};
  std::string first;
  for (int f = 0; f < 6; ++f) {
    const std::string v = pin->GetOrAddString("gravity", faces[f], "none");
    if (v != "periodic" && v != "multipole")
      throw std::runtime_error(std::string("<gravity>/") + faces[f] + " = " + v +
                               ": the device multigrid supports periodic or multipole faces only");
    if (f == 0) first = v;
    else if (v != first)
      throw std::runtime_error(std::string("<gravity>/") + faces[f] + " = " + v + " differs from ix1_bc = " +
                               first + ": all six faces must be periodic, or all multipole");
  }
  return first == "multipole" ? 1 : 0;
}

// <time>/xorder (reconstruction.cpp:46-71): "1", "2", "2c", "3", "3c".
inline void Reconstruction(const std::string &xorder, agx_mesh_desc *d) {
  if (xorder == "1" || xorder == "2" || xorder == "3") {
    d->xorder = xorder[0] - '0';
    d->char_proj = 0;
  } else if (xorder == "2c" || xorder == "3c") {
    d->xorder = xorder[0] - '0';
    d->char_proj = 1;
  } else {
    throw std::runtime_error("xorder=" + xorder + " has no device form (1, 2, 2c, 3, 3c)");
  }
}

template <class Pin>
agx_mesh_desc DescFromInput(Pin *pin, const MeshGeom &g, const BuildOptions &b, double four_pi_G) {
  agx_mesh_desc d = {};
  d.nx1 = g.nx1; d.nx2 = g.nx2; d.nx3 = g.nx3;
  d.bnx1 = pin->GetOrAddInteger("meshblock", "nx1", g.nx1);
  d.bnx2 = pin->GetOrAddInteger("meshblock", "nx2", g.nx2);
  d.bnx3 = pin->GetOrAddInteger("meshblock", "nx3", g.nx3);
  d.nghost = b.nghost;
  Reconstruction(pin->GetOrAddString("time", "xorder", "2"), &d);
  d.rsolver = RiemannSolver(b.flux);
  d.eos = b.adiabatic ? 0 : 1;
  d.gamma = b.adiabatic ? pin->GetOrAddReal("hydro", "gamma", 5.0 / 3.0) : 5.0 / 3.0;
  d.iso_cs = b.adiabatic ? 1.0 : pin->GetOrAddReal("hydro", "iso_sound_speed", 1.0);
  d.x1min = g.x1min; d.x1max = g.x1max;
  d.x2min = g.x2min; d.x2max = g.x2max;
  d.x3min = g.x3min; d.x3max = g.x3max;
  d.cfl = g.cfl_number;
  // EquationOfState defaults (eos/adiabatic_mhd.cpp:26-31)
  const double floor_default = std::sqrt(1024.0 * FLT_MIN);
  d.dfloor = pin->GetOrAddReal("hydro", "dfloor", floor_default);
  d.pfloor = pin->GetOrAddReal("hydro", "pfloor", floor_default);
  d.precision = pin->GetOrAddBoolean("gpu", "mixed_precision", false) ? AGX_PREC_MIXED : AGX_PREC_FP64;
  d.self_gravity = b.multigrid_gravity ? 1 : 0;
  d.mporder = 4;
  d.mp_auto_origin = 1;
  d.mg_threshold = -1.0;
  d.mg_niter = -1;
  if (b.multigrid_gravity) {
    if (!(four_pi_G > 0.0))   // mg_gravity.cpp:71-77
      throw std::runtime_error("4 pi G must be set (Mesh::SetFourPiG) and positive");
    d.four_pi_G = four_pi_G;
    // mg_gravity.cpp:45-70: mgmode FMG or MGI (any case); threshold or niteration
    std::string mode = pin->GetOrAddString("gravity", "mgmode", "none");
    for (char &ch : mode) ch = (char)std::tolower((unsigned char)ch);
    if (mode != "fmg" && mode != "mgi")
      throw std::runtime_error("<gravity>/mgmode = " + mode + ": FMG or MGI");
    d.mg_mode = (mode == "mgi") ? 1 : 0;
    d.mg_threshold = pin->GetOrAddReal("gravity", "threshold", -1.0);
    d.mg_niter = pin->GetOrAddInteger("gravity", "niteration", -1);
    if (d.mg_threshold < 0.0 && d.mg_niter < 0)
      throw std::runtime_error("<gravity>: set threshold or niteration");
    d.mg_bc = GravityBoundary(pin);
    if (d.mg_bc == 1) {
      d.mporder = pin->GetOrAddInteger("gravity", "mporder", 0);
      d.mp_auto_origin = pin->GetOrAddBoolean("gravity", "auto_mporigin", true) ? 1 : 0;
      d.mp_nodipole = pin->GetOrAddBoolean("gravity", "nodipole", false) ? 1 : 0;
      if (!d.mp_auto_origin) {
        d.mp_origin[0] = pin->GetOrAddReal("gravity", "mporigin_x1", 0.0);
        d.mp_origin[1] = pin->GetOrAddReal("gravity", "mporigin_x2", 0.0);
        d.mp_origin[2] = pin->GetOrAddReal("gravity", "mporigin_x3", 0.0);
      }
    }
  }
  return d;
}

}  // namespace agx_bridge

#endif  // AGX_BRIDGE_HPP_
