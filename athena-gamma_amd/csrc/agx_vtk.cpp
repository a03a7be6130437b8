// agx_vtk.cpp -- VTK legacy output of the MI355X library in the reference's format.
//
// VTKOutput::WriteOutputFile (outputs/vtk.cpp:54-220): one file per MeshBlock, named
// "<basename>.block<gid>.<file_id>.<NNNNN>.vtk", holding
//   "# vtk DataFile Version 2.0\n"
//   "# Athena++ data at time=%e  cycle=%d  variables=%s \n"
//   "BINARY\nDATASET RECTILINEAR_GRID\nDIMENSIONS n1+1 n2+1 n3+1\n"
//   the face coordinates along x1, x2, x3 ("X_COORDINATES n float\n" + big-endian
//   float32, then "\nY_COORDINATES ...", "\nZ_COORDINATES ..."),
//   "\nCELL_DATA n1*n2*n3", and per output quantity "\n<SCALARS|VECTORS> <name> float\n"
//   ("LOOKUP_TABLE default\n" for scalars) followed by big-endian float32 values, k
//   slowest, i fastest, a vector's three components together per cell.
// The quantities and their order are those of OutputType::LoadOutputData
// (outputs/outputs.cpp:355-660) for the Cartesian, non-relativistic MHD build: dens,
// rho, Etot, press, mom, mom1-3, vel, vel1-3, phi, Bcc, Bcc1-3, selected by the
// comma-separated tokens of <output>/variable (ContainVariable, outputs.cpp:1080-1091):
// "prim" = rho, press, vel, phi, Bcc; "cons" = dens, Etot, mom, phi, Bcc.  Tokens the
// reference maps to quantities this library does not hold (face fields, scalars,
// user variables) are refused rather than skipped.  Slices and sums
// (TransformOutputData) are not supported.
//
// Host code only in agx_vtk_write_host; agx_vtk_write downloads the state first.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "agx_internal.hpp"

namespace agx {

namespace {

struct Quantity {
  const char *name;
  bool vector;
  int arr;     // 0 u, 1 w, 2 bcc, 3 phi
  int first;   // first variable index in the array
};

bool contains(const std::string &hay, const std::string &needle) {   // outputs.cpp:1080-1091
  if (hay == needle) return true;
  if (hay.find(',' + needle + ',') != std::string::npos) return true;
  if (hay.find(needle + ',') == 0) return true;
  const size_t p = hay.find(',' + needle);
  return p != std::string::npos && p == hay.size() - needle.size() - 1;
}

// the quantities of `variable`, in LoadOutputData's order; -1 with the error set if a
// token is not one this writer knows
int select(const std::string &var, bool iso, bool grav, std::vector<Quantity> &q) {
  static const char *known[] = {"prim", "cons", "D", "d", "E", "p", "m", "m1", "m2", "m3", "v", "vx", "v1",
                                "vy", "v2", "vz", "v3", "phi", "bcc", "bcc1", "bcc2", "bcc3"};
  size_t a = 0;
  while (a <= var.size()) {
    size_t b = var.find(',', a);
    if (b == std::string::npos) b = var.size();
    const std::string tok = var.substr(a, b - a);
    bool ok = false;
    for (const char *k : known) ok = ok || tok == k;
    if (!ok) {
      set_error("agx_vtk_write: output variable '%s' is not one of prim, cons, D, d, E, p, m, m1-m3, v, "
                "v1-v3 (vx-vz), phi, bcc, bcc1-bcc3", tok.c_str());
      return -1;
    }
    a = b + 1;
  }
  const bool prim = contains(var, "prim"), cons = contains(var, "cons");
  if (contains(var, "D") || cons) q.push_back({"dens", false, 0, 0});
  if (contains(var, "d") || prim) q.push_back({"rho", false, 1, 0});
  if (!iso && (contains(var, "E") || cons)) q.push_back({"Etot", false, 0, 4});
  if (!iso && (contains(var, "p") || prim)) q.push_back({"press", false, 1, 4});
  if (contains(var, "m") || cons) q.push_back({"mom", true, 0, 1});
  if (contains(var, "m1")) q.push_back({"mom1", false, 0, 1});
  if (contains(var, "m2")) q.push_back({"mom2", false, 0, 2});
  if (contains(var, "m3")) q.push_back({"mom3", false, 0, 3});
  if (contains(var, "v") || prim) q.push_back({"vel", true, 1, 1});
  if (contains(var, "vx") || contains(var, "v1")) q.push_back({"vel1", false, 1, 1});
  if (contains(var, "vy") || contains(var, "v2")) q.push_back({"vel2", false, 1, 2});
  if (contains(var, "vz") || contains(var, "v3")) q.push_back({"vel3", false, 1, 3});
  if (grav && (contains(var, "phi") || prim || cons)) q.push_back({"phi", false, 3, 0});
  if (contains(var, "bcc") || prim || cons) q.push_back({"Bcc", true, 2, 0});
  if (contains(var, "bcc1")) q.push_back({"Bcc1", false, 2, 0});
  if (contains(var, "bcc2")) q.push_back({"Bcc2", false, 2, 1});
  if (contains(var, "bcc3")) q.push_back({"Bcc3", false, 2, 2});
  return 0;
}

void put_be(std::vector<unsigned char> &out, float v) {   // big-endian float32 (vtk.cpp:45-52)
  unsigned char b[4];
  std::memcpy(b, &v, 4);
  out.push_back(b[3]);
  out.push_back(b[2]);
  out.push_back(b[1]);
  out.push_back(b[0]);
}

void put_text(std::vector<unsigned char> &out, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void put_text(std::vector<unsigned char> &out, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  const int n = std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  out.insert(out.end(), buf, buf + (n < (int)sizeof(buf) ? n : (int)sizeof(buf) - 1));
}

}  // namespace

}  // namespace agx

using namespace agx;

extern "C" {

int agx_vtk_write_host(const char *basename, const char *file_id, int32_t file_number, const char *variable,
                       int32_t ghost, const agx_mesh_desc *d, double time, int32_t ncycle, const double *u,
                       const double *w, const double *bcc, const double *phi) {
  if (!basename || !file_id || !variable || !d || !u || !w || !bcc) {
    set_error("agx_vtk_write_host: null argument");
    return -1;
  }
  if (d->nx1 % d->bnx1 || d->nx2 % d->bnx2 || d->nx3 % d->bnx3 || d->nghost < 1) {
    set_error("agx_vtk_write_host: the mesh is not a lattice of MeshBlocks");
    return -2;
  }
  if (file_number < 0 || file_number > 99999) {
    set_error("agx_vtk_write_host: file number %d does not fit five digits", file_number);
    return -2;
  }
  const bool iso = d->eos == 1, grav = d->self_gravity != 0;
  std::vector<Quantity> qs;
  if (select(variable, iso, grav, qs)) return -3;
  if (grav && !phi) {
    for (const Quantity &q : qs)
      if (q.arr == 3) { set_error("agx_vtk_write_host: phi requested but not given"); return -1; }
  }
  const int ng = d->nghost, nh = iso ? 4 : 5;
  const int n1 = d->bnx1 + 2 * ng, n2 = d->bnx2 + 2 * ng, n3 = d->bnx3 + 2 * ng;
  const int nb1 = d->nx1 / d->bnx1, nb2 = d->nx2 / d->bnx2, nb3 = d->nx3 / d->bnx3;
  const int64_t ncell = (int64_t)n1 * n2 * n3;
  // output range: the interior, or with NGHOST ghosts on every side (vtk.cpp:63-71)
  const int o = ghost ? 0 : ng;
  const int is = o, ie = n1 - 1 - o, js = o, je = n2 - 1 - o, ks = o, ke = n3 - 1 - o;
  const int c1 = ie - is + 1, c2 = je - js + 1, c3 = ke - ks + 1;
  std::vector<unsigned char> out;
  std::vector<double> xf[3], xv[3];
  const std::vector<int64_t> lattice = gid_lattice_order(nb1, nb2, nb3);
  for (int gid = 0; gid < nb1 * nb2 * nb3; ++gid) {
    const int b = (int)lattice[gid];   // files are named by the reference's gid
    const int lx[3] = {b % nb1, (b / nb1) % nb2, b / (nb1 * nb2)};
    const int nm[3] = {d->nx1, d->nx2, d->nx3}, nbk[3] = {nb1, nb2, nb3}, nbx[3] = {d->bnx1, d->bnx2, d->bnx3};
    const double mn[3] = {d->x1min, d->x2min, d->x3min}, mx[3] = {d->x1max, d->x2max, d->x3max};
    for (int dd = 0; dd < 3; ++dd) {
      double dx;
      block_coords(nm[dd], nbk[dd], nbx[dd], ng, lx[dd], mn[dd], mx[dd], xf[dd], xv[dd], dx);
    }
    out.clear();
    put_text(out, "# vtk DataFile Version 2.0\n");
    put_text(out, "# Athena++ data at time=%e", time);
    put_text(out, "  cycle=%d", ncycle);
    put_text(out, "  variables=%s \n", variable);
    put_text(out, "BINARY\n");
    put_text(out, "DATASET RECTILINEAR_GRID\n");
    put_text(out, "DIMENSIONS %d %d %d\n", c1 + 1, c2 + 1, c3 + 1);
    const char *axis[3] = {"X", "Y", "Z"};
    const int lo[3] = {is, js, ks}, nc[3] = {c1, c2, c3};
    for (int dd = 0; dd < 3; ++dd) {
      put_text(out, "%s%s_COORDINATES %d float\n", dd ? "\n" : "", axis[dd], nc[dd] + 1);
      for (int i = 0; i <= nc[dd]; ++i) put_be(out, (float)xf[dd][lo[dd] + i]);
    }
    put_text(out, "\nCELL_DATA %d", c1 * c2 * c3);
    for (const Quantity &q : qs) {
      put_text(out, "\n%s %s float\n", q.vector ? "VECTORS" : "SCALARS", q.name);
      if (!q.vector) put_text(out, "LOOKUP_TABLE default\n");
      const double *src = q.arr == 0 ? u + (int64_t)b * nh * ncell
                        : q.arr == 1 ? w + (int64_t)b * nh * ncell
                        : q.arr == 2 ? bcc + (int64_t)b * 3 * ncell
                                     : phi + (int64_t)b * ncell;
      const int nv = q.vector ? 3 : 1;
      for (int k = ks; k <= ke; ++k)
        for (int j = js; j <= je; ++j)
          for (int i = is; i <= ie; ++i)
            for (int n = 0; n < nv; ++n)
              put_be(out, (float)src[((int64_t)(q.first + n) * n3 + k) * n2 * n1 + (int64_t)j * n1 + i]);
    }
    char name[64];
    std::snprintf(name, sizeof(name), ".block%d.", gid);
    char num[8];
    std::snprintf(num, sizeof(num), ".%05d", file_number);
    const std::string fname = std::string(basename) + name + file_id + num + ".vtk";
    FILE *f = std::fopen(fname.c_str(), "wb");
    if (!f) { set_error("agx_vtk_write_host: output file '%s' could not be opened", fname.c_str()); return -4; }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) != 0 || !ok) { set_error("agx_vtk_write_host: writing '%s' failed", fname.c_str()); return -4; }
  }
  return 0;
}

int agx_vtk_write(agx_mesh *m, const char *basename, const char *file_id, int32_t file_number,
                  const char *variable, int32_t ghost) {
  if (!m) { set_error("agx_vtk_write: null mesh"); return -1; }
  if (int rc = refuse_tree(m, "agx_vtk_write")) return rc;
  int64_t nb, du[4], db[4], dp[4];
  int rc = agx_array_dims(m, AGX_ARR_U, &nb, du);
  if (!rc) rc = agx_array_dims(m, AGX_ARR_BCC, &nb, db);
  if (!rc) rc = agx_array_dims(m, AGX_ARR_PHI, &nb, dp);
  if (rc) return rc;
  const int64_t nu = nb * du[0] * du[1] * du[2] * du[3], nbc = nb * db[0] * db[1] * db[2] * db[3];
  std::vector<double> u(nu), w(nu), bcc(nbc), phi;
  if ((rc = agx_download(m, AGX_ARR_U, u.data())) || (rc = agx_download(m, AGX_ARR_W, w.data())) ||
      (rc = agx_download(m, AGX_ARR_BCC, bcc.data())))
    return rc;
  if (m->d.self_gravity) {
    phi.resize(nb * dp[0] * dp[1] * dp[2] * dp[3]);
    if ((rc = agx_download(m, AGX_ARR_PHI, phi.data()))) return rc;
  }
  return agx_vtk_write_host(basename, file_id, file_number, variable, ghost, &m->d, m->time, m->ncycle, u.data(),
                            w.data(), bcc.data(), phi.empty() ? nullptr : phi.data());
}

}  // extern "C"
