// agx_restart.cpp -- restart files of the MI355X library in the reference's format.
//
// RestartOutput::WriteOutputFile (outputs/restart.cpp:37-207) writes, in one file:
//   1. the input parameters as text, ParameterInput::ParameterDump (parameter_input.cpp:
//      834-859), ending with the line "<par_end>";
//   2. the header: int nbtotal, int root_level, RegionSize mesh_size (athena.hpp:83-89:
//      9 doubles, 3 ints, tail-padded to 88 bytes), double time, double dt, int ncycle,
//      uint64 datasize (the bytes of one MeshBlock's data) -- fields written one after
//      another, no padding between them;
//   3. the ID list: per MeshBlock its LogicalLocation (athena.hpp:61-66: three int64 and
//      an int, padded to 32 bytes) and its load-balancing cost (double), in the Z order of
//      MeshBlockTree::GetMeshBlockList (mesh/meshblock_tree.cpp:336-352);
//   4. the data: per MeshBlock, in the same order, Hydro::u (NHYDRO, n3, n2, n1), then
//      Field::b.x1f (n3, n2, n1+1), b.x2f (n3, n2+1, n1), b.x3f (n3+1, n2, n1), ghosts
//      included (MeshBlock::GetBlockSizeInBytes, mesh/meshblock.cpp:442-463).
// The reader follows Mesh::Mesh(pin, resfile) (mesh/mesh.cpp:551-700) and the restart
// MeshBlock constructor: it skips the text up to "<par_end>\n", reads the header, and
// places each MeshBlock by its logical location.  The gravitational potential is not in
// the file: the reference solves for it again in Mesh::Initialize.
//
// Host code only (no device work in the *_host functions); agx_restart_write and
// agx_restart_read move the state between the device and these files.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "agx_internal.hpp"

namespace agx {

namespace {

struct RegionSizeF {   // athena.hpp:83-89, same member order and types
  double x1min, x2min, x3min;
  double x1max, x2max, x3max;
  double x1rat, x2rat, x3rat;
  int32_t nx1, nx2, nx3;
};
static_assert(sizeof(RegionSizeF) == 88, "RegionSize is 9 doubles + 3 ints, padded to 88 bytes");

struct LocF {          // athena.hpp:61-66
  int64_t lx1, lx2, lx3;
  int32_t level;
};
static_assert(sizeof(LocF) == 32, "LogicalLocation is 3 int64 + int, padded to 32 bytes");

constexpr size_t HEADER_BYTES = 3 * sizeof(int32_t) + sizeof(RegionSizeF) + 2 * sizeof(double) + sizeof(uint64_t);

int root_level_of(int nb1, int nb2, int nb3) {   // mesh.cpp:291-292
  int nbmax = std::max(nb1, std::max(nb2, nb3));
  int rl = 0;
  while ((1 << rl) < nbmax) ++rl;
  return rl;
}

// Leaves of the root grid in Z order: the recursive descent of MeshBlockTree::
// CreateRootGrid (children n = i + 2j + 4k, kept if inside the root grid,
// meshblock_tree.cpp:87-110) listed as GetMeshBlockList does.
void zorder(int64_t lx1, int64_t lx2, int64_t lx3, int level, int root, int nb1, int nb2, int nb3,
            std::vector<LocF> &out) {
  if (level == root) {
    LocF l;
    std::memset(&l, 0, sizeof(l));   // the padding bytes go to the file too: keep them zero
    l.lx1 = lx1; l.lx2 = lx2; l.lx3 = lx3; l.level = level;
    out.push_back(l);
    return;
  }
  const int64_t levfac = int64_t(1) << (root - level - 1);
  for (int n = 0; n < 8; ++n) {
    const int i = n & 1, j = (n >> 1) & 1, k = (n >> 2) & 1;
    if ((lx3 * 2 + k) * levfac < nb3 && (lx2 * 2 + j) * levfac < nb2 && (lx1 * 2 + i) * levfac < nb1)
      zorder(lx1 * 2 + i, lx2 * 2 + j, lx3 * 2 + k, level + 1, root, nb1, nb2, nb3, out);
  }
}

struct Layout {
  int nb1, nb2, nb3, nb, nh;
  int64_t ncell, nf1, nf2, nf3;      // per-block element counts (u per variable, faces)
  uint64_t datasize;
};

int layout_of(const agx_mesh_desc *d, Layout &L, const char *who) {
  if (!d) { set_error("%s: null descriptor", who); return -1; }
  if (d->bnx1 <= 0 || d->bnx2 <= 0 || d->bnx3 <= 0 || d->nx1 % d->bnx1 || d->nx2 % d->bnx2 || d->nx3 % d->bnx3) {
    set_error("%s: the mesh must divide into MeshBlocks", who);
    return -2;
  }
  L.nb1 = d->nx1 / d->bnx1; L.nb2 = d->nx2 / d->bnx2; L.nb3 = d->nx3 / d->bnx3;
  L.nb = L.nb1 * L.nb2 * L.nb3;
  L.nh = (d->eos == 1) ? 4 : 5;
  const int64_t n1 = d->bnx1 + 2 * d->nghost, n2 = d->bnx2 + 2 * d->nghost, n3 = d->bnx3 + 2 * d->nghost;
  L.ncell = n1 * n2 * n3;
  L.nf1 = n3 * n2 * (n1 + 1);
  L.nf2 = n3 * (n2 + 1) * n1;
  L.nf3 = (n3 + 1) * n2 * n1;
  L.datasize = sizeof(double) * (uint64_t)(L.nh * L.ncell + L.nf1 + L.nf2 + L.nf3);
  return 0;
}

// ParameterInput::ParameterDump of the inputs the descriptor carries, for files written
// without the caller's own parameter text.
std::string default_params(const agx_mesh_desc *d) {
  char b[4096];
  std::snprintf(b, sizeof(b),
                "#------------------------- PAR_DUMP -------------------------\n"
                "<mesh>\nnx1 = %d\nx1min = %.17g\nx1max = %.17g\nnx2 = %d\nx2min = %.17g\nx2max = %.17g\n"
                "nx3 = %d\nx3min = %.17g\nx3max = %.17g\n"
                "<meshblock>\nnx1 = %d\nnx2 = %d\nnx3 = %d\n"
                "<time>\ncfl_number = %.17g\nxorder = %d\nintegrator = vl2\n"
                "<hydro>\ngamma = %.17g\niso_sound_speed = %.17g\ndfloor = %.17g\npfloor = %.17g\n"
                "#------------------------- PAR_DUMP -------------------------\n",
                d->nx1, d->x1min, d->x1max, d->nx2, d->x2min, d->x2max, d->nx3, d->x3min, d->x3max, d->bnx1,
                d->bnx2, d->bnx3, d->cfl, d->xorder, d->gamma, d->iso_cs, d->dfloor, d->pfloor);
  return std::string(b);
}

struct File {
  FILE *f = nullptr;
  ~File() { if (f) std::fclose(f); }
};

}  // namespace

std::vector<int64_t> gid_lattice_order(int nb1, int nb2, int nb3) {
  std::vector<LocF> locs;
  const int root = root_level_of(nb1, nb2, nb3);
  zorder(0, 0, 0, 0, root, nb1, nb2, nb3, locs);
  std::vector<int64_t> out;
  out.reserve(locs.size());
  for (const LocF &l : locs) out.push_back((l.lx3 * nb2 + l.lx2) * nb1 + l.lx1);
  return out;
}

}  // namespace agx

using namespace agx;

extern "C" {

int agx_block_gids(const agx_mesh_desc *d, int64_t *lattice_of_gid) {
  Layout L;
  int rc = layout_of(d, L, "agx_block_gids");
  if (rc) return rc;
  if (!lattice_of_gid) { set_error("agx_block_gids: null output"); return -1; }
  const std::vector<int64_t> o = gid_lattice_order(L.nb1, L.nb2, L.nb3);
  for (size_t q = 0; q < o.size(); ++q) lattice_of_gid[q] = o[q];
  return 0;
}


int agx_restart_write_host(const char *path, const agx_mesh_desc *d, const char *params, double time, double dt,
                           int32_t ncycle, const double *u, const double *b1f, const double *b2f,
                           const double *b3f) {
  Layout L;
  int rc = layout_of(d, L, "agx_restart_write_host");
  if (rc) return rc;
  if (!path || !u || !b1f || !b2f || !b3f) { set_error("agx_restart_write_host: null argument"); return -1; }
  std::string text = params ? std::string(params) : default_params(d);
  if (text.find("<par_end>") != std::string::npos) {
    set_error("agx_restart_write_host: the parameter text must not contain <par_end> (it is appended)");
    return -3;
  }
  if (!text.empty() && text.back() != '\n') text.push_back('\n');
  text += "<par_end>\n";
  File F;
  F.f = std::fopen(path, "wb");
  if (!F.f) { set_error("agx_restart_write_host: cannot open %s", path); return -4; }
  // header (restart.cpp:104-113)
  const int32_t nbtotal = L.nb, root = root_level_of(L.nb1, L.nb2, L.nb3);
  RegionSizeF ms;
  std::memset(&ms, 0, sizeof(ms));
  ms.x1min = d->x1min; ms.x2min = d->x2min; ms.x3min = d->x3min;
  ms.x1max = d->x1max; ms.x2max = d->x2max; ms.x3max = d->x3max;
  ms.x1rat = ms.x2rat = ms.x3rat = 1.0;
  ms.nx1 = d->nx1; ms.nx2 = d->nx2; ms.nx3 = d->nx3;
  const uint64_t datasize = L.datasize;
  bool ok = std::fwrite(text.data(), 1, text.size(), F.f) == text.size();
  ok = ok && std::fwrite(&nbtotal, sizeof(int32_t), 1, F.f) == 1;
  ok = ok && std::fwrite(&root, sizeof(int32_t), 1, F.f) == 1;
  ok = ok && std::fwrite(&ms, sizeof(ms), 1, F.f) == 1;
  ok = ok && std::fwrite(&time, sizeof(double), 1, F.f) == 1;
  ok = ok && std::fwrite(&dt, sizeof(double), 1, F.f) == 1;
  ok = ok && std::fwrite(&ncycle, sizeof(int32_t), 1, F.f) == 1;
  ok = ok && std::fwrite(&datasize, sizeof(uint64_t), 1, F.f) == 1;
  // ID list and data in Z order (restart.cpp:127-200)
  std::vector<LocF> locs;
  zorder(0, 0, 0, 0, root, L.nb1, L.nb2, L.nb3, locs);
  if ((int)locs.size() != L.nb) { set_error("agx_restart_write_host: Z-order list size mismatch"); return -5; }
  for (const LocF &loc : locs) {
    unsigned char rec[sizeof(LocF) + sizeof(double)];
    std::memset(rec, 0, sizeof(rec));
    std::memcpy(rec, &loc, sizeof(LocF));
    const double cost = 1.0;   // MeshBlock::cost_ starts at 1 (meshblock.cpp)
    std::memcpy(rec + sizeof(LocF), &cost, sizeof(double));
    ok = ok && std::fwrite(rec, 1, sizeof(rec), F.f) == sizeof(rec);
  }
  for (const LocF &loc : locs) {
    const int64_t g = (loc.lx3 * L.nb2 + loc.lx2) * L.nb1 + loc.lx1;
    ok = ok && std::fwrite(u + g * L.nh * L.ncell, sizeof(double), L.nh * L.ncell, F.f) == (size_t)(L.nh * L.ncell);
    ok = ok && std::fwrite(b1f + g * L.nf1, sizeof(double), L.nf1, F.f) == (size_t)L.nf1;
    ok = ok && std::fwrite(b2f + g * L.nf2, sizeof(double), L.nf2, F.f) == (size_t)L.nf2;
    ok = ok && std::fwrite(b3f + g * L.nf3, sizeof(double), L.nf3, F.f) == (size_t)L.nf3;
  }
  if (!ok) { set_error("agx_restart_write_host: write to %s failed", path); return -6; }
  if (std::fclose(F.f) != 0) { F.f = nullptr; set_error("agx_restart_write_host: close failed"); return -6; }
  F.f = nullptr;
  return 0;
}

int agx_restart_read_host(const char *path, const agx_mesh_desc *d, double *time, double *dt, int32_t *ncycle,
                          double *u, double *b1f, double *b2f, double *b3f) {
  Layout L;
  int rc = layout_of(d, L, "agx_restart_read_host");
  if (rc) return rc;
  if (!path || !u || !b1f || !b2f || !b3f) { set_error("agx_restart_read_host: null argument"); return -1; }
  File F;
  F.f = std::fopen(path, "rb");
  if (!F.f) { set_error("agx_restart_read_host: cannot open %s", path); return -4; }
  // the parameter text ends at "<par_end>" + newline (parameter_input.cpp:193-221)
  std::string head;
  const std::string tag = "<par_end>";
  size_t hdr = std::string::npos;
  char buf[4096];
  while (hdr == std::string::npos) {
    const size_t n = std::fread(buf, 1, sizeof(buf), F.f);
    if (n == 0) break;
    head.append(buf, n);
    const size_t loc = head.find(tag);
    if (loc != std::string::npos) hdr = loc + tag.size() + 1;
    if (head.size() > 40960 && hdr == std::string::npos) break;
  }
  if (hdr == std::string::npos) {
    set_error("agx_restart_read_host: <par_end> not found in the first 40 KB of %s", path);
    return -7;
  }
  if (std::fseek(F.f, (long)hdr, SEEK_SET) != 0) { set_error("agx_restart_read_host: seek failed"); return -7; }
  unsigned char hb[HEADER_BYTES];
  if (std::fread(hb, 1, sizeof(hb), F.f) != sizeof(hb)) { set_error("agx_restart_read_host: header truncated"); return -7; }
  int32_t nbtotal, root, nc;
  RegionSizeF ms;
  double t, h;
  uint64_t datasize;
  size_t o = 0;
  std::memcpy(&nbtotal, hb + o, 4); o += 4;
  std::memcpy(&root, hb + o, 4); o += 4;
  std::memcpy(&ms, hb + o, sizeof(ms)); o += sizeof(ms);
  std::memcpy(&t, hb + o, 8); o += 8;
  std::memcpy(&h, hb + o, 8); o += 8;
  std::memcpy(&nc, hb + o, 4); o += 4;
  std::memcpy(&datasize, hb + o, 8);
  if (nbtotal != L.nb || root != root_level_of(L.nb1, L.nb2, L.nb3) || ms.nx1 != d->nx1 || ms.nx2 != d->nx2 ||
      ms.nx3 != d->nx3 || ms.x1min != d->x1min || ms.x1max != d->x1max || ms.x2min != d->x2min ||
      ms.x2max != d->x2max || ms.x3min != d->x3min || ms.x3max != d->x3max) {
    set_error("agx_restart_read_host: %s holds another mesh (%d blocks at root level %d, %dx%dx%d)", path, nbtotal,
              root, ms.nx1, ms.nx2, ms.nx3);
    return -8;
  }
  if (ms.x1rat != 1.0 || ms.x2rat != 1.0 || ms.x3rat != 1.0) {
    set_error("agx_restart_read_host: stretched meshes are not supported");
    return -8;
  }
  if (datasize != L.datasize) {
    set_error("agx_restart_read_host: MeshBlock data of %llu bytes, this mesh needs %llu (NHYDRO, NGHOST or block "
              "size differ)", (unsigned long long)datasize, (unsigned long long)L.datasize);
    return -8;
  }
  std::vector<LocF> locs(L.nb);
  std::vector<char> seen(L.nb, 0);
  for (int q = 0; q < L.nb; ++q) {
    unsigned char rec[sizeof(LocF) + sizeof(double)];
    if (std::fread(rec, 1, sizeof(rec), F.f) != sizeof(rec)) { set_error("agx_restart_read_host: ID list truncated"); return -7; }
    std::memcpy(&locs[q], rec, sizeof(LocF));
    const LocF &l = locs[q];
    if (l.level != root || l.lx1 < 0 || l.lx1 >= L.nb1 || l.lx2 < 0 || l.lx2 >= L.nb2 || l.lx3 < 0 || l.lx3 >= L.nb3) {
      set_error("agx_restart_read_host: MeshBlock %d is not a root-level block of this mesh (refined meshes are not "
                "read)", q);
      return -8;
    }
    const int64_t g = (l.lx3 * L.nb2 + l.lx2) * L.nb1 + l.lx1;
    if (seen[g]++) { set_error("agx_restart_read_host: MeshBlock listed twice"); return -8; }
  }
  for (int q = 0; q < L.nb; ++q) {
    const LocF &l = locs[q];
    const int64_t g = (l.lx3 * L.nb2 + l.lx2) * L.nb1 + l.lx1;
    bool ok = std::fread(u + g * L.nh * L.ncell, sizeof(double), L.nh * L.ncell, F.f) == (size_t)(L.nh * L.ncell);
    ok = ok && std::fread(b1f + g * L.nf1, sizeof(double), L.nf1, F.f) == (size_t)L.nf1;
    ok = ok && std::fread(b2f + g * L.nf2, sizeof(double), L.nf2, F.f) == (size_t)L.nf2;
    ok = ok && std::fread(b3f + g * L.nf3, sizeof(double), L.nf3, F.f) == (size_t)L.nf3;
    if (!ok) { set_error("agx_restart_read_host: data of MeshBlock %d truncated", q); return -7; }
  }
  if (time) *time = t;
  if (dt) *dt = h;
  if (ncycle) *ncycle = nc;
  return 0;
}

int agx_restart_write(agx_mesh *m, const char *path, const char *params) {
  if (!m) { set_error("agx_restart_write: null mesh"); return -1; }
  if (int rc = refuse_tree(m, "agx_restart_write")) return rc;
  Layout L;
  int rc = layout_of(&m->d, L, "agx_restart_write");
  if (rc) return rc;
  std::vector<double> u((size_t)L.nb * L.nh * L.ncell), b1((size_t)L.nb * L.nf1), b2((size_t)L.nb * L.nf2),
      b3((size_t)L.nb * L.nf3);
  if ((rc = agx_download(m, AGX_ARR_U, u.data())) || (rc = agx_download(m, AGX_ARR_B1F, b1.data())) ||
      (rc = agx_download(m, AGX_ARR_B2F, b2.data())) || (rc = agx_download(m, AGX_ARR_B3F, b3.data())))
    return rc;
  return agx_restart_write_host(path, &m->d, params, m->time, m->dt, m->ncycle, u.data(), b1.data(), b2.data(),
                                b3.data());
}

int agx_restart_read(agx_mesh *m, const char *path) {
  if (!m) { set_error("agx_restart_read: null mesh"); return -1; }
  if (int rc = refuse_tree(m, "agx_restart_read")) return rc;
  Layout L;
  int rc = layout_of(&m->d, L, "agx_restart_read");
  if (rc) return rc;
  std::vector<double> u((size_t)L.nb * L.nh * L.ncell), b1((size_t)L.nb * L.nf1), b2((size_t)L.nb * L.nf2),
      b3((size_t)L.nb * L.nf3);
  double t = 0.0, h = 0.0;
  int32_t nc = 0;
  rc = agx_restart_read_host(path, &m->d, &t, &h, &nc, u.data(), b1.data(), b2.data(), b3.data());
  if (rc) return rc;
  if ((rc = agx_upload(m, AGX_ARR_U, u.data())) || (rc = agx_upload(m, AGX_ARR_B1F, b1.data())) ||
      (rc = agx_upload(m, AGX_ARR_B2F, b2.data())) || (rc = agx_upload(m, AGX_ARR_B3F, b3.data())))
    return rc;
  m->time = t;
  m->dt = h;
  m->ncycle = nc;
  m->restarted = true;   // agx_initialize keeps this dt (the restart Mesh constructor's)
  return 0;
}

}  // extern "C"
