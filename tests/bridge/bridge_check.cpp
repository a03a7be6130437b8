// Builds agx_mesh_desc values from input decks through athena-gamma_amd/bridge/
// agx_bridge.hpp and prints them (or the refusal) one line per deck; linked against the
// library so that Check() resolves agx_last_error.  No GPU call is made.
#include <cstdio>
#include <map>
#include <string>

#include "agx_bridge.hpp"

// A deck with the four ParameterInput accessors the bridge uses.
struct Deck {
  std::map<std::string, std::string> kv;
  std::string Get(const std::string &b, const std::string &n) const {
    auto it = kv.find(b + "/" + n);
    return it == kv.end() ? std::string() : it->second;
  }
  std::string GetOrAddString(std::string b, std::string n, std::string v) {
    std::string s = Get(b, n);
    return s.empty() ? v : s;
  }
  int GetOrAddInteger(std::string b, std::string n, int v) {
    std::string s = Get(b, n);
    return s.empty() ? v : std::stoi(s);
  }
  double GetOrAddReal(std::string b, std::string n, double v) {
    std::string s = Get(b, n);
    return s.empty() ? v : std::stod(s);
  }
  bool GetOrAddBoolean(std::string b, std::string n, bool v) {
    std::string s = Get(b, n);
    return s.empty() ? v : (s == "true" || s == "1");
  }
};

static void show(const char *name, Deck deck, agx_bridge::BuildOptions b, double fpg) {
  agx_bridge::MeshGeom g = {64, 32, 32, 0.0, 3.0, 0.0, 1.5, 0.0, 1.5, 0.4};
  try {
    agx_mesh_desc d = agx_bridge::DescFromInput(&deck, g, b, fpg);
    std::printf("%s ok bnx=%d,%d,%d xorder=%d char=%d rs=%d eos=%d prec=%d sg=%d mode=%d bc=%d mporder=%d "
                "auto=%d fpg=%.17g thr=%.17g niter=%d gamma=%.17g cs=%.17g dfloor=%.17g cfl=%.17g\n",
                name, d.bnx1, d.bnx2, d.bnx3, d.xorder, d.char_proj, d.rsolver, d.eos, d.precision,
                d.self_gravity, d.mg_mode, d.mg_bc, d.mporder, d.mp_auto_origin, d.four_pi_G,
                d.mg_threshold, d.mg_niter, d.gamma, d.iso_cs, d.dfloor, d.cfl);
  } catch (const std::exception &e) {
    std::printf("%s error %s\n", name, e.what());
  }
}

int main() {
  agx_bridge::BuildOptions mhd = {2, true, false, "hlld"};
  agx_bridge::BuildOptions grav = {3, true, true, "lhlld"};
  agx_bridge::BuildOptions iso = {2, false, true, "hlle"};
  Deck lw;
  lw.kv = {{"meshblock/nx1", "64"}, {"meshblock/nx2", "32"}, {"meshblock/nx3", "32"}, {"hydro/gamma", "1.6666666666666667"}};
  show("linwave", lw, mhd, -1.0);
  Deck c3 = lw;
  c3.kv["time/xorder"] = "3c";
  c3.kv["gpu/mixed_precision"] = "true";
  show("ppm_char_mixed", c3, mhd, -1.0);
  Deck bad = lw;
  bad.kv["time/xorder"] = "4";
  show("xorder4", bad, mhd, -1.0);
  Deck sg = lw;
  sg.kv["gravity/mgmode"] = "fmg";
  sg.kv["gravity/threshold"] = "1e-10";
  for (const char *f : {"ix1_bc", "ox1_bc", "ix2_bc", "ox2_bc", "ix3_bc", "ox3_bc"}) sg.kv[std::string("gravity/") + f] = "periodic";
  show("selfgrav_periodic", sg, grav, 1.0);
  show("selfgrav_no_4piG", sg, grav, -1.0);
  Deck mixbc = sg;
  mixbc.kv["gravity/ox3_bc"] = "multipole";
  show("selfgrav_mixed_faces", mixbc, grav, 1.0);
  Deck mp = sg;
  for (const char *f : {"ix1_bc", "ox1_bc", "ix2_bc", "ox2_bc", "ix3_bc", "ox3_bc"}) mp.kv[std::string("gravity/") + f] = "multipole";
  mp.kv["gravity/mporder"] = "4";
  mp.kv["gravity/mgmode"] = "MGI";
  mp.kv["hydro/iso_sound_speed"] = "0.5";
  show("iso_multipole", mp, iso, 0.7);
  Deck zg = sg;
  zg.kv["gravity/ix1_bc"] = "zerograd";
  show("selfgrav_zerograd", zg, grav, 1.0);
  Deck nothr = sg;
  nothr.kv.erase("gravity/threshold");
  show("selfgrav_no_threshold", nothr, grav, 1.0);
  Deck flux = lw;
  show("flux_hllc", flux, {2, true, false, "hllc"}, -1.0);
  try {
    agx_bridge::Check(-7, "agx_example");
  } catch (const std::exception &e) {
    std::printf("check error %s\n", e.what());
  }
  return 0;
}
