"""The host shim's reference-independent half (athena-gamma_amd/bridge/agx_bridge.hpp)
compiled with g++ against include/athena_gamma_amd.h and linked with the library, then
run on decks covering each option it reads and each refusal it applies (CPU only)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "athena-gamma_amd", "lib")


@pytest.fixture(scope="module")
def lines(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libathena_gamma_amd.so")):
        pytest.fail("library not built (python athena-gamma_amd/build.py)")
    exe = str(tmp_path_factory.mktemp("bridge") / "bridge_check")
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-o", exe,
           os.path.join(ROOT, "tests", "bridge", "bridge_check.cpp"),
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "athena-gamma_amd", "bridge"),
           "-L", LIBDIR, "-lathena_gamma_amd", "-Wl,-rpath," + LIBDIR]
    assert shutil.which("g++"), "g++ missing"
    subprocess.run(cmd, check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {ln.split()[0]: ln for ln in out.splitlines()}


def fields(line):
    return dict(kv.split("=", 1) for kv in line.split()[2:])


def test_plain_deck(lines):
    f = fields(lines["linwave"])
    assert lines["linwave"].split()[1] == "ok"
    assert f["bnx"] == "64,32,32" and f["xorder"] == "2" and f["char"] == "0" and f["rs"] == "0"
    assert f["eos"] == "0" and f["prec"] == "0" and f["sg"] == "0"
    assert float(f["dfloor"]) == pytest.approx((1024 * 1.1754943508222875e-38) ** 0.5, rel=1e-15)
    assert float(f["cfl"]) == 0.4


def test_characteristic_ppm_and_mixed(lines):
    f = fields(lines["ppm_char_mixed"])
    assert f["xorder"] == "3" and f["char"] == "1" and f["prec"] == "1"
    assert "xorder=4" in lines["xorder4"] and lines["xorder4"].split()[1] == "error"


def test_gravity_options(lines):
    f = fields(lines["selfgrav_periodic"])
    assert f["sg"] == "1" and f["mode"] == "0" and f["bc"] == "0" and f["rs"] == "3"
    assert float(f["fpg"]) == 1.0 and float(f["thr"]) == 1e-10
    f = fields(lines["iso_multipole"])
    assert f["eos"] == "1" and f["bc"] == "1" and f["mode"] == "1" and f["mporder"] == "4"
    assert float(f["cs"]) == 0.5 and float(f["fpg"]) == 0.7 and f["rs"] == "1"


def test_refusals(lines):
    assert "4 pi G" in lines["selfgrav_no_4piG"]
    assert "differs from ix1_bc" in lines["selfgrav_mixed_faces"]
    assert "zerograd" in lines["selfgrav_zerograd"]
    assert "threshold or niteration" in lines["selfgrav_no_threshold"]
    assert "hllc" in lines["flux_hllc"]
    for k in ("selfgrav_no_4piG", "selfgrav_mixed_faces", "selfgrav_zerograd", "selfgrav_no_threshold", "flux_hllc"):
        assert lines[k].split()[1] == "error", k


def test_check_carries_the_code(lines):
    assert "agx_example failed (-7)" in lines["check"]
