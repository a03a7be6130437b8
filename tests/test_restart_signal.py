"""Checkpoint on SIGTERM and resume (CPU): a child process runs the main loop on the
oracle, receives SIGTERM mid-run, leaves the loop after the current cycle and writes
"<basename>.final.rst" in the reference's layout (main.cpp:511-533, restart.cpp:50-56).
A fresh mesh then reads that file (the "-r" path: Mesh::Mesh(pin, resfile),
mesh.cpp:551-700, then Mesh::Initialize) and runs on to 10 cycles: every array equals
an uninterrupted 10-cycle run bit for bit."""
import os
import signal
import subprocess
import sys

import numpy as np

from oracle.orc import Oracle
from tests.conftest import load_agx
from tests.restart_signal_case import case

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("u", "b1f", "b2f", "b3f")


def test_sigterm_final_restart_and_resume(tmp_path):
    agx = load_agx()
    base = str(tmp_path / "turb")
    child = subprocess.Popen([sys.executable, "-u", os.path.join(HERE, "restart_signal_child.py"), base, "40"],
                             stdout=subprocess.PIPE, text=True)
    lines = []
    try:
        for ln in child.stdout:
            lines.append(ln.strip())
            if ln.strip() == "cycle 3":
                child.send_signal(signal.SIGTERM)
                break
        rest, _ = child.communicate(timeout=120)
    finally:
        if child.poll() is None:
            child.kill()
    lines += rest.split()
    assert child.returncode == 0, lines
    path = base + ".final.rst"
    assert os.path.exists(path), lines
    desc, state = case()
    ref = Oracle(desc)
    for n, a in zip(FIELDS, state):
        ref.set(n, a)
    ref.initialize()
    shapes = [ref.get(f).shape for f in FIELDS]
    t, dt, ncyc, *arrs = agx.restart_read_host(path, desc, shapes)
    assert 3 <= ncyc < 10, ncyc
    # the resumed run ("-r"): state, time, dt and cycle from the file, then Initialize
    o = Oracle(desc)
    for n, a in zip(FIELDS, arrs):
        o.set(n, a)
    o.set_time(t, dt, ncyc)
    o.initialize(restarted=True)
    while o.time()[2] < 10:
        o.cycle(1e30)
    for _ in range(10):
        ref.cycle(1e30)
    assert o.time() == ref.time()
    for f in FIELDS + ("w", "bcc"):
        assert np.array_equal(o.get(f), ref.get(f)), f
