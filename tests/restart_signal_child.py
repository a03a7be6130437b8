"""Child process of tests/test_restart_signal.py: the main loop of main.cpp:430-515 on
the CPU oracle, with the reference's signal handling (SignalHandler: SIGTERM sets a flag,
checked after every cycle, main.cpp:511-514) and its final restart output
("<basename>.final.rst", outputs/restart.cpp:50-56) written by the library's host
writer (agx_restart_write_host).  Usage: restart_signal_child.py <basename> <max cycles>."""
import os
import signal
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle.orc import Oracle  # noqa: E402
from tests.conftest import load_agx  # noqa: E402
from tests.restart_signal_case import case  # noqa: E402


def main():
    base, nmax = sys.argv[1], int(sys.argv[2])
    flags = {"term": False}
    signal.signal(signal.SIGTERM, lambda *_: flags.update(term=True))
    agx = load_agx()
    desc, state = case()
    o = Oracle(desc)
    for n, a in zip(("u", "b1f", "b2f", "b3f"), state):
        o.set(n, a)
    o.initialize()
    print("ready", flush=True)
    for n in range(nmax):
        o.cycle(1e30)
        print(f"cycle {n + 1}", flush=True)
        if flags["term"]:
            break
        time.sleep(0.2)   # a long run: the parent signals while cycles are still going
    if flags["term"]:
        t, dt, ncyc = o.time()
        agx.restart_write_host(base + ".final.rst", desc, t, dt, ncyc, o.get("u"), o.get("b1f"), o.get("b2f"),
                               o.get("b3f"))
        print(f"final restart at cycle {ncyc}", flush=True)


if __name__ == "__main__":
    main()
