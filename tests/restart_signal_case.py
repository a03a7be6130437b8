"""The workload of the restart-on-signal test (shared by the test and its child process):
24^3 turbulence in 8 MeshBlocks of 12^3, VL2 + PLM + HLLD + CT."""
from oracle.orc import make_desc
from tests.states import turbulent_state


def case():
    desc = make_desc(nx=(24, 24, 24), bnx=(12, 12, 12), box=((0, 1),) * 3, cfl=0.4)
    return desc, turbulent_state(desc, seed=31, mach=1.5)
