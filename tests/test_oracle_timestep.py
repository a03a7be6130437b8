"""Pins the oracle's time step (Hydro::NewBlockTimeStep, hydro/new_blockdt.cpp:42-190,
and Mesh::NewTimeStep, mesh/mesh.cpp:1078-1119) against closed forms.

On a uniform MHD state every cell gives the same bound, so the mesh's time step is

    dt = cfl * min_d  dx_d / (|v_d| + c_f,d),

with c_f,d the fast magnetosonic speed along axis d (field component B_d normal,
the other two transverse):

    c_f,d^2 = ( a^2 + v_A^2 + sqrt((a^2 + v_A^2)^2 - 4 a^2 B_d^2 / rho) ) / 2,

a^2 = gamma P / rho (adiabatic) or c_s^2 (isothermal), v_A^2 = |B|^2 / rho.  The
formula here is the textbook one, written independently of the oracle's operand
order.  Cell widths and velocities differ per axis, and each axis is made the binding
one in turn, so an axis mix-up (a wrong velocity, width or field component for one
axis) changes the result.  Mesh::NewTimeStep is pinned by its two closed-form rules:
growth capped at twice the last step, and the last step clipped to reach tlim.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc

GAMMA = 5.0 / 3.0
CFL = 0.3


def fast_speed(rho, p, B, d, iso_cs=None):
    a2 = iso_cs ** 2 if iso_cs is not None else GAMMA * p / rho
    va2 = (B[0] ** 2 + B[1] ** 2 + B[2] ** 2) / rho
    s = a2 + va2
    return np.sqrt(0.5 * (s + np.sqrt(s * s - 4.0 * a2 * B[d] ** 2 / rho)))


def uniform_oracle(rho, v, p, B, nx=(16, 8, 8), box=((0.0, 1.0), (0.0, 0.25), (0.0, 0.5)),
                   bnx=(8, 8, 4), eos=0):
    o = Oracle(make_desc(nx=nx, bnx=bnx, box=box, cfl=CFL, eos=eos, iso_cs=1.3))
    u = o.view("u")
    u[:, 0] = rho
    for d in range(3):
        u[:, 1 + d] = rho * v[d]
    ke = 0.5 * rho * sum(x * x for x in v)
    u[:, 4] = (p / (GAMMA - 1.0) + ke + 0.5 * sum(x * x for x in B)) if eos == 0 else 0.0
    for d, name in enumerate(("b1f", "b2f", "b3f")):
        o.view(name)[...] = B[d]
    return o


def closed_form_dt(rho, v, p, B, dx, iso_cs=None):
    return CFL * min(dx[d] / (abs(v[d]) + fast_speed(rho, p, B, d, iso_cs)) for d in range(3))


# three velocity orderings: with the cell widths (1/16, 1/32, 1/16) each makes a
# different axis the binding one
CASES = [((3.0, -0.2, 0.1), 0), ((0.1, -4.0, 0.3), 1), ((0.2, 0.1, -5.0), 2)]


@pytest.mark.parametrize("v,axis", CASES)
@pytest.mark.parametrize("eos", [0, 1])
def test_block_dt_closed_form(v, axis, eos):
    rho, p, B = 1.7, 0.8, (0.9, -0.4, 0.6)
    o = uniform_oracle(rho, v, p, B, eos=eos)
    o.initialize()
    dx = (1.0 / 16, 0.25 / 8, 0.5 / 8)
    iso = 1.3 if eos == 1 else None
    dt_ref = closed_form_dt(rho, v, p, B, dx, iso)
    per_axis = [dx[d] / (abs(v[d]) + fast_speed(rho, p, B, d, iso)) for d in range(3)]
    assert int(np.argmin(per_axis)) == axis, "test setup: intended axis is not the binding one"
    dt = o.L.orc_new_block_dt(o.h)
    assert abs(dt - dt_ref) <= 1e-14 * dt_ref, (dt, dt_ref)
    # Mesh::Initialize's first step is the block bound itself (no previous step)
    assert abs(o.time()[1] - dt_ref) <= 1e-14 * dt_ref


def test_block_dt_uses_face_field_jump():
    """new_blockdt.cpp:78-97 takes the normal field as bcc + |b_face - bcc|: with
    unequal faces on each side of a cell the bound uses the larger magnitude."""
    rho, p, v = 1.0, 0.6, (0.0, 0.0, 0.0)
    B = (0.0, 0.0, 0.0)
    o = uniform_oracle(rho, v, p, B, nx=(8, 8, 8), box=((0.0, 1.0),) * 3, bnx=(8, 8, 8))
    # alternate x1 faces between 0 and 2: bcc1 = 1 everywhere, |b_face - bcc| = 1
    b1 = o.view("b1f")
    b1[..., 0::2] = 0.0
    b1[..., 1::2] = 2.0
    u = o.view("u")
    o.L.orc_cons2prim(o.h)   # bcc from the faces
    bcc1 = o.get("bcc")[:, 0, 2:-2, 2:-2, 2:-2]
    assert np.allclose(bcc1, 1.0)
    # the energy holds the magnetic energy of bcc so that P = p
    u[:, 4] = p / (GAMMA - 1.0) + 0.5 * 1.0
    o.L.orc_cons2prim(o.h)
    dt = o.L.orc_new_block_dt(o.h)
    # x1 bound with normal field bx = 1 + |0 - 1| = 2 at the lower face of even cells,
    # 1 + |2 - 1| = 2 at odd ones; transverse bound: bx = 0, transverse field (1, 0)
    dx = 1.0 / 8
    bx1 = fast_speed(rho, p, (2.0, 0.0, 0.0), 0)   # normal 2, no transverse field
    bx2 = fast_speed(rho, p, (1.0, 0.0, 0.0), 1)   # normal 0, transverse bcc1 = 1
    ref = CFL * dx / max(bx1, bx2)
    assert abs(dt - ref) <= 1e-14 * ref, (dt, ref)


def test_new_time_step_growth_cap_and_tlim():
    rho, p, v, B = 1.0, 0.6, (0.5, 0.2, -0.1), (0.3, 0.2, 0.1)
    o = uniform_oracle(rho, v, p, B)
    o.initialize()
    bdt = o.L.orc_new_block_dt(o.h)
    # growth capped at twice the previous step (mesh.cpp:1083)
    o.set_time(0.0, 0.25 * bdt, 3)
    o.L.orc_new_time_step(o.h, 1e9)
    assert o.time()[1] == 0.5 * bdt
    # no cap when the previous step was large
    o.set_time(0.0, 10.0 * bdt, 3)
    o.L.orc_new_time_step(o.h, 1e9)
    assert o.time()[1] == bdt
    # the last step reaches tlim exactly (mesh.cpp:1104-1106)
    o.set_time(1.0, bdt, 3)
    o.L.orc_new_time_step(o.h, 1.0 + 0.3 * bdt)
    t, dt, _ = o.time()
    assert dt == (1.0 + 0.3 * bdt) - 1.0
