"""agx_history_sum_noalloc (partials in a static device array) against agx_history_sum
(partials in memory it allocates), through the C ABI.

Both run the same kernels with the same work split on the same cells, so the sums must
be identical bit for bit.  Two calls in a row, the second on a smaller input (fewer
blocks, smaller blocks, no potential): the array keeps the first call's partials, and
the second call must not read any of them.  Inputs beyond the 32-bit row index are
rejected before any device work.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dense_state(nb, nx, ng, nh, seed):
    rs = np.random.default_rng(seed)
    n = tuple(x + 2 * ng for x in nx[::-1])
    u = rs.standard_normal((nb, nh) + n)
    u[:, 0] = 1.0 + rs.random((nb,) + n)
    bcc = rs.standard_normal((nb, 3) + n)
    phi = rs.standard_normal((nb,) + n)
    vol = 1.0 / (np.arange(nb) + 7.0) ** 3
    return u, bcc, phi, vol


def run(agx, fn, D, nb, nx, ng, nh, use_phi):
    L = agx.lib()
    out = agx.DeviceArray(np.full(16, np.nan))
    rc = getattr(L, fn)(D["u"].ptr, D["bcc"].ptr, D["phi"].ptr if use_phi else None, nb, nh, nx[0], nx[1], nx[2],
                        ng, D["vol"].ptr, out.ptr, None)
    assert rc == 0, (fn, L.agx_last_error())
    assert L.agx_ref_sync(None) == 0
    h = out.get()
    out.close()
    return h


def test_noalloc_equals_allocating_entry_point(agx):
    # 64 blocks of 32x16x16 (128 partial rows), then a smaller isothermal input without
    # phi, 8 blocks of 16x8x16 (8 partial rows): a final sum that read past this call's
    # rows would pick up the first call's partials
    cases = [(64, (32, 16, 16), 2, 5, True, 1), (8, (16, 8, 16), 3, 4, False, 2)]
    inputs = []
    for nb, nx, ng, nh, use_phi, seed in cases:
        u, bcc, phi, vol = dense_state(nb, nx, ng, nh, seed)
        inputs.append({k: agx.DeviceArray(v) for k, v in (("u", u), ("bcc", bcc), ("phi", phi), ("vol", vol))})
    # the allocation-free entry point twice in a row, then the allocating one on each input
    got = [run(agx, "agx_history_sum_noalloc", D, *c[:5]) for D, c in zip(inputs, cases)]
    ref = [run(agx, "agx_history_sum", D, *c[:5]) for D, c in zip(inputs, cases)]
    for (nb, nx, ng, nh, use_phi, _), g, r in zip(cases, got, ref):
        nout = nh + 3 + (1 if use_phi else 0) + 3
        print(f"nb {nb} nx {nx} ng {ng} nhydro {nh} phi {use_phi}: {g[:nout]}", flush=True)
        assert np.isfinite(g[:nout]).all()
        assert np.array_equal(g[:nout], r[:nout])
        assert np.isnan(g[nout:]).all()     # nothing written past the sums
    for D in inputs:
        for a in D.values():
            a.close()


def test_noalloc_rejects_oversized_input(agx):
    L = agx.lib()
    d = agx.DeviceArray(np.zeros(16))
    # 2^20 blocks of 64 x 64 rows: beyond the 32-bit row index; refused before any launch
    rc = L.agx_history_sum_noalloc(d.ptr, d.ptr, None, 1 << 20, 5, 8, 64, 64, 2, d.ptr, d.ptr, None)
    assert rc < 0
    assert b"2^31" in L.agx_last_error()
    assert L.agx_history_sum_noalloc(None, d.ptr, None, 1, 5, 8, 8, 8, 2, d.ptr, d.ptr, None) < 0
    d.close()
