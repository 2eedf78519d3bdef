"""g4acc_compute_G_noalloc: the M -> G entry point that allocates nothing (its intermediates live in
a static device array) against g4acc_compute_G, the entry point it is a variant of, through the C
ABI. Two calls in a row on one handle, the second with fewer vertices, so the second runs over
the first one's leftovers in the shared array; FP64 and FP32; the single-band (register-fused)
and three-band (four-stage) M -> G forms. The bar is the parity bar of the two precisions
(max-normalised l_inf 1e-12 / 1e-6); both calls run the same kernels, so equality is expected.
Measurements too large for the array are refused before any launch."""
import ctypes
import math

import numpy as np
import pytest

from dca_amd import geometry as geo
from dca_amd import g4acc, synth
from oracle import bruteforce as bf

pytestmark = pytest.mark.gpu

CASES = {
    # name: (nb, cluster, n_freq, w_ex, k_ex, n_range of the first call)
    "1band_4x4": (1, 4, 8, (0, 1, 2), ((0.0, 0.0), (math.pi, 0.0)), (40, 120)),
    "3band_2x2": (3, 2, 4, (0, 1), ((0.0, 0.0), (math.pi, math.pi)), (6, 24)),
}


def workload(name):
    nb, L, nf, wex, kex, nr = CASES[name]
    return synth.tiny_workload(nb=nb, cluster=geo.square_cluster(L), n_freq=nf, w_ex=wex, k_ex=kex,
                               n_range=nr, seed_cfg=700 + list(CASES).index(name))


def smaller(meas, keep):
    """The measurement restricted to its first `keep` vertices per spin sector."""
    n = [min(keep, x) for x in meas.n]
    return synth.Measurement(n=n, M=[np.asfortranarray(meas.M[s][:n[s], :n[s]]) for s in range(2)],
                             tau=[meas.tau[s][:n[s]].copy() for s in range(2)],
                             site=[meas.site[s][:n[s]].copy() for s in range(2)],
                             band=[meas.band[s][:n[s]].copy() for s in range(2)], sign=meas.sign)


def noalloc_G(acc, torch, meas):
    from dca_amd.devpool import DevicePool
    pool = DevicePool(torch, [meas])
    n = acc.lib.g4acc_g_size(acc._h)
    out = torch.full((2 * n,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()  # inputs written on torch's stream; the library runs on its own
    w = pool.walkers[0]
    acc._check(acc.lib.g4acc_compute_G_noalloc(acc._h, ctypes.byref(w), out.data_ptr(), n))
    acc._check(acc.lib.g4acc_sync(acc._h))
    g = acc.geom
    flat = out.cpu().numpy()
    return (flat[0::2] + 1j * flat[1::2]).reshape((g.nb, g.nb, 2, g.nc, g.nc, g.n_ext, g.n_ext), order="F")


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_noalloc_G_equals_compute_G(name, precision, tol):
    import torch
    wl = workload(name)
    first = wl.draw(1)[0]
    second = smaller(wl.draw(2)[1], max(2, min(first.n) // 3))
    assert max(second.n) < max(first.n)
    acc = g4acc.TpAccumulator(wl.geometry, precision=precision, max_vertices=wl.n_max)
    for call, meas in enumerate((first, second)):
        ref = acc.compute_G(meas)
        got = noalloc_G(acc, torch, meas)
        err = bf.difference(ref, got)["l_inf"]
        print(f"{name} {precision} call {call} (n = {meas.n}): l_inf = {err:.3e}", flush=True)
        assert np.all(np.isfinite(got))
        assert err < tol
    acc.close()


def test_noalloc_refuses_oversized_measurement():
    """A measurement whose intermediates exceed the static array is refused on the host (no launch)."""
    import torch
    L, nf = 8, 32
    wl = synth.tiny_workload(nb=1, cluster=geo.square_cluster(L), n_freq=nf, w_ex=(0,), k_ex=((0.0, 0.0),),
                             n_range=(200, 200), seed_cfg=799)
    acc = g4acc.TpAccumulator(wl.geometry, precision="fp64", max_vertices=wl.n_max)
    g = wl.geometry
    assert 2 * g.D ** 2 * 16 > g4acc_scratch_bytes()  # G alone (FP64, both spins) exceeds it
    with pytest.raises(g4acc.LogicError, match="workspace"):
        noalloc_G(acc, torch, wl.draw(1)[0])
    acc.close()


def g4acc_scratch_bytes():
    return 64 << 20
