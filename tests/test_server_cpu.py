"""Host side of the serving loop, without a device: request bookkeeping,
argument checks, the latency percentile and the request generator."""
import numpy as np
import pytest

from dmd_amd import _abi
from dmd_amd import workloads as W
from dmd_amd.server import Server, weighted_percentile


def has_device():
    try:
        import ctypes as C
        lib = _abi.load()
        n = C.c_int32(0)
        return lib.dmdx_device_info(C.byref(n), None, None) == 0 and n.value > 0
    except Exception:
        return False


def test_weighted_percentile():
    v = np.array([5.0, 1.0, 3.0])
    w = np.array([1, 1, 2])
    # sorted: 1 (w1), 3 (w2), 5 (w1): cumulative 1, 3, 4
    assert weighted_percentile(v, w, 25) == 1.0
    assert weighted_percentile(v, w, 50) == 3.0
    assert weighted_percentile(v, w, 75) == 3.0
    assert weighted_percentile(v, w, 99) == 5.0
    # equal weights reduce to the nearest-rank percentile
    x = np.arange(1, 101, dtype=float)
    assert weighted_percentile(x, np.ones(100), 99) == 99.0
    assert weighted_percentile(x, np.ones(100), 50) == 50.0


def test_serving_requests_ranges():
    reqs = W.serving_requests(40, seed=9)
    kinds = {"lj": 0, "water": 0}
    for spec, n in reqs:
        assert 1000 <= n <= 20000
        if spec.name.startswith("lj"):
            kinds["lj"] += 1
            assert 1000 <= spec.n <= 32000 and spec.coulomb_type == _abi.COULOMB_NONE
        else:
            kinds["water"] += 1
            assert spec.n % 3 == 0 and 216 <= spec.n // 3 <= 8192
            g = spec.pme_grid[0]
            k = g
            for p in (2, 3, 5, 7):
                while k % p == 0:
                    k //= p
            assert k == 1 and abs(g - spec.box[0] / 0.1) < 0.25 * g
        assert spec.precision == _abi.MIXED
    assert kinds["lj"] > 5 and kinds["water"] > 5
    # same seed, same stream
    again = W.serving_requests(40, seed=9)
    assert all(a.n == b.n and na == nb and np.array_equal(a.pos, b.pos)
               for (a, na), (b, nb) in zip(reqs, again))


@pytest.mark.skipif(has_device(), reason="checks the no-device error path")
def test_server_bookkeeping_without_device():
    srv = Server(hbm_budget_bytes=1 << 30)
    a = W.lj_fluid(500, 0.8, 1.0, 1)
    b = W.chain_melt(400, 2)
    ia, ib = srv.submit(a, 10), srv.submit(b, 20)
    assert ib == ia + 1
    for rid, n, steps in ((ia, 500, 10), (ib, 400, 20)):
        info = srv.info(rid)
        assert info["state"] == "queued" and info["n_atoms"] == n
        assert info["steps_target"] == steps and info["est_bytes"] > 0
    st = srv.stats()
    assert st["queued"] == 2 and st["running"] == 0 and st["bytes_budget"] == 1 << 30
    with pytest.raises(_abi.DmdxError):
        srv.submit(a, 0)
    with pytest.raises(_abi.DmdxError):
        srv.result(ia)  # not finished
    with pytest.raises(_abi.DmdxError):
        srv.release(ia)
    with pytest.raises(_abi.DmdxError):
        srv.pump()  # no device to build a batch on
    srv.close()
