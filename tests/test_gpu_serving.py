"""The serving path: many fits' data in one device pool, one batched launch per round of
requests, native BFGS per fit.

Bars: a batched request equals the single-set pass (tpa_eval) on the same events to 1e-12
relative (only the summation order differs); a request's result is bit-identical whether it
is evaluated alone or among other requests, in any order; a fit run alone and the same fit run
among others end bit-identically; the server's NLL and gradient equal FCN's mixed-precision
ones to 1e-9 relative; a served fit ends at least as low as scipy's BFGS on FCN from the same
start (to 1e-4), at a point from which scipy finds nothing lower.
"""

import numpy as np
import pytest

from helpers import make_case

pytestmark = pytest.mark.gpu

# sample sizes around the pool's 2048-event chunk and a single event
SIZES = [1, 777, 2048, 2049, 3100]


def _slice(p4, a, b):
    return {k: v[a:b] for k, v in p4.items()}


@pytest.fixture(scope="module")
def case6():
    return make_case("lc_pkpi_6res.yml", n_mc=12000, n_data=sum(SIZES), seed=31)


@pytest.fixture(scope="module")
def case2res():
    return make_case("lc_pkpi_2res.yml", n_mc=20000, n_data=12000, seed=41)


def _random_coefs(amp, vm, rng, m):
    x0 = np.array(vm.get_all_val())
    out = []
    for _ in range(m):
        vm.set_all(x0 + rng.normal(scale=0.2, size=len(x0)))
        out.append(amp.coefficients())
    vm.set_all(x0)
    return np.array(out)


def test_pool_batch_matches_single_sets(case6, gpu_available):
    from tf_pwa_amd import EventPool
    from tf_pwa_amd import _lib

    conf = case6["conf"]
    amp = conf.get_amplitude()
    pool = EventPool(amp, capacity=20000, max_items=16)
    rng = np.random.default_rng(5)
    slots, sets, a = [], [], 0
    for n in SIZES:
        p4 = _slice(case6["data"], a, a + n)
        w = rng.uniform(0.5, 1.5, n)
        slots.append(pool.add(p4, w, order=conf.dat_order))
        ev = conf.data_from_p4(p4)
        ev.set_weight(w * (w.sum() / (w * w).sum()))
        sets.append(ev)
        a += n
    coefs = _random_coefs(amp, conf.vm, rng, len(SIZES))
    val, G, _ = pool.eval(slots, coefs)
    for i, ev in enumerate(sets):
        v, g, _ = ev.eval(coefs[i], _lib.MODE_DATA, grad=True)
        assert abs(val[i] - v) <= 1e-12 * max(1.0, abs(v)), (i, val[i], v)
        np.testing.assert_allclose(G[i], g, rtol=1e-11, atol=1e-11 * np.abs(g).max())
    # alone == in the mix, and the batch order does not matter: bit-identical
    for i in range(len(SIZES)):
        v1, g1, _ = pool.eval([slots[i]], coefs[i:i + 1])
        assert v1[0] == val[i] and np.array_equal(g1[0], G[i])
    perm = rng.permutation(len(SIZES))
    vp, gp, _ = pool.eval([slots[j] for j in perm], coefs[perm])
    assert np.array_equal(vp, val[perm]) and np.array_equal(gp, G[perm])
    # the same sample twice in one batch with different coefficients
    v2, g2, _ = pool.eval([slots[3], slots[3]], coefs[[3, 0]])
    assert v2[0] == val[3] and np.array_equal(g2[0], G[3])
    pool.close()


def test_pool_fp32_batch_matches_fp32_single_sets(case6, gpu_available):
    """The pool's FP32 copy (tpa_pool_keep_f32), made for the ranges already allocated when it
    is switched on and by every later add: a batched request equals the single-set pass over
    the set's own FP32 copy (TPA_FLAG_F32) to 1e-12, alone or in the mix bit-identically, and
    the FP64 pool's result to the FP32 rounding (1e-6 relative at these sizes)."""
    import ctypes

    from tf_pwa_amd import EventPool
    from tf_pwa_amd import _lib

    conf = case6["conf"]
    amp = conf.get_amplitude()
    pool = EventPool(amp, capacity=20000, max_items=16)
    rng = np.random.default_rng(6)
    slots, sets, a = [], [], 0
    for i, n in enumerate(SIZES):
        if i == 2:  # switched on with two ranges allocated
            _lib.check(_lib.lib().tpa_pool_keep_f32(pool._h, 1))
        p4 = _slice(case6["data"], a, a + n)
        w = rng.uniform(0.5, 1.5, n)
        slots.append(pool.add(p4, w, order=conf.dat_order))
        ev = conf.data_from_p4(p4)
        ev.set_weight(w * (w.sum() / (w * w).sum()))
        sets.append(ev)
        a += n
    coefs = _random_coefs(amp, conf.vm, rng, len(SIZES))
    val, G, _ = pool.eval(slots, coefs)
    for i, ev in enumerate(sets):
        v, g, _ = ev.eval(coefs[i], _lib.MODE_DATA, grad=True, f32=True)
        assert abs(val[i] - v) <= 1e-12 * max(1.0, abs(v)), (i, val[i], v)
        np.testing.assert_allclose(G[i], g, rtol=1e-11, atol=1e-11 * np.abs(g).max())
        v64, g64, _ = ev.eval(coefs[i], _lib.MODE_DATA, grad=True)
        assert abs(val[i] - v64) <= 1e-6 * max(1.0, abs(v64)), (i, val[i], v64)
    for i in range(len(SIZES)):
        v1, g1, _ = pool.eval([slots[i]], coefs[i:i + 1])
        assert v1[0] == val[i] and np.array_equal(g1[0], G[i])
    # switched off: the batched passes read the FP64 cache again
    _lib.check(_lib.lib().tpa_pool_keep_f32(pool._h, 0))
    v0, _, _ = pool.eval(slots[:1], coefs[:1])
    ref, _, _ = sets[0].eval(coefs[0], _lib.MODE_DATA, grad=True)
    assert abs(v0[0] - ref) <= 1e-12 * max(1.0, abs(ref))
    pool.close()


def test_pool_admission_and_release(case6, gpu_available):
    from tf_pwa_amd import EventPool, PoolFull

    conf = case6["conf"]
    amp = conf.get_amplitude()
    pool = EventPool(amp, capacity=5000, max_items=4)
    cap = pool.capacity
    assert cap == 5056  # rounded up to the 64-event range granule (the row stride)
    p4 = case6["data"]
    # ranges are granted in multiples of 64 events (aligned chunk starts)
    a = pool.add(_slice(p4, 0, 2000), order=conf.dat_order)
    b = pool.add(_slice(p4, 2000, 4000), order=conf.dat_order)
    assert (a.offset, b.offset) == (0, 2048) and pool.free_events() == cap - 4096
    with pytest.raises(PoolFull):
        pool.add(_slice(p4, 0, 1500), order=conf.dat_order)
    pool.release(a)
    c = pool.add(_slice(p4, 0, 1500), order=conf.dat_order)
    assert c.offset == 0 and pool.free_events() == (2048 - 1536) + (cap - 4096)
    pool.release(b)
    pool.release(c)
    assert pool.free_events() == cap
    d = pool.add(_slice(p4, 0, 5000), order=conf.dat_order)  # free ranges merged again
    assert d.offset == 0
    with pytest.raises(Exception):
        pool.eval(None, np.zeros((1, amp.P)), ranges=[(4000, 2000)])  # past the allocation
    pool.close()


def _server(case, capacity=60000, **kw):
    from tf_pwa_amd import FitServer

    conf = case["conf"]
    amp = conf.get_amplitude()
    mc = conf.data_from_p4(case["mc"])
    return FitServer(amp, mc, capacity=capacity, **kw), mc


def test_server_nll_grad_matches_fcn(case6, gpu_available):
    from tf_pwa_amd import FCN

    conf = case6["conf"]
    srv, mc = _server(case6)
    p4 = _slice(case6["data"], 0, 5000)
    slot = srv.add_data(p4, order=conf.dat_order)
    data = conf.data_from_p4(p4)
    fcn = FCN(conf.get_model(), data, mc, precision="mixed")
    rng = np.random.default_rng(8)
    x0 = np.array(conf.vm.get_all_val())
    for _ in range(3):
        y = x0 + rng.normal(scale=0.1, size=len(x0))
        n1, g1 = srv.nll_grad(slot, y)
        n2, g2 = fcn.nll_grad(y)
        assert abs(n1 - n2) <= 1e-9 * abs(n2), (n1, n2)
        assert np.abs(g1 - g2).max() <= 1e-9 * np.abs(g2).max()
    conf.vm.set_all(x0)
    srv.close()


def _starts(conf, m, seed, scale=0.15):
    rng = np.random.default_rng(seed)
    x0 = np.array(conf.vm.get_all_val())
    return [x0 + rng.normal(scale=scale, size=len(x0)) for _ in range(m)]


def _wait(srv, n, timeout=120.0):
    import time

    done, t0 = [], time.time()
    while len(done) < n:
        done += srv.poll()
        assert time.time() - t0 < timeout, "fits did not end"
        time.sleep(0.001)
    return done


def test_fit_alone_equals_fit_in_the_mix(case2res, gpu_available):
    conf = case2res["conf"]
    srv, _ = _server(case2res, maxiter=300)
    sizes = [3000, 1200, 2500, 4000, 1300]
    slots, a = [], 0
    for n in sizes:
        slots.append(srv.add_data(_slice(case2res["data"], a, a + n), order=conf.dat_order))
        a += n
    starts = _starts(conf, len(sizes), 2)
    alone = srv.submit(slots[0], starts[0])
    _wait(srv, 1)
    r_alone = srv.raw_result(alone)
    ids = [srv.submit(s, y0) for s, y0 in zip(slots, starts)]
    _wait(srv, len(ids))
    r_mix = srv.raw_result(ids[0])
    assert r_alone[2] in (0, 1, 2) and r_alone[3] > 5
    assert np.array_equal(r_alone[0], r_mix[0]) and r_alone[1] == r_mix[1]
    assert r_alone[2:] == r_mix[2:]
    c = srv.counters()
    assert c["batches"] < c["requests"]  # requests of different fits shared launches
    srv.close()


def test_served_fit_reaches_scipy_minimum(case2res, gpu_available):
    from tf_pwa_amd import FCN
    from tf_pwa_amd.fit import fit_scipy

    conf = case2res["conf"]
    srv, mc = _server(case2res)
    p4 = _slice(case2res["data"], 0, 6000)
    slot = srv.add_data(p4, order=conf.dat_order)
    y0 = _starts(conf, 1, 9, scale=0.1)[0]
    fid = srv.submit(slot, y0)
    _wait(srv, 1)
    res = srv.result(fid)
    x_true = np.array(conf.vm.get_all_val())
    conf.vm.set_all(y0)
    fcn = FCN(conf.get_model(), conf.data_from_p4(p4), mc, precision="mixed")
    ref = fit_scipy(fcn)
    assert res.message in ("converged", "line search failed (precision loss)")
    # at least as deep as scipy's BFGS (which may stop early on precision loss), and a
    # minimum: scipy restarted from the served result finds nothing lower
    assert res.min_nll <= ref.min_nll + 1e-4, (res.min_nll, ref.min_nll)
    y = np.array([res.params[n] for n in conf.vm.trainable_vars])
    conf.vm.set_all(y)
    again = fit_scipy(fcn)
    assert again.min_nll >= res.min_nll - 1e-4, (again.min_nll, res.min_nll)
    conf.vm.set_all(x_true)
    # the reported minimum is the NLL at the reported parameters
    n_at, _ = srv.nll_grad(slot, y)
    assert abs(n_at - res.min_nll) < 1e-6
    srv.close()


def test_served_fit_with_fp32_pool_matches_fp64_pool(case2res, gpu_available):
    """A fit served from the pool's FP32 copy ends at the FP64 pool's minimum to the FP32
    rounding of the data waves (|dNLL| < 1e-4 at 6000 events)."""
    conf = case2res["conf"]
    p4 = _slice(case2res["data"], 0, 6000)
    y0 = _starts(conf, 1, 9, scale=0.1)[0]
    out = {}
    for storage in ("fp64", "fp32"):
        srv, _ = _server(case2res, data_storage=storage)
        slot = srv.add_data(p4, order=conf.dat_order)
        fid = srv.submit(slot, y0)
        _wait(srv, 1)
        out[storage] = srv.result(fid)
        srv.close()
    assert out["fp32"].message in ("converged", "line search failed (precision loss)")
    assert abs(out["fp32"].min_nll - out["fp64"].min_nll) < 1e-4, (
        out["fp32"].min_nll, out["fp64"].min_nll)


def test_serve_loop_admission_and_results(case2res, gpu_available):
    """Poisson arrivals into a pool too small for all samples at once: admission waits for
    fits to end, and every fit ends exactly as it does alone."""
    conf = case2res["conf"]
    sizes = [2500, 3000, 1800, 2200, 1500, 1000]
    starts = _starts(conf, len(sizes), 4)
    jobs, a = [], 0
    for n, y0 in zip(sizes, starts):
        jobs.append(dict(p4=_slice(case2res["data"], a, a + n), y0=y0))
        a += n
    srv, _ = _server(case2res, capacity=6000, maxiter=300)
    alone = []
    for j in jobs:
        s = srv.add_data(j["p4"])
        fid = srv.submit(s, j["y0"])
        _wait(srv, 1)
        alone.append(srv.raw_result(fid))
        srv.release(s)
    arrivals = np.cumsum(np.random.default_rng(1).exponential(0.002, len(jobs)))
    res, times = srv.serve(jobs, arrivals, timeout_s=120)
    assert srv.peak_events <= 6000
    assert any(t["admitted"] - t["arrival"] > 0.001 for t in times)  # someone had to wait
    for r, (y, nll, st, nfev, nit) in zip(res, alone):
        assert r.min_nll == nll and r.nfev == nfev
    lat = srv.latencies_us()
    assert len(lat) == srv.counters()["requests"] and np.all(lat > 0)
    srv.close()


def test_pool_add_events_equals_add_from_momenta(case2res, gpu_available):
    """A sample generated on the device enters the pool through tpa_pool_add_events (its angles
    copied on the device) and gives bit-identical batched sums to the same events uploaded as
    four-momenta (tpa_pool_add), for a whole generated set with weights and for a selected
    subset (the accept-reject path), from FP64 and FP32 pools. A set of another model is
    refused."""
    from conftest import config_path
    from tf_pwa_amd import ConfigLoader, EventPool, EventSet

    conf = case2res["conf"]
    amp = conf.get_amplitude()
    n = 3100
    ev, p4 = EventSet.generate(amp, n, seed=77, return_p4=True)
    idx = np.sort(np.random.default_rng(3).choice(n, 1234, replace=False))
    sub = ev.select(idx)
    w = np.random.default_rng(4).uniform(0.5, 1.5, n)
    c = _random_coefs(amp, conf.vm, np.random.default_rng(5), 2)
    other = ConfigLoader(config_path("lc_pkpi_2res.yml"), seed=1).get_amplitude()
    foreign = EventSet.generate(other, 100, seed=1)
    for storage in ("fp64", "fp32"):
        pool = EventPool(amp, capacity=20000, max_items=8, data_storage=storage)
        a = pool.add(p4, weight=w, engine_layout=True)
        b = pool.add_events(ev, weight=w)
        s1 = pool.add(np.ascontiguousarray(p4[:, idx]), engine_layout=True)
        s2 = pool.add_events(sub)
        assert (b.n, s2.n) == (n, idx.size) and b.sw == a.sw and s2.sw == s1.sw
        v, G, _ = pool.eval([a, b, s1, s2], np.array([c[0], c[0], c[1], c[1]]))
        assert v[0] == v[1] and np.array_equal(G[0], G[1]), storage
        assert v[2] == v[3] and np.array_equal(G[2], G[3]), storage
        assert v[0] != v[2]
        with pytest.raises(Exception, match="another model"):
            pool.add_events(foreign)
        pool.close()
