"""Host side of the device path, no GPU: the C ABI exports, Create/Connect/SetStatus
semantics and error behaviour of the kernel, and connectivity identical to the oracle's.

Nothing here calls Prepare/Run (those touch the device).
"""

import ctypes
import os
import re

import numpy as np
import pytest

import netbuild as nb
import nestgpu
from _oracle import ROOT, Oracle

HEADER = os.path.join(ROOT, "include", "nestgpu.h")


def _declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(nestgpu_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    names = _declared_symbols()
    assert len(names) >= 20
    lib = ctypes.CDLL(nestgpu.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _conns_gpu(k):
    c = k.GetConnections()
    return c["source"], c["target"], c["weight"], c["delay_steps"]


CASES = [
    # (n_vp, seed, pre_n, post_n, rule, param, weight, delay, autapses, multapses, same_pop)
    (1, 1, 50, 40, "fixed_indegree", 7, ("normal", 1.0, 0.3), ("uniform", 0.5, 2.0), True, True, False),
    (3, 11, 60, 60, "fixed_indegree", 9, ("sum", 2.0, ("normal", 0.0, 1.0)), 1.5, False, False, True),
    (4, 5, 30, 45, "pairwise_bernoulli", 0.2, ("max", ("normal", 0.5, 1.0), 0.0), ("redraw", ("normal", 2.0, 1.0), 0.1, 4.0), True, True, False),
    (2, 77, 25, 25, "pairwise_bernoulli", 0.3, 1.0, 1.0, False, True, True),
    (5, 3, 20, 20, "one_to_one", 0, ("uniform", -1.0, 1.0), ("sum", 1.0, ("product", ("uniform_int", 21), 0.1)), True, True, False),
    (3, 8, 12, 17, "all_to_all", 0, ("normal", -2.0, 0.5), 2.0, True, True, False),
    (2, 9, 15, 15, "all_to_all", 0, 0.5, 1.0, False, True, True),
    (6, 123, 40, 30, "fixed_indegree", 25, ("min", ("normal", 0.0, 2.0), 1.0), ("uniform", 0.1, 3.0), True, False, False),
    # many normal and uniform draws per VP stream: the canonical and polar-method draws
    (2, 31, 400, 300, "fixed_indegree", 50, ("normal", 0.0, 1.0), ("uniform", 0.1, 5.0), True, True, False),
    # lognormal weights, exponential delays
    (3, 41, 120, 90, "fixed_indegree", 20, ("lognormal", 0.5, 0.3), ("sum", 0.5, ("exponential", 1.5)), True, True,
     False),
    # nest.math and parameter arithmetic: exp, sin, cos, quotient, pow
    (2, 43, 80, 70, "fixed_indegree", 15, ("quotient", ("exp", ("normal", 0.0, 0.5)), ("sum", 2.0, ("sin", ("uniform", 0.0, 3.0)))),
     ("sum", 0.5, ("pow", ("cos", ("uniform", 0.0, 1.0)), 2.0)), True, True, False),
    # nest.logic.conditional on a comparison: only the chosen branch draws
    (3, 47, 60, 50, "fixed_indegree", 12,
     ("conditional", ("compare", ("uniform", 0.0, 1.0), 0.3, 0), ("normal", 2.0, 0.5), ("exponential", 0.5)),
     ("sum", 1.0, ("compare", ("uniform", 0.0, 1.0), 0.5, 5)), True, True, False),
]


@pytest.mark.parametrize("rng_type", ["mt19937_64", "Philox_32"])
@pytest.mark.parametrize("case", CASES, ids=[c[4] + "-vp%d" % c[0] for c in CASES])
def test_connectivity_identical_to_oracle(case, rng_type):
    # both of NEST's per-VP stream types: mt19937_64 (the default) and Philox_32
    n_vp, seed, npre, npost, rule, rp, w, d, aut, mult, same = case
    pops = [("iaf_psc_delta", npre, {})] if same else [("iaf_psc_delta", npre, {}), ("iaf_psc_exp", npost, {})]
    post = (0, 0, npre) if same else (1, 0, npost)
    desc = dict(n_vp=n_vp, seed=seed, pops=pops, conns=[((0, 0, npre), post, rule, rp, w, d, aut, mult)],
                rng_type=rng_type)
    k, kp, sr = nb.build_gpu(desc)
    o, op = nb.build_oracle(desc)
    s1, t1, w1, d1 = _conns_gpu(k)
    # drop the recorder connections (devices carry no synapses in the oracle)
    keep = t1 != sr.global_id
    s1, t1, w1, d1 = s1[keep], t1[keep], w1[keep], d1[keep]
    s2, t2, w2, d2 = o.connections()
    assert len(s1) == len(s2)
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2)
    assert np.array_equal(w1, w2)
    assert np.array_equal(d1, d2)
    if not aut:
        assert not np.any(s1 == t1)
    if not mult:
        pairs = s1 * 100000 + t1
        assert len(np.unique(pairs)) == len(pairs)


def test_connectivity_independent_of_call_granularity_per_vp():
    # the per-VP streams are consumed in connect order: two Connect calls equal one oracle build
    desc = dict(n_vp=3, seed=4, pops=[("iaf_psc_delta", 30, {})],
                conns=[((0, 0, 20), (0, 0, 30), "fixed_indegree", 3, ("normal", 1.0, 1.0), 1.0),
                       ((0, 20, 30), (0, 0, 30), "pairwise_bernoulli", 0.4, -2.0, ("uniform", 1.0, 2.0))])
    k, kp, sr = nb.build_gpu(desc)
    o, op = nb.build_oracle(desc)
    s1, t1, w1, d1 = _conns_gpu(k)
    keep = t1 != sr.global_id
    s2, t2, w2, d2 = o.connections()
    assert np.array_equal(s1[keep], s2) and np.array_equal(w1[keep], w2) and np.array_equal(d1[keep], d2)


def test_parameter_object_continues_across_connect_calls():
    # one normal parameter object in several Connect calls (as the 8-population circuit's
    # builder uses its weight and delay objects): its per-VP distributions, with the cached
    # second polar value, carry over from call to call (parameter.cpp:95-117); a second object
    # with the same mean and std starts its own
    w = ("normal", 1.0, 0.5, 7)
    d = ("max", ("normal", 1.5, 0.75, 8), 0.1)
    desc = dict(n_vp=3, seed=19, pops=[("iaf_psc_exp", 90, {})],
                conns=[((0, 0, 30), (0, 0, 30), "pairwise_bernoulli", 0.1, w, d),
                       ((0, 0, 30), (0, 30, 60), "pairwise_bernoulli", 0.13, w, d),
                       ((0, 0, 30), (0, 60, 90), "fixed_indegree", 3, ("normal", 1.0, 0.5), d)])
    k, kp, sr = nb.build_gpu(desc)
    o, op = nb.build_oracle(desc)
    s1, t1, w1, d1 = _conns_gpu(k)
    keep = t1 != sr.global_id
    s2, t2, w2, d2 = o.connections()
    assert np.array_equal(s1[keep], s2) and np.array_equal(t1[keep], t2)
    assert np.array_equal(w1[keep], w2) and np.array_equal(d1[keep], d2)
    # the same build through nestgpu.random objects held by Python variables
    import nestgpu
    k2 = nestgpu.Kernel(0)
    k2.SetKernelStatus({"total_num_virtual_procs": 3, "rng_seed": 19})
    g = k2.Create("iaf_psc_exp", 90)
    wo = nestgpu.random.normal(1.0, 0.5)
    do = nestgpu.math.max(nestgpu.random.normal(1.5, 0.75), 0.1)
    k2.Connect(g[0:30], g[0:30], {"rule": "pairwise_bernoulli", "p": 0.1}, {"weight": wo, "delay": do})
    k2.Connect(g[0:30], g[30:60], {"rule": "pairwise_bernoulli", "p": 0.13}, {"weight": wo, "delay": do})
    k2.Connect(g[0:30], g[60:90], {"rule": "fixed_indegree", "indegree": 3},
               {"weight": nestgpu.random.normal(1.0, 0.5), "delay": do})
    s3, t3, w3, d3 = _conns_gpu(k2)
    assert np.array_equal(s3, s2) and np.array_equal(w3, w2) and np.array_equal(d3, d2)


@pytest.mark.parametrize("n_vp, n", [(1, 20000), (3, 20000), (8, 70000)])
def test_connect_arrays_connectivity_identical_to_oracle(n_vp, n):
    # connect_arrays: every VP creates its targets' connections in array order
    # (connection_manager.cpp:624-760); a bad delay anywhere adds nothing. Arrays of 65,536
    # or more elements are filled by one thread per VP.
    N = 300
    rng = np.random.default_rng(n_vp)
    src = rng.integers(1, N + 1, size=n)
    tgt = rng.integers(1, N + 1, size=n)
    w = rng.normal(10.0, 5.0, size=n)
    d = np.round(rng.uniform(0.1, 4.0, size=n), 1)
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"total_num_virtual_procs": n_vp})
    k.Create("iaf_psc_exp", N)
    bad = d.copy()
    bad[n // 2] = 0.01
    with pytest.raises(nestgpu.NESTError, match="BadDelay"):
        k.ConnectArrays(src, tgt, w, bad)
    assert len(k.GetConnections()["source"]) == 0
    k.ConnectArrays(src, tgt, w, d)
    k.ConnectArrays(src[:50], tgt[:50])  # default weight 1 and delay 1 ms
    o = Oracle(n_vp=n_vp)
    o.create("iaf_psc_exp", N, {})
    steps = np.floor(d * 10 + 0.5).astype(np.int64)
    for vp in range(n_vp):
        sel = (tgt % n_vp) == vp
        o.add_connections(src[sel], tgt[sel], w[sel], steps[sel])
    for vp in range(n_vp):
        sel = (tgt[:50] % n_vp) == vp
        o.add_connections(src[:50][sel], tgt[:50][sel], np.ones(sel.sum()), np.full(sel.sum(), 10))
    s1, t1, w1, d1 = _conns_gpu(k)
    s2, t2, w2, d2 = o.connections()
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2)
    assert np.array_equal(w1, w2) and np.array_equal(d1, d2)


def _weighted_matrix(k, pre, post, field):
    c = k.GetConnections()
    vals = c["weight"] if field == "weight" else c["delay_steps"] * 0.1
    M = np.zeros((len(post), len(pre)))
    first_pre, first_post = pre.tolist()[0], post.tolist()[0]
    for s_, t_, v in zip(c["source"], c["target"], vals):
        if first_pre <= s_ < first_pre + len(pre) and first_post <= t_ < first_post + len(post):
            M[t_ - first_post, s_ - first_pre] += v
    return M


@pytest.mark.parametrize("n_vp", [1, 3])
def test_one_to_one_parameter_arrays(n_vp):
    # connect/test_connect_one_to_one.py:63-74: one value per pair
    N = 50
    for label, arr in [("weight", np.arange(N, dtype=float)), ("delay", np.arange(1, N + 1) * 0.1)]:
        k = nestgpu.Kernel(0)
        k.SetKernelStatus({"total_num_virtual_procs": n_vp})
        a = k.Create("iaf_psc_exp", N)
        b = k.Create("iaf_psc_exp", N)
        k.Connect(a, b, "one_to_one", {label: arr})
        np.testing.assert_allclose(_weighted_matrix(k, a, b, label), np.diag(arr), rtol=0, atol=1e-12)


@pytest.mark.parametrize("n_vp", [1, 4])
def test_all_to_all_parameter_arrays(n_vp):
    # connect/test_connect_all_to_all.py:56-71: [n_target, n_source] arrays
    N1, N2 = 500, 10
    for label, arr in [("weight", np.arange(N1 * N2, dtype=float).reshape(N2, N1)),
                       ("delay", np.arange(1, N1 * N2 + 1).reshape(N2, N1) * 0.1)]:
        k = nestgpu.Kernel(0)
        k.SetKernelStatus({"total_num_virtual_procs": n_vp})
        a = k.Create("iaf_psc_exp", N1)
        b = k.Create("iaf_psc_exp", N2)
        k.Connect(a, b, "all_to_all", {label: arr})
        np.testing.assert_allclose(_weighted_matrix(k, a, b, label), arr, rtol=0, atol=1e-9)


@pytest.mark.parametrize("n_vp", [1, 3])
def test_all_to_all_parameter_arrays_without_autapses(n_vp):
    # connect/test_connect_all_to_all.py:73-85: an autapse's value is passed over
    N = 20
    for label, arr in [("weight", np.arange(N * N, dtype=float).reshape(N, N)),
                       ("delay", np.arange(1, N * N + 1).reshape(N, N) * 0.1)]:
        k = nestgpu.Kernel(0)
        k.SetKernelStatus({"total_num_virtual_procs": n_vp})
        a = k.Create("iaf_psc_exp", N)
        k.Connect(a, a, {"rule": "all_to_all", "allow_autapses": False}, {label: arr})
        expected = arr.copy()
        np.fill_diagonal(expected, 0)
        np.testing.assert_allclose(_weighted_matrix(k, a, a, label), expected, rtol=0, atol=1e-9)


@pytest.mark.parametrize("n_vp", [1, 3])
def test_fixed_indegree_parameter_arrays(n_vp):
    # connect/test_connect_array_fixed_indegree.py: [n_target, indegree] arrays, each target
    # receives its row; delay = weight + 1 pairs the two arrays
    N, K = 20, 5
    Warr = [[y * K + x for x in range(K)] for y in range(N)]
    Darr = [[y * K + x + 1 for x in range(K)] for y in range(N)]
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"total_num_virtual_procs": n_vp})
    net1 = k.Create("iaf_psc_exp", N)
    net2 = k.Create("iaf_psc_exp", N)
    k.Connect(net1, net2, {"rule": "fixed_indegree", "indegree": K},
              {"synapse_model": "static_synapse", "weight": Warr, "delay": Darr})
    c = k.GetConnections()
    first = net2.tolist()[0]
    for i in range(N):
        sel = c["target"] == first + i
        w, d = c["weight"][sel], c["delay_steps"][sel] * 0.1
        np.testing.assert_allclose(d - w, 1.0, rtol=0, atol=1e-9)
        assert len(w) == K and sorted(w) == sorted(Warr[i])


def test_parameter_arrays_draw_nothing_from_the_streams():
    # an array value takes no random draw: sources and normal weights are the oracle's with a
    # constant delay, and every target's delays are its row of the array in draw order
    n_vp, N, K = 3, 60, 7
    darr = np.round(np.random.default_rng(4).uniform(0.1, 3.0, size=(N, K)), 1)
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"total_num_virtual_procs": n_vp, "rng_seed": 9})
    a = k.Create("iaf_psc_delta", N)
    k.Connect(a, a, {"rule": "fixed_indegree", "indegree": K},
              {"weight": nestgpu.random.normal(1.0, 0.5), "delay": darr})
    o = Oracle(n_vp=n_vp, seed=9)
    g = o.create("iaf_psc_delta", N, {})
    o.connect(g, g, "fixed_indegree", K, weight=("normal", 1.0, 0.5), delay=1.0)
    s1, t1, w1, d1 = _conns_gpu(k)
    s2, t2, w2, d2 = o.connections()
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2) and np.array_equal(w1, w2)
    first = a.tolist()[0]
    for i in range(N):
        np.testing.assert_array_equal(d1[t1 == first + i], np.floor(darr[i] * 10 + 0.5).astype(np.int64))


@pytest.mark.parametrize("bad, exc", [
    (lambda k, a: k.Connect(a, a, "one_to_one", {"weight": np.ones(4)}), ValueError),
    (lambda k, a: k.Connect(a, a, "all_to_all", {"weight": np.ones(5)}), TypeError),
    (lambda k, a: k.Connect(a, a, "all_to_all", {"weight": np.ones((5, 4))}), ValueError),
    (lambda k, a: k.Connect(a, a, {"rule": "pairwise_bernoulli", "p": 0.5}, {"weight": np.ones((5, 5))}), TypeError),
])
def test_parameter_array_shapes_checked(bad, exc):
    # hl_api_connection_helpers.py:48-130
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_exp", 5)
    with pytest.raises(exc):
        bad(k, a)
    assert len(k.GetConnections()["source"]) == 0


def test_get_connections_filters():
    # hl_api_connections.py GetConnections(source, target, synapse_model): filters keep order
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"total_num_virtual_procs": 3})
    a = k.Create("iaf_psc_exp", 30)
    k.Connect(a, a, {"rule": "fixed_indegree", "indegree": 4}, {"weight": nestgpu.random.normal(1.0, 0.2)})
    c = k.GetConnections()
    src, tgt = a[:10], a[5:20]
    f = k.GetConnections(source=src, target=tgt)
    keep = np.isin(c["source"], src.tolist()) & np.isin(c["target"], tgt.tolist())
    assert keep.any() and not keep.all()
    for key in c:
        np.testing.assert_array_equal(f[key], c[key][keep])
    assert np.array_equal(k.GetConnections(target=a[3])["target"], c["target"][c["target"] == a.tolist()[3]])
    assert len(k.GetConnections(synapse_model="static_synapse")["source"]) == len(c["source"])
    assert len(k.GetConnections(synapse_model="stdp_synapse")["source"]) == 0


def test_parameter_arithmetic_expressions():
    # hl_api_types.py Parameter: / and ** build quotient and pow; lhs / p is p**-1 * lhs
    p = nestgpu.random.uniform(1.0, 2.0)
    assert (p / 2.0).spec == {"quotient": [p.spec, 2.0]}
    assert (p ** 3).spec == {"pow": [p.spec, 3.0]}
    assert (2.0 / p).spec == {"product": [{"pow": [p.spec, -1.0]}, 2.0]}
    assert nestgpu.math.exp(p).spec == {"exp": p.spec}
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_exp", 4)
    k.Connect(a, a, "all_to_all", {"weight": nestgpu.math.cos(p) / 2.0, "delay": 1.0 + nestgpu.math.sin(p) ** 2})
    c = k.GetConnections()
    assert np.all(np.abs(c["weight"]) <= 0.5) and np.all(c["delay_steps"] >= 10)


def test_comparison_and_conditional_expressions():
    # hl_api_types.py comparisons (comparator 0..5) and hl_api_logic.py conditional
    p = nestgpu.random.uniform(0.0, 1.0)
    assert [(p < 0.5).spec["compare"][2], (p <= 0.5).spec["compare"][2], (p == 0.5).spec["compare"][2],
            (p != 0.5).spec["compare"][2], (p >= 0.5).spec["compare"][2], (p > 0.5).spec["compare"][2]] == [0, 1, 2, 3, 4, 5]
    c = nestgpu.logic.conditional(p < 0.5, 1.0, -1.0)
    assert c.spec == {"conditional": [(p < 0.5).spec, 1.0, -1.0]}
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_exp", 5)
    k.Connect(a, a, "all_to_all", {"weight": c})
    assert set(np.unique(k.GetConnections()["weight"])) <= {1.0, -1.0}
    with pytest.raises(nestgpu.NESTError, match="range 0-5"):
        k.Connect(a, a, "all_to_all", {"weight": nestgpu.Parameter({"compare": [p.spec, 0.5, 7]})})


def test_get_set_status_roundtrip_and_relative_voltages():
    k = nestgpu.Kernel(0)
    n = k.Create("iaf_psc_delta", 3, {"E_L": -65.0, "V_th": -50.0})
    # absolute V_m is kept when E_L changes without V_m (models/iaf_psc_delta.cpp:177-187)
    assert np.all(n.get("V_m") == -70.0)
    n.set(V_m=np.array([-60.0, -61.0, -62.0]))
    np.testing.assert_array_equal(n.get("V_m"), [-60.0, -61.0, -62.0])
    # voltages are stored relative to E_L but an E_L change alone keeps their absolute
    # values (models/iaf_psc_delta.cpp:109-142)
    np.testing.assert_array_equal(n.get("V_th"), -50.0)
    n.set({"E_L": -75.0})
    np.testing.assert_allclose(n.get("V_th"), -50.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(n.get("V_reset"), -70.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(n.get("V_m"), [-60.0, -61.0, -62.0], rtol=0, atol=1e-12)
    n.set({"E_L": -60.0, "V_th": -40.0})
    np.testing.assert_allclose(n.get("V_th"), -40.0, rtol=0, atol=1e-12)
    e = k.Create("aeif_cond_exp", 2, {"a": 2.0})
    np.testing.assert_array_equal(e.get("a"), [2.0, 2.0])
    np.testing.assert_array_equal(e.get("w"), [0.0, 0.0])


@pytest.mark.parametrize("bad, msg", [
    (lambda k: k.Create("no_such_model", 1), "UnknownModelName"),
    (lambda k: k.Create("iaf_psc_delta", 1, {"no_such_param": 1.0}), "UnaccessedDictionaryEntry"),
    (lambda k: k.Create("iaf_psc_delta", 1, {"C_m": -1.0}), "Capacitance must be > 0"),
    # messages and check order of each model's Parameters_::set
    (lambda k: k.Create("iaf_psc_exp", 1, {"C_m": 0.0, "tau_m": 0.0}), "Capacitance must be strictly positive"),
    (lambda k: k.Create("iaf_psc_exp", 1, {"tau_syn_in": 0.0, "t_ref": -1.0}), "time constants must be strictly"),
    (lambda k: k.Create("iaf_psc_delta", 1, {"t_ref": -1.0, "tau_m": 0.0}), "Refractory time must not be negative"),
    (lambda k: k.Create("iaf_psc_delta", 1, {"tau_m": 0.0}), "Membrane time constant must be > 0"),
    (lambda k: k.Create("aeif_cond_exp", 1, {"t_ref": -1.0}), "Refractory time cannot be negative"),
    (lambda k: k.Create("aeif_cond_exp", 1, {"C_m": 0.0}), "Ensure that C_m >0"),
    (lambda k: k.Create("iaf_psc_exp", 1, {"V_reset": -40.0, "V_th": -50.0}), "Reset potential"),
    (lambda k: k.Create("aeif_cond_exp", 1, {"V_peak": -60.0}), "V_peak"),
    # aeif_cond_exp.cpp:244-256 (exponential overflow guard) and :294-304 (state)
    (lambda k: k.Create("aeif_cond_exp", 1, {"V_peak": 1000.0, "Delta_T": 0.01}), "numerical overflow"),
    (lambda k: k.Create("aeif_cond_exp", 1, {"g_ex": -1.0}), "Conductances must not be negative"),
    (lambda k: k.SetKernelStatus({"rng_seed": 0}), "RNG seed"),
    (lambda k: k.SetKernelStatus({"no_such_key": 1}), "UnaccessedDictionaryEntry"),
])
def test_error_paths(bad, msg):
    k = nestgpu.Kernel(0)
    with pytest.raises(nestgpu.NESTError, match=msg):
        bad(k)


@pytest.mark.parametrize("case, msg", [
    # connect/test_connect_node_collection.py:55-69
    ("empty_pre", "IllegalConnection"),
    ("empty_post", "IllegalConnection"),
    ("empty_both", "IllegalConnection"),
    # connect/test_connect_fixed_indegree.py:52-62 (indegree above the population, no multapses)
    ("indegree_too_large", "BadProperty"),
    ("indegree_negative", "BadProperty"),
    ("p_above_one", "BadProperty"),
    ("one_to_one_sizes", "DimensionMismatch"),
    ("unknown_rule", "BadProperty"),
    ("unknown_conn_key", "UnaccessedDictionaryEntry"),
    ("unknown_synapse", "UnknownSynapseType"),
    ("generator_target", "IllegalConnection"),
    ("recorder_source", "IllegalConnection"),
    # iaf_psc_exp.h:466-474: receptor 0 only
    ("receptor_type", "UnknownReceptorType: Receptor type 1 is not available in iaf_psc_exp"),
    # connection_manager.cpp:470 (unread keys), conn_builder.cpp:1403,1861 (required keys)
    ("indegree_unread", "UnaccessedDictionaryEntry"),
    ("p_unread", "UnaccessedDictionaryEntry"),
    ("indegree_missing", "needs 'indegree'"),
    ("p_missing", "needs 'p'"),
])
def test_connect_error_paths(case, msg):
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_exp", 5)
    sg = k.Create("spike_generator", 5)
    sr = k.Create("spike_recorder")
    calls = {
        "empty_pre": lambda: k.Connect(a[0:0], a),
        "empty_post": lambda: k.Connect(a, a[0:0]),
        "empty_both": lambda: k.Connect(a[0:0], a[0:0]),
        "indegree_too_large": lambda: k.Connect(a, a, {"rule": "fixed_indegree", "indegree": 6,
                                                       "allow_autapses": True, "allow_multapses": False}),
        "indegree_negative": lambda: k.Connect(a, a, {"rule": "fixed_indegree", "indegree": -1}),
        "p_above_one": lambda: k.Connect(a, a, {"rule": "pairwise_bernoulli", "p": 1.5}),
        "one_to_one_sizes": lambda: k.Connect(a, a[:3], "one_to_one"),
        "unknown_rule": lambda: k.Connect(a, a, {"rule": "no_such_rule"}),
        "unknown_conn_key": lambda: k.Connect(a, a, {"rule": "all_to_all", "no_such_key": 1}),
        "unknown_synapse": lambda: k.Connect(a, a, "all_to_all", {"synapse_model": "no_such_synapse"}),
        "generator_target": lambda: k.Connect(a, sg, "one_to_one"),
        "recorder_source": lambda: k.Connect(sr, a),
        "receptor_type": lambda: k.Connect(a, a, "all_to_all", {"receptor_type": 1}),
        "indegree_unread": lambda: k.Connect(a, a, {"rule": "all_to_all", "indegree": 2}),
        "p_unread": lambda: k.Connect(a, a, {"rule": "fixed_indegree", "indegree": 2, "p": 0.5}),
        "indegree_missing": lambda: k.Connect(a, a, {"rule": "fixed_indegree"}),
        "p_missing": lambda: k.Connect(a, a, {"rule": "pairwise_bernoulli"}),
    }
    with pytest.raises(nestgpu.NESTError, match=msg):
        calls[case]()
    # a failed Connect adds nothing
    assert len(k.GetConnections()["source"]) == 0


def test_failed_set_changes_nothing():
    # Node::set_status sets temporaries and commits them only when every check passes
    k = nestgpu.Kernel(0)
    e = k.Create("aeif_cond_exp", 2)
    a0, g0 = e.get("a").copy(), e.get("g_ex").copy()
    with pytest.raises(nestgpu.NESTError, match="Conductances"):
        e.set({"a": 7.0, "g_ex": -1.0})
    np.testing.assert_array_equal(e.get("a"), a0)
    np.testing.assert_array_equal(e.get("g_ex"), g0)
    e.set({"Delta_T": 0.0})  # Delta_T = 0 is allowed: the model without the exponential term
    np.testing.assert_array_equal(e.get("Delta_T"), [0.0, 0.0])


def test_spike_recorder_window_parameters():
    # Device::Parameters_ (device.cpp:83-118): origin, start, stop on the time grid, stop >= start
    k = nestgpu.Kernel(0)
    sr = k.Create("spike_recorder", 1, {"start": 10.0, "stop": 50.0, "origin": 5.0})
    assert (sr.get("start"), sr.get("stop"), sr.get("origin")) == (10.0, 50.0, 5.0)
    assert sr.get("n_events") == 0
    sr.set({"stop": float("inf")})
    assert np.isinf(sr.get("stop"))
    d = k.Create("spike_recorder")
    assert d.get("start") == 0.0 and np.isinf(d.get("stop")) and d.get("origin") == 0.0
    for bad, msg in [({"start": 0.05}, "multiple of the simulation resolution"),
                     ({"start": 20.0, "stop": 10.0}, "stop >= start"),
                     ({"record_to": "ascii"}, "memory"),
                     ({"no_such_key": 1.0}, "UnaccessedDictionaryEntry")]:
        with pytest.raises(nestgpu.NESTError, match=msg):
            k.Create("spike_recorder", 1, bad)
    with pytest.raises(nestgpu.NESTError, match="stop >= start"):
        sr.set({"start": 100.0, "stop": 60.0})
    assert sr.get("start") == 10.0  # a failed set changes nothing


def test_spike_generator_window_parameters():
    # StimulationDevice window (device.cpp:83-118): checked like the recorder's, and a failed
    # set keeps the generator's spike times
    k = nestgpu.Kernel(0)
    g = k.Create("spike_generator", 1, {"spike_times": [1.0, 2.0], "origin": 1.0, "start": 0.5, "stop": 2.0})
    assert (g.get("origin"), g.get("start"), g.get("stop")) == (1.0, 0.5, 2.0)
    with pytest.raises(nestgpu.NESTError, match="stop >= start"):
        g.set({"spike_times": [5.0], "start": 3.0})
    with pytest.raises(nestgpu.NESTError, match="multiple of the simulation resolution"):
        g.set({"origin": -2.0})  # Time(ms) truncates toward zero: off the grid, as in the reference
    with pytest.raises(nestgpu.NESTError, match="sorted"):
        g.set({"spike_times": [3.0, 1.0]})
    assert g.get("start") == 0.5
    n = k.Create("iaf_psc_exp", 1)
    k.Connect(g, n, "one_to_one", {"weight": 1.0, "delay": 1.0})  # spike times kept: still valid


def test_create_with_per_node_parameters():
    # hl_api_nodes.py Create: a list of dicts or per-node arrays, applied after creation
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_exp", 3, [{"I_e": 1.0}, {"I_e": 2.0}, {"I_e": 3.0, "tau_m": 5.0}])
    np.testing.assert_array_equal(a.get("I_e"), [1.0, 2.0, 3.0])
    np.testing.assert_array_equal(a.get("tau_m"), [10.0, 10.0, 5.0])
    b = k.Create("iaf_psc_exp", 3, {"I_e": np.array([4.0, 5.0, 6.0]), "C_m": 200.0})
    np.testing.assert_array_equal(b.get("I_e"), [4.0, 5.0, 6.0])
    np.testing.assert_array_equal(b.get("C_m"), [200.0, 200.0, 200.0])
    c = k.Create("iaf_psc_delta", {"V_th": -50.0})  # a dict as second argument: one node
    assert len(c) == 1 and c.get("V_th") == -50.0
    with pytest.raises(TypeError, match="one dictionary per node"):
        k.Create("iaf_psc_exp", 2, [{"I_e": 1.0}])
    with pytest.raises(TypeError, match="dict of parameters"):
        k.Create("iaf_psc_exp", 2, 5.0)
    with pytest.raises(nestgpu.NESTError, match="must be a dictionary"):
        k._chk(nestgpu._lib.nestgpu_create(k._h, b"iaf_psc_exp", 1, b"[1, 2]", ctypes.byref(ctypes.c_int64(0))))


def test_node_placement_and_flags_readable():
    # nodes are placed on VP gid % n_vp (vp_manager_impl.h:43-52); refractory_input reads back
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"total_num_virtual_procs": 3})
    a = k.Create("iaf_psc_delta", 5, {"refractory_input": True})
    np.testing.assert_array_equal(a.get("vp"), np.array(a.tolist()) % 3)
    np.testing.assert_array_equal(a.get("thread"), a.get("vp"))
    np.testing.assert_array_equal(a.get("refractory_input"), 1.0)
    a.set({"refractory_input": False})
    np.testing.assert_array_equal(a.get("refractory_input"), 0.0)


def test_receptor_type_zero_accepted():
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_exp", 3)
    k.Connect(a, a, "one_to_one", {"receptor_type": 0, "weight": 2.0})
    k.Connect(a, a, "one_to_one", {"receptor_type": [0, 0, 0]})
    assert len(k.GetConnections()["source"]) == 6


def test_bad_delay_rejected():
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_delta", 2)
    with pytest.raises(nestgpu.NESTError, match="BadDelay"):
        k.Connect(a, a, "all_to_all", {"delay": 0.04})  # rounds to 0 steps at 0.1 ms


def test_delay_extrema_reported_from_connections():
    # connect/test_delay_check.py: min_delay/max_delay follow the delays connected so far
    k = nestgpu.Kernel(0)
    assert k.GetKernelStatus("min_delay") == 0.1 and k.GetKernelStatus("max_delay") == 0.1
    n1, n2 = k.Create("iaf_psc_delta", 2)
    k.Connect(n1, n2, syn_spec={"weight": 1.0, "delay": 2.0})
    assert k.GetKernelStatus("min_delay") == 2.0 and k.GetKernelStatus("max_delay") == 2.0
    k.Connect(n2, n1, syn_spec={"weight": 1.0, "delay": 6.0})
    assert k.GetKernelStatus("min_delay") == 2.0 and k.GetKernelStatus("max_delay") == 6.0
    k.Connect(n2, n1, syn_spec={"weight": 1.0, "delay": 1.0})
    assert k.GetKernelStatus("min_delay_steps") == 10 and k.GetKernelStatus("max_delay_steps") == 60


def test_biological_time_only_settable_to_zero():
    # simulation_manager.cpp:151-158
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"biological_time": 0.0})
    with pytest.raises(nestgpu.NESTError, match="can only be set to 0.0"):
        k.SetKernelStatus({"biological_time": 5.0})
    assert k.GetKernelStatus("biological_time") == 0.0


def test_get_kernel_status_keys():
    k = nestgpu.Kernel(0)
    assert k.GetKernelStatus(["resolution", "total_num_virtual_procs"]) == (0.1, 1)
    with pytest.raises(KeyError):
        k.GetKernelStatus("no_such_key")
    with pytest.raises(TypeError):
        k.GetKernelStatus(5)


def test_resolution_fixed_once_nodes_exist():
    k = nestgpu.Kernel(0)
    k.Create("iaf_psc_delta", 1)
    with pytest.raises(nestgpu.NESTError, match="time representation cannot be changed"):
        k.SetKernelStatus({"resolution": 0.01})


@pytest.mark.parametrize("res, msg", [(0.0005, "at least one tic|greater than or equal to one tic"),
                                      (0.1005, "multiple of the tic length"), (-0.1, "one tic")])
def test_resolution_checked_against_the_tic(res, msg):
    # simulation_manager.cpp:257-266: at least one tic (1 us) and a whole number of tics
    k = nestgpu.Kernel(0)
    with pytest.raises(nestgpu.NESTError, match=msg):
        k.SetKernelStatus({"resolution": res})
    assert k.GetKernelStatus("resolution") == 0.1
    for ok in (0.01, 0.25, 1.0):
        k.SetKernelStatus({"resolution": ok})
        assert k.GetKernelStatus("resolution") == ok


def test_reset_kernel_clears_network():
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_delta", 5)
    k.Connect(a, a, "all_to_all")
    assert len(k.GetConnections()["source"]) == 25
    k.ResetKernel()
    assert len(k.GetConnections()["source"]) == 0
    b = k.Create("iaf_psc_exp", 2)
    assert b.tolist() == [1, 2]


def test_poisson_counts_statistics_and_determinism():
    lam = 0.8
    c1 = nestgpu.poisson_counts(11, lam, 100, 500, 0, 400)
    c2 = nestgpu.poisson_counts(11, lam, 100, 500, 0, 400)
    assert c1.shape == (400, 500)
    assert np.array_equal(c1, c2)
    # stream/step addressing: a sub-window equals the same window of the full matrix
    sub = nestgpu.poisson_counts(11, lam, 250, 50, 120, 30)
    assert np.array_equal(sub, c1[120:150, 150:200])
    m, v = c1.mean(), c1.var()
    se = np.sqrt(lam / c1.size)
    assert abs(m - lam) < 5 * se
    assert abs(v - lam) < 0.05 * lam
    # frequency of k spikes follows the Poisson pmf
    from scipy import stats
    ks = np.arange(6)
    obs = np.array([(c1 == x).sum() for x in ks])
    exp = stats.poisson(lam).pmf(ks) * c1.size
    assert np.all(np.abs(obs - exp) < 6 * np.sqrt(exp) + 5)
    assert not np.array_equal(c1, nestgpu.poisson_counts(12, lam, 100, 500, 0, 400))


def test_layout_keys_validated_and_reported():
    # device-layout kernel-status keys (no counterpart in the reference): accepted values are
    # echoed back before a build, out-of-range values raise BadProperty
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"targets_per_block": 64, "delivery_mode": "window", "window_slots": 3,
                       "delivery_passes": 1, "workgroups_per_cu": 2})
    st = k.GetKernelStatus()
    assert st["targets_per_block"] == 64 and st["delivery_mode"] == "window"
    assert st["window_slots"] == 3 and st["workgroups_per_cu"] == 2
    for bad in ({"window_slots": -1}, {"targets_per_block": 4096}, {"delivery_mode": "ring"},
                {"delivery_passes": 9}, {"workgroups_per_cu": 9}):
        with pytest.raises(nestgpu.NESTError, match="BadProperty"):
            k.SetKernelStatus(bad)


# ---------------------------------------------------------------- multimeter (host checks)
def test_multimeter_parameters_and_checks():
    # models/multimeter.cpp:115-162 and universal_data_logger.h:280-295
    k = nestgpu.Kernel(0)
    mm = k.Create("multimeter", params={"interval": 0.5, "offset": 0.3, "record_from": ["V_m", "I_syn_ex"]})
    assert mm.get("record_from") == ["V_m", "I_syn_ex"]
    assert mm.get("interval") == pytest.approx(0.5) and mm.get("offset") == pytest.approx(0.3)
    assert mm.get("model") == "multimeter"
    with pytest.raises(nestgpu.NESTError, match="at least as long as the simulation resolution"):
        k.Create("multimeter", params={"interval": 0.05})
    with pytest.raises(nestgpu.NESTError, match="multiple of the simulation resolution"):
        k.Create("multimeter", params={"interval": 0.25})
    with pytest.raises(nestgpu.NESTError, match="offset for the sampling interval must be at least"):
        k.Create("multimeter", params={"offset": 0.01})
    exp = k.Create("iaf_psc_exp", 2)
    delta = k.Create("iaf_psc_delta", 1)
    with pytest.raises(nestgpu.NESTError, match="unknown recordable I_syn_ex"):
        k.Connect(mm, delta)
    k.Connect(mm, exp)
    with pytest.raises(nestgpu.NESTError, match="cannot be changed after the multimeter has been connected"):
        mm.set(interval=1.0)
    mm.set(start=2.0, stop=5.0)  # the window may change
    vm = k.Create("voltmeter")
    assert vm.get("record_from") == ["V_m"]
    k.Connect(vm, delta)
    with pytest.raises(nestgpu.NESTError, match="IllegalConnection"):
        k.Connect(exp, mm)


def test_connect_on_device_needs_counter_based_streams():
    # device draws need rng_type Philox_32; a call with random parameters draws on the host
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"connect_on_device": True})
    a = k.Create("iaf_psc_exp", 20)
    with pytest.raises(nestgpu.NESTError, match="Philox_32"):
        k.Connect(a, a, {"rule": "pairwise_bernoulli", "p": 0.5}, {"weight": 1.0, "delay": 1.0})
    assert len(k.GetConnections()["source"]) == 0
    k.SetKernelStatus({"rng_type": "Philox_32"})
    k.Connect(a, a, {"rule": "fixed_indegree", "indegree": 3}, {"weight": nestgpu.random.normal(0.0, 1.0)})
    k.Connect(a, a, {"rule": "fixed_indegree", "indegree": 3, "allow_multapses": False}, {"weight": 1.0})
    assert k.GetKernelStatus("device_connect_calls") == 0
    assert len(k.GetConnections()["source"]) == 120


def test_stdp_synapse_connect_checks():
    # stdp_synapse::set_status (stdp_synapse.h:303-314) and this path's limits
    k = nestgpu.Kernel(0)
    a = k.Create("iaf_psc_exp", 10, {"tau_minus": 25.0})
    assert np.all(a.get("tau_minus") == 25.0)
    with pytest.raises(nestgpu.NESTError, match="same sign"):
        k.Connect(a, a, "all_to_all", {"synapse_model": "stdp_synapse", "weight": -1.0, "Wmax": 50.0})
    with pytest.raises(nestgpu.NESTError, match="Kplus"):
        k.Connect(a, a, "all_to_all", {"synapse_model": "stdp_synapse", "Kplus": -1.0})
    with pytest.raises(nestgpu.NESTError, match="mu_plus"):
        k.Connect(a, a, "all_to_all", {"synapse_model": "stdp_synapse", "mu_plus": 0.5})
    with pytest.raises(nestgpu.NESTError, match="UnaccessedDictionaryEntry"):
        k.Connect(a, a, "all_to_all", {"weight": 1.0, "tau_plus": 10.0})
    with pytest.raises(nestgpu.NESTError, match="strictly positive"):
        a.set(tau_minus=0.0)
    assert len(k.GetConnections()["source"]) == 0
    k.Connect(a[:5], a[5:], "all_to_all", {"synapse_model": "stdp_synapse", "weight": 30.0, "Wmax": 60.0,
                                             "lambda": 0.02, "delay": 1.5})
    k.Connect(a[5:], a[:5], "all_to_all", {"weight": 2.0})
    c = k.GetConnections()
    assert len(c["source"]) == 50
    assert np.array_equal(np.sort(np.unique(c["weight"])), [2.0, 30.0])
    # the synapse model of each connection, and the filter on it
    assert np.array_equal(c["synapse_model"] == "stdp_synapse", c["weight"] == 30.0)
    s = k.GetConnections(synapse_model="stdp_synapse")
    assert len(s["source"]) == 25 and np.all(s["weight"] == 30.0) and np.all(np.isin(s["target"], a[5:].tolist()))
    assert len(k.GetConnections(synapse_model="static_synapse")["source"]) == 25
    # post-spike history capacity: a power of two in [2, 65536]
    assert k.GetKernelStatus("stdp_history_capacity") == 256
    for bad in (100, 1, 131072, 0.5):
        with pytest.raises(nestgpu.NESTError, match="power of two"):
            k.SetKernelStatus({"stdp_history_capacity": bad})
    k.SetKernelStatus({"stdp_history_capacity": 1024})
    assert k.GetKernelStatus("stdp_history_capacity") == 1024


def test_model_registry_status():
    # ModelManager::get_status (model_manager.cpp:116-131): the node and synapse models
    k = nestgpu.Kernel(0)
    st = k.GetKernelStatus()
    assert st["node_models"][:4] == ["iaf_psc_delta", "iaf_psc_exp", "aeif_cond_exp", "iaf_psc_alpha"]
    assert set(st["node_models"][4:]) == {"spike_generator", "spike_recorder", "multimeter", "voltmeter"}
    assert st["synapse_models"] == ["static_synapse", "stdp_synapse"]
    # every registered neuron model is created through the registry, with its own defaults
    # and keys
    for name, key, val in (("iaf_psc_delta", "V_min", -1.7976931348623157e308), ("iaf_psc_exp", "tau_syn_ex", 2.0),
                           ("aeif_cond_exp", "gsl_error_tol", 1e-6), ("iaf_psc_alpha", "V_min", -np.inf)):
        n = k.Create(name, 2)
        assert np.all(np.atleast_1d(n.get(key)) == val)
        assert np.all(np.atleast_1d(n.get("tau_minus")) == 20.0)
    with pytest.raises(nestgpu.NESTError, match="UnknownModelName|unknown"):
        k.Create("hh_psc_alpha", 1)


def test_iaf_psc_alpha_parameters():
    # iaf_psc_alpha::Parameters_::set and State_::set (iaf_psc_alpha.cpp:112-199): a change of
    # E_L keeps the absolute V_th, V_reset and V_m (stored relative to E_L), and the
    # reference's checks and messages
    k = nestgpu.Kernel(0)
    n = k.Create("iaf_psc_alpha", 2, {"V_m": -65.0, "tau_syn_ex": 0.5})
    assert np.all(n.get("V_th") == -55.0) and np.all(n.get("V_reset") == -70.0) and np.all(n.get("V_m") == -65.0)
    assert np.all(n.get("tau_syn_ex") == 0.5) and np.all(n.get("tau_syn_in") == 2.0)
    assert np.all(n.get("I_syn_ex") == 0.0) and np.all(n.get("I_syn_in") == 0.0)
    n.set(E_L=-60.0)
    assert np.all(n.get("V_th") == -55.0) and np.all(n.get("V_m") == -65.0) and np.all(n.get("V_min") == -np.inf)
    assert np.all(n.get("E_L") == -60.0)
    for bad, msg in (({"C_m": 0.0}, "Capacitance must be > 0"), ({"tau_m": -1.0}, "Membrane time constant"),
                     ({"tau_syn_in": 0.0}, "synaptic time constants"), ({"t_ref": -0.1}, "t_ref can't be negative"),
                     ({"V_reset": -40.0}, "Reset potential"), ({"g_ex": 1.0}, "UnaccessedDictionaryEntry")):
        with pytest.raises(nestgpu.NESTError, match=msg):
            n.set(**bad)
    assert np.all(n.get("C_m") == 250.0)
