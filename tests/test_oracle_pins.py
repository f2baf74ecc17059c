"""Pins of the CPU oracle (no GPU): reference golden values, closed forms, special cases.

Each check is chosen so that a plausible mistake in the restatement (a dropped term, a
wrong rounding rule, a swapped operand, a wrong seed word) fails it.
"""

import json
import os
import subprocess

import numpy as np
import pytest

from _oracle import ROOT, Oracle, lib

GOLDEN = os.path.join(ROOT, "tests", "golden")
PROBE = os.path.join(ROOT, "oracle", "_ref", "randutils_probe")


# ------------------------------------------------------------------ RNG seeding
def _seedseq(seed, n_words=8, n_draws=5):
    import ctypes

    L = lib()
    s = (ctypes.c_uint32 * len(seed))(*seed)
    w = (ctypes.c_uint32 * n_words)()
    L.or_seed_seq_generate(s, len(seed), w, n_words)
    d = (ctypes.c_uint64 * n_draws)()
    L.or_mt64_first(s, len(seed), d, n_draws)
    return list(w), list(d)


def test_seed_sequence_matches_reference_vectors():
    # vectors produced by the reference's own randutils.hpp (tests/golden/make_seedseq_vectors.py)
    with open(os.path.join(GOLDEN, "seedseq_fe128.json")) as f:
        cases = json.load(f)
    assert len(cases) >= 5
    for c in cases:
        words, draws = _seedseq(c["seed"])
        assert words == c["words"], c["seed"]
        assert draws == c["mt19937_64"], c["seed"]


@pytest.mark.skipif(not os.path.exists(PROBE), reason="reference probe not built (oracle/ref)")
def test_seed_sequence_matches_reference_probe_live():
    rng = np.random.default_rng(3)
    for _ in range(20):
        seed = [int(x) for x in rng.integers(0, 2 ** 32, size=rng.integers(1, 6))]
        txt = subprocess.run([PROBE, "8", "5"] + [str(x) for x in seed], check=True, capture_output=True,
                             text=True).stdout.split()
        words, draws = _seedseq(seed)
        assert words == [int(x) for x in txt[:8]]
        assert draws == [int(x) for x in txt[8:]]


# ------------------------------------------------------------------ iaf_psc_delta
def test_compare_delta_spike_at_4_1_ms():
    # testsuite/pytests/recording/test_compare_delta.py: first spike at exactly 4.1 ms
    o = Oracle(resolution=0.01)
    n = o.create("iaf_psc_delta", 1, {"E_L": -49.0, "V_m": -60.0, "V_th": -50.0, "V_reset": -60.0, "C_m": 200.0,
                                       "tau_m": 20.0, "t_ref": 5.0})
    counts = np.zeros((1201, 1), np.uint8)
    for t in (1.0, 2.0, 3.0, 4.0, 5.0, 10.5, 12.0):
        counts[int(round(t / 0.01)) - 1, 0] = 1
    o.add_drive(n[0], counts, 2.5, 0.1)
    o.simulate(200.0)
    g, s = o.spikes()
    assert len(s) > 0
    assert s[0] * 0.01 == pytest.approx(4.1, abs=1e-12)


def test_constant_potential_without_input():
    # testsuite/pytests/regressions/test_ticket_421.py: V_m stays at E_L without input
    for model in ("iaf_psc_delta", "iaf_psc_exp"):
        o = Oracle()
        n = o.create(model, 3, {"E_L": -62.0})
        o.set_values(n, "V_m", [-62.0] * 3)
        for _ in range(5):
            o.simulate(10.0)
            assert np.all(o.get(n, "V_m") == -62.0)


def test_delta_refractory_and_reset():
    # one strong input spike: crosses threshold, resets, stays refractory t_ref = 2 ms
    o = Oracle()
    n = o.create("iaf_psc_delta", 1, {"E_L": 0.0, "V_th": 20.0, "V_reset": 10.0, "t_ref": 2.0, "tau_m": 20.0,
                                       "C_m": 1.0})
    o.set_values(n, "V_m", [0.0])
    counts = np.zeros((200, 1), np.uint8)
    counts[9, 0] = 1   # emission step 9 -> arrives at step 9 + 15, spike stamp 25
    counts[20, 0] = 1  # arrives at step 35, within the 20 refractory steps: discarded
    counts[40, 0] = 1  # arrives at step 55 after refractoriness: V_reset decayed + 25 > V_th
    o.add_drive(n[0], counts, 25.0, 1.5)
    o.simulate(15.0)
    g, s = o.spikes()
    assert list(s) == [25, 56]


def test_delta_refractory_input_closed_form():
    # refractory_input (models/iaf_psc_delta.cpp:284-299): input arriving while refractory is
    # scaled by exp(-r h / tau_m), r the refractory steps left at its arrival, and added at the
    # first active step. The first input fires the neuron at step 24 (r = 20 from step 25 on);
    # the second arrives at step 35 with r = 10 and joins V at step 45 without firing it.
    o = Oracle()
    n = o.create("iaf_psc_delta", 1, {"E_L": 0.0, "V_th": 40.0, "V_reset": 10.0, "t_ref": 2.0, "tau_m": 20.0,
                                       "C_m": 1.0, "refractory_input": True})
    o.set_values(n, "V_m", [20.0])
    counts = np.zeros((200, 1), np.uint8)
    counts[9, 0] = 1   # arrives at step 24: 20 e^(-0.005 * 24) + 25 > 40
    counts[20, 0] = 1  # arrives at step 35, refractory
    o.add_drive(n[0], counts, 25.0, 1.5)
    o.simulate(4.6)    # through the update of step 45
    g, s = o.spikes()
    assert list(s) == [25]
    expected = 10.0 * np.exp(-0.1 / 20.0) + 25.0 * np.exp(-10 * 0.1 / 20.0)
    assert abs(o.get(n, "V_m")[0] - expected) < 1e-12
    # without refractory_input the refractory input is lost
    o2 = Oracle()
    n2 = o2.create("iaf_psc_delta", 1, {"E_L": 0.0, "V_th": 40.0, "V_reset": 10.0, "t_ref": 2.0, "tau_m": 20.0,
                                         "C_m": 1.0})
    o2.set_values(n2, "V_m", [20.0])
    o2.add_drive(n2[0], counts, 25.0, 1.5)
    o2.simulate(4.6)
    assert abs(o2.get(n2, "V_m")[0] - 10.0 * np.exp(-0.1 / 20.0)) < 1e-12


# ------------------------------------------------------------------ iaf_psc_exp
def test_iaf_psc_exp_dc_closed_form():
    # testsuite/pytests/neurons/test_iaf_psc_exp.py: DC response follows
    # V(t) = E_L + (1 - exp(-t / tau_m)) R I, reset and refractoriness at threshold
    dt, I, T = 0.1, 1000.0, 8.0
    o = Oracle(resolution=dt)
    n = o.create("iaf_psc_exp", 1, {"I_e": I})
    tau_m, C, E_L, V_th, V_reset, t_ref = 10.0, 250.0, -70.0, -55.0, -70.0, 2.0
    R = tau_m / C
    steps = int(round(T / dt))
    vm = np.empty(steps)
    for i in range(steps):
        o.simulate(dt)
        vm[i] = o.get(n, "V_m")[0]
    t = dt * np.arange(1, steps + 1)
    soln = E_L + (1.0 - np.exp(-t / tau_m)) * R * I
    expected = soln.copy()
    cross = int(np.argmax(soln >= V_th))  # first step at which V reaches threshold
    n_ref = int(round(t_ref / dt))
    expected[cross:cross + n_ref + 1] = V_reset
    rest = len(t) - (cross + n_ref + 1)
    expected[cross + n_ref + 1:] = soln[:rest]
    np.testing.assert_allclose(vm, expected, rtol=0, atol=1e-9)
    g, s = o.spikes()
    assert list(s) == [cross + 1]


def test_propagator_singular_limit():
    # testsuite/pytests/neurons/test_iaf_singularity.py / libnestutil/iaf_propagator.cpp:57-74
    L = lib()
    h, C, tm = 0.1, 250.0, 10.0
    p_sing = L.or_iaf_propagator(tm, tm, C, h)
    assert p_sing == pytest.approx(h / C * np.exp(-h / tm), rel=1e-15)
    for eps in (1e-6, 1e-9, 1e-12):
        p = L.or_iaf_propagator(tm + eps, tm, C, h)
        assert p == pytest.approx(p_sing, rel=1e-5)
    # regular case against the closed form of the exact-integration matrix element
    ts = 2.0
    exact = (ts * tm / (tm - ts) / C) * (np.exp(-h / tm) - np.exp(-h / ts))
    assert L.or_iaf_propagator(ts, tm, C, h) == pytest.approx(exact, rel=1e-12)


# ------------------------------------------------------------------ aeif_cond_exp
AEIF_LSODAR_PARAMS = {"V_reset": -58.0, "V_peak": 0.0, "V_th": -50.0, "I_e": 420.0, "g_L": 11.0, "tau_w": 300.0,
                      "E_L": -70.0, "Delta_T": 2.0, "a": 3.0, "b": 0.0, "C_m": 200.0, "V_m": -70.0, "w": 5.0}


@pytest.mark.parametrize("portable", [False, True])
def test_aeif_against_lsodar_reference(portable):
    # testsuite/pytests/neurons/test_aeif_lsodar.py: average relative deviation from the
    # LSODAR solution held in the reference tree (V_m < 5e-4, w < 1e-4), resolution 0.01 ms
    data = np.loadtxt(os.path.join(GOLDEN, "aeif_lsodar_trajectory.dat")).T
    o = Oracle(resolution=0.01, portable_math=portable)
    n = o.create("aeif_cond_exp", 1, AEIF_LSODAR_PARAMS)
    steps = 10000
    V = np.empty(steps)
    W = np.empty(steps)
    for i in range(steps):
        o.simulate(0.01)
        V[i] = o.get(n, "V_m")[0]
        W[i] = o.get(n, "w")[0]
    t = 0.01 * np.arange(1, steps + 1)
    keep = (t >= data[0, 0]) & (t <= data[0, -1])
    t, V, W = t[keep], V[keep], W[keep]
    refV = np.interp(t, data[0], data[1])
    refW = np.interp(t, data[0], data[2])
    V_lim = (AEIF_LSODAR_PARAMS["V_th"] + AEIF_LSODAR_PARAMS["V_peak"]) / 2.0
    rV = np.abs(refV - V) / np.abs(refV)
    rV = rV[V <= V_lim]
    rW = np.abs(refW - W) / np.abs(refW)
    assert np.average(rV) < 5e-4
    assert np.average(rW) < 1e-4


def test_aeif_equilibrium_exact():
    # Delta_T = 0, a = b = 0, no input: V = E_L and w = 0 are a fixed point of the dynamics
    o = Oracle()
    n = o.create("aeif_cond_exp", 2, {"Delta_T": 0.0, "a": 0.0, "b": 0.0})
    for _ in range(3):
        o.simulate(50.0)
        assert np.all(o.get(n, "V_m") == -70.6)
        assert np.all(o.get(n, "w") == 0.0)


def test_aeif_spike_adds_b_and_resets():
    o = Oracle()
    n = o.create("aeif_cond_exp", 1, {"I_e": 1000.0})
    o.simulate(100.0)
    g, s = o.spikes()
    assert len(s) >= 2
    # w decays between spikes and jumps by b at each: w stays positive and bounded
    assert 0 < o.get(n, "w")[0] < 80.5 * len(s)
    assert o.get(n, "V_m")[0] <= 0.0


# ------------------------------------------------------------------ portable exp/log
def test_portable_exp_log_within_one_ulp():
    L = lib()
    xs = np.concatenate([np.linspace(-700, 700, 2001), np.linspace(-1, 1, 2001), [0.0, 1e-300, -1e-300, 0.5 * np.log(2)]])
    for x in xs:
        y = L.or_fd_exp(float(x))
        ref = np.exp(x)
        assert abs(y - ref) <= np.spacing(ref) * 1.0 + 1e-300
    for x in np.concatenate([np.logspace(-300, 300, 601), np.linspace(0.1, 10, 1000)]):
        y = L.or_fd_log(float(x))
        ref = np.log(x)
        assert abs(y - ref) <= max(np.spacing(abs(ref)), 1e-300) * 1.0


# ------------------------------------------------------------------ invariants
def test_thread_count_does_not_change_results():
    import netbuild as nb

    desc = nb.exp_random(N=600, KE=60, KI=15, T=50.0, n_vp=3, seed=8)
    res = []
    for th in (1, 4):
        o, pops = nb.build_oracle(desc, threads=th)
        o.simulate(desc["T"])
        res.append((o.spikes(), o.get(pops[0], "V_m")))
    assert np.array_equal(res[0][0][0], res[1][0][0])
    assert np.array_equal(res[0][0][1], res[1][0][1])
    assert np.array_equal(res[0][1], res[1][1])


def test_fixed_indegree_degrees_and_uniform_sources():
    # connect/test_connect_fixed_indegree.py: every target has exactly K inputs; sources
    # are uniform (chi-square as in connect_test_base.py)
    from scipy import stats

    o = Oracle(n_vp=2, seed=99)
    pre = o.create("iaf_psc_delta", 200)
    post = o.create("iaf_psc_delta", 100)
    o.connect(pre, post, "fixed_indegree", 50)
    s, t, w, d = o.connections()
    assert np.all(np.bincount(t - post[0], minlength=100) == 50)
    counts = np.bincount(s - pre[0], minlength=200)
    chi2, p = stats.chisquare(counts)
    assert p > 1e-4


def test_bernoulli_degree_distribution():
    # connect/test_connect_pairwise_bernoulli.py: in-degrees are Binomial(n_pre, p)
    o = Oracle(n_vp=3, seed=5)
    pre = o.create("iaf_psc_delta", 400)
    post = o.create("iaf_psc_delta", 300)
    p = 0.1
    o.connect(pre, post, "pairwise_bernoulli", p)
    s, t, w, d = o.connections()
    deg = np.bincount(t - post[0], minlength=300)
    assert abs(deg.mean() - 400 * p) < 4 * np.sqrt(400 * p * (1 - p) / 300)
    assert abs(deg.var() - 400 * p * (1 - p)) < 0.35 * 400 * p * (1 - p)
    assert len(np.unique(np.stack([s, t]), axis=1)[0]) == len(s)  # no multapses


def test_delay_rounding_and_min_delay():
    # nest_time.h:539-542 delay_ms_to_steps = ld_round(ms * 10); min_delay over all delays
    o = Oracle()
    a = o.create("iaf_psc_delta", 2)
    o.add_connections([a[0]], [a[1]], [1.0], [3])
    o.connect([a[1]], [a[0]], "one_to_one", weight=1.0, delay=0.25)  # 2.5 steps -> 3 (round half up)
    o.connect([a[0]], [a[0]], "one_to_one", weight=1.0, delay=0.149999)  # -> 1
    s, t, w, d = o.connections()
    assert sorted(d) == [1, 3, 3]
    o.simulate(1.0)
    st = o.status()
    assert st["min_delay"] == 1 and st["max_delay"] == 3


def _weights(weight, n_vp=2, seed=17, N=200, K=50):
    o = Oracle(n_vp=n_vp, seed=seed)
    g = o.create("iaf_psc_delta", N)
    o.connect(g, g, "fixed_indegree", K, weight=weight, delay=1.0)
    s, t, w, d = o.connections()
    return s, t, w


def test_exponential_parameter_maps_one_uniform_draw():
    # ExponentialParameter::value (parameter.h:339-342): beta * -log(1 - u), one uniform draw
    # u per value: the same stream drawn as uniform(0, 1) gives the same sources and u
    import math
    beta = 2.5
    s1, t1, u = _weights(("uniform", 0.0, 1.0))
    s2, t2, w = _weights(("exponential", beta))
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2)
    expected = np.array([beta * -math.log(1 - x) for x in u])
    np.testing.assert_allclose(w, expected, rtol=1e-15, atol=0)
    assert abs(w.mean() / beta - 1) < 0.02  # 20,000 draws: mean beta within 4 sigma


def test_lognormal_parameter_is_exp_of_a_normal_draw():
    # std::lognormal_distribution: exp(s * z + m), z from its own N(0, 1) polar-method draws,
    # the same consumption as a normal(0, 1) parameter on the same stream
    import math
    m, sd = 0.3, 0.4
    s1, t1, z = _weights(("normal", 0.0, 1.0))
    s2, t2, w = _weights(("lognormal", m, sd))
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2)
    np.testing.assert_allclose(w, [math.exp(sd * x + m) for x in z], rtol=1e-15, atol=0)
    assert abs(w.mean() / math.exp(m + sd * sd / 2) - 1) < 0.01
    # std = 0 draws nothing and gives exp(mean) (parameter.cpp:149-152)
    s3, t3, w3 = _weights(("lognormal", m, 0.0))
    assert np.all(w3 == math.exp(m))


@pytest.mark.parametrize("op, f", [
    ("exp", lambda z: __import__("math").exp(z)), ("sin", lambda z: __import__("math").sin(z)),
    ("cos", lambda z: __import__("math").cos(z)), ("pow", lambda z: z ** 3.0), ("quotient", lambda z: z / 4.0),
])
def test_composite_parameters_apply_to_the_operand_draws(op, f):
    # ExpParameter, SinParameter, CosParameter, PowParameter, QuotientParameter
    # (parameter.h:546,991,1037,1082,1130): the function of the operand's value, no extra draw
    s1, t1, z = _weights(("normal", 0.0, 1.0))
    spec = {"pow": ("pow", ("normal", 0.0, 1.0), 3.0), "quotient": ("quotient", ("normal", 0.0, 1.0), 4.0)}.get(
        op, (op, ("normal", 0.0, 1.0)))
    s2, t2, w = _weights(spec)
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2)
    np.testing.assert_allclose(w, [f(x) for x in z], rtol=1e-15, atol=0)


def test_comparison_gives_one_or_zero_of_the_operand_draws():
    # ComparingParameter (parameter.h:714, 734-752): 1.0 where u < 0.4 for the same stream's u
    s1, t1, u = _weights(("uniform", 0.0, 1.0))
    for c, f in enumerate([np.less, np.less_equal, np.equal, np.not_equal, np.greater_equal, np.greater]):
        s2, t2, w = _weights(("compare", ("uniform", 0.0, 1.0), 0.4, c))
        assert np.array_equal(s1, s2) and np.array_equal(w, f(u, 0.4).astype(float))


def test_conditional_draws_only_the_chosen_branch():
    # ConditionalParameter::value (parameter.h:791-797): a branch not taken draws nothing, so a
    # constant condition gives exactly the chosen branch's stream
    s_n, t_n, w_n = _weights(("normal", 1.0, 0.5))
    s_e, t_e, w_e = _weights(("exponential", 2.0))
    s_t, t_t, w_t = _weights(("conditional", ("compare", 1.0, 0.0, 5), ("normal", 1.0, 0.5), ("exponential", 2.0)))
    s_f, t_f, w_f = _weights(("conditional", ("compare", 1.0, 0.0, 0), ("normal", 1.0, 0.5), ("exponential", 2.0)))
    assert np.array_equal(s_t, s_n) and np.array_equal(w_t, w_n)
    assert np.array_equal(s_f, s_e) and np.array_equal(w_f, w_e)
