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


# ------------------------------------------------------------------ draw order (conn_builder.cpp)
# NEST's per-VP engines are std::mt19937_64 seeded with seed_seq_fe128{rng_seed, 0xb84c9bae, vp}
# (random_manager.cpp:63-100). mt19937_64_streams.json holds the first outputs of such engines,
# written by the reference's own randutils.hpp (tests/golden/make_seedseq_vectors.py). The
# distributions map engine outputs by libstdc++'s published algorithms (GCC 11):
#   uniform_real / drand: generate_canonical<double, 53> with one 64-bit draw, double(x) / 2^64,
#   values >= 1 replaced by the largest double below 1 (bits/random.tcc generate_canonical);
#   uniform_int over [0, n): Lemire's nearly divisionless method with 128-bit products
#   (bits/uniform_int_dist.h _S_nd), which rejects a draw whose low word is below 2^64 mod n.
class _Stream:
    def __init__(self, seed):
        with open(os.path.join(GOLDEN, "mt19937_64_streams.json")) as f:
            self.x = next(c["mt19937_64"] for c in json.load(f) if c["seed"] == seed)
        self.i = 0

    def raw(self):
        v = self.x[self.i]
        self.i += 1
        return v

    def canonical(self):
        u = float(self.raw()) / 18446744073709551616.0
        return u if u < 1.0 else np.nextafter(1.0, 0.0)

    def ulrand(self, n):
        threshold = (1 << 64) % n
        while True:
            prod = self.raw() * n
            if (prod & ((1 << 64) - 1)) >= threshold:
                return prod >> 64


SEED_VP0 = [143202461, 0xb84c9bae, 0]
SEED_VP1 = [143202461, 0xb84c9bae, 1]


def _delay_steps(ms):
    # Time::delay_ms_to_steps (nest_time.h:539-542): ld_round(ms * 10) = floor(x + 0.5)
    return int(np.floor(ms * 10.0 + 0.5))


def test_fixed_indegree_draw_order_from_reference_engine_outputs():
    # FixedInDegreeBuilder::inner_connect_ (conn_builder.cpp:1526-1545): per target in order,
    # per connection: the source ulrand(n_sources), then single_connect_ (:419-420): the delay,
    # then the weight. Fails for: delay and weight drawn in the other order; all K sources drawn
    # before their weights; a source mapping other than Lemire's; uniform(min, max) computed as
    # max - u * range; targets of other VPs drawing from this VP's stream.
    for n_vp in (1, 2):
        o = Oracle(n_vp=n_vp, seed=143202461)
        g = o.create("iaf_psc_delta", 12)
        pre, post = g[:9], g[9:]
        o.connect(pre, post, "fixed_indegree", 4, weight=("uniform", 0.5, 1.5), delay=("uniform", 1.0, 3.0))
        s, t, w, d = o.connections()
        streams = {0: _Stream(SEED_VP0), 1: _Stream(SEED_VP1)}
        exp = {}
        for tgt in post:
            vp = int(tgt % n_vp)
            st = streams[vp]
            for _ in range(4):
                src = pre[st.ulrand(len(pre))]
                dms = 1.0 + st.canonical() * 2.0
                wv = 0.5 + st.canonical() * 1.0
                exp.setdefault(vp, []).append((src, tgt, wv, _delay_steps(dms)))
        got = list(zip(s.tolist(), t.tolist(), w.tolist(), d.tolist()))
        # GetConnections order: per VP in creation order
        want = [c for vp in sorted(exp) for c in exp[vp]]
        assert sorted(got) == sorted(want)
        for vp in exp:  # creation order within each VP
            mine = [c for c in got if c[1] % n_vp == vp]
            assert mine == exp[vp]


def test_pairwise_bernoulli_draw_order_from_reference_engine_outputs():
    # BernoulliBuilder::inner_connect_ (conn_builder.cpp:1937-1965): per target, per source in
    # order, one drand(); a connection when it is below p, whose delay then weight are drawn
    # at once (single_connect_). Fails for: the weight drawn before the Bernoulli trial of the
    # next source; sources looped outside targets; drand() <= p instead of < p (golden draws
    # are never equal to p, so that mistake shows only as a changed comparison on equal
    # values: the test checks the strict one on the draws it has).
    o = Oracle(n_vp=1, seed=143202461)
    g = o.create("iaf_psc_delta", 10)
    pre, post = g[:6], g[6:]
    o.connect(pre, post, "pairwise_bernoulli", 0.5, weight=("normal", 0.0, 1.0), delay=("uniform", 0.5, 2.5))
    s, t, w, d = o.connections()
    st = _Stream(SEED_VP0)
    want = []
    saved = None
    for tgt in post:
        for src in pre:
            if st.canonical() >= 0.5:
                continue
            dms = 0.5 + st.canonical() * 2.0
            # std::normal_distribution: Marsaglia's polar method, the second value cached
            if saved is not None:
                z, saved = saved, None
            else:
                while True:
                    x = 2.0 * st.canonical() - 1.0
                    y = 2.0 * st.canonical() - 1.0
                    r2 = x * x + y * y
                    if r2 <= 1.0 and r2 != 0.0:
                        break
                import math
                mult = math.sqrt(-2.0 * math.log(r2) / r2)
                saved = x * mult
                z = y * mult
            want.append((src, tgt, z * 1.0 + 0.0, _delay_steps(dms)))
    assert list(zip(s.tolist(), t.tolist(), w.tolist(), d.tolist())) == want
    assert len(want) > 5


def test_compare_draws_right_operand_first():
    # ComparingParameter::value (parameter.h:712-715) evaluates both operands as the arguments
    # of one call; g++ on x86-64 evaluates them right to left, so compare(uniform, uniform, <)
    # on consecutive draws u0, u1 of the stream gives u1 < u0. Fails for left-first evaluation.
    o = Oracle(n_vp=1, seed=143202461)
    g = o.create("iaf_psc_delta", 40)
    o.connect(g[:20], g[20:], "one_to_one", weight=("compare", ("uniform", 0.0, 1.0), ("uniform", 0.0, 1.0), 0),
              delay=1.0)
    s, t, w, d = o.connections()
    st = _Stream(SEED_VP0)
    want = []
    for _ in range(20):
        right = st.canonical()
        left = st.canonical()
        want.append(1.0 if left < right else 0.0)
    assert w.tolist() == want
    assert 0 < sum(want) < 20


# ------------------------------------------------------------------ spike arrival (a1, a7-a10)
def _relay_net(n_vp, n_pre, n_post, rule, d_steps, min_delay, max_delay, w=1.0, e_drive=4):
    """iaf_psc_delta relays restating other/test_spike_transmission.py: the pre neurons fire at
    stamp e_drive + 1 + min_delay (a drive event emitted at step e_drive with delay min_delay),
    their spikes reach the post neurons with delay d_steps. Two idle neurons connected with
    delays min_delay and max_delay fix the delay extrema."""
    o = Oracle(n_vp=n_vp)
    relay = {"E_L": 0.0, "V_th": 20.0, "V_reset": 0.0, "tau_m": 1e6, "C_m": 1.0, "t_ref": 2.0, "V_m": 0.0}
    pre = o.create("iaf_psc_delta", n_pre, relay)
    post = o.create("iaf_psc_delta", n_post, dict(relay, V_th=1e9))
    idle = o.create("iaf_psc_delta", 2, relay)
    o.add_connections([idle[0]], [idle[1]], [0.0], [min_delay])
    o.add_connections([idle[1]], [idle[0]], [0.0], [max_delay])
    if rule == "one_to_one":
        o.add_connections(pre, post, [w] * n_pre, [d_steps] * n_pre)
    else:
        src = np.repeat(pre, n_post)
        tgt = np.tile(post, n_pre)
        o.add_connections(src, tgt, [w] * len(src), [d_steps] * len(src))
    counts = np.zeros((e_drive + 1, n_pre), np.uint8)
    counts[e_drive, :] = 1
    o.add_drive(pre[0], counts, 25.0, min_delay * 0.1)
    return o, pre, post


@pytest.mark.parametrize("n_vp", [1, 2, 3, 4])
@pytest.mark.parametrize("rule", ["one_to_one", "all_to_all"])
@pytest.mark.parametrize("min_delay, max_delay, d", [(1, 1, 1), (1, 12, 10), (3, 3, 3), (3, 9, 7), (5, 20, 12),
                                                       (10, 10, 10)])
def test_spike_arrival_time(n_vp, rule, min_delay, max_delay, d):
    # other/test_spike_transmission.py: a spike of stamp S over a connection of delay d is
    # read in the update of step S + d - 1 (slot lag + d - min_delay of the delivering slice,
    # event.h:1006-1013), so V_m of the post neuron rises by the weight between times
    # (S + d - 1) h and (S + d) h and by exactly n_pre weights for all_to_all. Fails for an
    # off-by-one in the slot, a delivery in the wrong slice, or a lost or doubled event.
    n = 4 if rule == "one_to_one" else 3
    o, pre, post = _relay_net(n_vp, n, n, rule, d, min_delay, max_delay)
    S = 4 + 1 + min_delay  # pre spike stamp
    steps_before = S + d - 1
    o.run_steps(steps_before)
    g, s = o.spikes()
    assert sorted(g.tolist()) == sorted(pre.tolist()) and np.all(s == S)
    assert np.all(o.get(post, "V_m") == 0.0)
    o.run_steps(1)
    assert np.all(o.get(post, "V_m") == (1.0 if rule == "one_to_one" else float(n)))
    st = o.status()
    assert st["min_delay"] == min_delay and st["max_delay"] == max_delay


# ------------------------------------------------------------------ split Simulate (a1)
@pytest.mark.parametrize("pieces", [[3.0], [1.0, 1.0, 1.0], [0.3, 0.2, 1.7, 0.8], [0.1] * 30])
def test_split_simulate_equals_single_run(pieces):
    # other/test_stepped_simulation.py: Simulate split into several calls gives the same spikes
    # as one call; here also the same V_m, on a recurrent network whose min_delay (1.5 ms) is
    # cut by the splits. Fails for a split that drops or repeats a partial slice, or moves the
    # delivery of a slice split across calls.
    import netbuild as nb

    desc = nb.brunel_delta(N=200, K=20, T=3.0, n_vp=2, seed=31)
    ref, pr = nb.build_oracle(desc)
    ref.simulate(sum(pieces))
    o, po = nb.build_oracle(desc)
    for p in pieces:
        o.simulate(p)
    g1, s1 = ref.spikes()
    g2, s2 = o.spikes()
    assert sorted(zip(s1.tolist(), g1.tolist())) == sorted(zip(s2.tolist(), g2.tolist()))
    assert len(s1) > 0
    assert np.array_equal(ref.get(pr[0], "V_m"), o.get(po[0], "V_m"))


# ------------------------------------------------------------------ iaf_psc_exp synaptic current (a4)
@pytest.mark.parametrize("w", [300.0, -300.0])
def test_iaf_psc_exp_single_psc_closed_form(w):
    # One input of weight w (SYN_EX for w > 0, SYN_IN for w < 0, iaf_psc_exp.h:372-380) read in
    # the update of step k: the current is added after that step's V update
    # (iaf_psc_exp.cpp:313-330), so n steps later I_syn = w exp(-n h / tau_s) and
    # V_m - E_L = w / C_m * tau_s tau_m / (tau_m - tau_s) * (exp(-t/tau_m) - exp(-t/tau_s)),
    # t = n h, the exact solution. Fails for the channels swapped, the input added before the
    # decay, or the current reaching V_m in the step it arrives.
    tau_m, tau_s, C, E_L, h = 10.0, 2.0, 250.0, -70.0, 0.1
    o = Oracle()
    n = o.create("iaf_psc_exp", 1, {"tau_m": tau_m, "tau_syn_ex": tau_s, "tau_syn_in": tau_s, "C_m": C,
                                     "E_L": E_L, "V_m": E_L, "V_th": 1e9})
    counts = np.zeros((10, 1), np.uint8)
    counts[4, 0] = 1  # emission step 4, delay 1 step: read in the update of step 5
    o.add_drive(n[0], counts, w, 0.1)
    o.run_steps(5)
    assert o.get(n, "V_m")[0] == E_L
    assert o.get(n, "I_syn_ex")[0] == 0.0 and o.get(n, "I_syn_in")[0] == 0.0
    o.run_steps(1)  # through step 5
    ch, other = ("I_syn_ex", "I_syn_in") if w > 0 else ("I_syn_in", "I_syn_ex")
    assert o.get(n, ch)[0] == w and o.get(n, other)[0] == 0.0
    assert o.get(n, "V_m")[0] == E_L
    for k in range(1, 120):
        o.run_steps(1)
        t = k * h
        i_exp = w * np.exp(-t / tau_s)
        v_exp = w / C * tau_s * tau_m / (tau_m - tau_s) * (np.exp(-t / tau_m) - np.exp(-t / tau_s))
        assert abs(o.get(n, ch)[0] - i_exp) <= 1e-12 * abs(w)
        assert abs(o.get(n, "V_m")[0] - E_L - v_exp) <= 1e-12 * (abs(v_exp) + 1e-3)
        assert o.get(n, other)[0] == 0.0


@pytest.mark.parametrize("w", [3.5, -2.25])
def test_iaf_psc_alpha_single_psc_closed_form(w):
    # One input of weight w read in the update of step k: after that step's V update and
    # current propagation, dI += w e / tau_s (iaf_psc_alpha.cpp:329-346), so n steps later
    # I_syn = w (t / tau_s) e^(1 - t / tau_s) and, with a = w e / tau_s, k = 1/tau_s - 1/tau_m,
    # V_m - E_L = a / C_m e^(-t/tau_m) (1 - e^(-k t) (1 + k t)) / k^2, t = n h: the exact
    # solution of dV/dt = -V/tau_m + I/C_m. Fails for the channels swapped, a missing factor e,
    # the input added before the propagation, or P31 and P32 exchanged.
    tau_m, tau_s, C, E_L, h = 10.0, 2.0, 250.0, -70.0, 0.1
    o = Oracle()
    n = o.create("iaf_psc_alpha", 1, {"tau_m": tau_m, "tau_syn_ex": tau_s, "tau_syn_in": tau_s, "C_m": C,
                                       "E_L": E_L, "V_m": E_L, "V_th": 1e9})
    counts = np.zeros((10, 1), np.uint8)
    counts[4, 0] = 1  # emission step 4, delay 1 step: read in the update of step 5
    o.add_drive(n[0], counts, w, 0.1)
    o.run_steps(6)  # through step 5
    ch, other = ("I_syn_ex", "I_syn_in") if w > 0 else ("I_syn_in", "I_syn_ex")
    assert o.get(n, ch)[0] == 0.0 and o.get(n, "V_m")[0] == E_L
    a = w * np.e / tau_s
    k = 1.0 / tau_s - 1.0 / tau_m
    peak = 0.0
    for step in range(1, 150):
        o.run_steps(1)
        t = step * h
        i_exp = w * (t / tau_s) * np.exp(1.0 - t / tau_s)
        v_exp = a / C * np.exp(-t / tau_m) * (1.0 - np.exp(-k * t) * (1.0 + k * t)) / k ** 2
        assert abs(o.get(n, ch)[0] - i_exp) <= 1e-12 * abs(w)
        assert abs(o.get(n, "V_m")[0] - E_L - v_exp) <= 1e-11 * (abs(v_exp) + 1e-3)
        assert o.get(n, other)[0] == 0.0
        peak = max(peak, abs(o.get(n, ch)[0]))
    assert peak == pytest.approx(abs(w), rel=1e-12)  # the current peaks at w, at t = tau_s


def test_iaf_psc_alpha_dc_and_equal_time_constants():
    # I_e alone: V_m - E_L = I_e tau_m / C_m (1 - e^(-t/tau_m)) from the first step; with
    # tau_s = tau_m the propagators take their singular limits and one input still gives the
    # exact solution V = a / C t^2 / 2 e^(-t/tau_m) (iaf_propagator.cpp:94-116)
    tau_m, C, E_L, h, I_e = 10.0, 250.0, -70.0, 0.1, 300.0
    o = Oracle()
    n = o.create("iaf_psc_alpha", 1, {"tau_m": tau_m, "C_m": C, "E_L": E_L, "V_m": E_L, "V_th": 1e9, "I_e": I_e})
    for step in range(1, 40):
        o.run_steps(1)
        v_exp = I_e * tau_m / C * (1 - np.exp(-step * h / tau_m))
        assert abs(o.get(n, "V_m")[0] - E_L - v_exp) <= 1e-11 * v_exp
    w = 4.0
    o2 = Oracle()
    n2 = o2.create("iaf_psc_alpha", 1, {"tau_m": tau_m, "tau_syn_ex": tau_m, "C_m": C, "E_L": E_L, "V_m": E_L,
                                         "V_th": 1e9})
    counts = np.zeros((3, 1), np.uint8)
    counts[0, 0] = 1
    o2.add_drive(n2[0], counts, w, 0.1)
    o2.run_steps(2)
    a = w * np.e / tau_m
    for step in range(1, 60):
        o2.run_steps(1)
        t = step * h
        v_exp = a / C * t * t / 2 * np.exp(-t / tau_m)
        assert abs(o2.get(n2, "V_m")[0] - E_L - v_exp) <= 1e-10 * v_exp


# testsuite/pytests/neurons/test_iaf_psc_alpha.py: the reference's V_m traces (rtol 1e-6)
ALPHA_I0 = [(0.1, -69.602), (0.2, -69.2079), (0.3, -68.8178), (0.4, -68.4316), (0.5, -68.0492), (4.3, -56.0204),
            (4.4, -55.7615), (4.5, -55.5051), (4.6, -55.2513), (4.7, -55.0001), (4.8, -70), (4.9, -70), (5.0, -70)]
ALPHA_I0_REFR = [-69.4229, -68.8515, -68.2858, -67.7258, -67.1713, -66.6223, -66.0788, -65.5407, -65.008, -64.4806,
                 -63.9584, -63.4414, -62.9295, -62.4228, -61.9211, -61.4243, -60.9326, -60.4457, -59.9636, -59.4864,
                 -59.0139, -58.5461, -58.0829, -57.6244, -57.1704, -56.721, -56.276, -55.8355, -55.3993] + [-70.0] * 21 + [
                 -69.4229, -68.8515, -68.2858, -67.7258, -67.1713, -66.6223, -66.0788, -65.5407, -65.008, -64.4806,
                 -63.9584, -63.4414, -62.9295, -62.4228, -61.9211, -61.4243, -60.9326, -60.4457, -59.9636]


def alpha_i0_refractory_expected():
    """(time, V_m) of test_iaf_psc_alpha_i0_refractory, 0.1 ms to 6.9 ms."""
    return [(round(0.1 * (i + 1), 10), v) for i, v in enumerate(ALPHA_I0_REFR)]


def _alpha_trace(res, I_e, T=8.0):
    o = Oracle(resolution=res)
    n = o.create("iaf_psc_alpha", 1, {"I_e": I_e})
    o.set_sampling(n, 1, 0)
    o.simulate(T)
    g, st, f = o.samples()
    g2, s2 = o.spikes()
    return dict(zip(np.round(st * res, 10), f[:, 0])), s2 * res


def test_iaf_psc_alpha_reference_trace_i0():
    trace, spikes = _alpha_trace(0.1, 1000.0)
    for t, v in ALPHA_I0:
        assert trace[round(t, 10)] == pytest.approx(v, rel=1e-6), t
    assert list(np.round(spikes, 10)) == [4.8]


@pytest.mark.parametrize("res", [0.1, 0.2, 0.5, 1.0])
def test_iaf_psc_alpha_reference_trace_i0_refractory(res):
    trace, spikes = _alpha_trace(res, 1450.0)
    n = 0
    for t, v in alpha_i0_refractory_expected():
        k = round(t / res)
        if abs(k * res - t) < 1e-9:
            assert trace[round(k * res, 10)] == pytest.approx(v, rel=1e-6), (res, t)
            n += 1
    assert n >= 6
    assert np.round(spikes[0], 10) == 3.0


# ------------------------------------------------------------------ receive-order summation (a11)
B = 2.0 ** 53


def _sum_net(n_vp, pre_weights, creation=None):
    """Sources (iaf_psc_delta, one spike each, all in the same step) into one target whose V_m
    after the arrival step equals the sum of the cell: tau_m huge, V_th huge, E_L = 0."""
    o = Oracle(n_vp=n_vp)
    relay = {"E_L": 0.0, "V_th": 20.0, "V_reset": 0.0, "tau_m": 1e6, "C_m": 1.0, "t_ref": 2.0, "V_m": 0.0}
    pre = o.create("iaf_psc_delta", len(pre_weights), relay)
    post = o.create("iaf_psc_delta", 1, dict(relay, V_th=1e300, V_min=-1e300))
    for i, ws in (creation or enumerate(pre_weights)):
        for wv in (ws if isinstance(ws, (list, tuple)) else [ws]):
            o.add_connections([pre[i]], [post[0]], [wv], [2])
    counts = np.zeros((3, len(pre_weights)), np.uint8)
    counts[2, :] = 1
    o.add_drive(pre[0], counts, 25.0, 0.1)
    o.run_steps(3 + 1 + 2)  # pre spikes at stamp 4, read in the update of step 5
    return o.get(post, "V_m")[0]


def test_receive_order_summation_one_vp():
    # One VP: spikes of one slice are added in ascending source gid (event_delivery_manager.cpp:
    # 470-497, 599-776); 0 + a + b + c rounds differently depending on which weight comes
    # last: with (-2^53, 0.75, 2^53) in gid order the cell is 1.0; the reversed order gives 0.0,
    # 2^53 + 0.75 - 2^53 gives 0.0 and (2^53 - 2^53) + 0.75 gives 0.75.
    assert _sum_net(1, [-B, 0.75, B]) == 1.0
    assert _sum_net(1, [B, 0.75, -B]) == 0.0


def _fold(ws):
    acc = 0.0
    for x in ws:
        acc = acc + x
    return acc


def test_receive_order_summation_across_vps():
    # Two VPs: the receive buffer holds VP 0's spikes, then VP 1's, each in ascending gid
    # (collocate_spike_data_buffers_, event_delivery_manager.cpp:470-497; vp = gid % 2). Source
    # gids 1..4 sit on VPs 1, 0, 1, 0, so the cell adds gid 2, 4, 1, 3. The weights are chosen
    # so that gid order (1, 2, 3, 4) and VP-1-first order (1, 3, 2, 4) give other sums.
    for w, expected in (([0.0, -B, B, 0.75], 1.0), ([B, 0.0, 0.75, -B], 0.75)):
        receive = [w[1], w[3], w[0], w[2]]
        assert _fold(receive) == expected
        assert _fold(w) != expected and _fold([w[0], w[2], w[1], w[3]]) != expected
        assert _sum_net(2, w) == expected


def test_connector_run_order_within_a_source():
    # Connections of one source to one target are added in connector run order, which is
    # creation order here (connector_base.h:410-427): source gid 1 = -2^53, then gid 2 with two
    # connections created 0.75 first and 2^53 second gives (-2^53 + 0.75) + 2^53 = 1.0;
    # the reversed run gives (-2^53 + 2^53) + 0.75 = 0.75.
    assert _sum_net(1, [-B, [0.75, B]]) == 1.0
    assert _sum_net(1, [-B, [B, 0.75]]) == 0.75


def _polar_normals(st, n, mean=0.0, sd=1.0):
    """n values of one std::normal_distribution (libstdc++: Marsaglia's polar method, the
    second value of each pair cached) over the stream st."""
    import math
    out = []
    while len(out) < n:
        while True:
            x = 2.0 * st.canonical() - 1.0
            y = 2.0 * st.canonical() - 1.0
            r2 = x * x + y * y
            if r2 <= 1.0 and r2 != 0.0:
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        out.extend([y * mult * sd + mean, x * mult * sd + mean])
    return out[:n]


def test_normal_parameter_object_carries_its_state_across_connect_calls():
    # NormalParameter (parameter.cpp:95-117) owns one normal_distribution per thread: used in two
    # Connect calls, the second call continues with the value cached by the first. Three
    # one_to_one connections per call: the six weights are six consecutive values of one
    # distribution on the VP stream. Fails for a parameter whose state restarts per call (the
    # fourth weight would then come from a new pair).
    o = Oracle(n_vp=1, seed=143202461)
    g = o.create("iaf_psc_delta", 9)
    w = ("normal", 2.0, 0.5, 1)
    o.connect(g[:3], g[3:6], "one_to_one", weight=w, delay=1.0)
    o.connect(g[:3], g[6:9], "one_to_one", weight=w, delay=1.0)
    s, t, wv, d = o.connections()
    assert wv.tolist() == _polar_normals(_Stream(SEED_VP0), 6, 2.0, 0.5)
    # two objects: the second call starts a pair of its own and discards nothing on the stream
    o2 = Oracle(n_vp=1, seed=143202461)
    g = o2.create("iaf_psc_delta", 9)
    o2.connect(g[:3], g[3:6], "one_to_one", weight=("normal", 2.0, 0.5, 1), delay=1.0)
    o2.connect(g[:3], g[6:9], "one_to_one", weight=("normal", 2.0, 0.5, 2), delay=1.0)
    s, t, wv2, d = o2.connections()
    st = _Stream(SEED_VP0)
    first = _polar_normals(st, 4, 2.0, 0.5)[:3]
    assert wv2.tolist()[:3] == first and wv2.tolist()[3:] == _polar_normals(st, 3, 2.0, 0.5)
    assert wv2.tolist() != wv.tolist()


# ---------------------------------------------------------------- multimeter sampling
def test_sampling_stamps_follow_interval_and_offset():
    # testsuite/pytests/recording/test_multimeter_offset.py::test_recorded_times_relative_to_offset:
    # resolution 2^-3 ms, interval 3 ms, offset 5 ms, Simulate(15) -> times 5, 8, 11, 14
    o = Oracle(resolution=0.125)
    n = o.create("iaf_psc_exp", 1)
    o.set_sampling(n, interval_steps=24, offset_steps=40)
    o.simulate(15.0)
    g, st, f = o.samples()
    assert list(st * 0.125) == [5.0, 8.0, 11.0, 14.0]
    # the default interval of 1 ms and offset 0 sample at every full millisecond after the start
    o = Oracle()
    n = o.create("iaf_psc_delta", 2)
    o.set_sampling(n, interval_steps=10)
    o.simulate(3.5)
    g, st, f = o.samples()
    assert list(st) == [10, 10, 20, 20, 30, 30] and list(g) == [1, 2] * 3


def test_sampled_state_is_after_the_step():
    # the logger records after the step's update (iaf_psc_exp.cpp:346): a DC-driven neuron's
    # samples follow V(t) = E_L + (1 - exp(-t / tau_m)) R I at the sample stamps t; a sample taken
    # before the update would lag by one step and miss this by about 0.4 mV
    dt, I = 0.1, 200.0
    o = Oracle(resolution=dt)
    n = o.create("iaf_psc_exp", 1, {"I_e": I})
    o.set_sampling(n, interval_steps=3, offset_steps=2)
    o.simulate(6.0)
    g, st, f = o.samples()
    t = st * dt
    assert st[0] == 2 and np.all(np.diff(st) == 3)
    np.testing.assert_allclose(f[:, 0], -70.0 + (1.0 - np.exp(-t / 10.0)) * (10.0 / 250.0) * I, rtol=0, atol=1e-9)
    # the synaptic currents of an unstimulated neuron stay at zero
    assert np.all(f[:, 1:3] == 0.0)


# ------------------------------------------------------------------ Philox_32 streams
PHILOX_PROBE = os.path.join(ROOT, "oracle", "_ref", "philox_probe")


def _philox_draws(family, n, rng, seed):
    import ctypes

    L = lib()
    L.or_rng_draws.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int, ctypes.c_int,
                               ctypes.c_ulong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    s = (ctypes.c_uint32 * len(seed))(*seed)
    out = (ctypes.c_double * n)()
    assert L.or_rng_draws(1, s, len(seed), family, rng, n, out) == 0
    return list(out)


def test_philox4x32_known_answers():
    # Random123's own known-answer vectors (thirdparty/Random123/tests/kat_vectors), 7 and 10 rounds
    import ctypes

    L = lib()
    with open(os.path.join(GOLDEN, "philox4x32_kat.json")) as f:
        kat = json.load(f)
    assert len(kat) == 6
    for v in kat:
        c = (ctypes.c_uint32 * 4)(*v["ctr"])
        k = (ctypes.c_uint32 * 2)(*v["key"])
        o = (ctypes.c_uint32 * 4)()
        L.or_philox4x32(v["rounds"], c, k, o)
        assert list(o) == v["out"], v


def test_philox_engine_draws_match_reference_vectors():
    # engine words, uniform_real, uniform_int (a range rejecting about half of the words
    # included) and normal draws of r123::Engine<Philox4x32> seeded as NEST seeds a VP stream,
    # from the reference's own headers (tests/golden/make_philox_vectors.py)
    with open(os.path.join(GOLDEN, "philox_engine_draws.json")) as f:
        cases = json.load(f)
    assert len(cases) == 24
    for c in cases:
        got = _philox_draws(c["family"], len(c["values"]), c["range"], c["seed"])
        if c["family"] in (1, 3):
            assert got == c["values"], (c["seed"], c["family"])
        else:
            assert [int(x) for x in got] == c["values"], (c["seed"], c["family"], c["range"])


@pytest.mark.skipif(not os.path.exists(PHILOX_PROBE), reason="reference probe not built (oracle/ref)")
def test_philox_engine_draws_match_reference_probe_live():
    rng = np.random.default_rng(11)
    for _ in range(12):
        seed = [int(rng.integers(1, 2 ** 32)), 0xB84C9BAE, int(rng.integers(0, 64))]
        fam = int(rng.integers(0, 4))
        r = int(rng.choice([2, 1000, 77169, 3000000001]))
        txt = subprocess.run([PHILOX_PROBE, str(fam), "300", str(r)] + [str(x) for x in seed], check=True,
                             capture_output=True, text=True).stdout.split()
        got = _philox_draws(fam, 300, r, seed)
        want = [float(x) for x in txt]
        assert got == want, (seed, fam, r)


# ------------------------------------------------------------------ stdp_synapse
def _stdp_reproduce(pre, post, w0, Wmax, lam, alpha, tau_plus, tau_minus, mu_plus, mu_minus):
    """Final weight of one stdp_synapse from its pre- and postsynaptic spike trains (post times
    already shifted by the dendritic delay), processed event by event in time order: a
    presynaptic spike depresses with the postsynaptic trace and then raises its own trace; a
    postsynaptic spike facilitates with the presynaptic trace and then raises its own. This is
    the reproduction testsuite/pytests/synapses/test_stdp_synapse_basic.py:120-212 makes."""
    kp = km = 0.0
    last_pre = last_post = 0.0
    w = w0 / Wmax
    i = j = 0
    pre_t = pre[0] if len(pre) else np.inf
    post_t = post[0] if len(post) else np.inf

    def fac():
        nonlocal w
        w = min(1.0, w + lam * (1.0 - w) ** mu_plus * kp * np.exp((last_pre - post_t) / tau_plus))

    def dep():
        nonlocal w
        w = max(0.0, w - lam * alpha * w ** mu_minus * km * np.exp((last_post - pre_t) / tau_minus))

    while j < len(pre) or i < len(post):
        if pre_t == post_t:
            if last_post != post_t:
                fac()
            if last_pre != pre_t:
                dep()
            if j + 1 >= len(pre):
                break
            kp = kp * np.exp((last_pre - pre_t) / tau_plus) + 1.0
            j += 1
            last_pre, pre_t = pre_t, pre[j]
            if i + 1 < len(post):
                km = km * np.exp((last_post - post_t) / tau_minus) + 1.0
            i += 1
            last_post, post_t = post_t, (post[i] if i < len(post) else np.inf)
        elif pre_t < post_t:
            dep()
            kp = kp * np.exp((last_pre - pre_t) / tau_plus) + 1.0
            if j + 1 >= len(pre):
                break
            j += 1
            last_pre, pre_t = pre_t, pre[j]
        else:
            fac()
            km = km * np.exp((last_post - post_t) / tau_minus) + 1.0
            i += 1
            last_post, post_t = post_t, (post[i] if i < len(post) else np.inf)
    return w * Wmax


# the spike trains of testsuite/pytests/synapses/test_stdp_synapse.py:76-84: every order of pre
# and post spikes, ties included, then Poisson trains
STDP_PRE_MS = [1, 5, 6, 7, 9, 11, 12, 13, 14.5, 16.1, 21, 25, 26, 27, 29, 31, 32, 33, 34.5, 36.1]
STDP_POST_MS = [2, 3, 4, 8, 9, 10, 12, 12.2, 14.1, 15.4, 22, 23, 24, 28, 29, 30, 32, 33.2, 35.1, 36.4]
# its synapse: mu 0 (additive), tau_plus 16.8, tau_minus 33.7, Wmax 15, initial weight 0.5
STDP_REF_PRM = {"tau_plus": 16.8, "lambda": 0.01, "alpha": 0.85, "mu_plus": 0.0, "mu_minus": 0.0, "Wmax": 15.0}


def stdp_relay_counts(T=300.0, h=0.1, seed=11):
    """Dense drive counts that make two relays fire at the hard-coded times (shifted by the
    relay latency), then at Poisson times (about 50 Hz) up to T - 10 ms."""
    steps = int(round(T / h))
    counts = np.zeros((steps, 2), np.uint8)
    rng = np.random.default_rng(seed)
    for col, times in ((0, STDP_PRE_MS), (1, STDP_POST_MS)):
        for t in times:
            counts[int(round(t / h)), col] = 1
        tail = np.flatnonzero(rng.random(steps) < 0.005)
        tail = tail[(tail * h > 40.0) & (tail * h < T - 10.0)]
        counts[tail, col] = 1
    return counts


def stdp_relays(o_or_k, oracle, delay):
    """Two iaf_psc_delta relays that fire once per drive count (t_ref 0, a 1e4 mV kick; the
    postsynaptic relay's threshold is far above what the stdp synapse can add) joined by one
    stdp_synapse of the reference test's parameters."""
    relay = {"E_L": 0.0, "V_m": 0.0, "V_reset": 0.0, "t_ref": 0.0, "tau_m": 10.0, "C_m": 250.0, "tau_minus": 33.7}
    if oracle:
        pre = o_or_k.create("iaf_psc_delta", 1, dict(relay, V_th=20.0))
        post = o_or_k.create("iaf_psc_delta", 1, dict(relay, V_th=1000.0))
        o_or_k.connect(pre, post, "one_to_one", weight=0.5, delay=delay, stdp=STDP_REF_PRM)
        return pre, post
    pre = o_or_k.Create("iaf_psc_delta", 1, dict(relay, V_th=20.0))
    post = o_or_k.Create("iaf_psc_delta", 1, dict(relay, V_th=1000.0))
    o_or_k.Connect(pre, post, "one_to_one", dict(STDP_REF_PRM, synapse_model="stdp_synapse", weight=0.5,
                                                 delay=delay))
    return pre, post


@pytest.mark.parametrize("delay", [0.1, 1.0, 2.5])
def test_stdp_synapse_reference_trains(delay):
    # the reference test's hard-coded pre/post orders and ties, through relays, against the
    # event-driven reproduction of the final weight
    T, h = 300.0, 0.1
    o = Oracle(resolution=h)
    pre, post = stdp_relays(o, True, delay)
    o.add_drive(pre[0], stdp_relay_counts(T, h), 1e4, 0.1)
    o.simulate(T)
    g, s = o.spikes()
    t_pre, t_post = s[g == pre[0]] * h, s[g == post[0]] * h
    # spike times follow the drive one step later, with the reference test's ties intact
    assert np.allclose(t_pre[:20], np.array(STDP_PRE_MS) + 0.2) and np.allclose(t_post[:20], np.array(STDP_POST_MS) + 0.2)
    w_sim = o.connections()[2][0]
    w_rep = _stdp_reproduce(t_pre, t_post + delay, 0.5, 15.0, 0.01, 0.85, 16.8, 33.7, 0.0, 0.0)
    assert w_sim != 0.5
    assert w_sim == pytest.approx(w_rep, abs=1e-12), (w_sim, w_rep)


@pytest.mark.parametrize("delay,mu", [(1.0, (1.0, 1.0)), (2.3, (1.0, 1.0)), (1.0, (0.0, 1.0)), (1.5, (1.0, 0.0))])
def test_stdp_synapse_reproduced_from_spike_trains(delay, mu):
    # one stdp_synapse from an irregularly firing neuron onto a driven one (the reference's
    # test_stdp_synapse_basic setup, with iaf_psc_exp neurons and seeded Poisson drive)
    T, h = 10000.0, 0.1
    o = Oracle(resolution=h)
    pre = o.create("iaf_psc_exp", 1, {"I_e": 300.0})
    post = o.create("iaf_psc_exp", 1, {"I_e": 330.0, "tau_minus": 30.0})
    n_steps = int(round(T / h))
    rng = np.random.default_rng(int(delay * 10) + int(mu[0]) * 3 + int(mu[1]))
    o.add_drive(pre[0], rng.poisson(0.04, size=(n_steps, 1)).astype(np.uint8), 900.0, 1.0)
    o.add_drive(post[0], rng.poisson(0.03, size=(n_steps, 1)).astype(np.uint8), 700.0, 1.0)
    prm = {"tau_plus": 20.0, "lambda": 0.01, "alpha": 1.1, "mu_plus": mu[0], "mu_minus": mu[1], "Wmax": 90.0}
    o.connect(pre, post, "one_to_one", weight=45.0, delay=delay, stdp=prm)
    o.simulate(T)
    g, s = o.spikes()
    t_pre = s[g == pre[0]] * h
    t_post = s[g == post[0]] * h
    # the last slice's spikes are not delivered before the run ends
    t_pre = t_pre[t_pre <= T - 1.0]
    assert len(t_pre) > 50 and len(t_post) > 50
    w_sim = o.connections()[2][0]
    w_rep = _stdp_reproduce(t_pre, t_post + delay, 45.0, 90.0, 0.01, 1.1, 20.0, 30.0, mu[0], mu[1])
    assert w_sim != 45.0
    assert w_sim == pytest.approx(w_rep, abs=1e-12), (w_sim, w_rep)
