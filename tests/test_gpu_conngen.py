"""Connectivity generated on the device from the Philox_32 per-VP streams.

pairwise_bernoulli and fixed_indegree (autapses and multapses allowed) with constant weight
and delay draw on the device when connect_on_device is set. The connections, in creation
order, must equal those the host builder makes from the same streams and those of the oracle
(NEST's builders over r123::Engine<Philox4x32>, pinned on the reference's own headers), and
the streams must continue where the host's would: a later Connect with random parameters
(drawn on the host) gives the same connections either way.
"""

import numpy as np
import pytest

import nestgpu
import netbuild as nb
from _oracle import Oracle

pytestmark = pytest.mark.gpu

CASES = [
    # (n_vp, seed, n_pre, n_post (0: the same population), rule, param, autapses)
    (3, 17, 2000, 0, "pairwise_bernoulli", 0.1, False),
    (4, 5, 3000, 1500, "pairwise_bernoulli", 0.05, True),
    (1, 23, 700, 0, "pairwise_bernoulli", 1.0, False),
    # 200,000 sources: Lemire's method rejects about 4e-5 of the words, a few dozen here
    (2, 99, 200000, 2000, "fixed_indegree", 400, True),
    (5, 3, 5000, 0, "fixed_indegree", 100, True),
]


def _build(case, on_device):
    n_vp, seed, n_pre, n_post, rule, param, aut = case
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"total_num_virtual_procs": n_vp, "rng_seed": seed, "rng_type": "Philox_32",
                       "connect_on_device": on_device})
    a = k.Create("iaf_psc_exp", n_pre)
    b = a if n_post == 0 else k.Create("iaf_psc_exp", n_post)
    spec = {"rule": rule, "p" if rule == "pairwise_bernoulli" else "indegree": param}
    if rule == "pairwise_bernoulli":
        spec["allow_autapses"] = aut
    k.Connect(a, b, spec, {"weight": 2.5, "delay": 1.2})
    calls = k.GetKernelStatus("device_connect_calls")
    # a later call with random parameters draws on the host from where the streams stand
    k.Connect(b[:50], a[:40], {"rule": "fixed_indegree", "indegree": 5},
              {"weight": nestgpu.random.normal(0.0, 1.0), "delay": nestgpu.random.uniform(1.0, 3.0)})
    return k, calls


@pytest.mark.parametrize("case", CASES, ids=["%s-vp%d-%d" % (c[4], c[0], c[2]) for c in CASES])
def test_device_connectivity_equals_host_and_oracle(case):
    kd, calls_d = _build(case, True)
    kh, calls_h = _build(case, False)
    assert calls_d == 1 and calls_h == 0
    cd, ch = kd.GetConnections(), kh.GetConnections()
    n = len(ch["source"])
    print("%s: %d connections" % (case[4], n), flush=True)
    assert n > 0
    for key in ("source", "target", "weight", "delay_steps"):
        assert np.array_equal(cd[key], ch[key]), key
    if n <= 2_000_000:
        n_vp, seed, n_pre, n_post, rule, param, aut = case
        o = Oracle(n_vp=n_vp, seed=seed, rng_type="Philox_32")
        a = o.create("iaf_psc_exp", n_pre)
        b = a if n_post == 0 else o.create("iaf_psc_exp", n_post)
        o.connect(a, b, rule, param, weight=2.5, delay=1.2, allow_autapses=aut)
        o.connect(b[:50], a[:40], "fixed_indegree", 5, weight=("normal", 0.0, 1.0), delay=("uniform", 1.0, 3.0))
        s, t, w, d = o.connections()
        assert np.array_equal(cd["source"], s) and np.array_equal(cd["target"], t)
        assert np.array_equal(cd["weight"], w) and np.array_equal(cd["delay_steps"], d)


def test_device_connected_network_against_oracle():
    # a recurrent iaf_psc_exp network connected on the device, simulated against the oracle
    desc = dict(n_vp=4, seed=31, rng_type="Philox_32",
                pops=[("iaf_psc_exp", 1200, {"I_e": 380.0})],
                conns=[((0, 0, 960), (0, 0, 1200), "fixed_indegree", 80, 30.0, 1.5),
                       ((0, 960, 1200), (0, 0, 1200), "pairwise_bernoulli", 0.1, -90.0, 1.0)],
                v0={0: np.random.default_rng(3).uniform(-70.0, -55.0, 1200)})
    k, kp, sr = nb.build_gpu(desc, layout={"connect_on_device": True})
    assert k.GetKernelStatus("device_connect_calls") == 2
    o, op = nb.build_oracle(desc)
    k.Simulate(200.0)
    o.simulate(200.0)
    ev = sr.get("events")
    g, s = o.spikes()
    order = np.lexsort((g, s))
    print("device-connected network: %d spikes (oracle %d)" % (len(ev["senders"]), len(g)), flush=True)
    assert len(g) > 100
    assert np.array_equal(ev["senders"], g[order])
    assert np.array_equal(np.atleast_1d(kp[0].get("V_m")), o.get(op[0], "V_m"))
