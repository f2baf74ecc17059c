"""stdp_synapse on the device path vs the oracle.

The device applies stdp_synapse::send (models/stdp_synapse.h:222-283) to each event as it is
delivered and records every spike of a neuron with incoming stdp connections in its
post-spike history (ArchivingNode::set_spiketime, archiving_node.cpp:196-230). Both sides
compute the traces with the same restated exp (portable math in the oracle), so spike trains,
neuron states and every final weight must be bit-identical, on every delivery layout, across
split runs and in the batched form.

This path's limits raise instead of diverging: a neuron with outgoing stdp_synapse
connections that spikes twice within one min_delay slice, and a target whose post-spike
history between two presynaptic spikes outgrows the entries kept per neuron
(stdp_history_capacity).
"""

import numpy as np
import pytest

import nestgpu
import netbuild as nb
from _oracle import Oracle

pytestmark = pytest.mark.gpu

LAYOUTS = {"auto": {}, "win48": {"delivery_mode": "window", "targets_per_block": 48},
           "win2": {"delivery_mode": "window", "window_slots": 2},
           "glob32": {"delivery_mode": "global", "targets_per_block": 32},
           "glob64p3": {"delivery_mode": "global", "targets_per_block": 64, "delivery_passes": 3},
           "cache": {"delivery_mode": "cache"}}
FIELDS = ("V_m", "I_syn_ex", "I_syn_in")


def stdp_net(N=1500, KE=80, KI=20, T=300.0, n_vp=2, seed=21, drive_seed=4, lam=0.8, mu=(1.0, 1.0),
             lam_stdp=0.05):
    """iaf_psc_exp network: the excitatory connections are stdp_synapse (normal initial
    weights, Wmax 150), the inhibitory ones static; delays 1.0-3.0 ms, refractory time 2 ms
    (no neuron spikes twice within a 1 ms slice)."""
    NE = int(0.8 * N)
    params = {"tau_m": 10.0, "tau_syn_ex": 0.5, "tau_syn_in": 0.5, "C_m": 250.0, "V_th": -55.0, "V_reset": -70.0,
              "t_ref": 2.0, "E_L": -70.0, "tau_minus": 30.0}
    rng = np.random.default_rng(drive_seed)
    steps = int(round(T / 0.1))
    delay = ("sum", 1.0, ("product", ("uniform_int", 21), 0.1))
    stdp = {"tau_plus": 15.0, "lambda": lam_stdp, "alpha": 1.2, "mu_plus": mu[0], "mu_minus": mu[1],
            "Wmax": 150.0, "Kplus": 0.5}
    return dict(
        n_vp=n_vp, seed=seed,
        pops=[("iaf_psc_exp", N, params)],
        v0={0: rng.uniform(-70, -55, N)},
        conns=[((0, 0, NE), (0, 0, N), "fixed_indegree", KE, ("normal", 60.0, 6.0), delay, True, True, stdp),
               ((0, NE, N), (0, 0, N), "fixed_indegree", KI, ("product", ("normal", 60.0, 6.0), -4.0), delay)],
        drive=[(0, rng.poisson(lam, size=(steps, N)).astype(np.uint8), 87.8, 1.0)],
        T=T)


def _weights_gpu(k, pops):
    # neuron to neuron connections (the oracle holds the drive and the recorder elsewhere)
    c = k.GetConnections()
    ids = np.concatenate([np.atleast_1d(p.global_id) for p in pops])
    keep = np.isin(c["source"], ids) & np.isin(c["target"], ids)
    return c["source"][keep], c["target"][keep], c["weight"][keep], c["delay_steps"][keep]


def _check(k, pops, sr, o, opops, w0=None):
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    print("spikes gpu=%d oracle=%d" % (len(ev["senders"]), len(og)), flush=True)
    assert len(ev["senders"]) == len(og)
    assert np.array_equal(ev["senders"], og)
    assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9)
    for p, op in zip(pops, opops):
        for f in FIELDS:
            a = np.atleast_1d(p.get(f))
            b = o.get(op, f)
            assert np.array_equal(a, b), "%s differs: max abs diff %g" % (f, np.max(np.abs(a - b)))
    s1, t1, w1, d1 = _weights_gpu(k, pops)
    s2, t2, w2, d2 = o.connections()
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2) and np.array_equal(d1, d2)
    bad = np.flatnonzero(w1 != w2)
    assert len(bad) == 0, "%d weights differ, first at %d: %r vs %r" % (
        len(bad), bad[0], w1[bad[0]], w2[bad[0]])
    if w0 is not None:
        changed = np.count_nonzero(w1 != w0)
        print("weights changed by stdp: %d of %d (range %.3f .. %.3f)" % (changed, len(w1), w1.min(), w1.max()),
              flush=True)
        assert changed > len(w1) // 10
    return w1


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_stdp_network(layout):
    desc = stdp_net()
    k, pops, sr = nb.build_gpu(desc, layout=LAYOUTS[layout])
    o, opops = nb.build_oracle(desc, portable_math=True)
    w0 = _weights_gpu(k, pops)[2].copy()
    k.Simulate(desc["T"])
    o.simulate(desc["T"])
    print("layout=%s window=%s tb=%d" % (layout, k.GetKernelStatus("window_delivery"),
                                         k.GetKernelStatus("targets_per_block")), flush=True)
    _check(k, pops, sr, o, opops, w0)


@pytest.mark.parametrize("mu", [(0.0, 0.0), (0.0, 1.0)])
def test_stdp_split_runs_and_readback(mu):
    # runs ending inside slices; the weights read back between runs do not disturb the next
    desc = stdp_net(N=1000, n_vp=3, seed=8, drive_seed=13, mu=mu, lam_stdp=0.02)
    k, pops, sr = nb.build_gpu(desc)
    o, opops = nb.build_oracle(desc, portable_math=True)
    w0 = _weights_gpu(k, pops)[2].copy()
    for t in (47.3, 100.0, 0.1, 52.6):
        k.Simulate(t)
        o.simulate(t)
        _check(k, pops, sr, o, opops)
    _check(k, pops, sr, o, opops, w0)


def test_stdp_sources_with_several_parameter_sets():
    # every excitatory source has stdp synapses of two Connect calls with different tau_plus
    # and initial Kplus (two presynaptic traces per source), and static ones
    desc = stdp_net(N=1200, seed=41, drive_seed=6)
    N = 1200
    pre, w, d, stdp = desc["conns"][0][0], desc["conns"][0][4], desc["conns"][0][5], desc["conns"][0][8]
    other = dict(stdp, tau_plus=25.0, Kplus=0.0, **{"lambda": 0.03}, mu_plus=0.0)
    desc["conns"] = [(pre, (0, 0, N // 2), "fixed_indegree", 60, w, d, True, True, stdp),
                     (pre, (0, N // 2, N), "fixed_indegree", 60, w, d, True, True, other),
                     (pre, (0, 0, N), "fixed_indegree", 20, w, d),
                     desc["conns"][1]]
    k, pops, sr = nb.build_gpu(desc)
    o, opops = nb.build_oracle(desc, portable_math=True)
    w0 = _weights_gpu(k, pops)[2].copy()
    k.Simulate(desc["T"])
    o.simulate(desc["T"])
    _check(k, pops, sr, o, opops, w0)


def test_stdp_batched():
    # instances with and without stdp_synapse share the slice launches
    descs = [stdp_net(N=800, seed=31, drive_seed=1), nb.exp_random(N=900, T=200.0, seed=5, n_vp=2),
             stdp_net(N=1200, seed=32, drive_seed=2, mu=(0.0, 1.0))]
    built = [nb.build_gpu(d) for d in descs]
    nestgpu.batch_simulate([b[0] for b in built], 120.0)
    nestgpu.batch_simulate([b[0] for b in built], 80.7)
    for d, (k, pops, sr) in zip(descs, built):
        o, opops = nb.build_oracle(d, portable_math=True)
        o.simulate(200.7)
        _check(k, pops, sr, o, opops)


def test_stdp_source_spiking_twice_in_a_slice_raises():
    # min_delay 5 ms and a 0.5 ms refractory time: the driven source spikes several times per
    # slice
    k = nestgpu.Kernel(0)
    pre = k.Create("iaf_psc_exp", 4, {"I_e": 3000.0, "t_ref": 0.5})
    post = k.Create("iaf_psc_exp", 4)
    k.Connect(pre, post, "one_to_one", {"synapse_model": "stdp_synapse", "weight": 10.0, "delay": 5.0})
    with pytest.raises(nestgpu.NESTError, match="twice"):
        k.Simulate(50.0)


def _late_source(cap=None, T=200.0):
    """The source's first spike comes after about 105 ms; the target spikes about 80 times
    by then."""
    k = nestgpu.Kernel(0)
    if cap:
        k.SetKernelStatus({"stdp_history_capacity": cap})
    pre = k.Create("iaf_psc_exp", 1, {"I_e": 375.01, "V_m": -70.0})
    post = k.Create("iaf_psc_exp", 1, {"I_e": 5000.0, "t_ref": 0.5})
    k.Connect(pre, post, "one_to_one", {"synapse_model": "stdp_synapse", "weight": 10.0, "delay": 1.0})
    sr = k.Create("spike_recorder")
    k.Connect(pre + post, sr)
    return k, (pre, post), sr


def test_stdp_history_capacity():
    # 64 entries do not reach back to the synapse's range: the run fails
    k, _, _ = _late_source(cap=64)
    with pytest.raises(nestgpu.NESTError, match="history"):
        k.Simulate(200.0)
    # the default capacity holds them: same spikes, states and weight as the oracle's
    # unbounded history
    k, (pre, post), sr = _late_source()
    assert k.GetKernelStatus("stdp_history_capacity") == 256
    o = Oracle(portable_math=True)
    opre = o.create("iaf_psc_exp", 1, {"I_e": 375.01, "V_m": -70.0})
    opost = o.create("iaf_psc_exp", 1, {"I_e": 5000.0, "t_ref": 0.5})
    o.connect(opre, opost, "one_to_one", weight=10.0, delay=1.0, stdp={})
    k.Simulate(200.0)
    o.simulate(200.0)
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    print("spikes gpu=%d oracle=%d (source %d)" % (len(ev["senders"]), len(og), np.count_nonzero(og == opre[0])),
          flush=True)
    assert np.count_nonzero(og == opre[0]) >= 1 and np.count_nonzero(og == opost[0]) > 64
    assert np.array_equal(ev["senders"], og)
    assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9)
    for p, op in ((pre, opre), (post, opost)):
        for f in FIELDS:
            assert np.array_equal(np.atleast_1d(p.get(f)), o.get(op, f)), f
    c = k.GetConnections(source=pre, target=post)
    w = o.connections()[2]
    print("weight gpu %r oracle %r" % (c["weight"][0], w[0]), flush=True)
    assert c["weight"][0] == w[0] and w[0] != 10.0


def test_stdp_circuit_scaled():
    # the benchmark circuit at scale 0.1 with stdp_synapse between its excitatory populations
    # (min_delay one step, normal delays), against the oracle
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    T, n_vp = 100.0, 4
    k = nestgpu.Kernel(0)
    pops, n, _ = bench.build_c3(k, nestgpu, n_vp, 20241, 777, scale=0.1, stdp=bench.C3_STDP)
    sr = k.Create("spike_recorder")
    allp = pops[0]
    for p in pops[1:]:
        allp = allp + p
    k.Connect(allp, sr)
    w0 = _weights_gpu(k, pops)[2].copy()
    n_stdp = len(k.GetConnections(synapse_model="stdp_synapse")["source"])
    k.Simulate(T)
    print("circuit scale 0.1: %d neurons, %d connections (%d stdp_synapse), min_delay %d" % (
        n, k.GetKernelStatus("num_connections"), n_stdp, k.GetKernelStatus("min_delay_steps")), flush=True)
    assert k.GetKernelStatus("min_delay_steps") == 1 and n_stdp > 0
    o = Oracle(n_vp=n_vp, seed=20241, portable_math=True)
    opops = bench.build_c3_oracle(o, nestgpu, 777, int(round(T / 0.1)), scale=0.1, stdp=bench.C3_STDP)
    o.simulate(T)
    w1 = _check(k, pops, sr, o, opops)
    print("stdp weights changed: %d of %d" % (np.count_nonzero(w1 != w0), n_stdp), flush=True)
    assert np.count_nonzero(w1 != w0) > n_stdp // 10


@pytest.mark.parametrize("delay", [0.1, 1.0, 2.5])
def test_stdp_reference_trains(delay):
    # testsuite/pytests/synapses/test_stdp_synapse.py's hard-coded pre/post orders (ties
    # included) and Poisson tails through relays: the device's final weight equals the
    # oracle's (portable exp on both sides) and the event-driven reproduction
    import test_oracle_pins as pins

    T, h = 300.0, 0.1
    counts = pins.stdp_relay_counts(T, h)
    k = nestgpu.Kernel(0)
    pre, post = pins.stdp_relays(k, False, delay)
    sg = k.Create("spike_generator", 2)
    k.SetSpikeCounts(sg, counts)
    k.Connect(sg, pre + post, "one_to_one", {"weight": 1e4, "delay": 0.1})
    sr = k.Create("spike_recorder")
    k.Connect(pre + post, sr)
    k.Simulate(T)
    o = Oracle(resolution=h, portable_math=True)
    opre, opost = pins.stdp_relays(o, True, delay)
    o.add_drive(opre[0], counts, 1e4, 0.1)
    o.simulate(T)
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    assert np.array_equal(ev["senders"], og) and np.allclose(ev["times"], os_ * h, rtol=0, atol=1e-9)
    w_dev = k.GetConnections(source=pre, target=post)["weight"][0]
    w_orc = o.connections()[2][0]
    t_pre = os_[og == opre[0]] * h
    t_post = os_[og == opost[0]] * h
    w_rep = pins._stdp_reproduce(t_pre, t_post + delay, 0.5, 15.0, 0.01, 0.85, 16.8, 33.7, 0.0, 0.0)
    print("delay %g: device %r oracle %r reproduction %r" % (delay, w_dev, w_orc, w_rep), flush=True)
    assert w_dev == w_orc
    assert w_dev == pytest.approx(w_rep, abs=1e-12)


def test_stdp_circuit_full_size():
    # the benchmark circuit itself (77,169 neurons, about 3.2e8 synapses, 8 VPs) with every
    # connection between excitatory populations plastic, over 20 ms: spikes and the neurons'
    # state bit-identical to the oracle (the weights act on the state from the first plastic
    # event on; reading all 3.2e8 back is left to the scaled test)
    import os
    import sys
    import time
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    T, n_vp = 20.0, 8
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.time()
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"keep_host_connections": False})
    pops, n, _ = bench.build_c3(k, nestgpu, n_vp, 20241, 777, stdp=bench.C3_STDP)
    sr = k.Create("spike_recorder")
    allp = pops[0]
    for p in pops[1:]:
        allp = allp + p
    k.Connect(allp, sr)
    k.Simulate(T)
    print("device: %d connections, %.1f s" % (k.GetKernelStatus("num_connections"), time.time() - t0), flush=True)
    t0 = time.time()
    o = Oracle(n_vp=n_vp, seed=20241, threads=threads, portable_math=True)
    opops = bench.build_c3_oracle(o, nestgpu, 777, int(round(T / 0.1)), stdp=bench.C3_STDP)
    o.simulate(T)
    print("oracle: %.1f s" % (time.time() - t0), flush=True)
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    print("spikes gpu=%d oracle=%d" % (len(ev["senders"]), len(og)), flush=True)
    assert len(og) > 1000
    assert np.array_equal(ev["senders"], og)
    assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9)
    for p, op in zip(pops, opops):
        for f in FIELDS:
            a, b = np.atleast_1d(p.get(f)), o.get(op, f)
            assert np.array_equal(a, b), "%s differs: max abs diff %g" % (f, np.max(np.abs(a - b)))
