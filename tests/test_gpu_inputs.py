"""Device path inputs and full-size properties.

- spike_generator spike_times / multiplicities, the seeded Poisson drive generated on the
  device, and connect_arrays, each against the oracle fed the same input;
- at the full 1.25e8-synapse size, where the oracle is too slow to follow: the result is the
  same for every device layout (block size, window or global delivery) and run to run.
"""

import os
import sys

import numpy as np
import pytest

import nestgpu
import netbuild as nb
from _oracle import Oracle

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

EXP = {"tau_m": 10.0, "tau_syn_ex": 0.5, "tau_syn_in": 0.5, "C_m": 250.0, "V_th": -55.0, "V_reset": -70.0,
       "t_ref": 2.0, "E_L": -70.0}


def _small_net(k, N=300, seed=3):
    k.SetKernelStatus({"total_num_virtual_procs": 2, "rng_seed": seed})
    pop = k.Create("iaf_psc_exp", N, EXP)
    rng = np.random.default_rng(seed)
    v0 = rng.uniform(-70, -55, N)
    pop.set(V_m=v0)
    k.Connect(pop[:240], pop, {"rule": "fixed_indegree", "indegree": 30},
              {"weight": nestgpu.random.normal(87.8, 8.8), "delay": 1.5})
    k.Connect(pop[240:], pop, {"rule": "fixed_indegree", "indegree": 8},
              {"weight": -350.0, "delay": 1.5})
    return pop, v0


def _oracle_net(N=300, seed=3):
    o = Oracle(n_vp=2, seed=seed)
    g = o.create("iaf_psc_exp", N, EXP)
    rng = np.random.default_rng(seed)
    o.set_values(g, "V_m", rng.uniform(-70, -55, N))
    o.connect(g[:240], g, "fixed_indegree", 30, weight=("normal", 87.8, 8.8), delay=1.5)
    o.connect(g[240:], g, "fixed_indegree", 8, weight=-350.0, delay=1.5)
    return o, g


def _check(k, pop, sr, o, g):
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    print("spikes gpu=%d oracle=%d" % (len(ev["senders"]), len(og)), flush=True)
    assert len(og) > 0
    assert np.array_equal(ev["senders"], og)
    assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9)
    assert np.array_equal(np.atleast_1d(pop.get("V_m")), o.get(g, "V_m"))


def test_spike_times_generators():
    N, T = 300, 80.0
    rng = np.random.default_rng(11)
    steps = int(T / 0.1)
    counts = np.zeros((steps, N), np.uint8)
    k = nestgpu.Kernel(0)
    pop, _ = _small_net(k, N)
    for i in range(N):
        stamps = np.unique(rng.integers(1, steps, size=300))
        mult = rng.integers(1, 4, size=len(stamps))
        counts[stamps - 1, i] = mult
        sg = k.Create("spike_generator", 1, {"spike_times": list(np.round(stamps * 0.1, 1)),
                                              "spike_multiplicities": [int(m) for m in mult]})
        k.Connect(sg, pop[i:i + 1], "one_to_one", {"weight": 150.0, "delay": 1.0})
    sr = k.Create("spike_recorder")
    k.Connect(pop, sr)
    k.Simulate(T)
    o, g = _oracle_net(N)
    o.add_drive(g[0], counts, 150.0, 1.0)
    o.simulate(T)
    _check(k, pop, sr, o, g)


def test_device_poisson_drive_equals_host_counts():
    N, T, rate, seed = 300, 80.0, 9000.0, 4242
    k = nestgpu.Kernel(0)
    pop, _ = _small_net(k, N)
    sg = k.Create("spike_generator", N, {"poisson_rate": rate, "poisson_seed": seed})
    k.Connect(sg, pop, "one_to_one", {"weight": 87.8, "delay": 1.0})
    sr = k.Create("spike_recorder")
    k.Connect(pop, sr)
    k.Simulate(T)
    # the same counts generated on the host: streams are the generators' node ids
    counts = nestgpu.poisson_counts(seed, rate * 0.1e-3, sg.tolist()[0], N, 0, int(T / 0.1))
    o, g = _oracle_net(N)
    o.add_drive(g[0], counts, 87.8, 1.0)
    o.simulate(T)
    _check(k, pop, sr, o, g)


def test_connect_arrays_against_oracle():
    N, T = 400, 60.0
    rng = np.random.default_rng(2)
    src = rng.integers(1, N + 1, size=8000)
    tgt = rng.integers(1, N + 1, size=8000)
    w = rng.normal(60.0, 30.0, size=8000)
    d = np.round(rng.uniform(0.1, 3.0, size=8000), 1)
    counts = rng.poisson(0.9, size=(int(T / 0.1), N)).astype(np.uint8)
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"total_num_virtual_procs": 3})
    pop = k.Create("iaf_psc_exp", N, EXP)
    k.ConnectArrays(src, tgt, w, d)
    gen = k.Create("spike_generator", N)
    k.SetSpikeCounts(gen, counts)
    k.Connect(gen, pop, "one_to_one", {"weight": 87.8, "delay": 0.5})
    sr = k.Create("spike_recorder")
    k.Connect(pop, sr)
    k.Simulate(T)
    o = Oracle(n_vp=3)
    g = o.create("iaf_psc_exp", N, EXP)
    steps = np.array([int(np.floor(x * 10 + 0.5)) for x in d])
    # connect_arrays: every VP adds its targets' connections in array order
    for vp in range(3):
        sel = (tgt % 3) == vp
        o.add_connections(src[sel], tgt[sel], w[sel], steps[sel])
    o.add_drive(g[0], counts, 87.8, 0.5)
    o.simulate(T)
    s1 = k.GetConnections()
    keep = s1["target"] <= N
    s2, t2, w2, d2 = o.connections()
    assert np.array_equal(s1["source"][keep][: len(s2)], s2)
    _check(k, pop, sr, o, g)


def test_no_input_network_stays_at_rest():
    k = nestgpu.Kernel(0)
    pop = k.Create("iaf_psc_exp", 1000, EXP)
    sr = k.Create("spike_recorder")
    k.Connect(pop, sr)
    k.Simulate(50.0)
    assert len(sr.get("events")["senders"]) == 0
    assert np.all(np.atleast_1d(pop.get("V_m")) == -70.0)


def _c2_run(layout, T=100.0):
    k = nestgpu.Kernel(0)
    k.SetKernelStatus(layout)
    pop, sg = bench.build_c2(k, nestgpu, 100000, 1000, 250, 8, 20241, 777)
    sr = k.Create("spike_recorder")
    k.Connect(pop, sr)
    k.Simulate(T)
    ev = sr.get("events")
    return ev["senders"], ev["times"], np.atleast_1d(pop.get("V_m")), k.GetKernelStatus("synaptic_events")


def test_full_size_layout_invariance_and_repeatability():
    ref = _c2_run({})
    print("full size: %d spikes, %d events" % (len(ref[0]), ref[3]), flush=True)
    assert len(ref[0]) > 10000
    for layout in ({}, {"delivery_mode": "global", "targets_per_block": 160},
                   {"delivery_mode": "global", "delivery_passes": 3},
                   {"delivery_mode": "window", "targets_per_block": 0}):
        other = _c2_run(layout)
        assert other[3] == ref[3], layout
        assert np.array_equal(other[0], ref[0]) and np.array_equal(other[1], ref[1]), layout
        assert np.array_equal(other[2], ref[2]), layout


def test_brunel_full_size_full_run():
    # 5,000 iaf_psc_delta (4,000 E / 1,000 I), in-degree 400 E + 100 I, J = 0.1 mV, g = 5,
    # delay 1.5 ms, Poisson drive eta = 2 as pre-generated counts, T = 1000 ms: the whole spike
    # train and every V_m identical to the oracle
    desc = nb.brunel_delta(N=5000, K=500, T=1000.0, seed=143, n_vp=4, drive_seed=2024)
    k, pops, sr = nb.build_gpu(desc)
    o, opops = nb.build_oracle(desc, threads=8)
    k.Simulate(1000.0)
    o.simulate(1000.0)
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    print("Brunel full run: gpu %d spikes, oracle %d" % (len(ev["senders"]), len(og)), flush=True)
    assert len(og) > 10000
    assert np.array_equal(ev["senders"], og)
    assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9)
    assert np.array_equal(np.atleast_1d(pops[0].get("V_m")), o.get(opops[0], "V_m"))


def test_exponential_network_full_size_against_oracle():
    # the benchmark network itself (100,000 iaf_psc_exp, 1.25e8 synapses, 8 kHz seeded drive)
    # for 200 ms: every spike and every V_m identical to the oracle
    T = 200.0
    k = nestgpu.Kernel(0)
    pop, sg = bench.build_c2(k, nestgpu, 100000, 1000, 250, 8, 20241, 777)
    sr = k.Create("spike_recorder")
    k.Connect(pop, sr)
    k.Simulate(T)
    ev = sr.get("events")
    print("gpu: %d spikes" % len(ev["senders"]), flush=True)
    NE = 80000
    o = Oracle(n_vp=8, seed=20241, threads=16)
    g = o.create("iaf_psc_exp", 100000, bench.NEURON)
    rng = np.random.default_rng(777)
    o.set_values(g, "V_m", rng.uniform(-70.0, -55.0, 100000))
    delay = ("sum", ("product", ("uniform_int", 21), 0.1), 1.0)
    o.connect(g[:NE], g, "fixed_indegree", 1000, weight=("normal", 87.8, 8.8), delay=delay)
    o.connect(g[NE:], g, "fixed_indegree", 250, weight=("product", ("normal", 87.8, 8.8), -4.0), delay=delay)
    counts = nestgpu.poisson_counts(777, 8000.0 * 0.1e-3, sg.tolist()[0], 100000, 0, int(T / 0.1))
    o.add_drive(g[0], counts, 87.8, 1.0)
    o.simulate(T)
    og, os_ = nb.sorted_spikes_oracle(o)
    print("oracle: %d spikes" % len(og), flush=True)
    assert len(og) > 100000
    assert np.array_equal(ev["senders"], og)
    assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9)
    assert np.array_equal(np.atleast_1d(pop.get("V_m")), o.get(g, "V_m"))
