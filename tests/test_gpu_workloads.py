"""Device path vs oracle on the benchmark workloads themselves.

- the 8-population circuit (bench.build_c3: pairwise_bernoulli with the seeded 8x8
  probability matrix, parameter objects shared across Connect calls, normal delays clipped at
  dt so that min_delay is one step, per-population Poisson background), scaled to about 7,700
  neurons against the oracle, and at its full size (77,169 neurons, about 3.2e8 synapses)
  against the oracle over a short run and across device layouts;
- batched-sweep instances (bench.c4_specs, N in [1k, 5k], iaf_psc_exp and aeif_cond_exp)
  advanced together in one batch, each against the oracle;
- a served mix of such instances (continuous batching within an HBM budget, 1-4 Simulate
  calls per request), each against the oracle.

iaf_psc_exp: identical spikes and bit-identical state. aeif_cond_exp: identical senders,
spike times within one dt, states within 1e-6 relative (the oracle runs the same portable
exp/log as the device).
"""

import ctypes
import os
import sys
import time

import numpy as np
import pytest

import nestgpu
import netbuild as nb
from _oracle import Oracle

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
EXP_FIELDS = ("V_m", "I_syn_ex", "I_syn_in")
AEIF_FIELDS = ("V_m", "g_ex", "g_in", "w")


def _spikes_oracle(o):
    g, s = o.spikes()
    order = np.lexsort((g, s))
    return g[order], s[order]


def _check(ev, pops, o, opops, model, label):
    og, os_ = _spikes_oracle(o)
    print("%s: spikes gpu=%d oracle=%d" % (label, len(ev["senders"]), len(og)), flush=True)
    assert len(og) > 0
    assert np.array_equal(ev["senders"], og), label
    if model == "iaf_psc_exp":
        assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9), label
        for p, op in zip(pops, opops):
            for f in EXP_FIELDS:
                a, b = np.atleast_1d(p.get(f)), o.get(op, f)
                assert np.array_equal(a, b), "%s: %s differs, max abs diff %g" % (label, f, np.max(np.abs(a - b)))
    else:
        assert np.max(np.abs(ev["times"] - os_ * 0.1), initial=0.0) <= 0.1 + 1e-9, label
        for p, op in zip(pops, opops):
            for f in AEIF_FIELDS:
                a, b = np.atleast_1d(p.get(f)), o.get(op, f)
                err = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12))
                assert err <= 1e-6, "%s: %s max rel diff %g" % (label, f, err)


def _c3_gpu(scale, n_vp, layout=None):
    k = nestgpu.Kernel(0)
    if layout:
        k.SetKernelStatus(layout)
    pops, n, _ = bench.build_c3(k, nestgpu, n_vp, 20241, 777, scale=scale)
    sr = k.Create("spike_recorder")
    allp = pops[0]
    for p in pops[1:]:
        allp = allp + p
    k.Connect(allp, sr)
    return k, pops, sr


@pytest.mark.parametrize("layout", [None, {"delivery_mode": "global", "targets_per_block": 64},
                                    {"delivery_mode": "cache"}], ids=["auto", "global", "cache"])
def test_circuit_scaled_against_oracle(layout):
    T, n_vp = 100.0, 4
    k, pops, sr = _c3_gpu(0.1, n_vp, layout)
    k.Simulate(T)
    print("circuit scale 0.1: %d neurons, %d connections, min_delay %d steps, window %s, tb %d" % (
        sum(len(p) for p in pops), k.GetKernelStatus("num_connections"), k.GetKernelStatus("min_delay_steps"),
        k.GetKernelStatus("window_delivery"), k.GetKernelStatus("targets_per_block")), flush=True)
    assert k.GetKernelStatus("min_delay_steps") == 1
    o = Oracle(n_vp=n_vp, seed=20241, threads=THREADS)
    opops = bench.build_c3_oracle(o, nestgpu, 777, int(round(T / 0.1)), scale=0.1)
    o.simulate(T)
    _check(sr.get("events"), pops, o, opops, "iaf_psc_exp", "circuit scale 0.1")


def test_circuit_full_size_against_oracle():
    # the benchmark circuit itself (77,169 neurons, about 3.2e8 synapses, 8 VPs) over 20 ms
    T, n_vp = 20.0, 8
    t0 = time.time()
    k, pops, sr = _c3_gpu(1.0, n_vp, {"keep_host_connections": False})
    k.Simulate(T)
    print("device: %d connections, %.1f s" % (k.GetKernelStatus("num_connections"), time.time() - t0), flush=True)
    t0 = time.time()
    o = Oracle(n_vp=n_vp, seed=20241, threads=THREADS)
    opops = bench.build_c3_oracle(o, nestgpu, 777, int(round(T / 0.1)))
    o.simulate(T)
    print("oracle: %.1f s" % (time.time() - t0), flush=True)
    # the device also counts the generator connections, one per neuron
    n_rec = o.L.or_num_connections(ctypes.c_void_p(o.k))
    assert k.GetKernelStatus("num_connections") == n_rec + sum(len(p) for p in pops)
    _check(sr.get("events"), pops, o, opops, "iaf_psc_exp", "circuit full size")


def test_circuit_full_size_layouts_agree():
    # full size over 100 ms: the automatic layout (LDS window), global delivery in blocks of
    # 128 and cache delivery give identical spikes and bit-identical states
    T = 100.0
    res = []
    for layout in ({"keep_host_connections": False},
                   {"keep_host_connections": False, "delivery_mode": "global", "targets_per_block": 128},
                   {"keep_host_connections": False, "delivery_mode": "cache"}):
        k, pops, sr = _c3_gpu(1.0, 8, layout)
        k.Simulate(T)
        ev = sr.get("events")
        res.append([ev["senders"], ev["times"]] + [np.concatenate([np.atleast_1d(p.get(f)) for p in pops])
                                                    for f in EXP_FIELDS])
        print(layout, "spikes", len(ev["senders"]), "window", k.GetKernelStatus("window_delivery"), flush=True)
        del k, pops, sr
    assert len(res[0][0]) > 0
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert np.array_equal(a, b)


def _instance_gpu(sp):
    k = nestgpu.Kernel(0)
    pop = bench.build_instance(k, nestgpu, sp, n_vp=2)
    sr = k.Create("spike_recorder")
    k.Connect(pop, sr)
    return k, pop, sr


def test_batched_sweep_instances_against_oracle():
    specs = bench.c4_specs(6, seed=2024, n_min=1000, n_max=5000)
    assert {sp["model"] for sp in specs} == {"iaf_psc_exp", "aeif_cond_exp"}
    built = [_instance_gpu(sp) for sp in specs]
    times = [60.0, 45.5, 60.0, 30.0, 52.1, 60.0]
    nestgpu.batch_simulate([b[0] for b in built], times)
    for sp, (k, pop, sr), T in zip(specs, built, times):
        o, g = nb.build_instance_oracle(sp, 2, int(round(T / 0.1)), threads=THREADS)
        o.simulate(T)
        _check(sr.get("events"), [pop], o, [g], sp["model"], "%s N=%d" % (sp["model"], sp["N"]))


def test_served_mix_against_oracle():
    specs = bench.c4_specs(6, seed=303, n_min=1000, n_max=5000)
    rng = np.random.default_rng(8)
    calls = [[float(x) for x in np.round(rng.uniform(5.0, 30.0, size=rng.integers(1, 5)), 1)] for _ in specs]
    built = [_instance_gpu(sp) for sp in specs]
    est = [k.GetKernelStatus("device_bytes_estimate") for k, _, _ in built]
    srv = nestgpu.Server(0, hbm_budget_bytes=max(max(est), sum(est) // 2), slices_per_chunk=8)
    ids = []
    for i, (k, _, _) in enumerate(built):
        ids.append(srv.submit(k, calls[i]))
        time.sleep(0.01 * (i % 3))
    infos = [srv.wait(r, timeout=120.0) for r in ids]
    print("server", srv.stats(), flush=True)
    srv.close()
    for sp, (k, pop, sr), c, info in zip(specs, built, calls, infos):
        assert info["state"] == nestgpu.Server.DONE
        o, g = nb.build_instance_oracle(sp, 2, int(round(sum(c) / 0.1)) + 1, threads=THREADS)
        for t in c:
            o.simulate(t)
        _check(sr.get("events"), [pop], o, [g], sp["model"], "served %s N=%d calls %s" % (sp["model"], sp["N"], c))
