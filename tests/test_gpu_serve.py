"""Served form: requests admitted by HBM budget and batched continuously must give exactly
the results each network has when run alone on the plain path."""

import os
import sys
import time

import numpy as np
import pytest

import nestgpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def _build(sp):
    k = nestgpu.Kernel(0)
    pop = bench.build_instance(k, nestgpu, sp, n_vp=2)
    sr = k.Create("spike_recorder")
    k.Connect(pop, sr)
    return k, pop, sr


def _result(pop, sr):
    ev = sr.get("events")
    return ev["senders"], ev["times"], np.atleast_1d(pop.get("V_m"))


def test_served_equals_alone():
    specs = bench.c4_specs(8, seed=99, n_min=500, n_max=4000)
    rng = np.random.default_rng(3)
    calls = [list(np.round(rng.uniform(5.0, 40.0, size=rng.integers(1, 4)), 1)) for _ in specs]
    built = [_build(sp) for sp in specs]
    # a budget for about half of the networks at a time: the rest queue for memory
    est = [k.GetKernelStatus("device_bytes_estimate") for k, _, _ in built]
    srv = nestgpu.Server(0, hbm_budget_bytes=max(max(est), sum(est) // 2), slices_per_chunk=8)
    ids = []
    for i, ((k, pop, sr), c) in enumerate(zip(built, calls)):
        ids.append(srv.submit(k, c))
        time.sleep(0.01 * (i % 3))  # staggered arrivals: some join while others run
    infos = [srv.wait(r, timeout=120.0) for r in ids]
    stats = srv.stats()
    print("server", stats, flush=True)
    for i, (sp, c) in enumerate(zip(specs, calls)):
        assert infos[i]["state"] == nestgpu.Server.DONE
        k, pop, sr = _build(sp)
        for t in c:
            k.Simulate(t)
        a = _result(built[i][1], built[i][2])
        b = _result(pop, sr)
        print(sp["model"], sp["N"], c, "spikes", len(a[0]), "latency %.3f s" % (infos[i]["t_done"] - infos[i]["t_submit"]),
              flush=True)
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), "request %d differs from its run alone" % i
        assert infos[i]["events"] == k.GetKernelStatus("synaptic_events")
    assert stats["done"] == len(specs)
    assert stats["max_concurrent"] >= 2
    srv.close()


def test_budget_forces_queueing():
    specs = bench.c4_specs(4, seed=7, n_min=2000, n_max=3000)
    built = [_build(sp) for sp in specs]
    # room for one network at a time: requests run strictly one after another
    est = [k.GetKernelStatus("device_bytes_estimate") for k, _, _ in built]
    two_smallest = sorted(est)[0] + sorted(est)[1]
    assert two_smallest > max(est)
    srv = nestgpu.Server(0, hbm_budget_bytes=max(est), slices_per_chunk=8)
    ids = [srv.submit(k, [20.0]) for k, _, _ in built]
    infos = [srv.wait(r, timeout=120.0) for r in ids]
    for a, b in zip(infos, infos[1:]):
        assert b["t_admit"] >= a["t_done"] - 1e-3
    srv.close()


def test_oversized_request_rejected():
    k, pop, sr = _build(bench.c4_specs(1, seed=1, n_min=5000, n_max=5001)[0])
    srv = nestgpu.Server(0, hbm_budget_bytes=1 << 20)
    with pytest.raises(nestgpu.NESTError, match="budget"):
        srv.submit(k, [10.0])
    srv.close()


def test_http_sessions_through_served_form_equal_plain():
    """Two nest-server clients (NEST-Session a and b) build networks over HTTP and call
    Simulate twice each from two threads; the served form batches their calls. Each session's
    spikes and V_m equal the same network run alone on the plain path."""
    import json
    import threading
    import urllib.request

    from nestgpu import http_api

    srv = http_api.NESTServer(port=0, slices_per_chunk=8).start()

    def call(route, payload, session):
        req = urllib.request.Request("http://127.0.0.1:%d/api/%s" % (srv.port, route), json.dumps(payload).encode())
        req.add_header("Content-Type", "application/json")
        req.add_header("NEST-Session", session)
        with urllib.request.urlopen(req, timeout=120) as r:
            return json.loads(r.read().decode())

    nets = {"a": (700, 5), "b": (1300, 9)}
    results = {}

    def client(name):
        n, seed = nets[name]
        rng = np.random.default_rng(seed)
        v0 = rng.uniform(-70.0, -55.0, n).tolist()
        call("SetKernelStatus", {"params": {"rng_seed": seed, "total_num_virtual_procs": 2}}, name)
        pop = call("Create", {"model": "iaf_psc_exp", "n": n, "params": {"I_e": 450.0}}, name)
        call("SetStatus", {"nodes": pop, "params": [{"V_m": v} for v in v0]}, name)
        call("Connect", {"pre": pop, "post": pop, "conn_spec": {"rule": "fixed_indegree", "indegree": 50},
                         "syn_spec": {"weight": {"normal": {"mean": 40.0, "std": 5.0}}, "delay": 1.5}}, name)
        sr = call("Create", {"model": "spike_recorder"}, name)
        call("Connect", {"pre": pop, "post": sr}, name)
        call("Simulate", {"t": 30.0}, name)
        call("Simulate", {"t": 25.5}, name)
        ev = call("GetStatus", {"nodes": sr, "keys": "events"}, name)[0]
        vm = call("GetStatus", {"nodes": pop, "keys": "V_m"}, name)
        results[name] = (v0, ev, vm)

    threads = [threading.Thread(target=client, args=(nm,)) for nm in nets]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    srv.stop()
    for name, (n, seed) in nets.items():
        v0, ev, vm = results[name]
        k = nestgpu.Kernel(0)
        k.SetKernelStatus({"rng_seed": seed, "total_num_virtual_procs": 2})
        pop = k.Create("iaf_psc_exp", n, {"I_e": 450.0})
        pop.set(V_m=np.array(v0))
        k.Connect(pop, pop, {"rule": "fixed_indegree", "indegree": 50},
                  {"weight": nestgpu.Parameter({"normal": {"mean": 40.0, "std": 5.0}}), "delay": 1.5})
        sr = k.Create("spike_recorder")
        k.Connect(pop, sr)
        k.Simulate(30.0)
        k.Simulate(25.5)
        e2 = sr.get("events")
        print(name, "spikes", len(ev["senders"]), flush=True)
        assert len(ev["senders"]) > 0
        assert np.array_equal(np.asarray(ev["senders"]), e2["senders"])
        assert np.array_equal(np.asarray(ev["times"]), e2["times"])
        assert np.array_equal(np.asarray(vm), pop.get("V_m"))


def test_served_plastic_and_alpha_requests():
    # requests with stdp_synapse, with iaf_psc_alpha and with static iaf_psc_exp networks share
    # the served launches (the stdp instantiation then runs every request; static ones skip
    # the plastic code): each result, plastic weights included, equals its run alone
    import netbuild as nb
    from test_gpu_models import alpha_random
    from test_gpu_stdp import stdp_net

    descs = [stdp_net(N=600, seed=51, drive_seed=3, T=60.0), nb.exp_random(N=700, T=60.0, seed=6, n_vp=2),
             alpha_random(N=700, T=60.0, seed=9, n_vp=2), stdp_net(N=900, seed=52, drive_seed=4, mu=(0.0, 1.0),
                                                                  T=60.0)]
    calls = [[20.0, 40.0], [60.0], [33.3, 26.7], [10.0, 50.0]]

    def weights(k, pops):
        c = k.GetConnections()
        ids = np.concatenate([np.atleast_1d(p.global_id) for p in pops])
        keep = np.isin(c["source"], ids) & np.isin(c["target"], ids)
        return c["weight"][keep]

    built = [nb.build_gpu(d) for d in descs]
    srv = nestgpu.Server(0, slices_per_chunk=8)
    ids = [srv.submit(k, c) for (k, _, _), c in zip(built, calls)]
    infos = [srv.wait(r, timeout=120.0) for r in ids]
    stats = srv.stats()
    print("server", stats, flush=True)
    for i, (d, c) in enumerate(zip(descs, calls)):
        assert infos[i]["state"] == nestgpu.Server.DONE, infos[i]
        k, pops, sr = nb.build_gpu(d)
        for t in c:
            k.Simulate(t)
        ks, ps, srs = built[i]
        a = (srs.get("events")["senders"], srs.get("events")["times"], np.atleast_1d(ps[0].get("V_m")),
             weights(ks, ps))
        b = (sr.get("events")["senders"], sr.get("events")["times"], np.atleast_1d(pops[0].get("V_m")),
             weights(k, pops))
        print(d["pops"][0][0], "spikes", len(a[0]), flush=True)
        assert len(a[0]) > 0
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), "request %d differs from its run alone" % i
    srv.close()
