"""Target blocks balanced by measured delivery work.

With the block size left to the library, the first Run boundary that ends a slice after at
least 64 counted slices moves the network onto blocks balanced by the events each target
received (the ring buffers and pending spike lists move with their targets). Results must not
depend on the block layout: the same network on uniform blocks (balanced_blocks off) gives
identical spikes and bit-identical state, and the scaled benchmark circuit still matches the
oracle after the move.
"""

import os
import sys

import numpy as np
import pytest

import nestgpu
from _oracle import Oracle

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
FIELDS = ("V_m", "I_syn_ex", "I_syn_in")


def _three_populations(balanced):
    """A driven population A (fires), a silent population S and a target population B:
    A -> A in-degree 60, S -> A in-degree 300 (never active), A -> B in-degree 300. In-degree
    alone rates A and B alike; the events B receives are five times those of A. Delays
    1.0-2.0 ms (min_delay 10 steps)."""
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"total_num_virtual_procs": 4, "rng_seed": 91, "balanced_blocks": balanced})
    a = k.Create("iaf_psc_exp", 3000, {"I_e": 200.0})
    s = k.Create("iaf_psc_exp", 3000)
    b = k.Create("iaf_psc_exp", 3000)
    rng = np.random.default_rng(5)
    for p in (a, s, b):
        p.set(V_m=rng.uniform(-70.0, -55.0, len(p)))
    d = nestgpu.random.uniform(1.0, 2.0)
    k.Connect(a, a, {"rule": "fixed_indegree", "indegree": 60}, {"weight": 20.0, "delay": d})
    k.Connect(s, a, {"rule": "fixed_indegree", "indegree": 300}, {"weight": 5.0, "delay": d})
    k.Connect(a, b, {"rule": "fixed_indegree", "indegree": 300}, {"weight": 2.0, "delay": d})
    gen = k.Create("spike_generator", 3000, {"poisson_rate": 12000.0, "poisson_seed": 17})
    k.Connect(gen, a, "one_to_one", {"weight": 30.0, "delay": 1.0})
    sr = k.Create("spike_recorder")
    k.Connect(a + s + b, sr)
    mm = k.Create("multimeter", params={"record_from": ["V_m"], "interval": 0.5})
    k.Connect(mm, b[:50] + a[:50])
    return k, (a, s, b), sr, mm


def _state(pops, sr, mm):
    ev = sr.get("events")
    m = mm.get("events")
    return ([ev["senders"], ev["times"]] + [np.atleast_1d(p.get(f)) for p in pops for f in FIELDS]
            + [m["senders"], m["times"], m["V_m"]])


def test_measured_balance_keeps_results_split_runs():
    # 64 slices are counted first; the first call ends inside a slice, so the move waits for the
    # second
    runs = [64.3, 9.7, 15.0, 7.3]
    res = {}
    for balanced in (True, False):
        k, pops, sr, mm = _three_populations(balanced)
        moves = []
        for t in runs:
            k.Simulate(t)
            moves.append(k.GetKernelStatus("block_relayouts"))
        print("balanced %s: relayouts after each Simulate %s, blocks %d, tb %d, spikes %d" % (
            balanced, moves, k.GetKernelStatus("target_blocks"), k.GetKernelStatus("targets_per_block"),
            k.GetKernelStatus("local_spike_counter")), flush=True)
        if balanced:
            assert moves == [0, 1, 1, 1]
        else:
            assert moves == [0, 0, 0, 0]
        res[balanced] = _state(pops, sr, mm)
    assert len(res[True][0]) > 1000
    for x, y in zip(res[True], res[False]):
        assert np.array_equal(x, y)


def test_measured_balance_circuit_against_oracle():
    # scaled benchmark circuit (min_delay one step): the move happens after the first call
    T1, T2, n_vp = 40.0, 60.0, 4
    k = nestgpu.Kernel(0)
    pops, n, _ = bench.build_c3(k, nestgpu, n_vp, 20241, 777, scale=0.1)
    allp = pops[0]
    for p in pops[1:]:
        allp = allp + p
    sr = k.Create("spike_recorder")
    k.Connect(allp, sr)
    k.Simulate(T1)
    assert k.GetKernelStatus("block_relayouts") == 1
    k.Simulate(T2)
    o = Oracle(n_vp=n_vp, seed=20241, threads=THREADS)
    opops = bench.build_c3_oracle(o, nestgpu, 777, int(round((T1 + T2) / 0.1)), scale=0.1)
    o.simulate(T1 + T2)
    g, s = o.spikes()
    order = np.lexsort((g, s))
    ev = sr.get("events")
    print("circuit scale 0.1 after the move: spikes gpu=%d oracle=%d, tb %d" % (
        len(ev["senders"]), len(g), k.GetKernelStatus("targets_per_block")), flush=True)
    assert len(g) > 0
    assert np.array_equal(ev["senders"], g[order])
    assert np.allclose(ev["times"], s[order] * 0.1, rtol=0, atol=1e-9)
    for p, op in zip(pops, opops):
        for f in FIELDS:
            assert np.array_equal(np.atleast_1d(p.get(f)), o.get(op, f)), f
