"""Batched form: instances advanced together must give exactly their plain-path results.

Every instance of a batch shares the slice launches with the others; its spike trains and
states must be bit-identical to running it alone (iaf and aeif alike: same device code).
"""

import os
import sys

import numpy as np
import pytest

import nestgpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (instance builder of the batched sweep)

pytestmark = pytest.mark.gpu


def _instances(n=6, seed=77, n_min=400, n_max=3000):
    return bench.c4_specs(n, seed, n_min=n_min, n_max=n_max)


def _build(specs, layout=None):
    ks, pops, recs = [], [], []
    for sp in specs:
        k = nestgpu.Kernel(0)
        if layout:
            k.SetKernelStatus(layout)
        pop = bench.build_instance(k, nestgpu, sp, n_vp=2)
        sr = k.Create("spike_recorder")
        k.Connect(pop, sr)
        ks.append(k)
        pops.append(pop)
        recs.append(sr)
    return ks, pops, recs


def _result(pop, rec):
    ev = rec.get("events")
    return ev["senders"], ev["times"], np.atleast_1d(pop.get("V_m"))


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_batched_equals_plain():
    specs = _instances()
    print("instances", [(s["model"], s["N"]) for s in specs], flush=True)
    kb, pb, rb = _build(specs)
    kp, pp, rp = _build(specs)
    nestgpu.batch_simulate(kb, 60.0)
    nestgpu.batch_simulate(kb, 41.3)
    for k in kp:
        k.Simulate(60.0)
        k.Simulate(41.3)
    for i in range(len(specs)):
        a, b = _result(pb[i], rb[i]), _result(pp[i], rp[i])
        print(specs[i]["model"], specs[i]["N"], "spikes", len(a[0]), flush=True)
        assert len(a[0]) > 0
        assert _same(a, b), "instance %d differs between batched and plain runs" % i


def test_batched_uneven_durations():
    # instances stop at different times inside one batch call; each matches its own plain run
    specs = _instances(4, seed=5)
    kb, pb, rb = _build(specs)
    kp, pp, rp = _build(specs)
    times = [10.0, 37.2, 0.0, 55.5]
    nestgpu.batch_simulate(kb, times)
    for k, t in zip(kp, times):
        k.Simulate(t)
    for i in range(len(specs)):
        assert _same(_result(pb[i], rb[i]), _result(pp[i], rp[i])), i


def test_autotune_leaves_results_unchanged():
    specs = _instances(6, seed=11)
    ka, pa, ra = _build(specs)
    chosen = nestgpu.batch_autotune(ka, probe_ms=5.0, candidates=(32, 128, 512))
    print("chosen", chosen, flush=True)
    assert all(c in (32, 128, 512) for c in chosen)
    kp, pp, rp = _build(specs)
    nestgpu.batch_simulate(ka, 50.0)
    for k in kp:
        k.Simulate(50.0)
    for i in range(len(specs)):
        assert ka[i].GetKernelStatus("targets_per_block") == chosen[i]
        assert _same(_result(pa[i], ra[i]), _result(pp[i], rp[i])), i


def test_batch_mixes_window_and_large_global_instances():
    # a window-delivery instance and a global one with more blocks than a window instance may
    # have (20,000 neurons in blocks of 32: 625 blocks) share the window kernel's launches
    specs = _instances(2, seed=23, n_min=2000, n_max=3000)
    specs.append(dict(specs[0], N=20000))
    layouts = [None, None, {"delivery_mode": "global", "targets_per_block": 32}]
    built = [_build([sp], lay) for sp, lay in zip(specs, layouts)]
    plain = [_build([sp], lay) for sp, lay in zip(specs, layouts)]
    kb = [b[0][0] for b in built]
    nestgpu.batch_simulate(kb, 45.0)
    for p in plain:
        p[0][0].Simulate(45.0)
    assert [k.GetKernelStatus("window_delivery") for k in kb] == [True, True, False]
    assert kb[2].GetKernelStatus("targets_per_block") == 32
    for b, p in zip(built, plain):
        a, c = _result(b[1][0], b[2][0]), _result(p[1][0], p[2][0])
        assert len(a[0]) > 0
        assert _same(a, c)
