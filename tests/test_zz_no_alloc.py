"""The batched run that allocates nothing must give exactly what nestgpu_batch_simulate gives.

nestgpu_batch_simulate_noalloc (include/nestgpu.h) runs the same slice kernels as
nestgpu_batch_simulate, with its launch records and status words in a fixed device array of
the library. Each case builds every network twice from the same seeds (twin kernels): one
set advanced through nestgpu_batch_simulate, the other through the variant, called through
the C ABI. Spike trains and every state variable must be bit-identical (iaf and aeif alike:
the same device code runs). Two calls in a row, the second with fewer and smaller networks,
check that nothing left in the fixed array by the first call is used by the second.
"""

import ctypes
import os
import sys

import numpy as np
import pytest

import nestgpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (instance builder of the batched sweep)

pytestmark = pytest.mark.gpu


def _build(specs):
    ks, pops, recs = [], [], []
    for sp in specs:
        k = nestgpu.Kernel(0)
        pop = bench.build_instance(k, nestgpu, sp, n_vp=2)
        sr = k.Create("spike_recorder")
        k.Connect(pop, sr)
        k.Prepare()
        ks.append(k)
        pops.append(pop)
        recs.append(sr)
    return ks, pops, recs


def _noalloc(kernels, times):
    """Direct C-ABI call of nestgpu_batch_simulate_noalloc."""
    lib = nestgpu._lib
    arr = (ctypes.c_void_p * len(kernels))(*[k._h for k in kernels])
    t = np.ascontiguousarray(np.asarray(times, dtype=np.float64))
    rc = lib.nestgpu_batch_simulate_noalloc(arr, len(kernels), t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, lib.nestgpu_last_error(None).decode()


def _state(pop, rec, model):
    ev = rec.get("events")
    keys = ["V_m", "I_syn_ex", "I_syn_in"] if model == "iaf_psc_exp" else ["V_m", "g_ex", "g_in", "w"]
    return [ev["senders"], ev["times"]] + [np.atleast_1d(pop.get(k)) for k in keys]


def _check_case(specs, times):
    print("case", [(s["model"], s["N"]) for s in specs], list(times), flush=True)
    ka, pa, ra = _build(specs)
    kb, pb, rb = _build(specs)
    nestgpu.batch_simulate(ka, times)
    _noalloc(kb, times)
    n_spikes = 0
    for i, sp in enumerate(specs):
        a = _state(pa[i], ra[i], sp["model"])
        b = _state(pb[i], rb[i], sp["model"])
        for x, y in zip(a, b):
            assert np.array_equal(x, y), "instance %d (%s, N=%d) differs" % (i, sp["model"], sp["N"])
        n_spikes += len(a[0])
    assert n_spikes > 0
    return ka, kb


def test_noalloc_equals_batch_simulate_two_calls():
    specs = bench.c4_specs(6, 919, n_min=600, n_max=3000)
    big = sorted(specs, key=lambda s: -s["N"])
    # first call: four networks, including the two largest
    _check_case(big[:4], [60.0, 41.3, 75.0, 50.0])
    # second call right after: one smaller network, a different length
    _check_case(big[-1:], [33.3])


def test_noalloc_continues_resident_state():
    """Two noalloc calls on the same kernels equal two batch_simulate calls (state, ring
    buffers and pending spikes stay resident between calls, as in Run)."""
    specs = bench.c4_specs(3, 4711, n_min=800, n_max=2000)
    ka, pa, ra = _build(specs)
    kb, pb, rb = _build(specs)
    for t in ([40.0, 40.0, 40.0], [12.5, 7.1, 20.0]):
        nestgpu.batch_simulate(ka, t)
        _noalloc(kb, t)
    for i, sp in enumerate(specs):
        for x, y in zip(_state(pa[i], ra[i], sp["model"]), _state(pb[i], rb[i], sp["model"])):
            assert np.array_equal(x, y)


def test_noalloc_rejects_unprepared_and_oversized():
    k = nestgpu.Kernel(0)
    pop = k.Create("iaf_psc_exp", 10)
    k.Connect(pop, pop, "all_to_all", {"weight": 1.0, "delay": 1.0})
    with pytest.raises(AssertionError, match="prepared"):
        _noalloc([k], [10.0])
    k.Prepare()
    with pytest.raises(AssertionError, match="at most 64"):
        _noalloc([k] * 65, [10.0] * 65)
