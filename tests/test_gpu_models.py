"""A neuron model added through the model plug-in surface: iaf_psc_alpha.

iaf_psc_alpha (models/iaf_psc_alpha.cpp) joins the library as a NodeModel (host half:
defaults, set/get status, recordables, pre_run_hook with IAFPropagatorAlpha) registered with
ModelManager, and a DeviceNode (update over the lags of a slice, handle( SpikeEvent& )) in the
slice kernel's model list. Its networks must match the oracle's restatement bit for bit:
spikes, V_m, I_syn_ex and I_syn_in, on every delivery layout, alone, with the other models in
one network, in split runs and through the multimeter.
"""

import numpy as np
import pytest

import nestgpu
import netbuild as nb

pytestmark = pytest.mark.gpu

LAYOUTS = {"auto": {}, "win48": {"delivery_mode": "window", "targets_per_block": 48},
           "win2": {"delivery_mode": "window", "window_slots": 2},
           "glob32": {"delivery_mode": "global", "targets_per_block": 32},
           "glob64p3": {"delivery_mode": "global", "targets_per_block": 64, "delivery_passes": 3},
           "cache": {"delivery_mode": "cache"}}
FIELDS = ("V_m", "I_syn_ex", "I_syn_in")


def alpha_random(N=2000, KE=100, KI=25, T=200.0, seed=23, n_vp=3, drive_seed=29, lam=0.8):
    """iaf_psc_alpha random network (the exponential network's sizes, alpha currents of
    0.5 ms), normal weights, delays 1.0-3.0 ms."""
    NE = int(0.8 * N)
    params = {"tau_m": 10.0, "tau_syn_ex": 0.5, "tau_syn_in": 0.5, "C_m": 250.0, "V_th": -55.0, "V_reset": -70.0,
              "t_ref": 2.0, "E_L": -70.0, "V_min": -80.0}
    rng = np.random.default_rng(drive_seed)
    steps = int(round(T / 0.1))
    delay = ("sum", 1.0, ("product", ("uniform_int", 21), 0.1))
    return dict(
        n_vp=n_vp, seed=seed,
        pops=[("iaf_psc_alpha", N, params)],
        v0={0: rng.uniform(-70, -55, N)},
        conns=[((0, 0, NE), (0, 0, N), "fixed_indegree", KE, ("normal", 40.0, 4.0), delay),
               ((0, NE, N), (0, 0, N), "fixed_indegree", KI, ("product", ("normal", 40.0, 4.0), -4.0), delay)],
        drive=[(0, rng.poisson(lam, size=(steps, N)).astype(np.uint8), 40.0, 1.0)],
        T=T)


def _check(pops, sr, o, opops, fields_of):
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    print("spikes gpu=%d oracle=%d" % (len(ev["senders"]), len(og)), flush=True)
    assert len(og) > 0
    assert np.array_equal(ev["senders"], og)
    assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9)
    for p, op, fields in zip(pops, opops, fields_of):
        for f in fields:
            a, b = np.atleast_1d(p.get(f)), o.get(op, f)
            assert np.array_equal(a, b), "%s differs: max abs diff %g" % (f, np.max(np.abs(a - b)))


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_alpha_network(layout):
    desc = alpha_random()
    k, pops, sr = nb.build_gpu(desc, layout=LAYOUTS[layout])
    o, opops = nb.build_oracle(desc)
    k.Simulate(desc["T"])
    o.simulate(desc["T"])
    print("layout=%s window=%s tb=%d" % (layout, k.GetKernelStatus("window_delivery"),
                                         k.GetKernelStatus("targets_per_block")), flush=True)
    _check(pops, sr, o, opops, [FIELDS])


def test_alpha_split_runs():
    desc = alpha_random(N=1000, T=150.0, n_vp=2, seed=4)
    k, pops, sr = nb.build_gpu(desc)
    o, opops = nb.build_oracle(desc)
    for t in (33.3, 0.1, 66.6, 50.0):
        k.Simulate(t)
        o.simulate(t)
    _check(pops, sr, o, opops, [FIELDS])


@pytest.mark.parametrize("layout", ["auto", "glob32"])
def test_four_models_in_one_network(layout):
    # the mixed network with an iaf_psc_alpha population fed by and feeding the others: one
    # launch dispatches four models per neuron
    desc = nb.mixed_models(T=100.0)
    rng = np.random.default_rng(12)
    na = 300
    steps = 1000
    desc["pops"].append(("iaf_psc_alpha", na, {"tau_syn_ex": 1.0, "tau_syn_in": 3.0, "t_ref": 1.5}))
    desc["v0"][3] = rng.uniform(-70, -55, na)
    desc["conns"] += [((1, 0, 400), (3, 0, na), "fixed_indegree", 30, ("normal", 60.0, 6.0), ("uniform", 1.0, 2.0)),
                      ((3, 0, na), (1, 0, 400), "fixed_indegree", 10, -50.0, 1.3),
                      ((3, 0, na), (3, 0, na), "pairwise_bernoulli", 0.05, 20.0, 1.0)]
    desc["drive"].append((3, rng.poisson(1.0, size=(steps, na)).astype(np.uint8), 70.0, 1.0))
    k, pops, sr = nb.build_gpu(desc, layout=LAYOUTS[layout])
    o, opops = nb.build_oracle(desc, portable_math=True)
    k.Simulate(desc["T"])
    o.simulate(desc["T"])
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    print("spikes gpu=%d oracle=%d" % (len(ev["senders"]), len(og)), flush=True)
    assert np.array_equal(ev["senders"], og)
    assert np.max(np.abs(ev["times"] - os_ * 0.1), initial=0.0) <= 0.1 + 1e-9
    for i, fields in ((0, ("V_m",)), (1, FIELDS), (3, FIELDS)):
        for f in fields:
            assert np.array_equal(np.atleast_1d(pops[i].get(f)), o.get(opops[i], f)), (i, f)
    for f in ("V_m", "w"):
        a, b = np.atleast_1d(pops[2].get(f)), o.get(opops[2], f)
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12)) <= 1e-6
    assert np.count_nonzero(np.isin(ev["senders"], pops[3].global_id)) > 0


def test_alpha_multimeter():
    desc = alpha_random(N=800, T=40.0, n_vp=2, seed=8)
    k, pops, sr = nb.build_gpu(desc)
    mm = k.Create("multimeter", params={"interval": 0.5, "offset": 0.1, "record_from": list(FIELDS)})
    targets = np.arange(0, 800, 53)
    k.Connect(mm, nestgpu.NodeCollection(k, pops[0].global_id[targets]))
    o, opops = nb.build_oracle(desc)
    o.set_sampling(opops[0][targets], 5, 1)
    k.Simulate(desc["T"])
    o.simulate(desc["T"])
    ev = mm.get("events")
    g, st, f = o.samples()
    print("samples device %d oracle %d" % (len(ev["senders"]), len(g)), flush=True)
    assert len(g) == len(targets) * 80
    assert np.array_equal(ev["senders"], g)
    assert np.allclose(ev["times"], st * 0.1, rtol=0, atol=1e-9)
    for j, name in enumerate(FIELDS):
        assert np.array_equal(ev[name], f[:, j]), name


@pytest.mark.parametrize("res,I_e", [(0.1, 1000.0), (0.1, 1450.0), (0.2, 1450.0), (0.5, 1450.0), (1.0, 1450.0)])
def test_alpha_reference_traces(res, I_e):
    # testsuite/pytests/neurons/test_iaf_psc_alpha.py (test_iaf_psc_alpha_i0 and
    # test_iaf_psc_alpha_i0_refractory): the device's voltmeter trace against the reference's
    # values (rtol 1e-6, as there) and its spike time
    import test_oracle_pins as pins

    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"resolution": res})
    n = k.Create("iaf_psc_alpha", params={"I_e": I_e})
    vm = k.Create("voltmeter", params={"interval": res})
    sr = k.Create("spike_recorder")
    k.Connect(vm, n)
    k.Connect(n, sr)
    k.Simulate(8.0)
    ev = vm.get("events")
    trace = dict(zip(np.round(ev["times"], 10), ev["V_m"]))
    expected = pins.ALPHA_I0 if I_e == 1000.0 else pins.alpha_i0_refractory_expected()
    checked = 0
    for t, v in expected:
        kk = round(t / res)
        if abs(kk * res - t) < 1e-9:
            assert trace[round(kk * res, 10)] == pytest.approx(v, rel=1e-6), t
            checked += 1
    assert checked >= 6
    assert list(np.round(sr.get("events")["times"], 10))[:1] == [4.8 if I_e == 1000.0 else 3.0]


@pytest.mark.parametrize("model", ["iaf_psc_delta", "iaf_psc_exp", "aeif_cond_exp", "iaf_psc_alpha"])
def test_spike_multiplicity(model):
    # testsuite/pytests/neurons/test_neurons_handle_multiplicity.py:128-180 checks that a spike
    # of multiplicity 2 acts like two spikes. A neuron takes one spike_generator here, and one
    # generator's spike times must differ, so the restatement compares a spike of multiplicity
    # 2 and weight 1 with one spike of weight 2 (handle( SpikeEvent& ) adds weight *
    # multiplicity): equal V_m, changed by the input
    k = nestgpu.Kernel(0)
    n1 = k.Create(model)
    n2 = k.Create(model)
    g1 = k.Create("spike_generator", params={"spike_times": [5.0], "spike_multiplicities": [2]})
    g2 = k.Create("spike_generator", params={"spike_times": [5.0]})
    k.Connect(g1, n1, syn_spec={"weight": 1.0})
    k.Connect(g2, n2, syn_spec={"weight": 2.0})
    v0 = (float(np.atleast_1d(n1.get("V_m"))[0]), float(np.atleast_1d(n2.get("V_m"))[0]))
    assert v0[0] == v0[1]
    k.Simulate(8.0)
    v1, v2 = float(np.atleast_1d(n1.get("V_m"))[0]), float(np.atleast_1d(n2.get("V_m"))[0])
    print("%s: V_m %r -> %r, %r" % (model, v0[0], v1, v2), flush=True)
    assert v1 == v2
    assert v1 != v0[0]
