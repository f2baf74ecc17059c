"""On-device readout: multimeter sampling and the spike_recorder window against the oracle.

The multimeter samples the state of its targets after every step whose stamp is
offset + k * interval (UniversalDataLogger::record_data, nestkernel/universal_data_logger_impl.h:
343-375) and keeps the samples whose stamps lie in its window when they are read back
(models/multimeter.cpp:190-216). The oracle samples the same neurons at the same stamps.
iaf models: bit-identical values; aeif_cond_exp: within 1e-6 relative.
"""

import numpy as np
import pytest

import nestgpu
import netbuild as nb

pytestmark = pytest.mark.gpu


def _sampled(desc, targets, params, fields, layout=None, steps=None):
    k, pops, sr = nb.build_gpu(desc, layout=layout)
    mm = k.Create("multimeter", params=dict(params, record_from=list(fields)))
    ids = pops[0].global_id[targets]
    k.Connect(mm, nestgpu.NodeCollection(k, ids))
    o, opops = nb.build_oracle(desc, portable_math=True)
    res = desc.get("resolution", 0.1)
    iv = int(round(params.get("interval", 1.0) / res))
    off = int(round(params.get("offset", 0.0) / res))
    o.set_sampling(opops[0][targets], iv, off)
    for t in steps or [desc["T"]]:
        k.Simulate(t)
        o.simulate(t)
    return k, mm, o


def _oracle_events(o, fields, model):
    g, st, f = o.samples()
    col = {"V_m": 0, "I_syn_ex": 1, "I_syn_in": 2, "g_ex": 1, "g_in": 2, "w": 3}
    return g, st, {name: f[:, col[name]] for name in fields}


@pytest.mark.parametrize("layout", [None, {"delivery_mode": "global", "targets_per_block": 32}])
def test_multimeter_exp_network_against_oracle(layout):
    desc = nb.exp_random(N=1200, KE=80, KI=20, T=60.0, seed=17, n_vp=2)
    fields = ("V_m", "I_syn_ex", "I_syn_in")
    targets = np.arange(0, 1200, 37)
    k, mm, o = _sampled(desc, targets, {"interval": 0.3, "offset": 0.2}, fields, layout, steps=[25.1, 34.9])
    ev = mm.get("events")
    g, st, vals = _oracle_events(o, fields, "iaf_psc_exp")
    print("samples device %d oracle %d" % (len(ev["senders"]), len(g)), flush=True)
    assert len(g) == len(targets) * len(np.arange(2, 600, 3))
    assert np.array_equal(ev["senders"], g)
    assert np.allclose(ev["times"], st * 0.1, rtol=0, atol=1e-9)
    for f in fields:
        assert np.array_equal(ev[f], vals[f]), f


def test_multimeter_delta_network_with_window():
    desc = nb.brunel_delta(N=500, K=50, T=80.0, seed=3, n_vp=3)
    targets = np.arange(0, 500, 11)
    k, mm, o = _sampled(desc, targets, {"interval": 1.0, "start": 20.0, "stop": 50.0}, ("V_m",))
    ev = mm.get("events")
    g, st, vals = _oracle_events(o, ("V_m",), "iaf_psc_delta")
    keep = (st > 200) & (st <= 500)  # the window (start, stop] = (20, 50] ms
    assert np.array_equal(ev["senders"], g[keep])
    assert np.allclose(ev["times"], st[keep] * 0.1, rtol=0, atol=1e-9)
    assert np.array_equal(ev["V_m"], vals["V_m"][keep])
    assert mm.get("n_events") == keep.sum()


def test_multimeter_aeif_network_against_oracle():
    desc = nb.aeif_random(N=600, KE=60, KI=15, T=40.0, seed=8)
    fields = ("V_m", "g_ex", "g_in", "w")
    targets = np.arange(0, 600, 23)
    k, mm, o = _sampled(desc, targets, {"interval": 0.5}, fields)
    ev = mm.get("events")
    g, st, vals = _oracle_events(o, fields, "aeif_cond_exp")
    assert np.array_equal(ev["senders"], g)
    assert np.allclose(ev["times"], st * 0.1, rtol=0, atol=1e-9)
    for f in fields:
        err = np.max(np.abs(ev[f] - vals[f]) / np.maximum(np.abs(vals[f]), 1e-12))
        assert err <= 1e-6, (f, err)


def test_multimeter_reference_offset_cases():
    # testsuite/pytests/recording/test_multimeter_offset.py, on the device path
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"resolution": 0.125})
    nrn = k.Create("iaf_psc_exp")
    mm = k.Create("multimeter", params={"interval": 3.0, "offset": 5.0, "record_from": ["V_m"]})
    k.Connect(mm, nrn)
    k.Simulate(15.0)
    assert list(mm.get("events")["times"]) == [5.0, 8.0, 11.0, 14.0]
    # test_offset_wrt_origin_start_stop: a new origin shifts the window for later samples only
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"resolution": 0.125})
    nrn = k.Create("iaf_psc_exp")
    mm = k.Create("multimeter", params={"start": 3.0, "stop": 15.0, "interval": 3.0, "offset": 5.0,
                                        "record_from": ["V_m"]})
    k.Connect(mm, nrn)
    k.Simulate(20.0)
    assert list(mm.get("events")["times"]) == [5.0, 8.0, 11.0, 14.0]
    mm.set(origin=20.0)
    k.Simulate(20.0)
    assert list(mm.get("events")["times"]) == [5.0, 8.0, 11.0, 14.0, 26.0, 29.0, 32.0, 35.0]
    # test_correct_data_logger_initialization: a multimeter connected after a first run
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"resolution": 0.125})
    nrn = k.Create("iaf_psc_exp")  # the reference test uses iaf_psc_alpha
    p = {"start": 20.0, "stop": 30.0, "interval": 3.0, "offset": 5.0, "record_from": ["V_m"]}
    mm1 = k.Create("multimeter", params=p)
    k.Connect(mm1, nrn)
    k.Simulate(10.0)
    mm2 = k.Create("multimeter", params=p)
    k.Connect(mm2, nrn)
    k.Simulate(20.0)
    assert list(mm1.get("events")["times"]) == [23.0, 26.0, 29.0]
    assert list(mm2.get("events")["times"]) == [23.0, 26.0, 29.0]


def test_spike_recorder_window_against_oracle():
    # spike_recorder start/stop/origin (nestkernel/device.cpp:83-118): spikes with stamps in
    # (origin + start, origin + stop] are kept; a new origin applies to later spikes only
    desc = nb.exp_random(N=1000, KE=80, KI=20, T=60.0, seed=23, n_vp=2)
    k, pops, sr = nb.build_gpu(desc)
    sr.set(start=10.0, stop=25.0)
    o, opops = nb.build_oracle(desc)
    k.Simulate(30.0)
    o.simulate(30.0)
    sr.set(origin=30.0)
    k.Simulate(30.0)
    o.simulate(30.0)
    ev = sr.get("events")
    g, s = nb.sorted_spikes_oracle(o)
    keep = ((s > 100) & (s <= 250)) | ((s > 400) & (s <= 550))
    print("kept %d of %d spikes" % (keep.sum(), len(g)), flush=True)
    assert keep.sum() > 0
    assert np.array_equal(ev["senders"], g[keep])
    assert np.allclose(ev["times"], s[keep] * 0.1, rtol=0, atol=1e-9)
    assert sr.get("n_events") == keep.sum()
