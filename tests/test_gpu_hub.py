"""Hub targets: thousands of events to one target within one round of delivery.

A round buckets up to 1024 (window) or 2048 (global) events by target and ranks each event in
its bucket by (cell, key); a hub target's bucket can hold the whole round. Three hubs receive
every spike of 3,000 synchronously driven sources, so each volley brings 3,000 events per hub
into one cell (equal delays) or a few cells (uniform delays), spread over several rounds. The
sums must keep NEST's receive order (weights of mixed sign and magnitude make the order
visible), so V_m, I_syn_ex and I_syn_in of the hubs must be bit-identical to the oracle's.
"""

import numpy as np
import pytest

import netbuild as nb

pytestmark = pytest.mark.gpu

LAYOUTS = {"auto": {}, "glob32": {"delivery_mode": "global", "targets_per_block": 32}}


def hub_net(delay, T=60.0, n_src=3000, n_hub=3, seed=5):
    rng = np.random.default_rng(seed)
    steps = int(round(T / 0.1))
    src = {"tau_m": 10.0, "C_m": 250.0, "V_th": -55.0, "V_reset": -70.0, "E_L": -70.0, "t_ref": 2.0,
           "tau_syn_ex": 0.5, "tau_syn_in": 0.5}
    hub = dict(src, V_th=1e6)  # the hubs integrate without firing
    counts = np.zeros((steps, n_src), np.uint8)
    counts[::50, :] = 1  # a volley every 5 ms: every source fires in the same step
    return dict(
        n_vp=2, seed=seed,
        pops=[("iaf_psc_exp", n_src, src), ("iaf_psc_exp", n_hub, hub)],
        v0={0: rng.uniform(-70, -60, n_src)},
        conns=[((0, 0, n_src), (1, 0, n_hub), "all_to_all", 0, ("normal", 0.0, 1e3), delay)],
        drive=[(0, counts, 20000.0, 0.1)],
        T=T)


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("delay", [1.0, ("uniform", 1.0, 2.0)], ids=["one_cell", "spread"])
def test_hub_targets_keep_receive_order(layout, delay):
    desc = hub_net(delay)
    k, pops, sr = nb.build_gpu(desc, layout=LAYOUTS[layout])
    o, opops = nb.build_oracle(desc)
    k.Simulate(desc["T"])
    o.simulate(desc["T"])
    ev = sr.get("events")
    og, os_ = nb.sorted_spikes_oracle(o)
    n_hub_events = k.GetKernelStatus("synaptic_events")
    print("spikes gpu=%d oracle=%d, events %d, slice kernel %.3f ms" % (
        len(ev["senders"]), len(og), n_hub_events, k.GetKernelStatus("time_simulate_kernel_ms")), flush=True)
    assert len(og) >= 3000 * 10
    assert np.array_equal(ev["senders"], og)
    assert np.allclose(ev["times"], os_ * 0.1, rtol=0, atol=1e-9)
    for f in ("V_m", "I_syn_ex", "I_syn_in"):
        a, b = np.atleast_1d(pops[1].get(f)), o.get(opops[1], f)
        assert np.array_equal(a, b), "%s: %r vs %r" % (f, a, b)
    assert np.all(np.abs(o.get(opops[1], "I_syn_ex")) > 0)
