"""Builds one network description on the device path (nestgpu) and on the oracle.

A description is a dict:
  n_vp, seed, resolution
  pops:  [(model, n, params)]                  node ids are assigned in this order
  v0:    {pop_index: array}                     initial V_m per neuron (optional)
  conns: [(pre, post, rule, rule_param, weight, delay, autapses, multapses, stdp)]
         pre/post are (pop_index, start, stop) slices; weight/delay use the oracle's
         tuple encoding (numbers, ("normal", m, s), ("sum", a, b), ...); stdp (optional) is
         None for static_synapse or a dict of stdp_synapse parameters (and Kplus)
  drive: [(pop_index, counts[n_steps, n], weight, delay_ms)]
"""

import numpy as np

import nestgpu
from _oracle import Oracle


PARAM_ID_BASE = 1 << 40  # JSON numbers are doubles: ids stay below 2^53


def to_param(x):
    if isinstance(x, (int, float, np.floating, np.integer)):
        return float(x)
    op = x[0]
    if op == "uniform":
        return nestgpu.Parameter({"uniform": {"min": x[1], "max": x[2]}})
    if op == "uniform_int":
        return nestgpu.Parameter({"uniform_int": {"max": x[1]}})
    # ("normal", mean, std, id): one parameter object shared by every connection spec naming
    # the id (ids offset so that they never meet the ids nestgpu.random hands out)
    if op == "normal":
        d = {"mean": x[1], "std": x[2]}
        if len(x) > 3:
            d["id"] = PARAM_ID_BASE + int(x[3])
        return nestgpu.Parameter({"normal": d})
    if op == "lognormal":
        d = {"mean": x[1], "std": x[2]}
        if len(x) > 3:
            d["id"] = PARAM_ID_BASE + int(x[3])
        return nestgpu.Parameter({"lognormal": d})
    if op == "exponential":
        return nestgpu.Parameter({"exponential": {"beta": x[1]}})
    if op in ("exp", "sin", "cos"):
        return nestgpu.Parameter({op: nestgpu.Parameter._wrap(to_param(x[1]))})
    if op == "quotient":
        return nestgpu.Parameter({"quotient": [nestgpu.Parameter._wrap(to_param(x[1])),
                                               nestgpu.Parameter._wrap(to_param(x[2]))]})
    if op == "pow":
        return nestgpu.Parameter({"pow": [nestgpu.Parameter._wrap(to_param(x[1])), float(x[2])]})
    if op == "compare":
        return nestgpu.Parameter({"compare": [nestgpu.Parameter._wrap(to_param(x[1])),
                                              nestgpu.Parameter._wrap(to_param(x[2])), int(x[3])]})
    if op == "conditional":
        return nestgpu.logic.conditional(to_param(x[1]), to_param(x[2]), to_param(x[3]))
    if op in ("sum", "product"):
        return nestgpu.Parameter({op: [nestgpu.Parameter._wrap(to_param(x[1])), nestgpu.Parameter._wrap(to_param(x[2]))]})
    if op in ("max", "min"):
        return nestgpu.Parameter({op: [nestgpu.Parameter._wrap(to_param(x[1])), float(x[2])]})
    if op == "redraw":
        return nestgpu.Parameter({"redraw": [nestgpu.Parameter._wrap(to_param(x[1])), float(x[2]), float(x[3])]})
    raise ValueError(op)


def _ids(pops_ids, sl):
    p, a, b = sl
    return pops_ids[p][a:b]


def build_gpu(desc, device=0, kernel=None, layout=None):
    """layout: extra kernel status for the device layout (delivery_mode, targets_per_block)."""
    k = kernel or nestgpu.Kernel(device)
    st = {"total_num_virtual_procs": desc.get("n_vp", 1), "rng_seed": desc.get("seed", 143202461)}
    if "rng_type" in desc:
        st["rng_type"] = desc["rng_type"]
    st.update(layout or {})
    if "resolution" in desc:
        st["resolution"] = desc["resolution"]
    k.SetKernelStatus(st)
    pops = []
    for model, n, params in desc["pops"]:
        pops.append(k.Create(model, n, params))
    for pi, v in desc.get("v0", {}).items():
        pops[pi].set(V_m=v)
    ids = [p.global_id for p in pops]
    for c in desc.get("conns", []):
        pre, post, rule, rp, w, d = c[:6]
        aut = c[6] if len(c) > 6 else True
        mult = c[7] if len(c) > 7 else True
        cs = {"rule": rule, "allow_autapses": aut, "allow_multapses": mult}
        if rule == "fixed_indegree":
            cs["indegree"] = int(rp)
        elif rule == "pairwise_bernoulli":
            cs["p"] = float(rp)
        syn = {} if len(c) < 9 or c[8] is None else dict(c[8], synapse_model="stdp_synapse")
        if w is not None:
            syn["weight"] = to_param(w)
        if d is not None:
            syn["delay"] = to_param(d)
        k.Connect(_ids(ids, pre), _ids(ids, post), cs, syn)
    for pi, counts, w, d in desc.get("drive", []):
        sg = k.Create("spike_generator", len(ids[pi]))
        k.SetSpikeCounts(sg, counts)
        k.Connect(sg, pops[pi], "one_to_one", {"weight": w, "delay": d})
    sr = k.Create("spike_recorder")
    allp = pops[0]
    for p in pops[1:]:
        allp = allp + p
    k.Connect(allp, sr)
    return k, pops, sr


def build_oracle(desc, threads=1, portable_math=False):
    o = Oracle(n_vp=desc.get("n_vp", 1), seed=desc.get("seed", 143202461), resolution=desc.get("resolution", 0.1),
               threads=threads, portable_math=portable_math, rng_type=desc.get("rng_type", "mt19937_64"))
    pops = []
    for model, n, params in desc["pops"]:
        pops.append(o.create(model, n, params))
    for pi, v in desc.get("v0", {}).items():
        o.set_values(pops[pi], "V_m", v)
    for c in desc.get("conns", []):
        pre, post, rule, rp, w, d = c[:6]
        aut = c[6] if len(c) > 6 else True
        mult = c[7] if len(c) > 7 else True
        o.connect(_ids(pops, pre), _ids(pops, post), rule, rp, weight=w, delay=d, allow_autapses=aut,
                  allow_multapses=mult, stdp=c[8] if len(c) > 8 else None)
    for pi, counts, w, d in desc.get("drive", []):
        o.add_drive(pops[pi][0], counts, w, d)
    return o, pops


def sorted_spikes_oracle(o, res=0.1):
    g, s = o.spikes()
    order = np.lexsort((g, s))
    return g[order], s[order]


def brunel_delta(N=500, frac_e=0.8, K=50, g=5.0, J=0.1, delay=1.5, eta_lam=2.0, T=200.0, seed=1, n_vp=1, drive_seed=7):
    NE = int(N * frac_e)
    NI = N - NE
    KE = int(K * frac_e)
    KI = K - KE
    params = {"C_m": 1.0, "tau_m": 20.0, "V_th": 20.0, "V_reset": 10.0, "E_L": 0.0, "t_ref": 2.0}
    rng = np.random.default_rng(drive_seed)
    steps = int(round(T / 0.1))
    return dict(
        n_vp=n_vp, seed=seed,
        pops=[("iaf_psc_delta", N, params)],
        v0={0: rng.uniform(0, 20, N)},
        conns=[((0, 0, NE), (0, 0, N), "fixed_indegree", KE, J, delay),
               ((0, NE, N), (0, 0, N), "fixed_indegree", KI, -g * J, delay)],
        drive=[(0, rng.poisson(eta_lam, size=(steps, N)).astype(np.uint8), J, delay)],
        T=T)


def exp_random(N=2000, frac_e=0.8, KE=100, KI=25, T=200.0, seed=3, n_vp=1, drive_seed=9, lam=0.8, dmin=1.0):
    NE = int(N * frac_e)
    params = {"tau_m": 10.0, "tau_syn_ex": 0.5, "tau_syn_in": 0.5, "C_m": 250.0, "V_th": -55.0, "V_reset": -70.0,
              "t_ref": 2.0, "E_L": -70.0}
    rng = np.random.default_rng(drive_seed)
    steps = int(round(T / 0.1))
    delay = ("sum", dmin, ("product", ("uniform_int", 21), 0.1))
    return dict(
        n_vp=n_vp, seed=seed,
        pops=[("iaf_psc_exp", N, params)],
        v0={0: rng.uniform(-70, -55, N)},
        conns=[((0, 0, NE), (0, 0, N), "fixed_indegree", KE, ("normal", 87.8, 8.8), delay),
               ((0, NE, N), (0, 0, N), "fixed_indegree", KI, ("product", ("normal", 87.8, 8.8), -4.0), delay)],
        drive=[(0, rng.poisson(lam, size=(steps, N)).astype(np.uint8), 87.8, dmin)],
        T=T)


AEIF_C4 = {"a": 4.0, "b": 80.5, "tau_w": 144.0, "Delta_T": 2.0, "g_L": 30.0, "gsl_error_tol": 1e-3}


def aeif_random(N=1000, KE=80, KI=20, wE=2.0, g=5.0, lam=2.0, wD=6.0, T=100.0, seed=3, n_vp=2, drive_seed=1):
    """Brunel-style aeif_cond_exp network (batched-sweep neuron parameters), conductances in nS."""
    NE = int(0.8 * N)
    rng = np.random.default_rng(drive_seed)
    steps = int(round(T / 0.1))
    return dict(
        n_vp=n_vp, seed=seed,
        pops=[("aeif_cond_exp", N, dict(AEIF_C4))],
        v0={0: rng.uniform(-70.6, -55.0, N)},
        conns=[((0, 0, NE), (0, 0, N), "fixed_indegree", KE, wE, 1.5),
               ((0, NE, N), (0, 0, N), "fixed_indegree", KI, -g * wE, 1.5)],
        drive=[(0, rng.poisson(lam, size=(steps, N)).astype(np.uint8), wD, 1.5)],
        T=T)


def mixed_models(T=100.0, seed=12, n_vp=3, drive_seed=5):
    """iaf_psc_delta, iaf_psc_exp and aeif_cond_exp populations connected to each other."""
    rng = np.random.default_rng(drive_seed)
    steps = int(round(T / 0.1))
    nd, ne, na = 300, 400, 200
    delta = {"C_m": 1.0, "tau_m": 20.0, "V_th": 20.0, "V_reset": 10.0, "E_L": 0.0, "t_ref": 2.0}
    exp = {"tau_m": 10.0, "tau_syn_ex": 0.5, "tau_syn_in": 0.5, "C_m": 250.0, "V_th": -55.0, "V_reset": -70.0,
           "t_ref": 2.0, "E_L": -70.0}
    return dict(
        n_vp=n_vp, seed=seed,
        pops=[("iaf_psc_delta", nd, delta), ("iaf_psc_exp", ne, exp), ("aeif_cond_exp", na, dict(AEIF_C4))],
        v0={0: rng.uniform(0, 20, nd), 1: rng.uniform(-70, -55, ne), 2: rng.uniform(-70.6, -55, na)},
        conns=[((0, 0, nd), (0, 0, nd), "fixed_indegree", 30, ("normal", 0.15, 0.05), 1.5),
               ((0, 0, nd), (1, 0, ne), "fixed_indegree", 20, 60.0, ("uniform", 1.0, 2.0)),
               ((1, 0, ne), (2, 0, na), "pairwise_bernoulli", 0.05, ("normal", 1.0, 0.3), 1.2),
               ((2, 0, na), (0, 0, nd), "fixed_indegree", 10, -0.3, 1.0),
               ((2, 0, na), (1, 0, ne), "fixed_indegree", 10, -150.0, 1.0),
               ((1, 0, ne), (1, 0, ne), "fixed_indegree", 40, ("normal", 87.8, 8.8), ("uniform", 1.0, 3.0))],
        drive=[(0, rng.poisson(2.0, size=(steps, nd)).astype(np.uint8), 0.1, 1.0),
               (1, rng.poisson(0.8, size=(steps, ne)).astype(np.uint8), 87.8, 1.0),
               (2, rng.poisson(2.0, size=(steps, na)).astype(np.uint8), 6.0, 1.0)],
        T=T)


def build_instance_oracle(sp, n_vp, n_steps, portable_math=True, threads=4):
    """A batched-sweep instance (bench.build_instance) on the oracle: the same connections
    from the same per-VP streams and the library's seeded Poisson counts as dense drive for
    emission steps [0, n_steps) (the generators are created after the N neurons: node ids
    N+1 .. 2N)."""
    import bench

    ip = bench.instance_params(sp)
    N, K = sp["N"], sp["K"]
    KE = int(0.8 * K)
    KI = K - KE
    NE = int(0.8 * N)
    o = Oracle(n_vp=n_vp, seed=sp["seed"], threads=threads, portable_math=portable_math)
    rng = np.random.default_rng(sp["drive_seed"])
    g = o.create(sp["model"], N, ip["params"])
    o.set_values(g, "V_m", rng.uniform(*ip["v_range"], N))
    o.connect(g[:NE], g, "fixed_indegree", KE, weight=ip["J"], delay=1.5)
    o.connect(g[NE:], g, "fixed_indegree", KI, weight=-sp["g"] * ip["J"], delay=1.5)
    lam = ip["rate_hz"] * 0.1 * 1e-3
    o.add_drive(g[0], nestgpu.poisson_counts(sp["drive_seed"], lam, N + 1, N, 0, n_steps), ip["w_drive"], 1.5)
    return o, g
