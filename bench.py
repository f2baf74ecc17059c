"""Throughput of the MI355X simulation step (one GPU).

Default workload: the 8-population layered circuit -- 77,169 iaf_psc_exp neurons (tau_m 10 ms,
tau_syn 0.5 ms, C_m 250 pF, V_th -55 mV, V_reset -70 mV, t_ref 2 ms) in populations of 20683,
5834, 21915, 5479, 4850, 1065, 14395 and 2948 (excitatory and inhibitory alternating),
pairwise_bernoulli connections with an 8x8 probability matrix drawn from the seed (uniform in
[0, 0.14], about 3.2e8 synapses), weights normal(87.8, 8.78) pA and -4x from inhibitory
populations, delays normal(1.5, 0.75) ms (excitatory) and normal(0.75, 0.375) ms (inhibitory)
clipped at 0.1 ms, so min_delay is one step; seeded Poisson background of 8 Hz x K_ext per
neuron with K_ext drawn from [1500, 2500] per population; dt = 0.1 ms, FP64.

A step is Run(1000 ms) -- one biological second, 10,000 one-step slices -- continuing the same
network. value = recurrent synaptic events delivered per wall-clock second over the timed
steps; RTF (wall / biological time) is reported beside it. The exponential-synapse random
network (100,000 neurons, fixed_indegree 1000 E + 250 I, delays 1-3 ms) is measured after it
and reported under "c2". --workload c4 / c5 run the batched sweep and the served stream.

The roofline pass re-runs part of a step with HIP events around every slice launch on the
simulation stream and divides the algorithmic bytes by the average launch duration; the
workgroups' own clock splits the launch into delivery and the rest. The CPU baseline times
the oracle (C++/OpenMP restatement of the reference algorithm) on the same network, after its
start-up transient, on this host's cores.
"""

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nest-simulator_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "synaptic events/s + real-time factor (1 GPU); batched p50/p99 sim latency"

NEURON = {"tau_m": 10.0, "tau_syn_ex": 0.5, "tau_syn_in": 0.5, "C_m": 250.0, "V_th": -55.0, "V_reset": -70.0,
          "t_ref": 2.0, "E_L": -70.0}


def build_c2(k, nestgpu, N, KE, KI, n_vp, seed, drive_seed, rate_hz=8000.0, t_stop=1e9):
    NE = int(0.8 * N)
    k.SetKernelStatus({"total_num_virtual_procs": n_vp, "rng_seed": seed})
    pop = k.Create("iaf_psc_exp", N, NEURON)
    rng = np.random.default_rng(drive_seed)
    pop.set(V_m=rng.uniform(-70.0, -55.0, N))
    delay = 1.0 + 0.1 * nestgpu.random.uniform_int(21)
    k.Connect(pop[:NE], pop, {"rule": "fixed_indegree", "indegree": KE},
              {"weight": nestgpu.random.normal(87.8, 8.8), "delay": delay})
    k.Connect(pop[NE:], pop, {"rule": "fixed_indegree", "indegree": KI},
              {"weight": -4.0 * nestgpu.random.normal(87.8, 8.8), "delay": delay})
    sg = k.Create("spike_generator", N, {"poisson_rate": rate_hz, "poisson_seed": drive_seed, "poisson_stop": t_stop})
    k.Connect(sg, pop, "one_to_one", {"weight": 87.8, "delay": 1.0})
    return pop, sg


C3_POPS = [20683, 5834, 21915, 5479, 4850, 1065, 14395, 2948]  # E/I alternating
C3_W, C3_W_SD = 87.8, 8.78  # pA
C3_DELAY = ((1.5, 0.75), (0.75, 0.375))  # ms, (mean, sd) of excitatory and inhibitory sources
C3_DRIVE_W, C3_DRIVE_D = 87.8, 0.1


def c3_description(drive_seed, scale=1.0):
    """Draws of the 8-population circuit, in the order the builders use them: initial V_m of
    each population, the 8x8 connection probabilities (uniform in [0, 0.14]; prob[i, j] from
    population j to i) and the background in-degrees K_ext (uniform integers in [1500, 2500]).
    scale shrinks every population (at least one neuron) and keeps the probabilities."""
    rng = np.random.default_rng(drive_seed)
    sizes = [max(1, int(round(n * scale))) for n in C3_POPS]
    v0 = [rng.uniform(-70.0, -55.0, n) for n in sizes]
    prob = rng.uniform(0.0, 0.14, size=(8, 8))
    k_ext = rng.integers(1500, 2501, size=8)
    return dict(sizes=sizes, v0=v0, prob=prob, k_ext=k_ext, drive_seed=drive_seed)


# stdp_synapse parameters of the plastic variant of the circuit (excitatory to excitatory
# connections); Wmax twice the mean weight
C3_STDP = {"tau_plus": 20.0, "lambda": 0.01, "alpha": 1.0, "mu_plus": 1.0, "mu_minus": 1.0, "Wmax": 2 * C3_W}


def build_c3(k, nestgpu, n_vp, seed, drive_seed, t_stop=1e9, scale=1.0, stdp=None):
    """8-population layered circuit: 77,169 iaf_psc_exp (scale 1), pairwise_bernoulli with an 8x8
    probability matrix drawn from the seed (uniform in [0, 0.14]), weights normal(87.8, 10 %)
    and x -4 from inhibitory populations, delays normal (E 1.5 ms, I 0.75 ms, sd 50 %)
    clipped at dt, background Poisson 8 Hz x K_ext with K_ext uniform in [1500, 2500] per
    population. stdp: None, or stdp_synapse parameters for the connections between excitatory
    populations. Returns (populations, total neurons, expected drive events per step)."""
    c = c3_description(drive_seed, scale)
    k.SetKernelStatus({"total_num_virtual_procs": n_vp, "rng_seed": seed})
    pops = [k.Create("iaf_psc_exp", n, NEURON) for n in c["sizes"]]
    for p, v in zip(pops, c["v0"]):
        p.set(V_m=v)
    for j, pre in enumerate(pops):
        exc = j % 2 == 0
        w = nestgpu.random.normal(C3_W, C3_W_SD) if exc else -4.0 * nestgpu.random.normal(C3_W, C3_W_SD)
        d = nestgpu.math.max(nestgpu.random.normal(*C3_DELAY[0 if exc else 1]), 0.1)
        for i, post in enumerate(pops):
            syn = {"weight": w, "delay": d}
            if stdp is not None and exc and i % 2 == 0:
                syn.update(stdp, synapse_model="stdp_synapse")
            k.Connect(pre, post, {"rule": "pairwise_bernoulli", "p": float(c["prob"][i, j])}, syn)
    lam = 0.0
    for i, post in enumerate(pops):
        rate = 8.0 * float(c["k_ext"][i])
        sg = k.Create("spike_generator", len(post), {"poisson_rate": rate, "poisson_seed": drive_seed + i,
                                                     "poisson_stop": t_stop})
        k.Connect(sg, post, "one_to_one", {"weight": C3_DRIVE_W, "delay": C3_DRIVE_D})
        lam += len(post) * (1.0 - np.exp(-rate * 0.1e-3))
    return pops, sum(c["sizes"]), lam


def build_c3_oracle(o, nestgpu, drive_seed, n_steps, scale=1.0, stdp=None):
    """The same circuit on the oracle (connections from the same per-VP streams; the
    parameter objects that several Connect calls share carry their normal-draw state, as the
    library's do), with the library's seeded Poisson counts of every generator as dense drive
    for emission steps [0, n_steps). Returns the populations' node ids."""
    c = c3_description(drive_seed, scale)
    pops = [o.create("iaf_psc_exp", n, NEURON) for n in c["sizes"]]
    for p, v in zip(pops, c["v0"]):
        o.set_values(p, "V_m", v)
    for j, pre in enumerate(pops):
        exc = j % 2 == 0
        wn = ("normal", C3_W, C3_W_SD, 2 * j)
        w = wn if exc else ("product", wn, -4.0)
        d = ("max", ("normal",) + C3_DELAY[0 if exc else 1] + (2 * j + 1,), 0.1)
        for i, post in enumerate(pops):
            o.connect(pre, post, "pairwise_bernoulli", float(c["prob"][i, j]), weight=w, delay=d,
                      stdp=stdp if exc and i % 2 == 0 else None)
    n_total = sum(c["sizes"])
    gen0 = n_total + 1  # generators are created after every population, population by population
    for i, post in enumerate(pops):
        lam = 8.0 * float(c["k_ext"][i]) * 0.1e-3
        counts = nestgpu.poisson_counts(drive_seed + i, lam, gen0, len(post), 0, n_steps)
        o.add_drive(post[0], counts, C3_DRIVE_W, C3_DRIVE_D)
        gen0 += len(post)
    return pops


AEIF_C4 = {"a": 4.0, "b": 80.5, "tau_w": 144.0, "Delta_T": 2.0, "g_L": 30.0, "gsl_error_tol": 1e-3}


def c4_specs(n=256, seed=4242, n_min=1000, n_max=50000):
    """Batched sweep: Brunel-style instances, model iaf_psc_exp or aeif_cond_exp 50/50, N
    log-uniform in [n_min, n_max], in-degree 10 % of N capped at 2000 (80 % E), eta uniform
    [1, 3], g uniform [3, 7], delay 1.5 ms, one seed per instance."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        model = "aeif_cond_exp" if rng.random() < 0.5 else "iaf_psc_exp"
        N = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        specs.append(dict(model=model, N=N, K=min(int(0.1 * N), 2000), eta=float(rng.uniform(1.0, 3.0)),
                          g=float(rng.uniform(3.0, 7.0)), seed=int(rng.integers(1, 2 ** 31)),
                          drive_seed=int(rng.integers(1, 2 ** 31))))
    return specs


def instance_params(sp):
    """Model parameters, initial V_m range, synaptic weight J, drive weight and drive rate (Hz)
    of one batched-sweep instance. iaf_psc_exp: J = 87.8 pA x min(1, sqrt(250 / K)), drive at
    eta x nu_th x C_E (nu_th from the PSC amplitude); aeif_cond_exp: J = 6 nS x sqrt(0.8 / K),
    drive eta x 10 kHz of 6 nS. Inhibitory weights are -g J."""
    K = sp["K"]
    if sp["model"] == "iaf_psc_exp":
        J = 87.8 * min(1.0, np.sqrt(250.0 / K))
        return dict(params=NEURON, v_range=(-70.0, -55.0), J=J, w_drive=J,
                    rate_hz=sp["eta"] * 15.0 * 250.0 / (J * 0.5 * 10.0) * 1000.0)
    return dict(params=AEIF_C4, v_range=(-70.6, -55.0), J=6.0 * np.sqrt(0.8 / K), w_drive=6.0,
                rate_hz=sp["eta"] * 10000.0)


def build_instance(k, nestgpu, sp, n_vp=4, keep_host_connections=True, rng_type="mt19937_64", connect_on_device=False):
    """One batched-sweep instance: N neurons (80 % excitatory), fixed_indegree 0.8 K from the
    excitatory and 0.2 K from the inhibitory neurons, delay 1.5 ms, seeded Poisson drive through
    one spike_generator per neuron (instance_params). With rng_type Philox_32 and
    connect_on_device the two fixed_indegree calls draw on the device. Returns the population."""
    k.SetKernelStatus({"total_num_virtual_procs": n_vp, "rng_seed": sp["seed"],
                       "keep_host_connections": keep_host_connections, "rng_type": rng_type,
                       "connect_on_device": connect_on_device})
    N, K = sp["N"], sp["K"]
    KE = int(0.8 * K)
    KI = K - KE
    NE = int(0.8 * N)
    ip = instance_params(sp)
    rng = np.random.default_rng(sp["drive_seed"])
    pop = k.Create(sp["model"], N, ip["params"])
    pop.set(V_m=rng.uniform(*ip["v_range"], N))
    J = ip["J"]
    k.Connect(pop[:NE], pop, {"rule": "fixed_indegree", "indegree": KE}, {"weight": J, "delay": 1.5})
    k.Connect(pop[NE:], pop, {"rule": "fixed_indegree", "indegree": KI}, {"weight": -sp["g"] * J, "delay": 1.5})
    sg = k.Create("spike_generator", N, {"poisson_rate": ip["rate_hz"], "poisson_seed": sp["drive_seed"]})
    k.Connect(sg, pop, "one_to_one", {"weight": ip["w_drive"], "delay": 1.5})
    return pop


def run_batched(args, nestgpu):
    """Batched form: every instance advances in the same slice launches (one grid over all
    blocks of all instances, replayed from a hipGraph). value = aggregate recurrent synaptic
    events/s of the batch; every 4th instance is then built again with the library's default
    layout and run one by one on the plain path."""
    from concurrent.futures import ThreadPoolExecutor

    specs = c4_specs(args.instances, args.seed)
    t_build = time.perf_counter()
    kernels = [nestgpu.Kernel(0) for _ in specs]
    workers = max(1, min(16, (os.cpu_count() or 4) // 2))
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda a: build_instance(a[0], nestgpu, a[1], keep_host_connections=False, rng_type=args.rng_type,
                                   connect_on_device=bool(args.connect_on_device)), zip(kernels, specs)))
    n_neurons = sum(sp["N"] for sp in specs)
    n_syn = sum(k.GetKernelStatus("num_connections") for k in kernels) - n_neurons
    print("built %d instances: %d neurons, %d synapses in %.1f s" % (len(specs), n_neurons, n_syn,
                                                                    time.perf_counter() - t_build),
          file=sys.stderr, flush=True)
    chosen = None
    if args.autotune:
        t_at = time.perf_counter()
        caps = tuple(int(x) for x in args.autotune_windows.split(","))
        cands = tuple(int(x) for x in args.autotune_blocks.split(","))
        chosen = nestgpu.batch_autotune(kernels, probe_ms=args.probe_ms, candidates=cands, window_caps=caps,
                                        with_windows=True)
        print("autotune: %.1f s, (targets_per_block, window cap) per size class %s" % (
            time.perf_counter() - t_at, sorted(set(zip([sp["N"].bit_length() for sp in specs], chosen)))),
            file=sys.stderr, flush=True)
    with ThreadPoolExecutor(4) as ex:
        list(ex.map(lambda k: k.Prepare(), kernels))
    build_s = time.perf_counter() - t_build
    print("prepared in %.1f s total" % build_s, file=sys.stderr, flush=True)

    def events():
        return sum(k.GetKernelStatus("synaptic_events") for k in kernels)

    for i in range(args.warmup):
        nestgpu.batch_simulate(kernels, args.t_step)
        print("batched warmup %d done" % i, file=sys.stderr, flush=True)
    ev0 = events()
    t0 = time.perf_counter()
    for i in range(args.steps):
        nestgpu.batch_simulate(kernels, args.t_step)
        print("batched step %d done" % i, file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    ev_b = events() - ev0
    value = ev_b / wall
    rtf = wall / (args.steps * args.t_step * 1e-3)
    print("batched: %d x %g ms, %.3f s wall, RTF %.3f, %.3g events/s" % (args.steps, args.t_step, wall, rtf, value),
          file=sys.stderr, flush=True)
    # the plain path: every 4th instance, built again with the library's default layout (the
    # layout a single run would get), warmed up, then run one at a time
    kernels = None
    gc.collect()
    sub = specs[::4]
    plain = [nestgpu.Kernel(0) for _ in sub]
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda a: build_instance(a[0], nestgpu, a[1], keep_host_connections=False, rng_type=args.rng_type,
                                   connect_on_device=bool(args.connect_on_device)), zip(plain, sub)))
    for k in plain:
        k.SetKernelStatus({"graph_launch": bool(args.graph_launch)})
        k.Prepare()
        k.Run(args.plain_warm_ms)
    ev0 = sum(k.GetKernelStatus("synaptic_events") for k in plain)
    plain_wall = 0.0
    for k in plain:
        t0 = time.perf_counter()
        k.Run(args.plain_ms)
        plain_wall += time.perf_counter() - t0
    ev_p = sum(k.GetKernelStatus("synaptic_events") for k in plain) - ev0
    plain_value = ev_p / plain_wall
    print("plain one by one (%d instances, default layout, after %g ms warm-up): %g ms each, %.3f s wall, "
          "%.3g events/s; batched / plain = %.2f" % (len(plain), args.plain_warm_ms, args.plain_ms, plain_wall,
                                                    plain_value, value / plain_value), file=sys.stderr, flush=True)
    models = {m: sum(1 for sp in specs if sp["model"] == m) for m in ("iaf_psc_exp", "aeif_cond_exp")}
    line = {
        "metric": METRIC, "value": value, "unit": "synaptic events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic", "rtf": rtf,
        "config": {"workload": "batched sweep: %d Brunel-style instances (iaf_psc_exp %d, aeif_cond_exp %d), N "
                               "log-uniform [1k, 50k], in-degree min(0.1 N, 2000), delay 1.5 ms, one step = %g ms "
                               "of every instance" % (len(specs), models["iaf_psc_exp"], models["aeif_cond_exp"],
                                                       args.t_step),
                   "n_instances": len(specs), "n_neurons": n_neurons, "n_synapses": n_syn, "build_s": build_s,
                   "rng_type": args.rng_type, "connect_on_device": bool(args.connect_on_device),
                   "layout": sorted(set(chosen)) if chosen else "library default",
                   "layout_key": "(targets_per_block, cap on LDS ring slots, 0 = as many as fit)"},
        "batched_vs_plain": value / plain_value,
        "plain_path": {"value": plain_value, "wall_s": plain_wall, "ms_per_instance": args.plain_ms,
                       "instances": len(plain), "layout": "library default", "warm_ms": args.plain_warm_ms},
        "roofline": None, "cpu_baseline": None,
    }
    return line


def c5_requests(n, seed, qps):
    """Served workload: batched-sweep-style networks with N log-uniform in [2k, 100k], T
    uniform in [100, 2000] ms issued as 1-4 Simulate calls, arriving as a Poisson stream."""
    rng = np.random.default_rng(seed)
    specs = c4_specs(n, seed + 1, n_min=2000, n_max=100000)
    reqs = []
    t = 0.0
    for sp in specs:
        t += rng.exponential(1.0 / qps)
        T = round(float(rng.uniform(100.0, 2000.0)), 1)
        nc = int(rng.integers(1, 5))
        cuts = np.sort(np.round(rng.uniform(0.0, T, size=nc - 1), 1))
        calls = [float(x) for x in np.diff(np.concatenate([[0.0], cuts, [T]])) if x > 0]
        reqs.append(dict(spec=sp, arrival=t, calls=calls))
    return reqs


def run_served(args, nestgpu):
    """Served form: requests arrive open-loop at a fixed rate; a builder pool constructs each
    network on the host and submits it; the server admits by HBM budget and batches all
    running requests in the same slice launches. Reported: p50/p99 end-to-end latency and
    aggregate recurrent events/s, next to the same requests run one at a time on the plain
    path (measured service times, queued in arrival order)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    reqs = c5_requests(args.requests, args.seed, args.qps)
    srv = nestgpu.Server(0, hbm_budget_bytes=int(args.hbm_budget_gib * (1 << 30)), slices_per_chunk=16)
    lock = threading.Lock()
    results = [None] * len(reqs)
    t_origin = time.perf_counter()

    def handle(i):
        r = reqs[i]
        k = nestgpu.Kernel(0)
        build_instance(k, nestgpu, r["spec"], keep_host_connections=False, rng_type=args.rng_type,
                       connect_on_device=bool(args.connect_on_device))
        built = time.perf_counter() - t_origin
        rid = srv.submit(k, r["calls"])
        info = srv.wait(rid)
        done = time.perf_counter() - t_origin
        with lock:
            results[i] = dict(arrival=r["arrival"], built=built, done=done, events=info["events"], server=info)
        del k

    with ThreadPoolExecutor(args.builders) as ex:
        futs = []
        for i, r in enumerate(reqs):
            delay = r["arrival"] - (time.perf_counter() - t_origin)
            if delay > 0:
                time.sleep(delay)
            futs.append(ex.submit(handle, i))
        for f in futs:
            f.result()
    stats = srv.stats()
    srv.close()
    lat = np.array([x["done"] - x["arrival"] for x in results])
    makespan = max(x["done"] for x in results) - min(x["arrival"] for x in results)
    events = sum(x["events"] for x in results)
    value = events / makespan
    print("served: %d requests, makespan %.1f s, p50 %.2f s, p99 %.2f s, %.3g events/s, max concurrent %d" % (
        len(reqs), makespan, np.percentile(lat, 50), np.percentile(lat, 99), value, stats["max_concurrent"]),
        file=sys.stderr, flush=True)
    # where a request's latency goes (medians): host construction including the wait for a
    # builder, the wait for HBM budget, device preparation up to the next chunk boundary, run
    stages = {
        "host_build_s": float(np.median([x["built"] - x["arrival"] for x in results])),
        "admission_wait_s": float(np.median([x["server"]["t_admit"] - x["server"]["t_submit"] for x in results])),
        "prepare_s": float(np.median([x["server"]["t_start"] - x["server"]["t_admit"] for x in results])),
        "run_s": float(np.median([x["server"]["t_done"] - x["server"]["t_start"] for x in results])),
    }
    print("median stages: " + ", ".join("%s %.2f" % kv for kv in stages.items()), file=sys.stderr, flush=True)
    # serial plain path: each request alone (host build, then its Simulate calls)
    service = []
    for r in reqs:
        t0 = time.perf_counter()
        k = nestgpu.Kernel(0)
        build_instance(k, nestgpu, r["spec"], keep_host_connections=False, rng_type=args.rng_type,
                       connect_on_device=bool(args.connect_on_device))
        for c in r["calls"]:
            k.Simulate(c)
        service.append(time.perf_counter() - t0)
        del k
    done = 0.0
    lat_serial = []
    for r, sv in zip(reqs, service):
        done = max(done, r["arrival"]) + sv
        lat_serial.append(done - r["arrival"])
    lat_serial = np.array(lat_serial)
    serial_makespan = done - reqs[0]["arrival"]
    serial_value = events / serial_makespan
    print("serial plain path: makespan %.1f s, p50 %.2f s, p99 %.2f s, %.3g events/s" % (
        serial_makespan, np.percentile(lat_serial, 50), np.percentile(lat_serial, 99), serial_value),
        file=sys.stderr, flush=True)
    line = {
        "metric": METRIC, "value": value, "unit": "synaptic events/s", "n_gpus": 1, "steps": len(reqs),
        "warmup": 0, "ms_per_step": 1e3 * makespan / len(reqs), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "median_stages": stages},
        "config": {"workload": "served: %d requests at %g QPS (Poisson arrivals), batched-sweep-style networks "
                               "N log-uniform [2k, 100k], T uniform [100, 2000] ms in 1-4 Simulate calls, HBM "
                               "budget %g GiB, %d host builders" % (len(reqs), args.qps, args.hbm_budget_gib,
                                                                   args.builders),
                   "server": stats, "rng_type": args.rng_type, "connect_on_device": bool(args.connect_on_device)},
        "serial_plain_path": {"value": serial_value, "makespan_s": serial_makespan,
                              "latency_s": {"p50": float(np.percentile(lat_serial, 50)),
                                            "p99": float(np.percentile(lat_serial, 99))}},
        "served_vs_serial": value / serial_value,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def _poisson_counts_c2(nestgpu, N, drive_seed, steps):
    # the generators are created after the neurons: their node ids are N+1 .. 2N
    return nestgpu.poisson_counts(drive_seed, 8000.0 * 0.1 * 1e-3, N + 1, N, 0, steps)


def cpu_baseline(args, workload, threads):
    """The oracle (C++/OpenMP restatement of the reference algorithm, test infrastructure) on
    the same network with identical connections and drive, on this host's cores: the start-up
    transient is simulated first, untimed, then a sample of biological time is timed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import nestgpu
    from _oracle import Oracle

    t_warm, t_sample = args.cpu_transient_ms, args.cpu_sample_ms
    steps = int(round((t_warm + t_sample) / 0.1))
    o = Oracle(n_vp=args.n_vp, seed=args.seed, threads=threads)
    t0 = time.perf_counter()
    if workload == "c2":
        N, KE, KI = args.n, args.ke, args.ki
        NE = int(0.8 * N)
        g = o.create("iaf_psc_exp", N, NEURON)
        rng = np.random.default_rng(args.drive_seed)
        o.set_values(g, "V_m", rng.uniform(-70.0, -55.0, N))
        delay = ("sum", ("product", ("uniform_int", 21), 0.1), 1.0)
        o.connect(g[:NE], g, "fixed_indegree", KE, weight=("normal", 87.8, 8.8), delay=delay)
        o.connect(g[NE:], g, "fixed_indegree", KI, weight=("product", ("normal", 87.8, 8.8), -4.0), delay=delay)
        o.add_drive(g[0], _poisson_counts_c2(nestgpu, N, args.drive_seed, steps), 87.8, 1.0)
    else:
        build_c3_oracle(o, nestgpu, args.drive_seed, steps, stdp=C3_STDP if args.stdp else None)
    o.run_steps(0)  # prepare (store build) outside the timed window
    setup_s = time.perf_counter() - t0
    o.simulate(t_warm)
    ev0 = o.synaptic_events()
    t0 = time.perf_counter()
    o.simulate(t_sample)
    wall = time.perf_counter() - t0
    ev = o.synaptic_events() - ev0
    return {"value": ev / wall, "unit": "synaptic events/s", "cores": threads, "kind": "port",
            "sample": "same network (oracle connectivity from the same per-VP streams, identical drive), %d VPs; "
                      "%.0f ms of biological time after an untimed %.0f ms start-up transient: %d recurrent events in "
                      "%.2f s wall, RTF %.1f (oracle setup %.0f s, untimed)"
                      % (args.n_vp, t_sample, t_warm, ev, wall, wall / (t_sample * 1e-3), setup_s)}


def measured_traffic(workload):
    """HBM bytes per slice launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    of this workload, with the library source digest it was measured on; None when the built
    library's sources differ (the figure would describe another build)."""
    path = os.path.join(ROOT, "profiles", "k_slice_%s_hbm_traffic.txt" % workload)
    vals = {}
    try:
        for line in open(path):
            if ":" in line:
                key, val = line.split(":", 1)
                vals[key.strip()] = val.strip()
    except OSError:
        return None, None
    if "traffic_bytes_per_launch" not in vals:
        return None, None
    digest = vals.get("source_digest")
    return (float(vals["traffic_bytes_per_launch"]) if digest == source_digest() else None), digest


def source_digest():
    """sha256 (16 hex digits) over the library's kernel and host sources, in name order."""
    import hashlib

    h = hashlib.sha256()
    cdir = os.path.join(ROOT, "nest-simulator_amd", "csrc")
    for name in sorted(os.listdir(cdir)):
        with open(os.path.join(cdir, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def run_plain(args, nestgpu, workload, steps, warmup, with_cpu):
    """One plain-path measurement of the exponential-synapse random network (c2) or the
    8-population circuit (c3). Returns the bench line (without the outer keys)."""
    t_build = time.perf_counter()
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"delivery_mode": args.delivery_mode, "targets_per_block": args.targets_per_block,
                       "workgroups_per_cu": args.workgroups_per_cu, "window_slots": args.window_slots,
                       "graph_launch": bool(args.graph_launch)})
    if workload == "c2":
        build_c2(k, nestgpu, args.n, args.ke, args.ki, args.n_vp, args.seed, args.drive_seed)
        n_neurons = args.n
        drive_nz_per_step = args.n * (1.0 - np.exp(-8000.0 * 0.1e-3))
    else:
        if args.stdp:
            k.SetKernelStatus({"stdp_history_capacity": args.stdp_history})
        _, n_neurons, drive_nz_per_step = build_c3(k, nestgpu, args.n_vp, args.seed, args.drive_seed,
                                                   stdp=C3_STDP if args.stdp else None)
    n_syn = k.GetKernelStatus("num_connections") - n_neurons
    k.Prepare()
    build_s = time.perf_counter() - t_build
    print("%s built: %d neurons, %d synapses in %.1f s" % (workload, n_neurons, n_syn, build_s), file=sys.stderr,
          flush=True)

    for i in range(warmup):
        k.Run(args.t_step)
        print("%s warmup %d done" % (workload, i), file=sys.stderr, flush=True)
    st0 = k.GetKernelStatus()
    kernel_ms = 0.0  # GPU time of the slice launches: HIP events around every chunk on the run's stream
    t0 = time.perf_counter()
    for i in range(steps):
        k.Run(args.t_step)  # Run synchronizes the device before it returns
        kernel_ms += k.GetKernelStatus("time_simulate_kernel_ms")
    wall = time.perf_counter() - t0
    st = k.GetKernelStatus()
    events = st["synaptic_events"] - st0["synaptic_events"]
    spikes = st["local_spike_counter"] - st0["local_spike_counter"]
    graph_chunks = st["graph_chunks"] - st0["graph_chunks"]
    direct_chunks = st["direct_chunks"] - st0["direct_chunks"]
    bio_s = steps * args.t_step * 1e-3
    value = events / wall
    rtf = wall / bio_s
    rate = spikes / n_neurons / bio_s
    print("%s timed: %d steps, %.3f s wall, RTF %.4f, %.3g events/s, rate %.2f Hz" % (
        workload, steps, wall, rtf, value, rate), file=sys.stderr, flush=True)
    min_delay = st["min_delay_steps"]
    steps_t = steps * int(round(args.t_step / 0.1))
    slices_t = steps_t // min_delay
    # algorithmic bytes of the fused slice kernel, per unit:
    #   28 B per recurrent synaptic event (12 B synapse record, 16 B ring-buffer cell RMW)
    #   32 B per neuron-step (two ring-buffer channels read and cleared)
    #   56 B per neuron per slice (V, I_ex, I_in, r read once and written once)
    #   16 B per nonzero drive count (ring-buffer RMW; counts are generated in the kernel)
    def alg_bytes(ev, n_steps, n_slices):
        return 28.0 * ev + 32.0 * n_neurons * n_steps + 56.0 * n_neurons * n_slices + 16.0 * drive_nz_per_step * n_steps

    avg_us = 1e3 * kernel_ms / slices_t
    alg_per_launch = alg_bytes(events, steps_t, slices_t) / slices_t
    achieved = alg_per_launch / (avg_us * 1e-6) / 1e9

    # ---- second pass: HIP events around every single slice launch, and the workgroups' own
    # clock for the delivery phase
    k.SetKernelStatus({"kernel_timing": True, "phase_timing": True})
    evr0 = k.GetKernelStatus("synaptic_events")
    k.Run(args.roofline_ms)
    st2 = k.GetKernelStatus()
    k.SetKernelStatus({"kernel_timing": False, "phase_timing": False})
    events_r = st2["synaptic_events"] - evr0
    kms = st2["kernel_ms"]
    kl = st2["kernel_launches"]
    launches = kl["slice"]
    ticks = st2["phase_ticks"]
    deliv_share = ticks["deliver"] / ticks["total"] if ticks["total"] else 0.0
    deliv_us = avg_us * deliv_share
    deliv_bytes = 28.0 * events / slices_t
    deliv_achieved = deliv_bytes / (deliv_us * 1e-6) / 1e9 if deliv_us > 0 else 0.0
    per_kernel = {n: {"avg_us": 1e3 * ms / kl[n], "launches": kl[n], "total_ms": ms}
                  for n, ms in kms.items() if kl.get(n)}
    print("%s slice kernel: %.1f us per launch over the timed region (delivery %.1f us, %.0f %%); %.1f us between "
          "events around single launches" % (workload, avg_us, deliv_us, 100 * deliv_share,
                                            1e3 * kms["slice"] / launches), file=sys.stderr, flush=True)
    layout = {"delivery": "window" if st2["window_delivery"] else "cache" if st2["cache_delivery"] else "global",
              "targets_per_block": st2["targets_per_block"], "window_slots": st2["window_slots"]}
    del k
    gc.collect()

    cpu = None
    if with_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(args, workload, threads)
        print("%s cpu baseline: %.3g events/s on %d threads" % (workload, cpu["value"], threads), file=sys.stderr,
              flush=True)

    if workload == "c2":
        workload_text = ("exponential-synapse random network: %d iaf_psc_exp, fixed_indegree %d E + %d I (%d synapses), "
                         "delays 1.0-3.0 ms, seeded 8 kHz Poisson drive, dt 0.1 ms, one step = Run(%g ms)"
                         % (args.n, args.ke, args.ki, n_syn, args.t_step))
    else:
        workload_text = ("8-population layered circuit: 77169 iaf_psc_exp, pairwise_bernoulli (%d synapses), normal "
                         "delays clipped at dt (min_delay one step), Poisson background 8 Hz x K_ext, dt 0.1 ms, one "
                         "step = Run(%g ms)" % (n_syn, args.t_step))
        if args.stdp:
            workload_text += ("; connections between excitatory populations are stdp_synapse (tau_plus 20 ms, "
                              "lambda 0.01, alpha 1, mu 1, Wmax %g pA)" % C3_STDP["Wmax"])
    default_layout = (args.delivery_mode == "auto" and args.targets_per_block == 0 and args.workgroups_per_cu == 0
                      and args.window_slots == 0)
    traffic, traffic_digest = measured_traffic(workload) if default_layout else (None, None)
    return {
        "value": value, "unit": "synaptic events/s", "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * wall / steps, "rtf": rtf,
        "config": dict({"workload": workload_text, "n_neurons": n_neurons, "n_synapses": n_syn, "n_vp": args.n_vp,
                        "t_step_ms": args.t_step, "min_delay_steps": min_delay, "mean_rate_hz": rate,
                        "build_s": build_s, "graph_chunks": graph_chunks, "direct_chunks": direct_chunks}, **layout),
        "roofline": {"bound": "hbm", "kernel": "k_slice (fused deliver + update + gather, iaf_psc_exp)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source_digest": traffic_digest, "source_digest": source_digest(),
                     "avg_launch_us": avg_us, "alg_bytes_per_launch": alg_per_launch,
                     "events_per_launch": events / slices_t, "launches": slices_t,
                     "method": "timed region: GPU time of the slice launches (HIP events around every chunk on the "
                               "simulation stream) / launches",
                     "single_launch_pass": {"kernels": per_kernel, "events": events_r,
                                            "note": "HIP events around each launch, one at a time; includes the "
                                                    "event and launch overhead"},
                     "delivery": {"achieved": deliv_achieved, "frac": deliv_achieved / HBM_PEAK_GBS,
                                  "avg_us": deliv_us, "share_of_slice": deliv_share,
                                  "alg_bytes_per_launch": deliv_bytes,
                                  "method": "28 B per recurrent event / (slice launch time x the workgroups' mean "
                                            "delivery-phase share of their run time, 100 MHz clock)"}},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------- rocprofv3 summaries
def _csv_rows(path):
    import csv

    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def summarize_trace(path, skip, count, kernel="k_slice"):
    """Per-kernel totals of a rocprofv3 --kernel-trace CSV, and the mean duration of the
    dispatches [skip, skip + count) of the named kernel (the timed region of a bench run)."""
    rows = sorted(_csv_rows(path), key=lambda r: int(r["Start_Timestamp"]))
    per = {}
    for r in rows:
        per.setdefault(r["Kernel_Name"], []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
    total = sum(sum(v) for v in per.values())
    out = ["%-64s %7s %12s %10s %7s" % ("kernel", "calls", "total_ms", "avg_us", "pct")]
    for name, v in sorted(per.items(), key=lambda kv: -sum(kv[1]))[:6]:
        out.append("%-64s %7d %12.3f %10.2f %7.2f" % (name[:64], len(v), sum(v) * 1e-3, sum(v) / len(v),
                                                       100.0 * sum(v) / total))
    ks = [d for r, d in ((r, (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3) for r in rows)
          if kernel in r["Kernel_Name"]]
    sel = ks[skip:skip + count]
    out.append("")
    out.append("%s: dispatches %d-%d of %d (the timed region): avg %.2f us, median %.2f us" % (
        kernel, skip, skip + len(sel) - 1, len(ks), float(np.mean(sel)), float(np.median(sel))))
    return "\n".join(out), float(np.mean(sel))


def _counter(path, counter, kernel=None):
    rows = [r for r in _csv_rows(path) if r["Counter_Name"] == counter and (kernel is None or kernel in r["Kernel_Name"])]
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return np.array([float(r["Counter_Value"]) for r in rows])


def summarize_pmc(fetch, write, copy_fetch, copy_write, skip, count, kernel="k_slice"):
    """HBM bytes per dispatch of the named kernel over dispatches [skip, skip + count) from
    FETCH_SIZE and WRITE_SIZE runs, each scaled by what the counter reads for the plain 512 MiB
    device copy of bench.py --calibrate-copy (its three largest dispatches)."""
    copy_bytes = 512.0 * 1024 * 1024
    out, tot = [], 0.0
    for name, path, cpath in (("FETCH_SIZE", fetch, copy_fetch), ("WRITE_SIZE", write, copy_write)):
        c = np.sort(_counter(cpath, name))[-3:]
        scale = copy_bytes / (float(np.median(c)) * 1024.0)
        v = _counter(path, name, kernel)
        sel = v[skip:skip + count]
        per = float(np.mean(sel)) * 1024.0 * scale
        tot += per
        out.append("%s: copy reads %.0f KiB for 512 MiB (scale %.3f); %s dispatches %d-%d of %d: mean %.0f KiB, "
                   "median %.0f KiB -> %.1f MB per launch" % (name, float(np.median(c)), scale, kernel, skip,
                                                              skip + len(sel) - 1, len(v), float(np.mean(sel)),
                                                              float(np.median(sel)), per * 1e-6))
    out.append("traffic_bytes_per_launch: %d" % int(round(tot)))
    out.append("source_digest: %s" % source_digest())
    return "\n".join(out)


def summarize_counters(paths, skip, count, kernel="k_slice"):
    """Mean per dispatch of every counter in rocprofv3 --pmc CSVs over the named kernel's
    dispatches [skip, skip + count), and the ratios the SQ counters give: the share of wave
    cycles spent waiting and issuing, instructions per wave by kind, LDS bank-conflict cycles
    per LDS-active cycle."""
    vals = {}
    n_sel = {}
    for path in paths:
        names = sorted({r["Counter_Name"] for r in _csv_rows(path)})
        for name in names:
            v = _counter(path, name, kernel)
            sel = v[skip:skip + count]
            vals[name] = float(np.mean(sel))
            n_sel[name] = (len(sel), len(v))
    out = ["%s dispatches %d-%d" % (kernel, skip, skip + count - 1)]
    for name in sorted(vals):
        out.append("%-24s %16.1f   (%d of %d dispatches)" % (name, vals[name], n_sel[name][0], n_sel[name][1]))
    g = vals.get
    if g("SQ_WAVE_CYCLES"):
        for name in ("SQ_WAIT_ANY", "SQ_ACTIVE_INST_ANY", "SQ_WAIT_INST_ANY"):
            if name in vals:
                out.append("%s / SQ_WAVE_CYCLES = %.3f" % (name, vals[name] / vals["SQ_WAVE_CYCLES"]))
    if g("SQ_WAVES"):
        for name in sorted(vals):
            if name.startswith("SQ_INSTS_"):
                out.append("%s per wave = %.1f" % (name, vals[name] / vals["SQ_WAVES"]))
    if g("SQ_LDS_IDX_ACTIVE") and "SQ_LDS_BANK_CONFLICT" in vals:
        out.append("SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE = %.3f" % (vals["SQ_LDS_BANK_CONFLICT"] / vals["SQ_LDS_IDX_ACTIVE"]))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--t-step", type=float, default=1000.0, help="biological ms per step")
    ap.add_argument("--n", type=int, default=100000, help="c2: neurons")
    ap.add_argument("--ke", type=int, default=1000, help="c2: excitatory in-degree")
    ap.add_argument("--ki", type=int, default=250, help="c2: inhibitory in-degree")
    ap.add_argument("--n-vp", type=int, default=8)
    ap.add_argument("--seed", type=int, default=20241)
    ap.add_argument("--drive-seed", type=int, default=777)
    ap.add_argument("--roofline-ms", type=float, default=128.0,
                    help="biological ms of the per-launch timed pass (a multiple of 64 slices avoids a short chunk)")
    ap.add_argument("--cpu-transient-ms", type=float, default=100.0,
                    help="untimed biological ms of the CPU baseline before its sample")
    ap.add_argument("--cpu-sample-ms", type=float, default=100.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c2", action="store_true", help="c3: skip the extra c2 measurement")
    ap.add_argument("--no-c4", action="store_true", help="c3: skip the extra batched-sweep (c4) measurement")
    ap.add_argument("--stdp", action="store_true",
                    help="c3: connections between excitatory populations are stdp_synapse")
    ap.add_argument("--no-c3-stdp", action="store_true",
                    help="c3: skip the extra measurement of the circuit with stdp_synapse between excitatory populations")
    ap.add_argument("--stdp-history", type=int, default=256, help="c3 --stdp: post-spike history entries per neuron")
    ap.add_argument("--c2-steps", type=int, default=5, help="c3: timed steps of the extra c2 measurement")
    ap.add_argument("--workload", default="c3", choices=["c2", "c3", "c4", "c5"],
                    help="c3: 77k-neuron 8-population circuit with min_delay one step (default); c2: 100k-neuron "
                         "exponential-synapse random network; c4: batched sweep of many instances; c5: served "
                         "request stream")
    ap.add_argument("--instances", type=int, default=256, help="c4: number of instances")
    ap.add_argument("--plain-ms", type=float, default=100.0, help="c4: biological ms per instance on the plain path")
    ap.add_argument("--plain-warm-ms", type=float, default=100.0,
                    help="c4: untimed biological ms per instance before the plain-path timing")
    ap.add_argument("--autotune", action="store_true", help="c4: pick targets_per_block per size class first")
    ap.add_argument("--probe-ms", type=float, default=30.0, help="c4: autotuner probe length")
    ap.add_argument("--autotune-blocks", default="64,128,256,512",
                    help="c4: candidate targets_per_block for the autotuner")
    ap.add_argument("--autotune-windows", default="0,4",
                    help="c4: candidate caps on the LDS ring slots for the autotuner (0: as many as fit)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 512 MiB three times (scale for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--requests", type=int, default=32, help="c5: number of requests")
    ap.add_argument("--qps", type=float, default=1.0, help="c5: request arrival rate")
    ap.add_argument("--builders", type=int, default=6, help="c5: host threads building request networks")
    ap.add_argument("--hbm-budget-gib", type=float, default=128.0, help="c5: device memory budget of the server")
    ap.add_argument("--delivery-mode", default="auto", choices=["auto", "window", "global", "cache"])
    ap.add_argument("--rng-type", default="mt19937_64", choices=["mt19937_64", "Philox_32"],
                    help="c4/c5: per-VP stream type of the instances")
    ap.add_argument("--connect-on-device", type=int, default=0, choices=[0, 1],
                    help="c4/c5: draw the instances' connections on the device (needs --rng-type Philox_32)")
    ap.add_argument("--targets-per-block", type=int, default=0, help="0: chosen by the library")
    ap.add_argument("--window-slots", type=int, default=0, help="cap on ring slots in LDS; 0: as many as fit")
    ap.add_argument("--workgroups-per-cu", type=int, default=0, help="0: as many as fit")
    ap.add_argument("--graph-launch", type=int, default=1, choices=[0, 1],
                    help="1: replay each chunk of slice launches from a captured hipGraph (default); 0: launch "
                         "every slice directly")
    ap.add_argument("--summarize-trace", metavar="KERNEL_TRACE_CSV",
                    help="print the kernel-time summary of a rocprofv3 --kernel-trace CSV (with --skip/--count)")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="print the HBM traffic per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE CSVs of a "
                         "bench run and of --calibrate-copy (with --skip/--count)")
    ap.add_argument("--summarize-counters", nargs="+", metavar="COUNTER_CSV",
                    help="print the mean per k_slice dispatch of every counter in rocprofv3 --pmc CSVs of a bench "
                         "run, with the SQ ratios (with --skip/--count)")
    ap.add_argument("--skip", type=int, default=0, help="summaries: k_slice dispatches before the timed region")
    ap.add_argument("--count", type=int, default=1 << 30, help="summaries: k_slice dispatches of the timed region")
    args = ap.parse_args()

    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace, args.skip, args.count)[0])
        return
    if args.summarize_pmc:
        print(summarize_pmc(*args.summarize_pmc, args.skip, args.count))
        return
    if args.summarize_counters:
        print(summarize_counters(args.summarize_counters, args.skip, args.count))
        return

    if args.calibrate_copy:
        import torch

        n = 64 * 1024 * 1024  # float64 elements: 512 MiB
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copy bytes per launch: read %d, written %d" % (n * 8, n * 8), flush=True)
        return

    if args.gpus != 1:
        print("bench.py measures one GPU: the slice step runs on a single MI355X here and is not "
              "sharded across GPUs.", file=sys.stderr)
        sys.exit(2)

    import nestgpu

    if nestgpu.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload in ("c4", "c5"):
        if args.seed == 20241:
            args.seed = 4242
        if args.workload == "c4":
            print(json.dumps(run_batched(args, nestgpu)), flush=True)
        else:
            run_served(args, nestgpu)
        return

    res = run_plain(args, nestgpu, args.workload, args.steps, args.warmup, not args.no_cpu_baseline)
    line = {"metric": METRIC, "value": res["value"], "unit": res["unit"], "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic", "rtf": res["rtf"], "config": res["config"],
            "roofline": res["roofline"], "cpu_baseline": res["cpu_baseline"]}
    if args.workload == "c3" and not args.no_c2:
        c2 = run_plain(args, nestgpu, "c2", args.c2_steps, 1, False)
        line["c2"] = {k: c2[k] for k in ("value", "unit", "steps", "ms_per_step", "rtf", "config")}
        line["c2"]["roofline"] = {k: c2["roofline"][k] for k in ("achieved", "frac", "avg_launch_us", "traffic",
                                                                 "delivery")}
    if args.workload == "c3" and not args.stdp and not args.no_c3_stdp:
        # the same circuit with every connection between excitatory populations plastic
        a3 = argparse.Namespace(**vars(args))
        a3.stdp = True
        c3s = run_plain(a3, nestgpu, "c3", 1, 1, False)
        line["c3_stdp"] = {k: c3s[k] for k in ("value", "unit", "steps", "ms_per_step", "rtf", "config")}
        line["c3_stdp"]["avg_launch_us"] = c3s["roofline"]["avg_launch_us"]
    if args.workload == "c3" and not args.no_c4:
        # the batched sweep, autotuned, at 100 ms steps: its batched / plain ratio
        a4 = argparse.Namespace(**vars(args))
        a4.seed, a4.t_step, a4.steps, a4.warmup, a4.plain_ms = 4242, 100.0, 2, 1, 50.0
        a4.autotune, a4.autotune_blocks = True, "128,256,512"
        c4 = run_batched(a4, nestgpu)
        line["c4"] = {k: c4[k] for k in ("value", "unit", "steps", "ms_per_step", "rtf", "config", "batched_vs_plain",
                                         "plain_path")}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
