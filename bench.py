"""Throughput of the MI355X simulation step on the exponential-synapse random network.

Workload (one GPU): 100,000 iaf_psc_exp neurons (tau_m 10 ms, tau_syn 0.5 ms, C_m 250 pF,
V_th -55 mV, V_reset -70 mV, t_ref 2 ms), 80 % excitatory, fixed_indegree 1000 E + 250 I
(1.25e8 static synapses), weights normal(87.8, 8.8) pA and -4x for inhibitory sources,
delays uniform over the integer steps 1.0..3.0 ms, seeded Poisson drive of 8 kHz per
neuron through one spike_generator each (weight 87.8 pA, delay 1.0 ms), dt = 0.1 ms, FP64.

A step is Run(1000 ms) -- one biological second, 1,000 min_delay slices -- continuing the
same network. value = recurrent synaptic events delivered per wall-clock second over the
timed steps; RTF (wall / biological time) is reported beside it.

The roofline pass re-runs part of a step with per-kernel HIP events on the simulation
stream and divides each kernel's algorithmic bytes by its average launch duration.
The CPU baseline times the oracle (C++/OpenMP restatement of the reference algorithm)
on the same network for a short biological window on this host's cores.
"""

import argparse
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


def cpu_baseline(N, KE, KI, n_vp, seed, drive_seed, t_sample, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import nestgpu
    from _oracle import Oracle

    NE = int(0.8 * N)
    o = Oracle(n_vp=n_vp, seed=seed, threads=threads)
    g = o.create("iaf_psc_exp", N, NEURON)
    rng = np.random.default_rng(drive_seed)
    o.set_values(g, "V_m", rng.uniform(-70.0, -55.0, N))
    delay = ("sum", ("product", ("uniform_int", 21), 0.1), 1.0)
    o.connect(g[:NE], g, "fixed_indegree", KE, weight=("normal", 87.8, 8.8), delay=delay)
    o.connect(g[NE:], g, "fixed_indegree", KI, weight=("product", ("normal", 87.8, 8.8), -4.0), delay=delay)
    steps = int(round(t_sample / 0.1))
    # the generators are created after the neurons: their node ids are N+1 .. 2N
    counts = nestgpu.poisson_counts(drive_seed, 8000.0 * 0.1 * 1e-3, N + 1, N, 0, steps)
    o.add_drive(g[0], counts, 87.8, 1.0)
    o.run_steps(0)  # prepare (store build) outside the timed window
    t0 = time.perf_counter()
    o.simulate(t_sample)
    wall = time.perf_counter() - t0
    ev = o.synaptic_events()
    return {"value": ev / wall, "unit": "synaptic events/s", "cores": threads, "kind": "port",
            "sample": "same network (oracle connectivity, identical draws), first %.0f ms of biological time; "
                      "%d recurrent events in %.2f s wall, RTF %.1f" % (t_sample, ev, wall, wall / (t_sample * 1e-3))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--t-step", type=float, default=1000.0, help="biological ms per step")
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--ke", type=int, default=1000)
    ap.add_argument("--ki", type=int, default=250)
    ap.add_argument("--n-vp", type=int, default=8)
    ap.add_argument("--seed", type=int, default=20241)
    ap.add_argument("--drive-seed", type=int, default=777)
    ap.add_argument("--roofline-ms", type=float, default=200.0)
    ap.add_argument("--cpu-sample-ms", type=float, default=30.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--delivery-mode", default="auto", choices=["auto", "window", "global"])
    ap.add_argument("--targets-per-block", type=int, default=0, help="0: chosen by the library")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X and the step "
              "does not shard across GPUs here (see DESIGN.md for how it would).", file=sys.stderr)
        sys.exit(2)

    import nestgpu

    if nestgpu.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    t_build = time.perf_counter()
    k = nestgpu.Kernel(0)
    k.SetKernelStatus({"delivery_mode": args.delivery_mode, "targets_per_block": args.targets_per_block})
    pop, sg = build_c2(k, nestgpu, args.n, args.ke, args.ki, args.n_vp, args.seed, args.drive_seed)
    n_syn = k.GetKernelStatus("num_connections") - args.n
    k.Prepare()
    build_s = time.perf_counter() - t_build
    print("built: %d neurons, %d synapses in %.1f s" % (args.n, n_syn, build_s), file=sys.stderr, flush=True)

    for i in range(args.warmup):
        k.Run(args.t_step)
        print("warmup %d done" % i, file=sys.stderr, flush=True)
    ev0 = k.GetKernelStatus("synaptic_events")
    sp0 = k.GetKernelStatus("local_spike_counter")
    t0 = time.perf_counter()
    for i in range(args.steps):
        k.Run(args.t_step)  # Run synchronizes the device before it returns
    wall = time.perf_counter() - t0
    st = k.GetKernelStatus()
    events = st["synaptic_events"] - ev0
    spikes = st["local_spike_counter"] - sp0
    bio_s = args.steps * args.t_step * 1e-3
    value = events / wall
    rtf = wall / bio_s
    rate = spikes / args.n / bio_s
    print("timed: %d steps, %.3f s wall, RTF %.4f, %.3g events/s, rate %.2f Hz" % (args.steps, wall, rtf, value, rate),
          file=sys.stderr, flush=True)

    # ---- roofline pass: per-kernel HIP events on the simulation stream
    k.SetKernelStatus({"kernel_timing": True})
    evr0 = k.GetKernelStatus("synaptic_events")
    k.Run(args.roofline_ms)
    k.SetKernelStatus({"kernel_timing": False})
    st2 = k.GetKernelStatus()
    events_r = st2["synaptic_events"] - evr0
    kms = st2["kernel_ms"]
    kl = st2["kernel_launches"]
    steps_r = int(round(args.roofline_ms / 0.1))
    min_delay = st2["min_delay_steps"]
    slices_r = steps_r // min_delay
    # algorithmic bytes of the fused slice kernel (per-unit figures in DESIGN.md):
    #   28 B per recurrent synaptic event (12 B synapse word + weight, 16 B ring-buffer RMW)
    #   32 B per neuron-step (two ring-buffer channels read and cleared)
    #   56 B per neuron per slice (V, I_ex, I_in, r read once and written once)
    #   16 B per nonzero drive count (ring-buffer RMW; counts are generated in the kernel)
    lam = 8000.0 * 0.1e-3
    nz = args.n * steps_r * (1.0 - np.exp(-lam))
    launches = kl["slice"]
    alg_total = 28.0 * events_r + 32.0 * args.n * steps_r + 56.0 * args.n * slices_r + 16.0 * nz
    alg_per_launch = alg_total / launches
    avg_us = 1e3 * kms["slice"] / launches
    achieved = alg_per_launch / (avg_us * 1e-6) / 1e9
    dominant = "k_slice (fused deliver + update, iaf_psc_exp)"
    per_kernel = {n: {"avg_us": 1e3 * ms / kl[n], "launches": kl[n], "total_ms": ms}
                  for n, ms in kms.items() if kl.get(n)}
    print("slice kernel: %.1f us avg over %d launches, %.3g events in roofline pass" % (avg_us, launches, events_r),
          file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(args.n, args.ke, args.ki, args.n_vp, args.seed, args.drive_seed, args.cpu_sample_ms, threads)
        print("cpu baseline: %.3g events/s on %d threads" % (cpu["value"], threads), file=sys.stderr, flush=True)

    line = {
        "metric": METRIC, "value": value, "unit": "synaptic events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "rtf": rtf,
        "config": {"workload": "exponential-synapse random network: %d iaf_psc_exp, fixed_indegree %d E + %d I "
                               "(%d synapses), delays 1.0-3.0 ms, seeded 8 kHz Poisson drive, dt 0.1 ms, one "
                               "step = Run(%g ms)" % (args.n, args.ke, args.ki, n_syn, args.t_step),
                   "n_neurons": args.n, "n_synapses": n_syn, "n_vp": args.n_vp, "t_step_ms": args.t_step,
                   "mean_rate_hz": rate, "build_s": build_s,
                   "delivery": "window" if st2["window_delivery"] else "global",
                   "targets_per_block": st2["targets_per_block"]},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_us": avg_us, "alg_bytes_per_launch": alg_per_launch,
                     "events_per_launch": events_r / launches, "kernels": per_kernel},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
