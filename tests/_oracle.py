"""ctypes client for the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker the MI355X path is compared with.
Parameter specs use the prefix encoding documented in oracle/nest_oracle.cpp
(0 const, 1 uniform, 2 uniform_int, 3 normal, 4 sum, 5 product, 6 max, 7 min, 8 redraw, ...; 18/19
normal/lognormal parameter objects named by an id).
"""

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

MODELS = {"iaf_psc_delta": 0, "iaf_psc_exp": 1, "aeif_cond_exp": 2, "iaf_psc_alpha": 3}
RULES = {"one_to_one": 0, "all_to_all": 1, "fixed_indegree": 2, "pairwise_bernoulli": 3}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.or_new.restype = vp
        L.or_new.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_double]
        L.or_free.argtypes = [vp]
        L.or_error.restype = ctypes.c_char_p
        L.or_error.argtypes = [vp]
        L.or_create.restype = ctypes.c_long
        L.or_num_connections.restype = ctypes.c_long
        L.or_num_spikes.restype = ctypes.c_long
        L.or_synaptic_events.restype = ctypes.c_long
        L.or_fd_exp.restype = ctypes.c_double
        L.or_fd_exp.argtypes = [ctypes.c_double]
        L.or_fd_log.restype = ctypes.c_double
        L.or_fd_log.argtypes = [ctypes.c_double]
        L.or_iaf_propagator.restype = ctypes.c_double
        L.or_iaf_propagator.argtypes = [ctypes.c_double] * 4
        L.or_num_samples.restype = ctypes.c_long
        L.or_set_synapse.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double]
        for name in ("or_set_synapse", "or_set_rng_type", "or_set_sampling", "or_get_samples", "or_set_threads", "or_set_portable_math", "or_set_status", "or_get_state", "or_connect",
                     "or_add_connections", "or_get_connections", "or_add_drive", "or_simulate", "or_run_steps",
                     "or_get_spikes", "or_clear_spikes", "or_kernel_status"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def spec(x):
    """Encode a parameter spec: float -> const, or a nested tuple like ('normal', m, s)."""
    if x is None:
        return np.zeros(0)
    out = []

    def enc(v):
        if isinstance(v, (int, float, np.floating, np.integer)):
            out.extend([0, float(v)])
            return
        op = v[0]
        if op == "uniform":
            out.extend([1, v[1], v[2]])
        elif op == "uniform_int":
            out.extend([2, v[1]])
        elif op == "normal":
            # ("normal", mean, std[, id]): with an id, the same parameter object in every
            # connect call that names it (its per-VP distribution state carries over)
            out.extend([3, v[1], v[2]] if len(v) < 4 else [18, v[3], v[1], v[2]])
        elif op == "lognormal":
            out.extend([9, v[1], v[2]] if len(v) < 4 else [19, v[3], v[1], v[2]])
        elif op == "exponential":
            out.extend([10, v[1]])
        elif op in ("exp", "sin", "cos"):
            out.append({"exp": 11, "sin": 12, "cos": 13}[op])
            enc(v[1])
        elif op == "quotient":
            out.append(14)
            enc(v[1])
            enc(v[2])
        elif op == "pow":
            out.append(15)
            enc(v[1])
            out.append(v[2])
        elif op == "compare":
            out.append(16)
            enc(v[1])
            enc(v[2])
            out.append(v[3])
        elif op == "conditional":
            out.append(17)
            enc(v[1])
            enc(v[2])
            enc(v[3])
        elif op in ("sum", "product"):
            out.append(4 if op == "sum" else 5)
            enc(v[1])
            enc(v[2])
        elif op in ("max", "min"):
            out.append(6 if op == "max" else 7)
            enc(v[1])
            out.append(v[2])
        elif op == "redraw":
            out.append(8)
            enc(v[1])
            out.extend([v[2], v[3]])
        else:
            raise ValueError(op)

    enc(x)
    return np.asarray(out, dtype=np.float64)


class OracleError(RuntimeError):
    pass


class Oracle:
    RNG_TYPES = {"mt19937_64": 0, "Philox_32": 1}

    def __init__(self, n_vp=1, seed=143202461, resolution=0.1, threads=1, portable_math=False, rng_type="mt19937_64"):
        self.L = lib()
        self.k = self.L.or_new(n_vp, seed, resolution)
        if rng_type != "mt19937_64":
            self._chk(self.L.or_set_rng_type(ctypes.c_void_p(self.k), self.RNG_TYPES[rng_type]))
        if threads != 1:
            self._chk(self.L.or_set_threads(ctypes.c_void_p(self.k), threads))
        if portable_math:
            self._chk(self.L.or_set_portable_math(ctypes.c_void_p(self.k), 1))

    def __del__(self):
        if getattr(self, "k", None):
            self.L.or_free(ctypes.c_void_p(self.k))
            self.k = None

    def _chk(self, rc):
        if rc != 0:
            raise OracleError(self.L.or_error(ctypes.c_void_p(self.k)).decode())

    def create(self, model, n, params=None):
        params = params or {}
        keys = (ctypes.c_char_p * max(1, len(params)))(*[k.encode() for k in params])
        vals = np.asarray([float(v) for v in params.values()] or [0.0])
        g = self.L.or_create(ctypes.c_void_p(self.k), MODELS[model], ctypes.c_long(n), len(params), keys,
                             _p(vals, ctypes.c_double))
        if g < 0:
            raise OracleError(self.L.or_error(ctypes.c_void_p(self.k)).decode())
        return np.arange(g, g + n, dtype=np.int64)

    def set_status(self, gids, params):
        for g in np.atleast_1d(gids):
            keys = (ctypes.c_char_p * len(params))(*[k.encode() for k in params])
            vals = np.asarray([float(v) for v in params.values()])
            self._chk(self.L.or_set_status(ctypes.c_void_p(self.k), ctypes.c_long(int(g)), len(params), keys,
                                           _p(vals, ctypes.c_double)))

    def set_values(self, gids, key, values):
        for g, v in zip(np.atleast_1d(gids), np.atleast_1d(values)):
            self.set_status([g], {key: float(v)})

    def get(self, gids, field):
        gids = np.ascontiguousarray(np.atleast_1d(gids), dtype=np.int64)
        out = np.zeros(len(gids))
        self._chk(self.L.or_get_state(ctypes.c_void_p(self.k), _p(gids, ctypes.c_long), ctypes.c_long(len(gids)),
                                      field.encode(), _p(out, ctypes.c_double)))
        return out

    STDP_KEYS = ("tau_plus", "lambda", "alpha", "mu_plus", "mu_minus", "Wmax")
    STDP_DEFAULTS = {"tau_plus": 20.0, "lambda": 0.01, "alpha": 1.0, "mu_plus": 1.0, "mu_minus": 1.0, "Wmax": 100.0}

    def connect(self, pre, post, rule="all_to_all", rule_param=0.0, weight=None, delay=None,
                allow_autapses=True, allow_multapses=True, stdp=None):
        """stdp: None for static_synapse, else a dict of stdp_synapse parameters (and Kplus)."""
        if stdp is not None:
            prm = dict(self.STDP_DEFAULTS, **{k: v for k, v in stdp.items() if k != "Kplus"})
            arr = (ctypes.c_double * 6)(*[float(prm[k]) for k in self.STDP_KEYS])
            self._chk(self.L.or_set_synapse(ctypes.c_void_p(self.k), 1, arr, ctypes.c_double(stdp.get("Kplus", 0.0))))
        try:
            self._connect(pre, post, rule, rule_param, weight, delay, allow_autapses, allow_multapses)
        finally:
            if stdp is not None:
                self._chk(self.L.or_set_synapse(ctypes.c_void_p(self.k), 0, None, ctypes.c_double(0.0)))

    def _connect(self, pre, post, rule, rule_param, weight, delay, allow_autapses, allow_multapses):
        pre = np.ascontiguousarray(pre, dtype=np.int64)
        post = np.ascontiguousarray(post, dtype=np.int64)
        ws, ds = spec(weight), spec(delay)
        self._chk(self.L.or_connect(ctypes.c_void_p(self.k), _p(pre, ctypes.c_long), ctypes.c_long(len(pre)),
                                    _p(post, ctypes.c_long), ctypes.c_long(len(post)), RULES[rule],
                                    ctypes.c_double(rule_param), int(allow_autapses), int(allow_multapses),
                                    _p(ws, ctypes.c_double), len(ws), _p(ds, ctypes.c_double), len(ds)))

    def add_connections(self, src, tgt, w, d_steps):
        src = np.ascontiguousarray(src, dtype=np.int64)
        tgt = np.ascontiguousarray(tgt, dtype=np.int64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        d = np.ascontiguousarray(d_steps, dtype=np.int64)
        self._chk(self.L.or_add_connections(ctypes.c_void_p(self.k), _p(src, ctypes.c_long), _p(tgt, ctypes.c_long),
                                            _p(w, ctypes.c_double), _p(d, ctypes.c_long), ctypes.c_long(len(src))))

    def connections(self):
        n = self.L.or_num_connections(ctypes.c_void_p(self.k))
        s = np.zeros(n, np.int64)
        t = np.zeros(n, np.int64)
        w = np.zeros(n)
        d = np.zeros(n, np.int64)
        self._chk(self.L.or_get_connections(ctypes.c_void_p(self.k), _p(s, ctypes.c_long), _p(t, ctypes.c_long),
                                            _p(w, ctypes.c_double), _p(d, ctypes.c_long)))
        return s, t, w, d

    def add_drive(self, gid0, counts, weight, delay):
        """counts: uint8 [n_steps, n]; emission step e sends at spike time (e+1)*h."""
        counts = np.ascontiguousarray(counts, dtype=np.uint8)
        n_steps, n = counts.shape
        self._chk(self.L.or_add_drive(ctypes.c_void_p(self.k), ctypes.c_long(int(gid0)), ctypes.c_long(n),
                                      ctypes.c_long(n_steps), _p(counts, ctypes.c_uint8), ctypes.c_double(weight),
                                      ctypes.c_double(delay)))

    def simulate(self, t):
        self._chk(self.L.or_simulate(ctypes.c_void_p(self.k), ctypes.c_double(t)))

    def run_steps(self, n):
        self._chk(self.L.or_run_steps(ctypes.c_void_p(self.k), ctypes.c_long(n)))

    def spikes(self):
        n = self.L.or_num_spikes(ctypes.c_void_p(self.k))
        g = np.zeros(n, np.int64)
        s = np.zeros(n, np.int64)
        self._chk(self.L.or_get_spikes(ctypes.c_void_p(self.k), _p(g, ctypes.c_long), _p(s, ctypes.c_long)))
        return g, s

    def set_sampling(self, gids, interval_steps, offset_steps=0):
        """Sample the state of these neurons after every step whose stamp is
        offset + k * interval (the multimeter's UniversalDataLogger)."""
        gids = np.ascontiguousarray(np.atleast_1d(gids), dtype=np.int64)
        self._chk(self.L.or_set_sampling(ctypes.c_void_p(self.k), _p(gids, ctypes.c_long), ctypes.c_long(len(gids)),
                                         ctypes.c_long(int(interval_steps)), ctypes.c_long(int(offset_steps))))

    def samples(self):
        """(gid, stamp, fields[n, 4]) sorted by (stamp, gid); fields are V_m and the model's
        other state variables as a user reads them (iaf_psc_exp: I_syn_ex, I_syn_in;
        aeif_cond_exp: g_ex, g_in, w)."""
        n = self.L.or_num_samples(ctypes.c_void_p(self.k))
        g = np.zeros(n, np.int64)
        st = np.zeros(n, np.int64)
        f = np.zeros((n, 4))
        self._chk(self.L.or_get_samples(ctypes.c_void_p(self.k), _p(g, ctypes.c_long), _p(st, ctypes.c_long),
                                        _p(f, ctypes.c_double)))
        return g, st, f

    def synaptic_events(self):
        return int(self.L.or_synaptic_events(ctypes.c_void_p(self.k)))

    def status(self):
        out = np.zeros(5, np.int64)
        self._chk(self.L.or_kernel_status(ctypes.c_void_p(self.k), _p(out, ctypes.c_long)))
        return dict(min_delay=int(out[0]), max_delay=int(out[1]), clock=int(out[2]), from_step=int(out[3]),
                    slice=int(out[4]))
