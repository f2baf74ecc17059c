"""nestgpu -- PyNEST-shaped Python binding of the MI355X point-neuron simulation step.

The calls mirror PyNEST's (pynest/nest/lib/hl_api_*.py): ResetKernel, SetKernelStatus,
GetKernelStatus, Create, Connect, Simulate/Prepare/Run/Cleanup, GetConnections,
NodeCollection.get/set, spike_recorder events and the nest.random / nest.math parameter
objects. Every call goes through the C ABI declared in include/nestgpu.h
(libnestgpu.so, built in place by `make -C nest-simulator_amd`). The library must be
present: there is no CPU fallback.
"""

import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnestgpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libnestgpu.so is not built (%s); run `make -C nest-simulator_amd` or __graft_entry__.build()" % LIB_PATH)

_lib = ctypes.CDLL(LIB_PATH)
_vp = ctypes.c_void_p
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "nestgpu_kernel_new": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "nestgpu_kernel_free": (None, [_vp]),
    "nestgpu_last_error": (ctypes.c_char_p, [_vp]),
    "nestgpu_reset_kernel": (ctypes.c_int, [_vp]),
    "nestgpu_set_kernel_status": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "nestgpu_get_kernel_status": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t,
                                                 ctypes.POINTER(ctypes.c_size_t)]),
    "nestgpu_create": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int64, ctypes.c_char_p, _i64p]),
    "nestgpu_set_status": (ctypes.c_int, [_vp, _i64p, ctypes.c_int64, ctypes.c_char_p]),
    "nestgpu_set_values": (ctypes.c_int, [_vp, _i64p, ctypes.c_int64, ctypes.c_char_p, _f64p]),
    "nestgpu_get_values": (ctypes.c_int, [_vp, _i64p, ctypes.c_int64, ctypes.c_char_p, _f64p]),
    "nestgpu_connect": (ctypes.c_int, [_vp, _i64p, ctypes.c_int64, _i64p, ctypes.c_int64, ctypes.c_char_p,
                                       ctypes.c_char_p]),
    "nestgpu_connect_arrays": (ctypes.c_int, [_vp, _i64p, _i64p, _f64p, _f64p, ctypes.c_int64]),
    "nestgpu_set_spike_counts": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                ctypes.POINTER(ctypes.c_uint8)]),
    "nestgpu_num_connections": (ctypes.c_int64, [_vp]),
    "nestgpu_get_connections": (ctypes.c_int, [_vp, _i64p, _i64p, _f64p, _i64p, ctypes.POINTER(ctypes.c_int32)]),
    "nestgpu_get_connection_models": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "nestgpu_prepare": (ctypes.c_int, [_vp]),
    "nestgpu_run": (ctypes.c_int, [_vp, ctypes.c_double]),
    "nestgpu_cleanup": (ctypes.c_int, [_vp]),
    "nestgpu_simulate": (ctypes.c_int, [_vp, ctypes.c_double]),
    "nestgpu_recorder_events": (ctypes.c_int, [_vp, ctypes.c_int64, _i64p, _f64p, ctypes.c_int64, _i64p]),
    "nestgpu_multimeter_events": (ctypes.c_int, [_vp, ctypes.c_int64, _i64p, _f64p, _f64p, ctypes.c_int64, _i64p]),
    "nestgpu_node_status": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_char_p, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_size_t)]),
    "nestgpu_batch_simulate": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _f64p]),
    "nestgpu_batch_simulate_noalloc": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _f64p]),
    "nestgpu_batch_autotune": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                              ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_int)]),
    "nestgpu_server_new": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_vp)]),
    "nestgpu_server_free": (None, [_vp]),
    "nestgpu_server_submit": (ctypes.c_int, [_vp, _vp, _f64p, ctypes.c_int, _i64p]),
    "nestgpu_server_wait": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_int), _f64p,
                                           ctypes.POINTER(ctypes.c_uint64)]),
    "nestgpu_server_submit_open": (ctypes.c_int, [_vp, _vp, _f64p, ctypes.c_int, _i64p]),
    "nestgpu_server_extend": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_double]),
    "nestgpu_server_close": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "nestgpu_server_wait_calls": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                                 ctypes.POINTER(ctypes.c_int), _i64p]),
    "nestgpu_server_stats": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "nestgpu_poisson_counts": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, ctypes.c_int64,
                                              ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint8)]),
    "nestgpu_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class NESTError(RuntimeError):
    """Raised when a kernel call fails; the message is the kernel's error text."""


def _i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64).ravel())


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def device_count():
    n = ctypes.c_int(0)
    _lib.nestgpu_device_count(ctypes.byref(n))
    return n.value


def poisson_counts(seed, lam, stream0, n, step0, n_steps):
    """Seeded Poisson counts [n_steps, n] (uint8) -- the drive of spike_generator(poisson_rate=...)."""
    out = np.zeros((n_steps, n), dtype=np.uint8)
    if _lib.nestgpu_poisson_counts(int(seed), float(lam), int(stream0), int(n), int(step0), int(n_steps),
                                   _ptr(out, ctypes.c_uint8)) != 0:
        raise NESTError(_lib.nestgpu_last_error(None).decode())
    return out


# ------------------------------------------------------------------ parameters
class Parameter:
    """nest.Parameter expression (pynest/nest/lib/hl_api_types.py) serialised as JSON."""

    def __init__(self, spec):
        self.spec = spec

    @staticmethod
    def _wrap(x):
        return x.spec if isinstance(x, Parameter) else float(x)

    def __add__(self, other):
        return Parameter({"sum": [self.spec, Parameter._wrap(other)]})

    __radd__ = __add__

    def __mul__(self, other):
        return Parameter({"product": [self.spec, Parameter._wrap(other)]})

    __rmul__ = __mul__

    def __sub__(self, other):
        return self + (-1.0) * other if isinstance(other, Parameter) else self + (-float(other))

    def __neg__(self):
        return self * -1.0

    def __truediv__(self, other):
        return Parameter({"quotient": [self.spec, Parameter._wrap(other)]})

    def __rtruediv__(self, lhs):
        return self**-1 * lhs  # as hl_api_types.py Parameter.__rtruediv__

    def __pow__(self, exponent):
        return Parameter({"pow": [self.spec, float(exponent)]})

    # comparisons give 1 or 0 (hl_api_types.py Parameter.__lt__ ..., comparator 0..5)
    def _compare(self, other, comparator):
        return Parameter({"compare": [self.spec, Parameter._wrap(other), comparator]})

    def __lt__(self, other):
        return self._compare(other, 0)

    def __le__(self, other):
        return self._compare(other, 1)

    def __eq__(self, other):
        return self._compare(other, 2)

    def __ne__(self, other):
        return self._compare(other, 3)

    def __ge__(self, other):
        return self._compare(other, 4)

    def __gt__(self, other):
        return self._compare(other, 5)

    __hash__ = object.__hash__


_param_ids = iter(range(1, 1 << 62))


class _Random:
    # normal and lognormal parameters carry an id: the library keeps one set of per-VP
    # distributions per parameter object, so an object used in several Connect calls continues
    # its draws from call to call as NEST's NormalParameter does (parameter.cpp:95-157)
    @staticmethod
    def uniform(min=0.0, max=1.0):
        return Parameter({"uniform": {"min": float(min), "max": float(max)}})

    @staticmethod
    def uniform_int(max):
        return Parameter({"uniform_int": {"max": int(max)}})

    @staticmethod
    def normal(mean=0.0, std=1.0):
        return Parameter({"normal": {"mean": float(mean), "std": float(std), "id": next(_param_ids)}})

    @staticmethod
    def lognormal(mean=0.0, std=1.0):
        return Parameter({"lognormal": {"mean": float(mean), "std": float(std), "id": next(_param_ids)}})

    @staticmethod
    def exponential(beta=1.0):
        return Parameter({"exponential": {"beta": float(beta)}})


class _Math:
    @staticmethod
    def exp(p):
        return Parameter({"exp": Parameter._wrap(p)})

    @staticmethod
    def sin(p):
        return Parameter({"sin": Parameter._wrap(p)})

    @staticmethod
    def cos(p):
        return Parameter({"cos": Parameter._wrap(p)})

    @staticmethod
    def max(p, value):
        return Parameter({"max": [Parameter._wrap(p), float(value)]})

    @staticmethod
    def min(p, value):
        return Parameter({"min": [Parameter._wrap(p), float(value)]})

    @staticmethod
    def redraw(p, min, max):
        return Parameter({"redraw": [Parameter._wrap(p), float(min), float(max)]})


class _Logic:
    @staticmethod
    def conditional(condition, param_if_true, param_if_false):
        """hl_api_logic.py conditional: the first parameter where the condition is nonzero."""
        return Parameter({"conditional": [Parameter._wrap(condition), Parameter._wrap(param_if_true),
                                          Parameter._wrap(param_if_false)]})


random = _Random()
math = _Math()
logic = _Logic()


def _to_json(d):
    def conv(x):
        if isinstance(x, Parameter):
            return x.spec
        if isinstance(x, np.ndarray):
            return x.tolist()
        if isinstance(x, (np.floating, np.integer)):
            x = x.item()
        if isinstance(x, float) and np.isinf(x):
            return 1e308 if x > 0 else -1e308  # JSON has no infinity; the library reads these as infinite
        if isinstance(x, dict):
            return {k: conv(v) for k, v in x.items()}
        if isinstance(x, (list, tuple)):
            return [conv(v) for v in x]
        return x

    return json.dumps(conv(d if d is not None else {}))


# node parameters whose value is itself a list (not one value per node)
_LIST_PARAMS = ("spike_times", "spike_multiplicities", "record_from")


def _process_syn_spec(syn_spec, conn_spec, prelength, postlength):
    """Array-valued syn_spec entries checked and laid out as PyNEST does
    (pynest/nest/lib/hl_api_connection_helpers.py:48-130): 1-D with one value per pair for
    one_to_one; 2-D [n_target, n_source] for all_to_all and [n_target, indegree] for
    fixed_indegree, flattened target-major."""
    if not syn_spec:
        return syn_spec
    rule = conn_spec if isinstance(conn_spec, str) else conn_spec.get("rule", "all_to_all")
    out = dict(syn_spec)
    for key, value in syn_spec.items():
        if isinstance(value, (list, tuple)):
            value = np.asarray(value)
        if not isinstance(value, np.ndarray) or value.ndim == 0:
            continue
        if value.ndim == 1:
            if rule != "one_to_one":
                raise TypeError(f"'{key}' has the wrong type. One-dimensional parameter arrays can only be used in "
                                "conjunction with rule 'one_to_one'.")
            if value.shape[0] != prelength:
                raise ValueError(f"'{key}' has to be an array of dimension {prelength}, a scalar or a dictionary.")
            out[key] = value
        elif value.ndim == 2:
            if rule == "all_to_all":
                if value.shape != (postlength, prelength):
                    raise ValueError(f"'{key}' has to be an array of dimension {postlength}x{prelength} "
                                     "(n_target x n_sources), a scalar or a dictionary.")
            elif rule == "fixed_indegree":
                indegree = conn_spec["indegree"]
                if value.shape != (postlength, indegree):
                    raise ValueError(f"'{key}' has to be an array of dimension {postlength}x{indegree} "
                                     "(n_target x indegree), a scalar or a dictionary.")
            else:
                raise TypeError(f"'{key}' has the wrong type. Two-dimensional parameter arrays can only be used in "
                                "conjunction with rules 'all_to_all' or 'fixed_indegree'.")
            out[key] = value.flatten()
        else:
            raise TypeError(f"'{key}' has the wrong type: parameter arrays have one or two dimensions.")
    return out


# ------------------------------------------------------------------ nodes
class NodeCollection:
    """Ids of nodes of one kernel (pynest NodeCollection subset)."""

    def __init__(self, kernel, ids):
        self._k = kernel
        self._ids = _i64(ids)

    def __len__(self):
        return len(self._ids)

    def __iter__(self):
        for i in self._ids:
            yield NodeCollection(self._k, [i])

    def __getitem__(self, key):
        if isinstance(key, (int, np.integer)):
            return NodeCollection(self._k, [self._ids[key]])
        return NodeCollection(self._k, self._ids[key])

    def __add__(self, other):
        return NodeCollection(self._k, np.concatenate([self._ids, other._ids]))

    def tolist(self):
        return [int(i) for i in self._ids]

    @property
    def global_id(self):
        return self._ids.copy()

    def get(self, key, output=None):
        if key == "events":
            return self._k._events(self)
        if key in ("record_from", "model"):
            vals = [self._k._node_status(i)[key] for i in self._ids]
            return vals[0] if len(vals) == 1 else vals
        if isinstance(key, (list, tuple)):
            return {k: self.get(k) for k in key}
        out = np.zeros(len(self._ids))
        self._k._chk(_lib.nestgpu_get_values(self._k._h, _ptr(self._ids, ctypes.c_int64), len(self._ids),
                                             key.encode(), _ptr(out, ctypes.c_double)))
        return out[0] if len(out) == 1 else out

    def set(self, params=None, **kwargs):
        if isinstance(params, (list, tuple)):
            # one dictionary per node (hl_api_types.py NodeCollection.set)
            if len(params) != len(self._ids):
                raise TypeError("list of params must have one dictionary per node")
            for i, d in enumerate(params):
                self[i].set(d)
            return
        params = dict(params or {}, **kwargs)
        scalar = {}
        for k, v in params.items():
            if np.ndim(v) > 0 and k not in _LIST_PARAMS:
                vals = np.ascontiguousarray(np.asarray(v, dtype=np.float64).ravel())
                if len(vals) != len(self._ids):
                    raise NESTError("value array length does not match the NodeCollection")
                self._k._chk(_lib.nestgpu_set_values(self._k._h, _ptr(self._ids, ctypes.c_int64), len(self._ids),
                                                     k.encode(), _ptr(vals, ctypes.c_double)))
            elif isinstance(v, Parameter):
                raise NESTError("random node parameters: draw them in Python and pass an array")
            else:
                scalar[k] = v
        if scalar:
            self._k._chk(_lib.nestgpu_set_status(self._k._h, _ptr(self._ids, ctypes.c_int64), len(self._ids),
                                                 _to_json(scalar).encode()))

    def __repr__(self):
        return "NodeCollection(ids=%s)" % (self._ids.tolist() if len(self._ids) < 8 else
                                           "%d..%d (%d)" % (self._ids[0], self._ids[-1], len(self._ids)))


# ------------------------------------------------------------------ kernel
class Kernel:
    """One simulation kernel (NEST has a single global kernel; here any number can coexist)."""

    def __init__(self, device=0):
        h = _vp()
        if _lib.nestgpu_kernel_new(int(device), ctypes.byref(h)) != 0:
            raise NESTError(_lib.nestgpu_last_error(None).decode())
        self._h = h
        self.device = device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.nestgpu_kernel_free(h)
            self._h = None

    def _chk(self, rc):
        if rc != 0:
            raise NESTError(_lib.nestgpu_last_error(self._h).decode())

    def ResetKernel(self):
        self._chk(_lib.nestgpu_reset_kernel(self._h))

    def SetKernelStatus(self, params):
        self._chk(_lib.nestgpu_set_kernel_status(self._h, _to_json(params).encode()))

    def GetKernelStatus(self, key=None):
        # a served kernel changes between calls: retry until the text fits
        n = ctypes.c_size_t(0)
        self._chk(_lib.nestgpu_get_kernel_status(self._h, None, 0, ctypes.byref(n)))
        while True:
            cap = n.value + 64
            buf = ctypes.create_string_buffer(cap)
            self._chk(_lib.nestgpu_get_kernel_status(self._h, buf, cap, ctypes.byref(n)))
            if n.value < cap:
                break
        d = json.loads(buf.value.decode())
        # hl_api_simulation.py GetKernelStatus: one key, or a list of keys giving a tuple
        if key is None:
            return d
        if isinstance(key, str):
            return d[key]
        if isinstance(key, (list, tuple)):
            return tuple(d[k] for k in key)
        raise TypeError("keys should be either a string or an iterable")

    def Create(self, model, n=1, params=None):
        """hl_api_nodes.py Create: params is a dict (scalars, or arrays with one value per
        node) or a list of dicts, one per node."""
        if isinstance(n, dict):
            if params is not None:
                raise ValueError("A parameter dictionary can be passed as second argument only of Create() is "
                                 "called with two arguments.")
            params, n = n, 1
        if int(n) != n:
            raise TypeError("n must have an integer value")
        n = int(n)
        if isinstance(params, (list, tuple)) and len(params) != n:
            raise TypeError("list of params must have one dictionary per node")
        if params is not None and not (isinstance(params, dict) or (isinstance(params, (list, tuple))
                                                                    and all(isinstance(e, dict) for e in params))):
            raise TypeError("params must be either a dict of parameters or a list or tuple of dicts")
        per_node = isinstance(params, (list, tuple)) or (isinstance(params, dict) and any(
            np.ndim(v) > 0 and k not in _LIST_PARAMS for k, v in params.items()))
        first = ctypes.c_int64(0)
        self._chk(_lib.nestgpu_create(self._h, model.encode(), n,
                                      _to_json({} if per_node else (params or {})).encode(), ctypes.byref(first)))
        nc = NodeCollection(self, np.arange(first.value, first.value + n))
        if per_node:
            nc.set(params)
        return nc

    def Connect(self, pre, post, conn_spec=None, syn_spec=None):
        pre_ids = pre._ids if isinstance(pre, NodeCollection) else _i64(pre)
        post_ids = post._ids if isinstance(post, NodeCollection) else _i64(post)
        if conn_spec is None:
            conn_spec = "all_to_all"
        if isinstance(syn_spec, str):
            syn_spec = {"synapse_model": syn_spec}
        syn_spec = _process_syn_spec(syn_spec, conn_spec, len(pre_ids), len(post_ids))
        self._chk(_lib.nestgpu_connect(self._h, _ptr(pre_ids, ctypes.c_int64), len(pre_ids),
                                       _ptr(post_ids, ctypes.c_int64), len(post_ids),
                                       _to_json(conn_spec).encode(), _to_json(syn_spec or {}).encode()))

    def ConnectArrays(self, sources, targets, weights=None, delays=None):
        s, t = _i64(sources), _i64(targets)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        d = None if delays is None else np.ascontiguousarray(delays, dtype=np.float64)
        self._chk(_lib.nestgpu_connect_arrays(self._h, _ptr(s, ctypes.c_int64), _ptr(t, ctypes.c_int64),
                                              None if w is None else _ptr(w, ctypes.c_double),
                                              None if d is None else _ptr(d, ctypes.c_double), len(s)))

    def SetSpikeCounts(self, generators, counts, first_step=0):
        c = np.ascontiguousarray(counts, dtype=np.uint8)
        ids = generators._ids
        if c.ndim != 2 or c.shape[1] != len(ids) or np.any(np.diff(ids) != 1):
            raise NESTError("counts must be [n_steps, n_generators] for contiguous generator ids")
        self._chk(_lib.nestgpu_set_spike_counts(self._h, int(ids[0]), len(ids), int(first_step), c.shape[0],
                                                _ptr(c, ctypes.c_uint8)))

    def GetConnections(self, source=None, target=None, synapse_model=None):
        """Connections in creation order (VP, then creation within the VP), as arrays. source,
        target: node collections (or ids) that the connections' ends must belong to;
        synapse_model: "static_synapse" or "stdp_synapse" (hl_api_connections.py GetConnections);
        "synapse_model" holds each connection's model."""
        c = self._all_connections()
        keep = np.ones(len(c["source"]), bool)
        if source is not None:
            keep &= np.isin(c["source"], source._ids if isinstance(source, NodeCollection) else _i64(source))
        if target is not None:
            keep &= np.isin(c["target"], target._ids if isinstance(target, NodeCollection) else _i64(target))
        if synapse_model is not None:
            name = "static_synapse" if synapse_model == "static_synapse_hpc" else synapse_model
            keep &= c["synapse_model"] == name
        if keep.all():
            return c
        return {key: v[keep] for key, v in c.items()}

    def _all_connections(self):
        n = _lib.nestgpu_num_connections(self._h)
        s = np.zeros(n, np.int64)
        t = np.zeros(n, np.int64)
        w = np.zeros(n)
        d = np.zeros(n, np.int64)
        vp = np.zeros(n, np.int32)
        self._chk(_lib.nestgpu_get_connections(self._h, _ptr(s, ctypes.c_int64), _ptr(t, ctypes.c_int64),
                                               _ptr(w, ctypes.c_double), _ptr(d, ctypes.c_int64),
                                               _ptr(vp, ctypes.c_int32)))
        m = np.zeros(n, np.int32)
        self._chk(_lib.nestgpu_get_connection_models(self._h, _ptr(m, ctypes.c_int32)))
        return {"source": s, "target": t, "weight": w, "delay_steps": d, "vp": vp,
                "synapse_model": np.where(m == 1, "stdp_synapse", "static_synapse")}

    def Prepare(self):
        self._chk(_lib.nestgpu_prepare(self._h))

    def Run(self, t):
        self._chk(_lib.nestgpu_run(self._h, float(t)))

    def Cleanup(self):
        self._chk(_lib.nestgpu_cleanup(self._h))

    def Simulate(self, t):
        self._chk(_lib.nestgpu_simulate(self._h, float(t)))

    def _node_status(self, node_id):
        n = ctypes.c_size_t(0)
        self._chk(_lib.nestgpu_node_status(self._h, int(node_id), None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        self._chk(_lib.nestgpu_node_status(self._h, int(node_id), buf, n.value + 1, ctypes.byref(n)))
        return json.loads(buf.value.decode())

    def _events(self, rec):
        rid = int(rec._ids[0])
        st = self._node_status(rid)
        if st["model"] in ("multimeter", "voltmeter"):
            names = st["record_from"]
            n = ctypes.c_int64(0)
            self._chk(_lib.nestgpu_multimeter_events(self._h, rid, None, None, None, 0, ctypes.byref(n)))
            s = np.zeros(n.value, np.int64)
            tm = np.zeros(n.value)
            vals = np.zeros((n.value, max(1, len(names))))
            self._chk(_lib.nestgpu_multimeter_events(self._h, rid, _ptr(s, ctypes.c_int64), _ptr(tm, ctypes.c_double),
                                                     _ptr(vals, ctypes.c_double), n.value, ctypes.byref(n)))
            ev = {"senders": s, "times": tm}
            for j, name in enumerate(names):
                ev[name] = vals[:, j].copy()
            return ev
        n = ctypes.c_int64(0)
        self._chk(_lib.nestgpu_recorder_events(self._h, rid, None, None, 0, ctypes.byref(n)))
        s = np.zeros(n.value, np.int64)
        tm = np.zeros(n.value)
        self._chk(_lib.nestgpu_recorder_events(self._h, rid, _ptr(s, ctypes.c_int64), _ptr(tm, ctypes.c_double),
                                               n.value, ctypes.byref(n)))
        return {"senders": s, "times": tm}


def batch_simulate(kernels, times):
    """Advance several kernels on one device together (batched form)."""
    arr = (_vp * len(kernels))(*[k._h for k in kernels])
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(times, dtype=np.float64), (len(kernels),)))
    if _lib.nestgpu_batch_simulate(arr, len(kernels), _ptr(t, ctypes.c_double)) != 0:
        raise NESTError(_lib.nestgpu_last_error(None).decode())


def batch_simulate_noalloc(kernels, times):
    """batch_simulate for prepared kernels that allocates no device memory (the launch
    records and status words live in a fixed device array of the library, at most 64 kernels
    per call). Calls on one device must not overlap."""
    arr = (_vp * len(kernels))(*[k._h for k in kernels])
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(times, dtype=np.float64), (len(kernels),)))
    if _lib.nestgpu_batch_simulate_noalloc(arr, len(kernels), _ptr(t, ctypes.c_double)) != 0:
        raise NESTError(_lib.nestgpu_last_error(None).decode())


def batch_autotune(kernels, probe_ms=20.0, candidates=(64, 128, 256, 512), window_caps=(0,), with_windows=False):
    """Pick targets_per_block and the cap on LDS ring slots (0: as many as fit) per class
    (neuron models, size octave) of not yet simulated kernels from probe runs (discarded) of
    every pair, scored by block-time. Returns the targets_per_block chosen for each kernel, or
    with with_windows the list of (targets_per_block, window cap) pairs."""
    arr = (_vp * len(kernels))(*[k._h for k in kernels])
    cand = (ctypes.c_int * len(candidates))(*[int(c) for c in candidates])
    wc = (ctypes.c_int * len(window_caps))(*[int(c) for c in window_caps])
    out = (ctypes.c_int * len(kernels))()
    out_w = (ctypes.c_int * len(kernels))()
    if _lib.nestgpu_batch_autotune(arr, len(kernels), float(probe_ms), cand, len(candidates), wc, len(window_caps),
                                   out, out_w) != 0:
        raise NESTError(_lib.nestgpu_last_error(None).decode())
    return list(zip(out, out_w)) if with_windows else list(out)


class Server:
    """Served form: continuous batching of simulate requests on one device within an HBM
    budget. submit() takes a kernel built on the host (not yet simulated) and its Simulate
    durations; the kernel must not be used until wait() reports it done."""

    QUEUED, ADMITTED, RUNNING, DONE, FAILED = 0, 1, 2, 3, -1

    def __init__(self, device=0, hbm_budget_bytes=64 << 30, slices_per_chunk=16):
        h = _vp()
        if _lib.nestgpu_server_new(int(device), int(hbm_budget_bytes), int(slices_per_chunk), ctypes.byref(h)) != 0:
            raise NESTError(_lib.nestgpu_last_error(None).decode())
        self._h = h
        self._kernels = {}

    def close(self):
        if getattr(self, "_h", None):
            _lib.nestgpu_server_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, kernel, times):
        t = np.ascontiguousarray(np.atleast_1d(np.asarray(times, dtype=np.float64)))
        rid = ctypes.c_int64(0)
        if _lib.nestgpu_server_submit(self._h, kernel._h, _ptr(t, ctypes.c_double), len(t), ctypes.byref(rid)) != 0:
            raise NESTError(_lib.nestgpu_last_error(None).decode())
        self._kernels[rid.value] = kernel
        return rid.value

    def submit_open(self, kernel, times=()):
        """Open request (a session): the network stays resident after its calls until
        close_request(); extend() adds Simulate calls."""
        t = np.ascontiguousarray(np.atleast_1d(np.asarray(times, dtype=np.float64)))
        rid = ctypes.c_int64(0)
        if _lib.nestgpu_server_submit_open(self._h, kernel._h, _ptr(t, ctypes.c_double), len(t),
                                           ctypes.byref(rid)) != 0:
            raise NESTError(_lib.nestgpu_last_error(None).decode())
        self._kernels[rid.value] = kernel
        return rid.value

    def extend(self, rid, t):
        if _lib.nestgpu_server_extend(self._h, int(rid), float(t)) != 0:
            raise NESTError(_lib.nestgpu_last_error(None).decode())

    def close_request(self, rid):
        if _lib.nestgpu_server_close(self._h, int(rid)) != 0:
            raise NESTError(_lib.nestgpu_last_error(None).decode())

    def wait_calls(self, rid, n, timeout=-1.0):
        """Waits until the request has completed n Simulate calls; returns (state, calls done)."""
        st = ctypes.c_int(0)
        nd = ctypes.c_int64(0)
        if _lib.nestgpu_server_wait_calls(self._h, int(rid), int(n), float(timeout), ctypes.byref(st),
                                          ctypes.byref(nd)) != 0:
            raise NESTError(_lib.nestgpu_last_error(None).decode())
        return st.value, nd.value

    def wait(self, rid, timeout=-1.0):
        """Returns dict(state, t_submit, t_admit, t_start, t_done, events); raises if failed."""
        st = ctypes.c_int(0)
        tm = np.zeros(4)
        ev = ctypes.c_uint64(0)
        if _lib.nestgpu_server_wait(self._h, int(rid), float(timeout), ctypes.byref(st), _ptr(tm, ctypes.c_double),
                                    ctypes.byref(ev)) != 0:
            raise NESTError(_lib.nestgpu_last_error(None).decode())
        if st.value == self.FAILED:
            raise NESTError("request %d failed: %s" % (rid, _lib.nestgpu_last_error(None).decode()))
        return {"state": st.value, "t_submit": tm[0], "t_admit": tm[1], "t_start": tm[2], "t_done": tm[3],
                "events": ev.value}

    def stats(self):
        # the server thread changes the figures between calls: retry until the text fits
        n = ctypes.c_size_t(0)
        _lib.nestgpu_server_stats(self._h, None, 0, ctypes.byref(n))
        while True:
            cap = n.value + 64
            buf = ctypes.create_string_buffer(cap)
            if _lib.nestgpu_server_stats(self._h, buf, cap, ctypes.byref(n)) != 0:
                raise NESTError(_lib.nestgpu_last_error(None).decode())
            if n.value < cap:
                return json.loads(buf.value.decode())


# ------------------------------------------------------------------ module-level kernel (PyNEST style)
_default = None


def _k():
    global _default
    if _default is None:
        _default = Kernel(0)
    return _default


def ResetKernel():
    _k().ResetKernel()


def SetKernelStatus(params):
    _k().SetKernelStatus(params)


def GetKernelStatus(key=None):
    return _k().GetKernelStatus(key)


def Create(model, n=1, params=None):
    return _k().Create(model, n, params)


def Connect(pre, post, conn_spec=None, syn_spec=None):
    _k().Connect(pre, post, conn_spec, syn_spec)


def GetConnections(source=None, target=None, synapse_model=None):
    return _k().GetConnections(source, target, synapse_model)


def Simulate(t):
    _k().Simulate(t)


def Prepare():
    _k().Prepare()


def Run(t):
    _k().Run(t)


def Cleanup():
    _k().Cleanup()
