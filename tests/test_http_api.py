"""nest-server protocol (pynest/nest/server/hl_api_server.py) over the library: routes,
argument forms, node collections as id lists, serialize_data and the error reply. These
calls build networks on the host only; Simulate through the served form is a GPU test
(tests/test_gpu_serve.py)."""

import json
import urllib.error
import urllib.request

import numpy as np
import pytest

from nestgpu import http_api


@pytest.fixture
def server():
    srv = http_api.NESTServer(port=0).start()
    yield srv
    srv.stop()


def _call(srv, route, payload=None, session=None, method="POST"):
    url = "http://127.0.0.1:%d%s" % (srv.port, route)
    data = None if payload is None else json.dumps(payload).encode()
    req = urllib.request.Request(url, data=data, method=method if data is not None else "GET")
    if data is not None:
        req.add_header("Content-Type", "application/json")
    if session:
        req.add_header("NEST-Session", session)
    try:
        with urllib.request.urlopen(req) as r:
            return r.status, json.loads(r.read().decode())
    except urllib.error.HTTPError as e:
        return e.code, json.loads(e.read().decode())


def test_index_and_api_listing(server):
    code, body = _call(server, "/")
    assert code == 200 and body["mpi"] is False and "nest" in body
    code, calls = _call(server, "/api")
    assert code == 200
    for c in ("Create", "Connect", "Simulate", "GetStatus", "SetStatus", "GetKernelStatus", "ResetKernel"):
        assert c in calls
    assert calls == sorted(calls)


def test_create_connect_status_roundtrip(server):
    code, nodes = _call(server, "/api/Create", {"model": "iaf_psc_exp", "n": 5, "params": {"I_e": 300.0}})
    assert code == 200 and nodes == [1, 2, 3, 4, 5]
    # positional arguments in "args"; node collections travel as id lists
    code, _ = _call(server, "/api/Connect", {"args": [[1, 2], [3, 4, 5]], "conn_spec": {"rule": "all_to_all"},
                                             "syn_spec": {"weight": 2.5, "delay": 1.0}})
    assert code == 200
    code, conns = _call(server, "/api/GetKernelStatus", {"keys": "num_connections"})
    assert code == 200 and conns == 6
    code, vals = _call(server, "/api/GetStatus", {"nodes": [1, 3], "keys": "I_e"})
    assert code == 200 and vals == [300.0, 300.0]
    code, _ = _call(server, "/api/SetStatus", {"nodes": [3], "params": {"V_m": -60.0}})
    assert code == 200
    code, vals = _call(server, "/api/GetStatus", {"nodes": [3], "keys": ["V_m", "I_e"]})
    assert code == 200 and vals == [[-60.0, 300.0]]
    code, st = _call(server, "/api/GetStatus", {"nodes": [1]})
    assert code == 200 and st[0]["model"] == "iaf_psc_exp" and st[0]["global_id"] == 1


def test_errors_use_the_reference_reply(server):
    # get_or_error (hl_api_server.py:431-461): HTTP 400, "<name> (NESTError): <detail>"
    code, body = _call(server, "/api/Create", {"model": "no_such_model"})
    assert code == 400 and body["message"].startswith("UnknownModelName (NESTError):")
    code, body = _call(server, "/api/NoSuchCall", {})
    assert code == 400 and "has no attribute" in body["message"]
    code, body = _call(server, "/exec", {"source": "print(1)"})
    assert code == 403


def test_sessions_are_isolated(server):
    _call(server, "/api/Create", {"model": "iaf_psc_delta", "n": 3}, session="a")
    _call(server, "/api/Create", {"model": "iaf_psc_delta", "n": 7}, session="b")
    assert _call(server, "/api/GetKernelStatus", {"keys": "network_size"}, session="a")[1] == 3
    assert _call(server, "/api/GetKernelStatus", {"keys": "network_size"}, session="b")[1] == 7
    _call(server, "/api/ResetKernel", {}, session="a")
    assert _call(server, "/api/GetKernelStatus", {"keys": "network_size"}, session="a")[1] == 0


def test_serialize_data_forms():
    out = http_api.serialize_data({"a": np.arange(3), "b": np.float64(1.5), "c": (np.int64(2), np.bool_(True))})
    assert out == {"a": [0, 1, 2], "b": 1.5, "c": [2, True]}
