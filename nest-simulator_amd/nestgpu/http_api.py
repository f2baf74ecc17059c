"""nest-server wire protocol over the served form.

Clients of the reference's nest-server (pynest/nest/server/hl_api_server.py) call PyNEST
functions over HTTP: ``GET /`` (version info, :180-187), ``GET /api`` (the callable names,
:295-299), ``GET|POST /api/<call>`` with the arguments as JSON (:301-310, get_arguments
:369-390), node collections passed as lists of node ids and turned back into collections for
the parameters named nodes/source/target/pre/post (nestify :501-513), and every response
passed through serialize_data (pynest/nest/lib/hl_api_types.py:1260-1290). A failing call
answers HTTP 400 with {"message": "<error name> (NESTError): <detail>"} (get_or_error
:431-461). ``/exec`` runs scripts in the reference only when an administrator enables it; it is
disabled here (403), as by default there.

This server keeps that protocol and routes Simulate through the served form
(nestgpu.Server): each client session's network is built on the host by Create/Connect, then
admitted as an open request whose state stays resident in HBM across Simulate calls, and all
sessions' Simulate calls advance together in the same slice launches (continuous batching).
The reference serves one kernel; a request header ``NEST-Session: <name>`` selects an
isolated kernel here (without it every client shares the session "default", as in the
reference). After the first Simulate of a session its network can be read (GetStatus,
GetKernelStatus, events) but not changed; ResetKernel releases it.
"""

import json
import threading
import traceback
from http.server import BaseHTTPRequestHandler, ThreadingHTTPServer
from urllib.parse import parse_qs, urlparse

import numpy as np

from . import Kernel, NESTError, NodeCollection, Server

VERSION = "nestgpu-mi355x"


def serialize_data(data):
    """JSON-ready form of a response (hl_api_types.py:1260-1290)."""
    if isinstance(data, NodeCollection):
        return [int(i) for i in data.global_id]
    if isinstance(data, np.ndarray):
        return serialize_data(data.tolist())
    if isinstance(data, np.bool_):
        return bool(data)
    if isinstance(data, np.integer):
        return int(data)
    if isinstance(data, np.floating):
        return float(data)
    if isinstance(data, (list, tuple)):
        return [serialize_data(d) for d in data]
    if isinstance(data, dict):
        return {serialize_data(k): serialize_data(v) for k, v in data.items()}
    return data


class Session:
    """One client's kernel; after its first Simulate, an open request of the served form."""

    def __init__(self, front):
        self.front = front
        self.lock = threading.Lock()
        self.kernel = Kernel(front.device)
        self.rid = None
        self.calls = 0

    def release(self):
        if self.rid is not None:
            srv = self.front.served()
            srv.close_request(self.rid)
            srv.wait(self.rid)
            self.rid = None
            self.calls = 0

    # ---- the PyNEST calls a client can make (hl_api_*.py names and argument order)
    def _frozen(self, what):
        if self.rid is not None:
            raise NESTError("NotImplemented: %s after the first Simulate of a served session; ResetKernel starts a "
                            "new network" % what)

    def _nodes(self, x):
        if x is None or isinstance(x, NodeCollection):
            return x
        return NodeCollection(self.kernel, np.atleast_1d(np.asarray(x, dtype=np.int64)))

    def ResetKernel(self):
        self.release()
        self.kernel = Kernel(self.front.device)

    def SetKernelStatus(self, params):
        self._frozen("SetKernelStatus")
        self.kernel.SetKernelStatus(params)

    def GetKernelStatus(self, keys=None):
        return self.kernel.GetKernelStatus(keys)

    def Create(self, model, n=1, params=None, positions=None):
        self._frozen("Create")
        if positions is not None:
            raise NESTError("NotImplemented: spatial node positions")
        return self.kernel.Create(model, int(n), params)

    def Connect(self, pre, post, conn_spec=None, syn_spec=None, return_synapsecollection=False):
        self._frozen("Connect")
        self.kernel.Connect(self._nodes(pre), self._nodes(post), conn_spec, syn_spec)

    def SetStatus(self, nodes, params, val=None):
        self._frozen("SetStatus")
        nc = self._nodes(nodes)
        if isinstance(params, str):
            params = {params: val}
        if isinstance(params, (list, tuple)):
            for i, d in enumerate(params):
                nc[i].set(d)
        else:
            nc.set(params)

    def GetStatus(self, nodes, keys=None, output=""):
        nc = self._nodes(nodes)
        out = []
        for i in range(len(nc)):
            one = nc[i]
            if keys is None:
                out.append(self.kernel._node_status(int(one.global_id[0])))
            elif isinstance(keys, str):
                out.append(one.get(keys))
            else:
                out.append([one.get(k) for k in keys])
        return out

    def GetConnections(self, source=None, target=None, synapse_model=None, synapse_label=None):
        return self.kernel.GetConnections(self._nodes(source), self._nodes(target), synapse_model)

    def Simulate(self, t):
        t = float(t)
        srv = self.front.served()
        if self.rid is None:
            self.rid = srv.submit_open(self.kernel, [t])
        else:
            srv.extend(self.rid, t)
        self.calls += 1
        srv.wait_calls(self.rid, self.calls)


CALLS = sorted(name for name in dir(Session) if name[0].isupper())


class NESTServer:
    """HTTP front end (nest-server's routes) over one served-form Server."""

    def __init__(self, host="127.0.0.1", port=52425, device=0, hbm_budget_bytes=64 << 30, slices_per_chunk=16):
        self.device = device
        self._served_args = (device, hbm_budget_bytes, slices_per_chunk)
        self._served = None  # the served form starts with the first Simulate
        self.sessions = {}
        self.lock = threading.Lock()
        front = self

        class Handler(BaseHTTPRequestHandler):
            def log_message(self, fmt, *args):  # quiet
                pass

            def _reply(self, code, obj):
                body = json.dumps(obj).encode()
                self.send_response(code)
                self.send_header("Content-Type", "application/json")
                self.send_header("Content-Length", str(len(body)))
                self.end_headers()
                self.wfile.write(body)

            def _arguments(self):
                # get_arguments (hl_api_server.py:369-390): a JSON string, list or object body,
                # else form or query fields; "args" holds the positional arguments
                args, kwargs = [], {}
                n = int(self.headers.get("Content-Length") or 0)
                body = self.rfile.read(n) if n else b""
                ctype = self.headers.get("Content-Type", "")
                if body and "json" in ctype:
                    data = json.loads(body.decode())
                    if isinstance(data, str) and data:
                        args = [data]
                    elif isinstance(data, list):
                        args = data
                    elif isinstance(data, dict):
                        kwargs = data
                        if "args" in kwargs:
                            args = kwargs.pop("args")
                else:
                    fields = parse_qs(body.decode()) if body else parse_qs(urlparse(self.path).query)
                    if "args" in fields:
                        args = fields["args"]
                    else:
                        kwargs = {k: v[0] for k, v in fields.items()}
                return list(args), kwargs

            def _handle(self):
                path = urlparse(self.path).path.rstrip("/")
                if path == "":
                    return self._reply(200, {"nest": VERSION, "mpi": False})
                if path == "/api":
                    return self._reply(200, CALLS)
                if path == "/exec":
                    return self._reply(403, {"message": "The route `/exec` has been disabled. Please contact the "
                                                        "server administrator."})
                if not path.startswith("/api/"):
                    return self._reply(404, {"message": "unknown route %s" % path})
                call = path[len("/api/"):]
                if call not in CALLS:
                    return self._reply(400, {"message": "AttributeError: module 'nest' has no attribute '%s'" % call})
                try:
                    args, kwargs = self._arguments()
                    sess = front.session(self.headers.get("NEST-Session", "default"))
                    with sess.lock:
                        result = getattr(sess, call)(*args, **kwargs)
                    return self._reply(200, serialize_data(result))
                except NESTError as err:
                    text = str(err)
                    name, _, detail = text.partition(": ")
                    if not detail:
                        name, detail = "NESTError", text
                    return self._reply(400, {"message": "%s (NESTError): %s" % (name, detail)})
                except (KeyError, TypeError, ValueError) as err:
                    return self._reply(400, {"message": "%s: %s" % (err.__class__.__name__, err.args[0] if err.args
                                                                   else "")})
                except Exception as err:  # noqa: BLE001  (reported to the client, as the reference does)
                    traceback.print_exc()
                    return self._reply(400, {"message": "%s: %s" % (err.__class__.__name__, err)})

            do_GET = _handle
            do_POST = _handle

        self.httpd = ThreadingHTTPServer((host, port), Handler)
        self.httpd.daemon_threads = True
        self.port = self.httpd.server_address[1]
        self._thread = None

    def session(self, name):
        with self.lock:
            if name not in self.sessions:
                self.sessions[name] = Session(self)
            return self.sessions[name]

    def served(self):
        with self.lock:
            if self._served is None:
                self._served = Server(*self._served_args)
            return self._served

    def start(self):
        self._thread = threading.Thread(target=self.httpd.serve_forever, daemon=True)
        self._thread.start()
        return self

    def stop(self):
        self.httpd.shutdown()
        self.httpd.server_close()
        for s in list(self.sessions.values()):
            try:
                s.release()
            except NESTError:
                pass
        self.sessions.clear()
        if self._served is not None:
            self._served.close()
            self._served = None


def main():
    import argparse

    ap = argparse.ArgumentParser(description="nest-server protocol over the MI355X served form")
    ap.add_argument("--host", default="127.0.0.1")
    ap.add_argument("--port", type=int, default=52425)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--hbm-budget-gib", type=float, default=64.0)
    a = ap.parse_args()
    srv = NESTServer(a.host, a.port, a.device, int(a.hbm_budget_gib * (1 << 30)))
    print("nest-server protocol on http://%s:%d (device %d)" % (a.host, srv.port, a.device), flush=True)
    srv.httpd.serve_forever()


if __name__ == "__main__":
    main()
