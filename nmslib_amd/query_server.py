"""A QueryService server that speaks the reference's Thrift wire protocol.

The reference serves an index over Thrift (query_server/protocol.thrift:18-54;
the C++ server query_server/cpp_client_server/QueryService_server.cpp runs
TBinaryProtocol over a buffered, unframed TCP transport, :657-676).  This
module answers the same service on the same wire, so the reference's clients
(query_client.py, QueryClient.cpp, the Java client) talk to it unchanged:

  setQueryTimeParams(1: string)                       (:161-191)
  knnQuery(1: i32 k, 2: binary queryObj, 3: bool retExternId, 4: bool retObj)
                                                      (:320-405)
  knnQueryBatch(1: i32 k, 2: list<binary> queryObj, 3: bool retExternId,
                4: bool retObj, 5: i32 numThreads)    (:411-466)
  rangeQuery(...)  -> QueryException "Range search is not supported!", as the
                      reference's Hnsw::Search(RangeQuery) throws (hnsw.cc:714)
  getDistance(1: binary obj1, 2: binary obj2)         (:276-312)

Every failure comes back as the IDL's QueryException(1: string message).  A
query object is the text form the reference parses (VectorSpace::ReadVec,
space_vector.cc:124-137: an optional "label:N " prefix, numbers separated by
spaces, commas or colons, read with strtof).  Results are ascending by
distance; ``id`` is the object id, ``externId`` its external id (given to the
server, or the id in decimal), ``obj`` the vector's text form
(VectorSpace::CreateStrFromObj, :96-108: %.9g numbers joined by spaces).

Searches go to the device through the continuous-batching loop
(nmslib_amd.serving): queries arriving on any connection are collected by one
dispatcher thread into batches for the server's lanes, so concurrent clients
share device launches.  As in the reference, setQueryTimeParams waits until
no query is in flight (:161-191).

Run: ``python -m nmslib_amd.query_server -p 10000 -s l2 -L index.bin
[--load-data] [-t efSearch=100]``.
"""
from __future__ import annotations

import argparse
import ctypes
import errno
import io
import queue
import socketserver
import struct
import sys
import threading

import numpy as np

# Thrift binary protocol (big-endian): field / element type codes and message
# types
T_STOP, T_BOOL, T_BYTE, T_DOUBLE, T_I16, T_I32, T_I64, T_STRING, T_STRUCT = 0, 2, 3, 4, 6, 8, 10, 11, 12
T_MAP, T_SET, T_LIST = 13, 14, 15
M_CALL, M_REPLY, M_EXCEPTION, M_ONEWAY = 1, 2, 3, 4
VERSION_1 = 0x80010000
# TApplicationException types
APP_UNKNOWN_METHOD, APP_PROTOCOL_ERROR = 1, 7


class ProtocolError(Exception):
    pass


class Reader:
    """Reads Thrift binary-protocol values from a byte stream."""

    def __init__(self, stream):
        self.s = stream

    def _take(self, n: int) -> bytes:
        b = self.s.read(n)
        if len(b) != n:
            raise EOFError("connection closed")
        return b

    def i8(self) -> int:
        return struct.unpack(">b", self._take(1))[0]

    def i16(self) -> int:
        return struct.unpack(">h", self._take(2))[0]

    def i32(self) -> int:
        return struct.unpack(">i", self._take(4))[0]

    def i64(self) -> int:
        return struct.unpack(">q", self._take(8))[0]

    def double(self) -> float:
        return struct.unpack(">d", self._take(8))[0]

    def binary(self) -> bytes:
        n = self.i32()
        if n < 0:
            raise ProtocolError("negative string length")
        return self._take(n)

    def message_begin(self):
        """(name, type, seqid): the strict header, or the old unversioned one."""
        first = self.i32()
        if first < 0:
            if (first & 0xFFFF0000) != VERSION_1:
                raise ProtocolError("bad protocol version")
            mtype = first & 0xFF
            name = self.binary().decode()
            return name, mtype, self.i32()
        name = self._take(first).decode()
        mtype = self.i8()
        return name, mtype, self.i32()

    def value(self, ttype: int):
        """Any value, generically: structs become {field id: value}."""
        if ttype == T_BOOL:
            return self.i8() != 0
        if ttype == T_BYTE:
            return self.i8()
        if ttype == T_I16:
            return self.i16()
        if ttype == T_I32:
            return self.i32()
        if ttype == T_I64:
            return self.i64()
        if ttype == T_DOUBLE:
            return self.double()
        if ttype == T_STRING:
            return self.binary()
        if ttype == T_STRUCT:
            return self.struct()
        if ttype in (T_LIST, T_SET):
            et, n = self.i8(), self.i32()
            return [self.value(et) for _ in range(n)]
        if ttype == T_MAP:
            kt, vt, n = self.i8(), self.i8(), self.i32()
            return {self.value(kt): self.value(vt) for _ in range(n)}
        raise ProtocolError(f"unknown type {ttype}")

    def struct(self) -> dict:
        out = {}
        while True:
            ft = self.i8()
            if ft == T_STOP:
                return out
            fid = self.i16()
            out[fid] = (ft, self.value(ft))


class Writer:
    """Writes Thrift binary-protocol values into a buffer."""

    def __init__(self):
        self.b = io.BytesIO()

    def getvalue(self) -> bytes:
        return self.b.getvalue()

    def i8(self, v):
        self.b.write(struct.pack(">b", v))

    def i16(self, v):
        self.b.write(struct.pack(">h", v))

    def i32(self, v):
        self.b.write(struct.pack(">i", v))

    def double(self, v):
        self.b.write(struct.pack(">d", v))

    def binary(self, v: bytes):
        self.i32(len(v))
        self.b.write(v)

    def message_begin(self, name: str, mtype: int, seqid: int):
        self.i32((VERSION_1 | mtype) - (1 << 32))
        self.binary(name.encode())
        self.i32(seqid)

    def field(self, ftype: int, fid: int):
        self.i8(ftype)
        self.i16(fid)

    def stop(self):
        self.i8(T_STOP)

    def reply_entry(self, e):
        """ReplyEntry {1: i32 id, 2: double dist, 3: optional string externId,
        4: optional binary obj} (protocol.thrift:4-9)."""
        self.field(T_I32, 1)
        self.i32(e[0])
        self.field(T_DOUBLE, 2)
        self.double(e[1])
        if e[2] is not None:
            self.field(T_STRING, 3)
            self.binary(e[2].encode())
        if e[3] is not None:
            self.field(T_STRING, 4)
            self.binary(e[3].encode())
        self.stop()

    def entry_list(self, entries):
        self.i8(T_STRUCT)
        self.i32(len(entries))
        for e in entries:
            self.reply_entry(e)


class QueryException(Exception):
    """The IDL's QueryException (protocol.thrift:14-16)."""


# --- the reference's text forms ------------------------------------------------

_libc = ctypes.CDLL(None, use_errno=True)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]


def parse_vector(s: bytes) -> np.ndarray:
    """VectorSpace::ReadVec (space_vector.cc:124-137): an optional "label:N "
    prefix (Object::extractLabel, object.h:118-146), then ',' and ':' become
    spaces (ReplaceSomePunct, utils.h:341-344) and numbers are read with
    strtof until one fails to parse (ReadVecDataEfficiently)."""
    text = s.decode("latin-1")
    if text.startswith("label:") and len(text) > len("label:") + 1:
        rest = text[len("label:"):]
        for i, ch in enumerate(rest):
            if ch.isspace():
                text = rest[i:].lstrip()
                break
    text = text.replace(",", " ").replace(":", " ")
    buf = ctypes.create_string_buffer(text.encode("latin-1"))
    base = ctypes.addressof(buf)
    ptr = ctypes.c_char_p(base)
    end = ctypes.c_char_p()
    out = []
    pos = base
    ctypes.set_errno(0)
    while True:
        val = _libc.strtof(ctypes.c_char_p(pos), ctypes.byref(end))
        e = ctypes.cast(end, ctypes.c_void_p).value
        if e == pos:
            break
        if ctypes.get_errno() == errno.ERANGE:
            raise QueryException(f"Failed to parse the line: '{s.decode('latin-1')}'")
        out.append(val)
        pos = e
    return np.array(out, np.float32)


def vector_text(v) -> str:
    """VectorSpace::CreateStrFromObj (space_vector.cc:96-108): each float with
    max_digits10 (9) significant digits in the default float format."""
    return " ".join("%.9g" % float(x) for x in np.asarray(v, np.float32))


def split_params(s: str):
    """ParseArg (params.cc:51-59) through SplitStr (utils.h:210-231): ','
    becomes a space, then whitespace-separated tokens; a string ending in
    whitespace fails the stream read, as the reference's does."""
    if s == "":
        return []
    t = s.replace(",", " ")
    if t.strip() == "" or t[-1].isspace():
        raise QueryException(f"Cannot split arguments in: '{s}'")
    return t.split()


# --- the service ---------------------------------------------------------------------

class Dispatcher:
    """One thread that gathers the queries of concurrent calls into batches
    for the serving lanes (nmslib_amd.serving.Server) and hands each call its
    rows back."""

    def __init__(self, index, lanes: int, max_queries: int, max_k: int):
        self.index = index
        self.lanes, self.max_queries, self.max_k = lanes, max_queries, max_k
        self.q = queue.Queue()
        self.server = None
        self.server_ef = -1
        self.stop = False
        self.t = threading.Thread(target=self._run, daemon=True)
        self.t.start()

    def search(self, vecs: np.ndarray, k: int, ef: int):
        """Rows of (ids, dists, counts) for vecs [n][dim]; blocks the caller."""
        done = threading.Event()
        slot = {"vecs": vecs, "k": k, "ef": ef, "done": done}
        self.q.put(slot)
        done.wait()
        if "error" in slot:
            raise slot["error"]
        return slot["result"]

    def reopen(self):
        """Drops the lanes (after query-time parameters or the index changed)."""
        if self.server is not None:
            self.server.close()
            self.server = None

    def _server_for(self, ef: int):
        from . import serving
        if self.server is None or ef > self.server_ef:
            self.reopen()
            self.server_ef = max(ef, 16)
            self.server = serving.Server(self.index, lanes=self.lanes,
                                         max_queries=self.max_queries, max_ef=self.server_ef,
                                         max_k=self.max_k, use_graph=True)
        return self.server

    def _run(self):
        pending = []
        inflight = {}
        while not self.stop:
            try:
                timeout = 0.0005 if (pending or inflight) else 0.05
                pending.append(self.q.get(timeout=timeout))
                while True:
                    pending.append(self.q.get_nowait())
            except queue.Empty:
                pass
            for lane, batch in list(inflight.items()):
                if self.server.poll(lane):
                    try:
                        ids, dists, counts = self.server.fetch(lane)
                        off = 0
                        for slot in batch:
                            n = len(slot["vecs"])
                            slot["result"] = (ids[off:off + n, :slot["k"]],
                                              dists[off:off + n, :slot["k"]], counts[off:off + n])
                            off += n
                    except Exception as e:  # noqa: BLE001 -- handed to the callers
                        for slot in batch:
                            slot["error"] = QueryException(str(e))
                    for slot in batch:
                        slot["done"].set()
                    del inflight[lane]
            while pending:
                try:
                    srv = self._server_for(max(s["ef"] for s in pending))
                except Exception as e:  # noqa: BLE001
                    for slot in pending:
                        slot["error"] = QueryException(str(e))
                        slot["done"].set()
                    pending = []
                    break
                free = [ln for ln in range(self.lanes) if ln not in inflight]
                if not free:
                    break
                batch, nq = [], 0
                while pending and nq + len(pending[0]["vecs"]) <= self.max_queries:
                    s = pending.pop(0)
                    batch.append(s)
                    nq += len(s["vecs"])
                if not batch:  # a single call larger than a lane: split it
                    big = pending.pop(0)
                    try:
                        big["result"] = self.index.knnQueryBatchUneven(
                            big["vecs"], np.full(len(big["vecs"]), big["k"], np.int32),
                            np.full(len(big["vecs"]), big["ef"], np.int32))
                    except Exception as e:  # noqa: BLE001
                        big["error"] = QueryException(str(e))
                    big["done"].set()
                    continue
                q = np.concatenate([s["vecs"] for s in batch])
                ks = np.concatenate([np.full(len(s["vecs"]), s["k"], np.int32) for s in batch])
                efs = np.concatenate([np.full(len(s["vecs"]), s["ef"], np.int32) for s in batch])
                try:
                    srv.submit(free[0], q, ks, efs)
                    inflight[free[0]] = batch
                except Exception as e:  # noqa: BLE001
                    for slot in batch:
                        slot["error"] = QueryException(str(e))
                        slot["done"].set()


class QueryServiceHandler:
    """The QueryService methods over a loaded FloatIndex (the reference's
    QueryServiceHandler, QueryService_server.cpp:75-470)."""

    def __init__(self, index, extern_ids=None, lanes: int = 4, max_queries: int = 1024,
                 max_k: int = 100):
        self.index = index
        self.extern_ids = extern_ids
        self.max_k = max_k
        self.dim = index.info()["dim"]
        self.lock = threading.Condition()
        self.active = 0  # queries in flight (LockedCounterManager, :61-76)
        self.dispatcher = Dispatcher(index, lanes, max_queries, max_k)

    # the device path runs V1Merge in the serving lanes; SearchOld (algoType
    # old, or hybrid at efSearch >= 1000) and k beyond the lanes' limit take
    # the batch entry point
    def _search(self, vecs: np.ndarray, k: int):
        ef = self.index.info()["ef"]
        algo = self.index._algo
        if k <= self.max_k and (algo == "v1merge" or (algo == "hybrid" and ef < 1000)):
            return self.dispatcher.search(vecs, k, ef)
        ids, dists, counts = self.index.knnQueryBatchArrays(vecs, k)
        return ids, dists, counts

    def _enter(self):
        with self.lock:
            self.active += 1

    def _leave(self):
        with self.lock:
            self.active -= 1
            self.lock.notify_all()

    def _vector(self, obj: bytes) -> np.ndarray:
        v = parse_vector(obj)
        if len(v) != self.dim:
            raise QueryException(f"The query has {len(v)} elements, the index {self.dim}")
        return v

    def _entries(self, ids, dists, count, ret_extern, ret_obj):
        out = []
        for j in range(count):
            oid = int(ids[j])
            ext = None
            if ret_extern or ret_obj:
                ext = self.extern_ids[oid] if self.extern_ids is not None else str(oid)
            obj = None
            if ret_obj:
                if len(self.index) == 0:
                    raise QueryException("retObj needs the data points: load the index with "
                                         "its data (load_data)")
                obj = vector_text(self.index[oid])
            out.append((oid, float(dists[j]), ext if ret_extern or ret_obj else None, obj))
        return out

    def setQueryTimeParams(self, params: str):
        kv = split_params(params)
        with self.lock:
            while self.active:
                self.lock.wait()
            try:
                self.index.setQueryTimeParams(kv)
            except Exception as e:  # noqa: BLE001
                raise QueryException(str(e)) from None
            self.dispatcher.reopen()

    def knnQuery(self, k: int, obj: bytes, ret_extern: bool, ret_obj: bool):
        self._enter()
        try:
            v = self._vector(obj)[None, :]
            ids, dists, counts = self._search(v, k)
            return self._entries(ids[0], dists[0], int(counts[0]), ret_extern, ret_obj)
        except QueryException:
            raise
        except Exception as e:  # noqa: BLE001
            raise QueryException(str(e)) from None
        finally:
            self._leave()

    def knnQueryBatch(self, k: int, objs, ret_extern: bool, ret_obj: bool, num_threads: int):
        self._enter()
        try:
            if not objs:
                return []
            vecs = np.stack([self._vector(o) for o in objs])
            ids, dists, counts = self._search(vecs, k)
            return [self._entries(ids[i], dists[i], int(counts[i]), ret_extern, ret_obj)
                    for i in range(len(objs))]
        except QueryException:
            raise
        except Exception as e:  # noqa: BLE001
            raise QueryException(str(e)) from None
        finally:
            self._leave()

    def rangeQuery(self, r: float, obj: bytes, ret_extern: bool, ret_obj: bool):
        raise QueryException("Range search is not supported!")

    def getDistance(self, obj1: bytes, obj2: bytes) -> float:
        a, b = parse_vector(obj1), parse_vector(obj2)
        try:
            return float(self.index.vectorDistance(a, b))
        except Exception as e:  # noqa: BLE001
            raise QueryException(str(e)) from None

    def close(self):
        self.dispatcher.stop = True
        self.dispatcher.t.join(timeout=1.0)
        self.dispatcher.reopen()


def _arg(args: dict, fid: int, ttype: int, default=None):
    v = args.get(fid)
    if v is None:
        return default
    if v[0] != ttype:
        raise ProtocolError(f"field {fid} has type {v[0]}, expected {ttype}")
    return v[1]


def process(handler: QueryServiceHandler, r: Reader) -> bytes:
    """Reads one call and returns the reply's bytes."""
    name, mtype, seqid = r.message_begin()
    args = r.struct()
    w = Writer()
    try:
        if name == "setQueryTimeParams":
            handler.setQueryTimeParams(_arg(args, 1, T_STRING, b"").decode())
            result = None
        elif name == "knnQuery":
            result = handler.knnQuery(_arg(args, 1, T_I32, 0), _arg(args, 2, T_STRING, b""),
                                      _arg(args, 3, T_BOOL, False), _arg(args, 4, T_BOOL, False))
        elif name == "knnQueryBatch":
            result = handler.knnQueryBatch(_arg(args, 1, T_I32, 0), _arg(args, 2, T_LIST, []),
                                           _arg(args, 3, T_BOOL, False),
                                           _arg(args, 4, T_BOOL, False), _arg(args, 5, T_I32, 0))
        elif name == "rangeQuery":
            result = handler.rangeQuery(_arg(args, 1, T_DOUBLE, 0.0), _arg(args, 2, T_STRING, b""),
                                        _arg(args, 3, T_BOOL, False), _arg(args, 4, T_BOOL, False))
        elif name == "getDistance":
            result = handler.getDistance(_arg(args, 1, T_STRING, b""), _arg(args, 2, T_STRING, b""))
        else:
            w.message_begin(name, M_EXCEPTION, seqid)
            w.field(T_STRING, 1)
            w.binary(f"Invalid method name: '{name}'".encode())
            w.field(T_I32, 2)
            w.i32(APP_UNKNOWN_METHOD)
            w.stop()
            return w.getvalue()
    except QueryException as e:
        w.message_begin(name, M_REPLY, seqid)
        w.field(T_STRUCT, 1)  # the result struct's `err` field
        w.field(T_STRING, 1)
        w.binary(str(e).encode())
        w.stop()
        w.stop()
        return w.getvalue()
    w.message_begin(name, M_REPLY, seqid)
    if name == "knnQuery" or name == "rangeQuery":
        w.field(T_LIST, 0)
        w.entry_list(result)
    elif name == "knnQueryBatch":
        w.field(T_LIST, 0)
        w.i8(T_LIST)
        w.i32(len(result))
        for row in result:
            w.entry_list(row)
    elif name == "getDistance":
        w.field(T_DOUBLE, 0)
        w.double(result)
    w.stop()
    return w.getvalue()


class _Conn(socketserver.StreamRequestHandler):
    def handle(self):
        r = Reader(self.rfile)
        while True:
            try:
                out = process(self.server.handler, r)
            except (EOFError, ConnectionError):
                return
            except ProtocolError:
                return
            self.wfile.write(out)
            self.wfile.flush()


class ThriftServer(socketserver.ThreadingMixIn, socketserver.TCPServer):
    """TThreadPoolServer's role: a thread per connection (the reference pools
    threadQty threads, QueryService_server.cpp:664-676)."""
    daemon_threads = True
    allow_reuse_address = True

    def __init__(self, addr, handler: QueryServiceHandler):
        super().__init__(addr, _Conn)
        self.handler = handler


def serve(index, port: int, host: str = "127.0.0.1", **kw) -> ThriftServer:
    """Starts the server in a background thread and returns it (shutdown()
    and server_close() stop it)."""
    srv = ThriftServer((host, port), QueryServiceHandler(index, **kw))
    threading.Thread(target=srv.serve_forever, daemon=True).start()
    return srv


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("-p", "--port", type=int, required=True)
    ap.add_argument("-s", "--spaceType", required=True)
    ap.add_argument("-L", "--loadIndex", required=True)
    ap.add_argument("--load-data", action="store_true", help="also read the index's .dat file")
    ap.add_argument("-t", "--queryTimeParams", default="")
    ap.add_argument("--host", default="0.0.0.0")
    ap.add_argument("--lanes", type=int, default=4)
    args = ap.parse_args(argv)
    import nmslib_amd as nmslib
    idx = nmslib.init(method="hnsw", space=args.spaceType)
    idx.loadIndex(args.loadIndex, load_data=args.load_data)
    srv = ThriftServer((args.host, args.port), QueryServiceHandler(idx, lanes=args.lanes))
    if args.queryTimeParams:
        srv.handler.setQueryTimeParams(args.queryTimeParams)
    print(f"Started a server on port {args.port}", flush=True)
    try:
        srv.serve_forever()
    except KeyboardInterrupt:
        pass
    srv.handler.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
