"""Compact container for an nmslib optimized HNSW index file.

The reference writes an optimized index as (src/method/hnsw.cc:776-806,
layout include/method/hnsw.h:328-371, include/object.h:40-104):

    u32 optimIndexFlag(=1) | u32 N | u64 memoryPerObject | u64 offsetLevel0
    u64 offsetData(=0) | i32 maxlevel | u32 enterpointId | u64 maxM | u64 maxM0
    i32 dist_func_type | u64 searchMethod
    N x [ i32 id | i32 label | u64 datalength | f32 vec[dim] | i32 cnt | i32 ids[maxM0] ]
    N x [ u32 sizemass | sizemass bytes: level x (i32 cnt | i32 ids[maxM]) ]

Unused adjacency slots hold 0x01010101 (the block is memset to 1 before it is
filled, hnsw.cc:427).  A *pack* keeps everything but the vectors, which are
regenerated from their seeded generator and, for cosine, normalised exactly
as ``NormalizeVect`` (include/method/hnsw.h:486-497) does in float32.
``unpack`` writes the byte-identical reference file back.
"""
from __future__ import annotations

import struct

import numpy as np

HEADER = struct.Struct("<IIQQQiIQQiQ")  # 68 bytes
UNUSED = 0x01010101


def read_index(path: str) -> dict:
    with open(path, "rb") as f:
        buf = f.read(HEADER.size)
        (flag, n, elem, off_l0, off_data, maxlevel, enter, maxM, maxM0, func,
         method) = HEADER.unpack(buf)
        if flag != 1:
            raise ValueError("not an optimized HNSW index")
        dim = (off_l0 - 16) // 4
        block = np.frombuffer(f.read(n * elem), dtype=np.uint8).reshape(n, elem)
        levels = np.zeros(n, np.int32)
        uppers = []
        for i in range(n):
            (sm,) = struct.unpack("<I", f.read(4))
            if sm:
                uppers.append(np.frombuffer(f.read(sm), dtype=np.int32))
                levels[i] = sm // (4 * (maxM + 1))
        rest = f.read()
    if rest:
        raise ValueError("trailing bytes after the index")
    hdr = dict(n=n, elem=elem, off_l0=off_l0, off_data=off_data, maxlevel=maxlevel,
               enter=enter, maxM=maxM, maxM0=maxM0, func=func, method=method, dim=dim)
    return dict(header=hdr, block=block, levels=levels,
                upper=np.concatenate(uppers) if uppers else np.zeros(0, np.int32))


def pack(path: str) -> dict:
    ix = read_index(path)
    h, block = ix["header"], ix["block"]
    dim = h["dim"]
    head = block[:, :16].copy().view(np.int32)  # id, label, datalength(lo, hi)
    l0 = block[:, h["off_l0"]:h["off_l0"] + 4 * (h["maxM0"] + 1)].copy().view(np.int32)
    assert np.all(head[:, 2] == 4 * dim) and np.all(head[:, 3] == 0)
    return dict(hdr=np.array([h[k] for k in ("n", "elem", "off_l0", "off_data", "maxlevel",
                                            "enter", "maxM", "maxM0", "func", "method", "dim")],
                             np.int64),
                ext_id=head[:, 0].copy(), label=head[:, 1].copy(), l0=l0,
                levels=ix["levels"], upper=ix["upper"])


def save_pack(pk: dict, path: str) -> None:
    np.savez_compressed(path, **pk)


def load_pack(path: str) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def normalize_like_reference(x: np.ndarray) -> np.ndarray:
    """Row-wise NormalizeVect in float32: sequential sum of squares, then
    1/sqrt(sum) and a multiply (rows with zero sum are left as they are)."""
    x = np.asarray(x, np.float32)
    sq = np.cumsum(x * x, axis=1, dtype=np.float32)[:, -1]
    out = x.copy()
    nz = sq != 0
    inv = np.float32(1.0) / np.sqrt(sq[nz])
    out[nz] = x[nz] * inv[:, None]
    return out


def unpack(pk: dict, vectors: np.ndarray, path: str) -> None:
    """Write the reference-format index file for pack ``pk`` with ``vectors``
    (the raw base vectors in internal order; cosine ones get normalised)."""
    (n, elem, off_l0, off_data, maxlevel, enter, maxM, maxM0, func, method,
     dim) = (int(v) for v in pk["hdr"])
    vectors = np.ascontiguousarray(vectors, np.float32)
    if vectors.shape != (n, dim):
        raise ValueError(f"vectors shape {vectors.shape} != {(n, dim)}")
    if func == 3:
        vectors = normalize_like_reference(vectors)
    rec = np.empty((n, elem), np.uint8)
    rec[:] = 1
    head = np.zeros((n, 4), np.int32)
    head[:, 0] = pk["ext_id"]
    head[:, 1] = pk["label"]
    head[:, 2] = 4 * dim
    rec[:, 0:16] = head.view(np.uint8)
    rec[:, 16:16 + 4 * dim] = vectors.view(np.uint8)
    rec[:, off_l0:off_l0 + 4 * (maxM0 + 1)] = np.ascontiguousarray(pk["l0"]).view(np.uint8)
    with open(path, "wb") as f:
        f.write(HEADER.pack(1, n, elem, off_l0, off_data, maxlevel, enter, maxM, maxM0, func,
                            method))
        f.write(rec.tobytes())
        levels = pk["levels"]
        upper = pk["upper"]
        per = maxM + 1
        sizemass = (levels.astype(np.int64) * per * 4).astype(np.uint32)
        pos = 0
        chunks = []
        for i in range(n):
            sm = int(sizemass[i])
            chunks.append(struct.pack("<I", sm))
            if sm:
                cnt = sm // 4
                chunks.append(upper[pos:pos + cnt].tobytes())
                pos += cnt
            if len(chunks) > 65536:
                f.write(b"".join(chunks))
                chunks = []
        f.write(b"".join(chunks))
