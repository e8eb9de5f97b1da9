"""HONEE binary Vec files (src/honee-file.c:37-39, 136-217).

Layout, all big-endian: an optional HONEE header ``int32 token`` (0xceedf32 -> int32 step, 0xceedf64 -> int64 step,
0xceedf00 legacy) followed by ``int step`` and ``float64 time``; then PETSc's ``VecView`` binary record
``PetscInt VEC_FILE_CLASSID (1211214)``, ``PetscInt n`` and ``n`` float64 values, where PetscInt is int32 or, from a
PETSc built with 64-bit indices, int64 (VecLoad_Binary tells them apart the same way: an int32 0 followed by the
class id).  A file may also be a bare Vec record.
"""
from __future__ import annotations

import numpy as np

VEC_FILE_CLASSID = 1211214
TOKEN_LEGACY, TOKEN_32, TOKEN_64 = 0xCEEDF00, 0xCEEDF32, 0xCEEDF64


def read_vec(path: str) -> tuple[np.ndarray, int, float]:
    """Return (values, step, time); step and time are 0 for bare Vec files."""
    buf = open(path, "rb").read()
    tok = int(np.frombuffer(buf[:4], ">i4")[0])
    off, step, time = 0, 0, 0.0
    if tok in (TOKEN_32, TOKEN_64, TOKEN_LEGACY):
        isz = 8 if tok == TOKEN_64 else 4
        step = int(np.frombuffer(buf[4:4 + isz], ">i8" if isz == 8 else ">i4")[0])
        time = float(np.frombuffer(buf[4 + isz:12 + isz], ">f8")[0])
        off = 12 + isz
    w0, w1 = (int(v) for v in np.frombuffer(buf[off:off + 8], ">i4"))
    if w0 == 0 and w1 == VEC_FILE_CLASSID:  # 64-bit PetscInt header
        cid, n = (int(v) for v in np.frombuffer(buf[off:off + 16], ">i8"))
        off += 16
    else:
        cid, n = w0, w1
        off += 8
    if cid != VEC_FILE_CLASSID:
        raise ValueError(f"{path}: not a PETSc Vec record (class id {cid})")
    vals = np.frombuffer(buf[off:off + 8 * n], ">f8").astype(np.float64)
    if vals.size != n:
        raise ValueError(f"{path}: truncated Vec ({vals.size} of {n} values)")
    return vals, step, time


def write_vec(path: str, values: np.ndarray, step: int = 0, time: float = 0.0, index_bits: int = 32) -> None:
    """Write a HONEE-formatted file (HoneeWriteBinaryVec) as a PETSc with index_bits-wide PetscInt writes it."""
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    it = ">i8" if index_bits == 64 else ">i4"
    with open(path, "wb") as f:
        f.write(np.array([TOKEN_64 if index_bits == 64 else TOKEN_32], ">i4").tobytes() + np.array([step], it).tobytes())
        f.write(np.array([time], ">f8").tobytes())
        f.write(np.array([VEC_FILE_CLASSID, v.size], it).tobytes())
        f.write(v.astype(">f8").tobytes())


def index_bits(path: str) -> int:
    """PetscInt width (32 or 64) of a HONEE-formatted file's header."""
    tok = int(np.frombuffer(open(path, "rb").read(4), ">i4")[0])
    return 64 if tok == TOKEN_64 else 32


def global_indices(mesh, essential_dofs=()) -> np.ndarray:
    """L-vector entries of a PETSc global vector, in its order, for a degree-1 box mesh of this package.

    HONEE's state files hold the DMPlex global vector (src/honee-file.c:196-217): the section's points in chart order
    with their free components, constrained (essential) components left out.  At degree 1 the dofs sit on the
    vertices only, and the box mesh numbers its vertices x fastest, then y, then z, as this package's box_mesh numbers
    its nodes, so the global vector is the L-vector without the essential entries.  Higher degrees add edge, face and
    cell dofs, whose DMPlex order is not reproduced here: they raise."""
    if getattr(mesh, "degree", None) != 1:
        raise NotImplementedError("the DMPlex global order is reproduced for degree-1 box meshes only")
    return np.setdiff1d(np.arange(5 * mesh.nnodes), np.asarray(essential_dofs, dtype=np.int64))


def to_global(q_local, mesh, essential_dofs=()) -> np.ndarray:
    """The global vector of the L-vector q_local (5 components per node), for write_vec."""
    return np.asarray(q_local, dtype=np.float64).ravel()[global_indices(mesh, essential_dofs)]


def to_local(values, mesh, essential_dofs=(), essential_values=0.0) -> np.ndarray:
    """The L-vector of a global vector read by read_vec, with the essential entries set to essential_values."""
    idx = global_indices(mesh, essential_dofs)
    if np.size(values) != idx.size:
        raise ValueError(f"global vector of {np.size(values)} values, the mesh and constraints give {idx.size}")
    q = np.zeros(5 * mesh.nnodes)
    q[np.asarray(essential_dofs, dtype=np.int64)] = essential_values
    q[idx] = values
    return q
