"""PETSc binary files: the scratch-file format between petgem's stages.

The reference writes every preprocessing output and reads every solver input
through PETSc viewers (``petgem/parallel.py:65-147``) -- dense matrices via
``MatView`` (stored in the AIJ layout with every entry present), vectors via
``VecView``, solutions ``x{i}.dat`` via ``writePetscVector``
(``petgem/solver.py:593-594``).  petsc4py is not available here, so this
module reads and writes the same byte layout directly.

Layout (big-endian, 32-bit PetscInt, complex scalars as the reference's
complex PETSc build uses):
  Vec: int32 1211214, int32 n, n scalars
  Mat: int32 1211216, int32 M, int32 N, int32 nz, M int32 row lengths,
       nz int32 column indices, nz scalars
  (nz == -1: dense, M*N scalars row by row)
"""
import numpy as np

VEC_CLASSID = 1211214
MAT_CLASSID = 1211216
_I = np.dtype(">i4")
_Z = np.dtype(">c16")
_D = np.dtype(">f8")


class PetscFileError(ValueError):
    pass


def _scalar_dtype(nbytes_per, real):
    return _D if real else _Z


def read_vec(path, real=False):
    """Read a PETSc Vec; returns complex128 (or float64 when ``real``)."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(8), dtype=_I)
        if hdr.size != 2 or hdr[0] != VEC_CLASSID:
            raise PetscFileError(f"{path}: not a PETSc Vec file")
        n = int(hdr[1])
        dt = _scalar_dtype(None, real)
        data = np.frombuffer(f.read(n * dt.itemsize), dtype=dt)
        if data.size != n:
            raise PetscFileError(f"{path}: truncated ({data.size} of {n} entries)")
    return data.astype(np.float64 if real else np.complex128)


def write_vec(path, values, real=False):
    v = np.asarray(values)
    dt = _scalar_dtype(None, real)
    with open(path, "wb") as f:
        f.write(np.array([VEC_CLASSID, v.size], dtype=_I).tobytes())
        f.write(np.ascontiguousarray(v.ravel(), dtype=dt).tobytes())


def read_mat(path, real=False):
    """Read a PETSc Mat.  Returns (indptr, indices, data, shape) in CSR form."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(16), dtype=_I)
        if hdr.size != 4 or hdr[0] != MAT_CLASSID:
            raise PetscFileError(f"{path}: not a PETSc Mat file")
        M, N, nz = (int(x) for x in hdr[1:])
        dt = _scalar_dtype(None, real)
        if nz == -1:
            data = np.frombuffer(f.read(M * N * dt.itemsize), dtype=dt)
            if data.size != M * N:
                raise PetscFileError(f"{path}: truncated dense matrix")
            indptr = np.arange(0, M * N + 1, N, dtype=np.int64)
            indices = np.tile(np.arange(N, dtype=np.int32), M)
        else:
            rl = np.frombuffer(f.read(4 * M), dtype=_I).astype(np.int64)
            indices = np.frombuffer(f.read(4 * nz), dtype=_I).astype(np.int32)
            data = np.frombuffer(f.read(nz * dt.itemsize), dtype=dt)
            if rl.size != M or indices.size != nz or data.size != nz or rl.sum() != nz:
                raise PetscFileError(f"{path}: inconsistent AIJ matrix file")
            indptr = np.concatenate([[0], np.cumsum(rl)])
    return indptr, indices, data.astype(np.float64 if real else np.complex128), (M, N)


def read_dense(path, real=False):
    """Read a matrix written from a dense array (every row full) as a 2-D array."""
    indptr, indices, data, (M, N) = read_mat(path, real)
    if indices.size != M * N:
        raise PetscFileError(f"{path}: matrix is not dense ({indices.size} of {M*N} entries)")
    out = np.zeros((M, N), dtype=data.dtype)
    rows = np.repeat(np.arange(M), np.diff(indptr))
    out[rows, indices] = data
    return out


def write_dense(path, array, real=False):
    """Write a 2-D array the way ``MatView`` stores a dense matrix (AIJ layout, all entries)."""
    a = np.atleast_2d(np.asarray(array))
    M, N = a.shape
    dt = _scalar_dtype(None, real)
    with open(path, "wb") as f:
        f.write(np.array([MAT_CLASSID, M, N, M * N], dtype=_I).tobytes())
        f.write(np.full(M, N, dtype=_I).tobytes())
        f.write(np.tile(np.arange(N, dtype=_I), M).tobytes())
        f.write(np.ascontiguousarray(a.ravel(), dtype=dt).tobytes())


def write_aij(path, indptr, indices, data, shape, real=False):
    M, N = shape
    indptr = np.asarray(indptr, dtype=np.int64)
    dt = _scalar_dtype(None, real)
    with open(path, "wb") as f:
        f.write(np.array([MAT_CLASSID, M, N, int(indptr[-1])], dtype=_I).tobytes())
        f.write(np.diff(indptr).astype(_I).tobytes())
        f.write(np.asarray(indices).astype(_I).tobytes())
        f.write(np.ascontiguousarray(data, dtype=dt).tobytes())
