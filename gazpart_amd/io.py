"""The reference's on-disk formats either side of the step, through the
library's host-only C-ABI (gp_part_file_*, gp_field_file_*; no GPU needed).

Particle files (ISAVEPART = 2): one per class J, ``part_pJJ.end`` or
``part_pJJ_tNNNNNNNN.bin`` as SAVE_PARTICLE writes them
(save_particle.f90:116-185) and ``part_pJJ.ini`` as
INITIATION_PARTICLE_POSITION reads them (initiation_particle_position.f90:
430-514).  Field files: ``uf.end`` / ``vf.end`` / ``wf.end`` (or
``uf_tNNNNNNNN.bin`` ...) of SAVE_MPIIO / READ_MPIIO
(MPI_structures.f90:434-493, 571-621)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import GpError, load_library


def _chk(rc, what):
    if rc:
        raise GpError(rc, what)


def part_file_name(j: int, ncycle: int | None = None, restart: bool = False) -> str:
    """File name of class j (1-based): part_pJJ.ini (restart input),
    part_pJJ.end (final) or part_pJJ_tNNNNNNNN.bin (cycle ncycle >= 0), as
    save_particle.f90:161-165 and initiation_particle_position.f90:442 form
    them."""
    if restart:
        return f"part_p{j:02d}.ini"
    if ncycle is None or ncycle < 0:
        return f"part_p{j:02d}.end"
    return f"part_p{j:02d}_t{ncycle:08d}.bin"


def write_part_file(path: str, state: np.ndarray, idp: np.ndarray, npclass, rhop, emaj,
                    apr: float) -> None:
    """state (13, npart) in the library's layout, idp (npart,) 1-based size
    indices; npclass / rhop / emaj are the class's NPCLASS / RHOP /
    EMAJ_PART rows (length npoly)."""
    L = load_library()
    st = np.asarray(state, dtype=np.float64)
    npart = st.shape[1]
    cols = np.empty((14, npart))
    cols[:13] = st
    cols[13] = np.asarray(idp, dtype=np.float64)
    ncl = np.ascontiguousarray(npclass, dtype=np.int32)
    rh = np.ascontiguousarray(rhop, dtype=np.float64)
    em = np.ascontiguousarray(emaj, dtype=np.float64)
    assert ncl.size == rh.size == em.size
    _chk(L.gp_part_file_write(path.encode(), npart, ncl.size, ncl.ctypes.data, rh.ctypes.data,
                              em.ctypes.data, float(apr), cols.ctypes.data), path)


def read_part_file(path: str) -> dict:
    """{npart, npclass, rhop, emaj, apr, state (13, npart), idp (npart,) int}"""
    L = load_library()
    npart, npoly = C.c_int64(0), C.c_int(0)
    _chk(L.gp_part_file_info(path.encode(), C.byref(npart), C.byref(npoly)), path)
    n, m = npart.value, npoly.value
    ncl = np.zeros(m, dtype=np.int32)
    rh, em = np.zeros(m), np.zeros(m)
    apr = C.c_double(0.0)
    cols = np.zeros((14, max(n, 1)))
    _chk(L.gp_part_file_read(path.encode(), n, m, ncl.ctypes.data, rh.ctypes.data,
                             em.ctypes.data, C.byref(apr), cols.ctypes.data), path)
    cols = cols[:, :n]
    return {"npart": n, "npclass": ncl, "rhop": rh, "emaj": em, "apr": apr.value,
            "state": np.ascontiguousarray(cols[:13]), "idp": cols[13].astype(np.int64)}


def write_field_file(path: str, f: np.ndarray) -> None:
    """f (n, n, n) indexed [i, j, k] (x, y, z): written x fastest."""
    L = load_library()
    a = np.asarray(f, dtype=np.float64)
    n = a.shape[0]
    assert a.shape == (n, n, n)
    fo = np.asfortranarray(a)
    _chk(L.gp_field_file_write(path.encode(), n, fo.ctypes.data), path)


def read_field_file(path: str, n: int) -> np.ndarray:
    """(n, n, n) array indexed [i, j, k]."""
    L = load_library()
    out = np.zeros((n, n, n), order="F")
    _chk(L.gp_field_file_read(path.encode(), n, out.ctypes.data), path)
    return out
