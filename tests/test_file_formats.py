"""The reference's on-disk formats either side of the step, written and read
by the library (gazpart_amd/csrc/gp_io.h through gazpart_amd/io.py) and by
the reference's own SAVE_PARTICLE, INITIATION_PARTICLE_POSITION (ISAVEPART
= 2), SAVE_MPIIO and READ_MPIIO, compiled unmodified and run on one MPI rank
by oracle/_ref/part_io_ref.  Pinned: the files the library writes are byte
for byte the files the reference writes from the same data, and each side
reads what the other wrote, bit-exactly.  Host only (no GPU); skipped where
the reference build is absent."""
import os
import subprocess

import numpy as np
import pytest

from gazpart_amd import io as gio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "part_io_ref")
pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="reference build missing")


def _run(args, cwd):
    r = subprocess.run([EXE, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def _classes(seed, nig=2, npoly=2, npart=700, n=16):
    g = np.random.default_rng(seed)
    L = 2.0 * np.pi
    cls = []
    for j in range(nig):
        st = np.empty((13, npart))
        st[0:3] = g.uniform(0.05, L - 0.05, (3, npart))        # inside every rank filter
        st[3:6] = g.normal(size=(3, npart))
        q = g.normal(size=(4, npart))
        st[6:10] = q / np.linalg.norm(q, axis=0)
        st[10:13] = g.normal(size=(3, npart)) * 1e3
        idp = g.integers(1, npoly + 1, npart).astype(np.int32)
        ncl = np.bincount(idp, minlength=npoly + 1)[1:].astype(np.int32)
        cls.append(dict(state=st, idp=idp, npclass=ncl, rhop=g.uniform(500, 3000, npoly),
                        emaj=g.uniform(1e-3, 1e-2, npoly), apr=float(g.uniform(0.2, 5.0))))
    return cls, L, n


def _write_raw(path, cls, L, n):
    nig, npoly, npart = len(cls), len(cls[0]["rhop"]), cls[0]["state"].shape[1]
    with open(path, "wb") as f:
        np.array([nig, npoly, npart, n], dtype=np.int32).tofile(f)
        np.array([L]).tofile(f)
        np.array([c["npclass"] for c in cls], dtype=np.int32).tofile(f)
        np.array([c["rhop"] for c in cls]).tofile(f)
        np.array([c["emaj"] for c in cls]).tofile(f)
        np.array([c["apr"] for c in cls]).tofile(f)
        for c in cls:
            np.ascontiguousarray(c["state"]).tofile(f)       # = state(npart, 13) in Fortran
            c["idp"].astype(np.int32).tofile(f)


def _read_raw_out(path, nig, npoly):
    out = []
    with open(path, "rb") as f:
        for _ in range(nig):
            nloc = int(np.fromfile(f, np.int32, 1)[0])
            ncl = np.fromfile(f, np.int32, npoly)
            rh = np.fromfile(f, np.float64, npoly)
            em = np.fromfile(f, np.float64, npoly)
            apr = float(np.fromfile(f, np.float64, 1)[0])
            st = np.fromfile(f, np.float64, 13 * nloc).reshape(13, nloc)
            idp = np.fromfile(f, np.int32, nloc)
            out.append(dict(npclass=ncl, rhop=rh, emaj=em, apr=apr, state=st, idp=idp))
    return out


@pytest.mark.parametrize("ncycle", [-1, 7])
def test_particle_files_match_the_reference(tmp_path, ncycle):
    cls, L, n = _classes(3 + ncycle)
    raw = tmp_path / "in.raw"
    _write_raw(raw, cls, L, n)
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    _run(["psave", str(raw), str(ncycle)], ref_dir)
    for j, c in enumerate(cls, start=1):
        name = gio.part_file_name(j, ncycle)
        ref_file = ref_dir / name
        assert ref_file.exists(), sorted(os.listdir(ref_dir))
        # the library reads what the reference wrote
        d = gio.read_part_file(str(ref_file))
        assert d["npart"] == c["state"].shape[1]
        np.testing.assert_array_equal(d["state"], c["state"])
        np.testing.assert_array_equal(d["idp"], c["idp"])
        np.testing.assert_array_equal(d["npclass"], c["npclass"])
        np.testing.assert_array_equal(d["rhop"], c["rhop"])
        np.testing.assert_array_equal(d["emaj"], c["emaj"])
        assert d["apr"] == c["apr"]
        # and writes the same bytes from the same data
        ours = tmp_path / name
        gio.write_part_file(str(ours), c["state"], c["idp"], c["npclass"], c["rhop"], c["emaj"],
                            c["apr"])
        assert ours.read_bytes() == ref_file.read_bytes()


def test_reference_reads_library_particle_files(tmp_path):
    cls, L, n = _classes(11)
    raw = tmp_path / "in.raw"
    _write_raw(raw, cls, L, n)
    for j, c in enumerate(cls, start=1):
        gio.write_part_file(str(tmp_path / gio.part_file_name(j, restart=True)), c["state"],
                            c["idp"], c["npclass"], c["rhop"], c["emaj"], c["apr"])
    _run(["pread", str(raw), "out.raw"], tmp_path)
    got = _read_raw_out(tmp_path / "out.raw", len(cls), len(cls[0]["rhop"]))
    for g, c in zip(got, cls):
        np.testing.assert_array_equal(g["state"], c["state"])
        np.testing.assert_array_equal(g["idp"], c["idp"])
        np.testing.assert_array_equal(g["npclass"], c["npclass"])
        np.testing.assert_array_equal(g["rhop"], c["rhop"])
        np.testing.assert_array_equal(g["emaj"], c["emaj"])
        assert g["apr"] == c["apr"]


def test_field_files_match_the_reference(tmp_path):
    g = np.random.default_rng(5)
    n = 14
    f = g.normal(size=(n, n, n))
    raw = tmp_path / "f.raw"
    with open(raw, "wb") as fh:
        np.array([n], dtype=np.int32).tofile(fh)
        np.asfortranarray(f).ravel(order="F").tofile(fh)
    _run(["fsave", str(raw), "uf.end"], tmp_path)
    np.testing.assert_array_equal(gio.read_field_file(str(tmp_path / "uf.end"), n), f)
    gio.write_field_file(str(tmp_path / "vf.end"), f)
    assert (tmp_path / "vf.end").read_bytes() == (tmp_path / "uf.end").read_bytes()
    _run(["fread", "vf.end", str(n), "back.raw"], tmp_path)
    with open(tmp_path / "back.raw", "rb") as fh:
        assert int(np.fromfile(fh, np.int32, 1)[0]) == n
        back = np.fromfile(fh, np.float64, n ** 3).reshape((n, n, n), order="F")
    np.testing.assert_array_equal(back, f)


def test_format_errors_are_reported(tmp_path):
    from gazpart_amd import GpError
    p = tmp_path / "bad.end"
    p.write_bytes(b"\x04\x00\x00\x00\x05\x00\x00\x00\x08\x00\x00\x00")   # marker mismatch
    with pytest.raises(GpError):
        gio.read_part_file(str(p))
    with pytest.raises(GpError):
        gio.read_field_file(str(tmp_path / "missing.end"), 4)
    gio.write_field_file(str(tmp_path / "f.end"), np.zeros((4, 4, 4)))
    with pytest.raises(GpError):
        gio.read_field_file(str(tmp_path / "f.end"), 5)
