"""Checkpoint / restart files (cs_dump_write, cs_dump_read, cs_dump_info):
the host-array entry points, no GPU.

The format is restated here independently in numpy (header fields at their
byte offsets, sections at 4096-byte aligned offsets, the SplitMix64-based
checksum), so a file the library writes is parsed without the library, and
a file written by numpy is read by the library.  The checksum's mixing
function is pinned by the published first outputs of SplitMix64 seeded with
0 (Steele, Lea & Flood 2014): 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4,
0x06c45d188009454f."""
import os
import struct

import numpy as np
import pytest

import cubesphere_amd as cs

GOLDEN = np.uint64(0x9e3779b97f4a7c15)
NAMES = ("U", "ent", "Bxf", "Byf", "Bzf")


def _mix(x):
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
    return x ^ (x >> np.uint64(31))


def _checksum(words):
    w = np.frombuffer(np.ascontiguousarray(words).tobytes(), dtype="<u8")
    i = np.arange(1, w.size + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(_mix(w + i * GOLDEN).sum(dtype=np.uint64))


def _cfg(N=6, NR=3):
    return cs.make_config(N, NR, 4, cs.RECON_P1ADP, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2, cs.RBC_PAD,
                          cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0, imex=1)


def _state(cfg, seed=3):
    st = cs.init_problem(cfg, 3, seed=seed)
    rng = np.random.default_rng(seed)
    for a in st.arrays():
        a += 1e-3 * rng.normal(size=a.size)
    return st


def _np_read(path):
    """(header dict, arrays) by the format's own description, no library."""
    raw = open(path, "rb").read()
    h = raw[:4096]
    assert h[:8] == b"CSDUMP\0\0"
    version, hbytes, bom, nsec = struct.unpack_from("<4I", h, 8)
    ints = struct.unpack_from("<12i", h, 24)
    dbls = struct.unpack_from("<10d", h, 72)
    time, dtcur = struct.unpack_from("<2d", h, 152)
    ncc, zero = struct.unpack_from("<2i", h, 168)
    filenum, step = struct.unpack_from("<2q", h, 176)
    secs = []
    for q in range(5):
        nm, off, cnt, sm = struct.unpack_from("<8s3Q", h, 192 + 32 * q)
        secs.append((nm.rstrip(b"\0").decode(), off, cnt, sm))
    (hsum,) = struct.unpack_from("<Q", h, 352)
    arrays = [np.frombuffer(raw, dtype="<f8", count=cnt, offset=off) for _, off, cnt, _ in secs]
    return dict(version=version, hbytes=hbytes, bom=bom, nsec=nsec, ints=ints, dbls=dbls, time=time, dtcur=dtcur,
                ncc=ncc, zero=zero, filenum=filenum, step=step, secs=secs, hsum=hsum, head=h), arrays


def _np_write(path, cfg, arrays, time, dtcur, filenum, step, version=1):
    h = bytearray(4096)
    h[:8] = b"CSDUMP\0\0"
    struct.pack_into("<4I", h, 8, version, 4096, 0x01020304, 5)
    struct.pack_into("<12i", h, 24, cfg.ntp, cfg.nr, cfg.nghost, cfg.recon, cfg.riemann, cfg.emf, cfg.precision,
                     cfg.rk_order, cfg.radial_bc, cfg.radial_mesh, cfg.imex, cfg.bc_problem)
    struct.pack_into("<10d", h, 72, cfg.rmin, cfg.rmax, cfg.gamma, cfg.rhofloor, cfg.prsfloor, cfg.cfl, cfg.grav,
                     *cfg.omega)
    struct.pack_into("<2d2i2q", h, 152, time, dtcur, 8, 0, filenum, step)
    body, off = [], 4096
    for q, (nm, a) in enumerate(zip(NAMES, arrays)):
        struct.pack_into("<8s3Q", h, 192 + 32 * q, nm.encode(), off, a.size, _checksum(a))
        b = a.astype("<f8").tobytes()
        pad = (-len(b)) % 4096
        body.append(b + bytes(pad))
        off += len(b) + pad
    struct.pack_into("<Q", h, 352, _checksum(np.frombuffer(bytes(h[:352]), dtype="<u8")))
    with open(path, "wb") as f:
        f.write(bytes(h))
        for b in body:
            f.write(b)


def test_checksum_mixing_is_splitmix64():
    assert _checksum(np.zeros(1, np.uint64)) == 0xe220a8397b1dcdaf
    assert _checksum(np.zeros(2, np.uint64)) == (0xe220a8397b1dcdaf + 0x6e789e6aa1b965f4) % 2 ** 64
    assert _checksum(np.zeros(3, np.uint64)) == (0xe220a8397b1dcdaf + 0x6e789e6aa1b965f4 + 0x06c45d188009454f) % 2 ** 64
    # one changed word, or two words swapped, change the sum
    a = np.arange(1.0, 9.0)
    b = a.copy()
    b[3] = np.nextafter(b[3], 10.0)
    c = a.copy()
    c[[2, 5]] = c[[5, 2]]
    assert len({_checksum(a), _checksum(b), _checksum(c)}) == 3


def test_round_trip_and_header(tmp_path):
    cfg = _cfg()
    st = _state(cfg)
    p = tmp_path / "d.csd"
    cs.write_dump(p, st, time=1.25, dtcur=3e-3, filenum=7, step=42)
    assert not os.path.exists(str(p) + ".tmp")
    got, m = cs.read_dump(p, cfg)
    for a, b in zip(got.arrays(), st.arrays()):
        assert np.array_equal(a, b)
    assert (m.time, m.dtcur, m.filenum, m.step, m.n_cc_components, m.version) == (1.25, 3e-3, 7, 42, 8, 1)
    assert (m.cfg.ntp, m.cfg.nr, m.cfg.nghost, m.cfg.recon, m.cfg.imex) == (6, 3, 4, cs.RECON_P1ADP, 1)
    i = cs.dump_info(p)
    assert (i.time, i.step, i.cfg.rmax) == (1.25, 42, 1.5)


def test_numpy_reads_the_library_file(tmp_path):
    cfg = _cfg()
    st = _state(cfg, seed=5)
    p = tmp_path / "d.csd"
    cs.write_dump(p, st, time=0.5, dtcur=1e-3, filenum=2, step=9)
    hd, arrays = _np_read(str(p))
    assert (hd["version"], hd["hbytes"], hd["bom"], hd["nsec"], hd["ncc"], hd["zero"]) == (1, 4096, 0x01020304, 5, 8, 0)
    assert hd["ints"][:3] == (6, 3, 4) and hd["dbls"][:2] == (1.0, 1.5)
    assert (hd["time"], hd["dtcur"], hd["filenum"], hd["step"]) == (0.5, 1e-3, 2, 9)
    assert hd["hsum"] == _checksum(np.frombuffer(hd["head"][:352], dtype="<u8"))
    for (nm, off, cnt, sm), a, want in zip(hd["secs"], arrays, st.arrays()):
        assert off % 4096 == 0 and cnt == want.size
        assert sm == _checksum(a), nm
        assert np.array_equal(a, want), nm
    assert os.path.getsize(str(p)) % 4096 == 0


def test_library_reads_a_numpy_file(tmp_path):
    cfg = _cfg()
    st = _state(cfg, seed=7)
    p = tmp_path / "n.csd"
    _np_write(str(p), cfg, st.arrays(), 2.0, 4e-3, 3, 11)
    got, m = cs.read_dump(p, cfg)
    for a, b in zip(got.arrays(), st.arrays()):
        assert np.array_equal(a, b)
    assert (m.time, m.dtcur, m.filenum, m.step) == (2.0, 4e-3, 3, 11)


@pytest.mark.parametrize("where", ["header", "U", "ent", "Bxf", "Byf", "Bzf", "truncated", "magic", "version"])
def test_damaged_files_are_refused(tmp_path, where):
    cfg = _cfg()
    st = _state(cfg, seed=9)
    p = tmp_path / "d.csd"
    cs.write_dump(p, st, time=1.0, dtcur=1e-3)
    raw = bytearray(open(p, "rb").read())
    hd, _ = _np_read(str(p))
    if where == "header":
        raw[152] ^= 1                      # a bit of the time
    elif where == "truncated":
        raw = raw[:len(raw) - 8192]
    elif where == "magic":
        raw[0] = ord("X")
    elif where == "version":
        struct.pack_into("<I", raw, 8, 2)
    else:
        q = NAMES.index(where)
        off, cnt = hd["secs"][q][1], hd["secs"][q][2]
        raw[off + 8 * (cnt // 2) + 3] ^= 0x10   # one bit in the middle of the section
    open(p, "wb").write(bytes(raw))
    with pytest.raises(cs.CsError) as e:
        cs.read_dump(p, cfg)
    assert e.value.code == cs.E_FORMAT, str(e.value)
    if where in NAMES:
        assert f"section {where}" in str(e.value)


def test_grid_must_match_scheme_may_differ(tmp_path):
    cfg = _cfg()
    st = _state(cfg)
    p = tmp_path / "d.csd"
    cs.write_dump(p, st)
    with pytest.raises(cs.CsError) as e:
        cs.read_dump(p, _cfg(N=8))
    assert e.value.code == cs.E_INVALID and "grid" in str(e.value)
    other = cs.make_config(6, 3, 4, cs.RECON_P1ADP, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2, cs.RBC_PAD,
                           cs.RMESH_UNIFORM, 1.0, 1.5)
    got, m = cs.read_dump(p, other)
    assert np.array_equal(got.U, st.U) and m.cfg.riemann == cs.RS_HLLD


def test_missing_file_and_replacement(tmp_path):
    cfg = _cfg()
    with pytest.raises(cs.CsError) as e:
        cs.dump_info(tmp_path / "none.csd")
    assert e.value.code == cs.E_IO
    p = tmp_path / "d.csd"
    a, b = _state(cfg, seed=1), _state(cfg, seed=2)
    cs.write_dump(p, a, step=1)
    cs.write_dump(p, b, step=2)
    got, m = cs.read_dump(p, cfg)
    assert m.step == 2 and np.array_equal(got.U, b.U)
