"""File formats, host side (no GPU work):
  * the serial SaveToFile / ReadFromFile vector file
    (bridge/algebra_bridges/restart_bridge.cpp:49-75, 146-158): framing
    restated from the serial pcl stand-in of that file; the payload is pinned
    byte for byte to the reference's own Serialize(BinaryBuffer, Vector)
    compiled into oracle/_ref;
  * ConnectionViewer .mat/.vec text (connection_viewer_output.h:74-141,
    402-422): restated; the writer itself cannot be built here (its progress
    output pulls in common/log.cpp -> debug_id -> boost), so numbers are pinned
    to C's %g formatting (which std::ostream's default float format is) via
    Python's identical 'g' format: 6 significant digits for positions and
    matrix values, 16 for vector values, "0" for stored zeros."""
import struct

import numpy as np
import pytest

import oracle_lib as O

ref = O.ref_lib()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref not built (reference checkout absent)")


def U():
    import ugcore_amd
    return ugcore_amd


SPECIAL = np.array([0.0, -0.0, 1.0, -2.5, 1e-310, -1e308, np.inf, -np.inf, np.nan, np.pi, 1.0 / 3.0])


@pytest.mark.parametrize("n", [0, 1, 11, 1000])
def test_vector_file_layout_and_roundtrip(tmp_path, n):
    u = U()
    rng = np.random.default_rng(n)
    v = SPECIAL[:n].copy() if n <= SPECIAL.size else rng.normal(size=n) * 10.0 ** rng.integers(-300, 300, size=n)
    path = tmp_path / "v.ug4vector"
    u.write_vector_file(path, v)
    raw = path.read_bytes()
    assert len(raw) == 4 + 4 + 8 + 8 * n
    num_procs, next_offset = struct.unpack("<ii", raw[:8])
    assert num_procs == 1
    assert next_offset == len(raw)
    (count,) = struct.unpack("<Q", raw[8:16])
    assert count == n
    assert raw[16:] == v.astype("<f8").tobytes()
    back = u.read_vector_file(path)
    assert back.tobytes() == v.tobytes()


@needs_ref
@pytest.mark.parametrize("n", [0, 1, 11, 257])
def test_vector_file_payload_matches_reference_serialize(tmp_path, n):
    u = U()
    v = np.random.default_rng(3 + n).normal(size=n)
    if n >= SPECIAL.size:
        v[:SPECIAL.size] = SPECIAL
    path = tmp_path / "v.bin"
    u.write_vector_file(path, v)
    length = ref.ref_vec_serialize(n, v, None, 0)
    buf = C_buffer(length)
    ref.ref_vec_serialize(n, v, buf, length)
    assert path.read_bytes()[8:] == bytes(buf.raw[:length])


def C_buffer(n):
    import ctypes
    return ctypes.create_string_buffer(max(n, 1))


def test_vector_file_rejects_bad_files(tmp_path):
    u = U()
    p = tmp_path / "bad"
    p.write_bytes(struct.pack("<ii", 2, 24) + struct.pack("<Q", 1) + struct.pack("<d", 1.0))
    with pytest.raises(u.UGError):
        u.read_vector_file(p)       # two process parts
    p.write_bytes(struct.pack("<ii", 1, 40) + struct.pack("<Q", 1) + struct.pack("<d", 1.0))
    with pytest.raises(u.UGError):
        u.read_vector_file(p)       # offset does not match the length
    p.write_bytes(struct.pack("<ii", 1, 32) + struct.pack("<Q", 2) + struct.pack("<d", 1.0))
    with pytest.raises(u.UGError):
        u.read_vector_file(p)       # truncated values
    with pytest.raises(u.UGError):
        u.read_vector_file(tmp_path / "missing")


def _cv_lines(path):
    return path.read_text().splitlines()


def _g(x, digits):
    return format(float(x), f".{digits}g")


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_connection_viewer_matrix_format(tmp_path, dim):
    u = U()
    rng = np.random.default_rng(dim)
    n = 9
    ent = []
    for r in range(n):
        for c in sorted(rng.choice(n, size=3, replace=False)):
            ent.append((r, int(c), float(rng.choice([0.0, rng.normal() * 10.0 ** rng.integers(-8, 8)]))))
    M = O.builder_csr(ent, n, n, mode=0, add=True)
    pos = rng.uniform(-2, 2, size=(n, dim)) / 3.0
    path = tmp_path / "A.mat"
    u.write_connection_viewer_matrix(path, M.rowptr.astype(np.int32), M.col, M.val, pos)
    lines = _cv_lines(path)
    assert lines[0] == "1" and lines[1] == str(dim) and lines[2] == str(n)
    for i in range(n):
        want = [_g(x, 6) for x in pos[i]] + (["0.0"] if dim == 1 else [])
        assert lines[3 + i] == " ".join(want)
    assert lines[3 + n] == "1"
    body = lines[4 + n:]
    assert len(body) == len(M.col)
    k = 0
    for r in range(n):
        for p in range(M.rowptr[r], M.rowptr[r + 1]):
            v = M.val[p]
            assert body[k] == f"{r} {M.col[p]} " + ("0" if v == 0.0 else _g(v, 6))
            k += 1


def test_connection_viewer_vector_format(tmp_path):
    u = U()
    rng = np.random.default_rng(4)
    n = 12
    v = rng.normal(size=n) * 10.0 ** rng.integers(-20, 20, size=n)
    v[3] = 0.0
    pos = rng.uniform(0, 1, size=(n, 2))
    path = tmp_path / "b.vec"
    u.write_connection_viewer_vector(path, v, pos)
    lines = _cv_lines(path)
    assert lines[:3] == ["1", "2", str(n)]
    for i in range(n):
        assert lines[3 + i] == f"{_g(pos[i, 0], 6)} {_g(pos[i, 1], 6)}"
    assert lines[3 + n] == "1"
    for i in range(n):
        assert lines[4 + n + i] == f"{i} {i} {_g(v[i], 16)}"
    # 16 significant digits carry the value to within one unit in the 16th digit
    back = np.array([float(l.split()[2]) for l in lines[4 + n:]])
    assert np.all(np.abs(back - v) <= 1e-15 * np.abs(v))
