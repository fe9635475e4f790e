"""AGCSRv1 files on the host (no device): a file written here from the
documented layout (32-byte header: magic, N, E, section mask, reserved,
FNV-1a 64 of the payload; then uint32 sections in bit order) is accepted by
the library's checker, and corrupt, truncated or mislabelled files are
refused.  The GPU tests check that the library writes and loads the same
layout."""
import os
import struct

import numpy as np
import pytest

import accelgraph_express_amd as ag
import oracle

MASK_FULL = 0x3FF


def fnv1a64(b: bytes) -> int:
    h = 0xCBF29CE484222325
    for x in b:
        h ^= x
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def write_agcsr(path, csr, n, e, mask=MASK_FULL):
    # offsets follow the degrees unless the caller gives them (malformed files)
    off = csr.get("offsets")
    if off is None:
        off = np.concatenate([[0], np.cumsum(csr["vertices.out_degree"])]).astype(np.uint32)
    ioff = csr.get("inverse_offsets")
    if ioff is None:
        ioff = np.concatenate([[0], np.cumsum(csr["inverse_vertices.out_degree"])]).astype(np.uint32)
    secs = [csr["sorted_edges_array.edges_array_src"], csr["sorted_edges_array.edges_array_dest"],
            csr["vertices.out_degree"], csr["vertices.edges_idx"], off,
            csr["inverse_sorted_edges_array.edges_array_src"],
            csr["inverse_sorted_edges_array.edges_array_dest"],
            csr["inverse_vertices.out_degree"], csr["inverse_vertices.edges_idx"], ioff]
    payload = b"".join(np.ascontiguousarray(s, dtype="<u4").tobytes()
                       for i, s in enumerate(secs) if mask >> i & 1)
    hdr = b"AGCSRv1\0" + struct.pack("<IIIIQ", n, e, mask, 0, fnv1a64(payload))
    with open(path, "wb") as f:
        f.write(hdr + payload)
    return hdr + payload


@pytest.fixture
def small_csr():
    rng = np.random.RandomState(3)
    n, e = 40, 300
    src, dst = rng.randint(0, n, e), rng.randint(0, n, e)
    return oracle.build_csr(src, dst, n), n, e


def test_documented_layout_is_accepted(tmp_path, small_csr):
    csr, n, e = small_csr
    p = str(tmp_path / "g.agcsr")
    blob = write_agcsr(p, csr, n, e)
    assert len(blob) == 32 + 4 * (4 * e + 6 * n + 2)
    assert ag.csr_file_info(p) == {"num_vertices": n, "num_edges": e, "sections": MASK_FULL}


def test_corrupt_truncated_and_foreign_files_are_refused(tmp_path, small_csr):
    csr, n, e = small_csr
    p = str(tmp_path / "g.agcsr")
    blob = write_agcsr(p, csr, n, e)
    bad = bytearray(blob)
    bad[100] ^= 1  # one payload bit
    open(p, "wb").write(bytes(bad))
    with pytest.raises(ag.AGError):
        ag.csr_file_info(p)
    open(p, "wb").write(blob[:-4])  # truncated
    with pytest.raises(ag.AGError):
        ag.csr_file_info(p)
    open(p, "wb").write(b"XGCSRv1\0" + blob[8:])  # wrong magic
    with pytest.raises(ag.AGError):
        ag.csr_file_info(p)
    with pytest.raises(ag.AGError):
        ag.csr_file_info(str(tmp_path / "missing.agcsr"))


def test_fnv1a64_known_answers():
    # published FNV-1a 64 test vectors
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8
