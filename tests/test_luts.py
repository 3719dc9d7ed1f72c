"""Case tables: SHA-256 pins against the reference arrays and structural invariants.

The reference keeps the tables as literals (DMC_CUDA/DMC_CUDA/src/dmc.cu:45-620);
the pins below are the SHA-256 of those arrays as raw bytes in declaration
order (config_edge_lut1, config_edge_lut2, num_vertex_lut1, num_vertex_lut2)
and of the 2B/3B list sorted by config (with its faces in the same order).
"""
import hashlib

import numpy as np
import pytest

import oracle

PINS = {
    "edge_label_v0": "80c5320d8572a5f4bf2996ef96588fc96419361d437e36c01ef45181e75cdbfd",
    "edge_label_v1": "107ba88c5a5dfb68e118c2a5af0fafaf32197b4107171a2ce7cbd21a6faa2845",
    "num_vertex_v0": "df65cc5deb9afc1b4054a3a5eae0bb3449cca8fd74daaf6faeb91cd672c4836c",
    "num_vertex_v1": "6af715c52508b433a5f9f80bf2773135b3e9eed8f096f23a118e51fbe4db10ec",
    "amb_sorted": "4011e8ec448f44dde806bab21f80b6156dd1a2fbc73e5fad65892eb88c1a691a",
    "amb_faces_sorted": "99f3efa10920572f3fdea43023900b08bc42a649e966f8bac67cfac2f5063295",
}

EDGES = [(0, 1), (1, 2), (2, 3), (0, 3), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6), (6, 7), (4, 7)]
FACES = [(0, 1, 2, 3), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 3, 7, 4), (4, 5, 6, 7)]


@pytest.fixture(scope="module")
def tables():
    b = oracle.luts_bytes()
    lab = np.frombuffer(b[:6144], dtype=np.uint8).reshape(2, 256, 12)
    nv = np.frombuffer(b[6144:6656], dtype=np.uint8).reshape(2, 256)
    amb = np.frombuffer(b[6656:6912], dtype=np.uint8)
    return lab, nv, amb


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_sha256_pins(tables):
    lab, nv, amb = tables
    assert sha(lab[0]) == PINS["edge_label_v0"]
    assert sha(lab[1]) == PINS["edge_label_v1"]
    assert sha(nv[0]) == PINS["num_vertex_v0"]
    assert sha(nv[1]) == PINS["num_vertex_v1"]
    cfgs = np.nonzero(amb != 0xFF)[0].astype(np.uint8)
    assert sha(cfgs) == PINS["amb_sorted"]
    assert sha(amb[cfgs]) == PINS["amb_faces_sorted"]


def test_labels_mark_exactly_the_bipolar_edges(tables):
    lab, nv, _ = tables
    for v in range(2):
        for c in range(256):
            for e, (a, b) in enumerate(EDGES):
                bip = ((c >> a) & 1) != ((c >> b) & 1)
                assert (lab[v, c, e] != 0xFF) == bip


def test_labels_are_dense(tables):
    lab, nv, _ = tables
    for v in range(2):
        for c in range(256):
            used = sorted(set(int(x) for x in lab[v, c] if x != 0xFF))
            assert used == list(range(nv[v, c]))


def test_variants_differ_exactly_on_2b3b(tables):
    lab, nv, amb = tables
    differ = {c for c in range(256) if not np.array_equal(lab[0, c], lab[1, c])}
    assert differ == {c for c in range(256) if amb[c] != 0xFF}
    assert len(differ) == 36
    pc = {bin(c).count("1") for c in differ}
    assert pc == {2, 3}


def test_listed_face_is_the_unique_ambiguous_face(tables):
    _, _, amb = tables
    for c in range(256):
        if amb[c] == 0xFF:
            continue
        faces = []
        for f, cs in enumerate(FACES):
            b = [(c >> q) & 1 for q in cs]
            if b[0] == b[2] and b[1] == b[3] and b[0] != b[1]:
                faces.append(f)
        assert faces == [amb[c]]


def test_max_vertex_count(tables):
    _, nv, _ = tables
    assert nv.max() == 4
    assert sorted(np.nonzero(nv[0] == 4)[0].tolist()) == [0x5A, 0xA5]
    assert nv[0, 0] == 0 and nv[0, 255] == 0


def test_complement_symmetry_breaks_only_on_ambiguous_faces(tables):
    """Partitions of c and ~c coincide unless c has an ambiguous face (joining
    inside on the face is joining outside on the complement)."""
    lab, _, _ = tables

    def partition(row):
        groups = {}
        for e, m in enumerate(row):
            if m != 0xFF:
                groups.setdefault(int(m), []).append(e)
        return sorted(tuple(g) for g in groups.values())

    n_asym = 0
    for c in range(256):
        has_amb = any(((c >> cs[0]) & 1) == ((c >> cs[2]) & 1) and ((c >> cs[1]) & 1) == ((c >> cs[3]) & 1)
                      and ((c >> cs[0]) & 1) != ((c >> cs[1]) & 1) for cs in FACES)
        sym = partition(lab[0, c]) == partition(lab[0, 255 - c])
        assert sym != has_amb
        n_asym += has_amb
    assert n_asym == 120
