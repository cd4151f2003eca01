"""The strict OBJ loader (C ABI sie_obj_load, the C++ ObjMeshLoader) against
the reference's own ObjMeshLoader (src/sie_mom/mesh/obj_mesh_loader.cpp,
compiled into oracle/_ref): vertex and face order, RWG edge numbering and
face-to-edge maps bit for bit, and the same rejections with the same messages.
The edge numbering fixes the unknown order of currents.bin."""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.skipif(not ol.ref_available(), reason="oracle/_ref not built (reference tree absent)")


def ref_load(path):
    lib = ctypes.CDLL(ol.REF_SO)
    P = ctypes.c_void_p
    c = np.zeros(3, np.int32)
    err = ctypes.create_string_buffer(4096)
    if lib.ref_load_obj(str(path).encode(), c.ctypes.data_as(P), None, None, None, None, err, 4096):
        return err.value.decode()
    V, F = np.zeros((c[0], 3)), np.zeros((c[1], 3), np.int32)
    E, FE = np.zeros((c[2], 2), np.int32), np.zeros((c[1], 3), np.int32)
    lib.ref_load_obj(str(path).encode(), c.ctypes.data_as(P), V.ctypes.data_as(P), F.ctypes.data_as(P),
                     E.ctypes.data_as(P), FE.ctypes.data_as(P), err, 4096)
    return V, F, E, FE


def ours(path):
    import sie_mom_amd as sm
    try:
        m = sm.Mesh.from_obj(path)
    except sm.SieError as e:
        return str(e)
    return m.vertices, m.faces, m.edges, m.face_edges


def write(path, V, F, form="plain", extra=""):
    with open(path, "w") as f:
        f.write("# generated\nvn 0 0 1\nvt 0 0\n" + extra)
        for v in V:
            f.write("v %.17g %.17g %.17g\n" % tuple(v))
        for t in F:
            if form == "atn":
                f.write("f %d/1/1 %d//1 %d/1\n" % tuple(t + 1))
            elif form == "neg":
                f.write("f %d %d %d  # relative\n" % tuple(t - len(V)))
            else:
                f.write("f " + " ".join(str(int(i) + 1) for i in t) + "\n")


@pytest.fixture(scope="module")
def meshes():
    import sie_mom_amd as sm
    return sm.Mesh.icosphere(1.0, 2), sm.Mesh.spindle(2.0, 1.5, 12)


@pytest.mark.parametrize("case", ["icosphere", "spindle_shuffled_atn", "negative_indices"])
def test_obj_matches_reference_loader(tmp_path, meshes, case):
    ico, sp = meshes
    rng = np.random.default_rng(2)
    if case == "icosphere":
        V, F, form = ico.vertices, ico.faces, "plain"
    elif case == "spindle_shuffled_atn":
        V, F, form = sp.vertices, sp.faces[rng.permutation(sp.nf)], "atn"
    else:
        V, F, form = ico.vertices, ico.faces, "neg"
    p = tmp_path / f"{case}.obj"
    write(p, V, F, form)
    a, b = ours(p), ref_load(p)
    assert not isinstance(a, str) and not isinstance(b, str), (a, b)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(a[1], F)  # face order preserved


def test_obj_rejections_match_reference(tmp_path, meshes):
    ico, _ = meshes
    V, F = ico.vertices, ico.faces
    flipped = F.copy()
    flipped[5] = flipped[5][::-1]
    cases = {
        "open": (V, F[1:], ""),
        "flipped": (V, flipped, ""),
        "inward": (V, F[:, ::-1], ""),
        "unreferenced": (np.vstack([V, [[5.0, 5.0, 5.0]]]), F, ""),
        "duplicate_face": (V, np.vstack([F, F[:1]]), ""),
        "degenerate": (V, np.vstack([F, [[0, 0, 1]]]), ""),
        "quad": (V, F, "f 1 2 3 4\n"),
    }
    for name, (Vc, Fc, extra) in cases.items():
        p = tmp_path / f"{name}.obj"
        if name == "quad":
            write(p, Vc, Fc)
            with open(p, "a") as f:
                f.write(extra)
        else:
            write(p, Vc, Fc)
        a, b = ours(p), ref_load(p)
        assert isinstance(a, str) and isinstance(b, str), (name, a, b)
        assert a == b, (name, a, b)
    a, b = ours(tmp_path / "missing.obj"), ref_load(tmp_path / "missing.obj")
    assert a == b and "Failed to open" in a
