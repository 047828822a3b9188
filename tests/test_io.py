"""CPU: the reference's interchange formats (htool_amd/io.py).

The cluster-tree CSV files and the vector byte files below are the reference's
own test data (data/data_test/, read by its solver tests,
tests/functional_tests/solvers/test_solver_ddm.hpp:55-95), copied as fixtures.
They must read and write back byte for byte. A tree of ours must survive a
save/read round trip (radii and centres to the 6 digits the format keeps).
Where the reference tree is present, its dense 485 x 485 test matrix is read
and A sol = rhs checked (the solver's tolerance)."""
import os

import numpy as np
import pytest

from htool_amd import geometry as geo
from htool_amd import io

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_data")
REF_DATA = "/root/reference/data/data_test/symmetric"


@pytest.mark.parametrize("name", ["symmetric_cluster_1", "non_symmetric_cluster_3"])
def test_reference_cluster_files_round_trip(tmp_path, htg, name):
    perm_f, tree_f = os.path.join(DATA, name + "_permutation.csv"), os.path.join(DATA, name + "_tree.csv")
    t = io.read_cluster(perm_f, tree_f)
    assert t.offset[0] == 0 and t.size[0] == 485 and sorted(t.perm) == list(range(485))
    for i in range(len(t.offset)):  # sons tile their parent
        ks = list(t.sons(i))
        if ks:
            assert t.offset[ks[0]] == t.offset[i]
            assert sum(t.size[k] for k in ks) == t.size[i]
            assert all(t.offset[a] + t.size[a] == t.offset[b] for a, b in zip(ks, ks[1:]))
    out = str(tmp_path / "c")
    io.save_cluster(t, out)
    assert open(out + "_tree.csv").read() == open(tree_f).read()
    assert open(out + "_permutation.csv").read() == open(perm_f).read()


def test_cluster_tree_save_read(tmp_path, htg):
    x = geo.sphere(3000, seed=3)
    t = io.cluster_tree(x, 10)
    assert sorted(t.perm) == list(range(3000))
    out = str(tmp_path / "ours")
    io.save_cluster(t, out)
    r = io.read_cluster(out + "_permutation.csv", out + "_tree.csv")
    assert np.array_equal(r.perm, t.perm)
    a, b = t.preorder(), r.preorder()
    assert len(a) == len(b)
    for i, j in zip(a, b):
        assert (t.offset[i], t.size[i], t.depth[i], t.rank[i], t.nb_sons[i]) == \
               (r.offset[j], r.size[j], r.depth[j], r.rank[j], r.nb_sons[j])
        assert abs(t.rad[i] - r.rad[j]) <= 5e-6 * t.rad[i]
        assert np.allclose(t.ctr[i], r.ctr[j], rtol=5e-6, atol=1e-300)


def test_vector_and_matrix_bytes(tmp_path, htg):
    rhs = io.bytes_to_vector(os.path.join(DATA, "symmetric_rhs.bin"), np.complex128)
    sol = io.bytes_to_vector(os.path.join(DATA, "symmetric_sol.bin"), np.complex128)
    assert rhs.shape == sol.shape == (485,)
    io.vector_to_bytes(rhs, str(tmp_path / "r.bin"))
    assert open(tmp_path / "r.bin", "rb").read() == open(os.path.join(DATA, "symmetric_rhs.bin"), "rb").read()
    A = np.random.default_rng(1).standard_normal((7, 5))
    io.matrix_to_bytes(A, str(tmp_path / "A.bin"))
    raw = open(tmp_path / "A.bin", "rb").read()
    assert np.frombuffer(raw[:8], np.int32).tolist() == [7, 5]
    assert np.array_equal(np.frombuffer(raw[8:], np.float64), A.ravel(order="F"))
    assert np.array_equal(io.bytes_to_matrix(str(tmp_path / "A.bin")), A)


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_DATA, "matrix.bin")), reason="reference data not present")
def test_reference_dense_test_matrix_solves_its_system(htg):
    A = io.bytes_to_matrix(os.path.join(REF_DATA, "matrix.bin"), np.complex128)
    rhs = io.bytes_to_vector(os.path.join(DATA, "symmetric_rhs.bin"), np.complex128)
    sol = io.bytes_to_vector(os.path.join(DATA, "symmetric_sol.bin"), np.complex128)
    assert A.shape == (485, 485)
    assert np.linalg.norm(A @ sol - rhs) / np.linalg.norm(rhs) < 1e-6
    assert np.linalg.norm(A - A.T) / np.linalg.norm(A) < 1e-14  # the 'symmetric' case


def test_save_plot_format(tmp_path, htg):
    class Fake:  # the leaf-list writer needs only these three calls
        def nb_rows(self): return 100
        def nb_cols(self): return 90
        def get_leaves(self):
            return np.array([[0, 50, 0, 45, -1], [0, 50, 45, 45, 3], [50, 50, 0, 45, 2], [50, 50, 45, 45, -1]],
                            np.int32)
    path = io.save_plot(Fake(), str(tmp_path / "plot"))
    assert path.endswith("_0.csv")
    lines = open(path).read().splitlines()
    assert lines[0] == "100,90"
    assert lines[1:] == ["0,50,0,45,-1", "50,50,45,45,-1", "0,50,45,45,3", "50,50,0,45,2"]  # dense first
    nr, nc, leaves = io.load_plot(path)
    assert (nr, nc) == (100, 90) and np.array_equal(leaves, Fake().get_leaves())
