"""The PETSc global vector of a degree-1 box mesh (pyhonee.honee_file.global_indices / to_global / to_local) on the
reference's own state files: a file read, mapped to an L-vector and back, and written again is byte-identical; the
constrained entries of the L-vector carry the given boundary values; the gaussian-wave fixtures have the z-velocity
constrained at every node (4 values per node), the channel fixtures the velocity at the wall nodes as well (2 values).
The order itself is pinned by the GPU and oracle reproductions of these fixtures, compared in order
(tests/test_oracle_gaussianwave_implicit.py, tests/test_oracle_channel_golden.py)."""
import os

import numpy as np
import pytest

import channel_case as CH
import gaussianwave_case as GW
from pyhonee import honee_file
from pyhonee import mesh as M

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name,case", [("gaussianwave-IDL.bin", "gw"), ("gaussianwave-IDL-entropy.bin", "gw"), ("channel.bin", "ch"),
                                       ("channel-prim.bin", "ch")])
def test_global_local_round_trip(name, case, tmp_path):
    if case == "gw":
        mesh, _, _, _, ess, q0, _, _ = GW.setup_implicit(1)
    else:
        mesh, _, _, ess, q0, _, _ = CH.setup(0)
    path = os.path.join(GOLDEN, name)
    vals, step, t = honee_file.read_vec(path)
    q = honee_file.to_local(vals, mesh, ess, q0[ess])
    assert np.array_equal(q[ess], q0[ess])
    assert np.array_equal(honee_file.to_global(q, mesh, ess), vals)
    out = tmp_path / name
    honee_file.write_vec(str(out), honee_file.to_global(q, mesh, ess), step=step, time=t, index_bits=honee_file.index_bits(path))
    assert out.read_bytes() == open(path, "rb").read()


def test_constrained_record_sizes():
    mesh, _, _, _, ess, _, free, _ = GW.setup_implicit(1)
    assert free.size == 4 * mesh.nnodes  # z-velocity constrained everywhere
    mesh_c, _, _, ess_c, _, free_c, _ = CH.setup(0)
    per_node = np.bincount(free_c // 5, minlength=mesh_c.nnodes)
    assert sorted(set(per_node)) == [2, 4] and (per_node == 2).sum() == 20  # 20 wall nodes of 5 x 6 x 2


def test_higher_degree_order_is_refused():
    with pytest.raises(NotImplementedError):
        honee_file.global_indices(M.box_mesh((2, 2, 1), 3), [])
