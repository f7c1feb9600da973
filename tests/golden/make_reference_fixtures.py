"""Writes reference_fixtures.json: data held by the reference's own unit tests for this path.

Run in a container that has the reference tree (read as text only; nothing of it is executed):
    python tests/golden/make_reference_fixtures.py /path/to/coolfluid3
The JSON it writes is data (inputs and expected outputs of the reference tests), committed so that
the tests do not need the reference tree at run time.

Sources:
  test/mesh/utest-mesh-lagrangep1-tetra3d.cpp:36-175  (nodes, volume, Jacobian, adjoint)
  test/mesh/utest-mesh-lagrangep1-hexa3d.cpp:36-270   (nodes, shape values/gradients, Jacobian, adjoint, det)
  test/mesh/utest-mesh-lagrangep1-triag2d.cpp:36-178  (nodes, volumes, Jacobian, adjoint)
  test/mesh/utest-mesh-lagrangep1-prism3d.cpp:36-178  (nodes, volume, shape values/gradients, integrals,
                                                       det, Jacobian, adjoint)
  test/math/utest-lss-test-matrix.hpp:32-719          (4-rank Stokes system: per-rank element lists, node-graph
                                                       CSR, owned-row values before / after the conditions, rhs,
                                                       solution, Dirichlet and periodic conditions, result)
  test/math/utest-lss-symmetric-dirichlet.cpp:72-208  (2-rank 4-node system, merged to serial)
  test/math/utest-lss-vector.cpp:144-187               (apply y = 2*A*x + y with A == 1)
Also copies mesh files of the reference (data files, unchanged) into meshes/: the Gambit neutral files
plugins/UFEM/test/meshes/ring2d-tris.neu (669 nodes, 1204 triangles, boundary sets inner and outer),
plugins/UFEM/test/meshes/ring2d-quads.neu (560 nodes, 490 quadrilaterals; atest-ufem-heat2d-disk),
plugins/UFEM/test/meshes/kvs15.neu (16953 nodes, 33538 triangles of the channel around a cylinder; sets
in, out, symm and wall; atest-ufem-navier-stokes-cylinder2d) and
resources/hextet.neu (hexahedra and tetrahedra; test/mesh/utest-mesh-neu.cpp:123-125), and the Gmsh
files resources/rectangle-tg-p1.msh (507 nodes, 929 triangles; test/mesh/utest-mesh-gmsh.cpp:96),
resources/rectangle-qd-p1.msh (56 nodes, 42 quadrilaterals) and resources/sqduct_3000e.msh (3751 nodes,
3000 hexahedra of a square duct).
"""
import json
import os
import re
import shutil
import sys

MESHES = ["plugins/UFEM/test/meshes/ring2d-tris.neu", "plugins/UFEM/test/meshes/ring2d-quads.neu",
          "plugins/UFEM/test/meshes/kvs15.neu", "resources/hextet.neu", "resources/rectangle-tg-p1.msh",
          "resources/rectangle-qd-p1.msh", "resources/sqduct_3000e.msh"]


def copy_meshes(ref):
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "meshes")
    os.makedirs(dst, exist_ok=True)
    for rel in MESHES:
        shutil.copyfile(os.path.join(ref, rel), os.path.join(dst, os.path.basename(rel)))


def cpp_array(src, name):
    m = re.search(r"cf3::(?:Uint|Real) %s\[(\d+)\]=\{([^}]*)\}" % re.escape(name), src)
    body = m.group(2).replace("\\\n", "")
    vals = [x.strip() for x in body.split(",") if x.strip()]
    assert len(vals) == int(m.group(1)), name
    return vals


def main(ref):
    out = {}
    # ---- Tetra3D known answers (values typed in the reference test) ----
    out["tetra3d"] = {
        "nodes": [[0.830434, 0.885201, 0.188108], [0.89653, 0.899961, 0.297475],
                  [0.888273, 0.821744, 0.211428], [0.950439, 0.904872, 0.20736]],
        "volume": 1.451803461048456186e-4,
        "jacobian": [[6.609600000000004361e-02, 1.475999999999999535e-02, 1.093669999999999920e-01],
                     [5.783899999999997377e-02, -6.345699999999998564e-02, 2.332000000000000739e-02],
                     [1.200050000000000283e-01, 1.967099999999999405e-02, 1.925199999999999134e-02]],
        "jacobian_adjoint": [[-1.680401883999999273e-03, 1.867198736999999475e-03, 7.284304918999997755e-03],
                             [1.685000172000002431e-03, -1.185210664300000161e-02, 4.784319192999994877e-03],
                             [8.752908253999998334e-03, 4.710993839999994925e-04, -5.047957512000001042e-03]],
        "mapped_coords": [0.1, 0.8, 0.45],
        "shape_values": [-0.35, 0.1, 0.8, 0.45],
    }
    out["hexa3d"] = {
        "mapped_coords": [0.5, 0.4, 0.1],
        "unit_nodes": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
        "skewed_nodes": [[0.5, 0.5, 0.5], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1.5, 1.5, 1.5],
                         [0, 1, 1]],
        "shape_values": [0.03375, 0.10125, 0.23625, 0.07875, 0.04125, 0.12375, 0.28875, 0.09625],
        "shape_gradients": [[-0.06750, 0.06750, 0.1575, -0.1575, -0.08250, 0.08250, 0.1925, -0.1925],
                            [-0.05625, -0.16875, 0.16875, 0.05625, -0.06875, -0.20625, 0.20625, 0.06875],
                            [-0.03750, -0.1125, -0.2625, -0.08750, 0.03750, 0.1125, 0.2625, 0.08750]],
        "det_unit": 0.125, "det_skewed": 0.1875, "volume_unit": 1.0, "volume_skewed": 1.0,
        "jacobian_skewed": [[0.5625, 0.06250, 0.06250], [0.07500, 0.5750, 0.07500], [0.1125, 0.1125, 0.6125]],
        "adjoint_skewed": [[0.34375, -0.03125, -0.03125], [-0.03750, 0.3375, -0.03750],
                           [-0.05625, -0.05625, 0.31875]],
    }
    out["prism3d"] = {
        "mapped_coords": [0.5, 0.4, 0.1],
        "unit_nodes": [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]],
        "skewed_nodes": [[0.2, 0.3, 0], [1.2, 0.3, 0], [0.2, 1.3, 0], [0.2, 0.3, 1], [1.2, 0.3, 1], [0.2, 1.3, 1]],
        "volume_unit": 0.5, "volume_skewed": 0.5,
        "shape_values": [0.045, 0.225, 0.18, 0.055, 0.275, 0.22],
        "shape_gradients": [[-0.45, 0.45, 0, -0.55, 0.55, 0], [-0.45, 0, 0.45, -0.55, 0, 0.55],
                            [-0.05, -0.25, -0.2, 0.05, 0.25, 0.2]],
        "integral_det_order1": 0.5, "integral_det_order2": 0.5,
        "det_unit": 0.5, "det_skewed": 0.5,
        "jacobian_skewed": [[1, 0, 0], [0, 1, 0], [0, 0, 0.5]],
        "adjoint_skewed": [[0.5, 0, 0], [0, 0.5, 0], [0, 0, 1.0]],
    }
    out["triag2d"] = {
        "nodes": [[0.5, 0.3], [1.1, 1.2], [0.8, 2.1]],
        "jacobian": [[0.6, 0.9], [0.3, 1.8]],
        "jacobian_adjoint": [[1.8, -0.9], [-0.3, 0.6]],
        "volume_cases": [[[[0, 0], [1, 0], [1, 1]], 0.5], [[[15, 15], [40, 25], [25, 30]], 137.5]],
    }
    # ---- LSS test matrix (4-rank Stokes system): per-rank element lists and node-graph CSR ----
    src = open(os.path.join(ref, "test/math/utest-lss-test-matrix.hpp")).read()
    # the bc_value arrays use two macros of the file: P_ (fixed pressure) and U_ (fixed velocity)
    macros = {"P_": 0.0, "U_": 1.0}
    ranks = []
    for r in range(4):
        rk = {
            "elem_nodes": [int(x) for x in cpp_array(src, "elem_nodes%d" % r)],
            "column_indices": [int(x) for x in cpp_array(src, "column_indices%d" % r)],
            "rowstart_positions": [int(x) for x in cpp_array(src, "rowstart_positions%d" % r)],
            "periodic_pairs": [int(x) for x in cpp_array(src, "periodic_pairs%d" % r)],
        }
        # the assembled Stokes system of utest-lss-distributed-matrix.cpp:102-166: values of the owned
        # rows before and after the Dirichlet and periodic conditions, rhs / solution of all local
        # nodes, the conditions (local node indices) and the reference solution
        for name in ("global_numbering", "irank_updatable", "bc_node", "bc_eqn"):
            rk[name] = [int(x) for x in cpp_array(src, "%s%d" % (name, r))]
        for name in ("mat_prebc", "rhs_prebc", "sol_prebc", "mat_presolve", "rhs_presolve", "sol_presolve",
                     "result"):
            rk[name] = [float(x) for x in cpp_array(src, "%s%d" % (name, r))]
        rk["bc_value"] = [macros[x] if x in macros else float(x) for x in cpp_array(src, "bc_value%d" % r)]
        ranks.append(rk)
    out["lss_test_matrix"] = {"nbeqs": 3, "element_nodes_per_elem": 3, "ranks": ranks}
    # ---- symmetric Dirichlet: gids 0..3 chain, neq 1, set_row(i,0,2,1), BC on gid 1 value 10 ----
    out["symmetric_dirichlet"] = {
        "node_connectivity": [0, 1, 0, 1, 2, 1, 2, 3, 2, 3], "starting_indices": [0, 2, 5, 8, 10],
        "neq": 1, "set_row": {"diag": 2.0, "offdiag": 1.0}, "bc": {"node": 1, "eq": 0, "value": 10.0},
        "expected_matrix": [[2, 0, 0, 0], [0, 1, 0, 0], [0, 0, 2, 1], [0, 0, 1, 2]],
        "expected_rhs": [-10.0, 10.0, -10.0, 0.0],
    }
    out["vector_apply"] = {"reset_matrix": 1.0, "reset_rhs": 1.0, "alpha": 2.0, "beta": 1.0,
                           "expect": "y_i = 1 + 2 * sum_j x_j (rows of an all-ones matrix)"}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    copy_meshes(ref)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
