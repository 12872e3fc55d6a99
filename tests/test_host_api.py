"""Host side of the product library (no GPU needed): the crate's validation
order and messages, prefix-commit batched adds, the C-ABI surface, and that the
assembly refuses to run without a device (no CPU fallback)."""

import ctypes
import os
import re

import numpy as np
import pytest

import finite_element_method_amd as fa
from finite_element_method_amd import FEM, FemError, DeviceError
from oracle.oracle import OracleModel, OracleError
from tests.conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _both(fn):
    out = []
    for cls, err in ((FEM, FemError), (OracleModel, OracleError)):
        m = cls(1e-4, 1e-12, 5)
        try:
            fn(m)
            out.append("ok")
        except err as e:
            out.append(str(e))
    return out


QUAD = [(1, 1.0, 1.0, 0.0), (2, 0.0, 1.0, 0.0), (3, 0.0, 0.0, 0.0), (4, 1.0, 0.0, 0.0)]


def _quad(m):
    for n in QUAD:
        m.add_node(*n)


CASES = {
    # methods_for_node_data_handle.rs:39-79
    "node_number": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(1, 1, 0, 0)), "Node with number 1 already exists!"),
    "node_coords": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(2, 0, 0, 0)),
                    "Node with coordinates x: 0.0, y: 0.0, z: 0.0 already exists!"),
    "node_limit": (lambda m: [m.add_node(i, i, 0, 0) for i in range(1, 7)], "Nodes number could not be greater than 5!"),
    # methods_for_truss_data_handle.rs:49-67, truss.rs:44-64
    "truss_node": (lambda m: (m.add_node(1, 0, 0, 0), m.add_truss(1, 1, 9, 1e6, 1.0)), "Node with number 9 does not exist!"),
    "truss_E": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(2, 1, 0, 0), m.add_truss(1, 1, 2, 0.0, 1.0)),
                "Young's modulus 0.0 is less or equal to zero!"),
    "truss_A2": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(2, 1, 0, 0), m.add_truss(1, 1, 2, 1e6, 1.0, -2e-5)),
                 "Area2 -2e-5 is less or equal to zero!"),
    "truss_dup": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(2, 1, 0, 0), m.add_truss(1, 1, 2, 1e6, 1.0),
                             m.add_truss(1, 2, 1, 1e6, 1.0)), "Truss element with number 1 already exists!"),
    "truss_same": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(2, 1, 0, 0), m.add_truss(1, 1, 2, 1e6, 1.0),
                              m.add_truss(2, 2, 1, 1e6, 1.0)), "Truss element with node number 2 and 1 already exists!"),
    # methods_for_beam_data_handle.rs:49-68, beam.rs:63-117 (the crate prints E for nu)
    "beam_nu": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(2, 1, 0, 0),
                           m.add_beam(1, 1, 2, 2e11, 0.0, 1, 1, 1, 0, 1, 0.8, [0, 1, 0])),
                "Poisson's ratio 200000000000.0 is less or equal to zero!"),
    "beam_it": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(2, 1, 0, 0),
                           m.add_beam(1, 1, 2, 2e11, 0.3, 1, 1, 1, 0, -1, 0.8, [0, 1, 0])),
                "It -1.0 is less or equal to zero!"),
    "beam_axis": (lambda m: (m.add_node(1, 0, 0, 0), m.add_node(2, 1, 0, 0),
                             m.add_beam(7, 1, 2, 2e11, 0.3, 1, 1, 1, 0, 1, 0.8, [2, 0, 0])),
                  "Local axis 1 direction [2.0, 0.0, 0.0] parallel to element 7!"),
    # methods_for_plate_data_handle.rs:62-102, plate.rs:58-158 (the crate prints E for nu and t)
    "plate_t": (lambda m: (_quad(m), m.add_plate(1, 1, 2, 3, 4, 7e10, 0.3, 0.0, 0.8)),
                "Thickness 70000000000.0 is less or equal to zero!"),
    "plate_k": (lambda m: (_quad(m), m.add_plate(1, 1, 2, 3, 4, 7e10, 0.3, 0.01, 0.0)),
                "Shear factor 0.0 is less or equal to zero!"),
    "plate_line": (lambda m: (m.add_node(1, 1, 1, 0), m.add_node(2, 2, 2, 0), m.add_node(3, 0, 0, 0),
                              m.add_node(4, 1, 0, 0), m.add_plate(3, 1, 2, 3, 4, 7e10, 0.3, 0.01, 0.8)),
                   "Some nodes of 3 element lie on the line!"),
    "plate_plane": (lambda m: (m.add_node(1, 1, 1, 0), m.add_node(2, 0, 1, 0), m.add_node(3, 0, 0, 0),
                               m.add_node(4, 1, 0, 0.1), m.add_plate(1, 1, 2, 3, 4, 7e10, 0.3, 0.01, 0.8)),
                    "Not all nodes of element 1 lie on the plane!"),
    "plate_convex": (lambda m: (m.add_node(1, 1, 1, 0), m.add_node(2, 0, 1, 0), m.add_node(3, 0, 0, 0),
                                m.add_node(4, 0.2, 0.8, 0), m.add_plate(1, 1, 2, 3, 4, 7e10, 0.3, 0.01, 0.8)),
                     "Element 1 non-convex!"),
    "plate_same": (lambda m: (_quad(m), m.add_plate(1, 1, 2, 3, 4, 7e10, 0.3, 0.01, 0.8),
                              m.add_plate(2, 4, 3, 2, 1, 7e10, 0.3, 0.01, 0.8)),
                   "Plate element with nodes numbers [4, 3, 2, 1] already exists!"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_validation_matches_crate_and_oracle(name):
    fn, expected = CASES[name]
    got, ref = _both(fn)
    assert got == expected
    assert ref == expected


def test_batched_add_is_prefix_commit():
    m = FEM(1e-4, 1e-12, 10)
    with pytest.raises(FemError, match="already exists"):
        m.add_nodes([1, 2, 3, 2], [[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]])
    # nodes 1..3 were committed: adding them again fails on the first one
    with pytest.raises(FemError, match="Node with number 1 already exists!"):
        m.add_node(1, 9, 9, 9)
    m.add_node(4, 3, 0, 0)


def test_rotation_getter_errors_without_element():
    m = FEM(1e-4, 1e-12, 2)
    with pytest.raises(FemError, match="^Beam element with number 3 does not exist!$"):
        m.get_beam_rotation_matrix_elements(3)


def test_c_abi_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "fem_amd.h")).read()
    names = re.findall(r"^\s*(?:int|void|double|const char\*)\s+(fem_\w+)\s*\(", header, re.M)
    assert len(names) >= 18
    lib = ctypes.CDLL(fa.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU behaviour")
def test_assembly_refuses_without_gpu():
    m = FEM(1e-4, 1e-12, 2)
    m.add_node(1, 0, 0, 0)
    m.add_node(2, 1, 0, 0)
    m.add_truss(1, 1, 2, 1e6, 1.0)
    with pytest.raises(DeviceError, match="no HIP device"):
        m.assemble()
