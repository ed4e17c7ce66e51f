"""Random GP deciders for synthetic workloads: IR v1 programs (postfix,
depth-bounded, at most 63 nodes and a stack of 6, constants in [-2, 2]: the
limits of the reference's validate_program, cpp/src/program_cpu.cpp:26-62)
and conditional three-tree controllers (python/gp_faco/controller.py:61-82)."""
import struct

import numpy as np

ARITY = {2: 2, 3: 2, 4: 2, 5: 2, 6: 2, 7: 1, 8: 2}


def random_tree(rng, depth):
    """Postfix list of (opcode, operand, constant or None)."""
    if depth == 0 or rng.random() < 0.3:
        if rng.random() < 0.7:
            return [(0, int(rng.integers(0, 12)), None)]
        return [(1, 0, float(np.float32(rng.uniform(-2, 2))))]
    op = int(rng.choice(list(ARITY)))
    out = []
    for _ in range(ARITY[op]):
        out += random_tree(rng, depth - 1)
    return out + [(op, 0, None)]


def stack_peak(nodes):
    top = peak = 0
    for op, _, _ in nodes:
        top = top + 1 if op in (0, 1) else top - ARITY[op] + 1
        peak = max(peak, top)
    return peak


def random_program(rng, depth=5, feature_spec_id=1):
    while True:
        nodes = random_tree(rng, depth)
        if len(nodes) <= 63 and stack_peak(nodes) <= 6:
            break
    opcode, operand, constants = [], [], []
    for op, arg, const in nodes:
        if op == 1:
            arg = len(constants)
            constants.append(struct.unpack("<I", struct.pack("<f", const))[0])
        opcode.append(op)
        operand.append(arg)
    return {"ir_version": 1, "numeric_spec_id": 1, "feature_spec_id": feature_spec_id,
            "opcode": opcode, "operand": operand, "constant_bits": constants}


def random_controller_tree(rng, role, depth=3):
    """A feature-spec-2 tree for stage `role` of a conditional three-tree
    controller: the restart stage (0) reads neither the MNE level nor the
    regional features, the region stage (1) not the MNE level."""
    while True:
        p = {**random_program(rng, depth=depth), "feature_spec_id": 2}
        bad = any(op == 0 and ((role == 0 and (f == 5 or f >= 8)) or (role == 1 and f == 5))
                  for op, f in zip(p["opcode"], p["operand"]))
        if not bad:
            return p


def random_controller(rng, name):
    return {"controller_version": 1, "representation": "conditional_three", "controller_id": name,
            "trees": [random_controller_tree(rng, role) for role in range(3)]}
