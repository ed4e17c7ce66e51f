"""Host-side product code and the C ABI boundary (no GPU needed).

- libfaco_hip.so loads and exports every entry point include/faco_engine.h
  declares; record layouts match the reference ABI sizes;
- the product's host preparation equals the oracle's;
- the product's score_cpu equals the reference score_cpu golden vectors and
  the reference's known answers (tests/python/test_program_ir.py:12-36,74-80,
  tests/python/test_native_scoring.py:41-46);
- input validation raises ValueError/TypeError like the reference module.
"""
import ctypes
import subprocess
import json
import re
import struct
from pathlib import Path

import numpy as np
import pytest

from gplsaco_amd import _native, gp_faco_ext as ext
from oracle import oracle as O
from gplsaco_amd.programs import random_program

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "faco_engine.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(faco_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
        assert name in _native.SIGNATURES, name
    ext.lib()  # every ctypes signature binds


def test_record_layouts(tmp_path):
    assert ctypes.sizeof(_native.Settings) == 64        # FixedFacoSettings
    assert ctypes.sizeof(_native.Program) == 388        # Program IR v1
    assert ctypes.sizeof(_native.Controller) == 1168    # ControllerProgram
    assert ctypes.sizeof(_native.BaselinePolicy) == 48  # BaselinePolicy
    # every ctypes mirror has the size and field offsets the C compiler gives the header's record
    mirrors = {"faco_settings": _native.Settings, "faco_ant_info": _native.AntInfo, "faco_result": _native.Result,
               "faco_program": _native.Program, "faco_controller": _native.Controller,
               "faco_baseline_policy": _native.BaselinePolicy, "faco_control_state": _native.ControlState,
               "faco_control_spec": _native.ControlSpec, "faco_deadline_spec": _native.DeadlineSpec,
               "faco_deadline_result": _native.DeadlineResult, "faco_batch_profile": _native.BatchProfile,
               "faco_fork": _native.Fork, "faco_tune_entry": _native.TuneEntry}
    lines = ['#include <cstddef>', '#include <cstdio>', f'#include "{ROOT / "include" / "faco_engine.h"}"',
             'int main() {']
    for c_name, mirror in mirrors.items():
        lines.append(f'    std::printf("{c_name} %zu\\n", sizeof({c_name}));')
        for field, _ in mirror._fields_:
            lines.append(f'    std::printf("{c_name}.{field} %zu\\n", offsetof({c_name}, {field}));')
    lines.append('    return 0;\n}')
    source = tmp_path / "layout.cpp"
    source.write_text("\n".join(lines))
    binary = tmp_path / "layout"
    subprocess.run(["g++", "-std=c++17", str(source), "-o", str(binary)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(binary)], check=True, capture_output=True,
                                                               text=True).stdout.splitlines())
    for c_name, mirror in mirrors.items():
        assert int(got[c_name]) == ctypes.sizeof(mirror), c_name
        for field, _ in mirror._fields_:
            assert int(got[f"{c_name}.{field}"]) == getattr(mirror, field).offset, (c_name, field)


@pytest.mark.parametrize("n,grid,widths", [(5, 10000, (16, 64, 20)), (60, 9, (16, 64, 20)),
                                           (300, 10000, (16, 64, 20)), (200, 10000, (2, 1, 5)),
                                           (120, 30, (20, 40, 8))])
def test_host_preparation_equals_oracle(n, grid, widths):
    rng = np.random.default_rng(n + grid)
    xy = rng.integers(0, grid + 1, size=(n, 2))
    s = ext.FixedFacoSettings()
    s.primary_width, s.backup_width, s.ls_width = widths
    got = ext.prepare_rows(xy.astype(np.float64), s)
    want = O.prepare(xy, O.default_settings(primary_width=widths[0], backup_width=widths[1], ls_width=widths[2]))
    for key in ("primary", "backup", "ls", "local_scale", "initial_tour"):
        np.testing.assert_array_equal(got[key], want[key])
    assert got["initial_cost"] == want["initial_cost"]
    assert got["settings"].ants == want["settings"].ants
    common = ext.prepare_common_initial(xy.astype(np.float64), s)
    assert common["tour"] == want["initial_tour"].tolist() and common["cost"] == float(want["initial_cost"])


def test_host_start_nodes_equal_oracle():
    counts = np.array([3, 1000, 10000, 2**31 + 1, 0xFFFFFFFF], dtype=np.uint32)
    for key in range(40):
        got = ext.diagnostic_start_nodes(key * 0x9E3779B97F4A7C15 % 2**64, key, 7, counts)
        want = [O.start_node(key * 0x9E3779B97F4A7C15 % 2**64, key, 7, int(c)) for c in counts]
        assert got.tolist() == want


def test_preparation_rejects_non_integer_or_out_of_range_coordinates():
    s = ext.FixedFacoSettings()
    with pytest.raises(ValueError):
        ext.prepare_common_initial(np.array([[0.5, 1], [2, 3], [4, 5]], dtype=np.float64), s)
    with pytest.raises(ValueError):
        ext.prepare_common_initial(np.array([[0, 1], [2, 3], [40000, 5]], dtype=np.float64), s)
    with pytest.raises(TypeError):
        ext.prepare_common_initial(np.array([[0, 1], [2, 3], [4, 5]], dtype=np.int64), s)
    bad = ext.FixedFacoSettings()
    bad.retention = 1.0
    with pytest.raises(ValueError):
        ext.prepare_common_initial(np.array([[0, 1], [2, 3], [4, 5]], dtype=np.float64), bad)
    bad = ext.FixedFacoSettings()
    bad.ants = True
    with pytest.raises(TypeError):
        ext.prepare_common_initial(np.array([[0, 1], [2, 3], [4, 5]], dtype=np.float64), bad)


# ---------------------------------------------------------------------------
# scoring

def program(opcode, operand, constants=()):
    return {"ir_version": 1, "numeric_spec_id": 1, "feature_spec_id": 1, "opcode": list(opcode),
            "operand": list(operand), "constant_bits": [struct.unpack("<I", struct.pack("<f", c))[0]
                                                        for c in constants]}


def test_score_cpu_golden_vectors():
    data = json.loads((ROOT / "tests" / "golden" / "score_cpu.json").read_text())
    features = np.array(data["features_bits"], dtype=np.uint32).view(np.float32)
    masks = np.array(data["masks"], dtype=np.uint32)
    for case in data["cases"]:
        out = ext.score_cpu(case["program"], features, masks)
        assert out["scores"].view(np.uint32).tolist() == case["scores_bits"]
        assert out["actions"].tolist() == case["actions"]


def test_score_known_answers_from_reference_tests():
    # SUB(MUL(return_rate, restart), AQ(ls_work, mne_level)): opcodes
    # (0,0,4,0,0,8,3), operands (2,4,0,3,5,0,0); value 0.5 - 0.75/sqrt(1.25)
    features = np.zeros((12, 1, 32), dtype=np.float32)
    features[[2, 4, 3, 5], 0, :] = np.array([0.5, 1.0, 0.75, 0.5], dtype=np.float32)[:, None]
    out = ext.score_cpu(program((0, 0, 4, 0, 0, 8, 3), (2, 4, 0, 3, 5, 0, 0)), features,
                        np.array([0xFFFFFFFF], dtype=np.uint32))
    assert np.allclose(out["scores"], 0.5 - 0.75 / np.sqrt(1.25), atol=1e-7)
    # every intermediate clipped to [-8, 8]: SUB(MUL(MUL(2,2),MUL(2,2)),2) = 6
    clip = program((1, 1, 4, 1, 1, 4, 4, 1, 3), (0, 1, 0, 2, 3, 0, 0, 4, 0), (2.0, 2.0, 2.0, 2.0, 2.0))
    assert ext.score_cpu(clip, np.zeros((12, 1, 32), np.float32), np.array([1], np.uint32))["scores"][0, 0] == 6.0
    # masked ties: constant score, masks pick the lowest legal action
    masks = np.array([0xFFFFFFFF, 1 << 31, 1 << 7], dtype=np.uint32)
    out = ext.score_cpu(program((0,), (4,)), np.zeros((12, 3, 32), np.float32), masks)
    assert out["actions"].tolist() == [0, 31, 7]
    # global epsilon ties: within 1e-6 of the legal maximum, smallest id wins
    features = np.zeros((12, 3, 32), dtype=np.float32)
    features[0, 0, :3] = [0, 0.75e-6, 1.5e-6]
    out = ext.score_cpu(program((0,), (0,)), features, np.array([7, 1 << 31, 0xFFFFFFFF], dtype=np.uint32))
    assert out["actions"].tolist() == [1, 31, 0]


def test_score_input_validation():
    features = np.zeros((12, 3, 32), dtype=np.float32)
    masks = np.array([1, 2, 4], dtype=np.uint32)
    p = program((0,), (4,))
    with pytest.raises((TypeError, ValueError)):
        ext.score_cpu(p, features.astype(np.float64), masks)
    with pytest.raises((TypeError, ValueError)):
        ext.score_cpu(p, features[:, :, ::-1], masks)
    with pytest.raises(ValueError):
        ext.score_cpu(p, features, np.zeros(3, dtype=np.uint32))
    nan = features.copy()
    nan[0, 0, 0] = np.nan
    with pytest.raises(ValueError):
        ext.score_cpu(p, nan, masks)
    for malformed in ({**p, "ir_version": 4}, {**p, "ir_version": True}, {**p, "opcode": [False]},
                      {**p, "operand": [12]}, {**p, "extra": 1}, program((3,), (0,)),
                      program((0, 0), (0, 1)), program((0,) + (7,) * 6, (0,) * 7),
                      program((1,), (0,), (3.0,)), {**p, "numeric_spec_id": 2}):
        with pytest.raises(ValueError):
            ext.score_cpu(malformed, features, masks)


def test_score_cpu_matches_reference_core_live():
    if not O.ref_available():
        pytest.skip("reference checker oracle/_ref not built here")
    rng = np.random.default_rng(99)
    features = rng.random((12, 9, 32), dtype=np.float32)
    masks = rng.integers(1, 0xFFFFFFFF, size=9, dtype=np.uint32)
    for _ in range(300):
        p = random_program(rng, depth=5)
        consts = [struct.unpack("<f", struct.pack("<I", b))[0] for b in p["constant_bits"]]
        s, a = O.ref_score_cpu(p["opcode"], p["operand"], consts, features, masks)
        out = ext.score_cpu(p, features, masks)
        assert out["scores"].view(np.uint32).tolist() == s.view(np.uint32).tolist()
        assert out["actions"].tolist() == a.tolist()


def test_score_controller_host_matches_oracle():
    """The product's host controller scorer (the device kernel's twin) against
    the oracle's restatement of cuda/gp_score.cuh:37-62."""
    rng = np.random.default_rng(12)
    features = rng.random((12, 7, 32), dtype=np.float32)
    masks = rng.integers(1, 2**32, size=7, dtype=np.uint32) | 1
    masks[:3] = [0xFFFF, 1 << 16 | 1, 1 << 5]
    for _ in range(60):
        trees = []
        for role in range(3):
            while True:
                p = {**random_program(rng, depth=3), "feature_spec_id": 2}
                if not any(op == 0 and ((role == 0 and (f == 5 or f >= 8)) or (role == 1 and f == 5))
                           for op, f in zip(p["opcode"], p["operand"])):
                    break
            trees.append(p)
        d = {"controller_version": 1, "representation": "conditional_three", "controller_id": "c", "trees": trees}
        out = ext.score_controller_cpu(d, features, masks)
        c = O.Controller.from_buffer_copy(bytes(ext.read_controller(d)))
        for col in range(7):
            scores, action = O.score_controller(c, features[:, col, :], int(masks[col]))
            assert out["actions"][col] == action
            assert out["scores"][col].view(np.uint32).tolist() == scores.view(np.uint32).tolist()
    with pytest.raises(ValueError):  # a restart tree may not read region features
        bad = {"controller_version": 1, "representation": "conditional_three", "controller_id": "bad",
               "trees": [program((0,), (9,)), program((0,), (1,)), program((0,), (1,))]}
        bad["trees"] = [{**t, "feature_spec_id": 2} for t in bad["trees"]]
        ext.score_controller_cpu(bad, features, masks)
