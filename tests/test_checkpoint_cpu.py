"""OpenLB checkpoint layout, pinned on the CPU against OpenLB's own files.

tests/golden/checkpoints/ holds checkpoint files that the compiled reference
wrote with Serializable::save after k steps (gen_golden.py --checkpoints-only),
with its populations after k and m steps.  The oracle's restatement of the
layout (oracle/checkpoint.py) must rebuild every byte of each file from the
case set-up and the k-step populations: shifts from CyclicColumn::rotate,
population data in rotated order, the other fields' values on every padded
cell.  The product's .dat codec must read and write OpenLB's text exactly.
Where the reference binary is built, OpenLB itself must resume from a
checkpoint that carries only the core cells.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import golden_cases as G
from openlb_amd import checkpoint as CK
from oracle import checkpoint as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_oracle_layout_rebuilds_openlb_checkpoint(name):
    case, meta, arrays, text = G.load_checkpoint(name)
    k = meta["checkpoint_step"]
    real = meta["precision"]
    raw = K.decode_dat(text)
    assert len(raw) == K.serializable_size(case, real)
    shifts, pops, fields = K.unpack(case, raw, real)
    np.testing.assert_array_equal(shifts, K.rotate_shifts(case, k))
    # the k-step populations OpenLB dumped cell by cell are the checkpoint's core cells
    np.testing.assert_array_equal(K.core_of(case, pops).astype(np.float64), arrays[f"f_{k}"])
    for f in K.fields(case):
        np.testing.assert_array_equal(fields[f[0]], K.field_values(case, f).astype(fields[f[0]].dtype),
                                      err_msg=f[2])
    # and every byte, overlap included
    assert K.pack(case, pops, shifts, real) == raw


def test_fixtures_cover_the_field_kinds():
    kinds = set()
    for name in G.checkpoint_names():
        case, *_ = G.load_checkpoint(name)
        kinds |= {f[2] for f in K.fields(case)}
    assert kinds == {"force", "wall_velocity", "rho", "pressure_velocity", "velocity"}


def test_rotate_shift_sign_convention():
    # c_i . projection > 0 keeps OpenLB's shift in (-n, 0], < 0 in [0, n)
    case = G.load_checkpoint(G.checkpoint_names()[0])[0]
    n = int(np.prod(K.padded_shape(case)))
    for steps in (1, 7, n + 3):
        s = K.rotate_shifts(case, steps)
        for i, c in enumerate(K.VELOCITIES[case.q]):
            d = K.neighbor_distance(case, c)
            assert (s[i] % n) == ((-d * steps) % n)
            if d > 0:
                assert -n < s[i] <= 0
            else:
                assert 0 <= s[i] < n


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_dat_codec_matches_openlb(name):
    _, _, _, text = G.load_checkpoint(name)
    raw = CK.decode_dat(text)
    assert raw == K.decode_dat(text)
    assert CK.encode_dat(raw) == text


def test_dat_codec_enforce_uint_and_whitespace():
    raw = bytes(range(256)) * 3 + b"\x01"
    for enforce in (False, True):
        text = CK.encode_dat(raw, enforce)
        assert CK.decode_dat(text, enforce) == raw
        assert CK.decode_dat("\n".join(text[i:i + 60] for i in range(0, len(text), 60)), enforce) == raw
    assert CK.parallel_file_name("cavity3d") == "cavity3d_rank0_size1.dat"
    with pytest.raises(ValueError):
        CK.decode_dat(CK.encode_dat(raw)[:-40])


def _reference(precision):
    exe = os.path.join(ROOT, "oracle", "_ref", f"olb_ref_{precision}")
    if not os.path.exists(exe):
        pytest.skip("OpenLB reference binary not built (needs the reference sources)")
    return exe


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_openlb_resumes_from_core_only_checkpoint(name):
    """A checkpoint built from the core cells alone (zero overlap, as a writer
    that knows only the core would produce) lets OpenLB continue to the same
    populations as its uninterrupted run, bit for bit on the cells that carry
    the flow: the overlap is rewritten before any fluid cell reads it."""
    case, meta, arrays, _ = G.load_checkpoint(name)
    k, m = meta["steps"]
    real = meta["precision"]
    exe = _reference(real)
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import gen_golden
    pops = K.padded_of(case, arrays[f"f_{k}"])
    raw = K.pack(case, pops, K.rotate_shifts(case, k), real)
    with tempfile.TemporaryDirectory() as d:
        gen_golden.write_case_dir(d, case, [m - k])
        CK.write_file(os.path.join(d, CK.parallel_file_name("resume")), raw)
        subprocess.run([exe, d], check=True, cwd=d, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OLB_REF_LOAD=os.path.join(d, "resume")))
        got = np.fromfile(os.path.join(d, f"pops_{m - k}.bin")).reshape((case.q,) + tuple(case.shape))
    role, _ = K.cell_roles(case)
    flow = np.isin(role, ["bulk", "velocity", "pressure"])
    np.testing.assert_array_equal(got[:, flow], arrays[f"f_{m}"][:, flow])


LIBRARY = os.path.join(G.CHECKPOINTS, "library")


def _library_names():
    return sorted(p[:-4] for p in os.listdir(LIBRARY) if p.endswith(".npz")) if os.path.isdir(LIBRARY) else []


def test_library_checkpoints_cover_every_fixture():
    assert _library_names() == G.checkpoint_names()


@pytest.mark.parametrize("name", _library_names())
def test_openlb_resumes_from_library_checkpoint(name):
    """OpenLB loads the checkpoint the GPU library wrote mid-run
    (tests/golden/gen_library_checkpoints.py: OpenLB's state at step k, then
    library steps to k') and continues to step m; it must reach its own
    uninterrupted populations within the time-step bounds of the parity tests
    (FP64 1e-12, FP32 2e-5 of max |f_i|, cells that carry the flow, without
    the slots that stream in from beyond a non-periodic face)."""
    case, meta, arrays, _ = G.load_checkpoint(name)
    k, m = meta["steps"]
    real = meta["precision"]
    z = np.load(os.path.join(LIBRARY, name + ".npz"), allow_pickle=False)
    info = json.loads(str(z["meta"]))
    start = info["step"]
    assert info["precision"] == real and k < start < m
    raw = CK.decode_dat(z["dat"].tobytes().decode("ascii"))
    # the library's file has OpenLB's layout for that step
    shifts, _, fields = K.unpack(case, raw, real)
    np.testing.assert_array_equal(shifts, K.rotate_shifts(case, start))
    for f in K.fields(case):
        np.testing.assert_array_equal(fields[f[0]], K.field_values(case, f).astype(fields[f[0]].dtype))
    exe = _reference(real)
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import gen_golden
    with tempfile.TemporaryDirectory() as d:
        gen_golden.write_case_dir(d, case, [m - start])
        CK.write_file(os.path.join(d, CK.parallel_file_name("library")), raw)
        subprocess.run([exe, d], check=True, cwd=d, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OLB_REF_LOAD=os.path.join(d, "library")))
        got = np.fromfile(os.path.join(d, f"pops_{m - start}.bin")).reshape((case.q,) + tuple(case.shape))
    role, _ = K.cell_roles(case)
    flow = np.isin(role, ["bulk", "velocity", "pressure"])[None] & ~G.streamed_from_outside(case)
    err = G.rel_error(got, arrays[f"f_{m}"], case.q, flow)
    print(f"OpenLB resumed from the library's step-{start} checkpoint of {name}: rel {err:.2e}", flush=True)
    assert err <= {"f64": 1e-12, "f32": 2e-5}[real], err
