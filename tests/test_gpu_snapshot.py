"""GPU parity of snapshot load/save and the replay harness
(src/plane_wave_operator.cpp:520-653, src/solver_replay_cli.cpp): the device
operator restored from a snapshot applies bitwise as the original; oracle-
written PWO3 and the PWO2/PWO1 legacy forms agree with the oracle operator."""
import json
import os

import numpy as np
import pytest

import pysnapshot
from sdft_amd import native
from sdft_amd import replay
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def si8_device(hip_ctx, si8):
    return native.Operator(hip_ctx, si8)


def test_save_load_round_trip_bitwise(tmp_path, hip_ctx, si8, si8_oracle, si8_device):
    x = syn.initial_psi(si8, 13, 0)
    path = str(tmp_path / "dev.pwop")
    si8_device.save_snapshot(path, 13, 256, 1e-12, x)
    s = pysnapshot.read(path)
    assert s["form"] == 3 and s["roots"] == 13 and np.array_equal(s["initial"], x)
    assert np.array_equal(s["indices"], si8_oracle.indices)
    assert np.array_equal(s["kinetic"], si8_oracle.kinetic)
    assert np.abs(s["local_real"] - si8_oracle.local_real).max() < 1e-12
    ref = pysnapshot.from_problem(si8, si8_oracle, 13, 256, 1e-12)
    assert np.array_equal(s["all_b"], ref["all_b"])
    for a, b in zip(s["onsite_d"] + s["onsite_q"], ref["onsite_d"] + ref["onsite_q"]):
        assert np.array_equal(a, b)
    op, warm = native.Operator.from_snapshot(hip_ctx, path)
    assert np.array_equal(warm, x) and op.shape == si8_device.shape
    for which in ("h", "s"):
        assert np.array_equal(op.apply(x, which, precision=0), si8_device.apply(x, which, precision=0))
    with pytest.raises(native.HipError, match="refusing to overwrite"):
        si8_device.save_snapshot(path, 13, 256, 1e-12, x)


@pytest.mark.parametrize("form", [3, 2, 1])
def test_oracle_snapshot_forms_against_oracle(tmp_path, hip_ctx, si8, si8_oracle, form):
    base = pysnapshot.from_problem(si8, si8_oracle, 13, 256, 1e-12)
    s = base if form == 3 else pysnapshot.legacy(base, form)
    path = str(tmp_path / f"o{form}.pwop")
    pysnapshot.write(path, s)
    op, warm = native.Operator.from_snapshot(hip_ctx, path)
    assert warm is None
    x = syn.initial_psi(si8, 13, 0)
    rh, rs = si8_oracle.apply(x, "pair")
    gh, gs = op.apply(x, "pair", precision=0)
    tol = 1e-12 if form == 3 else 1e-10  # legacy forms: tests/native/plane_wave_operator_smoke.cpp:201
    assert np.abs(gh - rh).max() / np.abs(rh).max() < tol
    assert np.abs(gs - rs).max() / np.abs(rs).max() < tol
    # the joint pair equals the separate applications bitwise (smoke test code 27)
    assert np.array_equal(op.apply(x, "h", precision=0), gh)
    assert np.array_equal(op.apply(x, "s", precision=0), gs)
    fh = op.apply(x, "h", precision=1)
    assert np.abs(fh - rh).max() / np.abs(rh).max() < 2e-5


def test_replay_harness(tmp_path, hip_ctx, si8, si8_device):
    import scipy.linalg as sl
    init = syn.initial_psi(si8, 16, 0)
    snap = str(tmp_path / "replay.pwop")
    si8_device.save_snapshot(snap, 16, 256, 1e-12, init)
    report = str(tmp_path / "report.json")
    warm = str(tmp_path / "warm.pwop")
    out = replay.replay(snap, report, warm_snapshot=warm, ctx=hip_ctx)
    assert out["passed"] and out["solver_converged"] and out["fresh_max_residual"] <= 1e-12
    assert out["s_normalization_error"] <= 1e-12
    on_disk = json.loads(open(report).read())
    assert on_disk["schema_version"] == "sparse-dft-solver-replay-1" and on_disk["dimension"] == si8.n
    h, s = syn.dense_operator(si8)
    ref = sl.eigh(h, s, eigvals_only=True, subset_by_index=[0, 15])
    assert np.abs(np.array(on_disk["eigenvalues_hartree"]) - ref).max() < 1e-10
    assert os.path.exists(report + ".performance.json")
    with pytest.raises(RuntimeError, match="refusing to overwrite"):
        replay.replay(snap, report, ctx=hip_ctx)
    # the warm snapshot carries the certified vectors: replaying it converges at once
    out2 = replay.replay(warm, str(tmp_path / "report2.json"), ctx=hip_ctx)
    assert out2["passed"] and out2["iterations"] <= out["iterations"]


def test_replay_guards_identity(tmp_path, hip_ctx, si8, si8_device):
    snap = str(tmp_path / "g.pwop")
    si8_device.save_snapshot(snap, 8, 256, 1e-12, syn.initial_psi(si8, 8, 0))
    guards = syn.initial_psi(si8, 4, 1)
    replay.save_search_guards(snap + ".guards", "0" * 64, guards)
    with pytest.raises(RuntimeError, match="identity"):
        replay.replay(snap, str(tmp_path / "r1.json"), ctx=hip_ctx)
    os.remove(snap + ".guards")
    replay.save_search_guards(snap + ".guards", replay.file_sha256(snap), guards)
    out = replay.replay(snap, str(tmp_path / "r2.json"), ctx=hip_ctx)
    assert out["passed"]
    perf = json.loads(open(str(tmp_path / "r2.json") + ".performance.json").read())
    assert perf["guard_columns_validated"] == 4


def test_project_and_projector_metric(si8, si8_device):
    """project_into / projector_metric_into (src/plane_wave_operator.cpp:375-389):
    P = B X against the stacked projectors; Q P per atom bitwise equal to the
    oracle's compensated product (src/paw_compensated_product.hpp order)."""
    import pyoracle
    x = syn.initial_psi(si8, 9, 2)
    b = np.concatenate([np.asarray(a) for a in si8.amplitudes], axis=0)
    p = si8_device.project(x)
    ref = b @ x
    assert np.abs(p - ref).max() / np.abs(ref).max() < 1e-13
    p32 = si8_device.project(x, precision=1)
    assert np.abs(p32 - ref).max() / np.abs(ref).max() < 2e-5
    qp = si8_device.projector_metric(p)
    o = 0
    bits = lambda a: np.ascontiguousarray(a).view(np.uint64)  # noqa: E731
    for q in si8.q_blocks:
        r = q.shape[0]
        qh = 0.5 * (q + q.conj().T)
        expect = pyoracle.compensated_product(0, qh, p[o:o + r])
        assert np.array_equal(bits(qp[o:o + r]), bits(expect))
        o += r
