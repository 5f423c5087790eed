"""Restart files and checkpoints of a device-resident domain (GPU).

* a checkpoint (every carried field, ghosts included) restores a domain that
  continues bit for bit like the uninterrupted one;
* the reference's restart pair (data30/data31, interiors only) restores the
  interiors exactly, and a run continued from it on the GPU follows the C
  restatement continued from the same files (per-step relative L2 <= 1e-5,
  the trajectory bound).
"""
import numpy as np
import pytest

from backends import rel_l2
from lesamd import api, layout, restart, setup_gen
from oracle import oracle

pytestmark = pytest.mark.gpu


def _setup():
    return setup_gen.build(setup_gen.small_workload(36, 28, 18, seed=31))


def test_checkpoint_resume_is_bit_identical(tmp_path):
    s = _setup()
    print("\n[checkpoint] 36x28x18", flush=True)
    whole, first, resumed = api.Domain(s), api.Domain(s), api.Domain(s)
    for d in (whole, first, resumed):
        d.set_graph(True)
    whole.run(1, 7, 7)
    first.run(1, 4, 7)
    restart.save_checkpoint(first, tmp_path / "ck.npz", 4)
    n = restart.load_checkpoint(resumed, tmp_path / "ck.npz")
    assert n == 4
    resumed.run(n + 1, 7, 7)
    for nm in restart.CHECKPOINT_FIELDS:
        assert np.array_equal(whole.download(nm), resumed.download(nm)), nm
    for d in (whole, first, resumed):
        d.close()


def test_reference_restart_pair_continues_like_the_restatement(tmp_path):
    s = _setup()
    w = s.workload
    print("\n[restart pair] 36x28x18", flush=True)
    src = api.Domain(s)
    src.run(1, 4, 8)
    p30, p31 = restart.save_restart(src, tmp_path, 4, 3 * w.dt)
    dst = api.Domain(s)
    n, t = restart.load_restart(dst, p30, p31)
    assert n == 4 and abs(t - 3 * w.dt) < 1e-6
    for nm in restart.DATA30 + restart.DATA31:
        a = layout.interior(nm, dst.download(nm), w.im, w.jm, w.km)
        b = layout.interior(nm, src.download(nm), w.im, w.jm, w.km)
        assert np.array_equal(a, b), nm
    # the restatement started the way the reference's ical == 1 path starts:
    # interiors from the files, every other cell as at program start
    orc = oracle.OracleBackend(s, 32, redmode=1)
    _, _, fields = restart.read_restart(p30, p31, w.im, w.jm, w.km)
    state = {}
    for nm, a in fields.items():
        full = layout.new_field(nm, w.im, w.jm, w.km)
        full[layout.ref_slice(nm, (1, 1, 1), (w.im, w.jm, w.km))] = a
        state[nm] = full
    orc.load(state)
    for step in range(n + 1, n + 5):
        dst.step(step, 8)
        orc.step(step, 8)
        ref = orc.fetch(["u", "v", "w", "p", "sm"])
        for nm in ref:
            e = rel_l2(dst.download(nm), ref[nm])
            assert e <= 1e-5, (step, nm, e)
    src.close()
    dst.close()
