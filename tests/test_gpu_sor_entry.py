"""SOR iteration 1 reads the ghost values p had on entry to press: the
reference refreshes the ghosts (boundp1, boundp.f95:28-78) only between the
colour sweeps and (boundp2, boundp.f95:7-26) after them (press.f95:89-122).
When every ghost next to the interior equals the copy boundp would make, as
after any press, the library runs iteration 1 on the kernel of the later
iterations; otherwise on the entry-ghost kernel.  Either way the step must
give the same bits as the entry-ghost kernel alone (LES_SOR_ENTRY_FAST=0):
checked with the ghosts a press leaves, and with the ghosts of each face,
and of all six, perturbed by an upload of p.  The FP64 residual history is
summed per block, and the two kernels group iteration 1's squares
differently: it agrees to rounding."""
import os

import numpy as np
import pytest

from lesamd import api, setup_gen

pytestmark = pytest.mark.gpu
FIELDS = ["u", "v", "w", "p", "rhs", "f", "g", "h", "sm"]


def _domain(s, fast: bool):
    os.environ["LES_SOR_ENTRY_FAST"] = "1" if fast else "0"
    try:
        return api.Domain(s)
    finally:
        os.environ.pop("LES_SOR_ENTRY_FAST", None)


CASES = [((64, 48, 20), None), ((90, 40, 24), None), ((90, 40, 24), "all")] + [
    ((64, 48, 20), f) for f in ("i0", "im1", "j0", "jm1", "k0", "km1")]


@pytest.mark.parametrize("dims,perturb", CASES,
                         ids=[f"{'x'.join(map(str, d))}-{p or 'press-ghosts'}" for d, p in CASES])
def test_iteration_one_equals_the_entry_ghost_kernel(dims, perturb):
    im, jm, km = dims
    w = setup_gen.Workload("sor-entry", im, jm, km, seed=41)
    s = setup_gen.build(w, dense=False)
    print(f"\n[sor entry] {dims} perturb={perturb}", flush=True)
    fast, ref = _domain(s, True), _domain(s, False)
    for n in (1, 2):
        fast.step(n, 6)
        ref.step(n, 6)
    if perturb:
        p = fast.download("p")  # p(0:im+2, 0:jm+2, 0:km+1)
        assert np.array_equal(p, ref.download("p"))
        rng = np.random.default_rng(7)
        faces = {"i0": np.s_[0, 1:jm + 1, 1:km + 1], "im1": np.s_[im + 1, 1:jm + 1, 1:km + 1],
                 "j0": np.s_[1:im + 1, 0, 1:km + 1], "jm1": np.s_[1:im + 1, jm + 1, 1:km + 1],
                 "k0": np.s_[1:im + 1, 1:jm + 1, 0], "km1": np.s_[1:im + 1, 1:jm + 1, km + 1]}
        for f in (faces if perturb == "all" else [perturb]):
            p[faces[f]] += rng.uniform(-0.5, 0.5, p[faces[f]].shape).astype(np.float32)
        fast.upload("p", p)
        ref.upload("p", p)
    for n in (3, 4):
        fast.step(n, 6)
        ref.step(n, 6)
    for f in FIELDS:
        assert np.array_equal(fast.download(f), ref.download(f)), f
    sf, sr = fast.press_stats(), ref.press_stats()
    assert sf["iters"] == sr["iters"]
    np.testing.assert_allclose(sf["sor"], sr["sor"], rtol=1e-12, atol=0)
    fast.close()
    ref.close()
