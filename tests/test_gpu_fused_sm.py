"""sm (les.f95:44-79, with boundsm.f95:18-64's ghost cells) computed inside
the fused vel2+velfg+feedbf kernel, the default, against k_les_sm after it
(LES_FUSE_SM=0): every field bit for bit over eager and graph-replayed
steps, for odd and even extents, one to several k levels per thread (the
levels below a thread's first are recomputed there), and diu and sm fields
whose cells vel2 and les never store (diu2, diu3 at i=0; diu7, diu8 at k=0
and k=km; sm at j=-1) holding values of their own, which les reads."""
import os

import numpy as np
import pytest

from lesamd import api, setup_gen

pytestmark = pytest.mark.gpu
FIELDS = ("u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "sm") + tuple(f"diu{q}" for q in range(1, 10))


def _domain(s, **env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return api.Domain(s)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


CASES = [
    dict(im=40, jm=28, km=16, kc=0, graph=False),
    dict(im=37, jm=21, km=9, kc=1, graph=False),
    dict(im=65, jm=6, km=5, kc=3, graph=True),
    dict(im=5, jm=2, km=2, kc=0, graph=False),
    dict(im=128, jm=12, km=11, kc=4, graph=True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['im']}x{c['jm']}x{c['km']}-kc{c['kc']}-{'graph' if c['graph'] else 'eager'}")
def test_fused_sm_equals_les_sm(case):
    w = setup_gen.small_workload(im=case["im"], jm=case["jm"], km=case["km"], seed=51)
    s = setup_gen.build(w, dense=False)
    env = dict(LES_VELFG_KCHUNK=case["kc"])
    fused = _domain(s, **env)
    plain = _domain(s, LES_FUSE_SM=0, **env)
    rng = np.random.default_rng(7)
    # every cell of the diu fields and sm, the ones no kernel stores included
    for name in ("sm",) + tuple(f"diu{q}" for q in range(1, 10)):
        a = fused.download(name)
        r = rng.uniform(-0.5, 0.5, a.shape).astype(np.float32)
        fused.upload(name, r)
        plain.upload(name, r)
    for d in (fused, plain):
        d.set_graph(case["graph"])
    nmax = 20
    for n in range(1, 6):
        for d in (fused, plain):
            d.step(n, nmax)
        for d in (fused, plain):
            d.synchronize()
        for f in FIELDS:
            assert np.array_equal(fused.download(f), plain.download(f)), (n, f)
    assert fused.press_stats()["iters"] == plain.press_stats()["iters"]
    fused.close()
    plain.close()
