"""Building masks built on the device from the height map, where the domain
keeps zbm and the level thresholds and k_velfg_feedbf evaluates the masks
from them, against the same masks given as fields, which k_velfg_feedbf
reads (feedbf.f95:35-61 with the masks of feedbfm.f95:87-116): every field
bit for bit over eager and graph-replayed steps.  An upload of other masks
replaces the height map's: the domain then reads the uploaded fields, bit for
bit as a domain given those masks, and a captured step is captured again."""
import dataclasses

import numpy as np
import pytest

from lesamd import api, setup_gen

pytestmark = pytest.mark.gpu
MASKS = ("bmask1", "cmask1", "dmask1")
FIELDS = ("u", "v", "w", "p", "f", "g", "h", "usum", "vsum", "wsum", "sm") + tuple(f"diu{q}" for q in range(1, 10))


def _steps(doms, n0, n1, nmax):
    for n in range(n0, n1):
        for d in doms:
            d.step(n, nmax)
    for d in doms:
        d.synchronize()


def _same(a, b, tag):
    for f in FIELDS:
        assert np.array_equal(a.download(f), b.download(f)), (tag, f)


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_height_map_masks_equal_given_masks(graph):
    w = setup_gen.small_workload(im=40, jm=28, km=16, seed=41)
    light = setup_gen.build(w, dense=False)
    masks = setup_gen.build(w, dense=True).masks
    assert masks["bmask1"].any() and not masks["bmask1"].all()  # buildings inside the grid
    given = api.Domain(dataclasses.replace(light, masks=masks))
    built = api.Domain(light)
    kept = api.Domain(light)  # keeps the height map's masks throughout
    doms = (given, built, kept)
    for d in doms:
        d.set_graph(graph)
    nmax = 40
    _steps(doms, 1, 7, nmax)
    for n in MASKS:
        assert np.array_equal(given.download(n), built.download(n)), n
    _same(given, built, "height map")
    _same(given, kept, "height map")

    # taller buildings, uploaded to the given and to a built domain
    taller = setup_gen.building_masks(light.zbm * np.float32(2.5), light.metrics, w.im, w.jm, w.km)
    assert int(taller["bmask1"].sum()) > int(masks["bmask1"].sum())
    for d in (given, built):
        for n in MASKS:
            d.upload(n, taller[n])
    _steps(doms, 7, 12, nmax)
    _same(given, built, "uploaded masks")
    # the uploaded masks changed the flow: a domain still on the height map
    # (what an ignored upload would leave) differs
    assert any(not np.array_equal(built.download(f), kept.download(f)) for f in ("usum", "vsum", "wsum"))
    for d in doms:
        d.close()
