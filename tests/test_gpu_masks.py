"""The fused velfg+feedbf (feedbf.f95:35-61 with the masks of
feedbfm.f95:87-116) evaluates masks of buildings standing on the ground from
a value per column and a threshold per level instead of reading the three
mask fields: from the height map when the domain built its masks from it,
and from the top solid level of each column when the masks were given as
fields of that form.  Against the same domain reading the fields
(LES_MASK_FIELDS=1): every field bit for bit over eager and graph-replayed
steps.  An upload of other masks replaces the column form, after which the
fields are read (and a captured step is captured again); given masks of
another form (a floating block, unequal fields) are read as fields."""
import dataclasses
import os

import numpy as np
import pytest

from lesamd import api, layout, setup_gen

pytestmark = pytest.mark.gpu
MASKS = ("bmask1", "cmask1", "dmask1")
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
def test_column_masks_equal_mask_fields(graph):
    w = setup_gen.small_workload(im=40, jm=28, km=16, seed=41)
    light = setup_gen.build(w, dense=False)
    masks = setup_gen.build(w, dense=True).masks
    assert masks["bmask1"].any() and not masks["bmask1"].all()  # buildings inside the grid
    given = dataclasses.replace(light, masks=masks)
    fields = _domain(given, LES_MASK_FIELDS=1)  # reads the three fields
    built = _domain(light)                       # from the height map
    cols = _domain(given)                        # given fields of the column form
    kept = _domain(light)                        # keeps the height map's masks throughout
    doms = (fields, built, cols, kept)
    extra = (w.im * w.jm + w.km) * 4  # the column values and level thresholds
    assert cols.nbytes() - fields.nbytes() == extra  # the given masks were taken in column form
    for d in doms:
        d.set_graph(graph)
    nmax = 40
    _steps(doms, 1, 7, nmax)
    for n in MASKS:
        assert np.array_equal(fields.download(n), built.download(n)), n
    for d, tag in ((built, "height map"), (cols, "columns"), (kept, "height map, kept")):
        _same(fields, d, tag)

    # taller buildings, uploaded to all but one
    taller = setup_gen.building_masks(light.zbm * np.float32(2.5), light.metrics, w.im, w.jm, w.km)
    assert int(taller["bmask1"].sum()) > int(masks["bmask1"].sum())
    for d in (fields, built, cols):
        for n in MASKS:
            d.upload(n, taller[n])
    assert cols.nbytes() == fields.nbytes()  # the column form was dropped
    _steps(doms, 7, 12, nmax)
    _same(fields, built, "uploaded masks")
    _same(fields, cols, "uploaded masks")
    # the uploaded masks changed the flow: a domain still on the height map
    # (what an ignored upload would leave) differs
    assert any(not np.array_equal(built.download(f), kept.download(f)) for f in ("usum", "vsum", "wsum"))
    for d in doms:
        d.close()


@pytest.mark.parametrize("form", ["floating", "unequal"])
def test_other_given_masks_are_read(form):
    w = setup_gen.small_workload(im=36, jm=24, km=12, seed=43)
    light = setup_gen.build(w, dense=False)
    masks = {n: a.copy() for n, a in setup_gen.build(w, dense=True).masks.items()}
    if form == "floating":  # a solid block above open cells: not ground-standing
        for n in MASKS:
            masks[n][layout.ref_slice(n, (3, 4, 6), (8, 9, 7))] = np.float32(1.0)
            masks[n][layout.ref_slice(n, (3, 4, 1), (8, 9, 3))] = np.float32(0.0)
    else:  # a ground-standing block in bmask1 and dmask1 only
        for n in ("bmask1", "dmask1"):
            masks[n][layout.ref_slice(n, (3, 4, 1), (8, 9, 7))] = np.float32(1.0)
    given = dataclasses.replace(light, masks=masks)
    fields = _domain(given, LES_MASK_FIELDS=1)
    other = _domain(given)
    assert other.nbytes() == fields.nbytes()  # not taken in column form
    _steps((fields, other), 1, 6, 20)
    _same(fields, other, form)
    fields.close()
    other.close()
