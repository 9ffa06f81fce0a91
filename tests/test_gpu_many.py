"""fl3h_render_many: several renders sharing the GPU.

Small renders (below 2^24 plotted samples, one batch, production mode) run
their walkers in one launch, each into its own histogram, then their
density estimation and tone maps.  Every request must come out exactly as
fl3h_render makes it alone (same walkers, keys and integer deposits), and
the same alone and in any mix with other requests.  Requests that cannot
share (large, other bits) run one after another with the same result as
fl3h_render.
"""
import numpy as np
import pytest

from flam3_amd import genome as G

pytestmark = pytest.mark.gpu


def _preview(seed, size=None, quality=None):
    return {"genomes": G.synthetic_preview(seed, size=size, quality=quality), "seed": "req-%d" % seed}


def _animated(seed):
    a = G.synthetic_preview(seed, size=96, quality=20)
    b = G.synthetic_preview(seed, size=96, quality=20)
    b.rotate = 40.0
    a.time, b.time = 0.0, 1.0
    for g in (a, b):
        g.temporal_samples = 5
    return {"genomes": [a, b], "time": 0.3, "seed": "anim-%d" % seed}


def _alone(prod, it):
    return prod.render(it["genomes"], seed=it["seed"], time=it.get("time", 0.0), bits=it.get("bits", 33),
                       nchan=it.get("nchan", 3), transp=it.get("transp", 0), bpc=it.get("bpc", 1),
                       earlyclip=it.get("earlyclip", 0))


def _mix():
    items = [_preview(s) for s in (301, 302, 303, 304, 305, 306)]
    items.append(_animated(307))
    items.append(dict(_preview(308, size=64, quality=15), nchan=4, transp=1))
    items.append(dict(_preview(309, size=128, quality=30), earlyclip=1))
    items.append(dict(_preview(310, size=64, quality=40), bpc=2))
    return items


def test_shared_renders_equal_fl3h_render_alone(prod):
    items = _mix()
    got = prod.render_many(items)
    for it, (img, st) in zip(items, got):
        want, ws = _alone(prod, it)
        assert st["num_iters"] == ws["num_iters"]
        assert st["badvals"] == ws["badvals"]
        assert np.array_equal(img, want), (it["seed"], np.abs(img.astype(int) - want.astype(int)).max())
        assert img.max() > 0


@pytest.mark.parametrize("order", [0, 1, 2])
def test_result_is_the_same_alone_and_in_any_mix(prod, order):
    items = _mix()
    rng = np.random.default_rng(order)
    perm = rng.permutation(len(items))
    mixed = prod.render_many([items[i] for i in perm])
    for k, i in enumerate(perm[:4]):
        alone = prod.render_many([items[i]])[0][0]
        assert np.array_equal(mixed[k][0], alone), items[i]["seed"]


def test_requests_that_cannot_share_render_alone(prod):
    """A linear-palette render above 2^24 samples (binned deposit), one with
    bits 64 among bits-33 ones and a request with several batches each take
    fl3h_render's own path; a stepped-palette render above 2^24 samples
    shares the launch with saturating deposits and still equals
    fl3h_render's binned image; the shared ones around them are
    unaffected."""
    big = G.synthetic_preview(401, size=512, quality=100)   # 2.6e7 samples
    big_lin = G.synthetic_preview(406, size=512, quality=100)
    big_lin.palette_mode = "linear"
    b64 = _preview(402, size=64, quality=10)
    b64["bits"] = 64
    multi = G.synthetic_preview(403, size=64, quality=20)
    multi.passes = 2
    items = [_preview(404), {"genomes": big, "seed": "big"}, b64, {"genomes": multi, "seed": "multi"},
             {"genomes": big_lin, "seed": "big-linear"}, _preview(405)]
    got = prod.render_many(items)
    for it, (img, st) in zip(items, got):
        want, ws = _alone(prod, it)
        assert st["num_iters"] == ws["num_iters"]
        assert np.array_equal(img, want), it["seed"]


def test_many_small_requests_in_one_call(prod):
    items = [_preview(500 + k, size=64, quality=10) for k in range(64)]
    got = prod.render_many(items)
    for k in (0, 17, 63):
        want, _ = _alone(prod, items[k])
        assert np.array_equal(got[k][0], want)


def test_frames_sharing_a_palette_equal_fl3h_render_alone(prod):
    """Frames of one animation share a palette: the shared launch writes
    4-byte records (cell | palette index) into the frames' joint cell
    space; with a second palette in the group it writes job records.  Each
    frame equals fl3h_render's image either way."""
    frames = []
    for t in range(6):
        it = _animated(311)
        it["time"] = 0.15 * t
        it["seed"] = "frame-%d" % t
        frames.append(it)
    got = prod.render_many(frames)
    mixed = prod.render_many(frames[:3] + [_preview(312)] + frames[3:])
    for k, it in enumerate(frames):
        want, ws = _alone(prod, it)
        assert got[k][1]["num_iters"] == ws["num_iters"]
        assert np.array_equal(got[k][0], want), it["seed"]
        assert np.array_equal(mixed[k if k < 3 else k + 1][0], want), it["seed"]
