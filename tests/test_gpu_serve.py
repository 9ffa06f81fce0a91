"""Render server (fl3h_server_*): continuous batching of independent renders.

Requests submitted from several threads at different times are admitted in
arrival order as the GPU frees; each must come out exactly as fl3h_render
makes it alone.
"""
import threading
import time

import numpy as np
import pytest

from flam3_amd import genome as G
from flam3_amd.serve import Server

pytestmark = pytest.mark.gpu


def _req(seed, size=None, quality=None):
    return G.synthetic_preview(seed, size=size, quality=quality), "srv-%d" % seed


def test_server_results_equal_render_alone(prod):
    reqs = [_req(600 + k) for k in range(24)]
    out = [None] * len(reqs)
    with Server() as srv:
        def worker(ks):
            rng = np.random.default_rng(ks[0])
            tickets = []
            for k in ks:
                time.sleep(float(rng.uniform(0, 0.004)))
                g, seed = reqs[k]
                tickets.append((k, srv.submit(g, seed=seed)))
            for k, t in tickets:
                out[k] = srv.wait(t)
        th = [threading.Thread(target=worker, args=(list(range(i, len(reqs), 3)),)) for i in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        batches, renders = srv.counts()
    assert renders == len(reqs)
    for (g, seed), (img, st, lat) in zip(reqs, out):
        want, ws = prod.render(g, seed=seed)
        assert np.array_equal(img, want), seed
        assert st["num_iters"] == ws["num_iters"]
        assert lat > 0


def test_requests_arriving_while_busy_join_the_next_batch(prod):
    big, sbig = _req(700, size=512, quality=100)
    small = [_req(701 + k, size=64, quality=20) for k in range(30)]
    with Server() as srv:
        t0 = srv.submit(big, seed=sbig)
        ts = [srv.submit(g, seed=s) for g, s in small]
        res = [srv.wait(t) for t in ts]
        srv.wait(t0)
        batches, renders = srv.counts()
    assert renders == 31
    # the small requests arrive while the large one is on the GPU and are
    # admitted together as it frees (how many per batch depends on timing)
    assert batches <= 15, batches
    for (g, s), (img, _, _) in zip(small[:5], res[:5]):
        want, _ = prod.render(g, seed=s)
        assert np.array_equal(img, want)


def test_histogram_budget_limits_a_batch(prod):
    small = [_req(800 + k, size=128, quality=10) for k in range(12)]
    per = (128 + 40) ** 2 * 36.0    # an upper bound of one request's bytes
    with Server(hist_budget_bytes=3.5 * per) as srv:
        ts = [srv.submit(g, seed=s) for g, s in small]
        res = [srv.wait(t) for t in ts]
        batches, renders = srv.counts()
    assert renders == 12 and batches >= 4, batches
    for (g, s), (img, _, _) in zip(small[:3], res[:3]):
        want, _ = prod.render(g, seed=s)
        assert np.array_equal(img, want)


def test_batch_of_several_image_groups(prod):
    """A batch above the shared launch's cell limit (32e6 cells) runs as
    several groups; each group's images are downloaded into one of two
    pinned slots and copied out while the next group runs."""
    reqs = [_req(900 + k, size=512, quality=5) for k in range(240)]   # 240 x 0.29e6 cells: 3 groups
    with Server(hist_budget_bytes=1e12) as srv:
        ts = [srv.submit(g, seed=s) for g, s in reqs]
        res = [srv.wait(t) for t in ts]
        batches, renders = srv.counts()
    assert renders == len(reqs)
    for k in (0, 108, 110, 200, 239):
        want, _ = prod.render(reqs[k][0], seed=reqs[k][1])
        assert np.array_equal(res[k][0], want), k
