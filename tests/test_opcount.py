"""Operation count used for the walker kernel's roofline."""
import pytest

from flam3_amd import genome as G
from flam3_amd import opcount as O


def test_every_workload_variation_is_counted():
    for v in G.WORKLOAD_VARS:
        assert v in O.VARIATION_OPS


def test_hand_counted_genome():
    # xform A: linear only -> 16 + 4 = 20; xform B: swirl -> 16 + 12 = 28;
    # weights 1:3; final julia -> 16 + 9.5 + 1 (atan2) = 26.5; deposit 5
    a = G.Xform(weight=1.0, variations={"linear": 1.0})
    b = G.Xform(weight=3.0, variations={"swirl": 0.5})
    fin = G.Xform(weight=0.0, variations={"julia": 1.0})
    g = G.Genome(xforms=[a, b], final=fin)
    assert O.ops_per_sample(g) == pytest.approx(0.25 * 20 + 0.75 * 28 + 26.5 + 5)
    # precalc terms are paid once per xform, not per variation
    c = G.Xform(weight=1.0, variations={"spiral": 1.0, "diamond": 1.0, "polar": 1.0})
    assert O.xform_ops(c) == 16 + 10 + 8 + 6 + 2 + 1
    # a post affine and a rotated camera add their arithmetic
    c.post = (1.0, 0.5, 0.0, 1.0, 0.0, 0.0)
    assert O.xform_ops(c) == 16 + 10 + 8 + 6 + 2 + 1 + 8
    g2 = G.Genome(xforms=[c], rotate=30.0)
    assert O.ops_per_sample(g2) == O.xform_ops(c) + 5 + 10


def test_headline_genome_is_about_a_hundred_ops():
    ops = O.ops_per_sample(G.synthetic_1080p(seed=2))
    assert 60 < ops < 140


def test_unknown_variation_is_an_error():
    g = G.Genome(xforms=[G.Xform(weight=1.0, variations={"blob": 1.0})])
    with pytest.raises(KeyError):
        O.ops_per_sample(g)
