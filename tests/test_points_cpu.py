"""CPU: Points modification tracking and the lazy tree cache (host-side
mirror of pptk/points/points.py:86-166, expr.py:397-424), with the device
tree replaced by a recording stand-in so no GPU is needed."""
import numpy as np
import pytest

from pptk_amd import points as P


class FakeKd:
    def __init__(self):
        self.builds = 0

    def _build(self, a):
        self.builds += 1
        return ("tree", np.array(a, copy=True))

    def _query(self, tree, queries=None, k=1, dmax=None):
        pts = tree[1]
        q = np.arange(len(pts)) if queries is None else np.asarray(queries)
        out = []
        for i in q:
            d = ((pts - pts[i]) ** 2).sum(1)
            d[i] = np.inf
            out.append(np.argsort(d, kind="stable")[:k].astype(np.int32))
        return out


@pytest.fixture
def fake(monkeypatch):
    f = FakeKd()
    monkeypatch.setattr(P, "kdtree", f)
    return f


def test_write_locked_and_modifications_rebuild(fake):
    x = P.rand(200, 3)
    assert not x.flags.writeable
    with pytest.raises(ValueError):
        np.ndarray.__setitem__(x, 0, 1.0)       # bypassing Points is refused
    x.nbhds(k=4)
    x.nbhds(k=5)
    assert fake.builds == 1                      # unchanged points reuse the tree
    x[3] = [5, 5, 5]
    r = x.nbhds(k=2)
    assert fake.builds == 2
    assert len(r) == 200
    x += 1.0
    x.nbhds(queries=[0, 1], k=2)
    assert fake.builds == 3


def test_views_share_the_memory_stamp(fake):
    x = P.rand(100, 3)
    v = x[10:20]
    x.nbhds(k=2)
    v.nbhds(k=2)
    assert fake.builds == 2                      # each object holds its own tree
    v[0] = [9, 9, 9]                             # modify through the view
    x.nbhds(k=2)
    assert fake.builds == 3                      # the base sees the change
    assert np.array_equal(x[10], [9, 9, 9])
    assert not x.flags.writeable and not v.flags.writeable


def test_copy_is_independent(fake):
    x = P.rand(50, 3)
    y = x.copy()
    assert isinstance(y, P.Points) and not y.flags.writeable
    x.nbhds(k=1)
    y.nbhds(k=1)
    y[0] = [3, 3, 3]
    x.nbhds(k=1)
    assert fake.builds == 2                      # x did not change


def test_constructors():
    assert np.all(P.zeros((4, 3)) == 0) and np.all(P.ones((4, 3)) == 1)
    a = P.points([[1, 2, 3], [4, 5, 6]], dtype=np.float32)
    assert isinstance(a, P.Points) and a.dtype == np.float32
    assert np.all(P.ones_like(a) == 1) and P.zeros_like(a).shape == a.shape
    with pytest.raises(TypeError):
        np.zeros((3, 3)).view(P.Points)          # views of non-Points refused
    with pytest.raises(TypeError):
        P.points(np.zeros((3, 3)), copy=False)
    with pytest.raises(TypeError):
        class Sub(P.Points):                      # not subclass-able
            pass


def test_nbhds_op_lazy_evaluation(fake):
    x = P.rand(2500, 3)
    e = x.NBHDS(k=3)
    assert len(e) == 2500
    direct = x.nbhds(k=3)
    ev = e.evaluate()
    assert len(ev) == 2500
    assert all(np.array_equal(a, b) for a, b in zip(ev, direct))
    assert all(np.array_equal(a, b) for a, b in zip(list(e), direct))
    assert np.array_equal(e[7], direct[7]) and np.array_equal(e[-1], direct[-1])
    with pytest.raises(IndexError):
        e[2500]
    q = x.NBHDS(queries=np.array([5, 6, 7]), k=2)
    assert len(q) == 3 and "3 elements" in repr(q)
