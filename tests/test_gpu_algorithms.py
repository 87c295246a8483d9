"""GPU parity of the reference's remaining algorithms (transform, replace,
for_each, adjacent_difference, sequence, tabulate, gather / scatter, swap,
fill, equal, extrema, inner_product, all / any / none_of, find, mismatch,
bounds and binary searches, partition_point, is_partitioned,
is_sorted_until, merge, merge_by_key, DVTuple) against oracle/algorithms.py,
through the reference-shaped Python API.  Integer results bit-exact; the
float inner product within a relative 1e-6 of a float64 sum."""
import json
import os

import numpy as np
import pytest

from helpers import oracle_algorithms, workloads

pytestmark = pytest.mark.gpu

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "algorithms_known_answers.json")))


@pytest.fixture(scope="module")
def trtc():
    import thrustrtc_amd
    return thrustrtc_amd


@pytest.fixture(scope="module")
def oa():
    return oracle_algorithms()


@pytest.fixture(scope="module")
def wl():
    return workloads()


def say(*a):
    print(*a, flush=True)


def dv(trtc, x, elem="int32_t"):
    return trtc.device_vector_from_list(list(x), elem)


def test_known_answers_elementwise(trtc):
    odd = trtc.Functor({}, ["x"], "        return x % 2;\n")
    neg = trtc.Functor({}, ["x"], "        return x < 0;\n")
    even = trtc.Functor({}, ["x"], "        return ((x % 2) == 0);\n")
    c = KA["transform_negate"]
    d = dv(trtc, c["in"])
    trtc.Transform(d, d, trtc.Negate())
    assert d.to_host().tolist() == c["out"]
    c = KA["transform_binary_plus"]
    o = trtc.device_vector("int32_t", 6)
    trtc.Transform_Binary(dv(trtc, c["in1"]), dv(trtc, c["in2"]), o, trtc.Plus())
    assert o.to_host().tolist() == c["out"]
    c = KA["transform_if_negate_odd"]
    d = dv(trtc, c["in"])
    trtc.Transform_If(d, d, trtc.Negate(), odd)
    assert d.to_host().tolist() == c["out"]
    c = KA["transform_if_stencil"]
    d = dv(trtc, c["in"])
    trtc.Transform_If_Stencil(d, dv(trtc, c["stencil"]), d, trtc.Negate(), trtc.Identity())
    assert d.to_host().tolist() == c["out"]
    c = KA["transform_binary_if_stencil"]
    o = dv(trtc, c["init"])
    trtc.Transform_Binary_If_Stencil(dv(trtc, c["in1"]), dv(trtc, c["in2"]), dv(trtc, c["stencil"]), o, trtc.Plus(),
                                     trtc.Identity())
    assert o.to_host().tolist() == c["out"]
    c = KA["replace"]
    d = dv(trtc, c["in"])
    trtc.Replace(d, trtc.DVInt32(c["old"]), trtc.DVInt32(c["new"]))
    assert d.to_host().tolist() == c["out"]
    o = trtc.device_vector("int32_t", 5)
    trtc.Replace_Copy(dv(trtc, c["in"]), o, trtc.DVInt32(c["old"]), trtc.DVInt32(c["new"]))
    assert o.to_host().tolist() == c["out"]
    c = KA["replace_if_negative"]
    d = dv(trtc, c["in"])
    trtc.Replace_If(d, neg, trtc.DVInt32(c["new"]))
    assert d.to_host().tolist() == c["out"]
    trtc.Replace_Copy_If(dv(trtc, c["in"]), o, neg, trtc.DVInt32(c["new"]))
    assert o.to_host().tolist() == c["out"]
    c = KA["gather"]
    o = trtc.device_vector("int32_t", 10)
    trtc.Gather(dv(trtc, c["map"]), dv(trtc, c["values"]), o)
    assert o.to_host().tolist() == c["out"]
    for name, p in (("gather_if", None), ("gather_if_even", even)):
        c = KA[name]
        o = dv(trtc, [c["init"]] * 10)
        trtc.Gather_If(dv(trtc, c["map"]), dv(trtc, c["stencil"]), dv(trtc, c["values"]), o, p)
        assert o.to_host().tolist() == c["out"], name
    c = KA["scatter"]
    o = trtc.device_vector("int32_t", 10)
    trtc.Scatter(dv(trtc, c["values"]), dv(trtc, c["map"]), o)
    assert o.to_host().tolist() == c["out"]
    c = KA["scatter_if"]
    o = trtc.device_vector("int32_t", 8)
    trtc.Scatter_If(dv(trtc, c["values"]), dv(trtc, c["map"]), dv(trtc, c["stencil"]), o)
    assert o.to_host().tolist() == c["out"]
    c = KA["sequence"]
    d = trtc.device_vector("int32_t", 10)
    trtc.Sequence(d)
    assert d.to_host().tolist() == c["out"]
    trtc.Sequence(d, trtc.DVInt32(1))
    assert d.to_host().tolist() == c["out_init1"]
    trtc.Sequence(d, trtc.DVInt32(1), trtc.DVInt32(3))
    assert d.to_host().tolist() == c["out_init1_step3"]
    trtc.Sequence(d)
    trtc.Tabulate(d, trtc.Negate())  # reference test_tabulate: negated sequence
    assert d.to_host().tolist() == [-i for i in range(10)]
    c = KA["adjacent_difference"]
    o = trtc.device_vector("int32_t", 8)
    trtc.Adjacent_Difference(dv(trtc, c["in"]), o)
    assert o.to_host().tolist() == c["minus"]
    trtc.Adjacent_Difference(dv(trtc, c["in"]), o, trtc.Plus())
    assert o.to_host().tolist() == c["plus"]
    d = dv(trtc, c["in"])
    trtc.Adjacent_Difference(d, d)  # in place
    assert d.to_host().tolist() == c["minus"]
    c = KA["swap"]
    a, b = dv(trtc, c["a"]), dv(trtc, c["b"])
    trtc.Swap(a, b)
    assert a.to_host().tolist() == c["b"] and b.to_host().tolist() == c["a"]
    d = trtc.device_vector("int32_t", 5)
    trtc.Fill(d, trtc.DVInt32(-4))
    assert d.to_host().tolist() == [-4] * 5
    acc = trtc.device_vector("int32_t", 1)
    add = trtc.Functor({"acc": acc}, ["x"], "        atomicAdd(&acc[0], x);\n")
    trtc.For_Each(dv(trtc, range(100)), add)
    assert acc.to_host().tolist() == [4950]


def test_known_answers_reductions_searches_merge(trtc):
    c = KA["extrema"]
    d = dv(trtc, c["in"])
    assert trtc.Min_Element(d) == c["min"] and trtc.Max_Element(d) == c["max"]
    assert tuple(trtc.MinMax_Element(d)) == (c["min"], c["max"])
    c = KA["inner_product"]
    r = trtc.Inner_Product(dv(trtc, c["a"], "float"), dv(trtc, c["b"], "float"), trtc.DVFloat(c["init"]))
    assert r == c["out"]
    c = KA["logical"]
    d = trtc.device_vector_from_list(c["in"], "bool")
    assert trtc.All_Of(d, trtc.Identity()) == c["all"]
    assert trtc.Any_Of(d, trtc.Identity()) == c["any"]
    assert trtc.None_Of(d, trtc.Identity()) == c["none"]
    c = KA["find"]
    d = dv(trtc, c["in"])
    assert trtc.Find(d, trtc.DVInt32(3)) == c["find3"] and trtc.Find(d, trtc.DVInt32(5)) == c["find5"]
    assert trtc.Find(d, trtc.DVInt32(9)) is None
    c = KA["mismatch"]
    assert trtc.Mismatch(dv(trtc, c["a"]), dv(trtc, c["b"])) == c["out"]
    assert trtc.Mismatch(dv(trtc, c["a"]), dv(trtc, c["a"])) is None
    c = KA["bounds_v"]
    vi, vv, o = dv(trtc, c["input"]), dv(trtc, c["values"]), trtc.device_vector("int32_t", 6)
    trtc.Lower_Bound_V(vi, vv, o)
    assert o.to_host().tolist() == c["lower"]
    trtc.Upper_Bound_V(vi, vv, o)
    assert o.to_host().tolist() == c["upper"]
    trtc.Binary_Search_V(vi, vv, o)
    assert o.to_host().tolist() == c["found"]
    for q, lo, up, f in zip(c["values"], c["lower"], c["upper"], c["found"]):
        assert trtc.Lower_Bound(vi, trtc.DVInt32(q)) == lo and trtc.Upper_Bound(vi, trtc.DVInt32(q)) == up
        assert trtc.Binary_Search(vi, trtc.DVInt32(q)) == bool(f)
    c = KA["merge"]
    o = trtc.device_vector("int32_t", 13)
    trtc.Merge(dv(trtc, c["a"]), dv(trtc, c["b"]), o)
    assert o.to_host().tolist() == c["out"]
    for name, comp in (("merge_by_key", None), ("merge_by_key_greater", trtc.Greater())):
        c = KA[name]
        ko, vo = trtc.device_vector("int32_t", 13), trtc.device_vector("int32_t", 13)
        trtc.Merge_By_Key(dv(trtc, c["ka"]), dv(trtc, c["kb"]), dv(trtc, c["va"]), dv(trtc, c["vb"]), ko, vo, comp)
        assert ko.to_host().tolist() == c["keys"] and vo.to_host().tolist() == c["values"], name
    c = KA["equal"]
    assert trtc.Equal(dv(trtc, c["a"]), dv(trtc, c["b"])) == c["ab"]
    assert trtc.Equal(dv(trtc, c["a"]), dv(trtc, c["c"])) == c["ac"]
    mod2 = trtc.Functor({}, ["x", "y"], "        return (x % 2) == (y % 2);\n")
    assert trtc.Equal(dv(trtc, c["x"]), dv(trtc, c["y"]), mod2) == c["xy_mod2"]
    srt = dv(trtc, [1, 2, 4, 3, 5])
    assert trtc.Is_Sorted_Until(srt) == 3
    pos = trtc.Functor({}, ["x"], "        return x > 0;\n")
    part = dv(trtc, [3, 1, 2, -1, -5])
    assert trtc.Is_Partitioned(part, pos) is True and trtc.Partition_Point(part, pos) == 3
    assert trtc.Is_Partitioned(dv(trtc, [3, -1, 2]), pos) is False


@pytest.mark.parametrize("n", [8193, 1 << 20])
def test_parity_large(trtc, oa, wl, n):
    say("algorithms parity", n)
    x = wl.i32_range(n, 61, -1000, 1000)
    dx = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector("int32_t", n)
    trtc.Transform(dx, o, trtc.Negate())
    assert (o.to_host() == oa.transform(x, "Negate")).all()
    trtc.Adjacent_Difference(dx, o)
    assert (o.to_host() == oa.adjacent_difference_np(x)).all()
    m = (wl.u32_uniform(n, 62) % np.uint32(n)).astype(np.int32)
    trtc.Gather(trtc.device_vector_from_numpy(m), dx, o)
    assert (o.to_host() == x[m]).all()
    perm = np.argsort(wl.u32_uniform(n, 63), kind="stable").astype(np.int32)
    trtc.Scatter(dx, trtc.device_vector_from_numpy(perm), o)
    want = np.empty_like(x)
    want[perm] = x
    assert (o.to_host() == want).all()
    # first-index searches (a single match planted near the end)
    y = x.copy()
    y[n - 7] = 5000
    dy = trtc.device_vector_from_numpy(y)
    assert trtc.Find(dy, trtc.DVInt32(5000)) == n - 7
    assert trtc.Mismatch(dx, dy) == n - 7 and trtc.Equal(dx, dx) is True
    big = trtc.Functor({}, ["x"], "        return x > 2000;\n")
    assert trtc.Find_If(dy, big) == n - 7 and trtc.Any_Of(dy, big) is True and trtc.None_Of(dx, big) is True
    assert trtc.Find_If_Not(dx, trtc.Functor({}, ["x"], "        return x < 2000;\n")) is None
    # extrema: first occurrence of ties
    assert trtc.Min_Element(dx) == int(np.argmin(x)) and trtc.Max_Element(dx) == int(np.argmax(x))
    gt = trtc.Greater()
    assert trtc.Min_Element(dx, gt) == int(np.argmax(x))  # smallest under Greater = first largest
    f = wl.f32_unit(n, 64)
    g = wl.f32_unit(n, 65)
    r = trtc.Inner_Product(trtc.device_vector_from_numpy(f), trtc.device_vector_from_numpy(g), trtc.DVFloat(1.0))
    ref = 1.0 + float(np.dot(f.astype(np.float64), g.astype(np.float64)))
    assert abs(r - ref) <= 1e-6 * abs(ref) + 1e-3
    # sorted: searches, is_sorted_until, merge
    s = np.sort(x)
    ds = trtc.device_vector_from_numpy(s)
    q = wl.i32_range(4099, 66, -1100, 1100)
    res = trtc.device_vector("int64_t", len(q))
    trtc.Lower_Bound_V(ds, trtc.device_vector_from_numpy(q), res)
    assert (res.to_host() == np.searchsorted(s, q, "left")).all()
    trtc.Upper_Bound_V(ds, trtc.device_vector_from_numpy(q), res)
    assert (res.to_host() == np.searchsorted(s, q, "right")).all()
    assert trtc.Is_Sorted_Until(ds) == n and trtc.Is_Sorted_Until(dx) == oa.is_sorted_until(x)
    s2 = np.sort(wl.i32_range(n // 3 + 5, 67, -1000, 1000))
    va = np.arange(n, dtype=np.int32)
    vb = np.arange(n, n + len(s2), dtype=np.int32)
    ko, vo = trtc.device_vector("int32_t", n + len(s2)), trtc.device_vector("int32_t", n + len(s2))
    trtc.Merge_By_Key(ds, trtc.device_vector_from_numpy(s2), trtc.device_vector_from_numpy(va),
                      trtc.device_vector_from_numpy(vb), ko, vo)
    wk, wv = oa.merge_np(s, s2, False, va, vb)
    assert (ko.to_host() == wk).all() and (vo.to_host() == wv).all()  # payload pins tie order


def test_dvtuple_into_zipped(trtc):
    # reference test_zipped.py: a constant tuple copied into a zipped vector
    ai, af = trtc.device_vector("int32_t", 5), trtc.device_vector("float", 5)
    z = trtc.DVZipped([ai, af], ["a", "b"])
    t = trtc.DVTuple({"a": trtc.DVInt32(123), "b": trtc.DVFloat(456.0)})
    trtc.Copy(trtc.DVConstant(t, 5), z)
    assert ai.to_host().tolist() == [123] * 5 and af.to_host().tolist() == [456.0] * 5
