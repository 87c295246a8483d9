"""GPU parity of the scan / compaction consumers (Reduce_By_Key, Unique*,
Partition*, Remove*, Transform_*_Scan, Count*, Transform_Reduce) and of the
fake vectors (DVConstant, DVCounter, DVDiscard, DVPermutation, DVReverse,
DVTransform, DVZipped, DVCustomVector) against the oracle, through the
reference-shaped Python API over libPyThrustRTC.so.

Integer results are compared bit-exactly, including the elements past a
returned count (they must be untouched); float segment reductions within a
relative 1e-5 of the oracle's sequential float result (the GPU associates a
segment's additions differently).
"""
import json
import os

import numpy as np
import pytest

from helpers import Oracle, workloads

pytestmark = pytest.mark.gpu

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "consumer_known_answers.json")))
SIZES = [1, 7, 64, 1025, 8191, 8192, 8193, 100003, 1 << 20]  # 7, 64, 1025: waves past n
SENT = -7777


@pytest.fixture(scope="module")
def trtc():
    import thrustrtc_amd
    return thrustrtc_amd


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def wl():
    return workloads()


def say(*a):
    print(*a, flush=True)


def pred_functor(trtc, p):
    if p[0] == "mod":
        return trtc.Functor({}, ["x"], f"        return x % {p[1]} == {p[2]};\n")
    if p[0] == "nonzero":
        return trtc.Identity()
    raise ValueError(p)


def runs_keys(wl, n, seed, max_run):
    """Keys with runs of equal values (run lengths 1..max_run)."""
    sizes, _ = wl.segment_sizes(n, seed, 1, max_run)
    k = int(np.searchsorted(np.cumsum(sizes.astype(np.int64)), n)) + 1  # enough runs to cover n
    keys = np.repeat(np.arange(k, dtype=np.int64), sizes[:k].astype(np.int64))[:n]
    # neighbouring runs may repeat a value after a gap, like real data
    return ((keys * 7919) % 1000 - 500).astype(np.int32)


# ---------------------------------------------------------------- known answers
def test_consumer_known_answers(trtc):
    for c in KA["unique"]:
        d = trtc.device_vector_from_list(c["in"], "int32_t")
        n = trtc.Unique(d)
        assert d.to_host(0, n).tolist() == c["out"]
        o = trtc.device_vector("int32_t", len(c["in"]))
        n = trtc.Unique_Copy(trtc.device_vector_from_list(c["in"], "int32_t"), o, trtc.EqualTo())
        assert o.to_host(0, n).tolist() == c["out"]
    for c in KA["unique_by_key"]:
        k = trtc.device_vector_from_list(c["keys"], "int32_t")
        v = trtc.device_vector_from_list(c["values"], "int32_t")
        n = trtc.Unique_By_Key(k, v)
        assert k.to_host(0, n).tolist() == c["keys_out"] and v.to_host(0, n).tolist() == c["values_out"]
    for c in KA["reduce_by_key"]:
        k = trtc.device_vector_from_list(c["keys"], "int32_t")
        v = trtc.device_vector_from_list(c["values"], "int32_t")
        ko = trtc.device_vector("int32_t", 7)
        vo = trtc.device_vector("int32_t", 7)
        n = trtc.Reduce_By_Key(k, v, ko, vo)
        assert ko.to_host(0, n).tolist() == c["keys_out"] and vo.to_host(0, n).tolist() == c["values_out"]
    for c in KA["partition"]:
        d = trtc.device_vector_from_list(c["in"], "int32_t")
        p = pred_functor(trtc, c["pred"])
        if "stencil" in c:
            n = trtc.Partition_Stencil(d, trtc.device_vector_from_list(c["stencil"], "int32_t"), p)
        else:
            n = trtc.Partition(d, p)
        assert n == c["count"] and d.to_host().tolist() == c["out"]
    for c in KA["partition_copy"]:
        d = trtc.device_vector_from_list(c["in"], "int32_t")
        t = trtc.device_vector("int32_t", 10)
        f = trtc.device_vector("int32_t", 10)
        p = pred_functor(trtc, c["pred"])
        if "stencil" in c:
            n = trtc.Partition_Copy_Stencil(d, trtc.device_vector_from_list(c["stencil"], "int32_t"), t, f, p)
        else:
            n = trtc.Partition_Copy(d, t, f, p)
        assert t.to_host(0, n).tolist() == c["true"] and f.to_host(0, 10 - n).tolist() == c["false"]
    for c in KA["remove"]:
        d = trtc.device_vector_from_list(c["in"], "int32_t")
        o = trtc.device_vector("int32_t", len(c["in"]))
        st = trtc.device_vector_from_list(c["stencil"], "int32_t") if "stencil" in c else None
        p = pred_functor(trtc, c["pred"]) if "pred" in c else None
        fn = c["fn"]
        if fn == "Remove":
            n, res = trtc.Remove(d, trtc.DVInt32(c["value"])), d
        elif fn == "Remove_Copy":
            n, res = trtc.Remove_Copy(d, o, trtc.DVInt32(c["value"])), o
        elif fn == "Remove_If":
            n, res = trtc.Remove_If(d, p), d
        elif fn == "Remove_Copy_If":
            n, res = trtc.Remove_Copy_If(d, o, p), o
        elif fn == "Remove_If_Stencil":
            n, res = trtc.Remove_If_Stencil(d, st, p), d
        else:
            n, res = trtc.Remove_Copy_If_Stencil(d, st, o, p), o
        assert res.to_host(0, n).tolist() == c["out"], fn
    for c in KA["transform_scan"]:
        d = trtc.device_vector_from_list(c["in"], "int32_t")
        if c["exclusive"]:
            trtc.Transform_Exclusive_Scan(d, d, trtc.Negate(), trtc.DVInt32(c["init"]), trtc.Plus())
        else:
            trtc.Transform_Inclusive_Scan(d, d, trtc.Negate(), trtc.Plus())
        assert d.to_host().tolist() == c["out"]
    x = trtc.device_vector_from_numpy(np.arange(2000, dtype=np.int32) % 100)
    assert trtc.Count(x, trtc.DVInt32(47)) == 20
    assert trtc.Count_If(x, trtc.Functor({}, ["x"], "        return x == 47;\n")) == 20
    c = KA["transform_reduce"][0]
    absval = trtc.Functor({}, ["x"], "        return x < (decltype(x))0 ? -x : x;\n")
    r = trtc.Transform_Reduce(trtc.device_vector_from_list(c["in"], "int32_t"), absval, trtc.DVInt32(0),
                              trtc.Maximum())
    assert r == c["out"]


# ---------------------------------------------------------------- unique
@pytest.mark.parametrize("n", SIZES)
def test_unique_parity(trtc, orc, wl, n):
    say("unique", n)
    keys = runs_keys(wl, n, 21, 40)
    vals = wl.i32_range(n, 22, -1000, 1000)
    (want,) = orc.unique(keys, "int32_t")
    # copy form: output past the count untouched
    o = trtc.device_vector_from_numpy(np.full(n, SENT, dtype=np.int32))
    c = trtc.Unique_Copy(trtc.device_vector_from_numpy(keys), o)
    got = o.to_host()
    assert c == len(want) and (got[:c] == want).all() and (got[c:] == SENT).all()
    # in place: the tail keeps the original items
    d = trtc.device_vector_from_numpy(keys)
    c = trtc.Unique(d)
    got = d.to_host()
    assert c == len(want) and (got[:c] == want).all() and (got[c:] == keys[c:]).all()
    # by key, in place and copy
    wk, wv = orc.unique(keys, "int32_t", vals)
    dk, dv = trtc.device_vector_from_numpy(keys), trtc.device_vector_from_numpy(vals)
    c = trtc.Unique_By_Key(dk, dv)
    assert c == len(wk) and (dk.to_host(0, c) == wk).all() and (dv.to_host(0, c) == wv).all()
    ko, vo = trtc.device_vector("int32_t", n), trtc.device_vector("int32_t", n)
    c = trtc.Unique_By_Key_Copy(trtc.device_vector_from_numpy(keys), trtc.device_vector_from_numpy(vals), ko, vo)
    assert c == len(wk) and (ko.to_host(0, c) == wk).all() and (vo.to_host(0, c) == wv).all()


def test_unique_custom_predicate_and_types(trtc, orc, wl):
    n = 300007
    keys = np.sort(wl.i32_range(n, 23, -5000, 5000))
    pred = trtc.Functor({}, ["x", "y"], "        return x / 10 == y / 10;\n")
    (want,) = orc.unique(keys, "int32_t", div=10)
    d = trtc.device_vector_from_numpy(keys)
    c = trtc.Unique(d, pred)
    assert c == len(want) and (d.to_host(0, c) == want).all()
    # 64-bit keys with 1-byte values; float keys (-0.0 == +0.0 is a duplicate)
    k64 = runs_keys(wl, n, 24, 9).astype(np.int64) * 3000000000
    v8 = (np.arange(n) % 251).astype(np.int8)
    wk, wv = orc.unique(k64, "int64_t", v8)
    dk, dv = trtc.device_vector_from_numpy(k64), trtc.device_vector_from_numpy(v8)
    c = trtc.Unique_By_Key(dk, dv)
    assert c == len(wk) and (dk.to_host(0, c) == wk).all() and (dv.to_host(0, c) == wv).all()
    kf = np.array([0.0, -0.0, 0.0, 1.5, 1.5, -2.0, np.nan, np.nan, 3.0], dtype=np.float32)
    (wf,) = orc.unique(kf, "float")
    d = trtc.device_vector_from_numpy(kf)
    c = trtc.Unique(d)
    assert c == len(wf) and d.to_host(0, c).tobytes() == wf.tobytes()


# ---------------------------------------------------------------- reduce_by_key
@pytest.mark.parametrize("n", SIZES)
def test_reduce_by_key_parity(trtc, orc, wl, n):
    say("reduce_by_key", n)
    keys = runs_keys(wl, n, 25, 3000)
    vals = wl.i32_range(n, 26, -1000000000, 1000000000)
    wk, wv = orc.reduce_by_key(keys, vals, "int32_t", "int32_t")
    ko = trtc.device_vector_from_numpy(np.full(n, SENT, dtype=np.int32))
    vo = trtc.device_vector_from_numpy(np.full(n, SENT, dtype=np.int32))
    c = trtc.Reduce_By_Key(trtc.device_vector_from_numpy(keys), trtc.device_vector_from_numpy(vals), ko, vo)
    gk, gv = ko.to_host(), vo.to_host()
    assert c == len(wk) and (gk[:c] == wk).all() and (gv[:c] == wv).all()
    assert (gk[c:] == SENT).all() and (gv[c:] == SENT).all()
    # Maximum over the same runs
    wk, wm = orc.reduce_by_key(keys, vals, "int32_t", "int32_t", op="Maximum")
    c = trtc.Reduce_By_Key(trtc.device_vector_from_numpy(keys), trtc.device_vector_from_numpy(vals), ko, vo,
                           trtc.EqualTo(), trtc.Maximum())
    assert c == len(wk) and (ko.to_host(0, c) == wk).all() and (vo.to_host(0, c) == wm).all()


def test_reduce_by_key_extremes(trtc, orc, wl):
    n = (1 << 20) + 3
    vals = wl.i32_range(n, 27, -5, 5)
    for name, keys in [("one run", np.zeros(n, dtype=np.int32)),
                       ("all distinct", np.arange(n, dtype=np.int32)),
                       ("pairs", (np.arange(n) // 2).astype(np.int32))]:
        say("reduce_by_key", name)
        wk, wv = orc.reduce_by_key(keys, vals, "int32_t", "int32_t")
        ko, vo = trtc.device_vector("int32_t", n), trtc.device_vector("int32_t", n)
        c = trtc.Reduce_By_Key(trtc.device_vector_from_numpy(keys), trtc.device_vector_from_numpy(vals), ko, vo)
        assert c == len(wk) and (ko.to_host(0, c) == wk).all() and (vo.to_host(0, c) == wv).all(), name


def test_reduce_by_key_custom_pred_tail_key_and_float(trtc, orc, wl):
    n = 500009
    keys = np.sort(wl.i32_range(n, 28, -100000, 100000))
    vals = wl.f32_unit(n, 29)
    pred = trtc.Functor({}, ["x", "y"], "        return x / 10 == y / 10;\n")
    wk, wv = orc.reduce_by_key(keys, vals, "int32_t", "float", div=10)
    ko, vo = trtc.device_vector("int32_t", n), trtc.device_vector("float", n)
    c = trtc.Reduce_By_Key(trtc.device_vector_from_numpy(keys), trtc.device_vector_from_numpy(vals), ko, vo, pred)
    assert c == len(wk)
    assert (ko.to_host(0, c) == wk).all()  # the run's tail key
    gv = vo.to_host(0, c)
    assert np.allclose(gv, wv, rtol=1e-5, atol=0)
    # values converted to a wider output type (double) before reduction
    vd = trtc.device_vector("double", n)
    c = trtc.Reduce_By_Key(trtc.device_vector_from_numpy(keys), trtc.device_vector_from_numpy(vals), ko, vd, pred)
    wk, wvd = orc.reduce_by_key(keys, vals.astype(np.float64), "int32_t", "double", div=10)
    assert c == len(wk) and np.allclose(vd.to_host(0, c), wvd, rtol=1e-12, atol=0)


def test_reduce_by_key_custom_vectors_discard(trtc):
    # reference test_custom_vector.py: keys idx / 2000 and values "divisible
    # by group + 1" over ten groups of 0..1999, keys discarded
    darr = trtc.device_vector_from_numpy(np.arange(2000, dtype=np.int32))
    vals = trtc.DVCustomVector({"src": darr}, "idx", """
    unsigned group = idx / src.size();
    unsigned sub_idx = idx % src.size();
    return src[sub_idx] % (group+1) ==0 ? 1: 0;
""", "uint32_t", 20000)
    keys = trtc.DVCustomVector({"src": darr}, "idx", "        return idx / src.size();\n", "uint32_t", 20000)
    vo = trtc.device_vector("uint32_t", 10)
    c = trtc.Reduce_By_Key(keys, vals, trtc.DVDiscard("uint32_t", 10), vo)
    assert c == 10
    assert vo.to_host().tolist() == [1999 // (g + 1) + 1 for g in range(10)]


# ---------------------------------------------------------------- remove / partition
@pytest.mark.parametrize("n", SIZES)
def test_remove_family_parity(trtc, orc, wl, n):
    say("remove", n)
    x = wl.i32_range(n, 30, -20, 20)
    st = wl.i32_range(n, 31, 0, 2)
    even = trtc.Functor({}, ["x"], "        return x % 2 == 0;\n")
    cases = [
        ("Remove", lambda d, o: trtc.Remove(d, trtc.DVInt32(3)), orc.select(x, "int32_t", ("eq", 3), negate=True), True),
        ("Remove_Copy", lambda d, o: trtc.Remove_Copy(d, o, trtc.DVInt32(3)),
         orc.select(x, "int32_t", ("eq", 3), negate=True), False),
        ("Remove_If", lambda d, o: trtc.Remove_If(d, even), orc.select(x, "int32_t", ("mod", 2, 0), negate=True), True),
        ("Remove_Copy_If", lambda d, o: trtc.Remove_Copy_If(d, o, even),
         orc.select(x, "int32_t", ("mod", 2, 0), negate=True), False),
        ("Remove_If_Stencil", lambda d, o: trtc.Remove_If_Stencil(d, trtc.device_vector_from_numpy(st), trtc.Identity()),
         orc.select(x, "int32_t", ("nonzero",), stencil=st, negate=True), True),
        ("Remove_Copy_If_Stencil",
         lambda d, o: trtc.Remove_Copy_If_Stencil(d, trtc.device_vector_from_numpy(st), o, trtc.Identity()),
         orc.select(x, "int32_t", ("nonzero",), stencil=st, negate=True), False),
    ]
    for name, fn, want, inplace in cases:
        d = trtc.device_vector_from_numpy(x)
        o = trtc.device_vector_from_numpy(np.full(n, SENT, dtype=np.int32))
        c = fn(d, o)
        got = (d if inplace else o).to_host()
        assert c == len(want) and (got[:c] == want).all(), name
        assert (got[c:] == (x[c:] if inplace else SENT)).all(), name


@pytest.mark.parametrize("n", SIZES)
def test_partition_family_parity(trtc, orc, wl, n):
    say("partition", n)
    x = wl.i32_range(n, 32, -1000, 1000)
    st = wl.i32_range(n, 33, 0, 3)
    p3 = trtc.Functor({}, ["x"], "        return x % 3 == 1;\n")
    wt, wf = orc.partition(x, "int32_t", ("mod", 3, 1))
    d = trtc.device_vector_from_numpy(x)
    c = trtc.Partition(d, p3)
    assert c == len(wt) and d.to_host().tobytes() == np.concatenate([wt, wf]).tobytes()
    wt2, wf2 = orc.partition(x, "int32_t", ("nonzero",), stencil=st)
    d = trtc.device_vector_from_numpy(x)
    c = trtc.Partition_Stencil(d, trtc.device_vector_from_numpy(st), trtc.Identity())
    assert c == len(wt2) and d.to_host().tobytes() == np.concatenate([wt2, wf2]).tobytes()
    t = trtc.device_vector_from_numpy(np.full(n, SENT, dtype=np.int32))
    f = trtc.device_vector_from_numpy(np.full(n, SENT, dtype=np.int32))
    c = trtc.Partition_Copy(trtc.device_vector_from_numpy(x), t, f, p3)
    gt, gf = t.to_host(), f.to_host()
    assert c == len(wt) and (gt[:c] == wt).all() and (gf[:n - c] == wf).all()
    assert (gt[c:] == SENT).all() and (gf[n - c:] == SENT).all()
    c = trtc.Partition_Copy_Stencil(trtc.device_vector_from_numpy(x), trtc.device_vector_from_numpy(st), t, f,
                                    trtc.Identity())
    assert c == len(wt2) and (t.to_host(0, c) == wt2).all() and (f.to_host(0, n - c) == wf2).all()


def test_stencil_predicate_sees_only_the_stencil_type(trtc, wl):
    # float items, an integer-only predicate on an int32 stencil
    n = 100003
    x = wl.f32_unit(n, 34)
    st = wl.i32_range(n, 35, 0, 9)
    odd = trtc.Functor({}, ["s"], "        return s % 2 == 1;\n")
    o = trtc.device_vector("float", n)
    c = trtc.Copy_If_Stencil(trtc.device_vector_from_numpy(x), trtc.device_vector_from_numpy(st), o, odd)
    assert c == int(np.count_nonzero(st % 2 == 1)) and o.to_host(0, c).tobytes() == x[st % 2 == 1].tobytes()
    t, f = trtc.device_vector("float", n), trtc.device_vector("float", n)
    c = trtc.Partition_Copy_Stencil(trtc.device_vector_from_numpy(x), trtc.device_vector_from_numpy(st), t, f, odd)
    assert t.to_host(0, c).tobytes() == x[st % 2 == 1].tobytes()
    assert f.to_host(0, n - c).tobytes() == x[st % 2 == 0].tobytes()


# ---------------------------------------------------------------- transform scans, counts
@pytest.mark.parametrize("n", [1, 8193, 1 << 20])
def test_transform_scans_counts(trtc, orc, wl, n):
    say("transform scans", n)
    x = wl.i32_range(n, 36, -1000, 1000)
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector("int32_t", n)
    trtc.Transform_Inclusive_Scan(d, o, trtc.Negate(), trtc.Plus())
    assert (o.to_host() == orc.transform_scan(x, "int32_t", "Negate")).all()
    trtc.Transform_Exclusive_Scan(d, o, trtc.Negate(), trtc.DVInt32(4), trtc.Plus())
    assert (o.to_host() == orc.transform_scan(x, "int32_t", "Negate", exclusive=True, init=4)).all()
    sq = trtc.Functor({}, ["x"], "        return (int64_t)x * x;\n")
    o64 = trtc.device_vector("int64_t", n)
    trtc.Transform_Inclusive_Scan(d, o64, sq, trtc.Plus())
    assert (o64.to_host() == orc.transform_scan(x.astype(np.int64), "int64_t", "Square")).all()
    m7 = trtc.Functor({}, ["x"], "        return x % 7 == 0;\n")
    assert trtc.Count_If(d, m7) == orc.count(x, "int32_t", ("mod", 7, 0))
    assert trtc.Count(d, trtc.DVInt32(-3)) == orc.count(x, "int32_t", ("eq", -3))
    r = trtc.Transform_Reduce(d, trtc.Negate(), trtc.DVInt64(5), trtc.Plus())
    assert r == 5 - int(x.astype(np.int64).sum())


# ---------------------------------------------------------------- fake vectors
def test_fake_vectors(trtc, wl):
    n = 100003
    # DVCounter / DVConstant (docs: DVCounter(DVInt32(1), 10) is 1..10)
    o = trtc.device_vector("int32_t", 10)
    trtc.Copy(trtc.DVCounter(trtc.DVInt32(1), 10), o)
    assert o.to_host().tolist() == list(range(1, 11))
    assert trtc.Reduce(trtc.DVConstant(trtc.DVInt32(5), n)) == 5 * n
    o = trtc.device_vector("int64_t", n)
    trtc.Inclusive_Scan(trtc.DVCounter(trtc.DVInt64(0), n), o)
    assert (o.to_host() == np.cumsum(np.arange(n, dtype=np.int64))).all()
    # DVPermutation / DVReverse, read and written
    vals = wl.f32_unit(1000, 37)
    idx = wl.i32_range(n, 38, 0, 999)
    dvals, didx = trtc.device_vector_from_numpy(vals), trtc.device_vector_from_numpy(idx)
    o = trtc.device_vector("float", n)
    trtc.Copy(trtc.DVPermutation(dvals, didx), o)
    assert o.to_host().tobytes() == vals[idx].tobytes()
    x = wl.i32_range(n, 39, -100, 100)
    d = trtc.device_vector_from_numpy(x)
    o = trtc.device_vector("int32_t", n)
    trtc.Inclusive_Scan(trtc.DVReverse(d), o)
    assert (o.to_host() == np.cumsum(x[::-1].astype(np.int64)).astype(np.int32)).all()
    o2 = trtc.device_vector("int32_t", n)
    trtc.Inclusive_Scan(d, trtc.DVReverse(o2))
    assert (o2.to_host()[::-1] == np.cumsum(x.astype(np.int64)).astype(np.int32)).all()
    # DVTransform (reference test_transform_iter.py: square roots)
    sq = trtc.Functor({}, ["x"], "        return sqrtf(x);\n")
    src = trtc.device_vector_from_list([1.0, 4.0, 9.0, 16.0], "float")
    o = trtc.device_vector("float", 4)
    trtc.Copy(trtc.DVTransform(src, "float", sq), o)
    assert o.to_host().tolist() == [1.0, 2.0, 3.0, 4.0]
    # DVDiscard as a compaction target: the count only
    even = trtc.Functor({}, ["x"], "        return x % 2 == 0;\n")
    assert trtc.Copy_If(d, trtc.DVDiscard("int32_t", n), even) == int(np.count_nonzero(x % 2 == 0))
    # a range of a computed vector
    assert trtc.Reduce(trtc.DVRange(trtc.DVCounter(trtc.DVInt64(0), n), 10, 20)) == sum(range(10, 20))


def test_zipped_copy_and_sort(trtc, wl):
    # reference test_zipped.py: copy between zipped vectors
    ai = trtc.device_vector_from_list([0, 1, 2, 3, 4], "int32_t")
    af = trtc.device_vector_from_list([0.0, 10.0, 20.0, 30.0, 40.0], "float")
    bi, bf = trtc.device_vector("int32_t", 5), trtc.device_vector("float", 5)
    trtc.Copy(trtc.DVZipped([ai, af], ["a", "b"]), trtc.DVZipped([bi, bf], ["a", "b"]))
    assert bi.to_host().tolist() == [0, 1, 2, 3, 4] and bf.to_host().tolist() == [0.0, 10.0, 20.0, 30.0, 40.0]
    # sort of zipped keys by (a, then b descending): the merge path through
    # staged contiguous copies, written back through the zipped view
    n = 50021
    a = wl.i32_range(n, 40, 0, 50)
    b = wl.f32_unit(n, 41)
    da, db = trtc.device_vector_from_numpy(a), trtc.device_vector_from_numpy(b)
    z = trtc.DVZipped([da, db], ["a", "b"])
    comp = trtc.Functor({}, ["x", "y"], "        return x.a < y.a || (x.a == y.a && x.b > y.b);\n")
    trtc.Sort(z, comp)
    order = np.lexsort((-b.astype(np.float64), a))
    assert (da.to_host() == a[order]).all() and db.to_host().tobytes() == b[order].tobytes()
