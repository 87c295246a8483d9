"""ORACLE -- test infrastructure only.

CPU restatement of the reference's remaining algorithms (fynv/ThrustRTC),
used as the checker by tests/.  Nothing in the product (thrustrtc_amd/) may
import or call it.  Explicit loops for the semantics; numpy only for
element-wise arithmetic (with the element type's wrap-around).

Every function cites the reference file it restates.  Parity status: all
functions are pinned by tests/test_oracle_algorithms.py (the reference tests'
inputs with their results, and brute-force loops on random inputs).
"""
import numpy as np

NONE = None


# transform.cpp:4-60 -- out[i] = op(in[i]) (where pred holds, for _If)
def transform(x, unary):
    x = np.asarray(x)
    if unary == "Negate":
        return (np.zeros_like(x) - x).astype(x.dtype)
    if unary == "Square":
        return (x * x).astype(x.dtype)
    return x.copy()


def transform_if(x, unary, mask, out):
    r = np.array(out, copy=True)
    t = transform(x, unary)
    r[mask] = t[mask]
    return r


# replace.cpp:4-40
def replace_copy(x, old, new):
    x = np.asarray(x)
    r = x.copy()
    r[x == old] = new
    return r


def replace_copy_mask(x, mask, new):
    r = np.array(x, copy=True)
    r[mask] = new
    return r


# adjacent_difference.cpp:4-30: out[0] = in[0], out[i] = op(in[i], in[i-1])
def adjacent_difference(x, op="Minus"):
    x = np.asarray(x)
    r = x.copy()
    for i in range(1, len(x)):
        r[i] = x[i] - x[i - 1] if op == "Minus" else x[i] + x[i - 1]
    return r


def adjacent_difference_np(x, op="Minus"):
    """The same as adjacent_difference for large inputs (wrapping arithmetic)."""
    x = np.asarray(x)
    r = x.copy()
    r[1:] = (x[1:] - x[:-1]) if op == "Minus" else (x[1:] + x[:-1])
    return r


# sequence.cpp:4-30: vec[i] = init + i * step in the element type
def sequence(n, dtype, init=0, step=1):
    i = np.arange(n).astype(dtype)
    return (np.array(init, dtype=dtype) + i * np.array(step, dtype=dtype)).astype(dtype)


# gather.cpp / scatter.cpp
def gather(m, x, out=None, mask=None):
    m, x = np.asarray(m), np.asarray(x)
    r = np.array(out, copy=True) if out is not None else np.empty(len(m), dtype=x.dtype)
    for i in range(len(m)):
        if mask is None or mask[i]:
            r[i] = x[m[i]]
    return r


def scatter(x, m, out, mask=None):
    r = np.array(out, copy=True)
    for i in range(len(x)):
        if mask is None or mask[i]:
            r[m[i]] = x[i]
    return r


# find.cpp / mismatch.cpp (general_find.cpp): first index where the
# condition holds, None when none
def find_first(mask):
    for i, b in enumerate(mask):
        if b:
            return i
    return NONE


def find_first_np(mask):
    idx = np.flatnonzero(np.asarray(mask))
    return int(idx[0]) if len(idx) else NONE


# extrema.cpp:5-46: first smallest / first largest (ties keep the earlier)
def min_element(x):
    best = 0
    for i in range(1, len(x)):
        if x[i] < x[best]:
            best = i
    return best


def max_element(x):
    best = 0
    for i in range(1, len(x)):
        if x[best] < x[i]:
            best = i
    return best


# inner_product.cpp:4-20: init + sum a[i] * b[i] (float64 accumulation here)
def inner_product(a, b, init):
    s = float(init)
    for u, v in zip(a, b):
        s += float(u) * float(v)
    return s


# binary_search.cpp (std::lower_bound / upper_bound / binary_search)
def lower_bound(v, value):
    lo, hi = 0, len(v)
    while lo < hi:
        mid = (lo + hi) // 2
        if v[mid] < value:
            lo = mid + 1
        else:
            hi = mid
    return lo


def upper_bound(v, value):
    lo, hi = 0, len(v)
    while lo < hi:
        mid = (lo + hi) // 2
        if not value < v[mid]:
            lo = mid + 1
        else:
            hi = mid
    return lo


def binary_search(v, value):
    i = lower_bound(v, value)
    return i < len(v) and not value < v[i]


# merge.cpp: stable merge, equal keys take the first input first
def merge(a, b, greater=False, va=None, vb=None):
    before = (lambda x, y: x > y) if greater else (lambda x, y: x < y)
    i = j = 0
    keys, vals = [], []
    while i < len(a) or j < len(b):
        take_a = j >= len(b) or (i < len(a) and not before(b[j], a[i]))
        if take_a:
            keys.append(a[i])
            vals.append(None if va is None else va[i])
            i += 1
        else:
            keys.append(b[j])
            vals.append(None if vb is None else vb[j])
            j += 1
    return keys, vals


def merge_np(a, b, greater=False, va=None, vb=None):
    """merge for large inputs: a stable sort of a followed by b (ties keep
    a's items first), the textbook equivalent of std::merge."""
    a, b = np.asarray(a), np.asarray(b)
    k = np.concatenate([a, b])
    order = np.argsort(-k.astype(np.float64) if greater else k, kind="stable")
    if va is None:
        return k[order], None
    return k[order], np.concatenate([np.asarray(va), np.asarray(vb)])[order]


# sort.cpp:85-101 / partition.cpp:132-210
def is_sorted_until(x):
    for i in range(1, len(x)):
        if x[i] < x[i - 1]:
            return i
    return len(x)


def is_partitioned(mask):
    seen_false = False
    for b in mask:
        if not b:
            seen_false = True
        elif seen_false:
            return False
    return True


def partition_point(mask):
    lo, hi = 0, len(mask)
    while lo < hi:
        mid = (lo + hi) // 2
        if mask[mid]:
            lo = mid + 1
        else:
            hi = mid
    return lo
