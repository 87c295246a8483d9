// Device-side built-in types that user code strings may name.
//
// ThrustRTC user code refers to these by name (VectorView<T>, value_t, Pair,
// Plus, Less, ...), so the names and member meanings follow the reference's
// built_in.h:5-174; the definitions here are written for HIP/gfx950.  This
// header is compiled both by hipcc (ahead-of-time kernels) and by hipRTC
// (kernels instantiated at run time for user functors), where it is embedded
// into the library as text.
#pragma once

#if defined(__HIPCC_RTC__)
#include <stdint.h>
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#endif

#define TRTC_DEV __device__ __forceinline__

////////// Elements (built_in.h:5-10) ///////////
template <class T1, class T2>
struct Pair
{
    T1 first;
    T2 second;
};

////////// Functors (built_in.h:14-150) ///////////
// Empty structs: a built-in functor's view is one byte (an empty C++ struct).
struct Identity
{
    template <typename T> TRTC_DEV T operator()(const T& x) const { return x; }
};
struct Maximum
{
    template <typename T> TRTC_DEV T operator()(const T& x, const T& y) const { return x < y ? y : x; }
};
struct Minimum
{
    template <typename T> TRTC_DEV T operator()(const T& x, const T& y) const { return x > y ? y : x; }
};
struct EqualTo
{
    template <typename T> TRTC_DEV bool operator()(const T& x, const T& y) const { return x == y; }
};
struct NotEqualTo
{
    template <typename T> TRTC_DEV bool operator()(const T& x, const T& y) const { return x != y; }
};
struct Greater
{
    template <typename T> TRTC_DEV bool operator()(const T& x, const T& y) const { return x > y; }
};
struct Less
{
    template <typename T> TRTC_DEV bool operator()(const T& x, const T& y) const { return x < y; }
};
struct GreaterEqual
{
    template <typename T> TRTC_DEV bool operator()(const T& x, const T& y) const { return x >= y; }
};
struct LessEqual
{
    template <typename T> TRTC_DEV bool operator()(const T& x, const T& y) const { return x <= y; }
};
struct Plus
{
    template <typename T> TRTC_DEV T operator()(const T& x, const T& y) const { return x + y; }
};
struct Minus
{
    template <typename T> TRTC_DEV T operator()(const T& x, const T& y) const { return x - y; }
};
struct Multiplies
{
    template <typename T> TRTC_DEV T operator()(const T& x, const T& y) const { return x * y; }
};
struct Divides
{
    template <typename T> TRTC_DEV T operator()(const T& x, const T& y) const { return x / y; }
};
struct Modulus
{
    template <typename T> TRTC_DEV T operator()(const T& x, const T& y) const { return x % y; }
};
struct Negate
{
    template <typename T> TRTC_DEV T operator()(const T& x) const { return -x; }
};

////////// Vectors (built_in.h:154-174) ///////////
// Host layout: { T* _data; size_t _size; } — 16 bytes, 8-byte aligned.
template <class _T>
struct VectorView
{
    typedef _T value_t;
    typedef _T& ref_t;

    value_t* _data;
    size_t _size;

    TRTC_DEV size_t size() const { return _size; }
    TRTC_DEV ref_t operator[](size_t idx) const { return _data[idx]; }
};

// A sub-range of another view (DVRange).  Host layout: the inner view, then
// two size_t.
template <class _TVVALUE>
struct RangeView
{
    typedef typename _TVVALUE::value_t value_t;
    typedef typename _TVVALUE::ref_t ref_t;
    mutable _TVVALUE _view_vec_value;
    size_t _begin;
    size_t _end;

    TRTC_DEV size_t size() const { return _end - _begin; }
    TRTC_DEV ref_t operator[](size_t idx) const { return _view_vec_value[idx + _begin]; }
};

// Stencil argument of the stencil-less selections (Copy_If, Partition, ...):
// one byte, never read, so the predicate is only instantiated on the input.
struct _NoStencil
{
    typedef unsigned char value_t;
    TRTC_DEV value_t operator[](size_t) const { return 0; }
};
template <class S> struct _IsNoStencil { static constexpr bool value = false; };
template <> struct _IsNoStencil<_NoStencil> { static constexpr bool value = true; };

////////// Fake vectors (fake_vectors/*, built_in.h:196-334) ///////////
// Views computed from an index instead of read from memory.  Member names are
// the ones the reference's host classes fill in; host layouts follow C rules.
// operator[] is const (kernels may hold views by const reference); nested
// views and functors are mutable because their own operators need not be.

// Every element is the same value (DVConstant).  { size_t; T; }
template <class _T>
struct ConstantView
{
    typedef _T value_t;
    typedef const _T& ref_t;
    size_t _size;
    value_t _value;

    TRTC_DEV size_t size() const { return _size; }
    TRTC_DEV ref_t operator[](size_t) const { return _value; }
};

// Element i is init + i (DVCounter).  { size_t; T; }
template <class _T>
struct CounterView
{
    typedef _T value_t;
    typedef _T ref_t;
    size_t _size;
    value_t _value_init;

    TRTC_DEV size_t size() const { return _size; }
    TRTC_DEV ref_t operator[](size_t idx) const { return _value_init + (value_t)idx; }
};

// Assignments to a discard vector's elements go nowhere (DVDiscard).
template <class _T>
struct _Sink
{
    TRTC_DEV const _Sink& operator=(const _T&) const { return *this; }
};

// { size_t; _Sink<T>; } -- 16 bytes
template <class _T>
struct DiscardView
{
    typedef _T value_t;
    typedef const _Sink<_T>& ref_t;
    size_t _size;
    _Sink<_T> _sink;

    TRTC_DEV size_t size() const { return _size; }
    TRTC_DEV ref_t operator[](size_t) const { return _sink; }
};

// Element i is value[index[i]] (DVPermutation).  { VV; VI; }
template <class _TVVALUE, class _TVINDEX>
struct PermutationView
{
    typedef typename _TVVALUE::value_t value_t;
    typedef typename _TVVALUE::ref_t ref_t;
    mutable _TVVALUE _view_vec_value;
    mutable _TVINDEX _view_vec_index;

    TRTC_DEV size_t size() const { return _view_vec_index.size(); }
    TRTC_DEV ref_t operator[](size_t idx) const { return _view_vec_value[(size_t)_view_vec_index[idx]]; }
};

// Element i is value[size - 1 - i] (DVReverse).  { VV; }
template <class _TVVALUE>
struct ReverseView
{
    typedef typename _TVVALUE::value_t value_t;
    typedef typename _TVVALUE::ref_t ref_t;
    mutable _TVVALUE _view_vec_value;

    TRTC_DEV size_t size() const { return _view_vec_value.size(); }
    TRTC_DEV ref_t operator[](size_t idx) const { return _view_vec_value[_view_vec_value.size() - 1 - idx]; }
};

// Element i is op(in[i]) converted to T (DVTransform).  { VIN; OP; }
template <class _T, class _T_VIN, class _T_OP>
struct TransformView
{
    typedef _T value_t;
    typedef _T ref_t;
    mutable _T_VIN _view_vec_in;
    mutable _T_OP _view_op;  // user functors have a non-const operator()

    TRTC_DEV size_t size() const { return _view_vec_in.size(); }
    TRTC_DEV ref_t operator[](size_t idx) const { return (value_t)_view_op(_view_vec_in[idx]); }
};

////////// Binary searches over sorted views (built_in.h binary-search routines) ///////////
// The position of the first element of `vec` that is not before `value`
// (comp(vec[i], value) false), std::lower_bound.
template <class TVec, class T, class TComp>
TRTC_DEV size_t d_lower_bound(TVec& vec, const T& value, TComp& comp)
{
    size_t lo = 0, hi = vec.size();
    while (lo < hi) {
        size_t mid = (lo + hi) >> 1;
        if (comp(vec[mid], value)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The position of the first element that is after `value` (comp(value,
// vec[i]) true), std::upper_bound.
template <class TVec, class T, class TComp>
TRTC_DEV size_t d_upper_bound(TVec& vec, const T& value, TComp& comp)
{
    size_t lo = 0, hi = vec.size();
    while (lo < hi) {
        size_t mid = (lo + hi) >> 1;
        if (!comp(value, vec[mid])) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Whether an element equivalent to `value` is present, std::binary_search.
template <class TVec, class T, class TComp>
TRTC_DEV bool d_binary_search(TVec& vec, const T& value, TComp& comp)
{
    size_t i = d_lower_bound(vec, value, comp);
    return i < vec.size() && !comp(value, vec[i]);
}
