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
    _TVVALUE _view_vec_value;
    size_t _begin;
    size_t _end;

    TRTC_DEV size_t size() const { return _end - _begin; }
    TRTC_DEV ref_t operator[](size_t idx) const { return _view_vec_value[idx + _begin]; }
};
