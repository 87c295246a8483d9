/* Host-visible built-in types of the device code (reference: built_in.h:5-10,
 * 154-334, installed beside sort.h and the other API headers,
 * cpp/CMakeLists.txt:73,179).
 *
 * Host programs use these to lay out element data they hand to the library,
 * e.g. DVVector("Pair<int, int>", n, host_pairs) with Pair<int, int> host
 * arrays (cpp/test/test_extrema.cpp), or to read views back.  Each struct
 * has the member names, order and therefore the byte layout of the device
 * type with the same name in the library's device header trtc_builtin.hpp,
 * which user code strings name when they are compiled for the GPU; only the
 * data members are declared here.  The functors (Plus, Less, ...) and the
 * device-side binary searches exist only in device code: on the host they are
 * named through Functor("Plus") and friends (functor.h).
 *
 * With DEVICE_ONLY defined (code compiled for the GPU) nothing is declared:
 * the device definitions come from trtc_builtin.hpp. */
#ifndef _built_in_h
#define _built_in_h

#ifndef DEVICE_ONLY

#include <cstddef>

////////// Elements ///////////
template <class T1, class T2>
struct Pair
{
    T1 first;
    T2 second;
};

////////// Vectors ///////////
// DVVector / DVVectorAdaptor: { T* _data; size_t _size; }
template <class _T>
struct VectorView
{
    typedef _T value_t;
    typedef _T& ref_t;
    value_t* _data;
    size_t _size;
};

// DVRange: the inner view, then the half-open index range.
template <class _TVVALUE>
struct RangeView
{
    typedef typename _TVVALUE::value_t value_t;
    typedef typename _TVVALUE::ref_t ref_t;
    _TVVALUE _view_vec_value;
    size_t _begin;
    size_t _end;
};

////////// Fake vectors ///////////
// DVConstant: every element is _value.
template <class _T>
struct ConstantView
{
    typedef _T value_t;
    typedef const _T& ref_t;
    size_t _size;
    value_t _value;
};

// DVCounter: element i is _value_init + i.
template <class _T>
struct CounterView
{
    typedef _T value_t;
    typedef _T ref_t;
    size_t _size;
    value_t _value_init;
};

// DVDiscard: assignments go nowhere.
template <class _T>
struct _Sink
{
};

template <class _T>
struct DiscardView
{
    typedef _T value_t;
    typedef _Sink<_T>& ref_t;
    size_t _size;
    _Sink<_T> _sink;
};

// DVPermutation: element i is value[index[i]].
template <class _TVVALUE, class _TVINDEX>
struct PermutationView
{
    typedef typename _TVVALUE::value_t value_t;
    typedef typename _TVVALUE::ref_t ref_t;
    _TVVALUE _view_vec_value;
    _TVINDEX _view_vec_index;
};

// DVReverse: element i is value[size - 1 - i].
template <class _TVVALUE>
struct ReverseView
{
    typedef typename _TVVALUE::value_t value_t;
    typedef typename _TVVALUE::ref_t ref_t;
    _TVVALUE _view_vec_value;
};

// DVTransform: element i is op(in[i]) converted to _T.
template <class _T, class _T_VIN, class _T_OP>
struct TransformView
{
    typedef _T value_t;
    typedef _T ref_t;
    _T_VIN _view_vec_in;
    _T_OP _view_op;
};

#endif  // DEVICE_ONLY

#endif
