// C ABI (include/trtc_capi.h): thin wrappers over the C++ API with the
// reference's conventions (python/api_*.cpp): handles are heap objects, a NULL
// functor / init selects the default overload, ints return 0 / -1 / -2.
#include <algorithm>
#include <cstring>

#include "copy.h"
#include "DVTuple.h"
#include "adjacent_difference.h"
#include "binary_search.h"
#include "equal.h"
#include "extrema.h"
#include "fill.h"
#include "find.h"
#include "for_each.h"
#include "gather.h"
#include "inner_product.h"
#include "logical.h"
#include "merge.h"
#include "mismatch.h"
#include "replace.h"
#include "scatter.h"
#include "sequence.h"
#include "swap.h"
#include "tabulate.h"
#include "transform.h"
#include "count.h"
#include "partition.h"
#include "remove.h"
#include "transform_reduce.h"
#include "transform_scan.h"
#include "unique.h"
#include "fake_vectors/DVConstant.h"
#include "fake_vectors/DVCounter.h"
#include "fake_vectors/DVCustomVector.h"
#include "fake_vectors/DVDiscard.h"
#include "fake_vectors/DVPermutation.h"
#include "fake_vectors/DVReverse.h"
#include "fake_vectors/DVTransform.h"
#include "fake_vectors/DVZipped.h"
#include "reduce.h"
#include "scan.h"
#include "sort.h"
#include "batch.h"
#include "server.h"
#include "trtc_capi.h"
#include "trtc_internal.h"

using namespace trtc_host;

typedef std::vector<std::string> StrArray;
typedef std::vector<const DeviceViewable*> PtrArray;

// ---- utils
void* n_string_array_create(unsigned long long size, const char* const* strs)
{
    StrArray* r = new StrArray(size);
    for (size_t i = 0; i < size; i++) (*r)[i] = strs[i];
    return r;
}
void n_string_array_destroy(void* p) { delete (StrArray*)p; }
void* n_pointer_array_create(unsigned long long size, const void* const* ptrs)
{
    PtrArray* r = new PtrArray(size);
    if (size) memcpy(r->data(), ptrs, sizeof(void*) * size);
    return r;
}
void n_pointer_array_destroy(void* p) { delete (PtrArray*)p; }
void* n_dim3_create(unsigned x, unsigned y, unsigned z) { return new dim_type{x, y, z}; }
void n_dim3_destroy(void* p) { delete (dim_type*)p; }

// ---- context
void n_set_libnvrtc_path(const char* path) { set_libnvrtc_path(path); }
int n_trtc_try_init(void) { return TRTC_Try_Init() ? 0 : -100; }
int n_get_ptx_arch(void) { return TRTC_Get_PTX_Arch(); }
void n_set_verbose(unsigned v) { TRTC_Set_Verbose(v != 0); }
void n_set_kernel_debug(unsigned v) { TRTC_Set_Kernel_Debug(v != 0); }
void n_add_include_dir(const char* dir) { TRTC_Add_Include_Dir(dir); }
void n_add_built_in_header(const char* fn, const char* content) { TRTC_Add_Built_In_Header(fn, content); }
void n_add_inlcude_filename(const char* fn) { TRTC_Add_Inlcude_Filename(fn); }
void n_add_code_block(const char* line) { TRTC_Add_Code_Block(line); }
void n_add_constant_object(const char* name, void* cptr) { TRTC_Add_Constant_Object(name, *(DeviceViewable*)cptr); }
void n_wait(void) { TRTC_Wait(); }

static std::vector<const char*> cstrs(const StrArray& a)
{
    std::vector<const char*> r(a.size());
    for (size_t i = 0; i < a.size(); i++) r[i] = a[i].c_str();
    return r;
}

void* n_kernel_create(void* ptr_param_list, const char* body)
{
    return new TRTC_Kernel(cstrs(*(StrArray*)ptr_param_list), body);
}
void n_kernel_destroy(void* cptr) { delete (TRTC_Kernel*)cptr; }
int n_kernel_num_params(void* cptr) { return (int)((TRTC_Kernel*)cptr)->num_params(); }

static bool arg_count_ok(size_t want, const PtrArray* args)
{
    if (args->size() == want) return true;
    printf("Wrong number of arguments received. %d required, %d received.", (int)want, (int)args->size());
    return false;
}

int n_kernel_calc_optimal_block_size(void* ptr_kernel, void* ptr_arg_list, unsigned sharedMemBytes)
{
    TRTC_Kernel* k = (TRTC_Kernel*)ptr_kernel;
    PtrArray* a = (PtrArray*)ptr_arg_list;
    if (!arg_count_ok(k->num_params(), a)) return -2;
    int s;
    if (k->calc_optimal_block_size(a->data(), s, sharedMemBytes)) return s;
    return -1;
}

int n_kernel_calc_number_blocks(void* ptr_kernel, void* ptr_arg_list, int sizeBlock, unsigned sharedMemBytes)
{
    TRTC_Kernel* k = (TRTC_Kernel*)ptr_kernel;
    PtrArray* a = (PtrArray*)ptr_arg_list;
    if (!arg_count_ok(k->num_params(), a)) return -2;
    int nb;
    if (k->calc_number_blocks(a->data(), sizeBlock, nb, sharedMemBytes)) return nb;
    return -1;
}

int n_kernel_launch(void* ptr_kernel, void* ptr_gridDim, void* ptr_blockDim, void* ptr_arg_list, int sharedMemBytes)
{
    TRTC_Kernel* k = (TRTC_Kernel*)ptr_kernel;
    PtrArray* a = (PtrArray*)ptr_arg_list;
    if (!arg_count_ok(k->num_params(), a)) return -2;
    return k->launch(*(dim_type*)ptr_gridDim, *(dim_type*)ptr_blockDim, a->data(), (unsigned)sharedMemBytes) ? 0 : -1;
}

void* n_for_create(void* ptr_param_list, const char* name_iter, const char* body)
{
    return new TRTC_For(cstrs(*(StrArray*)ptr_param_list), name_iter, body);
}
void n_for_destroy(void* cptr) { delete (TRTC_For*)cptr; }
int n_for_num_params(void* cptr) { return (int)((TRTC_For*)cptr)->num_params(); }
int n_for_launch(void* cptr, int begin, int end, void* ptr_arg_list)
{
    TRTC_For* f = (TRTC_For*)cptr;
    PtrArray* a = (PtrArray*)ptr_arg_list;
    if (!arg_count_ok(f->num_params(), a)) return -2;
    return f->launch((size_t)begin, (size_t)end, a->data()) ? 0 : -1;
}
int n_for_launch_n(void* cptr, int n, void* ptr_arg_list)
{
    TRTC_For* f = (TRTC_For*)cptr;
    PtrArray* a = (PtrArray*)ptr_arg_list;
    if (!arg_count_ok(f->num_params(), a)) return -2;
    return f->launch_n((size_t)n, a->data()) ? 0 : -1;
}

// ---- device-viewable scalars
const char* n_dv_name_view_cls(void* cptr) { return ((DeviceViewable*)cptr)->name_view_cls().c_str(); }
void n_dv_destroy(void* cptr) { delete (DeviceViewable*)cptr; }

template <class T> static T dv_value(void* cptr)
{
    ViewBuf b = ((DeviceViewable*)cptr)->view();
    T v;
    memcpy(&v, b.data(), sizeof(T));
    return v;
}

void* n_dvint8_create(int v) { return new DVInt8((int8_t)v); }
int n_dvint8_value(void* c) { return dv_value<int8_t>(c); }
void* n_dvuint8_create(unsigned v) { return new DVUInt8((uint8_t)v); }
unsigned n_dvuint8_value(void* c) { return dv_value<uint8_t>(c); }
void* n_dvint16_create(int v) { return new DVInt16((int16_t)v); }
int n_dvint16_value(void* c) { return dv_value<int16_t>(c); }
void* n_dvuint16_create(unsigned v) { return new DVUInt16((uint16_t)v); }
unsigned n_dvuint16_value(void* c) { return dv_value<uint16_t>(c); }
void* n_dvint32_create(int v) { return new DVInt32(v); }
int n_dvint32_value(void* c) { return dv_value<int32_t>(c); }
void* n_dvuint32_create(unsigned v) { return new DVUInt32(v); }
unsigned n_dvuint32_value(void* c) { return dv_value<uint32_t>(c); }
void* n_dvint64_create(long long v) { return new DVInt64(v); }
long long n_dvint64_value(void* c) { return dv_value<int64_t>(c); }
void* n_dvuint64_create(unsigned long long v) { return new DVUInt64(v); }
unsigned long long n_dvuint64_value(void* c) { return dv_value<uint64_t>(c); }
void* n_dvfloat_create(float v) { return new DVFloat(v); }
float n_dvfloat_value(void* c) { return dv_value<float>(c); }
void* n_dvdouble_create(double v) { return new DVDouble(v); }
double n_dvdouble_value(void* c) { return dv_value<double>(c); }
void* n_dvbool_create(int v) { return new DVBool(v != 0); }
int n_dvbool_value(void* c) { return dv_value<bool>(c) ? 1 : 0; }
void* n_dvcomplex64_create(float re, float im) { return new DVComplex64(re, im); }
float n_dvcomplex64_real(void* c) { return dv_value<float>(c); }
float n_dvcomplex64_imag(void* c)
{
    ViewBuf b = ((DeviceViewable*)c)->view();
    float v;
    memcpy(&v, b.data() + 4, 4);
    return v;
}
void* n_dvcomplex128_create(double re, double im) { return new DVComplex128(re, im); }
double n_dvcomplex128_real(void* c) { return dv_value<double>(c); }
double n_dvcomplex128_imag(void* c)
{
    ViewBuf b = ((DeviceViewable*)c)->view();
    double v;
    memcpy(&v, b.data() + 8, 8);
    return v;
}

// ---- vectors
const char* n_dvvectorlike_name_elem_cls(void* cptr) { return ((DVVectorLike*)cptr)->name_elem_cls().c_str(); }
unsigned long long n_dvvectorlike_size(void* cptr) { return ((DVVectorLike*)cptr)->size(); }

void* n_dvvector_create(const char* elem_cls, unsigned long long size, void* hdata)
{
    DVVector* v = new DVVector(elem_cls, size, hdata);
    if (size > 0 && !v->data()) {
        delete v;
        return nullptr;
    }
    return v;
}

void n_dvvector_to_host(void* cptr, void* hdata, unsigned long long begin, unsigned long long end)
{
    ((DVVector*)cptr)->to_host(hdata, begin, end);
}

void* n_dvvector_from_dvs(void* ptr_dvs)
{
    PtrArray* dvs = (PtrArray*)ptr_dvs;
    size_t n = dvs->size();
    if (n < 1) return nullptr;
    std::string elem = (*dvs)[0]->name_view_cls();
    for (size_t i = 1; i < n; i++)
        if ((*dvs)[i]->name_view_cls() != elem) return nullptr;
    size_t es = (*dvs)[0]->view().size();
    std::vector<char> buf(es * n);
    for (size_t i = 0; i < n; i++) memcpy(buf.data() + es * i, (*dvs)[i]->view().data(), es);
    return n_dvvector_create(elem.c_str(), n, buf.data());
}

void* n_dvvectoradaptor_create(const char* elem_cls, unsigned long long size, void* data)
{
    return new DVVectorAdaptor(elem_cls, size, data);
}

void* n_dvrange_create(void* ptr_in, unsigned long long begin, unsigned long long end)
{
    DVVectorLike* v = (DVVectorLike*)ptr_in;
    if (auto p = dynamic_cast<DVVector*>(v)) return new DVVectorAdaptor(*p, begin, end);
    if (auto p = dynamic_cast<DVVectorAdaptor*>(v)) return new DVVectorAdaptor(*p, begin, end);
    return new DVRange(*v, begin, end);
}

// ---- fake vectors
void* n_dvconstant_create(void* cptr, unsigned long long size) { return new DVConstant(*(DeviceViewable*)cptr, size); }
void* n_dvcounter_create(void* cptr, unsigned long long size) { return new DVCounter(*(DeviceViewable*)cptr, size); }
void* n_dvdiscard_create(const char* elem_cls, unsigned long long size) { return new DVDiscard(elem_cls, size); }
void* n_dvpermutation_create(void* ptr_value, void* ptr_index)
{
    return new DVPermutation(*(DVVectorLike*)ptr_value, *(DVVectorLike*)ptr_index);
}
void* n_dvreverse_create(void* ptr_value) { return new DVReverse(*(DVVectorLike*)ptr_value); }
void* n_dvtransform_create(void* ptr_in, const char* elem_cls, void* ptr_op)
{
    return new DVTransform(*(DVVectorLike*)ptr_in, elem_cls, *(Functor*)ptr_op);
}
void* n_dvzipped_create(void* ptr_vecs, void* ptr_elem_names)
{
    const PtrArray& v = *(PtrArray*)ptr_vecs;
    const StrArray& names = *(StrArray*)ptr_elem_names;
    if (v.empty() || v.size() != names.size()) return nullptr;
    std::vector<DVVectorLike*> vecs;
    for (auto* p : v) vecs.push_back((DVVectorLike*)p);
    return new DVZipped(vecs, cstrs(names));
}
void* n_dvcustomvector_create(void* ptr_dvs, void* ptr_names, const char* name_idx, const char* body,
                              const char* elem_cls, unsigned long long size, unsigned read_only)
{
    const PtrArray& dvs = *(PtrArray*)ptr_dvs;
    const StrArray& names = *(StrArray*)ptr_names;
    if (dvs.size() != names.size()) return nullptr;
    std::vector<CapturedDeviceViewable> arg_map(dvs.size());
    for (size_t i = 0; i < dvs.size(); i++) arg_map[i] = {names[i].c_str(), dvs[i]};
    return new DVCustomVector(arg_map, name_idx, body, elem_cls, size, read_only != 0);
}

void* n_dvvectorlike_data(void* cptr) { return contiguous_data(*(DVVectorLike*)cptr); }

// ---- functors
void* n_functor_create(void* ptr_dvs, void* ptr_names, void* ptr_functor_params, const char* code_body)
{
    PtrArray* dvs = (PtrArray*)ptr_dvs;
    StrArray* names = (StrArray*)ptr_names;
    std::vector<CapturedDeviceViewable> arg_map(dvs->size());
    for (size_t i = 0; i < dvs->size(); i++) {
        arg_map[i].obj_name = (*names)[i].c_str();
        arg_map[i].obj = (*dvs)[i];
    }
    return new Functor(arg_map, cstrs(*(StrArray*)ptr_functor_params), code_body);
}

void* n_built_in_functor_create(const char* name) { return new Functor(name); }

// ---- algorithms
int n_fill(void* ptr_vec, void* ptr_value)
{
    return TRTC_Fill(*(DVVectorLike*)ptr_vec, *(DeviceViewable*)ptr_value) ? 0 : -1;
}

int n_copy(void* ptr_in, void* ptr_out)
{
    return TRTC_Copy(*(DVVectorLike*)ptr_in, *(DVVectorLike*)ptr_out) ? 0 : -1;
}

void* n_reduce(void* ptr_vec, void* ptr_init, void* ptr_bin_op)
{
    DVVectorLike* vec = (DVVectorLike*)ptr_vec;
    ViewBuf ret;
    bool ok;
    if (!ptr_init) ok = TRTC_Reduce(*vec, ret);
    else if (!ptr_bin_op) ok = TRTC_Reduce(*vec, *(DeviceViewable*)ptr_init, ret);
    else ok = TRTC_Reduce(*vec, *(DeviceViewable*)ptr_init, *(Functor*)ptr_bin_op, ret);
    if (!ok) return nullptr;
    return dv_from_viewbuf(ret, vec->name_elem_cls().c_str());
}

int n_is_sorted(void* ptr_vec, void* ptr_comp)
{
    bool res;
    bool ok = ptr_comp ? TRTC_Is_Sorted(*(DVVectorLike*)ptr_vec, *(Functor*)ptr_comp, res)
                       : TRTC_Is_Sorted(*(DVVectorLike*)ptr_vec, res);
    if (!ok) return -1;
    return res ? 1 : 0;
}

int n_inclusive_scan(void* ptr_vec_in, void* ptr_vec_out, void* ptr_binary_op)
{
    DVVectorLike* in = (DVVectorLike*)ptr_vec_in;
    DVVectorLike* out = (DVVectorLike*)ptr_vec_out;
    bool ok = ptr_binary_op ? TRTC_Inclusive_Scan(*in, *out, *(Functor*)ptr_binary_op) : TRTC_Inclusive_Scan(*in, *out);
    return ok ? 0 : -1;
}

int n_exclusive_scan(void* ptr_vec_in, void* ptr_vec_out, void* ptr_init, void* ptr_binary_op)
{
    DVVectorLike* in = (DVVectorLike*)ptr_vec_in;
    DVVectorLike* out = (DVVectorLike*)ptr_vec_out;
    bool ok;
    if (!ptr_init) ok = TRTC_Exclusive_Scan(*in, *out);
    else if (!ptr_binary_op) ok = TRTC_Exclusive_Scan(*in, *out, *(DeviceViewable*)ptr_init);
    else ok = TRTC_Exclusive_Scan(*in, *out, *(DeviceViewable*)ptr_init, *(Functor*)ptr_binary_op);
    return ok ? 0 : -1;
}

int n_inclusive_scan_by_key(void* ptr_vec_key, void* ptr_vec_value, void* ptr_vec_out, void* ptr_binary_pred,
                            void* ptr_binary_op)
{
    DVVectorLike* k = (DVVectorLike*)ptr_vec_key;
    DVVectorLike* v = (DVVectorLike*)ptr_vec_value;
    DVVectorLike* o = (DVVectorLike*)ptr_vec_out;
    bool ok;
    if (!ptr_binary_pred) ok = TRTC_Inclusive_Scan_By_Key(*k, *v, *o);
    else if (!ptr_binary_op) ok = TRTC_Inclusive_Scan_By_Key(*k, *v, *o, *(Functor*)ptr_binary_pred);
    else ok = TRTC_Inclusive_Scan_By_Key(*k, *v, *o, *(Functor*)ptr_binary_pred, *(Functor*)ptr_binary_op);
    return ok ? 0 : -1;
}

int n_exclusive_scan_by_key(void* ptr_vec_key, void* ptr_vec_value, void* ptr_vec_out, void* ptr_init,
                            void* ptr_binary_pred, void* ptr_binary_op)
{
    DVVectorLike* k = (DVVectorLike*)ptr_vec_key;
    DVVectorLike* v = (DVVectorLike*)ptr_vec_value;
    DVVectorLike* o = (DVVectorLike*)ptr_vec_out;
    bool ok;
    if (!ptr_init) ok = TRTC_Exclusive_Scan_By_Key(*k, *v, *o);
    else if (!ptr_binary_pred) ok = TRTC_Exclusive_Scan_By_Key(*k, *v, *o, *(DeviceViewable*)ptr_init);
    else if (!ptr_binary_op)
        ok = TRTC_Exclusive_Scan_By_Key(*k, *v, *o, *(DeviceViewable*)ptr_init, *(Functor*)ptr_binary_pred);
    else
        ok = TRTC_Exclusive_Scan_By_Key(*k, *v, *o, *(DeviceViewable*)ptr_init, *(Functor*)ptr_binary_pred,
                                        *(Functor*)ptr_binary_op);
    return ok ? 0 : -1;
}

unsigned n_copy_if(void* ptr_in, void* ptr_out, void* ptr_pred)
{
    return TRTC_Copy_If(*(DVVectorLike*)ptr_in, *(DVVectorLike*)ptr_out, *(Functor*)ptr_pred);
}

unsigned n_copy_if_stencil(void* ptr_in, void* ptr_stencil, void* ptr_out, void* ptr_pred)
{
    return TRTC_Copy_If_Stencil(*(DVVectorLike*)ptr_in, *(DVVectorLike*)ptr_stencil, *(DVVectorLike*)ptr_out,
                                *(Functor*)ptr_pred);
}

// ---- scan / compaction consumers (reference python/api_Reductions.cpp,
// api_PrefixSums.cpp, api_Reordering.cpp): a NULL functor selects the
// default overload
typedef DVVectorLike DVL;

unsigned long long n_count(void* ptr_vec, void* ptr_value)
{
    size_t r;
    return TRTC_Count(*(DVL*)ptr_vec, *(DeviceViewable*)ptr_value, r) ? (unsigned long long)r : ~0ull;
}

unsigned long long n_count_if(void* ptr_vec, void* ptr_pred)
{
    size_t r;
    return TRTC_Count_If(*(DVL*)ptr_vec, *(Functor*)ptr_pred, r) ? (unsigned long long)r : ~0ull;
}

unsigned n_reduce_by_key(void* ptr_key_in, void* ptr_value_in, void* ptr_key_out, void* ptr_value_out,
                         void* ptr_binary_pred, void* ptr_binary_op)
{
    DVL& ki = *(DVL*)ptr_key_in;
    DVL& vi = *(DVL*)ptr_value_in;
    DVL& ko = *(DVL*)ptr_key_out;
    DVL& vo = *(DVL*)ptr_value_out;
    if (!ptr_binary_pred) return TRTC_Reduce_By_Key(ki, vi, ko, vo);
    if (!ptr_binary_op) return TRTC_Reduce_By_Key(ki, vi, ko, vo, *(Functor*)ptr_binary_pred);
    return TRTC_Reduce_By_Key(ki, vi, ko, vo, *(Functor*)ptr_binary_pred, *(Functor*)ptr_binary_op);
}

void* n_transform_reduce(void* ptr_vec, void* ptr_unary_op, void* ptr_init, void* ptr_binary_op)
{
    DeviceViewable* init = (DeviceViewable*)ptr_init;
    ViewBuf ret;
    if (!TRTC_Transform_Reduce(*(DVL*)ptr_vec, *(Functor*)ptr_unary_op, *init, *(Functor*)ptr_binary_op, ret))
        return nullptr;
    return dv_from_viewbuf(ret, init->name_view_cls().c_str());
}

int n_transform_inclusive_scan(void* ptr_vec_in, void* ptr_vec_out, void* ptr_unary_op, void* ptr_binary_op)
{
    return TRTC_Transform_Inclusive_Scan(*(DVL*)ptr_vec_in, *(DVL*)ptr_vec_out, *(Functor*)ptr_unary_op,
                                         *(Functor*)ptr_binary_op)
               ? 0
               : -1;
}

int n_transform_exclusive_scan(void* ptr_vec_in, void* ptr_vec_out, void* ptr_unary_op, void* ptr_init,
                               void* ptr_binary_op)
{
    return TRTC_Transform_Exclusive_Scan(*(DVL*)ptr_vec_in, *(DVL*)ptr_vec_out, *(Functor*)ptr_unary_op,
                                         *(DeviceViewable*)ptr_init, *(Functor*)ptr_binary_op)
               ? 0
               : -1;
}

unsigned n_remove(void* ptr_vec, void* ptr_value) { return TRTC_Remove(*(DVL*)ptr_vec, *(DeviceViewable*)ptr_value); }
unsigned n_remove_copy(void* ptr_in, void* ptr_out, void* ptr_value)
{
    return TRTC_Remove_Copy(*(DVL*)ptr_in, *(DVL*)ptr_out, *(DeviceViewable*)ptr_value);
}
unsigned n_remove_if(void* ptr_vec, void* ptr_pred) { return TRTC_Remove_If(*(DVL*)ptr_vec, *(Functor*)ptr_pred); }
unsigned n_remove_copy_if(void* ptr_in, void* ptr_out, void* ptr_pred)
{
    return TRTC_Remove_Copy_If(*(DVL*)ptr_in, *(DVL*)ptr_out, *(Functor*)ptr_pred);
}
unsigned n_remove_if_stencil(void* ptr_vec, void* ptr_stencil, void* ptr_pred)
{
    return TRTC_Remove_If_Stencil(*(DVL*)ptr_vec, *(DVL*)ptr_stencil, *(Functor*)ptr_pred);
}
unsigned n_remove_copy_if_stencil(void* ptr_in, void* ptr_stencil, void* ptr_out, void* ptr_pred)
{
    return TRTC_Remove_Copy_If_Stencil(*(DVL*)ptr_in, *(DVL*)ptr_stencil, *(DVL*)ptr_out, *(Functor*)ptr_pred);
}

unsigned n_unique(void* ptr_vec, void* ptr_binary_pred)
{
    DVL& v = *(DVL*)ptr_vec;
    return ptr_binary_pred ? TRTC_Unique(v, *(Functor*)ptr_binary_pred) : TRTC_Unique(v);
}
unsigned n_unique_copy(void* ptr_vec_in, void* ptr_vec_out, void* ptr_binary_pred)
{
    DVL& i = *(DVL*)ptr_vec_in;
    DVL& o = *(DVL*)ptr_vec_out;
    return ptr_binary_pred ? TRTC_Unique_Copy(i, o, *(Functor*)ptr_binary_pred) : TRTC_Unique_Copy(i, o);
}
unsigned n_unique_by_key(void* ptr_keys, void* ptr_values, void* ptr_binary_pred)
{
    DVL& k = *(DVL*)ptr_keys;
    DVL& v = *(DVL*)ptr_values;
    return ptr_binary_pred ? TRTC_Unique_By_Key(k, v, *(Functor*)ptr_binary_pred) : TRTC_Unique_By_Key(k, v);
}
unsigned n_unique_by_key_copy(void* ptr_keys_in, void* ptr_values_in, void* ptr_key_out, void* ptr_values_out,
                              void* ptr_binary_pred)
{
    DVL& ki = *(DVL*)ptr_keys_in;
    DVL& vi = *(DVL*)ptr_values_in;
    DVL& ko = *(DVL*)ptr_key_out;
    DVL& vo = *(DVL*)ptr_values_out;
    return ptr_binary_pred ? TRTC_Unique_By_Key_Copy(ki, vi, ko, vo, *(Functor*)ptr_binary_pred)
                           : TRTC_Unique_By_Key_Copy(ki, vi, ko, vo);
}

unsigned n_partition(void* ptr_vec, void* ptr_pred) { return TRTC_Partition(*(DVL*)ptr_vec, *(Functor*)ptr_pred); }
unsigned n_partition_stencil(void* ptr_vec, void* ptr_stencil, void* ptr_pred)
{
    return TRTC_Partition_Stencil(*(DVL*)ptr_vec, *(DVL*)ptr_stencil, *(Functor*)ptr_pred);
}
unsigned n_partition_copy(void* ptr_vec_in, void* ptr_vec_true, void* ptr_vec_false, void* ptr_pred)
{
    return TRTC_Partition_Copy(*(DVL*)ptr_vec_in, *(DVL*)ptr_vec_true, *(DVL*)ptr_vec_false, *(Functor*)ptr_pred);
}
unsigned n_partition_copy_stencil(void* ptr_vec_in, void* ptr_stencil, void* ptr_vec_true, void* ptr_vec_false,
                                  void* ptr_pred)
{
    return TRTC_Partition_Copy_Stencil(*(DVL*)ptr_vec_in, *(DVL*)ptr_stencil, *(DVL*)ptr_vec_true,
                                       *(DVL*)ptr_vec_false, *(Functor*)ptr_pred);
}

// ---- the other algorithms (reference python/api_Transformations.cpp,
// api_Copying.cpp, api_Reductions.cpp, api_Searching.cpp, api_Merging.cpp,
// api_DVTuple.cpp)
static inline int ok_i(bool b) { return b ? 0 : -1; }
static inline int bool_i(bool ok, bool v) { return ok ? (v ? 1 : 0) : -1; }
#define F(p) (*(Functor*)(p))
#define V(p) (*(DVL*)(p))
#define D(p) (*(DeviceViewable*)(p))

void* n_dvtuple_create(void* ptr_dvs, void* ptr_names)
{
    const PtrArray& dvs = *(PtrArray*)ptr_dvs;
    const StrArray& names = *(StrArray*)ptr_names;
    if (dvs.size() != names.size()) return nullptr;
    std::vector<CapturedDeviceViewable> m(dvs.size());
    for (size_t i = 0; i < dvs.size(); i++) m[i] = {names[i].c_str(), dvs[i]};
    return new DVTuple(m);
}
int n_replace(void* v, void* o, void* nv) { return ok_i(TRTC_Replace(V(v), D(o), D(nv))); }
int n_replace_if(void* v, void* p, void* nv) { return ok_i(TRTC_Replace_If(V(v), F(p), D(nv))); }
int n_replace_copy(void* i, void* o, void* ov, void* nv) { return ok_i(TRTC_Replace_Copy(V(i), V(o), D(ov), D(nv))); }
int n_replace_copy_if(void* i, void* o, void* p, void* nv) { return ok_i(TRTC_Replace_Copy_If(V(i), V(o), F(p), D(nv))); }
int n_for_each(void* v, void* f) { return ok_i(TRTC_For_Each(V(v), F(f))); }
int n_adjacent_difference(void* i, void* o, void* op)
{
    return ok_i(op ? TRTC_Adjacent_Difference(V(i), V(o), F(op)) : TRTC_Adjacent_Difference(V(i), V(o)));
}
int n_sequence(void* v, void* init, void* step)
{
    if (!init) return ok_i(TRTC_Sequence(V(v)));
    if (!step) return ok_i(TRTC_Sequence(V(v), D(init)));
    return ok_i(TRTC_Sequence(V(v), D(init), D(step)));
}
int n_tabulate(void* v, void* op) { return ok_i(TRTC_Tabulate(V(v), F(op))); }
int n_transform(void* i, void* o, void* op) { return ok_i(TRTC_Transform(V(i), V(o), F(op))); }
int n_transform_binary(void* i1, void* i2, void* o, void* op) { return ok_i(TRTC_Transform_Binary(V(i1), V(i2), V(o), F(op))); }
int n_transform_if(void* i, void* o, void* op, void* p) { return ok_i(TRTC_Transform_If(V(i), V(o), F(op), F(p))); }
int n_transform_if_stencil(void* i, void* s, void* o, void* op, void* p)
{
    return ok_i(TRTC_Transform_If_Stencil(V(i), V(s), V(o), F(op), F(p)));
}
int n_transform_binary_if_stencil(void* i1, void* i2, void* s, void* o, void* op, void* p)
{
    return ok_i(TRTC_Transform_Binary_If_Stencil(V(i1), V(i2), V(s), V(o), F(op), F(p)));
}
int n_gather(void* m, void* i, void* o) { return ok_i(TRTC_Gather(V(m), V(i), V(o))); }
int n_gather_if(void* m, void* s, void* i, void* o, void* p)
{
    return ok_i(p ? TRTC_Gather_If(V(m), V(s), V(i), V(o), F(p)) : TRTC_Gather_If(V(m), V(s), V(i), V(o)));
}
int n_scatter(void* i, void* m, void* o) { return ok_i(TRTC_Scatter(V(i), V(m), V(o))); }
int n_scatter_if(void* i, void* m, void* s, void* o, void* p)
{
    return ok_i(p ? TRTC_Scatter_If(V(i), V(m), V(s), V(o), F(p)) : TRTC_Scatter_If(V(i), V(m), V(s), V(o)));
}
int n_swap(void* a, void* b) { return ok_i(TRTC_Swap(V(a), V(b))); }
int n_equal(void* a, void* b, void* p)
{
    bool r = false;
    bool ok = p ? TRTC_Equal(V(a), V(b), F(p), r) : TRTC_Equal(V(a), V(b), r);
    return bool_i(ok, r);
}
unsigned long long n_min_element(void* v, void* c)
{
    size_t r;
    return (c ? TRTC_Min_Element(V(v), F(c), r) : TRTC_Min_Element(V(v), r)) ? r : ~0ull;
}
unsigned long long n_max_element(void* v, void* c)
{
    size_t r;
    return (c ? TRTC_Max_Element(V(v), F(c), r) : TRTC_Max_Element(V(v), r)) ? r : ~0ull;
}
int n_minmax_element(void* v, void* c, unsigned long long* ret)
{
    size_t a, b;
    bool ok = c ? TRTC_MinMax_Element(V(v), F(c), a, b) : TRTC_MinMax_Element(V(v), a, b);
    if (!ok) return -1;
    ret[0] = a;
    ret[1] = b;
    return 0;
}
void* n_inner_product(void* a, void* b, void* init, void* op1, void* op2)
{
    ViewBuf ret;
    bool ok = (op1 && op2) ? TRTC_Inner_Product(V(a), V(b), D(init), ret, F(op1), F(op2))
                           : TRTC_Inner_Product(V(a), V(b), D(init), ret);
    if (!ok) return nullptr;
    return dv_from_viewbuf(ret, D(init).name_view_cls().c_str());
}
int n_all_of(void* v, void* p)
{
    bool r = false;
    bool ok = TRTC_All_Of(V(v), F(p), r);
    return bool_i(ok, r);
}
int n_any_of(void* v, void* p)
{
    bool r = false;
    bool ok = TRTC_Any_Of(V(v), F(p), r);
    return bool_i(ok, r);
}
int n_none_of(void* v, void* p)
{
    bool r = false;
    bool ok = TRTC_None_Of(V(v), F(p), r);
    return bool_i(ok, r);
}
int n_is_partitioned(void* v, void* p)
{
    bool r = false;
    bool ok = TRTC_Is_Partitioned(V(v), F(p), r);
    return bool_i(ok, r);
}
// not found: the vector's size (api_Searching.cpp:9-20)
static long long found_ll(bool ok, size_t r, size_t n) { return !ok ? -1 : (r == (size_t)-1 ? (long long)n : (long long)r); }
long long n_find(void* v, void* val)
{
    size_t r = 0;
    bool ok = TRTC_Find(V(v), D(val), r);
    return found_ll(ok, r, V(v).size());
}
long long n_find_if(void* v, void* p)
{
    size_t r = 0;
    bool ok = TRTC_Find_If(V(v), F(p), r);
    return found_ll(ok, r, V(v).size());
}
long long n_find_if_not(void* v, void* p)
{
    size_t r = 0;
    bool ok = TRTC_Find_If_Not(V(v), F(p), r);
    return found_ll(ok, r, V(v).size());
}
long long n_mismatch(void* a, void* b, void* p)
{
    size_t r = 0;
    bool ok = p ? TRTC_Mismatch(V(a), V(b), F(p), r) : TRTC_Mismatch(V(a), V(b), r);
    return found_ll(ok, r, V(a).size());
}
unsigned long long n_lower_bound(void* v, void* val, void* c)
{
    size_t r;
    return (c ? TRTC_Lower_Bound(V(v), D(val), F(c), r) : TRTC_Lower_Bound(V(v), D(val), r)) ? r : ~0ull;
}
unsigned long long n_upper_bound(void* v, void* val, void* c)
{
    size_t r;
    return (c ? TRTC_Upper_Bound(V(v), D(val), F(c), r) : TRTC_Upper_Bound(V(v), D(val), r)) ? r : ~0ull;
}
int n_binary_search(void* v, void* val, void* c)
{
    bool r = false;
    bool ok = c ? TRTC_Binary_Search(V(v), D(val), F(c), r) : TRTC_Binary_Search(V(v), D(val), r);
    return bool_i(ok, r);
}
int n_lower_bound_v(void* v, void* vals, void* res, void* c)
{
    return ok_i(c ? TRTC_Lower_Bound_V(V(v), V(vals), V(res), F(c)) : TRTC_Lower_Bound_V(V(v), V(vals), V(res)));
}
int n_upper_bound_v(void* v, void* vals, void* res, void* c)
{
    return ok_i(c ? TRTC_Upper_Bound_V(V(v), V(vals), V(res), F(c)) : TRTC_Upper_Bound_V(V(v), V(vals), V(res)));
}
int n_binary_search_v(void* v, void* vals, void* res, void* c)
{
    return ok_i(c ? TRTC_Binary_Search_V(V(v), V(vals), V(res), F(c)) : TRTC_Binary_Search_V(V(v), V(vals), V(res)));
}
unsigned long long n_partition_point(void* v, void* p)
{
    size_t r;
    return TRTC_Partition_Point(V(v), F(p), r) ? r : ~0ull;
}
unsigned long long n_is_sorted_until(void* v, void* c)
{
    size_t r;
    return (c ? TRTC_Is_Sorted_Until(V(v), F(c), r) : TRTC_Is_Sorted_Until(V(v), r)) ? r : ~0ull;
}
int n_merge(void* a, void* b, void* o, void* c)
{
    return ok_i(c ? TRTC_Merge(V(a), V(b), V(o), F(c)) : TRTC_Merge(V(a), V(b), V(o)));
}
int n_merge_by_key(void* k1, void* k2, void* v1, void* v2, void* ko, void* vo, void* c)
{
    return ok_i(c ? TRTC_Merge_By_Key(V(k1), V(k2), V(v1), V(v2), V(ko), V(vo), F(c))
                  : TRTC_Merge_By_Key(V(k1), V(k2), V(v1), V(v2), V(ko), V(vo)));
}
#undef F
#undef V
#undef D

int n_sort(void* ptr_vec, void* ptr_comp)
{
    DVVectorLike* v = (DVVectorLike*)ptr_vec;
    bool ok = ptr_comp ? TRTC_Sort(*v, *(Functor*)ptr_comp) : TRTC_Sort(*v);
    return ok ? 0 : -1;
}

int n_sort_by_key(void* ptr_keys, void* ptr_values, void* ptr_comp)
{
    DVVectorLike* k = (DVVectorLike*)ptr_keys;
    DVVectorLike* v = (DVVectorLike*)ptr_values;
    bool ok = ptr_comp ? TRTC_Sort_By_Key(*k, *v, *(Functor*)ptr_comp) : TRTC_Sort_By_Key(*k, *v);
    return ok ? 0 : -1;
}

// ---- extensions
int n_segmented_sort_by_key(void* ptr_keys, void* ptr_values, const unsigned long long* offsets,
                            unsigned long long num_segments, void* ptr_comp)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    Functor less("Less");
    const Functor& comp = ptr_comp ? *(Functor*)ptr_comp : less;
    return TRTC_Segmented_Sort_By_Key(*(DVVectorLike*)ptr_keys, *(DVVectorLike*)ptr_values, off.data(), num_segments,
                                      comp)
               ? 0
               : -1;
}

int n_segmented_inclusive_scan(void* ptr_vec_in, void* ptr_vec_out, const unsigned long long* offsets,
                               unsigned long long num_segments)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    return TRTC_Segmented_Inclusive_Scan(*(DVVectorLike*)ptr_vec_in, *(DVVectorLike*)ptr_vec_out, off.data(),
                                         num_segments)
               ? 0
               : -1;
}

int n_segmented_sort(void* ptr_keys, const unsigned long long* offsets, unsigned long long num_segments,
                     void* ptr_comp)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    Functor less("Less");
    const Functor& comp = ptr_comp ? *(Functor*)ptr_comp : less;
    return TRTC_Segmented_Sort(*(DVVectorLike*)ptr_keys, off.data(), num_segments, comp) ? 0 : -1;
}

int n_segmented_exclusive_scan(void* ptr_vec_in, void* ptr_vec_out, const unsigned long long* offsets,
                               unsigned long long num_segments)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    return TRTC_Segmented_Exclusive_Scan(*(DVVectorLike*)ptr_vec_in, *(DVVectorLike*)ptr_vec_out, off.data(),
                                         num_segments)
               ? 0
               : -1;
}

int n_segmented_reduce(void* ptr_vec_in, const unsigned long long* offsets, unsigned long long num_segments,
                       void* ptr_results, void* ptr_op)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    Functor plus("Plus");
    const Functor& op = ptr_op ? *(Functor*)ptr_op : plus;
    return TRTC_Segmented_Reduce(*(DVVectorLike*)ptr_vec_in, off.data(), num_segments, *(DVVectorLike*)ptr_results,
                                 op)
               ? 0
               : -1;
}

static unsigned long long copy_json(const std::string& js, char* out, unsigned long long capacity)
{
    if (out && capacity) {
        size_t n = std::min<size_t>(js.size(), capacity - 1);
        memcpy(out, js.data(), n);
        out[n] = 0;
    }
    return js.size();
}

void n_set_batch_autotune(int on) { TRTC_Set_Batch_Autotune(on != 0); }
int n_set_batch_schedule(const char* schedule) { return TRTC_Set_Batch_Schedule(schedule) ? 0 : -2; }
unsigned long long n_batch_tuning_report(char* json_out, unsigned long long capacity)
{
    return copy_json(TRTC_Batch_Tuning_Report(), json_out, capacity);
}

void* n_batch_create(void) { return new TRTC_Batch; }
void n_batch_destroy(void* batch) { delete (TRTC_Batch*)batch; }

int n_batch_add_sort(void* batch, void* ptr_keys, const unsigned long long* offsets, unsigned long long num_segments,
                     void* ptr_comp)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    Functor less("Less");
    const Functor& comp = ptr_comp ? *(Functor*)ptr_comp : less;
    return ((TRTC_Batch*)batch)->Add_Sort(*(DVVectorLike*)ptr_keys, off.data(), num_segments, comp) ? 0 : -1;
}

int n_batch_add_sort_by_key(void* batch, void* ptr_keys, void* ptr_values, const unsigned long long* offsets,
                            unsigned long long num_segments, void* ptr_comp)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    Functor less("Less");
    const Functor& comp = ptr_comp ? *(Functor*)ptr_comp : less;
    return ((TRTC_Batch*)batch)
                   ->Add_Sort_By_Key(*(DVVectorLike*)ptr_keys, *(DVVectorLike*)ptr_values, off.data(), num_segments,
                                     comp)
               ? 0
               : -1;
}

int n_batch_add_inclusive_scan(void* batch, void* ptr_vec_in, void* ptr_vec_out, const unsigned long long* offsets,
                               unsigned long long num_segments, void* ptr_op)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    Functor plus("Plus");
    const Functor& op = ptr_op ? *(Functor*)ptr_op : plus;
    return ((TRTC_Batch*)batch)
                   ->Add_Inclusive_Scan(*(DVVectorLike*)ptr_vec_in, *(DVVectorLike*)ptr_vec_out, off.data(),
                                        num_segments, op)
               ? 0
               : -1;
}

int n_batch_add_reduce(void* batch, void* ptr_vec_in, const unsigned long long* offsets,
                       unsigned long long num_segments, void* ptr_results, void* ptr_op)
{
    std::vector<size_t> off(offsets, offsets + num_segments + 1);
    Functor plus("Plus");
    const Functor& op = ptr_op ? *(Functor*)ptr_op : plus;
    return ((TRTC_Batch*)batch)
                   ->Add_Reduce(*(DVVectorLike*)ptr_vec_in, off.data(), num_segments, *(DVVectorLike*)ptr_results, op)
               ? 0
               : -1;
}

int n_batch_instantiate(void* batch) { return ((TRTC_Batch*)batch)->Instantiate() ? 0 : -1; }
int n_batch_launch(void* batch) { return ((TRTC_Batch*)batch)->Launch() ? 0 : -1; }
unsigned long long n_batch_describe(void* batch, char* json_out, unsigned long long capacity)
{
    return copy_json(((TRTC_Batch*)batch)->Describe(), json_out, capacity);
}

void* n_server_create(unsigned long long hbm_budget, int slots, int batching)
{
    if (!TRTC_Try_Init()) return nullptr;
    return new TRTC_Server((size_t)hbm_budget, slots, batching != 0);
}
void n_server_destroy(void* server) { delete (TRTC_Server*)server; }
void n_server_set_batch_limits(void* server, unsigned long long max_requests, unsigned long long max_elements)
{
    ((TRTC_Server*)server)->Set_Batch_Limits((size_t)max_requests, (size_t)max_elements);
}
long long n_server_submit(void* server, int op, const char* key_elem, const char* val_elem, void* in_a, void* in_b,
                          unsigned long long n)
{
    return ((TRTC_Server*)server)->Submit(op, key_elem, val_elem, in_a, in_b, (size_t)n);
}
int n_server_wait(void* server, long long ticket) { return ((TRTC_Server*)server)->Wait(ticket) ? 0 : -1; }
int n_server_done(void* server, long long ticket) { return ((TRTC_Server*)server)->Done(ticket) ? 1 : 0; }
void* n_server_output(void* server, long long ticket, int which)
{
    return ((TRTC_Server*)server)->Output(ticket, which);
}
double n_server_latency_us(void* server, long long ticket) { return ((TRTC_Server*)server)->Latency_Us(ticket); }
void n_server_release(void* server, long long ticket) { ((TRTC_Server*)server)->Release(ticket); }
unsigned long long n_server_stats(void* server, char* json_out, unsigned long long capacity)
{
    return copy_json(((TRTC_Server*)server)->Stats(), json_out, capacity);
}
double n_server_replay(void* server, unsigned long long count, const double* arrival_us, const int* ops,
                       const char* const* key_elem, const char* const* val_elem, void* const* in_a,
                       void* const* in_b, const unsigned long long* n, double* latency_us)
{
    std::vector<size_t> nn(n, n + count);
    return TRTC_Server_Replay(*(TRTC_Server*)server, (size_t)count, arrival_us, ops, key_elem, val_elem,
                              (const void* const*)in_a, (const void* const*)in_b, nn.data(), latency_us);
}

void n_set_stream(void* stream) { TRTC_Set_Stream(stream); }
void* n_get_stream(void) { return TRTC_Get_Stream(); }

int n_jit_compile_check(const char* arch, const char* name_expr)
{
    std::string log;
    if (compile_only(arch, name_expr, name_expr, &log)) return 0;
    return -1;
}

void n_jit_cache_stats(unsigned long long* out4)
{
    if (out4) jit_cache_stats(out4);
}

int n_jit_cache_entry_valid(const char* blob, unsigned long long size)
{
    if (!blob) return 0;
    std::string lowered, code;
    return parse_cache_entry(std::string(blob, (size_t)size), lowered, code) ? 1 : 0;
}

int n_profile_begin(void)
{
    profile_begin();
    return 0;
}

unsigned long long n_profile_end(char* json_out, unsigned long long capacity)
{
    std::string js = profile_end();
    if (json_out && capacity) {
        size_t n = std::min<size_t>(js.size(), capacity - 1);
        memcpy(json_out, js.data(), n);
        json_out[n] = 0;
    }
    return js.size();
}
