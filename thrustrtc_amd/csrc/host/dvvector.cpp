// Device vectors and functors (reference: DVVector.cpp:1-104, functor.cpp:1-60,
// fake_vectors/DVRange.cpp:1-27).  Struct layouts of functor captures and
// ranges are computed on the host from the members' sizes and alignments
// (C layout rules), where the reference compiled and ran a probe kernel.
#include <algorithm>

#include "trtc_internal.h"

using namespace trtc_host;

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ================================================================ DVVectorLike
DVVectorLike::DVVectorLike(const char* elem_cls, const char* ref_type, size_t size)
{
    m_elem_cls = elem_cls;
    m_ref_type = ref_type;
    const ScalarInfo* si = scalar_info(elem_cls);
    m_elem_size = si ? si->size : TRTC_Size_Of(elem_cls);
    m_size = size;
}

static ViewBuf vector_view(void* data, size_t size)
{
    ViewBuf buf(16);
    memcpy(buf.data(), &data, sizeof(void*));
    memcpy(buf.data() + 8, &size, sizeof(size_t));
    return buf;
}

// ================================================================ DVVector
DVVector::DVVector(const char* elem_cls, size_t size, void* hdata)
    : DVVectorLike(elem_cls, (std::string(elem_cls) + "&").c_str(), size), m_data(nullptr)
{
    m_name_view_cls = std::string("VectorView<") + m_elem_cls + ">";
    if (!ensure_init()) {
        TRTC_ERR("DVVector: no usable GPU");
        return;
    }
    size_t bytes = m_elem_size * m_size;
    if (bytes == 0) return;
    if (!check_hip(hipMalloc(&m_data, bytes), "hipMalloc")) {
        m_data = nullptr;
        return;
    }
    hipStream_t s = cur_stream();
    if (hdata)
        check_hip(hipMemcpyAsync(m_data, hdata, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    else
        check_hip(hipMemsetAsync(m_data, 0, bytes, s), "hipMemsetAsync");
    // host data may be freed right after the constructor returns
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
}

DVVector::~DVVector()
{
    if (m_data) {
        // the vector may still be in use by queued work on any stream
        hipDeviceSynchronize();
        hipFree(m_data);
    }
}

void DVVector::to_host(void* hdata, size_t begin, size_t end) const
{
    if (end == (size_t)(-1) || end > m_size) end = m_size;
    if (end <= begin || !m_data) return;
    hipStream_t s = cur_stream();
    check_hip(hipMemcpyAsync(hdata, (char*)m_data + begin * m_elem_size, (end - begin) * m_elem_size,
                             hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
}

ViewBuf DVVector::view() const { return vector_view(m_data, m_size); }

// ================================================================ DVVectorAdaptor
DVVectorAdaptor::DVVectorAdaptor(const char* elem_cls, size_t size, void* ddata)
    : DVVectorLike(elem_cls, (std::string(elem_cls) + "&").c_str(), size), m_data(ddata)
{
    m_name_view_cls = std::string("VectorView<") + m_elem_cls + ">";
}

DVVectorAdaptor::DVVectorAdaptor(const DVVector& vec, size_t begin, size_t end)
    : DVVectorLike(vec.name_elem_cls().c_str(), vec.name_ref_type().c_str(),
                   (end == (size_t)(-1) ? vec.size() : end) - begin)
{
    m_data = (char*)vec.data() + begin * vec.elem_size();
    m_name_view_cls = std::string("VectorView<") + m_elem_cls + ">";
}

DVVectorAdaptor::DVVectorAdaptor(const DVVectorAdaptor& vec, size_t begin, size_t end)
    : DVVectorLike(vec.name_elem_cls().c_str(), vec.name_ref_type().c_str(),
                   (end == (size_t)(-1) ? vec.size() : end) - begin)
{
    m_data = (char*)vec.data() + begin * vec.elem_size();
    m_name_view_cls = std::string("VectorView<") + m_elem_cls + ">";
}

ViewBuf DVVectorAdaptor::view() const { return vector_view(m_data, m_size); }

// ================================================================ DVRange
// Device layout: RangeView<Inner> { Inner _view_vec_value; size_t _begin; size_t _end; }
DVRange::DVRange(const DVVectorLike& vec_value, size_t begin, size_t end)
    : DVVectorLike(vec_value.name_elem_cls().c_str(), vec_value.name_ref_type().c_str(),
                   (end == (size_t)(-1) ? vec_value.size() : end) - begin)
{
    if (end == (size_t)(-1)) end = vec_value.size();
    m_readable = vec_value.is_readable();
    m_writable = vec_value.is_writable();
    m_view_vec_value = vec_value.view();
    m_begin = begin;
    m_end = end;
    m_name_view_cls = std::string("RangeView<") + vec_value.name_view_cls() + ">";
    void* base = contiguous_data(vec_value);
    m_contig = base ? (char*)base + begin * vec_value.elem_size() : nullptr;
}

ViewBuf DVRange::view() const
{
    size_t a = std::max<size_t>(8, view_align(m_name_view_cls.substr(10, m_name_view_cls.size() - 11),
                                              m_view_vec_value.size()));
    size_t o1 = align_up(m_view_vec_value.size(), 8);
    size_t total = align_up(o1 + 16, a);
    ViewBuf buf(total, 0);
    memcpy(buf.data(), m_view_vec_value.data(), m_view_vec_value.size());
    memcpy(buf.data() + o1, &m_begin, 8);
    memcpy(buf.data() + o1 + 8, &m_end, 8);
    return buf;
}

// ================================================================ Functor
// The struct the reference generates (functor.cpp:4-45):
//   struct _S_<hash> { typedef _S_<hash> CurType;
//       <view type> <name>; ...
//       template<class _T0, ...> __device__ inline auto operator()(const _T0& p0, ...) { body } };
// Members are laid out with C rules; the view is the members' views at their
// offsets (an empty struct is one byte, like C++).
Functor::Functor(const std::vector<CapturedDeviceViewable>& arg_map, const std::vector<const char*>& functor_params,
                 const char* code_body)
{
    std::string body;
    m_view_args.resize(arg_map.size());
    m_offsets.resize(arg_map.size() + 1);
    size_t off = 0, max_align = 1;
    for (size_t i = 0; i < arg_map.size(); i++) {
        const DeviceViewable* obj = arg_map[i].obj;
        body += std::string("    ") + obj->name_view_cls() + " " + arg_map[i].obj_name + ";\n";
        m_view_args[i] = obj->view();
        size_t a = view_align(obj->name_view_cls(), m_view_args[i].size());
        off = align_up(off, a);
        m_offsets[i] = off;
        off += m_view_args[i].size();
        max_align = std::max(max_align, a);
    }
    size_t total = arg_map.empty() ? 1 : align_up(off, max_align);
    m_offsets[arg_map.size()] = total;

    if (!functor_params.empty()) {
        body += "    template<";
        for (size_t i = 0; i < functor_params.size(); i++) {
            body += "class _T" + std::to_string(i);
            if (i + 1 < functor_params.size()) body += ",";
        }
        body += ">\n";
    }
    body += "    __device__ inline auto operator()(";
    for (size_t i = 0; i < functor_params.size(); i++) {
        body += "const _T" + std::to_string(i) + "& " + functor_params[i];
        if (i + 1 < functor_params.size()) body += ",";
    }
    body += ")\n    {\n";
    body += code_body;
    body += "\n    }\n";
    m_name_view_cls = add_struct(body);
    register_struct_layout(m_name_view_cls, total, max_align);
}

Functor::Functor(const char* name_built_in_view_cls)
{
    m_name_view_cls = name_built_in_view_cls;
    m_offsets.resize(1);
    m_offsets[0] = 1;
}

ViewBuf Functor::view() const
{
    ViewBuf ret(m_offsets[m_offsets.size() - 1], 0);
    for (size_t i = 0; i < m_view_args.size(); i++)
        memcpy(ret.data() + m_offsets[i], m_view_args[i].data(), m_view_args[i].size());
    return ret;
}
