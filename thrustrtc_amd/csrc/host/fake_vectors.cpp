// Fake vectors (reference: fake_vectors/*.cpp): vector-likes whose elements
// are computed from an index.  Each class names a device view struct
// (trtc_builtin.hpp, or a struct generated here for DVZipped / DVCustomVector)
// and fills its bytes.  Layouts are computed on the host with C rules from the
// members' view sizes and alignments and registered, so nested views and
// TRTC_Size_Of of the generated element structs need no probe kernel (the
// reference queried every layout with a compiled probe).
#include <algorithm>

#include "fake_vectors/DVConstant.h"
#include "fake_vectors/DVCounter.h"
#include "fake_vectors/DVCustomVector.h"
#include "fake_vectors/DVDiscard.h"
#include "fake_vectors/DVPermutation.h"
#include "fake_vectors/DVReverse.h"
#include "fake_vectors/DVTransform.h"
#include "fake_vectors/DVZipped.h"
#include "trtc_internal.h"

using namespace trtc_host;

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// C struct layout built member by member.
struct CLayout
{
    size_t end = 0, align = 1;
    size_t add(size_t size, size_t a)
    {
        end = align_up(end, a);
        size_t off = end;
        end += size;
        align = std::max(align, a);
        return off;
    }
    size_t add_view(const std::string& cls, const ViewBuf& v) { return add(v.size(), view_align(cls, v.size())); }
    size_t total() const { return align_up(std::max<size_t>(end, 1), align); }
};

// Layout { size_t _size; <view of cls> _v; } of ConstantView / CounterView.
void size_value_layout(const std::string& view_cls, const std::string& elem_cls, const ViewBuf& value,
                       size_t offsets[3])
{
    CLayout l;
    offsets[0] = l.add(8, 8);
    offsets[1] = l.add_view(elem_cls, value);
    offsets[2] = l.total();
    register_struct_layout(view_cls, offsets[2], l.align);
}

// Two consecutive views (PermutationView, TransformView).
void pair_layout(const std::string& view_cls, const std::string& cls_a, const ViewBuf& a, const std::string& cls_b,
                 const ViewBuf& b, size_t offsets[3])
{
    CLayout l;
    offsets[0] = l.add_view(cls_a, a);
    offsets[1] = l.add_view(cls_b, b);
    offsets[2] = l.total();
    register_struct_layout(view_cls, offsets[2], l.align);
}

ViewBuf fill(size_t total, const size_t* offsets, std::initializer_list<const ViewBuf*> parts)
{
    ViewBuf buf(total, 0);
    size_t i = 0;
    for (const ViewBuf* p : parts) {
        memcpy(buf.data() + offsets[i], p->data(), p->size());
        i++;
    }
    return buf;
}

ViewBuf size_buf(size_t n)
{
    ViewBuf b(8);
    memcpy(b.data(), &n, 8);
    return b;
}

}  // namespace

// ================================================================ DVConstant
DVConstant::DVConstant(const DeviceViewable& dvobj, size_t size)
    : DVVectorLike(dvobj.name_view_cls().c_str(), ("const " + dvobj.name_view_cls() + "&").c_str(), size)
{
    m_value = dvobj.view();
    m_name_view_cls = "ConstantView<" + m_elem_cls + ">";
    size_value_layout(m_name_view_cls, m_elem_cls, m_value, m_offsets);
}

ViewBuf DVConstant::view() const
{
    ViewBuf sz = size_buf(m_size);
    return fill(m_offsets[2], m_offsets, {&sz, &m_value});
}

// ================================================================ DVCounter
DVCounter::DVCounter(const DeviceViewable& dvobj_init, size_t size)
    : DVVectorLike(dvobj_init.name_view_cls().c_str(), dvobj_init.name_view_cls().c_str(), size)
{
    m_value_init = dvobj_init.view();
    m_name_view_cls = "CounterView<" + m_elem_cls + ">";
    size_value_layout(m_name_view_cls, m_elem_cls, m_value_init, m_offsets);
}

ViewBuf DVCounter::view() const
{
    ViewBuf sz = size_buf(m_size);
    return fill(m_offsets[2], m_offsets, {&sz, &m_value_init});
}

// ================================================================ DVDiscard
// DiscardView<T> { size_t _size; _Sink<T> _sink; }: 16 bytes.
DVDiscard::DVDiscard(const char* elem_cls, size_t size)
    : DVVectorLike(elem_cls, ("const _Sink<" + std::string(elem_cls) + ">&").c_str(), size)
{
    m_name_view_cls = "DiscardView<" + m_elem_cls + ">";
    register_struct_layout(m_name_view_cls, 16, 8);
}

ViewBuf DVDiscard::view() const
{
    ViewBuf buf(16, 0);
    memcpy(buf.data(), &m_size, 8);
    return buf;
}

// ================================================================ DVPermutation
DVPermutation::DVPermutation(const DVVectorLike& vec_value, const DVVectorLike& vec_index)
    : DVVectorLike(vec_value.name_elem_cls().c_str(), vec_value.name_ref_type().c_str(), vec_index.size())
{
    m_readable = vec_value.is_readable();
    m_writable = vec_value.is_writable();
    m_view_value = vec_value.view();
    m_view_index = vec_index.view();
    m_name_view_cls = "PermutationView<" + vec_value.name_view_cls() + ", " + vec_index.name_view_cls() + ">";
    pair_layout(m_name_view_cls, vec_value.name_view_cls(), m_view_value, vec_index.name_view_cls(), m_view_index,
                m_offsets);
}

ViewBuf DVPermutation::view() const { return fill(m_offsets[2], m_offsets, {&m_view_value, &m_view_index}); }

// ================================================================ DVReverse
DVReverse::DVReverse(const DVVectorLike& vec_value)
    : DVVectorLike(vec_value.name_elem_cls().c_str(), vec_value.name_ref_type().c_str(), vec_value.size())
{
    m_readable = vec_value.is_readable();
    m_writable = vec_value.is_writable();
    m_view_value = vec_value.view();
    m_name_view_cls = "ReverseView<" + vec_value.name_view_cls() + ">";
    register_struct_layout(m_name_view_cls, m_view_value.size(),
                           view_align(vec_value.name_view_cls(), m_view_value.size()));
}

ViewBuf DVReverse::view() const { return m_view_value; }

// ================================================================ DVTransform
DVTransform::DVTransform(const DVVectorLike& vec_in, const char* elem_cls, const Functor& op)
    : DVVectorLike(elem_cls, elem_cls, vec_in.size())
{
    m_view_in = vec_in.view();
    m_view_op = op.view();
    m_name_view_cls =
        "TransformView<" + m_elem_cls + ", " + vec_in.name_view_cls() + ", " + op.name_view_cls() + ">";
    pair_layout(m_name_view_cls, vec_in.name_view_cls(), m_view_in, op.name_view_cls(), m_view_op, m_offsets);
}

ViewBuf DVTransform::view() const { return fill(m_offsets[2], m_offsets, {&m_view_in, &m_view_op}); }

// ================================================================ DVZipped
namespace {

// struct { T0 name0; T1 name1; ... } -- the element type
std::string zipped_elem_struct(const std::vector<DVVectorLike*>& vecs, const std::vector<const char*>& names)
{
    std::string body;
    CLayout l;
    for (size_t i = 0; i < vecs.size(); i++) {
        body += "    " + vecs[i]->name_elem_cls() + " " + names[i] + ";\n";
        l.add(vecs[i]->elem_size(), view_align(vecs[i]->name_elem_cls(), vecs[i]->elem_size()));
    }
    std::string name = add_struct(body);
    register_struct_layout(name, l.total(), l.align);
    return name;
}

// struct { R0 name0; R1 name1; ... } of the zipped vectors' reference types,
// convertible to the element type when every vector is readable and
// assignable from it when every vector is writable.
std::string zipped_ref_struct(const std::vector<DVVectorLike*>& vecs, const std::vector<const char*>& names)
{
    std::string elem = zipped_elem_struct(vecs, names);
    bool readable = true, writable = true;
    for (auto* v : vecs) {
        readable = readable && v->is_readable();
        writable = writable && v->is_writable();
    }
    std::string body;
    for (size_t i = 0; i < vecs.size(); i++) body += "    " + vecs[i]->name_ref_type() + " " + names[i] + ";\n";
    if (readable) {
        body += "    __device__ operator " + elem + "() const\n    {\n        " + elem + " r;\n";
        for (size_t i = 0; i < vecs.size(); i++) body += std::string("        r.") + names[i] + " = " + names[i] + ";\n";
        body += "        return r;\n    }\n";
    }
    if (writable) {
        body += "    __device__ const CurType& operator=(const " + elem + "& v) const\n    {\n";
        for (size_t i = 0; i < vecs.size(); i++) body += std::string("        ") + names[i] + " = v." + names[i] + ";\n";
        body += "        return *this;\n    }\n";
    }
    return add_struct(body);
}

}  // namespace

DVZipped::DVZipped(const std::vector<DVVectorLike*>& vecs, const std::vector<const char*>& elem_names)
    : DVVectorLike(zipped_elem_struct(vecs, elem_names).c_str(),
                   zipped_ref_struct(vecs, elem_names).c_str(), vecs.empty() ? 0 : vecs[0]->size())
{
    m_readable = m_writable = true;
    for (auto* v : vecs) {
        m_readable = m_readable && v->is_readable();
        m_writable = m_writable && v->is_writable();
    }
    std::string body = "    typedef " + m_elem_cls + " value_t;\n    typedef " + m_ref_type + " ref_t;\n";
    CLayout l;
    m_view_elems.resize(vecs.size());
    m_offsets.resize(vecs.size() + 1);
    for (size_t i = 0; i < vecs.size(); i++) {
        m_view_elems[i] = vecs[i]->view();
        body += "    mutable " + vecs[i]->name_view_cls() + " " + elem_names[i] + ";\n";
        m_offsets[i] = l.add_view(vecs[i]->name_view_cls(), m_view_elems[i]);
    }
    m_offsets[vecs.size()] = l.total();
    body += std::string("    __device__ size_t size() const { return ") + (vecs.empty() ? "0" : elem_names[0]) +
            ".size(); }\n";
    body += "    __device__ ref_t operator[](size_t idx) const { return ref_t{";
    for (size_t i = 0; i < vecs.size(); i++) body += std::string(i ? ", " : "") + elem_names[i] + "[idx]";
    body += "}; }\n";
    m_name_view_cls = add_struct(body);
    register_struct_layout(m_name_view_cls, l.total(), l.align);
}

ViewBuf DVZipped::view() const
{
    ViewBuf buf(m_offsets.back(), 0);
    for (size_t i = 0; i < m_view_elems.size(); i++)
        memcpy(buf.data() + m_offsets[i], m_view_elems[i].data(), m_view_elems[i].size());
    return buf;
}

// ================================================================ DVCustomVector
DVCustomVector::DVCustomVector(const std::vector<CapturedDeviceViewable>& arg_map, const char* name_idx,
                               const char* code_body, const char* elem_cls, size_t size, bool read_only)
    : DVVectorLike(elem_cls, read_only ? elem_cls : (std::string(elem_cls) + "&").c_str(), size),
      m_read_only(read_only)
{
    // the element functor: captured views as members, operator()(index)
    std::string fbody;
    CLayout fl;
    m_view_args.resize(arg_map.size());
    m_arg_offsets.resize(arg_map.size() + 1);
    for (size_t i = 0; i < arg_map.size(); i++) {
        const DeviceViewable* obj = arg_map[i].obj;
        m_view_args[i] = obj->view();
        fbody += "    " + obj->name_view_cls() + " " + arg_map[i].obj_name + ";\n";
        m_arg_offsets[i] = fl.add_view(obj->name_view_cls(), m_view_args[i]);
    }
    m_arg_offsets[arg_map.size()] = fl.total();
    fbody += "    __device__ inline " + m_ref_type + " operator()(size_t " + name_idx + ")\n    {\n" + code_body +
             "\n    }\n";
    std::string fname = add_struct(fbody);
    register_struct_layout(fname, fl.total(), fl.align);

    // the view: { size_t _size; functor _op; }
    std::string body = "    typedef " + m_elem_cls + " value_t;\n    typedef " + m_ref_type + " ref_t;\n";
    body += "    size_t _size;\n    mutable " + fname + " _op;\n";
    body += "    __device__ size_t size() const { return _size; }\n";
    body += "    __device__ ref_t operator[](size_t idx) const { return _op(idx); }\n";
    m_name_view_cls = add_struct(body);
    CLayout l;
    m_offsets[0] = l.add(8, 8);
    m_offsets[1] = l.add(fl.total(), fl.align);
    m_offsets[2] = l.total();
    register_struct_layout(m_name_view_cls, m_offsets[2], l.align);
}

ViewBuf DVCustomVector::view() const
{
    ViewBuf op(m_arg_offsets.back(), 0);
    for (size_t i = 0; i < m_view_args.size(); i++)
        memcpy(op.data() + m_arg_offsets[i], m_view_args[i].data(), m_view_args[i].size());
    ViewBuf sz = size_buf(m_size);
    return fill(m_offsets[2], m_offsets, {&sz, &op});
}
