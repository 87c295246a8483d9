// Batched segmented primitives (extensions; the reference has no segmented
// sort).  Each segment [offsets[s], offsets[s+1]) is processed exactly as a
// separate call on that sub-range would process it.
#include "scan.h"
#include "sort.h"
#include "trtc_internal.h"

using namespace trtc_host;

bool TRTC_Segmented_Sort_By_Key(DVVectorLike& keys, DVVectorLike& values, const size_t* offsets,
                                size_t num_segments, const Functor& comp)
{
    if (!ensure_init()) return false;
    void* kp = contiguous_data(keys);
    void* vp = contiguous_data(values);
    if (!kp || !vp) {
        TRTC_ERR("Segmented_Sort_By_Key: keys and values must be contiguous vectors");
        return false;
    }
    for (size_t s = 0; s < num_segments; s++) {
        size_t b = offsets[s], e = offsets[s + 1];
        if (e < b || e > keys.size() || e > values.size()) {
            TRTC_ERR("Segmented_Sort_By_Key: bad segment %zu [%zu, %zu)", s, b, e);
            return false;
        }
        if (e - b < 2) continue;
        DVVectorAdaptor ks(keys.name_elem_cls().c_str(), e - b, (char*)kp + b * keys.elem_size());
        DVVectorAdaptor vs(values.name_elem_cls().c_str(), e - b, (char*)vp + b * values.elem_size());
        if (!TRTC_Sort_By_Key(ks, vs, comp)) return false;
    }
    return true;
}

bool TRTC_Segmented_Inclusive_Scan(const DVVectorLike& vec_in, DVVectorLike& vec_out, const size_t* offsets,
                                   size_t num_segments)
{
    if (!ensure_init()) return false;
    void* ip = contiguous_data(vec_in);
    void* op = contiguous_data(vec_out);
    if (!ip || !op) {
        TRTC_ERR("Segmented_Inclusive_Scan: vectors must be contiguous");
        return false;
    }
    for (size_t s = 0; s < num_segments; s++) {
        size_t b = offsets[s], e = offsets[s + 1];
        if (e < b || e > vec_in.size() || e > vec_out.size()) {
            TRTC_ERR("Segmented_Inclusive_Scan: bad segment %zu [%zu, %zu)", s, b, e);
            return false;
        }
        if (e == b) continue;
        DVVectorAdaptor is(vec_in.name_elem_cls().c_str(), e - b, (char*)ip + b * vec_in.elem_size());
        DVVectorAdaptor os(vec_out.name_elem_cls().c_str(), e - b, (char*)op + b * vec_out.elem_size());
        if (!TRTC_Inclusive_Scan(is, os)) return false;
    }
    return true;
}
