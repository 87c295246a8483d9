/* C ABI of the ThrustRTC primitive core on MI355X (libPyThrustRTC.so).
 *
 * Every entry point keeps the signature, argument meaning, ownership rule and
 * error convention of the reference's FFI header python/api.h (cited per
 * line), so the reference's Python (cffi/ctypes), JNI and C++/CLI bindings
 * bind it unchanged.  Plain pointers and sizes only.
 *
 * Handles are heap objects created by n_*_create and freed by the matching
 * n_*_destroy / n_dv_destroy.  Return conventions: int 0 = ok, -1 = internal
 * error, -2 = wrong number of arguments, -100 = cannot initialise; unsigned
 * 0xFFFFFFFF = error; NULL pointer = error.  A NULL functor or init argument
 * selects the reference's default overload (Less / Plus / EqualTo / zero).
 */
#ifndef TRTC_CAPI_H
#define TRTC_CAPI_H

#define TRTC_CAPI __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* ---- utils (api.h:12-17) ---- */
TRTC_CAPI void* n_string_array_create(unsigned long long size, const char* const* strs); /* api.h:12 */
TRTC_CAPI void n_string_array_destroy(void* ptr_arr);                                    /* api.h:13 */
TRTC_CAPI void* n_pointer_array_create(unsigned long long size, const void* const* ptrs); /* api.h:14 */
TRTC_CAPI void n_pointer_array_destroy(void* ptr_arr);                                    /* api.h:15 */
TRTC_CAPI void* n_dim3_create(unsigned x, unsigned y, unsigned z);                        /* api.h:16 */
TRTC_CAPI void n_dim3_destroy(void* cptr);                                                /* api.h:17 */

/* ---- context (api.h:20-43) ---- */
TRTC_CAPI void n_set_libnvrtc_path(const char* path);   /* api.h:20; sets the hipRTC library path */
TRTC_CAPI int n_trtc_try_init(void);                    /* api.h:21; 0 or -100 */
TRTC_CAPI int n_get_ptx_arch(void);                     /* api.h:22; gfx number, e.g. 950 */
TRTC_CAPI void n_set_verbose(unsigned verbose);         /* api.h:23 */
TRTC_CAPI void n_set_kernel_debug(unsigned kernel_debug); /* api.h:24 */
TRTC_CAPI void n_add_include_dir(const char* dir);      /* api.h:25 */
TRTC_CAPI void n_add_built_in_header(const char* filename, const char* filecontent); /* api.h:26 */
TRTC_CAPI void n_add_inlcude_filename(const char* fn);  /* api.h:27 (sic) */
TRTC_CAPI void n_add_code_block(const char* line);      /* api.h:28 */
TRTC_CAPI void n_add_constant_object(const char* name, void* cptr); /* api.h:29 */
TRTC_CAPI void n_wait(void);                            /* api.h:30 */

TRTC_CAPI void* n_kernel_create(void* ptr_param_list, const char* body);  /* api.h:32 */
TRTC_CAPI void n_kernel_destroy(void* cptr);                              /* api.h:33 */
TRTC_CAPI int n_kernel_num_params(void* cptr);                            /* api.h:34 */
TRTC_CAPI int n_kernel_calc_optimal_block_size(void* ptr_kernel, void* ptr_arg_list, unsigned sharedMemBytes); /* api.h:35 */
TRTC_CAPI int n_kernel_calc_number_blocks(void* ptr_kernel, void* ptr_arg_list, int sizeBlock, unsigned sharedMemBytes); /* api.h:36 */
TRTC_CAPI int n_kernel_launch(void* ptr_kernel, void* ptr_gridDim, void* ptr_blockDim, void* ptr_arg_list, int sharedMemBytes); /* api.h:37 */

TRTC_CAPI void* n_for_create(void* ptr_param_list, const char* name_iter, const char* body); /* api.h:39 */
TRTC_CAPI void n_for_destroy(void* cptr);                                 /* api.h:40 */
TRTC_CAPI int n_for_num_params(void* cptr);                               /* api.h:41 */
TRTC_CAPI int n_for_launch(void* cptr, int begin, int end, void* ptr_arg_list); /* api.h:42 */
TRTC_CAPI int n_for_launch_n(void* cptr, int n, void* ptr_arg_list);      /* api.h:43 */

/* ---- device-viewable scalars (api.h:46-75) ---- */
TRTC_CAPI const char* n_dv_name_view_cls(void* cptr);   /* api.h:46 */
TRTC_CAPI void n_dv_destroy(void* cptr);                /* api.h:47 */
TRTC_CAPI void* n_dvint8_create(int v);                 /* api.h:48 */
TRTC_CAPI int n_dvint8_value(void* cptr);               /* api.h:49 */
TRTC_CAPI void* n_dvuint8_create(unsigned v);           /* api.h:50 */
TRTC_CAPI unsigned n_dvuint8_value(void* cptr);         /* api.h:51 */
TRTC_CAPI void* n_dvint16_create(int v);                /* api.h:52 */
TRTC_CAPI int n_dvint16_value(void* cptr);              /* api.h:53 */
TRTC_CAPI void* n_dvuint16_create(unsigned v);          /* api.h:54 */
TRTC_CAPI unsigned n_dvuint16_value(void* cptr);        /* api.h:55 */
TRTC_CAPI void* n_dvint32_create(int v);                /* api.h:56 */
TRTC_CAPI int n_dvint32_value(void* cptr);              /* api.h:57 */
TRTC_CAPI void* n_dvuint32_create(unsigned v);          /* api.h:58 */
TRTC_CAPI unsigned n_dvuint32_value(void* cptr);        /* api.h:59 */
TRTC_CAPI void* n_dvint64_create(long long v);          /* api.h:60 */
TRTC_CAPI long long n_dvint64_value(void* cptr);        /* api.h:61 */
TRTC_CAPI void* n_dvuint64_create(unsigned long long v); /* api.h:62 */
TRTC_CAPI unsigned long long n_dvuint64_value(void* cptr); /* api.h:63 */
TRTC_CAPI void* n_dvfloat_create(float v);              /* api.h:64 */
TRTC_CAPI float n_dvfloat_value(void* cptr);            /* api.h:65 */
TRTC_CAPI void* n_dvdouble_create(double v);            /* api.h:66 */
TRTC_CAPI double n_dvdouble_value(void* cptr);          /* api.h:67 */
TRTC_CAPI void* n_dvbool_create(int v);                 /* api.h:68 */
TRTC_CAPI int n_dvbool_value(void* cptr);               /* api.h:69 */
TRTC_CAPI void* n_dvcomplex64_create(float real, float imag);   /* api.h:70 */
TRTC_CAPI float n_dvcomplex64_real(void* cptr);                 /* api.h:71 */
TRTC_CAPI float n_dvcomplex64_imag(void* cptr);                 /* api.h:72 */
TRTC_CAPI void* n_dvcomplex128_create(double real, double imag); /* api.h:73 */
TRTC_CAPI double n_dvcomplex128_real(void* cptr);               /* api.h:74 */
TRTC_CAPI double n_dvcomplex128_imag(void* cptr);               /* api.h:75 */

/* ---- vectors (api.h:78-84) ---- */
TRTC_CAPI const char* n_dvvectorlike_name_elem_cls(void* cptr);  /* api.h:78 */
TRTC_CAPI unsigned long long n_dvvectorlike_size(void* cptr);    /* api.h:79 */
TRTC_CAPI void* n_dvvector_create(const char* elem_cls, unsigned long long size, void* hdata); /* api.h:80 */
TRTC_CAPI void n_dvvector_to_host(void* cptr, void* hdata, unsigned long long begin, unsigned long long end); /* api.h:81 */
TRTC_CAPI void* n_dvvector_from_dvs(void* ptr_dvs);              /* api.h:82 */
TRTC_CAPI void* n_dvvectoradaptor_create(const char* elem_cls, unsigned long long size, void* data); /* api.h:83 */
TRTC_CAPI void* n_dvrange_create(void* ptr_in, unsigned long long begin, unsigned long long end); /* api.h:84 */

/* ---- fake vectors (api.h:90-97; fake_vectors/*.h) ---- */
TRTC_CAPI void* n_dvconstant_create(void* cptr, unsigned long long size);          /* api.h:90 */
TRTC_CAPI void* n_dvcounter_create(void* cptr, unsigned long long size);           /* api.h:91 */
TRTC_CAPI void* n_dvdiscard_create(const char* elem_cls, unsigned long long size); /* api.h:92 */
TRTC_CAPI void* n_dvpermutation_create(void* ptr_value, void* ptr_index);          /* api.h:93 */
TRTC_CAPI void* n_dvreverse_create(void* ptr_value);                               /* api.h:94 */
TRTC_CAPI void* n_dvtransform_create(void* ptr_in, const char* elem_cls, void* ptr_op); /* api.h:95 */
TRTC_CAPI void* n_dvzipped_create(void* ptr_vecs, void* ptr_elem_names);          /* api.h:96 */
TRTC_CAPI void* n_dvcustomvector_create(void* ptr_dvs, void* ptr_names, const char* name_idx, const char* body,
                                        const char* elem_cls, unsigned long long size, unsigned read_only); /* api.h:97 */

/* ---- functors (api.h:100-101) ---- */
TRTC_CAPI void* n_functor_create(void* ptr_dvs, void* ptr_names, void* ptr_functor_params, const char* code_body); /* api.h:100 */
TRTC_CAPI void* n_built_in_functor_create(const char* name_built_in_view_cls); /* api.h:101 */

/* ---- algorithms on the hot path ---- */
TRTC_CAPI int n_fill(void* ptr_vec, void* ptr_value);                 /* api.h:104 */
TRTC_CAPI int n_copy(void* ptr_in, void* ptr_out);                    /* api.h:124 */
TRTC_CAPI void* n_reduce(void* ptr_vec, void* ptr_init, void* ptr_bin_op); /* api.h:130; new DV, caller frees */
TRTC_CAPI int n_is_sorted(void* ptr_vec, void* ptr_comp);             /* api.h:142 */
TRTC_CAPI int n_inclusive_scan(void* ptr_vec_in, void* ptr_vec_out, void* ptr_binary_op); /* api.h:145 */
TRTC_CAPI int n_exclusive_scan(void* ptr_vec_in, void* ptr_vec_out, void* ptr_init, void* ptr_binary_op); /* api.h:146 */
TRTC_CAPI int n_inclusive_scan_by_key(void* ptr_vec_key, void* ptr_vec_value, void* ptr_vec_out, void* ptr_binary_pred, void* ptr_binary_op); /* api.h:147 */
TRTC_CAPI int n_exclusive_scan_by_key(void* ptr_vec_key, void* ptr_vec_value, void* ptr_vec_out, void* ptr_init, void* ptr_binary_pred, void* ptr_binary_op); /* api.h:148 */
TRTC_CAPI unsigned n_copy_if(void* ptr_in, void* ptr_out, void* ptr_pred); /* api.h:153 */
TRTC_CAPI unsigned n_copy_if_stencil(void* ptr_in, void* ptr_stencil, void* ptr_out, void* ptr_pred); /* api.h:154 */
TRTC_CAPI int n_sort(void* ptr_vec, void* ptr_comp);                  /* api.h:189 */
TRTC_CAPI int n_sort_by_key(void* ptr_keys, void* ptr_values, void* ptr_comp); /* api.h:190 */

/* ---- scan / compaction consumers (api.h:128-168) ---- */
TRTC_CAPI unsigned long long n_count(void* ptr_vec, void* ptr_value);  /* api.h:128; (unsigned long long)-1 on error */
TRTC_CAPI unsigned long long n_count_if(void* ptr_vec, void* ptr_pred); /* api.h:129 */
TRTC_CAPI unsigned n_reduce_by_key(void* ptr_key_in, void* ptr_value_in, void* ptr_key_out, void* ptr_value_out, void* ptr_binary_pred, void* ptr_binary_op); /* api.h:131 */
TRTC_CAPI void* n_transform_reduce(void* ptr_vec, void* ptr_unary_op, void* ptr_init, void* ptr_binary_op); /* api.h:137; new DV, caller frees */
TRTC_CAPI int n_transform_inclusive_scan(void* ptr_vec_in, void* ptr_vec_out, void* ptr_unary_op, void* ptr_binary_op); /* api.h:149 */
TRTC_CAPI int n_transform_exclusive_scan(void* ptr_vec_in, void* ptr_vec_out, void* ptr_unary_op, void* ptr_init, void* ptr_binary_op); /* api.h:150; -1 on error (the reference returns 1) */
TRTC_CAPI unsigned n_remove(void* ptr_vec, void* ptr_value);           /* api.h:155 */
TRTC_CAPI unsigned n_remove_copy(void* ptr_in, void* ptr_out, void* ptr_value); /* api.h:156 */
TRTC_CAPI unsigned n_remove_if(void* ptr_vec, void* ptr_pred);         /* api.h:157 */
TRTC_CAPI unsigned n_remove_copy_if(void* ptr_in, void* ptr_out, void* ptr_pred); /* api.h:158 */
TRTC_CAPI unsigned n_remove_if_stencil(void* ptr_vec, void* ptr_stencil, void* ptr_pred); /* api.h:159 */
TRTC_CAPI unsigned n_remove_copy_if_stencil(void* ptr_in, void* ptr_stencil, void* ptr_out, void* ptr_pred); /* api.h:160 */
TRTC_CAPI unsigned n_unique(void* ptr_vec, void* ptr_binary_pred);    /* api.h:161 */
TRTC_CAPI unsigned n_unique_copy(void* ptr_vec_in, void* ptr_vec_out, void* ptr_binary_pred); /* api.h:162 */
TRTC_CAPI unsigned n_unique_by_key(void* ptr_keys, void* ptr_values, void* ptr_binary_pred); /* api.h:163 */
TRTC_CAPI unsigned n_unique_by_key_copy(void* ptr_keys_in, void* ptr_values_in, void* ptr_key_out, void* ptr_values_out, void* ptr_binary_pred); /* api.h:164 */
TRTC_CAPI unsigned n_partition(void* ptr_vec, void* ptr_pred);        /* api.h:165 */
TRTC_CAPI unsigned n_partition_stencil(void* ptr_vec, void* ptr_stencil, void* ptr_pred); /* api.h:166 */
TRTC_CAPI unsigned n_partition_copy(void* ptr_vec_in, void* ptr_vec_true, void* ptr_vec_false, void* ptr_pred); /* api.h:167 */
TRTC_CAPI unsigned n_partition_copy_stencil(void* ptr_vec_in, void* ptr_stencil, void* ptr_vec_true, void* ptr_vec_false, void* ptr_pred); /* api.h:168 */

/* ---- the other algorithms (api.h:87-186) ---- */
TRTC_CAPI void* n_dvtuple_create(void* ptr_dvs, void* ptr_names); /* api.h:87 */
TRTC_CAPI int n_replace(void* ptr_vec, void* ptr_old_value, void* ptr_new_value); /* api.h:105 */
TRTC_CAPI int n_replace_if(void* ptr_vec, void* p_pred, void* ptr_new_value); /* api.h:106 */
TRTC_CAPI int n_replace_copy(void* ptr_in, void* ptr_out, void* ptr_old_value, void* ptr_new_value); /* api.h:107 */
TRTC_CAPI int n_replace_copy_if(void* ptr_in, void* ptr_out, void* p_pred, void* ptr_new_value); /* api.h:108 */
TRTC_CAPI int n_for_each(void* ptr_vec, void* ptr_f); /* api.h:109 */
TRTC_CAPI int n_adjacent_difference(void* ptr_in, void* ptr_out, void* ptr_binary_op); /* api.h:110 */
TRTC_CAPI int n_sequence(void* ptr_vec, void* ptr_value_init, void* ptr_value_step); /* api.h:111 */
TRTC_CAPI int n_tabulate(void* ptr_vec, void* ptr_op); /* api.h:112 */
TRTC_CAPI int n_transform(void* ptr_in, void* ptr_out, void* ptr_op); /* api.h:113 */
TRTC_CAPI int n_transform_binary(void* ptr_in1, void* ptr_in2, void* ptr_out, void* ptr_op); /* api.h:114 */
TRTC_CAPI int n_transform_if(void* ptr_in, void* ptr_out, void* ptr_op, void* ptr_pred); /* api.h:115 */
TRTC_CAPI int n_transform_if_stencil(void* ptr_in, void* ptr_stencil, void* ptr_out, void* ptr_op, void* ptr_pred); /* api.h:116 */
TRTC_CAPI int n_transform_binary_if_stencil(void* ptr_in1, void* ptr_in2, void* ptr_stencil, void* ptr_out, void* ptr_op, void* ptr_pred); /* api.h:117 */
TRTC_CAPI int n_gather(void* ptr_map, void* ptr_in, void* ptr_out); /* api.h:120 */
TRTC_CAPI int n_gather_if(void* ptr_map, void* ptr_stencil, void* ptr_in, void* ptr_out, void* ptr_pred); /* api.h:121 */
TRTC_CAPI int n_scatter(void* ptr_in, void* ptr_map, void* ptr_out); /* api.h:122 */
TRTC_CAPI int n_scatter_if(void* ptr_in, void* ptr_map, void* ptr_stencil, void* ptr_out, void* ptr_pred); /* api.h:123 */
TRTC_CAPI int n_swap(void* ptr_vec1, void* ptr_vec2); /* api.h:125 */
TRTC_CAPI int n_equal(void* ptr_vec1, void* ptr_vec2, void* ptr_binary_pred); /* api.h:132; 1 / 0, -1 on error */
TRTC_CAPI unsigned long long n_min_element(void* ptr_vec, void* ptr_comp); /* api.h:133 */
TRTC_CAPI unsigned long long n_max_element(void* ptr_vec, void* ptr_comp); /* api.h:134 */
TRTC_CAPI int n_minmax_element(void* ptr_vec, void* ptr_comp, unsigned long long* ret); /* api.h:135 */
TRTC_CAPI void* n_inner_product(void* ptr_vec1, void* ptr_vec2, void* ptr_init, void* ptr_binary_op1, void* ptr_binary_op2); /* api.h:136; new DV */
TRTC_CAPI int n_all_of(void* ptr_vec, void* ptr_pred); /* api.h:138 */
TRTC_CAPI int n_any_of(void* ptr_vec, void* ptr_pred); /* api.h:139 */
TRTC_CAPI int n_none_of(void* ptr_vec, void* ptr_pred); /* api.h:140 */
TRTC_CAPI int n_is_partitioned(void* ptr_vec, void* ptr_pred); /* api.h:141 */
TRTC_CAPI long long n_find(void* ptr_vec, void* ptr_value); /* api.h:171; vec size when absent, -1 on error */
TRTC_CAPI long long n_find_if(void* ptr_vec, void* ptr_pred); /* api.h:172 */
TRTC_CAPI long long n_find_if_not(void* ptr_vec, void* ptr_pred); /* api.h:173 */
TRTC_CAPI long long n_mismatch(void* ptr_vec1, void* ptr_vec2, void* ptr_pred); /* api.h:174 */
TRTC_CAPI unsigned long long n_lower_bound(void* ptr_vec, void* ptr_value, void* ptr_comp); /* api.h:175 */
TRTC_CAPI unsigned long long n_upper_bound(void* ptr_vec, void* ptr_value, void* ptr_comp); /* api.h:176 */
TRTC_CAPI int n_binary_search(void* ptr_vec, void* ptr_value, void* ptr_comp); /* api.h:177 */
TRTC_CAPI int n_lower_bound_v(void* ptr_vec, void* ptr_values, void* ptr_result, void* ptr_comp); /* api.h:178 */
TRTC_CAPI int n_upper_bound_v(void* ptr_vec, void* ptr_values, void* ptr_result, void* ptr_comp); /* api.h:179 */
TRTC_CAPI int n_binary_search_v(void* ptr_vec, void* ptr_values, void* ptr_result, void* ptr_comp); /* api.h:180 */
TRTC_CAPI unsigned long long n_partition_point(void* ptr_vec, void* ptr_pred); /* api.h:181 */
TRTC_CAPI unsigned long long n_is_sorted_until(void* ptr_vec, void* ptr_comp); /* api.h:182 */
TRTC_CAPI int n_merge(void* ptr_vec1, void* ptr_vec2, void* ptr_vec_out, void* ptr_comp); /* api.h:185 */
TRTC_CAPI int n_merge_by_key(void* ptr_keys1, void* ptr_keys2, void* ptr_value1, void* ptr_value2, void* ptr_keys_out, void* ptr_value_out, void* ptr_comp); /* api.h:186 */

/* ---- extensions: no counterpart in the reference ----
 * Batched segmented primitives (one launch per kernel stage for many
 * independent segments) and the graph-captured batch.  Callers: workloads
 * with many independent sorts / scans of uneven sizes. */
/* Sort each segment [offsets[s], offsets[s+1]) of keys/values independently
 * (offsets: host array, num_segments + 1 entries); comp NULL = Less. */
TRTC_CAPI int n_segmented_sort_by_key(void* ptr_keys, void* ptr_values, const unsigned long long* offsets,
                                      unsigned long long num_segments, void* ptr_comp);
/* Inclusive Plus scan of each segment independently. */
TRTC_CAPI int n_segmented_inclusive_scan(void* ptr_vec_in, void* ptr_vec_out, const unsigned long long* offsets,
                                         unsigned long long num_segments);
/* Keys-only segmented sort; exclusive Plus scan of each segment. */
TRTC_CAPI int n_segmented_sort(void* ptr_keys, const unsigned long long* offsets, unsigned long long num_segments,
                               void* ptr_comp);
TRTC_CAPI int n_segmented_exclusive_scan(void* ptr_vec_in, void* ptr_vec_out, const unsigned long long* offsets,
                                         unsigned long long num_segments);
/* results[s] = reduction of segment s with op (NULL = Plus); segments non-empty. */
TRTC_CAPI int n_segmented_reduce(void* ptr_vec_in, const unsigned long long* offsets, unsigned long long num_segments,
                                 void* ptr_results, void* ptr_op);
/* Tile-schedule autotuning of batched sorts (default on) and its JSON log. */
TRTC_CAPI void n_set_batch_autotune(int on);
/* Fixed schedule for every batched sort ("<=2048:256x8 rest:512x16"; NULL or
 * "" = automatic); -2 if the text is not a valid schedule. */
TRTC_CAPI int n_set_batch_schedule(const char* schedule);
TRTC_CAPI unsigned long long n_batch_tuning_report(char* json_out, unsigned long long capacity);
/* Graph-captured batch: record segmented operations, instantiate once (tune,
 * allocate, capture one HIP graph), launch many times (one hipGraphLaunch on
 * the calling thread's stream).  Recorded vectors must stay alive and in place. */
TRTC_CAPI void* n_batch_create(void);
TRTC_CAPI void n_batch_destroy(void* batch);
TRTC_CAPI int n_batch_add_sort(void* batch, void* ptr_keys, const unsigned long long* offsets,
                               unsigned long long num_segments, void* ptr_comp);
TRTC_CAPI int n_batch_add_sort_by_key(void* batch, void* ptr_keys, void* ptr_values, const unsigned long long* offsets,
                                      unsigned long long num_segments, void* ptr_comp);
TRTC_CAPI int n_batch_add_inclusive_scan(void* batch, void* ptr_vec_in, void* ptr_vec_out,
                                         const unsigned long long* offsets, unsigned long long num_segments,
                                         void* ptr_op);
TRTC_CAPI int n_batch_add_reduce(void* batch, void* ptr_vec_in, const unsigned long long* offsets,
                                 unsigned long long num_segments, void* ptr_results, void* ptr_op);
TRTC_CAPI int n_batch_instantiate(void* batch);
TRTC_CAPI int n_batch_launch(void* batch);
TRTC_CAPI unsigned long long n_batch_describe(void* batch, char* json_out, unsigned long long capacity);
/* Request server (include/server.h): HBM-budget admission and continuous
 * batching of mixed requests on `slots` streams; batching 0 = one request at
 * a time through the single-call path.  Ops: 0 Sort_By_Key, 1 Sort,
 * 2 Inclusive_Scan, 3 Reduce.  Submit returns a ticket or -1. */
TRTC_CAPI void* n_server_create(unsigned long long hbm_budget, int slots, int batching);
TRTC_CAPI void n_server_destroy(void* server);
// Batch limits of a batching server (TRTC_Server::Set_Batch_Limits); no counterpart in the reference.
TRTC_CAPI void n_server_set_batch_limits(void* server, unsigned long long max_requests, unsigned long long max_elements);
TRTC_CAPI long long n_server_submit(void* server, int op, const char* key_elem, const char* val_elem, void* in_a,
                                    void* in_b, unsigned long long n);
TRTC_CAPI int n_server_wait(void* server, long long ticket);
TRTC_CAPI int n_server_done(void* server, long long ticket);
TRTC_CAPI void* n_server_output(void* server, long long ticket, int which);
TRTC_CAPI double n_server_latency_us(void* server, long long ticket);
TRTC_CAPI void n_server_release(void* server, long long ticket);
TRTC_CAPI unsigned long long n_server_stats(void* server, char* json_out, unsigned long long capacity);
/* Load replay (TRTC_Server_Replay): returns the makespan in seconds. */
TRTC_CAPI double n_server_replay(void* server, unsigned long long count, const double* arrival_us, const int* ops,
                                 const char* const* key_elem, const char* const* val_elem, void* const* in_a,
                                 void* const* in_b, const unsigned long long* n, double* latency_us);
/* Stream used by the calling thread (a hipStream_t; NULL = null stream). */
TRTC_CAPI void n_set_stream(void* stream);
TRTC_CAPI void* n_get_stream(void);
/* Device pointer of a DVVector / DVVectorAdaptor (NULL for other kinds). */
TRTC_CAPI void* n_dvvectorlike_data(void* cptr);
/* Compile, without touching a device, the kernel template instantiation
 * `name_expr` (e.g. "trtc_copy_if_k<VectorView<int32_t>, ..., _S_...>") with
 * the definitions of every functor struct it names, for target `arch`
 * (e.g. "gfx950").  0 = compiled, -1 = compile error (log on stderr).
 * Callers: offline cache warm-up and build-time checks of functor code. */
TRTC_CAPI int n_jit_compile_check(const char* arch, const char* name_expr);
/* Code-object cache of run-time compiled kernels (in memory, and on disk
 * under $TRTC_CACHE_DIR/<arch>/<key>.co, else $XDG_CACHE_HOME/thrustrtc_amd
 * or ~/.cache/thrustrtc_amd): counts of in-memory hits, disk hits, hipRTC
 * compilations and disk entries rejected as damaged (out[0..3]).
 * Callers: cache tests and deployment checks that a warm cache is used. */
TRTC_CAPI void n_jit_cache_stats(unsigned long long* out4);
/* 1 if `blob` (size bytes) is a well-formed disk-cache entry (magic, code
 * length and checksum match), else 0; host only, needs no device.  The loader
 * sees only entries that pass this check. */
TRTC_CAPI int n_jit_cache_entry_valid(const char* blob, unsigned long long size);
/* No counterpart in the reference.  n_inclusive_scan (api.h:145) for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): same arguments, results
 * and return codes, but the scan's lookback state lives in a 1 MiB
 * per-GPU static device array, so the call allocates nothing (no hipMalloc,
 * no pool, no module load).  Limits: n <= 65535 * 16384 for elements of up to
 * 4 bytes, n <= 32767 * 16384 for 8-byte elements (larger inputs return -1
 * before anything is launched); the kernel must be built in (Plus / Maximum /
 * Minimum on plain vectors of built-in arithmetic types) or compiled by an
 * earlier n_inclusive_scan of the same views and functor, else -1.  Calls
 * that overlap on one GPU share the array and race: serialise them. */
TRTC_CAPI int n_inclusive_scan_noalloc(void* ptr_vec_in, void* ptr_vec_out, void* ptr_binary_op);
/* Kernel timing: between n_profile_begin and n_profile_end every kernel
 * launch is bracketed by HIP events on its stream; n_profile_end synchronises
 * and writes {"<kernel>": {"count": c, "total_ms": t}, ...} into json_out
 * (capacity bytes, NUL-terminated) and returns the full JSON length.
 * Callers: the benchmark's per-kernel roofline figures. */
TRTC_CAPI int n_profile_begin(void);
TRTC_CAPI unsigned long long n_profile_end(char* json_out, unsigned long long capacity);

#ifdef __cplusplus
}
#endif

#endif
