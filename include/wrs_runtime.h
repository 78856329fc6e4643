/*
 * wrs_runtime.h -- device-memory, stream and event plumbing exported by the
 * same library, so that a host with no HIP headers (the Python tests and the
 * benchmark drive the library through ctypes) can allocate device buffers,
 * move data and time launches.  Thin wrappers over the HIP runtime; none of
 * them has a counterpart in the reference, whose buffers come from the merian
 * Vulkan allocator (src/host/layout/BufferView.hpp:57-291).
 */
#ifndef WRS_RUNTIME_H
#define WRS_RUNTIME_H

#include <stddef.h>
#include <stdint.h>

#include "wrs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipEvent_t* wrs_event_t;

wrs_status wrs_rt_device_count(int* count);
/* name (up to len bytes), gcnArchName prefix, CU count, HBM bytes */
wrs_status wrs_rt_device_info(char* name, size_t len, char* arch, size_t arch_len,
                              int* cu_count, uint64_t* hbm_bytes);
wrs_status wrs_rt_malloc(void** ptr, size_t bytes);
wrs_status wrs_rt_free(void* ptr);
wrs_status wrs_rt_memcpy_htod(void* dst, const void* src, size_t bytes);
wrs_status wrs_rt_memcpy_dtoh(void* dst, const void* src, size_t bytes);
wrs_status wrs_rt_memcpy_dtod(void* dst, const void* src, size_t bytes, wrs_stream_t stream);
wrs_status wrs_rt_memset(void* dst, int value, size_t bytes, wrs_stream_t stream);
wrs_status wrs_rt_stream_create(wrs_stream_t* stream);
wrs_status wrs_rt_stream_destroy(wrs_stream_t stream);
wrs_status wrs_rt_stream_sync(wrs_stream_t stream);
wrs_status wrs_rt_device_sync(void);
wrs_status wrs_rt_event_create(wrs_event_t* ev);
wrs_status wrs_rt_event_destroy(wrs_event_t ev);
wrs_status wrs_rt_event_record(wrs_event_t ev, wrs_stream_t stream);
wrs_status wrs_rt_event_elapsed(float* ms, wrs_event_t start, wrs_event_t stop);
/* copy kernel of known size (bandwidth calibration): dst[i] = src[i], n floats */
wrs_status wrs_rt_copy_f32(float* dst, const float* src, uint64_t n, wrs_stream_t stream);

/* Baseline of the batched build: each segment built and sampled by its own
 * wrs_psa_build + wrs_alias_sample launches, issued from a host loop (what a
 * caller without the segmented path would write).  Same counters as
 * wrs_alias_sample_batched, so with identical tables the samples are identical. */
wrs_status wrs_rt_psa_build_sample_each(wrs_plan* plan, const float* weights,
                                        const uint64_t* offsets_host, uint32_t nseg,
                                        wrs_alias_entry* tables, uint32_t per_segment,
                                        uint64_t seed, uint64_t counter_base, uint32_t* samples,
                                        wrs_stream_t stream);

/* graph capture of everything issued on `stream` between begin and end */
typedef struct wrs_graph wrs_graph;
wrs_status wrs_rt_graph_begin(wrs_stream_t stream);
wrs_status wrs_rt_graph_end(wrs_stream_t stream, wrs_graph** graph);
wrs_status wrs_rt_graph_launch(wrs_graph* graph, wrs_stream_t stream);
wrs_status wrs_rt_graph_destroy(wrs_graph* graph);

#ifdef __cplusplus
}
#endif
#endif /* WRS_RUNTIME_H */
