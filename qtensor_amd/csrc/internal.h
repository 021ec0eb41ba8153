// internal.h -- shared by the library's translation units (not part of the C ABI).
#ifndef QTENSOR_AMD_INTERNAL_H
#define QTENSOR_AMD_INTERNAL_H

#include <stdint.h>

struct qb_exec_desc;

namespace qb_internal {
// Record `msg` as this thread's qb_last_error() and return `code`.
int fail(int code, const char* msg);
// The executor launch of qb_execute with an explicit workspace: `scratch`
// holds `n_slots_avail` slots of d->scratch_elems elements, `queue` one uint32.
int execute_impl(const qb_exec_desc* d, void* scratch, int32_t n_slots_avail, uint32_t* queue, void* stream);
// Items of d's launch shape resident at once (the scratch slots a launch uses).
int resident_items(const qb_exec_desc* d, int32_t* n_slots);
// Device selected by qb_init (-1 before).
int current_device();
// Resolve qb_execute_noalloc's static workspace on `device` (current), once.
int noalloc_init(int device);
}  // namespace qb_internal

#endif
