// internal.h -- shared by the library's translation units (not part of the C ABI).
#ifndef QTENSOR_AMD_INTERNAL_H
#define QTENSOR_AMD_INTERNAL_H

namespace qb_internal {
// Record `msg` as this thread's qb_last_error() and return `code`.
int fail(int code, const char* msg);
}  // namespace qb_internal

#endif
