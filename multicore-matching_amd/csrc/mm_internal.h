// Internal helpers shared by the host sources of the matcher library.
#ifndef MM_INTERNAL_H
#define MM_INTERNAL_H

#include <cstdint>
#include <functional>

namespace mm {

// Stores msg as the calling thread's last error and returns code.
int32_t set_error(int32_t code, const char *msg);
int32_t set_errorf(int32_t code, const char *fmt, ...);
void clear_error();

// Static partition of [0, total) over host worker threads (set_host_threads,
// else MM_GEN_THREADS, else all hardware threads); body(begin, end, worker).
void parallel_for(int64_t total, const std::function<void(int64_t, int64_t, unsigned)> &body);
// As parallel_for, with chunks of `grain` items handed out on demand (for
// loops whose cost per item is skewed, such as rows of a power-law graph).
void parallel_for_dynamic(int64_t total, int64_t grain,
                          const std::function<void(int64_t, int64_t, unsigned)> &body);
// Worker count of parallel_for; 0 restores the default.
void set_host_threads(unsigned n);

}  // namespace mm

#endif
