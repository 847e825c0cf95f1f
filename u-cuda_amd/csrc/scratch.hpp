// scratch.hpp — per-device address of the static device scratch used by
// uc_run_lle_noalloc (defined in scratch.hip). Internal to libucuda.so.
#ifndef UCUDA_SCRATCH_HPP
#define UCUDA_SCRATCH_HPP

#include <cstddef>

// 64 Ki doubles = 512 KiB per GPU: the tables of a map with nx + ny <= 65,440
// axis points (plus 64 parameter and 32 initial-state slots).
#define UC_SCRATCH_DOUBLES 65536

// With `device` current: looks up the scratch array's device address once
// per device. Returns 1 on success, else 0 with a message in msg.
__attribute__((visibility("hidden"))) int uc_scratch_setup(int device, char* msg, size_t cap);
// The address found by uc_scratch_setup, or nullptr.
__attribute__((visibility("hidden"))) void* uc_scratch_ptr(int device);

#endif
