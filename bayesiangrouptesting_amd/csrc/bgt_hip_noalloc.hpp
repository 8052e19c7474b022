// Selection on a caller-owned posterior with no device allocation (bgt_hip_noalloc.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "bgt_hip.h"

#include <cstddef>

namespace bgt_hip {

// Largest lattice the static scratch array serves: 2^20 states.
constexpr int kNoAllocMaxSubjects = BGT_HIP_NOALLOC_MAX_SUBJECTS;
// Bytes of the static scratch array each GPU holds (header plus the FP64 transform
// intermediate of 2^20 states).
constexpr size_t kNoAllocScratchBytes = BGT_HIP_NOALLOC_SCRATCH_BYTES;

// Op-BHA or binary Op-BBPA on `posterior` (2^subjects device entries, FP64 or FP32 per
// `precision`, each read as entry * scale), results into device memory; asynchronous on
// `stream`. Throws Error with a provider status code.
void select_noalloc(const void *posterior, int subjects, int precision, int selector, double scale,
					bgt_hip_selection *result, double *masses, hipStream_t stream);

} // namespace bgt_hip
