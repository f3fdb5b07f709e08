// The no-allocation count entry point (noalloc.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/gpujoin.h"

namespace cgj {

class Context;
struct DevTable;

// Hash-table bytes of the fixed per-GPU scratch array (key-only slots at load
// <= 0.5: build sides up to 4M rows).
constexpr uint64_t kNoAllocTableBytes = uint64_t(CGJ_NOALLOC_TABLE_BYTES);

// Device address of this GPU's scratch array (looked up once per device).
void* noalloc_scratch(int device);

// Rows of the join of the plan's two scans on their one shared key, computed
// in the scratch array (no device allocation).
uint64_t count_join_noalloc(Context& ctx, const cgj_plan& plan, const std::vector<const DevTable*>& inputs);

} // namespace cgj
