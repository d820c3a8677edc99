// groupby.hip -- sort-based group-by for gfx950 (placeholder until the
// segmented-reduction kernels land).
#include "vdf_internal.h"

namespace vdf {

size_t groupby_scratch(int key_type, uint64_t n) { return argsort_scratch(key_type, n, 0); }

void groupby_enqueue(Handle*, int, const void*, uint64_t, const vdf_value_col*, int, uint32_t,
                     const vdf_groupby_out*, void*, hipStream_t)
{
    throw Error{VDF_ERR_UNSUPPORTED, "vdf_groupby: not built yet"};
}

}  // namespace vdf
