// sah.cpp -- host binned SAH builder (binned_sah_builder, detail/bvh/sah.h:176-807).
#include "vr_internal.h"

namespace vr {

int sah_build_host(const vr_triangle*, size_t, int, vr_bvh_node*, size_t*, uint32_t*)
{
    return fail(VR_ERR_UNSUPPORTED, "vr_sah_build: not available in this build");
}

}  // namespace vr
