// Test infrastructure only: the emulation build has no rocPRIM, so the device BVH
// builder is left out and reports GP_ERR_UNSUPPORTED.
#include "gp_bvh_device.hpp"
#include "gp_runtime.hpp"
namespace gp {
gp_status buildBvhDevice(const float4*, const float4*, int64_t, const float*, const float*, DeviceBvh&) {
    return fail(GP_ERR_UNSUPPORTED, "device BVH not emulated");
}
gp_status deviceMinDistance2(const float4*, int64_t, const double*, double&) {
    return fail(GP_ERR_UNSUPPORTED, "device BVH not emulated");
}
}  // namespace gp
