// re_cl_algorithm_base.hpp -- device base of the algorithms.
//
// The reference's cl_base<FLT> (resaxslib/include/re_cl_algorithm_base.hpp:31-80)
// owns an OpenCL context and queue.  Here hip_base<FLT> selects the HIP device an
// engine will run on; cl_base<FLT> names the same type so callers that spell
// `algorithm::cl_base<float>` compile unchanged.
#ifndef RESAXS_AMD_RE_CL_ALGORITHM_BASE_HPP
#define RESAXS_AMD_RE_CL_ALGORITHM_BASE_HPP

#include <string>
#include <vector>

#include "re_algorithm.hpp"

namespace resaxs {
namespace algorithm {

template <typename FLT_T>
class hip_base : public algorithm {
public:
    // Picks the device from the first dev_id (an empty list means device 0): platform
    // must be 0 or -1, device a HIP ordinal or -1; a CPU device is SAXS_INVALID_DEVICE.
    void initialize(const std::vector<dev_id>& ids)
    {
        int count = 0;
        rsx_device_count(&count);
        long dev = 0;
        if (!ids.empty()) {
            const dev_id& d = ids.front();
            verify(d.type != CPU, error::SAXS_INVALID_DEVICE, "RESAXS::error -- no CPU device: the engine runs on HIP GPUs");
            verify(d.platform == -1 || d.platform == 0, error::SAXS_INVALID_DEVICE,
                   "RESAXS::error -- invalid platform " + std::to_string(d.platform));
            dev = d.device < 0 ? 0 : d.device;
        }
        verify(dev < count, error::SAXS_INVALID_DEVICE, "RESAXS::error -- no HIP device " + std::to_string(dev));
        device_ = (int)dev;
        this->state = this->alg_initialized;
    }

    int device() const { return device_; }

protected:
    int device_ = 0;
};

template <typename FLT_T>
using cl_base = hip_base<FLT_T>;

}  // namespace algorithm
}  // namespace resaxs

#endif
