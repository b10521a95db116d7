// Franck-Condon start vector (generate_initialvec with shiftref != 0,
// Source/secondary.f90:351-725).
#include <hip/hip_runtime.h>

#include "nv_pro.hpp"

namespace nv {

hipError_t build_fc_start(const Operator &op, const NvOpDev &dev, double *dst, hipStream_t s, std::string &err) {
    (void)op; (void)dev; (void)dst; (void)s;
    err = "Franck-Condon start vector not available in this build";
    return hipErrorNotSupported;
}

}  // namespace nv
