// Operator-specialised H*v kernels compiled at run time (see nv_jit.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "nv_kernels.hpp"
#include "nv_operator.hpp"

namespace nv {

struct JitKernels {
    hipModule_t module = nullptr;
    hipFunction_t iter = nullptr;  // nv_hv_iter: off-diagonal part of H q_i + alpha partials
    hipFunction_t full = nullptr;  // nv_hv_full: y = H x
    bool ready = false;
    bool failed = false;
};

std::string jit_source(const Operator &op, int prec);
bool jit_eligible(const Operator &op, int prec, std::string &why);
hipError_t jit_build(const Operator &op, int prec, JitKernels &out, std::string &log);
void jit_release(JitKernels &k);
// Launches nv_hv_iter (the caller adds the alpha finish).
hipError_t jit_launch_iter(const JitKernels &k, const NvOpDev &op, const LzState &st, hipStream_t s);
// y = H x; xin is x in the kernel's input precision (x itself for FP64).
hipError_t jit_launch_full(const JitKernels &k, const NvOpDev &op, const double *x, const void *xin, double *y,
                           hipStream_t s);

}  // namespace nv
