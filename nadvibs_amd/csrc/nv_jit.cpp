// Operator-specialised H*v kernels, generated as HIP source and compiled
// at run time with hiprtc for gfx950.
//
// The generic kernel reads the term/step lists at run time.  Here the
// structure of one operator becomes straight-line code: the mode digits of a
// row come from constant divisors, every factor h(p, n_k, Delta) of every code
// is looked up once per thread into a register, each step is one address
// add, one load per state and one FMA per state, and every coupling matrix
// entry is a literal.  Loads go through buffer resources (one per state,
// num_records = dimen): a step that leaves the basis has a zero factor and an
// offset outside [0, dimen), which the hardware range check turns into a zero
// load, so the code has no branches.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "nv_jit.hpp"

namespace nv {

static std::string lit(double v, bool f32) {
    char buf[64];
    if (f32) {
        std::snprintf(buf, sizeof(buf), "%af", (float)v);
    } else {
        std::snprintf(buf, sizeof(buf), "%a", v);
    }
    return buf;
}

std::string jit_source(const Operator &op, int prec) {
    const bool f32 = prec != 0;
    const char *T = f32 ? "float" : "double";
    const int ns = (int)op.ns;
    const int64_t D = op.D;
    std::ostringstream s;
    // hiprtc provides the HIP device runtime itself (no includes)
    s << "typedef " << T << " T;\n";
    // fused multiply-add in the accumulation type (the double builtin would
    // promote FP32 operands and convert back around every step)
    s << (f32 ? "#define NV_FMA __builtin_fmaf\n" : "#define NV_FMA __builtin_fma\n");
    s << "#define NV_D " << D << "LL\n#define NV_N " << op.N << "LL\n";
    // factor tables of the codes used by the steps
    // code -> (mode, offset into TAB)
    std::map<int, std::pair<int, int>> codes;
    std::vector<double> tab;
    for (const auto &st : op.dev_steps)
        for (int j = 0; j < st.nf; ++j) {
            const int c = st.code[j];
            if (codes.count(c)) continue;
            const int mode = st.mode[j];
            codes[c] = {mode, (int)tab.size()};
            for (int64_t n = 0; n < op.nfunc[mode]; ++n) tab.push_back(op.ftab[(size_t)c * op.nmax + n]);
        }
    // inter-state diagonal part D_ss'(n), s != s' (FP64): constant matrix
    // plus the Delta=0 factors of the all-even monomials
    bool offstate = false;
    for (int a = 0; a < ns; ++a)
        for (int c = 0; c < ns; ++c)
            if (a != c) {
                if (op.dconst[a * ns + c] != 0.0) offstate = true;
                for (size_t t = 0; t < op.dev_dterms.size(); ++t)
                    if (op.dcmat[t * ns * ns + a * ns + c] != 0.0) offstate = true;
            }
    std::map<int, std::pair<int, int>> dcodes;  // code -> (mode, offset into DTAB)
    std::vector<double> dtab;
    if (offstate)
        for (const auto &dt : op.dev_dterms)
            for (int j = 0; j < dt.nf; ++j) {
                const int c = dt.code[j];
                if (dcodes.count(c)) continue;
                dcodes[c] = {dt.mode[j], (int)dtab.size()};
                for (int64_t n = 0; n < op.nfunc[dt.mode[j]]; ++n) dtab.push_back(op.ftab[(size_t)c * op.nmax + n]);
            }
    s << "__constant__ double NV_DTAB[" << (dtab.empty() ? 1 : dtab.size()) << "] = {";
    for (size_t i = 0; i < dtab.size(); ++i) s << (i ? "," : "") << lit(dtab[i], false);
    if (dtab.empty()) s << "0";
    s << "};\n";
    s << "__constant__ T NV_TAB[" << (tab.empty() ? 1 : tab.size()) << "] = {";
    for (size_t i = 0; i < tab.size(); ++i) s << (i ? "," : "") << lit(tab[i], f32);
    if (tab.empty()) s << "0";
    s << "};\n";
    // helpers
    s << R"(
// workgroup b runs on XCD b % 8: give each XCD one contiguous range of rows
__device__ __forceinline__ int nv_xcd_block(int b, int g) {
    const int x = b & 7, j = b >> 3, q = g >> 3, rem = g & 7;
    return x * q + (x < rem ? x : rem) + j;
}
__device__ __forceinline__ double nv_block_sum(double v) {
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}
)";
    const int nsp = nv_lo_stride(ns);
    if (f32) {
        // the FP32 copy is state-interleaved, [D][nsp]: one vector load per step
        s << "typedef float nv_f2 __attribute__((ext_vector_type(2)));\n"
             "typedef float nv_f4 __attribute__((ext_vector_type(4)));\n"
             "__device__ __forceinline__ float nv_ld1(__amdgpu_buffer_rsrc_t r, unsigned o) { return "
             "__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)o, 0, 0)); }\n"
             "__device__ __forceinline__ nv_f2 nv_ld2(__amdgpu_buffer_rsrc_t r, unsigned o) { return "
             "__builtin_bit_cast(nv_f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)o, 0, 0)); }\n"
             "__device__ __forceinline__ nv_f4 nv_ld4(__amdgpu_buffer_rsrc_t r, unsigned o) { return "
             "__builtin_bit_cast(nv_f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)o, 0, 0)); }\n"
             "typedef float nv_f3 __attribute__((ext_vector_type(3)));\n"
             "__device__ __forceinline__ nv_f3 nv_ld3(__amdgpu_buffer_rsrc_t r, unsigned o) { return "
             "__builtin_bit_cast(nv_f3, __builtin_amdgcn_raw_buffer_load_b96(r, (int)o, 0, 0)); }\n";
    } else {
        s << "__device__ __forceinline__ T nv_ld(__amdgpu_buffer_rsrc_t r, unsigned o) { return "
             "__builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)o, 0, 0)); }\n";
    }
    // the body shared by both kernels: needs `x` (const T*), `r`, writes acc_s
    std::ostringstream b;
    b << "    __shared__ T tab[" << (tab.empty() ? 1 : tab.size()) << "];\n";
    b << "    for (int i = threadIdx.x; i < " << tab.size() << "; i += 256) tab[i] = NV_TAB[i];\n";
    b << "    __syncthreads();\n";
    b << "    const int r = nv_xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;\n";
    b << "    const bool live = r < NV_D;\n";
    b << "    unsigned t = live ? (unsigned)r : 0u;\n";
    for (int64_t k = op.M - 1; k >= 0; --k) {
        b << "    const unsigned d" << k << " = t % " << op.nfunc[k] << "u; t /= " << op.nfunc[k] << "u;\n";
    }
    for (auto &kv : codes)
        b << "    const T f" << kv.first << " = tab[" << kv.second.second << " + d" << kv.second.first << "];\n";
    if (f32) {
        b << "    const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, (int)(NV_D * "
          << nsp * 4 << "), 0x00020000);\n";
        b << "    const unsigned vo = (unsigned)r * " << nsp * 4 << "u;\n";
    } else {
        for (int q = 0; q < ns; ++q)
            b << "    const __amdgpu_buffer_rsrc_t rs" << q << " = __builtin_amdgcn_make_buffer_rsrc((void *)(x + "
              << q << "LL * NV_D), (short)0, (int)(NV_D * sizeof(T)), 0x00020000);\n";
        b << "    const unsigned vo = (unsigned)r * (unsigned)sizeof(T);\n";
    }
    for (int q = 0; q < ns; ++q) b << "    T acc" << q << " = 0;\n";
    const size_t el = f32 ? (size_t)nsp * 4 : 8;
    // the loads of one step: w_q += h * x[q][r + stride]
    auto step_loads = [&](std::ostringstream &o) {
        if (!f32) {
            for (int q = 0; q < ns; ++q) o << " w" << q << " = NV_FMA(h, nv_ld(rs" << q << ", o), w" << q << ");";
            return;
        }
        if (nsp == 1) {
            o << " w0 = NV_FMA(h, nv_ld1(rx, o), w0);";
        } else if (nsp == 2) {
            o << " { const nv_f2 v = nv_ld2(rx, o);";
            for (int q = 0; q < ns; ++q) o << " w" << q << " = NV_FMA(h, v[" << q << "], w" << q << ");";
            o << " }";
        } else if (nsp == 3) {
            o << " { const nv_f3 v = nv_ld3(rx, o);";
            for (int q = 0; q < ns; ++q) o << " w" << q << " = NV_FMA(h, v[" << q << "], w" << q << ");";
            o << " }";
        } else {
            for (int part = 0; part < nsp / 4; ++part) {
                if (part * 4 >= ns) break;
                o << " { const nv_f4 v = nv_ld4(rx, o + " << part * 16 << "u);";
                for (int q = part * 4; q < ns && q < part * 4 + 4; ++q)
                    o << " w" << q << " = NV_FMA(h, v[" << q - part * 4 << "], w" << q << ");";
                o << " }";
            }
        }
    };
    for (size_t ti = 0; ti < op.dev_terms.size(); ++ti) {
        const NvTerm &tm = op.dev_terms[ti];
        b << "    {\n";
        for (int q = 0; q < ns; ++q) b << "        T w" << q << " = 0;\n";
        for (int k = 0; k < tm.nsteps; ++k) {
            const NvStep &st = op.dev_steps[tm.first + k];
            b << "        { const unsigned o = vo + " << (uint32_t)((int64_t)st.stride * (int64_t)el) << "u; const T h = ";
            for (int j = 0; j < st.nf; ++j) b << (j ? " * " : "") << "f" << st.code[j];
            b << ";";
            step_loads(b);
            b << " }\n";
        }
        const double *C = op.cmat.data() + ti * ns * ns;
        for (int a = 0; a < ns; ++a)
            for (int c = 0; c < ns; ++c)
                if (C[a * ns + c] != 0.0)
                    b << "        acc" << a << " = NV_FMA((T)" << lit(C[a * ns + c], f32) << ", w" << c
                      << ", acc" << a << ");\n";
        b << "    }\n";
    }
    // ex_s = sum_{s' != s} D_ss'(n) v_s' (v_s' given as double expressions)
    auto offstate_code = [&](const char *vname) {
        std::ostringstream o;
        for (int a = 0; a < ns; ++a) o << "        double ex" << a << " = 0.0;\n";
        if (!offstate) return o.str();
        for (int a = 0; a < ns; ++a)
            for (int c = 0; c < ns; ++c)
                if (a != c && op.dconst[a * ns + c] != 0.0)
                    o << "        ex" << a << " += " << lit(op.dconst[a * ns + c], false) << " * " << vname << c << ";\n";
        for (size_t t = 0; t < op.dev_dterms.size(); ++t) {
            const NvDiagTerm &dt = op.dev_dterms[t];
            bool any = false;
            for (int a = 0; a < ns; ++a)
                for (int c = 0; c < ns; ++c)
                    if (a != c && op.dcmat[t * ns * ns + a * ns + c] != 0.0) any = true;
            if (!any) continue;
            o << "        { const double h = 1.0";
            for (int j = 0; j < dt.nf; ++j)
                o << " * NV_DTAB[" << dcodes[dt.code[j]].second << " + d" << dt.mode[j] << "]";
            o << ";\n";
            for (int a = 0; a < ns; ++a)
                for (int c = 0; c < ns; ++c) {
                    const double v = op.dcmat[t * ns * ns + a * ns + c];
                    if (a != c && v != 0.0) o << "          ex" << a << " += h * " << lit(v, false) << " * " << vname << c << ";\n";
                }
            o << "        }\n";
        }
        return o.str();
    };
    // iteration kernel: x = vlo (F32) or q (F64); y = off-diagonal part; alpha partials
    const char *sfx = f32 ? "_f32" : "_f64";
    s << "extern \"C\" __global__ __launch_bounds__(256) void nv_hv_iter" << sfx
      << "(const double *V, long long nslots, const T *vlo, T *y, const double *dg, double *partials, "
         "const long long *iter) {\n";
    s << "    const long long it = *iter;\n";
    s << "    const double *q = V + (it % nslots) * NV_N;\n";
    s << (f32 ? "    const T *x = vlo;\n" : "    const T *x = q;\n");
    s << b.str();
    s << "    double part = 0.0;\n    if (live) {\n";
    for (int q = 0; q < ns; ++q) s << "        const double v" << q << " = q[" << q << "LL * NV_D + r];\n";
    s << offstate_code("v");
    for (int q = 0; q < ns; ++q) {
        s << "        y[" << q << "LL * NV_D + r] = acc" << q << ";\n";
        s << "        part += v" << q << " * (dg[" << q << "LL * NV_D + r] * v" << q << " + ex" << q << " + (double)acc"
          << q << ");\n";
    }
    s << "    }\n    const double bsum = nv_block_sum(part);\n";
    s << "    if (threadIdx.x == 0) partials[blockIdx.x] = bsum;\n}\n";
    // standalone kernel: x double in, y = H x double out
    s << "extern \"C\" __global__ __launch_bounds__(256) void nv_hv_full" << sfx
      << "(const double *xd, const T *x, double *yd, const double *dg) {\n";
    s << b.str();
    s << "    if (live) {\n";
    for (int q = 0; q < ns; ++q) s << "        const double v" << q << " = xd[" << q << "LL * NV_D + r];\n";
    s << offstate_code("v");
    for (int q = 0; q < ns; ++q)
        s << "        yd[" << q << "LL * NV_D + r] = dg[" << q << "LL * NV_D + r] * v" << q << " + ex" << q
          << " + (double)acc" << q << ";\n";
    s << "    }\n}\n";
    return s.str();
}

bool jit_eligible(const Operator &op, int prec, std::string &why) {
    if (prec != 0 && prec != 1) { why = "precision"; return false; }
    if (op.dev_steps.size() > 1200) { why = "operator too large for straight-line code"; return false; }
    if (op.N * 8 >= (int64_t(1) << 31)) { why = "vector beyond buffer range"; return false; }
    return true;
}

hipError_t jit_build(const Operator &op, int prec, JitKernels &out, std::string &log) {
    const std::string src = jit_source(op, prec);
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "nv_hv_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return hipErrorInvalidValue;
    }
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops"};
    hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    size_t lsz = 0;
    hiprtcGetProgramLogSize(prog, &lsz);
    if (lsz > 1) {
        log.resize(lsz);
        hiprtcGetProgramLog(prog, &log[0]);
    }
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return hipErrorInvalidValue;
    }
    size_t csz = 0;
    hiprtcGetCodeSize(prog, &csz);
    std::vector<char> code(csz);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    hipError_t e = hipModuleLoadData(&out.module, code.data());
    if (e != hipSuccess) return e;
    const std::string sfx = prec != 0 ? "_f32" : "_f64";
    e = hipModuleGetFunction(&out.iter, out.module, ("nv_hv_iter" + sfx).c_str());
    if (e == hipSuccess) e = hipModuleGetFunction(&out.full, out.module, ("nv_hv_full" + sfx).c_str());
    out.ready = e == hipSuccess;
    return e;
}

void jit_release(JitKernels &k) {
    if (k.module) (void)hipModuleUnload(k.module);
    k = JitKernels{};
}

hipError_t jit_launch_iter(const JitKernels &k, const NvOpDev &op, const LzState &st, hipStream_t s) {
    const double *V = st.V;
    long long nslots = st.nslots;
    const void *vlo = st.vlo;
    void *y = st.y;
    const double *dg = op.dg;
    double *partials = st.partials;
    const int64_t *iter = st.iter;
    void *args[] = {(void *)&V, (void *)&nslots, (void *)&vlo, (void *)&y, (void *)&dg, (void *)&partials,
                    (void *)&iter};
    const unsigned g = (unsigned)((op.D + 255) / 256);
    return hipModuleLaunchKernel(k.iter, g, 1, 1, 256, 1, 1, 0, s, args, nullptr);
}

hipError_t jit_launch_full(const JitKernels &k, const NvOpDev &op, const double *x, const void *xin, double *y,
                           hipStream_t s) {
    const double *dg = op.dg;
    void *args[] = {(void *)&x, (void *)&xin, (void *)&y, (void *)&dg};
    const unsigned g = (unsigned)((op.D + 255) / 256);
    return hipModuleLaunchKernel(k.full, g, 1, 1, 256, 1, 1, 0, s, args, nullptr);
}

}  // namespace nv
