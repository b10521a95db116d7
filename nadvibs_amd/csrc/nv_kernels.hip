// HIP kernels of the Lanczos path for gfx950 (64-wide wavefronts).
//
//  * hv_*          matrix-free H*v over the state-major tensor x[s][n_1..n_M]
//                  (last mode fastest): one thread per vibrational index r
//                  computes all states; the mode digits of r and the factor
//                  table sit in LDS; coupling matrices are read as uniform
//                  (scalar) loads.
//  * normalize / lanczos_step   the three-term recurrence in FP64 with the
//                  dots reduced to per-block partials and summed by a
//                  one-block finish kernel (fixed summation order:
//                  reproducible run to run).
//  * tri_*         Sturm bisection and inverse iteration for the tridiagonal
//                  eigenproblem, one root per thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hip/hip_bf16.h>

#include <type_traits>

#include "nv_kernels.hpp"

namespace nv {

// Workgroups of one launch go to the 8 XCDs round-robin (workgroup b to XCD
// b % 8).  H*v gathers x at +-stride of every mode, so a workgroup reads the
// rows of its neighbours along the slow modes; mapping each XCD onto one
// contiguous range of rows keeps that window inside the XCD's own L2.
__device__ __forceinline__ int xcd_block(int b, int g) {
    const int x = b & 7, j = b >> 3, q = g >> 3, rem = g & 7;
    return x * q + (x < rem ? x : rem) + j;
}

// ------------------------------------------------------------------------
// reductions
// ------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum; every thread gets the result.  blockDim.x <= 1024.
__device__ double block_sum(double v) {
    __shared__ double red[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[i];
    return t;
}

// ------------------------------------------------------------------------
// element types
// ------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ float to_f(T v);
template <> __device__ __forceinline__ float to_f<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f<__hip_bfloat16>(__hip_bfloat16 v) { return __bfloat162float(v); }

template <typename TIn, typename TAcc> __device__ __forceinline__ TAcc load_acc(const TIn *p);
template <> __device__ __forceinline__ double load_acc<double, double>(const double *p) { return *p; }
template <> __device__ __forceinline__ float load_acc<float, float>(const float *p) { return *p; }
template <> __device__ __forceinline__ float load_acc<__hip_bfloat16, float>(const __hip_bfloat16 *p) {
    return __bfloat162float(*p);
}
template <> __device__ __forceinline__ float load_acc<double, float>(const double *p) { return (float)*p; }

template <typename T> __device__ __forceinline__ T from_d(double v);
template <> __device__ __forceinline__ double from_d<double>(double v) { return v; }
template <> __device__ __forceinline__ float from_d<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ __hip_bfloat16 from_d<__hip_bfloat16>(double v) {
    return __float2bfloat16((float)v);
}

template <typename T> __device__ __forceinline__ const T *ftab_of(const NvOpDev &op);
template <> __device__ __forceinline__ const double *ftab_of<double>(const NvOpDev &op) { return op.ftab64; }
template <> __device__ __forceinline__ const float *ftab_of<float>(const NvOpDev &op) { return op.ftab32; }
template <typename T> __device__ __forceinline__ const T *cmat_of(const NvOpDev &op);
template <> __device__ __forceinline__ const double *cmat_of<double>(const NvOpDev &op) { return op.cmat64; }
template <> __device__ __forceinline__ const float *cmat_of<float>(const NvOpDev &op) { return op.cmat32; }

// ------------------------------------------------------------------------
// mode digits and the diagonal part
// ------------------------------------------------------------------------
// Writes the digits n_k of r into dig[k*kBlock + tid]; returns E(r).
__device__ __forceinline__ double digits_energy(const NvOpDev &op, int r, uint8_t *dig, int tid) {
    unsigned rr = (unsigned)r;
    double e = 0.0;
    for (int k = op.M - 1; k >= 0; --k) {
        const unsigned nf = (unsigned)op.nfunc[k];
        const unsigned q = rr / nf;
        const unsigned d = rr - q * nf;
        rr = q;
        dig[k * kBlock + tid] = (uint8_t)d;
        e += ((double)d + 0.5) * op.omega[k];
    }
    return e;
}

// The vibrationally diagonal part of H (Source/top.f90:712-756) splits into
// the state-diagonal E(n) + D_ss(n), precomputed once into op.dg, and the
// inter-state part D_ss'(n), s != s', evaluated per row when present.
template <int NS>
__device__ __forceinline__ void diag_offstate(const NvOpDev &op, const uint8_t *dig, int tid,
                                              const double (&v)[NS], double (&out)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double a = 0.0;
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2)
            if (s2 != s) a += op.dconst[s * NS + s2] * v[s2];
        out[s] += a;
    }
    for (int t = 0; t < op.ndterms; ++t) {
        const NvDiagTerm dt = op.dterms[t];
        double h = 1.0;
#pragma unroll
        for (int j = 0; j < NV_MAXF; ++j)
            if (j < dt.nf) h *= op.ftab64[dt.code[j] * op.nmax + dig[dt.mode[j] * kBlock + tid]];
        const double *C = op.dcmat + (size_t)t * NS * NS;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double a = 0.0;
#pragma unroll
            for (int s2 = 0; s2 < NS; ++s2)
                if (s2 != s) a += C[s * NS + s2] * v[s2];
            out[s] += h * a;
        }
    }
}

// out[s] = sum_s' (E + D)(n)_{ss'} v[s'] in FP64.  dig is only read when
// states are coupled through D.
template <int NS>
__device__ __forceinline__ void diag_apply(const NvOpDev &op, int r, const uint8_t *dig, int tid,
                                           const double (&v)[NS], double (&out)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; ++s) out[s] = op.dg[(size_t)s * op.D + r] * v[s];
    if (op.offstate) diag_offstate<NS>(op, dig, tid, v, out);
}

// dg[s][r] = E(r) + D_ss(r): harmonic energy, constant term and the Delta=0
// part of the all-even monomials of the diagonal blocks.
template <int NS>
__global__ __launch_bounds__(kBlock) void diag_precompute_kernel(NvOpDev op, double *dg) {
    __shared__ uint8_t dig[NV_MAXM * kBlock];
    const int tid = threadIdx.x;
    const int r = blockIdx.x * kBlock + tid;
    if (r >= op.D) return;
    const double E = digits_energy(op, r, dig, tid);
    double d[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) d[s] = E + op.dconst[s * NS + s];
    for (int t = 0; t < op.ndterms; ++t) {
        const NvDiagTerm dt = op.dterms[t];
        double h = 1.0;
#pragma unroll
        for (int j = 0; j < NV_MAXF; ++j)
            if (j < dt.nf) h *= op.ftab64[dt.code[j] * op.nmax + dig[dt.mode[j] * kBlock + tid]];
        const double *C = op.dcmat + (size_t)t * NS * NS;
#pragma unroll
        for (int s = 0; s < NS; ++s) d[s] += h * C[s * NS + s];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) dg[(size_t)s * op.D + r] = d[s];
}

// ------------------------------------------------------------------------
// H*v: off-diagonal (in vibrational space) terms, Source/top.f90:758-808.
// acc[s] = sum_t sum_s' C_t[s][s'] sum_{steps} prod_j f_j(n) x[s'][r+stride]
// ------------------------------------------------------------------------
template <int NS, typename TIn, typename TAcc>
__device__ __forceinline__ void offdiag_accumulate(const NvOpDev &op, const TIn *__restrict__ x, int r,
                                                   const uint8_t *dig, const TAcc *ftab, int tid,
                                                   TAcc (&acc)[NS]) {
    const TAcc *__restrict__ cm = cmat_of<TAcc>(op);
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = (TAcc)0;
    for (int t = 0; t < op.nterms; ++t) {
        const NvTerm tm = op.terms[t];
        TAcc w[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) w[s] = (TAcc)0;
        for (int q = 0; q < tm.nsteps; ++q) {
            const NvStep st = op.steps[tm.first + q];
            TAcc h = ftab[st.code[0] * op.nmax + dig[st.mode[0] * kBlock + tid]];
#pragma unroll
            for (int j = 1; j < NV_MAXF; ++j)
                if (j < st.nf) h *= ftab[st.code[j] * op.nmax + dig[st.mode[j] * kBlock + tid]];
            if (h != (TAcc)0) {
                const int idx = r + st.stride;
                if constexpr (std::is_same<TIn, double>::value) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) w[s] += h * load_acc<TIn, TAcc>(x + (size_t)s * op.D + idx);
                } else {
                    // interleaved low-precision copy: the states of a row are adjacent
                    const TIn *xr = x + (size_t)idx * nv_lo_stride(NS);
#pragma unroll
                    for (int s = 0; s < NS; ++s) w[s] += h * load_acc<TIn, TAcc>(xr + s);
                }
            }
        }
        const TAcc *C = cm + (size_t)t * NS * NS;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            TAcc a = acc[s];
#pragma unroll
            for (int s2 = 0; s2 < NS; ++s2) a += C[s * NS + s2] * w[s2];
            acc[s] = a;
        }
    }
}

// The same for two rows of one thread (r0 and r1, digits at dig0 and dig1): the step descriptors are read once for both, and the two rows'
// table and vector loads are independent, so they overlap.  A row that is
// not live (past the end) has zero digits and is skipped.
template <int NS, typename TIn, typename TAcc>
__device__ __forceinline__ void offdiag_accumulate2(const NvOpDev &op, const TIn *__restrict__ x, int r0, int r1,
                                                    bool l0, bool l1, const uint8_t *dig0, const uint8_t *dig1,
                                                    const TAcc *ftab, int tid, TAcc (&acc0)[NS], TAcc (&acc1)[NS]) {
    const TAcc *__restrict__ cm = cmat_of<TAcc>(op);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        acc0[s] = (TAcc)0;
        acc1[s] = (TAcc)0;
    }
    for (int t = 0; t < op.nterms; ++t) {
        const NvTerm tm = op.terms[t];
        TAcc w0[NS], w1[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            w0[s] = (TAcc)0;
            w1[s] = (TAcc)0;
        }
        for (int q = 0; q < tm.nsteps; ++q) {
            const NvStep st = op.steps[tm.first + q];
            TAcc h0 = ftab[st.code[0] * op.nmax + dig0[st.mode[0] * kBlock + tid]];
            TAcc h1 = ftab[st.code[0] * op.nmax + dig1[st.mode[0] * kBlock + tid]];
#pragma unroll
            for (int j = 1; j < NV_MAXF; ++j)
                if (j < st.nf) {
                    h0 *= ftab[st.code[j] * op.nmax + dig0[st.mode[j] * kBlock + tid]];
                    h1 *= ftab[st.code[j] * op.nmax + dig1[st.mode[j] * kBlock + tid]];
                }
            if (l0 && h0 != (TAcc)0) {
                const int idx = r0 + st.stride;
                if constexpr (std::is_same<TIn, double>::value) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) w0[s] += h0 * load_acc<TIn, TAcc>(x + (size_t)s * op.D + idx);
                } else {
                    const TIn *xr = x + (size_t)idx * nv_lo_stride(NS);
#pragma unroll
                    for (int s = 0; s < NS; ++s) w0[s] += h0 * load_acc<TIn, TAcc>(xr + s);
                }
            }
            if (l1 && h1 != (TAcc)0) {
                const int idx = r1 + st.stride;
                if constexpr (std::is_same<TIn, double>::value) {
#pragma unroll
                    for (int s = 0; s < NS; ++s) w1[s] += h1 * load_acc<TIn, TAcc>(x + (size_t)s * op.D + idx);
                } else {
                    const TIn *xr = x + (size_t)idx * nv_lo_stride(NS);
#pragma unroll
                    for (int s = 0; s < NS; ++s) w1[s] += h1 * load_acc<TIn, TAcc>(xr + s);
                }
            }
        }
        const TAcc *C = cm + (size_t)t * NS * NS;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            TAcc a0 = acc0[s], a1 = acc1[s];
#pragma unroll
            for (int s2 = 0; s2 < NS; ++s2) {
                a0 += C[s * NS + s2] * w0[s2];
                a1 += C[s * NS + s2] * w1[s2];
            }
            acc0[s] = a0;
            acc1[s] = a1;
        }
    }
}

// ------------------------------------------------------------------------
// Shaped off-diagonal H*v.  The same sum as offdiag_accumulate, with the
// steps of a term generated from its shape (NvTermShape): the digits and the
// per-mode factors f_j(Delta_j) are read once per term, the step loop is
// unrolled, a step's factor is f_0 * f_1 * f_2 (the step kernel's product
// order) and its offset a scalar.  Neighbour rows come through buffer
// resources sized to the vector: an offset outside it (a step that leaves the
// basis, whose factor is 0; a negative offset wraps past 2^31) loads 0, so
// there is no branch per step.
// ------------------------------------------------------------------------
typedef float nv_f2 __attribute__((ext_vector_type(2)));
typedef float nv_f3 __attribute__((ext_vector_type(3)));
typedef float nv_f4 __attribute__((ext_vector_type(4)));

// Word 3 of the buffer resource: data format 32 (bits 15-18 = 4); the
// run-time compiled kernels use the same value.
constexpr int kBufFlags = 0x00020000;

template <int NS, typename TIn> struct XRes;

// One interleaved row as loaded (P = nv_lo_stride values), held apart from its
// use so that a loop can keep loads of the next iteration in flight.
template <int P> struct FRow { nv_f4 v[P / 4]; };
template <> struct FRow<1> { float v; };
template <> struct FRow<2> { nv_f2 v; };
template <> struct FRow<3> { nv_f3 v; };

// interleaved FP32 copy, x[r * nv_lo_stride(NS) + s]: one load per row up to
// four states, 16-byte loads of the states present beyond
template <int NS> struct XRes<NS, float> {
    static constexpr unsigned kEl = nv_lo_stride(NS) * 4;
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ XRes(const float *x, int D)
        : r(__builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, (int)((unsigned)D * kEl), kBufFlags)) {}
    template <typename TAcc>
    __device__ __forceinline__ void fma(unsigned o, TAcc h, TAcc (&w)[NS]) const { fma_so<TAcc>(o, 0u, h, w); }
    // offset o (per lane) plus so (wave-uniform, a scalar register)
    template <typename TAcc>
    __device__ __forceinline__ void fma_so(unsigned o, unsigned so, TAcc h, TAcc (&w)[NS]) const {
        constexpr int P = nv_lo_stride(NS);
        if constexpr (P == 1) {
            const float v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)o, (int)so, 0));
            w[0] += h * (TAcc)v;
        } else if constexpr (P == 2) {
            const nv_f2 v = __builtin_bit_cast(nv_f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)o, (int)so, 0));
#pragma unroll
            for (int s = 0; s < NS; ++s) w[s] += h * (TAcc)v[s];
        } else if constexpr (P == 3) {
            const nv_f3 v = __builtin_bit_cast(nv_f3, __builtin_amdgcn_raw_buffer_load_b96(r, (int)o, (int)so, 0));
#pragma unroll
            for (int s = 0; s < NS; ++s) w[s] += h * (TAcc)v[s];
        } else if constexpr (P == 4) {
            const nv_f4 v = __builtin_bit_cast(nv_f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)o, (int)so, 0));
#pragma unroll
            for (int s = 0; s < NS; ++s) w[s] += h * (TAcc)v[s];
        } else {
#pragma unroll
            for (int part = 0; part * 4 < NS; ++part) {
                const nv_f4 v = __builtin_bit_cast(
                    nv_f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(o + 16u * part), (int)so, 0));
#pragma unroll
                for (int s = part * 4; s < NS && s < part * 4 + 4; ++s) w[s] += h * (TAcc)v[s - part * 4];
            }
        }
    }
    using Row = FRow<nv_lo_stride(NS)>;
    __device__ __forceinline__ Row load_so(unsigned o, unsigned so) const {
        constexpr int P = nv_lo_stride(NS);
        Row x;
        if constexpr (P == 1) x.v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)o, (int)so, 0));
        else if constexpr (P == 2) x.v = __builtin_bit_cast(nv_f2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)o, (int)so, 0));
        else if constexpr (P == 3) x.v = __builtin_bit_cast(nv_f3, __builtin_amdgcn_raw_buffer_load_b96(r, (int)o, (int)so, 0));
        else {
#pragma unroll
            for (int part = 0; part < P / 4; ++part)
                x.v[part] = __builtin_bit_cast(
                    nv_f4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(o + 16u * part), (int)so, 0));
        }
        return x;
    }
    template <typename TAcc> static __device__ __forceinline__ void fma_row(const Row &x, TAcc h, TAcc (&w)[NS]) {
        constexpr int P = nv_lo_stride(NS);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if constexpr (P == 1) w[s] += h * (TAcc)x.v;
            else if constexpr (P <= 3) w[s] += h * (TAcc)x.v[s];
            else w[s] += h * (TAcc)x.v[s / 4][s % 4];
        }
    }
};

// state-major FP64 vector, x[s * D + r]: one resource per state
template <int NS> struct XRes<NS, double> {
    static constexpr unsigned kEl = 8;
    __amdgpu_buffer_rsrc_t r[NS];
    __device__ __forceinline__ XRes(const double *x, int D) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
            r[s] = __builtin_amdgcn_make_buffer_rsrc((void *)(x + (size_t)s * D), (short)0, (int)((unsigned)D * 8u),
                                                     kBufFlags);
    }
    template <typename TAcc>
    __device__ __forceinline__ void fma(unsigned o, TAcc h, TAcc (&w)[NS]) const {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r[s], (int)o, 0, 0));
            w[s] += h * (TAcc)v;
        }
    }
};

// interleaved FP64 copy, x[r * nv_lo_stride(NS) + s]: 16-byte loads of two states
struct F64I {};
template <int NS> struct XRes<NS, F64I> {
    static constexpr unsigned kEl = nv_lo_stride(NS) * 8;
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ XRes(const double *x, int D)
        : r(__builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, (int)((unsigned)D * kEl), kBufFlags)) {}
    template <typename TAcc>
    __device__ __forceinline__ void fma(unsigned o, TAcc h, TAcc (&w)[NS]) const { fma_so<TAcc>(o, 0u, h, w); }
    template <typename TAcc>
    __device__ __forceinline__ void fma_so(unsigned o, unsigned so, TAcc h, TAcc (&w)[NS]) const {
        typedef double nv_d2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int part = 0; 2 * part < NS; ++part) {
            if (2 * part + 1 < NS) {
                const nv_d2 v = __builtin_bit_cast(
                    nv_d2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(o + 16u * part), (int)so, 0));
                w[2 * part] += h * (TAcc)v[0];
                w[2 * part + 1] += h * (TAcc)v[1];
            } else {
                const double v = __builtin_bit_cast(
                    double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(o + 16u * part), (int)so, 0));
                w[2 * part] += h * (TAcc)v;
            }
        }
    }
    typedef double nv_rd2 __attribute__((ext_vector_type(2)));
    struct Row {
        nv_rd2 v[NS / 2];
        double t;  // the last state when NS is odd
    };
    __device__ __forceinline__ Row load_so(unsigned o, unsigned so) const {
        Row x;
#pragma unroll
        for (int part = 0; 2 * part + 1 < NS; ++part)
            x.v[part] = __builtin_bit_cast(nv_rd2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(o + 16u * part), (int)so, 0));
        if constexpr (NS % 2 == 1)
            x.t = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(o + 16u * (NS / 2)), (int)so, 0));
        else
            x.t = 0.0;
        return x;
    }
    template <typename TAcc> static __device__ __forceinline__ void fma_row(const Row &x, TAcc h, TAcc (&w)[NS]) {
#pragma unroll
        for (int s = 0; s + 1 < NS; s += 2) {
            w[s] += h * (TAcc)x.v[s / 2][0];
            w[s + 1] += h * (TAcc)x.v[s / 2][1];
        }
        if constexpr (NS % 2 == 1) w[NS - 1] += h * (TAcc)x.t;
    }
};

// The terms of shape group g; R rows per thread (digits of row q at
// dig + q * M * kBlock), row q's byte offset in the vector vo[q].  SKIP
// points the load of a step whose factor is 0 outside the buffer (it then
// moves no memory), where a plain out-of-basis offset may land on another row.
template <int NS, typename TIn, typename TAcc, int R, bool SKIP, int K0, int K1, int K2>
__device__ __forceinline__ void shaped_group(const NvOpDev &op, const XRes<NS, TIn> &xs, const unsigned (&vo)[R],
                                             const uint8_t *dig, const TAcc *ftab, int tid, int g,
                                             TAcc (&acc)[R][NS]) {
    constexpr int J1 = K1 ? K1 : 1, J2 = K2 ? K2 : 1;
    // all slots odd: the middle of the product is Delta = 0 (diagonal part)
    constexpr bool odd = (K0 & 1) && (K1 == 0 || (K1 & 1)) && (K2 == 0 || (K2 & 1));
    const TAcc *__restrict__ cm = cmat_of<TAcc>(op);
    const int t1 = op.tgrp[g + 1];
    for (int t = op.tgrp[g]; t < t1; ++t) {
        const NvTermShape d = op.tshape[t];
        TAcc f0[R][K0], f1[R][J1], f2[R][J2];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const uint8_t *dq = dig + q * op.M * kBlock + tid;
            const int n0 = dq[d.mode[0] * kBlock];
#pragma unroll
            for (int i = 0; i < K0; ++i) f0[q][i] = ftab[d.code[0][i] * op.nmax + n0];
            if constexpr (K1 > 0) {
                const int n1 = dq[d.mode[1] * kBlock];
#pragma unroll
                for (int i = 0; i < K1; ++i) f1[q][i] = ftab[d.code[1][i] * op.nmax + n1];
            }
            if constexpr (K2 > 0) {
                const int n2 = dq[d.mode[2] * kBlock];
#pragma unroll
                for (int i = 0; i < K2; ++i) f2[q][i] = ftab[d.code[2][i] * op.nmax + n2];
            }
        }
        TAcc w[R][NS];
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int s = 0; s < NS; ++s) w[q][s] = (TAcc)0;
#pragma unroll
        for (int i0 = 0; i0 < K0; ++i0)
#pragma unroll
            for (int i1 = 0; i1 < J1; ++i1)
#pragma unroll
                for (int i2 = 0; i2 < J2; ++i2) {
                    if (odd && 2 * i0 + 1 == K0 && (K1 == 0 || 2 * i1 + 1 == K1) && (K2 == 0 || 2 * i2 + 1 == K2))
                        continue;
                    int off = d.base + i0 * d.inc[0];
                    if constexpr (K1 > 0) off += i1 * d.inc[1];
                    if constexpr (K2 > 0) off += i2 * d.inc[2];
                    const unsigned ob = (unsigned)off * XRes<NS, TIn>::kEl;
#pragma unroll
                    for (int q = 0; q < R; ++q) {
                        TAcc h = f0[q][i0];
                        if constexpr (K1 > 0) h *= f1[q][i1];
                        if constexpr (K2 > 0) h *= f2[q][i2];
                        unsigned o = vo[q] + ob;
                        if constexpr (SKIP) o = h != (TAcc)0 ? o : 0x80000000u;
                        xs.template fma<TAcc>(o, h, w[q]);
                    }
                }
        const TAcc *C = cm + (size_t)d.cm * NS * NS;
#pragma unroll
        for (int q = 0; q < R; ++q)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                TAcc a = acc[q][s];
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) a += C[s * NS + s2] * w[q][s2];
                acc[q][s] = a;
            }
    }
}

template <int NS, typename TIn, typename TAcc, int R, bool SKIP>
__device__ __forceinline__ void offdiag_shaped(const NvOpDev &op, const XRes<NS, TIn> &xs, const unsigned (&vo)[R],
                                               const uint8_t *dig, const TAcc *ftab, int tid, TAcc (&acc)[R][NS]) {
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[q][s] = (TAcc)0;
    shaped_group<NS, TIn, TAcc, R, SKIP, 2, 0, 0>(op, xs, vo, dig, ftab, tid, 0, acc);
    shaped_group<NS, TIn, TAcc, R, SKIP, 3, 0, 0>(op, xs, vo, dig, ftab, tid, 1, acc);
    shaped_group<NS, TIn, TAcc, R, SKIP, 4, 0, 0>(op, xs, vo, dig, ftab, tid, 2, acc);
    shaped_group<NS, TIn, TAcc, R, SKIP, 2, 2, 0>(op, xs, vo, dig, ftab, tid, 3, acc);
    shaped_group<NS, TIn, TAcc, R, SKIP, 3, 2, 0>(op, xs, vo, dig, ftab, tid, 4, acc);
    shaped_group<NS, TIn, TAcc, R, SKIP, 2, 3, 0>(op, xs, vo, dig, ftab, tid, 5, acc);
    shaped_group<NS, TIn, TAcc, R, SKIP, 1, 2, 0>(op, xs, vo, dig, ftab, tid, 6, acc);
    shaped_group<NS, TIn, TAcc, R, SKIP, 2, 1, 0>(op, xs, vo, dig, ftab, tid, 7, acc);
    shaped_group<NS, TIn, TAcc, R, SKIP, 2, 2, 2>(op, xs, vo, dig, ftab, tid, 8, acc);
}

template <typename TAcc>
__device__ __forceinline__ TAcc *stage_tables(const NvOpDev &op, uint8_t *&dig) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    TAcc *ftab = reinterpret_cast<TAcc *>(lds_raw);
    const int nt = op.ncodes * op.nmax;
    const TAcc *g = ftab_of<TAcc>(op);
    for (int i = threadIdx.x; i < nt; i += blockDim.x) ftab[i] = g[i];
    size_t off = ((size_t)nt * sizeof(TAcc) + 15) & ~(size_t)15;
    dig = lds_raw + off;
    return ftab;
}

// Standalone y = H x (x, y double).  TAcc selects the precision of the
// off-diagonal part; the diagonal part is always FP64.
template <int NS, typename TAcc>
__global__ __launch_bounds__(kBlock) void hv_full_kernel(NvOpDev op, const double *__restrict__ x,
                                                          double *__restrict__ y) {
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    const int tid = threadIdx.x;
    const int r = xcd_block(blockIdx.x, gridDim.x) * kBlock + tid;
    const bool live = r < op.D;
    if (live) digits_energy(op, r, dig, tid);
    __syncthreads();
    if (!live) return;
    TAcc acc[NS];
    offdiag_accumulate<NS, double, TAcc>(op, x, r, dig, ftab, tid, acc);
    double v[NS], dv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) v[s] = x[(size_t)s * op.D + r];
    diag_apply<NS>(op, r, dig, tid, v, dv);
#pragma unroll
    for (int s = 0; s < NS; ++s) y[(size_t)s * op.D + r] = dv[s] + (double)acc[s];
}

__device__ __forceinline__ int64_t slot_of(int64_t i, int64_t nslots) { return i % nslots; }

// A vector element touched once per launch: non-temporal (NT) where the
// vectors are far larger than the caches, plain otherwise (a vector that fits
// the Infinity Cache is read again from it by the next kernel).
template <bool NT, typename T> __device__ __forceinline__ T ld_once(const T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T> __device__ __forceinline__ void st_once(T v, T *p) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Iteration H*v: y = offdiag(H) q_i in TAcc, alpha(i) = q_i . (H q_i) in FP64.
template <int NS, typename TIn, typename TAcc>
__global__ __launch_bounds__(kBlock) void hv_iter_kernel(NvOpDev op, LzState st) {
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    const int tid = threadIdx.x;
    const int r = xcd_block(blockIdx.x, gridDim.x) * kBlock + tid;
    const bool live = r < op.D;
    const int64_t it = *st.iter;
    const double *q = st.V + slot_of(it, st.nslots) * (int64_t)op.N;
    const TIn *x = st.vlo ? reinterpret_cast<const TIn *>(st.vlo) : reinterpret_cast<const TIn *>(q);
    if (live) digits_energy(op, r, dig, tid);
    __syncthreads();
    double part = 0.0;
    if (live) {
        TAcc acc[NS];
        offdiag_accumulate<NS, TIn, TAcc>(op, x, r, dig, ftab, tid, acc);
        double v[NS], dv[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r];
        diag_apply<NS>(op, r, dig, tid, v, dv);
        TAcc *y = reinterpret_cast<TAcc *>(st.y);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            y[(size_t)s * op.D + r] = acc[s];
            part += v[s] * (dv[s] + (double)acc[s]);
        }
    }
    const double b = block_sum(part);
    if (tid == 0) st.partials[blockIdx.x] = b;
}

// hv_iter_kernel with two rows per thread: block b covers the rows of the
// one-row kernel's blocks 2b and 2b+1 and writes their two partials, so alpha
// is the same sum in the same order.
template <int NS, typename TIn, typename TAcc>
__global__ __launch_bounds__(kBlock) void hv_iter2_kernel(NvOpDev op, LzState st) {
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    uint8_t *dig1 = dig + op.M * kBlock;
    const int tid = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = 2 * b * kBlock + tid, r1 = r0 + kBlock;
    const bool l0 = r0 < op.D, l1 = r1 < op.D;
    const int64_t it = *st.iter;
    const double *q = st.V + slot_of(it, st.nslots) * (int64_t)op.N;
    const TIn *x = st.vlo ? reinterpret_cast<const TIn *>(st.vlo) : reinterpret_cast<const TIn *>(q);
    if (l0) digits_energy(op, r0, dig, tid);
    else for (int k = 0; k < op.M; ++k) dig[k * kBlock + tid] = 0;
    if (l1) digits_energy(op, r1, dig1, tid);
    else for (int k = 0; k < op.M; ++k) dig1[k * kBlock + tid] = 0;
    __syncthreads();
    double part0 = 0.0, part1 = 0.0;
    if (l0) {
        TAcc acc0[NS], acc1[NS];
        offdiag_accumulate2<NS, TIn, TAcc>(op, x, r0, r1, l0, l1, dig, dig1, ftab, tid, acc0, acc1);
        TAcc *y = reinterpret_cast<TAcc *>(st.y);
        double v[NS], dv[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r0];
        diag_apply<NS>(op, r0, dig, tid, v, dv);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            y[(size_t)s * op.D + r0] = acc0[s];
            part0 += v[s] * (dv[s] + (double)acc0[s]);
        }
        if (l1) {
#pragma unroll
            for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r1];
            diag_apply<NS>(op, r1, dig1, tid, v, dv);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                y[(size_t)s * op.D + r1] = acc1[s];
                part1 += v[s] * (dv[s] + (double)acc1[s]);
            }
        }
    }
    const double s0 = block_sum(part0);
    const double s1 = block_sum(part1);
    if (tid == 0) {
        st.partials[2 * b] = s0;
        st.partials[2 * b + 1] = s1;
    }
}

// hv_iter_kernel over the shaped terms; R rows per thread.  Block b covers
// the rows of the one-row kernel's blocks R*b .. R*b+R-1 and writes their
// partials, so alpha is the same sum in the same order.  Rows past the end
// compute on zero digits and out-of-range loads; they are not written.
template <int NS, typename TIn, typename TAcc, int R, bool SKIP>
__global__ __launch_bounds__(kBlock) void hv_shaped_iter_kernel(NvOpDev op, LzState st) {
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    const int tid = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int64_t it = *st.iter;
    const double *q = st.V + slot_of(it, st.nslots) * (int64_t)op.N;
    using TX = typename std::conditional<std::is_same<TIn, F64I>::value, double, TIn>::type;
    const TX *x = std::is_same<TIn, F64I>::value ? reinterpret_cast<const TX *>(st.v64i)
                  : st.vlo                       ? reinterpret_cast<const TX *>(st.vlo)
                                                 : reinterpret_cast<const TX *>(q);
    int r[R];
    bool live[R];
    unsigned vo[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        r[k] = (R * b + k) * kBlock + tid;
        live[k] = r[k] < op.D;
        vo[k] = (unsigned)r[k] * XRes<NS, TIn>::kEl;
        uint8_t *dk = dig + k * op.M * kBlock;
        if (live[k]) digits_energy(op, r[k], dk, tid);
        else for (int m = 0; m < op.M; ++m) dk[m * kBlock + tid] = 0;
    }
    __syncthreads();
    const XRes<NS, TIn> xs(x, op.D);
    TAcc acc[R][NS];
    offdiag_shaped<NS, TIn, TAcc, R, SKIP>(op, xs, vo, dig, ftab, tid, acc);
    TAcc *y = reinterpret_cast<TAcc *>(st.y);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double part = 0.0;
        if (live[k]) {
            double v[NS], dv[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r[k]];
            diag_apply<NS>(op, r[k], dig + k * op.M * kBlock, tid, v, dv);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                y[(size_t)s * op.D + r[k]] = acc[k][s];
                part += v[s] * (dv[s] + (double)acc[k][s]);
            }
        }
        const double bs = block_sum(part);
        if (tid == 0) st.partials[R * b + k] = bs;
    }
}

// hv_full_kernel over the shaped terms.
template <int NS, typename TAcc>
__global__ __launch_bounds__(kBlock) void hv_shaped_full_kernel(NvOpDev op, const double *__restrict__ x,
                                                                 double *__restrict__ y) {
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    const int tid = threadIdx.x;
    const int r = xcd_block(blockIdx.x, gridDim.x) * kBlock + tid;
    const bool live = r < op.D;
    if (live) digits_energy(op, r, dig, tid);
    else for (int m = 0; m < op.M; ++m) dig[m * kBlock + tid] = 0;
    __syncthreads();
    const XRes<NS, double> xs(x, op.D);
    const unsigned vo[1] = {(unsigned)r * 8u};
    TAcc acc[1][NS];
    offdiag_shaped<NS, double, TAcc, 1, false>(op, xs, vo, dig, ftab, tid, acc);
    if (!live) return;
    double v[NS], dv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) v[s] = x[(size_t)s * op.D + r];
    diag_apply<NS>(op, r, dig, tid, v, dv);
#pragma unroll
    for (int s = 0; s < NS; ++s) y[(size_t)s * op.D + r] = dv[s] + (double)acc[0][s];
}

// ------------------------------------------------------------------------
// Two-pass H*v (tables in NvHv2Dev).  A term t = C_t (x) P_sigma (x) P_phi
// splits into sigma, its powers of the two slowest modes 0 and 1, and phi,
// its monomial over the other modes.  With x_r the row r of x (all states):
//   pass 1: Z_g[r] = sum over the terms with sigma = g of C_t (P_phi x)[r],
//           each distinct phi gathered once per row and fed to every
//           channel g it pairs with;
//   pass 2: y[r] = Z_0[r] + sum_{g > 0} sum_{Delta_0, Delta_1}
//           h_0 h_1 Z_g[r + Delta_0 shft_0 + Delta_1 shft_1].
// The step Delta = 0 of the whole term is diagonal (it lives in the FP64
// diagonal part): pass 1 leaves out phi's zero step for g = 0 only, and pass 2
// takes the pure q_a^2 term's C_t h(2, n_a, 0) x_r out again at Delta_a = 0.
// Pass 1 does the distinct phi's gathers (cubic 4-state operator on 10 modes:
// 936 of the one-pass form's 1760 per row) and all the coupling products;
// pass 2 streams the channels along the two slow modes, a workgroup covering
// every (n_0, n_1) of a few consecutive rows of the faster modes so that each
// channel row is fetched from memory once.
// ------------------------------------------------------------------------
typedef float nv_mf4 __attribute__((ext_vector_type(4)));

// Channel accumulators of pass 1.  VALU form: one row per lane, NS values per
// channel.  MFMA form (FP32, NS <= 4): the coupling products run on the
// matrix cores as v_mfma_f32_4x4x1 (16 blocks of 4 lanes): block b of a wave
// is its rows 4b..4b+3; A = the lane's own phi value of state s', B = row
// (lane & 3) of C_t at column s', so lane 4b+j, register i accumulates row
// 4b+i, state j.  The VALU is left with the gathers.
template <int NS, typename TAcc, bool MF> struct Hv2Acc {
    TAcc a[NV_HV2_NG][NS];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int g = 0; g < NV_HV2_NG; ++g)
#pragma unroll
            for (int s = 0; s < NS; ++s) a[g][s] = (TAcc)0;
    }
    template <int G> __device__ __forceinline__ void mix(const TAcc *C, const TAcc (&u)[NS]) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            TAcc v = a[G][s];
#pragma unroll
            for (int s2 = 0; s2 < NS; ++s2) v += C[s * NS + s2] * u[s2];
            a[G][s] = v;
        }
    }
};
template <int NS> struct Hv2Acc<NS, float, true> {
    nv_mf4 a[NV_HV2_NG];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int g = 0; g < NV_HV2_NG; ++g) a[g] = nv_mf4{0.f, 0.f, 0.f, 0.f};
    }
    // C: the lane's row (lane & 3) of the 4x4-padded coupling matrix
    template <int G> __device__ __forceinline__ void mix(const float *C, const float (&u)[NS]) {
        const nv_mf4 c = *reinterpret_cast<const nv_mf4 *>(C);
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) a[G] = __builtin_amdgcn_mfma_f32_4x4x1f32(u[s2], c[s2], a[G], 0, 0, 0);
    }
};

// The coupling rows a consumer reads: global C_t (VALU) or the lane's row of
// the padded copy staged in LDS (MFMA).
template <int NS, typename TAcc, bool MF>
__device__ __forceinline__ const TAcc *hv2_cmat(const NvOpDev &op, const float *ldsc, int cm) {
    if constexpr (MF) return reinterpret_cast<const TAcc *>(ldsc + cm * 16 + (threadIdx.x & 3) * 4);
    else return cmat_of<TAcc>(op) + (size_t)cm * NS * NS;
}

// acc[g] += C_g u over the channels: cm[g] is the term's cmat row, or -1
// (uniform over the wave); ucm0 is fed to channel 0, u to the others.
template <int NS, typename TAcc, bool MF>
__device__ __forceinline__ void hv2_feed(const NvOpDev &op, const float *ldsc, Hv2Acc<NS, TAcc, MF> &acc,
                                         const int32_t (&cm)[NV_HV2_NG], const TAcc (&u0)[NS], const TAcc (&u)[NS]) {
    if (cm[0] >= 0) acc.template mix<0>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[0]), u0);
    if (cm[1] >= 0) acc.template mix<1>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[1]), u);
    if (cm[2] >= 0) acc.template mix<2>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[2]), u);
    if (cm[3] >= 0) acc.template mix<3>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[3]), u);
    if (cm[4] >= 0) acc.template mix<4>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[4]), u);
    if (cm[5] >= 0) acc.template mix<5>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[5]), u);
    if (cm[6] >= 0) acc.template mix<6>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[6]), u);
    if (cm[7] >= 0) acc.template mix<7>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[7]), u);
    if (cm[8] >= 0) acc.template mix<8>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[8]), u);
    if (cm[9] >= 0) acc.template mix<9>(hv2_cmat<NS, TAcc, MF>(op, ldsc, cm[9]), u);
}

// The phis of shape group grp: steps gathered as in shaped_group (one row per
// thread, out-of-basis loads pointed outside the vector), then fed to their
// consumers.  u leaves out phi's zero step (for channel 0, where the whole
// step would be diagonal), uf keeps it.
template <int NS, typename TIn, typename TAcc, bool MF, bool SKIP, int K0, int K1, int K2>
__device__ __forceinline__ void hv2_phi_group(const NvOpDev &op, const XRes<NS, TIn> &xs, unsigned vo,
                                              const uint8_t *dig, const TAcc *ftab, const float *ldsc, int tid,
                                              int grp, const TAcc (&xr)[NS], Hv2Acc<NS, TAcc, MF> &acc) {
    constexpr int J1 = K1 ? K1 : 1, J2 = K2 ? K2 : 1;
    constexpr bool zs = (K0 & 1) && (K1 == 0 || (K1 & 1)) && (K2 == 0 || (K2 & 1));
    const NvHv2Dev &h = op.hv2;
    const int p1 = h.pgrp[grp + 1];
    for (int p = h.pgrp[grp]; p < p1; ++p) {
        const NvPhi &ph = h.phi[p];
        const NvTermShape d = ph.sh;
        TAcc f0[K0], f1[J1], f2[J2];
        const uint8_t *dq = dig + tid;
        const int n0 = dq[d.mode[0] * kBlock];
        TAcc fd = (TAcc)1;  // the factors of the modes whose only step is Delta = 0
        for (int j = 0; j < ph.nd; ++j) fd *= ftab[ph.dcode[j] * op.nmax + dq[ph.dmode[j] * kBlock]];
#pragma unroll
        for (int i = 0; i < K0; ++i) f0[i] = ftab[d.code[0][i] * op.nmax + n0] * fd;
        if constexpr (K1 > 0) {
            const int n1 = dq[d.mode[1] * kBlock];
#pragma unroll
            for (int i = 0; i < K1; ++i) f1[i] = ftab[d.code[1][i] * op.nmax + n1];
        } else {
            f1[0] = (TAcc)1;
        }
        if constexpr (K2 > 0) {
            const int n2 = dq[d.mode[2] * kBlock];
#pragma unroll
            for (int i = 0; i < K2; ++i) f2[i] = ftab[d.code[2][i] * op.nmax + n2];
        } else {
            f2[0] = (TAcc)1;
        }
        TAcc u[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) u[s] = (TAcc)0;
#pragma unroll
        for (int i0 = 0; i0 < K0; ++i0)
#pragma unroll
            for (int i1 = 0; i1 < J1; ++i1) {
                TAcc h01 = f0[i0];
                if constexpr (K1 > 0) h01 *= f1[i1];
#pragma unroll
                for (int i2 = 0; i2 < J2; ++i2) {
                    if (zs && 2 * i0 + 1 == K0 && (K1 == 0 || 2 * i1 + 1 == K1) && (K2 == 0 || 2 * i2 + 1 == K2))
                        continue;
                    int off = d.base + i0 * d.inc[0];
                    if constexpr (K1 > 0) off += i1 * d.inc[1];
                    if constexpr (K2 > 0) off += i2 * d.inc[2];
                    TAcc hf = h01;
                    if constexpr (K2 > 0) hf *= f2[i2];
                    // plain form: an out-of-basis step reads some row (or 0 past the ends) times 0
                    unsigned o = vo + (unsigned)off * XRes<NS, TIn>::kEl;
                    if constexpr (SKIP) o = hf != (TAcc)0 ? o : 0x80000000u;
                    xs.template fma<TAcc>(o, hf, u);
                }
            }
        TAcc uf[NS];
        if constexpr (zs) {
            TAcc z = f0[(K0 - 1) / 2];
            if constexpr (K1 > 0) z *= f1[(K1 - 1) / 2];
            if constexpr (K2 > 0) z *= f2[(K2 - 1) / 2];
#pragma unroll
            for (int s = 0; s < NS; ++s) uf[s] = u[s] + z * xr[s];
        } else {
#pragma unroll
            for (int s = 0; s < NS; ++s) uf[s] = u[s];
        }
        hv2_feed<NS, TAcc, MF>(op, ldsc, acc, ph.cm, u, uf);
    }
}

__device__ __forceinline__ size_t hv2_tables_bytes(const NvOpDev &op, size_t el) {
    return (((size_t)op.ncodes * op.nmax * el + 15) & ~(size_t)15) + 2 * (size_t)op.M * kBlock;
}

// Pass 1: Z_g[r] for every channel, x the interleaved copy of the vector
// (FP32, or FP64 as F64I).  Z is [NV_HV2_NG][D][nv_lo_stride(NS)].
template <int NS, typename TIn, typename TAcc, bool MF, bool SKIP>
__global__ __launch_bounds__(kBlock) void hv2_pass1_kernel(NvOpDev op, const void *__restrict__ xin,
                                                            TAcc *__restrict__ Z) {
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    float *ldsc = nullptr;
    if constexpr (MF) {
        extern __shared__ __align__(16) unsigned char lds_raw[];
        ldsc = reinterpret_cast<float *>(lds_raw + hv2_tables_bytes(op, sizeof(float)));
        for (int i = threadIdx.x; i < op.nterms * 16; i += blockDim.x) {
            const int t = i >> 4, j = (i >> 2) & 3, k = i & 3;
            ldsc[i] = (j < NS && k < NS) ? op.cmat32[(size_t)t * NS * NS + j * NS + k] : 0.f;
        }
    }
    const int tid = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r = b * kBlock + tid;
    const bool live = r < op.D;
    if (live) digits_energy(op, r, dig, tid);
    else for (int m = 0; m < op.M; ++m) dig[m * kBlock + tid] = 0;
    __syncthreads();
    using TX = typename std::conditional<std::is_same<TIn, F64I>::value, double, TIn>::type;
    const XRes<NS, TIn> xs(reinterpret_cast<const TX *>(xin), op.D);
    const unsigned vo = (unsigned)r * XRes<NS, TIn>::kEl;
    TAcc xr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) xr[s] = (TAcc)0;
    xs.template fma<TAcc>(vo, (TAcc)1, xr);  // row r itself (0 past the end)
    Hv2Acc<NS, TAcc, MF> acc;
    acc.zero();
    const NvHv2Dev &h = op.hv2;
    hv2_feed<NS, TAcc, MF>(op, ldsc, acc, h.idcm, xr, xr);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 2, 0, 0>(op, xs, vo, dig, ftab, ldsc, tid, 0, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 3, 0, 0>(op, xs, vo, dig, ftab, ldsc, tid, 1, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 4, 0, 0>(op, xs, vo, dig, ftab, ldsc, tid, 2, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 2, 2, 0>(op, xs, vo, dig, ftab, ldsc, tid, 3, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 3, 2, 0>(op, xs, vo, dig, ftab, ldsc, tid, 4, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 2, 3, 0>(op, xs, vo, dig, ftab, ldsc, tid, 5, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 1, 2, 0>(op, xs, vo, dig, ftab, ldsc, tid, 6, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 2, 1, 0>(op, xs, vo, dig, ftab, ldsc, tid, 7, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 2, 2, 2>(op, xs, vo, dig, ftab, ldsc, tid, 8, xr, acc);
    hv2_phi_group<NS, TIn, TAcc, MF, SKIP, 1, 0, 0>(op, xs, vo, dig, ftab, ldsc, tid, 9, xr, acc);
    constexpr int P = nv_lo_stride(NS);
    if constexpr (MF) {
        const int rw = b * kBlock + (tid & ~63) + 4 * ((tid & 63) >> 2);
        const int j = tid & 3;
#pragma unroll
        for (int g = 0; g < NV_HV2_NG; ++g) {
            if (!((h.gmask >> g) & 1)) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (rw + i < op.D && j < NS) Z[((size_t)g * op.D + rw + i) * P + j] = acc.a[g][i];
        }
    } else {
        if (!live) return;
#pragma unroll
        for (int g = 0; g < NV_HV2_NG; ++g) {
            if (!((h.gmask >> g) & 1)) continue;
#pragma unroll
            for (int s = 0; s < NS; ++s) Z[((size_t)g * op.D + r) * P + s] = acc.a[g][s];
        }
    }
}

// Pass 2, one signature channel g of powers (P0, P1) on modes 0 and 1: the
// R2 consecutive rows r0..r0+R2-1 of a thread share (n_0, n_1) and so every
// factor.  Steps leaving the basis have factor 0 (factor table) and their
// loads are pointed outside the buffer; no branch.  A signature whose step
// range was cut by a small basis (s < p) takes the generic loop.
template <int NS, typename TIn, typename TAcc, int R2, int P0, int P1>
__device__ __forceinline__ void hv2_sig_rows(const NvOpDev &op, const NvSigma &sg, const XRes<NS, TIn> &zg, int r0,
                                             int n0, int n1, int nf0, TAcc (&y)[R2][NS]) {
    const NvHv2Dev &h = op.hv2;
    const TAcc *ft = ftab_of<TAcc>(op);
    constexpr unsigned kEl = XRes<NS, TIn>::kEl;
    if (sg.s0 == (P0 ? P0 : -1) && sg.s1 == (P1 ? P1 : -1)) {
        constexpr int K0 = P0 ? P0 + 1 : 1, K1 = P1 ? P1 + 1 : 1;
        TAcc f0[K0], f1[K1];
#pragma unroll
        for (int i = 0; i < K0; ++i) f0[i] = P0 ? ft[sg.code0[i] * op.nmax + n0] : (TAcc)1;
#pragma unroll
        for (int i = 0; i < K1; ++i) f1[i] = P1 ? ft[sg.code1[i] * op.nmax + n1] : (TAcc)1;
#pragma unroll
        for (int i0 = 0; i0 < K0; ++i0)
#pragma unroll
            for (int i1 = 0; i1 < K1; ++i1) {
                const int d0 = P0 ? 2 * i0 - P0 : 0, d1 = P1 ? 2 * i1 - P1 : 0;
                const TAcc f = f0[i0] * f1[i1];
                const int off = (d0 * h.n1 + d1) * h.dr;
#pragma unroll
                for (int k = 0; k < R2; ++k) {
                    const unsigned o = f != (TAcc)0 ? (unsigned)(r0 + k + off) * kEl : 0x80000000u;
                    zg.template fma<TAcc>(o, f, y[k]);
                }
            }
        return;
    }
    const int k0 = sg.s0 < 0 ? 1 : sg.s0 + 1, k1 = sg.s1 < 0 ? 1 : sg.s1 + 1;
    for (int i0 = 0; i0 < k0; ++i0) {
        const int d0 = sg.s0 < 0 ? 0 : 2 * i0 - sg.s0;
        if (n0 + d0 < 0 || n0 + d0 >= nf0) continue;
        const TAcc f0 = sg.s0 < 0 ? (TAcc)1 : ft[sg.code0[i0] * op.nmax + n0];
        for (int i1 = 0; i1 < k1; ++i1) {
            const int d1 = sg.s1 < 0 ? 0 : 2 * i1 - sg.s1;
            if (n1 + d1 < 0 || n1 + d1 >= h.n1) continue;
            const TAcc f1 = sg.s1 < 0 ? (TAcc)1 : ft[sg.code1[i1] * op.nmax + n1];
            const int off = (d0 * h.n1 + d1) * h.dr;
#pragma unroll
            for (int k = 0; k < R2; ++k) zg.template fma<TAcc>((unsigned)(r0 + k + off) * kEl, f0 * f1, y[k]);
        }
    }
}

template <int NS, typename TIn, typename TAcc, int R2, int G, int P0, int P1>
__device__ __forceinline__ void hv2_channel(const NvOpDev &op, const TAcc *Z, int r0, int n0, int n1, int nf0,
                                            TAcc (&y)[R2][NS]) {
    if (!((op.hv2.gmask >> G) & 1)) return;
    using TX = typename std::conditional<std::is_same<TIn, F64I>::value, double, TIn>::type;
    const XRes<NS, TIn> zg(reinterpret_cast<const TX *>(Z) + (size_t)G * op.D * nv_lo_stride(NS), op.D);
    hv2_sig_rows<NS, TIn, TAcc, R2, P0, P1>(op, op.hv2.sig[G], zg, r0, n0, n1, nf0, y);
}

// Pass 2: y_off from the channels for R2 consecutive rows per thread; ITER:
// the iteration's epilogue (y in the accumulation type, state-major; alpha
// partials with the FP64 diagonal, one per workgroup), else y = diag x + y_off
// in FP64 (standalone H*v).  Workgroup: the h.nl planes (n_0, n_1) times h.c
// threads of R2 consecutive plane rows each, so that every channel row comes
// from memory once and is reused from the caches for its neighbours.
template <int NS, typename TIn, typename TAcc, bool ITER, int R2>
__global__ __launch_bounds__(kBlock) void hv2_pass2_kernel(NvOpDev op, LzState st, const void *__restrict__ xin,
                                                            const TAcc *__restrict__ Z,
                                                            const double *__restrict__ xfull,
                                                            double *__restrict__ yfull) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    uint8_t *dig = lds_raw;
    const NvHv2Dev &h = op.hv2;
    const int tid = threadIdx.x;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int lpos = tid / h.c;
    const int rest0 = (b * h.c + (tid - lpos * h.c)) * R2;
    const bool inl = lpos < h.nl && rest0 < h.dr;
    const int r0 = inl ? lpos * h.dr + rest0 : 0;
    const int n0 = lpos / h.n1, n1 = lpos - (lpos / h.n1) * h.n1;
    const int nf0 = h.nl / h.n1;
    using TX = typename std::conditional<std::is_same<TIn, F64I>::value, double, TIn>::type;
    constexpr unsigned kEl = XRes<NS, TIn>::kEl;
    TAcc y[R2][NS];
#pragma unroll
    for (int k = 0; k < R2; ++k)
#pragma unroll
        for (int s = 0; s < NS; ++s) y[k][s] = (TAcc)0;
    if (inl) {
        if (h.gmask & 1) {
            const XRes<NS, TIn> z0(reinterpret_cast<const TX *>(Z), op.D);
#pragma unroll
            for (int k = 0; k < R2; ++k)
                if (rest0 + k < h.dr) z0.template fma<TAcc>((unsigned)(r0 + k) * kEl, (TAcc)1, y[k]);
        }
        hv2_channel<NS, TIn, TAcc, R2, 1, 1, 0>(op, Z, r0, n0, n1, nf0, y);
        hv2_channel<NS, TIn, TAcc, R2, 2, 2, 0>(op, Z, r0, n0, n1, nf0, y);
        hv2_channel<NS, TIn, TAcc, R2, 3, 3, 0>(op, Z, r0, n0, n1, nf0, y);
        hv2_channel<NS, TIn, TAcc, R2, 4, 0, 1>(op, Z, r0, n0, n1, nf0, y);
        hv2_channel<NS, TIn, TAcc, R2, 5, 0, 2>(op, Z, r0, n0, n1, nf0, y);
        hv2_channel<NS, TIn, TAcc, R2, 6, 0, 3>(op, Z, r0, n0, n1, nf0, y);
        hv2_channel<NS, TIn, TAcc, R2, 7, 1, 1>(op, Z, r0, n0, n1, nf0, y);
        hv2_channel<NS, TIn, TAcc, R2, 8, 2, 1>(op, Z, r0, n0, n1, nf0, y);
        hv2_channel<NS, TIn, TAcc, R2, 9, 1, 2>(op, Z, r0, n0, n1, nf0, y);
        // the pure q_a^2 terms at Delta_a = 0 are diagonal: take them out again
        for (int g = 2; g <= 5; g += 3) {
            const NvSigma &sg = h.sig[g];
            if (!((h.gmask >> g) & 1) || sg.corr_cm < 0) continue;
            const TAcc hz = ftab_of<TAcc>(op)[sg.corr_code * op.nmax + (g <= 3 ? n0 : n1)];
            const TAcc *C = cmat_of<TAcc>(op) + (size_t)sg.corr_cm * NS * NS;
            const XRes<NS, TIn> xs(reinterpret_cast<const TX *>(xin), op.D);
#pragma unroll
            for (int k = 0; k < R2; ++k) {
                TAcc xr[NS];
#pragma unroll
                for (int s = 0; s < NS; ++s) xr[s] = (TAcc)0;
                xs.template fma<TAcc>((unsigned)(r0 + k) * kEl, (TAcc)1, xr);
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    TAcc a = (TAcc)0;
#pragma unroll
                    for (int s2 = 0; s2 < NS; ++s2) a += C[s * NS + s2] * xr[s2];
                    y[k][s] -= hz * a;
                }
            }
        }
    }
    double part = 0.0;
#pragma unroll
    for (int k = 0; k < R2; ++k) {
        const bool live = inl && rest0 + k < h.dr;
        const int r = r0 + k;
        if (op.offstate && live) digits_energy(op, r, dig, tid);
        if (!live) continue;
        double v[NS], dv[NS];
        if constexpr (ITER) {
            const int64_t it = *st.iter;
            const double *q = st.V + slot_of(it, st.nslots) * (int64_t)op.N;
#pragma unroll
            for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r];
            diag_apply<NS>(op, r, dig, tid, v, dv);
            TAcc *yo = reinterpret_cast<TAcc *>(st.y);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                yo[(size_t)s * op.D + r] = y[k][s];
                part += v[s] * (dv[s] + (double)y[k][s]);
            }
        } else {
#pragma unroll
            for (int s = 0; s < NS; ++s) v[s] = xfull[(size_t)s * op.D + r];
            diag_apply<NS>(op, r, dig, tid, v, dv);
#pragma unroll
            for (int s = 0; s < NS; ++s) yfull[(size_t)s * op.D + r] = dv[s] + (double)y[k][s];
        }
    }
    if constexpr (ITER) {
        const double bs = block_sum(part);
        if (tid == 0) st.partials[b] = bs;
    }
}

// ------------------------------------------------------------------------
// Fast-block H*v (tables in NvFbDev).  A wave owns one block of P <= 64
// consecutive rows, lane p = row row0 + p: the trailing ("fast") modes vary
// across the lanes and the slow digits are the same for the whole wave.
//   1. every distinct slow monomial S: w_S[p] = sum f_S(Delta_s) x[row0 +
//      Delta_s + p] (wave-uniform factors from a per-wave LDS table, offsets
//      in the scalar offset field of the buffer loads), mixed into the
//      channels V_F += C_t w_S of the fast monomials it pairs with (V_1 = y);
//   2. every channel F > 1: y[p] += sum h_F(p, Delta_f) V_F[p + Delta_f]
//      through a wave-private LDS row with zero margins;
//   3. the remaining terms without slow steps, one by one from x[row0 + p].
// An out-of-basis slow step has factor 0 and reads the first rows of the
// zero margin that the interleaved copy of x carries, which all such steps
// share; out-of-block fast neighbours carry factor 0.
// ------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ const T *fbtab_of(const NvOpDev &op);
template <> __device__ __forceinline__ const double *fbtab_of<double>(const NvOpDev &op) { return op.fb.tab64; }
template <> __device__ __forceinline__ const float *fbtab_of<float>(const NvOpDev &op) { return op.fb.tab32; }

// The four factors h(p, n, -s + 2i) of a slow slot at the wave's digit.
template <typename T> struct Fb4 { T v[4]; };

// V_c += C_{cm[c]} w for the channels c < NC that the slow monomial d feeds
// (channel 0 without the middle step when d.excl0).
template <int NS, typename TAcc, int NC, bool ODD>
__device__ __forceinline__ void fb_mix(const TAcc *__restrict__ cm, const NvFbS &d, const TAcc (&w)[NS],
                                       const TAcc (&wf)[NS], TAcc (&V)[NV_FB_NCH][NS]) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (d.cm[c] < 0) continue;
        const TAcc *C = cm + (size_t)d.cm[c] * NS * NS;
        const bool ex = c == 0 && ODD && d.excl0;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            TAcc a = V[c][s];
#pragma unroll
            for (int s2 = 0; s2 < NS; ++s2) a += C[s * NS + s2] * (ex ? w[s2] : wf[s2]);
            V[c][s] = a;
        }
    }
}

// The slow monomials [t0, t1) of shape (K0, K1, K2), gathered and mixed into
// the channels V; vl is the lane's byte offset in its block, rb the byte
// offset of the block's first row in the padded buffer, fw/fwm the wave's
// factor rows and their masks of slots inside the basis.  The loop is
// software-pipelined: monomial t+1's rows are loaded while monomial t's are
// summed and mixed, each load reusing the registers of the row just summed,
// so a wave keeps one monomial's loads in flight through the mixing.
template <int NS, typename TIn, typename TAcc, int NC, int K0, int K1, int K2>
__device__ __forceinline__ void fb_slow(const NvOpDev &op, const XRes<NS, TIn> &xs, unsigned vl, unsigned rb,
                                        const TAcc *fw, const int *fwm, int t0, int t1,
                                        TAcc (&V)[NV_FB_NCH][NS]) {
    using X = XRes<NS, TIn>;
    constexpr unsigned kEl = X::kEl;
    constexpr int J0 = K0 ? K0 : 1, J1 = K1 ? K1 : 1, J2 = K2 ? K2 : 1, J = J0 * J1 * J2;
    // all stepping slots odd: the middle step is Delta_s = 0
    constexpr bool odd = K0 > 0 && (K0 & 1) && (K1 == 0 || (K1 & 1)) && (K2 == 0 || (K2 & 1));
    if (t0 >= t1) return;
    const TAcc *__restrict__ cm = cmat_of<TAcc>(op);
    // scalar byte offsets of monomial t's steps, in (i0, i1, i2) order; a step
    // outside the basis reads the zero margin at offset 0
    auto offsets = [&](int t, unsigned (&so)[J]) {
        const NvFbS &d = op.fb.S[t];
        // slot i of slow slot j is inside the basis when bit i of m_j is set
        int m0 = 1, m1 = 1, m2 = 1;
        if constexpr (K0 > 0) m0 = __builtin_amdgcn_readfirstlane(fwm[d.tab[0]]);
        if constexpr (K1 > 0) m1 = __builtin_amdgcn_readfirstlane(fwm[d.tab[1]]);
        if constexpr (K2 > 0) m2 = __builtin_amdgcn_readfirstlane(fwm[d.tab[2]]);
        const unsigned b0 = rb + (unsigned)d.base * kEl;
        const unsigned ib0 = (unsigned)d.inc[0] * kEl, ib1 = (unsigned)d.inc[1] * kEl,
                       ib2 = (unsigned)d.inc[2] * kEl;
#pragma unroll
        for (int i0 = 0; i0 < J0; ++i0)
#pragma unroll
            for (int i1 = 0; i1 < J1; ++i1)
#pragma unroll
                for (int i2 = 0; i2 < J2; ++i2) {
                    unsigned o = rb;
                    if constexpr (K0 > 0) o = b0 + (unsigned)i0 * ib0;
                    if constexpr (K1 > 0) o += (unsigned)i1 * ib1;
                    if constexpr (K2 > 0) o += (unsigned)i2 * ib2;
                    const bool in = ((m0 >> i0) & (K1 > 0 ? m1 >> i1 : 1) & (K2 > 0 ? m2 >> i2 : 1) & 1) != 0;
                    so[(i0 * J1 + i1) * J2 + i2] = in ? o : 0u;
                }
    };
    typename X::Row xr[J];
    {
        unsigned so[J];
        offsets(t0, so);
#pragma unroll
        for (int k = 0; k < J; ++k) {
            xr[k] = xs.load_so(vl, so[k]);
            // loads leave in step order, the order the loop consumes them
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    for (int t = t0; t < t1; ++t) {
        const NvFbS &d = op.fb.S[t];
        Fb4<TAcc> f0, f1, f2;
        if constexpr (K0 > 0) f0 = *reinterpret_cast<const Fb4<TAcc> *>(fw + d.tab[0] * 4);
        if constexpr (K1 > 0) f1 = *reinterpret_cast<const Fb4<TAcc> *>(fw + d.tab[1] * 4);
        if constexpr (K2 > 0) f2 = *reinterpret_cast<const Fb4<TAcc> *>(fw + d.tab[2] * 4);
        // w: every step but the middle one of an all-odd shape, wc: that one
        TAcc w[NS], wc[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) w[s] = wc[s] = (TAcc)0;
        // sums the rows in flight; with next, each row's registers then take
        // the next monomial's row
        auto sum_rows = [&](auto next, const unsigned (&sn)[J]) {
#pragma unroll
            for (int i0 = 0; i0 < J0; ++i0)
#pragma unroll
                for (int i1 = 0; i1 < J1; ++i1)
#pragma unroll
                    for (int i2 = 0; i2 < J2; ++i2) {
                        const int k = (i0 * J1 + i1) * J2 + i2;
                        TAcc h = (TAcc)1;
                        if constexpr (K0 > 0) h = f0.v[i0];
                        if constexpr (K1 > 0) h *= f1.v[i1];
                        if constexpr (K2 > 0) h *= f2.v[i2];
                        const bool mid = odd && 2 * i0 + 1 == K0 && (K1 == 0 || 2 * i1 + 1 == K1) &&
                                         (K2 == 0 || 2 * i2 + 1 == K2);
                        if (mid) X::template fma_row<TAcc>(xr[k], h, wc);
                        else X::template fma_row<TAcc>(xr[k], h, w);
                        if constexpr (decltype(next)::value) {
                            xr[k] = xs.load_so(vl, sn[k]);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
        };
        unsigned sn[J];
        if (t + 1 < t1) {
            offsets(t + 1, sn);
            sum_rows(std::true_type{}, sn);
        } else {
            sum_rows(std::false_type{}, sn);
        }
        if (d.nd > 0) {
            TAcc fd = fw[d.dtab[0] * 4];
            if (d.nd > 1) fd *= fw[d.dtab[1] * 4];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                w[s] *= fd;
                if constexpr (odd) wc[s] *= fd;
            }
        }
        // wc stays zero unless the shape has a middle step
        TAcc wf[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) wf[s] = odd ? w[s] + wc[s] : w[s];
        fb_mix<NS, TAcc, NC, odd>(cm, d, w, wf, V);
    }
}

// y += sum_{steps [s0, s1)} h(p) src[p + stride]: the fast steps of one
// channel or term over the wave's LDS row vb, src published there first.
template <int NS, typename TAcc>
__device__ __forceinline__ void fb_fast(const NvOpDev &op, int s0, int s1, TAcc *vb, int lane, const uint8_t *dig,
                                        const TAcc *ftab, int tid, const TAcc (&src)[NS], TAcc (&y)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; ++s) vb[lane * NS + s] = src[s];
    __builtin_amdgcn_wave_barrier();
    const int nmax = op.nmax;
    for (int k = s0; k < s1; ++k) {
        const NvStep st = op.fb.steps[k];
        TAcc h = ftab[st.code[0] * nmax + dig[st.mode[0] * kBlock + tid]];
#pragma unroll
        for (int j = 1; j < NV_MAXF; ++j)
            if (j < st.nf) h *= ftab[st.code[j] * nmax + dig[st.mode[j] * kBlock + tid]];
        const TAcc *v = vb + (lane + st.stride) * NS;
#pragma unroll
        for (int s = 0; s < NS; ++s) y[s] += h * v[s];
    }
    __builtin_amdgcn_wave_barrier();
}

// The off-diagonal part of the wave's block.
template <int NS, typename TIn, typename TAcc>
__device__ __forceinline__ void fb_offdiag(const NvOpDev &op, const XRes<NS, TIn> &xs, unsigned vl, unsigned rb,
                                           int lane, const uint8_t *dig, const TAcc *fw, const int *fwm,
                                           const TAcc *ftab, int tid, TAcc *vb, TAcc (&y)[NS]) {
    TAcc V[NV_FB_NCH][NS];
#pragma unroll
    for (int c = 0; c < NV_FB_NCH; ++c)
#pragma unroll
        for (int s = 0; s < NS; ++s) V[c][s] = (TAcc)0;
    // phase 0: the slow monomials that feed a channel beyond the first four (degree <= 1);
    // those channels are complete afterwards and go through LDS, which frees their registers
    const int *g0 = op.fb.sgrp[0];
    fb_slow<NS, TIn, TAcc, NV_FB_NCH, 2, 0, 0>(op, xs, vl, rb, fw, fwm, g0[0], g0[1], V);
    fb_slow<NS, TIn, TAcc, NV_FB_NCH, 0, 0, 0>(op, xs, vl, rb, fw, fwm, g0[7], g0[8], V);
#pragma unroll
    for (int c = 4; c < NV_FB_NCH; ++c)
        if (c < op.fb.nch) fb_fast<NS, TAcc>(op, op.fb.chs[c], op.fb.chs[c + 1], vb, lane, dig, ftab, tid, V[c], y);
    // phase 1: channels 0..3 (slow degree <= 2)
    const int *g1 = op.fb.sgrp[1];
    fb_slow<NS, TIn, TAcc, 4, 2, 0, 0>(op, xs, vl, rb, fw, fwm, g1[0], g1[1], V);
    fb_slow<NS, TIn, TAcc, 4, 3, 0, 0>(op, xs, vl, rb, fw, fwm, g1[1], g1[2], V);
    fb_slow<NS, TIn, TAcc, 4, 2, 2, 0>(op, xs, vl, rb, fw, fwm, g1[3], g1[4], V);
    fb_slow<NS, TIn, TAcc, 4, 0, 0, 0>(op, xs, vl, rb, fw, fwm, g1[7], g1[8], V);
#pragma unroll
    for (int c = 1; c < 4; ++c)
        if (c < op.fb.nch) fb_fast<NS, TAcc>(op, op.fb.chs[c], op.fb.chs[c + 1], vb, lane, dig, ftab, tid, V[c], y);
    // phase 2: channel 0 only, where most of the slow steps are
    const int *g2 = op.fb.sgrp[2];
    fb_slow<NS, TIn, TAcc, 1, 2, 0, 0>(op, xs, vl, rb, fw, fwm, g2[0], g2[1], V);
    fb_slow<NS, TIn, TAcc, 1, 3, 0, 0>(op, xs, vl, rb, fw, fwm, g2[1], g2[2], V);
    fb_slow<NS, TIn, TAcc, 1, 4, 0, 0>(op, xs, vl, rb, fw, fwm, g2[2], g2[3], V);
    fb_slow<NS, TIn, TAcc, 1, 2, 2, 0>(op, xs, vl, rb, fw, fwm, g2[3], g2[4], V);
    fb_slow<NS, TIn, TAcc, 1, 3, 2, 0>(op, xs, vl, rb, fw, fwm, g2[4], g2[5], V);
    fb_slow<NS, TIn, TAcc, 1, 2, 3, 0>(op, xs, vl, rb, fw, fwm, g2[5], g2[6], V);
    fb_slow<NS, TIn, TAcc, 1, 2, 2, 2>(op, xs, vl, rb, fw, fwm, g2[6], g2[7], V);
    fb_slow<NS, TIn, TAcc, 1, 0, 0, 0>(op, xs, vl, rb, fw, fwm, g2[7], g2[8], V);
#pragma unroll
    for (int s = 0; s < NS; ++s) y[s] += V[0][s];
    if (op.fb.npure > 0) {
        TAcc xr[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) xr[s] = (TAcc)0;
        xs.template fma_so<TAcc>(vl, rb, (TAcc)1, xr);
        const TAcc *__restrict__ cm = cmat_of<TAcc>(op);
        for (int k = 0; k < op.fb.npure; ++k) {
            const NvFbPure u = op.fb.pure[k];
            TAcc fd = (TAcc)1;
            if (u.nd > 0) fd = fw[u.dtab[0] * 4];
            if (u.nd > 1) fd *= fw[u.dtab[1] * 4];
            const TAcc *C = cm + (size_t)u.cm * NS * NS;
            TAcc v[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                TAcc a = (TAcc)0;
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) a += C[s * NS + s2] * xr[s2];
                v[s] = fd * a;
            }
            fb_fast<NS, TAcc>(op, u.s0, u.s1, vb, lane, dig, ftab, tid, v, y);
        }
    }
}

// LDS of the fast-block kernels: factor table, the digits of one row per
// thread, four wave rows of (64 + 2 halo) lanes x NS values, four wave
// tables of ntab factor rows x 4 and four of ntab validity masks.
__host__ __device__ inline size_t fb_lds_layout(const NvOpDev &op, size_t el, size_t *dig_off, size_t *vb_off,
                                                size_t *fw_off) {
    const size_t a = ((size_t)op.ncodes * op.nmax * el + 15) & ~(size_t)15;
    const size_t b = a + (((size_t)op.M * kBlock + 15) & ~(size_t)15);
    const size_t c = b + ((4 * (size_t)(64 + 2 * op.fb.halo) * op.ns * el + 15) & ~(size_t)15);
    if (dig_off) *dig_off = a;
    if (vb_off) *vb_off = b;
    if (fw_off) *fw_off = c;
    // factor rows, then one validity mask per row
    return c + 4 * (size_t)op.fb.ntab * (4 * el + sizeof(int));
}

// Common prologue: factor table and digits in LDS, the wave's block, its LDS
// row and its factor rows.
template <typename TAcc> struct FbWave {
    const TAcc *ftab;
    uint8_t *dig;
    TAcc *vb;
    const TAcc *fw;
    const int *fwm;
    int lane, wave, blk, row0;
    bool inblk, live;
};
template <typename TAcc> __device__ __forceinline__ FbWave<TAcc> fb_prologue(const NvOpDev &op, int b) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    FbWave<TAcc> w;
    size_t dig_off, vb_off, fw_off;
    fb_lds_layout(op, sizeof(TAcc), &dig_off, &vb_off, &fw_off);
    TAcc *ft = reinterpret_cast<TAcc *>(lds_raw);
    const int nt = op.ncodes * op.nmax;
    const TAcc *g = ftab_of<TAcc>(op);
    for (int i = threadIdx.x; i < nt; i += blockDim.x) ft[i] = g[i];
    w.ftab = ft;
    w.dig = lds_raw + dig_off;
    const int tid = threadIdx.x;
    w.lane = tid & 63;
    w.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    w.blk = b * 4 + w.wave;
    w.inblk = w.blk < op.fb.nblk;
    {
        // block t of the sweep: its slow digits in the sweep's mode order
        int t = w.blk, row0 = 0;
        for (int j = 0; j < op.fb.nord; ++j) {
            const int nf = op.fb.ordnf[j];
            const int q = t / nf;
            row0 += (t - q * nf) * op.fb.ordshft[j];
            t = q;
        }
        w.row0 = row0;
    }
    w.live = w.inblk && w.lane < op.fb.P;
    if (w.live) digits_energy(op, w.row0 + w.lane, w.dig, tid);
    else for (int m = 0; m < op.M; ++m) w.dig[m * kBlock + tid] = 0;
    const int rowlen = (64 + 2 * op.fb.halo) * op.ns;
    TAcc *vrow = reinterpret_cast<TAcc *>(lds_raw + vb_off) + (size_t)w.wave * rowlen;
    for (int i = w.lane; i < rowlen; i += 64) vrow[i] = (TAcc)0;
    w.vb = vrow + (size_t)op.fb.halo * op.ns;
    __syncthreads();
    // the wave's factor rows at its slow digits (those of its first lane), and
    // their masks of slots inside the basis (factor nonzero)
    TAcc *fw = reinterpret_cast<TAcc *>(lds_raw + fw_off) + (size_t)w.wave * op.fb.ntab * 4;
    int *fwm = reinterpret_cast<int *>(lds_raw + fw_off + 4 * (size_t)op.fb.ntab * 4 * sizeof(TAcc)) +
               (size_t)w.wave * op.fb.ntab;
    const TAcc *tab = fbtab_of<TAcc>(op);
    for (int k = w.lane; k < op.fb.ntab; k += 64) {
        const int n = w.dig[op.fb.tabmode[k] * kBlock + w.wave * 64];
        int m = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const TAcc f = tab[((size_t)k * op.nmax + n) * 4 + i];
            fw[k * 4 + i] = f;
            m |= (f != (TAcc)0) << i;
        }
        fwm[k] = m;
    }
    __builtin_amdgcn_wave_barrier();
    w.fw = fw;
    w.fwm = fwm;
    return w;
}

// Iteration H*v in the fast-block form: y = offdiag(H) q_i, alpha partials.
// Waves per SIMD the compiler must fit (register budget): 8 for FP32 (60
// VGPRs) and 4 for FP64 (112).  Four states, 22M rows: FP32 10.7 ms at 6, 7
// or 8 waves; FP64 23.6 ms at 3 or 4 waves and 28.0 ms at 5, where it spills.
// (Before the slow monomials were split into phases, the ten channels held 40
// VGPRs across all of them and the FP32 kernel waited on every load.)
#ifndef NV_FB_WAVES32
#define NV_FB_WAVES32 8
#endif
#ifndef NV_FB_WAVES64
#define NV_FB_WAVES64 4
#endif
// (Above four states the ten channels of NS values need more registers: the
// budget follows the state count.)
constexpr int fb_waves(int ns, int bytes) {
    return ns <= 4 ? (bytes == 8 ? NV_FB_WAVES64 : NV_FB_WAVES32) : (bytes == 8 ? 1 : 2);
}
template <int NS, typename TIn, typename TAcc>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(fb_waves(NS, sizeof(TAcc))))) void
hv_fb_iter_kernel(NvOpDev op, LzState st) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const FbWave<TAcc> w = fb_prologue<TAcc>(op, b);
    const int tid = threadIdx.x;
    const int64_t it = *st.iter;
    const double *q = st.V + slot_of(it, st.nslots) * (int64_t)op.N;
    using TX = typename std::conditional<std::is_same<TIn, F64I>::value, double, TIn>::type;
    const TX *x = std::is_same<TIn, F64I>::value ? reinterpret_cast<const TX *>(st.v64i)
                                                 : reinterpret_cast<const TX *>(st.vlo);
    constexpr int kNsp = nv_lo_stride(NS);
    constexpr unsigned kEl = XRes<NS, TIn>::kEl;
    const XRes<NS, TIn> xs(x - (size_t)op.fb.pad * kNsp, op.D + 2 * op.fb.pad);
    TAcc y[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) y[s] = (TAcc)0;
    // a lane past the block's P rows loads nothing (offset outside the buffer)
    const unsigned vl = w.live ? (unsigned)w.lane * kEl : 0x80000000u;
    const unsigned rb = (unsigned)(w.row0 + op.fb.pad) * kEl;
    if (w.inblk) fb_offdiag<NS, TIn, TAcc>(op, xs, vl, rb, w.lane, w.dig, w.fw, w.fwm, w.ftab, tid, w.vb, y);
    double part = 0.0;
    // q_i, dg and y are touched once per launch (non-temporal on the largest vectors, so that
    // they do not displace the rows of x that neighbouring blocks gather from L2)
    auto epilogue = [&](auto nt) {
        constexpr bool NT = decltype(nt)::value;
        const int r = w.row0 + w.lane;
        double v[NS], dv[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = ld_once<NT>(&q[(size_t)s * op.D + r]);
#pragma unroll
        for (int s = 0; s < NS; ++s) dv[s] = ld_once<NT>(&op.dg[(size_t)s * op.D + r]) * v[s];
        if (op.offstate) diag_offstate<NS>(op, w.dig, tid, v, dv);
        TAcc *yo = reinterpret_cast<TAcc *>(st.y);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            st_once<NT>(y[s], &yo[(size_t)s * op.D + r]);
            part += v[s] * (dv[s] + (double)y[s]);
        }
    };
    if (w.live) {
        if (op.stream_nt) epilogue(std::true_type{});
        else epilogue(std::false_type{});
    }
    const double bs = block_sum(part);
    if (tid == 0) st.partials[b] = bs;
}

// Standalone y = H x in the fast-block form: xin is the interleaved copy of x
// (FP32 or FP64, with its margins), x the FP64 vector for the diagonal part.
template <int NS, typename TIn, typename TAcc>
__global__ __launch_bounds__(kBlock) void hv_fb_full_kernel(NvOpDev op, const void *__restrict__ xin,
                                                            const double *__restrict__ x, double *__restrict__ yfull) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const FbWave<TAcc> w = fb_prologue<TAcc>(op, b);
    const int tid = threadIdx.x;
    using TX = typename std::conditional<std::is_same<TIn, F64I>::value, double, TIn>::type;
    constexpr int kNsp = nv_lo_stride(NS);
    constexpr unsigned kEl = XRes<NS, TIn>::kEl;
    const XRes<NS, TIn> xs(reinterpret_cast<const TX *>(xin) - (size_t)op.fb.pad * kNsp, op.D + 2 * op.fb.pad);
    TAcc y[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) y[s] = (TAcc)0;
    const unsigned vl = w.live ? (unsigned)w.lane * kEl : 0x80000000u;
    const unsigned rb = (unsigned)(w.row0 + op.fb.pad) * kEl;
    if (w.inblk) fb_offdiag<NS, TIn, TAcc>(op, xs, vl, rb, w.lane, w.dig, w.fw, w.fwm, w.ftab, tid, w.vb, y);
    if (w.live) {
        const int r = w.row0 + w.lane;
        double v[NS], dv[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = x[(size_t)s * op.D + r];
        diag_apply<NS>(op, r, w.dig, tid, v, dv);
#pragma unroll
        for (int s = 0; s < NS; ++s) yfull[(size_t)s * op.D + r] = dv[s] + (double)y[s];
    }
}

// The finishing kernel's sum of parts[0, n) (finish_sum_kernel: strided
// per-thread sums, then the block tree), in every workgroup of kBlock threads.
__device__ __forceinline__ double fold_sum(const double *parts, int n) {
    double s = 0.0;
#pragma unroll 8
    for (int i = threadIdx.x; i < n; i += kBlock) s += parts[i];
    return block_sum(s);
}

// lanczos_step (Source/top.f90:820-842): r = (H q_i - alpha q_i) - beta(i-1) q_{i-1}
// stored in slot i+1, beta(i) = |r|.
template <int NS, typename TAcc, bool NT>
__global__ __launch_bounds__(kBlock) void lanczos_step_kernel(NvOpDev op, LzState st) {
    extern __shared__ __align__(16) unsigned char dig[];  // M*kBlock bytes when states couple through D
    const int tid = threadIdx.x;
    const int r = blockIdx.x * kBlock + tid;
    const int64_t it = *st.iter;
    const double *q = st.V + slot_of(it, st.nslots) * (int64_t)op.N;
    const double *qm = st.V + slot_of(it - 1, st.nslots) * (int64_t)op.N;
    double *res = st.V + slot_of(it + 1, st.nslots) * (int64_t)op.N;
    double a;
    if (st.fold_in) {
        a = fold_sum(st.fold_in, st.fold_n);
        if (blockIdx.x == 0 && tid == 0) st.alpha[it] = a;
    } else {
        a = st.alpha[it];
    }
    const double b = it > 1 ? st.beta[it - 1] : 0.0;
    double part = 0.0;
    if (r < op.D) {
        if (op.offstate) digits_energy(op, r, dig, tid);
        double v[NS], dv[NS];
        // every vector is streamed once
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = ld_once<NT>(&q[(size_t)s * op.D + r]);
#pragma unroll
        for (int s = 0; s < NS; ++s) dv[s] = ld_once<NT>(&op.dg[(size_t)s * op.D + r]) * v[s];
        if (op.offstate) diag_offstate<NS>(op, dig, tid, v, dv);
        const TAcc *y = reinterpret_cast<const TAcc *>(st.y);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const size_t o = (size_t)s * op.D + r;
            double w = dv[s] + (double)ld_once<NT>(&y[o]);
            w = w - a * v[s];
            if (it > 1) w = w - b * ld_once<NT>(&qm[o]);
            st_once<NT>(w, &res[o]);
            part += w * w;
        }
    }
    const double bs = block_sum(part);
    if (tid == 0) (st.fold_out ? st.fold_out : st.partials)[blockIdx.x] = bs;
    if (st.next_iter && blockIdx.x == 0 && tid == 0) *st.next_iter = it + st.next_inc;
}


// One interleaved row of K values (the H*v input copies), aligned for one
// wide store where its size allows.
__host__ __device__ constexpr size_t il_align(size_t b) { return b % 16 == 0 ? 16 : b % 8 == 0 ? 8 : 4; }
template <typename T, int K> struct alignas(il_align(sizeof(T) * K) < alignof(T) ? alignof(T)
                                                                                  : il_align(sizeof(T) * K)) IlRow {
    T v[K];
};

// update_lanczos (Source/top.f90:646-670): q_i = r / beta(i-1) in place, and
// the copies of q_i that H*v and re-orthogonalisation read.  A thread takes
// the NS states of row r, so an interleaved copy gets one row store.
template <int NS, typename TIn, bool NT>
__global__ __launch_bounds__(kBlock) void normalize_rows_kernel(int64_t N, int64_t D, LzState st) {
    constexpr int NSP = nv_lo_stride(NS);
    const int64_t it = *st.iter;
    double *q = st.V + slot_of(it, st.nslots) * N;
    double bprev;
    if (st.fold_in) {
        bprev = sqrt(fold_sum(st.fold_in, st.fold_n));
        if (blockIdx.x == 0 && threadIdx.x == 0) st.beta[it - 1] = bprev;
    } else {
        bprev = st.beta[it - 1];
    }
    const double sc = 1.0 / bprev;
    TIn *lo = reinterpret_cast<TIn *>(st.vlo);
    float *vf = st.Vf ? st.Vf + slot_of(it, st.nslots) * st.ldf : nullptr;
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < D; r += (int64_t)gridDim.x * kBlock) {
        double v[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            v[s] = ld_once<NT>(&q[s * D + r]) * sc;
            st_once<NT>(v[s], &q[s * D + r]);
            if (vf) vf[s * D + r] = (float)v[s];
        }
        if (lo) {
            IlRow<TIn, NSP> row;
#pragma unroll
            for (int k = 0; k < NSP; ++k) row.v[k] = from_d<TIn>(k < NS ? v[k] : 0.0);
            reinterpret_cast<IlRow<TIn, NSP> *>(lo)[r] = row;
        }
        if (st.v64i) {
            IlRow<double, NSP> row;
#pragma unroll
            for (int k = 0; k < NSP; ++k) row.v[k] = k < NS ? v[k] : 0.0;
            reinterpret_cast<IlRow<double, NSP> *>(st.v64i)[r] = row;
        }
    }
}

__global__ void set_iter_kernel(int64_t *it, int64_t v, int inc) {
    if (threadIdx.x == 0) *it = inc ? *it + 1 : v;
}

__global__ __launch_bounds__(kBlock) void fill_start_kernel(double *v, int64_t D, int ns, const double *vals) {
    const int64_t N = D * ns;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (int64_t)gridDim.x * kBlock) {
        const int64_t s = i / D, r = i - s * D;
        v[i] = r == 0 ? vals[s] : 0.0;
    }
}

__global__ __launch_bounds__(kBlock) void norm_kernel(const double *v, int64_t n, double *partials) {
    double p = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        p += v[i] * v[i];
    const double b = block_sum(p);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// Sums nparts per-block partials in block order (one block; the kernel
// boundary makes the partials visible) and stores the sum, or its square
// root, at out[*iter] (iter null: out[0]); then, with advance, sets
// *advance = *iter + 1 (the next iteration's counter, in place of a launch).
__global__ __launch_bounds__(kBlock) void finish_sum_kernel(const double *partials, int nparts, double *out,
                                                           const int64_t *iter, int take_sqrt, int64_t *advance) {
    double s = 0.0;
    // loads unrolled ahead of the adds, which keep their order (32 in flight per
    // thread: 86,400 partials are 338 per thread)
#pragma unroll 32
    for (int i = threadIdx.x; i < nparts; i += kBlock) s += partials[i];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        const int64_t k = iter ? *iter : 0;
        out[k] = take_sqrt ? sqrt(s) : s;
        if (advance) *advance = k + 1;
    }
}

hipError_t launch_finish_sum(const double *partials, int nparts, double *out, const int64_t *iter, int take_sqrt,
                             hipStream_t s, int64_t *advance) {
    hipLaunchKernelGGL(finish_sum_kernel, dim3(1), dim3(kBlock), 0, s, partials, nparts, out, iter, take_sqrt,
                       advance);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// host launchers for the above
// ------------------------------------------------------------------------
size_t hv_lds_bytes(const NvOpDev &op, int prec) {
    const size_t el = prec == 0 ? sizeof(double) : sizeof(float);
    size_t off = ((size_t)op.ncodes * op.nmax * el + 15) & ~(size_t)15;
    return off + 2 * (size_t)op.M * kBlock;  // digits of up to two rows per thread
}

int hv_blocks(const NvOpDev &op) { return (op.D + kBlock - 1) / kBlock; }

static int grid_stride_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    return (int)(b < 8192 ? b : 8192);
}

// Above op.skip_min_bytes of vector (set per context: kSkipMinBytes, or less
// for operators with many out-of-basis steps) the shaped kernels point the
// loads of out-of-basis steps outside the buffer (no memory traffic) instead
// of at a wrong row.
constexpr int64_t kSkipMinBytes = (int64_t)128 << 20;
constexpr int kShapedMaxStates = NV_MAXS;

// The shaped kernels serve operators whose terms all have a shape (ordr <= 3),
// any number of states (up to NV_MAXS), FP64 or FP32, with every row offset inside a 32-bit
// buffer range; op.tshape = nullptr selects the step-table kernels.
static bool shaped_ok(const NvOpDev &op, int64_t el) {
    if (!op.tshape || op.ns > kShapedMaxStates) return false;
    return ((int64_t)op.D + 2 * kBlock) * el < ((int64_t)1 << 31);
}

// The Lanczos iteration in FP64 reads the interleaved copy LzState::v64i (one
// 16-byte load per two states).  The standalone and batched H*v read the
// state-major vector, one load per state and step, and keep the step-table
// form above kSkipMinBytes, where its branch over out-of-basis steps wins
// (N = 88.5M, four states: 55.9 against 65.7 ms per H*v).
int64_t hv_skip_min_bytes() { return kSkipMinBytes; }

bool hv_shaped(const NvOpDev &op, int prec) {
    if (prec == 0) return shaped_ok(op, nv_lo_stride(op.ns) * 8);
    if (prec == 1) return shaped_ok(op, nv_lo_stride(op.ns) * 4);
    return false;
}

bool hv_shaped_skip(const NvOpDev &op, int prec) {
    const int64_t vbytes = (int64_t)op.D * nv_lo_stride(op.ns) * (prec == 0 ? 8 : 4);
    return hv_shaped(op, prec) && vbytes > op.skip_min_bytes;
}

// Standalone and batched H*v read the state-major FP64 vector itself.
static bool shaped_state_major(const NvOpDev &op) {
    return shaped_ok(op, 8) && (int64_t)op.D * 8 * op.ns <= kSkipMinBytes;
}

bool hv_shaped_batch(const NvOpDev &op, int prec) { return prec == 0 ? shaped_state_major(op) : hv_shaped(op, prec); }

template <int NS>
static hipError_t hv_full_ns(const NvOpDev &op, int prec, const double *x, double *y, hipStream_t s) {
    const size_t lds = hv_lds_bytes(op, prec);
    if constexpr (NS <= kShapedMaxStates) {
        if (shaped_state_major(op)) {
            if (prec == 0)
                hipLaunchKernelGGL((hv_shaped_full_kernel<NS, double>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op,
                                   x, y);
            else
                hipLaunchKernelGGL((hv_shaped_full_kernel<NS, float>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op,
                                   x, y);
            return hipGetLastError();
        }
    }
    if (prec == 0)
        hipLaunchKernelGGL((hv_full_kernel<NS, double>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op, x, y);
    else
        hipLaunchKernelGGL((hv_full_kernel<NS, float>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op, x, y);
    return hipGetLastError();
}

// Two rows per thread pays for the step-table kernel in FP32 on large
// operators (four states, 22M rows: 36.7 -> 29.7 ms per H*v); in FP64 (55.6 ->
// 62.2 ms) and below a few million rows (three states, 0.46M rows: 53.8 ->
// 57.4 us) the one-row kernel is faster.
constexpr int kTwoRowMinRows = 1 << 22;

template <int NS>
static hipError_t hv_iter_ns(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    const size_t lds = hv_lds_bytes(op, prec);
    if constexpr (NS <= kShapedMaxStates) {
        // FP64 runs shaped when the caller keeps the interleaved copy
        // one row per thread: the unrolled terms give the loads enough independence
        // (four states, 22M rows: FP32 17.1 ms against 17.8 ms with two rows; FP64
        // 40.4 against 46.8 ms)
        if (hv_shaped(op, prec) && (prec == 1 || st.v64i)) {
            const unsigned g = (unsigned)hv_blocks(op);
            if (hv_shaped_skip(op, prec)) {
                if (prec == 0)
                    hipLaunchKernelGGL((hv_shaped_iter_kernel<NS, F64I, double, 1, true>), dim3(g), dim3(kBlock), lds,
                                       s, op, st);
                else
                    hipLaunchKernelGGL((hv_shaped_iter_kernel<NS, float, float, 1, true>), dim3(g), dim3(kBlock), lds,
                                       s, op, st);
                return hipGetLastError();
            }
            if (prec == 0)
                hipLaunchKernelGGL((hv_shaped_iter_kernel<NS, F64I, double, 1, false>), dim3(g), dim3(kBlock), lds,
                                   s, op, st);
            else
                hipLaunchKernelGGL((hv_shaped_iter_kernel<NS, float, float, 1, false>), dim3(g), dim3(kBlock), lds,
                                   s, op, st);
            return hipGetLastError();
        }
    }
    if (prec != 0 && op.D >= kTwoRowMinRows) {
        const unsigned g = (unsigned)((hv_blocks(op) + 1) / 2);
        if (prec == 1)
            hipLaunchKernelGGL((hv_iter2_kernel<NS, float, float>), dim3(g), dim3(kBlock), lds, s, op, st);
        else
            hipLaunchKernelGGL((hv_iter2_kernel<NS, __hip_bfloat16, float>), dim3(g), dim3(kBlock), lds, s, op, st);
        return hipGetLastError();
    }
    if (prec == 0)
        hipLaunchKernelGGL((hv_iter_kernel<NS, double, double>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op, st);
    else if (prec == 1)
        hipLaunchKernelGGL((hv_iter_kernel<NS, float, float>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op, st);
    else
        hipLaunchKernelGGL((hv_iter_kernel<NS, __hip_bfloat16, float>), dim3(hv_blocks(op)), dim3(kBlock), lds, s,
                           op, st);
    return hipGetLastError();
}

// ---- two-pass H*v launchers ----
static size_t hv2_pass1_lds(const NvOpDev &op, int prec);

bool hv2_use(const NvOpDev &op, int prec) {
    if (!op.tshape || !op.hv2.on || (prec != 0 && prec != 1) || (prec == 1 && !op.hv2.f32)) return false;
    const int64_t el = (int64_t)nv_lo_stride(op.ns) * (prec == 0 ? 8 : 4);
    return (int64_t)op.D * el >= op.hv2_min_bytes && ((int64_t)op.D + 2 * kBlock) * el < ((int64_t)1 << 31) &&
           hv2_pass1_lds(op, prec) <= 64 * 1024;
}

size_t hv2_scratch_bytes(const NvOpDev &op, int prec) {
    return (size_t)NV_HV2_NG * op.D * nv_lo_stride(op.ns) * (prec == 0 ? 8 : 4);
}

// Consecutive plane rows per thread of the second pass.
constexpr int kHv2Rows = 4;

int hv2_parts(const NvOpDev &op) {
    const int per = op.hv2.c * kHv2Rows;
    return (op.hv2.dr + per - 1) / per;
}

// The FP32 path keeps the coupling products on the matrix cores up to four states.
static bool hv2_mfma(const NvOpDev &op, int prec) { return prec == 1 && op.ns <= 4 && op.hv2.mfma; }

static size_t hv2_pass1_lds(const NvOpDev &op, int prec) {
    return hv_lds_bytes(op, prec) + (hv2_mfma(op, prec) ? (size_t)op.nterms * 16 * sizeof(float) : 0);
}

template <int NS>
static hipError_t hv2_ns(const NvOpDev &op, int prec, const void *xin, void *Z, const LzState *st,
                         const double *xfull, double *yfull, hipStream_t s) {
    if constexpr (NS > 4) {
        return hipErrorInvalidValue;  // the tables are built for at most four states
    } else {
    const unsigned g1 = (unsigned)hv_blocks(op);
    const size_t l1 = hv2_pass1_lds(op, prec);
    // plain form in FP32 (18.1 against 19.2 ms), skip form in FP64 (37.9 against 40.3 ms)
    const bool skip = op.hv2.skip < 0 ? prec == 0 : op.hv2.skip != 0;
    if (prec == 1) {
        if (hv2_mfma(op, prec)) {
            if (skip)
                hipLaunchKernelGGL((hv2_pass1_kernel<NS, float, float, true, true>), dim3(g1), dim3(kBlock), l1, s,
                                   op, xin, (float *)Z);
            else
                hipLaunchKernelGGL((hv2_pass1_kernel<NS, float, float, true, false>), dim3(g1), dim3(kBlock), l1, s,
                                   op, xin, (float *)Z);
        } else {
            if (skip)
                hipLaunchKernelGGL((hv2_pass1_kernel<NS, float, float, false, true>), dim3(g1), dim3(kBlock), l1, s,
                                   op, xin, (float *)Z);
            else
                hipLaunchKernelGGL((hv2_pass1_kernel<NS, float, float, false, false>), dim3(g1), dim3(kBlock), l1,
                                   s, op, xin, (float *)Z);
        }
    } else {
        if (skip)
            hipLaunchKernelGGL((hv2_pass1_kernel<NS, F64I, double, false, true>), dim3(g1), dim3(kBlock), l1, s, op,
                               xin, (double *)Z);
        else
            hipLaunchKernelGGL((hv2_pass1_kernel<NS, F64I, double, false, false>), dim3(g1), dim3(kBlock), l1, s,
                               op, xin, (double *)Z);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned g2 = (unsigned)hv2_parts(op);
    const unsigned b2 = (unsigned)(op.hv2.nl * op.hv2.c);
    const size_t l2 = (size_t)op.M * kBlock;
    const LzState none{};
    if (st) {
        if (prec == 1)
            hipLaunchKernelGGL((hv2_pass2_kernel<NS, float, float, true, kHv2Rows>), dim3(g2), dim3(b2), l2, s, op, *st, xin,
                               (const float *)Z, nullptr, nullptr);
        else
            hipLaunchKernelGGL((hv2_pass2_kernel<NS, F64I, double, true, kHv2Rows>), dim3(g2), dim3(b2), l2, s, op, *st, xin,
                               (const double *)Z, nullptr, nullptr);
    } else {
        if (prec == 1)
            hipLaunchKernelGGL((hv2_pass2_kernel<NS, float, float, false, kHv2Rows>), dim3(g2), dim3(b2), l2, s, op, none, xin,
                               (const float *)Z, xfull, yfull);
        else
            hipLaunchKernelGGL((hv2_pass2_kernel<NS, F64I, double, false, kHv2Rows>), dim3(g2), dim3(b2), l2, s, op, none, xin,
                               (const double *)Z, xfull, yfull);
    }
    return hipGetLastError();
    }
}

template <int NS>
static hipError_t step_ns(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    const size_t lds = op.offstate ? (size_t)op.M * kBlock : 0;
    if (prec == 0 && op.stream_nt)
        hipLaunchKernelGGL((lanczos_step_kernel<NS, double, true>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op, st);
    else if (prec == 0)
        hipLaunchKernelGGL((lanczos_step_kernel<NS, double, false>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op, st);
    else if (op.stream_nt)
        hipLaunchKernelGGL((lanczos_step_kernel<NS, float, true>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op, st);
    else
        hipLaunchKernelGGL((lanczos_step_kernel<NS, float, false>), dim3(hv_blocks(op)), dim3(kBlock), lds, s, op, st);
    return hipGetLastError();
}

#define NV_DISPATCH_NS(fn, ...)                    \
    switch (op.ns) {                               \
    case 1: return fn<1>(__VA_ARGS__);             \
    case 2: return fn<2>(__VA_ARGS__);             \
    case 3: return fn<3>(__VA_ARGS__);             \
    case 4: return fn<4>(__VA_ARGS__);             \
    case 5: return fn<5>(__VA_ARGS__);             \
    case 6: return fn<6>(__VA_ARGS__);             \
    case 7: return fn<7>(__VA_ARGS__);             \
    case 8: return fn<8>(__VA_ARGS__);             \
    case 9: return fn<9>(__VA_ARGS__);             \
    case 10: return fn<10>(__VA_ARGS__);           \
    default: return hipErrorInvalidValue;          \
    }

template <int NS>
static hipError_t diag_ns(const NvOpDev &op, double *dg, hipStream_t s) {
    hipLaunchKernelGGL((diag_precompute_kernel<NS>), dim3(hv_blocks(op)), dim3(kBlock), 0, s, op, dg);
    return hipGetLastError();
}

hipError_t launch_diag_precompute(const NvOpDev &op, double *dg, hipStream_t s) {
    NV_DISPATCH_NS(diag_ns, op, dg, s)
}

hipError_t launch_hv_full(const NvOpDev &op, int prec, const double *x, double *y, hipStream_t s) {
    NV_DISPATCH_NS(hv_full_ns, op, prec, x, y, s)
}

hipError_t launch_hv2_full(const NvOpDev &op, int prec, const double *x, const void *xin, void *Z, double *y,
                           hipStream_t s) {
    NV_DISPATCH_NS(hv2_ns, op, prec, xin, Z, nullptr, x, y, s)
}

// ---- fast-block H*v launchers ----
static size_t hv_fb_lds(const NvOpDev &op, int prec) {
    return fb_lds_layout(op, prec == 0 ? 8 : 4, nullptr, nullptr, nullptr);
}

bool hv_fb_use(const NvOpDev &op, int prec) {
    if (!op.tshape || !op.fb.on || (prec != 0 && prec != 1)) return false;
    const int64_t el = (int64_t)nv_lo_stride(op.ns) * (prec == 0 ? 8 : 4);
    return (int64_t)op.D * el >= op.fb_min_bytes && shaped_ok(op, el) && hv_fb_lds(op, prec) <= 64 * 1024 &&
           ((int64_t)op.D + 2 * (int64_t)op.fb.pad) * el < ((int64_t)1 << 31);
}

int hv_fb_parts(const NvOpDev &op) { return (op.fb.nblk + 3) / 4; }

template <int NS>
static hipError_t fb_iter_ns(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    const size_t lds = hv_fb_lds(op, prec);
    const unsigned g = (unsigned)hv_fb_parts(op);
    if (prec == 0)
        hipLaunchKernelGGL((hv_fb_iter_kernel<NS, F64I, double>), dim3(g), dim3(kBlock), lds, s, op, st);
    else
        hipLaunchKernelGGL((hv_fb_iter_kernel<NS, float, float>), dim3(g), dim3(kBlock), lds, s, op, st);
    return hipGetLastError();
}

template <int NS>
static hipError_t fb_full_ns(const NvOpDev &op, int prec, const void *xin, const double *x, double *y,
                             hipStream_t s) {
    const size_t lds = hv_fb_lds(op, prec);
    const unsigned g = (unsigned)hv_fb_parts(op);
    if (prec == 0)
        hipLaunchKernelGGL((hv_fb_full_kernel<NS, F64I, double>), dim3(g), dim3(kBlock), lds, s, op, xin, x, y);
    else
        hipLaunchKernelGGL((hv_fb_full_kernel<NS, float, float>), dim3(g), dim3(kBlock), lds, s, op, xin, x, y);
    return hipGetLastError();
}

hipError_t launch_hv_fb_full(const NvOpDev &op, int prec, const double *x, const void *xin, double *y,
                             hipStream_t s) {
    NV_DISPATCH_NS(fb_full_ns, op, prec, xin, x, y, s)
}

// The iteration H*v runs in the fast-block form when the operator has its
// tables and the interleaved copy of q_i exists (FP32: LzState::vlo; FP64:
// LzState::v64i).
static bool fb_iter(const NvOpDev &op, int prec, const LzState &st) {
    return hv_fb_use(op, prec) && (prec == 1 ? st.vlo != nullptr : st.v64i != nullptr);
}

// The iteration H*v runs two-pass when the caller gave the channel scratch
// (LzState::z2) and, in FP64, the interleaved copy of q_i.
static bool hv2_iter(const NvOpDev &op, int prec, const LzState &st) {
    return st.z2 && hv2_use(op, prec) && (prec == 1 || st.v64i);
}

static hipError_t hv_iter_any(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    if (fb_iter(op, prec, st)) NV_DISPATCH_NS(fb_iter_ns, op, prec, st, s)
    if (hv2_iter(op, prec, st)) {
        const void *xin = prec == 1 ? st.vlo : (const void *)st.v64i;
        NV_DISPATCH_NS(hv2_ns, op, prec, xin, st.z2, &st, nullptr, nullptr, s)
    }
    NV_DISPATCH_NS(hv_iter_ns, op, prec, st, s)
}

int hv_iter_parts(const NvOpDev &op, int prec, const LzState &st) {
    if (fb_iter(op, prec, st)) return hv_fb_parts(op);
    return hv2_iter(op, prec, st) ? hv2_parts(op) : hv_blocks(op);
}

static hipError_t step_any(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    NV_DISPATCH_NS(step_ns, op, prec, st, s)
}

hipError_t launch_hv_iter_only(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    return hv_iter_any(op, prec, st, s);
}

hipError_t launch_hv_iter(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    hipError_t e = hv_iter_any(op, prec, st, s);
    if (e != hipSuccess) return e;
    return launch_finish_sum(st.partials, hv_iter_parts(op, prec, st), st.alpha, st.iter, 0, s);
}

hipError_t launch_lanczos_step_only(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    return step_any(op, prec, st, s);
}

hipError_t launch_lanczos_step(const NvOpDev &op, int prec, const LzState &st, hipStream_t s, int64_t *advance) {
    hipError_t e = step_any(op, prec, st, s);
    if (e != hipSuccess) return e;
    return launch_finish_sum(st.partials, hv_blocks(op), st.beta, st.iter, 1, s, advance);
}

template <int NS>
static hipError_t normalize_ns(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    const int g = grid_stride_blocks(op.D);
    const int64_t N = op.N, D = op.D;
    if (op.stream_nt) {
        if (prec == 0 || st.vlo == nullptr)
            hipLaunchKernelGGL((normalize_rows_kernel<NS, double, true>), dim3(g), dim3(kBlock), 0, s, N, D, st);
        else if (prec == 1)
            hipLaunchKernelGGL((normalize_rows_kernel<NS, float, true>), dim3(g), dim3(kBlock), 0, s, N, D, st);
        else
            hipLaunchKernelGGL((normalize_rows_kernel<NS, __hip_bfloat16, true>), dim3(g), dim3(kBlock), 0, s, N, D,
                               st);
    } else {
        if (prec == 0 || st.vlo == nullptr)
            hipLaunchKernelGGL((normalize_rows_kernel<NS, double, false>), dim3(g), dim3(kBlock), 0, s, N, D, st);
        else if (prec == 1)
            hipLaunchKernelGGL((normalize_rows_kernel<NS, float, false>), dim3(g), dim3(kBlock), 0, s, N, D, st);
        else
            hipLaunchKernelGGL((normalize_rows_kernel<NS, __hip_bfloat16, false>), dim3(g), dim3(kBlock), 0, s, N,
                               D, st);
    }
    return hipGetLastError();
}

hipError_t launch_normalize(const NvOpDev &op, int prec, const LzState &st, hipStream_t s) {
    NV_DISPATCH_NS(normalize_ns, op, prec, st, s)
}

// y[r * nsp + s] = (float)x[s * D + r] (the interleaved FP32 copy).
__global__ __launch_bounds__(kBlock) void cast_f32_kernel(const double *x, float *y, int64_t n, int64_t D, int nsp) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const int64_t s = i / D;
        y[(i - s * D) * nsp + s] = (float)x[i];
    }
}

// y[r * nsp + s] = x[s * D + r] (the interleaved FP64 copy).
__global__ __launch_bounds__(kBlock) void cast_f64i_kernel(const double *x, double *y, int64_t n, int64_t D, int nsp) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const int64_t s = i / D;
        y[(i - s * D) * nsp + s] = x[i];
    }
}

hipError_t launch_cast_f64i(const double *x, double *y, int64_t n, int64_t D, int ns, hipStream_t s) {
    hipLaunchKernelGGL(cast_f64i_kernel, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, x, y, n, D,
                       nv_lo_stride(ns));
    return hipGetLastError();
}

hipError_t launch_cast_f32(const double *x, float *y, int64_t n, int64_t D, int ns, hipStream_t s) {
    hipLaunchKernelGGL(cast_f32_kernel, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, x, y, n, D,
                       nv_lo_stride(ns));
    return hipGetLastError();
}

hipError_t launch_set_iter(int64_t *d_iter, int64_t value, int increment, hipStream_t s) {
    hipLaunchKernelGGL(set_iter_kernel, dim3(1), dim3(64), 0, s, d_iter, value, increment);
    return hipGetLastError();
}

hipError_t launch_fill_start(double *v, int64_t D, int ns, const double *vals_dev, hipStream_t s) {
    hipLaunchKernelGGL(fill_start_kernel, dim3(grid_stride_blocks(D * ns)), dim3(kBlock), 0, s, v, D, ns, vals_dev);
    return hipGetLastError();
}

hipError_t launch_norm_to(const double *v, int64_t n, double *out, double *partials, unsigned *counter,
                          hipStream_t s) {
    (void)counter;
    const int g = grid_stride_blocks(n);
    hipLaunchKernelGGL(norm_kernel, dim3(g), dim3(kBlock), 0, s, v, n, partials);
    return launch_finish_sum(partials, g, out, nullptr, 1, s);
}

// ------------------------------------------------------------------------
// Gaussian broadening of a stick spectrum (convolve.py:28-48):
// env[g] = sum_j y_j exp(-((grid_g - x_j)/sigma)^2 / 2), then env / max(env).
// One thread per grid point; the sticks are staged through LDS in tiles.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void convolve_kernel(const double *x, const double *y, int64_t n,
                                                          const double *grid, int64_t ng, double sigma, double *env,
                                                          double *part) {
    __shared__ double sx[kBlock], sy[kBlock];
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double gx = g < ng ? grid[g] : 0.0;
    const double inv = 1.0 / sigma;
    double acc = 0.0;
    for (int64_t j0 = 0; j0 < n; j0 += kBlock) {
        __syncthreads();
        if (j0 + threadIdx.x < n) {
            sx[threadIdx.x] = x[j0 + threadIdx.x];
            sy[threadIdx.x] = y[j0 + threadIdx.x];
        }
        __syncthreads();
        const int m = (int)(n - j0 < kBlock ? n - j0 : kBlock);
        for (int j = 0; j < m; ++j) {
            const double t = (gx - sx[j]) * inv;
            acc += sy[j] * exp(-0.5 * t * t);
        }
    }
    if (g < ng) env[g] = acc;
    // block maximum for the normalisation
    double mx = g < ng ? acc : -INFINITY;
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    __shared__ double red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m2 = red[0];
        for (int w = 1; w < kBlock / 64; ++w) m2 = fmax(m2, red[w]);
        part[blockIdx.x] = m2;
    }
}

__global__ __launch_bounds__(kBlock) void convolve_norm_kernel(double *env, int64_t ng, const double *part, int nb) {
    __shared__ double mx;
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int b = 1; b < nb; ++b) m = fmax(m, part[b]);
        mx = m;
    }
    __syncthreads();
    for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < ng; g += (int64_t)gridDim.x * kBlock)
        env[g] /= mx;
}

// Overlaps of up to kMaxProbes probe vectors with basis vectors j0..j0+nj-1
// (q_j in slot j % nslots of V): grid (chunks, nj), partial dots per chunk in
// part[(k * chunks + chunk) * kMaxProbes + p], summed on the host.
__global__ __launch_bounds__(kBlock) void probe_dots_kernel(ProbeSet ps, const double *__restrict__ V, int64_t nslots,
                                                            int64_t N, int64_t j0, int64_t chunk,
                                                            double *__restrict__ part) {
    const int64_t k = blockIdx.y;
    const double *q = V + ((j0 + k) % nslots) * N;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < N ? lo + chunk : N;
    double acc[kMaxProbes];
#pragma unroll
    for (int p = 0; p < kMaxProbes; ++p) acc[p] = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const double v = q[i];
#pragma unroll
        for (int p = 0; p < kMaxProbes; ++p)
            if (p < ps.np) acc[p] += v * ps.p[p][i];
    }
#pragma unroll
    for (int p = 0; p < kMaxProbes; ++p) {
        const double t = block_sum(acc[p]);
        if (threadIdx.x == 0) part[(k * gridDim.x + blockIdx.x) * kMaxProbes + p] = t;
    }
}

int probe_chunks(int64_t N) { return (int)std::min<int64_t>(64, (N + 65535) / 65536); }

hipError_t launch_probe_dots(const ProbeSet &ps, const double *V, int64_t nslots, int64_t N, int64_t j0, int64_t nj,
                             double *part, hipStream_t s) {
    const int nc = probe_chunks(N);
    const int64_t chunk = (N + nc - 1) / nc;
    hipLaunchKernelGGL(probe_dots_kernel, dim3(nc, (unsigned)nj), dim3(kBlock), 0, s, ps, V, nslots, N, j0, chunk,
                       part);
    return hipGetLastError();
}

hipError_t launch_convolve(const double *x, const double *y, int64_t n, const double *grid, int64_t ng,
                           double sigma, double *env, double *part, hipStream_t s) {
    const int nb = (int)((ng + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(convolve_kernel, dim3(nb), dim3(kBlock), 0, s, x, y, n, grid, ng, sigma, env, part);
    hipLaunchKernelGGL(convolve_norm_kernel, dim3(nb), dim3(kBlock), 0, s, env, ng, part, nb);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// Tridiagonal eigensolve (restates GIVENS, Source/yarkony.f90:91-543, for a
// tridiagonal input: the Householder stage and the back transformation are
// identities there).  The matrix arrives divided by its Frobenius norm.
// ------------------------------------------------------------------------
#define NV_EPS 1.1102230246251565e-16          /* 2^-53, determine_epsilon */
#define NV_HUGE 1.7976931348623157e308
#define NV_DEL1 (NV_EPS / 100.0)
#define NV_DELTA (NV_EPS * NV_EPS * 100.0)
#define NV_DELBIG (NV_HUGE * NV_DELTA / 1000.0)
#define NV_THETA1 (1000.0 / NV_HUGE)
#define NV_TOLER (NV_EPS * 100.0)

// Number of eigenvalues below x: negative pivots of the LDL^T of T - x,
// with a pivot below THETA1 in magnitude treated as in Source/yarkony.f90:360-389.
__device__ int64_t sturm_below(const double *d, const double *e2, int64_t n, double x) {
    int64_t nonneg = 0;
    int64_t j = 0;
    double f = d[0] - x;
    while (true) {
        if (fabs(f) < NV_THETA1) {
            j += 2;
            nonneg += 1;
            if (j >= n) break;
            f = d[j] - x;
            continue;
        }
        if (f >= 0.0) nonneg += 1;
        j += 1;
        if (j >= n) break;
        f = d[j] - x - e2[j - 1] / f;
    }
    return n - nonneg;
}

// One wave per root: multisection.  Each round the 64 lanes take Sturm counts
// at 64 interior points of [lo, hi] and the wave keeps the sub-interval where
// the count reaches i + 1 (count(lo) < i + 1 <= count(hi) throughout), so a
// round shrinks the interval 65-fold where bisection halves it.  When the
// points no longer fall strictly inside (a few ulps left) it bisects, and it
// stops at adjacent doubles; the root is hi, as in the bisection of GIVENS
// (Source/yarkony.f90:345-400).  T (d, e2) is staged in LDS when it fits
// (lds = 1): every Sturm step reads it.
constexpr int64_t kTriLdsMax = 4000;  // 2 * 4000 doubles, under 64 KB

__global__ __launch_bounds__(64) void tri_bisect_kernel(int64_t n, const double *dg, const double *e2g,
                                                         double *root, int lds) {
    extern __shared__ double tri_lds[];
    const double *d = dg, *e2 = e2g;
    if (lds) {
        for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
            tri_lds[k] = dg[k];
            tri_lds[n + k] = e2g[k];
        }
        __syncthreads();
        d = tri_lds;
        e2 = tri_lds + n;
    }
    const int64_t i = blockIdx.x;
    const int lane = threadIdx.x;
    double lo = -1.0, hi = 1.0;
    for (int guard = 0; guard < 4096; ++guard) {
        const double x = lo + (hi - lo) * ((double)(lane + 1) / 65.0);
        const bool inside = x > lo && x < hi;
        const bool ge = inside ? sturm_below(d, e2, n, x) >= i + 1 : x >= hi;
        const unsigned long long ball = __ballot(ge);
        const double xprev = __shfl_up(x, 1, 64);
        const int k = ball ? __ffsll(ball) - 1 : 64;  // first point at or above the root
        const double new_hi = k < 64 ? __shfl(x, k, 64) : hi;
        const double new_lo = k == 0 ? lo : __shfl(k < 64 ? xprev : x, k < 64 ? k : 63, 64);
        if ((new_lo > lo && new_lo < new_hi) || new_hi < hi) {
            lo = new_lo > lo ? new_lo : lo;
            hi = new_hi;
            continue;
        }
        const double mid = (lo + hi) * 0.5;
        if (mid == lo || mid == hi) break;
        if (sturm_below(d, e2, n, mid) >= i + 1) hi = mid;
        else lo = mid;
    }
    if (lane == 0) root[i] = hi;
}

// Back substitution through the pivoted factors of one root (Source/yarkony.f90:790-780).
// Returns the last pivot used (the value the forward sweep of the reference
// carries in TEMP).
__device__ double tri_backsolve(int64_t n, int64_t i, const double *U2, const double *U3, const double *U4,
                                double *v) {
    double t = 1.0;
    for (int64_t l = n - 1; l >= 0; --l) {
        while (true) {
            const double v1 = l + 1 < n ? v[(l + 1) * n + i] : 0.0;
            const double v2 = l + 2 < n ? v[(l + 2) * n + i] : 0.0;
            const double e = v[l * n + i] - v1 * U3[l * n + i] - v2 * U4[l * n + i];
            // overflow guard of the reference; a non-finite value cannot be
            // rescaled and ends the sweep instead of looping
            if (fabs(e) > NV_DELBIG && isfinite(e)) {
                for (int64_t k = 0; k < n; ++k) v[k * n + i] /= NV_DELBIG;
                continue;
            }
            t = U2[l * n + i];
            if (fabs(t) < NV_DELTA) t = NV_DELTA;
            v[l * n + i] = e / t;
            break;
        }
    }
    return t;
}

// Inverse iteration for root i: factor T - root (partial pivoting), start
// from ones (or the reference's multiplicative generator inside a cluster),
// back solve, forward sweep, back solve.  Vectors are stored component-major:
// vec[j*n + i].
__global__ void tri_inverse_kernel(int64_t n, const double *d, const double *e, const double *root,
                                   const int32_t *ia, const int64_t *rank, double *lu, double *vec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t nn = n * n;
    double *U2 = lu, *U3 = lu + nn, *U4 = lu + 2 * nn, *U5 = lu + 3 * nn;
    const double ar = root[i];
    double el1 = d[0] - ar;
    double el2 = n > 1 ? e[0] : 0.0;
    for (int64_t j = 0; j + 1 < n; ++j) {
        const double bj = e[j];
        const double bj1 = j + 2 < n ? e[j + 1] : 0.0;
        double t;
        if (fabs(el1) <= fabs(bj)) {
            U2[j * n + i] = bj;
            U3[j * n + i] = d[j + 1] - ar;
            U4[j * n + i] = bj1;
            t = 1.0;
            if (fabs(bj) > NV_THETA1) t = el1 / bj;
            el1 = el2 - t * (d[j + 1] - ar);
            el2 = -t * bj1;
        } else {
            U2[j * n + i] = el1;
            U3[j * n + i] = el2;
            U4[j * n + i] = 0.0;
            t = bj / el1;
            el1 = d[j + 1] - ar - t * el2;
            el2 = bj1;
        }
        U5[j * n + i] = t;
    }
    U2[(n - 1) * n + i] = el1;
    U3[(n - 1) * n + i] = 0.0;
    U4[(n - 1) * n + i] = 0.0;
    if (n >= 2) U4[(n - 2) * n + i] = 0.0;
    // start vector
    if (ia[i] == 0) {
        for (int64_t j = 0; j < n; ++j) vec[j * n + i] = 1.0;
    } else {
        // state of x <- 4099 x mod 2^23 after n*rank[i] draws from 2^23-3
        const uint64_t m = 1ull << 23;
        uint64_t a = 4099, x = m - 3;
        uint64_t k = (uint64_t)n * (uint64_t)rank[i];
        uint64_t p = a % m, acc = 1;
        while (k) {
            if (k & 1) acc = (acc * p) % m;
            p = (p * p) % m;
            k >>= 1;
        }
        x = (x * acc) % m;
        for (int64_t j = 0; j < n; ++j) {
            x = (4099ull * x) % m;
            vec[j * n + i] = (double)x / (double)(m / 2) - 1.0;
        }
    }
    const double t1 = tri_backsolve(n, i, U2, U3, U4, vec);
    // forward sweep (Source/yarkony.f90:890-830); a swapped row uses the last
    // pivot of the back substitution as multiplier, as the reference does.
    double el = vec[0 * n + i];
    for (int64_t j = 0; j + 1 < n; ++j) {
        const double nxt = vec[(j + 1) * n + i];
        if (U2[j * n + i] == e[j]) {
            vec[j * n + i] = nxt;
            el = el - nxt * t1;
        } else {
            vec[j * n + i] = el;
            el = nxt - el * U5[j * n + i];
        }
    }
    vec[(n - 1) * n + i] = el;
    tri_backsolve(n, i, U2, U3, U4, vec);
}

// Normalisation (Source/yarkony.f90:900-920) by one block: max |v|, sum of
// (v/max)^2, scale, flush components below eps/100 to zero.
__device__ void tri_normalize_block(int64_t n, int64_t i, double *vec) {
    __shared__ double red[16];
    double mx = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) mx = fmax(fabs(vec[j * n + i]), mx);
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mx = fmax(mx, red[w]);
    double s = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
        const double q = vec[j * n + i] / mx;
        s += q * q;
    }
    s = block_sum(s);
    const double t = 1.0 / (sqrt(s) * mx);
    for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
        double v = vec[j * n + i] * t;
        if (fabs(v) < NV_DEL1) v = 0.0;
        vec[j * n + i] = v;
    }
    __syncthreads();
}

// Gram-Schmidt inside clusters of near-equal roots (Source/yarkony.f90:800-870)
// and normalisation.  One block per root that starts a cluster (or stands
// alone); the block walks the cluster's members in order.
__global__ __launch_bounds__(256) void tri_finish_kernel(int64_t n, const int32_t *ia, double *vec) {
    const int64_t i = blockIdx.x;
    if (i >= n || ia[i] != 0) return;
    tri_normalize_block(n, i, vec);
    for (int64_t m = i + 1; m < n && ia[m] > 0; ++m) {
        for (int64_t j1 = 1; j1 <= ia[m]; ++j1) {
            const int64_t k = m - j1;
            double t = 0.0;
            for (int64_t j = threadIdx.x; j < n; j += blockDim.x) t += vec[j * n + m] * vec[j * n + k];
            t = block_sum(t);
            for (int64_t j = threadIdx.x; j < n; j += blockDim.x) vec[j * n + m] -= t * vec[j * n + k];
            __syncthreads();
        }
        tri_normalize_block(n, m, vec);
    }
}

__global__ void tri_cluster_kernel(int64_t n, const double *root, int32_t *ia, int64_t *rank) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int32_t c = 0;
    int64_t r = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (i > 0 && fabs(root[i - 1] - root[i]) < NV_TOLER) c += 1;
        else c = 0;
        ia[i] = c;
        rank[i] = r;
        if (c != 0) r += 1;
    }
}

hipError_t launch_tridiag(int64_t n, const TriWork &w, double anorm, hipStream_t s) {
    (void)anorm;
    const int tb = 64;
    const int g = (int)((n + tb - 1) / tb);
    const int lds = n <= kTriLdsMax;
    hipLaunchKernelGGL(tri_bisect_kernel, dim3((unsigned)n), dim3(64), lds ? (size_t)2 * n * sizeof(double) : 0, s,
                       n, w.d, w.e2, w.root, lds);
    hipLaunchKernelGGL(tri_cluster_kernel, dim3(1), dim3(64), 0, s, n, w.root, w.ia, w.rank);
    hipLaunchKernelGGL(tri_inverse_kernel, dim3(g), dim3(tb), 0, s, n, w.d, w.e, w.root, w.ia, w.rank, w.lu, w.vec);
    hipLaunchKernelGGL(tri_finish_kernel, dim3((unsigned)n), dim3(256), 0, s, n, w.ia, w.vec);
    return hipGetLastError();
}


// ------------------------------------------------------------------------
// Batched Lanczos: many operators of uneven size advance one iteration per
// launch sequence.  The rows of every problem are cut into tiles of
// blockDim.x rows; one workgroup per tile, so one launch covers the whole
// batch and the small problems cost no launches of their own.  Each tile
// reads its problem's operator tables and vector offsets.
// ------------------------------------------------------------------------
template <int NS, typename TIn, typename TAcc>
__device__ __forceinline__ void batch_hv_row(const NvOpDev &op, const TIn *x, const double *q, TAcc *y, int r,
                                             const uint8_t *dig, const TAcc *ftab, int tid, double &part) {
    TAcc acc[NS];
    offdiag_accumulate<NS, TIn, TAcc>(op, x, r, dig, ftab, tid, acc);
    double v[NS], dv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r];
    diag_apply<NS>(op, r, dig, tid, v, dv);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        y[(size_t)s * op.D + r] = acc[s];
        part += v[s] * (dv[s] + (double)acc[s]);
    }
}

template <typename TIn>
__global__ __launch_bounds__(kBlock) void batch_normalize_kernel(BatchDev b) {
    const BatchTile tl = b.tiles[blockIdx.x];
    const NvOpDev &opg = b.ops[tl.prob];
    const int D = opg.D, ns = opg.ns;
    const int r = tl.row0 + threadIdx.x;
    if (r >= D) return;
    const int64_t it = *b.iter;
    double *q = b.V + slot_of(it, 3) * b.Ntot + b.voff[tl.prob];
    const double sc = 1.0 / b.beta[(int64_t)tl.prob * b.ld + it - 1];
    TIn *lo = b.vlo ? reinterpret_cast<TIn *>(b.vlo) + b.loff[tl.prob] : nullptr;
    const int nsp = nv_lo_stride(ns);
    for (int s = 0; s < ns; ++s) {
        const double v = q[(size_t)s * D + r] * sc;
        q[(size_t)s * D + r] = v;
        if (lo) lo[(size_t)r * nsp + s] = from_d<TIn>(v);
    }
}

template <int NS, typename TIn, typename TAcc>
__global__ __launch_bounds__(kBlock) void batch_hv_kernel(BatchDev b, int t0) {
    const int t = t0 + blockIdx.x;
    const BatchTile tl = b.tiles[t];
    const NvOpDev op = b.ops[tl.prob];
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    const int tid = threadIdx.x;
    const int r = tl.row0 + tid;
    const bool live = r < op.D;
    const int64_t it = *b.iter;
    const double *q = b.V + slot_of(it, 3) * b.Ntot + b.voff[tl.prob];
    const TIn *x = b.vlo ? reinterpret_cast<const TIn *>(b.vlo) + b.loff[tl.prob] : reinterpret_cast<const TIn *>(q);
    TAcc *y = reinterpret_cast<TAcc *>(b.y) + b.voff[tl.prob];
    if (live) digits_energy(op, r, dig, tid);
    __syncthreads();
    double part = 0.0;
    if (live) batch_hv_row<NS, TIn, TAcc>(op, x, q, y, r, dig, ftab, tid, part);
    const double bs = block_sum(part);
    if (tid == 0) b.partials[t] = bs;
}

// batch_hv_kernel with two rows per thread: a workgroup of bs/2 threads
// covers the tile's bs rows as two halves.
template <int NS, typename TIn, typename TAcc>
__global__ __launch_bounds__(kBlock) void batch_hv2_kernel(BatchDev b, int t0) {
    const int t = t0 + blockIdx.x;
    const BatchTile tl = b.tiles[t];
    const NvOpDev op = b.ops[tl.prob];
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    uint8_t *dig1 = dig + op.M * kBlock;
    const int tid = threadIdx.x;
    const int half = b.bs / 2;
    const int r0 = tl.row0 + tid, r1 = r0 + half;
    const bool l0 = r0 < op.D, l1 = r1 < op.D;
    const int64_t it = *b.iter;
    const double *q = b.V + slot_of(it, 3) * b.Ntot + b.voff[tl.prob];
    const TIn *x = b.vlo ? reinterpret_cast<const TIn *>(b.vlo) + b.loff[tl.prob] : reinterpret_cast<const TIn *>(q);
    TAcc *y = reinterpret_cast<TAcc *>(b.y) + b.voff[tl.prob];
    if (l0) digits_energy(op, r0, dig, tid);
    else for (int k = 0; k < op.M; ++k) dig[k * kBlock + tid] = 0;
    if (l1) digits_energy(op, r1, dig1, tid);
    else for (int k = 0; k < op.M; ++k) dig1[k * kBlock + tid] = 0;
    __syncthreads();
    double part = 0.0;
    if (l0) {
        TAcc acc0[NS], acc1[NS];
        offdiag_accumulate2<NS, TIn, TAcc>(op, x, r0, r1, l0, l1, dig, dig1, ftab, tid, acc0, acc1);
        double v[NS], dv[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r0];
        diag_apply<NS>(op, r0, dig, tid, v, dv);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            y[(size_t)s * op.D + r0] = acc0[s];
            part += v[s] * (dv[s] + (double)acc0[s]);
        }
        if (l1) {
#pragma unroll
            for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r1];
            diag_apply<NS>(op, r1, dig1, tid, v, dv);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                y[(size_t)s * op.D + r1] = acc1[s];
                part += v[s] * (dv[s] + (double)acc1[s]);
            }
        }
    }
    const double bs = block_sum(part);
    if (tid == 0) b.partials[t] = bs;
}

// batch_hv2_kernel / batch_hv_kernel over the shaped terms: R rows per
// thread (row k of a thread at row0 + k * blockDim.x + tid), for batches
// whose every operator has shaped terms.
template <int NS, typename TIn, typename TAcc, int R>
__global__ __launch_bounds__(kBlock) void batch_shaped_kernel(BatchDev b, int t0) {
    const int t = t0 + blockIdx.x;
    const BatchTile tl = b.tiles[t];
    const NvOpDev op = b.ops[tl.prob];
    uint8_t *dig;
    const TAcc *ftab = stage_tables<TAcc>(op, dig);
    const int tid = threadIdx.x;
    const int64_t it = *b.iter;
    const double *q = b.V + slot_of(it, 3) * b.Ntot + b.voff[tl.prob];
    const TIn *x = b.vlo ? reinterpret_cast<const TIn *>(b.vlo) + b.loff[tl.prob] : reinterpret_cast<const TIn *>(q);
    TAcc *y = reinterpret_cast<TAcc *>(b.y) + b.voff[tl.prob];
    int r[R];
    bool live[R];
    unsigned vo[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        r[k] = tl.row0 + k * (int)blockDim.x + tid;
        live[k] = r[k] < op.D;
        vo[k] = (unsigned)r[k] * XRes<NS, TIn>::kEl;
        uint8_t *dk = dig + k * op.M * kBlock;
        if (live[k]) digits_energy(op, r[k], dk, tid);
        else for (int m = 0; m < op.M; ++m) dk[m * kBlock + tid] = 0;
    }
    __syncthreads();
    const XRes<NS, TIn> xs(x, op.D);
    TAcc acc[R][NS];
    offdiag_shaped<NS, TIn, TAcc, R, false>(op, xs, vo, dig, ftab, tid, acc);
    double part = 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (!live[k]) continue;
        double v[NS], dv[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r[k]];
        diag_apply<NS>(op, r[k], dig + k * op.M * kBlock, tid, v, dv);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            y[(size_t)s * op.D + r[k]] = acc[k][s];
            part += v[s] * (dv[s] + (double)acc[k][s]);
        }
    }
    const double bs = block_sum(part);
    if (tid == 0) b.partials[t] = bs;
}

template <int NS, typename TAcc>
__device__ __forceinline__ void batch_step_row(const NvOpDev &op, const double *q, const double *qm, double *res,
                                               const TAcc *y, double a, double bb, int64_t it, int r,
                                               const uint8_t *dig, int tid, double &part) {
    double v[NS], dv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) v[s] = q[(size_t)s * op.D + r];
    diag_apply<NS>(op, r, dig, tid, v, dv);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const size_t o = (size_t)s * op.D + r;
        double w = dv[s] + (double)y[o];
        w = w - a * v[s];
        if (it > 1) w = w - bb * qm[o];
        res[o] = w;
        part += w * w;
    }
}

template <int NS, typename TAcc>
__global__ __launch_bounds__(kBlock) void batch_step_kernel(BatchDev b, int t0) {
    extern __shared__ __align__(16) unsigned char dig[];
    const int t = t0 + blockIdx.x;
    const BatchTile tl = b.tiles[t];
    const NvOpDev op = b.ops[tl.prob];
    const int tid = threadIdx.x;
    const int r = tl.row0 + tid;
    const int64_t it = *b.iter;
    const int64_t off = b.voff[tl.prob];
    const double *q = b.V + slot_of(it, 3) * b.Ntot + off;
    const double *qm = b.V + slot_of(it - 1, 3) * b.Ntot + off;
    double *res = b.V + slot_of(it + 1, 3) * b.Ntot + off;
    const TAcc *y = reinterpret_cast<const TAcc *>(b.y) + off;
    const double a = b.alpha[(int64_t)tl.prob * b.ld + it];
    const double bb = it > 1 ? b.beta[(int64_t)tl.prob * b.ld + it - 1] : 0.0;
    double part = 0.0;
    if (r < op.D) {
        if (op.offstate) digits_energy(op, r, dig, tid);
        batch_step_row<NS, TAcc>(op, q, qm, res, y, a, bb, it, r, dig, tid, part);
    }
    const double bs = block_sum(part);
    if (tid == 0) b.partials[t] = bs;
}

// dst[p][it] = (sqrt of) the sum of problem p's tile partials, in tile order.
__global__ __launch_bounds__(256) void batch_finish_kernel(BatchDev b, double *dst, int take_sqrt) {
    const int p = blockIdx.x;
    const int t0 = b.tile0[2 * p], t1 = b.tile0[2 * p + 1];
    double s = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) s += b.partials[t];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        const int64_t it = *b.iter;
        dst[(int64_t)p * b.ld + it] = take_sqrt ? sqrt(s) : s;
    }
}

size_t batch_hv_lds(const NvOpDev &op, int prec) { return hv_lds_bytes(op, prec); }

hipError_t launch_batch_normalize(const BatchDev &b, int prec, hipStream_t s) {
    if (prec == 0 || b.vlo == nullptr)
        hipLaunchKernelGGL((batch_normalize_kernel<double>), dim3(b.T), dim3(b.bs), 0, s, b);
    else
        hipLaunchKernelGGL((batch_normalize_kernel<float>), dim3(b.T), dim3(b.bs), 0, s, b);
    return hipGetLastError();
}

// The tuned form of the batched H*v: shaped or step-table kernels, and in FP32
// one or two rows per thread (two rows need workgroups of at least one wave).
template <int NS>
static void batch_hv_ns(const BatchDev &b, int prec, size_t lds, int t0, int nt, hipStream_t s) {
    const bool two = prec != 0 && b.rows == 2 && b.bs >= 128;
    if constexpr (NS <= kShapedMaxStates) {
        if (b.shaped[prec == 0 ? 0 : 1] && !b.steps) {
            if (prec == 0)
                hipLaunchKernelGGL((batch_shaped_kernel<NS, double, double, 1>), dim3(nt), dim3(b.bs), lds, s, b, t0);
            else if (b.rows == 4 && b.bs >= 256)
                hipLaunchKernelGGL((batch_shaped_kernel<NS, float, float, 4>), dim3(nt), dim3(b.bs / 4), lds, s, b,
                                   t0);
            else if (two)
                hipLaunchKernelGGL((batch_shaped_kernel<NS, float, float, 2>), dim3(nt), dim3(b.bs / 2), lds, s, b,
                                   t0);
            else
                hipLaunchKernelGGL((batch_shaped_kernel<NS, float, float, 1>), dim3(nt), dim3(b.bs), lds, s, b, t0);
            return;
        }
    }
    if (prec == 0)
        hipLaunchKernelGGL((batch_hv_kernel<NS, double, double>), dim3(nt), dim3(b.bs), lds, s, b, t0);
    else if (two)
        hipLaunchKernelGGL((batch_hv2_kernel<NS, float, float>), dim3(nt), dim3(b.bs / 2), lds, s, b, t0);
    else
        hipLaunchKernelGGL((batch_hv_kernel<NS, float, float>), dim3(nt), dim3(b.bs), lds, s, b, t0);
}

template <int NS>
static void batch_step_ns(const BatchDev &b, int prec, size_t lds, int t0, int nt, hipStream_t s) {
    if (prec == 0)
        hipLaunchKernelGGL((batch_step_kernel<NS, double>), dim3(nt), dim3(b.bs), lds, s, b, t0);
    else
        hipLaunchKernelGGL((batch_step_kernel<NS, float>), dim3(nt), dim3(b.bs), lds, s, b, t0);
}

#define NV_BATCH_DISPATCH(fn, ns, ...)                                                                            \
    switch (ns) {                                                                                                 \
    case 1: fn<1>(__VA_ARGS__); break;                                                                            \
    case 2: fn<2>(__VA_ARGS__); break;                                                                            \
    case 3: fn<3>(__VA_ARGS__); break;                                                                            \
    case 4: fn<4>(__VA_ARGS__); break;                                                                            \
    case 5: fn<5>(__VA_ARGS__); break;                                                                            \
    case 6: fn<6>(__VA_ARGS__); break;                                                                            \
    case 7: fn<7>(__VA_ARGS__); break;                                                                            \
    case 8: fn<8>(__VA_ARGS__); break;                                                                            \
    case 9: fn<9>(__VA_ARGS__); break;                                                                            \
    case 10: fn<10>(__VA_ARGS__); break;                                                                          \
    default: return hipErrorInvalidValue;                                                                         \
    }

hipError_t launch_batch_hv(const BatchDev &b, const BatchGroup *groups, int ngroups, int prec, size_t lds,
                           hipStream_t s) {
    for (int g = 0; g < ngroups; ++g) {
        NV_BATCH_DISPATCH(batch_hv_ns, groups[g].ns, b, prec, lds, groups[g].t0, groups[g].nt, s)
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_finish_kernel, dim3(b.P), dim3(256), 0, s, b, b.alpha, 0);
    return hipGetLastError();
}

hipError_t launch_batch_step(const BatchDev &b, const BatchGroup *groups, int ngroups, int prec, size_t lds,
                             hipStream_t s) {
    for (int g = 0; g < ngroups; ++g) {
        NV_BATCH_DISPATCH(batch_step_ns, groups[g].ns, b, prec, lds, groups[g].t0, groups[g].nt, s)
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_finish_kernel, dim3(b.P), dim3(256), 0, s, b, b.beta, 1);
    return hipGetLastError();
}

}  // namespace nv
