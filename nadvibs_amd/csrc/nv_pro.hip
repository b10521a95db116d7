// Partial re-orthogonalisation on the device (Simon's omega recurrence,
// Source/top.f90:844-1005).  The scalar recurrence runs in one workgroup,
// the range detection in one thread; the dots against the Krylov basis and
// the classical Gram-Schmidt update are bandwidth-bound passes over the
// HBM-resident basis.  Nothing here synchronises with the host: whether a
// re-orthogonalisation happens is decided and consumed on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "nv_pro.hpp"

namespace nv {

#define PRO_EPS 1.1102230246251565e-16  /* 2^-53 (determine_epsilon, Source/secondary.f90:260-271) */

__device__ __forceinline__ double lz_beta(const LzState &st, int64_t i) { return st.beta[i]; }

__global__ void pro_init_kernel(ProDev p) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double *o1 = p.omega, *o2 = p.omega + p.ostride;
    // Source/top.f90:590-595
    o1[0] = 0.0;
    o1[1] = 1.0;
    o2[0] = 0.0;
    o2[1] = p.noise * noise_gauss(p.seed, 0, 0.6) * p.epsn;
    if (p.niter >= 2) o2[2] = 1.0;
    p.meta[0] = 1;  // firststep
    p.meta[1] = 0;  // northog
    p.meta[2] = 1;  // draws used
    p.meta[3] = 0;
    p.meta[4] = 0;
    p.meta[5] = 0;
    *p.njl = 0;
}

// omega recurrence (Source/top.f90:862-874) and the row swap (:875-879).
__global__ __launch_bounds__(kBlock) void pro_recur_kernel(ProDev p, LzState st) {
    const int64_t it = *st.iter;
    double *o1 = p.omega, *o2 = p.omega + p.ostride;
    const uint64_t base = (uint64_t)p.meta[2];
    const double bi = st.beta[it];
    if (threadIdx.x == 0) {
        if (it < p.niter) o2[it + 1] = 1.0;
        o1[it] = p.noise * noise_gauss(p.seed, base, 0.6) * p.epsn * st.beta[1] / bi;
    }
    for (int64_t k = 1 + threadIdx.x; k <= it - 1; k += blockDim.x) {
        o1[k] = (1 / bi) * (st.beta[k] * o2[k + 1] + (st.alpha[k] - st.alpha[it]) * o2[k] +
                            st.beta[k - 1] * o2[k - 1] - st.beta[it - 1] * o1[k]) +
                PRO_EPS * (st.beta[k] + bi) * (p.noise * noise_gauss(p.seed, base + (uint64_t)k, 0.3));
    }
    __syncthreads();
    for (int64_t i = 1 + threadIdx.x; i <= it; i += blockDim.x) {
        const double t = o2[i];
        o2[i] = o1[i];
        o1[i] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) p.meta[2] = (int64_t)base + it;
}

// Dots of a and b (either may be null) with the Krylov vectors q_j listed in
// jlist (or j = c+1 when jlist is null), c < count.  Each wave writes one
// partial per dot; pro_reduce_kernel sums them in a fixed order.
__global__ __launch_bounds__(kBlock) void pro_multidot_kernel(int64_t N, LzState st, const int64_t *jlist,
                                                              const int32_t *count_dev, int32_t count_host,
                                                              const double *a_fixed, int a_cur, int b_res,
                                                              double *part, int32_t nchunks) {
    const int64_t it = *st.iter;
    const int32_t count = count_dev ? *count_dev : count_host;
    const double *a = a_fixed ? a_fixed : (a_cur ? st.V + (it % st.nslots) * N : nullptr);
    const double *b = b_res ? st.V + ((it + 1) % st.nslots) * N : nullptr;
    const int64_t chunk = (N + nchunks - 1) / nchunks;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < N ? lo + chunk : N;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t g = (int64_t)blockIdx.x * (kBlock / 64) + wid;
    const int64_t nw = (int64_t)nchunks * (kBlock / 64);
    for (int32_t c = 0; c < count; ++c) {
        const int64_t j = jlist ? jlist[c] : (int64_t)c + 1;
        const double *q = st.V + (j % st.nslots) * N;
        double sa = 0.0, sb = 0.0;
        for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) {
            const double qi = q[i];
            if (a) sa += a[i] * qi;
            if (b) sb += b[i] * qi;
        }
        for (int off = 32; off > 0; off >>= 1) {
            sa += __shfl_xor(sa, off, 64);
            sb += __shfl_xor(sb, off, 64);
        }
        if (lane == 0) {
            part[((int64_t)c * nw + g) * 2 + 0] = sa;
            part[((int64_t)c * nw + g) * 2 + 1] = sb;
        }
    }
}

__global__ void pro_reduce_kernel(const int32_t *count_dev, int32_t count_host, const double *part, int64_t nw,
                                  double *out_a, double *out_b) {
    const int32_t count = count_dev ? *count_dev : count_host;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x) {
        double sa = 0.0, sb = 0.0;
        for (int64_t g = 0; g < nw; ++g) {
            sa += part[(c * nw + g) * 2 + 0];
            sb += part[(c * nw + g) * 2 + 1];
        }
        out_a[c] = sa;
        out_b[c] = sb;
    }
}

// Exact overlap check every chkorthog iterations (Source/top.f90:876-893):
// raise omega where the measured overlap is larger.
__global__ void pro_check_kernel(ProDev p, LzState st) {
    const int64_t it = *st.iter;
    double *o1 = p.omega, *o2 = p.omega + p.ostride;
    const double bi = st.beta[it];
    for (int64_t i = 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= it - 1;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double c1 = p.dots_a[i - 1];
        const double c2 = p.dots_b[i - 1] / bi;
        if (fabs(o1[i]) < fabs(c1)) o1[i] = c1;
        if (fabs(o2[i]) < fabs(c2)) o2[i] = c2;
    }
}

// Exact monitoring (reorthog > 1, Source/top.f90:854-861): omega(2,i) = q3 . q_i.
__global__ void pro_exact_kernel(ProDev p, LzState st) {
    const int64_t it = *st.iter;
    double *o2 = p.omega + p.ostride;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= it; i += (int64_t)gridDim.x * blockDim.x)
        o2[i] = i == 0 ? 0.0 : p.dots_b[i - 1];
}

// Range detection of partial_orthog (Source/top.f90:924-970), one thread.
__global__ void pro_decide_kernel(ProDev p, LzState st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t it = *st.iter;
    double *o2 = p.omega + p.ostride;
    const double seps = sqrt(PRO_EPS);
    const double eta = pow(PRO_EPS, 0.75);
    int64_t firststep = p.meta[0], northog = p.meta[1];
    if (firststep) {
        northog = 0;
        int64_t prev_lo = 0, prev_ro = 0;
        for (int64_t i = 1; i <= it; ++i) {
            if (!(fabs(o2[i]) >= seps)) continue;
            int64_t l = i;
            while (l >= 1 && !(fabs(o2[l]) < eta)) --l;
            if (l < 1) l = 1;
            int64_t r = i;
            while (r <= it && !(fabs(o2[r]) < eta)) ++r;
            if (r > it) r = it;
            if (northog == 0 || (l != prev_lo && r != prev_ro)) {
                if (northog < p.cap) {
                    p.lo[northog] = l;
                    p.ro[northog] = r;
                    ++northog;
                } else {
                    p.meta[5] = 1;
                }
                prev_lo = l;
                prev_ro = r;
            }
        }
        if (northog > 0) {
            // minval(loindex, mask=loindex>0) over the stored ranges, including
            // those left from earlier detections, as the reference's array keeps them
            const int64_t span = northog > p.extent ? northog : p.extent;
            int64_t mn = INT64_MAX;
            for (int64_t k = 0; k < span && k < p.cap; ++k)
                if (p.lo[k] > 0 && p.lo[k] < mn) mn = p.lo[k];
            if (mn > 1 && northog < p.cap) {
                p.lo[northog] = 1;
                p.ro[northog] = 1;
                ++northog;
            }
        }
    }
    int32_t n = 0;
    if (northog > 0) {
        for (int64_t k = 0; k < northog; ++k)
            for (int64_t j = p.lo[k]; j <= p.ro[k]; ++j) {
                if (n < p.jcap) p.jlist[n++] = j;
                else p.meta[5] = 1;
            }
        const uint64_t base = (uint64_t)p.meta[2];
        for (int32_t c = 0; c < n; ++c) o2[p.jlist[c]] = PRO_EPS * (p.noise * noise_gauss(p.seed, base + c, 2.0));
        p.meta[2] = (int64_t)base + n;
        firststep = !firststep;
        p.meta[3] += 1;
        p.meta[4] += n;
    }
    *p.njl = n;
    p.meta[0] = firststep;
    p.meta[1] = northog;
}

__device__ double pro_block_sum(double v) {
    __shared__ double red[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) t += red[i];
    return t;
}

// Classical Gram-Schmidt update (Source/top.f90:980-990): the residual loses
// its components along the listed vectors, then beta(i) = |residual|.
__global__ __launch_bounds__(kBlock) void pro_axpy_kernel(int64_t N, ProDev p, LzState st) {
    const int32_t n = *p.njl;
    if (n == 0) return;
    const int64_t it = *st.iter;
    double *w = st.V + ((it + 1) % st.nslots) * N;
    double part = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (int64_t)gridDim.x * kBlock) {
        double v = w[i];
        for (int32_t c = 0; c < n; ++c) v = v - p.dots_b[c] * st.V[(p.jlist[c] % st.nslots) * N + i];
        w[i] = v;
        part += v * v;
    }
    // last-block reduction of |w|^2 into beta(it)
    __shared__ bool last;
    const double b = pro_block_sum(part);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&st.partials[blockIdx.x], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        last = atomicAdd(st.counters + kCntPro, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double s = 0.0;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += kBlock)
        s += __hip_atomic_load(&st.partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = pro_block_sum(s);
    if (threadIdx.x == 0) {
        st.beta[it] = sqrt(s);
        __hip_atomic_store(st.counters + kCntPro, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

hipError_t pro_init(ProState &ps, const nv_lanczos_opts &o, const Operator &op, const NvOpDev &dev,
                    const LzState &st, std::vector<void *> &allocs, hipStream_t s) {
    (void)dev; (void)st;
    ProDev &p = ps.d;
    p.niter = o.niter;
    p.chkorthog = o.chkorthog < 1 ? 100 : o.chkorthog;
    p.exact = o.reorthog > 1;
    p.seed = o.seed;
    p.noise = o.noise_scale;
    p.epsn = PRO_EPS * (double)op.ns * (double)op.D;
    p.ostride = o.niter + 3;
    p.cap = o.niter + 2;
    p.extent = (o.niter + 99) / 100;
    p.jcap = 2 * (o.niter + 2);
    p.nchunks = 512;
    ps.nchunks = p.nchunks;
    ps.jcap = p.jcap;
    auto alloc = [&](size_t bytes, void **ptr) -> hipError_t {
        hipError_t e = hipMalloc(ptr, bytes);
        if (e == hipSuccess) allocs.push_back(*ptr);
        return e;
    };
    hipError_t e = alloc(2 * p.ostride * sizeof(double), (void **)&p.omega);
    if (e == hipSuccess) e = alloc(p.cap * sizeof(int64_t), (void **)&p.lo);
    if (e == hipSuccess) e = alloc(p.cap * sizeof(int64_t), (void **)&p.ro);
    if (e == hipSuccess) e = alloc(p.jcap * sizeof(int64_t), (void **)&p.jlist);
    if (e == hipSuccess) e = alloc(sizeof(int32_t), (void **)&p.njl);
    if (e == hipSuccess) e = alloc(8 * sizeof(int64_t), (void **)&p.meta);
    if (e == hipSuccess) e = alloc(p.jcap * sizeof(double), (void **)&p.dots_a);
    if (e == hipSuccess) e = alloc(p.jcap * sizeof(double), (void **)&p.dots_b);
    const int64_t nw = (int64_t)p.nchunks * (kBlock / 64);
    if (e == hipSuccess) e = alloc((size_t)p.jcap * nw * 2 * sizeof(double), (void **)&p.dpart);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(p.omega, 0, 2 * p.ostride * sizeof(double), s);
    if (e == hipSuccess) e = hipMemsetAsync(p.lo, 0, p.cap * sizeof(int64_t), s);
    if (e == hipSuccess) e = hipMemsetAsync(p.ro, 0, p.cap * sizeof(int64_t), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pro_init_kernel, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}

static hipError_t multidot(const ProState &ps, int64_t N, const LzState &st, const int64_t *jlist,
                           const int32_t *count_dev, int32_t count_host, int a_cur, int b_res, hipStream_t s) {
    const ProDev &p = ps.d;
    hipLaunchKernelGGL(pro_multidot_kernel, dim3(p.nchunks), dim3(kBlock), 0, s, N, st, jlist, count_dev,
                       count_host, (const double *)nullptr, a_cur, b_res, p.dpart, p.nchunks);
    const int64_t nw = (int64_t)p.nchunks * (kBlock / 64);
    hipLaunchKernelGGL(pro_reduce_kernel, dim3(64), dim3(64), 0, s, count_dev, count_host, p.dpart, nw, p.dots_a,
                       p.dots_b);
    return hipGetLastError();
}

hipError_t pro_set_omega(ProState &ps, const NvOpDev &dev, const LzState &st, int64_t it, hipStream_t s) {
    // `it` sizes the exact-overlap passes; the kernels read the iteration
    // from the device.
    hipError_t e = hipSuccess;
    const ProDev &p = ps.d;
    if (p.exact) {
        if (it > p.jcap) return hipErrorInvalidValue;
        e = multidot(ps, dev.N, st, nullptr, nullptr, (int32_t)it, 0, 1, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(pro_exact_kernel, dim3(16), dim3(kBlock), 0, s, p, st);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(pro_recur_kernel, dim3(1), dim3(kBlock), 0, s, p, st);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (it % p.chkorthog == 0 && it > 1) {
        if (it - 1 > p.jcap) return hipErrorInvalidValue;
        e = multidot(ps, dev.N, st, nullptr, nullptr, (int32_t)(it - 1), 1, 1, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(pro_check_kernel, dim3(16), dim3(kBlock), 0, s, p, st);
        e = hipGetLastError();
    }
    return e;
}

hipError_t pro_partial_orthog(ProState &ps, const NvOpDev &dev, const LzState &st, hipStream_t s) {
    const ProDev &p = ps.d;
    hipLaunchKernelGGL(pro_decide_kernel, dim3(1), dim3(64), 0, s, p, st);
    hipError_t e = multidot(ps, dev.N, st, p.jlist, p.njl, 0, 0, 1, s);
    if (e != hipSuccess) return e;
    const int64_t N = dev.N;
    int g = (int)std::min<int64_t>((N + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(pro_axpy_kernel, dim3(g), dim3(kBlock), 0, s, N, p, st);
    return hipGetLastError();
}

hipError_t pro_get_omega(const ProState &ps, double *omega, int64_t n) {
    const ProDev &p = ps.d;
    std::vector<double> h(2 * p.ostride);
    hipError_t e = hipMemcpy(h.data(), p.omega, h.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (int64_t j = 0; j <= n; ++j) {
        omega[2 * j] = h[j];
        omega[2 * j + 1] = h[p.ostride + j];
    }
    return hipSuccess;
}

void pro_stats(const ProState &ps, int64_t *events, int64_t *touched) {
    int64_t m[8] = {0};
    if (ps.d.meta && hipMemcpy(m, ps.d.meta, sizeof(m), hipMemcpyDeviceToHost) == hipSuccess) {
        *events = m[3];
        *touched = m[4];
    }
}

}  // namespace nv
