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

#define PRO_EPS64 1.1102230246251565e-16  /* 2^-53 (determine_epsilon, Source/secondary.f90:260-271) */

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

// omega recurrence (Source/top.f90:862-874) and the row swap (:875-879), one
// workgroup; bi = beta(it).
__device__ void recur_body(const ProDev &p, const LzState &st, double bi) {
    const int64_t it = *st.iter;
    double *o1 = p.omega, *o2 = p.omega + p.ostride;
    const uint64_t base = (uint64_t)p.meta[2];
    if (threadIdx.x == 0) {
        if (it < p.niter) o2[it + 1] = 1.0;
        o1[it] = p.noise * noise_gauss(p.seed, base, 0.6) * p.epsn * st.beta[1] / bi;
    }
    for (int64_t k = 1 + threadIdx.x; k <= it - 1; k += blockDim.x) {
        o1[k] = (1 / bi) * (st.beta[k] * o2[k + 1] + (st.alpha[k] - st.alpha[it]) * o2[k] +
                            st.beta[k - 1] * o2[k - 1] - st.beta[it - 1] * o1[k]) +
                p.eps * (st.beta[k] + bi) * (p.noise * noise_gauss(p.seed, base + (uint64_t)k, 0.3));
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

__global__ __launch_bounds__(kBlock) void pro_recur_kernel(ProDev p, LzState st) {
    recur_body(p, st, st.beta[*st.iter]);
}

// Dots of a and b (either may be null) with the Krylov vectors q_j listed in
// jlist (or j = c+1 when jlist is null), c < count: a block GEMV over the
// HBM-resident basis.  Work items are (group of G vectors) x (segment of the
// vectors).  The number of segments follows the device-side count, so that
// about kDotItems items exist for any count, each segment holding at least
// kDotMinSeg and at most kBlock * kDotMaxPasses load units.  A lane keeps its
// sums over the whole segment in registers and the block reduces them once
// per item; pro_reduce_kernel sums the segments of a dot in segment order.
// The a and b loads are shared by the G vectors of a group.
constexpr int kGemvCh = 2048;        // elements; dpart holds ceil(N / kGemvCh) segments per dot
constexpr int kDotItems = 2048;
constexpr int kDotMinSeg = kGemvCh;  // elements
constexpr int kDotMaxPasses = 64;

// TB: element type of the basis read (double: the archive itself, slot
// stride N; float: its FP32 copy in mixed precision, slot stride ldf, read in
// element pairs: ldf is even and the pad element is zero).
template <typename TB> struct DotBasis;
template <> struct DotBasis<float> { static constexpr int W = 2, G = 16; };
template <> struct DotBasis<double> { static constexpr int W = 1, G = 8; };

template <typename TB>
__device__ __forceinline__ const TB *basis_of(const LzState &st);
template <> __device__ __forceinline__ const double *basis_of<double>(const LzState &st) { return st.V; }
template <> __device__ __forceinline__ const float *basis_of<float>(const LzState &st) { return st.Vf; }
template <typename TB>
__device__ __forceinline__ int64_t stride_of(const LzState &st, int64_t N);
template <> __device__ __forceinline__ int64_t stride_of<double>(const LzState &, int64_t N) { return N; }
template <> __device__ __forceinline__ int64_t stride_of<float>(const LzState &st, int64_t) { return st.ldf; }

// count: *count_dev, or iter + count_off when count_dev is null.
__device__ __forceinline__ int32_t dot_count(const LzState &st, const int32_t *count_dev, int32_t count_off) {
    return count_dev ? *count_dev : (int32_t)(*st.iter + count_off);
}

// Segments of a multidot over `count` vectors; seg_units: load units (W
// elements) per segment, a multiple of kBlock.  At most ceil(N / kGemvCh).
template <typename TB>
__device__ __forceinline__ int64_t dot_segments(int64_t N, int32_t count, int64_t &seg_units) {
    constexpr int W = DotBasis<TB>::W, G = DotBasis<TB>::G;
    const int64_t units = (N + W - 1) / W;
    const int64_t ngroups = (count + G - 1) / G;
    const int64_t lo = (units + (int64_t)kBlock * kDotMaxPasses - 1) / ((int64_t)kBlock * kDotMaxPasses);
    const int64_t hi = (units + kDotMinSeg / W - 1) / (kDotMinSeg / W);
    int64_t nseg = (kDotItems + ngroups - 1) / ngroups;
    nseg = nseg < lo ? lo : (nseg > hi ? hi : nseg);
    int64_t per = (units + nseg - 1) / nseg;
    per = (per + kBlock - 1) / kBlock * kBlock;
    seg_units = per;
    return (units + per - 1) / per;
}

// One work item: the dots of a and b with the nk <= G vectors c0 .. c0+nk-1
// over segment sg.  Slots k >= nk repeat the last vector (their loads hit the
// cache and their sums are dropped): the loads stay unconditional, which
// keeps them batched.
template <typename TB, int G, int GMAX>
__device__ __forceinline__ void dot_item(int64_t N, const LzState &st, const double *a, const double *b,
                                         const int64_t *jlist, int64_t c0, int nk, int64_t sg, int64_t seg_units,
                                         int64_t nseg, double (*red)[2 * GMAX], double *part) {
    constexpr int W = DotBasis<TB>::W;
    const TB *basis = basis_of<TB>(st);
    const int64_t ld = stride_of<TB>(st, N);
    const int64_t units = (N + W - 1) / W;
    const TB *q[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        const int64_t c = c0 + (k < nk ? k : nk - 1);
        q[k] = basis + ((jlist ? jlist[c] : c + 1) % st.nslots) * ld;
    }
    double sa[G], sb[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
        sa[k] = 0.0;
        sb[k] = 0.0;
    }
    const int64_t u_end = (sg + 1) * seg_units < units ? (sg + 1) * seg_units : units;
    for (int64_t u = sg * seg_units + threadIdx.x; u < u_end; u += kBlock) {
        if constexpr (W == 2) {
            const int64_t i = 2 * u;
            const bool two = i + 1 < N;
            const double a0 = a ? a[i] : 0.0, a1 = (a && two) ? a[i + 1] : 0.0;
            const double b0 = b ? b[i] : 0.0, b1 = (b && two) ? b[i + 1] : 0.0;
            float2 v[G];
#pragma unroll
            for (int k = 0; k < G; ++k) v[k] = reinterpret_cast<const float2 *>(q[k])[u];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                sa[k] += a0 * (double)v[k].x + a1 * (double)v[k].y;
                sb[k] += b0 * (double)v[k].x + b1 * (double)v[k].y;
            }
        } else {
            const double av = a ? a[u] : 0.0, bv = b ? b[u] : 0.0;
            double v[G];
#pragma unroll
            for (int k = 0; k < G; ++k) v[k] = (double)q[k][u];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                sa[k] += av * v[k];
                sb[k] += bv * v[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            sa[k] += __shfl_xor(sa[k], off, 64);
            sb[k] += __shfl_xor(sb[k], off, 64);
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < G; ++k) {
            red[wid][k] = sa[k];
            red[wid][G + k] = sb[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * G) {
        const int k = threadIdx.x % G, which = threadIdx.x / G;
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += red[w][threadIdx.x];
        if (k < nk) part[((c0 + k) * nseg + sg) * 2 + which] = t;
    }
}

template <typename TB>
__global__ __launch_bounds__(kBlock) void pro_multidot_kernel(int64_t N, LzState st, const int64_t *jlist,
                                                              const int32_t *count_dev, int32_t count_off,
                                                              int a_cur, int b_res, double *part) {
    constexpr int G = DotBasis<TB>::G;
    __shared__ double red[kBlock / 64][2 * G];
    const int64_t it = *st.iter;
    const int32_t count = dot_count(st, count_dev, count_off);
    if (count <= 0) return;
    const double *a = a_cur ? st.V + (it % st.nslots) * N : nullptr;
    const double *b = b_res ? st.V + ((it + 1) % st.nslots) * N : nullptr;
    int64_t seg_units;
    const int64_t nseg = dot_segments<TB>(N, count, seg_units);
    const int64_t ngroups = (count + G - 1) / G;
    const int64_t items = ngroups * nseg;
    for (int64_t item = blockIdx.x; item < items; item += gridDim.x) {
        const int64_t g = item / nseg, sg = item - g * nseg;
        const int nk = (int)((count - g * G) < G ? (count - g * G) : G);
        // the last group runs at the narrowest width that holds it
        if (nk > G / 2)
            dot_item<TB, G, G>(N, st, a, b, jlist, g * G, nk, sg, seg_units, nseg, red, part);
        else if (G > 8 && nk > 4)
            dot_item<TB, (G > 8 ? 8 : G), G>(N, st, a, b, jlist, g * G, nk, sg, seg_units, nseg, red, part);
        else if (nk > 2)
            dot_item<TB, 4, G>(N, st, a, b, jlist, g * G, nk, sg, seg_units, nseg, red, part);
        else
            dot_item<TB, 2, G>(N, st, a, b, jlist, g * G, nk, sg, seg_units, nseg, red, part);
    }
}

// One wave per dot: sums its segment partials in segment order.
template <typename TB>
__global__ __launch_bounds__(kBlock) void pro_reduce_kernel(int64_t N, LzState st, const int32_t *count_dev,
                                                            int32_t count_off, const double *part, double *out_a,
                                                            double *out_b) {
    const int32_t count = dot_count(st, count_dev, count_off);
    if (count <= 0) return;
    int64_t seg_units;
    const int64_t nseg = dot_segments<TB>(N, count, seg_units);
    const int lane = threadIdx.x & 63;
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t c = w0; c < count; c += nw) {
        double sa = 0.0, sb = 0.0;
        for (int64_t k = lane; k < nseg; k += 64) {
            sa += part[(c * nseg + k) * 2 + 0];
            sb += part[(c * nseg + k) * 2 + 1];
        }
        for (int off = 32; off > 0; off >>= 1) {
            sa += __shfl_xor(sa, off, 64);
            sb += __shfl_xor(sb, off, 64);
        }
        if (lane == 0) {
            out_a[c] = sa;
            out_b[c] = sb;
        }
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

// Range detection of partial_orthog (Source/top.f90:924-970), block-parallel.
// The reference walks from every index with |omega| >= sqrt(eps) left and
// right to the nearest indices with |omega| < eps^(3/4) and keeps a range when
// both ends differ from the last kept one.  Both ends are monotone in the
// index, so the kept ranges are exactly the gaps between consecutive "small"
// indices (with 0 and iter+1 as sentinels, ends clipped to [1, iter]) that
// hold at least one "large" index: two scans and a compaction find them.
// Then the reference's extra range [1,1], the list of vectors and the omega
// resets (a vector on the boundary of two ranges keeps the later draw, as the
// reference's sequential loop leaves it).
constexpr int kDecide = 1024;

__device__ int64_t scan_excl(int64_t v, int64_t *total) {
    __shared__ int64_t wsum[kDecide / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    __syncthreads();
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int64_t before = 0, tot = 0;
    for (int w = 0; w < kDecide / 64; ++w) {
        if (w < wid) before += wsum[w];
        tot += wsum[w];
    }
    *total = tot;
    return before + x - v;
}

__device__ void decide_body(const ProDev &p, const LzState &st) {
    __shared__ int64_t sh_nr, sh_n, sh_first;
    const int64_t it = *st.iter;
    double *o2 = p.omega + p.ostride;
    int32_t *S = p.scr;                 // small indices in order
    uint8_t *act = reinterpret_cast<uint8_t *>(p.scr + (it + 3));  // gap holds a large index
    const int64_t seg = (it + kDecide - 1) / kDecide;  // indices 1..it, <= 32 per thread
    const int64_t i0 = 1 + (int64_t)threadIdx.x * seg;
    const int64_t i1 = i0 + seg - 1 < it ? i0 + seg - 1 : it;
    if (threadIdx.x == 0) sh_first = p.meta[0];
    __syncthreads();
    const bool firststep = sh_first != 0;
    int64_t tot;
    if (firststep) {
        uint32_t smallm = 0, bigm = 0;
        for (int64_t i = i0; i <= i1; ++i) {
            const double w = fabs(o2[i]);
            if (w < p.eta) smallm |= 1u << (i - i0);
            if (w >= p.seps) bigm |= 1u << (i - i0);
        }
        const int64_t sbase = scan_excl(__popc(smallm), &tot);
        const int64_t m = tot;  // number of small indices; gaps 0..m
        for (int64_t g = threadIdx.x; g <= m; g += kDecide) act[g] = 0;
        __syncthreads();
        int64_t k = sbase;
        for (int64_t i = i0; i <= i1; ++i) {
            const uint32_t bit = 1u << (i - i0);
            if (bigm & bit) act[k] = 1;
            if (smallm & bit) S[k++] = (int32_t)i;
        }
        __syncthreads();
        // compaction of the active gaps, in order
        const int64_t gseg = (m + 1 + kDecide - 1) / kDecide;
        const int64_t g0 = (int64_t)threadIdx.x * gseg;
        const int64_t g1 = g0 + gseg < m + 1 ? g0 + gseg : m + 1;
        int64_t cnt = 0;
        for (int64_t g = g0; g < g1; ++g) cnt += act[g];
        int64_t r0 = scan_excl(cnt, &tot);
        for (int64_t g = g0; g < g1; ++g) {
            if (!act[g]) continue;
            if (r0 < p.cap) {
                const int64_t a = g == 0 ? 0 : S[g - 1];
                const int64_t b = g == m ? it + 1 : S[g];
                p.lo[r0] = a > 1 ? a : 1;
                p.ro[r0] = b < it ? b : it;
            }
            ++r0;
        }
        int64_t northog = tot < p.cap ? tot : p.cap;
        __syncthreads();
        if (northog > 0) {
            // minval(loindex, mask=loindex>0) over the stored ranges, including
            // those left from earlier detections, as the reference's array keeps them
            const int64_t span = northog > p.extent ? northog : p.extent;
            int64_t mn = INT64_MAX;
            for (int64_t q = threadIdx.x; q < span && q < p.cap; q += kDecide)
                if (p.lo[q] > 0 && p.lo[q] < mn) mn = p.lo[q];
            for (int off = 32; off > 0; off >>= 1) {
                const int64_t y = __shfl_xor(mn, off, 64);
                mn = y < mn ? y : mn;
            }
            __shared__ int64_t wmin[kDecide / 64];
            if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = mn;
            __syncthreads();
            if (threadIdx.x == 0) {
                int64_t v = INT64_MAX;
                for (int w = 0; w < kDecide / 64; ++w) v = wmin[w] < v ? wmin[w] : v;
                if (v > 1 && northog < p.cap) {
                    p.lo[northog] = 1;
                    p.ro[northog] = 1;
                    ++northog;
                }
                if (tot > p.cap) p.meta[5] = 1;
                sh_nr = northog;
            }
        } else if (threadIdx.x == 0) {
            sh_nr = 0;
        }
    } else if (threadIdx.x == 0) {
        sh_nr = p.meta[1];
    }
    __syncthreads();
    const int64_t nr = sh_nr;
    // list offsets
    {
        const int64_t rseg = (nr + kDecide - 1) / kDecide;
        const int64_t k0 = (int64_t)threadIdx.x * rseg;
        const int64_t k1 = k0 + rseg < nr ? k0 + rseg : nr;
        int64_t len = 0;
        for (int64_t q = k0; q < k1; ++q) len += p.ro[q] - p.lo[q] + 1;
        int64_t off = scan_excl(len, &tot);
        for (int64_t q = k0; q < k1; ++q) {
            p.roff[q] = off;
            off += p.ro[q] - p.lo[q] + 1;
        }
        if (threadIdx.x == 0) {
            int64_t n = tot;
            if (n > p.jcap) {
                p.meta[5] = 1;
                n = p.jcap;
            }
            sh_n = n;
        }
    }
    __syncthreads();
    const int64_t n = sh_n;
    const uint64_t base = (uint64_t)p.meta[2];
    for (int64_t c = threadIdx.x; c < n; c += kDecide) {
        int64_t lo = 0, hi = nr - 1;  // last range with roff <= c
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (p.roff[mid] <= c) lo = mid;
            else hi = mid - 1;
        }
        const int64_t k = lo;
        const int64_t j = p.lo[k] + (c - p.roff[k]);
        p.jlist[c] = j;
        const bool shadowed = (j == p.ro[k]) && (k + 1 < nr) && (p.lo[k + 1] == j) && (p.roff[k + 1] < n);
        if (!shadowed) o2[j] = p.eps * (p.noise * noise_gauss(p.seed, base + (uint64_t)c, 2.0));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (nr > 0) {
            p.meta[0] = !firststep;
            p.meta[2] = (int64_t)base + n;
            p.meta[3] += 1;
            p.meta[4] += n;
        }
        p.meta[1] = nr;
        *p.njl = (int32_t)n;
    }
}

__global__ __launch_bounds__(kDecide) void pro_decide_kernel(ProDev p, LzState st) { decide_body(p, st); }

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
// its components along the listed vectors, subtracted in list order, then
// beta(i) = |residual|.  Eight basis vectors are loaded before they are used
// so that the loads overlap.
template <typename TB>
__global__ __launch_bounds__(kBlock) void pro_axpy_kernel(int64_t N, ProDev p, LzState st) {
    const int32_t n = *p.njl;
    if (n == 0) return;
    const int64_t it = *st.iter;
    double *w = st.V + ((it + 1) % st.nslots) * N;
    const TB *V = basis_of<TB>(st);
    const int64_t ld = stride_of<TB>(st, N);
    const int64_t ns = st.nslots;
    double part = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (int64_t)gridDim.x * kBlock) {
        double v = w[i];
        int32_t c = 0;
        for (; c + 8 <= n; c += 8) {
            double q[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] = (double)V[(p.jlist[c + k] % ns) * ld + i];
#pragma unroll
            for (int k = 0; k < 8; ++k) v = v - p.dots_b[c + k] * q[k];
        }
        for (; c < n; ++c) v = v - p.dots_b[c] * (double)V[(p.jlist[c] % ns) * ld + i];
        w[i] = v;
        part += v * v;
    }
    const double b = pro_block_sum(part);
    if (threadIdx.x == 0) st.partials[blockIdx.x] = b;
}

// beta(it) = |residual| after the update (only when vectors were listed).
__global__ __launch_bounds__(kBlock) void pro_beta_kernel(ProDev p, LzState st, int nparts) {
    if (*p.njl == 0) return;
    double s = 0.0;
#pragma unroll 8
    for (int i = threadIdx.x; i < nparts; i += kBlock) s += st.partials[i];
    s = pro_block_sum(s);
    if (threadIdx.x == 0) st.beta[*st.iter] = sqrt(s);
}

// An iteration without an exact check in one workgroup: beta(it) = |r| from
// lanczos_step's block partials (summed as finish_sum_kernel sums them: a
// 256-thread strided pass, then four wave sums in order), the omega
// recurrence and the range decision -- three launches of the per-call path.
__global__ __launch_bounds__(kDecide) void pro_step_kernel(ProDev p, LzState st, int nparts) {
    __shared__ double red[kBlock / 64];
    __shared__ double sh_beta;
    const int64_t it = *st.iter;
    double v = 0.0;
    if (threadIdx.x < kBlock) {
#pragma unroll 8
        for (int i = threadIdx.x; i < nparts; i += kBlock) v += st.partials[i];
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (threadIdx.x < kBlock && (threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kBlock / 64; ++i) t += red[i];
        const double b = sqrt(t);
        st.beta[it] = b;
        sh_beta = b;
    }
    __syncthreads();
    recur_body(p, st, sh_beta);
    __syncthreads();
    decide_body(p, st);
}

// Mixed modes with the default (FP64) roundoff model: FP32 rounding of H*v
// alone keeps the overlaps near 1e-8, at sqrt(2^-53) itself, so the omega
// recurrence (which models FP64 rounding) misses the loss between checks and
// every check would re-orthogonalise against the whole basis.  These modes
// check every min(chkorthog, 10) iterations and use thresholds above that
// level (NV_PRO_MIXED_TAU, NV_PRO_MIXED_ETA).  A caller that sets pro_eps
// chooses the period and gets sqrt(eps), eps^(3/4).
static bool mixed_model(const nv_lanczos_opts &o) { return o.prec != NV_PREC_F64 && !(o.pro_eps > 0); }

static int64_t chk_period(const nv_lanczos_opts &o) {
    int64_t c = o.chkorthog < 1 ? 100 : o.chkorthog;
    if (mixed_model(o) && c > 10) c = 10;
    return c;
}

// Thresholds and roundoff of the omega model (Source/top.f90:862-874, :922-950).
static void pro_thresholds(ProDev &p, const nv_lanczos_opts &o) {
    p.eps = o.pro_eps > 0 ? o.pro_eps : PRO_EPS64;
    p.seps = o.pro_tau > 0 ? o.pro_tau : (mixed_model(o) ? NV_PRO_MIXED_TAU : sqrt(p.eps));
    p.eta = o.pro_eta > 0 ? o.pro_eta : (mixed_model(o) ? NV_PRO_MIXED_ETA : pow(p.eps, 0.75));
}

void pro_effective_thresholds(const nv_lanczos_opts &o, double *tau, double *eta) {
    ProDev p{};
    pro_thresholds(p, o);
    *tau = p.seps;
    *eta = p.eta;
}

hipError_t pro_init(ProState &ps, const nv_lanczos_opts &o, const Operator &op, const NvOpDev &dev,
                    const LzState &st, std::vector<void *> &allocs, hipStream_t s) {
    (void)dev; (void)st;
    ProDev &p = ps.d;
    p.niter = o.niter;
    p.chkorthog = chk_period(o);
    p.exact = o.reorthog > 1;
    p.seed = o.seed;
    p.noise = o.noise_scale;
    pro_thresholds(p, o);
    p.epsn_scale = (double)op.ns * (double)op.D;
    p.epsn = p.eps * p.epsn_scale;
    p.ostride = o.niter + 3;
    p.cap = o.niter + 2;
    p.extent = (o.niter + 99) / 100;
    p.jcap = 2 * (o.niter + 2);
    p.nchunks = (int32_t)((op.N + kGemvCh - 1) / kGemvCh);
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
    if (e == hipSuccess) e = alloc(p.cap * sizeof(int64_t), (void **)&p.roff);
    if (e == hipSuccess) e = alloc((size_t)2 * (o.niter + 4) * sizeof(int32_t), (void **)&p.scr);
    if (e == hipSuccess) e = alloc((size_t)p.jcap * p.nchunks * 2 * sizeof(double), (void **)&p.dpart);
    if (e != hipSuccess) return e;
    e = hipMemset(p.omega, 0, 2 * p.ostride * sizeof(double));
    if (e == hipSuccess) e = hipMemset(p.lo, 0, p.cap * sizeof(int64_t));
    if (e == hipSuccess) e = hipMemset(p.ro, 0, p.cap * sizeof(int64_t));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pro_init_kernel, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t pro_reset(ProState &ps, const nv_lanczos_opts &o, hipStream_t s) {
    ProDev &p = ps.d;
    p.chkorthog = chk_period(o);
    p.seed = o.seed;
    p.noise = o.noise_scale;
    pro_thresholds(p, o);
    p.epsn = p.eps * (double)p.epsn_scale;
    hipError_t e = hipMemset(p.omega, 0, 2 * p.ostride * sizeof(double));
    if (e == hipSuccess) e = hipMemset(p.lo, 0, p.cap * sizeof(int64_t));
    if (e == hipSuccess) e = hipMemset(p.ro, 0, p.cap * sizeof(int64_t));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pro_init_kernel, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}

static hipError_t multidot(const ProState &ps, int64_t N, const LzState &st, const int64_t *jlist,
                           const int32_t *count_dev, int32_t count_off, int a_cur, int b_res, hipStream_t s) {
    const ProDev &p = ps.d;
    if (st.Vf) {
        hipLaunchKernelGGL(pro_multidot_kernel<float>, dim3(kDotItems), dim3(kBlock), 0, s, N, st, jlist, count_dev,
                           count_off, a_cur, b_res, p.dpart);
        hipLaunchKernelGGL(pro_reduce_kernel<float>, dim3(256), dim3(kBlock), 0, s, N, st, count_dev, count_off,
                           p.dpart, p.dots_a, p.dots_b);
    } else {
        hipLaunchKernelGGL(pro_multidot_kernel<double>, dim3(kDotItems), dim3(kBlock), 0, s, N, st, jlist,
                           count_dev, count_off, a_cur, b_res, p.dpart);
        hipLaunchKernelGGL(pro_reduce_kernel<double>, dim3(256), dim3(kBlock), 0, s, N, st, count_dev, count_off,
                           p.dpart, p.dots_a, p.dots_b);
    }
    return hipGetLastError();
}

size_t pro_decide_lds(int64_t niter) { return niter <= 32 * kDecide ? 0 : SIZE_MAX; }

int64_t pro_check_period(const ProState &ps) { return ps.d.chkorthog; }

hipError_t pro_set_omega(ProState &ps, const NvOpDev &dev, const LzState &st, int check, hipStream_t s) {
    hipError_t e = hipSuccess;
    const ProDev &p = ps.d;
    if (p.exact) {
        // omega(2, 1:iter) = q3 . q_j (count = iter)
        e = multidot(ps, dev.N, st, nullptr, nullptr, 0, 0, 1, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(pro_exact_kernel, dim3(16), dim3(kBlock), 0, s, p, st);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(pro_recur_kernel, dim3(1), dim3(kBlock), 0, s, p, st);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (check) {
        // q2 . q_j and q3 . q_j for j < iter (count = iter - 1)
        e = multidot(ps, dev.N, st, nullptr, nullptr, -1, 1, 1, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(pro_check_kernel, dim3(16), dim3(kBlock), 0, s, p, st);
        e = hipGetLastError();
    }
    return e;
}

// The dots and the update of the decided list, and beta after it.
static hipError_t orthog_apply(const ProState &ps, const NvOpDev &dev, const LzState &st, hipStream_t s) {
    const ProDev &p = ps.d;
    hipError_t e = multidot(ps, dev.N, st, p.jlist, p.njl, 0, 0, 1, s);  // count = list length
    if (e != hipSuccess) return e;
    const int64_t N = dev.N;
    int g = (int)std::min<int64_t>((N + kBlock - 1) / kBlock, 4096);
    if (st.Vf)
        hipLaunchKernelGGL(pro_axpy_kernel<float>, dim3(g), dim3(kBlock), 0, s, N, p, st);
    else
        hipLaunchKernelGGL(pro_axpy_kernel<double>, dim3(g), dim3(kBlock), 0, s, N, p, st);
    hipLaunchKernelGGL(pro_beta_kernel, dim3(1), dim3(kBlock), 0, s, p, st, g);
    return hipGetLastError();
}

hipError_t pro_partial_orthog(ProState &ps, const NvOpDev &dev, const LzState &st, hipStream_t s) {
    hipLaunchKernelGGL(pro_decide_kernel, dim3(1), dim3(kDecide), 0, s, ps.d, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return orthog_apply(ps, dev, st, s);
}

bool pro_fusable(const ProState &ps) { return !ps.d.exact; }

hipError_t pro_step_fused(ProState &ps, const NvOpDev &dev, const LzState &st, int nparts, hipStream_t s) {
    hipLaunchKernelGGL(pro_step_kernel, dim3(1), dim3(kDecide), 0, s, ps.d, st, nparts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return orthog_apply(ps, dev, st, s);
}

// ints: [0] n, [1] cap, [2] jcap, [3] njl, [4..11] meta, then lo[cap], ro[cap], jlist[jcap];
// reals: omega(1, 0..n+1), omega(2, 0..n+1).
hipError_t pro_export(const ProState &ps, int64_t n, std::vector<int64_t> &ints, std::vector<double> &reals) {
    const ProDev &p = ps.d;
    if (n + 1 >= p.ostride) return hipErrorInvalidValue;
    ints.assign(12 + 2 * p.cap + p.jcap, 0);
    ints[0] = n;
    ints[1] = p.cap;
    ints[2] = p.jcap;
    int32_t njl = 0;
    hipError_t e = hipMemcpy(&njl, p.njl, sizeof(int32_t), hipMemcpyDeviceToHost);
    ints[3] = njl;
    if (e == hipSuccess) e = hipMemcpy(ints.data() + 4, p.meta, 8 * sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ints.data() + 12, p.lo, p.cap * sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(ints.data() + 12 + p.cap, p.ro, p.cap * sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(ints.data() + 12 + 2 * p.cap, p.jlist, p.jcap * sizeof(int64_t), hipMemcpyDeviceToHost);
    reals.assign(2 * (n + 2), 0.0);
    if (e == hipSuccess) e = hipMemcpy(reals.data(), p.omega, (n + 2) * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(reals.data() + n + 2, p.omega + p.ostride, (n + 2) * sizeof(double), hipMemcpyDeviceToHost);
    return e;
}

hipError_t pro_import(ProState &ps, int64_t n, const std::vector<int64_t> &ints, const std::vector<double> &reals) {
    ProDev &p = ps.d;
    if (ints.size() < 12 || ints[0] != n || n + 1 >= p.ostride) return hipErrorInvalidValue;
    const int64_t cap = ints[1], jcap = ints[2];
    if (cap < 0 || jcap < 0 || (int64_t)ints.size() != 12 + 2 * cap + jcap || (int64_t)reals.size() != 2 * (n + 2))
        return hipErrorInvalidValue;
    const int32_t njl = (int32_t)ints[3];
    // ranges and the list past this run's capacity cannot be restored
    if (ints[5] > p.cap || njl > p.jcap) return hipErrorInvalidValue;
    const int64_t ncap = std::min(cap, p.cap), njc = std::min(jcap, p.jcap);
    hipError_t e = hipMemcpy(p.njl, &njl, sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p.meta, ints.data() + 4, 8 * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(p.lo, 0, p.cap * sizeof(int64_t));
    if (e == hipSuccess) e = hipMemset(p.ro, 0, p.cap * sizeof(int64_t));
    if (e == hipSuccess) e = hipMemcpy(p.lo, ints.data() + 12, ncap * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p.ro, ints.data() + 12 + cap, ncap * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p.jlist, ints.data() + 12 + 2 * cap, njc * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(p.omega, 0, 2 * p.ostride * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(p.omega, reals.data(), (n + 2) * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p.omega + p.ostride, reals.data() + n + 2, (n + 2) * sizeof(double), hipMemcpyHostToDevice);
    return e;
}

hipError_t pro_put_omega(ProState &ps, const double *omega, int64_t n) {
    ProDev &p = ps.d;
    if (n + 1 >= p.ostride) return hipErrorInvalidValue;
    std::vector<double> h(2 * p.ostride, 0.0);
    hipError_t e = hipMemcpy(h.data(), p.omega, h.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (int64_t j = 0; j <= n; ++j) {
        h[j] = omega[2 * j];
        h[p.ostride + j] = omega[2 * j + 1];
    }
    return hipMemcpy(p.omega, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
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
