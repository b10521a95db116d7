// Ritz vectors and their analysis (identify_roots, Source/top.f90:1122-1283).
//
// The reference forms each Ritz vector Y_i = sum_k Tvecs(k,i) q_k with one
// pass over the archived Krylov basis per root (:1177-1183), then for each
// root repeatedly selects the largest |Y_i|^2 element (:1191-1200), the
// per-state weights (:1203-1210) and, for the spin-orbit roots, the overlaps
// Y_i[:,k] . Y_j[:,l] (:1227-1243).  Here one pass over the HBM-resident basis
// forms up to kRitzChunk vectors at once, the selections run as block-level
// argmax reductions over all roots in one launch per rank, and every overlap
// comes from one Gram matrix of the (dimen x nstates*nroots) block of Ritz
// vectors.
#include <math.h>

#include "nv_kernels.hpp"

namespace nv {

// ---- Y = V S ----------------------------------------------------------------
// Y[c0+c][r] = sum_k S(k, c0+c) V_k[r], k = 0..n-1, in increasing k (the
// reference's GA_ADD order).  V_k at V + ((slot0 + k) % nslots) * N (the
// basis or a circular segment buffer); S component-major with row stride ldv,
// vec[k*ldv + i] = component k of root i.
__global__ __launch_bounds__(256) void ritz_vectors_kernel(const double *__restrict__ V, int64_t N, int64_t n,
                                                          const double *__restrict__ vec, int64_t ldv, int64_t c0,
                                                          int nc, int64_t slot0, int64_t nslots,
                                                          double *__restrict__ Y) {
    __shared__ double s[kRitzKB][kRitzChunk];
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double acc[kRitzChunk];
#pragma unroll
    for (int c = 0; c < kRitzChunk; ++c) acc[c] = 0.0;
    for (int64_t k0 = 0; k0 < n; k0 += kRitzKB) {
        const int kb = (int)(n - k0 < kRitzKB ? n - k0 : kRitzKB);
        __syncthreads();
        for (int t = threadIdx.x; t < kRitzKB * kRitzChunk; t += 256) {
            const int kk = t / kRitzChunk, cc = t % kRitzChunk;
            s[kk][cc] = (kk < kb && cc < nc) ? vec[(k0 + kk) * ldv + c0 + cc] : 0.0;
        }
        __syncthreads();
        if (r < N) {
            // the slots of this stage: one run up to the end of the buffer, then from its start
            const int64_t first = (slot0 + k0) % nslots;
            const int run = (int)(nslots - first < kb ? nslots - first : kb);
            const double *va = V + first * N + r;
#pragma unroll 4
            for (int kk = 0; kk < run; ++kk) {
                const double v = va[(int64_t)kk * N];
#pragma unroll
                for (int c = 0; c < kRitzChunk; ++c) acc[c] = fma(s[kk][c], v, acc[c]);
            }
            const double *vb = V + r - (int64_t)run * N;
#pragma unroll 4
            for (int kk = run; kk < kb; ++kk) {
                const double v = vb[(int64_t)kk * N];
#pragma unroll
                for (int c = 0; c < kRitzChunk; ++c) acc[c] = fma(s[kk][c], v, acc[c]);
            }
        }
    }
    if (r < N) {
#pragma unroll
        for (int c = 0; c < kRitzChunk; ++c)
            if (c < nc) Y[(c0 + c) * N + r] = acc[c];
    }
}

// ---- ranked selection of the largest |Y|^2 ---------------------------------
// Order: larger magnitude first, equal magnitudes by increasing element index
// (Fortran column-major order of the (dimen, nstates) array).  Pass k returns,
// per root, the first element after the one returned by pass k-1.
__device__ __forceinline__ bool nv_ranks_before(double m1, int64_t i1, double m2, int64_t i2) {
    return m1 > m2 || (m1 == m2 && i1 < i2);
}

__device__ void nv_block_argmax(double &m, int64_t &i) {
    __shared__ double sm[4];
    __shared__ int64_t si[4];
    for (int off = 32; off > 0; off >>= 1) {
        const double om = __shfl_xor(m, off, 64);
        const int64_t oi = __shfl_xor(i, off, 64);
        if (oi >= 0 && (i < 0 || nv_ranks_before(om, oi, m, i))) { m = om; i = oi; }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) { sm[threadIdx.x >> 6] = m; si[threadIdx.x >> 6] = i; }
    __syncthreads();
    m = sm[0]; i = si[0];
    for (int w = 1; w < 4; ++w)
        if (si[w] >= 0 && (i < 0 || nv_ranks_before(sm[w], si[w], m, i))) { m = sm[w]; i = si[w]; }
}

__global__ __launch_bounds__(256) void topk_partial_kernel(const double *__restrict__ Y, int64_t N,
                                                          const double *__restrict__ prev_m,
                                                          const int64_t *__restrict__ prev_i,
                                                          double *__restrict__ pm, int64_t *__restrict__ pi) {
    const int root = blockIdx.y;
    const double *y = Y + (int64_t)root * N;
    const double lm = prev_m[root];
    const int64_t li = prev_i[root];
    double bm = -1.0;
    int64_t bi = -1;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < N; e += (int64_t)gridDim.x * 256) {
        const double v = y[e];
        const double m = v * v;
        const bool after = m < lm || (m == lm && e > li);
        if (after && (bi < 0 || nv_ranks_before(m, e, bm, bi))) { bm = m; bi = e; }
    }
    nv_block_argmax(bm, bi);
    if (threadIdx.x == 0) {
        pm[(int64_t)root * gridDim.x + blockIdx.x] = bm;
        pi[(int64_t)root * gridDim.x + blockIdx.x] = bi;
    }
}

__global__ __launch_bounds__(256) void topk_finish_kernel(int nb, const double *__restrict__ pm,
                                                         const int64_t *__restrict__ pi,
                                                         const double *__restrict__ Y, int64_t N,
                                                         double *__restrict__ prev_m, int64_t *__restrict__ prev_i,
                                                         int k, int kmax, double *__restrict__ out_m,
                                                         int64_t *__restrict__ out_i, int32_t *__restrict__ out_ph) {
    const int root = blockIdx.x;
    double bm = -1.0;
    int64_t bi = -1;
    for (int b = threadIdx.x; b < nb; b += 256) {
        const double m = pm[(int64_t)root * nb + b];
        const int64_t i = pi[(int64_t)root * nb + b];
        if (i >= 0 && (bi < 0 || nv_ranks_before(m, i, bm, bi))) { bm = m; bi = i; }
    }
    nv_block_argmax(bm, bi);
    if (threadIdx.x == 0) {
        const int64_t o = (int64_t)root * kmax + k;
        out_m[o] = bi >= 0 ? bm : 0.0;
        out_i[o] = bi;
        out_ph[o] = (bi >= 0 && Y[(int64_t)root * N + bi] < 0.0) ? -1 : 1;
        if (bi >= 0) { prev_m[root] = bm; prev_i[root] = bi; }
    }
}

__global__ void topk_init_kernel(double *prev_m, int64_t *prev_i, int nroots) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nroots) { prev_m[t] = INFINITY; prev_i[t] = -1; }
}

// ---- Gram matrix G = A^T A of a D x R column block ---------------------------
// Column c of A starts at A + c*D.  Tile pairs (ti <= tj) of kGramT columns;
// blockIdx.y splits the rows into chunks whose partial tiles are summed by
// gram_finish_kernel.
__device__ __forceinline__ void nv_tile_pair(int p, int ntile, int &ti, int &tj) {
    ti = 0;
    while (p >= ntile - ti) { p -= ntile - ti; ++ti; }
    tj = ti + p;
}

__global__ __launch_bounds__(256) void gram_partial_kernel(const double *__restrict__ A, int64_t D, int R, int ntile,
                                                          int64_t rows_per_chunk, double *__restrict__ part) {
    __shared__ double As[kGramT][kGramT + 1];
    __shared__ double Bs[kGramT][kGramT + 1];
    int ti, tj;
    nv_tile_pair(blockIdx.x, ntile, ti, tj);
    const int npairs = gridDim.x;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t lo = (int64_t)blockIdx.y * rows_per_chunk;
    const int64_t hi = lo + rows_per_chunk < D ? lo + rows_per_chunk : D;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int64_t r0 = lo; r0 < hi; r0 += kGramT) {
        __syncthreads();
        for (int t = threadIdx.x; t < kGramT * kGramT; t += 256) {
            const int rr = t % kGramT, cc = t / kGramT;
            const int64_t row = r0 + rr;
            const int ca = ti * kGramT + cc, cb = tj * kGramT + cc;
            As[rr][cc] = (row < hi && ca < R) ? A[(int64_t)ca * D + row] : 0.0;
            Bs[rr][cc] = (row < hi && cb < R) ? A[(int64_t)cb * D + row] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int rr = 0; rr < kGramT; ++rr) {
            const double a0 = As[rr][ty * 2], a1 = As[rr][ty * 2 + 1];
            const double b0 = Bs[rr][tx * 2], b1 = Bs[rr][tx * 2 + 1];
            acc[0][0] = fma(a0, b0, acc[0][0]);
            acc[0][1] = fma(a0, b1, acc[0][1]);
            acc[1][0] = fma(a1, b0, acc[1][0]);
            acc[1][1] = fma(a1, b1, acc[1][1]);
        }
    }
    double *p = part + ((int64_t)blockIdx.y * npairs + blockIdx.x) * (kGramT * kGramT);
    for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v) p[(ty * 2 + u) * kGramT + tx * 2 + v] = acc[u][v];
}

__global__ __launch_bounds__(256) void gram_finish_kernel(const double *__restrict__ part, int nchunks, int ntile,
                                                         int R, double *__restrict__ G) {
    int ti, tj;
    nv_tile_pair(blockIdx.x, ntile, ti, tj);
    const int npairs = gridDim.x;
    for (int e = threadIdx.x; e < kGramT * kGramT; e += 256) {
        double s = 0.0;
        for (int ch = 0; ch < nchunks; ++ch) s += part[((int64_t)ch * npairs + blockIdx.x) * (kGramT * kGramT) + e];
        const int a = ti * kGramT + e / kGramT, b = tj * kGramT + e % kGramT;
        if (a < R && b < R) {
            G[(int64_t)a * R + b] = s;
            G[(int64_t)b * R + a] = s;
        }
    }
}

// ---- launchers ----------------------------------------------------------------
hipError_t launch_ritz_vectors(const double *V, int64_t N, int64_t n, const double *vec, int64_t nroots, double *Y,
                               hipStream_t s) {
    return launch_ritz_vectors(V, N, n, vec, n, nroots, 0, n, Y, s);
}

hipError_t launch_ritz_vectors(const double *V, int64_t N, int64_t n, const double *vec, int64_t ldv, int64_t nroots,
                               int64_t slot0, int64_t nslots, double *Y, hipStream_t s) {
    if (n < 1 || nslots < n || slot0 < 0 || ldv < 1) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((N + 255) / 256);
    for (int64_t c0 = 0; c0 < nroots; c0 += kRitzChunk) {
        const int nc = (int)(nroots - c0 < kRitzChunk ? nroots - c0 : kRitzChunk);
        ritz_vectors_kernel<<<nb, 256, 0, s>>>(V, N, n, vec, ldv, c0, nc, slot0 % nslots, nslots, Y);
    }
    return hipGetLastError();
}

int topk_blocks(int64_t N) {
    const int64_t nb = (N + 255) / 256;
    return (int)(nb < 1024 ? nb : 1024);
}

hipError_t launch_topk(const double *Y, int64_t N, int nroots, int kmax, double *prev_m, int64_t *prev_i, double *pm,
                       int64_t *pi, double *out_m, int64_t *out_i, int32_t *out_ph, hipStream_t s) {
    const int nb = topk_blocks(N);
    topk_init_kernel<<<(nroots + 255) / 256, 256, 0, s>>>(prev_m, prev_i, nroots);
    for (int k = 0; k < kmax; ++k) {
        topk_partial_kernel<<<dim3(nb, nroots), 256, 0, s>>>(Y, N, prev_m, prev_i, pm, pi);
        topk_finish_kernel<<<nroots, 256, 0, s>>>(nb, pm, pi, Y, N, prev_m, prev_i, k, kmax, out_m, out_i, out_ph);
    }
    return hipGetLastError();
}

void gram_shape(int64_t D, int R, int &npairs, int &nchunks, int64_t &rows_per_chunk) {
    const int ntile = (R + kGramT - 1) / kGramT;
    npairs = ntile * (ntile + 1) / 2;
    int64_t want = (2048 + npairs - 1) / npairs;
    const int64_t maxc = (D + 255) / 256;
    if (want > maxc) want = maxc;
    if (want < 1) want = 1;
    rows_per_chunk = (D + want - 1) / want;
    nchunks = (int)((D + rows_per_chunk - 1) / rows_per_chunk);
}

hipError_t launch_gram(const double *A, int64_t D, int R, double *part, double *G, hipStream_t s) {
    int npairs, nchunks;
    int64_t rows;
    gram_shape(D, R, npairs, nchunks, rows);
    const int ntile = (R + kGramT - 1) / kGramT;
    gram_partial_kernel<<<dim3(npairs, nchunks), 256, 0, s>>>(A, D, R, ntile, rows, part);
    gram_finish_kernel<<<npairs, 256, 0, s>>>(part, nchunks, ntile, R, G);
    return hipGetLastError();
}

}  // namespace nv
