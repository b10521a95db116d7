// Reverse-mode derivative of the Lanczos recurrence and of the broadened
// spectrum (the adjoint of Source/top.f90:646-670, :820-842 and of
// convolve.py:28-48 applied to the Ritz sticks of Source/secondary.f90:252-256).
//
// Forward:  u_i = H q_i, alpha_i = q_i.u_i, r_i = u_i - alpha_i q_i - beta_{i-1} q_{i-1},
//           beta_i = |r_i|, q_{i+1} = r_i / beta_i.
// Reverse (i = n..1), with qbar_{i+1} complete:
//   rbar  = (qbar_{i+1} - (qbar_{i+1}.q_{i+1}) q_{i+1}) / beta_i + bbar_i q_{i+1}
//   atot  = abar_i - rbar.q_i,   ubar = rbar + atot q_i
//   qbar_i += -alpha_i rbar + atot u_i + H^T ubar,  u_i = beta_i q_{i+1} + alpha_i q_i + beta_{i-1} q_{i-1}
//   qbar_{i-1} += -beta_{i-1} rbar,  bbar_{i-1} -= rbar.q_{i-1}
//   dL/dtheta += ubar . (dH/dtheta) q_i
// Every dot of one step comes from one pass (adj_dots_kernel): rbar is a
// combination of qbar_{i+1} and q_{i+1}, so its dots follow from five base
// dots.  The parameter derivatives are bilinear forms
//   G_t[a][c] = sum_r ubar_a[r] (P_t q_c)[r]
// of every monomial P_t, accumulated over the iterations by
// grad_contract_kernel; the host maps them onto the nadvibs.in
// coefficients and the frequencies.
#include <math.h>

#include "nv_kernels.hpp"

namespace nv {

__device__ __forceinline__ double adj_block_sum(double v) {
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// qbn <- qbn + hub (when hub != null); partial sums of
// d0 = qbn.qn, d1 = qbn.qi, d2 = qbn.qp, d3 = qn.qi, d4 = qn.qp  (qn scaled by sn).
__global__ __launch_bounds__(256) void adj_dots_kernel(int64_t N, double *__restrict__ qbn,
                                                      const double *__restrict__ hub, const double *__restrict__ qn,
                                                      double sn, const double *__restrict__ qi,
                                                      const double *__restrict__ qp, double *__restrict__ part) {
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < N; e += (int64_t)gridDim.x * 256) {
        double b = qbn[e];
        if (hub) {
            b += hub[e];
            qbn[e] = b;
        }
        const double v = qn[e] * sn, x = qi[e], y = qp ? qp[e] : 0.0;
        d0 = fma(b, v, d0);
        d1 = fma(b, x, d1);
        d2 = fma(b, y, d2);
        d3 = fma(v, x, d3);
        d4 = fma(v, y, d4);
    }
    double d[5] = {d0, d1, d2, d3, d4};
    for (int k = 0; k < 5; ++k) {
        const double s = adj_block_sum(d[k]);
        if (threadIdx.x == 0) part[(int64_t)k * gridDim.x + blockIdx.x] = s;
    }
}

// One thread: the five dots, then the scalars of the update
// scal = {1/beta_i, bbar_i - d0/beta_i, atot}; bbar_{i-1} -= rbar.q_{i-1}.
__global__ __launch_bounds__(256) void adj_scalars_kernel(int nb, const double *__restrict__ part, int64_t i,
                                                         const double *__restrict__ beta,
                                                         const double *__restrict__ abar, double *__restrict__ bbar,
                                                         double *__restrict__ scal) {
    __shared__ double dots[5];
    for (int k = 0; k < 5; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) s += part[(int64_t)k * nb + b];
        s = adj_block_sum(s);
        if (threadIdx.x == 0) dots[k] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double bi = beta[i], bb = bbar[i];
    const double c1 = 1.0 / bi, c2 = bb - dots[0] / bi;
    const double rq = c1 * dots[1] + c2 * dots[3];   // rbar . q_i
    const double rp = c1 * dots[2] + c2 * dots[4];   // rbar . q_{i-1}
    const double atot = abar[i - 1] - rq;
    if (i > 1) bbar[i - 1] -= rp;
    scal[0] = c1;
    scal[1] = c2;
    scal[2] = atot;
}

__global__ __launch_bounds__(256) void adj_update_kernel(int64_t N, const double *__restrict__ qbn,
                                                        const double *__restrict__ qn, double sn,
                                                        const double *__restrict__ qi, const double *__restrict__ qp,
                                                        const double *__restrict__ alpha,
                                                        const double *__restrict__ beta, int64_t i,
                                                        const double *__restrict__ scal, double *__restrict__ ub,
                                                        double *__restrict__ qbi, double *__restrict__ qbp) {
    const double c1 = scal[0], c2 = scal[1], at = scal[2];
    const double ai = alpha[i], bi = beta[i], bp = i > 1 ? beta[i - 1] : 0.0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < N; e += (int64_t)gridDim.x * 256) {
        const double v = qn[e] * sn, x = qi[e], y = qp ? qp[e] : 0.0;
        const double rb = c1 * qbn[e] + c2 * v;
        ub[e] = rb + at * x;
        const double u = bi * v + ai * x + bp * y;
        qbi[e] += -ai * rb + at * u;
        if (qp) qbp[e] += -bp * rb;
    }
}

__global__ __launch_bounds__(256) void adj_zero_kernel(int64_t N, double *__restrict__ v) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < N; e += (int64_t)gridDim.x * 256) v[e] = 0.0;
}

// ---- bilinear forms of the monomials ------------------------------------------
// blockIdx.y selects the group: [0, nterms) Delta != 0 part of term t,
// [nterms, nterms+ndterms) Delta = 0 part of diagonal term t,
// [.., +M) harmonic energy of mode k (entry [0][0] only), last: identity.
template <int NS>
__global__ __launch_bounds__(256) void grad_contract_kernel(NvOpDev g, const double *__restrict__ ub,
                                                           const double *__restrict__ q, double *__restrict__ part) {
    extern __shared__ __align__(16) double lds[];
    double *ftab = lds;                                   // ncodes * nmax
    int *shft = reinterpret_cast<int *>(ftab + g.ncodes * g.nmax);
    int *nf = shft + NV_MAXM;
    for (int i = threadIdx.x; i < g.ncodes * g.nmax; i += 256) ftab[i] = g.ftab64[i];
    if (threadIdx.x == 0) {
        int s = 1;
        for (int k = g.M - 1; k >= 0; --k) {
            shft[k] = s;
            nf[k] = g.nfunc[k];
            s *= g.nfunc[k];
        }
    }
    __syncthreads();
    const int grp = blockIdx.y;
    const int D = g.D;
    double acc[NS * NS];
#pragma unroll
    for (int e = 0; e < NS * NS; ++e) acc[e] = 0.0;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < D; r += gridDim.x * 256) {
        double u[NS];
#pragma unroll
        for (int a = 0; a < NS; ++a) u[a] = ub[(int64_t)a * D + r];
        if (grp < g.nterms) {
            const NvTerm tm = g.terms[grp];
            double w[NS];
#pragma unroll
            for (int c = 0; c < NS; ++c) w[c] = 0.0;
            for (int k = 0; k < tm.nsteps; ++k) {
                const NvStep st = g.steps[tm.first + k];
                double h = 1.0;
#pragma unroll
                for (int j = 0; j < NV_MAXF; ++j)
                    if (j < st.nf) h *= ftab[st.code[j] * g.nmax + (r / shft[st.mode[j]]) % nf[st.mode[j]]];
                if (h != 0.0) {
#pragma unroll
                    for (int c = 0; c < NS; ++c) w[c] = fma(h, q[(int64_t)c * D + r + st.stride], w[c]);
                }
            }
#pragma unroll
            for (int a = 0; a < NS; ++a)
#pragma unroll
                for (int c = 0; c < NS; ++c) acc[a * NS + c] = fma(u[a], w[c], acc[a * NS + c]);
        } else if (grp < g.nterms + g.ndterms) {
            const NvDiagTerm dt = g.dterms[grp - g.nterms];
            double h = 1.0;
#pragma unroll
            for (int j = 0; j < NV_MAXF; ++j)
                if (j < dt.nf) h *= ftab[dt.code[j] * g.nmax + (r / shft[dt.mode[j]]) % nf[dt.mode[j]]];
#pragma unroll
            for (int a = 0; a < NS; ++a)
#pragma unroll
                for (int c = 0; c < NS; ++c) acc[a * NS + c] = fma(u[a] * h, q[(int64_t)c * D + r], acc[a * NS + c]);
        } else if (grp < g.nterms + g.ndterms + g.M) {
            const int k = grp - g.nterms - g.ndterms;
            const double e = (double)((r / shft[k]) % nf[k]) + 0.5;
            double p = 0.0;
#pragma unroll
            for (int a = 0; a < NS; ++a) p = fma(u[a], q[(int64_t)a * D + r], p);
            acc[0] = fma(e, p, acc[0]);
        } else {
#pragma unroll
            for (int a = 0; a < NS; ++a)
#pragma unroll
                for (int c = 0; c < NS; ++c) acc[a * NS + c] = fma(u[a], q[(int64_t)c * D + r], acc[a * NS + c]);
        }
    }
#pragma unroll
    for (int e = 0; e < NS * NS; ++e) {
        const double s = adj_block_sum(acc[e]);
        if (threadIdx.x == 0) part[((int64_t)grp * gridDim.x + blockIdx.x) * NS * NS + e] = s;
    }
}

// gacc[grp][e] += sum over the chunks of part.
__global__ void grad_finish_kernel(int nchunks, int nss, const double *__restrict__ part, double *__restrict__ gacc) {
    const int grp = blockIdx.x;
    for (int e = threadIdx.x; e < nss; e += blockDim.x) {
        double s = 0.0;
        for (int x = 0; x < nchunks; ++x) s += part[((int64_t)grp * nchunks + x) * nss + e];
        gacc[(int64_t)grp * nss + e] += s;
    }
}

// H^T - H: only the inter-state vibrationally diagonal part can be
// unsymmetric (the lower-only quirk masks D_ss'(n) above the state diagonal).
// hub[s] += sum_{s' != s} (D_s's(n) - D_ss'(n)) ub[s'], D = dconst + sum_t h_t(n) dcmat_t.
template <int NS>
__global__ __launch_bounds__(256) void adj_transpose_kernel(NvOpDev op, const double *__restrict__ ub,
                                                           double *__restrict__ hub) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= op.D) return;
    double Dm[NS * NS];
#pragma unroll
    for (int e = 0; e < NS * NS; ++e) Dm[e] = op.dconst[e];
    for (int t = 0; t < op.ndterms; ++t) {
        const NvDiagTerm dt = op.dterms[t];
        double h = 1.0;
        for (int j = 0; j < dt.nf; ++j) {
            int sh = 1;
            for (int k = op.M - 1; k > dt.mode[j]; --k) sh *= op.nfunc[k];
            h *= op.ftab64[dt.code[j] * op.nmax + (r / sh) % op.nfunc[dt.mode[j]]];
        }
#pragma unroll
        for (int e = 0; e < NS * NS; ++e) Dm[e] = fma(h, op.dcmat[(size_t)t * NS * NS + e], Dm[e]);
    }
    double u[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) u[s] = ub[(int64_t)s * op.D + r];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double a = 0.0;
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2)
            if (s2 != s) a = fma(Dm[s2 * NS + s] - Dm[s * NS + s2], u[s2], a);
        hub[(int64_t)s * op.D + r] += a;
    }
}

// ---- launchers ----------------------------------------------------------------
hipError_t launch_adj_transpose(const NvOpDev &op, const double *ub, double *hub, hipStream_t s) {
    const unsigned g = (unsigned)((op.D + 255) / 256);
    switch (op.ns) {
#define NV_T(n) case n: adj_transpose_kernel<n><<<g, 256, 0, s>>>(op, ub, hub); break;
        NV_T(1) NV_T(2) NV_T(3) NV_T(4) NV_T(5) NV_T(6) NV_T(7) NV_T(8) NV_T(9) NV_T(10)
#undef NV_T
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int adj_blocks(int64_t N) {
    const int64_t nb = (N + 255) / 256;
    return (int)(nb < 2048 ? nb : 2048);
}

hipError_t launch_adj_step(int64_t N, int64_t i, double *qbn, const double *hub, const double *qn, double sn,
                           const double *qi, const double *qp, const double *alpha, const double *beta,
                           const double *abar, double *bbar, double *part, double *scal, double *ub, double *qbi,
                           double *qbp, hipStream_t s) {
    const int nb = adj_blocks(N);
    adj_dots_kernel<<<nb, 256, 0, s>>>(N, qbn, hub, qn, sn, qi, qp, part);
    adj_scalars_kernel<<<1, 256, 0, s>>>(nb, part, i, beta, abar, bbar, scal);
    adj_update_kernel<<<nb, 256, 0, s>>>(N, qbn, qn, sn, qi, qp, alpha, beta, i, scal, ub, qbi, qbp);
    return hipGetLastError();
}

hipError_t launch_adj_zero(int64_t N, double *v, hipStream_t s) {
    adj_zero_kernel<<<adj_blocks(N), 256, 0, s>>>(N, v);
    return hipGetLastError();
}

int grad_groups(const NvOpDev &g) { return g.nterms + g.ndterms + g.M + 1; }

int grad_chunks(const NvOpDev &g) {
    const int groups = grad_groups(g);
    int ch = (4096 + groups - 1) / groups;
    const int maxc = (g.D + 255) / 256;
    if (ch > maxc) ch = maxc;
    return ch < 1 ? 1 : ch;
}

size_t grad_lds_bytes(const NvOpDev &g) {
    return (size_t)g.ncodes * g.nmax * sizeof(double) + 2 * NV_MAXM * sizeof(int);
}

template <int NS>
static hipError_t contract_ns(const NvOpDev &g, const double *ub, const double *q, double *part, double *gacc,
                              hipStream_t s) {
    const int groups = grad_groups(g), ch = grad_chunks(g);
    grad_contract_kernel<NS><<<dim3(ch, groups), 256, grad_lds_bytes(g), s>>>(g, ub, q, part);
    grad_finish_kernel<<<groups, 64, 0, s>>>(ch, NS * NS, part, gacc);
    return hipGetLastError();
}

hipError_t launch_grad_contract(const NvOpDev &g, const double *ub, const double *q, double *part, double *gacc,
                                hipStream_t s) {
    switch (g.ns) {
        case 1: return contract_ns<1>(g, ub, q, part, gacc, s);
        case 2: return contract_ns<2>(g, ub, q, part, gacc, s);
        case 3: return contract_ns<3>(g, ub, q, part, gacc, s);
        case 4: return contract_ns<4>(g, ub, q, part, gacc, s);
        case 5: return contract_ns<5>(g, ub, q, part, gacc, s);
        case 6: return contract_ns<6>(g, ub, q, part, gacc, s);
        case 7: return contract_ns<7>(g, ub, q, part, gacc, s);
        case 8: return contract_ns<8>(g, ub, q, part, gacc, s);
        case 9: return contract_ns<9>(g, ub, q, part, gacc, s);
        case 10: return contract_ns<10>(g, ub, q, part, gacc, s);
        default: return hipErrorInvalidValue;
    }
}

// ---- broadened spectrum of the tridiagonal eigenpairs -------------------------
// env[g] = sum_i a_i^2 G(grid[g] - (theta_i - ref) c),  a_i = vec[i] (first
// components), G(u) = exp(-u^2 / (2 sigma^2)).
__global__ __launch_bounds__(256) void spec_env_kernel(int64_t n, const double *__restrict__ theta,
                                                      const double *__restrict__ vec, double ref, double c,
                                                      double sigma, int64_t ng, const double *__restrict__ grid,
                                                      double *__restrict__ env) {
    const int64_t gi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (gi >= ng) return;
    const double x = grid[gi];
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double a = vec[i];
        const double u = (x - (theta[i] - ref) * c) / sigma;
        s = fma(a * a, exp(-0.5 * u * u), s);
    }
    env[gi] = s;
}

// W[i] = sum_g gbar_g G(u_gi), Wp[i] = sum_g gbar_g u_gi G(u_gi) c / sigma: one block per root.
__global__ __launch_bounds__(256) void spec_w_kernel(const double *__restrict__ theta, double ref, double c,
                                                    double sigma, int64_t ng, const double *__restrict__ grid,
                                                    const double *__restrict__ gbar, double *__restrict__ W,
                                                    double *__restrict__ Wp) {
    const int64_t i = blockIdx.x;
    const double xi = (theta[i] - ref) * c;
    double w = 0.0, wp = 0.0;
    for (int64_t g = threadIdx.x; g < ng; g += 256) {
        const double u = (grid[g] - xi) / sigma;
        const double G = exp(-0.5 * u * u) * gbar[g];
        w += G;
        wp = fma(u, G, wp);
    }
    w = adj_block_sum(w);
    wp = adj_block_sum(wp);
    if (threadIdx.x == 0) {
        W[i] = w;
        Wp[i] = wp * c / sigma;
    }
}

// M[i][k] = a_i a_k D_ik, D the divided differences of W (derivative average
// for |theta_i - theta_k| <= tol).  M[0][0] gains ref0 (the dependence of the
// origin on theta_1).
__global__ __launch_bounds__(256) void spec_m_kernel(int64_t n, const double *__restrict__ theta,
                                                    const double *__restrict__ vec, const double *__restrict__ W,
                                                    const double *__restrict__ Wp, double tol, double ref0,
                                                    double *__restrict__ M) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n * n) return;
    const int64_t i = t / n, k = t % n;
    const double d = theta[i] - theta[k];
    const double Dik = fabs(d) <= tol ? 0.5 * (Wp[i] + Wp[k]) : (W[i] - W[k]) / d;
    double m = vec[i] * vec[k] * Dik;
    if (t == 0) m += ref0;
    M[t] = m;
}

// ref0 = -sum_i a_i^2 Wp_i (one block).
__global__ __launch_bounds__(256) void spec_ref_kernel(int64_t n, const double *__restrict__ vec,
                                                      const double *__restrict__ Wp, double *__restrict__ out) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s = fma(vec[i] * vec[i], Wp[i], s);
    s = adj_block_sum(s);
    if (threadIdx.x == 0) *out = -s;
}

// P = S M for row-major n x n S (S[j][i] = vec[j*n + i]) and M; 32x32 tiles.
__global__ __launch_bounds__(256) void spec_gemm_kernel(int64_t n, const double *__restrict__ S,
                                                       const double *__restrict__ M, double *__restrict__ P) {
    __shared__ double As[32][33];
    __shared__ double Bs[32][33];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t row0 = (int64_t)blockIdx.y * 32, col0 = (int64_t)blockIdx.x * 32;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int64_t k0 = 0; k0 < n; k0 += 32) {
        __syncthreads();
        for (int t = threadIdx.x; t < 1024; t += 256) {
            const int r = t / 32, cc = t % 32;
            As[r][cc] = (row0 + r < n && k0 + cc < n) ? S[(row0 + r) * n + k0 + cc] : 0.0;
            Bs[r][cc] = (k0 + r < n && col0 + cc < n) ? M[(k0 + r) * n + col0 + cc] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < 32; ++kk) {
            const double a0 = As[ty * 2][kk], a1 = As[ty * 2 + 1][kk];
            const double b0 = Bs[kk][tx * 2], b1 = Bs[kk][tx * 2 + 1];
            acc[0][0] = fma(a0, b0, acc[0][0]);
            acc[0][1] = fma(a0, b1, acc[0][1]);
            acc[1][0] = fma(a1, b0, acc[1][0]);
            acc[1][1] = fma(a1, b1, acc[1][1]);
        }
    }
    for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v) {
            const int64_t r = row0 + ty * 2 + u, cc = col0 + tx * 2 + v;
            if (r < n && cc < n) P[r * n + cc] = acc[u][v];
        }
}

// Tb = P S^T (the full dL/dT), 32x32 tiles.
__global__ __launch_bounds__(256) void spec_gemm_nt_kernel(int64_t n, const double *__restrict__ P,
                                                          const double *__restrict__ S, double *__restrict__ Tb) {
    __shared__ double As[32][33];
    __shared__ double Bs[32][33];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t row0 = (int64_t)blockIdx.y * 32, col0 = (int64_t)blockIdx.x * 32;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int64_t k0 = 0; k0 < n; k0 += 32) {
        __syncthreads();
        for (int t = threadIdx.x; t < 1024; t += 256) {
            const int r = t / 32, cc = t % 32;
            As[r][cc] = (row0 + r < n && k0 + cc < n) ? P[(row0 + r) * n + k0 + cc] : 0.0;
            Bs[r][cc] = (col0 + r < n && k0 + cc < n) ? S[(col0 + r) * n + k0 + cc] : 0.0;  // Bs[col][k]
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < 32; ++kk) {
            const double a0 = As[ty * 2][kk], a1 = As[ty * 2 + 1][kk];
            const double b0 = Bs[tx * 2][kk], b1 = Bs[tx * 2 + 1][kk];
            acc[0][0] = fma(a0, b0, acc[0][0]);
            acc[0][1] = fma(a0, b1, acc[0][1]);
            acc[1][0] = fma(a1, b0, acc[1][0]);
            acc[1][1] = fma(a1, b1, acc[1][1]);
        }
    }
    for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v) {
            const int64_t r = row0 + ty * 2 + u, cc = col0 + tx * 2 + v;
            if (r < n && cc < n) Tb[r * n + cc] = acc[u][v];
        }
}

// abar[j] = sum_k P[j][k] S[j][k]; bbar[j+1] = 2 sum_k P[j][k] S[j+1][k]: one block per row j.
__global__ __launch_bounds__(256) void spec_rows_kernel(int64_t n, const double *__restrict__ S,
                                                       const double *__restrict__ P, double *__restrict__ abar,
                                                       double *__restrict__ bbar) {
    const int64_t j = blockIdx.x;
    double d = 0.0, o = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += 256) {
        const double p = P[j * n + k];
        d = fma(p, S[j * n + k], d);
        if (j + 1 < n) o = fma(p, S[(j + 1) * n + k], o);
    }
    d = adj_block_sum(d);
    o = adj_block_sum(o);
    if (threadIdx.x == 0) {
        abar[j] = d;
        if (j + 1 < n) bbar[j + 1] = 2.0 * o;
    }
}

hipError_t launch_spec_env(int64_t n, const double *theta, const double *vec, double ref, double c, double sigma,
                           int64_t ng, const double *grid, double *env, hipStream_t s) {
    spec_env_kernel<<<(unsigned)((ng + 255) / 256), 256, 0, s>>>(n, theta, vec, ref, c, sigma, ng, grid, env);
    return hipGetLastError();
}

hipError_t launch_spec_adjoint(int64_t n, const double *theta, const double *vec, double ref, int lowest, double c,
                               double sigma, int64_t ng, const double *grid, const double *gbar, double *work,
                               double *abar, double *bbar, double *tbar, hipStream_t s) {
    double *W = work, *Wp = work + n, *ref0 = work + 2 * n, *M = work + 2 * n + 8, *P = M + n * n;
    spec_w_kernel<<<(unsigned)n, 256, 0, s>>>(theta, ref, c, sigma, ng, grid, gbar, W, Wp);
    if (lowest) spec_ref_kernel<<<1, 256, 0, s>>>(n, vec, Wp, ref0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    double r0 = 0.0;
    if (lowest) {
        e = hipMemcpyAsync(&r0, ref0, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
    }
    const double tol = 1e-7 * sigma / c;
    spec_m_kernel<<<(unsigned)((n * n + 255) / 256), 256, 0, s>>>(n, theta, vec, W, Wp, tol, r0, M);
    const unsigned nt = (unsigned)((n + 31) / 32);
    spec_gemm_kernel<<<dim3(nt, nt), 256, 0, s>>>(n, vec, M, P);
    spec_rows_kernel<<<(unsigned)n, 256, 0, s>>>(n, vec, P, abar, bbar);
    if (tbar) spec_gemm_nt_kernel<<<dim3(nt, nt), 256, 0, s>>>(n, P, vec, tbar);
    return hipGetLastError();
}

}  // namespace nv
