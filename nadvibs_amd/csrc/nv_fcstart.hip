// Franck-Condon start vector (generate_initialvec with shiftref != 0,
// Source/secondary.f90:351-809).
//
// The overlaps <m|n> of the initial-state vibrational function |m> (ground
// state, or one mode excited) with the final-state basis functions |n> follow
// from the recursion of Source/secondary.f90:487-570 (in n, level by level in
// total quanta) and :602-689 (in m for one excited mode), scaled at the end by
// G / sqrt(prod 2^n n! 2^m m!) (:694-725).  The small matrices (nmodes^2) are
// formed on the host in FP64; the recursions run on the device, one launch per
// level, every basis function of the level in parallel.  A function of level
// L+1 is reached from level L through its first excited mode.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "nv_pro.hpp"

namespace nv {

struct FcDev {
    int32_t M, D;
    const int32_t *nfunc;   // [M]
    const int64_t *shft;    // [M]
    const int32_t *irrep;   // [M] irreducible representation of each mode
    const double *A1;       // [M][M] (diagonal already reduced by 1)
    const double *A2;       // [M][M]
    const double *A2d;      // [M]
};

__device__ __forceinline__ void fc_digits(const FcDev &f, int r, int *n) {
    unsigned rr = (unsigned)r;
    for (int k = f.M - 1; k >= 0; --k) {
        const unsigned nf = (unsigned)f.nfunc[k];
        n[k] = (int)(rr % nf);
        rr /= nf;
    }
}

// C(n) for the basis functions of total quanta `level`:
// C(n) = 2 sum_{k in irrep(j)} A1[j][k] n'_k C(n' - e_k) - A2d[j] C(n'), n' = n - e_j
__global__ __launch_bounds__(256) void fc_level_kernel(FcDev f, int level, double *C) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= f.D) return;
    int n[NV_MAXM];
    fc_digits(f, r, n);
    int L = 0, j = -1;
    for (int k = 0; k < f.M; ++k) {
        L += n[k];
        if (j < 0 && n[k] > 0) j = k;
    }
    if (L != level || j < 0) return;
    const int64_t rp = r - f.shft[j];  // n' = n - e_j
    double s = 0.0;
    for (int k = 0; k < f.M; ++k) {
        if (f.irrep[k] != f.irrep[j]) continue;
        const int nk = n[k] - (k == j ? 1 : 0);
        if (nk > 0) s += f.A1[j * f.M + k] * (double)nk * C[rp - f.shft[k]];
    }
    C[r] = 2.0 * s - f.A2d[j] * C[rp];
}

// One step of the recursion in m (Source/secondary.f90:602-689):
// Cn(m+1) = 2 m (A3_ii - 1) C(m-1) - A3d_i C(m) + 2 sum_k A2[k][i] n_k C(m, n - e_k)
__global__ __launch_bounds__(256) void fc_mstep_kernel(FcDev f, int imode, double twoma, double a3d, const double *Cm1,
                                                       const double *Cm, double *Cp) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= f.D) return;
    int n[NV_MAXM];
    fc_digits(f, r, n);
    double v = twoma * Cm1[r] - a3d * Cm[r];
    for (int k = 0; k < f.M; ++k)
        if (n[k] > 0) v += 2.0 * f.A2[k * f.M + imode] * (double)n[k] * Cm[r - f.shft[k]];
    Cp[r] = v;
}

// C(n) * G / sqrt(prod_j 2^n_j n_j! * 2^m_j m_j!), sum of squares per block.
__global__ __launch_bounds__(256) void fc_scale_kernel(FcDev f, double G, double mfac, double *C, double *part) {
    __shared__ double red[4];
    const int r = blockIdx.x * 256 + threadIdx.x;
    double sq = 0.0;
    if (r < f.D) {
        int n[NV_MAXM];
        fc_digits(f, r, n);
        double d = mfac;
        for (int k = 0; k < f.M; ++k) {
            double fk = 1.0;
            for (int t = 2; t <= n[k]; ++t) fk *= t;
            d *= exp2((double)n[k]) * fk;
        }
        const double v = G * C[r] / sqrt(d);
        C[r] = v;
        sq = v * v;
    }
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// dst[s][r] = sc[s] * C[r]
__global__ __launch_bounds__(256) void fc_spread_kernel(const double *C, int64_t D, int ns, const double *sc,
                                                        double *dst) {
    const int64_t N = D * ns;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (int64_t)gridDim.x * 256) {
        const int64_t s = i / D;
        dst[i] = sc[s] * C[i - s * D];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct FcHost {
    std::vector<double> A1, A2, A3, A2d, A3d;
    double G = 1.0, detT = 1.0, detA = 1.0;
    std::vector<int32_t> irrep;
};

// |det T| by Gaussian elimination with partial pivoting (the reference's
// gauss_elim, Source/secondary.f90:311-348, then the product of |pivots|).
static double abs_det(std::vector<double> a, int M) {
    double det = 1.0;
    for (int i = 0; i < M; ++i) {
        int p = i;
        for (int j = i + 1; j < M; ++j)
            if (std::fabs(a[j * M + i]) > std::fabs(a[p * M + i])) p = j;
        if (p != i)
            for (int k = 0; k < M; ++k) std::swap(a[i * M + k], a[p * M + k]);
        for (int j = i + 1; j < M; ++j) {
            if (a[j * M + i] == 0.0) continue;
            const double v = -a[j * M + i] / a[i * M + i];
            for (int k = 0; k < M; ++k) a[j * M + k] += v * a[i * M + k];
        }
        det *= std::fabs(a[i * M + i]);
    }
    return det;
}

// A^{-1} and det A of a symmetric positive definite A by Cholesky.  (The
// reference diagonalises A with GIVENS and forms V diag(1/lambda) V^T with
// single-precision accumulators in EBC/EBCT, Source/yarkony.f90:23-42,69-87;
// that rounding, ~1e-7 relative, is not reproduced.)
static bool spd_inverse(const std::vector<double> &A, int M, std::vector<double> &inv, double &det) {
    std::vector<double> L(M * M, 0.0);
    det = 1.0;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = A[i * M + j];
            for (int k = 0; k < j; ++k) s -= L[i * M + k] * L[j * M + k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[i * M + i] = std::sqrt(s);
                det *= s;
            } else {
                L[i * M + j] = s / L[j * M + j];
            }
        }
    inv.assign(M * M, 0.0);
    for (int c = 0; c < M; ++c) {
        std::vector<double> y(M, 0.0), x(M, 0.0);
        for (int i = 0; i < M; ++i) {
            double s = i == c ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= L[i * M + k] * y[k];
            y[i] = s / L[i * M + i];
        }
        for (int i = M - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < M; ++k) s -= L[k * M + i] * x[k];
            x[i] = s / L[i * M + i];
        }
        for (int i = 0; i < M; ++i) inv[i * M + c] = x[i];
    }
    return true;
}

// Matrices of Source/secondary.f90:402-469 and the prefactor G of :694-707.
static bool fc_setup(const Operator &op, FcHost &h, std::string &err) {
    const int M = (int)op.M;
    const std::vector<double> &a = op.omega, &b = op.bomega, &d = op.dvec;
    // tmatrix(j,i) = tmat((i-1)*M + j): T[j][i]
    std::vector<double> T(M * M);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) T[j * M + i] = op.tmat[i * M + j];
    h.detT = abs_det(T, M);
    std::vector<double> A(M * M, 0.0);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            double s = 0.0;
            for (int k = 0; k < M; ++k) s += 0.5 * b[k] * T[k * M + i] * T[k * M + j];
            if (i == j) s += 0.5 * a[i];
            A[i * M + j] = s;
        }
    std::vector<double> Ainv;
    if (!spd_inverse(A, M, Ainv, h.detA)) {
        err = "Franck-Condon A matrix is not positive definite";
        return false;
    }
    h.A1.assign(M * M, 0.0);
    h.A2.assign(M * M, 0.0);
    h.A3.assign(M * M, 0.0);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            h.A1[i * M + j] = Ainv[i * M + j] * std::sqrt(a[i]) * std::sqrt(a[j]);
            double s2 = 0.0, s3 = 0.0;
            for (int k = 0; k < M; ++k) {
                s2 += Ainv[i * M + k] * T[j * M + k];
                for (int l = 0; l < M; ++l) s3 += T[i * M + k] * Ainv[k * M + l] * T[j * M + l];
            }
            h.A2[i * M + j] = s2 * std::sqrt(a[i]) * std::sqrt(b[j]);
            h.A3[i * M + j] = s3 * std::sqrt(b[i]) * std::sqrt(b[j]);
        }
    h.A2d.assign(M, 0.0);
    h.A3d.assign(M, 0.0);
    for (int i = 0; i < M; ++i) {
        h.A1[i * M + i] -= 1.0;
        for (int j = 0; j < M; ++j) {
            double d1 = h.A3[j * M + i];
            if (j == i) d1 -= 2.0;
            h.A2d[i] += h.A2[i * M + j] * d[j] * std::sqrt(b[j]);
            h.A3d[i] += d1 * d[j] * std::sqrt(b[j]);
        }
    }
    double e = 0.0, pab = 1.0;
    for (int i = 0; i < M; ++i) e -= 0.5 * b[i] * d[i] * d[i];
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) e += 0.25 * std::sqrt(b[i]) * d[i] * h.A3[i * M + j] * d[j] * std::sqrt(b[j]);
    for (int i = 0; i < M; ++i) pab *= b[i] * a[i];
    h.G = std::sqrt(std::sqrt(pab) * h.detT / h.detA) * std::exp(e);
    // irreducible representation of every mode (npirrep consecutive modes each)
    h.irrep.assign(M, 0);
    int m0 = 0;
    for (size_t g = 0; g < op.npirrep.size(); ++g)
        for (int64_t q = 0; q < op.npirrep[g]; ++q) h.irrep[m0++] = (int32_t)g;
    return true;
}

hipError_t build_fc_start(const Operator &op, const NvOpDev &dev, double *dst, hipStream_t s, std::string &err) {
    FcHost h;
    if (!fc_setup(op, h, err)) return hipErrorInvalidValue;
    const int M = (int)op.M;
    const int64_t D = op.D;
    std::vector<void *> tmp;
    auto up = [&](const void *src, size_t bytes, void **out) -> hipError_t {
        hipError_t e = hipMalloc(out, bytes);
        if (e != hipSuccess) return e;
        tmp.push_back(*out);
        return hipMemcpyAsync(*out, src, bytes, hipMemcpyHostToDevice, s);
    };
    FcDev f;
    f.M = M;
    f.D = (int32_t)D;
    f.nfunc = dev.nfunc;
    std::vector<int64_t> shft(op.shft.begin(), op.shft.end());
    hipError_t e = up(shft.data(), M * sizeof(int64_t), (void **)&f.shft);
    if (e == hipSuccess) e = up(h.irrep.data(), M * sizeof(int32_t), (void **)&f.irrep);
    if (e == hipSuccess) e = up(h.A1.data(), M * M * sizeof(double), (void **)&f.A1);
    if (e == hipSuccess) e = up(h.A2.data(), M * M * sizeof(double), (void **)&f.A2);
    if (e == hipSuccess) e = up(h.A2d.data(), M * sizeof(double), (void **)&f.A2d);
    double *C[3] = {nullptr, nullptr, nullptr}, *part = nullptr, *sc = nullptr;
    const unsigned g = (unsigned)((D + 255) / 256);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
        e = hipMalloc((void **)&C[k], D * sizeof(double));
        if (e == hipSuccess) tmp.push_back(C[k]);
    }
    if (e == hipSuccess) e = hipMalloc((void **)&part, g * sizeof(double));
    if (e == hipSuccess) tmp.push_back(part);
    if (e == hipSuccess) e = hipMalloc((void **)&sc, op.ns * sizeof(double));
    if (e == hipSuccess) tmp.push_back(sc);
    // C(0) = 1; every other entry is written before it is read
    const double one = 1.0;
    if (e == hipSuccess) e = hipMemcpyAsync(C[0], &one, sizeof(double), hipMemcpyHostToDevice, s);
    int nlev = 0;
    for (int k = 0; k < M; ++k) nlev += (int)op.nfunc[k] - 1;
    for (int L = 1; L <= nlev && e == hipSuccess; ++L) {
        hipLaunchKernelGGL(fc_level_kernel, dim3(g), dim3(256), 0, s, f, L, C[0]);
        e = hipGetLastError();
    }
    // one excited mode of the initial state: recursion in m
    int imode = 0;
    while (imode < M - 1 && op.istate[imode] <= 0) ++imode;
    const int ninit = (int)op.istate[imode];
    int cur = 0;  // buffer holding C(m)
    if (ninit > 0 && e == hipSuccess) {
        int prev = 1, next = 2;
        e = hipMemsetAsync(C[prev], 0, D * sizeof(double), s);
        for (int i = 1; i <= ninit && e == hipSuccess; ++i) {
            const double twoma = 2.0 * (i - 1) * (h.A3[imode * M + imode] - 1.0);
            hipLaunchKernelGGL(fc_mstep_kernel, dim3(g), dim3(256), 0, s, f, imode, twoma, h.A3d[imode], C[prev],
                               C[cur], C[next]);
            e = hipGetLastError();
            const int t = prev;
            prev = cur;
            cur = next;
            next = t;
        }
    }
    double mfac = 1.0;
    for (int k = 0; k < M; ++k) {
        double fk = 1.0;
        for (int t = 2; t <= op.istate[k]; ++t) fk *= t;
        mfac *= std::exp2((double)op.istate[k]) * fk;
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fc_scale_kernel, dim3(g), dim3(256), 0, s, f, h.G, mfac, C[cur], part);
        e = hipGetLastError();
    }
    std::vector<double> hp(g);
    if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), part, g * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) {
        // weights and normalisation as in Source/secondary.f90:727-783: every
        // state holds C, divided by sqrt(q.q/nstates) = |C|, weighted by
        // statew/sum(statew), then the whole vector normalised
        double cc = 0.0;
        for (double v : hp) cc += v;
        const double d1 = std::sqrt(cc);
        double W = 0.0;
        for (int64_t q = 0; q < op.ns; ++q) W += op.statew[q];
        if (W == 0.0) W = 1.0;
        std::vector<double> w(op.ns);
        double tot = 0.0;
        for (int64_t q = 0; q < op.ns; ++q) {
            w[q] = op.statew[q] / (W * d1);
            tot += w[q] * w[q] * cc;
        }
        for (int64_t q = 0; q < op.ns; ++q) w[q] /= std::sqrt(tot);
        e = hipMemcpyAsync(sc, w.data(), op.ns * sizeof(double), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            const int gs = (int)std::min<int64_t>((D * op.ns + 255) / 256, 8192);
            hipLaunchKernelGGL(fc_spread_kernel, dim3(gs), dim3(256), 0, s, C[cur], D, (int)op.ns, sc, dst);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    for (void *p : tmp) hipFree(p);
    if (e != hipSuccess && err.empty()) err = std::string("Franck-Condon start: ") + hipGetErrorString(e);
    return e;
}

}  // namespace nv
