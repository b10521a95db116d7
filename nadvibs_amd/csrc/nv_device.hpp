// Structures shared by the host operator builder and the HIP kernels.
//
// The operator H = E(n) + D(n) + sum_t C_t (x) prod_k Q_k^{p_k} is stored as
//   * a diagonal part: harmonic energy E(n) = sum_k (n_k + 1/2) w_k, a constant
//     state matrix (concoef) and "diagonal terms" (all-even monomials at
//     Delta = 0) -- Source/top.f90:712-756;
//   * off-diagonal terms: for each stored monomial t an nstates x nstates
//     coupling matrix C_t and a list of steps Delta != 0, each a flat-index
//     stride plus the per-mode factors h(p_k, min(n_k, n_k+Delta_k), |Delta_k|)
//     -- Source/top.f90:758-808.
// A factor is addressed by a "code" (mode, power, Delta); ftab[code][n] holds
// the factor for quantum n of that mode, 0 where n + Delta leaves the basis.
#pragma once
#include <stdint.h>

#define NV_MAXF 4          // factors per step (ordr <= 4)
#define NV_MAXM 64         // modes
#define NV_MAXS 10         // states (basis.in weights has 10 slots)

// The FP32/BF16 copy of a Lanczos vector that H*v reads is state-interleaved,
// x_lo[r * nv_lo_stride(ns) + s], padded so that the states of one row are
// one 4/8/12/16-byte load (padding lanes hold 0; three states take 12 bytes).
#ifdef __HIPCC__
__host__ __device__
#endif
inline constexpr int nv_lo_stride(int ns) { return ns <= 3 ? ns : ns <= 4 ? 4 : ns <= 8 ? 8 : 16; }

struct NvStep {
    int32_t stride;        // flat-index offset of the column, sum Delta_k * shft_k
    int32_t nf;            // number of factors
    int32_t mode[NV_MAXF]; // 0-based mode of each factor
    int32_t code[NV_MAXF]; // ftab row of each factor
};

struct NvTerm {
    int32_t first;         // first step in the step list
    int32_t nsteps;        // number of steps (Delta != 0)
};

struct NvDiagTerm {
    int32_t nf;
    int32_t mode[NV_MAXF];
    int32_t code[NV_MAXF];
};

// Everything the H*v kernels read, passed by value.
struct NvOpDev {
    int32_t ns, M, D, N, nmax, ncodes, nterms, ndterms;
    const NvTerm *terms;
    const NvStep *steps;
    const double *cmat64;  // [nterms][ns][ns]  C_t[s][s'] (row s receives)
    const float *cmat32;
    const double *ftab64;  // [ncodes][nmax]
    const float *ftab32;
    const NvDiagTerm *dterms;
    const double *dcmat;   // [ndterms][ns][ns] masked (lower-only quirk applied)
    const double *dconst;  // [ns][ns] constant block matrix, masked
    const double *omega;   // [M]
    const int32_t *nfunc;  // [M]
    const double *dg;      // [ns][D] state-diagonal part E(n) + D_ss(n), precomputed
    int32_t offstate;      // 1 if D couples different states (then computed per row)
};
