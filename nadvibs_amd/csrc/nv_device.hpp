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
#define NV_MAXM 100        // modes (basis(100) of the reference's namelist)
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

// A term's steps are the product of Delta_j in {-s_j, -s_j+2, .., s_j} over its
// modes (slot 0 slowest), less Delta = 0.  Terms of at most three modes fall
// into NV_NSHAPES shapes by (s_0+1, s_1+1, s_2+1); the shaped H*v kernel walks
// each shape's product as unrolled code with the per-mode factors loaded once
// per term.  Shapes, in group order: (2) (3) (4) (2,2) (3,2) (2,3) (1,2)
// (2,1) (2,2,2).
#define NV_NSHAPES 9
struct NvTermShape {
    int32_t cm;            // coupling matrix (row of cmat), the term's NvTerm index
    int32_t base;          // flat stride of the first step, -sum_j s_j shft_j
    int32_t inc[3];        // 2 shft_j: stride between consecutive Delta_j
    int32_t mode[3];       // 0-based modes of the slots
    int32_t code[3][4];    // ftab rows of Delta_j = -s_j + 2i
};

// Two-pass H*v of large operators.  The two slowest modes (0 and 1) are
// split off: a term t = C_t (x) P_sigma(modes 0,1) (x) P_phi(modes >= 2).
// Pass 1 gathers each distinct phi over the fast modes once per row and mixes
// it into one channel Z_g per sigma signature g; pass 2 applies P_sigma along
// modes 0 and 1 to the channels.  Signatures g (powers of modes 0 and 1):
// 0 = (0,0), 1..3 = (1..3, 0), 4..6 = (0, 1..3), 7 = (1,1), 8 = (2,1), 9 = (1,2).
#define NV_HV2_NG 10
// Shapes of phi, in group order: the NV_NSHAPES term shapes, then (1) (a phi
// whose every mode keeps Delta = 0 only).
#define NV_HV2_NSHAPES (NV_NSHAPES + 1)
struct NvPhi {
    NvTermShape sh;        // steps of phi over its modes with more than one step (sh.cm unused)
    int32_t nd;            // modes of phi whose only step is Delta = 0 (at most 2):
    int32_t dmode[2];      //   their factor h(p, n, 0) multiplies every step
    int32_t dcode[2];
    int32_t cm[NV_HV2_NG]; // per channel g: coupling matrix (row of cmat) of the
                           // term (sigma = g, phi), or -1
};
struct NvSigma {
    int32_t s0, s1;        // largest |Delta| along modes 0 and 1 (-1: mode absent)
    int32_t code0[4], code1[4];  // ftab rows of Delta = -s + 2i
    int32_t corr_cm;       // g = q_a^2 only: cmat row of the pure q_a^2 term (its
                           // Delta = 0 part is diagonal and is taken out again), or -1
    int32_t corr_code;     // ftab row of h(2, n_a, 0)
};
struct NvHv2Dev {
    int32_t on;            // tables valid for this operator
    int32_t nl, n1;        // nfunc_0 * nfunc_1, nfunc_1
    int32_t dr;            // rows per (n_0, n_1) plane: D / nl
    int32_t c;             // pass 2: consecutive plane rows per workgroup
    int32_t gmask;         // signatures present
    int32_t mfma;          // FP32: coupling products on the matrix cores (else the VALU)
    int32_t skip;          // pass 1 points the loads of out-of-basis steps outside the vector:
                           // 1 always, 0 never, -1 in FP64 only
    int32_t f32;           // the two-pass form also serves FP32
    int32_t idcm[NV_HV2_NG];  // phi = 1 (terms on modes 0 and 1 only): cmat row per channel, or -1
    const NvPhi *phi;      // grouped by shape (NV_HV2_NSHAPES groups)
    int32_t pgrp[NV_HV2_NSHAPES + 1];
    NvSigma sig[NV_HV2_NG];
};

// Fast-block H*v.  The trailing modes mf..M-1 ("fast") span P <= 64 rows, so
// a wave of 64 lanes holds one block of P consecutive rows (lane p = row
// row0 + p) whose slow digits n_0..n_{mf-1} are the same for every lane.  A
// term t = C_t (x) P_S (x) P_F splits into its slow monomial S and its fast
// monomial F; every distinct S is gathered once per wave,
//   w_S[p] = sum_{Delta_s} f_S(Delta_s) x[row0 + Delta_s + p],
// with wave-uniform factors and offsets (an out-of-basis slow step has factor
// 0 and reads the zero margin of pad rows before the interleaved copy of x),
// and mixed into one channel per fast monomial, V_F += C_t w_S.  Channel 0
// (F = 1) is y itself; each other channel is applied once through LDS,
//   y[p] += sum_{Delta_f} h_F(p, Delta_f) V_F[p + Delta_f].
// Terms with no slow step whose fast monomial is no channel, or whose
// Delta = 0 step is diagonal, are applied one by one from x[row0 + p].
// Slow shapes (stepping slots, slowest first): (2) (3) (4) (2,2) (3,2) (2,3)
// (2,2,2) and () (no stepping slow slot).
#define NV_FB_NSLOW 8
#define NV_FB_NCH 10       // channels, F = 1 included
struct NvFbS {
    int32_t base;          // row offset of the first slow step, -sum_j s_j shft_j
    int32_t inc[3];        // 2 shft_j of the stepping slow slots
    int32_t tab[3];        // their factor rows: tab[(row * nmax + n) * 4 + i] = h(p, n, -s + 2i), 0 outside the basis
    int32_t nd;            // slow slots whose only step is Delta = 0 (at most 2)
    int32_t dtab[2];       // their factor rows; entry 0 is h(p, n, 0)
    int32_t cm[NV_FB_NCH]; // per channel: coupling matrix (row of cmat) of the term (S, F), or -1
    int32_t excl0;         // channel 0 leaves out the middle step Delta_s = 0 (all stepping slots even)
    int32_t nchk;          // 1 + the last channel c with cm[c] >= 0 (channels are ordered by the
                           // degree of their fast monomial)
};
struct NvFbPure {
    int32_t cm;            // coupling matrix
    int32_t nd;            // slow factors at Delta = 0
    int32_t dtab[2];
    int32_t s0, s1;        // its fast steps [s0, s1) of the step list
};
struct NvFbDev {
    int32_t on;            // tables valid for this operator
    int32_t mf;            // first fast mode
    int32_t P;             // rows per block, prod_{k >= mf} nfunc_k (33..64)
    int32_t nblk;          // D / P
    int32_t halo;          // largest |lane offset| of a fast step
    int32_t pad;           // zero rows before and after the interleaved copy of x (>= 64, and
                           // the largest |row offset| of a slow step)
    int32_t nord;          // slow modes in the order blocks are swept, the first varying fastest:
    int32_t ordnf[16];     //   block t of the sweep decodes to digits of these sizes
    int32_t ordshft[16];   //   whose rows are these strides (row0 = sum digit * stride)
    int32_t ntab;          // factor rows
    const int32_t *tabmode;  // [ntab] mode of each factor row
    const double *tab64;   // [ntab][nmax][4]
    const float *tab32;
    const NvFbS *S;        // grouped by phase, then by slow shape
    int32_t sgrp[3][NV_FB_NSLOW + 1];  // phase 0: nchk > 4, phase 1: 1 < nchk <= 4, phase 2: nchk == 1
    int32_t nch;           // channels in use
    int32_t chs[NV_FB_NCH + 1];  // fast steps of channel c: [chs[c], chs[c+1]) (c >= 1)
    const NvStep *steps;   // fast steps: stride = lane offset; factors (mode, ftab row)
    int32_t npure;
    const NvFbPure *pure;
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
    const NvTermShape *tshape;      // [nterms] grouped by shape, term order within a group
    int32_t tgrp[NV_NSHAPES + 1];   // first descriptor of each shape group
    int64_t skip_min_bytes;         // shaped H*v: vector size above which out-of-basis loads are skipped
    NvHv2Dev hv2;                   // two-pass H*v tables (hv2.on = 0: not available)
    int64_t hv2_min_bytes;          // vector size from which the two-pass H*v serves the iteration
    NvFbDev fb;                     // fast-block H*v tables (fb.on = 0: not available)
    int64_t fb_min_bytes;           // vector size from which the fast-block H*v serves
    int32_t stream_nt;              // vectors far beyond the caches: the step, the normalisation and
                                    // the fast-block H*v's epilogue stream them with non-temporal accesses
};
