// Host launchers of the HIP kernels (implemented in nv_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nv_device.hpp"

namespace nv {

constexpr int kBlock = 256;

// Device-side state of one Lanczos run.  Vector q_j lives in slot j % nslots
// of V (nslots = niter+2 when the Krylov basis is kept, else 3); the slot of
// iteration i+1 holds the unnormalised residual until update_lanczos(i+1).
struct LzState {
    double *V = nullptr;
    int64_t nslots = 0;
    void *vlo = nullptr;         // q_i in the H*v input precision (F32/BF16), interleaved
                                 // [D][nv_lo_stride(ns)], padding zero
    void *y = nullptr;           // off-diagonal part of H q_i (accumulation type)
    double *alpha = nullptr;     // [niter+2], alpha[i]
    double *beta = nullptr;      // [niter+2], beta[i], beta[0] = |start|
    double *partials = nullptr;  // per-block partial sums
    unsigned *counters = nullptr;
    const int64_t *iter = nullptr;  // current iteration (device)
    float *Vf = nullptr;         // mixed precision with re-orthogonalisation: FP32 copy of every
                                 // archived q_j (slot j % nslots, state-major, slot stride ldf),
                                 // read by the overlap dots and the Gram-Schmidt updates
    int64_t ldf = 0;             // N rounded up to even (pairs of FP32 elements load as one)
    double *v64i = nullptr;      // FP64 runs of the shaped generic H*v: q_i interleaved
                                 // [D][nv_lo_stride(ns)], padding zero, written by normalize
    void *z2 = nullptr;          // channel scratch of the two-pass H*v (hv2_scratch_bytes), or null
    // Finishing sums folded into their consumers (iterations without re-orthogonalisation on
    // operators with few partials): every workgroup sums fold_in[0, fold_n) in the finishing
    // kernel's order, so the result has the same bits.
    const double *fold_in = nullptr;  // normalize: the previous step's beta partials (block 0 stores
                                      // beta[it-1]); step: this H*v's alpha partials (stores alpha[it])
    int32_t fold_n = 0;
    double *fold_out = nullptr;       // step: its beta partials go here instead of partials
    int64_t *next_iter = nullptr;     // step: block 0 stores it + next_inc there (a counter that no
    int32_t next_inc = 0;             //   workgroup of the same launch reads)
};

// counters[] slots
enum { kCntAlpha = 0, kCntBeta = 1, kCntPro = 2, kCntChk = 3, kCntMisc = 4, kNumCounters = 8 };

size_t hv_lds_bytes(const NvOpDev &op, int prec);
// Whether the generic H*v of precision prec runs the shaped kernels (terms
// walked by step shape) rather than the step-table kernels.  In FP64 the
// Lanczos iteration then reads the interleaved copy LzState::v64i, which the
// caller allocates (D * nv_lo_stride(ns) doubles, zeroed).
bool hv_shaped(const NvOpDev &op, int prec);
// Whether the shaped iteration H*v of precision prec points the loads of
// out-of-basis steps outside the vector (vector above op.skip_min_bytes).
bool hv_shaped_skip(const NvOpDev &op, int prec);
// Two-pass H*v (NvHv2Dev): whether it serves precision prec (tables built,
// vector at least op.hv2_min_bytes), its channel scratch, and the number of
// alpha partials the iteration H*v then writes.
bool hv2_use(const NvOpDev &op, int prec);
size_t hv2_scratch_bytes(const NvOpDev &op, int prec);
int hv2_parts(const NvOpDev &op);
int hv_iter_parts(const NvOpDev &op, int prec, const LzState &st);
// y = H x (x, y state-major FP64) by the two passes; xin is the interleaved
// copy of x (FP32 for prec 1, FP64 for prec 0), Z the channel scratch.
hipError_t launch_hv2_full(const NvOpDev &op, int prec, const double *x, const void *xin, void *Z, double *y,
                           hipStream_t s);
hipError_t launch_cast_f64i(const double *x, double *y, int64_t n, int64_t D, int ns, hipStream_t s);
// Fast-block H*v (NvFbDev): whether it serves precision prec (tables built,
// vector at least op.fb_min_bytes), the number of alpha partials its
// iteration kernel writes, and y = H x with xin the interleaved copy of x.
bool hv_fb_use(const NvOpDev &op, int prec);
int hv_fb_parts(const NvOpDev &op);
hipError_t launch_hv_fb_full(const NvOpDev &op, int prec, const double *x, const void *xin, double *y,
                             hipStream_t s);
// The same for the batched H*v (which reads the state-major FP64 vector).
bool hv_shaped_batch(const NvOpDev &op, int prec);
// Default of NvOpDev::skip_min_bytes (128 MB).
int64_t hv_skip_min_bytes();
int hv_blocks(const NvOpDev &op);

// dg[s][r] = E(r) + D_ss(r) (once per operator).
hipError_t launch_diag_precompute(const NvOpDev &op, double *dg, hipStream_t s);
// y = H x, x and y double, full operator (standalone H*v).
hipError_t launch_hv_full(const NvOpDev &op, int prec, const double *x, double *y, hipStream_t s);
// The interleaved FP32 copy: y[r * nv_lo_stride(ns) + s] = (float)x[s * D + r].
hipError_t launch_cast_f32(const double *x, float *y, int64_t n, int64_t D, int ns, hipStream_t s);
// Sets (or increments) the device iteration counter.
hipError_t launch_set_iter(int64_t *d_iter, int64_t value, int increment, hipStream_t s);
// update_lanczos: q_i = r_{i-1} / beta(i-1), plus the low-precision copy.
hipError_t launch_normalize(const NvOpDev &op, int prec, const LzState &st, hipStream_t s);
// Hv: off-diagonal part of H q_i and alpha(i) = q_i . H q_i.
hipError_t launch_hv_iter(const NvOpDev &op, int prec, const LzState &st, hipStream_t s);
// The H*v kernel of launch_hv_iter alone (its alpha partials are not summed).
hipError_t launch_hv_iter_only(const NvOpDev &op, int prec, const LzState &st, hipStream_t s);
// lanczos_step: r_i = H q_i - alpha q_i - beta(i-1) q_{i-1}, beta(i) = |r_i|.
hipError_t launch_lanczos_step(const NvOpDev &op, int prec, const LzState &st, hipStream_t s,
                               int64_t *advance = nullptr);
// Its kernel alone: the residual and the block partials of |r|^2 (hv_blocks(op) of them).
hipError_t launch_lanczos_step_only(const NvOpDev &op, int prec, const LzState &st, hipStream_t s);
// Writes the start vector: zero except vib index 0 of each state, from vals[ns].
hipError_t launch_fill_start(double *v, int64_t D, int ns, const double *vals_dev, hipStream_t s);
// out[*iter] (or out[0]) = sum (or sqrt of the sum) of nparts partials; with
// advance, then *advance = *iter + 1.
hipError_t launch_finish_sum(const double *partials, int nparts, double *out, const int64_t *iter, int take_sqrt,
                             hipStream_t s, int64_t *advance = nullptr);
// beta[0] = |v| over n doubles.
hipError_t launch_norm_to(const double *v, int64_t n, double *out, double *partials, unsigned *counter,
                          hipStream_t s);

// Normalised Gaussian envelope of sticks (x, y) on a grid; part holds
// ceil(ng/256) doubles.
// Dots of up to kMaxProbes probe vectors (length N) with the basis vectors
// q_j0..q_{j0+nj-1} (slot j % nslots of V); part holds nj * probe_chunks(N) *
// kMaxProbes partial sums, summed on the host per (j, probe).
constexpr int kMaxProbes = 8;
struct ProbeSet {
    const double *p[kMaxProbes];
    int np;
};
int probe_chunks(int64_t N);
hipError_t launch_probe_dots(const ProbeSet &ps, const double *V, int64_t nslots, int64_t N, int64_t j0, int64_t nj,
                             double *part, hipStream_t s);
hipError_t launch_convolve(const double *x, const double *y, int64_t n, const double *grid, int64_t ng,
                           double sigma, double *env, double *part, hipStream_t s);

// ---- tridiagonal eigensolve ----
struct TriWork {
    double *d = nullptr;     // [n] normalised diagonal
    double *e = nullptr;     // [n] normalised off-diagonal, e[j] couples j, j+1
    double *e2 = nullptr;    // [n] squares
    double *root = nullptr;  // [n]
    double *vec = nullptr;   // [n][n], vec[j*n + i] = component j of root i
    double *lu = nullptr;    // [5][n][n] per-root elimination rows, interleaved by root
    int32_t *ia = nullptr;   // [n] position inside a cluster of near-equal roots
    int64_t *rank = nullptr; // [n] number of cluster members (ia>0) before root i
};
hipError_t launch_tridiag(int64_t n, const TriWork &w, double anorm, hipStream_t s);

// ---- Ritz vectors and root analysis (nv_roots.hip) ----
constexpr int kRitzChunk = 16;  // Ritz vectors formed per pass over the basis
constexpr int kRitzKB = 64;     // basis vectors per LDS stage of the eigenvectors
constexpr int kGramT = 32;      // Gram tile edge
// Y[i][r] = sum_k vec[k*n + i] V[k*N + r], i < nroots, k < n.
hipError_t launch_ritz_vectors(const double *V, int64_t N, int64_t n, const double *vec, int64_t nroots, double *Y,
                               hipStream_t s);
// The same with V_k at V + ((slot0 + k) % nslots) * N (nslots >= n) and
// coefficients vec[k*ldv + i].
hipError_t launch_ritz_vectors(const double *V, int64_t N, int64_t n, const double *vec, int64_t ldv, int64_t nroots,
                               int64_t slot0, int64_t nslots, double *Y, hipStream_t s);
// Partial-result blocks per root of launch_topk (pm, pi hold nroots * topk_blocks(N)).
int topk_blocks(int64_t N);
// For each of nroots vectors Y + root*N: the kmax largest squared elements in
// rank order (ties by index), their indices and signs.  out_* are [nroots][kmax].
hipError_t launch_topk(const double *Y, int64_t N, int nroots, int kmax, double *prev_m, int64_t *prev_i, double *pm,
                       int64_t *pi, double *out_m, int64_t *out_i, int32_t *out_ph, hipStream_t s);
// Partial-tile layout of launch_gram: part holds nchunks * npairs * kGramT^2 doubles.
void gram_shape(int64_t D, int R, int &npairs, int &nchunks, int64_t &rows_per_chunk);
// G (R x R) = A^T A, column c of A at A + c*D.
hipError_t launch_gram(const double *A, int64_t D, int R, double *part, double *G, hipStream_t s);

// ---- batched Lanczos (batch_* kernels in nv_kernels.hip) ----
struct BatchTile {
    int32_t prob;  // problem index
    int32_t row0;  // first vibrational row of the tile
};
struct BatchDev {
    const NvOpDev *ops = nullptr;     // [P] operator tables
    const BatchTile *tiles = nullptr; // [T] tiles, grouped by number of states, problem-major within
    const int32_t *tile0 = nullptr;   // [P][2] first and one-past-last tile of each problem
    const int64_t *voff = nullptr;    // [P] offset of problem p in the state-major vectors
    const int64_t *loff = nullptr;    // [P] offset of problem p in the interleaved low-precision copy
    double *V = nullptr;              // 3 slots of Ntot doubles (slot i % 3)
    int64_t Ntot = 0;
    void *vlo = nullptr;              // interleaved FP32 copy (mixed precision)
    void *y = nullptr;                // off-diagonal part of H q_i, state-major
    double *alpha = nullptr, *beta = nullptr;  // [P][ld], alpha[p][i], beta[p][i]
    int64_t ld = 0;
    double *partials = nullptr;       // [T]
    const int64_t *iter = nullptr;    // current iteration (device)
    int32_t P = 0, T = 0, bs = 256;   // problems, tiles, rows per tile (= workgroup size)
    int32_t shaped[2] = {0, 0};       // every operator runs the shaped H*v (FP64, FP32)
    int32_t rows = 2;                 // FP32 H*v: rows per thread (1, 2 or 4; 2 needs bs >= 128, 4 the shaped
                                      // form, bs >= 256 and launch LDS for four rows' digits)
    int32_t steps = 0;                // 1: the step-table H*v even where the shaped one serves
};
// Tiles [t0, t0+nt) all belong to problems with ns states (one launch each).
struct BatchGroup {
    int ns, t0, nt;
};
size_t batch_hv_lds(const NvOpDev &op, int prec);
// update_lanczos, Hv (+ alpha finish) and lanczos_step (+ beta finish) for every problem.
hipError_t launch_batch_normalize(const BatchDev &b, int prec, hipStream_t s);
hipError_t launch_batch_hv(const BatchDev &b, const BatchGroup *groups, int ngroups, int prec, size_t lds,
                           hipStream_t s);
hipError_t launch_batch_step(const BatchDev &b, const BatchGroup *groups, int ngroups, int prec, size_t lds,
                             hipStream_t s);

// ---- adjoint (nv_adjoint.hip) ----
int adj_blocks(int64_t N);
// One reverse step of the recurrence at iteration i (see nv_adjoint.hip).
// part holds 5 * adj_blocks(N) doubles, scal 3.
hipError_t launch_adj_step(int64_t N, int64_t i, double *qbn, const double *hub, const double *qn, double sn,
                           const double *qi, const double *qp, const double *alpha, const double *beta,
                           const double *abar, double *bbar, double *part, double *scal, double *ub, double *qbi,
                           double *qbp, hipStream_t s);
hipError_t launch_adj_zero(int64_t N, double *v, hipStream_t s);
// hub += (H^T - H) ub: the unsymmetric inter-state diagonal part (lower-only quirk).
hipError_t launch_adj_transpose(const NvOpDev &op, const double *ub, double *hub, hipStream_t s);
// Bilinear forms ubar.(P_t q) of every monomial of g, added to gacc
// [grad_groups][ns*ns]; part holds grad_groups * grad_chunks * ns*ns doubles.
int grad_groups(const NvOpDev &g);
int grad_chunks(const NvOpDev &g);
size_t grad_lds_bytes(const NvOpDev &g);
hipError_t launch_grad_contract(const NvOpDev &g, const double *ub, const double *q, double *part, double *gacc,
                                hipStream_t s);
// Broadened spectrum of tridiagonal eigenpairs (theta on the device, vec as
// TriWork::vec) and its adjoint; work holds 2n + 8 + 2n^2 doubles.
hipError_t launch_spec_env(int64_t n, const double *theta, const double *vec, double ref, double c, double sigma,
                           int64_t ng, const double *grid, double *env, hipStream_t s);
hipError_t launch_spec_adjoint(int64_t n, const double *theta, const double *vec, double ref, int lowest, double c,
                               double sigma, int64_t ng, const double *grid, const double *gbar, double *work,
                               double *abar, double *bbar, double *tbar, hipStream_t s);

}  // namespace nv
