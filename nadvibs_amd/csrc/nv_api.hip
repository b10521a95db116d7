// C ABI of the engine (include/nadvibs.h).  Owns the context: operator
// tables on the device, the Krylov basis and every scratch buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "nadvibs.h"
#include <cstdlib>

#include "nv_jit.hpp"
#include "nv_kernels.hpp"
#include "nv_operator.hpp"
#include "nv_pro.hpp"

struct nv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    nv::Operator op;
    NvOpDev dev{};
    std::vector<void *> allocs;  // operator tables

    // standalone H*v buffers
    double *hx = nullptr, *hy = nullptr;

    // Lanczos run
    bool lz_ready = false;
    nv_lanczos_opts opts{};
    nv::LzState st{};
    int64_t *d_iter = nullptr;
    double *d_startvals = nullptr;
    int64_t last_iter = 0;       // last iteration whose lanczos_step was issued
    int64_t archive_iter = 0;    // iteration up to which nv_load_archive restored the kept basis
    bool pro_loaded = false;     // nv_load_archive restored the re-orthogonalisation state
    int64_t next_update = 1;     // iteration expected by update_lanczos
    std::vector<void *> lz_allocs;
    nv::ProState pro{};

    // graphs of one iteration: [0] plain, [1] with re-orthogonalisation,
    // [2] re-orthogonalisation with the exact overlap check
    hipGraphExec_t graphs[4] = {nullptr, nullptr, nullptr, nullptr};  // kinds 0-2, and kIterBatch of kind 0
    hipGraphExec_t graph = nullptr;

    // operator-specialised H*v kernels per precision (F64, F32)
    bool jit_on = true;
    const NvTermShape *tshape = nullptr;  // the shaped-kernel descriptors (dev.tshape when in use)
    double *v64i = nullptr;               // interleaved FP64 copy of q_i (shaped generic H*v in FP64)
    nv::JitKernels jit[2];
    std::string jit_log;
    float *hx32 = nullptr;
    double *hx64i = nullptr;              // interleaved FP64 copy of a standalone H*v input (two-pass form)
    void *z2 = nullptr;                   // channel scratch of the two-pass H*v
    size_t z2_bytes = 0;

    // raw nadvibs.in coefficients (for the frequency chain rule of the adjoint)
    std::vector<double> potraw;
    double ztoler = 0.0;
    int64_t ntmax = 0;
    // checkpoints (q_{cK}, r_{cK}) of a run without the archive, c = 0..ncp-1
    double *cp = nullptr;
    int64_t ncp = 0;
    // gradient operator: every monomial with a unit coefficient (adjoint)
    bool gop_ready = false;
    nv::Operator gop;
    NvOpDev gdev{};
    std::vector<void *> gallocs;

    // H*v timing
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;
    size_t lz_bytes_key = 0;
};

static std::string g_create_error;

// ---- static scratch of the no-allocation entry point ------------------------
// nv_convolve_noalloc keeps its per-workgroup maxima (one double per kBlock grid
// points) here instead of in an allocation: 32 KiB per GPU, held until the
// process ends.  In host code `scratch` only names the symbol; its address on
// each device comes from hipGetSymbolAddress, taken once per device in
// nv_create with that device current.
constexpr int64_t kScratchDoubles = NV_CONVOLVE_MAX_GRID / nv::kBlock;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices] = {nullptr};
static std::mutex g_scratch_mutex;

// The address of `scratch` on the current device `device` (once per device).
static hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipSuccess;  // the entry point then refuses
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[device] = static_cast<double *>(p);
    return e;
}

static double *scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

namespace {

int32_t fail(nv_ctx *c, int32_t code, const std::string &msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

int32_t hip_fail(nv_ctx *c, hipError_t e, const char *where) {
    return fail(c, NV_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define NV_HIP(c, call)                                      \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
    } while (0)

template <typename T>
hipError_t upload(nv_ctx *c, const std::vector<T> &h, const T **out, std::vector<void *> &list) {
    *out = nullptr;
    if (h.empty()) return hipSuccess;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, h.size() * sizeof(T));
    if (e != hipSuccess) return e;
    list.push_back(p);
    e = hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    *out = reinterpret_cast<const T *>(p);
    return e;
}

void free_list(std::vector<void *> &l) {
    for (void *p : l) hipFree(p);
    l.clear();
}

void release_graphs(nv_ctx *c) {
    for (auto &g : c->graphs)
        if (g) { hipGraphExecDestroy(g); g = nullptr; }
}

void release_lanczos(nv_ctx *c) {
    release_graphs(c);
    if (c->graph) { hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    free_list(c->lz_allocs);
    c->cp = nullptr;
    c->ncp = 0;
    c->st = nv::LzState{};
    c->pro = nv::ProState{};
    c->lz_ready = false;
}

bool valid_prec(int32_t p) { return p == NV_PREC_F64 || p == NV_PREC_F32 || p == NV_PREC_BF16; }

}  // namespace

extern "C" {

int32_t nv_create(const nv_problem *prob, int32_t device, nv_ctx **out) {
    if (!out) return fail(nullptr, NV_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!prob) return fail(nullptr, NV_ERR_ARG, "problem is NULL");
    nv_ctx *c = new (std::nothrow) nv_ctx();
    if (!c) return fail(nullptr, NV_ERR_NOMEM, "host allocation failed");
    std::string err;
    if (!nv::build_operator(*prob, c->op, err)) {
        delete c;
        return fail(nullptr, NV_ERR_ARG, err);
    }
    c->device = device;
    c->ntmax = nv::noterms(prob->nmodes, prob->ordr);
    c->potraw.assign(prob->potterms, prob->potterms + c->ntmax * prob->ordr * c->op.nstblks);
    c->ztoler = prob->ztoler;
    {
        const char *k = std::getenv("NV_HV_KERNEL");
        c->jit_on = !(k && (std::strcmp(k, "generic") == 0 || std::strcmp(k, "steps") == 0));
    }
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = scratch_setup(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    const nv::Operator &op = c->op;
    NvOpDev &d = c->dev;
    d.ns = (int32_t)op.ns; d.M = (int32_t)op.M; d.D = (int32_t)op.D; d.N = (int32_t)op.N;
    d.nmax = (int32_t)op.nmax; d.ncodes = op.ncodes;
    d.nterms = (int32_t)op.dev_terms.size(); d.ndterms = (int32_t)op.dev_dterms.size();
    std::vector<float> cm32(op.cmat.begin(), op.cmat.end()), ft32(op.ftab.begin(), op.ftab.end());
    std::vector<int32_t> nf32(op.nfunc.begin(), op.nfunc.end());
    if (e == hipSuccess) e = upload(c, op.dev_terms, &d.terms, c->allocs);
    if (e == hipSuccess) e = upload(c, op.dev_steps, &d.steps, c->allocs);
    if (e == hipSuccess) e = upload(c, op.cmat, &d.cmat64, c->allocs);
    if (e == hipSuccess) e = upload(c, cm32, &d.cmat32, c->allocs);
    if (e == hipSuccess) e = upload(c, op.ftab, &d.ftab64, c->allocs);
    if (e == hipSuccess) e = upload(c, ft32, &d.ftab32, c->allocs);
    if (e == hipSuccess) e = upload(c, op.dev_dterms, &d.dterms, c->allocs);
    if (e == hipSuccess) e = upload(c, op.dcmat, &d.dcmat, c->allocs);
    if (e == hipSuccess) e = upload(c, op.dconst, &d.dconst, c->allocs);
    if (e == hipSuccess) e = upload(c, op.omega, &d.omega, c->allocs);
    if (e == hipSuccess) e = upload(c, nf32, &d.nfunc, c->allocs);
    if (e == hipSuccess && op.shaped) e = upload(c, op.dev_shapes, &d.tshape, c->allocs);
    for (int q = 0; q <= NV_NSHAPES; ++q) d.tgrp[q] = d.tshape ? op.shape_grp[q] : 0;
    c->tshape = d.tshape;
    // Two-pass H*v: by default in FP64 from 128 MB of vector (the 88.5M-row cubic operator:
    // 37.9 against 40.5 ms per H*v); in FP32 the one-pass shaped kernel stays ahead there
    // (17.05 against 18.1 ms), so FP32 takes it only with NV_HV_TWOPASS=1, which also
    // drops the size threshold; NV_HV_TWOPASS=0 turns it off.
    d.hv2 = NvHv2Dev{};
    d.hv2_min_bytes = (int64_t)128 << 20;
    {
        const char *k = std::getenv("NV_HV_TWOPASS");
        const bool force = k && std::strcmp(k, "1") == 0;
        if (force) d.hv2_min_bytes = 0;
        d.hv2.f32 = force ? 1 : 0;
        if (op.hv2 && !(k && std::strcmp(k, "0") == 0)) {
            NvHv2Dev &h = d.hv2;
            h.nl = (int32_t)(op.nfunc[0] * op.nfunc[1]);
            h.n1 = (int32_t)op.nfunc[1];
            h.dr = (int32_t)(op.D / h.nl);
            h.c = std::max(1, nv::kBlock / h.nl);
            h.gmask = op.hv2_gmask;
            // FP32 coupling products on v_mfma_f32_4x4x1 measured slower than the VALU
            // (20.3 against 18.1 ms on the 88.5M-row operator): opt-in with NV_HV2_MFMA=1
            const char *m = std::getenv("NV_HV2_MFMA");
            h.mfma = m && std::strcmp(m, "1") == 0;
            const char *sk = std::getenv("NV_HV2_SKIP");
            h.skip = sk ? (std::strcmp(sk, "1") == 0 ? 1 : 0) : -1;
            for (int q = 0; q <= NV_HV2_NSHAPES; ++q) h.pgrp[q] = op.hv2_pgrp[q];
            for (int g = 0; g < NV_HV2_NG; ++g) {
                h.sig[g] = op.hv2_sig[g];
                h.idcm[g] = op.hv2_idcm[g];
            }
            if (e == hipSuccess) e = upload(c, op.hv2_phi, &h.phi, c->allocs);
            h.on = e == hipSuccess ? 1 : 0;
        }
    }
    // Fast-block H*v: by default from 16 MB of interleaved vector, where the
    // vector leaves the caches (four states, ordr 3, 2.1M rows: FP32 633 against
    // 3286 us for the run-time compiled kernel, FP64 1578 against 1916 us; 22M
    // rows: 13.2 against 17.05 ms and 30.3 against 38.0 ms for the two-pass
    // form); below, the cache-resident vector favours the other kernels (three
    // states, 0.46M rows: 40.9 against 26.0 us).  NV_HV_FB=1 takes it at any
    // size, NV_HV_FB=0 turns it off.
    // Non-temporal streaming of the vectors that the step, the normalisation and the fast-block
    // epilogue touch once, from 512 MB of FP64 vector (twice the Infinity Cache): 88.5M elements,
    // 11.62 -> 11.49 ms per FP32 iteration; a cache-resident vector is better read again from the
    // cache (1.38M elements: 48.5 -> 54.5 us with it).  NV_STREAM_NT=1 / 0 forces it on / off.
    d.stream_nt = (int64_t)op.N * 8 >= ((int64_t)512 << 20) ? 1 : 0;
    if (const char *k = std::getenv("NV_STREAM_NT")) d.stream_nt = std::strcmp(k, "1") == 0 ? 1 : 0;
    d.fb = NvFbDev{};
    d.fb_min_bytes = (int64_t)16 << 20;
    {
        const char *k = std::getenv("NV_HV_FB");
        if (k && std::strcmp(k, "1") == 0) d.fb_min_bytes = 0;
        if (op.fb && !(k && std::strcmp(k, "0") == 0)) {
            NvFbDev &f = d.fb;
            f.mf = op.fb_mf;
            f.P = op.fb_P;
            f.nblk = (int32_t)(op.D / op.fb_P);
            f.halo = op.fb_halo;
            f.pad = op.fb_pad;
            f.nch = op.fb_nch;
            for (int ph = 0; ph < 3; ++ph)
                for (int q = 0; q <= NV_FB_NSLOW; ++q) f.sgrp[ph][q] = op.fb_sgrp[ph][q];
            for (int q = 0; q <= NV_FB_NCH; ++q) f.chs[q] = op.fb_chs[q];
            f.npure = (int32_t)op.fb_pure.size();
            f.ntab = (int32_t)op.fb_tabmode.size();
            // sweep order: rows in order (one digit of nblk blocks of P rows), or NV_FB_ORDER, a
            // permutation of the slow modes listed from the fastest-varying in the sweep
            f.nord = 1;
            f.ordnf[0] = f.nblk;
            f.ordshft[0] = f.P;
            if (const char *o = std::getenv("NV_FB_ORDER")) {
                std::vector<int> ms;
                for (const char *q = o; *q;) {
                    char *end = nullptr;
                    const long v = std::strtol(q, &end, 10);
                    if (end == q) break;
                    ms.push_back((int)v);
                    q = *end == ',' ? end + 1 : end;
                }
                std::vector<int> seen(op.fb_mf, 0);
                bool ok = (int)ms.size() == op.fb_mf && op.fb_mf <= 16;
                for (int m : ms) ok = ok && m >= 0 && m < op.fb_mf && !seen[m]++;
                if (ok) {
                    f.nord = op.fb_mf;
                    for (int j = 0; j < op.fb_mf; ++j) {
                        f.ordnf[j] = (int32_t)op.nfunc[ms[j]];
                        f.ordshft[j] = (int32_t)op.shft[ms[j]];
                    }
                }
            }
            std::vector<float> t32(op.fb_tab.begin(), op.fb_tab.end());
            if (e == hipSuccess) e = upload(c, op.fb_S, &f.S, c->allocs);
            if (e == hipSuccess) e = upload(c, op.fb_steps, &f.steps, c->allocs);
            if (e == hipSuccess && !op.fb_pure.empty()) e = upload(c, op.fb_pure, &f.pure, c->allocs);
            if (e == hipSuccess) e = upload(c, op.fb_tab, &f.tab64, c->allocs);
            if (e == hipSuccess) e = upload(c, op.fb_tabmode, &f.tabmode, c->allocs);
            if (e == hipSuccess) e = upload(c, t32, &f.tab32, c->allocs);
            f.on = e == hipSuccess ? 1 : 0;
        }
    }
    {
        const char *k = std::getenv("NV_HV_KERNEL");
        if (k && std::strcmp(k, "steps") == 0) d.tshape = nullptr;
        // The shaped H*v points the loads of out-of-basis steps outside the
        // vector (no memory traffic) above a vector size: 128 MB, or 16 MB
        // when under 75% of the (row, step) pairs lie inside the basis (small
        // bases, cubic terms: a 4-state ordr-3 operator with eight functions
        // per mode and a 34 MB FP32 copy runs 5% faster, 13% in FP64).  Below,
        // the vector sits in the caches and the compare only costs (1-6%).
        // NV_HV_SKIP_MIN_BYTES sets the size instead (0: always).
        const char *m = std::getenv("NV_HV_SKIP_MIN_BYTES");
        d.skip_min_bytes = m ? std::strtoll(m, nullptr, 10)
                             : (op.step_valid < 0.75 ? ((int64_t)16 << 20) : nv::hv_skip_min_bytes());
    }
    if (e != hipSuccess) {
        g_create_error = std::string("device setup failed: ") + hipGetErrorString(e);
        free_list(c->allocs);
        if (c->stream) hipStreamDestroy(c->stream);
        delete c;
        return NV_ERR_HIP;
    }
    // state-diagonal part of H, and whether D couples different states
    d.offstate = 0;
    for (int64_t a = 0; a < op.ns; ++a)
        for (int64_t b = 0; b < op.ns; ++b) {
            if (a == b) continue;
            if (op.dconst[a * op.ns + b] != 0.0) d.offstate = 1;
            for (size_t t = 0; t < op.dev_dterms.size(); ++t)
                if (op.dcmat[t * op.ns * op.ns + a * op.ns + b] != 0.0) d.offstate = 1;
        }
    if (e == hipSuccess) {
        void *p = nullptr;
        e = hipMalloc(&p, (size_t)op.N * sizeof(double));
        if (e == hipSuccess) {
            c->allocs.push_back(p);
            d.dg = (const double *)p;
            e = nv::launch_diag_precompute(d, (double *)p, c->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        g_create_error = std::string("device setup failed: ") + hipGetErrorString(e);
        free_list(c->allocs);
        hipStreamDestroy(c->stream);
        delete c;
        return NV_ERR_HIP;
    }
    if (nv::hv_lds_bytes(d, NV_PREC_F64) > 64 * 1024) {
        free_list(c->allocs);
        hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, NV_ERR_ARG,
                    "factor table and mode digits exceed the 64 KiB LDS budget of the H*v kernel");
    }
    *out = c;
    return NV_OK;
}

int32_t nv_destroy(nv_ctx *c) {
    if (!c) return NV_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    release_lanczos(c);
    for (auto &k : c->jit) nv::jit_release(k);
    free_list(c->gallocs);
    if (c->z2) hipFree(c->z2);
    free_list(c->allocs);
    for (auto &p : c->ev) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
    if (c->d_startvals) hipFree(c->d_startvals);
    if (c->hx) hipFree(c->hx);
    if (c->hy) hipFree(c->hy);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return NV_OK;
}

const char *nv_last_error(const nv_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int32_t nv_dims(const nv_ctx *c, int64_t *dimen, int64_t *n) {
    if (!c) return NV_ERR_ARG;
    if (dimen) *dimen = c->op.D;
    if (n) *n = c->op.N;
    return NV_OK;
}

int32_t nv_term_summary(const nv_ctx *c, int64_t *nztrms) {
    if (!c || !nztrms) return NV_ERR_ARG;
    for (int64_t i = 0; i < c->op.ordr; ++i) nztrms[i] = c->op.nztrms[i];
    return NV_OK;
}

int32_t nv_term_counts(const nv_ctx *c, int64_t *out) {
    if (!c || !out) return NV_ERR_ARG;
    for (size_t i = 0; i < c->op.ntcom.size(); ++i) out[i] = c->op.ntcom[i];
    return NV_OK;
}

// Compiles the operator-specialised kernels of one precision on first use;
// when that is not possible the generic kernels serve (both are HIP).
static const nv::JitKernels *ensure_jit(nv_ctx *c, int32_t prec) {
    if (!c->jit_on || (prec != NV_PREC_F64 && prec != NV_PREC_F32)) return nullptr;
    // in FP64 the shaped generic kernel over the interleaved copy of q_i is the
    // faster one (three states, N = 1.38M: 49.7 against 58.6 us); the specialised FP64 kernel serves
    // the operators it does not take
    if (prec == NV_PREC_F64 && nv::hv_shaped(c->dev, NV_PREC_F64)) return nullptr;
    // large vectors: the fast-block or the two-pass H*v
    if (c->dev.tshape && (nv::hv_fb_use(c->dev, prec) || nv::hv2_use(c->dev, prec))) return nullptr;
    nv::JitKernels &k = c->jit[prec];
    if (k.ready) return &k;
    if (k.failed) return nullptr;
    std::string why;
    if (!nv::jit_eligible(c->op, prec, why)) {
        k.failed = true;
        c->jit_log = why;
        return nullptr;
    }
    if (nv::jit_build(c->op, prec, k, c->jit_log) != hipSuccess) {
        k.failed = true;
        return nullptr;
    }
    return &k;
}

static hipError_t launch_hv_iter_st(nv_ctx *c, const nv::LzState &st) {
    const nv::JitKernels *k = ensure_jit(c, c->opts.prec);
    if (!k) return nv::launch_hv_iter(c->dev, c->opts.prec, st, c->stream);
    hipError_t e = nv::jit_launch_iter(*k, c->dev, st, c->stream);
    if (e != hipSuccess) return e;
    return nv::launch_finish_sum(st.partials, nv::hv_blocks(c->dev), st.alpha, st.iter, 0, c->stream);
}

static hipError_t launch_hv_iter_any(nv_ctx *c) { return launch_hv_iter_st(c, c->st); }

// The channel scratch of the two-pass H*v, at least `bytes` (a larger
// buffer replaces the old one; captured iterations name it, so they go).
static hipError_t ensure_z2(nv_ctx *c, size_t bytes) {
    if (c->z2 && c->z2_bytes >= bytes) return hipSuccess;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    release_graphs(c);
    if (c->z2) hipFree(c->z2);
    c->z2 = nullptr;
    c->z2_bytes = 0;
    e = hipMalloc(&c->z2, bytes);
    if (e == hipSuccess) c->z2_bytes = bytes;
    else c->z2 = nullptr;
    return e;
}

// Rows of zero margin before and after every interleaved copy of a vector:
// the fast-block H*v reads up to op.fb_pad rows beyond either end.
static size_t il_pad(const nv_ctx *c) { return c->op.fb ? (size_t)c->op.fb_pad : 0; }

// An interleaved copy (D rows of nv_lo_stride(ns) values of el bytes) with
// its margins, zeroed on stream s; *first is its first row and the
// allocation joins owner.
static hipError_t alloc_interleaved(nv_ctx *c, size_t el, std::vector<void *> &owner, void **first, hipStream_t s) {
    const size_t nsp = (size_t)nv_lo_stride((int)c->op.ns), pad = il_pad(c);
    const size_t bytes = ((size_t)c->op.D + 2 * pad) * nsp * el;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return e;
    owner.push_back(p);
    e = hipMemsetAsync(p, 0, bytes, s);
    *first = (char *)p + pad * nsp * el;
    return e;
}

// Whether the fast-block H*v serves precision prec on this context (the
// shaped kernels selected and the tables built for a large enough vector).
static bool use_fb(const nv_ctx *c, int32_t prec) { return c->dev.tshape && nv::hv_fb_use(c->dev, prec); }

// Whether the two-pass H*v serves precision prec (when the fast-block form does not).
static bool use_hv2(const nv_ctx *c, int32_t prec) {
    return c->dev.tshape && !use_fb(c, prec) && nv::hv2_use(c->dev, prec);
}

// The interleaved copy of x that the fast-block and two-pass forms read:
// FP32 for prec NV_PREC_F32, FP64 otherwise.
static hipError_t interleaved_input(nv_ctx *c, int32_t prec, const double *x, const void **xin, hipStream_t s) {
    hipError_t e = hipSuccess;
    if (prec == NV_PREC_F32) {
        if (!c->hx32) {
            e = alloc_interleaved(c, sizeof(float), c->allocs, (void **)&c->hx32, s);
            if (e != hipSuccess) return e;
        }
        e = nv::launch_cast_f32(x, c->hx32, c->op.N, c->op.D, (int)c->op.ns, s);
        *xin = c->hx32;
    } else {
        if (!c->hx64i) {
            e = alloc_interleaved(c, sizeof(double), c->allocs, (void **)&c->hx64i, s);
            if (e != hipSuccess) return e;
        }
        e = nv::launch_cast_f64i(x, c->hx64i, c->op.N, c->op.D, (int)c->op.ns, s);
        *xin = c->hx64i;
    }
    return e;
}

// y = H x (double in and out) on stream s, couplings in precision prec.
static hipError_t hv_apply(nv_ctx *c, int32_t prec, const double *x, double *y, hipStream_t s) {
    if (use_fb(c, prec)) {
        const void *xin = nullptr;
        hipError_t e = interleaved_input(c, prec, x, &xin, s);
        if (e != hipSuccess) return e;
        return nv::launch_hv_fb_full(c->dev, prec, x, xin, y, s);
    }
    if (use_hv2(c, prec)) {
        hipError_t e = ensure_z2(c, nv::hv2_scratch_bytes(c->dev, prec));
        if (e != hipSuccess) return e;
        const void *xin = nullptr;
        e = interleaved_input(c, prec, x, &xin, s);
        if (e != hipSuccess) return e;
        return nv::launch_hv2_full(c->dev, prec, x, xin, c->z2, y, s);
    }
    const nv::JitKernels *k = ensure_jit(c, prec);
    if (!k) return nv::launch_hv_full(c->dev, prec, x, y, s);
    const void *xin = x;
    if (prec == NV_PREC_F32) {
        if (!c->hx32) {
            hipError_t e = alloc_interleaved(c, sizeof(float), c->allocs, (void **)&c->hx32, s);
            if (e != hipSuccess) return e;
        }
        hipError_t e = nv::launch_cast_f32(x, c->hx32, c->op.N, c->op.D, (int)c->op.ns, s);
        if (e != hipSuccess) return e;
        xin = c->hx32;
    }
    return nv::jit_launch_full(*k, c->dev, x, xin, y, s);
}

int32_t nv_hv_device(nv_ctx *c, int32_t prec, const double *x, double *y, void *stream) {
    if (!c) return NV_ERR_ARG;
    if (!x || !y) return fail(c, NV_ERR_ARG, "x or y is NULL");
    if (prec != NV_PREC_F64 && prec != NV_PREC_F32)
        return fail(c, NV_ERR_ARG, "nv_hv supports NV_PREC_F64 and NV_PREC_F32");
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hv_apply(c, prec, x, y, s));
    return NV_OK;
}

int32_t nv_select_hv_kernel(nv_ctx *c, int32_t generic) {
    if (!c) return NV_ERR_ARG;
    const bool on = generic == 0;
    const NvTermShape *ts = generic == 2 ? nullptr : c->tshape;
    // captured iterations name the kernel
    if (on != c->jit_on || ts != c->dev.tshape) release_graphs(c);
    c->jit_on = on;
    c->dev.tshape = ts;
    return NV_OK;
}

int32_t nv_hv(nv_ctx *c, int32_t prec, const double *x, double *y) {
    if (!c) return NV_ERR_ARG;
    if (!x || !y) return fail(c, NV_ERR_ARG, "x or y is NULL");
    NV_HIP(c, hipSetDevice(c->device));
    const size_t bytes = (size_t)c->op.N * sizeof(double);
    if (!c->hx) {
        NV_HIP(c, hipMalloc(&c->hx, bytes));
        NV_HIP(c, hipMalloc(&c->hy, bytes));
    }
    NV_HIP(c, hipMemcpyAsync(c->hx, x, bytes, hipMemcpyHostToDevice, c->stream));
    int32_t rc = nv_hv_device(c, prec, c->hx, c->hy, c->stream);
    if (rc != NV_OK) return rc;
    NV_HIP(c, hipMemcpyAsync(y, c->hy, bytes, hipMemcpyDeviceToHost, c->stream));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    return NV_OK;
}

// ---- start vector -------------------------------------------------------
// Default start (shiftref = 0, Source/secondary.f90:727-783): |0> in every
// state, weighted by statew / sum(statew), normalised.  The Franck-Condon
// start (shiftref != 0) is built by nv::build_fc_start.
static int32_t build_start(nv_ctx *c, double *dst) {
    const nv::Operator &op = c->op;
    if (op.neworigin) {
        hipError_t e = nv::build_fc_start(op, c->dev, dst, c->stream, c->err);
        if (e != hipSuccess) return c->err.empty() ? hip_fail(c, e, "fc start") : NV_ERR_HIP;
        return NV_OK;
    }
    std::vector<double> vals(op.ns);
    // q3 = 1 at index 1 of every state; d1 = sqrt(q3.q3/nstates) = 1
    double d2 = 0.0;
    for (int64_t s = 0; s < op.ns; ++s) d2 += op.statew[s];
    if (d2 == 0.0) d2 = 1.0;
    double tot = 0.0;
    for (int64_t s = 0; s < op.ns; ++s) {
        vals[s] = 1.0 * (op.statew[s] / (d2 * 1.0));
        tot += vals[s] * vals[s];
    }
    const double sc = 1.0 / std::sqrt(tot);
    for (int64_t s = 0; s < op.ns; ++s) vals[s] *= sc;
    if (!c->d_startvals) NV_HIP(c, hipMalloc(&c->d_startvals, NV_MAXS * sizeof(double)));
    NV_HIP(c, hipMemcpyAsync(c->d_startvals, vals.data(), op.ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    NV_HIP(c, nv::launch_fill_start(dst, op.D, (int)op.ns, c->d_startvals, c->stream));
    return NV_OK;
}

int32_t nv_start_vector(nv_ctx *c, double *q) {
    if (!c) return NV_ERR_ARG;
    NV_HIP(c, hipSetDevice(c->device));
    const size_t bytes = (size_t)c->op.N * sizeof(double);
    double *tmp = nullptr;
    NV_HIP(c, hipMalloc(&tmp, bytes));
    int32_t rc = build_start(c, tmp);
    if (rc == NV_OK && q) {
        hipError_t e = hipMemcpyAsync(q, tmp, bytes, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "start vector copy");
    }
    hipStreamSynchronize(c->stream);
    hipFree(tmp);
    return rc;
}

// ---- Lanczos --------------------------------------------------------------
int32_t nv_lanczos_init(nv_ctx *c, const nv_lanczos_opts *o) {
    if (!c || !o) return NV_ERR_ARG;
    if (o->niter < 1) return fail(c, NV_ERR_ARG, "niter must be >= 1");
    if (!valid_prec(o->prec)) return fail(c, NV_ERR_ARG, "unknown precision");
    if (o->reorthog > 0 && nv::pro_decide_lds(o->niter) != 0)
        return fail(c, NV_ERR_ARG, "re-orthogonalisation supports up to 32768 iterations");
    if (o->reorthog > 0 && o->prec == NV_PREC_BF16)
        return fail(c, NV_ERR_ARG,
                    "NV_PREC_BF16 does not support re-orthogonalisation: the 2^-9 input rounding outgrows it "
                    "(non-finite coefficients on long runs); use NV_PREC_F32");
    if (o->reorthog > 0) {
        // the range detection assumes that no index is both small (|omega| < eta) and large (>= tau)
        double tau = 0.0, eta = 0.0;
        nv::pro_effective_thresholds(*o, &tau, &eta);
        if (!(eta < tau))
            return fail(c, NV_ERR_ARG, "re-orthogonalisation needs pro_eta < pro_tau (after the defaults: tau " +
                                           std::to_string(tau) + ", eta " + std::to_string(eta) + ")");
    }
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t N = c->op.N;
    const bool keep = o->reorthog > 0 || o->savevecs != 0;
    // buffers of an earlier run with the same shape are reused
    if (o->checkpoint < 0) return fail(c, NV_ERR_ARG, "checkpoint must be >= 0");
    if (o->checkpoint > 0 && keep)
        return fail(c, NV_ERR_ARG, "checkpoints apply to runs without re-orthogonalisation or savevecs: those "
                                   "keep every Lanczos vector (re-orthogonalisation reads them)");
    const int64_t K = keep ? 0 : o->checkpoint;
    const bool same = c->lz_ready && c->opts.niter == o->niter && c->opts.prec == o->prec &&
                      (keep ? 0 : c->opts.checkpoint) == K &&
                      (c->opts.reorthog > 0 || c->opts.savevecs != 0) == keep &&
                      (c->opts.reorthog > 0) == (o->reorthog > 0) && (c->opts.reorthog > 1) == (o->reorthog > 1);
    nv::LzState &st = c->st;
    // captured iteration graphs hold the options by value
    if (!same || std::memcmp(&c->opts, o, sizeof(*o)) != 0) release_graphs(c);
    if (!same) {
        release_lanczos(c);
        st.nslots = keep ? o->niter + 2 : 3;
        auto alloc = [&](size_t bytes, void **p) -> hipError_t {
            hipError_t e = hipMalloc(p, bytes);
            if (e == hipSuccess) c->lz_allocs.push_back(*p);
            return e;
        };
        NV_HIP(c, alloc((size_t)st.nslots * N * sizeof(double), (void **)&st.V));
        const size_t el = o->prec == NV_PREC_F64 ? 8 : (o->prec == NV_PREC_F32 ? 4 : 2);
        if (o->prec != NV_PREC_F64) {
            NV_HIP(c, alloc_interleaved(c, el, c->lz_allocs, &st.vlo, c->stream));
            // re-orthogonalisation reads the FP32 copy of the archive (half the bytes)
            if (o->reorthog > 0) {
                st.ldf = N + (N & 1);
                NV_HIP(c, alloc((size_t)st.nslots * st.ldf * sizeof(float), (void **)&st.Vf));
                NV_HIP(c, hipMemsetAsync(st.Vf, 0, (size_t)st.nslots * st.ldf * sizeof(float), c->stream));
            }
        }
        NV_HIP(c, alloc((size_t)N * (o->prec == NV_PREC_F64 ? 8 : 4), &st.y));
        NV_HIP(c, alloc((o->niter + 3) * sizeof(double), (void **)&st.alpha));
        NV_HIP(c, alloc((o->niter + 3) * sizeof(double), (void **)&st.beta));
        // per-workgroup partial sums: the one-pass H*v and the step kernel write hv_blocks, the
        // two-pass H*v one per workgroup of its second pass
        int64_t nparts = std::max<int64_t>(nv::hv_blocks(c->dev), 8192) + 64;
        if (c->dev.hv2.on) nparts = std::max<int64_t>(nparts, nv::hv2_parts(c->dev) + 64);
        if (c->dev.fb.on) nparts = std::max<int64_t>(nparts, nv::hv_fb_parts(c->dev) + 64);
        NV_HIP(c, alloc(nparts * sizeof(double), (void **)&st.partials));
        NV_HIP(c, alloc(nv::kNumCounters * sizeof(unsigned), (void **)&st.counters));
        // the iteration counter, and a second one that folded graph iterations alternate with
        NV_HIP(c, alloc(2 * sizeof(int64_t), (void **)&c->d_iter));
        st.iter = c->d_iter;
        if (K > 0) {
            c->ncp = (o->niter + K - 1) / K;
            NV_HIP(c, alloc((size_t)c->ncp * 2 * N * sizeof(double), (void **)&c->cp));
        }
    }
    c->opts = *o;
    if (c->opts.chkorthog < 1) c->opts.chkorthog = 100;
    NV_HIP(c, hipMemsetAsync(st.counters, 0, nv::kNumCounters * sizeof(unsigned), c->stream));
    NV_HIP(c, hipMemsetAsync(st.alpha, 0, (o->niter + 3) * sizeof(double), c->stream));
    NV_HIP(c, hipMemsetAsync(st.beta, 0, (o->niter + 3) * sizeof(double), c->stream));
    NV_HIP(c, hipMemsetAsync(st.V, 0, (size_t)std::min<int64_t>(st.nslots, 3) * N * sizeof(double), c->stream));
    if (o->reorthog > 0) {
        hipError_t e = same ? nv::pro_reset(c->pro, c->opts, c->stream)
                            : nv::pro_init(c->pro, c->opts, c->op, c->dev, st, c->lz_allocs, c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "pro_init");
    }
    // compile before any capture; the FP64 generic H*v reads an interleaved copy of q_i
    const bool jit = ensure_jit(c, o->prec) != nullptr;
    st.v64i = nullptr;
    if (o->prec == NV_PREC_F64 && !jit && nv::hv_shaped(c->dev, NV_PREC_F64)) {
        if (!c->v64i) NV_HIP(c, alloc_interleaved(c, sizeof(double), c->allocs, (void **)&c->v64i, c->stream));
        st.v64i = c->v64i;
    }
    st.z2 = nullptr;
    if (use_hv2(c, o->prec) && (o->prec == NV_PREC_F32 || st.v64i)) {
        NV_HIP(c, ensure_z2(c, nv::hv2_scratch_bytes(c->dev, o->prec)));
        st.z2 = c->z2;
    }
    // start vector in the slot of iteration 1, beta(0) = |start|
    double *s1 = st.V + (1 % st.nslots) * N;
    int32_t rc = build_start(c, s1);
    if (rc != NV_OK) return rc;
    NV_HIP(c, nv::launch_norm_to(s1, N, st.beta, st.partials, st.counters + nv::kCntMisc, c->stream));
    if (c->cp) {
        // checkpoint 0: q_0 = 0 and the unnormalised start vector
        NV_HIP(c, hipMemsetAsync(c->cp, 0, (size_t)N * sizeof(double), c->stream));
        NV_HIP(c, hipMemcpyAsync(c->cp + N, s1, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    c->lz_ready = true;
    c->last_iter = 0;
    c->next_update = 1;
    c->archive_iter = 0;
    c->pro_loaded = false;
    return NV_OK;
}

static int32_t check_iter(nv_ctx *c, int64_t it) {
    if (!c) return NV_ERR_ARG;
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "nv_lanczos_init has not been called");
    if (it < 1 || it > c->opts.niter) return fail(c, NV_ERR_ARG, "iteration out of range");
    return NV_OK;
}

int32_t nv_update_lanczos(nv_ctx *c, int64_t it) {
    int32_t rc = check_iter(c, it);
    if (rc) return rc;
    if (it != c->next_update) return fail(c, NV_ERR_STATE, "update_lanczos called out of order");
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, nv::launch_set_iter(c->d_iter, it, 0, c->stream));
    NV_HIP(c, nv::launch_normalize(c->dev, c->opts.prec, c->st, c->stream));
    c->next_update = it + 1;
    return NV_OK;
}

int32_t nv_hv_iter(nv_ctx *c, int64_t it) {
    int32_t rc = check_iter(c, it);
    if (rc) return rc;
    if (it != c->next_update - 1) return fail(c, NV_ERR_STATE, "Hv called before update_lanczos");
    if (c->prof) {
        if (c->ev_used == c->ev.size()) {
            hipEvent_t a, b;
            NV_HIP(c, hipEventCreate(&a));
            NV_HIP(c, hipEventCreate(&b));
            c->ev.emplace_back(a, b);
        }
        NV_HIP(c, hipEventRecord(c->ev[c->ev_used].first, c->stream));
    }
    NV_HIP(c, launch_hv_iter_any(c));
    if (c->prof) {
        NV_HIP(c, hipEventRecord(c->ev[c->ev_used].second, c->stream));
        c->ev_used++;
    }
    return NV_OK;
}

int32_t nv_hv_kernel_info(nv_ctx *c, int32_t prec, char *buf, int64_t len) {
    if (!c || !buf || len < 1) return NV_ERR_ARG;
    NV_HIP(c, hipSetDevice(c->device));
    const nv::JitKernels *k = ensure_jit(c, prec);
    std::string t = k ? "operator-specialised (hiprtc)"
                      : (std::string(nv::hv_shaped(c->dev, prec) ? "generic shaped" : "generic step-table") +
                         (c->jit_log.empty() ? std::string() : ": " + c->jit_log));
    if (!k && use_fb(c, prec)) {
        char fbs[160];
        std::snprintf(fbs, sizeof fbs,
                      "generic fast-block (%d rows per wave over modes %d..%lld; slow steps scalar, fast steps "
                      "through LDS)",
                      (int)c->dev.fb.P, (int)c->dev.fb.mf, (long long)c->op.M - 1);
        t = fbs;
    } else if (!k && use_hv2(c, prec))
        t = std::string("generic two-pass (modes 0 and 1 split off; coupling products on ") +
            (prec == NV_PREC_F32 && c->dev.hv2.mfma ? "MFMA" : "the VALU") + ")";
    else if (!k && nv::hv_shaped(c->dev, prec)) {
        char extra[96];
        std::snprintf(extra, sizeof extra, " [%s; steps inside the basis %.6f]",
                      nv::hv_shaped_skip(c->dev, prec) ? "skip form" : "plain form", c->op.step_valid);
        t += extra;
    }
    std::snprintf(buf, (size_t)len, "%s", t.c_str());
    return NV_OK;
}

int32_t nv_profile_enable(nv_ctx *c, int32_t on) {
    if (!c) return NV_ERR_ARG;
    c->prof = on != 0;
    return NV_OK;
}

int32_t nv_hv_flops(const nv_ctx *c, int64_t *flops) {
    if (!c || !flops) return NV_ERR_ARG;
    *flops = 2 * c->op.nnz_fma;
    return NV_OK;
}

int32_t nv_profile_hv(nv_ctx *c, int64_t launches, double *avg_us) {
    if (!c || !avg_us || launches < 1) return NV_ERR_ARG;
    if (!c->lz_ready || c->next_update < 2) return fail(c, NV_ERR_STATE, "update_lanczos(1) must come first");
    NV_HIP(c, hipSetDevice(c->device));
    const nv::JitKernels *k = ensure_jit(c, c->opts.prec);
    auto one = [&]() {
        return k ? nv::jit_launch_iter(*k, c->dev, c->st, c->stream)
                 : nv::launch_hv_iter_only(c->dev, c->opts.prec, c->st, c->stream);
    };
    hipEvent_t a, b;
    NV_HIP(c, hipEventCreate(&a));
    NV_HIP(c, hipEventCreate(&b));
    hipError_t e = one();  // warm
    if (e == hipSuccess) e = hipEventRecord(a, c->stream);
    for (int64_t i = 0; i < launches && e == hipSuccess; ++i) e = one();
    if (e == hipSuccess) e = hipEventRecord(b, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (e != hipSuccess) return hip_fail(c, e, "profile_hv");
    *avg_us = 1000.0 * ms / (double)launches;
    return NV_OK;
}

int32_t nv_profile_read(nv_ctx *c, double *ms, int64_t *count) {
    if (!c) return NV_ERR_ARG;
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    double t = 0.0;
    for (size_t i = 0; i < c->ev_used; ++i) {
        float f = 0.0f;
        NV_HIP(c, hipEventElapsedTime(&f, c->ev[i].first, c->ev[i].second));
        t += f;
    }
    if (ms) *ms = t;
    if (count) *count = (int64_t)c->ev_used;
    c->ev_used = 0;
    return NV_OK;
}

// After iteration it of a run with checkpoints: keep (q_it, r_it) every K iterations.
static hipError_t save_checkpoint(nv_ctx *c, int64_t it) {
    const int64_t K = c->opts.checkpoint;
    if (!c->cp || it % K != 0 || it / K >= c->ncp) return hipSuccess;
    const int64_t N = c->op.N;
    double *dst = c->cp + (size_t)(it / K) * 2 * N;
    hipError_t e = hipMemcpyAsync(dst, c->st.V + (it % c->st.nslots) * N, N * sizeof(double),
                                  hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst + N, c->st.V + ((it + 1) % c->st.nslots) * N, N * sizeof(double),
                           hipMemcpyDeviceToDevice, c->stream);
    return e;
}

int32_t nv_lanczos_step(nv_ctx *c, int64_t it) {
    int32_t rc = check_iter(c, it);
    if (rc) return rc;
    if (it != c->next_update - 1) return fail(c, NV_ERR_STATE, "lanczos_step called before update_lanczos");
    NV_HIP(c, nv::launch_lanczos_step(c->dev, c->opts.prec, c->st, c->stream));
    c->last_iter = it;
    NV_HIP(c, save_checkpoint(c, it));
    return NV_OK;
}

int32_t nv_set_omega(nv_ctx *c, int64_t it) {
    int32_t rc = check_iter(c, it);
    if (rc) return rc;
    if (c->opts.reorthog <= 0 || it <= 1) return NV_OK;
    const int check = (it % nv::pro_check_period(c->pro)) == 0;
    NV_HIP(c, nv::pro_set_omega(c->pro, c->dev, c->st, check, c->stream));
    return NV_OK;
}

int32_t nv_partial_orthog(nv_ctx *c, int64_t it) {
    int32_t rc = check_iter(c, it);
    if (rc) return rc;
    if (c->opts.reorthog <= 0 || it <= 1) return NV_OK;
    NV_HIP(c, nv::pro_partial_orthog(c->pro, c->dev, c->st, c->stream));
    return NV_OK;
}

// The kernels of one iteration, in the reference's order; the iteration
// number is advanced on the device so that a captured sequence replays for
// every iteration.
// One iteration on the stream.  set: advance the device counter first (a
// separate launch); advance: the beta finish advances it for the next
// iteration instead (runs of several iterations in one graph).
static hipError_t enqueue_iteration(nv_ctx *c, int pro, int check, bool set = true, bool advance = false) {
    hipError_t e = set ? nv::launch_set_iter(c->d_iter, 0, 1, c->stream) : hipSuccess;
    if (e == hipSuccess) e = nv::launch_normalize(c->dev, c->opts.prec, c->st, c->stream);
    if (e == hipSuccess) e = launch_hv_iter_any(c);
    if (pro && !check && nv::pro_fusable(c->pro)) {
        // beta, the omega recurrence and the range decision in one workgroup
        if (e == hipSuccess) e = nv::launch_lanczos_step_only(c->dev, c->opts.prec, c->st, c->stream);
        if (e == hipSuccess) e = nv::pro_step_fused(c->pro, c->dev, c->st, nv::hv_blocks(c->dev), c->stream);
        return e;
    }
    if (e == hipSuccess)
        e = nv::launch_lanczos_step(c->dev, c->opts.prec, c->st, c->stream, advance ? c->d_iter : nullptr);
    if (pro) {
        if (e == hipSuccess) e = nv::pro_set_omega(c->pro, c->dev, c->st, check, c->stream);
        if (e == hipSuccess) e = nv::pro_partial_orthog(c->pro, c->dev, c->st, c->stream);
    }
    return e;
}

// Iterations per graph launch in runs without re-orthogonalisation (graph kind
// 3): the host launch and the gaps around it were about a third of a 68 us
// iteration on a 1.38M-element operator with one iteration per graph.
constexpr int64_t kIterBatch = 16;

// Up to kFoldMax partials per sum, graph kind 3 leaves out the two finishing
// kernels of an iteration: the step kernel sums the H*v's alpha partials
// itself (every workgroup, in the finishing kernel's order) and stores alpha,
// the next iteration's normalisation sums the step's beta partials and stores
// beta.  The iteration counter then cannot advance between kernels, so
// iterations alternate between two counters: the step of iteration k sets the
// counter that iteration k+1 reads (c->d_iter for even k, c->d_iter + 1 for
// odd).  The first iteration of a launch normalises with the stored beta, and
// the last one's beta goes through the finishing kernel, so a launch starts
// and ends in the state the per-call path leaves.
constexpr int kFoldMax = 4096;

// The iteration H*v without its alpha finish, and the partials it writes.
static hipError_t launch_hv_iter_nofinish(nv_ctx *c, const nv::LzState &st) {
    const nv::JitKernels *k = ensure_jit(c, c->opts.prec);
    if (k) return nv::jit_launch_iter(*k, c->dev, st, c->stream);
    return nv::launch_hv_iter_only(c->dev, c->opts.prec, st, c->stream);
}
static int hv_iter_nparts(nv_ctx *c) {
    return ensure_jit(c, c->opts.prec) ? nv::hv_blocks(c->dev) : nv::hv_iter_parts(c->dev, c->opts.prec, c->st);
}
static bool fold_ok(nv_ctx *c) {
    return hv_iter_nparts(c) <= kFoldMax && nv::hv_blocks(c->dev) <= kFoldMax;
}

static hipError_t enqueue_folded_batch(nv_ctx *c) {
    hipStream_t s = c->stream;
    int64_t *ctr[2] = {c->d_iter, c->d_iter + 1};
    double *bparts = c->st.partials + kFoldMax;  // the partials buffer holds at least 8192 + 64
    const int nhv = hv_iter_nparts(c), nstep = nv::hv_blocks(c->dev);
    hipError_t e = nv::launch_set_iter(ctr[0], 0, 1, s);
    for (int64_t k = 0; k < kIterBatch && e == hipSuccess; ++k) {
        nv::LzState st = c->st;
        st.iter = ctr[k & 1];
        nv::LzState sn = st;
        if (k > 0) {
            sn.fold_in = bparts;
            sn.fold_n = nstep;
        }
        e = nv::launch_normalize(c->dev, c->opts.prec, sn, s);
        if (e == hipSuccess) e = launch_hv_iter_nofinish(c, st);
        const bool last = k + 1 == kIterBatch;
        nv::LzState ss = st;
        ss.fold_in = st.partials;
        ss.fold_n = nhv;
        ss.fold_out = bparts;
        ss.next_iter = last ? ctr[0] : ctr[(k + 1) & 1];
        ss.next_inc = last ? 0 : 1;
        if (e == hipSuccess) e = nv::launch_lanczos_step_only(c->dev, c->opts.prec, ss, s);
        if (e == hipSuccess && last) e = nv::launch_finish_sum(bparts, nstep, c->st.beta, st.iter, 1, s);
    }
    return e;
}

static hipError_t iteration_graph(nv_ctx *c, int kind, hipGraphExec_t *out) {
    if (c->graphs[kind]) {
        *out = c->graphs[kind];
        return hipSuccess;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t e2 = hipSuccess;
    if (kind == 3 && fold_ok(c))
        e2 = enqueue_folded_batch(c);
    else if (kind == 3)
        for (int64_t k = 0; k < kIterBatch && e2 == hipSuccess; ++k)
            e2 = enqueue_iteration(c, 0, 0, k == 0, k + 1 < kIterBatch);
    else
        e2 = enqueue_iteration(c, kind >= 1, kind == 2);
    e = hipStreamEndCapture(c->stream, &g);
    if (e2 != hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&c->graphs[kind], g, nullptr, nullptr, 0);
    if (g) hipGraphDestroy(g);
    *out = c->graphs[kind];
    return e;
}

int32_t nv_lanczos_run(nv_ctx *c, int64_t iiter, int64_t niter) {
    if (!c) return NV_ERR_ARG;
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "nv_lanczos_init has not been called");
    if (iiter != c->next_update) return fail(c, NV_ERR_STATE, "run must start at the next iteration");
    if (niter > c->opts.niter) return fail(c, NV_ERR_ARG, "niter beyond the initialised run");
    NV_HIP(c, hipSetDevice(c->device));
    const bool orthog = c->opts.reorthog > 0;
    for (int64_t i = iiter; i <= niter; ++i) {
        // the first iteration of a call and profiled runs go through the
        // per-call path, which sets the device iteration counter that the
        // captured graphs then advance (a resumed run starts past 1)
        if (i == iiter || c->prof) {
            int32_t rc = nv_update_lanczos(c, i);
            if (!rc) rc = nv_hv_iter(c, i);
            if (!rc) rc = nv_lanczos_step(c, i);
            if (!rc && orthog && i > 1) {
                rc = nv_set_omega(c, i);
                if (!rc) rc = nv_partial_orthog(c, i);
            }
            if (rc) return rc;
            continue;
        }
        if (!orthog && !c->cp && i + kIterBatch - 1 <= niter) {
            // kIterBatch iterations in one launch (the device counter advances in each)
            hipGraphExec_t gb = nullptr;
            NV_HIP(c, iteration_graph(c, 3, &gb));
            NV_HIP(c, hipGraphLaunch(gb, c->stream));
            i += kIterBatch - 1;
            c->next_update = i + 1;
            c->last_iter = i;
            continue;
        }
        int kind = 0;
        if (orthog) kind = (c->opts.reorthog == 1 && i % nv::pro_check_period(c->pro) == 0) ? 2 : 1;
        hipGraphExec_t ge = nullptr;
        NV_HIP(c, iteration_graph(c, kind, &ge));
        NV_HIP(c, hipGraphLaunch(ge, c->stream));
        c->next_update = i + 1;
        c->last_iter = i;
        NV_HIP(c, save_checkpoint(c, i));
    }
    return NV_OK;
}

int32_t nv_get_coefficients(nv_ctx *c, double *alpha, double *beta, double *omega) {
    if (!c) return NV_ERR_ARG;
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "no Lanczos run");
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t n = c->last_iter;
    if (alpha && n > 0) NV_HIP(c, hipMemcpy(alpha, c->st.alpha + 1, n * sizeof(double), hipMemcpyDeviceToHost));
    if (beta) NV_HIP(c, hipMemcpy(beta, c->st.beta, (n + 1) * sizeof(double), hipMemcpyDeviceToHost));
    if (omega) {
        if (c->opts.reorthog <= 0) return fail(c, NV_ERR_STATE, "omega is only kept with reorthog > 0");
        hipError_t e = nv::pro_get_omega(c->pro, omega, n);
        if (e != hipSuccess) return hip_fail(c, e, "omega copy");
    }
    return NV_OK;
}

int32_t nv_pro_stats(const nv_ctx *c, int64_t *events, int64_t *touched) {
    if (!c) return NV_ERR_ARG;
    int64_t ev = 0, tv = 0;
    if (c->lz_ready && c->opts.reorthog > 0) {
        hipStreamSynchronize(c->stream);
        nv::pro_stats(c->pro, &ev, &tv);
    }
    if (events) *events = ev;
    if (touched) *touched = tv;
    return NV_OK;
}

int32_t nv_get_lanczos_vector(nv_ctx *c, int64_t j, double *q) {
    if (!c || !q) return NV_ERR_ARG;
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "no Lanczos run");
    const int64_t last = c->last_iter;
    if (j < 1 || j > last + 1) return fail(c, NV_ERR_ARG, "vector index out of range");
    if (c->st.nslots == 3 && j < last - 1) return fail(c, NV_ERR_STATE, "Krylov basis not kept (reorthog = 0)");
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    const double *src = c->st.V + (j % c->st.nslots) * c->op.N;
    NV_HIP(c, hipMemcpy(q, src, c->op.N * sizeof(double), hipMemcpyDeviceToHost));
    return NV_OK;
}

// ---- tridiagonal eigensolve ---------------------------------------------
// Eigenvalues (ascending) and eigenvectors of the tridiagonal T(alpha,
// beta(1:n-1)) on the device.  *vec_dev (n*n doubles, vec[j*n + i] =
// component j of root i) belongs to the caller, who frees it with hipFree.
static int32_t tridiag_solve(nv_ctx *c, int64_t n, const double *alpha, const double *beta,
                             std::vector<double> &root, double **vec_dev) {
    *vec_dev = nullptr;
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(alpha[i]) || !std::isfinite(beta[i + 1]))
            return fail(c, NV_ERR_ARG, "alpha/beta hold a non-finite value at iteration " + std::to_string(i + 1));
    NV_HIP(c, hipSetDevice(c->device));
    root.assign(n, 0.0);
    // Frobenius scaling in the order GIVENS accumulates it (Source/yarkony.f90:221-247)
    double factor = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        factor = std::fmax(factor, std::fabs(alpha[i]));
        if (i > 0) factor = std::fmax(factor, std::fabs(beta[i]));
    }
    double *vec = nullptr;
    NV_HIP(c, hipMalloc(&vec, (size_t)n * n * sizeof(double)));
    if (n == 1 || factor == 0.0) {
        std::vector<double> eye((size_t)n * n, 0.0);
        for (int64_t i = 0; i < n; ++i) {
            eye[(size_t)i * n + i] = 1.0;
            root[i] = n == 1 ? alpha[0] : 0.0;
        }
        hipError_t er = hipMemcpy(vec, eye.data(), (size_t)n * n * sizeof(double), hipMemcpyHostToDevice);
        if (er != hipSuccess) {
            hipFree(vec);
            return hip_fail(c, er, "tridiagonal eigenvectors");
        }
        *vec_dev = vec;
        return NV_OK;
    }
    double anorm = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (i > 0) anorm += (beta[i] / factor) * (beta[i] / factor);
        anorm += (alpha[i] / factor) * (alpha[i] / factor) / 2.0;
    }
    anorm = std::sqrt(anorm * 2.0) * factor;
    std::vector<double> d(n), e(n, 0.0), e2(n, 0.0);
    for (int64_t i = 0; i < n; ++i) d[i] = alpha[i] / anorm;
    for (int64_t i = 0; i + 1 < n; ++i) { e[i] = beta[i + 1] / anorm; e2[i] = e[i] * e[i]; }
    nv::TriWork w;
    w.vec = vec;
    std::vector<void *> tmp;
    auto alloc = [&](size_t bytes, void **p) -> hipError_t {
        hipError_t er = hipMalloc(p, bytes);
        if (er == hipSuccess) tmp.push_back(*p);
        return er;
    };
    hipError_t er = alloc(n * 8, (void **)&w.d);
    if (er == hipSuccess) er = alloc(n * 8, (void **)&w.e);
    if (er == hipSuccess) er = alloc(n * 8, (void **)&w.e2);
    if (er == hipSuccess) er = alloc(n * 8, (void **)&w.root);
    if (er == hipSuccess) er = alloc((size_t)4 * n * n * 8, (void **)&w.lu);
    if (er == hipSuccess) er = alloc(n * 4, (void **)&w.ia);
    if (er == hipSuccess) er = alloc(n * 8, (void **)&w.rank);
    if (er == hipSuccess) er = hipMemcpyAsync(w.d, d.data(), n * 8, hipMemcpyHostToDevice, c->stream);
    if (er == hipSuccess) er = hipMemcpyAsync(w.e, e.data(), n * 8, hipMemcpyHostToDevice, c->stream);
    if (er == hipSuccess) er = hipMemcpyAsync(w.e2, e2.data(), n * 8, hipMemcpyHostToDevice, c->stream);
    if (er == hipSuccess) er = nv::launch_tridiag(n, w, anorm, c->stream);
    if (er == hipSuccess) er = hipMemcpyAsync(root.data(), w.root, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (er == hipSuccess) er = hipStreamSynchronize(c->stream);
    for (void *p : tmp) hipFree(p);
    if (er != hipSuccess) {
        hipFree(vec);
        return hip_fail(c, er, "tridiagonal eigensolve");
    }
    for (int64_t i = 0; i < n; ++i) root[i] *= anorm;
    *vec_dev = vec;
    return NV_OK;
}

int32_t nv_ritz(nv_ctx *c, int64_t n, const double *alpha, const double *beta, double *eval, double *inten,
                double *bji, int32_t *sign) {
    if (!c) return NV_ERR_ARG;
    if (n < 1 || !alpha || !beta || !eval) return fail(c, NV_ERR_ARG, "bad tridiagonal arguments");
    std::vector<double> root, v1(n), vn(n);
    double *vec = nullptr;
    int32_t rc = tridiag_solve(c, n, alpha, beta, root, &vec);
    if (rc != NV_OK) return rc;
    hipError_t er = hipMemcpyAsync(v1.data(), vec, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (er == hipSuccess) er = hipMemcpyAsync(vn.data(), vec + (n - 1) * n, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (er == hipSuccess) er = hipStreamSynchronize(c->stream);
    hipFree(vec);
    if (er != hipSuccess) return hip_fail(c, er, "tridiagonal eigenvectors");
    const double bn = beta[n];
    for (int64_t i = 0; i < n; ++i) {
        eval[i] = root[i];
        if (inten) inten[i] = v1[i] * v1[i];
        if (bji) bji[i] = std::fabs(bn * vn[i]);
        if (sign) sign[i] = v1[i] < 0 ? -1 : 1;
    }
    return NV_OK;
}

// ---- Ritz vectors and root analysis -------------------------------------------
int32_t nv_identify_roots(nv_ctx *c, int64_t n, const double *alpha, const double *beta, int64_t nroots,
                          int64_t idroots, double *eval, double *inten, int32_t *sign, int64_t *top_index,
                          int32_t *top_state, int32_t *top_phase, double *top_mag2, double *percent, double *overlap,
                          double *vectors) {
    if (!c) return NV_ERR_ARG;
    if (!alpha || !beta) return fail(c, NV_ERR_ARG, "alpha/beta missing");
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "nv_lanczos_init has not been called");
    if (c->opts.reorthog <= 0 && c->opts.savevecs == 0)
        return fail(c, NV_ERR_STATE, "Krylov basis not kept (set savevecs or reorthog)");
    if (n < 1 || n > c->last_iter) return fail(c, NV_ERR_ARG, "n must lie in 1..last iteration");
    if (nroots < 1 || nroots > n) return fail(c, NV_ERR_ARG, "nroots must lie in 1..n");
    const int64_t ns = c->op.ns, D = c->op.D, N = c->op.N;
    if (nroots * ns > 4096) return fail(c, NV_ERR_ARG, "nroots * nstates must not exceed 4096");
    if (idroots < 0 || idroots > N) return fail(c, NV_ERR_ARG, "idroots must lie in 0..n");
    if (idroots > 0 && (!top_index || !top_state || !top_phase || !top_mag2))
        return fail(c, NV_ERR_ARG, "top_* outputs missing");
    std::vector<double> root;
    double *vec = nullptr;
    int32_t rc = tridiag_solve(c, n, alpha, beta, root, &vec);
    if (rc != NV_OK) return rc;
    std::vector<void *> tmp{vec};
    auto alloc = [&](size_t bytes, void **p) -> hipError_t {
        hipError_t er = hipMalloc(p, bytes);
        if (er == hipSuccess) tmp.push_back(*p);
        return er;
    };
    const int R = (int)(nroots * ns);
    int npairs = 0, nchunks = 0;
    int64_t rows = 0;
    nv::gram_shape(D, R, npairs, nchunks, rows);
    const int nb = nv::topk_blocks(N);
    const int64_t kk = idroots > 0 ? idroots : 1;
    double *Y = nullptr, *G = nullptr, *part = nullptr, *pm = nullptr, *prev_m = nullptr, *om = nullptr;
    int64_t *pi = nullptr, *prev_i = nullptr, *oi = nullptr;
    int32_t *oph = nullptr;
    std::vector<double> v1(n), gram((size_t)R * R);
    // q_1 .. q_n are slots 1..n of the archive (nslots = niter + 2)
    const double *V = c->st.V + N;
    hipError_t er = alloc((size_t)nroots * N * 8, (void **)&Y);
    if (er == hipSuccess) er = alloc((size_t)R * R * 8, (void **)&G);
    if (er == hipSuccess) er = alloc((size_t)nchunks * npairs * nv::kGramT * nv::kGramT * 8, (void **)&part);
    if (er == hipSuccess) er = alloc((size_t)nroots * nb * 8, (void **)&pm);
    if (er == hipSuccess) er = alloc((size_t)nroots * nb * 8, (void **)&pi);
    if (er == hipSuccess) er = alloc((size_t)nroots * 8, (void **)&prev_m);
    if (er == hipSuccess) er = alloc((size_t)nroots * 8, (void **)&prev_i);
    if (er == hipSuccess) er = alloc((size_t)nroots * kk * 8, (void **)&om);
    if (er == hipSuccess) er = alloc((size_t)nroots * kk * 8, (void **)&oi);
    if (er == hipSuccess) er = alloc((size_t)nroots * kk * 4, (void **)&oph);
    if (er == hipSuccess) er = hipMemcpyAsync(v1.data(), vec, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (er == hipSuccess) er = nv::launch_ritz_vectors(V, N, n, vec, nroots, Y, c->stream);
    if (er == hipSuccess && idroots > 0)
        er = nv::launch_topk(Y, N, (int)nroots, (int)idroots, prev_m, prev_i, pm, pi, om, oi, oph, c->stream);
    if (er == hipSuccess) er = nv::launch_gram(Y, D, R, part, G, c->stream);
    if (er == hipSuccess) er = hipMemcpyAsync(gram.data(), G, (size_t)R * R * 8, hipMemcpyDeviceToHost, c->stream);
    std::vector<double> hm((size_t)nroots * kk);
    std::vector<int64_t> hi((size_t)nroots * kk);
    std::vector<int32_t> hph((size_t)nroots * kk);
    if (er == hipSuccess && idroots > 0) {
        er = hipMemcpyAsync(hm.data(), om, hm.size() * 8, hipMemcpyDeviceToHost, c->stream);
        if (er == hipSuccess) er = hipMemcpyAsync(hi.data(), oi, hi.size() * 8, hipMemcpyDeviceToHost, c->stream);
        if (er == hipSuccess) er = hipMemcpyAsync(hph.data(), oph, hph.size() * 4, hipMemcpyDeviceToHost, c->stream);
    }
    if (er == hipSuccess && vectors)
        er = hipMemcpyAsync(vectors, Y, (size_t)nroots * N * 8, hipMemcpyDeviceToHost, c->stream);
    if (er == hipSuccess) er = hipStreamSynchronize(c->stream);
    for (void *p : tmp) hipFree(p);
    if (er != hipSuccess) return hip_fail(c, er, "identify_roots");
    // intensities relative to the strongest of the nroots roots (Source/top.f90:1165-1168)
    double intfactor = 0.0;
    for (int64_t i = 0; i < nroots; ++i) intfactor = std::fmax(intfactor, v1[i] * v1[i]);
    for (int64_t i = 0; i < nroots; ++i) {
        if (eval) eval[i] = root[i];
        if (inten) inten[i] = intfactor > 0.0 ? v1[i] * v1[i] / intfactor : 0.0;
        if (sign) sign[i] = v1[i] < 0 ? -1 : 1;
    }
    for (int64_t i = 0; i < nroots && idroots > 0; ++i)
        for (int64_t k = 0; k < idroots; ++k) {
            const size_t o = (size_t)i * idroots + k;
            const int64_t e = hi[o];
            top_index[o] = e >= 0 ? e % D + 1 : 0;
            top_state[o] = e >= 0 ? (int32_t)(e / D + 1) : 0;
            top_phase[o] = hph[o];
            top_mag2[o] = hm[o];
        }
    if (percent)
        for (int64_t i = 0; i < nroots; ++i) {
            double tot = 0.0;
            for (int64_t k = 0; k < ns; ++k) tot += gram[(size_t)(i * ns + k) * R + i * ns + k];
            for (int64_t k = 0; k < ns; ++k)
                percent[i * ns + k] = 100.0 * std::fabs(gram[(size_t)(i * ns + k) * R + i * ns + k] / tot);
        }
    if (overlap)
        for (int64_t i = 0; i < nroots; ++i)
            for (int64_t j = 0; j < nroots; ++j)
                for (int64_t k = 0; k < ns; ++k)
                    for (int64_t l = 0; l < ns; ++l)
                        overlap[((i * nroots + j) * ns + k) * ns + l] = gram[(size_t)(i * ns + k) * R + j * ns + l];
    return NV_OK;
}

// ---- derivatives -----------------------------------------------------------------
// The gradient operator: every monomial of the basis with a unit coefficient
// in every block, so that its term list covers each nadvibs.in slot.
static int32_t ensure_gop(nv_ctx *c) {
    if (c->gop_ready) return NV_OK;
    const nv::Operator &op = c->op;
    std::vector<double> ones((size_t)c->ntmax * op.ordr * op.nstblks, 1.0), cc(op.nstblks, 1.0);
    nv_problem pr{};
    pr.nstates = op.ns; pr.nmodes = op.M; pr.ordr = op.ordr;
    pr.nfunc = op.nfunc.data(); pr.aomega = op.omega.data(); pr.concoef = cc.data(); pr.potterms = ones.data();
    pr.ztoler = 0.0; pr.quirks = op.quirks; pr.neworigin = 0;
    std::string err;
    if (!nv::build_operator(pr, c->gop, err)) return fail(c, NV_ERR_ARG, "gradient operator: " + err);
    const nv::Operator &g = c->gop;
    NvOpDev &d = c->gdev;
    d = NvOpDev{};
    d.ns = (int32_t)g.ns; d.M = (int32_t)g.M; d.D = (int32_t)g.D; d.N = (int32_t)g.N;
    d.nmax = (int32_t)g.nmax; d.ncodes = g.ncodes;
    d.nterms = (int32_t)g.dev_terms.size(); d.ndterms = (int32_t)g.dev_dterms.size();
    std::vector<int32_t> nf32(g.nfunc.begin(), g.nfunc.end());
    hipError_t e = upload(c, g.dev_terms, &d.terms, c->gallocs);
    if (e == hipSuccess) e = upload(c, g.dev_steps, &d.steps, c->gallocs);
    if (e == hipSuccess) e = upload(c, g.ftab, &d.ftab64, c->gallocs);
    if (e == hipSuccess) e = upload(c, g.dev_dterms, &d.dterms, c->gallocs);
    if (e == hipSuccess) e = upload(c, g.omega, &d.omega, c->gallocs);
    if (e == hipSuccess) e = upload(c, nf32, &d.nfunc, c->gallocs);
    if (e != hipSuccess) return hip_fail(c, e, "gradient operator upload");
    if (nv::grad_lds_bytes(d) > 64 * 1024) return fail(c, NV_ERR_ARG, "gradient factor table exceeds the LDS budget");
    c->gop_ready = true;
    return NV_OK;
}

// Maps the accumulated bilinear forms onto the nadvibs.in slots and the
// frequencies (chain rule through c * prod_k w_k^{-p_k/2}, Source/top.f90:212-217).
static void map_gradients(nv_ctx *c, const std::vector<double> &gacc, double *g_omega, double *g_concoef,
                          double *g_pot) {
    const nv::Operator &op = c->op, &g = c->gop;
    const int ns = (int)op.ns, nss = ns * ns;
    const int nT = c->gdev.nterms, nD = c->gdev.ndterms;
    const bool lower = (op.quirks & NV_QUIRK_LOWERDIAG) != 0;
    std::vector<int> bi(op.nstblks), bj(op.nstblks);
    for (int i = 1; i <= ns; ++i)
        for (int j = 1; j <= i; ++j) { bi[i * (i - 1) / 2 + j - 1] = i - 1; bj[i * (i - 1) / 2 + j - 1] = j - 1; }
    auto GT = [&](int t, int a, int b) { return gacc[(size_t)t * nss + a * ns + b]; };
    auto GD = [&](int t, int a, int b) { return gacc[(size_t)(nT + t) * nss + a * ns + b]; };
    auto GC = [&](int a, int b) { return gacc[(size_t)(nT + nD + op.M) * nss + a * ns + b]; };
    std::vector<double> gw(op.M, 0.0);
    for (int64_t k = 0; k < op.M; ++k) gw[k] = gacc[(size_t)(nT + nD + k) * nss];
    // signature -> gradient-operator term
    std::map<std::vector<int>, int> sig;
    for (size_t t = 0; t < g.terms.size(); ++t) {
        std::vector<int> key;
        for (int j = 0; j < g.terms[t].cnt; ++j) { key.push_back(g.terms[t].mode[j]); key.push_back(g.terms[t].pow[j]); }
        sig[key] = (int)t;
    }
    for (int64_t p = 1; p <= op.ordr; ++p) {
        auto mons = nv::monomials(op.M, p);
        for (size_t k = 0; k < mons.size(); ++k) {
            int um[NV_MAXF], up[NV_MAXF], cnt = 0;
            for (int v : mons[k]) {
                int f = -1;
                for (int j = 0; j < cnt; ++j) if (um[j] == v) f = j;
                if (f < 0) { um[cnt] = v; up[cnt] = 1; ++cnt; } else up[f]++;
            }
            std::vector<int> key;
            for (int j = 0; j < cnt; ++j) { key.push_back(um[j]); key.push_back(up[j]); }
            const int t = sig.count(key) ? sig[key] : -1;
            const int dt = t >= 0 ? g.term_dev[t] : -1, dd = t >= 0 ? g.term_ddev[t] : -1;
            const double dval = t >= 0 ? g.terms[t].dval : 0.0;
            for (int64_t b = 0; b < op.nstblks; ++b) {
                const int i = bi[b], j = bj[b];
                double v = 0.0;
                if (dt >= 0) v += GT(dt, i, j) + (i != j ? GT(dt, j, i) : 0.0);
                if (dd >= 0) v += GD(dd, i, j) + (i != j && !lower ? GD(dd, j, i) : 0.0);
                v *= dval;
                const size_t slot = k + (p - 1) * c->ntmax + b * c->ntmax * op.ordr;
                if (g_pot) g_pot[slot] = v;
                const double craw = c->potraw[slot];
                if (std::fabs(craw) >= c->ztoler)
                    for (int jj = 0; jj < cnt; ++jj) gw[um[jj]] += -0.5 * up[jj] / op.omega[um[jj]] * craw * v;
            }
        }
    }
    if (g_concoef)
        for (int64_t b = 0; b < op.nstblks; ++b) {
            const int i = bi[b], j = bj[b];
            g_concoef[b] = i == j ? GC(i, i) : GC(i, j) + (lower ? 0.0 : GC(j, i));
        }
    if (g_omega)
        for (int64_t k = 0; k < op.M; ++k) g_omega[k] = gw[k];
}

// Recomputes q_{s0-1} .. q_{s1} and r_{s1} of the stored run (s0 = sg*K + 1,
// s1 = min((sg+1)*K, n)) from checkpoint sg into seg, q_j in slot j % (K+2),
// with the forward's own kernels, and checks the recomputed alpha/beta bit
// for bit against the stored ones.  ta, tb: niter + 3 doubles of scratch.
static int32_t recompute_segment(nv_ctx *c, int64_t sg, int64_t K, int64_t n, double *seg, double *ta, double *tb) {
    const int64_t N = c->op.N, niter = c->opts.niter;
    const int64_t s0 = sg * K + 1, s1 = std::min((sg + 1) * K, n);
    hipStream_t s = c->stream;
    nv::LzState ts = c->st;
    ts.V = seg;
    ts.nslots = K + 2;
    ts.alpha = ta;
    ts.beta = tb;
    hipError_t e = hipMemcpyAsync(ta, c->st.alpha, (niter + 3) * 8, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(tb, c->st.beta, (niter + 3) * 8, hipMemcpyDeviceToDevice, s);
    const double *cpk = c->cp + (size_t)sg * 2 * N;
    if (e == hipSuccess) e = hipMemcpyAsync(seg + ((s0 - 1) % (K + 2)) * N, cpk, N * 8, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(seg + (s0 % (K + 2)) * N, cpk + N, N * 8, hipMemcpyDeviceToDevice, s);
    for (int64_t j = s0; j <= s1 && e == hipSuccess; ++j) {
        e = nv::launch_set_iter(c->d_iter, j, 0, s);
        if (e == hipSuccess) e = nv::launch_normalize(c->dev, c->opts.prec, ts, s);
        if (e == hipSuccess) e = launch_hv_iter_st(c, ts);
        if (e == hipSuccess) e = nv::launch_lanczos_step(c->dev, c->opts.prec, ts, s);
    }
    std::vector<double> ra(s1 - s0 + 1), rb(s1 - s0 + 1), oa(s1 - s0 + 1), ob(s1 - s0 + 1);
    if (e == hipSuccess) e = hipMemcpyAsync(ra.data(), ta + s0, ra.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(rb.data(), tb + s0, rb.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(oa.data(), c->st.alpha + s0, oa.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ob.data(), c->st.beta + s0, ob.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "checkpoint recomputation");
    if (std::memcmp(ra.data(), oa.data(), ra.size() * 8) != 0 || std::memcmp(rb.data(), ob.data(), rb.size() * 8) != 0)
        return fail(c, NV_ERR_STATE, "checkpoint recomputation differs from the stored run at iterations " +
                                         std::to_string(s0) + ".." + std::to_string(s1));
    return NV_OK;
}

int32_t nv_lanczos_adjoint(nv_ctx *c, int64_t n, const double *abar, const double *bbar, int32_t prec,
                           double *g_omega, double *g_concoef, double *g_potterms) {
    if (!c) return NV_ERR_ARG;
    if (!abar || !bbar) return fail(c, NV_ERR_ARG, "abar/bbar missing");
    if (prec != NV_PREC_F64 && prec != NV_PREC_F32) return fail(c, NV_ERR_ARG, "adjoint H*v supports F64 and F32");
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "nv_lanczos_init has not been called");
    if (c->opts.reorthog > 0)
        return fail(c, NV_ERR_STATE, "the adjoint differentiates the plain recurrence (reorthog = 0)");
    if (c->op.neworigin) return fail(c, NV_ERR_STATE, "the adjoint supports the default start vector only");
    if (n < 1 || n > c->last_iter) return fail(c, NV_ERR_ARG, "n must lie in 1..last iteration");
    const bool archive = c->opts.savevecs != 0;
    if (!archive && !c->cp) return fail(c, NV_ERR_STATE, "the adjoint needs savevecs or checkpoints");
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    int32_t rc = ensure_gop(c);
    if (rc != NV_OK) return rc;
    const int64_t N = c->op.N, niter = c->opts.niter;
    const int64_t K = archive ? n : c->opts.checkpoint;
    const int nss = (int)(c->op.ns * c->op.ns);
    const int groups = nv::grad_groups(c->gdev), chunks = nv::grad_chunks(c->gdev);
    std::vector<double> beta_h(niter + 3);
    NV_HIP(c, hipMemcpy(beta_h.data(), c->st.beta, (niter + 3) * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<void *> tmp;
    auto alloc = [&](size_t bytes, void **p) -> hipError_t {
        hipError_t er = hipMalloc(p, bytes);
        if (er == hipSuccess) tmp.push_back(*p);
        return er;
    };
    double *qb[3], *ub = nullptr, *hub = nullptr, *part = nullptr, *gpart = nullptr, *scal = nullptr;
    double *abar_d = nullptr, *bbar_d = nullptr, *gacc = nullptr, *seg = nullptr, *ta = nullptr, *tb = nullptr;
    hipError_t e = alloc((size_t)3 * N * 8, (void **)&qb[0]);
    qb[1] = qb[0] + N;
    qb[2] = qb[0] + 2 * N;
    if (e == hipSuccess) e = alloc((size_t)N * 8, (void **)&ub);
    if (e == hipSuccess) e = alloc((size_t)N * 8, (void **)&hub);
    if (e == hipSuccess) e = alloc((size_t)5 * nv::adj_blocks(N) * 8, (void **)&part);
    if (e == hipSuccess) e = alloc((size_t)groups * chunks * nss * 8, (void **)&gpart);
    if (e == hipSuccess) e = alloc(8 * 8, (void **)&scal);
    if (e == hipSuccess) e = alloc((size_t)(n + 1) * 8, (void **)&abar_d);
    if (e == hipSuccess) e = alloc((size_t)(n + 2) * 8, (void **)&bbar_d);
    if (e == hipSuccess) e = alloc((size_t)groups * nss * 8, (void **)&gacc);
    if (!archive) {
        if (e == hipSuccess) e = alloc((size_t)(K + 2) * N * 8, (void **)&seg);
        if (e == hipSuccess) e = alloc((size_t)(niter + 3) * 8, (void **)&ta);
        if (e == hipSuccess) e = alloc((size_t)(niter + 3) * 8, (void **)&tb);
    }
    auto done = [&](int32_t code) {
        for (void *p : tmp) hipFree(p);
        return code;
    };
    if (e != hipSuccess) return done(hip_fail(c, e, "adjoint allocation"));
    hipStream_t s = c->stream;
    e = hipMemcpyAsync(abar_d, abar, n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(bbar_d, bbar, (n + 1) * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(qb[0], 0, (size_t)3 * N * 8, s);
    if (e == hipSuccess) e = hipMemsetAsync(gacc, 0, (size_t)groups * nss * 8, s);
    int ia = 0, ib = 1, ic = 2;  // qbar_{i+1}, qbar_i, qbar_{i-1}
    bool have_hub = false;
    const int64_t nseg = (n + K - 1) / K;
    for (int64_t sg = nseg - 1; sg >= 0 && e == hipSuccess; --sg) {
        const int64_t s0 = sg * K + 1, s1 = std::min((sg + 1) * K, n);
        const double *V = c->st.V;
        int64_t nslots = c->st.nslots;
        bool end_unnormalised = s1 == c->last_iter;
        if (!archive) {
            const int32_t rc = recompute_segment(c, sg, K, n, seg, ta, tb);
            if (rc != NV_OK) return done(rc);
            V = seg;
            nslots = K + 2;
            end_unnormalised = true;
        }
        for (int64_t i = s1; i >= s0 && e == hipSuccess; --i) {
            const double *qi = V + (i % nslots) * N, *qn = V + ((i + 1) % nslots) * N;
            const double *qp = i > 1 ? V + ((i - 1) % nslots) * N : nullptr;
            const double sn = (i == s1 && end_unnormalised) ? 1.0 / beta_h[i] : 1.0;
            e = nv::launch_adj_step(N, i, qb[ia], have_hub ? hub : nullptr, qn, sn, qi, qp, c->st.alpha, c->st.beta,
                                    abar_d, bbar_d, part, scal, ub, qb[ib], qb[ic], s);
            if (e == hipSuccess) e = hv_apply(c, prec, ub, hub, s);
            if (e == hipSuccess && c->dev.offstate) e = nv::launch_adj_transpose(c->dev, ub, hub, s);
            if (e == hipSuccess) e = nv::launch_grad_contract(c->gdev, ub, qi, gpart, gacc, s);
            if (e == hipSuccess) e = nv::launch_adj_zero(N, qb[ia], s);
            have_hub = true;
            const int t = ia;
            ia = ib; ib = ic; ic = t;
        }
    }
    std::vector<double> gh((size_t)groups * nss);
    if (e == hipSuccess) e = hipMemcpyAsync(gh.data(), gacc, gh.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return done(hip_fail(c, e, "adjoint"));
    map_gradients(c, gh, g_omega, g_concoef, g_potterms);
    return done(NV_OK);
}

static int32_t spectrum_common(nv_ctx *c, int64_t n, const double *alpha, const double *beta, int64_t ng,
                               const double *grid, double sigma, int32_t origin, double e0, const double *gbar,
                               double *env, double *abar, double *bbar, double *tbar) {
    if (!c) return NV_ERR_ARG;
    if (n < 1 || ng < 1 || !alpha || !beta || !grid || !(sigma > 0) || (origin != 0 && origin != 1))
        return fail(c, NV_ERR_ARG, "bad spectrum arguments");
    std::vector<double> root;
    double *vec = nullptr;
    int32_t rc = tridiag_solve(c, n, alpha, beta, root, &vec);
    if (rc != NV_OK) return rc;
    std::vector<void *> tmp{vec};
    auto alloc = [&](size_t bytes, void **p) -> hipError_t {
        hipError_t er = hipMalloc(p, bytes);
        if (er == hipSuccess) tmp.push_back(*p);
        return er;
    };
    const double ref = origin == 0 ? root[0] : e0;
    const double cw = 219474.6313705;  // AU2WAVE (Source/progdata.f90:7)
    double *th = nullptr, *gr = nullptr, *gb = nullptr, *out = nullptr, *work = nullptr, *ab = nullptr, *bb = nullptr;
    double *tb = nullptr;
    hipStream_t s = c->stream;
    hipError_t e = alloc(n * 8, (void **)&th);
    if (e == hipSuccess) e = alloc(ng * 8, (void **)&gr);
    if (e == hipSuccess) e = hipMemcpyAsync(th, root.data(), n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(gr, grid, ng * 8, hipMemcpyHostToDevice, s);
    if (env) {
        if (e == hipSuccess) e = alloc(ng * 8, (void **)&out);
        if (e == hipSuccess) e = nv::launch_spec_env(n, th, vec, ref, cw, sigma, ng, gr, out, s);
        if (e == hipSuccess) e = hipMemcpyAsync(env, out, ng * 8, hipMemcpyDeviceToHost, s);
    } else {
        if (e == hipSuccess) e = alloc(ng * 8, (void **)&gb);
        if (e == hipSuccess) e = alloc(((size_t)2 * n + 8 + (size_t)2 * n * n) * 8, (void **)&work);
        if (e == hipSuccess) e = alloc(n * 8, (void **)&ab);
        if (e == hipSuccess) e = alloc((n + 1) * 8, (void **)&bb);
        if (tbar && e == hipSuccess) e = alloc((size_t)n * n * 8, (void **)&tb);
        if (e == hipSuccess) e = hipMemsetAsync(bb, 0, (n + 1) * 8, s);
        if (e == hipSuccess) e = hipMemcpyAsync(gb, gbar, ng * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = nv::launch_spec_adjoint(n, th, vec, ref, origin == 0, cw, sigma, ng, gr, gb, work, ab, bb, tb, s);
        if (abar && e == hipSuccess) e = hipMemcpyAsync(abar, ab, n * 8, hipMemcpyDeviceToHost, s);
        if (bbar && e == hipSuccess) e = hipMemcpyAsync(bbar, bb, (n + 1) * 8, hipMemcpyDeviceToHost, s);
        if (tbar && e == hipSuccess) e = hipMemcpyAsync(tbar, tb, (size_t)n * n * 8, hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (void *p : tmp) hipFree(p);
    if (e != hipSuccess) return hip_fail(c, e, "spectrum");
    return NV_OK;
}

int32_t nv_spectrum(nv_ctx *c, int64_t n, const double *alpha, const double *beta, int64_t ng, const double *grid,
                    double sigma, int32_t origin, double e0, double *env) {
    if (!env) return fail(c, NV_ERR_ARG, "env is NULL");
    return spectrum_common(c, n, alpha, beta, ng, grid, sigma, origin, e0, nullptr, env, nullptr, nullptr, nullptr);
}

int32_t nv_spectrum_adjoint(nv_ctx *c, int64_t n, const double *alpha, const double *beta, int64_t ng,
                            const double *grid, double sigma, int32_t origin, double e0, const double *gbar,
                            double *abar, double *bbar, double *tbar) {
    if (!gbar) return fail(c, NV_ERR_ARG, "gbar missing");
    return spectrum_common(c, n, alpha, beta, ng, grid, sigma, origin, e0, gbar, nullptr, abar, bbar, tbar);
}

int32_t nv_krylov_gradient(nv_ctx *c, int64_t n, const double *tbar, double *g_omega, double *g_concoef,
                           double *g_potterms) {
    if (!c) return NV_ERR_ARG;
    if (!tbar) return fail(c, NV_ERR_ARG, "tbar missing");
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "nv_lanczos_init has not been called");
    const bool archive = c->opts.reorthog > 0 || c->opts.savevecs != 0;
    if (!archive && !c->cp)
        return fail(c, NV_ERR_STATE, "the Krylov gradient needs the kept basis (savevecs or reorthog) or checkpoints");
    if (n < 1 || n > c->last_iter) return fail(c, NV_ERR_ARG, "n must lie in 1..last iteration");
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    int32_t rc = ensure_gop(c);
    if (rc != NV_OK) return rc;
    const int64_t N = c->op.N, niter = c->opts.niter;
    const int64_t K = archive ? n : c->opts.checkpoint;
    const int nss = (int)(c->op.ns * c->op.ns);
    const int groups = nv::grad_groups(c->gdev), chunks = nv::grad_chunks(c->gdev);
    const int64_t jc = std::min<int64_t>(K, 4 * nv::kRitzChunk);
    std::vector<void *> tmp;
    auto alloc = [&](size_t bytes, void **p) -> hipError_t {
        hipError_t er = hipMalloc(p, bytes);
        if (er == hipSuccess) tmp.push_back(*p);
        return er;
    };
    auto done = [&](int32_t code) {
        for (void *p : tmp) hipFree(p);
        return code;
    };
    double *tb = nullptr, *W = nullptr, *gpart = nullptr, *gacc = nullptr;
    double *segA = nullptr, *segB = nullptr, *ta = nullptr, *tbb = nullptr;
    hipError_t e = alloc((size_t)n * n * 8, (void **)&tb);
    if (e == hipSuccess) e = alloc((size_t)jc * N * 8, (void **)&W);
    if (e == hipSuccess) e = alloc((size_t)groups * chunks * nss * 8, (void **)&gpart);
    if (e == hipSuccess) e = alloc((size_t)groups * nss * 8, (void **)&gacc);
    if (!archive) {
        if (e == hipSuccess) e = alloc((size_t)2 * (K + 2) * N * 8, (void **)&segA);
        segB = segA + (K + 2) * N;
        if (e == hipSuccess) e = alloc((size_t)(niter + 3) * 8, (void **)&ta);
        if (e == hipSuccess) e = alloc((size_t)(niter + 3) * 8, (void **)&tbb);
    }
    if (e != hipSuccess) return done(hip_fail(c, e, "krylov gradient allocation"));
    hipStream_t s = c->stream;
    e = hipMemcpyAsync(tb, tbar, (size_t)n * n * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(gacc, 0, (size_t)groups * nss * 8, s);
    // Sum over j in [o0, o1] of q_j^T (dH/dtheta) w_j, w_j = sum_{k=i0..i1} Tbar[j][k] q_k: q_j in
    // slot (j % ons) of obuf, q_k in slot (k % ins) of ibuf.  Tbar is symmetric, so its rows
    // serve as the component-major coefficient table of launch_ritz_vectors.
    auto contract = [&](const double *obuf, int64_t ons, int64_t o0, int64_t o1, const double *ibuf, int64_t ins,
                        int64_t i0, int64_t i1) {
        for (int64_t j0 = o0; j0 <= o1 && e == hipSuccess; j0 += jc) {
            const int64_t m = std::min(jc, o1 - j0 + 1);
            e = nv::launch_ritz_vectors(ibuf, N, i1 - i0 + 1, tb + (i0 - 1) * n + (j0 - 1), n, m, i0 % ins, ins, W,
                                        s);
            for (int64_t j = 0; j < m && e == hipSuccess; ++j)
                e = nv::launch_grad_contract(c->gdev, W + j * N, obuf + ((j0 + j) % ons) * N, gpart, gacc, s);
        }
    };
    // Orthogonality of the basis the projection assumes: probe vectors against every later
    // vector of a buffer, the largest |overlap| kept.
    double *ppart = nullptr;
    const int pch = nv::probe_chunks(N);
    if (e == hipSuccess) e = alloc((size_t)std::max<int64_t>(n, 1) * pch * nv::kMaxProbes * 8, (void **)&ppart);
    double worst = 0.0;
    int64_t worst_j = 0, worst_k = 0;
    auto probe = [&](const double *pbuf, int64_t pns, const std::vector<int64_t> &pj, const double *vbuf,
                     int64_t vns, int64_t v0, int64_t v1) {
        if (e != hipSuccess || v1 < v0) return;
        nv::ProbeSet ps{};
        ps.np = (int)pj.size();
        for (int q = 0; q < ps.np; ++q) ps.p[q] = pbuf + (pj[q] % pns) * N;
        const int64_t nj = v1 - v0 + 1;
        e = nv::launch_probe_dots(ps, vbuf, vns, N, v0, nj, ppart, s);
        std::vector<double> h((size_t)nj * pch * nv::kMaxProbes);
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), ppart, h.size() * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return;
        for (int64_t k = 0; k < nj; ++k)
            for (int q = 0; q < ps.np; ++q) {
                if (v0 + k <= pj[q]) continue;  // later vectors only (and not the probe itself)
                double d = 0.0;
                for (int ch = 0; ch < pch; ++ch) d += h[((size_t)k * pch + ch) * nv::kMaxProbes + q];
                if (std::fabs(d) > worst) {
                    worst = std::fabs(d);
                    worst_j = pj[q];
                    worst_k = v0 + k;
                }
            }
    };
    if (archive) {
        // q_j in slot j of the archive
        contract(c->st.V, c->st.nslots, 1, n, c->st.V, c->st.nslots, 1, n);
        std::vector<int64_t> pj;
        for (int64_t j : {int64_t(1), int64_t(2), (n + 2) / 3, (2 * n + 2) / 3, n - 1})
            if (j >= 1 && j < n && std::find(pj.begin(), pj.end(), j) == pj.end()) pj.push_back(j);
        probe(c->st.V, c->st.nslots, pj, c->st.V, c->st.nslots, 2, n);
    } else {
        // Segment pairs (A, B), B >= A, from the checkpoints: segment A stays resident while
        // every later segment is recomputed beside it; an off-diagonal pair contributes both
        // orders, so the sum covers every (j, k).
        const int64_t nseg = (n + K - 1) / K;
        for (int64_t A = 0; A < nseg && e == hipSuccess; ++A) {
            rc = recompute_segment(c, A, K, n, segA, ta, tbb);
            if (rc != NV_OK) return done(rc);
            const int64_t a0 = A * K + 1, a1 = std::min((A + 1) * K, n);
            for (int64_t B = A; B < nseg && e == hipSuccess; ++B) {
                const int64_t b0 = B * K + 1, b1 = std::min((B + 1) * K, n);
                const std::vector<int64_t> pj = a1 > a0 ? std::vector<int64_t>{a0, a1} : std::vector<int64_t>{a0};
                if (B == A) {
                    contract(segA, K + 2, a0, a1, segA, K + 2, a0, a1);
                    probe(segA, K + 2, pj, segA, K + 2, a0 + 1, a1);
                    continue;
                }
                rc = recompute_segment(c, B, K, n, segB, ta, tbb);
                if (rc != NV_OK) return done(rc);
                contract(segA, K + 2, a0, a1, segB, K + 2, b0, b1);
                contract(segB, K + 2, b0, b1, segA, K + 2, a0, a1);
                probe(segA, K + 2, pj, segB, K + 2, b0, b1);
            }
        }
    }
    std::vector<double> gh((size_t)groups * nss);
    if (e == hipSuccess) e = hipMemcpyAsync(gh.data(), gacc, gh.size() * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return done(hip_fail(c, e, "krylov gradient"));
    map_gradients(c, gh, g_omega, g_concoef, g_potterms);
    if (worst > 1e-4) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "the Krylov basis has lost orthogonality: |q_%lld . q_%lld| = %.3e > 1e-4; the projected "
                      "gradient is not the derivative of the computed loss (re-orthogonalise, or shorten the run)",
                      (long long)worst_j, (long long)worst_k, worst);
        return done(fail(c, NV_WARN_NONORTHOGONAL, msg));
    }
    return done(NV_OK);
}

// ---- Gaussian broadening ----------------------------------------------------
int32_t nv_convolve(nv_ctx *c, int64_t n, const double *x, const double *y, int64_t ng, const double *grid,
                    double sigma, double *env) {
    if (!c) return NV_ERR_ARG;
    if (n < 1 || ng < 1 || !x || !y || !grid || !env || !(sigma > 0))
        return fail(c, NV_ERR_ARG, "bad convolve arguments");
    NV_HIP(c, hipSetDevice(c->device));
    const int nb = (int)((ng + nv::kBlock - 1) / nv::kBlock);
    double *d = nullptr;
    const size_t bytes = (size_t)(2 * n + 2 * ng + nb) * sizeof(double);
    NV_HIP(c, hipMalloc(&d, bytes));
    double *dx = d, *dy = d + n, *dg = d + 2 * n, *de = d + 2 * n + ng, *dp = d + 2 * n + 2 * ng;
    hipError_t e = hipMemcpyAsync(dx, x, n * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dy, y, n * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dg, grid, ng * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = nv::launch_convolve(dx, dy, n, dg, ng, sigma, de, dp, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(env, de, ng * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d);
    if (e != hipSuccess) return hip_fail(c, e, "convolve");
    return NV_OK;
}

int32_t nv_convolve_noalloc(nv_ctx *c, int64_t n, const double *x, const double *y, int64_t ng, const double *grid,
                            double sigma, double *env, void *stream) {
    if (!c) return NV_ERR_ARG;
    if (n < 1 || ng < 1 || !x || !y || !grid || !env || !(sigma > 0))
        return fail(c, NV_ERR_ARG, "bad convolve arguments");
    if (ng > NV_CONVOLVE_MAX_GRID)
        return fail(c, NV_ERR_ARG, "nv_convolve_noalloc takes at most NV_CONVOLVE_MAX_GRID (" +
                                       std::to_string(NV_CONVOLVE_MAX_GRID) + ") grid points, got " +
                                       std::to_string(ng));
    double *part = scratch_of(c->device);
    if (!part) return fail(c, NV_ERR_STATE, "the static scratch has no address on this device");
    NV_HIP(c, hipSetDevice(c->device));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    NV_HIP(c, nv::launch_convolve(x, y, n, grid, ng, sigma, env, part, s));
    return NV_OK;
}

// ---- restart vectors ------------------------------------------------------
int32_t nv_save_vectors(nv_ctx *c, const char *path) {
    if (!c || !path) return NV_ERR_ARG;
    if (!c->lz_ready || c->last_iter < 1) return fail(c, NV_ERR_STATE, "no completed iteration");
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t N = c->op.N, it = c->last_iter;
    std::vector<double> buf(3 * N, 0.0);
    const int64_t js[3] = {it - 1, it, it + 1};
    for (int k = 0; k < 3; ++k) {
        if (js[k] < 1) continue;
        const double *src = c->st.V + (js[k] % c->st.nslots) * N;
        NV_HIP(c, hipMemcpy(buf.data() + k * N, src, N * sizeof(double), hipMemcpyDeviceToHost));
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(c, NV_ERR_IO, std::string("cannot open ") + path);
    size_t w = std::fwrite(buf.data(), sizeof(double), buf.size(), f);
    std::fclose(f);
    if (w != buf.size()) return fail(c, NV_ERR_IO, std::string("short write to ") + path);
    return NV_OK;
}

int32_t nv_load_vectors(nv_ctx *c, const char *path, int64_t iter, const double *alpha, const double *beta,
                        const double *omega) {
    if (!c || !path || !alpha || !beta) return NV_ERR_ARG;
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "nv_lanczos_init has not been called");
    if (iter < 1 || iter >= c->opts.niter) return fail(c, NV_ERR_ARG, "restart iteration out of range");
    if (c->cp) return fail(c, NV_ERR_STATE, "a restart cannot rebuild the checkpoints of earlier Lanczos vectors");
    if ((c->opts.reorthog > 0 || c->opts.savevecs != 0) && c->archive_iter != iter)
        return fail(c, NV_ERR_STATE, "a run that keeps every Lanczos vector restarts from its archive: "
                                     "call nv_load_archive for the same iteration first");
    const int64_t N = c->op.N;
    std::vector<double> buf(3 * N);
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(c, NV_ERR_IO, std::string("cannot open ") + path);
    size_t r = std::fread(buf.data(), sizeof(double), buf.size(), f);
    std::fclose(f);
    if (r != buf.size()) return fail(c, NV_ERR_IO, std::string("short read from ") + path);
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t js[3] = {iter - 1, iter, iter + 1};
    for (int k = 0; k < 3; ++k) {
        double *dst = c->st.V + (js[k] % c->st.nslots) * N;
        NV_HIP(c, hipMemcpy(dst, buf.data() + k * N, N * sizeof(double), hipMemcpyHostToDevice));
    }
    NV_HIP(c, hipMemcpy(c->st.alpha + 1, alpha, iter * sizeof(double), hipMemcpyHostToDevice));
    NV_HIP(c, hipMemcpy(c->st.beta, beta, (iter + 1) * sizeof(double), hipMemcpyHostToDevice));
    // the reference's restart.log carries omega (Source/top.f90:564-568); a PRO state
    // loaded by nv_load_archive already holds it
    if (omega && c->opts.reorthog > 0 && !c->pro_loaded) {
        hipError_t e = nv::pro_put_omega(c->pro, omega, iter);
        if (e != hipSuccess) return hip_fail(c, e, "omega restore");
    }
    c->last_iter = iter;
    c->next_update = iter + 1;
    return NV_OK;
}

static __global__ void to_f32_kernel(const double *__restrict__ x, float *__restrict__ y, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (float)x[i];
}

static const char kProMagic[8] = {'N', 'V', 'P', 'R', 'O', '0', '0', '1'};

int32_t nv_save_archive(nv_ctx *c, const char *archive_path, const char *pro_path) {
    if (!c || !archive_path) return NV_ERR_ARG;
    if (!c->lz_ready || c->last_iter < 1) return fail(c, NV_ERR_STATE, "no completed iteration");
    if (c->opts.reorthog <= 0 && c->opts.savevecs == 0)
        return fail(c, NV_ERR_STATE, "the archive needs a run that keeps every Lanczos vector (savevecs or reorthog)");
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t N = c->op.N, n = c->last_iter;
    FILE *f = std::fopen(archive_path, "wb");
    if (!f) return fail(c, NV_ERR_IO, std::string("cannot open ") + archive_path);
    std::vector<double> buf(N);
    for (int64_t j = 1; j <= n; ++j) {
        hipError_t e = hipMemcpy(buf.data(), c->st.V + (j % c->st.nslots) * N, N * sizeof(double),
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            std::fclose(f);
            return hip_fail(c, e, "archive copy");
        }
        if (std::fwrite(buf.data(), sizeof(double), N, f) != (size_t)N) {
            std::fclose(f);
            return fail(c, NV_ERR_IO, std::string("short write to ") + archive_path);
        }
    }
    std::fclose(f);
    if (!pro_path) return NV_OK;
    if (c->opts.reorthog <= 0) return fail(c, NV_ERR_STATE, "no re-orthogonalisation state in this run");
    std::vector<int64_t> ints;
    std::vector<double> reals;
    hipError_t e = nv::pro_export(c->pro, n, ints, reals);
    if (e != hipSuccess) return hip_fail(c, e, "PRO state");
    f = std::fopen(pro_path, "wb");
    if (!f) return fail(c, NV_ERR_IO, std::string("cannot open ") + pro_path);
    const int64_t sizes[2] = {(int64_t)ints.size(), (int64_t)reals.size()};
    bool ok = std::fwrite(kProMagic, 1, 8, f) == 8 && std::fwrite(sizes, sizeof(int64_t), 2, f) == 2 &&
              std::fwrite(ints.data(), sizeof(int64_t), ints.size(), f) == ints.size() &&
              std::fwrite(reals.data(), sizeof(double), reals.size(), f) == reals.size();
    std::fclose(f);
    if (!ok) return fail(c, NV_ERR_IO, std::string("short write to ") + pro_path);
    return NV_OK;
}

int32_t nv_load_archive(nv_ctx *c, const char *archive_path, const char *pro_path, int64_t iter) {
    if (!c || !archive_path) return NV_ERR_ARG;
    if (!c->lz_ready) return fail(c, NV_ERR_STATE, "nv_lanczos_init has not been called");
    if (c->opts.reorthog <= 0 && c->opts.savevecs == 0)
        return fail(c, NV_ERR_STATE, "the archive needs a run that keeps every Lanczos vector (savevecs or reorthog)");
    if (iter < 1 || iter >= c->opts.niter) return fail(c, NV_ERR_ARG, "restart iteration out of range");
    if (pro_path && c->opts.reorthog <= 0) return fail(c, NV_ERR_STATE, "no re-orthogonalisation in this run");
    NV_HIP(c, hipSetDevice(c->device));
    NV_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t N = c->op.N;
    FILE *f = std::fopen(archive_path, "rb");
    if (!f) return fail(c, NV_ERR_IO, std::string("cannot open ") + archive_path);
    std::vector<double> buf(N);
    for (int64_t j = 1; j <= iter; ++j) {
        if (std::fread(buf.data(), sizeof(double), N, f) != (size_t)N) {
            std::fclose(f);
            return fail(c, NV_ERR_IO, std::string("short read from ") + archive_path);
        }
        double *dst = c->st.V + (j % c->st.nslots) * N;
        hipError_t e = hipMemcpy(dst, buf.data(), N * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess && c->st.Vf) {
            hipLaunchKernelGGL(to_f32_kernel, dim3(1024), dim3(256), 0, c->stream, dst,
                               c->st.Vf + (j % c->st.nslots) * c->st.ldf, N);
            e = hipStreamSynchronize(c->stream);
        }
        if (e != hipSuccess) {
            std::fclose(f);
            return hip_fail(c, e, "archive load");
        }
    }
    std::fclose(f);
    c->pro_loaded = false;
    if (pro_path) {
        f = std::fopen(pro_path, "rb");
        if (!f) return fail(c, NV_ERR_IO, std::string("cannot open ") + pro_path);
        char magic[8];
        int64_t sizes[2] = {0, 0};
        bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kProMagic, 8) == 0 &&
                  std::fread(sizes, sizeof(int64_t), 2, f) == 2 && sizes[0] >= 0 && sizes[1] >= 0 &&
                  sizes[0] < (int64_t)1 << 32 && sizes[1] < (int64_t)1 << 32;
        std::vector<int64_t> ints(ok ? sizes[0] : 0);
        std::vector<double> reals(ok ? sizes[1] : 0);
        ok = ok && std::fread(ints.data(), sizeof(int64_t), ints.size(), f) == ints.size() &&
             std::fread(reals.data(), sizeof(double), reals.size(), f) == reals.size();
        std::fclose(f);
        if (!ok) return fail(c, NV_ERR_IO, std::string("not a re-orthogonalisation state file: ") + pro_path);
        hipError_t e = nv::pro_import(c->pro, iter, ints, reals);
        if (e == hipErrorInvalidValue)
            return fail(c, NV_ERR_ARG, "the re-orthogonalisation state is of another iteration or does not fit this run");
        if (e != hipSuccess) return hip_fail(c, e, "PRO state load");
        c->pro_loaded = true;
    }
    c->archive_iter = iter;
    return NV_OK;
}

}  // extern "C"

// ---- batched runs ----------------------------------------------------------------
// Many problems advance together: one launch sequence per iteration covers
// every problem (tiles of bs rows, one workgroup each), captured once as a
// graph and replayed.  Each problem keeps its own context for its operator
// tables, start vector and eigensolve.
struct nv_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<nv_ctx *> ctx;
    std::vector<int64_t> voff, loff;
    int64_t Ntot = 0, Ltot = 0;
    int prec = NV_PREC_F64;
    int64_t niter = 0, last = 0, next = 1;
    bool ready = false;
    nv::BatchDev d{};
    NvOpDev *ops_d = nullptr;
    int64_t *voff_d = nullptr, *loff_d = nullptr, *iter_d = nullptr;
    nv::BatchTile *tiles_d = nullptr;
    int32_t *tile0_d = nullptr;
    std::vector<void *> allocs, run_allocs;
    size_t lds_hv = 0, lds_step = 0;
    bool rows4 = false;  // lds_hv holds the digits of four rows per thread (FP32 shaped H*v)
    hipGraphExec_t graph = nullptr;
    hipGraphExec_t graph16 = nullptr;  // 16 iterations
    int bs = 256;
    std::vector<nv::BatchGroup> groups;
};

namespace {

int32_t bfail(nv_batch *b, int32_t code, const std::string &msg) {
    if (b) b->err = msg;
    else g_create_error = msg;
    return code;
}

#define NV_BHIP(b, call)                                                                            \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return bfail((b), NV_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Tile table for bs rows per tile: problems grouped by their number of
// states (one specialised launch per group), each problem's tiles contiguous
// so that its partial sums add in tile order.
hipError_t batch_tiles(nv_batch *b, int bs) {
    // launches already queued may still read the current tables
    if (b->stream) {
        hipError_t e0 = hipStreamSynchronize(b->stream);
        if (e0 != hipSuccess) return e0;
    }
    const int64_t P = (int64_t)b->ctx.size();
    std::vector<int64_t> order(P);
    for (int64_t p = 0; p < P; ++p) order[p] = p;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return b->ctx[x]->op.ns < b->ctx[y]->op.ns; });
    std::vector<nv::BatchTile> t;
    std::vector<int32_t> t0(2 * P);
    b->groups.clear();
    for (int64_t p : order) {
        const int ns = (int)b->ctx[p]->op.ns;
        if (b->groups.empty() || b->groups.back().ns != ns) b->groups.push_back({ns, (int)t.size(), 0});
        t0[2 * p] = (int32_t)t.size();
        const int64_t D = b->ctx[p]->op.D;
        for (int64_t r = 0; r < D; r += bs) t.push_back({(int32_t)p, (int32_t)r});
        t0[2 * p + 1] = (int32_t)t.size();
        b->groups.back().nt = (int)t.size() - b->groups.back().t0;
    }
    if (b->tiles_d) hipFree(b->tiles_d);
    if (b->tile0_d) hipFree(b->tile0_d);
    b->tiles_d = nullptr;
    b->tile0_d = nullptr;
    hipError_t e = hipMalloc(&b->tiles_d, t.size() * sizeof(nv::BatchTile));
    if (e == hipSuccess) e = hipMalloc(&b->tile0_d, t0.size() * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpy(b->tiles_d, t.data(), t.size() * sizeof(nv::BatchTile), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->tile0_d, t0.data(), t0.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    b->d.tiles = b->tiles_d;
    b->d.tile0 = b->tile0_d;
    b->d.T = (int32_t)t.size();
    b->d.bs = bs;
    b->bs = bs;
    // captured iterations name the tiles
    if (b->graph) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    if (b->graph16) {
        hipGraphExecDestroy(b->graph16);
        b->graph16 = nullptr;
    }
    return e;
}

hipError_t batch_enqueue(nv_batch *b, bool absolute, int64_t it) {
    hipStream_t s = b->stream;
    hipError_t e = nv::launch_set_iter(b->iter_d, absolute ? it : 0, absolute ? 0 : 1, s);
    if (e == hipSuccess) e = nv::launch_batch_normalize(b->d, b->prec, s);
    if (e == hipSuccess) e = nv::launch_batch_hv(b->d, b->groups.data(), (int)b->groups.size(), b->prec, b->lds_hv, s);
    if (e == hipSuccess)
        e = nv::launch_batch_step(b->d, b->groups.data(), (int)b->groups.size(), b->prec, b->lds_step, s);
    return e;
}

void batch_release_run(nv_batch *b) {
    if (b->graph) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    if (b->graph16) {
        hipGraphExecDestroy(b->graph16);
        b->graph16 = nullptr;
    }
    for (void *p : b->run_allocs) hipFree(p);
    b->run_allocs.clear();
    b->ready = false;
}

}  // namespace

extern "C" {

int32_t nv_batch_create(const nv_problem *probs, int64_t count, int32_t device, nv_batch **out) {
    if (!out) return bfail(nullptr, NV_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!probs || count < 1) return bfail(nullptr, NV_ERR_ARG, "no problems");
    if (count > (int64_t)1 << 20) return bfail(nullptr, NV_ERR_ARG, "too many problems");
    nv_batch *b = new (std::nothrow) nv_batch();
    if (!b) return bfail(nullptr, NV_ERR_NOMEM, "host allocation failed");
    b->device = device;
    for (int64_t i = 0; i < count; ++i) {
        nv_ctx *c = nullptr;
        int32_t rc = nv_create(&probs[i], device, &c);
        if (rc != NV_OK) {
            const std::string msg = "problem " + std::to_string(i) + ": " + g_create_error;
            nv_batch_destroy(b);
            return bfail(nullptr, rc, msg);
        }
        c->jit_on = false;  // the batch runs the table-driven kernels
        b->ctx.push_back(c);
    }
    std::vector<NvOpDev> ops;
    for (nv_ctx *c : b->ctx) {
        b->voff.push_back(b->Ntot);
        b->loff.push_back(b->Ltot);
        b->Ntot += c->op.N;
        b->Ltot += c->op.D * nv_lo_stride((int)c->op.ns);
        ops.push_back(c->dev);
        b->lds_hv = std::max(b->lds_hv, nv::hv_lds_bytes(c->dev, NV_PREC_F64));
        b->lds_step = std::max(b->lds_step, (size_t)c->op.M * nv::kBlock);
    }
    auto fail_out = [&](int32_t rc) {
        const std::string msg = b->err;
        nv_batch_destroy(b);
        return bfail(nullptr, rc, msg);
    };
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&b->ops_d, ops.size() * sizeof(NvOpDev));
    if (e == hipSuccess) e = hipMalloc(&b->voff_d, b->voff.size() * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&b->loff_d, b->loff.size() * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&b->iter_d, sizeof(int64_t));
    if (e == hipSuccess) e = hipMemcpy(b->ops_d, ops.data(), ops.size() * sizeof(NvOpDev), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->voff_d, b->voff.data(), b->voff.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->loff_d, b->loff.data(), b->loff.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = batch_tiles(b, b->bs);
    if (e != hipSuccess) {
        b->err = std::string("batch setup: ") + hipGetErrorString(e);
        return fail_out(NV_ERR_HIP);
    }
    b->d.ops = b->ops_d;
    b->d.voff = b->voff_d;
    b->d.loff = b->loff_d;
    b->d.iter = b->iter_d;
    b->d.P = (int32_t)b->ctx.size();
    b->d.Ntot = b->Ntot;
    b->d.shaped[0] = b->d.shaped[1] = 1;
    for (nv_ctx *c : b->ctx) {
        if (!nv::hv_shaped_batch(c->dev, NV_PREC_F64)) b->d.shaped[0] = 0;
        if (!nv::hv_shaped_batch(c->dev, NV_PREC_F32)) b->d.shaped[1] = 0;
    }
    *out = b;
    return NV_OK;
}

int32_t nv_batch_destroy(nv_batch *b) {
    if (!b) return NV_OK;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    batch_release_run(b);
    for (void *p : {(void *)b->ops_d, (void *)b->voff_d, (void *)b->loff_d, (void *)b->iter_d, (void *)b->tiles_d,
                    (void *)b->tile0_d})
        if (p) hipFree(p);
    for (nv_ctx *c : b->ctx) nv_destroy(c);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return NV_OK;
}

const char *nv_batch_last_error(const nv_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int32_t nv_batch_init(nv_batch *b, int64_t niter, int32_t prec) {
    if (!b) return NV_ERR_ARG;
    if (niter < 1) return bfail(b, NV_ERR_ARG, "niter must be >= 1");
    if (prec != NV_PREC_F64 && prec != NV_PREC_F32) return bfail(b, NV_ERR_ARG, "batch runs support F64 and F32");
    NV_BHIP(b, hipSetDevice(b->device));
    NV_BHIP(b, hipStreamSynchronize(b->stream));
    batch_release_run(b);
    const int64_t P = (int64_t)b->ctx.size(), ld = niter + 2;
    int64_t maxtiles = 0;
    for (nv_ctx *c : b->ctx) maxtiles += (c->op.D + 63) / 64;  // the smallest tile size
    auto alloc = [&](size_t bytes, void **p) -> hipError_t {
        hipError_t er = hipMalloc(p, bytes);
        if (er == hipSuccess) b->run_allocs.push_back(*p);
        return er;
    };
    double *V = nullptr, *al = nullptr, *be = nullptr, *parts = nullptr;
    void *vlo = nullptr, *y = nullptr;
    NV_BHIP(b, alloc((size_t)3 * b->Ntot * sizeof(double), (void **)&V));
    NV_BHIP(b, alloc((size_t)b->Ntot * (prec == NV_PREC_F64 ? 8 : 4), &y));
    if (prec == NV_PREC_F32) {
        NV_BHIP(b, alloc((size_t)b->Ltot * sizeof(float), &vlo));
        NV_BHIP(b, hipMemsetAsync(vlo, 0, (size_t)b->Ltot * sizeof(float), b->stream));
    }
    NV_BHIP(b, alloc((size_t)P * ld * sizeof(double), (void **)&al));
    NV_BHIP(b, alloc((size_t)P * ld * sizeof(double), (void **)&be));
    NV_BHIP(b, alloc((size_t)maxtiles * sizeof(double), (void **)&parts));
    NV_BHIP(b, hipMemsetAsync(V, 0, (size_t)3 * b->Ntot * sizeof(double), b->stream));
    NV_BHIP(b, hipMemsetAsync(al, 0, (size_t)P * ld * sizeof(double), b->stream));
    NV_BHIP(b, hipMemsetAsync(be, 0, (size_t)P * ld * sizeof(double), b->stream));
    // the start vectors below are plain copies: the clears must have landed first
    NV_BHIP(b, hipStreamSynchronize(b->stream));
    // start vectors (each problem's own: default or Franck-Condon) into slot 1, beta(0) = |start|
    std::vector<double> beta0(P * ld, 0.0);
    for (int64_t p = 0; p < P; ++p) {
        nv_ctx *c = b->ctx[p];
        std::vector<double> q(c->op.N);
        int32_t rc = nv_start_vector(c, q.data());
        if (rc != NV_OK) return bfail(b, rc, "problem " + std::to_string(p) + ": " + c->err);
        double nn = 0.0;
        for (double v : q) nn += v * v;
        beta0[p * ld] = std::sqrt(nn);
        NV_BHIP(b, hipMemcpy(V + b->Ntot + b->voff[p], q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    NV_BHIP(b, hipMemcpy(be, beta0.data(), beta0.size() * sizeof(double), hipMemcpyHostToDevice));
    b->d.V = V;
    b->d.vlo = vlo;
    b->d.y = y;
    b->d.alpha = al;
    b->d.beta = be;
    b->d.ld = ld;
    b->d.partials = parts;
    b->prec = prec;
    b->niter = niter;
    b->last = 0;
    b->next = 1;
    b->lds_hv = 0;
    size_t lds4 = 0;
    for (nv_ctx *c : b->ctx) {
        const size_t l = nv::hv_lds_bytes(c->dev, prec);
        b->lds_hv = std::max(b->lds_hv, l);
        lds4 = std::max(lds4, l + 2 * (size_t)c->dev.M * 256);  // two more rows' digits per thread
    }
    // four rows per thread (FP32, shaped form) where their digits fit in 64 KB of LDS
    b->rows4 = prec != NV_PREC_F64 && b->d.shaped[1] != 0 && lds4 <= 65536;
    if (b->rows4) b->lds_hv = lds4;
    if (const char *r = std::getenv("NV_BATCH_ROWS")) {
        const int v = std::atoi(r);
        if (v != 1 && v != 2 && v != 4) return bfail(b, NV_ERR_ARG, "NV_BATCH_ROWS must be 1, 2 or 4");
        if (v > 1 && prec == NV_PREC_F64) return bfail(b, NV_ERR_ARG, "NV_BATCH_ROWS above 1 needs FP32 H*v");
        if (v == 4 && (!b->rows4 || b->bs < 256))
            return bfail(b, NV_ERR_ARG, "four rows per thread need the shaped form, 256-row tiles and LDS room");
        b->d.rows = v;
    }
    b->ready = true;
    return NV_OK;
}

int32_t nv_batch_run(nv_batch *b, int64_t iiter, int64_t niter) {
    if (!b) return NV_ERR_ARG;
    if (!b->ready) return bfail(b, NV_ERR_STATE, "nv_batch_init has not been called");
    if (iiter != b->next) return bfail(b, NV_ERR_STATE, "run must start at the next iteration");
    if (niter > b->niter) return bfail(b, NV_ERR_ARG, "niter beyond the initialised run");
    NV_BHIP(b, hipSetDevice(b->device));
    // one captured iteration, or kBatchIters of them in one launch
    constexpr int64_t kBatchIters = 16;
    auto capture = [&](int64_t count, hipGraphExec_t *out) -> hipError_t {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t e2 = hipSuccess;
        for (int64_t k = 0; k < count && e2 == hipSuccess; ++k) e2 = batch_enqueue(b, false, 0);
        e = hipStreamEndCapture(b->stream, &g);
        if (e2 != hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        if (g) hipGraphDestroy(g);
        return e;
    };
    for (int64_t i = iiter; i <= niter; ++i) {
        if (i == iiter) {
            NV_BHIP(b, batch_enqueue(b, true, i));
        } else if (i + kBatchIters - 1 <= niter) {
            if (!b->graph16) {
                hipError_t e = capture(kBatchIters, &b->graph16);
                if (e != hipSuccess) return bfail(b, NV_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(e));
            }
            NV_BHIP(b, hipGraphLaunch(b->graph16, b->stream));
            i += kBatchIters - 1;
        } else {
            if (!b->graph) {
                hipError_t e = capture(1, &b->graph);
                if (e != hipSuccess) return bfail(b, NV_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(e));
            }
            NV_BHIP(b, hipGraphLaunch(b->graph, b->stream));
        }
        b->last = i;
        b->next = i + 1;
    }
    return NV_OK;
}

int32_t nv_batch_coefficients(nv_batch *b, int64_t p, double *alpha, double *beta) {
    if (!b) return NV_ERR_ARG;
    if (!b->ready) return bfail(b, NV_ERR_STATE, "no batch run");
    if (p < 0 || p >= (int64_t)b->ctx.size()) return bfail(b, NV_ERR_ARG, "problem index out of range");
    NV_BHIP(b, hipSetDevice(b->device));
    NV_BHIP(b, hipStreamSynchronize(b->stream));
    const int64_t n = b->last;
    if (alpha && n > 0)
        NV_BHIP(b, hipMemcpy(alpha, b->d.alpha + p * b->d.ld + 1, n * sizeof(double), hipMemcpyDeviceToHost));
    if (beta) NV_BHIP(b, hipMemcpy(beta, b->d.beta + p * b->d.ld, (n + 1) * sizeof(double), hipMemcpyDeviceToHost));
    return NV_OK;
}

nv_ctx *nv_batch_context(nv_batch *b, int64_t p) {
    if (!b || p < 0 || p >= (int64_t)b->ctx.size()) return nullptr;
    return b->ctx[p];
}

int32_t nv_batch_tune(nv_batch *b, char *report, int64_t len) {
    if (!b) return NV_ERR_ARG;
    if (!b->ready || b->last != 0) return bfail(b, NV_ERR_STATE, "tune after nv_batch_init, before the run");
    NV_BHIP(b, hipSetDevice(b->device));
    // the search space: rows per workgroup (tile size) x rows per thread (FP32) x kernel form
    struct Cand {
        int bs, rows, steps;
    };
    std::vector<Cand> cands;
    const bool shaped = b->d.shaped[b->prec == NV_PREC_F64 ? 0 : 1] != 0;
    for (int steps = 0; steps <= 1; ++steps) {
        if (steps == 0 && !shaped) continue;
        for (int bs : {64, 128, 256})
            for (int rows : {1, 2, 4}) {
                if (rows == 2 && (b->prec == NV_PREC_F64 || bs < 128)) continue;
                if (rows == 4 && (steps || !b->rows4 || bs < 256)) continue;
                cands.push_back({bs, rows, steps});
            }
    }
    float best_ms = 0.0f;
    Cand best = {b->bs, b->d.rows, b->d.steps};
    std::string rep = "batched H*v (rows per workgroup / rows per thread / form):";
    hipEvent_t e0, e1;
    NV_BHIP(b, hipEventCreate(&e0));
    NV_BHIP(b, hipEventCreate(&e1));
    hipError_t e = hipSuccess;
    for (size_t k = 0; k < cands.size() && e == hipSuccess; ++k) {
        const Cand &cd = cands[k];
        e = batch_tiles(b, cd.bs);
        b->d.rows = cd.rows;
        b->d.steps = cd.steps;
        if (e == hipSuccess) e = nv::launch_set_iter(b->iter_d, 1, 0, b->stream);
        auto hv = [&]() {
            return nv::launch_batch_hv(b->d, b->groups.data(), (int)b->groups.size(), b->prec, b->lds_hv, b->stream);
        };
        for (int w = 0; w < 2 && e == hipSuccess; ++w) e = hv();
        if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
        for (int w = 0; w < 10 && e == hipSuccess; ++w) e = hv();
        if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        ms /= 10.0f;
        char buf[96];
        std::snprintf(buf, sizeof(buf), " %d/%d/%s %.1f us;", cd.bs, cd.rows, cd.steps ? "steps" : "shaped",
                      1000.0 * ms);
        rep += buf;
        if (k == 0 || ms < best_ms) {
            best_ms = ms;
            best = cd;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (e == hipSuccess) e = batch_tiles(b, best.bs);
    b->d.rows = best.rows;
    b->d.steps = best.steps;
    // alpha(1) of the timing launches is recomputed by the first iteration
    if (e == hipSuccess) e = hipMemsetAsync(b->d.alpha, 0, (size_t)b->ctx.size() * b->d.ld * sizeof(double), b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bfail(b, NV_ERR_HIP, std::string("batch tune: ") + hipGetErrorString(e));
    char buf[96];
    std::snprintf(buf, sizeof(buf), " chosen %d/%d/%s", best.bs, best.rows, best.steps ? "steps" : "shaped");
    rep += buf;
    if (report && len > 0) std::snprintf(report, (size_t)len, "%s", rep.c_str());
    return NV_OK;
}

}  // extern "C"
