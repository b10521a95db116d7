// Partial re-orthogonalisation (set_omega / partial_orthog,
// Source/top.f90:844-1005) and the Franck-Condon start vector
// (generate_initialvec, Source/secondary.f90:351-809) on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "nadvibs.h"
#include "nv_kernels.hpp"
#include "nv_operator.hpp"

namespace nv {

// Noise draw `idx` of the omega estimates: N(0, sigma) from a counter-based
// generator (splitmix64 + Box-Muller), identical on host and device.  The
// reference draws from an unseeded random_number stream
// (Source/function.f90:260-267); a counter-based stream makes runs
// reproducible and lets the device draw in parallel.
__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline double noise_gauss(uint64_t seed, uint64_t idx, double sigma) {
    const uint64_t a = splitmix64(seed ^ (2 * idx + 0x632BE59BD9B4E019ull));
    const uint64_t b = splitmix64(seed ^ (2 * idx + 1 + 0x632BE59BD9B4E019ull));
    const double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    return sigma * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct ProDev {
    double *omega = nullptr;     // [2][niter+2]: omega(1,j) at j, omega(2,j) at stride+j
    int64_t ostride = 0;
    int64_t niter = 0, chkorthog = 100;
    int32_t exact = 0;
    uint64_t seed = 0;
    double noise = 1.0;
    double eps = 0.0;            // unit roundoff of the recurrence
    double seps = 0.0, eta = 0.0;  // sqrt(eps), eps^(3/4)
    double epsn = 0.0;           // epsilon * nstates * dimen
    double epsn_scale = 0.0;     // nstates * dimen
    int64_t *lo = nullptr, *ro = nullptr;  // ranges [cap]
    int64_t *roff = nullptr;     // first list entry of each range [cap]
    int32_t *scr = nullptr;      // detection scratch: small indices, gap flags
    int64_t cap = 0;             // range slots
    int64_t extent = 0;          // ceil(niter/100): the reference's range array size
    int64_t *jlist = nullptr;    // vectors to orthogonalise against, in order
    int64_t jcap = 0;
    int32_t *njl = nullptr;      // entries in jlist
    int64_t *meta = nullptr;     // [0]=firststep [1]=northog [2]=draws [3]=events [4]=touched [5]=overflow
    double *dots_a = nullptr, *dots_b = nullptr;  // [jcap]
    double *dpart = nullptr;     // multidot partials
    int32_t nchunks = 0;
};

struct ProState {
    ProDev d;
    int32_t nchunks = 0;
    int64_t jcap = 0;
};

hipError_t pro_init(ProState &p, const nv_lanczos_opts &o, const Operator &op, const NvOpDev &dev,
                    const LzState &st, std::vector<void *> &allocs, hipStream_t s);
size_t pro_decide_lds(int64_t niter);
// The trigger tau and range end eta that a run with options o uses, after
// the defaults (sqrt(eps), eps^(3/4); NV_PRO_MIXED_TAU / _ETA in the mixed
// modes) are applied.
void pro_effective_thresholds(const nv_lanczos_opts &o, double *tau, double *eta);
hipError_t pro_reset(ProState &p, const nv_lanczos_opts &o, hipStream_t s);
int64_t pro_check_period(const ProState &p);
// check != 0: also the exact overlap check of this iteration.
hipError_t pro_set_omega(ProState &p, const NvOpDev &dev, const LzState &st, int check, hipStream_t s);
hipError_t pro_partial_orthog(ProState &p, const NvOpDev &dev, const LzState &st, hipStream_t s);
// Whether an iteration without an exact check may run as lanczos_step's
// kernel (without its beta finish) followed by pro_step_fused: beta, the omega
// recurrence and the range decision in one workgroup (nparts: the step's block
// partials), then the update.  Not in exact mode (reorthog > 1).
bool pro_fusable(const ProState &p);
hipError_t pro_step_fused(ProState &p, const NvOpDev &dev, const LzState &st, int nparts, hipStream_t s);
hipError_t pro_get_omega(const ProState &p, double *omega, int64_t n);
// The state of partial re-orthogonalisation after iteration n as a flat
// record (omega rows 0..n+1, ranges, list, counters, noise position), and its
// restoration into a run initialised for at least n + 1 iterations.
hipError_t pro_export(const ProState &p, int64_t n, std::vector<int64_t> &ints, std::vector<double> &reals);
// Returns hipErrorInvalidValue when the record does not fit this run.
hipError_t pro_import(ProState &p, int64_t n, const std::vector<int64_t> &ints, const std::vector<double> &reals);
// omega rows 0..n from the host (omega[2j] = omega(1,j), omega[2j+1] = omega(2,j)).
hipError_t pro_put_omega(ProState &p, const double *omega, int64_t n);
void pro_stats(const ProState &p, int64_t *events, int64_t *touched);

// Franck-Condon start vector into dst (n doubles), normalised.
hipError_t build_fc_start(const Operator &op, const NvOpDev &dev, double *dst, hipStream_t s, std::string &err);

}  // namespace nv
