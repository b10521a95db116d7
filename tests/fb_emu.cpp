// CPU emulation of the fast-block H*v's table walk, for the CPU test suite.
//
// Builds the operator with the library's own host code (nv_operator.cpp), then
// runs the walk the fast-block kernel does over the tables it gets (fb_S, its
// three phases and shape groups, the channels, the fast step lists and the pure
// terms) one block of P rows at a time: slow steps read the padded interleaved
// copy of x through the byte offsets and validity masks the kernel computes,
// fast steps read a row of 64 + 2 halo lanes with zero margins.  Every read is
// bounds-checked against the buffers the kernel has (the padded copy, the LDS
// row), and the phase rule (a monomial in phase 1 feeds channels 0..3 only, in
// phase 2 channel 0 only) is checked too.  The result is compared with the
// operator's off-diagonal step list applied row by row.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "nv_operator.hpp"

using namespace nv;

namespace {

// Stepping slots (K0, K1, K2) of the NV_FB_NSLOW slow shapes, in group order.
const int kShape[NV_FB_NSLOW][3] = {{2, 0, 0}, {3, 0, 0}, {4, 0, 0}, {2, 2, 0},
                                    {3, 2, 0}, {2, 3, 0}, {2, 2, 2}, {0, 0, 0}};

double xval(int64_t i) { return std::sin(0.37 * (double)i + 0.1) + 0.01 * (double)(i % 7); }

}  // namespace

// Returns the largest |emulated - step list| over the sampled rows relative to
// the largest |step list| value, or -1 if the operator does not build.  *on is
// the operator's fast-block flag, *P its rows per block, *bad the count of
// out-of-bounds reads and phase violations.  At most maxblocks blocks, equally
// spaced, are walked.
extern "C" double fb_emulate(const nv_problem *prob, long maxblocks, int *on, int *P, int *bad) {
    Operator op;
    std::string err;
    *on = 0;
    *P = 0;
    *bad = 0;
    if (!build_operator(*prob, op, err)) return -1.0;
    *on = op.fb;
    *P = op.fb_P;
    if (!op.fb) return 0.0;
    const int ns = (int)op.ns, nsp = nv_lo_stride(ns);
    const int64_t D = op.D, pad = op.fb_pad;
    const int Pb = op.fb_P, halo = op.fb_halo;
    const int64_t nblk = D / Pb;
    // the interleaved copy with its zero margins, x[(pad + r) * nsp + s] = x_s(r)
    std::vector<double> xb((size_t)(D + 2 * pad) * nsp, 0.0);
    for (int64_t r = 0; r < D; ++r)
        for (int s = 0; s < ns; ++s) xb[(size_t)(r + pad) * nsp + s] = xval(s * D + r);
    auto digit = [&](int64_t r, int m) { return (int)((r / op.shft[m]) % op.nfunc[m]); };
    auto ftab = [&](int code, int n) { return op.ftab[(size_t)code * op.nmax + n]; };
    const int ntab = (int)op.fb_tabmode.size();
    // a slow load: row so_rows + lane of the padded copy into w (h times)
    auto load = [&](int lane, int64_t so_rows, double h, double *w) {
        if (lane >= Pb) return;
        const int64_t rr = so_rows + lane;
        if (rr < 0 || rr >= D + 2 * pad) {
            ++*bad;
            return;
        }
        for (int s = 0; s < ns; ++s) w[s] += h * xb[(size_t)rr * nsp + s];
    };
    double maxdiff = 0.0, maxref = 0.0;
    const int64_t stride = nblk > maxblocks ? nblk / maxblocks : 1;
    for (int64_t blk = 0; blk < nblk; blk += stride) {
        const int64_t row0 = blk * Pb;
        // the wave's factor rows at its slow digits, and their masks of in-basis slots
        std::vector<double> fw((size_t)ntab * 4);
        std::vector<int> fwm(ntab);
        for (int k = 0; k < ntab; ++k) {
            const int n = digit(row0, op.fb_tabmode[k]);
            int m = 0;
            for (int i = 0; i < 4; ++i) {
                fw[k * 4 + i] = op.fb_tab[((size_t)k * op.nmax + n) * 4 + i];
                m |= (fw[k * 4 + i] != 0.0) << i;
            }
            fwm[k] = m;
        }
        // channels V[c][lane][s]
        std::vector<double> V((size_t)NV_FB_NCH * 64 * ns, 0.0);
        for (int lane = 0; lane < 64; ++lane)
            for (int ph = 0; ph < 3; ++ph)
                for (int q = 0; q < NV_FB_NSLOW; ++q) {
                    const int K0 = kShape[q][0], K1 = kShape[q][1], K2 = kShape[q][2];
                    const int J0 = K0 ? K0 : 1, J1 = K1 ? K1 : 1, J2 = K2 ? K2 : 1;
                    const bool odd = K0 > 0 && (K0 & 1) && (K1 == 0 || (K1 & 1)) && (K2 == 0 || (K2 & 1));
                    for (int t = op.fb_sgrp[ph][q]; t < op.fb_sgrp[ph][q + 1]; ++t) {
                        const NvFbS &d = op.fb_S[t];
                        const int cfirst = ph == 0 ? NV_FB_NCH : ph == 1 ? 4 : 1;
                        if (lane == 0)
                            for (int c = cfirst; c < NV_FB_NCH; ++c)
                                if (d.cm[c] >= 0) ++*bad;
                        double w[NV_MAXS] = {0}, wc[NV_MAXS] = {0};
                        for (int i0 = 0; i0 < J0; ++i0)
                            for (int i1 = 0; i1 < J1; ++i1)
                                for (int i2 = 0; i2 < J2; ++i2) {
                                    double h = 1.0;
                                    int64_t off = 0;
                                    bool in = true;
                                    if (K0) {
                                        h = fw[d.tab[0] * 4 + i0];
                                        off = d.base + (int64_t)i0 * d.inc[0];
                                        in = (fwm[d.tab[0]] >> i0) & 1;
                                    }
                                    if (K1) {
                                        h *= fw[d.tab[1] * 4 + i1];
                                        off += (int64_t)i1 * d.inc[1];
                                        in = in && ((fwm[d.tab[1]] >> i1) & 1);
                                    }
                                    if (K2) {
                                        h *= fw[d.tab[2] * 4 + i2];
                                        off += (int64_t)i2 * d.inc[2];
                                        in = in && ((fwm[d.tab[2]] >> i2) & 1);
                                    }
                                    const bool mid = odd && 2 * i0 + 1 == K0 && (K1 == 0 || 2 * i1 + 1 == K1) &&
                                                     (K2 == 0 || 2 * i2 + 1 == K2);
                                    // out-of-basis steps read the zero margin at offset 0
                                    load(lane, in ? row0 + pad + off : 0, h, mid ? wc : w);
                                }
                        if (d.nd > 0) {
                            double fd = fw[d.dtab[0] * 4];
                            if (d.nd > 1) fd *= fw[d.dtab[1] * 4];
                            for (int s = 0; s < ns; ++s) {
                                w[s] *= fd;
                                wc[s] *= fd;
                            }
                        }
                        for (int c = 0; c < NV_FB_NCH; ++c) {
                            if (d.cm[c] < 0) continue;
                            const bool ex = c == 0 && odd && d.excl0;
                            const double *C = &op.cmat[(size_t)d.cm[c] * ns * ns];
                            for (int s = 0; s < ns; ++s) {
                                double a = 0.0;
                                for (int s2 = 0; s2 < ns; ++s2) a += C[s * ns + s2] * (ex ? w[s2] : w[s2] + wc[s2]);
                                V[((size_t)c * 64 + lane) * ns + s] += a;
                            }
                        }
                    }
                }
        // channel 0 is y; the others and the pure terms go through a row with zero margins
        std::vector<double> y(V.begin(), V.begin() + 64 * ns);
        std::vector<double> vrow((size_t)(64 + 2 * halo) * ns, 0.0);
        auto fast = [&](int s0, int s1, const std::vector<double> &src) {
            for (int lane = 0; lane < 64; ++lane)
                for (int s = 0; s < ns; ++s) vrow[(size_t)(halo + lane) * ns + s] = src[(size_t)lane * ns + s];
            for (int lane = 0; lane < 64; ++lane) {
                const int64_t r = row0 + lane;
                const bool live = lane < Pb;
                for (int k = s0; k < s1; ++k) {
                    const NvStep &st = op.fb_steps[k];
                    double h = 1.0;
                    for (int j = 0; j < st.nf; ++j) h *= ftab(st.code[j], live ? digit(r, st.mode[j]) : 0);
                    const int idx = halo + lane + st.stride;
                    if (idx < 0 || idx >= 64 + 2 * halo) {
                        ++*bad;
                        continue;
                    }
                    for (int s = 0; s < ns; ++s) y[(size_t)lane * ns + s] += h * vrow[(size_t)idx * ns + s];
                }
            }
        };
        for (int c = 1; c < op.fb_nch; ++c) {
            std::vector<double> src(V.begin() + (size_t)c * 64 * ns, V.begin() + (size_t)(c + 1) * 64 * ns);
            fast(op.fb_chs[c], op.fb_chs[c + 1], src);
        }
        for (const NvFbPure &u : op.fb_pure) {
            double fd = 1.0;
            if (u.nd > 0) fd = fw[u.dtab[0] * 4];
            if (u.nd > 1) fd *= fw[u.dtab[1] * 4];
            const double *C = &op.cmat[(size_t)u.cm * ns * ns];
            std::vector<double> src((size_t)64 * ns, 0.0);
            for (int lane = 0; lane < 64; ++lane) {
                double xr[NV_MAXS] = {0};
                load(lane, row0 + pad, 1.0, xr);
                for (int s = 0; s < ns; ++s) {
                    double a = 0.0;
                    for (int s2 = 0; s2 < ns; ++s2) a += C[s * ns + s2] * xr[s2];
                    src[(size_t)lane * ns + s] = fd * a;
                }
            }
            fast(u.s0, u.s1, src);
        }
        // the reference: every off-diagonal step of every term, row by row
        for (int lane = 0; lane < Pb; ++lane) {
            const int64_t r = row0 + lane;
            double yref[NV_MAXS] = {0};
            for (size_t t = 0; t < op.dev_terms.size(); ++t) {
                const NvTerm &dt = op.dev_terms[t];
                const double *C = &op.cmat[t * ns * ns];
                for (int q = 0; q < dt.nsteps; ++q) {
                    const NvStep &st = op.dev_steps[dt.first + q];
                    double f = 1.0;
                    for (int j = 0; j < st.nf; ++j) f *= ftab(st.code[j], digit(r, st.mode[j]));
                    if (f == 0.0) continue;
                    const int64_t r2 = r + st.stride;
                    for (int s = 0; s < ns; ++s)
                        for (int s2 = 0; s2 < ns; ++s2) yref[s] += C[s * ns + s2] * f * xval(s2 * D + r2);
                }
            }
            for (int s = 0; s < ns; ++s) {
                maxdiff = std::fmax(maxdiff, std::fabs(y[(size_t)lane * ns + s] - yref[s]));
                maxref = std::fmax(maxref, std::fabs(yref[s]));
            }
        }
    }
    return maxdiff / (maxref > 0.0 ? maxref : 1.0);
}
