// Operator construction from the raw nadvibs.in data.  See nv_operator.hpp.
#include "nv_operator.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <tuple>

namespace nv {

// Matrix element <m|q^p|m+delta>, q = (a + a^dagger)/sqrt(2), for the
// (p, delta) pairs the steps can reach.  The arithmetic form of every
// expression is the one of Source/function.f90:198-246 (m there is the
// 0-based quantum of the lower function) so that the table rounds the same
// way.  With NV_QUIRK_Q3ELEM the p=3, delta=1 element keeps the reference's
// missing factor 3 (Source/function.f90:225).  Powers above 4 are zero in the
// reference (case default); the exact mode evaluates them by ladder algebra.
static double ladder_power_element(int p, int64_t m, int delta) {
    // <m|q^p|m+delta> by repeated application of q on a basis big enough to
    // hold every intermediate state.
    const int64_t L = m + delta + p + 2;
    std::vector<double> v(L, 0.0), w(L, 0.0);
    v[m + delta] = 1.0;  // |m+delta>
    for (int s = 0; s < p; ++s) {
        std::fill(w.begin(), w.end(), 0.0);
        for (int64_t n = 0; n < L; ++n) {
            if (v[n] == 0.0) continue;
            if (n + 1 < L) w[n + 1] += v[n] * std::sqrt((n + 1) / 2.0);
            if (n >= 1) w[n - 1] += v[n] * std::sqrt(n / 2.0);
        }
        v.swap(w);
    }
    return v[m];
}

double hoelem(int p, int64_t mi, int delta, int32_t quirks) {
    const double m = (double)mi;
    switch (p) {
    case 1:
        if (delta == 1) return std::sqrt((m + 1) / 2);
        return 0.0;
    case 2:
        if (delta == 2) return std::sqrt((m + 1) * (m + 2) / 4);
        if (delta == 0) return (2 * m + 1) / 2;
        return 0.0;
    case 3:
        if (delta == 3) return std::sqrt((m + 1) * (m + 2) * (m + 3) / 8);
        if (delta == 1) {
            const double e = std::sqrt(((m + 1) * (m + 1) * (m + 1)) / 8);
            return (quirks & NV_QUIRK_Q3ELEM) ? e : 3.0 * e;
        }
        return 0.0;
    case 4:
        if (delta == 4) return std::sqrt((m + 1) * (m + 2) * (m + 3) * (m + 4)) / 4;
        if (delta == 2) return (2 * m + 3) * std::sqrt((m + 1) * (m + 2)) / 2;
        if (delta == 0) return 3 * (2 * m * m + 2 * m + 1) / 4;
        return 0.0;
    default:
        if (quirks & NV_QUIRK_Q3ELEM) return 0.0;  // reference: powers > 4 vanish
        return ladder_power_element(p, mi, delta);
    }
}

int64_t noterms(int64_t nmodes, int64_t p) {
    // C(nmodes+p-1, p) by exact integer products (Source/function.f90:253-258).
    int64_t r = 1;
    for (int64_t i = 1; i <= p; ++i) r = r * (nmodes - 1 + i) / i;
    return r;
}

std::vector<std::vector<int>> monomials(int64_t nmodes, int64_t p) {
    // nadvibs.in lists a monomial as a non-increasing mode tuple with the
    // first entry running fastest; reversed, these are the non-decreasing
    // tuples in lexicographic order, generated here directly.
    std::vector<std::vector<int>> out;
    std::vector<int> b(p, 0);
    while (true) {
        out.emplace_back(b.rbegin(), b.rend());
        int i = (int)p - 1;
        while (i >= 0 && b[i] == nmodes - 1) --i;
        if (i < 0) break;
        int v = b[i] + 1;
        for (int j = i; j < p; ++j) b[j] = v;
    }
    return out;
}

static int64_t ipow_prod(const std::vector<int64_t> &v) {
    int64_t r = 1;
    for (auto x : v) r *= x;
    return r;
}

// Signature channel of the powers (p0, p1) of modes 0 and 1 (NvHv2Dev).
static int hv2_sig_id(int p0, int p1) {
    if (p1 == 0) return p0;                 // 0..3
    if (p0 == 0) return 3 + p1;             // 4..6
    if (p0 == 1 && p1 == 1) return 7;
    if (p0 == 2 && p1 == 1) return 8;
    if (p0 == 1 && p1 == 2) return 9;
    return -1;
}

// Tables of the two-pass H*v: every term with off-diagonal steps is split
// into its part on modes 0 and 1 (sigma) and its part on the other modes
// (phi).  Each distinct phi gets the shaped descriptor of its steps and, per
// sigma channel, the coupling matrix of the term it forms with that channel.
// Needs the shaped form (ordr <= 3), at least three modes, at most four
// states, nfunc_0 * nfunc_1 <= 256 and a known step shape for every phi.
template <typename CodeOf>
static void build_two_pass(Operator &op, CodeOf &code_of) {
    static const int kShapes[NV_NSHAPES][3] = {{2, 0, 0}, {3, 0, 0}, {4, 0, 0}, {2, 2, 0}, {3, 2, 0},
                                               {2, 3, 0}, {1, 2, 0}, {2, 1, 0}, {2, 2, 2}};
    op.hv2 = false;
    op.hv2_phi.clear();
    op.hv2_gmask = 0;
    for (int g = 0; g < NV_HV2_NG; ++g) op.hv2_idcm[g] = -1;
    for (auto &g : op.hv2_sig) {
        g.s0 = g.s1 = -1;
        for (int i = 0; i < 4; ++i) g.code0[i] = g.code1[i] = 0;
        g.corr_cm = -1;
        g.corr_code = 0;
    }
    if (!op.shaped || op.M < 3 || op.ordr > 3 || op.ns > 4 || op.nfunc[0] * op.nfunc[1] > 256) return;
    struct Tmp {
        int shape;
        NvPhi ph;
    };
    std::vector<Tmp> phis;
    std::map<std::vector<std::pair<int, int>>, int> index;
    for (size_t ti = 0; ti < op.terms.size(); ++ti) {
        if (op.term_dev[ti] < 0) continue;
        const Term &t = op.terms[ti];
        int p0 = 0, p1 = 0, s0 = -1, s1 = -1;
        std::vector<int> rs;  // slots of t on modes >= 2, in t's order
        for (int j = 0; j < t.cnt; ++j) {
            if (t.mode[j] == 0) { p0 = t.pow[j]; s0 = t.smax[j]; }
            else if (t.mode[j] == 1) { p1 = t.pow[j]; s1 = t.smax[j]; }
            else rs.push_back(j);
        }
        const int g = hv2_sig_id(p0, p1);
        if (g < 0 || rs.size() > 3) return;
        NvSigma &sg = op.hv2_sig[g];
        sg.s0 = p0 ? s0 : -1;
        sg.s1 = p1 ? s1 : -1;
        for (int i = 0; p0 && i <= s0; ++i) sg.code0[i] = code_of(0, p0, -s0 + 2 * i);
        for (int i = 0; p1 && i <= s1; ++i) sg.code1[i] = code_of(1, p1, -s1 + 2 * i);
        op.hv2_gmask |= 1 << g;
        const int cm = op.term_dev[ti];
        if (rs.empty()) {
            op.hv2_idcm[g] = cm;
            if ((p0 == 2 && p1 == 0) || (p0 == 0 && p1 == 2)) {
                sg.corr_cm = cm;
                sg.corr_code = code_of(p0 ? 0 : 1, 2, 0);
            }
            continue;
        }
        std::vector<std::pair<int, int>> key;
        for (int j : rs) key.push_back({t.mode[j], t.pow[j]});
        auto it = index.find(key);
        int pi;
        if (it == index.end()) {
            // slots with more than one step form the shape; the others (Delta = 0 only)
            // become diagonal factors, except that a phi with no stepping slot keeps
            // its first one as the shape (1)
            std::vector<int> act, dia;
            for (int j : rs) (t.smax[j] > 0 ? act : dia).push_back(j);
            if (act.empty()) {
                act.push_back(dia.front());
                dia.erase(dia.begin());
            }
            if (dia.size() > 2) return;
            int k[3] = {0, 0, 0};
            for (size_t j = 0; j < act.size(); ++j) k[j] = t.smax[act[j]] + 1;
            int sid = -1;
            for (int q = 0; q < NV_NSHAPES; ++q)
                if (k[0] == kShapes[q][0] && k[1] == kShapes[q][1] && k[2] == kShapes[q][2]) sid = q;
            if (k[0] == 1 && k[1] == 0) sid = NV_NSHAPES;
            if (sid < 0) return;
            Tmp tp;
            tp.shape = sid;
            for (int q = 0; q < NV_HV2_NG; ++q) tp.ph.cm[q] = -1;
            tp.ph.nd = (int32_t)dia.size();
            for (int q = 0; q < 2; ++q) {
                tp.ph.dmode[q] = q < tp.ph.nd ? t.mode[dia[q]] : 0;
                tp.ph.dcode[q] = q < tp.ph.nd ? code_of(t.mode[dia[q]], t.pow[dia[q]], 0) : 0;
            }
            rs = act;
            NvTermShape &sh = tp.ph.sh;
            sh.cm = 0;
            int64_t base = 0;
            for (int j = 0; j < 3; ++j) {
                sh.inc[j] = 0;
                sh.mode[j] = 0;
                for (int i = 0; i < 4; ++i) sh.code[j][i] = 0;
            }
            for (size_t j = 0; j < rs.size(); ++j) {
                const int m = t.mode[rs[j]], pw = t.pow[rs[j]], sm = t.smax[rs[j]];
                base -= (int64_t)sm * op.shft[m];
                sh.inc[j] = (int32_t)(2 * op.shft[m]);
                sh.mode[j] = m;
                for (int i = 0; i <= sm; ++i) sh.code[j][i] = code_of(m, pw, -sm + 2 * i);
            }
            sh.base = (int32_t)base;
            pi = (int)phis.size();
            index[key] = pi;
            phis.push_back(tp);
        } else {
            pi = it->second;
        }
        phis[pi].ph.cm[g] = cm;
    }
    for (int q = 0; q < NV_HV2_NSHAPES; ++q) {
        op.hv2_pgrp[q] = (int32_t)op.hv2_phi.size();
        for (auto &tp : phis)
            if (tp.shape == q) op.hv2_phi.push_back(tp.ph);
    }
    op.hv2_pgrp[NV_HV2_NSHAPES] = (int32_t)op.hv2_phi.size();
    op.hv2 = true;
}

// Tables of the fast-block H*v (NvFbDev).  The fast modes are the trailing
// modes whose basis product P is at most 64 (one block of P rows per wave);
// the form needs 33 <= P <= 64, a slow mode, ordr <= 3, at most NV_FB_NCH
// channels and a known shape for every slow part.  fb_tab is filled from
// ftab once the codes are final: tabcodes[k] lists the ftab rows of
// Delta = -s + 2i of factor row k.
template <typename CodeOf>
static void build_fast_block(Operator &op, CodeOf &code_of, std::vector<std::array<int, 4>> &tabcodes) {
    // slow shapes of NvFbS, in NV_FB_NSLOW order
    static const int kSlow[NV_FB_NSLOW - 1][3] = {{2, 0, 0}, {3, 0, 0}, {4, 0, 0}, {2, 2, 0},
                                                  {3, 2, 0}, {2, 3, 0}, {2, 2, 2}};
    op.fb = false;
    op.fb_S.clear();
    op.fb_steps.clear();
    op.fb_pure.clear();
    op.fb_tabmode.clear();
    tabcodes.clear();
    if (!op.shaped || op.ordr > 3 || op.M < 2) return;
    int mf = (int)op.M;
    int64_t P = 1;
    // as few fast modes as give 33..64 rows (the channels grow with them)
    while (mf > 0 && P < 33 && P * op.nfunc[mf - 1] <= 64) P *= op.nfunc[--mf];
    if (mf < 1 || P < 33) return;
    std::map<std::pair<int, int>, int> tabs;  // (mode, pow) -> factor row
    auto tab_of = [&](int m, int p, int s) {
        auto key = std::make_pair(m, p);
        auto it = tabs.find(key);
        if (it != tabs.end()) return it->second;
        std::array<int, 4> cs = {-1, -1, -1, -1};
        for (int i = 0; i <= s; ++i) cs[i] = code_of(m, p, -s + 2 * i);
        const int k = (int)tabcodes.size();
        tabcodes.push_back(cs);
        op.fb_tabmode.push_back(m);
        tabs[key] = k;
        return k;
    };
    typedef std::vector<std::pair<int, int>> Key;  // (mode, pow) of the slots
    struct Split {
        Key skey, fkey;
        std::vector<int> sact, sdia, fslots;
        bool pure;
    };
    std::vector<Split> sp(op.terms.size());
    // channels: the fast monomials that meet a slow step (channel 0: no fast slot)
    std::map<Key, int> chan;
    chan[Key()] = 0;
    std::vector<Key> chkeys(1);
    std::vector<int> chterm(1, -1);  // a term that carries each channel's fast slots
    for (size_t ti = 0; ti < op.terms.size(); ++ti) {
        if (op.term_dev[ti] < 0) continue;
        const Term &t = op.terms[ti];
        Split &x = sp[ti];
        for (int j = 0; j < t.cnt; ++j) {
            if (t.mode[j] >= mf) {
                x.fkey.push_back({t.mode[j], t.pow[j]});
                x.fslots.push_back(j);
            } else {
                x.skey.push_back({t.mode[j], t.pow[j]});
                (t.smax[j] > 0 ? x.sact : x.sdia).push_back(j);
            }
        }
        x.pure = x.sact.empty();
        if (x.sdia.size() > 2 || x.sact.size() > 3 || x.fslots.size() > NV_MAXF) return;
        // a Delta = 0 step that is diagonal needs no slow step or no fast slot (else degree 4)
        if (t.diag && !x.fkey.empty() && !x.pure) return;
        if (!x.pure && !chan.count(x.fkey)) {
            if ((int)chkeys.size() == NV_FB_NCH) return;
            chan[x.fkey] = (int)chkeys.size();
            chkeys.push_back(x.fkey);
            chterm.push_back((int)ti);
        }
    }
    // channels in order of the degree of their fast monomial, so that a slow monomial of high
    // degree (which pairs with low-degree fast monomials only) visits a short prefix
    {
        std::vector<int> ord(chkeys.size());
        for (size_t c = 0; c < ord.size(); ++c) ord[c] = (int)c;
        auto degree = [&](const Key &k) {
            int d = 0;
            for (auto &mp : k) d += mp.second;
            return d;
        };
        std::stable_sort(ord.begin() + 1, ord.end(),
                         [&](int a, int b) { return degree(chkeys[a]) < degree(chkeys[b]); });
        std::vector<Key> k2;
        std::vector<int> t2;
        for (int c : ord) {
            k2.push_back(chkeys[c]);
            t2.push_back(chterm[c]);
        }
        chkeys.swap(k2);
        chterm.swap(t2);
        for (size_t c = 0; c < chkeys.size(); ++c) chan[chkeys[c]] = (int)c;
    }
    // the fast steps of a term's fast slots (all of them; the Delta_f = 0 one left out when
    // skip0), appended to fb_steps
    int halo = 0;
    auto fast_steps = [&](const Term &t, const std::vector<int> &fs, bool skip0) {
        int d[NV_MAXF];
        const int nfs = (int)fs.size();
        for (int j = 0; j < nfs; ++j) d[j] = -t.smax[fs[j]];
        while (true) {
            bool zero = true;
            int64_t off = 0;
            for (int j = 0; j < nfs; ++j) {
                off += (int64_t)d[j] * op.shft[t.mode[fs[j]]];
                zero = zero && d[j] == 0;
            }
            if (!(skip0 && zero)) {
                NvStep st;
                st.stride = (int32_t)off;
                st.nf = nfs;
                for (int j = 0; j < NV_MAXF; ++j) { st.mode[j] = 0; st.code[j] = 0; }
                for (int j = 0; j < nfs; ++j) {
                    st.mode[j] = t.mode[fs[j]];
                    st.code[j] = code_of(t.mode[fs[j]], t.pow[fs[j]], d[j]);
                }
                op.fb_steps.push_back(st);
                halo = std::max<int>(halo, (int)std::abs(off));
            }
            int j = nfs - 1;
            while (j >= 0 && d[j] >= t.smax[fs[j]]) { d[j] = -t.smax[fs[j]]; --j; }
            if (j < 0) break;
            d[j] += 2;
        }
    };
    op.fb_chs[0] = op.fb_chs[1] = 0;
    for (size_t c = 1; c < chkeys.size(); ++c) {
        fast_steps(op.terms[chterm[c]], sp[chterm[c]].fslots, false);
        op.fb_chs[c + 1] = (int32_t)op.fb_steps.size();
    }
    for (size_t c = chkeys.size(); c < NV_FB_NCH; ++c) op.fb_chs[c + 1] = op.fb_chs[c];
    // slow monomials
    std::vector<std::vector<NvFbS>> byshape(NV_FB_NSLOW);
    std::map<Key, std::pair<int, int>> sindex;  // slow key -> (shape, position)
    int64_t pad = 0;
    for (size_t ti = 0; ti < op.terms.size(); ++ti) {
        if (op.term_dev[ti] < 0) continue;
        const Term &t = op.terms[ti];
        const Split &x = sp[ti];
        const auto ch = chan.find(x.fkey);
        if (x.pure && (t.diag || ch == chan.end())) {
            // applied on its own from x[row0 + p]
            NvFbPure u{};
            u.cm = op.term_dev[ti];
            u.nd = (int32_t)x.sdia.size();
            for (size_t j = 0; j < x.sdia.size(); ++j) u.dtab[j] = tab_of(t.mode[x.sdia[j]], t.pow[x.sdia[j]], 0);
            u.s0 = (int32_t)op.fb_steps.size();
            fast_steps(t, x.fslots, t.diag);
            u.s1 = (int32_t)op.fb_steps.size();
            op.fb_pure.push_back(u);
            continue;
        }
        auto it = sindex.find(x.skey);
        if (it == sindex.end()) {
            NvFbS S{};
            int64_t base = 0;
            int k[3] = {0, 0, 0};
            for (size_t j = 0; j < x.sact.size(); ++j) {
                const int jj = x.sact[j], m = t.mode[jj];
                base -= (int64_t)t.smax[jj] * op.shft[m];
                S.inc[j] = (int32_t)(2 * op.shft[m]);
                S.tab[j] = tab_of(m, t.pow[jj], t.smax[jj]);
                k[j] = t.smax[jj] + 1;
            }
            S.base = (int32_t)base;
            pad = std::max<int64_t>(pad, -base);
            S.nd = (int32_t)x.sdia.size();
            for (size_t j = 0; j < x.sdia.size(); ++j) S.dtab[j] = tab_of(t.mode[x.sdia[j]], t.pow[x.sdia[j]], 0);
            for (int c = 0; c < NV_FB_NCH; ++c) S.cm[c] = -1;
            int q = NV_FB_NSLOW - 1;
            if (!x.sact.empty()) {
                q = -1;
                for (int z = 0; z < NV_FB_NSLOW - 1; ++z)
                    if (k[0] == kSlow[z][0] && k[1] == kSlow[z][1] && k[2] == kSlow[z][2]) q = z;
                if (q < 0) return;
            }
            sindex[x.skey] = {q, (int)byshape[q].size()};
            byshape[q].push_back(S);
            it = sindex.find(x.skey);
        }
        NvFbS &S = byshape[it->second.first][it->second.second];
        S.cm[ch->second] = op.term_dev[ti];
        S.nchk = std::max<int32_t>(S.nchk, ch->second + 1);
        if (ch->second == 0 && t.diag) S.excl0 = 1;
    }
    // phases by the channels a slow monomial feeds: 0 beyond the first four (a fast monomial of
    // degree 2: the slow one has degree <= 1, shapes (2) or ()), 1 within 1..3 (degree <= 2:
    // (2), (3), (2,2) or ()), 2 only channel 0; each phase instantiates the shapes it can hold
    static const bool kPhaseShape[3][NV_FB_NSLOW] = {{1, 0, 0, 0, 0, 0, 0, 1},
                                                     {1, 1, 0, 1, 0, 0, 0, 1},
                                                     {1, 1, 1, 1, 1, 1, 1, 1}};
    for (int ph = 0; ph < 3; ++ph) {
        for (int q = 0; q < NV_FB_NSLOW; ++q) {
            op.fb_sgrp[ph][q] = (int32_t)op.fb_S.size();
            for (auto &S : byshape[q]) {
                const int sp_ = S.nchk > 4 ? 0 : (S.nchk > 1 ? 1 : 2);
                if (sp_ != ph) continue;
                if (!kPhaseShape[ph][q]) return;
                op.fb_S.push_back(S);
            }
        }
        op.fb_sgrp[ph][NV_FB_NSLOW] = (int32_t)op.fb_S.size();
    }
    op.fb_nch = (int32_t)chkeys.size();
    op.fb_mf = mf;
    op.fb_P = (int32_t)P;
    op.fb_halo = halo;
    // out-of-basis slow steps read the first 64 rows of the margin
    op.fb_pad = (int32_t)std::max<int64_t>(pad, 64);
    op.fb = true;
}

bool build_operator(const nv_problem &pr, Operator &op, std::string &err) {
    if (pr.nstates < 1 || pr.nstates > NV_MAXS) { err = "nstates must be in 1..10"; return false; }
    if (pr.nmodes < 1 || pr.nmodes > NV_MAXM) { err = "nmodes must be in 1.." + std::to_string(NV_MAXM); return false; }
    if (pr.ordr < 1 || pr.ordr > NV_MAXF) { err = "ordr must be in 1..4"; return false; }
    if (!pr.nfunc || !pr.aomega || !pr.concoef || !pr.potterms) { err = "null problem array"; return false; }
    op.ns = pr.nstates; op.M = pr.nmodes; op.ordr = pr.ordr; op.quirks = pr.quirks;
    op.nstblks = op.ns * (op.ns + 1) / 2;
    op.nfunc.assign(pr.nfunc, pr.nfunc + op.M);
    op.omega.assign(pr.aomega, pr.aomega + op.M);
    op.concoef.assign(pr.concoef, pr.concoef + op.nstblks);
    op.nmax = 1;
    for (int64_t k = 0; k < op.M; ++k) {
        if (op.nfunc[k] < 1 || op.nfunc[k] > 255) { err = "basis sizes must be in 1..255"; return false; }
        if (!(op.omega[k] > 0)) { err = "frequencies must be positive"; return false; }
        op.nmax = std::max(op.nmax, op.nfunc[k]);
    }
    op.D = ipow_prod(op.nfunc);
    op.N = op.D * op.ns;
    if (op.N >= (int64_t(1) << 31)) { err = "nstates*dimen must be below 2^31"; return false; }
    // shft_k = prod_{j>k} nfunc_j: last mode fastest (Source/top.f90:379-386)
    op.shft.assign(op.M, 1);
    for (int64_t k = op.M - 2; k >= 0; --k) op.shft[k] = op.shft[k + 1] * op.nfunc[k + 1];

    // ---- term numbering (Source/top.f90:150-201) ----
    const int64_t ntmax = noterms(op.M, op.ordr);
    std::vector<std::vector<Term>> bycnt(op.ordr + 1);
    // (modes, powers) -> position in bycnt[cnt], in first-seen order (the reference's
    // grouping; a map keeps it O(T log T) for many modes)
    std::vector<std::map<uint64_t, size_t>> seen(op.ordr + 1);
    for (int64_t b = 0; b < op.nstblks; ++b) {
        for (int64_t p = 1; p <= op.ordr; ++p) {
            auto mons = monomials(op.M, p);
            for (size_t k = 0; k < mons.size(); ++k) {
                double c = pr.potterms[k + (p - 1) * ntmax + b * ntmax * op.ordr];
                if (!(std::fabs(c) >= pr.ztoler)) continue;
                // unique modes in first-appearance order and their powers
                int um[NV_MAXF], up[NV_MAXF], cnt = 0;
                for (int v : mons[k]) {
                    int f = -1;
                    for (int j = 0; j < cnt; ++j) if (um[j] == v) f = j;
                    if (f < 0) { um[cnt] = v; up[cnt] = 1; ++cnt; } else up[f]++;
                }
                auto &list = bycnt[cnt];
                uint64_t key = 0;
                for (int j = 0; j < cnt; ++j) key = (key << 11) | ((uint64_t)um[j] << 3) | (uint64_t)up[j];
                auto found = seen[cnt].find(key);
                Term *t = nullptr;
                if (found != seen[cnt].end()) {
                    t = &list[found->second];
                } else {
                    seen[cnt][key] = list.size();
                    list.emplace_back();
                    t = &list.back();
                    t->cnt = cnt;
                    for (int j = 0; j < cnt; ++j) { t->mode[j] = um[j]; t->pow[j] = up[j]; }
                }
                t->blocks.push_back((int)b);
                t->raw.push_back(c);
            }
        }
    }
    op.nztrms.assign(op.ordr, 0);
    for (int64_t c = 1; c <= op.ordr; ++c) {
        op.nztrms[c - 1] = (int64_t)bycnt[c].size();
        for (auto &t : bycnt[c]) op.terms.push_back(t);
    }

    // block index <-> state pair (cpindx, Source/top.f90:80-85): b = i(i-1)/2+j-1
    std::vector<int> bi(op.nstblks), bj(op.nstblks);
    for (int i = 1; i <= op.ns; ++i)
        for (int j = 1; j <= i; ++j) { bi[i * (i - 1) / 2 + j - 1] = i - 1; bj[i * (i - 1) / 2 + j - 1] = j - 1; }

    // ---- dval scaling, steps, codes ----
    std::map<std::tuple<int, int, int>, int> codes;  // (mode, pow, delta) -> row
    auto code_of = [&](int mode, int p, int delta) {
        auto key = std::make_tuple(mode, p, delta);
        auto it = codes.find(key);
        if (it != codes.end()) return it->second;
        int c = (int)codes.size();
        codes[key] = c;
        return c;
    };
    const bool lower = (op.quirks & NV_QUIRK_LOWERDIAG) != 0;
    const int ns = (int)op.ns;
    op.ntcom.assign(op.ordr * op.nstblks, 0);
    std::vector<char> active(op.nstblks, 0);
    for (int64_t b = 0; b < op.nstblks; ++b) active[b] = std::fabs(op.concoef[b]) > 1e-16;
    op.nnz_fma = 0;
    for (auto &t : op.terms) {
        t.dval = 1.0;
        for (int j = 0; j < t.cnt; ++j) {
            double w = op.omega[t.mode[j]], wp = 1.0;
            for (int e = 0; e < t.pow[j]; ++e) wp *= w;
            t.dval = t.dval / std::sqrt(wp);
        }
        // smax (Source/secondary.f90:29-39)
        t.ndir = 1;
        bool even = true;
        for (int j = 0; j < t.cnt; ++j) {
            int s = t.pow[j];
            while (s > op.nfunc[t.mode[j]] - 1) s -= 2;
            t.smax[j] = s;
            if (s < 0) t.ndir = 0;
            else t.ndir *= (s + 1);
            if (s % 2 != 0) even = false;
        }
        t.diag = even && t.ndir > 0;
        if (t.diag)
            for (int b : t.blocks) active[b] = 1;
    }
    // coupling matrices, steps and the term-count check
    int64_t trm_rows = 0;
    op.term_dev.assign(op.terms.size(), -1);
    op.term_ddev.assign(op.terms.size(), -1);
    // step shapes (s_0+1, s_1+1, s_2+1), in the group order of NvTermShape
    static const int kShapes[NV_NSHAPES][3] = {{2, 0, 0}, {3, 0, 0}, {4, 0, 0}, {2, 2, 0}, {3, 2, 0},
                                               {2, 3, 0}, {1, 2, 0}, {2, 1, 0}, {2, 2, 2}};
    std::vector<std::vector<NvTermShape>> by_shape(NV_NSHAPES);
    op.shaped = true;
    double valid_rows = 0.0;
    int64_t all_steps = 0;
    for (size_t ti = 0; ti < op.terms.size(); ++ti) {
        Term &t = op.terms[ti];
        std::vector<double> C(ns * ns, 0.0);
        int npairs = 0;
        for (size_t q = 0; q < t.blocks.size(); ++q) {
            int b = t.blocks[q];
            double c = t.raw[q] * t.dval;
            C[bi[b] * ns + bj[b]] += c; ++npairs;
            if (bi[b] != bj[b]) { C[bj[b] * ns + bi[b]] += c; ++npairs; }
        }
        int64_t fixed = op.D;
        for (int j = 0; j < t.cnt; ++j) fixed /= op.nfunc[t.mode[j]];
        // enumerate Delta vectors, slot 0 slowest (Source/secondary.f90:93-106)
        NvTerm dt;
        dt.first = (int32_t)op.dev_steps.size();
        dt.nsteps = 0;
        if (t.ndir > 0) {
            int d[NV_MAXF];
            for (int j = 0; j < t.cnt; ++j) d[j] = -t.smax[j];
            while (true) {
                int64_t rows = fixed, stride = 0;
                bool zero = true;
                for (int j = 0; j < t.cnt; ++j) {
                    rows *= op.nfunc[t.mode[j]] - std::abs(d[j]);
                    stride += (int64_t)d[j] * op.shft[t.mode[j]];
                    zero = zero && d[j] == 0;
                }
                for (int b : t.blocks) op.ntcom[(t.cnt - 1) * op.nstblks + b] += rows;
                if (!zero) {
                    NvStep s;
                    s.stride = (int32_t)stride;
                    s.nf = t.cnt;
                    for (int j = 0; j < NV_MAXF; ++j) { s.mode[j] = 0; s.code[j] = 0; }
                    for (int j = 0; j < t.cnt; ++j) {
                        s.mode[j] = t.mode[j];
                        s.code[j] = code_of(t.mode[j], t.pow[j], d[j]);
                    }
                    op.dev_steps.push_back(s);
                    dt.nsteps++;
                    op.nnz_fma += rows * npairs;
                    valid_rows += (double)rows;
                    ++all_steps;
                }
                int j = t.cnt - 1;
                while (j >= 0 && d[j] >= t.smax[j]) { d[j] = -t.smax[j]; --j; }
                if (j < 0) break;
                d[j] += 2;
            }
        }
        if (dt.nsteps > 0) {
            int k[3] = {0, 0, 0};
            for (int j = 0; j < t.cnt && j < 3; ++j) k[j] = t.smax[j] + 1;
            int sid = -1;
            if (t.cnt <= 3)
                for (int q = 0; q < NV_NSHAPES; ++q)
                    if (k[0] == kShapes[q][0] && k[1] == kShapes[q][1] && k[2] == kShapes[q][2]) sid = q;
            if (sid < 0) {
                op.shaped = false;
            } else {
                NvTermShape sh;
                sh.cm = (int32_t)op.dev_terms.size();
                int64_t base = 0;
                for (int j = 0; j < 3; ++j) {
                    sh.inc[j] = 0;
                    sh.mode[j] = 0;
                    for (int i = 0; i < 4; ++i) sh.code[j][i] = 0;
                }
                for (int j = 0; j < t.cnt; ++j) {
                    base -= (int64_t)t.smax[j] * op.shft[t.mode[j]];
                    sh.inc[j] = (int32_t)(2 * op.shft[t.mode[j]]);
                    sh.mode[j] = t.mode[j];
                    for (int i = 0; i <= t.smax[j]; ++i)
                        sh.code[j][i] = code_of(t.mode[j], t.pow[j], -t.smax[j] + 2 * i);
                }
                sh.base = (int32_t)base;
                by_shape[sid].push_back(sh);
            }
            op.term_dev[ti] = (int32_t)op.dev_terms.size();
            op.dev_terms.push_back(dt);
            op.cmat.insert(op.cmat.end(), C.begin(), C.end());
        }
        if (t.diag) {
            NvDiagTerm dd;
            dd.nf = t.cnt;
            for (int j = 0; j < NV_MAXF; ++j) { dd.mode[j] = 0; dd.code[j] = 0; }
            for (int j = 0; j < t.cnt; ++j) { dd.mode[j] = t.mode[j]; dd.code[j] = code_of(t.mode[j], t.pow[j], 0); }
            op.term_ddev[ti] = (int32_t)op.dev_dterms.size();
            op.dev_dterms.push_back(dd);
            std::vector<double> Dm(ns * ns, 0.0);
            for (int s = 0; s < ns; ++s)
                for (int s2 = 0; s2 < ns; ++s2)
                    if (s == s2 || !lower || s > s2) Dm[s * ns + s2] = C[s * ns + s2];
            op.dcmat.insert(op.dcmat.end(), Dm.begin(), Dm.end());
        }
    }
    (void)trm_rows;
    op.step_valid = all_steps > 0 ? valid_rows / ((double)all_steps * (double)op.D) : 1.0;
    op.dev_shapes.clear();
    for (int q = 0; q < NV_NSHAPES; ++q) {
        op.shape_grp[q] = (int32_t)op.dev_shapes.size();
        op.dev_shapes.insert(op.dev_shapes.end(), by_shape[q].begin(), by_shape[q].end());
    }
    op.shape_grp[NV_NSHAPES] = (int32_t)op.dev_shapes.size();
    if (!op.shaped) {
        op.dev_shapes.clear();
        for (int q = 0; q <= NV_NSHAPES; ++q) op.shape_grp[q] = 0;
    }
    build_two_pass(op, code_of);
    std::vector<std::array<int, 4>> fb_tabcodes;
    build_fast_block(op, code_of, fb_tabcodes);
    // constant block matrix, gated by the diagonal-block list (Source/top.f90:528-553)
    op.dconst.assign(ns * ns, 0.0);
    int napplied = 0;
    for (int64_t b = 0; b < op.nstblks; ++b) {
        int i = bi[b], j = bj[b];
        if (i == j) { op.dconst[i * ns + i] = op.concoef[b]; continue; }
        if (!active[b]) continue;
        op.dconst[i * ns + j] = op.concoef[b];
        ++napplied;
        if (!lower) { op.dconst[j * ns + i] = op.concoef[b]; ++napplied; }
    }
    op.nnz_fma += op.D * (op.ns + napplied);
    // factor table ftab[code][n]
    op.ncodes = (int32_t)codes.size();
    op.ftab.assign((size_t)op.ncodes * op.nmax, 0.0);
    for (auto &kv : codes) {
        int mode = std::get<0>(kv.first), p = std::get<1>(kv.first), delta = std::get<2>(kv.first);
        int64_t nf = op.nfunc[mode];
        for (int64_t n = 0; n < nf; ++n) {
            int64_t n2 = n + delta;
            if (n2 < 0 || n2 >= nf) continue;
            op.ftab[(size_t)kv.second * op.nmax + n] = hoelem(p, std::min(n, n2), std::abs(delta), op.quirks);
        }
    }
    op.fb_tab.assign(fb_tabcodes.size() * op.nmax * 4, 0.0);
    for (size_t k = 0; k < fb_tabcodes.size(); ++k)
        for (int64_t n = 0; n < op.nmax; ++n)
            for (int i = 0; i < 4; ++i)
                if (fb_tabcodes[k][i] >= 0)
                    op.fb_tab[(k * op.nmax + n) * 4 + i] = op.ftab[(size_t)fb_tabcodes[k][i] * op.nmax + n];

    // ---- start-vector data ----
    op.neworigin = pr.neworigin;
    op.statew.assign(op.ns, 1.0);
    if (pr.statew) op.statew.assign(pr.statew, pr.statew + op.ns);
    op.istate.assign(op.M, 0);
    if (pr.istate) op.istate.assign(pr.istate, pr.istate + op.M);
    if (pr.nirreps >= 1 && pr.npirrep) {
        op.npirrep.assign(pr.npirrep, pr.npirrep + pr.nirreps);
        int64_t tot = 0;
        for (auto v : op.npirrep) tot += v;
        if (tot != op.M) op.npirrep.assign(1, op.M);  // symmetry disabled (Source/secondary.f90:848-852)
    } else {
        op.npirrep.assign(1, op.M);
    }
    if (op.neworigin) {
        if (!pr.bomega || !pr.dvec || !pr.tmat) { err = "neworigin needs bomega, dvec and tmat"; return false; }
        op.bomega.assign(pr.bomega, pr.bomega + op.M);
        op.dvec.assign(pr.dvec, pr.dvec + op.M);
        op.tmat.assign(pr.tmat, pr.tmat + op.M * op.M);
    }
    return true;
}

}  // namespace nv
