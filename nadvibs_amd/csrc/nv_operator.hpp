// Host-side construction of the vibronic operator from the raw nadvibs.in
// data (replaces read_constants' numbering, Source/top.f90:150-234, the
// hoelem table, Source/top.f90:501-527 / Source/function.f90:186-250, the
// diagonal block list, Source/top.f90:528-553, and the step enumeration of
// compute_allIndex, Source/secondary.f90:3-226).
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include "nadvibs.h"
#include "nv_device.hpp"

namespace nv {

// One stored potential term (a column of nzindx/nzblks/nzcoef).
struct Term {
    int cnt = 0;                 // distinct modes
    int mode[NV_MAXF] = {0};     // 0-based, descending (the order union() yields)
    int pow[NV_MAXF] = {0};
    int smax[NV_MAXF] = {0};     // largest |Delta| reachable in the basis
    std::vector<int> blocks;     // unique blocks (0-based packed index), first-seen order
    std::vector<double> raw;     // raw coefficient per block
    double dval = 1.0;           // prod_k w_k^{-p_k/2}
    bool diag = false;           // all smax even: has a Delta = 0 part
    int64_t ndir = 0;            // steps including Delta = 0
};

struct Operator {
    int64_t ns = 0, M = 0, ordr = 0, D = 0, N = 0, nstblks = 0, nmax = 0;
    int32_t quirks = NV_QUIRKS_REF;
    std::vector<int64_t> nfunc, shft;
    std::vector<double> omega;
    std::vector<double> concoef;
    std::vector<Term> terms;             // grouped by cnt = 1..ordr
    std::vector<int64_t> nztrms;         // [ordr] terms per cnt
    std::vector<int64_t> ntcom;          // [ordr][nstblks] term-count check

    // device tables (host copies)
    std::vector<NvTerm> dev_terms;
    std::vector<NvStep> dev_steps;
    std::vector<double> cmat;            // [nterms][ns][ns]
    std::vector<double> ftab;            // [ncodes][nmax]
    std::vector<NvDiagTerm> dev_dterms;
    std::vector<double> dcmat;           // [ndterms][ns][ns]
    std::vector<double> dconst;          // [ns][ns]
    int32_t ncodes = 0;
    int64_t nnz_fma = 0;                 // FMAs of one H*v (nonzero structure)
    double step_valid = 1.0;             // share of (row, step) pairs inside the basis
    std::vector<int32_t> term_dev;       // [terms] index into dev_terms, -1 if the term has no Delta != 0 step
    std::vector<int32_t> term_ddev;      // [terms] index into dev_dterms, -1 if it has no Delta = 0 part
    // the off-diagonal terms by step shape (valid when every term has one)
    bool shaped = false;
    std::vector<NvTermShape> dev_shapes;  // [dev_terms] grouped by shape
    int32_t shape_grp[NV_NSHAPES + 1] = {0};
    // two-pass H*v tables (modes 0 and 1 split off; see NvHv2Dev)
    bool hv2 = false;
    std::vector<NvPhi> hv2_phi;
    int32_t hv2_pgrp[NV_HV2_NSHAPES + 1] = {0};
    int32_t hv2_idcm[NV_HV2_NG] = {0};
    int32_t hv2_gmask = 0;
    NvSigma hv2_sig[NV_HV2_NG];
    // fast-block H*v tables (see NvFbDev)
    bool fb = false;
    int32_t fb_mf = 0, fb_P = 0, fb_halo = 0, fb_pad = 0, fb_nch = 0;
    std::vector<int32_t> fb_tabmode;       // [ntab] mode of each factor row
    std::vector<double> fb_tab;            // [ntab][nmax][4]
    std::vector<NvFbS> fb_S;               // grouped by phase, then by slow shape
    int32_t fb_sgrp[3][NV_FB_NSLOW + 1] = {};
    int32_t fb_chs[NV_FB_NCH + 1] = {0};
    std::vector<NvStep> fb_steps;          // fast steps of the channels, then of the pure terms
    std::vector<NvFbPure> fb_pure;

    // start vector data
    int32_t neworigin = 0;
    std::vector<double> statew, bomega, dvec, tmat;
    std::vector<int64_t> npirrep, istate;
};

// <m|q^p|m+delta> in the reference's conventions (delta >= 0, m 0-based).
double hoelem(int p, int64_t m, int delta, int32_t quirks);

// noterms(p) = C(nmodes+p-1, p).
int64_t noterms(int64_t nmodes, int64_t p);

// The monomials of order p in nadvibs.in order (non-increasing mode tuples,
// first index fastest, Source/top.f90:153-168); 0-based modes.
std::vector<std::vector<int>> monomials(int64_t nmodes, int64_t p);

// Builds the operator; returns false and sets err on invalid input.
bool build_operator(const nv_problem &prob, Operator &op, std::string &err);

}  // namespace nv
