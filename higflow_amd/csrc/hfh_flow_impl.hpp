// Flow<D> methods: viscoelastic recording, device layout and the step.
#pragma once

namespace hfh {

// hig-flow-step-generalized-newtonian.c:14-93 (velocity gradient rows) and the
// point evaluations of hig-flow-step-viscoelastic.c:352-544 (first tensor
// component; the other components revisit the same cached points).
template <int D>
void Flow<D>::record_ve() {
    Domain<D> &se = *sdED_;
    int ne = se.size();
    // velocity-derivative tensor, one combined row per (cell, i, j)
    for (int c = 0; c < ne; c++) {
        if ((c & 1023) == 0) progress("tensor rows", c, ne);
        Vec<D> cc = se.center_of(c), cd = se.delta_of(c);
        for (int dim = 0; dim < D; dim++)
            for (int dim2 = 0; dim2 < D; dim2++) {
                Domain<D> &fu = *sfdu_[dim];
                Combo cb;
                double k = 1.0 / (2.0 * 0.5 * cd[dim2]);
                if (dim == dim2) {
                    Vec<D> pl = cc, pr = cc;
                    pl[dim] = cc[dim] - 0.5 * cd[dim];
                    pr[dim] = cc[dim] + 0.5 * cd[dim];
                    Stencil sl = stencil_at(fu, cc, pl), sr = stencil_at(fu, cc, pr);
                    cb.add(sr, oU_[dim], k);
                    cb.add(sl, oU_[dim], -k);
                } else {
                    // hig-flow-eval.c:160-186 then :95-121 (rr, r, l, ll)
                    for (int side = 0; side < 2; side++) {
                        double sg = side == 0 ? -1.0 : 1.0;
                        Vec<D> p = cc;
                        Stencil s4[4];
                        p[dim] = cc[dim] + 0.5 * 1.0 * cd[dim];
                        p[dim2] = side == 0 ? cc[dim2] - 1.0 * cd[dim2] : cc[dim2] + 1.0 * cd[dim2];
                        s4[0] = stencil_at(fu, cc, p);
                        p[dim] = cc[dim] + 0.5 * 1.0 * cd[dim];
                        p[dim2] = cc[dim2];
                        s4[1] = stencil_at(fu, cc, p);
                        p[dim] = cc[dim] - 0.5 * 1.0 * cd[dim];
                        p[dim2] = cc[dim2];
                        s4[2] = stencil_at(fu, cc, p);
                        p[dim] = cc[dim] - 0.5 * 1.0 * cd[dim];
                        p[dim2] = side == 0 ? cc[dim2] - 1.0 * cd[dim2] : cc[dim2] + 1.0 * cd[dim2];
                        s4[3] = stencil_at(fu, cc, p);
                        for (int q = 0; q < 4; q++) cb.add(s4[q], oU_[dim], sg * 0.25 * k);
                    }
                }
                combo_row(gDu_, cb, oDv_ + (dim * D + dim2) * ne + c);
            }
    }
    // constitutive point evaluations
    veSlots_.assign((size_t)(6 * D) * ne, oZero_placeholder());
    veFlags_.assign(ne, 0);
    veDelta_.assign((size_t)D * ne, 0.0);
    for (int c = 0; c < ne; c++) {
        Vec<D> cc = se.center_of(c), cd = se.delta_of(c);
        for (int d = 0; d < D; d++) veDelta_[(size_t)d * ne + c] = cd[d];
        auto S = [&](int k) -> int & { return veSlots_[(size_t)k * ne + c]; };
        for (int dim = 0; dim < D; dim++) {  // hig_flow_velocity_at_center_cell
            Vec<D> pl = cc, pr = cc;
            pl[dim] = cc[dim] - 0.5 * cd[dim];
            pr[dim] = cc[dim] + 0.5 * cd[dim];
            S(2 * dim) = point_slot(*sfdu_[dim], oU_[dim], cc, pl, gCv_);
            S(2 * dim + 1) = point_slot(*sfdu_[dim], oU_[dim], cc, pr, gCv_);
        }
        unsigned fl = 0;
        for (int dim = 0; dim < D; dim++) {
            const int kb = 2 * D + 4 * dim;
            double al[4] = {1.0, 1.0, 2.0, 2.0};
            double sg[4] = {-1.0, 1.0, -1.0, 1.0};
            int nk = par_.ve_convec == CELL_CUBISTA ? 4 : 2;
            for (int q = 0; q < nk; q++) {
                Vec<D> p = cc;
                p[dim] = sg[q] > 0 ? cc[dim] + al[q] * cd[dim] : cc[dim] - al[q] * cd[dim];
                if (se.cell_with_point(p) >= 0) fl |= 1u << (4 * dim + q);
                S(kb + q) = point_slot(se, oK_, cc, p, gK_, true);
            }
            if (par_.ve_convec == CELL_CUBISTA) {
                Vec<D> pl = cc, pr = cc;
                pl[dim] = cc[dim] - 0.5 * cd[dim];
                pr[dim] = cc[dim] + 0.5 * cd[dim];
                (void)point_slot(*sfdu_[dim], oU_[dim], cc, pl, gCv_);
                (void)point_slot(*sfdu_[dim], oU_[dim], cc, pr, gCv_);
            }
        }
        veFlags_[c] = fl;
    }
}

template <int D>
void Flow<D>::finalize_layout() {
    int ne = sdED_->size();
    gK_.ncomp = NSYM;
    gK_.in_stride = ne;
    oPK_ = nX_;
    nX_ += NSYM * gK_.nrows();
    oPcv_ = nX_;
    nX_ += gCv_.nrows();
    oPmom_ = nX_;
    nX_ += gMom_.nrows();
    oZero_ = nX_;
    nX_ += 1;
    for (int r = 0; r < gK_.nrows(); r++) gK_.out[r] = oPK_ + r;
    for (int r = 0; r < gCv_.nrows(); r++) gCv_.out[r] = oPcv_ + r;
    for (int r = 0; r < gMom_.nrows(); r++) gMom_.out[r] = oPmom_ + r;
    auto resolve = [&](std::vector<int> &v, int base) {
        for (auto &s : v) {
            if (s == UNUSED) s = oZero_;
            else if (s < 0) s = base + (-s - 1);
        }
    };
    for (int d = 0; d < D; d++) resolve(momSlots_[d], oPmom_);
    for (size_t i = 0; i < veSlots_.size(); i++) {
        int k = (int)(i / (ne > 0 ? ne : 1));
        int &s = veSlots_[i];
        if (s == UNUSED) s = (k < 2 * D) ? oZero_ : oPK_;  // K slots stay valid for every component
        else if (s < 0) s = (k < 2 * D ? oPcv_ : oPK_) + (-s - 1);
    }
    validate();
    // algorithmic bytes per launch (fields read once + operator streamed)
    auto gbytes = [&](const Group &g) {
        return (long)g.col.size() * 12 + (long)g.nrows() * (8 + 8 + 4 + 4) * g.ncomp;
    };
    bytes_[ST_INTERP] = gbytes(gMom_) + gbytes(gDu_) + gbytes(gK_) + gbytes(gCv_);
    long mb = 0;
    for (int d = 0; d < D; d++) mb += (long)Nf_[d] * ((long)momNs_ * (4 + 8) + 4 + D * 8 + 8 + 8);
    bytes_[ST_MOMENTUM] = mb;
    bytes_[ST_DIV] = gbytes(gDiv_);
    bytes_[ST_PROJECT] = gbytes(gGrad_) + (long)gGrad_.nrows() * 12 + (long)Nc_ * 24;
    bytes_[ST_SOLVE] = (long)Aci_.size() * 12 + (long)Nc_ * 8 * 9;  // one CG iteration
    bytes_[ST_VE] = (long)ne * (D * D * 8 * 2 + NSYM * 8 * 2 + 6 * D * (4 + 8) + 4 + D * 8);
    if (par_.tempdiscrtype == SEMI_IMPLICIT_EULER) {
        for (int d = 0; d < D; d++) {
            std::vector<std::map<int, double>> m(Nf_[d]);
            for (int r = 0; r < Nf_[d]; r++)
                for (int k = Hrp_[d][r]; k < Hrp_[d][r + 1]; k++) m[r][Hci_[d][k]] = Hv_[d][k];
            Hsym_[d] = true;
            for (int r = 0; r < Nf_[d] && Hsym_[d]; r++)
                for (auto &e : m[r]) {
                    auto it = m[e.first].find(r);
                    double o = it == m[e.first].end() ? 0.0 : it->second;
                    if (std::fabs(o - e.second) > 1e-12 * std::max(1.0, std::fabs(e.second))) {
                        Hsym_[d] = false;
                        break;
                    }
                }
        }
    }
    if (host_only_) return;
    upload_device();
}

// Device tables from the host tables (single run or merged batch).
template <int D>
void Flow<D>::upload_device() {
    // device upload
    std::vector<double> hx(nX_, 0.0);
    X_.alloc(nX_);
    X_.upload(hx.data(), nX_, st_);
    bvec_.alloc(Nc_);
    // the divergence rows write every cell's entry; the padding entries of a
    // batch (between member blocks) are never written and must read zero
    k_fill<double><<<nblocks(Nc_), BLOCK, 0, st_>>>(Nc_, bvec_.p, 0.0);
    for (int d = 0; d < D; d++) {
        dMomFlags_[d].upload(momFlags_[d].data(), momFlags_[d].size(), st_);
        dMomDelta_[d].upload(momDelta_[d].data(), momDelta_[d].size(), st_);
    }
    if (!veSlots_.empty()) {
        dVeFlags_.upload(veFlags_.data(), veFlags_.size(), st_);
        dVeDelta_.upload(veDelta_.data(), veDelta_.size(), st_);
    }
    {
        // before the groups: the FP32 form uploads FP32 weights
        const char *mk = std::getenv("HFH_MIXED_KERNELS");
        k32_ = par_.precision == 1 && mk && std::strcmp(mk, "1") == 0;
        // HFH_VERBOSE=1: print each step's velocity range as the reference's
        // final-velocity step does (hig-flow-step.c:302-306)
        const char *ev = std::getenv("HFH_VERBOSE");
        verbose_ = ev && std::strcmp(ev, "1") == 0;
    }
    upload_group(dDu_, gDu_, 0, 0, false, k32_);
    upload_group(dK_, gK_, 0, gK_.nrows(), false, k32_);
    upload_group(dCv_, gCv_, 0, 0, false, k32_);
    upload_group(dMom_, gMom_, 0, 0, false, k32_);
    upload_group(dDiv_, gDiv_, 0, 0, false);
    // The point-value regions (momentum points, K at CUBISTA points,
    // cell-centre velocities) are laid out in the device row order of their
    // groups, so each group writes its region contiguously; the slot tables
    // of the momentum and constitutive kernels are relabelled to match
    // (rows of one template hold one point kind of consecutive elements, so
    // a wave's slot reads stay contiguous too).
    {
        std::vector<int> ms[D], vs = veSlots_;
        for (int d = 0; d < D; d++) ms[d] = momSlots_[d];
        auto relabel = [&](DGroup &dg, int oP, std::vector<int> *tabs, int ntab) {
            const int n = dg.nrow;
            if (n == 0 || (int)dg.order.size() != n) return;
            std::vector<int> inv(n), outp(n);
            for (int k = 0; k < n; k++) {
                inv[dg.order[k]] = k;
                outp[k] = oP + k;
            }
            for (int q = 0; q < ntab; q++)
                for (int &v : tabs[q])
                    if (v >= oP && v < oP + n) v = oP + inv[v - oP];
            dg.out.upload(outp.data(), outp.size(), st_);
            HFH_CHECK(hipStreamSynchronize(st_));  // outp is a local
        };
        relabel(dMom_, oPmom_, ms, D);
        relabel(dK_, oPK_, &vs, 1);
        relabel(dCv_, oPcv_, &vs, 1);
        for (int d = 0; d < D; d++) dMomSlots_[d].upload(ms[d].data(), ms[d].size(), st_);
        if (!vs.empty()) dVeSlots_.upload(vs.data(), vs.size(), st_);
        HFH_CHECK(hipStreamSynchronize(st_));  // the tables are locals
    }
    // projection rows: u = us - dt * grad
    {
        std::vector<int> y(gGrad_.nrows());
        for (int r = 0; r < gGrad_.nrows(); r++) {
            int o = gGrad_.out[r];
            for (int d = D - 1; d >= 0; d--)
                if (o >= oU_[d]) {
                    y[r] = oUS_[d] + (o - oU_[d]);
                    break;
                }
        }
        upload_group(dGrad_, gGrad_, 0, 0, true);
        std::vector<int> yp(y.size());
        for (size_t k = 0; k < yp.size(); k++) yp[k] = y[dGrad_.order[k]];
        y.swap(yp);
        dGrad_.y.upload(y.data(), y.size(), st_);
        HFH_CHECK(hipStreamSynchronize(st_));  // y is a local
    }
    // algorithmic bytes of the interpolation stages with the operators as
    // uploaded (template rows or CSR)
    bytes_[ST_INTERP] = gdbytes(dMom_) + gdbytes(dDu_) + gdbytes(dK_) + gdbytes(dCv_);
    bytes_[ST_DIV] = gdbytes(dDiv_);
    bytes_[ST_PROJECT] = gdbytes(dGrad_) + (long)dGrad_.nrow * 12 + (long)Nc_ * 24;
    {
        const char *ex = std::getenv("HFH_EXTRAP");
        extrap_ = ex ? std::max(0, std::min(2, std::atoi(ex))) : 2;
        extrap_hist_ = 0;
        pPrev_.alloc(Nc_);
        pPrev2_.alloc(Nc_);
        k_fill<double><<<nblocks(Nc_), BLOCK, 0, st_>>>(Nc_, pPrev_.p, 0.0);
        k_fill<double><<<nblocks(Nc_), BLOCK, 0, st_>>>(Nc_, pPrev2_.p, 0.0);
    }
    for (int d = 0; d < D; d++)
        if (!momRe_[d].empty()) dMomRe_[d].upload(momRe_[d].data(), momRe_[d].size(), st_);
    if (!vePar_.empty()) dVePar_.upload(vePar_.data(), vePar_.size(), st_);
    psys_.set_matrix(Arp_, Aci_, Av_, st_, true, pseg_);
    if (par_.tempdiscrtype == SEMI_IMPLICIT_EULER) {
        const char *ek = std::getenv("HFH_KRYLOV");
        hbatched_ = !(ek && std::strcmp(ek, "classic") == 0);
        for (int d = 0; d < D; d++)
            if (Nf_[d] == 0) hbatched_ = false;
        for (int d = 0; d < D; d++) hcst_[d].alloc(Nf_[d]);
        if (hbatched_) {
            // block-diagonal system of the D components
            std::vector<int> rp(1, 0), ci, segs(1, 0);
            std::vector<double> v;
            for (int d = 0; d < D; d++) {
                hoff_[d] = (int)rp.size() - 1;
                // batched runs: one segment per member inside the component
                for (size_t q = 1; q < hseg_[d].size(); q++) segs.push_back(hoff_[d] + hseg_[d][q]);
                for (int i = 0; i < Nf_[d]; i++) {
                    for (int k = Hrp_[d][i]; k < Hrp_[d][i + 1]; k++) {
                        ci.push_back(Hci_[d][k] + hoff_[d]);
                        v.push_back(Hv_[d][k]);
                    }
                    rp.push_back((int)ci.size());
                }
                while ((rp.size() - 1) % 1024 != 0) rp.push_back((int)ci.size());
                segs.push_back((int)rp.size() - 1);
            }
            hoff_[D] = (int)rp.size() - 1;
            // symmetric velocity systems (uniform meshes) get an AMG-PCG when
            // their hierarchy stays sparse (operator complexity <= 2; the WLS
            // rows of AMR meshes give about 6, where the V-cycle costs more
            // than the Jacobi iterations it saves); HFH_VEL_AMG=0 turns it off
            bool vamg = true;
            for (int d = 0; d < D; d++) vamg = vamg && Hsym_[d];
            if (const char *e = std::getenv("HFH_VEL_AMG"))
                if (std::strcmp(e, "0") == 0) vamg = false;
            hsysB_.set_matrix(rp, ci, v, st_, vamg, segs, 2.0);
            hbB_.alloc(hoff_[D]);
            hxB_.alloc(hoff_[D]);
            hxPrev_.alloc(hoff_[D]);
            hxPrev2_.alloc(hoff_[D]);
            k_fill<double><<<nblocks(hoff_[D]), BLOCK, 0, st_>>>(hoff_[D], hxPrev_.p, 0.0);
            k_fill<double><<<nblocks(hoff_[D]), BLOCK, 0, st_>>>(hoff_[D], hxPrev2_.p, 0.0);
            k_fill<double><<<nblocks(hoff_[D]), BLOCK, 0, st_>>>(hoff_[D], hbB_.p, 0.0);
            k_fill<double><<<nblocks(hoff_[D]), BLOCK, 0, st_>>>(hoff_[D], hxB_.p, 0.0);
        } else {
            for (int d = 0; d < D; d++) {
                hsys_[d].set_matrix(Hrp_[d], Hci_[d], Hv_[d], st_);
                hb_[d].alloc(Nf_[d]);
            }
        }
    }
    refold_constants();
    HFH_CHECK(hipStreamSynchronize(st_));
}

// Every index a kernel dereferences is checked on the host before upload.
template <int D>
void Flow<D>::validate() const {
    auto fail = [](const char *what) { throw std::runtime_error(std::string("hfh layout check: ") + what); };
    for (int d = 0; d < D; d++)
        for (int s : momSlots_[d])
            if (s < 0 || s >= nX_) fail("momentum slot out of range");
    int ne = ne_;
    for (size_t i = 0; i < veSlots_.size(); i++) {
        int k = (int)(i / (ne > 0 ? ne : 1));
        int s = veSlots_[i];
        long hi = (k < 2 * D) ? s : s + (long)(NSYM - 1) * gK_.nrows();
        if (s < 0 || hi >= nX_) fail("constitutive slot out of range");
    }
    struct G {
        const Group *g;
        long in_hi, out_hi;
        int ostride;
    } gs[] = {{&gDu_, nX_, nX_, 0},
              {&gK_, (long)nX_, (long)nX_, gK_.nrows()},
              {&gCv_, nX_, nX_, 0},
              {&gMom_, nX_, nX_, 0},
              {&gDiv_, nX_, Nc_, 0},
              {&gGrad_, nX_, nX_, 0}};
    for (auto &e : gs) {
        const Group &g = *e.g;
        if ((int)g.rowptr.size() != g.nrows() + 1 || g.rowptr.back() != (int)g.col.size()) fail("row pointers");
        for (int c : g.col)
            if (c < 0 || c + (long)(g.ncomp - 1) * g.in_stride >= e.in_hi) fail("row column out of range");
        for (int o : g.out)
            if (o < 0 || o + (long)(g.ncomp - 1) * e.ostride >= e.out_hi) fail("row destination out of range");
    }
    for (size_t k = 0; k < Aci_.size(); k++)
        if (Aci_[k] < 0 || Aci_[k] >= Nc_) fail("pressure matrix column out of range");
}

static long group_templates(const Group &g, std::vector<int> &code, std::vector<int> &base, std::vector<int> &tcnt,
                            std::vector<int> &toff, std::vector<double> &tw, long &gen_nnz);

template <int D>
void Flow<D>::upload_group(DGroup &dg, const Group &g, int, int out_stride, bool has_y, bool f32) {
    dg.nrow = g.nrows();
    dg.ncomp = g.ncomp;
    dg.in_stride = g.in_stride;
    dg.out_stride = out_stride;
    dg.has_y = has_y;
    dg.f32 = f32;
    dg.order.clear();
    if (dg.nrow == 0) return;
    // template rows: the operator bytes of repeated stencils shrink to a
    // template id and a column base per row
    std::vector<int> code, base, tcnt, toff;
    std::vector<double> tw;
    long gen_nnz = 0;
    const long covered = group_templates(g, code, base, tcnt, toff, tw, gen_nnz);
    if (verbose_)
        std::fprintf(stderr, "hfh: group of %d rows (%zu entries): %ld template rows, %zu templates\n", dg.nrow,
                     g.col.size(), covered, tcnt.size());
    // device row order: the rows of each template together, in recording
    // order (so with increasing column bases), then the general rows.  A
    // wave then reads one template's entries and gathers neighbouring state
    // values; each row still sums its own entries in its own order, and the
    // rows write to their own output slots.  (HFH_GROUP_ORDER=0: recording
    // order.)
    dg.order.resize(dg.nrow);
    bool grouped = false;
    {
        const char *eo = std::getenv("HFH_GROUP_ORDER");
        const bool sort = covered > 0 && !(eo && std::strcmp(eo, "0") == 0);
        grouped = sort;
        if (sort) {
            const int nt = (int)tcnt.size();
            std::vector<long> start(nt + 2, 0);
            for (int r = 0; r < dg.nrow; r++) start[(code[r] >= 0 ? code[r] : nt) + 1]++;
            for (int t = 0; t <= nt; t++) start[t + 1] += start[t];
            for (int r = 0; r < dg.nrow; r++) dg.order[start[code[r] >= 0 ? code[r] : nt]++] = r;
        } else {
            for (int r = 0; r < dg.nrow; r++) dg.order[r] = r;
        }
    }
    const std::vector<int> &ord = dg.order;
    std::vector<int> rp(dg.nrow + 1, 0), col, outp(dg.nrow);
    std::vector<double> val;
    col.reserve(g.col.size());
    val.reserve(g.val.size());
    for (int k = 0; k < dg.nrow; k++) {
        const int r = ord[k];
        for (int q = g.rowptr[r]; q < g.rowptr[r + 1]; q++) {
            col.push_back(g.col[q]);
            val.push_back(g.val[q]);
        }
        rp[k + 1] = (int)col.size();
        outp[k] = g.out[r];
    }
    if (col.empty()) col.push_back(0);
    if (val.empty()) val.push_back(0.0);
    dg.rp.upload(rp.data(), rp.size(), st_);
    dg.col.upload(col.data(), col.size(), st_);
    std::vector<float> v32;
    if (f32) {
        v32.assign(val.begin(), val.end());
        dg.val32.upload(v32.data(), v32.size(), st_);
        dg.cst32.alloc(dg.nrow);
    } else {
        dg.val.upload(val.data(), val.size(), st_);
    }
    dg.out.upload(outp.data(), outp.size(), st_);
    dg.cst.alloc(dg.nrow);
    dg.ntempl = 0;
    dg.op_bytes = (long)g.col.size() * (4 + (f32 ? 4 : 8)) + (long)(dg.nrow + 1) * 4;
    double avg = (double)g.col.size() / dg.nrow;
    std::vector<int> pcode, pbase;
    std::vector<float> tw32;
    if (covered > 0) {
        dg.ntempl = (int)tcnt.size();
        pcode.resize(dg.nrow);
        pbase.resize(dg.nrow);
        for (int k = 0; k < dg.nrow; k++) {
            pcode[k] = code[ord[k]];
            pbase[k] = base[ord[k]];
        }
        dg.tcode.upload(pcode.data(), pcode.size(), st_);
        dg.tbase.upload(pbase.data(), pbase.size(), st_);
        dg.tcnt.upload(tcnt.data(), tcnt.size(), st_);
        dg.toff.upload(toff.data(), toff.size(), st_);
        tw32.assign(tw.begin(), tw.end());
        if (f32)
            dg.tw32.upload(tw32.data(), tw32.size(), st_);
        else
            dg.tw.upload(tw.data(), tw.size(), st_);
        const long gen_rows = dg.nrow - covered;
        dg.op_bytes = (long)dg.nrow * 8 + gen_nnz * (4 + (f32 ? 4 : 8)) + gen_rows * 8 +
                      (long)dg.ntempl * (4 + GT_MAX * (4 + (f32 ? 4 : 8)));
        // lanes by the entries a row actually reads
        long entries = gen_nnz;
        for (int r = 0; r < dg.nrow; r++)
            if (code[r] >= 0) entries += tcnt[code[r]];
        avg = (double)entries / dg.nrow;
    }
    HFH_CHECK(hipStreamSynchronize(st_));  // the host copies above are locals
    // lanes per row from the mean row length (HFH_INTERP_LANES overrides).
    // Rows grouped by template gather neighbouring values across the rows
    // of a wave, so fewer lanes per row (more rows per wave) pay there:
    // momentum rows of about 15 entries take 790 us on 2 lanes against 916
    // on 4 (the 64-cavity sweep)
    dg.lanes = grouped ? (avg < 4 ? 1 : avg < 32 ? 2 : avg < 64 ? 4 : 8)
                       : (avg < 4 ? 1 : avg < 8 ? 2 : avg < 16 ? 4 : avg < 32 ? 8 : 16);
    if (const char *e = std::getenv("HFH_INTERP_LANES")) {
        int f = std::atoi(e);
        if (f == 1 || f == 2 || f == 4 || f == 8 || f == 16) dg.lanes = f;
    }
}

// Template rows of an interpolation group.  Rows of at most GT_MAX entries
// are keyed by their column offsets from the first column and by their
// weights rounded to 20 mantissa bits; a row joins the template of its key
// when every weight is within 1e-11 of the row's largest weight, and the
// template takes the weights of the first row that made it.  The WLS
// weights of congruent point sets are equal in exact arithmetic; computed
// from absolute coordinates, as the reference computes them, they differ
// from position to position by rounding of 1e-12 to 1e-11 relative, so a
// template row's weights match its own to that level (HFH_TEMPLATES=0 keeps
// every row's own weights).  Templates shared by at least 8 rows are kept, up to
// 16,384; with fewer than a quarter of the rows covered, or
// HFH_TEMPLATES=0, the group stays CSR.  Returns the rows covered.
static long group_templates(const Group &g, std::vector<int> &code, std::vector<int> &base, std::vector<int> &tcnt,
                            std::vector<int> &toff, std::vector<double> &tw, long &gen_nnz) {
    constexpr int GT_MIN = 8, GT_TABLE = 16384, QBITS = 20;
    constexpr double TOL = 1e-11;
    if (const char *e = std::getenv("HFH_TEMPLATES"))
        if (std::strcmp(e, "0") == 0) return 0;
    const int nr = g.nrows();
    struct Key {
        std::vector<int> off;
        std::vector<int64_t> q;
        bool operator==(const Key &o) const { return off == o.off && q == o.q; }
    };
    struct Hash {
        size_t operator()(const Key &k) const {
            uint64_t h = 1469598103934665603ull;
            for (int v : k.off) h = (h ^ (uint64_t)(uint32_t)v) * 1099511628211ull;
            for (int64_t v : k.q) h = (h ^ (uint64_t)v) * 1099511628211ull;
            return (size_t)h;
        }
    };
    auto quant = [](double w) -> int64_t {
        if (w == 0.0) return 0;
        int ex;
        const double m = std::frexp(w, &ex);
        return (int64_t)std::llround(std::ldexp(m, QBITS)) * 4096 + ex;
    };
    std::unordered_map<Key, int, Hash> ids;
    std::vector<int> first, cnt, rowkey(nr, -1);
    for (int r = 0; r < nr; r++) {
        const int b = g.rowptr[r], e = g.rowptr[r + 1];
        if (e - b < 1 || e - b > GT_MAX) continue;
        Key k;
        double mx = 0;
        for (int q = b; q < e; q++) mx = std::max(mx, std::fabs(g.val[q]));
        for (int q = b; q < e; q++) {
            k.off.push_back(g.col[q] - g.col[b]);
            k.q.push_back(quant(g.val[q]));
        }
        auto it = ids.find(k);
        int id;
        if (it == ids.end()) {
            id = (int)first.size();
            ids.emplace(std::move(k), id);
            first.push_back(r);
            cnt.push_back(0);
        } else {
            id = it->second;
            const int f = first[id], fb = g.rowptr[f];
            bool near = true;
            for (int q = b; q < e && near; q++) near = std::fabs(g.val[q] - g.val[fb + (q - b)]) <= TOL * mx;
            if (!near) continue;
        }
        cnt[id]++;
        rowkey[r] = id;
    }
    std::vector<int> order(cnt.size());
    for (size_t q = 0; q < order.size(); q++) order[q] = (int)q;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
    std::vector<int> tid(cnt.size(), -1);
    long covered = 0;
    for (int q : order) {
        if (cnt[q] < GT_MIN || (int)tcnt.size() >= GT_TABLE) break;
        tid[q] = (int)tcnt.size();
        covered += cnt[q];
        const int f = first[q], fb = g.rowptr[f], fe = g.rowptr[f + 1];
        tcnt.push_back(fe - fb);
        for (int e = 0; e < GT_MAX; e++) {
            const bool has = e < fe - fb;
            toff.push_back(has ? g.col[fb + e] - g.col[fb] : 0);
            tw.push_back(has ? g.val[fb + e] : 0.0);
        }
    }
    if (covered * 4 < nr) {
        tcnt.clear();
        toff.clear();
        tw.clear();
        return 0;
    }
    code.assign(nr, -1);
    base.assign(nr, 0);
    gen_nnz = 0;
    for (int r = 0; r < nr; r++) {
        const int t = rowkey[r] >= 0 ? tid[rowkey[r]] : -1;
        code[r] = t;
        if (t >= 0)
            base[r] = g.col[g.rowptr[r]];
        else
            gen_nnz += g.rowptr[r + 1] - g.rowptr[r];
    }
    return covered;
}

template <int D>
void Flow<D>::refold_constants() {
    struct P {
        DGroup *d;
        Group *g;
    } gs[] = {{&dDu_, &gDu_}, {&dK_, &gK_}, {&dCv_, &gCv_}, {&dMom_, &gMom_}, {&dDiv_, &gDiv_}, {&dGrad_, &gGrad_}};
    // the host copies stay alive until the stream has consumed them
    std::vector<std::vector<double>> keep;
    std::vector<std::vector<float>> keep32;
    for (auto &p : gs) {
        if (p.d->nrow == 0) continue;
        {
            // in the device row order (upload_group)
            const std::vector<double> c0 = p.g->constants(bcval_);
            std::vector<double> c(c0.size());
            for (size_t k = 0; k < c.size(); k++) c[k] = c0[p.d->order[k]];
            keep.push_back(std::move(c));
        }
        const std::vector<double> &c = keep.back();
        if (p.d->f32) {
            keep32.emplace_back(c.begin(), c.end());
            p.d->cst32.upload(keep32.back().data(), keep32.back().size(), st_);
        } else {
            p.d->cst.upload(c.data(), c.size(), st_);
        }
    }
    if (par_.tempdiscrtype == SEMI_IMPLICIT_EULER) {
        for (int d = 0; d < D; d++) {
            keep.emplace_back(Nf_[d]);
            std::vector<double> &c = keep.back();
            for (int f = 0; f < Nf_[d]; f++) {
                double r = Hrhs0_[d][f];
                for (auto &b : Hbc_[d][f]) r += b.second * bcval_[b.first];
                c[f] = r;
            }
            if (Nf_[d] > 0) hcst_[d].upload(c.data(), c.size(), st_);
        }
    }
    HFH_CHECK(hipStreamSynchronize(st_));
    bc_dirty_ = false;
}

// hig-flow-step.c:402-512: boundaries marked time-dependent take the
// driver's values at t + dt (the pressure's increment over the step for the
// incremental projection).  The source-term domains carry no boundary
// trees here, so the source-boundary refresh (:515-606) has nothing to do.
template <int D>
void Flow<D>::boundary_refresh() {
    if (timedep_.empty()) return;
    const size_t ns = specs_.size();
    const double t = par_.t, t1 = par_.t + par_.dt;
    double c[3] = {0, 0, 0};
    auto fill = [&](Boundary<D> &b, auto &&value) {
        for (size_t k = 0; k < b.cells.size(); k++) {
            Vec<D> x = b.tree->center(b.cells[k]);
            for (int d = 0; d < D; d++) c[d] = x[d];
            bcval_[b.base + k] = value();
        }
        bc_dirty_ = true;
    };
    for (int dim = 0; dim < D && fn_.get_boundary_velocity; dim++)
        for (size_t i = 0; i < ns; i++)
            if (timedep_[(1 + dim) * ns + i])
                fill(*bcu_[dim][i], [&] { return fn_.get_boundary_velocity(specs_[i].userid, c, dim, t1); });
    for (size_t i = 0; i < ns && fn_.get_boundary_pressure; i++)
        if (timedep_[i])
            fill(*bcp_[i], [&] {
                double v = fn_.get_boundary_pressure(specs_[i].userid, c, t1);
                if (par_.projtype != NON_INCREMENTAL) v -= fn_.get_boundary_pressure(specs_[i].userid, c, t);
                return v;
            });
}

// hig-flow-step.c:343-399: the cell and facet source terms at t, every step,
// for the callbacks the driver registered; uploaded into F and FU.
template <int D>
void Flow<D>::source_refresh() {
    if (!fn_.get_source_term && !fn_.get_facet_source_term) return;
    const double t = par_.t;
    double c[3] = {0, 0, 0};
    auto run = [&](int field, int comp, auto &&value) {
        const int n = field_size(field, comp);
        if (n <= 0) return;
        srcC_.resize((size_t)n * D);
        srcV_.resize(n);
        field_centers(field, comp, srcC_.data());
        for (int i = 0; i < n; i++) {
            for (int d = 0; d < D; d++) c[d] = srcC_[(size_t)i * D + d];
            srcV_[i] = value();
        }
        HFH_CHECK(hipMemcpyAsync(X_.p + field_offset(field, comp), srcV_.data(), sizeof(double) * n,
                                 hipMemcpyHostToDevice, st_));
        HFH_CHECK(hipStreamSynchronize(st_));
    };
    if (fn_.get_source_term) run(4, 0, [&] { return fn_.get_source_term(c, t); });
    for (int dim = 0; dim < D && fn_.get_facet_source_term; dim++)
        run(5, dim, [&] { return fn_.get_facet_source_term(c, dim, t); });
}

template <int D>
void Flow<D>::run_group(DGroup &dg, double alpha, double *dst, int stage) {
    if (dg.nrow == 0) return;
    (void)stage;
    if (dg.f32)
        launch_rows<float>(dg, dg.val32.p, dg.cst32.p, (float)alpha, dst);
    else
        launch_rows<double>(dg, dg.val.p, dg.cst.p, alpha, dst);
}

template <int D>
template <typename W>
void Flow<D>::launch_rows(DGroup &dg, const W *val, const W *cst, W alpha, double *dst) {
    long n = (long)dg.nrow * dg.ncomp;
    const int *y = dg.has_y ? dg.y.p : nullptr;
    GTempl<W> tp;
    if (dg.ntempl > 0) {
        tp.code = dg.tcode.p;
        tp.base = dg.tbase.p;
        tp.tcnt = dg.tcnt.p;
        tp.toff = dg.toff.p;
        if constexpr (sizeof(W) == 4)
            tp.tw = dg.tw32.p;
        else
            tp.tw = dg.tw.p;
    }
#define HFH_INTERP_L(L)                                                                                        \
    k_interp_l<double, L, W><<<nblocks(n * L), BLOCK, 0, st_>>>(dg.nrow, dg.ncomp, dg.rp.p, dg.col.p, val, cst, \
                                                                dg.out.p, y, alpha, dg.in_stride, dg.out_stride, \
                                                                X_.p, dst, tp)
    switch (dg.lanes) {
    case 1:
        k_interp<double, W><<<nblocks(n), BLOCK, 0, st_>>>(dg.nrow, dg.ncomp, dg.rp.p, dg.col.p, val, cst, dg.out.p,
                                                           y, alpha, dg.in_stride, dg.out_stride, X_.p, dst, tp);
        break;
    case 2: HFH_INTERP_L(2); break;
    case 4: HFH_INTERP_L(4); break;
    case 8: HFH_INTERP_L(8); break;
    default: HFH_INTERP_L(16); break;
    }
#undef HFH_INTERP_L
}

// hig-flow-step.c:1363-1413 / hig-flow-step-viscoelastic.c:2091-2163
template <int D>
int Flow<D>::step(bool ve) {
    if (!ready_ || host_only_) return -1;
    if (par_.tempdiscrtype != EXPLICIT_EULER && par_.tempdiscrtype != SEMI_IMPLICIT_EULER) return -3;
    const bool semi = par_.tempdiscrtype == SEMI_IMPLICIT_EULER;
    if (ve != (par_.flowtype == VISCOELASTIC)) return -4;
    if (!ve) {
        boundary_refresh();
        source_refresh();
    }
    if (bc_dirty_) refold_constants();
    double *X = X_.p;
    int ne = ne_;
    if (ve) {
        tbegin(ST_VE);
        run_group(dDu_, 1.0, X, ST_VE);
        run_group(dK_, 1.0, X, ST_VE);
        run_group(dCv_, 1.0, X, ST_VE);
        VeParams vp;
        vp.ne = ne;
        vp.nrowK = gK_.nrows();
        vp.oDv = oDv_;
        vp.oK = oK_;
        vp.oKN = oKN_;
        vp.oS = oS_;
        vp.dt = par_.dt;
        vp.Re = par_.Re;
        vp.De = par_.De;
        vp.beta = par_.beta;
        vp.tol = par_.kernel_tol;
        vp.galpha = par_.giesekus_alpha;
        vp.model = par_.ve_model;
        vp.convec = par_.ve_convec;
        vp.transform = par_.kernel_transform;
        vp.par = vePar_.empty() ? nullptr : dVePar_.p;
        if (k32_)
            k_constitutive<D, float><<<nblocks(ne), BLOCK, 0, st_>>>(vp, dVeSlots_.p, dVeFlags_.p, dVeDelta_.p, X);
        else
            k_constitutive<D><<<nblocks(ne), BLOCK, 0, st_>>>(vp, dVeSlots_.p, dVeFlags_.p, dVeDelta_.p, X);
        HFH_CHECK(hipMemcpyAsync(X + oK_, X + oKN_, sizeof(double) * NSYM * ne, hipMemcpyDeviceToDevice, st_));
        tend(ST_VE);
        if (verbose_ && mem_.empty())
            // the polymeric-tensor step's stress range (hig-flow-step-
            // viscoelastic.c:326-335), lower triangle
            for (int i = 0; i < D; i++)
                for (int j = 0; j <= i; j++) {
                    double mn, mx;
                    field_range(9, i * D + j, &mn, &mx);
                    std::printf("===> %d %d: Tmin = %15.10lf <===> Tmax = %15.10lf <===\n", i, j, mn, mx);
                    std::fflush(stdout);
                }
        // the viscoelastic step refreshes boundary values and sources after
        // the tensor updates (hig-flow-step-viscoelastic.c:2112-2125)
        boundary_refresh();
        source_refresh();
        if (bc_dirty_) refold_constants();
    }
    tbegin(ST_INTERP);
    run_group(dMom_, 1.0, X, ST_INTERP);
    tend(ST_INTERP);
    tbegin(ST_MOMENTUM);
    for (int d = 0; d < D; d++) {
        MomParams mp;
        mp.nf = Nf_[d];
        mp.ns = momNs_;
        mp.u_off = oU_[d];
        mp.us_off = oUS_[d];
        mp.dt = par_.dt;
        mp.Re = par_.Re;
        mp.incremental = par_.projtype == INCREMENTAL;
        mp.second = par_.secondconvecdiscrtype;
        mp.ve = ve;
        mp.dim = d;
        mp.semi = semi;
        mp.bout = semi ? (hbatched_ ? hbB_.p + hoff_[d] : hb_[d].p) : nullptr;
        mp.bcst = semi ? hcst_[d].p : nullptr;
        mp.re = momRe_[d].empty() ? nullptr : dMomRe_[d].p;
        if (mp.nf > 0 && k32_)
            k_momentum<D, float><<<nblocks(mp.nf), BLOCK, 0, st_>>>(mp, dMomSlots_[d].p, dMomFlags_[d].p,
                                                                    dMomDelta_[d].p, X);
        else if (mp.nf > 0)
            k_momentum<D><<<nblocks(mp.nf), BLOCK, 0, st_>>>(mp, dMomSlots_[d].p, dMomFlags_[d].p, dMomDelta_[d].p,
                                                             X);
    }
    if (semi && hbatched_) {
        // the D Helmholtz systems in the same launches, warm-started from
        // the last u*; one stopping test per component
        bool spd = true;
        ExtrapSet es{};
        CopySet cs{};
        int nmax = 0;
        for (int d = 0; d < D; d++) {
            spd = spd && Hsym_[d];
            es.n[d] = cs.n[d] = Nf_[d];
            es.cur[d] = X + oUS_[d];
            es.p1[d] = hxPrev_.p + hoff_[d];
            es.p2[d] = hxPrev2_.p + hoff_[d];
            es.out[d] = hxB_.p + hoff_[d];
            cs.src[d] = hxB_.p + hoff_[d];
            cs.dst[d] = X + oUS_[d];
            nmax = std::max(nmax, Nf_[d]);
        }
        // one launch for the D components: initial guesses, then (after the
        // solve) the copy back into u*
        if (nmax > 0) k_extrap_set<<<dim3(nblocks(nmax), D), BLOCK, 0, st_>>>(es, std::min(extrap_, extrap_hist_));
        if (par_.precision == 1)
            hsysB_.fused_mixed(spd, hbB_.p, hxB_.p, par_.rtol, par_.atol, par_.inner_rtol, par_.maxit, st_);
        else
            hsysB_.fused64(spd, hbB_.p, hxB_.p, par_.rtol, par_.atol, par_.maxit, st_);
        if (nmax > 0) k_copy_set<<<dim3(nblocks(nmax), D), BLOCK, 0, st_>>>(cs);
        for (int d = 0; d < D; d++)
            if ((int)hsysB_.seg_stats.size() == D) hstats_[d] = hsysB_.seg_stats[d];
    } else if (semi) {  // one Helmholtz solve per component, warm-started from the last u*
        for (int d = 0; d < D; d++) {
            if (Nf_[d] == 0) continue;
            double *x = X + oUS_[d];
            if (hsys_[d].fused && par_.precision == 1)
                hstats_[d] = hsys_[d].fused_mixed(Hsym_[d], hb_[d].p, x, par_.rtol, par_.atol, par_.inner_rtol,
                                                  par_.maxit, st_);
            else if (hsys_[d].fused)
                hstats_[d] = hsys_[d].fused64(Hsym_[d], hb_[d].p, x, par_.rtol, par_.atol, par_.maxit, st_);
            else if (par_.precision == 1)
                hstats_[d] = hsys_[d].mixed(hb_[d].p, x, par_.rtol, par_.atol, par_.inner_rtol, par_.maxit, st_);
            else if (Hsym_[d])
                hstats_[d] = hsys_[d].pcg(hb_[d].p, x, par_.rtol, par_.atol, par_.maxit, st_);
            else
                hstats_[d] = hsys_[d].bicgstab(hb_[d].p, x, par_.rtol, par_.atol, par_.maxit, st_);
        }
    }
    tend(ST_MOMENTUM);
    tbegin(ST_DIV);
    run_group(dDiv_, 1.0, bvec_.p, ST_DIV);
    tend(ST_DIV);
    tbegin(ST_SOLVE);
    double *x = X + (par_.projtype == INCREMENTAL ? oDP_ : oP_);
    if (extrap_)
        k_extrap<double><<<nblocks(Nc_), BLOCK, 0, st_>>>(Nc_, x, pPrev_.p, pPrev2_.p, x, std::min(extrap_, extrap_hist_));
    if (psys_.fused && par_.precision == 1) {
        stats_ = psys_.fused_mixed(symmetric_, bvec_.p, x, par_.rtol, par_.atol, par_.inner_rtol, par_.maxit, st_);
    } else if (psys_.fused) {
        stats_ = psys_.fused64(symmetric_, bvec_.p, x, par_.rtol, par_.atol, par_.maxit, st_);
    } else if (par_.precision == 1) {
        stats_ = psys_.mixed(bvec_.p, x, par_.rtol, par_.atol, par_.inner_rtol, par_.maxit, st_);
    } else if (symmetric_) {
        stats_ = psys_.pcg(bvec_.p, x, par_.rtol, par_.atol, par_.maxit, st_);
    } else {
        stats_ = psys_.bicgstab(bvec_.p, x, par_.rtol, par_.atol, par_.maxit, st_);
    }
    tend(ST_SOLVE);
    tbegin(ST_PROJECT);
    run_group(dGrad_, -par_.dt, X, ST_PROJECT);
    if (par_.projtype == INCREMENTAL) k_axpy_field<double><<<nblocks(Nc_), BLOCK, 0, st_>>>(Nc_, X, oP_, oDP_);
    tend(ST_PROJECT);
    par_.t += par_.dt;
    extrap_hist_ = std::min(extrap_hist_ + 1, 2);
    if (!mem_.empty()) collect_member_stats();
    if (verbose_ && mem_.empty())
        for (int d = 0; d < D; d++) {
            double mn, mx;
            field_range(0, d, &mn, &mx);
            std::printf("===> %d: Vmin = %15.10lf <===> Vmax = %15.10lf <===\n", d, mn, mx);
            std::fflush(stdout);
        }
    return stats_.converged ? 0 : 1;
}

// Launches one kernel of the step `reps` times back to back on the library
// stream between two hipEvents (inputs as left by the last step):
// the probed step kernels recompute their outputs from inputs they do not
// write, so repeated launches leave the state unchanged; the solver probes
// change only the solver's work vectors, which the next solve rewrites.
//   0 pressure SpMV + partial dot (one CG iteration's operator pass)
//   1 fused momentum kernel, first velocity component
//   2 momentum interpolation rows
//   3 cell divergence rows
//   4 projection rows
//   5 fused constitutive kernel (viscoelastic cases)
//   6 / 7 fused pressure BiCGStab launch B, FP64 / FP32
//   8 plain copy of 64 MiB
//   9 / 10 fused FP64 BiCGStab launch A / B of the semi-implicit velocity
//     block system (of the pressure system when there is none)
template <int D>
int Flow<D>::probe(int which, int reps, double *avg_ms, long *bytes) {
    if (!ready_ || host_only_ || reps <= 0) return -1;
    hipEvent_t a, b;
    HFH_CHECK(hipEventCreate(&a));
    HFH_CHECK(hipEventCreate(&b));
    double *X = X_.p;
    long by = 0;
    int n = Nc_;
    MomParams mp;
    mp.nf = Nf_[0];
    mp.ns = momNs_;
    mp.u_off = oU_[0];
    mp.us_off = oUS_[0];
    mp.dt = par_.dt;
    mp.Re = par_.Re;
    mp.incremental = par_.projtype == INCREMENTAL;
    mp.second = par_.secondconvecdiscrtype;
    mp.ve = par_.flowtype == VISCOELASTIC;
    mp.dim = 0;
    mp.semi = par_.tempdiscrtype == SEMI_IMPLICIT_EULER;
    mp.bout = mp.semi ? (hbatched_ ? hbB_.p : hb_[0].p) : nullptr;
    mp.bcst = mp.semi ? hcst_[0].p : nullptr;
    HFH_CHECK(hipEventRecord(a, st_));
    for (int r = 0; r < reps; r++) {
        switch (which) {
        case 0:
            k_spmv_dot<double><<<psys_.nb, BLOCK, 0, st_>>>(n, psys_.rp.p, psys_.ci.p, psys_.v.p, psys_.p.p,
                                                            psys_.q.p, psys_.part1.p);
            by = (long)psys_.nnz * 12 + (long)(n + 1) * 4 + (long)n * 16 + (long)psys_.nb * 8;
            break;
        case 6: {
            // the fused pressure BiCGStab t-kernel, timed by the solver itself
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            float ms = psys_.k64.probe(reps, &by, st_);
            *avg_ms = ms;
            *bytes = by;
            return 0;
        }
        case 7: {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            float ms = psys_.k32.probe(reps, &by, st_);
            *avg_ms = ms;
            *bytes = by;
            return 0;
        }
        case 9:
        case 10: {
            // the FP64 fused BiCGStab launches A (9) and B (10) of the
            // semi-implicit velocity block system (the pressure system's
            // when the step has none)
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            FusedKrylov<double> &k = hbatched_ ? hsysB_.k64 : psys_.k64;
            if (k.n == 0) return -1;
            float ms = which == 9 ? k.probe_a(reps, &by, st_) : k.probe(reps, &by, st_);
            *avg_ms = ms;
            *bytes = by;
            return 0;
        }
        case 11:
        case 12: {
            // PCG launch 1 (SpMV) of the semi-implicit velocity block system
            // in FP64 (11) / FP32 (12): the symmetric velocity systems of
            // uniform meshes solve by Jacobi-PCG
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            if (!hbatched_) return -1;
            float ms = which == 11 ? hsysB_.k64.probe_cg(reps, &by, st_) : hsysB_.k32.probe_cg(reps, &by, st_);
            *avg_ms = ms;
            *bytes = by;
            return 0;
        }
        case 13: {
            // fine-level prolongation + post-smoothing launch of the FP32
            // V-cycle of the semi-implicit velocity block system (AMG-PCG of
            // the symmetric velocity systems of uniform meshes)
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            if (!hbatched_ || !hsysB_.k32.amg_on()) return -1;
            const float ms = hsysB_.k32.amg.probe_pp(reps, &by, st_);
            if (ms < 0) return -1;
            *avg_ms = ms;
            *bytes = by;
            return 0;
        }
        case 14:
        case 15: {
            // the same launch of the pressure system's FP32 V-cycle (14);
            // the fine-level residual launch of its FP64 V-cycle (15)
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            const float ms = which == 14 ? (psys_.k32.amg_on() ? psys_.k32.amg.probe_pp(reps, &by, st_) : -1.0f)
                                         : (psys_.k64.amg_on() ? psys_.k64.amg.probe_resid(reps, &by, st_) : -1.0f);
            if (ms < 0) return -1;
            *avg_ms = ms;
            *bytes = by;
            return 0;
        }
        case 8: {
            // plain copy of 8 Mi doubles (64 MiB in, 64 MiB out): the
            // reference point that scales the hardware byte counters
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            const int nc = 8 << 20;
            DevBuf<double> src, dst;
            src.alloc(nc);
            dst.alloc(nc);
            k_fill<double><<<nblocks(nc), BLOCK, 0, st_>>>(nc, src.p, 1.0);
            hipEvent_t c0, c1;
            HFH_CHECK(hipEventCreate(&c0));
            HFH_CHECK(hipEventCreate(&c1));
            HFH_CHECK(hipEventRecord(c0, st_));
            for (int q = 0; q < reps; q++) k_copy<double><<<nblocks(nc), BLOCK, 0, st_>>>(nc, src.p, dst.p);
            HFH_CHECK(hipEventRecord(c1, st_));
            HFH_CHECK(hipEventSynchronize(c1));
            float ms = 0;
            HFH_CHECK(hipEventElapsedTime(&ms, c0, c1));
            (void)hipEventDestroy(c0);
            (void)hipEventDestroy(c1);
            *avg_ms = reps > 0 ? ms / reps : 0.0;
            *bytes = 2L * nc * (long)sizeof(double);
            return 0;
        }
        case 1:
            k_momentum<D><<<nblocks(mp.nf), BLOCK, 0, st_>>>(mp, dMomSlots_[0].p, dMomFlags_[0].p,
                                                             dMomDelta_[0].p, X);
            by = (long)mp.nf * ((long)momNs_ * (4 + 8) + 4 + D * 8 + 8 + 8);
            break;
        case 2:
            run_group(dMom_, 1.0, X, ST_INTERP);
            by = gdbytes(dMom_);
            break;
        case 3:
            run_group(dDiv_, 1.0, bvec_.p, ST_DIV);
            by = gdbytes(dDiv_);
            break;
        case 4:
            // recomputes u = ustar - dt*grad(dp) from unchanged inputs: idempotent
            run_group(dGrad_, -par_.dt, X, ST_PROJECT);
            by = gdbytes(dGrad_) + (long)dGrad_.nrow * 12;
            break;
        default: {
            if (par_.flowtype != VISCOELASTIC) return -1;
            VeParams vp;
            vp.ne = ne_;
            vp.nrowK = gK_.nrows();
            vp.oDv = oDv_;
            vp.oK = oK_;
            vp.oKN = oKN_;
            vp.oS = oS_;
            vp.dt = par_.dt;
            vp.Re = par_.Re;
            vp.De = par_.De;
            vp.beta = par_.beta;
            vp.tol = par_.kernel_tol;
            vp.galpha = par_.giesekus_alpha;
            vp.model = par_.ve_model;
            vp.convec = par_.ve_convec;
            vp.transform = par_.kernel_transform;
            k_constitutive<D><<<nblocks(vp.ne), BLOCK, 0, st_>>>(vp, dVeSlots_.p, dVeFlags_.p, dVeDelta_.p, X);
            by = bytes_[ST_VE];
        }
        }
    }
    HFH_CHECK(hipEventRecord(b, st_));
    HFH_CHECK(hipEventSynchronize(b));
    float ms = 0;
    HFH_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *avg_ms = ms / reps;
    *bytes = by;
    return 0;
}

// field ids as the oracle: 0 u, 1 p, 2 dp, 3 ustar, 4 F, 5 FU, 6 kernel
// (symmetric storage, comp = i*D+j), 7 S, 8 Du
template <int D>
int Flow<D>::field_size(int field, int comp) {
    if (!ready_ || !mem_.empty()) return -1;  // batched runs: per member
    int ne = ne_;
    switch (field) {
    case 0:
    case 3:
        return comp >= 0 && comp < D ? Nf_[comp] : -1;
    case 1:
    case 2:
        return Nc_;
    case 4:
        return sdF_->size();
    case 5:
        return comp >= 0 && comp < D ? sfdF_[comp]->size() : -1;
    case 6:
    case 7:
    case 8:
        return comp >= 0 && comp < D * D ? ne : -1;
    case 9:  // polymeric stress, computed on request
        return par_.flowtype == VISCOELASTIC && comp >= 0 && comp < D * D ? ne : -1;
    }
    return -1;
}

// field 9 component `comp` into tT_ (T = S + (1 - beta) Du / Re)
template <int D>
const double *Flow<D>::poly_stress(int comp) {
    if ((int)tT_.n < ne_) tT_.alloc(ne_);
    k_poly_stress<<<nblocks(ne_), BLOCK, 0, st_>>>(ne_, X_.p + oS_ + comp * ne_, X_.p + oDv_ + comp * ne_, par_.beta,
                                                   par_.Re, tT_.p);
    return tT_.p;
}

template <int D>
int Flow<D>::field_centers(int field, int comp, double *xyz) {
    int n = field_size(field, comp);
    if (n < 0) return -1;
    Domain<D> *dm = field == 0 || field == 3 ? sfdu_[comp].get()
                    : field == 5            ? sfdF_[comp].get()
                    : field == 4            ? sdF_.get()
                    : field >= 6            ? sdED_.get()
                                            : sdp_.get();
    for (int i = 0; i < n; i++) {
        Vec<D> c = dm->center_of(i);
        for (int d = 0; d < D; d++) xyz[(size_t)i * D + d] = c[d];
    }
    return n;
}

static inline int sym_index(int D, int i, int j) {
    if (i > j) std::swap(i, j);
    int m = 0;
    for (int a = 0; a < D; a++)
        for (int b = a; b < D; b++, m++)
            if (a == i && b == j) return m;
    return -1;
}

template <int D>
int Flow<D>::field_offset(int field, int comp) const {
    switch (field) {
    case 0: return oU_[comp];
    case 1: return oP_;
    case 2: return oDP_;
    case 3: return oUS_[comp];
    case 4: return oF_;
    case 5: return oFU_[comp];
    case 6: return oK_ + sym_index(D, comp / D, comp % D) * ne_;
    case 7: return oS_ + comp * ne_;
    default: return oDv_ + comp * ne_;
    }
}

template <int D>
int Flow<D>::get_field(int field, int comp, double *out) {
    int n = field_size(field, comp);
    if (n < 0 || host_only_ || !mem_.empty()) return -1;
    const double *src = field == 9 ? poly_stress(comp) : X_.p + field_offset(field, comp);
    HFH_CHECK(hipMemcpyAsync(out, src, sizeof(double) * n, hipMemcpyDeviceToHost, st_));
    HFH_CHECK(hipStreamSynchronize(st_));
    return n;
}

// min and max over the unknowns of a field (for velocity: the reference's
// Vmin/Vmax over the unknown facets, hig-flow-step.c:250-306)
template <int D>
int Flow<D>::field_range(int field, int comp, double *mn, double *mx) {
    int n = field_size(field, comp);
    if (n < 0 || host_only_ || !mem_.empty()) return -1;
    const int g = mm_groups(n);
    if ((long)mm_.n < 2 + 2L * g) mm_.alloc(2 + 2 * (size_t)g);  // min, max, then the partials
    if (!mmt_.p) mmt_.alloc(1);
    k_zero_u32<<<1, 64, 0, st_>>>(mmt_.p);
    const double *src = field == 9 ? poly_stress(comp) : X_.p + field_offset(field, comp);
    k_minmax<<<g, 1024, 0, st_>>>(n, src, mm_.p + 2, mmt_.p, mm_.p);
    double h[2];
    HFH_CHECK(hipMemcpyAsync(h, mm_.p, sizeof h, hipMemcpyDeviceToHost, st_));
    HFH_CHECK(hipStreamSynchronize(st_));
    *mn = h[0];
    *mx = h[1];
    return n;
}

template <int D>
int Flow<D>::set_field(int field, int comp, const double *in) {
    int n = field_size(field, comp);
    if (n < 0 || host_only_ || !mem_.empty() || field == 9) return -1;
    if (field == 6 && comp / D > comp % D) return n;  // symmetric storage keeps the upper triangle
    int off = field_offset(field, comp);
    extrap_hist_ = 0;
    HFH_CHECK(hipMemcpyAsync(X_.p + off, in, sizeof(double) * n, hipMemcpyHostToDevice, st_));
    HFH_CHECK(hipStreamSynchronize(st_));
    return n;
}

template <int D>
int Flow<D>::stencil(int domain, int dim, const double *center, const double *x, double alpha, int maxn, int *ids,
                     double *w, double *rhs) {
    if (!ready_ || !mem_.empty()) return -1;
    Domain<D> *dm = domain == 0 ? sdp_.get() : domain == 2 ? sdED_.get() : nullptr;
    if (domain == 1) {
        if (dim < 0 || dim >= D) return -1;
        dm = sfdu_[dim].get();
    }
    if (!dm) return -1;
    Vec<D> c, p;
    for (int d = 0; d < D; d++) {
        c[d] = center[d];
        p[d] = x[d];
    }
    Stencil s;
    dm->get_stencil(c, p, alpha, s);
    int n = (int)s.ids.size();
    for (int i = 0; i < n && i < maxn; i++) {
        ids[i] = s.ids[i];
        w[i] = s.w[i];
    }
    *rhs = s.rhs(bcval_);
    return n;
}

}  // namespace hfh
