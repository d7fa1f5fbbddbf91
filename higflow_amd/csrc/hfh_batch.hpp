// Batched runs: many independent simulations (parameter sweeps over Re and
// Wi on meshes of different sizes) advanced in the same launches.
//
// Each member is set up on the host exactly as a single run (same recorder,
// same operator tables), then the member tables are merged: every state
// offset is remapped into the batched state, every row keeps its member's
// coefficients, and the linear systems become block-diagonal with one
// segment per member, so each member keeps its own Krylov scalars and its
// own stopping test (a member's solve sees the same arithmetic as its
// single run).  The reference runs such a sweep as one driver process per
// simulation (higflow/example2d_Oldroyd/ns-example-2d.c:465-471 per run);
// nothing in it batches.
#pragma once
#include <algorithm>
#include <cstring>

namespace hfh {

template <int D>
std::vector<typename Flow<D>::Region> Flow<D>::regions() const {
    std::vector<Region> r;
    for (int d = 0; d < D; d++) r.push_back({oU_[d], 1, Nf_[d], true});
    for (int d = 0; d < D; d++) r.push_back({oUS_[d], 1, Nf_[d], true});
    for (int d = 0; d < D; d++) r.push_back({oFU_[d], 1, sfdF_[d]->size(), false});
    r.push_back({oP_, 1, Nc_, true});
    r.push_back({oDP_, 1, Nc_, true});
    r.push_back({oF_, 1, sdF_->size(), false});
    r.push_back({oK_, NSYM, ne_, false});
    r.push_back({oKN_, NSYM, ne_, false});
    r.push_back({oS_, D * D, ne_, false});
    r.push_back({oDv_, D * D, ne_, false});
    r.push_back({oPK_, NSYM, gK_.nrows(), false});
    r.push_back({oPcv_, 1, gCv_.nrows(), false});
    r.push_back({oPmom_, 1, gMom_.nrows(), false});
    r.push_back({oZero_, 1, 1, false});
    return r;
}

template <int D>
int Flow<D>::remap(int m, int x) const {
    const std::vector<Region> &rg = mreg_[m];
    // regions are consecutive and in offset order
    int k = (int)(std::upper_bound(rg.begin(), rg.end(), x, [](int v, const Region &r) { return v < r.off; }) -
                  rg.begin()) -
            1;
    if (k < 0) throw std::runtime_error("hfh batch: state offset below the first region");
    const Region &r = rg[k];
    int rel = x - r.off;
    if (r.len <= 0 || rel >= r.ncomp * r.len) throw std::runtime_error("hfh batch: state offset outside its region");
    int comp = rel / r.len, j = rel % r.len;
    return bbase_[k] + comp * btot_[k] + bmoff_[k][m] + j;
}

static inline int pad1024(int n) { return (n + 1023) / 1024 * 1024; }

template <int D>
int Flow<D>::setup_batch(const std::vector<FlowBase *> &members) {
    if (ready_ || !mem_.empty() || members.empty()) return -1;
    const char *ek = std::getenv("HFH_KRYLOV");
    if (ek && std::strcmp(ek, "classic") == 0)
        throw std::runtime_error("hfh batch: the batched solves need the fused Krylov solver");
    for (FlowBase *b : members) {
        auto *f = dynamic_cast<Flow<D> *>(b);
        if (!f || !f->ready_ || !f->host_only_ || !f->mem_.empty())
            throw std::runtime_error("hfh batch: members must be set up host-only, with the batch's dimension");
        mem_.push_back(f);
    }
    const int nm = (int)mem_.size();
    const Params &p0 = mem_[0]->par_;
    for (Flow<D> *f : mem_) {
        const Params &q = f->par_;
        bool same = q.dt == p0.dt && q.projtype == p0.projtype && q.flowtype == p0.flowtype &&
                    q.tempdiscrtype == p0.tempdiscrtype && q.convecdiscrtype == p0.convecdiscrtype &&
                    q.secondconvecdiscrtype == p0.secondconvecdiscrtype && q.ve_model == p0.ve_model &&
                    q.ve_convec == p0.ve_convec && q.kernel_transform == p0.kernel_transform &&
                    q.kernel_tol == p0.kernel_tol && q.giesekus_alpha == p0.giesekus_alpha && q.rtol == p0.rtol &&
                    q.atol == p0.atol && q.maxit == p0.maxit && q.precision == p0.precision &&
                    q.inner_rtol == p0.inner_rtol && f->momNs_ == mem_[0]->momNs_;
        if (!same)
            throw std::runtime_error("hfh batch: members differ in time step, schemes or solver settings "
                                     "(only Re, De and beta may differ)");
    }
    par_ = p0;
    par_.t = p0.t;
    const bool ve = par_.flowtype == VISCOELASTIC, semi = par_.tempdiscrtype == SEMI_IMPLICIT_EULER;

    // ---- batched layout
    mreg_.clear();
    for (Flow<D> *f : mem_) mreg_.push_back(f->regions());
    const int nr = (int)mreg_[0].size();
    for (int m = 0; m < nm; m++) {
        int o = 0;
        for (auto &r : mreg_[m]) {
            if (r.off != o) throw std::runtime_error("hfh batch: member state regions are not consecutive");
            o += r.ncomp * r.len;
        }
        if (o != mem_[m]->nX_) throw std::runtime_error("hfh batch: member state regions do not cover the state");
    }
    bbase_.assign(nr, 0);
    btot_.assign(nr, 0);
    bmoff_.assign(nr, std::vector<int>(nm, 0));
    int o = 0;
    for (int k = 0; k < nr; k++) {
        int t = 0;
        for (int m = 0; m < nm; m++) {
            bmoff_[k][m] = t;
            const Region &r = mreg_[m][k];
            t += r.pad ? pad1024(r.len) : r.len;
        }
        bbase_[k] = o;
        btot_[k] = t;
        o += mreg_[0][k].ncomp * t;
    }
    // region indices (order of regions())
    const int kU = 0, kUS = D, kP = 3 * D, kK = 3 * D + 3, kPK = 3 * D + 7, kZero = 3 * D + 10;
    oZero_ = bbase_[kZero];
    nX_ = o;
    for (int d = 0; d < D; d++) {
        oU_[d] = bbase_[kU + d];
        oUS_[d] = bbase_[kUS + d];
        oFU_[d] = bbase_[2 * D + d];
        Nf_[d] = btot_[kU + d];
    }
    oP_ = bbase_[kP];
    oDP_ = bbase_[kP + 1];
    oF_ = bbase_[kP + 2];
    oK_ = bbase_[kK];
    oKN_ = bbase_[kK + 1];
    oS_ = bbase_[kK + 2];
    oDv_ = bbase_[kK + 3];
    oPK_ = bbase_[kPK];
    oPcv_ = bbase_[kPK + 1];
    oPmom_ = bbase_[kPK + 2];
    Nc_ = btot_[kP];
    ne_ = btot_[kK];
    momNs_ = mem_[0]->momNs_;
    symmetric_ = true;
    for (Flow<D> *f : mem_) symmetric_ = symmetric_ && f->symmetric_;

    // ---- boundary values
    bcoff_.assign(nm, 0);
    bcval_.clear();
    for (int m = 0; m < nm; m++) {
        bcoff_[m] = (int)bcval_.size();
        bcval_.insert(bcval_.end(), mem_[m]->bcval_.begin(), mem_[m]->bcval_.end());
    }

    // ---- interpolation groups
    auto merge = [&](Group Flow<D>::*gp, bool div_out) {
        Group &g = this->*gp;
        g = Group();
        g.ncomp = (mem_[0]->*gp).ncomp;
        for (int m = 0; m < nm; m++) {
            const Group &s = mem_[m]->*gp;
            if (s.ncomp != g.ncomp) throw std::runtime_error("hfh batch: group shapes differ");
            for (int r = 0; r < s.nrows(); r++) {
                std::vector<std::pair<int, double>> cols, bcs;
                for (int q = s.rowptr[r]; q < s.rowptr[r + 1]; q++) cols.emplace_back(remap(m, s.col[q]), s.val[q]);
                for (auto &b : s.bc[r]) bcs.emplace_back(b.first + bcoff_[m], b.second);
                int out = div_out ? bmoff_[kP][m] + s.out[r] : remap(m, s.out[r]);
                g.add_row(cols, s.rhs0[r], bcs, out);
            }
        }
    };
    merge(&Flow<D>::gDu_, false);
    merge(&Flow<D>::gK_, false);
    merge(&Flow<D>::gCv_, false);
    merge(&Flow<D>::gMom_, false);
    merge(&Flow<D>::gDiv_, true);
    merge(&Flow<D>::gGrad_, false);
    gK_.in_stride = ne_;
    if (gK_.nrows() != btot_[kPK] || gCv_.nrows() != btot_[kPK + 1] || gMom_.nrows() != btot_[kPK + 2])
        throw std::runtime_error("hfh batch: merged row counts do not match the row regions");

    // ---- momentum slots (slot-major: slot k of facet f at k * nf + f);
    // the padding facets read the zero entry and compute zeros
    for (int d = 0; d < D; d++) {
        const int nf = Nf_[d];
        momSlots_[d].assign((size_t)momNs_ * nf, oZero_);
        momFlags_[d].assign(nf, 0u);
        momDelta_[d].assign((size_t)D * nf, 1.0);
        momRe_[d].assign(nf, 1.0);
        for (int m = 0; m < nm; m++) {
            const Flow<D> &f = *mem_[m];
            const int mf = f.Nf_[d], b = bmoff_[kU + d][m];
            for (int k = 0; k < momNs_; k++)
                for (int i = 0; i < mf; i++)
                    momSlots_[d][(size_t)k * nf + b + i] = remap(m, f.momSlots_[d][(size_t)k * mf + i]);
            for (int i = 0; i < mf; i++) {
                momFlags_[d][b + i] = f.momFlags_[d][i];
                momRe_[d][b + i] = f.par_.Re;
            }
            for (int k = 0; k < D; k++)
                for (int i = 0; i < mf; i++) momDelta_[d][(size_t)k * nf + b + i] = f.momDelta_[d][(size_t)k * mf + i];
        }
    }

    // ---- constitutive slots (slot-major over cells), per-cell parameters
    if (ve) {
        const int nsl = 6 * D;
        veSlots_.assign((size_t)nsl * ne_, oZero_);
        veFlags_.assign(ne_, 0u);
        veDelta_.assign((size_t)D * ne_, 1.0);
        vePar_.assign((size_t)3 * ne_, 1.0);
        for (int m = 0; m < nm; m++) {
            const Flow<D> &f = *mem_[m];
            const int me = f.ne_, b = bmoff_[kK][m];
            if ((int)f.veSlots_.size() != nsl * me) throw std::runtime_error("hfh batch: constitutive slots");
            for (int k = 0; k < nsl; k++)
                for (int c = 0; c < me; c++) veSlots_[(size_t)k * ne_ + b + c] = remap(m, f.veSlots_[(size_t)k * me + c]);
            for (int c = 0; c < me; c++) {
                veFlags_[b + c] = f.veFlags_[c];
                vePar_[b + c] = f.par_.Re;
                vePar_[(size_t)ne_ + b + c] = f.par_.De;
                vePar_[2 * (size_t)ne_ + b + c] = f.par_.beta;
            }
            for (int k = 0; k < D; k++)
                for (int c = 0; c < me; c++) veDelta_[(size_t)k * ne_ + b + c] = f.veDelta_[(size_t)k * me + c];
        }
    }

    // ---- pressure system: member blocks on the padded cell layout
    Arp_.assign(1, 0);
    Aci_.clear();
    Av_.clear();
    pseg_.assign(1, 0);
    for (int m = 0; m < nm; m++) {
        const Flow<D> &f = *mem_[m];
        const int b = bmoff_[kP][m];
        for (int i = 0; i < f.Nc_; i++) {
            for (int q = f.Arp_[i]; q < f.Arp_[i + 1]; q++) {
                Aci_.push_back(b + f.Aci_[q]);
                Av_.push_back(f.Av_[q]);
            }
            Arp_.push_back((int)Aci_.size());
        }
        while ((int)Arp_.size() - 1 < b + pad1024(f.Nc_)) Arp_.push_back((int)Aci_.size());
        pseg_.push_back((int)Arp_.size() - 1);
    }

    // ---- semi-implicit velocity systems: member blocks per component
    if (semi) {
        for (int d = 0; d < D; d++) {
            Hrp_[d].assign(1, 0);
            Hci_[d].clear();
            Hv_[d].clear();
            Hrhs0_[d].assign(Nf_[d], 0.0);
            Hbc_[d].assign(Nf_[d], {});
            Hsym_[d] = true;
            hseg_[d].clear();
            for (int m = 0; m < nm; m++) {
                const Flow<D> &f = *mem_[m];
                const int b = bmoff_[kU + d][m];
                hseg_[d].push_back(b);
                Hsym_[d] = Hsym_[d] && f.Hsym_[d];
                for (int i = 0; i < f.Nf_[d]; i++) {
                    for (int q = f.Hrp_[d][i]; q < f.Hrp_[d][i + 1]; q++) {
                        Hci_[d].push_back(b + f.Hci_[d][q]);
                        Hv_[d].push_back(f.Hv_[d][q]);
                    }
                    Hrp_[d].push_back((int)Hci_[d].size());
                    Hrhs0_[d][b + i] = f.Hrhs0_[d][i];
                    for (auto &e : f.Hbc_[d][i]) Hbc_[d][b + i].emplace_back(e.first + bcoff_[m], e.second);
                }
                while ((int)Hrp_[d].size() - 1 < b + pad1024(f.Nf_[d])) Hrp_[d].push_back((int)Hci_[d].size());
            }
        }
    }

    ready_ = true;
    validate();
    // algorithmic bytes per stage launch, as for a single run
    auto gbytes = [&](const Group &g) {
        return (long)g.col.size() * 12 + (long)g.nrows() * (8 + 8 + 4 + 4) * g.ncomp;
    };
    bytes_[ST_INTERP] = gbytes(gMom_) + gbytes(gDu_) + gbytes(gK_) + gbytes(gCv_);
    long mb = 0;
    for (int d = 0; d < D; d++) mb += (long)Nf_[d] * ((long)momNs_ * (4 + 8) + 4 + D * 8 + 8 + 8 + 8);
    bytes_[ST_MOMENTUM] = mb;
    bytes_[ST_DIV] = gbytes(gDiv_);
    bytes_[ST_PROJECT] = gbytes(gGrad_) + (long)gGrad_.nrows() * 12 + (long)Nc_ * 24;
    bytes_[ST_SOLVE] = (long)Aci_.size() * 12 + (long)Nc_ * 8 * 9;
    bytes_[ST_VE] = (long)ne_ * (D * D * 8 * 2 + NSYM * 8 * 2 + 6 * D * (4 + 8) + 4 + D * 8 + 3 * 8);
    mstats_.assign((size_t)nm * (D + 1), SolveStats());
    if (host_only_) return 0;
    upload_device();
    return 0;
}

template <int D>
void Flow<D>::collect_member_stats() {
    const int nm = (int)mem_.size();
    for (int m = 0; m < nm; m++) {
        if ((int)psys_.seg_stats.size() == nm) mstats_[(size_t)m * (D + 1)] = psys_.seg_stats[m];
        if (hbatched_ && (int)hsysB_.seg_stats.size() == D * nm)
            for (int d = 0; d < D; d++) mstats_[(size_t)m * (D + 1) + 1 + d] = hsysB_.seg_stats[(size_t)d * nm + m];
    }
}

template <int D>
int Flow<D>::member_field(int m, int field, int comp, double *buf, bool set) {
    if (m < 0 || m >= (int)mem_.size() || host_only_) return -1;
    Flow<D> &f = *mem_[m];
    int n = f.field_size(field, comp);
    if (n < 0) return -1;
    if (!buf || n == 0) return n;  // size query
    if (set && field == 6 && comp / D > comp % D) return n;  // upper triangle stored
    int off = remap(m, f.field_offset(field, comp));
    if (set) extrap_hist_ = 0;
    if (set)
        HFH_CHECK(hipMemcpyAsync(X_.p + off, buf, sizeof(double) * n, hipMemcpyHostToDevice, st_));
    else
        HFH_CHECK(hipMemcpyAsync(buf, X_.p + off, sizeof(double) * n, hipMemcpyDeviceToHost, st_));
    HFH_CHECK(hipStreamSynchronize(st_));
    return n;
}

template <int D>
int Flow<D>::member_bc_set_values(int m, int kind, int i, const double *vals) {
    if (m < 0 || m >= (int)mem_.size()) return -1;
    Boundary<D> *b = mem_[m]->bc_of(kind, i);
    if (!b) return -1;
    for (size_t k = 0; k < b->cells.size(); k++) bcval_[bcoff_[m] + b->base + k] = vals[k];
    bc_dirty_ = true;
    return (int)b->cells.size();
}

template <int D>
int Flow<D>::member_stats(int m, int which, SolveStats *out) const {
    if (m < 0 || m >= (int)mem_.size() || which < 0 || which > D) return -1;
    *out = mstats_[(size_t)m * (D + 1) + which];
    return 0;
}

}  // namespace hfh
