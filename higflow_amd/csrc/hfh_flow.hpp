// The HiG-Flow projection step on one GPU.
//
// Setup (host): build the HiG-Tree domains exactly as the reference does
// (higflow/src/hig-flow-kernel.c:389-507, 823-864; hig-flow-bc.c:27-122),
// then record one step symbolically (see hfh_plan.hpp) into device operator
// tables.  Step (device): interpolation rows -> fused per-facet momentum ->
// cell divergence rows -> Krylov pressure solve -> fused projection rows and
// pressure update; the viscoelastic step adds the velocity-gradient rows and
// the fused per-cell log-conformation + polymeric-stress kernel.
#pragma once
#include <chrono>

#include "hfh_kernels.hpp"
#include "hfh_linsolve.hpp"
#include "hfh_plan.hpp"
#include "hfh_ve.hpp"

namespace hfh {

// Mirrors the fields of higflow_parameters/controllers/ve_parameters the
// step reads (higflow/src/hig-flow-kernel.h:108-214, 425-482).
struct Params {
    double Re, dt, t;
    int projtype, flowtype, tempdiscrtype, convecdiscrtype, secondconvecdiscrtype;
    int ve_model, ve_convec, kernel_transform;
    double De, beta, kernel_tol, giesekus_alpha;
    double rtol, atol;
    int maxit;
    double pin_pressure;
    int order;
    int precision;      // 0 FP64 throughout, 1 mixed (FP32 inner Poisson in FP64 refinement)
    double inner_rtol;  // mixed: FP32 inner tolerance
};

enum Stage { ST_INTERP = 0, ST_MOMENTUM, ST_DIV, ST_SOLVE, ST_PROJECT, ST_VE, ST_NSTAGE };

struct BcSpec {
    std::string amr;
    int userid, ptype;
    int utypes[3];
};

// The driver callbacks a step re-evaluates on the host (higflow_functions,
// higflow/src/hig-flow-kernel.h:216-233); a null member is not evaluated.
struct Functions {
    double (*get_boundary_pressure)(int id, double *center, double t) = nullptr;
    double (*get_boundary_velocity)(int id, double *center, int dim, double t) = nullptr;
    double (*get_source_term)(double *center, double t) = nullptr;
    double (*get_facet_source_term)(double *center, int dim, double t) = nullptr;
};

class FlowBase {
  public:
    virtual ~FlowBase() {}
    virtual int dim() const = 0;
    virtual int add_domain(const std::string &amr) = 0;
    virtual int add_bc(const BcSpec &b) = 0;
    virtual int setup(const Params &p) = 0;
    virtual int field_size(int field, int comp) = 0;
    virtual int field_centers(int field, int comp, double *xyz) = 0;
    virtual int get_field(int field, int comp, double *out) = 0;
    virtual int set_field(int field, int comp, const double *in) = 0;
    // min and max of a field on the device (the reference's Vmin/Vmax log)
    virtual int field_range(int field, int comp, double *mn, double *mx) = 0;
    // lowpoint then highpoint of every pressure cell, in local-id order
    virtual int cell_bounds(double *lohi) = 0;
    virtual int bc_count(int kind, int i) = 0;
    virtual int bc_centers(int kind, int i, double *xyz) = 0;
    virtual int bc_set_values(int kind, int i, const double *vals) = 0;
    // time-dependent boundaries (kind 0 pressure, 1+d velocity component d)
    // take their values from the callbacks at every step
    virtual int set_functions(const Functions &f) = 0;
    virtual int bc_set_valuetype(int kind, int i, int timedependent) = 0;
    virtual int step(bool ve) = 0;
    virtual int stencil(int domain, int dim, const double *center, const double *x, double alpha, int maxn,
                        int *ids, double *w, double *rhs) = 0;
    virtual void set_time(double t) = 0;
    virtual double time() const = 0;
    virtual int pinned() const = 0;
    virtual SolveStats last_solve() const = 0;
    virtual int solve_stats(int which, SolveStats *out) const = 0;
    virtual void timing(double *ms_per_stage, long *calls) = 0;
    virtual void reset_timing() = 0;
    virtual long bytes(int stage) const = 0;
    virtual void set_timing(bool on) = 0;
    virtual int probe(int which, int reps, double *avg_ms, long *bytes) = 0;
    virtual void sync() = 0;
    virtual int operator_csr(int which, int *n, int *nnz, int *sym, int *rp, int *ci, double *v) = 0;
    virtual int tables_digest(unsigned long long *out) const = 0;
    // Batched runs: a Flow made by setup_batch steps independent member
    // simulations (set up host-only, same time step and schemes; Re, De and
    // beta may differ) in the same launches.  Member m's fields, boundary
    // values and solve statistics go through these.
    virtual int setup_batch(const std::vector<FlowBase *> &members) = 0;
    virtual int batch_size() const = 0;
    virtual int member_field(int m, int field, int comp, double *buf, bool set) = 0;
    virtual int member_bc_set_values(int m, int kind, int i, const double *vals) = 0;
    virtual int member_stats(int m, int which, SolveStats *out) const = 0;
};

template <int D>
class Flow : public FlowBase {
  public:
    static constexpr int NSYM = D * (D + 1) / 2;
    // host_only: build the mesh, stencils and operator tables without a
    // device (setup inspection and stencil queries); stepping then fails.
    explicit Flow(bool host_only = false) : host_only_(host_only) {
        if (!host_only_) HFH_CHECK(hipStreamCreate(&st_));
    }
    ~Flow() override {
        if (host_only_) return;
        for (int s = 0; s < ST_NSTAGE; s++) {
            if (ev0_[s]) (void)hipEventDestroy(ev0_[s]);
            if (ev1_[s]) (void)hipEventDestroy(ev1_[s]);
        }
        (void)hipStreamDestroy(st_);
    }
    int dim() const override { return D; }

    int add_domain(const std::string &amr) override {
        if (ready_) return -1;
        auto t = std::make_unique<Tree<D>>();
        if (!read_amr<D>(ids_, amr, *t)) return -2;
        trees_.push_back(std::move(t));
        return (int)trees_.size() - 1;
    }

    int add_bc(const BcSpec &b) override {
        if (ready_) return -1;
        specs_.push_back(b);
        return (int)specs_.size() - 1;
    }

    int setup(const Params &p) override;
    int field_size(int field, int comp) override;
    int field_centers(int field, int comp, double *xyz) override;
    int get_field(int field, int comp, double *out) override;
    int set_field(int field, int comp, const double *in) override;
    int field_range(int field, int comp, double *mn, double *mx) override;
    int cell_bounds(double *lohi) override {
        if (!ready_ || !mem_.empty()) return -1;
        const int n = field_size(1, 0);
        for (int i = 0; i < n; i++) sdp_->bounds_of(i, lohi + (size_t)i * 2 * D, lohi + (size_t)i * 2 * D + D);
        return n;
    }
    int bc_count(int kind, int i) override {
        Boundary<D> *b = bc_of(kind, i);
        return b ? (int)b->cells.size() : -1;
    }
    int bc_centers(int kind, int i, double *xyz) override {
        Boundary<D> *b = bc_of(kind, i);
        if (!b) return -1;
        for (size_t k = 0; k < b->cells.size(); k++) {
            Vec<D> c = b->tree->center(b->cells[k]);
            for (int d = 0; d < D; d++) xyz[k * D + d] = c[d];
        }
        return (int)b->cells.size();
    }
    int bc_set_values(int kind, int i, const double *vals) override {
        Boundary<D> *b = bc_of(kind, i);
        if (!b) return -1;
        for (size_t k = 0; k < b->cells.size(); k++) bcval_[b->base + k] = vals[k];
        bc_dirty_ = true;
        return (int)b->cells.size();
    }
    int set_functions(const Functions &f) override {
        if (!ready_ || !mem_.empty()) return -1;
        fn_ = f;
        return 0;
    }
    int bc_set_valuetype(int kind, int i, int timedependent) override {
        if (!bc_of(kind, i) || !mem_.empty()) return -1;
        if (timedep_.empty()) timedep_.assign((size_t)(D + 1) * specs_.size(), 0);
        timedep_[(size_t)kind * specs_.size() + i] = timedependent != 0;
        return 0;
    }
    int step(bool ve) override;
    int stencil(int domain, int dim, const double *center, const double *x, double alpha, int maxn, int *ids,
                double *w, double *rhs) override;
    int setup_batch(const std::vector<FlowBase *> &members) override;
    int batch_size() const override { return (int)mem_.size(); }
    int member_field(int m, int field, int comp, double *buf, bool set) override;
    int member_bc_set_values(int m, int kind, int i, const double *vals) override;
    int member_stats(int m, int which, SolveStats *out) const override;
    void set_time(double t) override { par_.t = t; }
    double time() const override { return par_.t; }
    int pinned() const override { return desing_ ? pin_ : -1; }
    SolveStats last_solve() const override { return stats_; }
    // 0 pressure, 1+d semi-implicit velocity component d
    int solve_stats(int which, SolveStats *out) const override {
        if (which == 0) {
            *out = stats_;
            return 0;
        }
        if (which >= 1 && which <= D) {
            *out = hstats_[which - 1];
            return 0;
        }
        return -1;
    }
    void timing(double *ms, long *calls) override {
        for (int s = 0; s < ST_NSTAGE; s++) {
            ms[s] = ms_[s];
            calls[s] = calls_[s];
        }
    }
    void reset_timing() override {
        for (int s = 0; s < ST_NSTAGE; s++) {
            ms_[s] = 0;
            calls_[s] = 0;
        }
    }
    long bytes(int stage) const override { return stage >= 0 && stage < ST_NSTAGE ? bytes_[stage] : 0; }
    void set_timing(bool on) override { timing_on_ = on; }
    int probe(int which, int reps, double *avg_ms, long *bytes) override;
    void sync() override { HFH_CHECK(hipStreamSynchronize(st_)); }
    // the assembled linear systems: 0 pressure (pinned unknown eliminated,
    // rows negated), 1+d semi-implicit Helmholtz system of component d
    int operator_csr(int which, int *n, int *nnz, int *sym, int *rp, int *ci, double *v) override {
        if (!ready_) return -1;
        const std::vector<int> *R, *Cc;
        const std::vector<double> *V;
        bool sy;
        if (which == 0) {
            R = &Arp_, Cc = &Aci_, V = &Av_, sy = symmetric_;
        } else if (which >= 1 && which <= D && !Hrp_[which - 1].empty()) {
            R = &Hrp_[which - 1], Cc = &Hci_[which - 1], V = &Hv_[which - 1], sy = Hsym_[which - 1];
        } else {
            return -1;
        }
        *n = (int)R->size() - 1;
        *nnz = (int)Cc->size();
        *sym = sy ? 1 : 0;
        if (rp) std::copy(R->begin(), R->end(), rp);
        if (ci) std::copy(Cc->begin(), Cc->end(), ci);
        if (v) std::copy(V->begin(), V->end(), v);
        return 0;
    }

    // FNV-1a over the recorded host tables: the interpolation groups (rows,
    // constants' symbolic parts, destinations), the momentum and tensor slot
    // tables, the pressure matrix and the Helmholtz systems, as held now
    // (a device solver relabels some of them at upload).  Equal digests mean
    // the same recorded step, e.g. for set-ups on different thread counts.
    int tables_digest(unsigned long long *out) const override {
        if (!ready_) return -1;
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](const void *p, size_t n) {
            const unsigned char *b = static_cast<const unsigned char *>(p);
            for (size_t i = 0; i < n; i++) {
                h ^= b[i];
                h *= 1099511628211ull;
            }
        };
        auto vec = [&](const auto &v) {
            const uint64_t n = v.size();
            mix(&n, sizeof n);
            if (n) mix(v.data(), n * sizeof(v[0]));
        };
        auto pairs = [&](const std::vector<std::pair<int, double>> &v) {
            const uint64_t n = v.size();
            mix(&n, sizeof n);
            for (const auto &e : v) {  // field by field: a pair has padding bytes
                mix(&e.first, sizeof e.first);
                mix(&e.second, sizeof e.second);
            }
        };
        for (const Group *g : {&gDu_, &gK_, &gCv_, &gMom_, &gDiv_, &gGrad_}) {
            vec(g->rowptr);
            vec(g->col);
            vec(g->val);
            vec(g->rhs0);
            vec(g->out);
            for (const auto &b : g->bc) pairs(b);
        }
        for (int d = 0; d < D; d++) {
            vec(momSlots_[d]);
            vec(momFlags_[d]);
            vec(momDelta_[d]);
            vec(Hrp_[d]);
            vec(Hci_[d]);
            vec(Hv_[d]);
            vec(Hrhs0_[d]);
            for (const auto &b : Hbc_[d]) pairs(b);
        }
        vec(veSlots_);
        vec(veFlags_);
        vec(veDelta_);
        vec(Arp_);
        vec(Aci_);
        vec(Av_);
        *out = (unsigned long long)h;
        return 0;
    }

  private:
    // ------------------------------------------------------------ host
    bool host_only_ = false;
    IdGen ids_;
    std::vector<std::unique_ptr<Tree<D>>> trees_, bct_p_, bct_u_;
    std::vector<BcSpec> specs_;
    Params par_{};
    bool ready_ = false, desing_ = true, bc_dirty_ = true;
    int pin_ = -1;
    std::unique_ptr<Domain<D>> sdp_, sdF_, sdED_, sfdu_[D], sfdF_[D];
    std::vector<std::unique_ptr<Boundary<D>>> bcp_, bcu_[D];
    std::vector<double> bcval_;
    Functions fn_;
    std::vector<char> timedep_;                  // [kind * nspecs + i]
    std::vector<double> srcC_, srcV_;            // source-term centres and values (host)
    void boundary_refresh();
    void source_refresh();

    // state layout
    int Nc_ = 0, Nf_[D] = {0}, ne_ = 0;
    int oU_[D], oUS_[D], oFU_[D], oP_, oDP_, oF_, oK_, oKN_, oS_, oDv_, oZero_, oPK_, oPcv_, oPmom_, nX_;
    Group gDu_, gK_, gCv_, gMom_, gDiv_, gGrad_;
    std::vector<int> momSlots_[D];
    std::vector<unsigned> momFlags_[D];
    std::vector<double> momDelta_[D];
    int momNs_ = 0;
    std::vector<int> veSlots_;  // per cell: D*(vl,vr) then D*(kl,kr,kll,krr)
    std::vector<unsigned> veFlags_;
    std::vector<double> veDelta_;
    bool symmetric_ = true;
    std::vector<int> Arp_, Aci_;
    std::vector<double> Av_;
    // semi-implicit Euler: Helmholtz rows per velocity component
    std::vector<int> Hrp_[D], Hci_[D];
    std::vector<double> Hv_[D];
    std::vector<double> Hrhs0_[D];
    std::vector<std::vector<std::pair<int, double>>> Hbc_[D];
    bool Hsym_[D] = {};

    // ---------------------------------------------------------- device
    hipStream_t st_ = nullptr;
    DevBuf<double> X_, bvec_;
    struct DGroup {
        DevBuf<int> rp, col, out, y;
        DevBuf<double> val, cst;
        DevBuf<float> val32, cst32;  // FP32 rows (mixed form)
        // template rows (GTempl): id and column base per row, the table
        DevBuf<int> tcode, tbase, tcnt, toff;
        DevBuf<double> tw;
        DevBuf<float> tw32;
        int ntempl = 0;
        long op_bytes = 0;  // the operator as a launch streams it
        int nrow = 0, ncomp = 1, in_stride = 0, out_stride = 0;
        int lanes = 1;  // lanes per row (from the mean row length)
        bool has_y = false, f32 = false;
        std::vector<int> order;  // device row k is the group's row order[k]
    } dDu_, dK_, dCv_, dMom_, dDiv_, dGrad_;
    // mixed form with HFH_MIXED_KERNELS=1: momentum, velocity-gradient and
    // constitutive evaluations in FP32 arithmetic (FP64 state, eigen-
    // decompositions, divergence and projection).  Off by default: at these
    // sizes the step is latency-bound, and the FP32 right-hand sides cost
    // solver iterations.
    bool k32_ = false;
    DevBuf<int> dMomSlots_[D], dVeSlots_;
    DevBuf<unsigned> dMomFlags_[D], dVeFlags_;
    DevBuf<double> dMomDelta_[D], dVeDelta_;
    DeviceSystem psys_;
    DeviceSystem hsys_[D];  // per component (HFH_KRYLOV=classic)
    DevBuf<double> hb_[D], hcst_[D];
    // the D components as one block-diagonal system (fused solves): block d
    // holds rows [hoff_[d], hoff_[d] + Nf_[d]), blocks padded to 1024 rows
    DeviceSystem hsysB_;
    DevBuf<double> hbB_, hxB_;
    int hoff_[D + 1] = {};
    bool hbatched_ = false;
    SolveStats stats_, hstats_[D];
    // initial guesses extrapolated from the last solutions: HFH_EXTRAP=0 the
    // last solution (the reference's warm start), 1 linear, 2 quadratic in
    // time (default); extrap_hist_ counts the solutions kept
    int extrap_ = 1, extrap_hist_ = 0;
    DevBuf<double> hxPrev_, hxPrev2_, pPrev_, pPrev2_;

    // ---------------------------------------------------------- batch
    // A member's state vector is a sequence of regions (offset, components,
    // length per component).  The batched state keeps each region
    // component-major with the members side by side inside a component,
    // each member's block padded to a multiple of 1024 where the region
    // holds the unknowns of a solve (velocity, pressure), so every member
    // is one segment of the block-diagonal systems.
    struct Region {
        int off, ncomp, len;
        bool pad;
    };
    std::vector<Region> regions() const;
    int remap(int m, int x) const;  // member m's state offset -> batched offset
    int field_offset(int field, int comp) const;
    std::vector<Flow<D> *> mem_;
    std::vector<std::vector<Region>> mreg_;   // per member
    std::vector<std::vector<int>> bmoff_;     // [region][member]: offset inside one component block
    std::vector<int> bbase_, btot_;           // per region: batched start, component block length
    std::vector<int> bcoff_;                  // per member: first boundary value
    std::vector<int> pseg_, hseg_[D];         // member block starts (pressure, velocity component d)
    std::vector<double> momRe_[D], vePar_;    // per-row Re; per-cell Re, De, beta
    DevBuf<double> dMomRe_[D], dVePar_;
    std::vector<SolveStats> mstats_;          // per member: pressure, then velocity components
    void collect_member_stats();
    void upload_device();
    hipEvent_t ev0_[ST_NSTAGE] = {}, ev1_[ST_NSTAGE] = {};
    bool timing_on_ = false;
    bool verbose_ = false;  // HFH_VERBOSE=1: velocity range per step
    DevBuf<double> tT_;     // polymeric stress (field 9), on request
    const double *poly_stress(int comp);
    DevBuf<double> mm_;     // k_minmax output and workgroup partials
    DevBuf<unsigned> mmt_;  // k_minmax ticket
    double ms_[ST_NSTAGE] = {0};
    long calls_[ST_NSTAGE] = {0};
    long bytes_[ST_NSTAGE] = {0};

    Boundary<D> *bc_of(int kind, int i) {
        if (!ready_ || i < 0 || i >= (int)specs_.size()) return nullptr;
        if (kind == 0) return bcp_[i].get();
        if (kind >= 1 && kind <= D) return bcu_[kind - 1][i].get();
        return nullptr;
    }

    // ---------------------------------------------------------- record
    static constexpr int UNUSED = -0x7fffffff;  // slot resolved to the zero entry
    static int oZero_placeholder() { return UNUSED; }
    static bool is_identity(const Stencil &s) {
        return s.ids.size() == 1 && s.w[0] == 1.0 && s.bc.empty() && s.rhs0 == 0.0;
    }
    Stencil stencil_at(Domain<D> &dom, const Vec<D> &c, const Vec<D> &x) {
        Stencil s;
        dom.get_stencil(c, x, 1.0, s);
        return s;
    }
    // slot for a point value: state offset, or -(1+row) of group g
    int point_slot(Domain<D> &dom, int field, const Vec<D> &c, const Vec<D> &x, Group &g, bool force_row = false) {
        Stencil s = stencil_at(dom, c, x);
        if (!force_row && is_identity(s)) return field + s.ids[0];
        auto k = point_key<D>(x);
        DedupKey key{(const void *)&dom, field, {k[0], k[1], k[2]}};
        auto it = g.dedup.find(key);
        if (it != g.dedup.end()) return -(1 + it->second);
        Combo cb;
        cb.add(s, field, 1.0);
        int r = g.add_row(cb.cols, cb.rhs0, cb.bc, -1);
        g.dedup[key] = r;
        return -(1 + r);
    }
    int combo_row(Group &g, const Combo &cb, int out) { return g.add_row(cb.cols, cb.rhs0, cb.bc, out); }

    void record();
    void reset_records();
    void record_momentum(bool ve);
    // Helmholtz rows of consecutive facets (entries per row, then rows)
    struct HRows {
        std::vector<int> ci, len;
        std::vector<double> v, rhs0;
        std::vector<std::vector<std::pair<int, double>>> bc;
    };
    void momentum_facet(int dim, int f, bool ve, Group &g, HRows &hr);
    void append_h(int dim, HRows &hr);
    bool record_momentum_parallel(bool ve, long total);
    template <class F>
    bool for_cells_concurrently(int n, const char *what, F &&fn);
    int record_threads_ = 1;  // threads of a recording pass of a deferred set-up
    bool record_logging_ = false;  // first pass: the threads only log point requests
    std::vector<Domain<D> *> all_domains() {
        std::vector<Domain<D> *> doms{sdp_.get(), sdF_.get(), sdED_.get()};
        for (int d = 0; d < D; d++) {
            doms.push_back(sfdu_[d].get());
            doms.push_back(sfdF_[d].get());
        }
        return doms;
    }
    void record_pressure();
    void record_projection();
    void record_ve();
    void finalize_layout();
    void validate() const;
    void upload_group(DGroup &dg, const Group &g, int pbase, int out_stride, bool has_y, bool f32 = false);
    // one launch of a group: the operator as uploaded, then per (row,
    // component) the constant, the destination index and the write
    static long gdbytes(const DGroup &dg) {
        return dg.nrow ? dg.op_bytes + (long)dg.nrow * dg.ncomp * (8 + 4 + 8) : 0;
    }
    void refold_constants();
    void run_group(DGroup &dg, double alpha, double *dst, int stage);
    template <typename W>
    void launch_rows(DGroup &dg, const W *val, const W *cst, W alpha, double *dst);
    void tbegin(int s) {
        if (!timing_on_) return;
        if (!ev0_[s]) {
            HFH_CHECK(hipEventCreate(&ev0_[s]));
            HFH_CHECK(hipEventCreate(&ev1_[s]));
        }
        HFH_CHECK(hipEventRecord(ev0_[s], st_));
    }
    void tend(int s) {
        if (!timing_on_) return;
        HFH_CHECK(hipEventRecord(ev1_[s], st_));
        HFH_CHECK(hipEventSynchronize(ev1_[s]));
        float m = 0;
        HFH_CHECK(hipEventElapsedTime(&m, ev0_[s], ev1_[s]));
        ms_[s] += m;
        calls_[s]++;
    }
};

// ================================================================= setup
template <int D>
int Flow<D>::setup(const Params &p) {
    if (ready_ || trees_.empty()) return -1;
    // HFH_VERBOSE=1: seconds spent in each set-up phase
    const char *evs = std::getenv("HFH_VERBOSE");
    const bool vset = evs && std::strcmp(evs, "1") == 0;
    auto t_last = std::chrono::steady_clock::now();
    progress_start();
    auto phase = [&](const char *name) {
        auto now = std::chrono::steady_clock::now();
        if (vset) std::fprintf(stderr, "hfh: set-up %s %.3f s\n", name, std::chrono::duration<double>(now - t_last).count());
        t_last = now;
    };
    par_ = p;
    if (par_.order <= 0) par_.order = 2;
    if (par_.maxit <= 0) par_.maxit = 20000;
    if (par_.inner_rtol <= 0) par_.inner_rtol = 1e-5;
    int order = par_.order;
    std::vector<const Tree<D> *> ct;
    std::vector<Tree<D> *> mt;
    for (auto &t : trees_) {
        ct.push_back(t.get());
        mt.push_back(t.get());
    }
    sdp_.reset(new Domain<D>(order, -1));
    sdF_.reset(new Domain<D>(order, -1));
    sdED_.reset(new Domain<D>(order, -1));
    for (auto *d : {sdp_.get(), sdF_.get(), sdED_.get()}) d->trees = ct;
    for (int d = 0; d < D; d++) {
        sfdu_[d].reset(new Domain<D>(order, d));
        sfdF_[d].reset(new Domain<D>(order, d));
        sfdu_[d]->trees = ct;
        sfdF_[d]->trees = ct;
        Domain<D>::adjust_facet_ids(mt, d);
        Domain<D>::adjust_facet_ids(mt, d);
    }
    for (auto &s : specs_) {
        auto tp = std::make_unique<Tree<D>>(), tu = std::make_unique<Tree<D>>();
        if (!read_amr<D>(ids_, s.amr, *tp) || !read_amr<D>(ids_, s.amr, *tu)) return -2;
        bct_p_.push_back(std::move(tp));
        bct_u_.push_back(std::move(tu));
    }
    int maxid = ids_.next + 1;
    sdp_->build_cell_map(maxid);
    sdF_->build_cell_map(maxid);
    sdED_->build_cell_map(maxid);
    for (int d = 0; d < D; d++) {
        sfdu_[d]->compute_sfbi();
        sfdu_[d]->build_facet_map(maxid);
        sfdF_[d]->compute_sfbi();
        sfdF_[d]->build_facet_map(maxid);
    }
    int base = 0;
    desing_ = true;
    for (size_t i = 0; i < specs_.size(); i++) {
        auto b = std::make_unique<Boundary<D>>();
        b->tree = bct_p_[i].get();
        b->type = specs_[i].ptype;
        b->userid = specs_[i].userid;
        b->build(maxid);
        b->base = base;
        base += (int)b->cells.size();
        if (b->type == DIRICHLET) {
            sdp_->dir.push_back(b.get());
            desing_ = false;
        } else {
            sdp_->neu.push_back(b.get());
        }
        bcp_.push_back(std::move(b));
    }
    for (int d = 0; d < D; d++)
        for (size_t i = 0; i < specs_.size(); i++) {
            auto b = std::make_unique<Boundary<D>>();
            b->tree = bct_u_[i].get();
            b->type = specs_[i].utypes[d];
            b->userid = specs_[i].userid;
            b->build(maxid);
            b->base = base;
            base += (int)b->cells.size();
            if (b->type == DIRICHLET) sfdu_[d]->dir.push_back(b.get());
            else sfdu_[d]->neu.push_back(b.get());
            bcu_[d].push_back(std::move(b));
        }
    bcval_.assign(base + 1, 0.0);
    phase("maps and boundaries");
    Nc_ = sdp_->size();
    for (int d = 0; d < D; d++) Nf_[d] = sfdu_[d]->size();
    {
        const Tree<D> &t0 = *trees_[0];
        int c2 = t0.cell_with_point(t0.center(0));
        pin_ = sdp_->map.lookup(t0.n[c2].cid);  // hig-flow-step.c:22-28
    }
    // state layout
    int o = 0;
    for (int d = 0; d < D; d++) {
        oU_[d] = o;
        o += Nf_[d];
    }
    for (int d = 0; d < D; d++) {
        oUS_[d] = o;
        o += Nf_[d];
    }
    for (int d = 0; d < D; d++) {
        oFU_[d] = o;
        o += sfdF_[d]->size();
    }
    oP_ = o;
    o += Nc_;
    oDP_ = o;
    o += Nc_;
    oF_ = o;
    o += sdF_->size();
    int ne = sdED_->size();
    ne_ = ne;
    oK_ = o;
    o += NSYM * ne;
    oKN_ = o;
    o += NSYM * ne;
    oS_ = o;
    o += D * D * ne;
    oDv_ = o;
    o += D * D * ne;
    nX_ = o;  // P regions appended by finalize_layout
    ready_ = true;
    int nth = setup_threads();
    if (nth > 1) {
        // pass 1 records every WLS request, the requests are computed in
        // parallel, pass 2 records the rows from the filled caches
        std::vector<Domain<D> *> doms{sdp_.get(), sdF_.get(), sdED_.get()};
        for (int d = 0; d < D; d++) {
            doms.push_back(sfdu_[d].get());
            doms.push_back(sfdF_[d].get());
        }
        for (auto *dm : doms) dm->begin_deferred();
        record_threads_ = nth;
        record_logging_ = true;
        record();
        record_logging_ = false;
        record_threads_ = 1;
        phase("stencil requests");
        for (auto *dm : doms) dm->resolve_pending(nth);
        phase("least-squares solves");
        reset_records();
        record_threads_ = nth;
    }
    record();
    record_threads_ = 1;
    phase("row recording");
    finalize_layout();
    phase("layout");
    return 0;
}

template <int D>
void Flow<D>::reset_records() {
    for (Group *g : {&gDu_, &gK_, &gCv_, &gMom_, &gDiv_, &gGrad_}) *g = Group();
    for (int d = 0; d < D; d++) {
        momSlots_[d].clear();
        momFlags_[d].clear();
        momDelta_[d].clear();
        Hrp_[d].clear();
        Hci_[d].clear();
        Hv_[d].clear();
        Hrhs0_[d].clear();
        Hbc_[d].clear();
        Hsym_[d] = false;
    }
    veSlots_.clear();
    veFlags_.clear();
    veDelta_.clear();
    Arp_.clear();
    Aci_.clear();
    Av_.clear();
    symmetric_ = true;
}


template <int D>
void Flow<D>::record() {
    bool ve = par_.flowtype == VISCOELASTIC;
    // HFH_VERBOSE=1: seconds of each recorder
    const char *evs = std::getenv("HFH_VERBOSE");
    const bool vset = evs && std::strcmp(evs, "1") == 0;
    auto t0 = std::chrono::steady_clock::now();
    auto lap = [&](const char *name) {
        auto now = std::chrono::steady_clock::now();
        if (vset) std::fprintf(stderr, "hfh: set-up   %s %.3f s\n", name, std::chrono::duration<double>(now - t0).count());
        t0 = now;
    };
    if (ve) record_ve();
    lap("tensor rows");
    record_momentum(ve);
    lap("momentum rows");
    record_pressure();
    lap("pressure rows");
    record_projection();
    lap("projection rows");
}

// hig-flow-discret.c:86-259 (+541-569 for the tensor divergence), in the
// same evaluation order as the reference.
template <int D>
void Flow<D>::record_momentum(bool ve) {
    momNs_ = 2 + 6 * D + (D - 1) + (ve ? D : 0);
    const int ns = momNs_;
    const bool semi = par_.tempdiscrtype == SEMI_IMPLICIT_EULER;
    long total = 0;
    for (int dim = 0; dim < D; dim++) {
        const int nf = Nf_[dim];
        momSlots_[dim].assign((size_t)ns * nf, UNUSED);
        momFlags_[dim].assign(nf, 0);
        momDelta_[dim].assign((size_t)D * nf, 0.0);
        if (semi && nf > 0) {
            Hrp_[dim].assign(1, 0);
            Hci_[dim].clear();
            Hv_[dim].clear();
            Hrhs0_[dim].clear();
            Hbc_[dim].clear();
        }
        total += nf;
    }
    if (record_threads_ > 1 && record_momentum_parallel(ve, total)) return;
    for (int dim = 0; dim < D; dim++) {
        HRows hr;
        for (int f = 0; f < Nf_[dim]; f++) {
            if ((f & 1023) == 0) progress("momentum rows", f, Nf_[dim]);
            momentum_facet(dim, f, ve, gMom_, hr);
        }
        append_h(dim, hr);
    }
}

// The rows of one facet (hig-flow-discret.c:86-259, +541-569), into group g
// and the Helmholtz rows hr; the facet's slots, flags and spacings go to
// their own entries of the slot tables.
template <int D>
void Flow<D>::momentum_facet(int dim, int f, bool ve, Group &g, HRows &hr) {
    Domain<D> &fu = *sfdu_[dim];
    const int nf = Nf_[dim];
    Vec<D> fc = fu.center_of(f), fd = fu.delta_of(f);
    for (int k = 0; k < D; k++) momDelta_[dim][(size_t)k * nf + f] = fd[k];
    auto S = [&](int k) -> int & { return momSlots_[dim][(size_t)k * nf + f]; };
    int ctype = par_.convecdiscrtype;
    // F = FU(fc) + (F(fc-) + F(fc+))/2
    {
        Combo cb;
        cb.add(stencil_at(*sfdF_[dim], fc, fc), oFU_[dim], 1.0);
        Vec<D> pl = fc, pr = fc;
        pl[dim] = fc[dim] - 0.5 * fd[dim];
        pr[dim] = fc[dim] + 0.5 * fd[dim];
        // pressure points are evaluated between the two source reads
        Stencil sl = stencil_at(*sdp_, fc, pl);
        Stencil sr = stencil_at(*sdp_, fc, pr);
        Stencil fl = stencil_at(*sdF_, fc, pl);
        Stencil fr = stencil_at(*sdF_, fc, pr);
        cb.add(fl, oF_, 0.5);
        cb.add(fr, oF_, 0.5);
        S(0) = -(1 + combo_row(g, cb, -1));
        Combo cp;
        double k = 1.0 / (2.0 * 0.5 * fd[dim]);
        cp.add(sr, oP_, k);
        cp.add(sl, oP_, -k);
        S(1) = -(1 + combo_row(g, cp, -1));
    }
    unsigned flags = 0;
    int vk = 2 + 6 * D;
    for (int d2 = 0; d2 < D; d2++) {
        const int b = 2 + 6 * d2;
        if (d2 != dim) {  // hig-flow-eval.c:60-92
            Combo cv;
            double sg[4][2] = {{+1, +1}, {+1, -1}, {-1, +1}, {-1, -1}};
            for (int q = 0; q < 4; q++) {
                Vec<D> p = fc;
                p[dim] = fc[dim] + sg[q][0] * (0.5 * 1.0 * fd[dim]);
                p[d2] = fc[d2] + sg[q][1] * (0.5 * 1.0 * fd[d2]);
                if (sg[q][0] < 0) p[dim] = fc[dim] - 0.5 * 1.0 * fd[dim];
                else p[dim] = fc[dim] + 0.5 * 1.0 * fd[dim];
                if (sg[q][1] < 0) p[d2] = fc[d2] - 0.5 * 1.0 * fd[d2];
                else p[d2] = fc[d2] + 0.5 * 1.0 * fd[d2];
                cv.add(stencil_at(*sfdu_[d2], fc, p), oU_[d2], 0.25);
            }
            S(vk++) = -(1 + combo_row(g, cv, -1));
        }
        Vec<D> xl = fc, xr = fc, xll = fc, xrr = fc;
        xl[d2] = fc[d2] - 1.0 * fd[d2];
        xr[d2] = fc[d2] + 1.0 * fd[d2];
        xrr[d2] = fc[d2] + 2.0 * fd[d2];
        xll[d2] = fc[d2] - 2.0 * fd[d2];
        S(b + 0) = point_slot(fu, oU_[dim], fc, xl, g);
        if (!fu.facet_with_point(xl) && ctype == SECOND_ORDER) ctype = FIRST_ORDER;
        S(b + 1) = point_slot(fu, oU_[dim], fc, xr, g);
        if (!fu.facet_with_point(xr) && ctype == SECOND_ORDER) ctype = FIRST_ORDER;
        S(b + 2) = point_slot(fu, oU_[dim], fc, xrr, g);
        bool irr = fu.facet_with_point(xrr);
        S(b + 3) = point_slot(fu, oU_[dim], fc, xll, g);
        bool ill = fu.facet_with_point(xll);
        if (irr) flags |= 1u << (2 + d2);
        if (ill) flags |= 1u << (2 + D + d2);
        if (ctype == SECOND_ORDER && par_.secondconvecdiscrtype == CUBISTA) {
            (void)point_slot(fu, oU_[dim], fc, xl, g);
            (void)point_slot(fu, oU_[dim], fc, xll, g);
            if (!fu.facet_with_point(xl)) ctype = FIRST_ORDER;
            (void)point_slot(fu, oU_[dim], fc, xr, g);
            (void)point_slot(fu, oU_[dim], fc, xrr, g);
            if (!fu.facet_with_point(xr)) ctype = FIRST_ORDER;
            Vec<D> pr = fc, pl = fc;
            pr[d2] = fc[d2] + 0.5 * fd[d2];
            pl[d2] = fc[d2] - 0.5 * fd[d2];
            S(b + 4) = point_slot(*sfdu_[d2], oU_[d2], fc, pr, g);
            S(b + 5) = point_slot(*sfdu_[d2], oU_[d2], fc, pl, g);
        }
    }
    flags |= (unsigned)ctype & 3u;
    momFlags_[dim][f] = flags;
    bool semi = par_.tempdiscrtype == SEMI_IMPLICIT_EULER;
    if (ve) {  // hig-flow-discret.c:541-569
        int tk = 2 + 6 * D + (D - 1);
        for (int d2 = 0; d2 < D; d2++) {
            int fS = oS_ + (dim * D + d2) * sdED_->size();
            Combo cs;
            if (d2 == dim) {
                Vec<D> pl = fc, pr = fc;
                pl[d2] = fc[d2] - 0.5 * fd[d2];
                pr[d2] = fc[d2] + 0.5 * fd[d2];
                Stencil a = stencil_at(*sdED_, fc, pl), bb = stencil_at(*sdED_, fc, pr);
                double k = 1.0 / (2.0 * 0.5 * fd[d2]);
                cs.add(bb, fS, k);
                cs.add(a, fS, -k);
            } else {
                Stencil l1, l2, r1, r2;
                Vec<D> p = fc;
                p[d2] = fc[d2] - 1.0 * fd[d2];
                p[dim] = fc[dim] - 0.5 * 1.0 * fd[dim];
                l1 = stencil_at(*sdED_, fc, p);
                p[d2] = fc[d2] - 1.0 * fd[d2];
                p[dim] = fc[dim] + 0.5 * 1.0 * fd[dim];
                l2 = stencil_at(*sdED_, fc, p);
                p[d2] = fc[d2] + 1.0 * fd[d2];
                p[dim] = fc[dim] - 0.5 * 1.0 * fd[dim];
                r1 = stencil_at(*sdED_, fc, p);
                p[d2] = fc[d2] + 1.0 * fd[d2];
                p[dim] = fc[dim] + 0.5 * 1.0 * fd[dim];
                r2 = stencil_at(*sdED_, fc, p);
                double k = 1.0 / (2.0 * 1.0 * fd[d2]);
                cs.add(r1, fS, 0.5 * k);
                cs.add(r2, fS, 0.5 * k);
                cs.add(l1, fS, -0.5 * k);
                cs.add(l2, fS, -0.5 * k);
            }
            S(tk + d2) = -(1 + combo_row(g, cs, -1));
        }
    }
    if (semi) {  // hig-flow-step.c:826-853: (I - dt/Re L) row at this facet
        Stencil H;
        double alpha = 0.0;
        for (int d2 = 0; d2 < D; d2++) {
            double w = -par_.dt / (par_.Re * fd[d2] * fd[d2]);
            alpha -= 2.0 * w;
            Vec<D> p = fc;
            p[d2] = fc[d2] + fd[d2];
            fu.get_stencil(fc, p, w, H);
            p[d2] = fc[d2] - fd[d2];
            fu.get_stencil(fc, p, w, H);
        }
        alpha = 1.0 + alpha;
        fu.get_stencil(fc, fc, alpha, H);
        std::vector<std::pair<int, double>> row;
        for (size_t k = 0; k < H.ids.size(); k++) row.emplace_back(H.ids[k], H.w[k]);
        std::sort(row.begin(), row.end());
        for (auto &e : row) {
            hr.ci.push_back(e.first);
            hr.v.push_back(e.second);
        }
        hr.len.push_back((int)row.size());
        hr.rhs0.push_back(H.rhs0);
        hr.bc.push_back(H.bc);
    }
}

template <int D>
void Flow<D>::append_h(int dim, HRows &hr) {
    size_t q = 0;
    for (size_t i = 0; i < hr.len.size(); i++) {
        for (int k = 0; k < hr.len[i]; k++, q++) {
            Hci_[dim].push_back(hr.ci[q]);
            Hv_[dim].push_back(hr.v[q]);
        }
        Hrp_[dim].push_back((int)Hci_[dim].size());
        Hrhs0_[dim].push_back(hr.rhs0[i]);
        Hbc_[dim].push_back(std::move(hr.bc[i]));
    }
    hr = HRows();
}

// Second pass of a deferred set-up: the facets are cut into chunks recorded
// on worker threads, each into a group of its own, then merged in facet
// order.  A chunk's row that the serial walk would have found in the
// group's point dedup (a point evaluated by an earlier chunk) maps to that
// row and is dropped; every other row is appended in the chunk's order, so
// the group, its dedup and the slot tables come out as the serial walk
// writes them.  The point caches are read-only here; a request they cannot
// answer (none is expected after the first pass) sends the whole pass back
// to the serial walk.  Returns false when it did.
template <int D>
bool Flow<D>::record_momentum_parallel(bool ve, long total) {
    struct Chunk {
        int dim, f0, f1;
        Group g;
        HRows hr;
    };
    std::vector<Chunk> chunks;
    const int per = (int)std::max<long>(64, total / ((long)record_threads_ * 8) + 1);
    for (int dim = 0; dim < D; dim++)
        for (int f0 = 0; f0 < Nf_[dim]; f0 += per) chunks.push_back(Chunk{dim, f0, std::min(Nf_[dim], f0 + per), {}, {}});
    std::vector<Domain<D> *> doms = all_domains();
    // first pass: per chunk and domain, the logged point requests
    std::vector<std::vector<std::vector<typename Domain<D>::Req>>> logs(record_logging_ ? chunks.size() : 0);
    for (auto *dm : doms) {
        if (record_logging_) dm->begin_logging();
        else dm->begin_concurrent();
    }
    std::atomic<size_t> next{0};
    std::atomic<long> done{0};
    // a set-up error on a worker (a malformed boundary) stops every worker
    // and is rethrown on the calling thread
    std::exception_ptr err;
    std::mutex err_mu;
    auto run = [&]() {
        try {
            for (size_t c; (c = next.fetch_add(1)) < chunks.size();) {
                Chunk &ch = chunks[c];
                for (int f = ch.f0; f < ch.f1; f++) momentum_facet(ch.dim, f, ve, ch.g, ch.hr);
                if (record_logging_)
                    for (auto *dm : doms) logs[c].push_back(dm->take_log());
                const long d = done.fetch_add(ch.f1 - ch.f0) + (ch.f1 - ch.f0);
                progress("momentum rows", d, total);
            }
        } catch (...) {
            std::lock_guard<std::mutex> g(err_mu);
            if (!err) err = std::current_exception();
            next = chunks.size();
        }
    };
    const int nt = (int)std::min<size_t>((size_t)record_threads_, chunks.size());
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(run);
    run();
    for (auto &t : th) t.join();
    if (err) {
        for (auto *dm : doms) {
            dm->end_logging();
            (void)dm->end_concurrent();
            (void)dm->take_log();
        }
        std::rethrow_exception(err);
    }
    if (record_logging_) {
        // the first pass keeps only the requests: replayed in facet order,
        // domain by domain, they fill the caches as the serial walk does
        for (auto *dm : doms) dm->end_logging();
        for (size_t c = 0; c < chunks.size(); c++)
            for (size_t k = 0; k < doms.size(); k++) doms[k]->replay(logs[c][k]);
        return true;
    }
    bool ok = true;
    for (auto *dm : doms) ok = dm->end_concurrent() && ok;
    if (!ok) {
        if (const char *ev = std::getenv("HFH_VERBOSE"))
            if (std::strcmp(ev, "1") == 0)
                std::fprintf(stderr, "hfh: a point request missed the cache; momentum rows recorded on one thread\n");
        for (int dim = 0; dim < D; dim++) {
            std::fill(momSlots_[dim].begin(), momSlots_[dim].end(), UNUSED);
            std::fill(momFlags_[dim].begin(), momFlags_[dim].end(), 0u);
        }
        return false;
    }
    for (Chunk &ch : chunks) {
        Group &cg = ch.g;
        const int nr = cg.nrows();
        std::vector<const DedupKey *> key(nr, nullptr);
        for (const auto &kv : cg.dedup) key[kv.second] = &kv.first;
        std::vector<int> map(nr);
        for (int r = 0; r < nr; r++) {
            if (key[r]) {
                auto it = gMom_.dedup.find(*key[r]);
                if (it != gMom_.dedup.end()) {
                    map[r] = it->second;
                    continue;
                }
            }
            gMom_.col.insert(gMom_.col.end(), cg.col.begin() + cg.rowptr[r], cg.col.begin() + cg.rowptr[r + 1]);
            gMom_.val.insert(gMom_.val.end(), cg.val.begin() + cg.rowptr[r], cg.val.begin() + cg.rowptr[r + 1]);
            gMom_.rowptr.push_back((int)gMom_.col.size());
            gMom_.rhs0.push_back(cg.rhs0[r]);
            gMom_.bc.push_back(std::move(cg.bc[r]));
            gMom_.out.push_back(cg.out[r]);
            map[r] = gMom_.nrows() - 1;
            if (key[r]) gMom_.dedup[*key[r]] = map[r];
        }
        const int nf = Nf_[ch.dim];
        for (int k = 0; k < momNs_; k++)
            for (int f = ch.f0; f < ch.f1; f++) {
                int &v = momSlots_[ch.dim][(size_t)k * nf + f];
                if (v < 0 && v != UNUSED) v = -(1 + map[-(v + 1)]);
            }
        append_h(ch.dim, ch.hr);
    }
    return true;
}

// Runs fn(i) for i in [0, n) on the row-recording threads, with the point
// caches read-only (each i must write only its own outputs).  Returns false,
// with the outputs to be recomputed, when a request missed a cache.
template <int D>
template <class F>
bool Flow<D>::for_cells_concurrently(int n, const char *what, F &&fn) {
    std::vector<Domain<D> *> doms = all_domains();
    const int per = std::max(64, n / (record_threads_ * 8) + 1);
    const int nchunk = (n + per - 1) / per;
    std::vector<std::vector<std::vector<typename Domain<D>::Req>>> logs(record_logging_ ? nchunk : 0);
    for (auto *dm : doms) {
        if (record_logging_) dm->begin_logging();
        else dm->begin_concurrent();
    }
    std::atomic<int> next{0};
    std::atomic<long> done{0};
    std::exception_ptr err;
    std::mutex err_mu;
    auto run = [&]() {
        try {
            for (int i0; (i0 = next.fetch_add(per)) < n;) {
                const int i1 = std::min(n, i0 + per);
                for (int i = i0; i < i1; i++) fn(i);
                if (record_logging_)
                    for (auto *dm : doms) logs[i0 / per].push_back(dm->take_log());
                progress(what, done.fetch_add(i1 - i0) + (i1 - i0), n);
            }
        } catch (...) {
            std::lock_guard<std::mutex> g(err_mu);
            if (!err) err = std::current_exception();
            next = n;
        }
    };
    const int nt = std::max(1, std::min(record_threads_, nchunk));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(run);
    run();
    for (auto &t : th) t.join();
    if (err) {
        for (auto *dm : doms) {
            dm->end_logging();
            (void)dm->end_concurrent();
            (void)dm->take_log();
        }
        std::rethrow_exception(err);
    }
    if (record_logging_) {
        for (auto *dm : doms) dm->end_logging();
        for (int c = 0; c < nchunk; c++)
            for (size_t k = 0; k < doms.size(); k++) doms[k]->replay(logs[c][k]);
        return true;
    }
    bool ok = true;
    for (auto *dm : doms) ok = dm->end_concurrent() && ok;
    return ok;
}

// hig-flow-step.c:48-131 with the pinned unknown eliminated symmetrically
template <int D>
void Flow<D>::record_pressure() {
    Domain<D> &sp = *sdp_;
    int n = Nc_;
    std::vector<std::vector<std::pair<int, double>>> rows(n);
    std::vector<Combo> bcombo(n);
    // each cell writes only its own row and right-hand-side combination
    auto cell = [&](int c) {
        Vec<D> cc = sp.center_of(c), cd = sp.delta_of(c);
        Combo &cb = bcombo[c];
        for (int dim = 0; dim < D; dim++) {
            Vec<D> pl = cc, pr = cc;
            pl[dim] = cc[dim] - 0.5 * cd[dim];
            pr[dim] = cc[dim] + 0.5 * cd[dim];
            Stencil sl = stencil_at(*sfdu_[dim], cc, pl), sr = stencil_at(*sfdu_[dim], cc, pr);
            double k = 1.0 / (2.0 * 0.5 * cd[dim]);
            cb.add(sr, oUS_[dim], k);
            cb.add(sl, oUS_[dim], -k);
        }
        cb.scale(1.0 / par_.dt);
        Stencil L;
        double alpha = 0.0;
        for (int dim = 0; dim < D; dim++) {
            double w = 1.0 / (cd[dim] * cd[dim]);
            alpha -= 2.0 * w;
            Vec<D> p = cc;
            p[dim] = cc[dim] + cd[dim];
            sp.get_stencil(cc, p, w, L);
            p[dim] = cc[dim] - cd[dim];
            sp.get_stencil(cc, p, w, L);
        }
        sp.get_stencil(cc, cc, alpha, L);
        // b = div/dt + rhs(L): fold rhs(L) with opposite sign (value = sum - rhs)
        cb.rhs0 -= L.rhs0;
        for (auto &bb : L.bc) cb.bc.emplace_back(bb.first, -bb.second);
        for (size_t k = 0; k < L.ids.size(); k++) rows[c].emplace_back(L.ids[k], L.w[k]);
    };
    if (!(record_threads_ > 1 && for_cells_concurrently(n, "pressure rows", cell))) {
        rows.assign(n, {});
        bcombo.assign(n, Combo());
        for (int c = 0; c < n; c++) {
            if ((c & 1023) == 0) progress("pressure rows", c, n);
            cell(c);
        }
    }
    // pinned row (solver.c:142-149): x_pin = p_exact - p[pin] (incremental)
    bool incr = par_.projtype == INCREMENTAL;
    std::vector<double> colpin(n, 0.0);
    if (desing_) {
        for (int c = 0; c < n; c++) {
            if (c == pin_) continue;
            for (auto &e : rows[c])
                if (e.first == pin_) colpin[c] += e.second;
        }
    }
    symmetric_ = true;
    std::vector<std::map<int, double>> M(n);
    for (int c = 0; c < n; c++) {
        if (desing_ && c == pin_) {
            M[c][c] = 1.0;
            continue;
        }
        for (auto &e : rows[c]) {
            if (desing_ && e.first == pin_) continue;
            M[c][e.first] += -e.second;
        }
    }
    for (int c = 0; c < n && symmetric_; c++)
        for (auto &e : M[c]) {
            auto it = M[e.first].find(c);
            double o = it == M[e.first].end() ? 0.0 : it->second;
            if (std::fabs(o - e.second) > 1e-12 * std::max(1.0, std::fabs(e.second))) {
                symmetric_ = false;
                break;
            }
        }
    Arp_.assign(1, 0);
    Aci_.clear();
    Av_.clear();
    for (int c = 0; c < n; c++) {
        for (auto &e : M[c]) {
            Aci_.push_back(e.first);
            Av_.push_back(e.second);
        }
        Arp_.push_back((int)Aci_.size());
    }
    for (int c = 0; c < n; c++) {
        Combo cb = bcombo[c];
        if (desing_ && c == pin_) {
            Combo pc;
            pc.rhs0 = -par_.pin_pressure;  // value = pin_p - p[pin]
            if (incr) pc.add_state(oP_ + pin_, -1.0);
            combo_row(gDiv_, pc, c);
            continue;
        }
        if (desing_ && colpin[c] != 0.0) {
            // b_c - A_c,pin * x_pin
            cb.rhs0 += colpin[c] * par_.pin_pressure;
            if (incr) cb.add_state(oP_ + pin_, colpin[c]);
        }
        cb.scale(-1.0);
        combo_row(gDiv_, cb, c);
    }
}

// hig-flow-step.c:242-308
template <int D>
void Flow<D>::record_projection() {
    bool incr = par_.projtype == INCREMENTAL;
    int src = incr ? oDP_ : oP_;
    for (int dim = 0; dim < D; dim++) {
        Domain<D> &fu = *sfdu_[dim];
        for (int f = 0; f < Nf_[dim]; f++) {
            if ((f & 1023) == 0) progress("projection rows", f, Nf_[dim]);
            Vec<D> fc = fu.center_of(f), fd = fu.delta_of(f);
            Vec<D> pl = fc, pr = fc;
            pl[dim] = fc[dim] - 0.5 * fd[dim];
            pr[dim] = fc[dim] + 0.5 * fd[dim];
            Stencil sl = stencil_at(*sdp_, fc, pl), sr = stencil_at(*sdp_, fc, pr);
            Combo cb;
            double k = 1.0 / (2.0 * 0.5 * fd[dim]);
            cb.add(sr, src, k);
            cb.add(sl, src, -k);
            int r = combo_row(gGrad_, cb, oU_[dim] + f);
            (void)r;
        }
    }
}

}  // namespace hfh

#include "hfh_flow_impl.hpp"
#include "hfh_batch.hpp"
