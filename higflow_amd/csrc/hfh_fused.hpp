// Fused Krylov solver over one CSR matrix: Jacobi-preconditioned PCG and
// BiCGStab in two launches per iteration (hfh_krylov.hpp), or BiCGStab with
// a smoothed-aggregation V-cycle as right preconditioner (hfh_amg.hpp).
#pragma once
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "hfh_amg.hpp"

#include <cstdint>
#include <cstdio>
#include <unordered_map>
#include "hfh_krylov.hpp"

namespace hfh {

// One precision's fused solver over a CSR matrix whose structure (rp, ci)
// the owner keeps on the device.
template <typename T>
class FusedKrylov {
  public:
    int n = 0, nb = 0, lanes = 1;
    int rows = 1;  // rows per lane group (R of the Krylov kernels)
    const int *rp = nullptr, *ci = nullptr;
    KBuf<T> val, vhat, dinv;  // A, A D^-1, D^-1
    KBuf<T> r, rh, z, p, p2, q, v, v2, s, t, part, ks;
    KBuf<int> ki;
    int pred = 0;  // iterations of the last solve (first chunk)
    // BiCGStab right-preconditioned by a smoothed-aggregation V-cycle
    // (requested before set(); used when the hierarchy builds)
    bool want_amg = false;
    double amg_max_cx = 0;  // > 0: no hierarchy above this operator complexity
    AmgPrecond<T> amg;
    // symmetric systems with a hierarchy: AMG-PCG (HFH_AMG_PCG=0: AMG-BiCGStab)
    bool amg_pcg = true;
    // an FP32 hierarchy of the same matrix that an FP64 solve may use as its
    // preconditioner (the mixed form's FP64 candidate); set by the owner
    AmgPrecond<float> *amg32 = nullptr;
    bool use_amg32 = false;
    KBuf<T> ph, sh;
    // segments (block-diagonal systems): row offsets, multiples of 1024
    int nseg = 1;
    std::vector<int> seg_off;
    KBuf<int> d_wseg, d_wbeg;
    KBuf<int> ndone_;  // {stopped segments, segments} (KSeg::ndone)
    KSeg S;  // device view; S.wseg == nullptr for one segment
    // template rows (see KRows): one id per row, offsets shared by A and
    // A D^-1, one weight table each
    KBuf<int> tcode, tcnt, toff;
    KBuf<T> twv, twh;
    int ntempl = 0;
    long templ_rows = 0;
    long op_bytes_ = 0;  // the operator as one SpMV launch streams it (one matrix)
    // diagonal form of the template rows (see KD_MAX), used with one lane
    // per row
    KBuf<int> dblk, dmask;
    KBuf<T> dwv, dwh;
    bool diag_ = false;
    long op_bytes_diag_ = 0;
    long op_bytes() const { return (diag_ && lanes == 1) ? op_bytes_diag_ : op_bytes_; }
    KRows<T> Av, Ah;  // A and A D^-1 as the kernels read them

    FusedKrylov() = default;
    FusedKrylov(const FusedKrylov &) = delete;
    FusedKrylov &operator=(const FusedKrylov &) = delete;

    // `segs`: row offsets of independent diagonal blocks (first 0, last n,
    // each a multiple of 1024); empty for one system
    void set(int n_, const int *drp, const int *dci, const std::vector<int> &hrp, const std::vector<int> &hci,
             const std::vector<double> &hv, hipStream_t st, const std::vector<int> &segs = {}) {
        n = n_;
        seg_off = segs.empty() ? std::vector<int>{0, n_} : segs;
        nseg = (int)seg_off.size() - 1;
        rp = drp;
        ci = dci;
        std::vector<double> hd(n, 1.0);
        for (int i = 0; i < n; i++)
            for (int k = hrp[i]; k < hrp[i + 1]; k++)
                if (hci[k] == i && hv[k] != 0.0) hd[i] = 1.0 / hv[k];
        std::vector<T> tv(hv.size()), th(hv.size()), td(n);
        for (size_t k = 0; k < hv.size(); k++) {
            tv[k] = (T)hv[k];
            th[k] = (T)(hv[k] * hd[hci[k]]);
        }
        for (int i = 0; i < n; i++) td[i] = (T)hd[i];
        up(val, tv, st);
        up(vhat, th, st);
        up(dinv, td, st);
        Av = KRows<T>();
        Av.rp = rp;
        Av.ci = ci;
        Av.val = val.p;
        Ah = Av;
        Ah.val = vhat.p;
        build_templates(hrp, hci, tv, th, st);
        lane_rule_ = lanes_by_rule(hrp);
        for (KBuf<T> *b : {&r, &rh, &z, &p, &p2, &q, &v, &v2, &s, &t}) b->alloc(n);
        ks.alloc((size_t)KS_N * nseg);
        ki.alloc((size_t)KI_N * nseg);
        pred = 0;
        tune(st);
        // defined contents before the first solve (a cached lane choice skips
        // the timing launches that would otherwise have written them)
        for (KBuf<T> *b : {&r, &rh, &v}) k_fill<T><<<nblocks(n), BLOCK, 0, st>>>(n, b->p, T(0));
        k_fill<T><<<nblocks((long)KP_N * nb), BLOCK, 0, st>>>(KP_N * nb, part.p, T(0));
        k_fill<T><<<nblocks((long)KS_N * nseg), BLOCK, 0, st>>>(KS_N * nseg, ks.p, T(0));
        k_fill<int><<<nblocks((long)KI_N * nseg), BLOCK, 0, st>>>(KI_N * nseg, ki.p, 0);
        amg.ready = false;
        {
            const char *e = std::getenv("HFH_AMG_PCG");
            amg_pcg = !(e && std::strcmp(e, "0") == 0);
        }
        amg.A0 = KRows<T>();
        if (want_amg && amg.build(n, hrp, hci, hv, st, amg_max_cx)) {
            if (diag_ && !(std::getenv("HFH_AMG_DIAG") && std::strcmp(std::getenv("HFH_AMG_DIAG"), "0") == 0))
                amg.A0 = Av;
            ph.alloc(n);
            sh.alloc(n);
        }
    }
    bool amg_on() const { return amg.ready; }

    // y = V(b): this solver's hierarchy, or the FP32 one when use_amg32
    void precondition(const T *b, T *y, const int *kc, hipStream_t st) {
        if (use_amg32 && amg32 && amg32->ready && amg32->nlev > 0)
            amg32->apply<T, T>(b, y, kc, st);
        else
            amg.template apply<T, T>(b, y, kc, st);
    }

    // workgroup -> segment tables for L lanes per row (workgroups of
    // KNT / L rows; segment offsets are multiples of KNT rows)
    void set_segments(int L, hipStream_t st) {
        S = KSeg();
        {
            ndone_.alloc(2);
            const int h[2] = {0, nseg};
            if (hipMemcpyAsync(ndone_.p, h, sizeof h, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                throw std::runtime_error("segment upload failed");
            S.ndone = ndone_.p;
        }
        if (nseg <= 1) return;
        const int rpw = KNT / L;
        std::vector<int> ws(nb), wb(nseg + 1);
        for (int q = 0; q <= nseg; q++) wb[q] = seg_off[q] / rpw;
        for (int q = 0; q < nseg; q++)
            for (int w = wb[q]; w < wb[q + 1]; w++) ws[w] = q;
        d_wseg.alloc(nb);
        d_wbeg.alloc(nseg + 1);
        if (hipMemcpyAsync(d_wseg.p, ws.data(), nb * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d_wbeg.p, wb.data(), (nseg + 1) * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            throw std::runtime_error("segment upload failed");
        S.wseg = d_wseg.p;
        S.wbeg = d_wbeg.p;
    }

    // calls f(integral_constant L, integral_constant R) for the current
    // lanes per row and rows per lane group
    template <typename F>
    void with_shape(F &&f) const {
        using std::integral_constant;
        switch (lanes * 10 + rows) {
        case 11: f(integral_constant<int, 1>{}, integral_constant<int, 1>{}); break;
        case 12: f(integral_constant<int, 1>{}, integral_constant<int, 2>{}); break;
        case 14: f(integral_constant<int, 1>{}, integral_constant<int, 4>{}); break;
        case 21: f(integral_constant<int, 2>{}, integral_constant<int, 1>{}); break;
        case 22: f(integral_constant<int, 2>{}, integral_constant<int, 2>{}); break;
        case 24: f(integral_constant<int, 2>{}, integral_constant<int, 4>{}); break;
        case 41: f(integral_constant<int, 4>{}, integral_constant<int, 1>{}); break;
        default: f(integral_constant<int, 8>{}, integral_constant<int, 1>{}); break;
        }
    }

    // Rows whose (column offset, weight of A, weight of A D^-1) lists
    // repeat become template rows: a list is kept as a template when at
    // least KT_MIN rows share it and it has at most KT_MAX entries, up to
    // KT_TABLE templates.  When fewer than a quarter of the rows would be
    // template rows, or HFH_TEMPLATES=0, the matrix stays plain CSR.
    void build_templates(const std::vector<int> &hrp, const std::vector<int> &hci, const std::vector<T> &tv,
                         const std::vector<T> &th, hipStream_t st) {
        constexpr int KT_MIN = 16, KT_TABLE = 8192;
        ntempl = 0;
        templ_rows = 0;
        op_bytes_ = (long)hci.size() * (4 + (long)sizeof(T)) + (long)(n + 1) * 4;
        if (const char *e = std::getenv("HFH_TEMPLATES"))
            if (std::strcmp(e, "0") == 0) return;
        struct Key {
            std::vector<int> off;
            std::vector<T> a, b;
            bool operator==(const Key &o) const {
                if (off != o.off || a.size() != o.a.size()) return false;
                for (size_t k = 0; k < a.size(); k++)  // bitwise: -0.0 and 0.0 differ
                    if (std::memcmp(&a[k], &o.a[k], sizeof(T)) || std::memcmp(&b[k], &o.b[k], sizeof(T)))
                        return false;
                return true;
            }
        };
        struct Hash {
            size_t operator()(const Key &k) const {
                uint64_t h = 1469598103934665603ull;
                auto mix = [&](const void *p, size_t nb) {
                    const unsigned char *c = static_cast<const unsigned char *>(p);
                    for (size_t q = 0; q < nb; q++) h = (h ^ c[q]) * 1099511628211ull;
                };
                mix(k.off.data(), k.off.size() * sizeof(int));
                mix(k.a.data(), k.a.size() * sizeof(T));
                mix(k.b.data(), k.b.size() * sizeof(T));
                return (size_t)h;
            }
        };
        std::unordered_map<Key, int, Hash> ids;
        std::vector<int> cnt, rowkey(n, -1);
        std::vector<const Key *> keys;
        for (int i = 0; i < n; i++) {
            const int len = hrp[i + 1] - hrp[i];
            if (len > KT_MAX) continue;
            Key k;
            for (int q = hrp[i]; q < hrp[i + 1]; q++) {
                k.off.push_back(hci[q] - i);
                k.a.push_back(tv[q]);
                k.b.push_back(th[q]);
            }
            auto it = ids.find(k);
            if (it == ids.end()) {
                it = ids.emplace(std::move(k), (int)cnt.size()).first;
                cnt.push_back(0);
                keys.push_back(&it->first);
            }
            cnt[it->second]++;
            rowkey[i] = it->second;
        }
        // the most shared lists first, up to the table size
        std::vector<int> order(cnt.size());
        for (size_t q = 0; q < order.size(); q++) order[q] = (int)q;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
        std::vector<int> tid(cnt.size(), -1);
        std::vector<int> hcnt, hoff;
        std::vector<T> hwv, hwh;
        long covered = 0;
        for (int q : order) {
            if (cnt[q] < KT_MIN || ntempl >= KT_TABLE) break;
            tid[q] = ntempl++;
            covered += cnt[q];
            const Key &k = *keys[q];
            hcnt.push_back((int)k.off.size());
            for (int e = 0; e < KT_MAX; e++) {
                const bool has = e < (int)k.off.size();
                hoff.push_back(has ? k.off[e] : 0);
                hwv.push_back(has ? k.a[e] : T(0));
                hwh.push_back(has ? k.b[e] : T(0));
            }
        }
        if (const char *ev = std::getenv("HFH_VERBOSE"))
            if (std::strcmp(ev, "1") == 0)
                std::fprintf(stderr, "hfh: %zu-byte system of %d rows: %ld template rows, %d templates\n",
                             sizeof(T), n, covered, ntempl);
        if (covered * 4 < n) {
            ntempl = 0;
            return;
        }
        templ_rows = covered;
        std::vector<int> code(n);
        long gen_nnz = 0, gen_rows = 0;
        for (int i = 0; i < n; i++) {
            code[i] = rowkey[i] >= 0 ? tid[rowkey[i]] : -1;
            if (code[i] < 0) {
                gen_nnz += hrp[i + 1] - hrp[i];
                gen_rows++;
            }
        }
        // row ids, the general rows' CSR (two row pointers each) and the
        // template table
        op_bytes_ = (long)n * 4 + gen_nnz * (4 + (long)sizeof(T)) + gen_rows * 8 +
                    (long)ntempl * (4 + KT_MAX * (4 + (long)sizeof(T)));
        up_int(tcode, code, st);
        up_int(tcnt, hcnt, st);
        up_int(toff, hoff, st);
        up(twv, hwv, st);
        up(twh, hwh, st);
        if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("template upload failed");
        Av.code = Ah.code = tcode.p;
        Av.tcnt = Ah.tcnt = tcnt.p;
        Av.toff = Ah.toff = toff.p;
        Av.tw = twv.p;
        Ah.tw = twh.p;
        build_diagonals(code, hcnt, hoff, hwv, hwh, gen_nnz, gen_rows, st);
    }

    // The diagonal form (KRows::dblk): each segment's diagonals are the
    // offsets of its longest template; every template must use them in
    // order in every segment it occurs in, or the matrix keeps the plain
    // template form.  HFH_DIAG=0 turns it off.
    void build_diagonals(const std::vector<int> &code, const std::vector<int> &hcnt, const std::vector<int> &hoff,
                         const std::vector<T> &hwv, const std::vector<T> &hwh, long gen_nnz, long gen_rows,
                         hipStream_t st) {
        diag_ = false;
        if (const char *e = std::getenv("HFH_DIAG"))
            if (std::strcmp(e, "0") == 0) return;
        std::vector<int> rseg(n);
        for (int g = 0; g < nseg; g++)
            for (int i = seg_off[g]; i < seg_off[g + 1]; i++) rseg[i] = g;
        for (int g = 1; g < nseg; g++)
            if (seg_off[g] % KNT) return;  // blocks must not straddle segments
        std::vector<int> best(nseg, -1);
        for (int i = 0; i < n; i++) {
            const int t = code[i], g = rseg[i];
            if (t >= 0 && (best[g] < 0 || hcnt[t] > hcnt[best[g]])) best[g] = t;
        }
        std::vector<std::vector<int>> dg(nseg);
        for (int g = 0; g < nseg; g++) {
            if (best[g] < 0) continue;
            if (hcnt[best[g]] > KD_MAX) return;
            dg[g].assign(hoff.begin() + (size_t)best[g] * KT_MAX, hoff.begin() + (size_t)best[g] * KT_MAX + hcnt[best[g]]);
        }
        // slot of each template entry, checked against every segment the
        // template occurs in
        std::vector<int> slot((size_t)ntempl * KT_MAX, -1), seen(ntempl, -1);
        for (int i = 0; i < n; i++) {
            const int t = code[i], g = rseg[i];
            if (t < 0 || seen[t] == g) continue;
            int pos = 0;
            for (int e = 0; e < hcnt[t]; e++) {
                const int off = hoff[(size_t)t * KT_MAX + e];
                while (pos < (int)dg[g].size() && dg[g][pos] != off) pos++;
                if (pos == (int)dg[g].size()) return;  // not an in-order subsequence
                if (seen[t] >= 0 && slot[(size_t)t * KT_MAX + e] != pos) return;
                slot[(size_t)t * KT_MAX + e] = pos++;
            }
            seen[t] = g;
        }
        const int nblk = (n + KNT - 1) / KNT;
        std::vector<int> hb((size_t)nblk * (KD_MAX + 1), 0), hm(ntempl, 0);
        for (int b = 0; b < nblk; b++) {
            const std::vector<int> &d = dg[rseg[(size_t)b * KNT]];
            for (size_t e = 0; e < d.size(); e++) hb[(size_t)b * (KD_MAX + 1) + e] = d[e];
            hb[(size_t)b * (KD_MAX + 1) + KD_MAX] = (int)d.size();
        }
        std::vector<T> wv((size_t)ntempl * KD_MAX, T(0)), wh((size_t)ntempl * KD_MAX, T(0));
        for (int t = 0; t < ntempl; t++)
            for (int e = 0; e < hcnt[t]; e++) {
                const int q = slot[(size_t)t * KT_MAX + e];
                if (q < 0) continue;  // template unused (cannot happen: every template has rows)
                hm[t] |= 1 << q;
                wv[(size_t)t * KD_MAX + q] = hwv[(size_t)t * KT_MAX + e];
                wh[(size_t)t * KD_MAX + q] = hwh[(size_t)t * KT_MAX + e];
            }
        up_int(dblk, hb, st);
        up_int(dmask, hm, st);
        up(dwv, wv, st);
        up(dwh, wh, st);
        if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("diagonal upload failed");
        Av.dblk = Ah.dblk = dblk.p;
        Av.dmask = Ah.dmask = reinterpret_cast<const unsigned *>(dmask.p);
        Av.dw = dwv.p;
        Ah.dw = dwh.p;
        Av.nrows = Ah.nrows = n;
        diag_ = true;
        op_bytes_diag_ = (long)n * 4 + gen_nnz * (4 + (long)sizeof(T)) + gen_rows * 8 +
                         (long)ntempl * (4 + KD_MAX * (long)sizeof(T)) + (long)nblk * (KD_MAX + 1) * 4;
        if (const char *ev = std::getenv("HFH_VERBOSE"))
            if (std::strcmp(ev, "1") == 0)
                std::fprintf(stderr, "hfh: %zu-byte system of %d rows: diagonal form, %zu diagonals in segment 0\n",
                             sizeof(T), n, dg[0].size());
    }

    // r = b - A x and the flags; x is the initial guess.  `cg` asks for
    // PCG (symmetric positive definite systems): Jacobi-PCG, or with a
    // V-cycle AMG-PCG (one symmetric V(1,1) cycle per iteration); without
    // `cg` the solve is Jacobi- or AMG-BiCGStab.  AMG-PCG starts as
    // BiCGStab does (r, r.r, b.b); its z comes from the first V-cycle.
    template <typename TB>
    void init(bool cg, const TB *b, T *x, hipStream_t st, bool xzero = false) {
        cg = cg && !amg.ready;
        with_shape([&](auto Lc, auto Rc) { launch_init<TB, Lc.value, Rc.value>(cg, xzero, b, x, st); });
    }
    // iterations [it0, it0 + count)
    void iterate(bool cg, int it0, int count, int maxit, T rtol, T atol, T *x, hipStream_t st) {
        cg = cg && (!amg.ready || amg_pcg);
        with_shape([&](auto Lc, auto Rc) {
            launch_iter<Lc.value, Rc.value>(cg, it0, count, maxit, rtol, atol, x, st);
        });
    }

    // Times `reps` launches of BiCGStab launch B, t = (A D^-1)(r - alpha v)
    // (the solve's SpMV launch shape), on the current vectors with the flags
    // cleared; call between solves only.  Returns the average ms and the
    // algorithmic bytes of one launch: the operator as stored (template row
    // ids, or CSR structure and values), r, v and rh read once, s and t
    // written once, the partials.
    float probe(int reps, long *bytes, hipStream_t st) {
        k_fill<int><<<nblocks((long)KI_N * nseg), BLOCK, 0, st>>>(KI_N * nseg, ki.p, 0);
        float ms = 0;
        with_shape([&](auto Lc, auto Rc) { ms = time_b<Lc.value, Rc.value>(reps, st, S); });
        *bytes = op_bytes() + (long)n * 5 * sizeof(T) + (long)nb * 6 * sizeof(T);
        return ms;
    }

    // Times `reps` launches of BiCGStab launch A at iteration 1, the solve's
    // longest launch: x += alpha D^-1 p + omega D^-1 s, r = s - omega t,
    // p_new = r + beta (p - omega v), v_new = (A D^-1) p_new.  The launches
    // run on the current s, t, p and v with fixed scalars (omega = 1,
    // beta = alpha = 1/2, no breakdown) and write r, p2, v2 and, in place
    // of the caller's x, the CG-only buffer q; call between solves only
    // (the next solve's init rewrites all of them).  Returns the average ms
    // and the algorithmic bytes of one launch: the operator as stored; s, t,
    // p, v, x, D^-1 and rh read once; x, r, p_new and v_new written once;
    // the partials.
    float probe_a(int reps, long *bytes, hipStream_t st) {
        k_fill<int><<<nblocks((long)KI_N * nseg), BLOCK, 0, st>>>(KI_N * nseg, ki.p, 0);
        // every partial of t.s, t.t, rh.s, rh.t set: omega = 1, rho > 0
        const T slotv[4] = {T(1), T(1), T(2), T(1)};
        const int slot[4] = {KP_TS, KP_TT, KP_RS, KP_RT};
        for (int k = 0; k < 4; k++)
            k_fill<T><<<nblocks(nb), BLOCK, 0, st>>>(nb, part.p + (size_t)slot[k] * nb, slotv[k]);
        std::vector<T> hk((size_t)KS_N * nseg, T(0));
        for (int g = 0; g < nseg; g++) {
            hk[(size_t)g * KS_N + KS_RHO] = T(1);
            hk[(size_t)g * KS_N + KS_ALPHA] = T(0.5);
        }
        if (hipStreamSynchronize(st) != hipSuccess ||
            hipMemcpy(ks.p, hk.data(), hk.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("probe set-up failed");
        float ms = 0;
        with_shape([&](auto Lc, auto Rc) { ms = time_a<Lc.value, Rc.value>(reps, st, S); });
        *bytes = op_bytes() + (long)n * 11 * sizeof(T) + (long)nb * 6 * sizeof(T);
        k_fill<T><<<nblocks((long)KS_N * nseg), BLOCK, 0, st>>>(KS_N * nseg, ks.p, T(0));
        return ms;
    }

    // Times `reps` launches of PCG launch 1 at iteration 1 (stopping test,
    // p = z + beta p_old, q = A z + beta A p_old; the solve's SpMV launch)
    // with the system's own segments, on the current z and p with scalars
    // that neither stop nor break down (r.r = 1 per workgroup above a zero
    // tolerance, rho_old = 1); writes p2, q and the partials; call between
    // solves only.  Returns the average ms and the algorithmic bytes of one
    // launch: the operator as stored, z and p_old read once, p_new and q
    // written once, the partials.
    float probe_cg(int reps, long *bytes, hipStream_t st) {
        k_fill<int><<<nblocks((long)KI_N * nseg), BLOCK, 0, st>>>(KI_N * nseg, ki.p, 0);
        k_fill<T><<<nblocks((long)KP_N * nb), BLOCK, 0, st>>>(KP_N * nb, part.p, T(1));
        std::vector<T> hk((size_t)KS_N * nseg, T(0));
        for (int g = 0; g < nseg; g++) {
            hk[(size_t)g * KS_N + KS_RR0] = T(1);
            hk[(size_t)g * KS_N + KS_RHO] = T(1);
            hk[(size_t)g * KS_N + KS_RHO + 1] = T(1);
        }
        if (hipStreamSynchronize(st) != hipSuccess ||
            hipMemcpy(ks.p, hk.data(), hk.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("probe set-up failed");
        float ms = 0;
        with_shape([&](auto Lc, auto Rc) { ms = time_cg<Lc.value, Rc.value>(reps, st); });
        *bytes = op_bytes() + (long)n * 4 * sizeof(T) + (long)nb * 4 * sizeof(T);
        k_fill<T><<<nblocks((long)KS_N * nseg), BLOCK, 0, st>>>(KS_N * nseg, ks.p, T(0));
        k_fill<int><<<nblocks((long)KI_N * nseg), BLOCK, 0, st>>>(KI_N * nseg, ki.p, 0);
        return ms;
    }

  private:
    template <int L, int R>
    float time_cg(int reps, hipStream_t st) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            throw std::runtime_error("hipEventCreate failed");
        (void)hipEventRecord(a, st);
        for (int rep = 0; rep < reps; rep++)
            k_kcg_spmv<T, L, R><<<nb, KNT / R, 0, st>>>(n, 1, 1 << 30, T(0), T(0), Av, z.p, p.p, p2.p, q.p,
                                                        part.p, nb, ks.p, ki.p, S);
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return reps > 0 ? ms / reps : 0.0f;
    }
    template <int L, int R>
    float time_a(int reps, hipStream_t st, KSeg sg) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            throw std::runtime_error("hipEventCreate failed");
        (void)hipEventRecord(a, st);
        for (int rep = 0; rep < reps; rep++)
            k_kbcg_a<T, L, R><<<nb, KNT / R, 0, st>>>(n, 1, 1 << 30, T(0), T(0), Ah, dinv.p, q.p, r.p, rh.p,
                                                      s.p, t.p, p.p, v.p, p2.p, v2.p, part.p, nb, ks.p, ki.p, sg);
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return reps > 0 ? ms / reps : 0.0f;
    }
    static void up_int(KBuf<int> &d, const std::vector<int> &h, hipStream_t st) {
        d.alloc(h.size());
        if (!h.empty() &&
            hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
            throw std::runtime_error("hipMemcpyAsync failed");
    }
    static void up(KBuf<T> &d, const std::vector<T> &h, hipStream_t st) {
        d.alloc(h.size());
        if (!h.empty() &&
            hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess)
            throw std::runtime_error("hipMemcpyAsync failed");
    }
    template <int L>
    int blocks() const {
        return std::max(1, (n + KNT / L - 1) / (KNT / L));
    }

    template <typename TB, int L, int R>
    void launch_init(bool cg, bool xzero, const TB *b, T *x, hipStream_t st) {
        k_kinit<T, TB, L, R><<<nb, KNT / R, 0, st>>>(n, cg ? 1 : 0, xzero ? 1 : 0, Av, x, b, dinv.p, r.p, z.p,
                                                     rh.p, part.p, nb, ki.p, S);
    }
    // the V-cycle path keeps R = 1 (its launches are bound by the cycle)
    template <int L, int R>
    void launch_iter(bool cg, int it0, int count, int maxit, T rtol, T atol, T *x, hipStream_t st) {
        if (amg.ready && cg) {
            // AMG-PCG: z = V(r), rho = r.z, then the two PCG launches
            const int *kc = S.ndone;
            for (int it = it0; it < it0 + count; it++) {
                T *pold = (it & 1) ? p.p : p2.p, *pnew = (it & 1) ? p2.p : p.p;
                precondition(r.p, z.p, kc, st);
                k_krz<T, L><<<nb, KNT, 0, st>>>(n, r.p, z.p, part.p, nb, ki.p, S);
                k_kcg_spmv<T, L, R><<<nb, KNT / R, 0, st>>>(n, it, maxit, rtol, atol, Av, z.p, pold, pnew, q.p,
                                                            part.p, nb, ks.p, ki.p, S);
                k_kcg_update<T, L, R><<<nb, KNT / R, 0, st>>>(n, it, x, r.p, pnew, q.p, dinv.p, z.p, part.p, nb,
                                                              ks.p, ki.p, S, 0);
            }
            return;
        }
        if (amg.ready) {
            // the cycle's launches return at once when every segment has stopped
            const int *kc = S.ndone;
            for (int it = it0; it < it0 + count; it++) {
                k_abcg_a<T, L><<<nb, KNT, 0, st>>>(n, it, x, r.p, s.p, t.p, ph.p, sh.p, p.p, v.p, part.p, nb, ks.p,
                                                   ki.p, S);
                precondition(p.p, ph.p, kc, st);
                k_abcg_b<T, L><<<nb, KNT, 0, st>>>(n, it, maxit, rtol, atol, Av, ph.p, rh.p, v.p, part.p,
                                                   nb, ks.p, ki.p, S);
                k_abcg_c<T, L><<<nb, KNT, 0, st>>>(n, it, r.p, v.p, s.p, part.p, nb, ks.p, ki.p, S);
                precondition(s.p, sh.p, kc, st);
                k_abcg_d<T, L><<<nb, KNT, 0, st>>>(n, it, Av, sh.p, s.p, rh.p, t.p, part.p, nb, ki.p, S);
            }
            return;
        }
        for (int it = it0; it < it0 + count; it++) {
            T *pold = (it & 1) ? p.p : p2.p, *pnew = (it & 1) ? p2.p : p.p;
            if (cg) {
                k_kcg_spmv<T, L, R><<<nb, KNT / R, 0, st>>>(n, it, maxit, rtol, atol, Av, z.p, pold, pnew, q.p,
                                                            part.p, nb, ks.p, ki.p, S);
                k_kcg_update<T, L, R><<<nb, KNT / R, 0, st>>>(n, it, x, r.p, pnew, q.p, dinv.p, z.p, part.p, nb,
                                                              ks.p, ki.p, S);
            } else {
                T *vold = (it & 1) ? v.p : v2.p, *vnew = (it & 1) ? v2.p : v.p;
                k_kbcg_a<T, L, R><<<nb, KNT / R, 0, st>>>(n, it, maxit, rtol, atol, Ah, dinv.p, x, r.p, rh.p,
                                                          s.p, t.p, pold, vold, pnew, vnew, part.p, nb, ks.p, ki.p,
                                                          S);
                k_kbcg_b<T, L, R><<<nb, KNT / R, 0, st>>>(n, it, maxit, Ah, r.p, rh.p, vnew, s.p, t.p, part.p,
                                                          nb, ks.p, ki.p, S);
            }
        }
    }
    template <int L, int R>
    float time_b(int reps, hipStream_t st, KSeg sg) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            throw std::runtime_error("hipEventCreate failed");
        (void)hipEventRecord(a, st);
        for (int rep = 0; rep < reps; rep++)
            k_kbcg_b<T, L, R><<<nb, KNT / R, 0, st>>>(n, 0, 1 << 30, Ah, r.p, rh.p, v.p, s.p, t.p, part.p, nb,
                                                      ks.p, ki.p, sg);
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return reps > 0 ? ms / reps : 0.0f;
    }
    // one BiCGStab iteration's two launches (A and B) with L lanes per row
    // and R rows per lane group, on zero vectors with the flags cleared and
    // launch A's scalars set so that every segment runs its update and
    // SpMV (as probe_a): reads and writes only buffers of this solver.
    // Best of two timings of 25 launches each.
    template <int L, int R>
    float time_shape(hipStream_t st) {
        lanes = L;
        rows = R;
        nb = blocks<L>();
        part.alloc((size_t)KP_N * nb);
        set_segments(L, st);
        for (KBuf<T> *b : {&r, &rh, &v, &v2, &s, &t, &p, &p2, &q}) k_fill<T><<<nblocks(n), BLOCK, 0, st>>>(n, b->p, T(0));
        k_fill<T><<<nblocks((long)KP_N * nb), BLOCK, 0, st>>>(KP_N * nb, part.p, T(0));
        k_fill<T><<<nblocks((long)KS_N * nseg), BLOCK, 0, st>>>(KS_N * nseg, ks.p, T(0));
        k_fill<int><<<nblocks((long)KI_N * nseg), BLOCK, 0, st>>>(KI_N * nseg, ki.p, 0);
        time_b<L, R>(2, st, S);
        const float tb = std::min(time_b<L, R>(25, st, S), time_b<L, R>(25, st, S));
        // every partial of t.s, t.t, rh.s, rh.t set: omega = 1, rho > 0
        const T slotv[4] = {T(1), T(1), T(2), T(1)};
        const int slot[4] = {KP_TS, KP_TT, KP_RS, KP_RT};
        for (int k = 0; k < 4; k++)
            k_fill<T><<<nblocks(nb), BLOCK, 0, st>>>(nb, part.p + (size_t)slot[k] * nb, slotv[k]);
        std::vector<T> hk((size_t)KS_N * nseg, T(0));
        for (int g = 0; g < nseg; g++) {
            hk[(size_t)g * KS_N + KS_RHO] = T(1);
            hk[(size_t)g * KS_N + KS_ALPHA] = T(0.5);
        }
        if (hipStreamSynchronize(st) != hipSuccess ||
            hipMemcpy(ks.p, hk.data(), hk.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("lane timing set-up failed");
        time_a<L, R>(2, st, S);
        const float ta = std::min(time_a<L, R>(25, st, S), time_a<L, R>(25, st, S));
        k_fill<T><<<nblocks((long)KS_N * nseg), BLOCK, 0, st>>>(KS_N * nseg, ks.p, T(0));
        return ta + tb;
    }
    // Lanes per row by rule: the 99th percentile of the row lengths over 12
    // entries per lane, as a power of two up to 8.  A wave waits for its
    // longest row, so the long tail (rows of an AMR interface or of a 3D
    // least-squares stencil) sets the lanes, not the mean.
    int lane_rule_ = 1;
    static int lanes_by_rule(const std::vector<int> &hrp) {
        const int n = (int)hrp.size() - 1;
        if (n <= 0) return 1;
        std::vector<int> len(n);
        for (int i = 0; i < n; i++) len[i] = hrp[i + 1] - hrp[i];
        const size_t k = std::min<size_t>((size_t)n - 1, (size_t)(0.99 * n));
        std::nth_element(len.begin(), len.begin() + k, len.end());
        const int p99 = len[k];
        return p99 <= 12 ? 1 : p99 <= 24 ? 2 : p99 <= 48 ? 4 : 8;
    }
    // Picks the lanes per row L and rows per lane group R.  The row sums
    // and partial sums depend on L, not on R, so L comes from the rule above
    // (the same on every run and machine) and only R is timed (launches A
    // and B of BiCGStab, with the system's segments; R > 1 only for L <= 2).
    // HFH_KRYLOV_TIMED_LANES=1 times L as well (the fastest shape of this
    // machine; results then depend on the timing).  The choice is kept per
    // matrix shape (rows, entries, segments, precision) for the life of the
    // process.  HFH_KRYLOV_LANES / HFH_KRYLOV_ROWS override the choice.
    void tune(hipStream_t st) {
        static std::mutex mu;
        static std::map<std::tuple<int, long, int, int>, int> chosen;
        long nz = 0;
        {
            int last = 0;
            if (hipMemcpy(&last, rp + n, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess) nz = last;
        }
        const char *tl = std::getenv("HFH_KRYLOV_TIMED_LANES");
        const bool timed_lanes = tl && std::strcmp(tl, "1") == 0;
        int lrule = lane_rule_;
        if (const char *e = std::getenv("HFH_KRYLOV_LANES")) {
            int f = std::atoi(e);
            if (f == 1 || f == 2 || f == 4 || f == 8) lrule = f;
        }
        const auto key = std::make_tuple(n, nz, nseg, (int)sizeof(T) + (timed_lanes ? 0 : 16 * lrule));
        int best = 10 * lrule + 1;
        bool known = false;
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = chosen.find(key);
            if (it != chosen.end()) {
                best = it->second;
                known = true;
            }
        }
        if (!known) {
            const int shapes[8] = {11, 12, 14, 21, 22, 24, 41, 81};
            float tb = 1e30f;
            for (int sh : shapes) {
                if (!timed_lanes && sh / 10 != lrule) continue;
                if (!timed_lanes && lrule > 2) break;  // one shape: nothing to time
                lanes = sh / 10;
                rows = sh % 10;
                float tm = 0;
                with_shape([&](auto Lc, auto Rc) { tm = time_shape<Lc.value, Rc.value>(st); });
                if (tm < tb) {
                    tb = tm;
                    best = sh;
                }
            }
            std::lock_guard<std::mutex> g(mu);
            chosen[key] = best;
        }
        lanes = best / 10;
        rows = best % 10;
        if (const char *e = std::getenv("HFH_KRYLOV_LANES")) {
            int f = std::atoi(e);
            if (f == 1 || f == 2 || f == 4 || f == 8) lanes = f;
        }
        if (const char *e = std::getenv("HFH_KRYLOV_ROWS")) {
            int f = std::atoi(e);
            if (f == 1 || f == 2 || f == 4) rows = f;
        }
        if (lanes > 2) rows = 1;
        nb = lanes == 1 ? blocks<1>() : lanes == 2 ? blocks<2>() : lanes == 4 ? blocks<4>() : blocks<8>();
        part.alloc((size_t)KP_N * nb);
        set_segments(lanes, st);
        if (const char *ev = std::getenv("HFH_VERBOSE"))
            if (std::strcmp(ev, "1") == 0)
                std::fprintf(stderr, "hfh: %zu-byte system of %d rows: %d lanes per row, %d rows per lane group\n",
                             sizeof(T), n, lanes, rows);
    }
};

}  // namespace hfh
