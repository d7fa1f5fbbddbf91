// Fused Krylov iterations: Jacobi-PCG in two launches per iteration and
// right-Jacobi BiCGStab in three, with the convergence test on the device.
//
// * Every launch begins by summing the per-workgroup partials of the
//   previous launch in a fixed order, redundantly in every workgroup, so each
//   workgroup derives the same scalars (alpha, beta, omega, the stopping
//   decision) without a reduction launch or a host round trip.
// * SpMV operands are formed on the fly at each column from two gathered
//   vectors: CG q = A (z + beta p_old); BiCGStab v = (A D^-1)(r + beta y) with
//   y = p - omega v kept by the previous update, and t = (A D^-1)(r - alpha v).
//   The right Jacobi scaling is folded into the matrix values (A D^-1).
// * Each row is processed by L lanes (L = 1, 2, 4, 8, picked per matrix by
//   timing at set-up); workgroups of 1024 threads hold 1024/L rows, so the
//   partial count stays small while enough waves are in flight to hide the
//   gather latency of short rows.
//
// Stopping: ||r|| <= max(rtol ||b||, atol) on the recursively updated
// residual, a non-finite residual, breakdown, or `maxit` iterations.  The
// launch that sees the test pass sets `done`; later launches return at once,
// so the host enqueues a predicted number of iterations and checks once.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "hfh_kernels.hpp"

namespace hfh {

constexpr int KNT = 1024;  // threads per workgroup of the Krylov kernels
constexpr int KNW = KNT / 64;

// scalar slots (ks, type T) and integer slots (ki)
enum { KS_TOL = 0, KS_BN, KS_RR, KS_RHO, KS_ALPHA = KS_RHO + 2, KS_OMEGA = KS_ALPHA + 2, KS_N = KS_OMEGA + 2 };
enum { KI_DONE = 0, KI_ITERS, KI_CONV, KI_N };
// partial-sum slots (each nb long)
enum { KP_RZ = 0, KP_RR, KP_BB, KP_PQ, KP_RV, KP_TS, KP_TT, KP_N };

// Sums K per-thread values over the workgroup; every thread gets the totals
// (fixed order: wave tree, then the wave sums in sequence).
template <typename T, int K>
__device__ inline void kblock_allsum(T (&v)[K]) {
    __shared__ T sh[K][KNW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        T a = v[k];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
        if (lane == 0) sh[k][w] = a;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        T s = 0;
        for (int j = 0; j < KNW; j++) s += sh[k][j];
        v[k] = s;
    }
    __syncthreads();
}

// totals of K partial arrays (slots `slot[k]` of `part`, nb entries each)
template <typename T, int K>
__device__ inline void sum_partials(const T *part, int nb, const int (&slot)[K], T (&out)[K]) {
#pragma unroll
    for (int k = 0; k < K; k++) {
        T a = 0;
        for (int i = threadIdx.x; i < nb; i += KNT) a += part[(size_t)slot[k] * nb + i];
        out[k] = a;
    }
    kblock_allsum<T, K>(out);
}

template <typename T, int K>
__device__ inline void write_partials(T (&v)[K], T *part, int nb, const int (&slot)[K]) {
    kblock_allsum<T, K>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; k++) part[(size_t)slot[k] * nb + blockIdx.x] = v[k];
}

// the workgroup's view of the done flag, uniform over the workgroup
__device__ inline bool krylov_done(const int *ki) {
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = ki[KI_DONE];
    __syncthreads();
    return s_done != 0;
}

// stopping test of the first launch of an iteration; true = stop
template <typename T>
__device__ inline bool krylov_stop(int it, int maxit, T rr, T bb, T rtol, T atol, bool breakdown, T *ks, int *ki) {
    T tol;
    if (it == 0) {
        tol = rtol * sqrt(bb);
        if (tol < atol) tol = atol;
    } else {
        tol = ks[KS_TOL];
    }
    const bool finite = isfinite(rr);
    const bool conv = finite && sqrt(rr) <= tol;
    const bool stop = conv || !finite || breakdown || it >= maxit;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (it == 0) {
            ks[KS_TOL] = tol;
            ks[KS_BN] = sqrt(bb);
        }
        if (stop) {
            ks[KS_RR] = rr;
            ki[KI_ITERS] = it;
            ki[KI_CONV] = conv ? 1 : 0;
            ki[KI_DONE] = 1;
        }
    }
    return stop;
}

// row `i` of a workgroup of 1024/L rows, lane within the row
template <int L>
__device__ inline int krow() {
    return blockIdx.x * (KNT / L) + (int)threadIdx.x / L;
}
template <int L>
__device__ inline int klane() {
    return (int)threadIdx.x % L;
}

// sum_k val[k] * f(ci[k]) over the row's entries, split over L lanes and
// combined so that every lane holds the row total (fixed order)
template <typename T, int L, typename F>
__device__ inline T row_sum(int b, int e, const int *__restrict__ ci, const T *__restrict__ val, F f) {
    T s = 0;
    for (int k = b + klane<L>(); k < e; k += L) s += val[k] * f(ci[k]);
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

// ------------------------------------------------------------------ init
// r = b - A x; CG: z = dinv r, partials r.z; BiCGStab: rh = r, partial r.r.
// Partials r.r and b.b; clears the flags.
template <typename T, typename TB, int L>
__global__ void __launch_bounds__(KNT) k_kinit(int n, int cg, const int *__restrict__ rp, const int *__restrict__ ci,
                                               const T *__restrict__ val, const T *__restrict__ x,
                                               const TB *__restrict__ b, const T *__restrict__ dinv,
                                               T *__restrict__ r, T *__restrict__ z, T *__restrict__ rh,
                                               T *__restrict__ part, int nb, int *ki) {
    const int i = krow<L>();
    const bool in = i < n;
    const int rb = in ? rp[i] : 0, re = in ? rp[i + 1] : 0;
    T s = row_sum<T, L>(rb, re, ci, val, [&](int j) { return x[j]; });
    T acc[3] = {0, 0, 0};
    if (in && klane<L>() == 0) {
        T bi = (T)b[i];
        T ri = bi - s;
        r[i] = ri;
        if (cg) {
            T zi = dinv[i] * ri;
            z[i] = zi;
            acc[0] = ri * zi;
        } else {
            rh[i] = ri;
            acc[0] = ri * ri;
        }
        acc[1] = ri * ri;
        acc[2] = bi * bi;
    }
    const int sl[3] = {KP_RZ, KP_RR, KP_BB};
    write_partials<T, 3>(acc, part, nb, sl);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ki[KI_DONE] = 0;
        ki[KI_ITERS] = 0;
        ki[KI_CONV] = 0;
    }
}

// ------------------------------------------------------------------- PCG
// launch 1: stopping test, beta = rho / rho_old, p = z + beta p_old (own rows
// to pnew), q = A p with p formed at each column; partial p.q
template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_kcg_spmv(int n, int it, int maxit, T rtol, T atol,
                                                  const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const T *__restrict__ val, const T *__restrict__ z,
                                                  const T *__restrict__ pold, T *__restrict__ pnew,
                                                  T *__restrict__ q, T *__restrict__ part, int nb, T *ks, int *ki) {
    if (krylov_done(ki)) return;
    T s3[3];
    const int sl[3] = {KP_RZ, KP_RR, KP_BB};
    sum_partials<T, 3>(part, nb, sl, s3);
    const T rho = s3[0];
    if (krylov_stop<T>(it, maxit, s3[1], s3[2], rtol, atol, false, ks, ki)) return;
    T beta = 0;
    if (it > 0) {
        T rho_old = ks[KS_RHO + ((it + 1) & 1)];
        beta = rho_old != T(0) ? rho / rho_old : T(0);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ks[KS_RHO + (it & 1)] = rho;
    const int i = krow<L>();
    const bool in = i < n;
    const int rb = in ? rp[i] : 0, re = in ? rp[i + 1] : 0;
    T s;
    if (it == 0)
        s = row_sum<T, L>(rb, re, ci, val, [&](int j) { return z[j]; });
    else
        s = row_sum<T, L>(rb, re, ci, val, [&](int j) { return z[j] + beta * pold[j]; });
    T acc[1] = {0};
    if (in && klane<L>() == 0) {
        T pi = it == 0 ? z[i] : z[i] + beta * pold[i];
        pnew[i] = pi;
        q[i] = s;
        acc[0] = pi * s;
    }
    const int so[1] = {KP_PQ};
    write_partials<T, 1>(acc, part, nb, so);
}

// launch 2: alpha = rho / p.q; x += alpha p; r -= alpha q; z = dinv r;
// partials r.z, r.r
template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_kcg_update(int n, int it, T *__restrict__ x, T *__restrict__ r,
                                                    const T *__restrict__ p, const T *__restrict__ q,
                                                    const T *__restrict__ dinv, T *__restrict__ z,
                                                    T *__restrict__ part, int nb, const T *ks, const int *ki) {
    if (krylov_done(ki)) return;
    T s1[1];
    const int sl[1] = {KP_PQ};
    sum_partials<T, 1>(part, nb, sl, s1);
    const T rho = ks[KS_RHO + (it & 1)];
    const T alpha = s1[0] != T(0) ? rho / s1[0] : T(0);
    const int i = blockIdx.x * (KNT / L) + (int)threadIdx.x;
    T acc[2] = {0, 0};
    if ((int)threadIdx.x < KNT / L && i < n) {
        x[i] = x[i] + alpha * p[i];
        T ri = r[i] - alpha * q[i];
        r[i] = ri;
        T zi = dinv[i] * ri;
        z[i] = zi;
        acc[0] = ri * zi;
        acc[1] = ri * ri;
    }
    const int so[2] = {KP_RZ, KP_RR};
    write_partials<T, 2>(acc, part, nb, so);
}

// -------------------------------------------------------------- BiCGStab
// launch 1: stopping test; rho = rh.r; beta = (rho/rho_old)(alpha/omega);
// p = r + beta y (y = p_old - omega v_old); v = (A D^-1) p; partial rh.v
template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_kbcg_v(int n, int it, int maxit, T rtol, T atol,
                                                const int *__restrict__ rp, const int *__restrict__ ci,
                                                const T *__restrict__ vhat, const T *__restrict__ r,
                                                const T *__restrict__ rh, const T *__restrict__ y,
                                                T *__restrict__ p, T *__restrict__ v, T *__restrict__ part, int nb,
                                                T *ks, int *ki) {
    if (krylov_done(ki)) return;
    T s3[3];
    const int sl[3] = {KP_RZ, KP_RR, KP_BB};
    sum_partials<T, 3>(part, nb, sl, s3);
    const T rho = s3[0];
    T beta = 0;
    bool breakdown = false;
    if (it > 0) {
        const int o = (it + 1) & 1;
        const T rho_old = ks[KS_RHO + o], om = ks[KS_OMEGA + o];
        breakdown = rho == T(0) || rho_old == T(0) || om == T(0);
        if (!breakdown) beta = (rho / rho_old) * (ks[KS_ALPHA + o] / om);
    }
    if (krylov_stop<T>(it, maxit, s3[1], s3[2], rtol, atol, breakdown, ks, ki)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) ks[KS_RHO + (it & 1)] = rho;
    const int i = krow<L>();
    const bool in = i < n;
    const int rb = in ? rp[i] : 0, re = in ? rp[i + 1] : 0;
    T s;
    if (it == 0)
        s = row_sum<T, L>(rb, re, ci, vhat, [&](int j) { return r[j]; });
    else
        s = row_sum<T, L>(rb, re, ci, vhat, [&](int j) { return r[j] + beta * y[j]; });
    T acc[1] = {0};
    if (in && klane<L>() == 0) {
        p[i] = it == 0 ? r[i] : r[i] + beta * y[i];
        v[i] = s;
        acc[0] = rh[i] * s;
    }
    const int so[1] = {KP_RV};
    write_partials<T, 1>(acc, part, nb, so);
}

// launch 2: alpha = rho / rh.v; s = r - alpha v; t = (A D^-1) s;
// partials t.s, t.t
template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_kbcg_t(int n, int it, const int *__restrict__ rp,
                                                const int *__restrict__ ci, const T *__restrict__ vhat,
                                                const T *__restrict__ r, const T *__restrict__ v, T *__restrict__ s,
                                                T *__restrict__ t, T *__restrict__ part, int nb, T *ks,
                                                const int *ki) {
    if (krylov_done(ki)) return;
    T s1[1];
    const int sl[1] = {KP_RV};
    sum_partials<T, 1>(part, nb, sl, s1);
    const T rho = ks[KS_RHO + (it & 1)];
    const T alpha = s1[0] != T(0) ? rho / s1[0] : T(0);
    if (blockIdx.x == 0 && threadIdx.x == 0) ks[KS_ALPHA + (it & 1)] = alpha;
    const int i = krow<L>();
    const bool in = i < n;
    const int rb = in ? rp[i] : 0, re = in ? rp[i + 1] : 0;
    T a = row_sum<T, L>(rb, re, ci, vhat, [&](int j) { return r[j] - alpha * v[j]; });
    T acc[2] = {0, 0};
    if (in && klane<L>() == 0) {
        T si = r[i] - alpha * v[i];
        s[i] = si;
        t[i] = a;
        acc[0] = a * si;
        acc[1] = a * a;
    }
    const int so[2] = {KP_TS, KP_TT};
    write_partials<T, 2>(acc, part, nb, so);
}

// launch 3: omega = t.s / t.t; x += alpha dinv p + omega dinv s;
// r = s - omega t; y = p - omega v; partials rh.r, r.r
template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_kbcg_x(int n, int it, T *__restrict__ x, T *__restrict__ r,
                                                const T *__restrict__ rh, const T *__restrict__ dinv,
                                                const T *__restrict__ p, const T *__restrict__ v,
                                                const T *__restrict__ s, const T *__restrict__ t,
                                                T *__restrict__ y, T *__restrict__ part, int nb, T *ks,
                                                const int *ki) {
    if (krylov_done(ki)) return;
    T s2[2];
    const int sl[2] = {KP_TS, KP_TT};
    sum_partials<T, 2>(part, nb, sl, s2);
    const T omega = s2[1] > T(0) ? s2[0] / s2[1] : T(0);
    const T alpha = ks[KS_ALPHA + (it & 1)];
    if (blockIdx.x == 0 && threadIdx.x == 0) ks[KS_OMEGA + (it & 1)] = omega;
    const int i = blockIdx.x * (KNT / L) + (int)threadIdx.x;
    T acc[2] = {0, 0};
    if ((int)threadIdx.x < KNT / L && i < n) {
        const T di = dinv[i], pi = p[i], si = s[i];
        x[i] = x[i] + alpha * (di * pi) + omega * (di * si);
        T ri = si - omega * t[i];
        r[i] = ri;
        y[i] = pi - omega * v[i];
        acc[0] = rh[i] * ri;
        acc[1] = ri * ri;
    }
    const int so[2] = {KP_RZ, KP_RR};
    write_partials<T, 2>(acc, part, nb, so);
}

// ---------------------------------------------------------------- driver
template <typename T>
struct KBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t k) {
        if (k <= n && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        if (k == 0) k = 1;
        if (hipMalloc((void **)&p, k * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        n = k;
    }
    ~KBuf() {
        if (p) (void)hipFree(p);
    }
};

// One precision's fused solver over a CSR matrix whose structure (rp, ci)
// the owner keeps on the device.
template <typename T>
class FusedKrylov {
  public:
    int n = 0, nb = 0, lanes = 1;
    const int *rp = nullptr, *ci = nullptr;
    KBuf<T> val, vhat, dinv;                        // A, A D^-1, D^-1
    KBuf<T> r, rh, z, p, p2, q, v, y, s, t, part, ks;
    KBuf<int> ki;
    int pred = 0;  // iterations of the last solve (first chunk)

    FusedKrylov() = default;
    FusedKrylov(const FusedKrylov &) = delete;
    FusedKrylov &operator=(const FusedKrylov &) = delete;

    void set(int n_, const int *drp, const int *dci, const std::vector<int> &hrp, const std::vector<int> &hci,
             const std::vector<double> &hv, hipStream_t st) {
        n = n_;
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
        val.alloc(tv.size());
        vhat.alloc(th.size());
        dinv.alloc(n);
        up(val, tv, st);
        up(vhat, th, st);
        up(dinv, td, st);
        for (KBuf<T> *b : {&r, &rh, &z, &p, &p2, &q, &v, &y, &s, &t}) b->alloc(n);
        ks.alloc(KS_N);
        ki.alloc(KI_N);
        pred = 0;
        tune(st);
    }

    // r = b - A x and the flags; x is the initial guess
    template <typename TB>
    void init(bool cg, const TB *b, T *x, hipStream_t st) {
        switch (lanes) {
        case 1: launch_init<TB, 1>(cg, b, x, st); break;
        case 2: launch_init<TB, 2>(cg, b, x, st); break;
        case 4: launch_init<TB, 4>(cg, b, x, st); break;
        default: launch_init<TB, 8>(cg, b, x, st); break;
        }
    }
    // iterations [it0, it0 + count)
    void iterate(bool cg, int it0, int count, int maxit, T rtol, T atol, T *x, hipStream_t st) {
        switch (lanes) {
        case 1: launch_iter<1>(cg, it0, count, maxit, rtol, atol, x, st); break;
        case 2: launch_iter<2>(cg, it0, count, maxit, rtol, atol, x, st); break;
        case 4: launch_iter<4>(cg, it0, count, maxit, rtol, atol, x, st); break;
        default: launch_iter<8>(cg, it0, count, maxit, rtol, atol, x, st); break;
        }
    }

  private:
    static void up(KBuf<T> &d, const std::vector<T> &h, hipStream_t st) {
        if (!h.empty() &&
            hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess)
            throw std::runtime_error("hipMemcpyAsync failed");
    }
    template <int L>
    int blocks() const {
        return std::max(1, (n + KNT / L - 1) / (KNT / L));
    }
    void size_partials() { part.alloc((size_t)KP_N * nb); }

    template <typename TB, int L>
    void launch_init(bool cg, const TB *b, T *x, hipStream_t st) {
        k_kinit<T, TB, L><<<nb, KNT, 0, st>>>(n, cg ? 1 : 0, rp, ci, val.p, x, b, dinv.p, r.p, z.p, rh.p, part.p, nb,
                                              ki.p);
    }
    template <int L>
    void launch_iter(bool cg, int it0, int count, int maxit, T rtol, T atol, T *x, hipStream_t st) {
        for (int it = it0; it < it0 + count; it++) {
            if (cg) {
                T *pold = (it & 1) ? p.p : p2.p, *pnew = (it & 1) ? p2.p : p.p;
                k_kcg_spmv<T, L><<<nb, KNT, 0, st>>>(n, it, maxit, rtol, atol, rp, ci, val.p, z.p, pold, pnew, q.p,
                                                     part.p, nb, ks.p, ki.p);
                k_kcg_update<T, L><<<nb, KNT, 0, st>>>(n, it, x, r.p, pnew, q.p, dinv.p, z.p, part.p, nb, ks.p, ki.p);
            } else {
                k_kbcg_v<T, L><<<nb, KNT, 0, st>>>(n, it, maxit, rtol, atol, rp, ci, vhat.p, r.p, rh.p, y.p, p.p, v.p,
                                                   part.p, nb, ks.p, ki.p);
                k_kbcg_t<T, L><<<nb, KNT, 0, st>>>(n, it, rp, ci, vhat.p, r.p, v.p, s.p, t.p, part.p, nb, ks.p, ki.p);
                k_kbcg_x<T, L><<<nb, KNT, 0, st>>>(n, it, x, r.p, rh.p, dinv.p, p.p, v.p, s.p, t.p, y.p, part.p, nb,
                                                   ks.p, ki.p);
            }
        }
    }
    template <int L>
    float time_lanes(hipStream_t st) {
        lanes = L;
        nb = blocks<L>();
        size_partials();
        // t = (A D^-1)(r - alpha v) on zero vectors, flags cleared: reads and
        // writes only buffers of this solver
        k_fill<T><<<nblocks(n), BLOCK, 0, st>>>(n, r.p, T(0));
        k_fill<T><<<nblocks(n), BLOCK, 0, st>>>(n, v.p, T(0));
        k_fill<T><<<nblocks((long)KP_N * nb), BLOCK, 0, st>>>(KP_N * nb, part.p, T(0));
        k_fill<T><<<1, BLOCK, 0, st>>>(KS_N, ks.p, T(0));
        k_fill<int><<<1, BLOCK, 0, st>>>(KI_N, ki.p, 0);
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            throw std::runtime_error("hipEventCreate failed");
        for (int rep = 0; rep < 2; rep++)
            k_kbcg_t<T, L><<<nb, KNT, 0, st>>>(n, 0, rp, ci, vhat.p, r.p, v.p, s.p, t.p, part.p, nb, ks.p, ki.p);
        (void)hipEventRecord(a, st);
        for (int rep = 0; rep < 20; rep++)
            k_kbcg_t<T, L><<<nb, KNT, 0, st>>>(n, 0, rp, ci, vhat.p, r.p, v.p, s.p, t.p, part.p, nb, ks.p, ki.p);
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ms;
    }
    // picks the lanes per row with the fastest SpMV launch
    void tune(hipStream_t st) {
        int best = 1;
        float tb = 1e30f;
        const float t1 = time_lanes<1>(st), t2 = time_lanes<2>(st), t4 = time_lanes<4>(st), t8 = time_lanes<8>(st);
        const float tl[4] = {t1, t2, t4, t8};
        const int ll[4] = {1, 2, 4, 8};
        for (int k = 0; k < 4; k++)
            if (tl[k] < tb) {
                tb = tl[k];
                best = ll[k];
            }
        if (const char *e = std::getenv("HFH_KRYLOV_LANES")) {
            int f = std::atoi(e);
            if (f == 1 || f == 2 || f == 4 || f == 8) best = f;
        }
        lanes = best;
        nb = best == 1 ? blocks<1>() : best == 2 ? blocks<2>() : best == 4 ? blocks<4>() : blocks<8>();
        size_partials();
    }
};

}  // namespace hfh
