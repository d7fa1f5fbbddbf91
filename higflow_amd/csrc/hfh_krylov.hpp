// Fused Krylov iterations: Jacobi-PCG and right-Jacobi BiCGStab in two
// launches per iteration each, with the convergence test on the device.
//
// * Every launch sums the per-workgroup partials of the previous launch in a
//   fixed order, redundantly in every workgroup, so each workgroup derives
//   the same scalars (alpha, beta, omega, the stopping decision) without a
//   reduction launch or a host round trip.  One wave sums the partials
//   while the others gather their rows; one barrier then publishes the
//   totals (the `done` flag rides along).
// * Gathered vectors are summed separately per row and combined once the
//   scalars are known (CG: q = A z + beta A p_old; BiCGStab:
//   v = (A D^-1)(s - omega t) + beta (A D^-1)(p_old - omega v_old)).
//   The right Jacobi scaling is folded into the matrix values (A D^-1).
// * BiCGStab takes two launches: launch A applies the previous iteration's
//   update (x, r = s - omega t) with omega from the t.s, t.t partials and
//   rho = rh.s - omega rh.t, then forms p and v = (A D^-1) p; launch B forms
//   s = r - alpha v and t = (A D^-1) s.  The stopping test runs in launch B
//   on the exact r.r of the residual launch A wrote.
// * Each row is processed by L lanes (L = 1, 2, 4, 8, picked per matrix by
//   timing at set-up); a workgroup holds 1024/L rows.  With R rows per lane
//   group (R = 1, 2, 4) the workgroup has 1024/R threads and each lane
//   group gathers R rows at once, which multiplies the loads in flight per
//   compute unit when the row count, not the bytes, bounds the launch.
//   The rows of a workgroup, the per-lane entry order and the partial-sum
//   tree are the same for every R, so the iterates do not depend on R.
//
// Stopping: ||r|| <= max(rtol ||b||, atol) on the recursively updated
// residual, a non-finite residual, breakdown, divergence, or `maxit`
// iterations.  The launch that sees the test pass sets `done`; later
// launches return at once, so the host enqueues a predicted number of
// iterations and checks once.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "hfh_kernels.hpp"

namespace hfh {

constexpr int KNT = 1024;  // threads per workgroup of the Krylov kernels
constexpr int KNW = KNT / 64;

// scalar slots (ks, type T) and integer slots (ki)
enum {
    KS_TOL = 0,
    KS_BN,
    KS_RR,   // r.r when the solve stopped
    KS_RR0,  // max(r.r, b.b) at the start (divergence scale)
    KS_RHO,
    KS_ALPHA = KS_RHO + 2,
    KS_N = KS_ALPHA + 2
};
// KI_DONE is set by a launch that stops without updating anything after the
// decision (every workgroup takes the same decision, so one that starts late
// and reads the flag already set skips nothing it should have done).
// Launch A of BiCGStab updates x and r before it can detect breakdown, so it
// reports breakdown in KI_BRK + (it & 1) and reads only the slot of the
// previous iteration: no workgroup skips its update on a flag its own launch
// set.
enum { KI_DONE = 0, KI_ITERS, KI_CONV, KI_BRK, KI_N = KI_BRK + 2 };
// partial-sum slots (each nb long)
enum { KP_RZ = 0, KP_RR, KP_BB, KP_PQ, KP_RV, KP_TS, KP_TT, KP_RS, KP_RT, KP_N };

// Block-diagonal systems: rows are split into segments padded to multiples
// of 1024 rows, each solved as an independent system in the same launches.
// wseg[w] is the segment of workgroup w, [wbeg[s], wbeg[s+1]) its
// workgroups; each segment has its own scalars (ks + s KS_N), flags
// (ki + s KI_N) and partial range.  wseg == nullptr: one segment.
struct KSeg {
    const int *wseg = nullptr;
    const int *wbeg = nullptr;
    // stopped segments and the segment count: {count, nseg}; the V-cycle
    // skips its launches once every segment has stopped
    int *ndone = nullptr;
};
struct KCtx {
    int w0, w1;  // the segment's workgroups
    bool lead;   // first workgroup of the segment (writes its scalars)
};
template <typename T>
__device__ inline KCtx kctx(const KSeg &S, int nb, T *&ks, int *&ki) {
    KCtx c;
    if (S.wseg) {
        const int sg = S.wseg[blockIdx.x];
        c.w0 = S.wbeg[sg];
        c.w1 = S.wbeg[sg + 1];
        ks += (size_t)sg * KS_N;
        ki += (size_t)sg * KI_N;
    } else {
        c.w0 = 0;
        c.w1 = nb;
    }
    c.lead = (int)blockIdx.x == c.w0;
    return c;
}

// Wave 0 sums K partial arrays (nb entries each, fixed order: lane strides,
// then a butterfly) plus the done flag (slot K) and leaves the totals in
// `sh`; the other waves go on with their row gathers meanwhile.  Every
// workgroup computes the same totals.  Read them after kblock_sync().
// The loads of a chunk (C partials of every slot per lane, 8 values in
// flight) are issued together before any is summed, so the wave waits for
// one memory round trip per chunk rather than one per partial; the sums
// keep the order of a plain loop over the partials.
template <typename T, int K>
__device__ inline void reduce_partials(const T *part, int nb, const KCtx &c, const int (&slot)[K], const int *ki,
                                       T *sh, int brk = -1) {
    if (threadIdx.x >= 64) return;
    const int lane = threadIdx.x;
    constexpr int C = K >= 8 ? 1 : 8 / K;
    T a[K];
#pragma unroll
    for (int k = 0; k < K; k++) a[k] = 0;
#pragma unroll 1
    for (int i0 = c.w0 + lane; i0 < c.w1; i0 += 64 * C) {
        T v[K][C];
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int q = 0; q < C; q++) {
                const int i = i0 + q * 64;
                v[k][q] = i < c.w1 ? part[(size_t)slot[k] * nb + i] : T(0);
            }
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int q = 0; q < C; q++)
                if (i0 + q * 64 < c.w1) a[k] += v[k][q];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        T t = a[k];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (lane == 0) sh[k] = t;
    }
    if (lane == 0) sh[K] = (T)(ki[KI_DONE] + (brk >= 0 ? ki[brk] : 0));
}

__device__ inline void kblock_sync() { __syncthreads(); }

// Sums K values over the workgroup's rows and writes the totals to
// part[slot[k] * nb + blockIdx.x].  The workgroup has KNT/R threads; v[s]
// holds this thread's values for slab s, the rows a workgroup of KNT
// threads would give to its threads s KNT/R .. (s+1) KNT/R - 1.  Fixed
// order, the same for every R: a tree within each wave of 64 of those
// threads, then the KNW wave sums in sequence by one thread.  Last
// shared-memory use of a launch.
template <typename T, int K, int R = 1>
__device__ inline void write_partials(T (&v)[R][K], T *part, int nb, const int (&slot)[K]) {
    constexpr int NW = KNW / R;  // waves of this workgroup
    __shared__ T sw[K][KNW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int sl = 0; sl < R; sl++)
#pragma unroll
        for (int k = 0; k < K; k++) {
            T a = v[sl][k];
            for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
            if (lane == 0) sw[k][sl * NW + w] = a;
        }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        T a = 0;
        for (int j = 0; j < KNW; j++) a += sw[k][j];
        part[(size_t)slot[k] * nb + blockIdx.x] = a;
    }
}
template <typename T, int K>
__device__ inline void write_partials(T (&v)[K], T *part, int nb, const int (&slot)[K]) {
    T w[1][K];
#pragma unroll
    for (int k = 0; k < K; k++) w[0][k] = v[k];
    write_partials<T, K, 1>(w, part, nb, slot);
}

// stopping test; `first` computes the tolerance from ||b||.  Besides
// convergence: a non-finite residual, breakdown, `maxit`, or divergence
// (||r|| above 1e5 max(||r0||, ||b||), PETSc's default dtol).
template <typename T>
__device__ inline bool krylov_stop(bool lead, bool first, int it, int maxit, T rr, T bb, T rtol, T atol, bool breakdown,
                                   T *ks, int *ki, int *ndone = nullptr) {
    T tol, rr0;
    if (first) {
        tol = rtol * sqrt(bb);
        if (tol < atol) tol = atol;
        rr0 = rr > bb ? rr : bb;
    } else {
        tol = ks[KS_TOL];
        rr0 = ks[KS_RR0];
    }
    const bool finite = isfinite(rr);
    const bool conv = finite && sqrt(rr) <= tol;
    const bool diverged = rr > T(1e10) * rr0;
    const bool stop = conv || !finite || breakdown || diverged || it >= maxit;
    if (lead && threadIdx.x == 0) {
        if (first) {
            ks[KS_TOL] = tol;
            ks[KS_BN] = sqrt(bb);
            ks[KS_RR0] = rr0;
        }
        if (stop) {
            ks[KS_RR] = rr;
            ki[KI_ITERS] = it;
            ki[KI_CONV] = conv ? 1 : 0;
            ki[KI_DONE] = 1;
            if (ndone) atomicAdd(ndone, 1);
        }
    }
    return stop;
}

// row of this thread (workgroups of 1024/L rows) and its lane
template <int L>
__device__ inline int krow() {
    return blockIdx.x * (KNT / L) + (int)threadIdx.x / L;
}
// row of this thread's slab s in a workgroup of KNT/R threads: the row
// thread s KNT/R + threadIdx.x of a KNT-thread workgroup would take
template <int L, int R>
__device__ inline int krow_s(int s) {
    return blockIdx.x * (KNT / L) + (s * (KNT / R) + (int)threadIdx.x) / L;
}
template <int L>
__device__ inline int klane() {
    return (int)threadIdx.x % L;
}

// A matrix as the Krylov kernels read it: CSR (rp, ci, val) for every row,
// and optionally template rows.  Rows of a uniform mesh repeat a few
// stencils; such a row stores only its template id code[i] >= 0, and
// template t holds tcnt[t] <= KT_MAX entries (column offset from the row,
// weight) at t * KT_MAX in toff / tw, in the row's CSR order.  code[i] < 0,
// or code == nullptr, reads the CSR row.  Template rows then cost 4 bytes
// of operator instead of 4 + 12 nnz (FP64), and sum in the same order.
constexpr int KT_MAX = 16;
// Diagonal form of the template rows (uniform-mesh systems): when every
// template's column offsets are an in-order subsequence of one list of at
// most KD_MAX offsets per 1024-row block (the block's "diagonals"), a
// template keeps only its weights in the diagonal slots and a mask of the
// slots it uses.  The kernels then know where to gather before the row's
// template id arrives: the offsets come from the block (one uniform load
// per workgroup), so the vector loads and the id load are in flight
// together instead of one after the other.  The entries are summed in the
// same order as the template (and the CSR row).  Used with one lane per
// row, whose workgroups hold exactly one 1024-row block, by the launches
// that gather one or two vectors.
constexpr int KD_MAX = 7;
template <typename T>
struct KRows {
    const int *rp = nullptr, *ci = nullptr;
    const T *val = nullptr;
    const int *code = nullptr, *tcnt = nullptr, *toff = nullptr;
    const T *tw = nullptr;
    // diagonal form: per block KD_MAX offsets and their count (dblk[b *
    // (KD_MAX + 1) + KD_MAX]); per template a slot mask and KD_MAX weights
    const int *dblk = nullptr;
    const unsigned *dmask = nullptr;
    const T *dw = nullptr;
    int nrows = 0;
};

// sums over row i[s]'s entries of val[k] * vec[m][col[k]] for the R rows
// of this lane group, split over L lanes and combined so that every lane
// holds the row totals (fixed order: lane l takes entries l, l + L, ... of
// each row in sequence; then a butterfly).  The entries are taken U at a
// time for all R rows together: their column indices are loaded, then all
// their vector values, so a lane group waits for one dependent round trip
// per U entries of all its rows.  A template row sums the same entries in
// the same order as its CSR form, and the order does not depend on R or U.
template <typename T, int L, int M, int R>
__device__ inline void row_sums_r(const KRows<T> &A, const int (&i)[R], const bool (&in)[R],
                                  const T *const (&vec)[M], T (&out)[R][M]) {
    constexpr int U = R >= 4 ? 2 : 4;
    int rb[R], re[R], base[R], t[R];
    const int *cp[R];
    const T *vp[R];
    bool diag = false;
    // (M <= 2: four gathered vectors in seven slots would cost the FP64
    // BiCGStab launch A half its occupancy)
    if constexpr (L == 1 && M <= 2) {
        if (A.dw) {
            // diagonal form: gathers first (offsets of this workgroup's
            // block), then the template ids, then weights and masks
            diag = true;
            const int *db = A.dblk + (size_t)blockIdx.x * (KD_MAX + 1);
            const int nd = db[KD_MAX];
            T xv[R][KD_MAX][M];
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int e = 0; e < KD_MAX; e++) {
                    int j = i[r] + db[e];
                    j = j < 0 ? 0 : (j >= A.nrows ? A.nrows - 1 : j);
#pragma unroll
                    for (int m = 0; m < M; m++) xv[r][e][m] = (in[r] && e < nd) ? vec[m][j] : T(0);
                }
#pragma unroll
            for (int r = 0; r < R; r++) t[r] = in[r] ? A.code[i[r]] : -1;
#pragma unroll
            for (int r = 0; r < R; r++) {
#pragma unroll
                for (int m = 0; m < M; m++) out[r][m] = 0;
                if (t[r] >= 0) {
                    const unsigned msk = A.dmask[t[r]];
                    const T *w = A.dw + (size_t)t[r] * KD_MAX;
#pragma unroll
                    for (int e = 0; e < KD_MAX; e++) {
                        const T a = w[e];
#pragma unroll
                        for (int m = 0; m < M; m++)
                            if ((msk >> e) & 1u) out[r][m] += a * xv[r][e][m];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
        if (!diag) t[r] = (in[r] && A.code) ? A.code[i[r]] : -1;
    int len = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (!diag) {
#pragma unroll
            for (int m = 0; m < M; m++) out[r][m] = 0;
        }
        if (diag && t[r] >= 0) {  // summed above
            rb[r] = re[r] = base[r] = 0;
            cp[r] = A.ci;
            vp[r] = A.val;
        } else if (t[r] >= 0) {
            rb[r] = t[r] * KT_MAX;
            re[r] = rb[r] + A.tcnt[t[r]];
            cp[r] = A.toff;
            vp[r] = A.tw;
            base[r] = i[r];
        } else {
            rb[r] = in[r] ? A.rp[i[r]] : 0;
            re[r] = in[r] ? A.rp[i[r] + 1] : 0;
            cp[r] = A.ci;
            vp[r] = A.val;
            base[r] = 0;
        }
        len = max(len, re[r] - rb[r]);
    }
    if constexpr (R == 1) {
        // one row: the loop bounds are the row's own
#pragma unroll 1
        for (int k0 = rb[0] + klane<L>(); k0 < re[0]; k0 += U * L) {
            int j[U];
            T a[U];
            bool ok[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int k = k0 + q * L;
                ok[q] = k < re[0];
                j[q] = ok[q] ? base[0] + cp[0][k] : 0;
                a[q] = ok[q] ? vp[0][k] : T(0);
            }
#pragma unroll
            for (int q = 0; q < U; q++)
#pragma unroll
                for (int m = 0; m < M; m++) {
                    const T x = vec[m][j[q]];
                    if (ok[q]) out[0][m] += a[q] * x;
                }
        }
    } else {
#pragma unroll 1
    for (int k0 = klane<L>(); k0 < len; k0 += U * L) {
        int j[R][U];
        T a[R][U];
        bool ok[R][U];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int k = rb[r] + k0 + q * L;
                ok[r][q] = k < re[r];
                j[r][q] = ok[r][q] ? base[r] + cp[r][k] : 0;
                a[r][q] = ok[r][q] ? vp[r][k] : T(0);
            }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < U; q++)
#pragma unroll
                for (int m = 0; m < M; m++) {
                    const T x = vec[m][j[r][q]];
                    if (ok[r][q]) out[r][m] += a[r][q] * x;
                }
    }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int m = 0; m < M; m++)
#pragma unroll
            for (int o = L / 2; o > 0; o >>= 1) out[r][m] += __shfl_xor(out[r][m], o, 64);
}
template <typename T, int L, int M>
__device__ inline void row_sums(const KRows<T> &A, int i, bool in, const T *const (&vec)[M], T (&out)[M]) {
    const int ii[1] = {i};
    const bool inn[1] = {in};
    T o[1][M];
    row_sums_r<T, L, M, 1>(A, ii, inn, vec, o);
#pragma unroll
    for (int m = 0; m < M; m++) out[m] = o[0][m];
}

// ------------------------------------------------------------------ init
// r = b - A x; CG: z = dinv r, partial r.z; BiCGStab: rh = r, partial r.r.
// Partials r.r and b.b; clears the flags.  xzero: x is set to zero here and
// r = b (no product; the corrections of a refinement start from zero).
// Workgroups of KNT/R threads, R rows per lane group (slabs, see krow_s).
template <typename T, typename TB, int L, int R = 1>
__global__ void __launch_bounds__(KNT / R) k_kinit(int n, int cg, int xzero, KRows<T> A,
                                                   T *__restrict__ x, const TB *__restrict__ b,
                                                   const T *__restrict__ dinv, T *__restrict__ r, T *__restrict__ z,
                                                   T *__restrict__ rh, T *__restrict__ part, int nb, int *ki, KSeg S) {
    T *ksd = nullptr;
    const KCtx c = kctx<T>(S, nb, ksd, ki);
    int i[R];
    bool in[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
        i[s] = krow_s<L, R>(s);
        in[s] = i[s] < n;
    }
    const bool l0 = klane<L>() == 0;
    T sx[R][1];
#pragma unroll
    for (int s = 0; s < R; s++) sx[s][0] = 0;
    if (!xzero) {  // uniform over the launch
        const T *const vx[1] = {x};
        row_sums_r<T, L, 1, R>(A, i, in, vx, sx);
    } else {
#pragma unroll
        for (int s = 0; s < R; s++)
            if (in[s] && l0) x[i[s]] = T(0);
    }
    T acc[R][3];
#pragma unroll
    for (int s = 0; s < R; s++) {
        acc[s][0] = acc[s][1] = acc[s][2] = 0;
        if (in[s] && l0) {
            const int ii = i[s];
            T bi = (T)b[ii];
            T ri = bi - sx[s][0];
            r[ii] = ri;
            if (cg) {
                T zi = dinv[ii] * ri;
                z[ii] = zi;
                acc[s][0] = ri * zi;
            } else {
                rh[ii] = ri;
                acc[s][0] = ri * ri;
            }
            acc[s][1] = ri * ri;
            acc[s][2] = bi * bi;
        }
    }
    const int sl[3] = {KP_RZ, KP_RR, KP_BB};
    write_partials<T, 3, R>(acc, part, nb, sl);
    if (S.ndone && blockIdx.x == 0 && threadIdx.x == 0) S.ndone[0] = 0;
    if (c.lead && threadIdx.x == 0) {
        ki[KI_DONE] = 0;
        ki[KI_ITERS] = 0;
        ki[KI_CONV] = 0;
        ki[KI_BRK] = 0;
        ki[KI_BRK + 1] = 0;
    }
}

// ------------------------------------------------------------------- PCG
// launch 1: stopping test, beta = rho / rho_old, p = z + beta p_old (own rows
// to pnew), q = A z + beta A p_old; partial p.q
template <typename T, int L, int R = 1>
__global__ void __launch_bounds__(KNT / R) k_kcg_spmv(int n, int it, int maxit, T rtol, T atol,
                                                      KRows<T> A, const T *__restrict__ z,
                                                      const T *__restrict__ pold, T *__restrict__ pnew,
                                                      T *__restrict__ q, T *__restrict__ part, int nb, T *ks, int *ki,
                                                      KSeg S) {
    const KCtx c = kctx<T>(S, nb, ks, ki);
    const int sl[3] = {KP_RZ, KP_RR, KP_BB};
    __shared__ T pv[4];
    reduce_partials<T, 3>(part, nb, c, sl, ki, pv);
    int i[R];
    bool in[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
        i[s] = krow_s<L, R>(s);
        in[s] = i[s] < n;
    }
    const T *const vz[2] = {z, it == 0 ? z : pold};
    T sz[R][2];
    row_sums_r<T, L, 2, R>(A, i, in, vz, sz);
    T zi[R], poi[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
        zi[s] = in[s] ? z[i[s]] : T(0);
        poi[s] = (in[s] && it > 0) ? pold[i[s]] : T(0);
    }
    kblock_sync();
    if (pv[3] != T(0)) return;
    const T rho = pv[0];
    if (krylov_stop<T>(c.lead, it == 0, it, maxit, pv[1], pv[2], rtol, atol, false, ks, ki, S.ndone)) return;
    T beta = 0;
    if (it > 0) {
        T rho_old = ks[KS_RHO + ((it + 1) & 1)];
        beta = rho_old != T(0) ? rho / rho_old : T(0);
    }
    if (c.lead && threadIdx.x == 0) ks[KS_RHO + (it & 1)] = rho;
    T acc[R][1];
#pragma unroll
    for (int s = 0; s < R; s++) {
        acc[s][0] = 0;
        if (in[s] && klane<L>() == 0) {
            const T pi = it == 0 ? zi[s] : zi[s] + beta * poi[s];
            const T qi = it == 0 ? sz[s][0] : sz[s][0] + beta * sz[s][1];
            pnew[i[s]] = pi;
            q[i[s]] = qi;
            acc[s][0] = pi * qi;
        }
    }
    const int so[1] = {KP_PQ};
    write_partials<T, 1, R>(acc, part, nb, so);
}

// launch 2: alpha = rho / p.q; x += alpha p; r -= alpha q; z = dinv r;
// partials r.z, r.r.  The workgroup's KNT/L rows go to its threads in
// order (row j of the workgroup to thread j mod KNT/R, slab j / (KNT/R)).
// jacobi = 0 (AMG-PCG): z is left to the V-cycle and the r.z partial to
// k_krz (the slot is written with zeros here).
template <typename T, int L, int R = 1>
__global__ void __launch_bounds__(KNT / R) k_kcg_update(int n, int it, T *__restrict__ x, T *__restrict__ r,
                                                        const T *__restrict__ p, const T *__restrict__ q,
                                                        const T *__restrict__ dinv, T *__restrict__ z,
                                                        T *__restrict__ part, int nb, T *ks, int *ki, KSeg S,
                                                        int jacobi = 1) {
    constexpr int NT = KNT / R, ROWS = KNT / L;
    const KCtx c = kctx<T>(S, nb, ks, ki);
    const int sl[1] = {KP_PQ};
    __shared__ T pv[2];
    reduce_partials<T, 1>(part, nb, c, sl, ki, pv);
    int i[R];
    bool own[R];
    T xi[R], ri[R], pi[R], qi[R], di[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
        const int j = s * NT + (int)threadIdx.x;
        i[s] = blockIdx.x * ROWS + j;
        own[s] = j < ROWS && i[s] < n;
        xi[s] = ri[s] = pi[s] = qi[s] = di[s] = 0;
        if (own[s]) {
            xi[s] = x[i[s]];
            ri[s] = r[i[s]];
            pi[s] = p[i[s]];
            qi[s] = q[i[s]];
            di[s] = jacobi ? dinv[i[s]] : T(0);
        }
    }
    kblock_sync();
    if (pv[1] != T(0)) return;
    const T rho = ks[KS_RHO + (it & 1)];
    const T alpha = pv[0] != T(0) ? rho / pv[0] : T(0);
    T acc[R][2];
#pragma unroll
    for (int s = 0; s < R; s++) {
        acc[s][0] = acc[s][1] = 0;
        if (own[s]) {
            x[i[s]] = xi[s] + alpha * pi[s];
            const T rn = ri[s] - alpha * qi[s];
            r[i[s]] = rn;
            if (jacobi) {
                T zn = di[s] * rn;
                z[i[s]] = zn;
                acc[s][0] = rn * zn;
            }
            acc[s][1] = rn * rn;
        }
    }
    const int so[2] = {KP_RZ, KP_RR};
    write_partials<T, 2, R>(acc, part, nb, so);
}

// AMG-PCG: partial r.z of the preconditioned residual z = V(r), the rho of
// the next k_kcg_spmv (rows as in k_kcg_update); returns at once when the
// segment's solve has stopped.
template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_krz(int n, const T *__restrict__ r, const T *__restrict__ z,
                                             T *__restrict__ part, int nb, const int *ki, KSeg S) {
    if (S.wseg) ki += (size_t)S.wseg[blockIdx.x] * KI_N;
    if (ki[KI_DONE]) return;
    const int i = blockIdx.x * (KNT / L) + (int)threadIdx.x;
    const bool own = (int)threadIdx.x < KNT / L && i < n;
    T acc[1] = {own ? r[i] * z[i] : T(0)};
    const int so[1] = {KP_RZ};
    write_partials<T, 1>(acc, part, nb, so);
}

// -------------------------------------------------------------- BiCGStab
// launch A of iteration `it`.  it = 0: stopping test on the initial
// residual, rho = rh.r, p = r, v = (A D^-1) r.  it > 0: omega = t.s / t.t of
// iteration it-1, x += alpha dinv p_old + omega dinv s, r = s - omega t,
// rho = rh.s - omega rh.t, beta = (rho / rho_old)(alpha / omega),
// p = r + beta (p_old - omega v_old), v = (A D^-1) p.  Partials rh.v, r.r.
template <typename T, int L, int R = 1>
__global__ void __launch_bounds__(KNT / R) k_kbcg_a(int n, int it, int maxit, T rtol, T atol,
                                                    KRows<T> A, const T *__restrict__ dinv,
                                                    T *__restrict__ x, T *__restrict__ r, const T *__restrict__ rh,
                                                    const T *__restrict__ s, const T *__restrict__ t,
                                                    const T *__restrict__ pold, const T *__restrict__ vold,
                                                    T *__restrict__ pnew, T *__restrict__ vnew, T *__restrict__ part,
                                                    int nb, T *ks, int *ki, KSeg S) {
    const KCtx c = kctx<T>(S, nb, ks, ki);
    int i[R];
    bool in[R], own[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        i[q] = krow_s<L, R>(q);
        in[q] = i[q] < n;
        own[q] = in[q] && klane<L>() == 0;
    }
    T acc[R][2];
#pragma unroll
    for (int q = 0; q < R; q++) acc[q][0] = acc[q][1] = 0;
    if (it == 0) {
        const int sl[3] = {KP_RZ, KP_RR, KP_BB};
        __shared__ T pv[4];
        reduce_partials<T, 3>(part, nb, c, sl, ki, pv);
        const T *const vr[1] = {r};
        T sr[R][1];
        row_sums_r<T, L, 1, R>(A, i, in, vr, sr);
        T ri[R], rhi[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            ri[q] = in[q] ? r[i[q]] : T(0);
            rhi[q] = in[q] ? rh[i[q]] : T(0);
        }
        kblock_sync();
        if (pv[3] != T(0)) return;
        if (krylov_stop<T>(c.lead, true, 0, maxit, pv[1], pv[2], rtol, atol, false, ks, ki, S.ndone)) return;
        if (c.lead && threadIdx.x == 0) ks[KS_RHO] = pv[0];
#pragma unroll
        for (int q = 0; q < R; q++)
            if (own[q]) {
                pnew[i[q]] = ri[q];
                vnew[i[q]] = sr[q][0];
                acc[q][0] = rhi[q] * sr[q][0];
                acc[q][1] = ri[q] * ri[q];
            }
    } else {
        const int sl[4] = {KP_TS, KP_TT, KP_RS, KP_RT};
        __shared__ T pv[5];
        reduce_partials<T, 4>(part, nb, c, sl, ki, pv, KI_BRK + ((it + 1) & 1));
        const T *const vv[4] = {s, t, pold, vold};
        T sv[R][4];
        row_sums_r<T, L, 4, R>(A, i, in, vv, sv);
        T si[R], ti[R], poi[R], voi[R], xi[R], di[R], rhi[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            si[q] = ti[q] = poi[q] = voi[q] = xi[q] = di[q] = rhi[q] = 0;
            if (own[q]) {
                si[q] = s[i[q]];
                ti[q] = t[i[q]];
                poi[q] = pold[i[q]];
                voi[q] = vold[i[q]];
                xi[q] = x[i[q]];
                di[q] = dinv[i[q]];
                rhi[q] = rh[i[q]];
            }
        }
        kblock_sync();
        if (pv[4] != T(0)) return;
        const int o = (it + 1) & 1;  // scalars of iteration it-1
        const T omega = pv[1] > T(0) ? pv[0] / pv[1] : T(0);
        const T alpha = ks[KS_ALPHA + o], rho_old = ks[KS_RHO + o];
        const T rho = pv[2] - omega * pv[3];
        const bool breakdown = rho == T(0) || rho_old == T(0) || omega == T(0);
        const T beta = breakdown ? T(0) : (rho / rho_old) * (alpha / omega);
        T ri[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            ri[q] = 0;
            if (own[q]) {
                x[i[q]] = xi[q] + alpha * (di[q] * poi[q]) + omega * (di[q] * si[q]);
                ri[q] = si[q] - omega * ti[q];
                r[i[q]] = ri[q];
            }
        }
        if (breakdown) {
            // x holds this iteration's update; the host checks the true residual
            if (c.lead && threadIdx.x == 0) {
                ki[KI_ITERS] = it;
                ki[KI_CONV] = 0;
                ki[KI_BRK + (it & 1)] = 1;
            }
            return;
        }
        if (c.lead && threadIdx.x == 0) ks[KS_RHO + (it & 1)] = rho;
#pragma unroll
        for (int q = 0; q < R; q++)
            if (own[q]) {
                const T vi = (sv[q][0] - omega * sv[q][1]) + beta * (sv[q][2] - omega * sv[q][3]);
                pnew[i[q]] = ri[q] + beta * (poi[q] - omega * voi[q]);
                vnew[i[q]] = vi;
                acc[q][0] = rhi[q] * vi;
                acc[q][1] = ri[q] * ri[q];
            }
    }
    // iteration 0 reads the r.r partials this launch would overwrite (with
    // the same values): it writes its r.r to an unused slot instead
    const int so[2] = {KP_RV, it == 0 ? KP_PQ : KP_RR};
    write_partials<T, 2, R>(acc, part, nb, so);
}

// launch B: stopping test on r.r; alpha = rho / rh.v; s = r - alpha v;
// t = (A D^-1) s; partials t.s, t.t, rh.s, rh.t
template <typename T, int L, int R = 1>
__global__ void __launch_bounds__(KNT / R) k_kbcg_b(int n, int it, int maxit, KRows<T> A,
                                                    const T *__restrict__ r, const T *__restrict__ rh,
                                                    const T *__restrict__ v, T *__restrict__ s, T *__restrict__ t,
                                                    T *__restrict__ part, int nb, T *ks, int *ki, KSeg S) {
    const KCtx c = kctx<T>(S, nb, ks, ki);
    const int sl[2] = {KP_RV, KP_RR};
    __shared__ T pv[3];
    reduce_partials<T, 2>(part, nb, c, sl, ki, pv, KI_BRK + (it & 1));
    int i[R];
    bool in[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        i[q] = krow_s<L, R>(q);
        in[q] = i[q] < n;
    }
    const T *const vv[2] = {r, v};
    T sv[R][2];
    row_sums_r<T, L, 2, R>(A, i, in, vv, sv);
    T ri[R], vi[R], rhi[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        ri[q] = in[q] ? r[i[q]] : T(0);
        vi[q] = in[q] ? v[i[q]] : T(0);
        rhi[q] = in[q] ? rh[i[q]] : T(0);
    }
    kblock_sync();
    if (pv[2] != T(0)) return;
    if (it > 0 && krylov_stop<T>(c.lead, false, it, maxit, pv[1], T(0), T(0), T(0), false, ks, ki, S.ndone)) return;
    const T rho = ks[KS_RHO + (it & 1)];
    const T alpha = pv[0] != T(0) ? rho / pv[0] : T(0);
    if (c.lead && threadIdx.x == 0) ks[KS_ALPHA + (it & 1)] = alpha;
    T acc[R][4];
#pragma unroll
    for (int q = 0; q < R; q++) {
        acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = 0;
        if (in[q] && klane<L>() == 0) {
            const T si = ri[q] - alpha * vi[q];
            const T ti = sv[q][0] - alpha * sv[q][1];
            s[i[q]] = si;
            t[i[q]] = ti;
            acc[q][0] = ti * si;
            acc[q][1] = ti * ti;
            acc[q][2] = rhi[q] * si;
            acc[q][3] = rhi[q] * ti;
        }
    }
    const int so[4] = {KP_TS, KP_TT, KP_RS, KP_RT};
    write_partials<T, 4, R>(acc, part, nb, so);
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

}  // namespace hfh
