// Smoothed-aggregation algebraic multigrid for the pressure systems:
// host set-up (aggregation, prolongator smoothing, Galerkin products, dense
// coarsest inverse) and the device V-cycle used as the right preconditioner
// of BiCGStab.
//
// Set-up (once per matrix):
//   strength   |a_ij| >= theta sqrt(|a_ii a_jj|) on (A + A^T)/2, theta 0.08
//   aggregates greedy: a node whose strong neighbours are all free seeds an
//              aggregate with them; left-over nodes join a neighbouring
//              aggregate (or form their own)
//   P          (I - 4/(3 lambda) D^-1 A) P_tent, P_tent piecewise constant,
//              lambda the spectral radius of D^-1 A by 15 power steps
//   R = P^T, A_c = R A P, until at most `coarse_max` unknowns remain; the
//   coarsest level is inverted densely (Gauss-Jordan, partial pivoting).
// V(1,1) cycle with damped Jacobi (omega 0.8):
//   x = omega D^-1 b; r = b - A x; b_c = R r; x_c = V(b_c); x += P x_c;
//   x += omega D^-1 (b - A x)
// Each kernel launch handles one such step for one level, except the last
// two, which are one launch through A P (k_amg_prolong_post); rows run on L
// lanes as in the Krylov kernels.  Every launch first reads the solve's done
// flag and returns when a previous launch has set it.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "hfh_krylov.hpp"
#include "hfh_progress.hpp"

namespace hfh {

// ------------------------------------------------------------- host CSR
struct HCsr {
    int nrow = 0, ncol = 0;
    std::vector<int> rp{0}, ci;
    std::vector<double> v;
};

// C = A B (row-wise with a dense accumulator over B's columns)
inline HCsr csr_mul(const HCsr &A, const HCsr &B) {
    HCsr C;
    C.nrow = A.nrow;
    C.ncol = B.ncol;
    C.rp.assign(1, 0);
    std::vector<double> acc(B.ncol, 0.0);
    std::vector<int> mark(B.ncol, -1), cols;
    for (int i = 0; i < A.nrow; i++) {
        cols.clear();
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int j = A.ci[k];
            const double a = A.v[k];
            for (int q = B.rp[j]; q < B.rp[j + 1]; q++) {
                const int c = B.ci[q];
                if (mark[c] != i) {
                    mark[c] = i;
                    acc[c] = 0.0;
                    cols.push_back(c);
                }
                acc[c] += a * B.v[q];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int c : cols) {
            C.ci.push_back(c);
            C.v.push_back(acc[c]);
        }
        C.rp.push_back((int)C.ci.size());
    }
    return C;
}

inline HCsr csr_transpose(const HCsr &A) {
    HCsr T;
    T.nrow = A.ncol;
    T.ncol = A.nrow;
    T.rp.assign(A.ncol + 1, 0);
    for (int c : A.ci) T.rp[c + 1]++;
    for (int i = 0; i < A.ncol; i++) T.rp[i + 1] += T.rp[i];
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    std::vector<int> pos(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < A.nrow; i++)
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int p = pos[A.ci[k]]++;
            T.ci[p] = i;
            T.v[p] = A.v[k];
        }
    return T;
}

inline std::vector<double> csr_diag_inv(const HCsr &A) {
    std::vector<double> d(A.nrow, 1.0);
    for (int i = 0; i < A.nrow; i++)
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
            if (A.ci[k] == i && A.v[k] != 0.0) d[i] = 1.0 / A.v[k];
    return d;
}

// greedy aggregation on the symmetrised strength graph; returns the
// aggregate of each row and the aggregate count
inline int sa_aggregate(const HCsr &A, double theta, std::vector<int> &agg) {
    const int n = A.nrow;
    std::vector<double> dg(n, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++)
            if (A.ci[k] == i) dg[i] = std::sqrt(std::fabs(A.v[k]));
    // strong neighbours: |(a_ij + a_ji)/2| >= theta sqrt(|a_ii a_jj|)
    HCsr At = csr_transpose(A);
    std::vector<std::vector<int>> S(n);
    std::vector<double> acc(n, 0.0);
    std::vector<int> mark(n, -1), cols;
    for (int i = 0; i < n; i++) {
        cols.clear();
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int j = A.ci[k];
            if (mark[j] != i) {
                mark[j] = i;
                acc[j] = 0.0;
                cols.push_back(j);
            }
            acc[j] += 0.5 * A.v[k];
        }
        for (int k = At.rp[i]; k < At.rp[i + 1]; k++) {
            const int j = At.ci[k];
            if (mark[j] != i) {
                mark[j] = i;
                acc[j] = 0.0;
                cols.push_back(j);
            }
            acc[j] += 0.5 * At.v[k];
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols)
            if (j != i && std::fabs(acc[j]) >= theta * dg[i] * dg[j]) S[i].push_back(j);
    }
    // rows without off-diagonal entries (the padding between the blocks of
    // a block-diagonal system, a pinned unknown's identity row) join no
    // aggregate: the smoother alone handles them.  As singleton aggregates
    // their smoothed prolongation would shrink them by (1 - w)^2 per level
    // until the coarsest matrix is numerically singular.
    agg.assign(n, -1);
    for (int i = 0; i < n; i++) {
        bool iso = true;
        for (int k = A.rp[i]; k < A.rp[i + 1] && iso; k++) iso = A.ci[k] == i;
        if (iso) agg[i] = -2;
    }
    // rows with no strong neighbour (a diagonally dominant row, as in the
    // Helmholtz rows of a high-Re, coarse-mesh simulation) are left to the
    // smoother too: as singleton aggregates they would keep every such row
    // on every level and widen its stencil each time
    for (int i = 0; i < n; i++)
        if (agg[i] == -1 && S[i].empty()) agg[i] = -2;
    int na = 0;
    for (int i = 0; i < n; i++) {
        if (agg[i] != -1) continue;
        bool free = true;
        for (int j : S[i])
            if (agg[j] >= 0) {
                free = false;
                break;
            }
        if (!free) continue;
        agg[i] = na;
        for (int j : S[i]) agg[j] = na;
        na++;
    }
    for (int i = 0; i < n; i++) {
        if (agg[i] != -1) continue;
        int a = -1;
        for (int j : S[i])
            if (agg[j] >= 0) {
                a = agg[j];
                break;
            }
        agg[i] = a >= 0 ? a : na++;
    }
    for (int i = 0; i < n; i++)
        if (agg[i] == -2) agg[i] = -1;
    return na;
}

struct AmgHostLevel {
    HCsr A, P, R, AP;  // AP = A P (the fused prolongation + post-smoothing)
    std::vector<double> dinv;
};

// Builds the hierarchy; returns false (no hierarchy) when the coarsening
// stalls or the coarsest matrix cannot be inverted.
// max_cx > 0: give up as soon as the operator complexity (sum over levels of
// the matrix entries / the fine matrix's) exceeds it.
inline bool sa_build(const HCsr &A0, std::vector<AmgHostLevel> &lv, std::vector<double> &cinv, int &nc,
                     int coarse_max = 400, double theta = 0.08, int max_levels = 10, double max_cx = 0) {
    lv.clear();
    HCsr A = A0;
    double nz = 0;
    while (A.nrow > coarse_max && (int)lv.size() < max_levels) {
        progress("AMG hierarchy levels", (long)lv.size(), (long)max_levels);
        nz += (double)A.v.size();
        if (max_cx > 0 && nz > max_cx * (double)std::max<size_t>(A0.v.size(), 1)) return false;
        AmgHostLevel L;
        L.A = A;
        L.dinv = csr_diag_inv(A);
        std::vector<int> agg;
        const int na = sa_aggregate(A, theta, agg);
        if (na >= A.nrow || na == 0) return false;  // no coarsening
        // spectral radius of D^-1 A (power iteration from a fixed start)
        std::vector<double> x(A.nrow), y(A.nrow);
        for (int i = 0; i < A.nrow; i++) x[i] = 1.0 + 0.5 * std::sin(0.7 * i);
        double lam = 1.0;
        for (int it = 0; it < 15; it++) {
            double nn = 0;
            for (int i = 0; i < A.nrow; i++) {
                double s = 0;
                for (int k = A.rp[i]; k < A.rp[i + 1]; k++) s += A.v[k] * x[A.ci[k]];
                y[i] = L.dinv[i] * s;
                nn += y[i] * y[i];
            }
            lam = std::sqrt(nn);
            if (!(lam > 0)) return false;
            for (int i = 0; i < A.nrow; i++) x[i] = y[i] / lam;
        }
        // P = (I - w D^-1 A) P_tent, w = 4 / (3 lambda)
        const double w = 4.0 / (3.0 * lam);
        HCsr Pt;
        Pt.nrow = A.nrow;
        Pt.ncol = na;
        Pt.rp.assign(A.nrow + 1, 0);
        Pt.ci.clear();
        Pt.v.clear();
        for (int i = 0; i < A.nrow; i++) {
            if (agg[i] >= 0) {
                Pt.ci.push_back(agg[i]);
                Pt.v.push_back(1.0);
            }
            Pt.rp[i + 1] = (int)Pt.ci.size();
        }
        HCsr S = A;  // S = I - w D^-1 A
        for (int i = 0; i < A.nrow; i++)
            for (int k = A.rp[i]; k < A.rp[i + 1]; k++) S.v[k] = -w * L.dinv[i] * A.v[k] + (A.ci[k] == i ? 1.0 : 0.0);
        L.P = csr_mul(S, Pt);
        L.R = csr_transpose(L.P);
        L.AP = csr_mul(A, L.P);
        A = csr_mul(L.R, L.AP);
        lv.push_back(std::move(L));
    }
    if (lv.empty() || A.nrow > 2000) return false;  // coarsening stalled
    // dense inverse of the coarsest matrix
    nc = A.nrow;
    std::vector<double> M((size_t)nc * nc, 0.0);
    for (int i = 0; i < nc; i++)
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) M[(size_t)i * nc + A.ci[k]] += A.v[k];
    cinv.assign((size_t)nc * nc, 0.0);
    for (int i = 0; i < nc; i++) cinv[(size_t)i * nc + i] = 1.0;
    double amax = 0;
    for (double a : M) amax = std::max(amax, std::fabs(a));
    for (int c = 0; c < nc; c++) {
        int pr = c;
        for (int r = c + 1; r < nc; r++)
            if (std::fabs(M[(size_t)r * nc + c]) > std::fabs(M[(size_t)pr * nc + c])) pr = r;
        if (!(std::fabs(M[(size_t)pr * nc + c]) > 1e-14 * amax)) return false;
        if (pr != c)
            for (int j = 0; j < nc; j++) {
                std::swap(M[(size_t)pr * nc + j], M[(size_t)c * nc + j]);
                std::swap(cinv[(size_t)pr * nc + j], cinv[(size_t)c * nc + j]);
            }
        const double d = 1.0 / M[(size_t)c * nc + c];
        for (int j = 0; j < nc; j++) {
            M[(size_t)c * nc + j] *= d;
            cinv[(size_t)c * nc + j] *= d;
        }
        for (int r = 0; r < nc; r++) {
            if (r == c) continue;
            const double f = M[(size_t)r * nc + c];
            if (f == 0.0) continue;
            for (int j = 0; j < nc; j++) {
                M[(size_t)r * nc + j] -= f * M[(size_t)c * nc + j];
                cinv[(size_t)r * nc + j] -= f * cinv[(size_t)c * nc + j];
            }
        }
    }
    for (double a : cinv)
        if (!std::isfinite(a)) return false;
    return true;
}

// ------------------------------------------------------------ V-cycle kernels
// Lanes per row AL (1, 2 or 4) of the V-cycle kernels: chosen per level
// and operator from its mean row length (amg_lanes).
constexpr int ANT = 256;  // threads per workgroup
constexpr long AMG_LANES_MIN_ROWS = 200000;
inline int amg_lanes(long nnz, long rows) {
    // short rows take fewer lanes on large levels; a level of fewer rows is
    // bound by its launch latency, and more threads per row hide it better
    // (47,800-row contraction: the pressure solve 1.82 ms with 1-2 lanes
    // against 1.69 with 4).  Long rows (3D least-squares stencils, the
    // Galerkin products of 3D aggregation) take 8 to 64 lanes, about 3 to 6
    // entries per lane.
    const double avg = rows > 0 ? (double)nnz / (double)rows : 0.0;
    if (avg > 192.0) return 64;
    if (avg > 96.0) return 32;
    if (avg > 48.0) return 16;
    if (avg > 24.0) return 8;
    if (rows < AMG_LANES_MIN_ROWS) return 4;
    return avg <= 3.0 ? 1 : avg <= 6.0 ? 2 : 4;
}

// `nd` is the solve's {stopped segments, segments} pair (KSeg::ndone): the
// cycle's launches return at once when every segment has stopped
__device__ inline bool amg_done(const int *nd) { return nd && nd[0] >= nd[1]; }

template <typename T, int AL>
__device__ inline T arow(int rb, int re, const int *__restrict__ ci, const T *__restrict__ v, const T *__restrict__ x) {
    T s = 0;
    for (int k = rb + (int)(threadIdx.x % AL); k < re; k += AL) s += v[k] * x[ci[k]];
#pragma unroll
    for (int o = AL / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

// r = b - A (omega D^-1 b).  The fine level's b may be of another type TB
// (an FP64 Krylov solve preconditioned by the FP32 cycle); the cycle
// computes in T.
template <typename T, typename TB, int AL>
__global__ void __launch_bounds__(ANT) k_amg_resid(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                   const T *__restrict__ v, const T *__restrict__ dinv, T omega,
                                                   const TB *__restrict__ b, T *__restrict__ r, const int *ki) {
    if (amg_done(ki)) return;
    const int i = (blockIdx.x * ANT + threadIdx.x) / AL;
    const bool in = i < n;
    const int rb = in ? rp[i] : 0, re = in ? rp[i + 1] : 0;
    T s = 0;
    for (int k = rb + (int)(threadIdx.x % AL); k < re; k += AL) {
        const int j = ci[k];
        s += v[k] * (omega * (dinv[j] * (T)b[j]));
    }
#pragma unroll
    for (int o = AL / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (in && threadIdx.x % AL == 0) r[i] = (T)b[i] - s;
}

// The fine level's residual launch when the matrix has the diagonal form
// of the Krylov kernels (KRows::dblk): one row per thread in 1024-row
// workgroups, the gathers of omega D^-1 b at the block's diagonals issued
// before the row's template id arrives; rows without a template read CSR.
template <typename T, typename TB = T>
__global__ void __launch_bounds__(KNT) k_amg_resid_d(int n, KRows<T> A, const T *__restrict__ dinv, T omega,
                                                     const TB *__restrict__ b, T *__restrict__ r, const int *ki) {
    if (amg_done(ki)) return;
    const int i = blockIdx.x * KNT + (int)threadIdx.x;
    const bool in = i < n;
    const int *db = A.dblk + (size_t)blockIdx.x * (KD_MAX + 1);
    const int nd = db[KD_MAX];
    T xv[KD_MAX];
#pragma unroll
    for (int e = 0; e < KD_MAX; e++) {
        int j = i + db[e];
        j = j < 0 ? 0 : (j >= n ? n - 1 : j);
        xv[e] = (in && e < nd) ? omega * (dinv[j] * (T)b[j]) : T(0);
    }
    const int t = in ? A.code[i] : -1;
    T s = 0;
    if (t >= 0) {
        const unsigned msk = A.dmask[t];
        const T *w = A.dw + (size_t)t * KD_MAX;
#pragma unroll
        for (int e = 0; e < KD_MAX; e++) {
            const T a = w[e];
            if ((msk >> e) & 1u) s += a * xv[e];
        }
    } else if (in) {
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int j = A.ci[k];
            s += A.val[k] * (omega * (dinv[j] * (T)b[j]));
        }
    }
    if (in) r[i] = (T)b[i] - s;
}

// b_c = R r
template <typename T, int AL>
__global__ void __launch_bounds__(ANT) k_amg_restrict(int nc, const int *__restrict__ rp, const int *__restrict__ ci,
                                                      const T *__restrict__ v, const T *__restrict__ r,
                                                      T *__restrict__ bc, const int *ki) {
    if (amg_done(ki)) return;
    const int i = (blockIdx.x * ANT + threadIdx.x) / AL;
    const bool in = i < nc;
    const T s = arow<T, AL>(in ? rp[i] : 0, in ? rp[i + 1] : 0, ci, v, r);
    if (in && threadIdx.x % AL == 0) bc[i] = s;
}

// coarsest level, one workgroup: b_c = R r, x_c = C^-1 b_c (a wave per
// row: lanes stride the row, then a butterfly)
template <typename T>
__global__ void __launch_bounds__(1024) k_amg_coarse(int nc, const int *__restrict__ rp, const int *__restrict__ ci,
                                                     const T *__restrict__ v, const T *__restrict__ r,
                                                     const T *__restrict__ cinv, T *__restrict__ xc, const int *ki) {
    if (amg_done(ki)) return;
    extern __shared__ unsigned char amg_sh[];
    T *bs = reinterpret_cast<T *>(amg_sh);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int i = w; i < nc; i += nw) {
        T s = 0;
        for (int k = rp[i] + lane; k < rp[i + 1]; k += 64) s += v[k] * r[ci[k]];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) bs[i] = s;
    }
    __syncthreads();
    for (int i = w; i < nc; i += nw) {
        const T *row = cinv + (size_t)i * nc;
        T s = 0;
        for (int j = lane; j < nc; j += 64) s += row[j] * bs[j];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) xc[i] = s;
    }
}

// coarsest level, many workgroups: x_c = C^-1 b_c with b_c already
// restricted (a wave per row)
template <typename T, typename TC>
__global__ void __launch_bounds__(ANT) k_amg_dense(int nc, const TC *__restrict__ cinv, const T *__restrict__ bc,
                                                   T *__restrict__ xc, const int *ki) {
    if (amg_done(ki)) return;
    const int lane = threadIdx.x & 63;
    const int i = (blockIdx.x * ANT + threadIdx.x) >> 6;
    if (i >= nc) return;
    const TC *row = cinv + (size_t)i * nc;
    T s = 0;
    for (int j = lane; j < nc; j += 64) s += (T)row[j] * bc[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) xc[i] = s;
}

// x = omega D^-1 b + P x_c
template <typename T, typename TB, int AL>
__global__ void __launch_bounds__(ANT) k_amg_prolong(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                     const T *__restrict__ v, const T *__restrict__ dinv, T omega,
                                                     const TB *__restrict__ b, const T *__restrict__ xc,
                                                     T *__restrict__ x, const int *ki) {
    if (amg_done(ki)) return;
    const int i = (blockIdx.x * ANT + threadIdx.x) / AL;
    const bool in = i < n;
    const T s = arow<T, AL>(in ? rp[i] : 0, in ? rp[i + 1] : 0, ci, v, xc);
    if (in && threadIdx.x % AL == 0) x[i] = omega * (dinv[i] * (T)b[i]) + s;
}

// y = x + omega D^-1 (b - A x)
template <typename T, typename TB, typename TY, int AL>
__global__ void __launch_bounds__(ANT) k_amg_post(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const T *__restrict__ v, const T *__restrict__ dinv, T omega,
                                                  const TB *__restrict__ b, const T *__restrict__ x,
                                                  TY *__restrict__ y, const int *ki) {
    if (amg_done(ki)) return;
    const int i = (blockIdx.x * ANT + threadIdx.x) / AL;
    const bool in = i < n;
    const T s = arow<T, AL>(in ? rp[i] : 0, in ? rp[i + 1] : 0, ci, v, x);
    if (in && threadIdx.x % AL == 0) y[i] = (TY)(x[i] + omega * (dinv[i] * ((T)b[i] - s)));
}

// y = omega D^-1 b + P x_c + omega D^-1 (r - (A P) x_c): prolongation and
// post-smoothing of one level in one launch.  r = b - A (omega D^-1 b) from
// the level's residual launch, so for x = omega D^-1 b + P x_c the smoother's
// b - A x is r - (A P) x_c, with A P stored at set-up.
template <typename T, typename TB, typename TY, int AL>
__global__ void __launch_bounds__(ANT) k_amg_prolong_post(int n, const int *__restrict__ prp,
                                                          const int *__restrict__ pci, const T *__restrict__ pv,
                                                          const int *__restrict__ qrp, const int *__restrict__ qci,
                                                          const T *__restrict__ qv, const T *__restrict__ dinv,
                                                          T omega, const TB *__restrict__ b, const T *__restrict__ r,
                                                          const T *__restrict__ xc, TY *__restrict__ y,
                                                          const int *ki) {
    if (amg_done(ki)) return;
    const int i = (blockIdx.x * ANT + threadIdx.x) / AL;
    const bool in = i < n;
    const T s1 = arow<T, AL>(in ? prp[i] : 0, in ? prp[i + 1] : 0, pci, pv, xc);
    const T s2 = arow<T, AL>(in ? qrp[i] : 0, in ? qrp[i + 1] : 0, qci, qv, xc);
    if (in && threadIdx.x % AL == 0) {
        const T di = dinv[i];
        y[i] = (TY)((omega * (di * (T)b[i]) + s1) + omega * (di * (r[i] - s2)));
    }
}

// The hierarchy on the device.  apply(b, y): y = V(b); b and y have the
// fine size, y must differ from b.
template <typename T>
class AmgPrecond {
  public:
    bool ready = false;
    int nlev = 0, ncoarse = 0;
    double complexity = 0;
    // the fine matrix as the Krylov kernels read it, when it has the
    // diagonal form (set by the owner; the fine residual launch then uses it)
    KRows<T> A0;
    T omega = T(0.8);
    // prolongation and post-smoothing fused through A P (HFH_AMG_FUSE=0:
    // two launches per level, as x += P x_c, x += omega D^-1 (b - A x))
    bool fuse_pp = true;

    // max_cx > 0: give up (no hierarchy) when the operator complexity
    // sum_l nnz(A_l) / nnz(A_0) exceeds it
    bool build(int n, const std::vector<int> &rp, const std::vector<int> &ci, const std::vector<double> &v,
               hipStream_t st, double max_cx = 0) {
        HCsr A;
        A.nrow = A.ncol = n;
        A.rp = rp;
        A.ci = ci;
        A.v = v;
        std::vector<AmgHostLevel> hl;
        std::vector<double> cinv;
        ready = false;
        // stop coarsening at <= 1000 unknowns: one dense solve of up to
        // 1000 rows (two launches) is cheaper than further levels (four
        // launches each)
        // HFH_AMG_THETA / HFH_AMG_OMEGA: strength threshold and Jacobi
        // damping for comparison runs (defaults 0.08 and 0.8)
        double theta = 0.08;
        if (const char *e = std::getenv("HFH_AMG_THETA")) theta = std::atof(e);
        if (const char *e = std::getenv("HFH_AMG_OMEGA")) omega = (T)std::atof(e);
        if (!sa_build(A, hl, cinv, ncoarse, 1000, theta, 10, max_cx)) {
            if (const char *ev = std::getenv("HFH_VERBOSE"))
                if (std::strcmp(ev, "1") == 0)
                    std::fprintf(stderr, "hfh: AMG over %d rows: no hierarchy (coarsening stalled or operator "
                                 "complexity above %.1f)\n", n, max_cx);
            return false;
        }
        {
            double nz = 0;
            for (const AmgHostLevel &L : hl) nz += (double)L.A.v.size();
            complexity = hl.empty() || hl[0].A.v.empty() ? 0.0 : nz / (double)hl[0].A.v.size();
            if (const char *ev = std::getenv("HFH_VERBOSE"))
                if (std::strcmp(ev, "1") == 0)
                    std::fprintf(stderr, "hfh: AMG over %d rows: %zu levels, coarsest %d, operator complexity %.2f%s\n",
                                 n, hl.size(), ncoarse, complexity,
                                 max_cx > 0 && complexity > max_cx ? " (not used)" : "");
            if (max_cx > 0 && complexity > max_cx) return false;
        }
        {
            const char *e = std::getenv("HFH_AMG_FUSE");
            fuse_pp = !(e && std::strcmp(e, "0") == 0);
        }
        nlev = (int)hl.size();
        lv_.clear();
        lv_.resize(nlev);
        for (int l = 0; l < nlev; l++) {
            Dev &d = lv_[l];
            d.n = hl[l].A.nrow;
            d.nc = hl[l].P.ncol;
            upi(d.Arp, hl[l].A.rp, st);
            upi(d.Aci, hl[l].A.ci, st);
            upt(d.Av, hl[l].A.v, st);
            upt(d.dinv, hl[l].dinv, st);
            upi(d.Prp, hl[l].P.rp, st);
            upi(d.Pci, hl[l].P.ci, st);
            upt(d.Pv, hl[l].P.v, st);
            upi(d.Rrp, hl[l].R.rp, st);
            upi(d.Rci, hl[l].R.ci, st);
            upt(d.Rv, hl[l].R.v, st);
            if (fuse_pp) {
                upi(d.Qrp, hl[l].AP.rp, st);
                upi(d.Qci, hl[l].AP.ci, st);
                upt(d.Qv, hl[l].AP.v, st);
            }
            d.la = amg_lanes((long)hl[l].A.v.size(), d.n);
            d.lr = amg_lanes((long)hl[l].R.v.size(), d.nc);
            d.lp = amg_lanes((long)hl[l].P.v.size(), d.n);
            d.lq = amg_lanes((long)std::max(hl[l].P.v.size(), hl[l].AP.v.size()), d.n);
            if (const char *e = std::getenv("HFH_AMG_LANES")) {
                const int f = std::atoi(e);
                if (f == 1 || f == 2 || f == 4 || f == 8 || f == 16 || f == 32 || f == 64)
                    d.la = d.lr = d.lp = d.lq = f;
            }
            if (const char *ev = std::getenv("HFH_VERBOSE"))
                if (std::strcmp(ev, "1") == 0)
                    std::fprintf(stderr, "hfh: AMG level %d: %d rows, %.1f entries per row; lanes A %d R %d P %d AP %d\n",
                                 l, d.n, d.n ? (double)hl[l].A.v.size() / d.n : 0.0, d.la, d.lr, d.lp, d.lq);
            d.r.alloc(d.n);
            d.x.alloc(d.n);
            d.bnext.alloc(d.nc);
            d.xnext.alloc(d.nc);
        }
        upt(cinv_, cinv, st);
        coarse_r_nnz_ = hl.empty() ? 0 : (long)hl.back().R.v.size();
        {
            std::vector<float> c32(cinv.begin(), cinv.end());
            cinv32_.alloc(c32.size());
            if (!c32.empty() &&
                (hipMemcpyAsync(cinv32_.p, c32.data(), c32.size() * sizeof(float), hipMemcpyHostToDevice, st) !=
                     hipSuccess ||
                 hipStreamSynchronize(st) != hipSuccess))
                throw std::runtime_error("amg upload failed");
        }
        if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("amg upload failed");
        ready = true;
        return true;
    }

    // y = V(b) on the fine level; `ki` is the solve's {stopped segments,
    // segments} pair (KSeg::ndone) or null.  b and y
    // may be of another precision than the cycle (TB, TY): the fine
    // level's first and last launches convert.
    template <typename TB, typename TY>
    void apply(const TB *b, TY *y, const int *ki, hipStream_t st) {
        cycle<TB, TY>(0, b, y, ki, st);
    }

    // Times `reps` launches of the fine level's prolongation and
    // post-smoothing (k_amg_prolong_post) on the level's current vectors,
    // writing a buffer of its own; returns the average ms (negative without
    // a fused hierarchy) and the algorithmic bytes of one launch: the P and
    // A P rows as stored (two row pointers each), b, r, dinv read and y
    // written once per row, the coarse correction once.
    float probe_pp(int reps, long *bytes, hipStream_t st) {
        if (!ready || nlev == 0 || !fuse_pp) return -1.0f;
        Dev &d = lv_[0];
        probe_y_.alloc(d.n);
        k_fill<T><<<nblocks(d.n), BLOCK, 0, st>>>(d.n, d.x.p, T(0));
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            throw std::runtime_error("hipEventCreate failed");
        (void)hipEventRecord(a, st);
        with_lanes(d.lq, [&](auto C) {
            constexpr int AL = decltype(C)::value;
            for (int q = 0; q < reps; q++)
                k_amg_prolong_post<T, T, T, AL><<<grid(d.n, AL), ANT, 0, st>>>(
                    d.n, d.Prp.p, d.Pci.p, d.Pv.p, d.Qrp.p, d.Qci.p, d.Qv.p, d.dinv.p, omega, d.x.p, d.r.p,
                    d.xnext.p, probe_y_.p, nullptr);
        });
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *bytes = (long)(d.Pv.n + d.Qv.n) * (4 + (long)sizeof(T)) + 2L * (d.n + 1) * 4 + (long)d.n * 4 * sizeof(T) +
                 (long)d.nc * sizeof(T);
        return reps > 0 ? ms / reps : 0.0f;
    }
    int fine_lanes_pp() const { return lv_.empty() ? 0 : lv_[0].lq; }
    // the same for the fine level's first launch, r = b - A (omega D^-1 b)
    // on CSR rows (-1 when the fine level runs in the diagonal form)
    float probe_resid(int reps, long *bytes, hipStream_t st) {
        if (!ready || nlev == 0 || A0.dw) return -1.0f;
        Dev &d = lv_[0];
        probe_y_.alloc(d.n);
        k_fill<T><<<nblocks(d.n), BLOCK, 0, st>>>(d.n, probe_y_.p, T(1));
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            throw std::runtime_error("hipEventCreate failed");
        (void)hipEventRecord(a, st);
        with_lanes(d.la, [&](auto C) {
            constexpr int AL = decltype(C)::value;
            for (int q = 0; q < reps; q++)
                k_amg_resid<T, T, AL><<<grid(d.n, AL), ANT, 0, st>>>(d.n, d.Arp.p, d.Aci.p, d.Av.p, d.dinv.p, omega,
                                                                     probe_y_.p, d.r.p, nullptr);
        });
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        // A's rows as stored with their row pointers; b and D^-1 read and r
        // written once per row
        *bytes = (long)d.Av.n * (4 + (long)sizeof(T)) + (long)(d.n + 1) * 4 + 3L * d.n * (long)sizeof(T);
        return reps > 0 ? ms / reps : 0.0f;
    }

    long bytes_per_cycle() const {
        long by = 0;
        for (const Dev &d : lv_) {
            const int apass = fuse_pp ? 2 : 3;  // A rows read per level (the fused launch reads A P instead)
            by += apass * (d.Arp.n * 4 + d.Av.n * (4 + sizeof(T))) + (d.Prp.n + d.Rrp.n) * 4;
            if (fuse_pp) by += d.Qrp.n * 4 + d.Qv.n * (4 + sizeof(T));
        }
        return by;
    }

  private:
    struct Dev {
        int n = 0, nc = 0;
        int la = 4, lr = 4, lp = 4, lq = 4;  // lanes per row: A, R, P, and P with A P (amg_lanes)
        KBuf<int> Arp, Aci, Prp, Pci, Rrp, Rci, Qrp, Qci;  // Q = A P
        KBuf<T> Av, dinv, Pv, Rv, Qv, r, x, bnext, xnext;
    };
    std::vector<Dev> lv_;
    KBuf<T> cinv_;
    KBuf<float> cinv32_;  // the dense coarsest inverse as read by k_amg_dense
    long coarse_r_nnz_ = 0;  // entries of the last restriction
    KBuf<T> probe_y_;

    static void upi(KBuf<int> &d, const std::vector<int> &h, hipStream_t st) {
        d.alloc(h.size());
        if (!h.empty() && hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
            throw std::runtime_error("hipMemcpyAsync failed");
    }
    static void upt(KBuf<T> &d, const std::vector<double> &h, hipStream_t st) {
        std::vector<T> t(h.begin(), h.end());
        d.alloc(t.size());
        if (!t.empty() && hipMemcpyAsync(d.p, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess)
            throw std::runtime_error("hipMemcpyAsync failed");
        if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("hipStreamSynchronize failed");
    }
    static int grid(int rows, int al) { return std::max(1, (rows * al + ANT - 1) / ANT); }

    // y = V(b) at level l (y, b of that level's size)
    template <typename TB, typename TY>
    void cycle(int l, const TB *b, TY *y, const int *ki, hipStream_t st) {
        Dev &d = lv_[l];
        if (l == 0 && A0.dw)
            k_amg_resid_d<T, TB><<<(d.n + KNT - 1) / KNT, KNT, 0, st>>>(d.n, A0, d.dinv.p, omega, b, d.r.p, ki);
        else
            with_lanes(d.la, [&](auto C) {
                constexpr int AL = decltype(C)::value;
                k_amg_resid<T, TB, AL><<<grid(d.n, AL), ANT, 0, st>>>(d.n, d.Arp.p, d.Aci.p, d.Av.p, d.dinv.p, omega,
                                                                      b, d.r.p, ki);
            });
        if (l + 1 < nlev) {
            restrict_(d, ki, st);
            cycle<T, T>(l + 1, d.bnext.p, d.xnext.p, ki, st);
        } else if (ncoarse <= 256 && coarse_r_nnz_ <= 32768) {
            k_amg_coarse<T><<<1, 1024, ncoarse * sizeof(T), st>>>(ncoarse, d.Rrp.p, d.Rci.p, d.Rv.p, d.r.p, cinv_.p,
                                                                   d.xnext.p, ki);
        } else if (ncoarse <= 256) {
            // a small coarsest level under a long last restriction (3D
            // aggregation): restrict on many workgroups, then the same dense
            // solve as above, a wave per row
            restrict_(d, ki, st);
            k_amg_dense<T, T><<<std::max(1, (ncoarse * 64 + ANT - 1) / ANT), ANT, 0, st>>>(
                ncoarse, cinv_.p, d.bnext.p, d.xnext.p, ki);
        } else {
            // larger coarsest level: restrict, then a wave per dense row
            restrict_(d, ki, st);
            // the dense inverse is read in FP32 (half the bytes of the
            // coarsest solve; it is part of a preconditioner, and the Krylov
            // iteration around it keeps its precision)
            k_amg_dense<T, float><<<std::max(1, (ncoarse * 64 + ANT - 1) / ANT), ANT, 0, st>>>(
                ncoarse, cinv32_.p, d.bnext.p, d.xnext.p, ki);
        }
        if (fuse_pp) {
            with_lanes(d.lq, [&](auto C) {
                constexpr int AL = decltype(C)::value;
                k_amg_prolong_post<T, TB, TY, AL><<<grid(d.n, AL), ANT, 0, st>>>(
                    d.n, d.Prp.p, d.Pci.p, d.Pv.p, d.Qrp.p, d.Qci.p, d.Qv.p, d.dinv.p, omega, b, d.r.p, d.xnext.p, y, ki);
            });
            return;
        }
        with_lanes(d.lp, [&](auto C) {
            constexpr int AL = decltype(C)::value;
            k_amg_prolong<T, TB, AL><<<grid(d.n, AL), ANT, 0, st>>>(d.n, d.Prp.p, d.Pci.p, d.Pv.p, d.dinv.p, omega, b,
                                                                    d.xnext.p, d.x.p, ki);
        });
        with_lanes(d.la, [&](auto C) {
            constexpr int AL = decltype(C)::value;
            k_amg_post<T, TB, TY, AL><<<grid(d.n, AL), ANT, 0, st>>>(d.n, d.Arp.p, d.Aci.p, d.Av.p, d.dinv.p, omega, b,
                                                                     d.x.p, y, ki);
        });
    }
    void restrict_(Dev &d, const int *ki, hipStream_t st) {
        with_lanes(d.lr, [&](auto C) {
            constexpr int AL = decltype(C)::value;
            k_amg_restrict<T, AL><<<grid(d.nc, AL), ANT, 0, st>>>(d.nc, d.Rrp.p, d.Rci.p, d.Rv.p, d.r.p, d.bnext.p,
                                                                  ki);
        });
    }
    template <typename F>
    static void with_lanes(int al, F &&f) {
        if (al == 1)
            f(std::integral_constant<int, 1>{});
        else if (al == 2)
            f(std::integral_constant<int, 2>{});
        else if (al == 4)
            f(std::integral_constant<int, 4>{});
        else if (al == 8)
            f(std::integral_constant<int, 8>{});
        else if (al == 16)
            f(std::integral_constant<int, 16>{});
        else if (al == 32)
            f(std::integral_constant<int, 32>{});
        else
            f(std::integral_constant<int, 64>{});
    }
};

}  // namespace hfh

namespace hfh {

// ------------------------------------------- BiCGStab with the V-cycle
// Right preconditioning x = M^-1 u with M^-1 one V-cycle:
//   A(it):  it = 0: rho = rh.r, p = r.  it > 0: omega = t.s / t.t,
//           x += alpha ph + omega sh, r = s - omega t,
//           rho = rh.s - omega rh.t, beta = (rho / rho_old)(alpha / omega),
//           p = r + beta (p - omega v)                    (pointwise)
//   ph = V(p)
//   B(it):  stopping test on r.r; v = A ph; partial rh.v
//   C(it):  alpha = rho / rh.v; s = r - alpha v            (pointwise)
//   sh = V(s)
//   D(it):  t = A sh; partials t.s, t.t, rh.s, rh.t
// Rows of the SpMV launches run on L lanes; the pointwise launches use the
// same rows per workgroup so the partial arrays line up.

template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_abcg_a(int n, int it, T *__restrict__ x, T *__restrict__ r,
                                                const T *__restrict__ s, const T *__restrict__ t,
                                                const T *__restrict__ ph, const T *__restrict__ sh,
                                                T *__restrict__ p, const T *__restrict__ v, T *__restrict__ part,
                                                int nb, T *ks, int *ki, KSeg S) {
    const KCtx c = kctx<T>(S, nb, ks, ki);
    const int i = blockIdx.x * (KNT / L) + (int)threadIdx.x;
    const bool own = (int)threadIdx.x < KNT / L && i < n;
    if (it == 0) {
        const int sl[1] = {KP_RZ};
        __shared__ T pv[2];
        reduce_partials<T, 1>(part, nb, c, sl, ki, pv);
        const T ri = own ? r[i] : T(0);
        kblock_sync();
        if (pv[1] != T(0)) return;
        if (c.lead && threadIdx.x == 0) ks[KS_RHO] = pv[0];
        if (own) p[i] = ri;
        return;
    }
    const int sl[4] = {KP_TS, KP_TT, KP_RS, KP_RT};
    __shared__ T pv[5];
    reduce_partials<T, 4>(part, nb, c, sl, ki, pv, KI_BRK + ((it + 1) & 1));
    T xi = 0, si = 0, ti = 0, phi = 0, shi = 0, pi = 0, vi = 0;
    if (own) {
        xi = x[i];
        si = s[i];
        ti = t[i];
        phi = ph[i];
        shi = sh[i];
        pi = p[i];
        vi = v[i];
    }
    kblock_sync();
    if (pv[4] != T(0)) return;
    const int o = (it + 1) & 1;
    const T omega = pv[1] > T(0) ? pv[0] / pv[1] : T(0);
    const T alpha = ks[KS_ALPHA + o], rho_old = ks[KS_RHO + o];
    const T rho = pv[2] - omega * pv[3];
    const bool breakdown = rho == T(0) || rho_old == T(0) || omega == T(0);
    const T beta = breakdown ? T(0) : (rho / rho_old) * (alpha / omega);
    T acc[1] = {0};
    T ri = 0;
    if (own) {
        x[i] = xi + alpha * phi + omega * shi;
        ri = si - omega * ti;
        r[i] = ri;
    }
    if (breakdown) {
        if (c.lead && threadIdx.x == 0) {
            ki[KI_ITERS] = it;
            ki[KI_CONV] = 0;
            ki[KI_BRK + (it & 1)] = 1;
        }
        return;
    }
    if (c.lead && threadIdx.x == 0) ks[KS_RHO + (it & 1)] = rho;
    if (own) {
        p[i] = ri + beta * (pi - omega * vi);
        acc[0] = ri * ri;
    }
    const int so[1] = {KP_RR};
    write_partials<T, 1>(acc, part, nb, so);
}

template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_abcg_b(int n, int it, int maxit, T rtol, T atol,
                                                KRows<T> A, const T *__restrict__ ph,
                                                const T *__restrict__ rh, T *__restrict__ v, T *__restrict__ part,
                                                int nb, T *ks, int *ki, KSeg S) {
    const KCtx c = kctx<T>(S, nb, ks, ki);
    const int sl[2] = {KP_RR, KP_BB};
    __shared__ T pv[3];
    reduce_partials<T, 2>(part, nb, c, sl, ki, pv, KI_BRK + (it & 1));
    const int i = krow<L>();
    const bool in = i < n;
    const T *const vv[1] = {ph};
    T sv[1];
    row_sums<T, L, 1>(A, i, in, vv, sv);
    const T rhi = in ? rh[i] : T(0);
    kblock_sync();
    if (pv[2] != T(0)) return;
    if (krylov_stop<T>(c.lead, it == 0, it, maxit, pv[0], pv[1], rtol, atol, false, ks, ki, S.ndone)) return;
    T acc[1] = {0};
    if (in && klane<L>() == 0) {
        v[i] = sv[0];
        acc[0] = rhi * sv[0];
    }
    const int so[1] = {KP_RV};
    write_partials<T, 1>(acc, part, nb, so);
}

template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_abcg_c(int n, int it, const T *__restrict__ r, const T *__restrict__ v,
                                                T *__restrict__ s, T *__restrict__ part, int nb, T *ks, int *ki,
                                                KSeg S) {
    const KCtx c = kctx<T>(S, nb, ks, ki);
    const int sl[1] = {KP_RV};
    __shared__ T pv[2];
    reduce_partials<T, 1>(part, nb, c, sl, ki, pv, KI_BRK + (it & 1));
    const int i = blockIdx.x * (KNT / L) + (int)threadIdx.x;
    const bool own = (int)threadIdx.x < KNT / L && i < n;
    const T ri = own ? r[i] : T(0), vi = own ? v[i] : T(0);
    kblock_sync();
    if (pv[1] != T(0)) return;
    const T rho = ks[KS_RHO + (it & 1)];
    const T alpha = pv[0] != T(0) ? rho / pv[0] : T(0);
    if (c.lead && threadIdx.x == 0) ks[KS_ALPHA + (it & 1)] = alpha;
    if (own) s[i] = ri - alpha * vi;
}

template <typename T, int L>
__global__ void __launch_bounds__(KNT) k_abcg_d(int n, int it, KRows<T> A, const T *__restrict__ sh,
                                                const T *__restrict__ s, const T *__restrict__ rh,
                                                T *__restrict__ t, T *__restrict__ part, int nb, const int *ki,
                                                KSeg S) {
    __shared__ int s_done;
    if (S.wseg) ki += (size_t)S.wseg[blockIdx.x] * KI_N;
    if (threadIdx.x == 0) s_done = ki[KI_DONE] + ki[KI_BRK + (it & 1)];
    const int i = krow<L>();
    const bool in = i < n;
    const T *const vv[1] = {sh};
    T sv[1];
    row_sums<T, L, 1>(A, i, in, vv, sv);
    const T si = in ? s[i] : T(0), rhi = in ? rh[i] : T(0);
    kblock_sync();
    if (s_done) return;
    T acc[4] = {0, 0, 0, 0};
    if (in && klane<L>() == 0) {
        const T ti = sv[0];
        t[i] = ti;
        acc[0] = ti * si;
        acc[1] = ti * ti;
        acc[2] = rhi * si;
        acc[3] = rhi * ti;
    }
    const int so[4] = {KP_TS, KP_TT, KP_RS, KP_RT};
    write_partials<T, 4>(acc, part, nb, so);
}

}  // namespace hfh
