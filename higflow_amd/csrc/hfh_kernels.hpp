// Device kernels of the projection step (gfx950, wave64).
//
// All FP64 arithmetic mirrors the reference's operation order
// (hig-flow-discret.c, hig-flow-step.c, hig-flow-step-viscoelastic.c) and the
// library is built with -ffp-contract=off, so a kernel and the CPU reference
// round identically whenever they see identical inputs.
#pragma once
#include <hip/hip_runtime.h>

namespace hfh {

constexpr int BLOCK = 256;

#define HFH_FLT_EQ(x, y) (fabs((x) - (y)) < 1.0e-12)

// ---------------------------------------------------------------- interp
// Template rows of an interpolation group: rows whose entries repeat a
// pattern store its id code[r] >= 0 and the column base[r] of their first
// entry; template t holds tcnt[t] <= GT_MAX entries (column offset from the
// base, weight) at t * GT_MAX in toff / tw, in the row's order.  code[r] < 0,
// or code == nullptr, reads the CSR row.
constexpr int GT_MAX = 64;
template <typename W>
struct GTempl {
    const int *code = nullptr, *base = nullptr, *tcnt = nullptr, *toff = nullptr;
    const W *tw = nullptr;
};

// entries [*b, *e) of row r, read through *cp / *vp with column base *cb
template <typename W>
__device__ inline void grow(const GTempl<W> &tp, int r, const int *rowptr, const int *col, const W *val, int &b,
                            int &e, const int *&cp, const W *&vp, int &cb) {
    const int t = tp.code ? tp.code[r] : -1;
    if (t >= 0) {
        b = t * GT_MAX;
        e = b + tp.tcnt[t];
        cp = tp.toff;
        vp = tp.tw;
        cb = tp.base[r];
    } else {
        b = rowptr[r];
        e = rowptr[r + 1];
        cp = col;
        vp = val;
        cb = 0;
    }
}

// Affine rows over the state vector X:
//   v = cst[r] + sum_k val[k] * X[col[k] + c*in_stride]
//   dst[out[r] + c*out_stride] = alpha * v + (yoff ? X[yoff[r]] : 0)
// One thread per (row, component); rows are short (identity/ghost rows) or
// WLS rows of up to 45/120 entries kept contiguous.  W is the precision of
// the weights and of the arithmetic (the state T stays FP64): FP32 rows
// serve the mixed form's momentum and constitutive evaluations.
template <typename T, typename W = T>
__global__ void __launch_bounds__(BLOCK) k_interp(int nrow, int ncomp, const int *__restrict__ rowptr,
                                                  const int *__restrict__ col, const W *__restrict__ val,
                                                  const W *__restrict__ cst, const int *__restrict__ out,
                                                  const int *__restrict__ yoff, W alpha, int in_stride,
                                                  int out_stride, const T *X, T *dst, GTempl<W> tp) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrow * ncomp) return;
    int r = t % nrow, c = t / nrow;
    int b, e, cb;
    const int *cp;
    const W *vp;
    grow<W>(tp, r, rowptr, col, val, b, e, cp, vp, cb);
    W v = cst[r];
    const T *Xc = X + (size_t)c * in_stride + cb;
    for (int k = b; k < e; k++) v += (W)Xc[cp[k]] * vp[k];
    W o = alpha * v;
    if (yoff) o = (W)X[yoff[r]] + o;
    dst[out[r] + (size_t)c * out_stride] = (T)o;
}

// The same rows with L lanes per (row, component): lane l sums entries
// l, l+L, ...; the lane sums combine by a butterfly in fixed order.  Long
// WLS rows (up to 45/120 entries) then run L-way parallel.
template <typename T, int L, typename W = T>
__global__ void __launch_bounds__(BLOCK) k_interp_l(int nrow, int ncomp, const int *__restrict__ rowptr,
                                                    const int *__restrict__ col, const W *__restrict__ val,
                                                    const W *__restrict__ cst, const int *__restrict__ out,
                                                    const int *__restrict__ yoff, W alpha, int in_stride,
                                                    int out_stride, const T *X, T *dst, GTempl<W> tp) {
    const long g = ((long)blockIdx.x * blockDim.x + threadIdx.x) / L;
    const int lane = threadIdx.x % L;
    const bool in = g < (long)nrow * ncomp;
    const int r = in ? (int)(g % nrow) : 0, c = in ? (int)(g / nrow) : 0;
    int b = 0, e = 0, cb = 0;
    const int *cp = col;
    const W *vp = val;
    if (in) grow<W>(tp, r, rowptr, col, val, b, e, cp, vp, cb);
    const T *Xc = X + (size_t)c * in_stride + cb;
    W v = 0;
    // four entries per pass: indices first, then the gathered values
    constexpr int U = 4;
#pragma unroll 1
    for (int k0 = b + lane; k0 < e; k0 += U * L) {
        int j[U];
        W a[U];
        bool ok[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int k = k0 + q * L;
            ok[q] = k < e;
            j[q] = ok[q] ? cp[k] : 0;
            a[q] = ok[q] ? vp[k] : W(0);
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            const W x = (W)Xc[j[q]];
            if (ok[q]) v += x * a[q];
        }
    }
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (in && lane == 0) {
        v = cst[r] + v;
        W o = alpha * v;
        if (yoff) o = (W)X[yoff[r]] + o;
        dst[out[r] + (size_t)c * out_stride] = (T)o;
    }
}

// ------------------------------------------------------------- momentum
struct MomParams {
    int nf, ns;
    int u_off, us_off;
    double dt, Re;
    int incremental, second, ve, dim;
    int semi;            // semi-implicit Euler: write the Helmholtz right-hand side
    double *bout;        // semi: b[f] = explicit part + boundary constant
    const double *bcst;  // semi: boundary constant of each Helmholtz row
    const double *re = nullptr;  // per-facet Reynolds number (batched runs), else Re
};

// Per-facet slot layout (produced by the recorder):
//   0 F, 1 dpdx, 2+6*d2+{0..5}: ul, ur, ull, urr, vcR, vcL (neighbours in
//   direction d2 and the d2-velocity at fc -/+ delta/2), then D-1 transverse
//   velocities, then D tensor divergences (viscoelastic).
// flags: bits 0-1 final convective type, bit 2+d2 infacet_rr, bit 2+D+d2
// infacet_ll (hig-flow-discret.c:131-251).
template <int D, typename C = double>
__global__ void __launch_bounds__(BLOCK) k_momentum(MomParams P, const int *__restrict__ slots,
                                                    const unsigned *__restrict__ flags,
                                                    const double *__restrict__ fdelta, double *X) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.nf) return;
    const int nf = P.nf;
    const int dim = P.dim;
    const C a = 1.7500, b = 0.3750, c = 0.7500, d = 0.1250, e = 0.2500;
    unsigned fl = flags[f];
    int ctype = fl & 3;
    C fd[D];
    for (int k = 0; k < D; k++) fd[k] = (C)fdelta[(size_t)k * nf + f];
    C uf = (C)X[P.u_off + f];
    C F = (C)X[slots[f]];
    C dpdx = (C)X[slots[(size_t)nf + f]];
    C conv = 0.0, diff = 0.0, tens = 0.0;
    int vk = 2 + 6 * D;
    for (int d2 = 0; d2 < D; d2++) {
        const int base = 2 + 6 * d2;
        C ul = (C)X[slots[(size_t)(base + 0) * nf + f]];
        C ur = (C)X[slots[(size_t)(base + 1) * nf + f]];
        C v;
        if (d2 == dim) {
            v = uf;
        } else {
            v = (C)X[slots[(size_t)vk * nf + f]];
            vk++;
        }
        C dudxc = (ur - ul) / (C(2) * C(1) * fd[d2]);
        C dudxr = (ur - uf) / (C(1) * fd[d2]);
        C dudxl = (uf - ul) / (C(1) * fd[d2]);
        if (ctype == 2) {
            if (P.second == 1) {
                C ull = (C)X[slots[(size_t)(base + 3) * nf + f]];
                C urr = (C)X[slots[(size_t)(base + 2) * nf + f]];
                int irr = (fl >> (2 + d2)) & 1, ill = (fl >> (2 + D + d2)) & 1;
                C vc = (C)X[slots[(size_t)(base + 4) * nf + f]];
                C conv1 = 0.0, conv2 = 0.0, fi;
                if (vc > 0.0) {
                    if (HFH_FLT_EQ(ur, ul)) {
                        conv1 = vc * uf;
                    } else {
                        fi = (uf - ul) / (ur - ul);
                        if ((fi <= 0.0) || (fi >= 1.0)) {
                            conv1 = vc * uf;
                        } else {
                            if (fi < b) conv1 = vc * (a * uf - c * ul);
                            if ((fi >= b) && (fi <= c)) conv1 = vc * (c * uf + b * ur - d * ul);
                            if (fi > c) conv1 = vc * (e * uf + c * ur);
                        }
                    }
                } else {
                    if (HFH_FLT_EQ(uf, urr)) {
                        conv1 = vc * ur;
                    } else {
                        fi = (ur - urr) / (uf - urr);
                        if ((fi <= 0.0) || (fi >= 1.0)) {
                            conv1 = vc * ur;
                        } else {
                            if (fi < b) conv1 = irr ? vc * (a * ur - c * urr) : vc * ur;
                            if ((fi >= b) && (fi <= c)) conv1 = irr ? vc * (c * ur + b * uf - d * urr) : vc * ur;
                            if (fi > c) conv1 = vc * (c * uf + e * ur);
                        }
                    }
                }
                vc = (C)X[slots[(size_t)(base + 5) * nf + f]];
                if (vc > 0.0) {
                    if (HFH_FLT_EQ(uf, ull)) {
                        conv2 = vc * ul;
                    } else {
                        fi = (ul - ull) / (uf - ull);
                        if ((fi <= 0.0) || (fi >= 1.0)) {
                            conv2 = vc * ul;
                        } else {
                            if (fi < b) conv2 = ill ? vc * (a * ul - c * ull) : vc * ul;
                            if ((fi >= b) && (fi <= c)) conv2 = ill ? vc * (b * uf + c * ul - d * ull) : vc * ul;
                            if (fi > c) conv2 = vc * (c * uf + e * ul);
                        }
                    }
                } else {
                    if (HFH_FLT_EQ(ul, ur)) {
                        conv2 = vc * uf;
                    } else {
                        fi = (uf - ur) / (ul - ur);
                        if ((fi <= 0.0) || (fi >= 1.0)) {
                            conv2 = vc * uf;
                        } else {
                            if (fi < b) conv2 = vc * (a * uf - c * ur);
                            if ((fi >= b) && (fi <= c)) conv2 = vc * (b * ul + c * uf - d * ur);
                            if (fi > c) conv2 = vc * (c * ul + e * uf);
                        }
                    }
                }
                conv += ((conv1 - conv2) / fd[dim]);
            } else if (P.second == 2) {
                C urr = (C)X[slots[(size_t)(base + 2) * nf + f]];
                C ull = (C)X[slots[(size_t)(base + 3) * nf + f]];
                C dudxrr = (urr - ur) / (C(1) * fd[d2]);
                C dudxll = (uf - ull) / (C(1) * fd[d2]);
                (void)dudxll;
                C dll = (ul - ull) / (C(1) * fd[d2]);
                conv += (v + fabs(v)) * C(0.1875) * dudxl;
                conv += (v + fabs(v)) * C(0.3750) * dudxr;
                conv -= (v + fabs(v)) * C(0.0625) * dudxrr;
                conv += (v - fabs(v)) * C(0.3750) * dudxl;
                conv += (v - fabs(v)) * C(0.1875) * dudxr;
                conv -= (v - fabs(v)) * C(0.0625) * dll;
            } else {
                C du2 = (ul - C(2) * uf + ur) / (fd[d2] * fd[d2] * C(1) * C(1));
                conv += C(0.5) * (v + fabs(v)) * dudxl;
                conv += C(0.5) * (v - fabs(v)) * dudxr;
                if (v > 0.0) {
                    if (fabs(dudxl) > 1.0e-8) conv += C(0.5) * v * fd[d2] * du2 / dudxl;
                } else {
                    if (fabs(dudxr) > 1.0e-8) conv -= C(0.5) * v * fd[d2] * du2 / dudxr;
                }
            }
        } else if (ctype == 1) {
            conv += C(0.5) * (v + fabs(v)) * dudxl;
            conv += C(0.5) * (v - fabs(v)) * dudxr;
        } else {
            conv += v * dudxc;
        }
        diff += (ul - C(2) * uf + ur) / (fd[d2] * fd[d2] * C(1) * C(1));
    }
    diff /= (C)(P.re ? P.re[f] : P.Re);
    C rhs = 0.0;
    C pt = P.incremental ? dpdx : C(0);
    if (P.semi) {  // hig-flow-step.c:815-824, hig-flow-step-viscoelastic.c:1137-1148
        rhs += F;
        rhs -= pt;
        if (P.ve) {
            int tk = 2 + 6 * D + (D - 1);
            for (int d2 = 0; d2 < D; d2++) tens += (C)X[slots[(size_t)(tk + d2) * nf + f]];
            rhs += tens;
        }
        rhs -= conv;
        rhs *= (C)P.dt;
        rhs += uf;
        P.bout[f] = (double)(rhs + (C)P.bcst[f]);
        return;
    }
    if (!P.ve) {
        rhs -= pt;
        rhs -= conv;
        rhs += diff;
        rhs += F;
    } else {
        int tk = 2 + 6 * D + (D - 1);
        for (int d2 = 0; d2 < D; d2++) tens += (C)X[slots[(size_t)(tk + d2) * nf + f]];
        rhs += F;
        rhs -= pt;
        rhs += tens;
        rhs -= conv;
        rhs += diff;
    }
    X[P.us_off + f] = (double)(uf + (C)P.dt * rhs);
}

// Initial guess of a time step's solve from the last solutions: `order`
// 0 out = cur; 1 out = 2 cur - p1; 2 out = 3 cur - 3 p1 + p2 (polynomial
// extrapolation in time); then p2 = p1, p1 = cur.  `out` may alias `cur`.
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_extrap(int n, const T *cur, T *p1, T *p2, T *out, int order) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T v = cur[i], a = p1[i];
    out[i] = order == 0 ? v : order == 1 ? T(2) * v - a : T(3) * (v - a) + p2[i];
    p2[i] = a;
    p1[i] = v;
}

// the same for the D velocity components in one launch (blockIdx.y is the
// component): cur[d], p1[d], p2[d], out[d] of length n[d]
struct ExtrapSet {
    int n[3];
    const double *cur[3];
    double *p1[3], *p2[3], *out[3];
};
__global__ void __launch_bounds__(BLOCK) k_extrap_set(ExtrapSet q, int order) {
    const int d = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n[d]) return;
    const double v = q.cur[d][i], a = q.p1[d][i];
    q.out[d][i] = order == 0 ? v : order == 1 ? 2.0 * v - a : 3.0 * (v - a) + q.p2[d][i];
    q.p2[d][i] = a;
    q.p1[d][i] = v;
}

// dst[d] = src[d] for the D velocity components in one launch
struct CopySet {
    int n[3];
    const double *src[3];
    double *dst[3];
};
__global__ void __launch_bounds__(BLOCK) k_copy_set(CopySet q) {
    const int d = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < q.n[d]) q.dst[d][i] = q.src[d][i];
}

// p += dp over all cells (hig-flow-step.c:311-340)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_axpy_field(int n, T *X, int dst, int src) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) X[dst + i] = X[dst + i] + X[src + i];
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_copy(int n, const T *src, T *dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_fill(int n, T *x, T v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

inline int nblocks(long n) { return (int)((n + BLOCK - 1) / BLOCK); }

// Min and max of x[0, n) from +-1e16, as the reference's final-velocity
// loop starts them (hig-flow-step.c:250-251, 290-291).  Workgroup w reduces
// elements [w MM_CHUNK, (w + 1) MM_CHUNK) into part[2w], part[2w + 1]; the
// workgroup that finishes last (an agent-scope ticket, no spinning) reduces
// the partials into out[0], out[1] and rearms the ticket.  Launch
// mm_groups(n) workgroups of 1024 threads with *ticket == 0.  Min and max
// are exact, so the result does not depend on the partition.
constexpr int MM_CHUNK = 8192;
inline int mm_groups(long n) { return n <= 0 ? 1 : (int)((n + MM_CHUNK - 1) / MM_CHUNK); }

__device__ inline void mm_block_reduce(double &mn, double &mx, double *smn, double *smx) {
    for (int o = 32; o > 0; o >>= 1) {
        const double a = __shfl_down(mn, o, 64), b = __shfl_down(mx, o, 64);
        if (a < mn) mn = a;
        if (b > mx) mx = b;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        smn[w] = mn;
        smx[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < 16; k++) {
            if (smn[k] < mn) mn = smn[k];
            if (smx[k] > mx) mx = smx[k];
        }
}

__global__ void __launch_bounds__(1024) k_minmax(long n, const double *__restrict__ x, double *part,
                                                   unsigned *ticket, double *__restrict__ out) {
    __shared__ double smn[16], smx[16];
    __shared__ int last;
    const long lo = (long)blockIdx.x * MM_CHUNK;
    const long hi = lo + MM_CHUNK < n ? lo + MM_CHUNK : n;
    double mn = 1.0e16, mx = -1.0e16;
    for (long i = lo + threadIdx.x; i < hi; i += 1024) {
        const double v = x[i];
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    }
    mm_block_reduce(mn, mx, smn, smx);
    if (threadIdx.x == 0) {
        __hip_atomic_store(&part[2 * blockIdx.x], mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&part[2 * blockIdx.x + 1], mx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    mn = 1.0e16;
    mx = -1.0e16;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += 1024) {
        const double a = __hip_atomic_load(&part[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const double b = __hip_atomic_load(&part[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a < mn) mn = a;
        if (b > mx) mx = b;
    }
    __syncthreads();
    mm_block_reduce(mn, mx, smn, smx);
    if (threadIdx.x == 0) {
        out[0] = mn;
        out[1] = mx;
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Polymeric stress of the reference's Tmin/Tmax log (hig-flow-step-
// viscoelastic.c:319-324): T = S + (1 - beta) Du / Re, in its order of
// operations
__global__ void __launch_bounds__(BLOCK) k_poly_stress(int n, const double *__restrict__ S,
                                                       const double *__restrict__ Du, double beta, double Re,
                                                       double *__restrict__ T) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) T[i] = S[i] + (1.0 - beta) * Du[i] / Re;
}

__global__ void k_zero_u32(unsigned *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

}  // namespace hfh
