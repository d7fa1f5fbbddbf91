// Fused per-cell viscoelastic kernel: explicit log-conformation (or
// conformation) update followed by the polymeric stress of the new kernel
// tensor, in FP64 (eigen-decomposition and Omega are precision critical).
// Restates higflow/src/hig-flow-step-viscoelastic.c:212-544, 1517-1734,
// 1879-1909 and hig-flow-step.c:1185-1358.
#pragma once
#include <hip/hip_runtime.h>

#include "hfh_kernels.hpp"

namespace hfh {

struct VeParams {
    int ne, nrowK;
    int oDv, oK, oKN, oS;
    double dt, Re, De, beta, tol, galpha;
    int model, convec, transform;
    // per-cell Re, De and beta (three arrays of ne; batched runs), else the
    // scalars above
    const double *par = nullptr;
};

template <int D>
__device__ inline void dev_jacobi(double A[D][D], double d[D], double V[D][D]) {
    double bb[D], z[D];
    for (int ip = 0; ip < D; ip++) {
        for (int iq = 0; iq < D; iq++) V[ip][iq] = 0.0;
        V[ip][ip] = 1.0;
        bb[ip] = A[ip][ip];
        d[ip] = bb[ip];
        z[ip] = 0.0;
    }
    for (int i = 1; i <= 500; i++) {
        double tresh, sm = 0.0;
        for (int ip = 0; ip < D - 1; ip++)
            for (int iq = ip + 1; iq < D; iq++) sm += fabs(A[ip][iq]);
        if (sm == 0.0) return;
        tresh = (i < 4) ? 0.2 * sm / (D * D) : 0.0;
        for (int ip = 0; ip < D - 1; ip++)
            for (int iq = ip + 1; iq < D; iq++) {
                double g = 100.0 * fabs(A[ip][iq]);
                if ((i > 4) && (fabs(d[ip]) + g == fabs(d[ip])) && (fabs(d[iq]) + g == fabs(d[iq]))) {
                    A[ip][iq] = 0.0;
                } else if (fabs(A[ip][iq]) > tresh) {
                    double t, theta;
                    double h = d[iq] - d[ip];
                    if (fabs(h) + g == fabs(h)) {
                        t = A[ip][iq] / h;
                    } else {
                        theta = 0.5 * h / A[ip][iq];
                        t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
                        if (theta < 0.0) t = -t;
                    }
                    double cc = 1.0 / sqrt(1.0 + t * t);
                    double ss = t * cc;
                    double tau = ss / (1.0 + cc);
                    h = t * A[ip][iq];
                    z[ip] = z[ip] - h;
                    z[iq] = z[iq] + h;
                    d[ip] = d[ip] - h;
                    d[iq] = d[iq] + h;
                    A[ip][iq] = 0.0;
                    for (int j = 0; j <= ip - 1; j++) {
                        g = A[j][ip];
                        h = A[j][iq];
                        A[j][ip] = g - ss * (h + g * tau);
                        A[j][iq] = h + ss * (g - h * tau);
                    }
                    for (int j = ip + 1; j <= iq - 1; j++) {
                        g = A[ip][j];
                        h = A[j][iq];
                        A[ip][j] = g - ss * (h + g * tau);
                        A[j][iq] = h + ss * (g - h * tau);
                    }
                    for (int j = iq + 1; j < D; j++) {
                        g = A[ip][j];
                        h = A[iq][j];
                        A[ip][j] = g - ss * (h + g * tau);
                        A[iq][j] = h + ss * (g - h * tau);
                    }
                    for (int j = 0; j < D; j++) {
                        g = V[j][ip];
                        h = V[j][iq];
                        V[j][ip] = g - ss * (h + g * tau);
                        V[j][iq] = h + ss * (g - h * tau);
                    }
                }
            }
        for (int ip = 0; ip < D; ip++) {
            bb[ip] = bb[ip] + z[ip];
            d[ip] = bb[ip];
            z[ip] = 0.0;
        }
    }
}

template <int D>
__device__ inline void rt_a_r(const double A[D][D], const double R[D][D], double B[D][D]) {
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) {
            double s = 0.0;
            for (int k = 0; k < D; k++)
                for (int l = 0; l < D; l++) s += R[k][i] * A[k][l] * R[l][j];
            B[i][j] = s;
        }
}

template <int D>
__device__ inline void r_a_rt(const double A[D][D], const double R[D][D], double B[D][D]) {
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) {
            double s = 0.0;
            for (int k = 0; k < D; k++)
                for (int l = 0; l < D; l++) s += R[i][k] * A[k][l] * R[j][l];
            B[i][j] = s;
        }
}

__device__ inline double kinv(const VeParams &P, double l) { return P.transform == 1 ? exp(l) : l; }
__device__ inline double kjac(const VeParams &P, double l) {
    if (P.transform != 1) return 1.0;
    return (l < P.tol) ? 1.0 / P.tol : 1.0 / l;
}

// hig-flow-step.c:1185-1358 (CUBISTA cell convection of one component)
template <typename C>
__device__ inline C cubista_cell(C vl, C vr, C kl, C kr, C kll, C krr, C kc, int il, int ir, int ill, int irr,
                                 C cd) {
    const C a = 1.7500, b = 0.3750, c = 0.7500, d = 0.1250, e = 0.2500;
    C vbar, fi, conv1 = 0.0, conv2 = 0.0;
    vbar = vr;
    if (vbar > 0.0) {
        if (HFH_FLT_EQ(kr, kl)) {
            conv1 = vbar * kc;
        } else {
            fi = (kc - kl) / (kr - kl);
            if ((fi <= 0.0) || (fi >= 1.0)) {
                conv1 = vbar * kc;
            } else {
                if (fi < b) conv1 = il ? vbar * (a * kc - c * kl) : vbar * kc;
                if ((fi >= b) && (fi <= c)) conv1 = (il && ir) ? vbar * (c * kc + b * kr - d * kl) : vbar * kc;
                if (fi > c) conv1 = ir ? vbar * (e * kc + c * kr) : vbar * kc;
            }
        }
    } else {
        if (ir && irr) {
            if (HFH_FLT_EQ(kc, krr)) {
                conv1 = vbar * kr;
            } else {
                fi = (kr - krr) / (kc - krr);
                if ((fi <= 0.0) || (fi >= 1.0)) {
                    conv1 = vbar * kr;
                } else {
                    if (fi < b) conv1 = vbar * (a * kr - c * krr);
                    if ((fi >= b) && (fi <= c)) conv1 = vbar * (c * kr + b * kc - d * krr);
                    if (fi > c) conv1 = vbar * (c * kc + e * kr);
                }
            }
        } else if (ir && !irr) {
            if (HFH_FLT_EQ(kc, krr)) {
                conv1 = vbar * kr;
            } else {
                fi = (kr - krr) / (kc - krr);
                if ((fi <= 0.0) || (fi >= 1.0)) {
                    conv1 = vbar * kr;
                } else {
                    if (fi <= c) conv1 = vbar * kr;
                    if (fi > c) conv1 = vbar * (c * kc + e * kr);
                }
            }
        } else {
            vbar = vr;
            conv1 = vbar * kc;
            vbar = vl;
            conv2 = (vbar > 0.0) ? vbar * kl : vbar * kc;
            return ((conv1 - conv2) / cd);
        }
    }
    vbar = vl;
    if (vbar > 0.0) {
        if (il && ill) {
            if (HFH_FLT_EQ(kc, kll)) {
                conv2 = vbar * kl;
            } else {
                fi = (kl - kll) / (kc - kll);
                if ((fi <= 0.0) || (fi >= 1.0)) {
                    conv2 = vbar * kl;
                } else {
                    if (fi < b) conv2 = vbar * (a * kl - c * kll);
                    if ((fi >= b) && (fi <= c)) conv2 = vbar * (b * kc + c * kl - d * kll);
                    if (fi > c) conv2 = vbar * (c * kc + e * kl);
                }
            }
        } else if (il && !ill) {
            if (HFH_FLT_EQ(kc, kll)) {
                conv2 = vbar * kl;
            } else {
                fi = (kl - kll) / (kc - kll);
                if ((fi <= 0.0) || (fi >= 1.0)) {
                    conv2 = vbar * kl;
                } else {
                    if (fi <= c) conv2 = vbar * kl;
                    if (fi > c) conv2 = vbar * (c * kc + e * kl);
                }
            }
        } else {
            vbar = vr;
            conv1 = (vbar > 0.0) ? vbar * kc : vbar * kr;
            vbar = vl;
            conv2 = vbar * kc;
            return ((conv1 - conv2) / cd);
        }
    } else {
        if (HFH_FLT_EQ(kl, kr)) {
            conv2 = vbar * kc;
        } else {
            fi = (kc - kr) / (kl - kr);
            if ((fi <= 0.0) || (fi >= 1.0)) {
                conv2 = vbar * kc;
            } else {
                if (fi < b) conv2 = ir ? vbar * (a * kc - c * kr) : vbar * kc;
                if ((fi >= b) && (fi <= c)) conv2 = (il && ir) ? vbar * (c * kc + b * kl - d * kr) : vbar * kc;
                if (fi > c) conv2 = il ? vbar * (e * kc + c * kl) : vbar * kc;
            }
        }
    }
    return ((conv1 - conv2) / cd);
}

// slots (slot-major, ne cells): 0..2D-1 (vl_d, vr_d); 2D+4d+{0..3} (kl, kr,
// kll, krr) of dimension d, component 0 (component m at +m*nrowK).
// flags: bit 4d+{0..3} = incell of kl, kr, kll, krr.
template <int D, typename C = double>
__global__ void __launch_bounds__(BLOCK) k_constitutive(VeParams P, const int *__restrict__ slots,
                                                        const unsigned *__restrict__ flags,
                                                        const double *__restrict__ cdelta, double *X) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.ne) return;
    const int ne = P.ne;
    const double Re = P.par ? P.par[c] : P.Re, De = P.par ? P.par[(size_t)ne + c] : P.De,
                 beta = P.par ? P.par[2 * (size_t)ne + c] : P.beta;
    constexpr int NS = D * (D + 1) / 2;
    double Du[D][D], Kn[D][D], Kc[D][D];
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) Du[i][j] = X[P.oDv + (size_t)(i * D + j) * ne + c];
    {
        int m = 0;
        for (int i = 0; i < D; i++)
            for (int j = i; j < D; j++, m++) {
                double v = X[P.oK + (size_t)m * ne + c];
                Kn[i][j] = v;
                Kn[j][i] = v;
            }
    }
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) Kc[i][j] = Kn[i][j];
    double R[D][D], Kl[D], lam[D];
    dev_jacobi<D>(Kc, Kl, R);
    for (int k = 0; k < D; k++) lam[k] = kinv(P, Kl[k]);
    double M[D][D];
    rt_a_r<D>(Du, R, M);
    double Oa[D][D], Om[D][D];
    for (int i = 0; i < D - 1; i++) {
        for (int j = i + 1; j < D; j++) {
            Oa[i][j] = (M[i][j] * lam[j] + M[j][i] * lam[i]) / (lam[j] - lam[i] + 1.0e-12);
            Oa[j][i] = -Oa[i][j];
        }
        Oa[i][i] = 0.0;
    }
    Oa[D - 1][D - 1] = 0.0;
    r_a_rt<D>(Oa, R, Om);
    double jl[D];
    for (int k = 0; k < D; k++) jl[k] = kjac(P, lam[k]);
    double Ba[D][D], BB[D][D], Ma[D][D], MM[D][D];
    for (int i = 0; i < D; i++) {
        for (int j = i + 1; j < D; j++) {
            Ba[i][j] = Ba[j][i] = 0.0;
            Ma[i][j] = Ma[j][i] = 0.0;
        }
        Ba[i][i] = M[i][i] * lam[i] * jl[i];
        double aux = 1.0 - lam[i];
        Ma[i][i] = (P.model == 1) ? (aux - P.galpha * aux * aux) * jl[i] : (1.0 - lam[i]) * jl[i];
    }
    r_a_rt<D>(Ba, R, BB);
    r_a_rt<D>(Ma, R, MM);
    double RHS[D][D];
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) {
            double ok = 0.0, ko = 0.0;
            for (int k = 0; k < D; k++) {
                ok += Om[i][k] * Kn[k][j];
                ko += Kn[i][k] * Om[k][j];
            }
            RHS[i][j] = ok - ko + 2.0 * BB[i][j] + MM[i][j] / De;
        }
    // convection, the kernel update and the stress in C (the eigen-
    // decompositions and Omega above stay FP64)
    C vl[D], vr[D], cd[D];
    unsigned fl = flags[c];
    for (int d = 0; d < D; d++) {
        vl[d] = (C)X[slots[(size_t)(2 * d) * ne + c]];
        vr[d] = (C)X[slots[(size_t)(2 * d + 1) * ne + c]];
        cd[d] = (C)cdelta[(size_t)d * ne + c];
    }
    double Knew[D][D];
    int m = 0;
    for (int i = 0; i < D; i++)
        for (int j = i; j < D; j++, m++) {
            C rhs = 0.0;
            const C kij = (C)Kn[i][j];
            if (P.convec == 0) {
                C dK[D];
                for (int d = 0; d < D; d++) {
                    const int kb = 2 * D + 4 * d;
                    C Kl2 = (C)X[slots[(size_t)(kb + 0) * ne + c] + (size_t)m * P.nrowK];
                    C Kr2 = (C)X[slots[(size_t)(kb + 1) * ne + c] + (size_t)m * P.nrowK];
                    int il = (fl >> (4 * d)) & 1, ir = (fl >> (4 * d + 1)) & 1;
                    if (il && ir) dK[d] = (Kr2 - Kl2) / (C(2) * C(1) * cd[d]);
                    else if (ir) dK[d] = (Kr2 - kij) / (C(1) * cd[d]);
                    else dK[d] = (kij - Kl2) / (C(1) * cd[d]);
                }
                for (int d = 0; d < D; d++) rhs -= (C(0.5) * (vl[d] + vr[d])) * dK[d];
            } else {
                for (int d = 0; d < D; d++) {
                    const int kb = 2 * D + 4 * d;
                    C kl = (C)X[slots[(size_t)(kb + 0) * ne + c] + (size_t)m * P.nrowK];
                    C kr = (C)X[slots[(size_t)(kb + 1) * ne + c] + (size_t)m * P.nrowK];
                    C kll = (C)X[slots[(size_t)(kb + 2) * ne + c] + (size_t)m * P.nrowK];
                    C krr = (C)X[slots[(size_t)(kb + 3) * ne + c] + (size_t)m * P.nrowK];
                    unsigned f4 = fl >> (4 * d);
                    rhs -= cubista_cell<C>(vl[d], vr[d], kl, kr, kll, krr, kij, f4 & 1, (f4 >> 1) & 1,
                                        (f4 >> 2) & 1, (f4 >> 3) & 1, cd[d]);
                }
            }
            rhs += (C)RHS[i][j];
            const double kernel = (double)(kij + (C)P.dt * rhs);
            X[P.oKN + (size_t)m * ne + c] = kernel;
            Knew[i][j] = kernel;
            Knew[j][i] = kernel;
        }
    // polymeric tensor of the new kernel (hig-flow-step-viscoelastic.c:212-347)
    double R2[D][D], l2[D], B[D][D], A[D][D];
    dev_jacobi<D>(Knew, l2, R2);
    for (int i = 0; i < D; i++) {
        for (int j = i + 1; j < D; j++) B[i][j] = B[j][i] = 0.0;
        B[i][i] = kinv(P, l2[i]);
    }
    r_a_rt<D>(B, R2, A);
    const double fA = 1.0, a = 1.0, xi = 0.0;
    for (int i = 0; i < D; i++) {
        for (int j = 0; j < D; j++) {
            const C cb = (C)beta, cre = (C)Re, cde = (C)De;
            C s = (C(1) - cb) * (C(fA) * (C)A[i][j]) / (cre * cde * (C(1) - C(xi))) - (C(1) - cb) * (C)Du[i][j] / cre;
            if (i == j) s += -C(a) * (C(1) - cb) / (cre * cde * (C(1) - C(xi)));
            X[P.oS + (size_t)(i * D + j) * ne + c] = (double)s;
        }
    }
    (void)NS;
}

}  // namespace hfh
