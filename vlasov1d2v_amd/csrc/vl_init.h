// Initial condition of the reference driver on the device: the ring
// distribution plus the Bernstein-mode eigenfunction perturbation
// (Distribution.initialize_gpu / eigenfunction, main/grid.py:207-279).
//
// One thread per velocity node (u, v) forms the ring value and the complex
// velocity mode
//   vel(u, v) = -i exp(i beta sin phi) df_dv Upsilon / k1^2,
//   Upsilon   = sum_{n=-T..T} n / (n - om) J_n(beta) exp(-i n phi),
//   beta = -k1 r om_pc, phi = atan2(v, u), x = (r / vt)^2 / 2,
//   df_dv = f0 (j / (x + 1e-16) - 1) / vt^2, f0 = x^j e^-x / (2 pi vt^2 j!),
// then writes f = ring + 2 delta_n Re(exp(i k1 x) vel) for every x node of
// the ghosted state (reference layout, FP64).
#pragma once
#include "vl_common.h"

namespace vl {

constexpr int kEigenMaxTerms = 24;   // Bessel orders 0..T, T <= kEigenMaxTerms

struct EigenArgs {
    int nx, nu, nv, n;
    int ring_j, terms, perturbation, pert_only;
    double vt, k1, om_re, om_im, om_pc, delta_n;
    double ring_factor;   // 1 / (2^j 2 pi vt^2 j!)   (main/grid.py:253)
    double f0_factor;     // 1 / (2 pi vt^2 j!)       (main/grid.py:213)
};

// J_k(b), k = 0..T, of real b, J_k(-b) = (-1)^k J_k(b).
//   |b| < 1e-2: three terms of the power series
//     J_k(x) = (x/2)^k / k! (1 - (x/2)^2 / (k+1) + (x/2)^4 / (2 (k+1)(k+2)) - ...)
//     (next term below 1e-15 relative);
//   otherwise Miller's backward recurrence J_{k-1} = (2k / x) J_k - J_{k+1}
//     from an even start order well above max(T, x), normalised by
//     J_0 + 2 sum_{m >= 1} J_{2m} = 1.  Orders above T run in a plain loop
//     (values kept below 1e150 by rescaling), orders T+1 .. 1 in an unrolled
//     one so that J stays in registers (they grow by at most 1e68 there for
//     x >= 1e-2, T <= 24).
template <int T>
__device__ __forceinline__ void bessel_j_orders(double b, double (&J)[T + 1]) {
    const double x = fabs(b);
    if (x < 1e-2) {
        const double h = 0.5 * x, h2 = h * h;
        double p = 1.0;                          // h^k / k!
#pragma unroll
        for (int k = 0; k <= T; ++k) {
            J[k] = p * (1.0 - h2 / (k + 1) + (h2 * h2) / (2.0 * (k + 1) * (k + 2)));
            p *= h / (k + 1);
        }
    } else {
        const int top = T > (int)x ? T : (int)x;
        int m = top + 20 + (int)sqrt(40.0 * (double)top);
        m += m & 1;                              // even start order
        double jp = 0.0, jc = 1e-30, even = 0.0; // J_{k+1}, J_k (unnormalised)
        for (int k = m; k > T + 1; --k) {        // produces J_{k-1}, k-1 > T
            const double jm = (2.0 * k / x) * jc - jp;
            jp = jc;
            jc = jm;
            if (((k - 1) & 1) == 0) even += jc;
            if (fabs(jc) > 1e150) {              // rescale (only ratios matter)
                jc *= 1e-150;
                jp *= 1e-150;
                even *= 1e-150;
            }
        }
#pragma unroll
        for (int k = T + 1; k >= 1; --k) {       // produces J_{k-1}
            const double jm = (2.0 * k / x) * jc - jp;
            jp = jc;
            jc = jm;
            J[k - 1] = jc;
            if ((k - 1) >= 2 && ((k - 1) & 1) == 0) even += jc;
        }
        const double norm = J[0] + 2.0 * even;
#pragma unroll
        for (int k = 0; k <= T; ++k) J[k] /= norm;
    }
    if (b < 0.0) {
#pragma unroll
        for (int k = 1; k <= T; k += 2) J[k] = -J[k];
    }
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_init_eigen(double *__restrict__ f, const double *__restrict__ xnodes,
                                                       const double *__restrict__ unodes,
                                                       const double *__restrict__ vnodes, EigenArgs a) {
    const int NUg = (a.nu + 2) * a.n, NVg = (a.nv + 2) * a.n;
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= NUg * NVg) return;
    const int gu = idx / NVg, gv = idx - (idx / NVg) * NVg;
    const double u = unodes[gu], v = vnodes[gv];
    // ring (main/grid.py:253-268)
    const double rsq = (u * u + v * v) / (a.vt * a.vt);
    const double ring = a.ring_factor * (pow(rsq, (double)a.ring_j) * exp(-0.5 * rsq));
    double vre = 0.0, vim = 0.0;     // velocity mode
    if (a.perturbation) {
        const double r = sqrt(u * u + v * v);
        const double phi = atan2(v, u);
        const double beta = -a.k1 * r * a.om_pc;
        const double x = 0.5 * ((r / a.vt) * (r / a.vt));
        const double f0 = a.f0_factor * (pow(x, (double)a.ring_j) * exp(-x));
        const double df_dv = (f0 * ((double)a.ring_j / (x + 1.0e-16) - 1.0)) / (a.vt * a.vt);
        double J[T + 1];
        bessel_j_orders<T>(beta, J);
        double ure = 0.0, uim = 0.0;
        for (int nn = -a.terms; nn <= a.terms; ++nn) {
            const int an = nn < 0 ? -nn : nn;
            double jn = 0.0;
#pragma unroll
            for (int k = 0; k <= T; ++k)
                if (k == an) jn = J[k];
            if (nn < 0 && (an & 1)) jn = -jn;               // J_{-n} = (-1)^n J_n
            // c = n / (n - om), om = om_re + i om_im
            const double dre = (double)nn - a.om_re, dim = -a.om_im;
            const double den = dre * dre + dim * dim;
            const double cre = (double)nn * dre / den, cim = -(double)nn * dim / den;
            double sn, cs;
            sincos((double)nn * phi, &sn, &cs);              // exp(-i n phi) = cs - i sn
            const double tre = cre * jn, tim = cim * jn;
            ure += tre * cs + tim * sn;
            uim += tim * cs - tre * sn;
        }
        const double k2 = a.k1 * a.k1;
        const double ere = df_dv * ure / k2, eim = df_dv * uim / k2;
        double sb, cb;
        sincos(beta * sin(phi), &sb, &cb);
        // -i (cb + i sb) (ere + i eim)
        const double pre = cb * ere - sb * eim, pim = cb * eim + sb * ere;
        vre = pim;
        vim = -pre;
    }
    // layout (ex, ix, eu, iu, ev, iv) = (gx, gu, gv) with gx = ex n + ix,
    // gu = eu n + iu, gv = ev n + iv
    const int NXg = (a.nx + 2) * a.n;
    const int64_t sx = (int64_t)NUg * NVg;
    const int64_t vo = (int64_t)gu * NVg + gv;
    for (int gx = 0; gx < NXg; ++gx) {
        double val = a.pert_only ? 0.0 : ring;
        if (a.perturbation) {
            double sk, ck;
            sincos(a.k1 * xnodes[gx], &sk, &ck);
            const double p = 2.0 * a.delta_n * (ck * vre - sk * vim);
            val = a.pert_only ? p : val + p;
        }
        f[(int64_t)gx * sx + vo] = val;
    }
}

}  // namespace vl
