// gauss_tables.hpp -- host construction of the Gauss-point tables uploaded to
// each translation unit's __constant__ memory: weights, reference-space shape
// derivatives and shape values per integration point.
//
// Gauss-Legendre points: the reference's Newton construction (gauss.cpp:14-60);
// tensor order xi outer, zeta inner (gauss.cpp:303-365, 825-877); hex8 and
// quad4 shape functions and derivatives (shape.cpp:147-260).
#pragma once

#include <cmath>
#include <vector>

namespace bhip_gauss {

inline void rule(int k, double *c, double *w) {
    const double EPS = 3e-15;
    int m = (k + 1) / 2;
    for (int i = 0; i < m; i++) {
        double z = cos(M_PI * ((double)i + 1 - .25) / ((double)k + .5)), z1, P1, P2, P3, PP;
        do {
            P1 = 1;
            P2 = 0;
            for (int j = 0; j < k; j++) {
                P3 = P2;
                P2 = P1;
                P1 = ((2 * j + 1) * z * P2 - j * P3) / (j + 1);
            }
            PP = k * (z * P1 - P2) / (z * z - 1);
            z1 = z;
            z = z1 - P1 / PP;
        } while (fabs(z - z1) > EPS);
        c[i] = -z;
        c[k - i - 1] = z;
        w[i] = 2 / ((1 - z * z) * PP * PP);
        w[k - i - 1] = w[i];
    }
}

// Fills hw[maxip], hdN[maxip*24], hN[maxip*8]; returns nip (0: unsupported).
inline int tables(int nnpe, int order, int maxip, std::vector<double> &hw, std::vector<double> &hdN,
                  std::vector<double> &hN) {
    double c[16], w[16];
    if (order < 1 || order > 16) return 0;
    rule(order, c, w);
    hw.assign(maxip, 0.0);
    hdN.assign((size_t)maxip * 24, 0.0);
    hN.assign((size_t)maxip * 8, 0.0);
    int ip = 0;
    if (nnpe == 8) {
        if (order * order * order > maxip) return 0;
        for (int i = 0; i < order; i++)
            for (int j = 0; j < order; j++)
                for (int k = 0; k < order; k++, ip++) {
                    double xi = c[i], eta = c[j], zeta = c[k];
                    hw[ip] = w[i] * w[j] * w[k];
                    double *dxi = &hdN[ip * 24], *deta = dxi + 8, *dzeta = dxi + 16;
                    double t1 = eta * zeta;
                    dxi[0] = -1.0 / 8.0 + eta / 8 + zeta / 8 - t1 / 8;
                    dxi[1] = -1.0 / 8.0 - eta / 8 + zeta / 8 + t1 / 8;
                    dxi[2] = -1.0 / 8.0 - eta / 8 - zeta / 8 - t1 / 8;
                    dxi[3] = -1.0 / 8.0 + eta / 8 - zeta / 8 + t1 / 8;
                    dxi[4] = 1.0 / 8.0 - eta / 8 - zeta / 8 + t1 / 8;
                    dxi[5] = 1.0 / 8.0 + eta / 8 - zeta / 8 - t1 / 8;
                    dxi[6] = 1.0 / 8.0 + eta / 8 + zeta / 8 + t1 / 8;
                    dxi[7] = 1.0 / 8.0 - eta / 8 + zeta / 8 - t1 / 8;
                    t1 = xi * zeta;
                    deta[0] = -1.0 / 8.0 + xi / 8 + zeta / 8 - t1 / 8;
                    deta[1] = 1.0 / 8.0 - xi / 8 - zeta / 8 + t1 / 8;
                    deta[2] = 1.0 / 8.0 - xi / 8 + zeta / 8 - t1 / 8;
                    deta[3] = -1.0 / 8.0 + xi / 8 - zeta / 8 + t1 / 8;
                    deta[4] = -1.0 / 8.0 - xi / 8 + zeta / 8 + t1 / 8;
                    deta[5] = 1.0 / 8.0 + xi / 8 - zeta / 8 - t1 / 8;
                    deta[6] = 1.0 / 8.0 + xi / 8 + zeta / 8 + t1 / 8;
                    deta[7] = -1.0 / 8.0 - xi / 8 - zeta / 8 - t1 / 8;
                    t1 = xi * eta;
                    dzeta[0] = -1.0 / 8.0 + xi / 8 + eta / 8 - t1 / 8;
                    dzeta[1] = -1.0 / 8.0 + xi / 8 - eta / 8 + t1 / 8;
                    dzeta[2] = 1.0 / 8.0 - xi / 8 + eta / 8 - t1 / 8;
                    dzeta[3] = 1.0 / 8.0 - xi / 8 - eta / 8 + t1 / 8;
                    dzeta[4] = -1.0 / 8.0 - xi / 8 + eta / 8 + t1 / 8;
                    dzeta[5] = -1.0 / 8.0 - xi / 8 - eta / 8 - t1 / 8;
                    dzeta[6] = 1.0 / 8.0 + xi / 8 + eta / 8 + t1 / 8;
                    dzeta[7] = 1.0 / 8.0 + xi / 8 - eta / 8 - t1 / 8;
                    double t2 = xi * zeta, t3 = eta * zeta, t4 = xi * eta * zeta;
                    double *S = &hN[ip * 8];
                    S[0] = 1.0 / 8.0 - xi / 8 - eta / 8 + t1 / 8 - zeta / 8 + t2 / 8 + t3 / 8 - t4 / 8;
                    S[1] = 1.0 / 8.0 - xi / 8 + eta / 8 - t1 / 8 - zeta / 8 + t2 / 8 - t3 / 8 + t4 / 8;
                    S[2] = 1.0 / 8.0 - xi / 8 + eta / 8 - t1 / 8 + zeta / 8 - t2 / 8 + t3 / 8 - t4 / 8;
                    S[3] = 1.0 / 8.0 - xi / 8 - eta / 8 + t1 / 8 + zeta / 8 - t2 / 8 - t3 / 8 + t4 / 8;
                    S[4] = 1.0 / 8.0 + xi / 8 - eta / 8 - t1 / 8 - zeta / 8 - t2 / 8 + t3 / 8 + t4 / 8;
                    S[5] = 1.0 / 8.0 + xi / 8 + eta / 8 + t1 / 8 - zeta / 8 - t2 / 8 - t3 / 8 - t4 / 8;
                    S[6] = 1.0 / 8.0 + xi / 8 + eta / 8 + t1 / 8 + zeta / 8 + t2 / 8 + t3 / 8 + t4 / 8;
                    S[7] = 1.0 / 8.0 + xi / 8 - eta / 8 - t1 / 8 + zeta / 8 + t2 / 8 - t3 / 8 - t4 / 8;
                }
    } else if (nnpe == 4) {
        if (order * order > maxip) return 0;
        for (int i = 0; i < order; i++)
            for (int k = 0; k < order; k++, ip++) {
                double xi = c[i], eta = c[k];
                hw[ip] = w[i] * w[k];
                double *dxi = &hdN[ip * 24], *deta = dxi + 4;
                dxi[0] = -.25 * (1 - eta);
                dxi[1] = .25 * (1 - eta);
                dxi[2] = .25 * (1 + eta);
                dxi[3] = -.25 * (1 + eta);
                deta[0] = -.25 * (1 - xi);
                deta[1] = -.25 * (1 + xi);
                deta[2] = .25 * (1 + xi);
                deta[3] = .25 * (1 - xi);
                double *S = &hN[ip * 8];
                S[0] = .25 * (1 - xi) * (1 - eta);
                S[1] = .25 * (1 + xi) * (1 - eta);
                S[2] = .25 * (1 + xi) * (1 + eta);
                S[3] = .25 * (1 - xi) * (1 + eta);
            }
    } else {
        return 0;
    }
    return ip;
}

}  // namespace bhip_gauss
