/*
 * beta_oracle.c -- serial CPU restatement of BETA's FE hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for libbeta_hip.so.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function names the reference routine (file:line, paths relative to
 * the reference root) whose arithmetic and loop order it follows.  Where the
 * reference's headers are missing the assumption taken is stated inline.
 */
#include "beta_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ======================================================================
 * Quadrature
 * ====================================================================== */

/* GAUSS::GAUSS, beta/math/gaussQuadrature/gauss.cpp:14-60.
 * Newton iteration on the Legendre polynomial P_k, tolerance 3e-15,
 * ascending abscissae. */
void orc_gauss(int k, double *coords, double *weights)
{
    const double EPS = 3e-15, xm = 0.0, xl = 1.0;
    int m = (k + 1) / 2;
    for (int i = 0; i < m; i++) {
        double z = cos(M_PI * ((double)i + 1 - .25) / ((double)k + .5));
        double z1, P1, P2, P3, PP;
        do {
            P1 = 1;
            P2 = 0;
            for (int j = 0; j < k; j++) {
                P3 = P2;
                P2 = P1;
                P1 = ((2 * j + 1) * z * P2 - (j)*P3) / (j + 1);
            }
            PP = k * (z * P1 - P2) / (z * z - 1);
            z1 = z;
            z = z1 - P1 / PP;
        } while (fabs(z - z1) > EPS);
        coords[i] = xm - xl * z;
        coords[k - i - 1] = xm + xl * z;
        weights[i] = 2 * xl / ((1 - z * z) * PP * PP);
        weights[k - i - 1] = weights[i];
    }
}

/* HexQuadPoints::getQuadraturePointsForHex, gauss.cpp:303-365.
 * ip = (i*k + j)*k + l with i over xi (outer), j over eta, l over zeta. */
int orc_hex_quad_points(int order, double *xi, double *eta, double *zeta, double *w)
{
    double c[16], wt[16];
    if (order < 1 || order > 16) return ORC_ERR_ARG;
    orc_gauss(order, c, wt);
    int ip = -1;
    for (int i = 0; i < order; i++)
        for (int j = 0; j < order; j++)
            for (int k = 0; k < order; k++) {
                ip++;
                xi[ip] = c[i];
                eta[ip] = c[j];
                zeta[ip] = c[k];
                w[ip] = wt[i] * wt[j] * wt[k];
            }
    return ip + 1;
}

/* QuadQuadPoints::getQuadraturePointsForQuad, gauss.cpp:825-877. */
int orc_quad_quad_points(int order, double *xi, double *eta, double *w)
{
    double c[16], wt[16];
    if (order < 1 || order > 16) return ORC_ERR_ARG;
    orc_gauss(order, c, wt);
    int ip = -1;
    for (int i = 0; i < order; i++)
        for (int k = 0; k < order; k++) {
            ip++;
            xi[ip] = c[i];
            eta[ip] = c[k];
            w[ip] = wt[i] * wt[k];
        }
    return ip + 1;
}

/* ======================================================================
 * Shape functions, beta/elements/shape.cpp
 * ====================================================================== */

/* shapeHex8, shape.cpp:210-224.  Node order (xi,eta,zeta):
 * (---),(-+-),(-++),(--+),(+--),(++-),(+++),(+-+). */
void orc_shape_hex8(double *S, double xi, double eta, double zeta)
{
    double t1 = xi * eta, t2 = xi * zeta, t3 = eta * zeta, t4 = t1 * zeta;
    S[0] = 1.0 / 8.0 - xi / 8 - eta / 8 + t1 / 8 - zeta / 8 + t2 / 8 + t3 / 8 - t4 / 8;
    S[1] = 1.0 / 8.0 - xi / 8 + eta / 8 - t1 / 8 - zeta / 8 + t2 / 8 - t3 / 8 + t4 / 8;
    S[2] = 1.0 / 8.0 - xi / 8 + eta / 8 - t1 / 8 + zeta / 8 - t2 / 8 + t3 / 8 - t4 / 8;
    S[3] = 1.0 / 8.0 - xi / 8 - eta / 8 + t1 / 8 + zeta / 8 - t2 / 8 - t3 / 8 + t4 / 8;
    S[4] = 1.0 / 8.0 + xi / 8 - eta / 8 - t1 / 8 - zeta / 8 - t2 / 8 + t3 / 8 + t4 / 8;
    S[5] = 1.0 / 8.0 + xi / 8 + eta / 8 + t1 / 8 - zeta / 8 - t2 / 8 - t3 / 8 - t4 / 8;
    S[6] = 1.0 / 8.0 + xi / 8 + eta / 8 + t1 / 8 + zeta / 8 + t2 / 8 + t3 / 8 + t4 / 8;
    S[7] = 1.0 / 8.0 + xi / 8 - eta / 8 - t1 / 8 + zeta / 8 + t2 / 8 - t3 / 8 - t4 / 8;
}

/* derivHex8, shape.cpp:226-260. */
void orc_deriv_hex8(double *Sxi, double *Seta, double *Szeta, double xi, double eta, double zeta)
{
    double t1 = eta * zeta;
    Sxi[0] = -1.0 / 8.0 + eta / 8 + zeta / 8 - t1 / 8;
    Sxi[1] = -1.0 / 8.0 - eta / 8 + zeta / 8 + t1 / 8;
    Sxi[2] = -1.0 / 8.0 - eta / 8 - zeta / 8 - t1 / 8;
    Sxi[3] = -1.0 / 8.0 + eta / 8 - zeta / 8 + t1 / 8;
    Sxi[4] = 1.0 / 8.0 - eta / 8 - zeta / 8 + t1 / 8;
    Sxi[5] = 1.0 / 8.0 + eta / 8 - zeta / 8 - t1 / 8;
    Sxi[6] = 1.0 / 8.0 + eta / 8 + zeta / 8 + t1 / 8;
    Sxi[7] = 1.0 / 8.0 - eta / 8 + zeta / 8 - t1 / 8;
    t1 = xi * zeta;
    Seta[0] = -1.0 / 8.0 + xi / 8 + zeta / 8 - t1 / 8;
    Seta[1] = 1.0 / 8.0 - xi / 8 - zeta / 8 + t1 / 8;
    Seta[2] = 1.0 / 8.0 - xi / 8 + zeta / 8 - t1 / 8;
    Seta[3] = -1.0 / 8.0 + xi / 8 - zeta / 8 + t1 / 8;
    Seta[4] = -1.0 / 8.0 - xi / 8 + zeta / 8 + t1 / 8;
    Seta[5] = 1.0 / 8.0 + xi / 8 - zeta / 8 - t1 / 8;
    Seta[6] = 1.0 / 8.0 + xi / 8 + zeta / 8 + t1 / 8;
    Seta[7] = -1.0 / 8.0 - xi / 8 - zeta / 8 - t1 / 8;
    t1 = xi * eta;
    Szeta[0] = -1.0 / 8.0 + xi / 8 + eta / 8 - t1 / 8;
    Szeta[1] = -1.0 / 8.0 + xi / 8 - eta / 8 + t1 / 8;
    Szeta[2] = 1.0 / 8.0 - xi / 8 + eta / 8 - t1 / 8;
    Szeta[3] = 1.0 / 8.0 - xi / 8 - eta / 8 + t1 / 8;
    Szeta[4] = -1.0 / 8.0 - xi / 8 + eta / 8 + t1 / 8;
    Szeta[5] = -1.0 / 8.0 - xi / 8 - eta / 8 - t1 / 8;
    Szeta[6] = 1.0 / 8.0 + xi / 8 + eta / 8 + t1 / 8;
    Szeta[7] = 1.0 / 8.0 + xi / 8 - eta / 8 - t1 / 8;
}

/* shapeQuad4 / derivQuad4, shape.cpp:147-168 (CCW from (-1,-1)). */
void orc_shape_quad4(double *S, double xi, double eta)
{
    S[0] = .25 * (1 - xi) * (1 - eta);
    S[1] = .25 * (1 + xi) * (1 - eta);
    S[2] = .25 * (1 + xi) * (1 + eta);
    S[3] = .25 * (1 - xi) * (1 + eta);
}

void orc_deriv_quad4(double *Sxi, double *Seta, double xi, double eta)
{
    Sxi[0] = -.25 * (1 - eta);
    Sxi[1] = .25 * (1 - eta);
    Sxi[2] = .25 * (1 + eta);
    Sxi[3] = -.25 * (1 + eta);
    Seta[0] = -.25 * (1 - xi);
    Seta[1] = -.25 * (1 + xi);
    Seta[2] = .25 * (1 + xi);
    Seta[3] = .25 * (1 - xi);
}

/* ======================================================================
 * Jacobian, IsoElement::jacobian, beta/elements/isoElement.cpp:250-389
 * ====================================================================== */

/* 3-direction branch, isoElement.cpp:324-372. */
double orc_jacobian3d(int n, const double *dxi, const double *deta, const double *dzeta,
                      const double *x, const double *y, const double *z,
                      double *dx, double *dy, double *dz)
{
    double px1, px2, px3, py1, py2, py3, pz1, pz2, pz3;
    px1 = px2 = px3 = py1 = py2 = py3 = pz1 = pz2 = pz3 = 0.;
    for (int i = 0; i < n; i++) {
        px1 += dxi[i] * x[i];
        px2 += deta[i] * x[i];
        px3 += dzeta[i] * x[i];
        py1 += dxi[i] * y[i];
        py2 += deta[i] * y[i];
        py3 += dzeta[i] * y[i];
        pz1 += dxi[i] * z[i];
        pz2 += deta[i] * z[i];
        pz3 += dzeta[i] * z[i];
    }
    double dett1 = py2 * pz3 - py3 * pz2;
    double dett2 = px2 * pz3 - px3 * pz2;
    double dett3 = px2 * py3 - px3 * py2;
    double dett4 = py1 * pz3 - py3 * pz1;
    double dett5 = px1 * pz3 - px3 * pz1;
    double dett6 = px1 * py3 - px3 * py1;
    double dett7 = py1 * pz2 - py2 * pz1;
    double dett8 = px1 * pz2 - px2 * pz1;
    double dett9 = px1 * py2 - px2 * py1;
    double detJ = px1 * dett1 - py1 * dett2 + pz1 * dett3;
    double inv = 1. / detJ;
    double a11 = dett1 * inv, a12 = -dett4 * inv, a13 = dett7 * inv;
    double a21 = -dett2 * inv, a22 = dett5 * inv, a23 = -dett8 * inv;
    double a31 = dett3 * inv, a32 = -dett6 * inv, a33 = dett9 * inv;
    for (int j = 0; j < n; j++) {
        dx[j] = a11 * dxi[j] + a12 * deta[j] + a13 * dzeta[j];
        dy[j] = a21 * dxi[j] + a22 * deta[j] + a23 * dzeta[j];
        dz[j] = a31 * dxi[j] + a32 * deta[j] + a33 * dzeta[j];
    }
    return detJ;
}

/* 2-direction branch, isoElement.cpp:290-322. */
double orc_jacobian2d(int n, const double *dxi, const double *deta,
                      const double *x, const double *y, double *dx, double *dy)
{
    double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
    for (int i = 0; i < n; i++) {
        J00 += dxi[i] * x[i];
        J01 += dxi[i] * y[i];
        J10 += deta[i] * x[i];
        J11 += deta[i] * y[i];
    }
    double detJ = J00 * J11 - J01 * J10;
    double t = 1 / detJ;
    double i00 = J11 * t, i01 = -J01 * t, i10 = -J10 * t, i11 = J00 * t;
    for (int i = 0; i < n; i++) {
        dx[i] = i00 * dxi[i] + i01 * deta[i];
        dy[i] = i10 * dxi[i] + i11 * deta[i];
    }
    return detJ;
}

/* ======================================================================
 * Dense LU, Matrix::ludcmp / lubksb / invert, beta/math/matrix.cpp:272-430
 * (no skipped equations: invert() clears skip[] first, matrix.cpp:354).
 * a is n x n row-major.
 * ====================================================================== */
int orc_ludcmp(int n, double *a, int *indx, double *d)
{
    const double TINY = 1.0e-20;
    double *vv = (double *)malloc(sizeof(double) * n);
    int imax = 0, last_eqn = n - 1;
    *d = 1.0;
    for (int i = 0; i < n; i++) {
        double big = 0.0, temp;
        for (int j = 0; j < n; j++)
            if ((temp = fabs(a[i * n + j])) > big) big = temp;
        if (big == 0.0) {
            free(vv);
            return ORC_ERR_ARG; /* reference: "Singular matrix!! Exiting." */
        }
        vv[i] = 1.0 / big;
    }
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < j; i++) {
            double sum = a[i * n + j];
            for (int k = 0; k < i; k++) sum -= a[i * n + k] * a[k * n + j];
            a[i * n + j] = sum;
        }
        double big = 0.0, dum;
        for (int i = j; i < n; i++) {
            double sum = a[i * n + j];
            for (int k = 0; k < j; k++) sum -= a[i * n + k] * a[k * n + j];
            a[i * n + j] = sum;
            if ((dum = vv[i] * fabs(sum)) >= big) {
                big = dum;
                imax = i;
            }
        }
        if (j != imax) {
            for (int k = 0; k < n; k++) {
                dum = a[imax * n + k];
                a[imax * n + k] = a[j * n + k];
                a[j * n + k] = dum;
            }
            *d = -(*d);
            vv[imax] = vv[j];
        }
        indx[j] = imax;
        if (a[j * n + j] == 0.0) a[j * n + j] = TINY;
        if (j != last_eqn) {
            dum = 1.0 / a[j * n + j];
            for (int i = j + 1; i < n; i++) a[i * n + j] *= dum;
        }
    }
    free(vv);
    return ORC_OK;
}

void orc_lubksb(int n, const double *a, const int *indx, double *b)
{
    int ii = -1;
    for (int i = 0; i < n; i++) {
        int ip = indx[i];
        double sum = b[ip];
        b[ip] = b[i];
        if (ii >= 0)
            for (int j = ii; j < i; j++) sum -= a[i * n + j] * b[j];
        else if (sum)
            ii = i;
        b[i] = sum;
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = b[i];
        for (int j = i + 1; j < n; j++) sum -= a[i * n + j] * b[j];
        b[i] = sum / a[i * n + i];
    }
}

/* Matrix::invert, matrix.cpp:350-370: columns of the identity are solved as
 * rows B[i] and the result is transposed back. */
int orc_invert(int n, double *a)
{
    int *indx = (int *)malloc(sizeof(int) * n);
    double d;
    int st = orc_ludcmp(n, a, indx, &d);
    if (st != ORC_OK) {
        free(indx);
        return st;
    }
    double *B = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; i++) B[i * n + i] = 1;
    for (int i = 0; i < n; i++) orc_lubksb(n, a, indx, B + (size_t)i * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) a[i * n + j] = B[j * n + i];
    free(B);
    free(indx);
    return ORC_OK;
}

/* ======================================================================
 * Materials
 * ====================================================================== */

/* ElasticMaterial::read "readModuli", beta/materials/3D/elastic.cpp:335-376:
 * compliance S from (E11 E22 E33 nu12 nu23 nu13 G12 G23 G13), C = S^-1 by
 * NR-LU.  Voigt order xx yy zz xy yz xz with engineering shear. */
int orc_elastic_cmat(const double *E, double *C)
{
    memset(C, 0, sizeof(double) * 36);
    C[0 * 6 + 0] = 1 / E[0];
    C[0 * 6 + 1] = C[1 * 6 + 0] = -E[3] / E[0];
    C[0 * 6 + 2] = C[2 * 6 + 0] = -E[5] / E[0];
    C[1 * 6 + 1] = 1 / E[1];
    C[2 * 6 + 1] = C[1 * 6 + 2] = -E[4] / E[1];
    C[2 * 6 + 2] = 1 / E[2];
    C[3 * 6 + 3] = 1 / E[6];
    C[4 * 6 + 4] = 1 / E[7];
    C[5 * 6 + 5] = 1 / E[8];
    return orc_invert(6, C);
}

/* Isotropic input to readModuli: E11=E22=E33=E, nu12=nu23=nu13=nu,
 * G12=G23=G13=E/(2(1+nu)). */
int orc_elastic_iso_cmat(double E, double nu, double *C)
{
    double G = E / (2 * (1 + nu));
    double m[9] = {E, E, E, nu, nu, nu, G, G, G};
    return orc_elastic_cmat(m, C);
}

/* PlaneStress::read, beta/materials/2D/pStress.cpp:68-85. */
void orc_plane_stress_cmat(double e11, double e22, double nu12, double G12, double *C)
{
    double s11 = 1 / e11, s12 = -nu12 / e11, s22 = 1. / e22, s66 = 1. / G12;
    double den = s11 * s22 - s12 * s12;
    C[0] = s22 / den;
    C[1] = -s12 / den;
    C[2] = 0.;
    C[3] = -s12 / den;
    C[4] = s11 / den;
    C[5] = 0.;
    C[6] = 0.;
    C[7] = 0.;
    C[8] = 1. / s66;
}

/* ======================================================================
 * Element stiffness
 * ====================================================================== */

/* ElasticityElement3D: calculateBmatrix (ElasticityElement3D.cpp:57-87),
 * addToKe (560-641), initializeArrays (643-651), loopOverIntegrationPoints
 * (983-1030), checkForNegativeDiagonal (BasicElement.cpp:116-130).
 * getIntegrationFactor() is taken as 1.0 for 3D (header missing).
 * Rotation is the identity (isotropic, unrotated materials).
 * Returns ORC_NEG_DETJ if any detJ < 0 (geometryCheck,
 * ElasticityElement3D.cpp:653-670), ORC_NEG_DIAG for Ke(i,i) <= 0. */
int orc_ke_hex8_elastic(const double *xyz, const double *C, int order, double *Ke)
{
    enum { NN = 8, ND = 24 };
    double gxi[64], geta[64], gzeta[64], gw[64];
    int nip = orc_hex_quad_points(order, gxi, geta, gzeta, gw);
    if (nip < 0) return ORC_ERR_ARG;
    double x[NN], y[NN], z[NN];
    for (int a = 0; a < NN; a++) {
        x[a] = xyz[3 * a];
        y[a] = xyz[3 * a + 1];
        z[a] = xyz[3 * a + 2];
    }
    for (int i = 0; i < ND; i++)
        for (int j = i; j < ND; j++) Ke[i * ND + j] = 0.;
    int bad = 0;
    for (int ip = 0; ip < nip; ip++) {
        double S[NN], dxi[NN], deta[NN], dzeta[NN], dx[NN], dy[NN], dz[NN];
        orc_shape_hex8(S, gxi[ip], geta[ip], gzeta[ip]);
        orc_deriv_hex8(dxi, deta, dzeta, gxi[ip], geta[ip], gzeta[ip]);
        double detJ = orc_jacobian3d(NN, dxi, deta, dzeta, x, y, z, dx, dy, dz);
        if (detJ < 0) bad = 1;
        double f = detJ * gw[ip] * 1.0;
        double b[6][ND];
        for (int i = 0, c1 = 0, c2 = 1, c3 = 2; i < NN; i++, c1 += 3, c2 += 3, c3 += 3) {
            b[0][c1] = dx[i]; b[0][c2] = 0.;    b[0][c3] = 0.;
            b[1][c1] = 0.;    b[1][c2] = dy[i]; b[1][c3] = 0.;
            b[2][c1] = 0.;    b[2][c2] = 0.;    b[2][c3] = dz[i];
            b[3][c1] = dy[i]; b[3][c2] = dx[i]; b[3][c3] = 0.;
            b[4][c1] = 0.;    b[4][c2] = dz[i]; b[4][c3] = dy[i];
            b[5][c1] = dz[i]; b[5][c2] = 0.;    b[5][c3] = dx[i];
        }
        double CxB[6][ND];
        for (int ii = 0; ii < ND; ii++)
            for (int jj = 0; jj < 6; jj++)
                CxB[jj][ii] = b[0][ii] * C[0 * 6 + jj] + b[1][ii] * C[1 * 6 + jj] +
                              b[2][ii] * C[2 * 6 + jj] + b[3][ii] * C[3 * 6 + jj] +
                              b[4][ii] * C[4 * 6 + jj] + b[5][ii] * C[5 * 6 + jj];
        for (int ii = 0; ii < ND; ii++)
            for (int jj = ii; jj < ND; jj++)
                Ke[ii * ND + jj] += (CxB[0][ii] * b[0][jj] + CxB[1][ii] * b[1][jj] +
                                     CxB[2][ii] * b[2][jj] + CxB[3][ii] * b[3][jj] +
                                     CxB[4][ii] * b[4][jj] + CxB[5][ii] * b[5][jj]) * f;
    }
    if (bad) return ORC_NEG_DETJ;
    for (int i = 0; i < ND; i++)
        if (Ke[i * ND + i] <= 0) return ORC_NEG_DIAG;
    return ORC_OK;
}

/* Internal force of a hex8 for displacement ue (node-major u,v,w):
 * Fe = sum_ip B^T (C B ue) detJ w  (ElasticityElement3D::addToFe,
 * ElasticityElement3D.cpp:41-52, with calculateStress = C*strain,
 * elastic.cpp:411-451, no hygrothermal strain). */
int orc_fe_hex8_elastic(const double *xyz, const double *C, int order, const double *ue,
                        double *Fe)
{
    enum { NN = 8, ND = 24 };
    double gxi[64], geta[64], gzeta[64], gw[64];
    int nip = orc_hex_quad_points(order, gxi, geta, gzeta, gw);
    if (nip < 0) return ORC_ERR_ARG;
    double x[NN], y[NN], z[NN];
    for (int a = 0; a < NN; a++) {
        x[a] = xyz[3 * a];
        y[a] = xyz[3 * a + 1];
        z[a] = xyz[3 * a + 2];
    }
    for (int i = 0; i < ND; i++) Fe[i] = 0;
    for (int ip = 0; ip < nip; ip++) {
        double dxi[NN], deta[NN], dzeta[NN], dx[NN], dy[NN], dz[NN];
        orc_deriv_hex8(dxi, deta, dzeta, gxi[ip], geta[ip], gzeta[ip]);
        double detJ = orc_jacobian3d(NN, dxi, deta, dzeta, x, y, z, dx, dy, dz);
        double f = detJ * gw[ip];
        double e[6] = {0, 0, 0, 0, 0, 0};
        for (int a = 0; a < NN; a++) {
            double u = ue[3 * a], v = ue[3 * a + 1], w = ue[3 * a + 2];
            e[0] += dx[a] * u;
            e[1] += dy[a] * v;
            e[2] += dz[a] * w;
            e[3] += dy[a] * u + dx[a] * v;
            e[4] += dz[a] * v + dy[a] * w;
            e[5] += dz[a] * u + dx[a] * w;
        }
        double s[6];
        for (int i = 0; i < 6; i++) {
            s[i] = 0;
            for (int j = 0; j < 6; j++) s[i] += C[i * 6 + j] * e[j];
        }
        for (int a = 0; a < NN; a++) {
            Fe[3 * a + 0] += (dx[a] * s[0] + dy[a] * s[3] + dz[a] * s[5]) * f;
            Fe[3 * a + 1] += (dy[a] * s[1] + dx[a] * s[3] + dz[a] * s[4]) * f;
            Fe[3 * a + 2] += (dz[a] * s[2] + dy[a] * s[4] + dx[a] * s[5]) * f;
        }
    }
    return ORC_OK;
}

/* ElasticityElement2D::calculateBmatrix (ElasticityElement2D.cpp:88-104) and
 * addToKe (22-86); factor = getIntegrationFactor() (thickness; header missing,
 * the caller passes it). */
int orc_ke_quad4_elastic(const double *xy, const double *C, int order, double factor, double *Ke)
{
    enum { NN = 4, ND = 8 };
    double gxi[64], geta[64], gw[64];
    int nip = orc_quad_quad_points(order, gxi, geta, gw);
    if (nip < 0) return ORC_ERR_ARG;
    double x[NN], y[NN];
    for (int a = 0; a < NN; a++) {
        x[a] = xy[2 * a];
        y[a] = xy[2 * a + 1];
    }
    for (int i = 0; i < ND; i++)
        for (int j = i; j < ND; j++) Ke[i * ND + j] = 0.;
    int bad = 0;
    for (int ip = 0; ip < nip; ip++) {
        double dxi[NN], deta[NN], dx[NN], dy[NN];
        orc_deriv_quad4(dxi, deta, gxi[ip], geta[ip]);
        double detJ = orc_jacobian2d(NN, dxi, deta, x, y, dx, dy);
        if (detJ < 0) bad = 1;
        double f = detJ * gw[ip] * factor;
        double b[3][ND];
        for (int i = 0, c1 = 0, c2 = 1; i < NN; i++, c1 += 2, c2 += 2) {
            b[0][c1] = dx[i]; b[0][c2] = 0.;
            b[1][c1] = 0.;    b[1][c2] = dy[i];
            b[2][c1] = dy[i]; b[2][c2] = dx[i];
        }
        double CxB[3][ND];
        for (int ii = 0; ii < ND; ii++)
            for (int jj = 0; jj < 3; jj++)
                CxB[jj][ii] = b[0][ii] * C[0 * 3 + jj] + b[1][ii] * C[1 * 3 + jj] +
                              b[2][ii] * C[2 * 3 + jj];
        for (int ii = 0; ii < ND; ii++)
            for (int jj = ii; jj < ND; jj++)
                Ke[ii * ND + jj] += (CxB[0][ii] * b[0][jj] + CxB[1][ii] * b[1][jj] +
                                     CxB[2][ii] * b[2][jj]) * f;
    }
    if (bad) return ORC_NEG_DETJ;
    for (int i = 0; i < ND; i++)
        if (Ke[i * ND + i] <= 0) return ORC_NEG_DIAG;
    return ORC_OK;
}

/* HeatTransferElement3Dtransient::addToKe (HeatTransferElement3DTransient.cpp:
 * 153-197) and addToMe (60-77), loopOverIntegrationPoints (665-704).
 * k9 is the 3x3 conductivity (readConductivity fills the diagonal,
 * heatTransferMaterial.cpp:89-92); capacity is getSaturationValue().
 * The loop zeroes the full Ke/Me (681-686) but only the upper triangle is
 * accumulated. */
int orc_ke_hex8_heat(const double *xyz, const double *k9, double capacity, int order,
                     double *Ke, double *Me)
{
    enum { NN = 8 };
    double gxi[64], geta[64], gzeta[64], gw[64];
    int nip = orc_hex_quad_points(order, gxi, geta, gzeta, gw);
    if (nip < 0) return ORC_ERR_ARG;
    double x[NN], y[NN], z[NN];
    for (int a = 0; a < NN; a++) {
        x[a] = xyz[3 * a];
        y[a] = xyz[3 * a + 1];
        z[a] = xyz[3 * a + 2];
    }
    for (int i = 0; i < NN * NN; i++) Ke[i] = Me[i] = 0.;
    double k11 = k9[0], k22 = k9[4], k33 = k9[8], k12 = k9[1], k13 = k9[2], k23 = k9[5];
    int bad = 0;
    for (int ip = 0; ip < nip; ip++) {
        double S[NN], dxi[NN], deta[NN], dzeta[NN], dx[NN], dy[NN], dz[NN];
        orc_shape_hex8(S, gxi[ip], geta[ip], gzeta[ip]);
        orc_deriv_hex8(dxi, deta, dzeta, gxi[ip], geta[ip], gzeta[ip]);
        double detJ = orc_jacobian3d(NN, dxi, deta, dzeta, x, y, z, dx, dy, dz);
        if (detJ < 0) bad = 1;
        double f = detJ * gw[ip] * 1.0;
        for (int i = 0; i < NN; i++) {
            double a1 = dx[i], a2 = dy[i], a3 = dz[i];
            for (int j = i; j < NN; j++)
                Ke[i * NN + j] += ((dx[j] * k11 + dy[j] * k12 + dz[j] * k13) * a1 +
                                   (dx[j] * k12 + dy[j] * k22 + dz[j] * k23) * a2 +
                                   (dx[j] * k13 + dy[j] * k23 + dz[j] * k33) * a3) * f;
        }
        for (int i = 0; i < NN; i++) {
            double Si = S[i];
            for (int j = i; j < NN; j++) Me[i * NN + j] += Si * S[j] * f * capacity;
        }
    }
    return bad ? ORC_NEG_DETJ : ORC_OK;
}

/* ======================================================================
 * Equations + SymmetricSparseMatrix
 * ====================================================================== */

struct orc_eq {
    int n;
    int **col;       /* per-row sorted column list (SortedList<int>) */
    int *len, *cap;
    double **val;    /* per-row values, allocated by allocate() */
    unsigned char *rest;
    double *solution, *load, *pseudo;
    double zero_diag_replacement;
    int allocated;
};

orc_eq *orc_eq_create(int n)
{
    orc_eq *eq = (orc_eq *)calloc(1, sizeof(orc_eq));
    eq->n = n;
    eq->col = (int **)calloc(n, sizeof(int *));
    eq->len = (int *)calloc(n, sizeof(int));
    eq->cap = (int *)calloc(n, sizeof(int));
    eq->val = (double **)calloc(n, sizeof(double *));
    eq->rest = (unsigned char *)calloc(n, 1);
    eq->solution = (double *)calloc(n, sizeof(double));
    eq->load = (double *)calloc(n, sizeof(double));
    eq->pseudo = (double *)calloc(n, sizeof(double));
    return eq;
}

void orc_eq_destroy(orc_eq *eq)
{
    if (!eq) return;
    for (int i = 0; i < eq->n; i++) {
        free(eq->col[i]);
        free(eq->val[i]);
    }
    free(eq->col); free(eq->len); free(eq->cap); free(eq->val);
    free(eq->rest); free(eq->solution); free(eq->load); free(eq->pseudo);
    free(eq);
}

/* Equations::setRestraint + setSolution, equation.cpp:306-341. */
void orc_eq_set_restraint(orc_eq *eq, int eqn, double value)
{
    eq->rest[eqn] = 1;
    eq->solution[eqn] = value;
}

/* SortedList<int>::add: unique sorted insertion (header missing; the
 * reference relies on it to hold sorted unique columns, sparse.cpp:144). */
static void row_add(orc_eq *eq, int I, int J)
{
    int lo = 0, hi = eq->len[I];
    int *c = eq->col[I];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (c[mid] < J) lo = mid + 1;
        else hi = mid;
    }
    if (lo < eq->len[I] && c[lo] == J) return;
    if (eq->len[I] == eq->cap[I]) {
        eq->cap[I] = eq->cap[I] ? 2 * eq->cap[I] : 8;
        eq->col[I] = (int *)realloc(eq->col[I], sizeof(int) * eq->cap[I]);
        c = eq->col[I];
    }
    memmove(c + lo + 1, c + lo, sizeof(int) * (eq->len[I] - lo));
    c[lo] = J;
    eq->len[I]++;
}

/* Equations::specifyNonZeroLocation, equation.cpp:1216-1267 (no additional
 * equations): J restrained -> (I,I) when I==J; I restrained, J free -> none;
 * both free and I>=J -> (I,J). */
void orc_eq_specify(orc_eq *eq, const int *map, int size)
{
    for (int i = 0; i < size; i++) {
        int I = map[i];
        for (int j = 0; j < size; j++) {
            int J = map[j];
            if (eq->rest[J]) {
                if (I == J) row_add(eq, I, I);
            } else if (eq->rest[I]) {
                /* additional equations only */
            } else if (I >= J) {
                row_add(eq, I, J);
            }
        }
    }
}

/* SymmetricSparseMatrix::allocate, sparse.cpp:482-495: empty rows get
 * their diagonal. */
void orc_eq_allocate(orc_eq *eq)
{
    for (int i = 0; i < eq->n; i++) {
        if (eq->len[i] == 0) row_add(eq, i, i);
        eq->val[i] = (double *)calloc(eq->len[i], sizeof(double));
    }
    eq->allocated = 1;
}

long long orc_eq_nnz(const orc_eq *eq)
{
    long long s = 0;
    for (int i = 0; i < eq->n; i++) s += eq->len[i];
    return s;
}

void orc_eq_get_pattern(const orc_eq *eq, int *rowptr, int *colind)
{
    long long p = 0;
    rowptr[0] = 0;
    for (int i = 0; i < eq->n; i++) {
        if (colind) memcpy(colind + p, eq->col[i], sizeof(int) * eq->len[i]);
        p += eq->len[i];
        rowptr[i + 1] = (int)p;
    }
}

void orc_eq_get_values(const orc_eq *eq, double *v)
{
    long long p = 0;
    for (int i = 0; i < eq->n; i++) {
        memcpy(v + p, eq->val[i], sizeof(double) * eq->len[i]);
        p += eq->len[i];
    }
}

void orc_eq_set_values(orc_eq *eq, const double *v)
{
    long long p = 0;
    for (int i = 0; i < eq->n; i++) {
        memcpy(eq->val[i], v + p, sizeof(double) * eq->len[i]);
        p += eq->len[i];
    }
}

/* Equations::zeroMatrix: zeroes K and the pseudo-load vector
 * (SerialAssembler.cpp:39 comment; sparse.cpp:244-249). */
void orc_eq_zero(orc_eq *eq)
{
    for (int i = 0; i < eq->n; i++) {
        memset(eq->val[i], 0, sizeof(double) * eq->len[i]);
        eq->pseudo[i] = 0;
    }
}

/* SymmetricSparseMatrix::operator(), sparse.cpp:80-99: binary search. */
static double *entry(orc_eq *eq, int I, int J)
{
    int lo = 0, hi = eq->len[I];
    const int *c = eq->col[I];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (c[mid] < J) lo = mid + 1;
        else hi = mid;
    }
    if (lo == eq->len[I] || c[lo] != J) return NULL;
    return &eq->val[I][lo];
}

/* Equations::addMatrix, equation.cpp:445-562 (no additional equations).
 * Ke is size x size row-major with the upper triangle valid. */
int orc_eq_add_matrix(orc_eq *eq, const double *Ke, const int *map, int size, const double *force)
{
    const unsigned char *rest = eq->rest;
    double *F = eq->pseudo;
    for (int i = 0; i < size; i++) {
        int I = map[i];
        if (force != 0 && !rest[I]) F[I] += force[i];
        for (int j = 0; j < size; j++) {
            double m = (i < j) ? Ke[i * size + j] : Ke[j * size + i];
            int J = map[j];
            if (rest[J]) {
                if (I == J) {
                    double *e = entry(eq, I, I);
                    if (!e) return ORC_ERR_ARG;
                    *e = 1;
                }
                if (!rest[I]) F[I] -= m * eq->solution[J];
            } else if (rest[I]) {
                /* additional equations only */
            } else if (I >= J) {
                double *e = entry(eq, I, J);
                if (!e) return ORC_ERR_ARG;
                *e += m;
            }
        }
    }
    return ORC_OK;
}

/* Equations::addLoadToEqn, equation.cpp:851-868 (load on a restrained DOF is
 * a fatal error there; here it is reported). */
int orc_eq_add_load(orc_eq *eq, int eqn, double load)
{
    if (eq->rest[eqn]) return ORC_ERR_ARG;
    eq->load[eqn] += load;
    return ORC_OK;
}

/* Equations::solve RHS build, equation.cpp:1324-1342. */
void orc_eq_build_rhs(const orc_eq *eq, double *b)
{
    for (int i = 0; i < eq->n; i++)
        b[i] = eq->rest[i] ? eq->solution[i] : eq->load[i] + eq->pseudo[i];
}

void orc_eq_get_pseudo(const orc_eq *eq, double *f)
{
    memcpy(f, eq->pseudo, sizeof(double) * eq->n);
}

void orc_eq_set_replace_zero_diag(orc_eq *eq, double value)
{
    eq->zero_diag_replacement = value;
}

/* SymmetricSparseMatrix::vectorProduct, sparse.cpp:452-478. */
void orc_eq_spmv(const orc_eq *eq, const double *x, double *y)
{
    const int L = eq->n;
    for (int i = 0; i < L; i++) y[i] = 0;
    for (int i = 0; i < L; i++) {
        int nt = eq->len[i] - 1;
        if (nt >= 0) {
            const int *rc = eq->col[i];
            const double *m = eq->val[i];
            double xi = x[i], t = 0;
            for (int j = 0; j < nt; j++) {
                t += m[j] * x[rc[j]];
                y[rc[j]] += m[j] * xi;
            }
            y[i] += m[nt] * x[rc[nt]] + t;
            if (rc[nt] != i) y[rc[nt]] += m[nt] * xi;
        }
    }
}

/* LargeMatrix::CheckAndFixZeroDiagonal, largeMatrix.cpp:43-79. */
int orc_eq_check_fix_diag(orc_eq *eq)
{
    int zero = 0;
    for (int i = 0; i < eq->n; i++) {
        double *d = entry(eq, i, i);
        double v = d ? *d : 0.0;
        if (v == 0) { zero = 1; break; }
        if (v < 0) return ORC_NEG_DIAG;
    }
    if (zero) {
        if (eq->zero_diag_replacement == 0) return ORC_ZERO_DIAG;
        for (int i = 0; i < eq->n; i++) {
            double *d = entry(eq, i, i);
            if (d && *d == 0) *d = eq->zero_diag_replacement;
        }
    }
    return ORC_OK;
}

/* SymmetricSparseMatrix::solve, sparse.cpp:114-242: Jacobi-PCG.
 * b holds the RHS on entry and the solution on exit.  stop_rule 0 is the
 * reference rule sqrt(|r|^2/|x|^2) < eps (sparse.cpp:223); stop_rule 1 is
 * |r|/|b| <= eps on the recursive residual.  Non-convergence is reported
 * when the iteration count reaches maxit (sparse.cpp:225-229, which exits). */
int orc_eq_pcg(orc_eq *eq, double *b, const double *x0, double eps, int maxit,
               int stop_rule, int *iters, double *final_norm)
{
    const int L = eq->n;
    int st = orc_eq_check_fix_diag(eq);
    if (st != ORC_OK) return st;
    double *x = (double *)malloc(sizeof(double) * L);
    double *r = (double *)malloc(sizeof(double) * L);
    double *diag = (double *)malloc(sizeof(double) * L);
    for (int i = 0; i < L; i++) diag[i] = eq->val[i][eq->len[i] - 1];
    if (x0 == 0) for (int i = 0; i < L; i++) x[i] = 0;
    else for (int i = 0; i < L; i++) x[i] = x0[i];
    orc_eq_spmv(eq, x, r);
    free(x);
    double *p = (double *)malloc(sizeof(double) * L);
    double *q = (double *)malloc(sizeof(double) * L);
    double *temp = (double *)malloc(sizeof(double) * L);
    double bnorm = 0;
    for (int i = 0; i < L; i++) bnorm += b[i] * b[i];
    for (int i = 0; i < L; i++) r[i] = b[i] - r[i];
    for (int i = 0; i < L; i++) p[i] = q[i] = r[i] / diag[i];
    int it = 0, conv = 0;
    double rnorm = 0, xnorm = 0;
    if (x0 == 0) for (int i = 0; i < L; i++) b[i] = 0;
    else for (int i = 0; i < L; i++) b[i] = x0[i];
    while (!conv && it < maxit) {
        double pAp = 0, rq = 0;
        orc_eq_spmv(eq, p, temp);
        for (int i = 0; i < L; i++) {
            pAp += p[i] * temp[i];
            rq += r[i] * q[i];
        }
        double a = rq / pAp;
        for (int i = 0; i < L; i++) b[i] += a * p[i];
        for (int i = 0; i < L; i++) r[i] -= a * temp[i];
        for (int i = 0; i < L; i++) q[i] = r[i] / diag[i];
        double newrq = 0;
        for (int i = 0; i < L; i++) newrq += r[i] * q[i];
        double bk = newrq / rq;
        for (int i = 0; i < L; i++) p[i] = q[i] + bk * p[i];
        rnorm = 0;
        for (int i = 0; i < L; i++) rnorm += r[i] * r[i];
        xnorm = 0;
        for (int i = 0; i < L; i++) xnorm += b[i] * b[i];
        it++;
        double nrm = stop_rule == 0 ? sqrt(rnorm / xnorm) : sqrt(rnorm / bnorm);
        if (stop_rule == 0 ? (nrm < eps) : (nrm <= eps)) conv = 1;
        if (final_norm) *final_norm = nrm;
    }
    free(q); free(p); free(temp); free(r); free(diag);
    if (iters) *iters = it;
    if (it == maxit) return ORC_NOT_CONVERGED;
    return ORC_OK;
}

/* ParCGMatrix::vectorProduct (ParCG.cpp:504-616): the symmetric product from
 * lower rows split at L1 = L/2 over two threads.  Section 1: thread 0 takes
 * rows < L1 (all their columns), thread 1 rows >= L1 with columns >= L1;
 * section 2: thread 0 scatters the transposed products of rows >= L1 into
 * columns < L1, thread 1 adds the rows' own products with those columns.
 * No two threads write one entry within a section. */
static void parcg_spmv(const orc_eq *eq, const double *v, double *res)
{
    const int L = eq->n, L1 = L / 2;
    for (int i = 0; i < L; i++) res[i] = 0;
#pragma omp parallel num_threads(2)
    {
        const int t = omp_get_thread_num();
        const int i0 = t == 0 ? 0 : L1, i1 = t == 0 ? L1 : L;
        for (int i = i0; i < i1; i++) {
            const int nt = eq->len[i] - 1;
            if (nt < 0) continue;
            const double vi = v[i];
            const double *m = eq->val[i];
            const int *c = eq->col[i];
            double acc = 0;
            for (int j = 0; j < nt; j++) {
                if (t == 1 && c[j] < L1) continue;
                acc += m[j] * v[c[j]];
                res[c[j]] += m[j] * vi;
            }
            res[i] += m[nt] * v[c[nt]] + acc;
        }
    }
#pragma omp parallel num_threads(2)
    {
        const int t = omp_get_thread_num();
        for (int i = L1; i < L; i++) {
            const int nt = eq->len[i] - 1;
            if (nt < 0) continue;
            const double vi = v[i];
            const double *m = eq->val[i];
            const int *c = eq->col[i];
            double acc = 0;
            for (int j = 0; j < nt; j++) {
                if (c[j] >= L1) continue;
                if (t == 0) res[c[j]] += m[j] * vi;
                else acc += m[j] * v[c[j]];
            }
            if (t == 1) res[i] += acc;
        }
    }
}

/* ParCGMatrix::solve (ParCG.cpp:93-347): the Jacobi-PCG of sparse.cpp with
 * its vector loops fused into three passes, each split at L1 = L/2 over two
 * OpenMP threads with (+) reductions, and the ||r||/||x|| stop rule.  The
 * CPU baseline's two-thread row; same iterates as orc_eq_pcg up to the
 * order of the partial sums. */
int orc_eq_pcg_parcg(orc_eq *eq, double *b, const double *x0, double eps, int maxit, int *iters,
                     double *final_norm)
{
    const int L = eq->n, L1 = L / 2;
    int st = orc_eq_check_fix_diag(eq);
    if (st != ORC_OK) return st;
    double *x = (double *)malloc(sizeof(double) * L);
    double *r = (double *)malloc(sizeof(double) * L);
    double *diag = (double *)malloc(sizeof(double) * L);
    for (int i = 0; i < L; i++) diag[i] = eq->val[i][eq->len[i] - 1];
    if (x0 == 0) for (int i = 0; i < L; i++) x[i] = 0;
    else for (int i = 0; i < L; i++) x[i] = x0[i];
    parcg_spmv(eq, x, r);
    free(x);
    double *p = (double *)malloc(sizeof(double) * L);
    double *q = (double *)malloc(sizeof(double) * L);
    double *temp = (double *)malloc(sizeof(double) * L);
    for (int i = 0; i < L; i++) {
        r[i] = b[i] - r[i];
        p[i] = q[i] = r[i] / diag[i];
    }
    if (x0 == 0) for (int i = 0; i < L; i++) b[i] = 0;
    else for (int i = 0; i < L; i++) b[i] = x0[i];
    int it = 0, conv = 0;
    while (!conv && it < maxit) {
        parcg_spmv(eq, p, temp);
        double pAp = 0, rq = 0;
#pragma omp parallel num_threads(2) reduction(+ : pAp, rq)
        {
            const int t = omp_get_thread_num();
            for (int i = t == 0 ? 0 : L1; i < (t == 0 ? L1 : L); i++) {
                pAp = pAp + p[i] * temp[i];
                rq = rq + r[i] * q[i];
            }
        }
        const double a = rq / pAp;
        double newrq = 0;
#pragma omp parallel num_threads(2) reduction(+ : newrq)
        {
            const int t = omp_get_thread_num();
            for (int i = t == 0 ? 0 : L1; i < (t == 0 ? L1 : L); i++) {
                b[i] += a * p[i];
                r[i] -= a * temp[i];
                q[i] = r[i] / diag[i];
                newrq = newrq + r[i] * q[i];
            }
        }
        const double bk = newrq / rq;
        double rnorm = 0, xnorm = 0;
#pragma omp parallel num_threads(2) reduction(+ : rnorm, xnorm)
        {
            const int t = omp_get_thread_num();
            for (int i = t == 0 ? 0 : L1; i < (t == 0 ? L1 : L); i++) {
                p[i] = q[i] + bk * p[i];
                rnorm = rnorm + r[i] * r[i];
                xnorm = xnorm + b[i] * b[i];
            }
        }
        it++;
        const double nrm = sqrt(rnorm / xnorm);
        if (nrm < eps) conv = 1;
        if (final_norm) *final_norm = nrm;
    }
    free(q); free(p); free(temp); free(r); free(diag);
    if (iters) *iters = it;
    if (it == maxit && !conv) return ORC_NOT_CONVERGED;
    return ORC_OK;
}

/* ProfileMatrix::solve (beta/math/equation/promat.cpp:111-204) is a skyline
 * LDL^T; restated here on the lower-triangle storage expanded to a dense
 * symmetric matrix, for small systems only (direct-solve oracle). */
int orc_eq_skyline_solve(const orc_eq *eq, double *b)
{
    const int n = eq->n;
    double *A = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; i++)
        for (int k = 0; k < eq->len[i]; k++) {
            int j = eq->col[i][k];
            A[(size_t)i * n + j] = eq->val[i][k];
            A[(size_t)j * n + i] = eq->val[i][k];
        }
    /* LDL^T, column by column */
    for (int j = 0; j < n; j++) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; k++)
            d -= A[(size_t)j * n + k] * A[(size_t)j * n + k] * A[(size_t)k * n + k];
        if (d == 0) { free(A); return ORC_ZERO_DIAG; }
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; i++) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++)
                s -= A[(size_t)i * n + k] * A[(size_t)j * n + k] * A[(size_t)k * n + k];
            A[(size_t)i * n + j] = s / d;
        }
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < i; k++) b[i] -= A[(size_t)i * n + k] * b[k];
    for (int i = 0; i < n; i++) b[i] /= A[(size_t)i * n + i];
    for (int i = n - 1; i >= 0; i--)
        for (int k = i + 1; k < n; k++) b[i] -= A[(size_t)k * n + i] * b[k];
    free(A);
    return ORC_OK;
}

/* ======================================================================
 * Element DOF map and serial assembler drivers
 * ====================================================================== */

/* BasicElement::calculateDofList, BasicElement.cpp:49-68. */
int orc_dof_list(int nnpe, const int *conn, const int *dof_first, const int *ndof, int *map)
{
    int k = 0;
    for (int i = 0; i < nnpe; i++)
        for (int j = 0; j < ndof[conn[i]]; j++) map[k++] = dof_first[conn[i]] + j;
    return k;
}

/* BasicModel::allocateEquationsMatrix, BasicModel.cpp:1022-1033, followed by
 * Equations::allocate (equation.cpp:1269-1280). */
void orc_specify_mesh(orc_eq *eq, int n_el, int nnpe, const int *conn,
                      const int *dof_first, const int *ndof)
{
    int map[256];
    for (int e = 0; e < n_el; e++) {
        int size = orc_dof_list(nnpe, conn + (size_t)e * nnpe, dof_first, ndof, map);
        orc_eq_specify(eq, map, size);
    }
    orc_eq_allocate(eq);
}

/* SerialAssembler::assembleK, SerialAssembler.cpp:35-59: zero, then per
 * element in order: update("K"), addMatrix(Ke, map, size, NULL).
 * cmats: 36 doubles per material group; elem_group NULL -> group = element. */
int orc_assemble_elastic3d(orc_eq *eq, int n_el, const int *conn, const double *xyz,
                           const int *dof_first, const int *ndof,
                           const double *cmats, const int *elem_group, int order)
{
    double Ke[24 * 24], ex[24];
    int map[24];
    orc_eq_zero(eq);
    for (int e = 0; e < n_el; e++) {
        const int *c = conn + (size_t)e * 8;
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) ex[3 * a + d] = xyz[3 * (size_t)c[a] + d];
        int g = elem_group ? elem_group[e] : e;
        int st = orc_ke_hex8_elastic(ex, cmats + (size_t)g * 36, order, Ke);
        if (st != ORC_OK) return st;
        int size = orc_dof_list(8, c, dof_first, ndof, map);
        st = orc_eq_add_matrix(eq, Ke, map, size, NULL);
        if (st != ORC_OK) return st;
    }
    return ORC_OK;
}

int orc_assemble_elastic2d(orc_eq *eq, int n_el, const int *conn, const double *xy,
                           const int *dof_first, const int *ndof,
                           const double *cmats, const int *elem_group, int order,
                           double thickness)
{
    double Ke[8 * 8], ex[8];
    int map[8];
    orc_eq_zero(eq);
    for (int e = 0; e < n_el; e++) {
        const int *c = conn + (size_t)e * 4;
        for (int a = 0; a < 4; a++)
            for (int d = 0; d < 2; d++) ex[2 * a + d] = xy[2 * (size_t)c[a] + d];
        int g = elem_group ? elem_group[e] : e;
        int st = orc_ke_quad4_elastic(ex, cmats + (size_t)g * 9, order, thickness, Ke);
        if (st != ORC_OK) return st;
        int size = orc_dof_list(4, c, dof_first, ndof, map);
        st = orc_eq_add_matrix(eq, Ke, map, size, NULL);
        if (st != ORC_OK) return st;
    }
    return ORC_OK;
}

/* DiffusionModel::assembleK_F_s_alpha_family, DiffusionModel.cpp:322-347,
 * with formK_s (HeatTransferElement3DTransient.cpp:323-350):
 *   Mbar_e = (a1*Ke + Me)/dt over the full 8x8 (Ke, Me upper-only, zero
 *   below, so the scatter's i<j selection reads the upper triangle),
 * and formF_s_new (372-405): Fbar_e = -(a1+a2) * Ke_sym * u_s / dt,
 * a1 = alpha*dt, a2 = (1-alpha)*dt (DiffusionModel.cpp:745-746).
 * kcap: 4 doubles per group (k11 k22 k33 capacity). */
int orc_assemble_heat_transient(orc_eq *eq, int n_el, const int *conn, const double *xyz,
                                const int *dof_first, const int *ndof,
                                const double *kcap, const int *elem_group, int order,
                                double alpha, double dt, const double *u_s)
{
    double Ke[64], Me[64], Mb[64], Fe[8], ex[24];
    int map[8];
    double a1 = alpha * dt, a2 = (1 - alpha) * dt;
    orc_eq_zero(eq);
    for (int e = 0; e < n_el; e++) {
        const int *c = conn + (size_t)e * 8;
        for (int a = 0; a < 8; a++)
            for (int d = 0; d < 3; d++) ex[3 * a + d] = xyz[3 * (size_t)c[a] + d];
        int g = elem_group ? elem_group[e] : e;
        const double *kc = kcap + (size_t)g * 4;
        double k9[9] = {kc[0], 0, 0, 0, kc[1], 0, 0, 0, kc[2]};
        int st = orc_ke_hex8_heat(ex, k9, kc[3], order, Ke, Me);
        if (st != ORC_OK) return st;
        int size = orc_dof_list(8, c, dof_first, ndof, map);
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) Mb[i * 8 + j] = (a1 * Ke[i * 8 + j] + Me[i * 8 + j]) / dt;
        for (int i = 0; i < 8; i++) {
            double dump = 0.0;
            for (int j = 0; j < 8; j++) {
                double Ku = (j >= i) ? Ke[i * 8 + j] * u_s[map[j]] : Ke[j * 8 + i] * u_s[map[j]];
                dump += -(a1 + a2) * Ku;
            }
            Fe[i] = dump / dt;
        }
        st = orc_eq_add_matrix(eq, Mb, map, size, Fe);
        if (st != ORC_OK) return st;
    }
    return ORC_OK;
}

/* ======================================================================
 * J2 plasticity: PlasticMaterial::getStressIncrement ("version 2.0",
 * beta_ext/plasticity/material/plastic.cpp:368-465) with the Hill function
 * (172-194), yieldFunction (216-229), m_getYieldStress (197-208), dYield
 * (716-735), hardeningModulus (685-690), m_dsdep/m_Stress for the POWER
 * master curve (612-660).  J2 is Hill with F=G=H=1/3, L=M=N=1; linear
 * isotropic hardening is the POWER curve with m_E=H, m_n=1, m_sigma0=sigma0.
 * Associated flow (dPlastic == dYield), numHardParam = 1, originState and
 * yStress zero (never set on the active code path).
 * Inputs: stress = converged stress (GaussISV::stress), eps_bar = converged
 * effective plastic strain, dstrain = strain increment since the converged
 * state.  Outputs: dstress, deps_bar, Cep (getTangentStiffness, 583-610),
 * plastic flag (inPlasticRegion).  Returns the return-map iteration count.
 * ====================================================================== */
typedef struct {
    double C[36];
    double s0[6], ds[6];
    double ep, dep;
    const orc_j2_params *p;
} j2_state;

static double j2_hill(const j2_state *g)
{
    const double F = 1.0 / 3.0, G = 1.0 / 3.0, H = 1.0 / 3.0, L = 1, M = 1, N = 1;
    double s[6];
    for (int i = 0; i < 6; i++) s[i] = (g->s0[i] - 0.0 + g->ds[i]) - 0.0;
    return 0.5 * (F * (s[1] - s[2]) * (s[1] - s[2]) + G * (s[2] - s[0]) * (s[2] - s[0]) +
                  H * (s[0] - s[1]) * (s[0] - s[1]) + 2 * L * s[4] * s[4] +
                  2 * M * s[5] * s[5] + 2 * N * s[3] * s[3]);
}

static double j2_mstress(const j2_state *g)
{
    double ep = g->ep + g->dep;
    return g->p->H * pow(ep, 1.0) + g->p->sigma0;
}

static double j2_dsdep(const j2_state *g)
{
    double ep = g->ep + g->dep;
    if (fabs(ep) < 1e-12) return g->p->H;
    return g->p->H * pow(ep, 1.0 - 1.0);
}

static double j2_yield_stress(const j2_state *g)
{
    double y = j2_mstress(g);
    return (y > g->p->sigma0) ? y : g->p->sigma0;
}

static double j2_yield_fn(const j2_state *g)
{
    double sy = j2_yield_stress(g);
    return j2_hill(g) / (sy * sy / 3.0) - 1.0;
}

static double j2_eff_stress(const j2_state *g) { return sqrt(3.0 * j2_hill(g)); }

static double j2_hard(const j2_state *g)
{
    double mm = j2_eff_stress(g);
    return 4.0 / 9.0 * mm * mm * j2_dsdep(g);
}

/* dYield is linear in the stress: nf = P s.  j2_flow applies it to any
 * 6-vector (the derivative of the loop below applies it to columns). */
static void j2_flow(const double *s, double *nf)
{
    const double F = 1.0 / 3.0, G = 1.0 / 3.0, H = 1.0 / 3.0, L = 1, M = 1, N = 1;
    nf[0] = -G * (s[2] - s[0]) + H * (s[0] - s[1]);
    nf[1] = F * (s[1] - s[2]) - H * (s[0] - s[1]);
    nf[2] = -F * (s[1] - s[2]) + G * (s[2] - s[0]);
    nf[3] = 2 * N * s[3];
    nf[4] = 2 * L * s[4];
    nf[5] = 2 * M * s[5];
}

static void j2_dyield(const j2_state *g, double *nf)
{
    const double F = 1.0 / 3.0, G = 1.0 / 3.0, H = 1.0 / 3.0, L = 1, M = 1, N = 1;
    double s[6];
    for (int i = 0; i < 6; i++) s[i] = (g->s0[i] - 0.0 + g->ds[i]) - 0.0;
    nf[0] = -G * (s[2] - s[0]) + H * (s[0] - s[1]);
    nf[1] = F * (s[1] - s[2]) - H * (s[0] - s[1]);
    nf[2] = -F * (s[1] - s[2]) + G * (s[2] - s[0]);
    nf[3] = 2 * N * s[3];
    nf[4] = 2 * L * s[4];
    nf[5] = 2 * M * s[5];
}

static void j2_cmul(const double *C, const double *v, double *out)
{
    for (int i = 0; i < 6; i++) {
        out[i] = 0;
        for (int j = 0; j < 6; j++) out[i] += C[i * 6 + j] * v[j];
    }
}

static double dot6(const double *a, const double *b)
{
    double s = 0;
    for (int i = 0; i < 6; i++) s += a[i] * b[i];
    return s;
}

/* The return map with, when Dalg is given, its exact derivative
 * Dalg = d(dstress)/d(dstrain) (row-major 6x6, engineering-shear strain in,
 * stress out): forward-mode differentiation through the iterations the loop
 * actually executes (the same branch, the same iteration count).  Per pass,
 * with s = s0 + ds and D = d(ds)/d(dstrain):
 *   nf = P s, sp = C nf, hard = (4/3) H hill(s)  (d hill/ds = nf, m_dsdep = H)
 *   dl = nf.dse / (hard + nf.sp)
 *   d nf = P D, d num = dse^T P D + nf^T C, d den = (4/3) H nf^T D + 2 sp^T P D
 *   ds' = dse - dl sp,  D' = C - sp (d dl) - dl C P D.
 * This is not the reference's tangent (getTangentStiffness, the continuum
 * C_ep); it is what a Newton iteration on the reference's stress update needs
 * for quadratic convergence.  It is not symmetric in general. */
static int j2_increment_impl(const orc_j2_params *p, const double *stress, double eps_bar,
                             const double *dstrain, double *dstress, double *deps_bar,
                             double *Cep, double *Dalg, int *plastic);

int orc_j2_stress_increment(const orc_j2_params *p, const double *stress, double eps_bar,
                            const double *dstrain, double *dstress, double *deps_bar,
                            double *Cep, int *plastic)
{
    return j2_increment_impl(p, stress, eps_bar, dstrain, dstress, deps_bar, Cep, NULL, plastic);
}

int orc_j2_stress_increment_alg(const orc_j2_params *p, const double *stress, double eps_bar,
                                const double *dstrain, double *dstress, double *deps_bar,
                                double *Cep, double *Dalg, int *plastic)
{
    return j2_increment_impl(p, stress, eps_bar, dstrain, dstress, deps_bar, Cep, Dalg, plastic);
}

static int j2_increment_impl(const orc_j2_params *p, const double *stress, double eps_bar,
                             const double *dstrain, double *dstress, double *deps_bar,
                             double *Cep, double *Dalg, int *plastic)
{
    j2_state g;
    g.p = p;
    orc_elastic_iso_cmat(p->E, p->nu, g.C);
    for (int i = 0; i < 6; i++) g.s0[i] = stress[i];
    g.ep = eps_bar;
    g.dep = 0.0;
    double dse[6], nf[6], sp[6], se[6];
    j2_cmul(g.C, dstrain, dse);
    for (int i = 0; i < 6; i++) g.ds[i] = dse[i];
    memcpy(Cep, g.C, sizeof(double) * 36);
    double D[36];  /* d(ds)/d(dstrain), columns j */
    memcpy(D, g.C, sizeof(D));
    int iter = 0;
    *plastic = 0;
    if (j2_yield_fn(&g) > 0.0) {
        *plastic = 1;
        double denom;
        do {
            double hard = j2_hard(&g);
            j2_dyield(&g, nf);
            j2_cmul(g.C, nf, sp);
            double num = dot6(nf, dse), den = hard + dot6(nf, sp);
            double dl = num / den;
            if (Dalg) {
                const double Hh = p->H; /* m_dsdep of the POWER curve with m_n = 1 */
                double PD[36], CPD[36], Cnf[6];
                for (int j = 0; j < 6; j++) {
                    double col[6], pc[6], cpc[6];
                    for (int i = 0; i < 6; i++) col[i] = D[i * 6 + j];
                    j2_flow(col, pc);
                    j2_cmul(g.C, pc, cpc);
                    for (int i = 0; i < 6; i++) {
                        PD[i * 6 + j] = pc[i];
                        CPD[i * 6 + j] = cpc[i];
                    }
                }
                j2_cmul(g.C, nf, Cnf);
                double ddl[6];
                for (int j = 0; j < 6; j++) {
                    double dnum = 0, dden = 0;
                    for (int i = 0; i < 6; i++) {
                        dnum += dse[i] * PD[i * 6 + j] + nf[i] * g.C[i * 6 + j];
                        dden += 4.0 / 3.0 * Hh * nf[i] * D[i * 6 + j] + 2.0 * sp[i] * PD[i * 6 + j];
                    }
                    ddl[j] = (dnum - dl * dden) / den;
                }
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) D[i * 6 + j] = g.C[i * 6 + j] - sp[i] * ddl[j] - dl * CPD[i * 6 + j];
            }
            for (int i = 0; i < 6; i++) g.ds[i] = dse[i] - dl * sp[i];
            g.dep = 2.0 * dl * j2_eff_stress(&g) / 3.0;
            denom = j2_yield_fn(&g);
            iter++;
        } while ((denom > 1e-6) && (iter < 5000));
        j2_dyield(&g, nf);
        j2_cmul(g.C, nf, sp);
        j2_cmul(g.C, nf, se);
        denom = j2_hard(&g) + dot6(nf, sp);
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++) Cep[i * 6 + j] = g.C[i * 6 + j] - sp[i] * se[j] / denom;
    }
    for (int i = 0; i < 6; i++) dstress[i] = g.ds[i];
    *deps_bar = g.dep;
    if (Dalg) memcpy(Dalg, D, sizeof(D));
    return iter;
}

/* ======================================================================
 * Plastic hex8 element: PlasticElement::update (plasticElement.cpp:39-250)
 * ====================================================================== */

/* Global gradients of all 8 nodes at Gauss point ip; returns detJ*w. */
static double hex8_ip_grads(const double *xyz, int order, int ip, double *dx, double *dy, double *dz)
{
    double gxi[64], geta[64], gzeta[64], gw[64];
    orc_hex_quad_points(order, gxi, geta, gzeta, gw);
    double x[8], y[8], z[8], dxi[8], deta[8], dzeta[8];
    for (int a = 0; a < 8; a++) {
        x[a] = xyz[3 * a];
        y[a] = xyz[3 * a + 1];
        z[a] = xyz[3 * a + 2];
    }
    orc_deriv_hex8(dxi, deta, dzeta, gxi[ip], geta[ip], gzeta[ip]);
    double detJ = orc_jacobian3d(8, dxi, deta, dzeta, x, y, z, dx, dy, dz);
    return detJ * gw[ip];
}

/* "V": strain increment from the element increment displacement
 * (ElasticityElement3D::calculateStrain, engineering shear) and
 * PlasticMaterial::updateISV -> getStressIncrement at every Gauss point. */
int orc_plastic_update(const double *xyz, const orc_j2_params *p, int order, const double *ue, double *isv)
{
    return orc_plastic_update_tangent(xyz, p, order, ue, isv, ORC_TANGENT_CONTINUUM);
}

/* As orc_plastic_update; tangent = ORC_TANGENT_ALGORITHMIC stores the
 * symmetric part (D + D^T)/2 of the return map's exact derivative in the
 * record's tangent slot instead of the reference's continuum C_ep (the
 * assembled stiffness must stay symmetric for PCG; the two agree for
 * proportional loading, where D is symmetric). */
int orc_plastic_update_tangent(const double *xyz, const orc_j2_params *p, int order, const double *ue,
                               double *isv, int tangent)
{
    int nip = order * order * order;
    for (int ip = 0; ip < nip; ip++) {
        double dx[8], dy[8], dz[8];
        hex8_ip_grads(xyz, order, ip, dx, dy, dz);
        double de[6] = {0, 0, 0, 0, 0, 0};
        for (int a = 0; a < 8; a++) {
            double u = ue[3 * a], v = ue[3 * a + 1], w = ue[3 * a + 2];
            de[0] += dx[a] * u;
            de[1] += dy[a] * v;
            de[2] += dz[a] * w;
            de[3] += dy[a] * u + dx[a] * v;
            de[4] += dz[a] * v + dy[a] * w;
            de[5] += dz[a] * u + dx[a] * w;
        }
        double *r = isv + (size_t)ip * ORC_PISV;
        double dep;
        int pl;
        double D[36];
        int it = orc_j2_stress_increment_alg(p, r, r[6], de, r + 7, &dep, r + 16, D, &pl);
        if (tangent == ORC_TANGENT_ALGORITHMIC)
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) r[16 + i * 6 + j] = 0.5 * (D[i * 6 + j] + D[j * 6 + i]);
        r[13] = dep;
        r[14] = pl;
        r[15] = it;
    }
    return ORC_OK;
}

/* "KI": Ke = sum_ip B^T Cep B detJ w with the tangent of the last return map
 * (getTangentStiffness, plastic.cpp:583-610; the elastic C where the point is
 * not plastic); "F": Fe = sum_ip B^T dstress detJ w (addToFe with
 * isv.dStress).  Ke is 24x24 row-major, upper triangle. */
int orc_plastic_ke_fe(const double *xyz, const orc_j2_params *p, int order, const double *isv, double *Ke,
                      double *Fe)
{
    int nip = order * order * order;
    for (int i = 0; i < 24 * 24; i++) Ke[i] = 0;
    for (int i = 0; i < 24; i++) Fe[i] = 0;
    for (int ip = 0; ip < nip; ip++) {
        double dx[8], dy[8], dz[8];
        double f = hex8_ip_grads(xyz, order, ip, dx, dy, dz);
        const double *r = isv + (size_t)ip * ORC_PISV;
        const double *C = r + 16;
        double b[6][24];
        for (int i = 0, c1 = 0, c2 = 1, c3 = 2; i < 8; i++, c1 += 3, c2 += 3, c3 += 3) {
            b[0][c1] = dx[i]; b[0][c2] = 0.;    b[0][c3] = 0.;
            b[1][c1] = 0.;    b[1][c2] = dy[i]; b[1][c3] = 0.;
            b[2][c1] = 0.;    b[2][c2] = 0.;    b[2][c3] = dz[i];
            b[3][c1] = dy[i]; b[3][c2] = dx[i]; b[3][c3] = 0.;
            b[4][c1] = 0.;    b[4][c2] = dz[i]; b[4][c3] = dy[i];
            b[5][c1] = dz[i]; b[5][c2] = 0.;    b[5][c3] = dx[i];
        }
        double CxB[6][24];
        for (int ii = 0; ii < 24; ii++)
            for (int jj = 0; jj < 6; jj++) {
                double s = 0;
                for (int t = 0; t < 6; t++) s += b[t][ii] * C[t * 6 + jj];
                CxB[jj][ii] = s;
            }
        for (int ii = 0; ii < 24; ii++)
            for (int jj = ii; jj < 24; jj++) {
                double s = 0;
                for (int t = 0; t < 6; t++) s += CxB[t][ii] * b[t][jj];
                Ke[ii * 24 + jj] += s * f;
            }
        const double *ds = r + 7;
        for (int i = 0; i < 24; i++) {
            double s = 0;
            for (int t = 0; t < 6; t++) s += b[t][i] * ds[t];
            Fe[i] += s * f;
        }
    }
    return ORC_OK;
}

/* "U": GaussISV::updateISV (plastic.cpp:769-785). */
void orc_plastic_commit(double *isv, int nip)
{
    for (int ip = 0; ip < nip; ip++) {
        double *r = isv + (size_t)ip * ORC_PISV;
        for (int i = 0; i < 6; i++) {
            r[i] += r[7 + i];
            r[7 + i] = 0;
        }
        r[6] += r[13];
        r[13] = 0;
    }
}
