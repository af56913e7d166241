// FP64 3x3 symmetric eigen-decomposition and the PCA basis convention of the
// reference's preprocessing (SVD_MFS.py:64-75, 98-103; the sign rule of
// SVDPreprocessKernels.cu:167-206).
#pragma once

#include <cmath>

namespace fip {

// Cyclic Jacobi on a symmetric 3x3 matrix a (destroyed).  evec columns are
// the eigenvectors of eval.
inline void jacobi3(double a[3][3], double evec[3][3], double eval[3])
{
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        if (off == 0.0) break;
        for (int p = 0; p < 2; ++p) {
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
        }
    }
    for (int i = 0; i < 3; ++i) {
        eval[i] = a[i][i];
        for (int k = 0; k < 3; ++k) evec[k][i] = v[k][i];
    }
}

// From a covariance matrix: V0, V1 = eigenvectors of the two largest
// eigenvalues, each with its largest-magnitude entry made positive (first
// index on ties), and V2 = V0 x V1.  V is row-major [3][3].
inline void pca_basis(const double cov_in[3][3], double V[3][3])
{
    double a[3][3], evec[3][3], eval[3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = cov_in[i][j];
    jacobi3(a, evec, eval);
    int order[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (eval[order[j]] > eval[order[i]]) {
                const int t = order[i];
                order[i] = order[j];
                order[j] = t;
            }
    for (int c = 0; c < 2; ++c) {
        for (int k = 0; k < 3; ++k) V[c][k] = evec[k][order[c]];
        int am = 0;
        for (int k = 1; k < 3; ++k)
            if (std::fabs(V[c][k]) > std::fabs(V[c][am])) am = k;
        if (V[c][am] < 0)
            for (int k = 0; k < 3; ++k) V[c][k] = -V[c][k];
    }
    V[2][0] = V[0][1] * V[1][2] - V[0][2] * V[1][1];
    V[2][1] = V[0][2] * V[1][0] - V[0][0] * V[1][2];
    V[2][2] = V[0][0] * V[1][1] - V[0][1] * V[1][0];
}

}  // namespace fip
