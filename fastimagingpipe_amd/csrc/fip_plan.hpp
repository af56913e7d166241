// Host-side plan math: grid sizing, grid-correction table and FFT tables.
#pragma once

#include <cstddef>
#include <vector>

namespace fip {

// ceil(1.5 * I) evaluated as the reference does (FIPkernels.cu:31-37, 583).
size_t grid_size_for(size_t image);
// Image sizes with compiled FFT plans for both the length-g and g/2 passes.
bool image_size_supported(size_t image);
// ES kernel exp(beta (sqrt(1-x^2) - 1)) (FIPkernels.cu:67-71), FP64.
double es_kernel(double x);
// Positive half of 28-point Gauss-Legendre, nodes descending (FIPkernels.cu:14-23).
void gauss_legendre28_half(double nodes[14], double weights[14]);
// 4 * integral_{-1}^{1} ES (the constant at FIPkernels.cu:585).
double es_norm();
// c(t) for t = 0..I/2 (convolveKernel, FIPkernels.cu:173-186), FP64.
std::vector<double> conv_corr(size_t image, size_t grid);
// 1 / (c(t) * norm) as float: the per-axis correction applied by the two
// FFT passes, whose product is 1 / (c(|x|) c(|y|) norm^2) (FIPkernels.cu:284).
std::vector<float> corr_table(size_t image, size_t grid);

}  // namespace fip
