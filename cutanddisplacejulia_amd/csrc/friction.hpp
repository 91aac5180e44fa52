// Frictional slip (SlipCalculator3D.jl's MixedBoundaryConditionsFriction method): the linear complementarity
// problem built from the inverse collocation matrix, solved by a Fischer-Newton method on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace cdj {

// A[0:count] *= s
void scale_matrix(double* A, int64_t count, double s, hipStream_t st);

// LCP [y] = [a][x] + [b] (a: 5n x 5n, leading dimension 5n) of Kaven et al. (2012) eq. 28 as modified in
// SlipCalculator3D.jl:90-167, from C = A^-1 (3n x 3n, leading dimension 3n), D = C b (3n), the friction
// coefficients mu (n) and cohesions Sf (n); D is added to the first 3n entries of bb.
void build_friction_lcp(const double* C, int64_t n, const double* D, const double* mu, const double* Sf, double* a,
                        double* bb, hipStream_t st);

// bb alone (as build_friction_lcp): the vector for another D with the same C and Sf
void build_friction_lcp_vector(const double* C, int64_t n, const double* D, const double* Sf, double* bb,
                               hipStream_t st);

// y[0:N] += x[0:N]
void add_vector(double* y, const double* x, int64_t N, hipStream_t st);

// The LCP unknowns kept (ascending indices into 0 .. 5n - 1): for an element with mu = 0 and Sf = 0 its two traction
// parts (x2, x3; rows 3n + i and 4n + i read y = -x, so both are 0) and its two negative slip parts (x4, x5; they
// enter only its own slip rows, so y2, y3 carry the slip's sign) are removed.  The LCP on the rest has the same
// solution for every kept unknown and no degenerate pair (x = y = 0), on which Fischer-Newton converges only
// linearly.  All 5n are kept when no element is frictionless.
std::vector<int32_t> lcp_keep(const double* mu, const double* Sf, int64_t n);

// ar (Nr x Nr, leading dimension Nr) = a[keep, keep] (a: N x N, leading dimension N), br = bb[keep]
void lcp_gather(const double* a, const double* bb, int64_t N, const int32_t* keep, int64_t Nr, double* ar,
                double* br, hipStream_t st);

// x[0:N] = 0, then x[keep[k]] = xr[k]
void lcp_scatter(const double* xr, const int32_t* keep, int64_t Nr, double* x, int64_t N, hipStream_t st);

struct FnParams {
  int max_iter = 100;
  double tol_abs = 1e-26;  // on 0.5 ||phi||^2
  double alpha = 0.5;      // Armijo step reduction
  double beta = 1e-3;      // Armijo sufficient-decrease constant
  double gamma = 1e-28;    // perturbation size and smallest tau^2
  double rho = 2.220446049250313e-16;  // sufficient-descent constant (machine epsilon)
};

struct FnResult {
  int iterations = 0;
  double err = 0.0;        // final 0.5 ||phi(a x + b, x)||^2
  int gradient_steps = 0;  // Newton systems that were singular or gave no descent (gradient direction taken)
};

// x >= 0 with y = a x + b >= 0 and x'y = 0 (a: N x N, leading dimension N; x: N, output).  Newton systems are solved
// by the GPU LU (ws: lu_workspace_bytes(N); J: N x N scratch; ipiv: N ints; info: one int; vec: 8 N doubles;
// red: 8 doubles; part: gemv_part_bytes(N, N) for the products a x; iwork: fn_iwork_ints(N) ints), the
// Fischer-Burmeister merit function is reduced on the device and read back once per trial.
hipError_t fischer_newton(const double* a, const double* bb, int64_t N, double* x, const FnParams& p, FnResult& r,
                          double* J, int32_t* ipiv, int* info, void* ws, double* vec, double* red, double* part,
                          int* iwork, hipStream_t st);
inline int64_t fn_iwork_ints(int64_t N) { return N + (N + 63) / 64 + 2; }

// Slip from the LCP solution: Dn = y1 s, Dss = (y2 - x4) s, Dds = (y3 - x5) s with y = a x + b and s the scale
// applied to A (D3: 3n doubles, [Dn; Dss; Dds]; y: 5n doubles; part: gemv_part_bytes(5n, 5n)).
void friction_slip(const double* a, const double* bb, const double* x, int64_t n, double scale, double* y, double* D3,
                   double* part, hipStream_t st);

}  // namespace cdj
