// qcp_hooks.cpp -- host builds of the device QCP routines, exported for CPU
// tests of the algebra.  Built into the diagnostics library
// (libbioshell_diag.so), not the product library.
#include "prep_device.hpp"
#include "qcp.hpp"

extern "C" {

double bsi_qcp_rmsd_f64(const double* M, double Ga, double Gb, int n)
{
    return bs::qcp_rmsd<double>(M, Ga, Gb, n);
}

// Rotation (row-major) that moves b onto a, from M = sum b a^T, as the
// preparation kernels compute it.
void bsi_rotation_fast(const double* M, double* R) { bs::rotation_fast(M, R); }

// The same with Newton started from an upper bound of lambda_max (E0).
void bsi_rotation_fast_from(const double* M, double upper, double* R) { bs::rotation_fast(M, R, upper); }

float bsi_qcp_rmsd_f32(const float* M, float Ga, float Gb, int n)
{
    return bs::qcp_rmsd<float>(M, Ga, Gb, n);
}

}  // extern "C"

extern "C" {

// The secular-form solve as the Gram epilogue runs it: out[0] = 2 (H - mu),
// out[1] = H, out[2] = steps taken (1 or 2), out[3] = converged (1/0; an
// unconverged pair is re-evaluated in FP64 on the device).
void bsi_qcp_secular(const double* E, const double* ci, const double* cj, double tol2L, double* out)
{
    int steps = 0;
    const bs::PairSolve o = bs::qcp_secular(E, ci, cj, tol2L, &steps);
    out[0] = o.rmsd2_L;
    out[1] = o.H;
    out[2] = steps;
    out[3] = o.converged ? 1.0 : 0.0;
}

}  // extern "C"
