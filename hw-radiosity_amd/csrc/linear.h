// Linear radiosity on an explicit form-factor matrix (the gathering solvers).
//
//   B = E + diag(rho) F B        per right-hand side (column) c
//
// F (n x n, row i = receiver) is built on the device from a scene
// (DeviceEngine::form_factors) or given by the caller.  Solvers:
//   gather   Jacobi sweeps with F stored in FP16 or BF16 and multiplied on
//            MFMA with FP32 accumulation, inside an iterative refinement
//            whose residual E + rho o (F x) - x and energy sums are FP64;
//   shoot    plain FP32 progressive shooting (one source per step, the
//            largest unshot power), the reference method on the same system;
//   adjoint  the transposed system (I - diag(rho) F)^T lambda = g in FP32
//            with FP64 refinement, and the gradients dL/dE = lambda,
//            dL/drho = lambda o (F B).
#pragma once

#include <string>

namespace rgpu {

enum Precision { PREC_F32 = 0, PREC_F16 = 1, PREC_BF16 = 2 };

// info[] slots written by the solvers
enum {
	INFO_OUTER = 0,       // refinement steps (gather/adjoint) or shots (shoot)
	INFO_SWEEPS = 1,      // low-precision sweeps
	INFO_REL_RES = 2,     // final FP64 relative residual ||r|| / ||rhs||
	INFO_MS = 3,          // device time of the solve
	INFO_MS_TARGET = 4,   // time when the error against the reference first met the target (-1: never/no reference)
	INFO_ERR = 5,         // final relative L2 error against the reference (-1: no reference)
	INFO_POWER_E = 6,     // sum_i area_i E_i (FP64), over all columns
	INFO_POWER_B = 7,     // sum_i area_i B_i (FP64)
	INFO_N = 8
};

class LinearSystem {
public:
	~LinearSystem();
	static LinearSystem *create(int n, std::string &err);
	int n() const { return n_; }
	int ld() const { return np_; }
	float *matrix32();  // device, row-major, leading dimension ld(), zero padding
	int set_matrix(const float *F, std::string &err);
	int get_matrix(float *F, std::string &err);
	int set_areas(const double *area, std::string &err);  // for the power sums and shooting's selection
	int autotune(int nrhs, int prec, std::string &err);
	// average device time of one low-precision product F X (reps launches)
	int gemm_ms(int nrhs, int prec, int reps, double *ms, std::string &err);
	// Y = F X with the stored matrix: prec FP32 multiplies in FP64 (exact
	// products of the FP32 entries), FP16/BF16 use the MFMA path (X rounded
	// to that precision, FP32 accumulation)
	int multiply(int prec, int nrhs, const double *X, double *Y, std::string &err);
	void *stream() const;  // the solver's HIP stream (hipStream_t)
	int gather(int prec, int nrhs, const double *rho, const double *E, double tol, int max_outer, int sweeps,
	           const double *Bref, double target, double *B, double *info, std::string &err);
	int shoot(int nrhs, const double *rho, const double *E, const double *Bref, double target, int max_shots,
	          double *B, double *info, std::string &err);
	int adjoint(int nrhs, const double *rho, const double *B, const double *g, double tol, int max_outer, int sweeps,
	            double *lambda, double *drho, double *info, std::string &err);
	// the tile configuration in use for (nrhs, prec): rows per wave (16 x rt),
	// split-K factor (0, 0 before autotune or set_tile)
	void tile_config(int nrhs, int prec, int *rt, int *splitk) const;
	int set_tile(int nrhs, int prec, int rt, int splitk, std::string &err);

private:
	LinearSystem() = default;
	struct Impl;
	Impl *d_ = nullptr;
	int n_ = 0, np_ = 0;
};

}  // namespace rgpu
