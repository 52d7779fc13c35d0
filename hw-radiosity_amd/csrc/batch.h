// Many independent radiosity systems of uneven size solved together:
//   B_s = E_s + diag(rho_s) F_s B_s,   s = 0 .. count-1, three columns (RGB)
// by the mixed-precision gather of linear.h (FP16/BF16 Jacobi sweeps on MFMA
// inside FP64-residual iterative refinement), every step of every scene in
// one kernel: the kernels walk a tile list spanning all scenes with a grid
// sized to the machine, a scene that meets its tolerance drops out of the
// tile walk, and the whole fixed schedule (refinements x sweeps) is captured
// once into a HIP graph and replayed with a single launch.
#pragma once

#include <string>

namespace rgpu {

// Schedule of the batched sweep product: each wave owns 16*rt rows (a tile
// is 64*rt rows of one scene), a tile spans kc of k, and grid persistent
// blocks walk the tile list.
struct BatchTile {
	int rt, kc, grid;
};

class SceneBatch {
public:
	~SceneBatch();
	static SceneBatch *create(int count, const int *n, std::string &err);
	int count() const;
	int n(int s) const;
	int ld(int s) const;
	int total() const;       // sum of n
	float *matrix32(int s);  // device, ld(s) x ld(s) row-major, zero padded
	int set_matrix(int s, const float *F, std::string &err);
	// Times one sweep for every candidate schedule on the current matrices
	// and keeps the fastest for precision prec; solve does it on first use
	// unless set_tile fixed the schedule.
	int autotune(int prec, BatchTile &out, std::string &err);
	int set_tile(const BatchTile &t, std::string &err);
	// rho, E, B: (total() x 3) row-major, scenes concatenated in order.
	// graph = true: one HIP graph launch for everything; false: the same
	// kernels launched scene after scene (the unbatched reference schedule,
	// bit-identical results).  rel_res[count] = each scene's final FP64
	// relative residual.  info[4]: device ms of the solve, ms to build the
	// graph (capture + instantiate, 0 when cached), kernel launches (graph
	// nodes), scenes still above tol.
	int solve(int prec, bool graph, const double *rho, const double *E, double tol, int outer, int sweeps, double *B,
	          double *rel_res, double *info, std::string &err);

private:
	SceneBatch() = default;
	struct Impl;
	Impl *d_ = nullptr;
};

}  // namespace rgpu
