// afm_internal.h -- device-side data structures and launchers shared by the
// HIP translation units of libafm.  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include "afm_delta.h"

namespace afm {

// ----------------------------------------------------------- errors ----
void set_error(const std::string& msg);
const char* last_error();

#define AFM_HIP(call)                                                                    \
	do {                                                                                 \
		hipError_t e_ = (call);                                                          \
		if (e_ != hipSuccess) {                                                          \
			::afm::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
			return -3;                                                                   \
		}                                                                                \
	} while (0)

#define AFM_LAUNCH_CHECK()                                                               \
	do {                                                                                 \
		hipError_t e_ = hipGetLastError();                                               \
		if (e_ != hipSuccess) {                                                          \
			::afm::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));    \
			return -3;                                                                   \
		}                                                                                \
	} while (0)

#define AFM_TRY(expr)                                                                    \
	do {                                                                                 \
		int r_ = (expr);                                                                 \
		if (r_ != 0) return r_;                                                          \
	} while (0)

// ---------------------------------------------------- solve control ----
// One iterative solve's scalars, resident in device memory so that the
// stop test (mg/solver.h:228, alg/cg.h:42) runs on the GPU and the host only
// synchronises once per batch of iterations.
struct SolveCtl {
	double tol;       // absolute threshold (RMS for MG, 2-norm for CG)
	double res;       // current residual norm
	double n0, nu, mu;
	double n0p[2];    // CG: r.r after k updates in n0p[k & 1] (read and written by consecutive kernels)
	double mean;      // mean removed from the MG right-hand side
	double sum2;      // scratch second sum
	// flexible CG around the multigrid preconditioner: (z.r, alpha) of
	// iteration k in fg[k & 1]; p_{k-1}.r_k and p_{k-1}.L p_{k-1} after the
	// stop test of iteration k-1
	double fg[2][2];
	double frho, fdel;
	int count;        // completed iterations
	int done;         // stop test has fired
	int max_iter;
	int err;          // kernel-detected error flag
	unsigned ticket;  // last-block counter of the fused reductions
	int pad[3];
};

constexpr int NCTL = 8;          // ctl[0] MG, ctl[1..2] CG fields, ctl[3] misc
constexpr int RED_THREADS = 256;
constexpr int MAX_RED_BLOCKS = 1024;   // grid of the reducing kernels (grid-stride loops)

// Per-marker, per-component transfer record used by spreading.
struct MarkerRec {
	int ix0, iy0;           // first tap cell
	int tile, pad;
	double dx[4], dy[4];    // delta factors of the 4x4 window
	double ds, F;           // source volume and value
};

// The same record in bin order, as the spreading gather reads it: marker
// index, window base (ix0 | iy0 << 16) and the terms' factors.  96 bytes,
// 16-aligned: the gather loads a record as six 16-byte vectors.
struct alignas(16) BinRec {
	int key, base, pad0, pad1;
	double ds, F;
	double dx[4], dy[4];
};
static_assert(sizeof(BinRec) == 96, "BinRec is read as six 16-byte vectors");

// ------------------------------------------------------------- context --
struct LevelBuf {
	unsigned n = 0, nx = 0, ny = 0;
	double *Xa = nullptr, *Xb = nullptr, *B = nullptr, *CP = nullptr;
	float *Xa32 = nullptr, *Xb32 = nullptr, *B32 = nullptr;
};

struct Ctx {
	int device = 0;
	hipStream_t stream = nullptr;          // stream all work is ordered on
	hipStream_t own_stream = nullptr;      // created with the context
	unsigned n = 0, nx = 0, ny = 0, W = 1, D = 1, levels = 0, lc = 0;
	size_t N = 0;
	std::vector<LevelBuf> lv;          // by level, 0 = coarsest
	double *XSa = nullptr, *XSb = nullptr, *RES = nullptr, *RESb = nullptr;
	float *RES32 = nullptr, *RES32b = nullptr;
	double* upartials = nullptr;       // per-block sums of the multigrid update
	double* P64[2] = {};               // flexible CG search directions (ping-pong; FP32 in the mixed form)
	double* fpartials = nullptr;       // per-block z.r, z.r_prev, z.Lz [3 * fstride]
	unsigned fstride = 0;
	double* gpartials[2] = {};         // per-block r.r, p.r, p.Lp (3 columns), by iteration parity
	SolveCtl* ctl = nullptr;           // device [NCTL]
	SolveCtl* hctl = nullptr;          // pinned host mirror
	double* partials = nullptr;        // [MAX_RED_BLOCKS * 8]
	// spreading
	MarkerRec* rec = nullptr;          // [2 * mcap]
	int* bins = nullptr;               // counts | offsets | cursors | chunk sums, 2 comps
	int* gorder = nullptr;             // gather order: class counts | tile lists, 2 comps
	int* tflag = nullptr;              // per gather tile and comp: a record reaches it (set by the
	                                   // window kernel, cleared by the gather that reads it)
	int* bin_list = nullptr;           // [2 * mcap]
	BinRec* srec = nullptr;            // [2 * mcap], records in bin order
	long mcap = 0;
	int sbx = 0, sby = 0;              // spread bins per axis (4x4 base cells each)
	int last_mg_cycles = 0;            // predicted batch size for the next solve
	// the V-cycle of levels 3..0 as a 64 x 64 matrix, stored by columns
	// (op3[j * 64 + i] = M[i][j]), one per (nu1, nu2) at index 2 (nu1 - 1) +
	// nu2 - 1, built by mg_setup for the FP32-cycle forms' coarse kernel
	double* op3[4] = {};
	// level 6's V(2, 2) cycle (levels 6..0) as a 4096 x 4096 matrix, stored by
	// rows in FP16, row i scaled by 1 / op6_inv[i], built by mg_setup for the
	// FP32-cycle forms (k_coarse_mv6)
	_Float16* op6 = nullptr;
	float* op6_inv = nullptr;
	int n_cu = 256;                    // compute units of the device (persistent grids)
	// opt-in HIP-event timing of three finest-level launches of mg_solve:
	// kind 0 the outer update (forms 0, 2: fused with the last post-smoothing;
	// 1, 3: the CG update), kind 1 the last post-smoothing of the CG forms'
	// preconditioner cycle, kind 2 the pre-smoothing + residual + restriction
	static constexpr int KT_SLOTS = 64, KT_KINDS = 3;
	bool kt_on = false;
	hipEvent_t kt_ev[KT_KINDS][2 * KT_SLOTS] = {};
	double kt_ms[KT_KINDS] = {};
	long kt_launches[KT_KINDS] = {};
	int last_cg_iters = -1;   // the last Helmholtz CG's iterations (-1: none yet)
	hipEvent_t cg_read_ev = nullptr;   // recorded after a CG's control-block read-back (op_cg2)
	std::vector<void*> owned;
	~Ctx();
	int alloc(void** p, size_t bytes);
};

int ensure_markers(Ctx* c, long m);
int spread_setup(Ctx* c);

// -------------------------------------------------------- launchers ----
// IB transfer (afm_ib.hip)
int launch_delta(Ctx* c, const double* r, long count, unsigned n, double* out);
int launch_interp(Ctx* c, int comp, const double* field, const double* X, long m, double* U);
int launch_spread(Ctx* c, int comp_mask, const double* Fx, const double* Fy, const double* ds,
                  const double* X, const double* R, const double* k, long m, double* fx, double* fy,
                  int accumulate = 0);
int launch_ib_move_half(Ctx* c, const double* u, const double* v, const double* X, long m, double half, double* Xh);
int launch_ib_move_full(Ctx* c, const double* u_old, const double* v_old, const double* u, const double* v,
                        const double* Xh, long m, double dt, double* X);

// Multigrid (afm_mg.hip)
int mg_setup(Ctx* c);
// sol != nullptr: rhs is left as the mean-free right-hand side and *sol
// points at the solution inside the context (valid until its next solve).
// x0: the CG forms' (precision 1, 3) initial guess (nullptr: zero); the
// reference's iteration always starts from its nested iteration.
// sums_ready: the caller's launch has already summed rhs into the partials
// (grid red_blocks(N)) and started the control block, as k_mean does.
int mg_solve(Ctx* c, double* rhs, double tol, int tol_mode, int nu1, int nu2, int precision, int max_cycles,
             int* cycles, double* final_rms, const double** sol = nullptr, const double* x0 = nullptr,
             bool sums_ready = false);
int mg_apply_laplacian(Ctx* c, unsigned level, const double* x, double* y);
int mg_smooth(Ctx* c, unsigned level, const double* x_in, const double* b, double* x_out, int sweeps, int zero_init);
int mg_restrict(Ctx* c, unsigned level, const double* fine, double* coarse);
int mg_prolong(Ctx* c, unsigned level, const double* coarse, double* fine);
int mg_vcycle(Ctx* c, double* rhs_inout, int nu1, int nu2);
// The 3-D mode (afm_mg3.hip): cubes of W 2^L cells per side (W a power of
// two), laplacian scale 2^L, walls in y
int mg3_laplacian(Ctx* c, unsigned L, unsigned W, const double* x, double* y);
int mg3_vcycle(Ctx* c, unsigned L, unsigned W, double* rhs, int nu1, int nu2, double tol);
int mg3_solve(Ctx* c, unsigned L, unsigned W, double* rhs, double tol, int tol_mode, int nu1, int nu2,
              int max_cycles, int* cycles, double* final_rms);
int mg3_restrict(Ctx* c, unsigned L, unsigned W, const double* fine, double* coarse);
int mg3_prolong(Ctx* c, unsigned L, unsigned W, const double* coarse, double* fine);
int mg3_smooth(Ctx* c, unsigned L, unsigned W, double* x, const double* b, int sweeps);
int mg3_cg(Ctx* c, unsigned L, unsigned W, const double* b, double* x, double tol, int max_iters, int* iters,
           double* final_norm);
// a persistent 3-D multigrid workspace (level buffers, iterate, residual)
struct Mg3;
int mg3_create(Ctx* c, unsigned L, unsigned W, int nu1, int nu2, Mg3** out);
int mg3_solve_with(Mg3* m, double* rhs, double tol, int tol_mode, int max_cycles, int* cycles, double* final_rms);
void mg3_destroy(Mg3* m);
// the 3-D Navier-Stokes solver (afm_ns3.hip)
struct NS3;
int ns3_create(Ctx* c, unsigned L, double mu, double rho, double dt, double tol, double* u, double* v, double* w,
               NS3** out);
int ns3_step(NS3* s, double* ut, double* ub, double* wt, double* wb, const double* fx, const double* fy,
             const double* fz);
int ns3_get_p(NS3* s, double* p);
void ns3_stats(const NS3* s, int* cycles, int* cg, int* half);
void ns3_destroy(NS3* s);
int ns3_div(Ctx* c, unsigned L, const double* u, const double* v, const double* w, const double* top,
            const double* bottom, double* d);
int ns3_grad(Ctx* c, unsigned L, const double* p, double* px, double* py, double* pz);
int ns3_advect(Ctx* c, unsigned L, const double* u, const double* v, const double* w, double* ua, double* va,
               double* wa);
int ns3_helmholtz_apply(Ctx* c, unsigned L, int kind, double scale, const double* x, double* y);
int ns3_helmholtz_boundary(Ctx* c, unsigned L, int kind, double scale, const double* top, const double* bottom,
                           double* out);
int ns3_projection_boundary(Ctx* c, unsigned L, const double* phx, const double* phz, double* ut, double* ub,
                            double* wt, double* wb);
// device CSR matrices and vectors (afm_csr.hip)
int csr_spmv(Ctx* c, unsigned long rows, const unsigned long* rs, const unsigned long* col, const double* val,
             const double* x, double* y);
int csr_smooth(Ctx* c, unsigned long rows, const unsigned long* rs, const unsigned long* col, const double* val,
               double* x, const double* b, int sweeps);
int csr_cg(Ctx* c, unsigned long rows, const unsigned long* rs, const unsigned long* col, const double* val,
           const double* b, double* x, double tol, int max_iters, int* iters, double* final_norm);
int vec_axpy(Ctx* c, unsigned long N, double a, const double* x, double* y);
int ns3_helmholtz_cg(Ctx* c, unsigned L, int kind, double scale, const double* b, double* x, double tol,
                     int max_iters, int* iters, double* final_norm);
// the IB transfer on the 3-D MAC grid (afm_ib3.hip)
int ib3_interp(Ctx* c, unsigned L, int comp, const double* field, const double* X, long m, double* U);
int ib3_spread(Ctx* c, unsigned L, int comp, const double* F, const double* ds, const double* X, long m, double* f);

// Reductions / helpers (afm_ns.hip)
int launch_sum(Ctx* c, const double* x, size_t N, double* out_host);

// Adjoint building blocks (afm_adj.hip)
int adj_axpy(Ctx* c, double* y, const double* x, double a, size_t N);
int adj_scale_copy(Ctx* c, double* y, const double* x, double a, size_t N);
int adj_advect_T(Ctx* c, const double* ab, const double* bb, const double* u, const double* v, double* ub, double* vb);
int adj_interp_dX(Ctx* c, int comp, const double* f, const double* X, const double* coef, long m, double* Xb);
int adj_split_scale(Ctx* c, const double* Xb, double scale, long m, double* Ux, double* Uy);
int adj_spring(Ctx* c, const double* P, const double* R, const double* k, const double* ds, const double* Jx,
               const double* Jy, long m, double* kb, double* Rb, double* Pb, double* cx, double* cy);
int adj_centroids(Ctx* c, const double* X, long nm, long msize, const double* c0, double* cout, double* Xb,
                  double* part);
int adj_kinetic(Ctx* c, const double* u, const double* v, double* ub, double* vb, double* out_host);

}  // namespace afm

struct afm_ctx : afm::Ctx {};
