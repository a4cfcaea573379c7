// afm_capi.hip -- the extern "C" boundary of libafm (declared in include/afm.h).
#include "afm_device.h"
#include "../../include/afm.h"

namespace afm {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

Ctx::~Ctx()
{
	if (device >= 0) (void) hipSetDevice(device);
	for (void* p : owned) (void) hipFree(p);
	if (rec) (void) hipFree(rec);
	if (bin_list) (void) hipFree(bin_list);
	if (srec) (void) hipFree(srec);
	if (hctl) (void) hipHostFree(hctl);
	for (auto& kind : kt_ev)
		for (hipEvent_t e : kind)
			if (e) (void) hipEventDestroy(e);
	if (cg_read_ev) (void) hipEventDestroy(cg_read_ev);
	if (own_stream) (void) hipStreamDestroy(own_stream);
}

int Ctx::alloc(void** p, size_t bytes)
{
	AFM_HIP(hipMalloc(p, bytes));
	owned.push_back(*p);
	return 0;
}

// defined in afm_ns.hip
struct NS;
struct IB;
int ns_create(Ctx*, double, double, double, double, int, int, double*, double*, NS**);
int ns_step(NS*, const double*, const double*);
int ns_project(NS*, double*, double*, int, bool div_ready);
int ns_get_start(NS*, int, double*, int*);
int ns_set_start(NS*, int, const double*);
int ns_get_p(NS*, double*);
int ib_create(NS*, long, const double*, const double*, const double*, const double*, IB**);
int ib_step(IB*);
int op_div(Ctx*, const double*, const double*, double*);
int op_grad_sub(Ctx*, const double*, double*, double*);
int op_advect(Ctx*, const double*, const double*, double*, double*);
int op_helm_apply(Ctx*, double, const double*, double*);
struct CGWork;
int op_cg_single(Ctx*, double, const double*, double*, double, int, int, int*, double*);
int op_cg_single_noalloc(Ctx*, double, const double*, double*, double, int, int, int*, double*);
int cg_scratch_setup(int);
int ns_stats(NS*, int*, int*, int*, int*);
int ns_last_residual(NS*, double*, double*);
int ns_old(NS*, const double**, const double**);
Ctx* ns_ctx(NS*);
void ns_free(NS*);
Ctx* ib_ctx(IB*);
int ib_markers(IB*, double*);
int ib_grad(IB*, int, int, long, double*, double*, double*);
void ib_free(IB*);

}  // namespace afm

using namespace afm;

struct afm_ns { NS* s; };
struct afm_ib { IB* b; };
struct afm_ns3 { NS3* s; afm_ctx* c; };

#define CHECK_ARG(cond, msg) do { if (!(cond)) { set_error(msg); return -1; } } while (0)

// the grid kernels move two cells per vector access
template<typename... P>
static bool aligned16(P... p) { return ((((uintptr_t) (const void*) p & 15) == 0) && ...); }
#define CHECK_ALIGNED(...) CHECK_ARG(aligned16(__VA_ARGS__), "field arrays must be 16-byte aligned")

static int use(Ctx* c)
{
	AFM_HIP(hipSetDevice(c->device));
	return 0;
}

extern "C" {

const char* afm_last_error(void) { return last_error(); }

int afm_ctx_create(const afm_geom* g, int device, afm_ctx** out)
{
	CHECK_ARG(g && out, "null argument");
	CHECK_ARG(g->dim == 2, "only dim == 2 (doubly periodic plane) is built");
	CHECK_ARG(g->width == 1 && g->depth == 1, "only the unit square (width = depth = 1) is built");
	CHECK_ARG(g->n >= 8 && (g->n & (g->n - 1)) == 0, "n must be a power of two >= 8");
	CHECK_ARG(g->n <= 16384, "n must be <= 16384");
	int ndev = 0;
	AFM_HIP(hipGetDeviceCount(&ndev));
	CHECK_ARG(device >= 0 && device < ndev, "device index out of range");
	AFM_HIP(hipSetDevice(device));
	AFM_TRY(cg_scratch_setup(device));
	afm_ctx* c = new afm_ctx();
	c->device = device;
	c->n = g->n; c->W = g->width; c->D = g->depth;
	c->nx = g->width * g->n; c->ny = g->depth * g->n;
	c->N = (size_t) c->nx * c->ny;
	int rc = 0;
	auto fail = [&](int r) { delete c; return r; };
	if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) { set_error("hipStreamCreate failed"); return fail(-3); }
	c->stream = c->own_stream;
	if (hipEventCreateWithFlags(&c->cg_read_ev, hipEventDisableTiming) != hipSuccess) { set_error("hipEventCreate failed"); return fail(-3); }
	if ((rc = c->alloc((void**) &c->ctl, sizeof(SolveCtl) * NCTL))) return fail(rc);
	if (hipHostMalloc((void**) &c->hctl, sizeof(SolveCtl) * NCTL, hipHostMallocDefault) != hipSuccess) { set_error("hipHostMalloc failed"); return fail(-3); }
	if ((rc = c->alloc((void**) &c->partials, sizeof(double) * MAX_RED_BLOCKS * 8))) return fail(rc);
	if (hipMemsetAsync(c->ctl, 0, sizeof(SolveCtl) * NCTL, c->stream) != hipSuccess) { set_error("memset failed"); return fail(-3); }
	c->sbx = (int) (c->nx / 4); c->sby = (int) (c->ny / 4);
	{
		// counts, offsets (+1), cursors and up to 1024 scan-chunk sums for both components
		const size_t nb = 2 * (size_t) c->sbx * c->sby;
		if ((rc = c->alloc((void**) &c->bins, sizeof(int) * (3 * nb + 1 + 1024)))) return fail(rc);
		// the bin counts start at zero once; each spreading's scan returns
		// them to zero after reading them
		if (hipMemsetAsync(c->bins, 0, sizeof(int) * nb, c->stream) != hipSuccess) { set_error("memset failed"); return fail(-3); }
		if ((rc = spread_setup(c))) return fail(rc);
	}
	if ((rc = mg_setup(c))) return fail(rc);
	if (hipStreamSynchronize(c->stream) != hipSuccess) { set_error("sync failed"); return fail(-3); }
	*out = c;
	return 0;
}

int afm_ctx_destroy(afm_ctx* c)
{
	if (!c) return 0;
	hipSetDevice(c->device);
	hipStreamSynchronize(c->stream);
	delete c;
	return 0;
}

int afm_ctx_info(afm_ctx* c, unsigned* nx, unsigned* ny, unsigned* levels, unsigned* coarse_level)
{
	CHECK_ARG(c, "null context");
	if (nx) *nx = c->nx;
	if (ny) *ny = c->ny;
	if (levels) *levels = c->levels;
	if (coarse_level) *coarse_level = c->lc;
	return 0;
}

int afm_set_stream(afm_ctx* c, void* stream)
{
	CHECK_ARG(c, "null context");
	AFM_TRY(use(c));
	AFM_HIP(hipStreamSynchronize(c->stream));
	c->stream = (hipStream_t) stream;     // NULL: HIP's null stream
	return 0;
}

void* afm_own_stream(afm_ctx* c) { return c ? (void*) c->own_stream : nullptr; }

int afm_update_timing(afm_ctx* c, int enable)
{
	CHECK_ARG(c, "null context");
	AFM_TRY(use(c));
	if (enable && !c->kt_ev[0][0])
		for (auto& kind : c->kt_ev)
			for (hipEvent_t& e : kind) AFM_HIP(hipEventCreate(&e));
	c->kt_on = enable != 0;
	for (int k = 0; k < Ctx::KT_KINDS; ++k) {
		c->kt_ms[k] = 0.0;
		c->kt_launches[k] = 0;
	}
	return 0;
}

int afm_update_timing_read(afm_ctx* c, double* avg_us, long* launches)
{
	CHECK_ARG(c && avg_us && launches, "null argument");
	return afm_kernel_timing_read(c, 0, avg_us, launches);
}

int afm_kernel_timing_read(afm_ctx* c, int kind, double* avg_us, long* launches)
{
	CHECK_ARG(c && avg_us && launches, "null argument");
	CHECK_ARG(kind >= 0 && kind < Ctx::KT_KINDS, "kind must be 0, 1 or 2");
	*launches = c->kt_launches[kind];
	*avg_us = c->kt_launches[kind] ? 1e3 * c->kt_ms[kind] / (double) c->kt_launches[kind] : 0.0;
	return 0;
}

int afm_sync(afm_ctx* c)
{
	CHECK_ARG(c, "null context");
	AFM_TRY(use(c));
	AFM_HIP(hipStreamSynchronize(c->stream));
	hipDeviceSynchronize();
	int err = 0;
	AFM_HIP(hipMemcpy(&err, &c->ctl[3].err, sizeof(int), hipMemcpyDeviceToHost));
	if (err) {
		int zero = 0;
		hipMemcpy(&c->ctl[3].err, &zero, sizeof(int), hipMemcpyHostToDevice);
		set_error("a kernel reported an error");
		return -1;
	}
	return 0;
}

int afm_alloc(afm_ctx* c, unsigned long bytes, void** out)
{
	CHECK_ARG(c && out, "null argument");
	AFM_TRY(use(c));
	AFM_HIP(hipMalloc(out, bytes ? bytes : 8));
	return 0;
}

int afm_free(afm_ctx* c, void* p)
{
	CHECK_ARG(c, "null argument");
	AFM_TRY(use(c));
	AFM_HIP(hipStreamSynchronize(c->stream));
	if (p) AFM_HIP(hipFree(p));
	return 0;
}

int afm_copy_to_device(afm_ctx* c, void* dst, const void* src, unsigned long bytes)
{
	CHECK_ARG(c && (bytes == 0 || (dst && src)), "null argument");
	AFM_TRY(use(c));
	AFM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
	AFM_HIP(hipStreamSynchronize(c->stream));
	return 0;
}

int afm_copy_to_host(afm_ctx* c, void* dst, const void* src, unsigned long bytes)
{
	CHECK_ARG(c && (bytes == 0 || (dst && src)), "null argument");
	AFM_TRY(use(c));
	AFM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
	AFM_HIP(hipStreamSynchronize(c->stream));
	return 0;
}

int afm_delta(afm_ctx* c, const double* r, long count, double* out)
{
	CHECK_ARG(c && (count == 0 || (r && out)), "null argument");
	AFM_TRY(use(c));
	return launch_delta(c, r, count, c->n, out);
}

int afm_delta_host(const double* r, long count, unsigned n, double* out)
{
	CHECK_ARG(count == 0 || (r && out), "null argument");
	CHECK_ARG(n >= 1 && (n & (n - 1)) == 0, "n must be a power of two");
	for (long i = 0; i < count; ++i) out[i] = delta(r[i], n);
	return 0;
}

int afm_interp(afm_ctx* c, int comp, const double* field, const double* X, long m, double* U)
{
	CHECK_ARG(c && (m == 0 || (field && X && U)), "null argument");
	CHECK_ARG(comp == 0 || comp == 1, "comp must be 0 (u) or 1 (v)");
	CHECK_ARG(m >= 0, "m must be >= 0");
	AFM_TRY(use(c));
	return launch_interp(c, comp, field, X, m, U);
}

int afm_spread(afm_ctx* c, int comp, const double* F, const double* ds, const double* X, long m, double* f)
{
	CHECK_ARG(c && f && (m == 0 || (F && ds && X)), "null argument");
	CHECK_ARG(comp == 0 || comp == 1, "comp must be 0 (u) or 1 (v)");
	CHECK_ARG(m >= 0 && m < (1l << 30), "m out of range");
	AFM_TRY(use(c));
	return launch_spread(c, 1 << comp, comp == 0 ? F : nullptr, comp == 1 ? F : nullptr, ds, X, nullptr, nullptr, m,
	                     comp == 0 ? f : nullptr, comp == 1 ? f : nullptr);
}

int afm_laplacian_apply(afm_ctx* c, unsigned level, const double* x, double* y)
{
	CHECK_ARG(c && x && y, "null argument");
	AFM_TRY(use(c));
	return mg_apply_laplacian(c, level, x, y);
}

int afm_mg_smooth(afm_ctx* c, unsigned level, const double* x_in, const double* b, double* x_out, int sweeps,
                  int zero_init)
{
	CHECK_ARG(c && b && x_out && (zero_init || x_in), "null argument");
	CHECK_ARG(x_in != x_out, "x_out must differ from x_in");
	CHECK_ALIGNED(x_in, b, x_out);
	AFM_TRY(use(c));
	return mg_smooth(c, level, x_in, b, x_out, sweeps, zero_init);
}

int afm_mg_restrict(afm_ctx* c, unsigned level, const double* fine, double* coarse)
{
	CHECK_ARG(c && fine && coarse, "null argument");
	AFM_TRY(use(c));
	return mg_restrict(c, level, fine, coarse);
}

int afm_mg_prolong(afm_ctx* c, unsigned level, const double* coarse, double* fine)
{
	CHECK_ARG(c && fine && coarse, "null argument");
	AFM_TRY(use(c));
	return mg_prolong(c, level, coarse, fine);
}

int afm_mg_vcycle(afm_ctx* c, double* rhs, int nu1, int nu2)
{
	CHECK_ARG(c && rhs, "null argument");
	CHECK_ARG(nu1 >= 1 && nu1 <= 2 && nu2 >= 1 && nu2 <= 2, "nu1, nu2 must be 1 or 2");
	CHECK_ALIGNED(rhs);
	AFM_TRY(use(c));
	return mg_vcycle(c, rhs, nu1, nu2);
}

// the 3-D mode: geometry<levels, width, width, width>, width 2^levels cells
// per axis; width a power of two, at most 512 cells per axis
#define CHECK_MG3(c, levels, width, nu1, nu2)                                                           \
	CHECK_ARG(c, "null context");                                                                       \
	CHECK_ARG((levels) <= 9 && (width) >= 1 && ((width) & ((width) - 1)) == 0,                         \
	          "levels must be 0..9 and width a power of two");                                         \
	CHECK_ARG(((size_t) (width) << (levels)) <= 512 && ((size_t) (width) << (levels)) >= 2,            \
	          "the cube must have 2..512 cells per axis");                                             \
	CHECK_ARG((nu1) >= 0 && (nu1) <= 16 && (nu2) >= 0 && (nu2) <= 16, "nu1, nu2 must be 0..16")

int afm_laplacian3_apply(afm_ctx* c, unsigned levels, unsigned width, const double* x, double* y)
{
	CHECK_MG3(c, levels, width, 0, 0);
	CHECK_ARG(x && y && x != y, "null or aliased argument");
	AFM_TRY(use(c));
	return mg3_laplacian(c, levels, width, x, y);
}

int afm_mg3_vcycle(afm_ctx* c, unsigned levels, unsigned width, double* rhs, int nu1, int nu2, double tol)
{
	CHECK_MG3(c, levels, width, nu1, nu2);
	CHECK_ARG(rhs, "null argument");
	CHECK_ARG(tol >= 0, "tol must be >= 0");
	AFM_TRY(use(c));
	return mg3_vcycle(c, levels, width, rhs, nu1, nu2, tol);
}

int afm_mg3_solve(afm_ctx* c, unsigned levels, unsigned width, double* rhs, double tol, int tol_mode, int nu1,
                  int nu2, int max_cycles, int* cycles, double* final_rms)
{
	CHECK_MG3(c, levels, width, nu1, nu2);
	CHECK_ARG(rhs, "null argument");
	CHECK_ARG(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
	CHECK_ARG(tol >= 0 && max_cycles >= 0, "tol and max_cycles must be >= 0");
	AFM_TRY(use(c));
	return mg3_solve(c, levels, width, rhs, tol, tol_mode, nu1, nu2, max_cycles, cycles, final_rms);
}

int afm_mg3_restrict(afm_ctx* c, unsigned levels, unsigned width, const double* fine, double* coarse)
{
	CHECK_MG3(c, levels, width, 0, 0);
	CHECK_ARG(levels >= 1, "restriction needs levels >= 1");
	CHECK_ARG(fine && coarse && fine != coarse, "null or aliased argument");
	AFM_TRY(use(c));
	return mg3_restrict(c, levels, width, fine, coarse);
}

int afm_mg3_prolong(afm_ctx* c, unsigned levels, unsigned width, const double* coarse, double* fine)
{
	CHECK_MG3(c, levels, width, 0, 0);
	CHECK_ARG(levels >= 1, "prolongation needs levels >= 1");
	CHECK_ARG(fine && coarse && fine != coarse, "null or aliased argument");
	AFM_TRY(use(c));
	return mg3_prolong(c, levels, width, coarse, fine);
}

int afm_mg3_smooth(afm_ctx* c, unsigned levels, unsigned width, double* x, const double* b, int sweeps)
{
	CHECK_MG3(c, levels, width, 0, 0);
	CHECK_ARG(x && b && x != b, "null or aliased argument");
	CHECK_ARG(sweeps >= 0, "sweeps must be >= 0");
	AFM_TRY(use(c));
	return mg3_smooth(c, levels, width, x, b, sweeps);
}

int afm_laplacian3_cg(afm_ctx* c, unsigned levels, unsigned width, const double* b, double* x, double tol,
                      int max_iters, int* iters, double* final_norm)
{
	CHECK_MG3(c, levels, width, 0, 0);
	CHECK_ARG(x && b, "null argument");
	CHECK_ARG(tol >= 0 && max_iters >= 0, "tol and max_iters must be >= 0");
	AFM_TRY(use(c));
	return mg3_cg(c, levels, width, b, x, tol, max_iters, iters, final_norm);
}

#define CHECK_NS3_LEVELS(levels) CHECK_ARG((levels) >= 1 && (levels) <= 8, "levels must be 1..8 (n = 2^levels)")

int afm_ns3_create(afm_ctx* c, unsigned levels, double mu, double rho, double dt, double tol, double* u0, double* v0,
                   double* w0, afm_ns3** out)
{
	CHECK_ARG(c && u0 && v0 && w0 && out, "null argument");
	CHECK_NS3_LEVELS(levels);
	CHECK_ARG(rho > 0 && dt > 0 && mu >= 0 && tol > 0, "need rho > 0, dt > 0, mu >= 0, tol > 0");
	AFM_TRY(use(c));
	NS3* s = nullptr;
	AFM_TRY(ns3_create(c, levels, mu, rho, dt, tol, u0, v0, w0, &s));
	*out = new afm_ns3{s, c};
	return 0;
}

int afm_ns3_step(afm_ns3* s, double* ut, double* ub, double* wt, double* wb, const double* fx, const double* fy,
                 const double* fz)
{
	CHECK_ARG(s && ut && ub && wt && wb && fx && fy && fz, "null argument");
	AFM_TRY(use(s->c));
	return ns3_step(s->s, ut, ub, wt, wb, fx, fy, fz);
}

int afm_ns3_get_p(afm_ns3* s, double* p)
{
	CHECK_ARG(s && p, "null argument");
	AFM_TRY(use(s->c));
	return ns3_get_p(s->s, p);
}

int afm_ns3_stats(afm_ns3* s, int* last_cycles, int* last_cg, int* have_half_stepped)
{
	CHECK_ARG(s, "null argument");
	ns3_stats(s->s, last_cycles, last_cg, have_half_stepped);
	return 0;
}

int afm_ns3_destroy(afm_ns3* s)
{
	if (!s) return 0;
	(void) use(s->c);
	ns3_destroy(s->s);
	delete s;
	return 0;
}

int afm_div3(afm_ctx* c, unsigned levels, const double* u, const double* v, const double* w, const double* top,
             const double* bottom, double* d)
{
	CHECK_ARG(c && u && v && w && top && bottom && d, "null argument");
	CHECK_NS3_LEVELS(levels);
	AFM_TRY(use(c));
	return ns3_div(c, levels, u, v, w, top, bottom, d);
}

int afm_grad3(afm_ctx* c, unsigned levels, const double* p, double* px, double* py, double* pz)
{
	CHECK_ARG(c && p && px && py && pz, "null argument");
	CHECK_NS3_LEVELS(levels);
	AFM_TRY(use(c));
	return ns3_grad(c, levels, p, px, py, pz);
}

int afm_advect3(afm_ctx* c, unsigned levels, const double* u, const double* v, const double* w, double* u_adv,
                double* v_adv, double* w_adv)
{
	CHECK_ARG(c && u && v && w && u_adv && v_adv && w_adv, "null argument");
	CHECK_NS3_LEVELS(levels);
	AFM_TRY(use(c));
	return ns3_advect(c, levels, u, v, w, u_adv, v_adv, w_adv);
}

int afm_helmholtz3_apply(afm_ctx* c, unsigned levels, int kind, double scale, const double* x, double* y)
{
	CHECK_ARG(c && x && y && x != y, "null or aliased argument");
	CHECK_NS3_LEVELS(levels);
	CHECK_ARG(kind == 1 || kind == 3, "kind must be 1 (helmholtz_periodic) or 3 (helmholtz_dirichlet)");
	AFM_TRY(use(c));
	return ns3_helmholtz_apply(c, levels, kind, scale, x, y);
}

int afm_helmholtz3_cg(afm_ctx* c, unsigned levels, int kind, double scale, const double* b, double* x, double tol,
                      int max_iters, int* iters, double* final_norm)
{
	CHECK_ARG(c && b && x, "null argument");
	CHECK_NS3_LEVELS(levels);
	CHECK_ARG(kind == 1 || kind == 3, "kind must be 1 (helmholtz_periodic) or 3 (helmholtz_dirichlet)");
	CHECK_ARG(tol >= 0 && max_iters >= 0, "tol and max_iters must be >= 0");
	AFM_TRY(use(c));
	return ns3_helmholtz_cg(c, levels, kind, scale, b, x, tol, max_iters, iters, final_norm);
}

int afm_helmholtz3_boundary(afm_ctx* c, unsigned levels, int kind, double scale, const double* top,
                            const double* bottom, double* out)
{
	CHECK_ARG(c && top && bottom && out, "null argument");
	CHECK_NS3_LEVELS(levels);
	CHECK_ARG(kind == 1 || kind == 3, "kind must be 1 (helmholtz_periodic) or 3 (helmholtz_dirichlet)");
	AFM_TRY(use(c));
	return ns3_helmholtz_boundary(c, levels, kind, scale, top, bottom, out);
}

int afm_projection3_boundary(afm_ctx* c, unsigned levels, const double* phx, const double* phz, double* ut,
                             double* ub, double* wt, double* wb)
{
	CHECK_ARG(c && phx && phz && ut && ub && wt && wb, "null argument");
	CHECK_NS3_LEVELS(levels);
	AFM_TRY(use(c));
	return ns3_projection_boundary(c, levels, phx, phz, ut, ub, wt, wb);
}

int afm_csr_spmv(afm_ctx* c, unsigned long rows, const unsigned long* row_start, const unsigned long* column,
                 const double* values, const double* x, double* y)
{
	CHECK_ARG(c && (rows == 0 || (row_start && x && y && x != y)), "null or aliased argument");
	AFM_TRY(use(c));
	return csr_spmv(c, rows, row_start, column, values, x, y);
}

int afm_csr_cg(afm_ctx* c, unsigned long rows, const unsigned long* row_start, const unsigned long* column,
               const double* values, const double* b, double* x, double tol, int max_iters, int* iters,
               double* final_norm)
{
	CHECK_ARG(c && (rows == 0 || (row_start && b && x)), "null argument");
	CHECK_ARG(tol >= 0 && max_iters >= 0, "tol and max_iters must be >= 0");
	AFM_TRY(use(c));
	return csr_cg(c, rows, row_start, column, values, b, x, tol, max_iters, iters, final_norm);
}

int afm_csr_smooth(afm_ctx* c, unsigned long rows, const unsigned long* row_start, const unsigned long* column,
                   const double* values, double* x, const double* b, int sweeps)
{
	CHECK_ARG(c && (rows == 0 || (row_start && x && b && x != b)), "null or aliased argument");
	CHECK_ARG(sweeps >= 0, "sweeps must be >= 0");
	AFM_TRY(use(c));
	return csr_smooth(c, rows, row_start, column, values, x, b, sweeps);
}

int afm_axpy(afm_ctx* c, unsigned long count, double a, const double* x, double* y)
{
	CHECK_ARG(c && (count == 0 || (x && y)), "null argument");
	AFM_TRY(use(c));
	return vec_axpy(c, count, a, x, y);
}

#define CHECK_IB3(c, levels, comp, m)                                                                 \
	CHECK_ARG(c, "null context");                                                                     \
	CHECK_NS3_LEVELS(levels);                                                                         \
	CHECK_ARG((comp) >= 0 && (comp) <= 3, "comp must be 0 (u), 1 (v), 2 (w) or 3 (nodes)");           \
	CHECK_ARG((m) >= 0 && (m) <= (1L << 24), "m must be 0..2^24 markers")

int afm_interp3(afm_ctx* c, unsigned levels, int comp, const double* field, const double* X, long m, double* U)
{
	CHECK_IB3(c, levels, comp, m);
	CHECK_ARG(m == 0 || (field && X && U), "null argument");
	AFM_TRY(use(c));
	return ib3_interp(c, levels, comp, field, X, m, U);
}

int afm_spread3(afm_ctx* c, unsigned levels, int comp, const double* F, const double* ds, const double* X, long m,
                double* f)
{
	CHECK_IB3(c, levels, comp, m);
	CHECK_ARG(f && (m == 0 || (F && ds && X)), "null argument");
	AFM_TRY(use(c));
	return ib3_spread(c, levels, comp, F, ds, X, m, f);
}

int afm_mg_solve(afm_ctx* c, double* rhs, double tol, int tol_mode, int nu1, int nu2, int precision, int max_cycles,
                 int* cycles, double* final_rms)
{
	CHECK_ARG(c && rhs, "null argument");
	CHECK_ARG(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
	CHECK_ARG(tol >= 0 && max_cycles >= 0, "tol and max_cycles must be >= 0");
	CHECK_ALIGNED(rhs);
	AFM_TRY(use(c));
	return mg_solve(c, rhs, tol, tol_mode, nu1, nu2, precision, max_cycles, cycles, final_rms);
}

int afm_helmholtz_apply(afm_ctx* c, double scale, const double* x, double* y)
{
	CHECK_ARG(c && x && y, "null argument");
	AFM_TRY(use(c));
	return op_helm_apply(c, scale, x, y);
}

int afm_helmholtz_cg(afm_ctx* c, double scale, const double* b, double* x, double tol, int tol_mode, int max_iters,
                     int* iters, double* final_norm)
{
	CHECK_ARG(c && b && x, "null argument");
	CHECK_ARG(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
	AFM_TRY(use(c));
	return op_cg_single(c, scale, b, x, tol, tol_mode, max_iters, iters, final_norm);
}

int afm_helmholtz_cg_noalloc(afm_ctx* c, double scale, const double* b, double* x, double tol, int tol_mode,
                             int max_iters, int* iters, double* final_norm)
{
	CHECK_ARG(c && b && x, "null argument");
	CHECK_ARG(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
	AFM_TRY(use(c));
	return op_cg_single_noalloc(c, scale, b, x, tol, tol_mode, max_iters, iters, final_norm);
}

int afm_div(afm_ctx* c, const double* u, const double* v, double* d)
{
	CHECK_ARG(c && u && v && d, "null argument");
	AFM_TRY(use(c));
	return op_div(c, u, v, d);
}

int afm_grad_sub(afm_ctx* c, const double* p, double* u, double* v)
{
	CHECK_ARG(c && p && u && v, "null argument");
	AFM_TRY(use(c));
	return op_grad_sub(c, p, u, v);
}

int afm_advect(afm_ctx* c, const double* u, const double* v, double* ua, double* va)
{
	CHECK_ARG(c && u && v && ua && va, "null argument");
	AFM_TRY(use(c));
	return op_advect(c, u, v, ua, va);
}

int afm_ns_create(afm_ctx* c, double mu, double rho, double dt, double tol, int tol_mode, int precision, double* u0,
                  double* v0, afm_ns** out)
{
	CHECK_ARG(c && u0 && v0 && out, "null argument");
	CHECK_ARG(rho > 0 && dt > 0 && mu >= 0 && tol >= 0, "need rho > 0, dt > 0, mu >= 0, tol >= 0");
	CHECK_ARG(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
	CHECK_ARG(precision >= 0 && precision <= 3, "precision must be 0, 1, 2 or 3 (as for afm_mg_solve)");
	AFM_TRY(use(c));
	NS* s = nullptr;
	AFM_TRY(ns_create(c, mu, rho, dt, tol, tol_mode, precision, u0, v0, &s));
	*out = new afm_ns{s};
	return 0;
}

int afm_ns_step(afm_ns* s, const double* fx, const double* fy)
{
	CHECK_ARG(s && fx && fy, "null argument");
	AFM_TRY(use(ns_ctx(s->s)));
	return ns_step(s->s, fx, fy);
}

int afm_ns_project(afm_ns* s, double* u, double* v)
{
	CHECK_ARG(s && u && v, "null argument");
	AFM_TRY(use(ns_ctx(s->s)));
	return ns_project(s->s, u, v, -1, false);
}

int afm_ns_get_p(afm_ns* s, double* p)
{
	CHECK_ARG(s && p, "null argument");
	AFM_TRY(use(ns_ctx(s->s)));
	return ns_get_p(s->s, p);
}

int afm_ns_stats(afm_ns* s, int* cyc, int* cgu, int* cgv, int* half)
{
	CHECK_ARG(s, "null argument");
	return ns_stats(s->s, cyc, cgu, cgv, half);
}

int afm_ns_last_residual(afm_ns* s, double* final_rms, double* threshold)
{
	CHECK_ARG(s && final_rms && threshold, "null argument");
	return ns_last_residual(s->s, final_rms, threshold);
}

int afm_ns_get_start(afm_ns* s, int kind, double* phi, int* valid)
{
	CHECK_ARG(s && phi && valid, "null argument");
	CHECK_ALIGNED(phi);
	AFM_TRY(use(ns_ctx(s->s)));
	return ns_get_start(s->s, kind, phi, valid);
}

int afm_ns_set_start(afm_ns* s, int kind, const double* phi)
{
	CHECK_ARG(s, "null argument");
	if (phi) CHECK_ALIGNED(phi);
	AFM_TRY(use(ns_ctx(s->s)));
	return ns_set_start(s->s, kind, phi);
}

int afm_ns_old(afm_ns* s, const double** uo, const double** vo)
{
	CHECK_ARG(s && uo && vo, "null argument");
	return ns_old(s->s, uo, vo);
}

int afm_ns_destroy(afm_ns* s)
{
	if (!s) return 0;
	ns_free(s->s);
	delete s;
	return 0;
}

int afm_ib_create(afm_ns* s, long m, const double* X, const double* R, const double* k, const double* ds,
                  afm_ib** out)
{
	CHECK_ARG(s && out && (m == 0 || (X && R && k && ds)), "null argument");
	CHECK_ARG(m >= 0 && m < (1l << 30), "m out of range");
	AFM_TRY(use(ns_ctx(s->s)));
	IB* b = nullptr;
	AFM_TRY(ib_create(s->s, m, X, R, k, ds, &b));
	*out = new afm_ib{b};
	return 0;
}

int afm_ib_step(afm_ib* b, int steps)
{
	CHECK_ARG(b && steps >= 0, "bad argument");
	AFM_TRY(use(ib_ctx(b->b)));
	for (int i = 0; i < steps; ++i) AFM_TRY(ib_step(b->b));
	return 0;
}

int afm_ib_grad(afm_ib* b, int steps, int every, long msize, double* J, double* gk, double* gR)
{
	CHECK_ARG(b, "null argument");
	AFM_TRY(use(ib_ctx(b->b)));
	return ib_grad(b->b, steps, every, msize, J, gk, gR);
}

int afm_ib_markers(afm_ib* b, double* X_out)
{
	CHECK_ARG(b && X_out, "null argument");
	AFM_TRY(use(ib_ctx(b->b)));
	return ib_markers(b->b, X_out);
}

int afm_ib_destroy(afm_ib* b)
{
	if (!b) return 0;
	ib_free(b->b);
	delete b;
	return 0;
}

}  // extern "C"
