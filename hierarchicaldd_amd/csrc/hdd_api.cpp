// extern "C" boundary of the library (declarations and reference citations in include/hdd.h).
#include <omp.h>
#include <cstdlib>
#include "../../include/hdd.h"
#include "hdd_runtime.h"
#include "hdd_server.h"
#include "hdd_sigma_device.h"
#include "hdd_spectral.h"
#include "hdd_triplan.h"
#include "hdd_parallel.h"
#include "hdd_factor.h"
#include <cstring>
#include <mutex>
#include <string>

using namespace hdd;

struct hdd_ctx {
    std::unique_ptr<Solver> solver;
    std::string err;
};

static thread_local std::string g_create_error;

#define HDD_TRY(ctx, ...)                                   \
    try {                                                   \
        __VA_ARGS__;                                        \
        return HDD_OK;                                      \
    } catch (std::exception & e) {                          \
        if (ctx) (ctx)->err = e.what();                     \
        return HDD_ERR_DEVICE;                              \
    }

extern "C" {

static bool desc_ok(const hdd_problem_desc* d, std::string& why) {
    if (!d || d->n_pages < 0 || (d->n_pages > 0 && (!d->page_offsets || !d->flags || !d->b))) { why = "invalid arguments"; return false; }
    for (int a = 0; a < 3; ++a)
        if (d->size[a] < 2 || d->sub_size[a] < 2) { why = "size and sub_size must be >= 2"; return false; }
    if (d->vlevels < 1 || d->sigma_levels < 1 || d->nVcycles < 0 || d->interior_smoothing < 0 || d->boundary_smoothing < 0 ||
        d->sigma_smoothing < 0 || d->sigma_mg < 0) { why = "invalid solver parameters"; return false; }
    return true;
}

static SolverParams params_of(const hdd_problem_desc* d) {
    SolverParams p;
    p.size = {d->size[0], d->size[1], d->size[2]};
    p.sub = {d->sub_size[0], d->sub_size[1], d->sub_size[2]};
    p.vlevels = d->vlevels; p.sigma_levels = d->sigma_levels; p.nV = d->nVcycles;
    p.int_s = d->interior_smoothing; p.bnd_s = d->boundary_smoothing;
    p.sig_s = d->sigma_smoothing; p.sig_mg = d->sigma_mg;
    return p;
}

static hdd_status create_impl(const hdd_problem_desc* d, hdd_ctx** out, bool host_only) {
    std::string why;
    if (!out || !desc_ok(d, why)) { g_create_error = "hdd_create: " + (out ? why : std::string("invalid arguments")); return HDD_ERR_ARG; }
    try {
        auto* c = new hdd_ctx();
        try {
            c->solver.reset(new Solver(params_of(d), d->n_pages, d->page_offsets, d->flags, d->b, host_only));
        } catch (...) { delete c; throw; }
        *out = c;
        return HDD_OK;
    } catch (std::exception& e) {
        g_create_error = e.what();
        return HDD_ERR_SETUP;
    }
}

hdd_status hdd_create(const hdd_problem_desc* d, hdd_ctx** out) { return create_impl(d, out, false); }
hdd_status hdd_create_host_only(const hdd_problem_desc* d, hdd_ctx** out) { return create_impl(d, out, true); }

hdd_status hdd_create_batch(const hdd_problem_desc* descs, int count, hdd_ctx** out) {
    if (!descs || !out || count < 1) { g_create_error = "hdd_create_batch: invalid arguments"; return HDD_ERR_ARG; }
    for (int i = 0; i < count; ++i) {
        std::string why;
        if (!desc_ok(descs + i, why)) { g_create_error = "hdd_create_batch: problem " + std::to_string(i) + ": " + why; return HDD_ERR_ARG; }
    }
    try {
        std::vector<std::shared_ptr<ProblemHost>> probs;
        for (int i = 0; i < count; ++i) {
            const hdd_problem_desc* d = descs + i;
            probs.push_back(std::make_shared<ProblemHost>(params_of(d), d->n_pages, d->page_offsets, d->flags, d->b));
        }
        auto* c = new hdd_ctx();
        try {
            c->solver.reset(new Solver(probs));
        } catch (...) { delete c; throw; }
        *out = c;
        return HDD_OK;
    } catch (std::exception& e) {
        g_create_error = e.what();
        return HDD_ERR_SETUP;
    }
}

hdd_status hdd_create_batch_from(hdd_ctx* const* ctxs, int count, hdd_ctx** out) {
    if (!ctxs || !out || count < 1) { g_create_error = "hdd_create_batch_from: invalid arguments"; return HDD_ERR_ARG; }
    try {
        std::vector<std::shared_ptr<ProblemHost>> probs;
        for (int i = 0; i < count; ++i) {
            if (!ctxs[i]) throw std::runtime_error("hdd_create_batch_from: null context");
            for (auto& q : ctxs[i]->solver->prob) probs.push_back(q);
        }
        auto* c = new hdd_ctx();
        try {
            c->solver.reset(new Solver(probs));
        } catch (...) { delete c; throw; }
        *out = c;
        return HDD_OK;
    } catch (std::exception& e) {
        g_create_error = e.what();
        return HDD_ERR_SETUP;
    }
}

hdd_status hdd_batch_count(const hdd_ctx* ctx, int* count) {
    if (!ctx || !count) return HDD_ERR_ARG;
    *count = ctx->solver->nprob();
    return HDD_OK;
}

hdd_status hdd_batch_num_pages(const hdd_ctx* ctx, int problem, int* n) {
    if (!ctx || !n || problem < 0 || problem >= ctx->solver->nprob()) return HDD_ERR_ARG;
    *n = ctx->solver->prob[problem]->lv[0].npages();
    return HDD_OK;
}

hdd_status hdd_batch_page_offsets(const hdd_ctx* ctx, int problem, uint64_t* out) {
    if (!ctx || !out || problem < 0 || problem >= ctx->solver->nprob()) return HDD_ERR_ARG;
    const PagedLevel& L = ctx->solver->prob[problem]->lv[0];
    for (int p = 0; p < L.npages(); ++p) out[p] = page_key(L.page_base[p].x >> 2, L.page_base[p].y >> 3, L.page_base[p].z >> 3) << 12;
    return HDD_OK;
}

hdd_status hdd_batch_pcg_solve(hdd_ctx* ctx, const hdd_pcg_params* prm, hdd_pcg_report* rep, double* hist, int hist_len) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || !prm || hist_len < 0) return HDD_ERR_ARG;
    const int np = ctx->solver->nprob();
    for (int p = 0; p < np; ++p)
        if (prm[p].max_iterations < 0) return HDD_ERR_ARG;
    try {
        std::vector<double> tol(np);
        std::vector<int> mit(np);
        for (int p = 0; p < np; ++p) { tol[p] = prm[p].tolerance; mit[p] = prm[p].max_iterations; }
        std::vector<PcgResult> res;
        ctx->solver->pcg_batch(tol, mit, res);
        bool all = true;
        for (int p = 0; p < np; ++p) {
            const PcgResult& r = res[p];
            if (rep) {
                rep[p].iterations = r.iterations; rep[p].seconds = r.seconds;
                rep[p].final_norm = r.hist.empty() ? 0.0 : r.hist.back(); rep[p].converged = r.converged;
            }
            if (hist) for (int i = 0; i < hist_len; ++i) hist[(size_t)p * hist_len + i] = i < (int)r.hist.size() ? r.hist[i] : 0.0;
            all = all && r.converged;
        }
        return all ? HDD_OK : HDD_ERR_NOT_CONVERGED;
    } catch (std::exception& e) {
        ctx->err = e.what();
        return HDD_ERR_DEVICE;
    }
}

hdd_status hdd_batches_pcg_solve(hdd_ctx* const* ctxs, int count, const hdd_pcg_params* prm, hdd_pcg_report* rep) {
    if (!ctxs || count < 1 || !prm) { g_create_error = "hdd_batches_pcg_solve: invalid arguments"; return HDD_ERR_ARG; }
    std::vector<int> first(count + 1, 0);
    for (int c = 0; c < count; ++c) {
        if (!ctxs[c] || ctxs[c]->solver->host_only) { g_create_error = "hdd_batches_pcg_solve: null or host-only context"; return HDD_ERR_ARG; }
        for (int d = 0; d < c; ++d)
            if (ctxs[d] == ctxs[c]) { g_create_error = "hdd_batches_pcg_solve: a context is listed twice"; return HDD_ERR_ARG; }
        first[c + 1] = first[c] + ctxs[c]->solver->nprob();
    }
    for (int p = 0; p < first[count]; ++p)
        if (prm[p].max_iterations < 0) { g_create_error = "hdd_batches_pcg_solve: negative iteration limit"; return HDD_ERR_ARG; }
    int c = 0;
    try {
        // every context's PCG advanced one iteration per round, each on its own stream, so their
        // kernels overlap on the device (the serving loop's driver, for contexts built in advance)
        std::vector<PcgRun> runs(count);
        for (c = 0; c < count; ++c) {
            std::vector<double> tol;
            std::vector<int> mit;
            for (int p = first[c]; p < first[c + 1]; ++p) { tol.push_back(prm[p].tolerance); mit.push_back(prm[p].max_iterations); }
            ctxs[c]->solver->pcg_start(runs[c], tol, mit);
        }
        for (bool all = false; !all;) {
            all = true;
            for (c = 0; c < count; ++c)
                if (!runs[c].done()) all = ctxs[c]->solver->pcg_advance(runs[c]) && all;
        }
        bool conv = true;
        for (c = 0; c < count; ++c)
            for (int p = first[c]; p < first[c + 1]; ++p) {
                const PcgResult& r = runs[c].out[p - first[c]];
                if (rep) {
                    rep[p].iterations = r.iterations; rep[p].seconds = r.seconds;
                    rep[p].final_norm = r.hist.empty() ? 0.0 : r.hist.back(); rep[p].converged = r.converged;
                }
                conv = conv && r.converged;
            }
        return conv ? HDD_OK : HDD_ERR_NOT_CONVERGED;
    } catch (std::exception& e) {
        if (c < count) ctxs[c]->err = e.what();
        g_create_error = e.what();
        // leave no iteration of any context in flight (its scalars land in the context's pinned buffers)
        for (int d = 0; d < count; ++d) (void)hipStreamSynchronize(ctxs[d]->solver->stream);
        return HDD_ERR_DEVICE;
    }
}

hdd_status hdd_batch_get_x(hdd_ctx* ctx, int problem, float* host) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || !host || problem < 0 || problem >= ctx->solver->nprob()) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->download_x(problem, host));
}

hdd_status hdd_set_graph(hdd_ctx* ctx, int enable) {
    if (!ctx) return HDD_ERR_ARG;
    ctx->solver->use_graph = enable != 0;
    return HDD_OK;
}

hdd_status hdd_set_vcycle_storage(hdd_ctx* ctx, int fp16) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->set_half_storage(fp16 != 0));
}

hdd_status hdd_set_dot_order(hdd_ctx* ctx, int serial) {
    if (!ctx) return HDD_ERR_ARG;
    ctx->solver->serial_dots = serial != 0;
    return HDD_OK;
}

hdd_status hdd_set_host_threads(int n) {
    if (n < 0) { g_create_error = "hdd_set_host_threads: n must be >= 0"; return HDD_ERR_ARG; }
    if (n == 0) {
        // OpenMP's own default: OMP_NUM_THREADS (its first value) if set, else the processors available
        const char* e = std::getenv("OMP_NUM_THREADS");
        const int v = e ? std::atoi(e) : 0;
        n = v > 0 ? v : omp_get_num_procs();
    }
    omp_set_num_threads(n);
    return HDD_OK;
}

hdd_status hdd_set_sigma_refinement(hdd_ctx* ctx, int steps) {
    if (!ctx) return HDD_ERR_ARG;
    if (ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (steps < 0 || steps > 3) { ctx->err = "sigma refinement steps must be 0..3"; return HDD_ERR_ARG; }
    HDD_TRY(ctx, ctx->solver->set_sigma_refinement(steps));
}

hdd_status hdd_set_pcg_precision(hdd_ctx* ctx, int fp64, int replace_every) {
    if (!ctx) return HDD_ERR_ARG;
    if (ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (replace_every < 0) { ctx->err = "replace_every must be >= 0"; return HDD_ERR_ARG; }
    ctx->solver->outer_fp64 = fp64 != 0;
    ctx->solver->replace_every = replace_every;
    return HDD_OK;
}

hdd_status hdd_get_x_f64(hdd_ctx* ctx, int problem, double* host) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || !host || problem < 0 || problem >= ctx->solver->nprob()) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->download_x_f64(problem, host));
}

hdd_status hdd_destroy(hdd_ctx* ctx) {
    if (!ctx) return HDD_ERR_ARG;
    delete ctx;
    return HDD_OK;
}

const char* hdd_last_error(const hdd_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

hdd_status hdd_num_pages(const hdd_ctx* ctx, int* n) {
    if (!ctx || !n) return HDD_ERR_ARG;
    *n = ctx->solver->lv[0].host.npages();
    return HDD_OK;
}

hdd_status hdd_page_offsets(const hdd_ctx* ctx, uint64_t* out) {
    if (!ctx || !out) return HDD_ERR_ARG;
    const PagedLevel& L = ctx->solver->lv[0].host;
    for (int p = 0; p < L.npages(); ++p) out[p] = page_key(L.page_base[p].x >> 2, L.page_base[p].y >> 3, L.page_base[p].z >> 3) << 12;
    return HDD_OK;
}

hdd_status hdd_set_channel(hdd_ctx* ctx, int ch, const float* host) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || !host || ch < 0 || ch > 6) return HDD_ERR_ARG;
    HDD_TRY(ctx, {
        Solver& S = *ctx->solver;
        const size_t n = (size_t)S.lv[0].host.npages() * kPageCells;
        hdd_check(hipMemcpyAsync(S.channel(ch), host, n * sizeof(float), hipMemcpyHostToDevice, S.stream), "set_channel");
        hdd_check(hipStreamSynchronize(S.stream), "set_channel");
    })
}

hdd_status hdd_get_channel(hdd_ctx* ctx, int ch, float* host) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || !host || ch < 0 || ch > 6) return HDD_ERR_ARG;
    HDD_TRY(ctx, {
        Solver& S = *ctx->solver;
        hdd_check(hipStreamSynchronize(S.stream), "sync");
        const size_t n = (size_t)S.lv[0].host.npages() * kPageCells;
        hdd_check(hipMemcpyAsync(host, S.channel(ch), n * sizeof(float), hipMemcpyDeviceToHost, S.stream), "get_channel");
        hdd_check(hipStreamSynchronize(S.stream), "get_channel");
    })
}

hdd_status hdd_level_pages(const hdd_ctx* ctx, int level, int* n) {
    if (!ctx || !n || level < 0 || level >= (int)ctx->solver->lv.size()) return HDD_ERR_ARG;
    *n = ctx->solver->lv[level].host.npages();
    return HDD_OK;
}

hdd_status hdd_level_flags(const hdd_ctx* ctx, int level, uint64_t* offsets, uint32_t* flags) {
    if (!ctx || level < 0 || level >= (int)ctx->solver->lv.size()) return HDD_ERR_ARG;
    const PagedLevel& L = ctx->solver->lv[level].host;
    for (int p = 0; p < L.npages(); ++p) {
        if (offsets) offsets[p] = page_key(L.page_base[p].x >> 2, L.page_base[p].y >> 3, L.page_base[p].z >> 3) << 12;
        if (flags) std::memcpy(flags + (size_t)p * kPageCells, &L.flags[(size_t)p * kPageCells], kPageCells * sizeof(uint32_t));
    }
    return HDD_OK;
}

hdd_status hdd_smooth(hdd_ctx* ctx, int mode) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || mode < 0 || mode > 2) return HDD_ERR_ARG;
    HDD_TRY(ctx, { ctx->solver->smooth0(mode); hdd_check(hipStreamSynchronize(ctx->solver->stream), "smooth"); })
}

hdd_status hdd_vcycle(hdd_ctx* ctx) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx) return HDD_ERR_ARG;
    HDD_TRY(ctx, {
        Solver& S = *ctx->solver;
        launch_fill(S.lv[0].u[S.lv[0].cur].p, 0.f, S.lv[0].cells(), S.stream);
        S.vcycle();
        hdd_check(hipStreamSynchronize(S.stream), "vcycle");
    })
}

hdd_status hdd_apply_preconditioner(hdd_ctx* ctx) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx) return HDD_ERR_ARG;
    HDD_TRY(ctx, { ctx->solver->apply_preconditioner(); hdd_check(hipStreamSynchronize(ctx->solver->stream), "apply"); })
}

// ---- Krylov plugin surface (KRYLOV_SYSTEM_BASE / CG_SYSTEM / CG_VECTOR)

static bool dev_ctx(hdd_ctx* ctx) {
    if (!ctx) return false;
    if (ctx->solver->host_only) { ctx->err = "context was created host-only"; return false; }
    return true;
}
static bool ch_ok(int c) { return c >= 0 && c <= 6; }

hdd_status hdd_set_stream(hdd_ctx* ctx, void* stream) {
    if (!dev_ctx(ctx)) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->set_stream((hipStream_t)stream));
}

hdd_status hdd_synchronize(hdd_ctx* ctx) {
    if (!dev_ctx(ctx)) return HDD_ERR_ARG;
    HDD_TRY(ctx, hdd_check(hipStreamSynchronize(ctx->solver->stream), "synchronize"));
}

hdd_status hdd_multiply(hdd_ctx* ctx, int ch_x, int ch_result) {
    if (!dev_ctx(ctx) || !ch_ok(ch_x) || !ch_ok(ch_result) || ch_x == ch_result) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->multiply(ch_x, ch_result));
}

hdd_status hdd_inner_product(hdd_ctx* ctx, int ch_a, int ch_b, double* out) {
    if (!dev_ctx(ctx) || !ch_ok(ch_a) || !ch_ok(ch_b) || !out) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->inner_product(ch_a, ch_b, out));
}

hdd_status hdd_inner_product_noalloc(const float* a, const float* b, int64_t n_pages, double* out, void* stream) {
    if (!a || !b || !out || n_pages < 1 || n_pages > HDD_NOALLOC_MAX_PAGES) {
        g_create_error = "hdd_inner_product_noalloc: null pointer, or n_pages outside [1, HDD_NOALLOC_MAX_PAGES]";
        return HDD_ERR_ARG;
    }
    try {
        launch_inner_product_noalloc(a, b, n_pages, out, static_cast<hipStream_t>(stream));
        return HDD_OK;
    } catch (std::exception& e) {
        g_create_error = e.what();
        return HDD_ERR_DEVICE;
    }
}

hdd_status hdd_convergence_norm(hdd_ctx* ctx, int ch, float* out) {
    if (!dev_ctx(ctx) || !ch_ok(ch) || !out) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->convergence_norm(ch, out));
}

hdd_status hdd_project(hdd_ctx* ctx, int ch) {
    if (!dev_ctx(ctx) || !ch_ok(ch)) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->project(ch));
}

hdd_status hdd_vector_copy(hdd_ctx* ctx, int ch_out, float c, int ch_in) {
    if (!dev_ctx(ctx) || !ch_ok(ch_out) || !ch_ok(ch_in)) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->vector_op(0, ch_out, ch_in, -1, c));
}

hdd_status hdd_vector_copy2(hdd_ctx* ctx, int ch_out, float c, int ch_in1, int ch_in2) {
    if (!dev_ctx(ctx) || !ch_ok(ch_out) || !ch_ok(ch_in1) || !ch_ok(ch_in2)) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->vector_op(1, ch_out, ch_in1, ch_in2, c));
}

hdd_status hdd_vector_plus_equal(hdd_ctx* ctx, int ch_out, int ch_in) {
    if (!dev_ctx(ctx) || !ch_ok(ch_out) || !ch_ok(ch_in)) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->vector_op(2, ch_out, ch_in, -1, 0.f));
}

hdd_status hdd_vector_minus_equal(hdd_ctx* ctx, int ch_out, int ch_in) {
    if (!dev_ctx(ctx) || !ch_ok(ch_out) || !ch_ok(ch_in)) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->vector_op(3, ch_out, ch_in, -1, 0.f));
}

hdd_status hdd_vector_scale(hdd_ctx* ctx, int ch_out, float c) {
    if (!dev_ctx(ctx) || !ch_ok(ch_out)) return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->vector_op(4, ch_out, -1, -1, c));
}

hdd_status hdd_apply_preconditioner_channels(hdd_ctx* ctx, int ch_result, int ch_rhs, int ch_tmp) {
    if (!dev_ctx(ctx) || !ch_ok(ch_result) || !ch_ok(ch_rhs) || !ch_ok(ch_tmp) || ch_result == ch_rhs || ch_result == ch_tmp ||
        ch_rhs == ch_tmp)
        return HDD_ERR_ARG;
    HDD_TRY(ctx, ctx->solver->apply_preconditioner_channels(ch_result, ch_rhs, ch_tmp));
}

hdd_status hdd_sigma_info(const hdd_ctx* ctx, int* levels, int* nsub) {
    if (!ctx) return HDD_ERR_ARG;
    if (levels) *levels = ctx->solver->sigma_host().levels;
    if (nsub) *nsub = ctx->solver->sigma_host().nsub;
    return HDD_OK;
}

hdd_status hdd_interface_dofs(const hdd_ctx* ctx, int level, int* n) {
    if (!ctx || !n || level < 0 || level >= ctx->solver->sigma_host().levels) return HDD_ERR_ARG;
    *n = ctx->solver->sigma_host().n_if[level];
    return HDD_OK;
}

hdd_status hdd_interface_coords(const hdd_ctx* ctx, int level, int* xyz) {
    if (!ctx || !xyz || level < 0 || level >= ctx->solver->sigma_host().levels) return HDD_ERR_ARG;
    const auto& v = ctx->solver->sigma_host().if_coords[level];
    for (size_t i = 0; i < v.size(); ++i) { xyz[3 * i] = v[i].x; xyz[3 * i + 1] = v[i].y; xyz[3 * i + 2] = v[i].z; }
    return HDD_OK;
}

hdd_status hdd_interface_matrix(const hdd_ctx* ctx, int which, int level, int sd, int* rows, int* cols, int* nnz, int* ptr, int* idx, float* val) {
    if (!ctx) return HDD_ERR_ARG;
    const SigmaHost& S = ctx->solver->sigma_host();
    if (level < 0 || level >= S.levels || which < 0 || which > 3) return HDD_ERR_ARG;
    if ((which == 1 || which == 2) && (sd < 0 || sd >= S.nsub)) return HDD_ERR_ARG;
    if (which == 3 && level >= S.levels - 1) return HDD_ERR_ARG;
    const Csr& M = which == 0 ? S.Arr[level] : which == 1 ? S.Aii[level][sd] : which == 2 ? S.Air[level][sd] : S.Prr[level];
    if (rows) *rows = M.rows;
    if (cols) *cols = M.cols;
    if (nnz) *nnz = (int)M.nnz();
    if (ptr) std::memcpy(ptr, M.ptr.data(), M.ptr.size() * sizeof(int));
    if (idx) std::memcpy(idx, M.idx.data(), M.idx.size() * sizeof(int));
    if (val) std::memcpy(val, M.val.data(), M.val.size() * sizeof(float));
    return HDD_OK;
}

hdd_status hdd_sigma_vcycle(hdd_ctx* ctx, const float* b, float* u) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || !b || !u) return HDD_ERR_ARG;
    HDD_TRY(ctx, { sigma_vcycle_host(*ctx->solver->sigma, b, u, ctx->solver->stream); })
}

hdd_status hdd_sigma_smooth(hdd_ctx* ctx, int level, const float* b, float* u) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || !b || !u || level < 0 || level >= ctx->solver->sigma_host().levels) return HDD_ERR_ARG;
    HDD_TRY(ctx, { sigma_smooth_host(*ctx->solver->sigma, level, b, u, ctx->solver->stream); })
}

hdd_status hdd_pcg_solve(hdd_ctx* ctx, const hdd_pcg_params* prm, hdd_pcg_report* rep, double* hist, int hist_len) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || !prm || prm->max_iterations < 0) return HDD_ERR_ARG;
    try {
        std::vector<double> h;
        double secs = 0;
        bool conv = false;
        int it = ctx->solver->pcg(prm->tolerance, prm->max_iterations, h, &secs, &conv);
        if (rep) { rep->iterations = it; rep->seconds = secs; rep->final_norm = h.empty() ? 0.0 : h.back(); rep->converged = conv; }
        if (hist) for (int i = 0; i < (int)h.size() && i < hist_len; ++i) hist[i] = h[i];
        return conv ? HDD_OK : HDD_ERR_NOT_CONVERGED;
    } catch (std::exception& e) {
        ctx->err = e.what();
        return HDD_ERR_DEVICE;
    }
}

hdd_status hdd_setup_seconds(const hdd_ctx* ctx, double* s) {
    if (!ctx || !s) return HDD_ERR_ARG;
    *s = ctx->solver->setup_seconds;
    return HDD_OK;
}

hdd_status hdd_profile_sweep(hdd_ctx* ctx, int reps, double* avg_ms, double* bytes) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || reps <= 0 || !avg_ms) return HDD_ERR_ARG;
    HDD_TRY(ctx, {
        Solver& S = *ctx->solver;
        hipEvent_t a, b;
        hdd_check(hipEventCreate(&a)); hdd_check(hipEventCreate(&b));
        S.smooth0(2);  // warm
        hdd_check(hipEventRecord(a, S.stream));
        for (int i = 0; i < reps; ++i) S.smooth0(2);
        hdd_check(hipEventRecord(b, S.stream));
        hdd_check(hipEventSynchronize(b));
        float ms = 0.f;
        hdd_check(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a); (void)hipEventDestroy(b);
        *avg_ms = ms / reps;
        if (bytes) *bytes = (S.half_storage ? 10.0 : 16.0) * kPageCells * S.lv[0].host.npages();
    })
}

hdd_status hdd_profile_sweep_grid(const hdd_problem_desc* d, int reps, double* avg_ms, double* bytes) {
    std::string why;
    if (!d || reps <= 0 || !avg_ms || !desc_ok(d, why)) {
        g_create_error = "hdd_profile_sweep_grid: " + (d && reps > 0 && avg_ms ? why : std::string("invalid arguments"));
        return HDD_ERR_ARG;
    }
    try {
        *avg_ms = profile_sweep_pages({d->size[0], d->size[1], d->size[2]}, (int)d->n_pages, d->page_offsets, d->flags, reps, bytes);
        return HDD_OK;
    } catch (std::exception& e) {
        g_create_error = e.what();
        return HDD_ERR_DEVICE;
    }
}

hdd_status hdd_profile_copy(hdd_ctx* ctx, int reps, double* avg_ms, double* bytes) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || reps <= 0 || !avg_ms) return HDD_ERR_ARG;
    HDD_TRY(ctx, {
        Solver& S = *ctx->solver;
        LevelBuffers& L = S.lv[0];
        const size_t n = L.cells();
        // the smoother's spare ping-pong buffer is overwritten by the next sweep before it is read
        float* dst = L.u[1 - L.cur].p;
        const float* src = L.u[L.cur].p;
        hipEvent_t a, b;
        hdd_check(hipEventCreate(&a)); hdd_check(hipEventCreate(&b));
        launch_copy(dst, src, n, S.stream);
        hdd_check(hipEventRecord(a, S.stream));
        for (int i = 0; i < reps; ++i) launch_copy(dst, src, n, S.stream);
        hdd_check(hipEventRecord(b, S.stream));
        hdd_check(hipEventSynchronize(b));
        float ms = 0.f;
        hdd_check(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a); (void)hipEventDestroy(b);
        *avg_ms = ms / reps;
        if (bytes) *bytes = 8.0 * n;
    })
}

hdd_status hdd_stats(const hdd_ctx* ctx, double* out, int n) {
    if (!ctx || !out) return HDD_ERR_ARG;
    const Solver& S = *ctx->solver;
    const SigmaDevice* G = S.sigma.get();
    int nsub = 0;
    for (auto& q : S.prob) nsub += q->sigma.nsub;
    // fused heights of every triangular-solve plan, and how many use the wavefront-per-run form
    double fused = 0, waved = 0;
    if (G) {
        auto count = [&](const TriPlan& T) {
            for (size_t h = 0; h < T.fused.size(); ++h)
                if (T.fused[h]) { fused += 1; waved += (h < T.wave.size() && T.wave[h]) ? 1 : 0; }
        };
        for (const SigmaLevel& L : G->lv) { count(L.Lii); count(L.Lrr); }
        count(G->Lbottom);
    }
    double v[18] = {S.setup_seconds, G ? G->factor_seconds : 0.0, G ? (double)G->plan_bytes : 0.0,
                    (double)S.lv[0].host.npages(), (G && G->levels) ? (double)G->lv[0].n_if : (S.sigma_host().levels ? (double)S.sigma_host().n_if[0] : 0.0),
                    (G && G->levels) ? (double)G->lv[0].n_int : 0.0, (double)nsub,
                    G ? G->bottom_rho : -1.0, G ? (double)G->bottom_direct : 0.0, G ? (double)G->bottom_k : 0.0,
                    G ? G->bottom_seconds : 0.0, G ? (double)G->bottom_steps : 0.0, G ? G->bottom_check : -1.0,
                    G ? (double)G->tuned_choices : 0.0, G ? (double)G->dense_top_rows : 0.0, S.bottom_local ? 1.0 : 0.0,
                    fused, waved};
    for (int i = 0; i < n && i < 18; ++i) out[i] = v[i];
    return HDD_OK;
}

hdd_status hdd_profile_solve(hdd_ctx* ctx, int level, int which, int reps, double* avg_ms, double* bytes) {
    if (ctx && ctx->solver->host_only) { ctx->err = "context was created host-only"; return HDD_ERR_ARG; }
    if (!ctx || reps <= 0 || !avg_ms || which < 0 || which > 2) return HDD_ERR_ARG;
    HDD_TRY(ctx, {
        Solver& S = *ctx->solver;
        if (!S.sigma || level < 0 || level >= S.sigma->levels) throw std::runtime_error("no such sigma level");
        SigmaLevel& L = S.sigma->lv[level];
        TriPlan* P = which == 0 ? &L.Lii : which == 1 ? &L.Lrr : &S.sigma->Lbottom;
        if (which == 2 && (level != S.sigma->levels - 1 || !S.sigma->bottom_direct)) throw std::runtime_error("no combined plan at this level");
        if (P->n == 0) throw std::runtime_error("empty plan");
        DBuf<float> bv, xv;  // zero right-hand side: the solve's cost does not depend on the values
        bv.alloc(P->n); xv.alloc(P->n);
        const float* b = bv.p;
        float* x = xv.p;
        hipEvent_t a, e;
        hdd_check(hipEventCreate(&a)); hdd_check(hipEventCreate(&e));
        P->solve(b, x, S.stream);
        hdd_check(hipEventRecord(a, S.stream));
        for (int i = 0; i < reps; ++i) P->solve(b, x, S.stream);
        hdd_check(hipEventRecord(e, S.stream));
        hdd_check(hipEventSynchronize(e));
        float ms = 0.f;
        hdd_check(hipEventElapsedTime(&ms, a, e));
        (void)hipEventDestroy(a); (void)hipEventDestroy(e);
        *avg_ms = ms / reps;
        if (bytes) *bytes = 8.0 * (double)P->factor_nnz;
    })
}

static constexpr int kPlanStatsPerHeight = 8;

static void plan_stats_impl(const SigmaHost& H, int level, int which, double* out, int cap, int* heights) {
    if (level < 0 || level >= H.levels) throw std::runtime_error("no such sigma level");
    std::vector<SupernodalFactor> f;
    std::vector<int> offs;
    if (which == 0) {
        f.resize(H.nsub);
        ParallelErrors errs;
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < H.nsub; ++s)
            if (H.Aii[level][s].rows > 0) errs.run([&] { factor_spd(H.Aii[level][s], H.Aii_coords[level][s], 1e-5, f[s]); });
        errs.rethrow();
        int off = 0;
        for (int s = 0; s < H.nsub; ++s) { offs.push_back(off); off += H.Aii[level][s].rows; }
    } else {
        f.resize(1);
        std::vector<Int3> rc(H.if_coords[level].size());
        for (size_t i = 0; i < rc.size(); ++i) rc[i] = {H.if_coords[level][i].x << level, H.if_coords[level][i].y << level, H.if_coords[level][i].z << level};
        factor_spd(H.Arr[level], rc, 1e-5, f[0]);
        offs.push_back(0);
    }
    std::vector<const SupernodalFactor*> fp;
    std::vector<int> fo;
    for (size_t s = 0; s < f.size(); ++s) if (f[s].n) { fp.push_back(&f[s]); fo.push_back(offs[s]); }
    TriPlanHost P;
    build_triplan(fp, fo, P);
    *heights = P.H + 1;
    constexpr int K = kPlanStatsPerHeight;
    std::vector<double> v(K * (size_t)(P.H + 1), 0.0);
    for (int h = 0; h <= P.H; ++h) {
        for (int q = P.sn_begin[h]; q < P.sn_begin[h + 1]; ++q) {
            const int g = P.sn_list[q];
            v[K * h + 0] += 1; v[K * h + 1] += P.sz[g];
            v[K * h + 2] += 0.5 * P.sz[g] * (P.sz[g] + 1.0); v[K * h + 3] += (double)P.sz[g] * P.nr[g];
            v[K * h + 5] = std::max(v[K * h + 5], (double)P.nr[g]);
            for (int i = P.s0[g]; i < P.s0[g] + P.sz[g]; ++i) {
                const int inc = P.inc_ptr[i + 1] - P.inc_ptr[i];
                v[K * h + 6] = std::max(v[K * h + 6], (double)inc);
                v[K * h + 7] += inc;
            }
        }
        v[K * h + 4] = P.max_sz[h];
    }
    for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
}

hdd_status hdd_sigma_plan_stats(hdd_ctx* ctx, int level, int which, double* out, int cap, int* heights) {
    if (!ctx || !heights || cap < 0 || (cap > 0 && !out) || (which != 0 && which != 1)) return HDD_ERR_ARG;
    HDD_TRY(ctx, plan_stats_impl(ctx->solver->sigma_host(), level, which, out, cap, heights));
}

hdd_status hdd_sigma_factor_solve(hdd_ctx* ctx, int level, int which, int subdomain, const double* b, double* x) {
    if (!ctx || !b || !x || (which != 0 && which != 1)) return HDD_ERR_ARG;
    HDD_TRY(ctx, {
        const SigmaHost& H = ctx->solver->sigma_host();
        if (level < 0 || level >= H.levels) throw std::runtime_error("no such sigma level");
        if (which == 1 && (subdomain < 0 || subdomain >= H.nsub)) throw std::runtime_error("no such subdomain");
        ProblemHost& P = *ctx->solver->prob[0];
        P.factorize();
        const SigmaFactors::Level& F = P.factors->lv[level];
        const SupernodalFactor& f = which == 0 ? F.Arr : F.Aii[F.rep[subdomain]];
        if (which == 1 && H.Aii[level][subdomain].rows == 0) return HDD_OK;
        factor_solve(f, b, x);
    })
}

hdd_status hdd_bottom_spectrum(hdd_ctx* ctx, double lambda_cut, double* lambdas, int cap, int* k, double* lambda_max) {
    if (!ctx || !k || cap < 0 || (cap > 0 && !lambdas) || !(lambda_cut > 0.0 && lambda_cut < 1.0)) return HDD_ERR_ARG;
    HDD_TRY(ctx, {
        const SigmaHost& H = ctx->solver->sigma_host();
        if (H.levels == 0) throw std::runtime_error("no sigma levels");
        const int l = H.levels - 1;
        Csr C; std::vector<Int3> xyz;
        combined_bottom(H, l, C, xyz);
        SupernodalFactor fc;
        factor_spd(C, xyz, 1e-5, fc);
        BottomSpectrum sp;
        bottom_spectrum(H.Arr[l], 1e-5, fc, C.rows - H.n_if[l], lambda_cut, size_t(2) << 30, sp);
        if (!sp.resolved) throw std::runtime_error("bottom spectrum not resolved (too many slow modes or basis exhausted)");
        *k = (int)sp.lambda.size();
        for (int j = 0; j < *k && j < cap; ++j) lambdas[j] = sp.lambda[j];
        if (lambda_max) *lambda_max = sp.lambda_max;
    })
}

// ---- serving

struct hdd_server {
    std::unique_ptr<Server> srv;
    mutable std::mutex mu;
    std::string err;
};

hdd_status hdd_server_create(const hdd_server_params* p, hdd_server** out) {
    if (!p || !out || p->setup_threads < 1 || p->groups < 1 || p->max_batch < 1 || p->hbm_budget_bytes < 0) {
        g_create_error = "hdd_server_create: invalid arguments";
        return HDD_ERR_ARG;
    }
    try {
        ServerParams sp;
        sp.setup_threads = p->setup_threads; sp.groups = p->groups; sp.max_batch = p->max_batch; sp.hbm_budget = p->hbm_budget_bytes;
        auto* s = new hdd_server();
        try {
            s->srv.reset(new Server(sp));
        } catch (...) { delete s; throw; }
        *out = s;
        return HDD_OK;
    } catch (std::exception& e) {
        g_create_error = e.what();
        return HDD_ERR_DEVICE;
    }
}

hdd_status hdd_server_submit(hdd_server* s, const hdd_problem_desc* d, const hdd_pcg_params* prm, int64_t* id) {
    if (!s || !prm || !id || prm->max_iterations < 0) return HDD_ERR_ARG;
    std::string why;
    if (!desc_ok(d, why)) { std::lock_guard<std::mutex> lk(s->mu); s->err = "hdd_server_submit: " + why; return HDD_ERR_ARG; }
    try {
        *id = s->srv->submit(params_of(d), d->n_pages, d->page_offsets, d->flags, d->b, prm->tolerance, prm->max_iterations);
        return HDD_OK;
    } catch (std::exception& e) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = e.what();
        return HDD_ERR_ARG;
    }
}

hdd_status hdd_server_wait(hdd_server* s, int64_t id, float* x, hdd_request_report* rep) {
    if (!s) return HDD_ERR_ARG;
    try {
        std::shared_ptr<ServerRequest> r = s->srv->wait(id);
        s->srv->forget(id);
        const bool ran = r->t_admit > 0;
        if (rep) {
            rep->iterations = r->iterations; rep->final_norm = r->final_norm;
            rep->latency_s = r->t_done - r->t_submit;
            rep->setup_s = r->t_ready > 0 ? r->t_ready - r->t_submit : 0.0;
            rep->queue_s = ran ? r->t_admit - r->t_ready : 0.0;
            rep->solve_s = ran ? r->t_done - r->t_admit : 0.0;
            rep->batch_size = r->batch_size;
            rep->concurrent = r->concurrent;
            rep->cancelled = r->cancelled;
        }
        if (r->failed) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->err = r->error;
            return ran ? HDD_ERR_DEVICE : HDD_ERR_SETUP;
        }
        if (x) {
            if (r->x.empty()) std::memset(x, 0, r->offsets.size() * kPageCells * sizeof(float));
            else std::memcpy(x, r->x.data(), r->x.size() * sizeof(float));
        }
        if (r->cancelled) return HDD_ERR_CANCELLED;
        return r->converged ? HDD_OK : HDD_ERR_NOT_CONVERGED;
    } catch (std::exception& e) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = e.what();
        return HDD_ERR_ARG;
    }
}

hdd_status hdd_server_cancel(hdd_server* s, int64_t id) {
    if (!s) return HDD_ERR_ARG;
    if (s->srv->cancel(id)) return HDD_OK;
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = "hdd_server_cancel: no such request, or it is already done";
    return HDD_ERR_ARG;
}

hdd_status hdd_server_destroy(hdd_server* s) {
    if (!s) return HDD_ERR_ARG;
    delete s;
    return HDD_OK;
}

// The server's message is shared by the threads that submit and wait; the
// caller gets a copy taken under the lock, kept per calling thread.
const char* hdd_server_last_error(const hdd_server* s) {
    if (!s) return g_create_error.c_str();
    static thread_local std::string copy;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        copy = s->err;
    }
    return copy.c_str();
}

hdd_status hdd_server_budget(const hdd_server* s, double* bytes) {
    if (!s || !bytes) return HDD_ERR_ARG;
    *bytes = s->srv->budget();
    return HDD_OK;
}

}  // extern "C"
