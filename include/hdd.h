/*
 * hdd.h — C ABI of the MI355X hierarchical domain-decomposition PCG solver.
 *
 * The library stands in for the reference's Domain_Decomposition PCG path
 * (demos/Domain_Decomposition_CPU_Only). The reference has no C ABI; its
 * interface is the C++ solver facade and the Krylov plugin classes. Each entry
 * point below names the reference function it replaces.
 *
 * Conventions: plain pointers and sizes only; every call returns hdd_status
 * (0 = success) and never throws; hdd_last_error() returns the message of the
 * last failing call on a context (or of hdd_create when ctx is NULL). One
 * context per host thread; calls on one context are serialised on the
 * context's HIP stream. All device memory is owned by the context.
 *
 * Data layout: SPGrid pages of 4(x) x 8(y) x 8(z) cells (z fastest inside a
 * page). `page_offsets` are SPGrid linear offsets of the pages (Morton page
 * index << 12), as stored in the reference's `block_offset` file
 * (DDS/SPGRID_READ_WRITE.h:95-106); `flags` and `b` hold 256 values per page
 * in page-local order, as in its `flags` and `data` files.
 */
#ifndef HDD_H
#define HDD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int hdd_status;
enum { HDD_OK = 0, HDD_ERR_ARG = 1, HDD_ERR_SETUP = 2, HDD_ERR_DEVICE = 3, HDD_ERR_NOT_CONVERGED = 4, HDD_ERR_CANCELLED = 5 };

typedef struct hdd_ctx hdd_ctx;

/* Problem description: the reference's PARSE_ARGS flags
 * (CPU/Test_Bench_3D/main.cpp:235-252) plus the SPGrid input arrays. */
typedef struct hdd_problem_desc {
    int size[3];            /* -size (solver size = generator size + 1) */
    int sub_size[3];        /* -sub_size */
    int vlevels;            /* -vlevels */
    int sigma_levels;       /* -sigma_levels */
    int nVcycles;           /* -nVcycles */
    int interior_smoothing; /* -interior_smoothing */
    int boundary_smoothing; /* -boundary_smoothing */
    int sigma_smoothing;    /* -sigma_smoothing */
    int sigma_mg;           /* -sigma_mg */
    int n_pages;
    const uint64_t* page_offsets; /* [n_pages] */
    const uint32_t* flags;        /* [n_pages*256] */
    const float* b;               /* [n_pages*256] right-hand side */
} hdd_problem_desc;

typedef struct hdd_pcg_params {
    double tolerance;  /* stop when max|r| <= tolerance (absolute, L-infinity) */
    int max_iterations;
} hdd_pcg_params;

typedef struct hdd_pcg_report {
    int iterations;        /* negative if not converged within max_iterations */
    double seconds;        /* wall time of the solve on the context stream */
    double final_norm;     /* max|r| at exit */
    int converged;         /* 1 if max|r| <= tolerance at exit (the status is HDD_OK exactly then) */
} hdd_pcg_report;

/* Replaces SPGrid_Domain_Decomposition_Solver::Read_SPGrid + Initialize_Preconditioner
 * (DDS/SPGrid_Domain_Decomposition_Solver.h:73-83, :121-161): reads the pages,
 * splits subdomains, builds the interface operators and their factors, the
 * V-cycle topology and flags, and uploads everything to the current device. */
hdd_status hdd_create(const hdd_problem_desc* desc, hdd_ctx** out);

/* Same setup as hdd_create, host side only: no device memory is touched and
 * every device entry point on the context returns HDD_ERR_ARG. For inspecting
 * flags and interface operators on machines without a GPU. */
hdd_status hdd_create_host_only(const hdd_problem_desc* desc, hdd_ctx** out);

/* Batch of the problems already set up in other contexts (hdd_create or
 * hdd_create_host_only), in order: their host setup and factors are shared,
 * not repeated, and the source contexts stay valid. Otherwise as
 * hdd_create_batch (no reference counterpart). */
hdd_status hdd_create_batch_from(hdd_ctx* const* ctxs, int count, hdd_ctx** out);

/* Batch of independent problems in one context (no reference counterpart:
 * the reference solves one problem per process). Each problem is set up as
 * by hdd_create; their pages share one device grid and their interface
 * operators one block-diagonal interface solver, so every V-cycle stage,
 * triangular-solve height and PCG vector operation is one launch for all of
 * them. Every problem needs the same vlevels, sigma_levels and smoothing
 * counts; sizes, sub_size and geometry may differ. Each problem keeps its own
 * PCG scalars and stopping test (hdd_batch_pcg_solve). The single-problem
 * entry points act on the combined grid (channels, level flags, statistics);
 * interface-operator introspection refers to problem 0. */
hdd_status hdd_create_batch(const hdd_problem_desc* descs, int count, hdd_ctx** out);

/* Number of problems in a context (1 for hdd_create). */
hdd_status hdd_batch_count(const hdd_ctx* ctx, int* count);

/* Level-0 pages of one problem and their SPGrid offsets (Morton order,
 * the problem's own coordinates). */
hdd_status hdd_batch_num_pages(const hdd_ctx* ctx, int problem, int* n_pages);
hdd_status hdd_batch_page_offsets(const hdd_ctx* ctx, int problem, uint64_t* out);

/* PCG from x = 0 for every problem at once: params[p], reports[p] per
 * problem; residual_history (may be NULL) receives max|r| per iteration,
 * history_len values per problem (row p at p*history_len, zero-padded).
 * Per problem the iteration is CONJUGATE_GRADIENT_DUMPING::Solve
 * (DDS/CONJUGATE_GRADIENT_DUMPING.cpp:146-212); a problem that has met its
 * tolerance keeps its x while the others continue. Returns
 * HDD_ERR_NOT_CONVERGED if any problem stopped at its iteration limit. */
hdd_status hdd_batch_pcg_solve(hdd_ctx* ctx, const hdd_pcg_params* params, hdd_pcg_report* reports,
                               double* residual_history, int history_len);

/* PCG from x = 0 for every problem of several batch (or single-problem)
 * contexts at once: each context's iterations run on its own stream, one
 * iteration per context per round, so the contexts' kernels overlap on the
 * device. params and reports hold one entry per problem, the contexts in
 * order and each context's problems in order. Each problem's result is the
 * one hdd_batch_pcg_solve gives it on its context. HDD_ERR_NOT_CONVERGED if
 * any problem stopped at its limit; on failure hdd_last_error(NULL) and the
 * failing context's hdd_last_error hold the message. No reference
 * counterpart. */
hdd_status hdd_batches_pcg_solve(hdd_ctx* const* ctxs, int count, const hdd_pcg_params* params, hdd_pcg_report* reports);

/* x of one problem: n_pages(problem)*256 floats in the problem's page order. */
hdd_status hdd_batch_get_x(hdd_ctx* ctx, int problem, float* host);

/* enable = 1: the preconditioner apply is replayed from a captured HIP graph
 * (one per state of the smoothing ping-pong buffers); 0 (default): every
 * kernel is launched from the host (same kernels in the same order, results
 * bit-identical). Capture needs no other thread of the process to use the
 * null stream meanwhile. No reference counterpart. */
hdd_status hdd_set_graph(hdd_ctx* ctx, int enable);

/* V-cycle storage of the preconditioner (no reference counterpart; the
 * reference stores everything in float): fp16 = 0 (default) keeps the
 * reference's FP32 V-cycles; fp16 = 1 stores the V-cycle vectors u, b and r of
 * every level in FP16 with FP32 arithmetic (the smoothing sweep then moves
 * 10 B per cell instead of 16). The preconditioner is linear, so it runs on
 * r / max|r| of each problem and scales its result back, keeping the stored
 * values inside FP16's normal range; the interface solver, the PCG vectors and
 * reductions stay FP32/FP64. hdd_smooth and hdd_profile_sweep then act on the
 * FP16 buffers; hdd_vcycle stays FP32. */
hdd_status hdd_set_vcycle_storage(hdd_ctx* ctx, int fp16);

/* Accuracy mode of the interface solver (no reference counterpart): steps = 0
 * (default) solves every (A + 1e-5 I) system of SIGMA_SMOOTHER::Smooth /
 * Compute_Residual (IFS/SIGMA_SMOOTHER.cpp:292-352) once with FP32 factors, the
 * precision of the reference's single-precision PARDISO (iparm[27] = 1,
 * IFS/SIGMA_SMOOTHER.cpp:151); steps = 1..3 follows each solve with that many
 * steps of iterative refinement, x += (A + 1e-5 I)^{-1} (b - (A + 1e-5 I) x),
 * the residual accumulated in FP64, so the solves are accurate to FP32
 * rounding of their result (each step costs one more solve). */
hdd_status hdd_set_sigma_refinement(hdd_ctx* ctx, int steps);

/* Summation order of the inner products of the PCG loop and hdd_inner_product:
 * serial = 0 (default) sums each problem's products in double over a fixed
 * tree of blocks (deterministic, full bandwidth); serial = 1 sums them in one
 * running double sum in page order and cell order, the order of the
 * reference's CG_SYSTEM::Inner_Product loop (DDS/CG_SYSTEM.cpp:62-72), so the
 * float values of rho and s.q round exactly as the reference's do (one
 * workgroup per problem: milliseconds per product on large grids; for parity
 * checks against the reference's own rounding). */
hdd_status hdd_set_dot_order(hdd_ctx* ctx, int serial);

/* Host threads of the set-up work (interface generator, factorisations, plan
 * building) that the calling thread starts from now on: n >= 1 threads, or
 * n = 0 for all the host's (the default). Per calling thread, so a caller
 * that sets several problems up side by side can give each its share (the
 * server's set-up workers take theirs themselves). No reference counterpart. */
hdd_status hdd_set_host_threads(int n);

/* Precision of the PCG loop (no reference counterpart): fp64 = 0 (default)
 * keeps CONJUGATE_GRADIENT_DUMPING's float rho, s.q, alpha and beta
 * (DDS/CONJUGATE_GRADIENT_DUMPING.cpp:193-205), with its true residual only at
 * the start (:156-162); fp64 = 1 forms those scalars in FP64, accumulates x in
 * FP64 (the x channel receives its float copy), and replaces the recursive
 * residual by the true residual b - L x of the FP64 iterate every
 * `replace_every` iterations (0: never on a schedule) and whenever it passes
 * the stopping test, which is then taken on the true residual, so a solve
 * that stops has max|b - L x| <= tolerance. Vectors r, s, q, z and the
 * preconditioner stay FP32. */
hdd_status hdd_set_pcg_precision(hdd_ctx* ctx, int fp64, int replace_every);

/* The FP64 iterate of the last solve with fp64 = 1: n_pages(problem)*256
 * doubles in the problem's page order. */
hdd_status hdd_get_x_f64(hdd_ctx* ctx, int problem, double* host);

/* Releases all device and host memory of a context. */
hdd_status hdd_destroy(hdd_ctx* ctx);

/* Message of the last failure (ctx may be NULL for hdd_create failures). */
const char* hdd_last_error(const hdd_ctx* ctx);

/* Number of level-0 pages (page order = Morton order of page_offsets). */
hdd_status hdd_num_pages(const hdd_ctx* ctx, int* n_pages);

/* Copies the context's level-0 page offsets (Morton order) into out[n_pages]. */
hdd_status hdd_page_offsets(const hdd_ctx* ctx, uint64_t* out);

/* Channel ids of the level-0 grid, matching the reference's field assignment
 * in Read_SPGrid (DDS/SPGrid_Domain_Decomposition_Solver.h:75-82). */
enum { HDD_CH_Z = 0, HDD_CH_R = 1, HDD_CH_K = 2, HDD_CH_X = 3, HDD_CH_B = 4, HDD_CH_Q = 5, HDD_CH_S = 6 };

/* Host <-> device copies of a level-0 channel, n_pages*256 floats in page order. */
hdd_status hdd_set_channel(hdd_ctx* ctx, int channel, const float* host);
hdd_status hdd_get_channel(hdd_ctx* ctx, int channel, float* host);

/* Level-l flags after setup (n = pages of level l, out[n*256]); pages of
 * level l in Morton order, offsets into offsets_out[n] when not NULL. */
hdd_status hdd_level_pages(const hdd_ctx* ctx, int level, int* n_pages);
hdd_status hdd_level_flags(const hdd_ctx* ctx, int level, uint64_t* offsets_out, uint32_t* flags_out);

/* One damped-Jacobi sweep on level 0 of z (channel 0) with rhs r (channel 1):
 * mode 0 interior, 1 boundary, 2 all Active. Replaces V_Cycle_Helper::
 * {interior,boundary,bottom}_smoothing (LIB/SPGrid_V_Cycle_Helper.h:69-112). */
hdd_status hdd_smooth(hdd_ctx* ctx, int mode);

/* z = 0, then one V-cycle for z from r. Replaces V_Cycle_Helper::v_cycle
 * (LIB/SPGrid_V_Cycle_Helper.h:156-192). */
hdd_status hdd_vcycle(hdd_ctx* ctx);

/* z = M^{-1} r followed by Project(z). Replaces SPGrid_Domain_Decomposition_
 * Preconditioner::Apply_Preconditioner (DDS/SPGrid_Domain_Decomposition_Preconditioner.h:88-128)
 * as called through KRYLOV_SYSTEM_BASE::Precondition. */
hdd_status hdd_apply_preconditioner(hdd_ctx* ctx);

/* ---- Krylov plugin surface on device channels. The reference's PCG
 * (CONJUGATE_GRADIENT_DUMPING::Solve, DDS/CONJUGATE_GRADIENT_DUMPING.cpp:146-212)
 * drives its problem through KRYLOV_SYSTEM_BASE (PhysBAM_Tools/Krylov_Solvers/
 * KRYLOV_SYSTEM_BASE.h:15-43), implemented by CG_SYSTEM (DDS/CG_SYSTEM.cpp:29-144),
 * and KRYLOV_VECTOR_BASE, implemented by CG_VECTOR (DDS/CG_VECTOR.cpp:18-74), whose
 * vectors are SPGrid channels. The entry points below are those operations on the
 * context's level-0 channels (HDD_CH_*), resident in device memory: a Krylov
 * loop written against them moves no vector across PCIe. They are enqueued on
 * the context's stream (hdd_set_stream) and return before the work is done,
 * except hdd_inner_product and hdd_convergence_norm, which wait for their
 * values. On a batch context each reduction returns one value per problem and
 * a scalar argument applies to every problem. include/hdd_krylov.hpp wraps them
 * in CG_SYSTEM- and CG_VECTOR-shaped C++ classes. */

/* Work of later calls goes to `stream` (a hipStream_t; NULL: the context's
 * own non-blocking stream, the default). Waits for the work already queued on
 * the previous stream. The caller keeps ownership of its stream. */
hdd_status hdd_set_stream(hdd_ctx* ctx, void* stream);

/* Waits for all work queued on the context's stream. */
hdd_status hdd_synchronize(hdd_ctx* ctx);

/* result = L x on Active|Interface cells, double accumulation, other cells
 * unchanged. Replaces CG_SYSTEM::Multiply -> SPGrid_Laplace
 * (DDS/CG_SYSTEM.cpp:29-44, DDS/SPGrid_Gemini_Arithmetic_Helpers.h:210-263). */
hdd_status hdd_multiply(hdd_ctx* ctx, int ch_x, int ch_result);

/* out[p] = sum over every cell of problem p of (float)(a*b), summed in double.
 * Replaces CG_SYSTEM::Inner_Product (DDS/CG_SYSTEM.cpp:50-73). */
hdd_status hdd_inner_product(hdd_ctx* ctx, int ch_a, int ch_b, double* out);

/* The same inner product with no allocation and no context. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). a and b are device arrays of n_pages*256 floats (the pages of a
 * level-0 channel in page order); *out (device memory) receives the sum over
 * every cell of (float)(a*b), summed in double, with the bits hdd_inner_product
 * gives on a one-problem context holding the same pages. Enqueued on `stream`
 * (a hipStream_t; NULL: the null stream) and returns before it runs. The
 * partial sums live in a file-scope device array of 1024 doubles (8 KiB), one
 * per device, held until the process ends; nothing is allocated or freed.
 * Calls that overlap on one device race on that array. n_pages must be in
 * [1, HDD_NOALLOC_MAX_PAGES]; hdd_last_error(NULL) gives the message of a
 * failure. */
#define HDD_NOALLOC_MAX_PAGES (1LL << 24)
hdd_status hdd_inner_product_noalloc(const float* a, const float* b, int64_t n_pages, double* out, void* stream);

/* out[p] = max |v| over the cells of problem p. Replaces
 * CG_SYSTEM::Convergence_Norm (DDS/CG_SYSTEM.cpp:78-96). */
hdd_status hdd_convergence_norm(hdd_ctx* ctx, int ch, float* out);

/* Zero every cell that is neither Active nor Interface. Replaces
 * CG_SYSTEM::Project (DDS/CG_SYSTEM.cpp:101-119). (Set_Boundary_Conditions and
 * Project_Nullspace are no-ops in CG_SYSTEM, :124-133, and have no entry point.) */
hdd_status hdd_project(hdd_ctx* ctx, int ch);

/* CG_VECTOR operations (DDS/CG_VECTOR.cpp:18-74 -> SPGrid_Copy, SPGrid_Copy2,
 * SPGrid_Plus_Equal, SPGrid_Minus_Equal, SPGrid_Scale of
 * DDS/SPGrid_Gemini_Arithmetic_Helpers.h:31-207), over every cell:
 *   copy: out = in * c        copy2: out = in1 * c + in2
 *   plus_equal: out += in     minus_equal: out -= in     scale: out *= c */
hdd_status hdd_vector_copy(hdd_ctx* ctx, int ch_out, float c, int ch_in);
hdd_status hdd_vector_copy2(hdd_ctx* ctx, int ch_out, float c, int ch_in1, int ch_in2);
hdd_status hdd_vector_plus_equal(hdd_ctx* ctx, int ch_out, int ch_in);
hdd_status hdd_vector_minus_equal(hdd_ctx* ctx, int ch_out, int ch_in);
hdd_status hdd_vector_scale(hdd_ctx* ctx, int ch_out, float c);

/* result = M^{-1} rhs followed by Project(result): the DD preconditioner
 * with its result, rhs and tmp fields chosen by channel, as
 * SPGrid_Domain_Decomposition_Preconditioner::Apply_Preconditioner(result, rhs, tmp)
 * (DDS/SPGrid_Domain_Decomposition_Preconditioner.h:88-128) called through
 * CG_SYSTEM::Apply_Preconditioner (DDS/CG_SYSTEM.cpp:138-144) and
 * KRYLOV_SYSTEM_BASE::Precondition. The three channels must differ; tmp is
 * overwritten; every other channel keeps its contents. hdd_apply_preconditioner
 * is this call with (HDD_CH_Z, HDD_CH_R, HDD_CH_K), synchronous. */
hdd_status hdd_apply_preconditioner_channels(hdd_ctx* ctx, int ch_result, int ch_rhs, int ch_tmp);

/* Interface solver introspection: sigma levels, subdomains, interface dofs. */
hdd_status hdd_sigma_info(const hdd_ctx* ctx, int* levels, int* n_subdomains);
hdd_status hdd_interface_dofs(const hdd_ctx* ctx, int level, int* n);
hdd_status hdd_interface_coords(const hdd_ctx* ctx, int level, int* xyz);

/* Interface operator export (which: 0 Arr, 1 Aii, 2 Air, 3 Prr) as CSR with
 * sorted columns; first call with ptr=NULL to get rows/cols/nnz. Replaces the
 * matrices of INTERFACE_MATRIX_GENERATOR::Convert_To_Eigen
 * (IFS/INTERFACE_MATRIX_GENERATOR.h:660-734). */
hdd_status hdd_interface_matrix(const hdd_ctx* ctx, int which, int level, int subdomain,
                                int* rows, int* cols, int* nnz, int* ptr, int* idx, float* val);

/* One sigma multigrid V-cycle on a level-0 interface vector: u = V(b), u0 = 0.
 * Replaces SIGMA_MULTIGRID_SOLVER::V_Cycle (IFS/SIGMA_MULTIGRID_SOLVER.cpp:36-47). */
hdd_status hdd_sigma_vcycle(hdd_ctx* ctx, const float* b_host, float* u_host);

/* One sigma smoothing step on level l: u <- Smooth(u, b). Replaces
 * SIGMA_SMOOTHER::Smooth (IFS/SIGMA_SMOOTHER.cpp:292-324). */
hdd_status hdd_sigma_smooth(hdd_ctx* ctx, int level, const float* b_host, float* u_host);

/* Preconditioned CG from x = 0 on the context's b. Replaces
 * SPGrid_Domain_Decomposition_Solver::Solve -> CONJUGATE_GRADIENT_DUMPING::Solve
 * (DDS/SPGrid_Domain_Decomposition_Solver.h:165-197, DDS/CONJUGATE_GRADIENT_DUMPING.cpp:146-212).
 * residual_history (may be NULL) receives max|r| per iteration, up to history_len. */
hdd_status hdd_pcg_solve(hdd_ctx* ctx, const hdd_pcg_params* params, hdd_pcg_report* report,
                         double* residual_history, int history_len);

/* Setup wall time in seconds. */
hdd_status hdd_setup_seconds(const hdd_ctx* ctx, double* seconds);

/* Measurement: runs `reps` level-0 smoothing sweeps (all Active cells) on the
 * context stream bracketed by HIP events; returns the mean duration of one
 * launch and the algorithmic HBM bytes of one launch (16 B per page slot:
 * flags, u and b read, u written; 10 B with FP16 V-cycle storage). No
 * reference counterpart. */
hdd_status hdd_profile_sweep(hdd_ctx* ctx, int reps, double* avg_ms, double* bytes_per_launch);

/* Measurement without a context: the same FP32 sweep on the level-0 grid of a
 * page set alone (desc's size, pages and flags, plus their face bits; no
 * divider, interface operators or factorisation), `reps` back-to-back launches
 * on a stream of its own bracketed by HIP events: mean duration of one launch
 * and 16 B per page slot. Lets a grid larger than the Infinity Cache (512^3
 * sparse: 572 MB) be timed without its set-up. No reference counterpart. */
hdd_status hdd_profile_sweep_grid(const hdd_problem_desc* desc, int reps, double* avg_ms, double* bytes_per_launch);

/* Measurement: runs `reps` plain copies of one level-0 float channel into the
 * other smoothing buffer (the calibration kernel for hardware byte counters:
 * it reads and writes 4 B per page slot); returns the mean duration and the
 * bytes read plus written by one launch. No reference counterpart. */
hdd_status hdd_profile_copy(hdd_ctx* ctx, int reps, double* avg_ms, double* bytes_per_launch);

/* Setup statistics: out[0] setup s, out[1] factorisation s, out[2] device
 * bytes of the interface solver, out[3] level-0 pages, out[4] interface dofs
 * (level 0), out[5] interior dofs (level 0, all subdomains), out[6] subdomains,
 * out[7] largest eigenvalue of the bottom sigma smoother's iteration matrix
 * (-1 if not computed), out[8] 1 if the bottom level's 200 smooths are
 * evaluated as one exact solve plus a low-rank correction, else 0, out[9]
 * eigenmodes in that correction, out[10] seconds of its eigen-solve, out[11]
 * block Krylov steps of it (0: dense), out[12] relative max difference of that
 * evaluation against the 200 smooths on a random right-hand side (-1: not run),
 * out[13] launch shapes of the triangular-solve plans that the setup's
 * autotuner changed from their defaults (HDD_AUTOTUNE=0 in the environment
 * turns it off; results do not depend on it), out[14] rows of the dense top
 * blocks of the triangular-solve plans (each tall tree's top heights solved as
 * two dense triangular matvecs; HDD_DENSE_TOP=0 in the environment turns them
 * off, HDD_DENSE_TOP=k gives a dense top to factors of at least k rows, 4096
 * by default), out[15] 1 if the interior V-cycle's 200 coarsest-level sweeps
 * run page by page in one launch (bit-identical to 200 launches; set-up picks
 * it when that level's Active cells couple only within their pages,
 * HDD_BOTTOM_LOCAL=0 turns it off), else 0, out[16] tree heights of the
 * triangular-solve plans done by one fused launch per height (a workgroup or
 * a wavefront per run of supernodes), out[17] how many of those use the
 * wavefront form. By default the fused heights and the heights with big
 * supernodes use the W form (their sub-diagonal blocks stored as
 * W = L21 Linv: one dependent phase less per fused workgroup, two launches
 * instead of three per big height and pass) with a workgroup per run;
 * HDD_FUSED_WAVE=1 / 0 in the environment forces the L21 form with a
 * wavefront / a workgroup per run (the same bits as each other; the W form
 * rounds differently), and refined solves (hdd_set_sigma_refinement) use the
 * L21 form.
 * n = number of slots provided (up to 18). No reference counterpart. */
hdd_status hdd_stats(const hdd_ctx* ctx, double* out, int n);

/* Measurement: runs `reps` exact solves of one triangular-solve plan of sigma
 * level `level` (which = 0: all subdomains' (Aii + eps I), 1: (Arr + eps I),
 * 2: the bottom level's combined system) on the context stream, bracketed by
 * HIP events; returns the mean duration of one solve and its algorithmic
 * bytes (4 B per factor entry per pass, two passes, every subdomain counted).
 * The right-hand side is zero (the cost does not depend on the values). No
 * reference counterpart. */
hdd_status hdd_profile_solve(hdd_ctx* ctx, int level, int which, int reps, double* avg_ms, double* bytes_per_solve);

/* Shape of the triangular-solve plan of sigma level `level`: which = 0 for
 * the batch of all subdomains' (Aii + eps I), 1 for (Arr + eps I). Factors on
 * the host (works on host-only contexts) and writes 8 values per tree height
 * h into out[8h .. 8h+7]: supernodes, rows, packed diagonal-inverse entries,
 * sub-diagonal (L21) entries, largest supernode, most L21 rows of one
 * supernode, most incidences (contributions of descendants) of one row, all
 * incidences of the height's rows; *heights = tree heights (only the first
 * cap/8 are written). Used to state the algorithmic bytes and the largest work
 * items of the solve kernels. No reference counterpart. */
hdd_status hdd_sigma_plan_stats(hdd_ctx* ctx, int level, int which, double* out, int cap, int* heights);

/* x = (A + 1e-5 I)^{-1} b in double precision with the host factor of sigma
 * level `level` (which = 0: Arr, 1: Aii of `subdomain`), the factorisation
 * that replaces PARDISO phases 11/22 (IFS/SIGMA_SMOOTHER.cpp:115-288); b and x
 * hold the matrix's rows (hdd_interface_matrix numbering). Works on host-only
 * contexts. For checking the factors; no reference counterpart. */
hdd_status hdd_sigma_factor_solve(hdd_ctx* ctx, int level, int which, int subdomain, const double* b, double* x);

/* Eigenvalues >= lambda_cut (0 < lambda_cut < 1) of the bottom sigma
 * smoother's iteration matrix (Arr+eI)^-1 sum_s Air_s^T (Aii_s+eI)^-1 Air_s,
 * descending, computed on the host in double precision (works on host-only
 * contexts). *k = their count (only the first cap are written), *lambda_max =
 * the largest eigenvalue. These are the modes that survive the reference's 200
 * bottom smooths of SIGMA_MULTIGRID_SOLVER::V_Cycle
 * (IFS/SIGMA_MULTIGRID_SOLVER.cpp:36-47). No reference counterpart. */
hdd_status hdd_bottom_spectrum(hdd_ctx* ctx, double lambda_cut, double* lambdas, int cap, int* k, double* lambda_max);

/* ---- Serving (no reference counterpart: the reference solves one problem per
 * process). Requests of different sizes are submitted as they arrive and solved
 * with continuous batching: worker threads set each one up on the host (as
 * hdd_create); a builder admits the ready requests, within an HBM budget, as a
 * unit (one batched context, as hdd_create_batch, on its own stream); a driver
 * thread advances every resident unit one PCG iteration at a time, so units
 * admitted while others are mid-solve join at the next iteration and their
 * kernels overlap on the device. A request's device state stays resident until
 * it stops; its result is returned as soon as it does, and a unit's memory and
 * slot are freed when all its requests have stopped. Requests in one unit need
 * the same vlevels, sigma_levels and smoothing counts (others wait for a later
 * unit). */
typedef struct hdd_server hdd_server;

typedef struct hdd_server_params {
    int setup_threads;        /* host setup workers */
    int groups;               /* units resident and iterating on the device at once */
    int max_batch;            /* problems per unit */
    double hbm_budget_bytes;  /* device bytes for resident problems; 0 = 90% of the free device memory */
} hdd_server_params;

typedef struct hdd_request_report {
    int iterations;           /* negative if it stopped before converging */
    double final_norm;        /* max|r| at exit */
    double latency_s;         /* submit -> result */
    double setup_s;           /* submit -> host setup done (includes waiting for a worker) */
    double queue_s;           /* host setup done -> its unit's first iteration (includes the device build) */
    double solve_s;           /* its unit's first iteration -> result */
    int batch_size;           /* problems in its unit */
    int concurrent;           /* units iterating on the device when its unit started, its own included */
    int cancelled;            /* 1 if it stopped because of hdd_server_cancel */
} hdd_request_report;

hdd_status hdd_server_create(const hdd_server_params* params, hdd_server** out);

/* Queues one problem (copied; the caller's arrays may be freed on return). */
hdd_status hdd_server_submit(hdd_server* srv, const hdd_problem_desc* desc, const hdd_pcg_params* params, int64_t* request_id);

/* Blocks until the request is done; x (may be NULL) receives n_pages*256
 * floats in the page order of its desc. Returns HDD_ERR_SETUP / HDD_ERR_DEVICE
 * with hdd_server_last_error() set if the request failed, HDD_ERR_NOT_CONVERGED
 * if it stopped at its iteration limit, HDD_ERR_CANCELLED if it was cancelled
 * (x then holds the iterate it stopped at, or zeros if it never ran). The
 * request's record is released. */
hdd_status hdd_server_wait(hdd_server* srv, int64_t request_id, float* x, hdd_request_report* report);

/* Asks a request to stop: one still queued is dropped before it reaches the
 * device; one that is running stops at its unit's next residual test (the
 * other requests of its unit continue). HDD_ERR_ARG if the request is unknown
 * or already done. */
hdd_status hdd_server_cancel(hdd_server* srv, int64_t request_id);

/* Finishes every submitted request, then stops the workers and frees the server. */
hdd_status hdd_server_destroy(hdd_server* srv);

/* Message of the last failure on this server (or of hdd_server_create when srv is NULL). */
const char* hdd_server_last_error(const hdd_server* srv);

/* HBM budget of the server in bytes. */
hdd_status hdd_server_budget(const hdd_server* srv, double* bytes);

#ifdef __cplusplus
}
#endif
#endif
