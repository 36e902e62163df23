/*
 * C ABI of the MI355X-native SPOCK solver (libspock_amd.so).
 *
 * The reference exposes its solver as C++ classes over CUDA (src/tree.cuh:102,
 * src/problem.cuh:126, src/cache.cuh:1138-1245); there is no C FFI in the reference.
 * The entry points below replace those classes one for one and are what the
 * C++ drop-in header (spock_amd.hpp) and the Python bindings call:
 *
 *   spk_problem_load      <- ScenarioTree<T>(path) + ProblemData<T>(tree)
 *                            (src/tree.cuh:102-179, src/problem.cuh:126-153)
 *   spk_batch_create      <- CacheBuilder<T>::build() / make_unique()  (src/cache.cuh:1215-1244)
 *   spk_batch_set_x0      <- Cache::initialiseState (src/cache.cuh:592-605)
 *   spk_batch_set_iterate <- runSpock/runCp's previousSolution argument (src/cache.cuh:611-632)
 *   spk_batch_solve       <- Cache::runSpock / Cache::runCp (src/cache.cuh:967-1132)
 *   spk_batch_results     <- status(), solveIter(), solveTime() (src/cache.cuh:357-361)
 *   spk_batch_get_iterate <- solution() (src/cache.cuh:378)
 *   spk_batch_reset       <- Cache::reset (src/cache.cuh:436-479)
 *   spk_batch_set_debug   <- CacheBuilder::enableDebug (src/cache.cuh:1204-1209) and the debug
 *   spk_batch_debug_records  caches printToJson writes (src/cache.cuh:947-959, :1274-1311)
 *   spk_batch_debug_op    <- the operator pieces the reference's gtest suite reaches through
 *                            friend functions (LinearOperator::op/adj, Dynamics::project,
 *                            projectPrimalWorkspaceOnKernels, iter(), dotM;
 *                            tests/cuda/testOperator.cu, testProblem.cu, testCache.cu)
 *
 * Conventions: every function returns 0 on success and a negative code on error
 * (enum spk_error); spk_last_error(ctx) returns the message.  No exception crosses the ABI.  A context
 * owns one HIP stream (default priority); solves are enqueued on it and
 * spk_batch_results() waits for them.  One context per host thread.
 * Lifetimes: a context lives while the caller's handle or any problem, batch or network made on
 * it does, and a problem while the caller's handle or any batch holding an instance of it does;
 * spk_*_destroy drops the caller's handle, so handles may be destroyed in any order.  (The
 * reference's Cache and ProblemData hold references that must not dangle, src/cache.cuh:97-98.)
 * A batch holds `count` instances; each instance refers to a loaded problem (several
 * instances may share one problem, e.g. many initial states of one tree).
 */
#ifndef SPOCK_AMD_H
#define SPOCK_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spk_ctx spk_ctx;
typedef struct spk_problem spk_problem;
typedef struct spk_batch spk_batch;
typedef struct spk_mlp spk_mlp;

enum spk_dtype { SPK_F64 = 0, SPK_F32 = 1 };
/* error codes: SPK_ERR for every failure but one; SPK_ERR_COOPERATIVE_LAUNCH when the device
 * refused the cooperative launch of a grid-mode solve (spk_batch_set_workgroups > 1), which
 * queued nothing and leaves the batch as it was, so the solve can be repeated with one workgroup */
enum spk_error { SPK_OK = 0, SPK_ERR = -1, SPK_ERR_COOPERATIVE_LAUNCH = -2 };
enum spk_method { SPK_SPOCK = 0, SPK_CP = 1 };
/* solve status codes, identical to enum CacheStatus (src/cache.cuh:68-73) */
enum spk_status { SPK_NOT_RUN = -1, SPK_CONVERGED = 0, SPK_OUT_OF_ITER = 1, SPK_OUT_OF_TIME = 2,
                  /* debug traces on: a candidate entry was not finite and the solve stopped there,
                     where the reference's isFinite throws (src/cache.cuh:52-62, :912) */
                  SPK_NONFINITE = 3 };
/* Anderson least squares: Householder QR as the reference (geqrf/ormqr/trsm), or the Gram
 * matrix kept incrementally and solved by Cholesky in FP64 (one fused pass per iteration). */
enum spk_ls_method { SPK_LS_HOUSEHOLDER = 0, SPK_LS_GRAM = 1 };
enum spk_debug_op { SPK_OP_L = 0, SPK_OP_LADJ = 1, SPK_OP_DYN = 2, SPK_OP_KERNEL = 3, SPK_OP_T = 4, SPK_OP_ELL = 5,
                    SPK_OP_DUAL = 6 };

/* Builder options (src/cache.cuh:1154-1210).  tol_abs/tol_rel are multiplied by
 * min(scaling) for preconditioned problems (src/cache.cuh:302-306).  max_iters = 0 means
 * no iteration cap from the caller (a device-side cap of 2e6 outer iterations still applies).
 * check_period is the error-check period (25 in the reference, src/cache.cuh:112). */
typedef struct spk_options {
    double tol_abs;
    double tol_rel;
    double max_time_secs;
    long max_iters;
    int max_inner;
    int anderson_buffer;
    int check_period;
    int method;
    int block_size;  /* threads per workgroup; 0 (default): for per-instance batches the size chosen by
                        spk_batch_autotune_uneven, else 512 (256 in grid mode); for shared batches the
                        size chosen by spk_batch_autotune, else 1024 */
    int ls_method;   /* enum spk_ls_method */
    int dyn_inverse; /* 1: apply (L L')^-1 as one matvec; 0: two triangular solves as the reference */
    long chunk_iters; /* shared batches: > 0 ends a launch after this many outer iterations with the
                         whole solver state kept on the device; spk_batch_resume continues it exactly */
} spk_options;

typedef struct spk_result {
    int status;
    long iters;
    double err;
    double tol;
    double w;
    long k0, k1, k2, k2bt, k3;
    long num_T;   /* applications of the CP operator T */
    long num_ell; /* applications of [L*; L] inside norms and error checks */
    /* device wall-clock seconds per phase: L* + axpy, dynamics/kernel projections, L + dual prox,
     * residual, [L*; L] reductions, Anderson column, Anderson least squares, Anderson direction,
     * line-search updates, other */
    double phase_seconds[10];
} spk_result;

void spk_default_options(spk_options *opts);

int spk_ctx_create(int device, spk_ctx **out);
int spk_ctx_destroy(spk_ctx *ctx);
const char *spk_last_error(spk_ctx *ctx);
/* the context's HIP stream, as a hipStream_t stored in *stream */
int spk_ctx_stream(spk_ctx *ctx, void **stream);
int spk_ctx_synchronize(spk_ctx *ctx);

int spk_problem_load(spk_ctx *ctx, const char *folder, spk_problem **out);
int spk_problem_destroy(spk_problem *p);
/* out[0..15] = N, nl, S, E, nx, nu, size_prim, size_dual, size_iterate, preconditioned,
 * con3, con5, #distinct dynamics factors (ABKTr), #distinct cost matrices (sqrtQ), 0, 0 */
int spk_problem_sizes(const spk_problem *p, long *out);
/* preconditioning scaling (nx + nu values; all ones when not preconditioned) */
int spk_problem_scaling(const spk_problem *p, double *out);
double spk_problem_step_size(const spk_problem *p);

int spk_batch_create(spk_ctx *ctx, int count, spk_problem *const *problems, int dtype, spk_batch **out);
/* `count` instances of ONE problem (e.g. many initial states of one tree) in a block-interleaved
 * layout solved in lockstep: instances form blocks of 16 (the default tile engine) or 64 (the lane
 * engine, see spk_batch_set_schedule) and element e of instance k is stored at
 * ((k / w) * n + e) * w + k % w (n = size_iterate, w the block width); one workgroup serves one
 * block and the factor tables are shared by its instances.  Every spk_batch_* call below accepts
 * such a batch; the Anderson least squares is the FP64 Gram solve (ls_method is ignored), the
 * dynamics sweep applies the precomputed (L L')^-1 (dyn_inverse is ignored), anderson_buffer <= 4,
 * numInputs <= 32, numEvents <= 7.  block_size is the workgroup size (when <= 0 the size chosen by
 * spk_batch_autotune, else the engine's default).  No counterpart in the reference, which solves
 * one instance per Cache. */
int spk_batch_create_shared(spk_ctx *ctx, spk_problem *problem, int count, int dtype, spk_batch **out);
int spk_batch_is_shared(const spk_batch *b);
int spk_batch_destroy(spk_batch *b);
int spk_batch_count(const spk_batch *b);
/* workgroups that solve one instance together (per-instance batches of ONE instance; a batch of
 * several instances always runs one workgroup per instance).  0 or 1 (default): one workgroup;
 * n > 1: n workgroups in a cooperative launch with grid barriers between the stages of the tree
 * sweeps and grid-wide reductions, so one large tree is spread over the GPU.  The count must fit
 * co-resident on the device (the solve fails with a message otherwise).  The reference runs one
 * instance on the whole GPU through its many small launches (src/cache.cuh:1008-1132). */
int spk_batch_set_workgroups(spk_batch *b, int workgroups);
/* zero every iterate and workspace (Cache::reset, src/cache.cuh:436-479); the initial states set by
 * spk_batch_set_x0 are inputs and stay, where the reference re-uploads them on every runSpock */
int spk_batch_reset(spk_batch *b);
/* debug traces (per-instance batches): max_records > 0 makes every later solve record, at each
 * error check (every check_period outer iterations, from 0), one record of SPK_DEBUG_WORDS doubles
 *   {outer iteration, calls to L and L* of this solve (two per application of T), err0, err1, err2, err3}
 * as the reference's debug caches (src/cache.cuh:947-959): err1 and err2 the max-norms of the
 * primal and dual parts of r/gamma - [L*; L] r, err3 = ||z - T(z)||_2, and err0 as the reference
 * computes it, e <- cand_prim - e then ||e||_2 with e zeroed only by spk_batch_reset (the L* product
 * next to it in the reference never reaches e).  Every outer iteration also checks the candidate
 * for non-finite entries and stops the solve with status SPK_NONFINITE at the first one.  Records
 * beyond max_records are counted but not kept.  0 turns the traces off. */
#define SPK_DEBUG_WORDS 6
int spk_batch_set_debug(spk_batch *b, long max_records);
/* the records of instance inst's last solve: up to cap of them into records (cap x SPK_DEBUG_WORDS),
 * *count = the records the solve made (may exceed what was kept), *nonfinite = the index of the
 * first non-finite candidate entry when the solve ended with SPK_NONFINITE, else -1 */
int spk_batch_debug_records(spk_batch *b, int inst, long cap, double *records, long *count, long *nonfinite);
/* initial states, unscaled, concatenated over instances (nx_k values each) */
int spk_batch_set_x0(spk_batch *b, const double *x0);
/* per-instance acceptance tolerances in the solver's scaled units (spk_result.tol of an earlier
 * solve): when set (> 0), a solve tests convergence from its first error check on against this
 * value instead of calibrating max(tol_abs, tol_rel * err0) there (src/cache.cuh:932-936).  Used for
 * warm starts and FP64 polish so they accept against the cold solve's threshold.  tol == NULL
 * clears; 0 entries calibrate as the reference. */
int spk_batch_set_tolerances(spk_batch *b, const double *tol);
/* host iterate (size_iterate_k values) of instance k, used as the warm start of the next solve */
int spk_batch_set_iterate(spk_batch *b, int k, const double *z);
int spk_batch_get_iterate(spk_batch *b, int k, double *z);
/* all iterates at once: host arrays concatenated over instances */
int spk_batch_set_iterates(spk_batch *b, const double *z);
int spk_batch_get_iterates(spk_batch *b, double *z);
/* the accepted iterates as device rows: FP32, instance-major, leading dimension ld >= size_iterate
 * (all instances of one problem size).  set also zeroes the residual history, as a warm start. */
int spk_batch_set_iterates_device(spk_batch *b, const float *src, long ld);
int spk_batch_get_iterates_device(spk_batch *b, float *dst, long ld);
/* device pointer to instance k's accepted iterate (dtype of the batch); contiguous for
 * per-instance batches, stride w elements (the block width, 16 or 64) for shared batches */
void *spk_batch_iterate_ptr(spk_batch *b, int k);
/* copy iterates between batches of the same problems (converting dtype) on the device */
int spk_batch_copy_iterates(spk_batch *dst, const spk_batch *src);
/* enqueue the solve of every instance; warm != 0 keeps the current iterates as previous solution */
int spk_batch_solve(spk_batch *b, const spk_options *opts, int warm);
/* Allocation-free variant of spk_batch_solve (no counterpart in the reference).  For callers that
 * cannot allocate device memory: code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc.  It calls no hipMalloc, hipFree or other allocating routine.  Per-instance
 * batches (spk_batch_create) only.  The solve's instance descriptors, Anderson rings and grid-mode
 * partials live in one fixed scratch array per device of SPK_NOALLOC_SCRATCH_BYTES, held until the
 * process ends; a batch needing more is rejected with SPK_ERR and a message naming both sizes
 * (the scratch holds sizeof descriptor x count + (3m+1) x size_iterate words per instance for
 * anderson_buffer m, + 256 KiB of partials for a single instance).  The iterates, results and
 * timing are the batch's as after spk_batch_solve, and equal to them.  Every call on a device shares
 * that device's scratch: calls that overlap on one GPU (batches of different contexts, or threads)
 * race and must be serialised by the caller; calls on one context are ordered by its stream. */
#define SPK_NOALLOC_SCRATCH_BYTES (64u << 20)
int spk_batch_solve_noalloc(spk_batch *b, const spk_options *opts, int warm);
/* continue a chunked solve of a shared batch for another chunk_iters outer iterations; the
 * continued iterates, branch decisions and counters are those of one uninterrupted solve.  Results
 * after a chunk report SPK_NOT_RUN (and the iterations so far) for instances still running. */
int spk_batch_resume(spk_batch *b);
/* Continuous batching of a shared batch (no counterpart in the reference, which solves one
 * instance per Cache): between the launches of a chunked solve, start new solves in `count`
 * distinct instance slots `lanes` (normally slots whose solve has ended; a running solve in a
 * listed slot is abandoned).  Each gets the unscaled initial state x0[i] (nx values), a fresh
 * SuperMann state and its own iteration count, and starts at the next spk_batch_resume while the
 * other slots continue; its checks, caps and results count from its own first iteration.  The slot's
 * previous result (spk_batch_get_iterate, spk_batch_results) must be read before the refill.
 *   spk_batch_refill:            cold starts (the zero iterate; previous iterate and residual ones)
 *   spk_batch_refill_from_batch: warm starts from the accepted iterates of slots src_lanes of the
 *                                shared batch src (same problem size, either dtype), e.g. the FP64
 *                                polish of FP32 solutions
 *   spk_batch_refill_rows:       warm starts from FP32 device rows z (count rows, leading dimension ld)
 * (a warm start's previous iterate is the starting iterate and its previous residual zero, as
 * initialisePrev with a previous solution, src/cache.cuh:611-632). */
int spk_batch_refill(spk_batch *b, int count, const int *lanes, const double *x0);
/* an empty pool: a chunked solve (opts->chunk_iters > 0) in which every slot is idle (its result
 * stays SPK_NOT_RUN) until a refill starts a solve in it; spk_batch_resume then runs the slots that
 * have one.  opts (tolerances, caps, ...) apply to every solve of the pool. */
int spk_batch_pool_open(spk_batch *b, const spk_options *opts);
/* outer iterations per launch of the following resumes of a chunked solve */
int spk_batch_set_chunk(spk_batch *b, long chunk_iters);
/* A time slice for the launches of the chunked solves of a shared batch (no counterpart in the
 * reference; for continuous batching): each workgroup also leaves a launch, its state saved as at
 * the end of a chunk, once it has run `seconds` of wall-clock time and at least one outer iteration
 * in it, so fast workgroups keep iterating while slow ones finish instead of idling until the
 * slowest has done chunk_iters.  Where a launch ends never changes the iterates, branch decisions
 * or counters (those of one uninterrupted solve); only the iterations done per launch vary.
 * 0 (the default) turns it off.  Applies from the next launch. */
int spk_batch_set_chunk_time(spk_batch *b, double seconds);
int spk_batch_refill_from_batch(spk_batch *b, int count, const int *lanes, const double *x0, spk_batch *src,
                                const int *src_lanes);
int spk_batch_refill_rows(spk_batch *b, int count, const int *lanes, const double *x0, const float *z, long ld);
/* the accepted iterates of `count` slots `lanes` of a shared batch as FP32 device rows (leading
 * dimension ld >= size_iterate) */
int spk_batch_get_lanes_device(spk_batch *b, int count, const int *lanes, float *dst, long ld);
/* shared batches: the engine and schedule.  mv = 17 (default) or 16: the tile engine, 16 instances
 * per workgroup in blocks of 16 (element e of instance k at ((k / 16) * n + e) * 16 + k % 16) with the
 * per-node products on the matrix cores (v_mfma 16x16x4, the 16 instances as the B operand), built
 * for workgroups of up to 1024 threads (17; block 0: 1024) or 512 (16; block 0: 256).
 * mv = 0 or 8: the lane engine, 64 instances per workgroup in blocks of 64 with
 * the factor entries read by scalar loads (0) or staged through LDS in 8-column chunks (8), which
 * give bitwise identical iterates, block 64..1024 (0: 1024).  Switching between the engines re-lays
 * out the initial states and accepted iterates and clears the workspace.  No counterpart in the
 * reference. */
int spk_batch_set_schedule(spk_batch *b, int mv, int block);
/* shared batches: time `iters` cold outer iterations (from the batch's initial states) under each of
 * the SPK_AUTOTUNE_CANDIDATES schedules (mv, block) = (0, 1024), (8, 1024), (0, 512), (8, 512),
 * (16, 256), (16, 512), (17, 1024), keep the fastest for the batch's later solves and return it (mv, block:
 * nullable; seconds: nullable, the kernel times in that order).  The iterates are reset (initial
 * states and given tolerances are kept).  opts supplies the solver options of the timed runs
 * (nullable: defaults). */
enum { SPK_AUTOTUNE_CANDIDATES = 7 };
int spk_batch_autotune(spk_batch *b, const spk_options *opts, int iters, int *mv, int *block, double *seconds);
/* shared batches: a chunked solve (opts->chunk_iters > 0) captured as one HIP graph of `launches`
 * kernel launches (the first and launches - 1 resumes) and replayed with no host round trip between
 * chunks; workgroups whose instances have all finished return at once from the remaining launches.
 * The iterates are those of one uninterrupted solve.  If launches x chunk_iters is too few, results
 * report SPK_NOT_RUN for unfinished instances and spk_batch_resume continues.
 * Per-instance (uneven) batches: the solve is one persistent launch; launches must be 1 and
 * chunk_iters 0, and that launch is captured as a graph with one kernel node and launched (not
 * in grid mode: a cooperative launch is not captured).  spk_batch_replay_graph launches it again.
 * No counterpart in the reference. */
int spk_batch_solve_graph(spk_batch *b, const spk_options *opts, int warm, int launches);
/* per-instance batches: launch the solve captured by spk_batch_solve_graph again, with its options
 * and warm flag, on whatever the batch's buffers now hold: initial states written by
 * spk_batch_set_x0, iterates by spk_batch_set_iterate(s) or spk_batch_reset, tolerances by
 * spk_batch_set_tolerances (the descriptors are rewritten in place first).  Repeated solves of the
 * same problems (a receding-horizon loop) then cost one graph launch.  A later spk_batch_solve,
 * debug operator, spk_batch_set_debug or larger Anderson memory reallocates buffers the graph points
 * to and drops it; replaying then fails with a message.  No counterpart in the reference. */
int spk_batch_replay_graph(spk_batch *b);
/* continue a chunked solve of a shared batch with `launches` resume launches captured as one graph
 * (as spk_batch_resume, without host round trips between the launches). */
int spk_batch_resume_graph(spk_batch *b, int launches);
/* per-instance (uneven) batches: time `iters` cold outer iterations under each of
 * SPK_AUTOTUNE_UNEVEN_CANDIDATES schedules of the persistent launch -- workgroups of 128, 256 and
 * 512 threads, with the instances in batch order and then largest first (descending iterate size,
 * so the longest solves start in the first dispatch wave) -- keep the fastest for later solves
 * that leave spk_options.block_size at 0, and return it (block, largest_first, seconds: nullable;
 * seconds gets the kernel times in that order).  The order never changes an instance's result;
 * the workgroup size fixes the order of its sums.  The iterates are reset (initial states and given
 * tolerances are kept).  No counterpart in the reference, which solves one instance per Cache
 * (src/cache.cuh:1008-1132). */
#define SPK_AUTOTUNE_UNEVEN_CANDIDATES 6
int spk_batch_autotune_uneven(spk_batch *b, const spk_options *opts, int iters, int *block, int *largest_first,
                              double *seconds);
/* set that schedule directly (block 0: 256 threads). */
int spk_batch_set_uneven_schedule(spk_batch *b, int block, int largest_first);
/* per-instance batches of ONE instance: time `iters` cold outer iterations with 1, 8, 32, 64 and
 * 128 cooperating workgroups (grid mode above 1), keep the fastest for later solves and return it
 * (workgroups, seconds: nullable; seconds gets the 5 kernel times in that order).  The iterates are
 * reset (initial states and given tolerances are kept).  No counterpart in the reference. */
int spk_batch_autotune_workgroups(spk_batch *b, const spk_options *opts, int iters, int *workgroups,
                                  double *seconds);
/* wait for the last solve; fills count results and the kernel time in seconds (nullable) */
int spk_batch_results(spk_batch *b, spk_result *out, double *kernel_seconds);
/* single operator on every instance: in/out are host arrays concatenated over instances
 * (size_iterate_k values each).  L reads the primal part of `in` and writes the dual part of
 * `out`; LADJ the reverse; DYN/KERNEL project the primal part; T writes T(in); ELL writes
 * scalars[2k] = <in, M out_in> and scalars[2k+1] = max|out_in/gamma - ell(out_in)| where out_in
 * is taken from `out` on entry.  DUAL (per-instance batches): the primal part of `out` on entry is a
 * primal point p; writes the dual part of `out` as the CP dual step from the iterate `in` with
 * p in place of the primal prox, i.e. Cache::getDual(p) (src/cache.cuh:371-376, 770-789). */
int spk_batch_debug_op(spk_batch *b, int op, const double *in, double *out, double *scalars);

/* ------------------------------------------------------------------ learned warm start
 * An MLP x0 -> primal-dual iterate (no counterpart in the reference, whose solves start cold or
 * from a caller-given previous solution, src/cache.cuh:611-632).  in -> `layers` hidden layers of
 * `hidden` units with GELU -> out; bf16 weights feed the matrix cores with FP32 accumulation and FP32
 * master weights take the Adam steps.  Outputs are standardized per component (mean and standard
 * deviation of the training targets); spk_mlp_predict writes mean + std * output into a batch as its
 * warm-start iterates.  Parameters are concatenated per layer: W (d_in x d_out, row-major), then b. */
int spk_mlp_create(spk_ctx *ctx, int in, int hidden, int layers, int out, unsigned long long seed, spk_mlp **out_mlp);
int spk_mlp_destroy(spk_mlp *m);
long spk_mlp_num_params(const spk_mlp *m);
int spk_mlp_get_params(spk_mlp *m, float *params);
int spk_mlp_set_params(spk_mlp *m, const float *params);
int spk_mlp_get_grads(spk_mlp *m, float *grads);
int spk_mlp_get_normalization(spk_mlp *m, float *in_mean, float *in_std, float *out_mean, float *out_std);
int spk_mlp_set_normalization(spk_mlp *m, const float *in_mean, const float *in_std, const float *out_mean,
                              const float *out_std);
/* forward of host rows (raw inputs, normalized inside); y (nullable) gets the standardized outputs */
int spk_mlp_forward(spk_mlp *m, const float *x, int rows, float *y);
/* after spk_mlp_forward of the same rows: mean squared error against standardized targets (host),
 * gradients of every parameter into the gradient buffer */
int spk_mlp_loss_backward(spk_mlp *m, const float *target, int rows, double *loss);
int spk_mlp_adam(spk_mlp *m, float lr, float beta1, float beta2, float eps);
/* device-resident training set from a solved batch (targets = its accepted iterates, standardized)
 * and the x0 of its instances (host, count x in); sets the normalization */
int spk_mlp_dataset(spk_mlp *m, spk_batch *solved, const double *x0);
/* the same from `count` slots `lanes` of a solved shared batch only (e.g. its converged solves), x0
 * holding their initial states (host, count x in) */
int spk_mlp_dataset_lanes(spk_mlp *m, spk_batch *solved, int count, const int *lanes, const double *x0);
/* epochs of shuffled minibatch Adam on the training set; epoch_loss (nullable) gets the mean loss */
int spk_mlp_train(spk_mlp *m, int epochs, int minibatch, float lr, unsigned long long seed, double *epoch_loss);
/* warm-start iterates of `rows` initial states (host, rows x in) written into dst (rows = its count) */
int spk_mlp_predict(spk_mlp *m, const double *x0, int rows, spk_batch *dst);
/* continuous batching: new solves in slots `lanes` of the shared batch dst (see spk_batch_refill),
 * warm-started from the network's prediction for their initial states x0 (host, rows x in) */
int spk_mlp_predict_lanes(spk_mlp *m, const double *x0, int rows, spk_batch *dst, const int *lanes);

#ifdef __cplusplus
}
#endif
#endif
