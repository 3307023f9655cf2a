/* l2g.h — C ABI of the MI355X (gfx950) two-electron Coulomb integral library.
 *
 * This is the drop-in boundary for libint's
 *   libint2::Engine(Operator::coulomb, max_nprim, max_l) / Engine::compute(s1,s2,s3,s4)
 * path (include/libint2/engine.h:502-526, :747-749; implementation
 * include/libint2/engine.impl.h:1151-2113).  Every entry point takes plain
 * pointers and sizes, returns 0 on success and a negative L2G_E* code on
 * failure (message in l2g_last_error()), and never lets a C++ exception
 * cross the boundary.  Results are bitwise laid out exactly as libint lays
 * out Engine::results()[0]: row-major (s1 s2|s3 s4) in the caller's shell
 * order, Cartesian functions in libint's standard order
 * (include/libint2/cgshell_ordering.h:125-130), real solid harmonics in the
 * Standard (m=-l..l) or Gaussian order (include/libint2/shgshell_ordering.h).
 *
 * Host-only utilities (no GPU needed): l2g_shell_renorm,
 * l2g_solid_harmonics_coeff, l2g_boys_table, l2g_version.
 */
#ifndef L2G_H
#define L2G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define L2G_OK 0
#define L2G_E_INVALID -1      /* bad argument (std::invalid_argument in libint) */
#define L2G_E_LMAX -2         /* angular momentum above the compiled limit (Engine::lmax_exceeded) */
#define L2G_E_NODEVICE -3     /* no HIP device / HIP runtime error */
#define L2G_E_OOM -4          /* device or host allocation failed */
#define L2G_E_LAUNCH -5       /* kernel launch or execution error */
#define L2G_E_STATE -6        /* object used in the wrong state (Engine::using_default_initialized) */

/* screening methods, values as libint2::ScreeningMethod (include/libint2/shell.h:1041-1059) */
#define L2G_SCREEN_ORIGINAL 0x0001
#define L2G_SCREEN_CONSERVATIVE 0x0010
#define L2G_SCREEN_SCHWARZ 0x0100
#define L2G_SCREEN_SCHWARZINF 0x1000

/* solid-harmonic orderings (include/libint2/shgshell_ordering.h) */
#define L2G_SHG_STANDARD 1
#define L2G_SHG_GAUSSIAN 2

typedef struct l2g_context_s* l2g_context;
typedef struct l2g_basis_s* l2g_basis;
typedef struct l2g_eri_plan_s* l2g_eri_plan;
typedef struct l2g_fock_s* l2g_fock;

/* ---- library / device ------------------------------------------------- */

/* Version string. */
const char* l2g_version(void);

/* Message of the last failed call on this thread. */
const char* l2g_last_error(void);

/* Largest angular momentum per shell the compiled kernels cover
 * (libint: LIBINT2_MAX_AM_eri, engine.impl.h:620-622). */
int l2g_max_l(void);

/* Largest l for which FP32 kernel variants exist: in mixed precision a
 * quartet with a shell above it runs in FP64 whatever its bounds.  No
 * counterpart in the reference (FP64 only). */
int l2g_fp32_max_l(void);

/* Which quartets mixed batches route to FP32 variants (plans and Fock
 * builders created afterwards on this context).  L2G_FP32_MEASURED (the
 * default), so that mixed precision is not slower than FP64: in ERI plans,
 * the classes whose FP32 kernel the autotuner measured faster than their
 * FP64 kernel on FP32-eligible quartets (libint_amd/tuned.cfg), when the
 * plan holds at least 4 x 256 lanes per CU of such quartets of the class
 * (a smaller FP32 launch cannot fill the chip); in Fock builds, none (their
 * FP32-eligible quartets are light, and mixed builds measured slower).
 * L2G_FP32_ALWAYS: every eligible quartet of every class with an FP32
 * variant (the FP32 path's tests and error tables).  No counterpart in the
 * reference. */
#define L2G_FP32_MEASURED 0
#define L2G_FP32_ALWAYS 1
int l2g_set_fp32_policy(l2g_context ctx, int policy);

/* The classes (la lb lc ld as the decimal code la*1000+lb*100+lc*10+ld)
 * whose FP32 ERI variant was measured faster than FP64, up to cap of them;
 * *n receives how many there are. */
int l2g_fp32_paying_classes(int cap, int* n, int* codes);

/* Opens a device context: selects the HIP device, creates a stream, builds
 * and uploads the Boys-function table.  Replaces libint2::initialize()
 * (include/libint2/initialize.h:76-90) for this path. */
int l2g_init(int device, l2g_context* ctx);

/* Routes every class through the generic kernel (l2g_generic.hip), which
 * otherwise serves only the classes with an f shell; on = 0 restores the
 * generated register-resident kernels for s, p and d classes.  A testing
 * aid with no counterpart in the reference: it lets the generic kernel be
 * checked on every class.  Affects plans, Schwarz factors and Fock builds
 * created afterwards. */
int l2g_set_generic_kernels(l2g_context ctx, int on);

/* Releases the context (libint2::finalize, initialize.h:105-113).  Returns
 * L2G_E_STATE, and keeps the context, while any basis, plan or Fock builder
 * created on it has not been destroyed. */
int l2g_finalize(l2g_context ctx);

/* Solid-harmonic ordering of pure shells for results produced after the call
 * (libint2::set_solid_harmonics_ordering, initialize.h:117-119). */
int l2g_set_solid_harmonics_ordering(l2g_context ctx, int ordering);

/* ---- host utilities ---------------------------------------------------- */

/* Embeds primitive normalization into contraction coefficients and, if
 * enforce_unit, rescales the contraction to unit self-overlap; also returns
 * ln|c| per primitive.  Replaces Shell::renorm + update_max_ln_coeff
 * (include/libint2/shell.h:958-1011).  coeff is updated in place. */
int l2g_shell_renorm(int l, int nprim, const double* alpha, double* coeff, double* max_ln_coeff,
                     int enforce_unit);

/* Coefficient of Cartesian (lx,ly,lz) in real solid harmonic (l,m)
 * (SolidHarmonicsCoefficients::coeff, include/libint2/solidharmonics.h:114-180). */
double l2g_solid_harmonics_coeff(int l, int m, int lx, int ly, int lz);

/* Copies the host Boys table, layout [m][interval][8] (monomial coefficients
 * of xd = 7T - iv - 1/2), m in [0, *mmax], interval in [0, *nint).
 * Pass out = NULL to query the sizes. */
int l2g_boys_table(double* out, int* nint, int* mmax);

/* ---- basis ------------------------------------------------------------- */

/* Uploads a basis.  Arrays are per shell (l, pure, nprim, origin[3*nshell],
 * bohr) and per primitive in shell order (alpha, coeff).  Coefficients must
 * already be renormalised (l2g_shell_renorm), exactly as libint2::Shell
 * stores them.  Only single-contraction shells (libint asserts the same,
 * engine.impl.h:1168-1171). */
int l2g_basis_create(l2g_context ctx, int nshell, const int* l, const int* pure, const int* nprim,
                     const double* alpha, const double* coeff, const double* origin, l2g_basis* basis);
int l2g_basis_destroy(l2g_basis basis);

/* ---- ERI batches ------------------------------------------------------- */

/* Plans a batch of shell quartets given as caller-order index 4-tuples
 * (quartets[4*q+0..3]).  Each quartet is canonicalised as
 * Engine::compute2 does (engine.impl.h:1181-1208), bucketed by class and
 * contraction depth, its shell-pair data is built with the requested
 * precision and primitive screening (Original or Conservative, as Engine
 * allows without precomputed pairs, engine.impl.h:1212-1216), and the
 * work lists are uploaded.  scale = Engine::prescale_by (engine.h:857),
 * uniform_cart = CartesianShellNormalization::uniform. */
int l2g_eri_plan_create(l2g_basis basis, int64_t nquartets, const int32_t* quartets, double precision,
                        int screening, double scale, int uniform_cart, l2g_eri_plan* plan);

/* Mixed-precision variant of l2g_eri_plan_create: quartets whose Schwarz
 * bound K12 K34 (K = sqrt(max|(ab|ab)|), computed on the GPU for the pairs of
 * the batch) is below fp32_threshold (1e-6 in the mixed-precision workload)
 * run the recursion, HRR and transforms in FP32; the Boys function and the
 * prefactors stay FP64, outputs are FP64.  fp32_threshold = 0 is the plain
 * FP64 plan.  No libint counterpart: libint evaluates everything in FP64. */
int l2g_eri_plan_create_mixed(l2g_basis basis, int64_t nquartets, const int32_t* quartets, double precision,
                              int screening, double scale, int uniform_cart, double fp32_threshold,
                              l2g_eri_plan* plan);

/* Number of doubles the batch produces. */
int l2g_eri_plan_size(l2g_eri_plan plan, int64_t* total);

/* Per-quartet offsets (in doubles) into the batch output, caller order. */
int l2g_eri_plan_offsets(l2g_eri_plan plan, int64_t* offsets);

/* Runs the batch on the context stream.  d_out = device buffer of
 * l2g_eri_plan_size doubles, or NULL to use a buffer owned by the plan.
 * Asynchronous: returns after the launches are queued. */
int l2g_eri_plan_execute(l2g_eri_plan plan, double* d_out);

/* Waits for the context stream. */
int l2g_synchronize(l2g_context ctx);

/* Copies the plan-owned output to the host; nonzero (optional, nquartets
 * bytes) receives 0 for quartets whose every primitive was screened out,
 * where Engine::compute returns results()[0] == nullptr
 * (engine.impl.h:1781-1784). */
int l2g_eri_plan_download(l2g_eri_plan plan, double* host_out, int8_t* nonzero);

/* Device pointer of the plan-owned output (valid until destroy). */
int l2g_eri_plan_device_output(l2g_eri_plan plan, double** d_out);

/* Per-class breakdown of the plan: for each of the *nclasses classes
 * (up to cap), the class (la,lb,lc,ld) packed as la*1000+lb*100+lc*10+ld,
 * the number of quartets, the number of primitive quartets (before the
 * per-primitive screen) and the duration in ms of its last execution
 * (HIP events on the context stream).  The code adds 10000 for an FP32
 * kernel, 20000 for the generic kernel and 100000 for a wave per quartet
 * (l2g_eri_plan_set_wave_split). */
int l2g_eri_plan_stats(l2g_eri_plan plan, int cap, int* nclasses, int* class_code, int64_t* nquartets,
                       int64_t* nprim_quartets, float* ms);

/* Graph mode (on != 0): the next execution captures the plan's class
 * launches on the context stream into a hipGraph, which later executions
 * into the same output pointer replay with one hipGraphLaunch (a new output
 * pointer re-captures).  The graph holds kernel nodes only.  Per-class
 * timings in l2g_eri_plan_stats are then -1; l2g_eri_plan_elapsed gives the
 * whole execution.  No counterpart in the reference, whose callers loop over
 * Engine::compute per shell quartet (engine.impl.h:140-178); it removes the
 * per-class launch overhead of repeated executions of one plan. */
int l2g_eri_plan_set_graph(l2g_eri_plan plan, int on);

/* Operator of the plan: L2G_OPER_COULOMB (default), L2G_OPER_ERF_COULOMB
 * (erf(omega r)/r) or L2G_OPER_ERFC_COULOMB (erfc(omega r)/r), libint's
 * Operator::coulomb / erf_coulomb / erfc_coulomb (engine.h:158-167) with
 * omega its parameter.  The core integrals follow erf_coulomb_gm_eval and
 * erfc_coulomb_gm_eval (boys.h:1913-1985); the primitive screen stays on the
 * Coulomb prefactor, as in engine.impl.h:1362-1386.  erfc is evaluated as
 * the Coulomb integrals minus the erf ones (two passes per class). */
#define L2G_OPER_COULOMB 0
#define L2G_OPER_ERF_COULOMB 1
#define L2G_OPER_ERFC_COULOMB 2
int l2g_eri_plan_set_operator(l2g_eri_plan plan, int oper, double omega);

/* Wave split (on by default; on = 0 turns it off): classes of few, deeply
 * contracted quartets run a wave per quartet whose 64 lanes split the
 * primitive loop, instead of one lane per quartet, so that a handful of
 * long contractions no longer serialise on a few lanes.  With on = 1 a
 * class measured faster there on the generic kernel (its quartets split
 * over blocks by primitive range, summed per quartet in a fixed order) runs
 * on it instead when it holds at least half of the plan's quartets; on = 2
 * keeps every such class on its register kernel.  Each contraction
 * is then summed as 64 serial partial sums joined by a fixed butterfly,
 * another order than libint's serial primitive loop
 * (engine.impl.h:1306-1766).  Well-conditioned integrals agree with the
 * lane-per-quartet order to rounding; on ill-conditioned ones (large
 * cancelling terms) both orders stay within the error the reference's own
 * FP64 algorithm makes against the exact integral.  Off: such classes of
 * total angular momentum 7 and more run on the generic kernel (a block per
 * quartet, serial order), the rest a lane per quartet.  Whether a class
 * runs a wave per quartet depends on how many quartets of it the plan holds,
 * so the same quartet in plans of another composition can differ in its last
 * bits; with the split off a quartet's values depend on the quartet only.
 * No counterpart in the reference. */
int l2g_eri_plan_set_wave_split(l2g_eri_plan plan, int on);

/* Duration in ms of the plan's last execution (all classes; HIP events on
 * the context stream; waits for it to finish). */
int l2g_eri_plan_elapsed(l2g_eri_plan plan, float* ms);

int l2g_eri_plan_destroy(l2g_eri_plan plan);

/* ---- one-body integrals ------------------------------------------------ */

/* Overlap S, kinetic energy T and point-charge attraction
 * V = -sum_C q_C <a|1/|r - C||b> over every pair of shells of the basis
 * (charges[c] at xyz[3c..3c+2], bohr), as Engine(Operator::overlap /
 * kinetic / nuclear) computes them per shell pair (engine.impl.h:181-561,
 * charges as Engine::set_params for nuclear, engine.h:685-690) and the
 * reference's driver assembles them (tests/hartree-fock/hartree-fock.cc:
 * 480-535).  Each of S, T, V (nbf x nbf, row-major, the basis's function
 * order; may be NULL) is written whole.  Shells up to l = 3. */
int l2g_onebody(l2g_basis basis, int ncharges, const double* charges, const double* xyz, double* S, double* T,
                double* V);

/* ---- first geometric derivatives (deriv_order = 1) --------------------- */

/* Plans the first derivatives of a batch of quartets, replacing
 * Engine(Operator::coulomb, max_nprim, max_l, 1)::compute
 * (engine.h:502-526, 747-749; compute2<coulomb, xx_xx, 1>, engine.impl.h:
 * 1151-2113): per quartet 12 shell sets, d = 3 * centre + axis for the
 * shells' centres in caller order (CartesianDerivIterator<4>,
 * util/intpart_iter.h:89-103), each laid out like the integrals of
 * l2g_eri_plan_create.  Derivative set d of quartet q is at
 * out[d * total + offsets[q]].  Every shell needs l + 1 <= l2g_max_l()
 * (L2G_E_LMAX otherwise).  The solid-harmonics ordering is the context's
 * at execution. */
typedef struct l2g_deriv1_plan_s* l2g_deriv1_plan;
int l2g_eri_deriv1_plan_create(l2g_basis basis, int64_t nquartets, const int32_t* quartets, double precision,
                               int screening, l2g_deriv1_plan* plan);
/* doubles per derivative set (the output holds 12 * total) */
int l2g_eri_deriv1_plan_size(l2g_deriv1_plan plan, int64_t* total);
int l2g_eri_deriv1_plan_offsets(l2g_deriv1_plan plan, int64_t* offsets);
/* d_out: 12 * total device doubles, or NULL for the plan's own buffer */
int l2g_eri_deriv1_plan_execute(l2g_deriv1_plan plan, double* d_out);
/* host_out: 12 * total doubles (plan's own buffer only); nonzero[q] = 0 when
 * every primitive of the quartet's raised and lowered ERIs was screened out */
int l2g_eri_deriv1_plan_download(l2g_deriv1_plan plan, double* host_out, int8_t* nonzero);
int l2g_eri_deriv1_plan_destroy(l2g_deriv1_plan plan);

/* Schwarz factors K(s1,s2) = sqrt(max|(s1 s2|s1 s2)|) (norm = 0) or
 * sqrt(||(s1 s2|s1 s2)||_2) (norm = 1) for npairs shell pairs
 * (pairs[2*i], pairs[2*i+1]), on the GPU.  Replaces the per-pair loop of
 * compute_schwarz_ints (tests/hartree-fock/hartree-fock++.cc:1231-1298),
 * which uses precision 0 (no primitive screening); pass 0 for the same. */
int l2g_schwarz(l2g_basis basis, int64_t npairs, const int32_t* pairs, double precision, int norm, double* K);

/* The two per-pair factors that select FP32 in mixed batches
 * (l2g_eri_plan_create_mixed, l2g_fock_set_fp32_threshold): K as
 * l2g_schwarz with norm 0, and S = sqrt(max_e |[e0|e0]|) (1 + max_x|AB_x|)^lb
 * from the contracted pre-HRR integrals of the pair in its canonical
 * orientation.  A quartet runs in FP32 when K_ab K_cd < threshold and
 * S_ab S_cd < 16 threshold.  No counterpart in the reference, which has no
 * mixed precision; it is here so callers can see which quartets a
 * threshold sends to FP32. */
int l2g_precision_bounds(l2g_basis basis, int64_t npairs, const int32_t* pairs, double* K, double* S);

/* ---- Fock build ---------------------------------------------------------- */

/* Set-up of a screened two-electron Fock build for a basis, restating
 * compute_shellpairs and compute_schwarz_ints of
 * tests/hartree-fock/hartree-fock++.cc:1231-1436: the significant shell-pair
 * list (same centre, or overlap-block Frobenius norm >= pair_threshold,
 * 1e-12 there), the Schwarz matrix sqrt(max|(ab|ab)|) and shell-pair data at
 * max_engine_precision (epsilon/1e10 there) with the given primitive
 * screening (SchwarzInf there, :67).  Schwarz factors run on the GPU. */
int l2g_fock_create(l2g_basis basis, double pair_threshold, double max_engine_precision, int screening,
                    l2g_fock* fock);

/* Number of basis functions of the Fock build's basis. */
int l2g_fock_nbf(l2g_fock fock, int* nbf);

/* Copies the nshell x nshell Schwarz matrix (zero for insignificant pairs). */
int l2g_fock_schwarz(l2g_fock fock, double* K);

/* One screened Fock build for density D (nbf x nbf, row-major, symmetric),
 * replacing compute_2body_fock (hartree-fock++.cc:1574-1772): quartets with
 * Dnorm1234 K12 K34 < precision are skipped, each quartet is evaluated at
 * precision / Dnorm1234, integrals are digested on the GPU in FP64 with the
 * permutational degeneracy.  Outputs (any may be NULL, nbf x nbf row-major):
 * J_ab = sum (ab|cd) D_cd, K_ab = sum (ac|bd) D_cd, G = 2J - K. */
int l2g_fock_build(l2g_fock fock, const double* D, double precision, double* J, double* K, double* G);

/* Deterministic J/K accumulation (on != 0): every contribution is added as
 * a 128-bit fixed-point number (resolution 2^-80) with integer atomics, so
 * the sum does not depend on the order in which the GPU's threads arrive
 * and a build returns the same bits on every run with the same settings.
 * The reference sums per-thread Fock matrices in a fixed order
 * (tests/hartree-fock/hartree-fock++.cc:1752-1755); FP64 atomics (the
 * default, on = 0) round in a run-dependent order.  Costs a second word
 * plane of the J and K accumulators and about twice the atomics. */
int l2g_fock_set_deterministic(l2g_fock fock, int on);

/* Screened unique quartets a build with this density and precision would
 * digest (the count pass of l2g_fock_build, without the build). */
int l2g_fock_count(l2g_fock fock, const double* D, double precision, int64_t* nquartets);

/* Progress of long builds: after each window, fn(done, total, user) with
 * the quartets digested so far and the build's total (fn = NULL: none).
 * Called on the calling thread, between windows.  No counterpart in the
 * reference. */
typedef void (*l2g_progress_fn)(int64_t done, int64_t total, void* user);
int l2g_fock_set_progress(l2g_fock fock, l2g_progress_fn fn, void* user);

/* A build produces and digests its screened quartet list in windows of at
 * most `items` quartets (0: the default, 2^26), so that the list of a large
 * molecule never has to be held in device memory at once.  No counterpart
 * in the reference, whose CPU loop visits the quartets one at a time. */
int l2g_fock_set_window(l2g_fock fock, int64_t items);

/* Windows the last build's quartet list came in. */
int l2g_fock_stats_windows(l2g_fock fock, int64_t* windows);

/* Counters of the last build: quartets evaluated, primitive quartets (after
 * pair screening) and the GPU time of the digest kernels in ms. */
int l2g_fock_stats(l2g_fock fock, int64_t* nquartets, int64_t* nprim_quartets, float* kernel_ms);

/* Quartets of later builds whose Schwarz bound K12 K34 is below threshold
 * (and whose pre-HRR bound S12 S34 is below 16 threshold, see
 * l2g_precision_bounds) are evaluated in FP32; Boys function, prefactors,
 * digest and accumulation stay FP64; 0 = all FP64.  No libint counterpart. */
int l2g_fock_set_fp32_threshold(l2g_fock fock, double threshold);

/* Number of quartets the last build evaluated in FP32. */
int l2g_fock_stats_mixed(l2g_fock fock, int64_t* nquartets_fp32);

/* Per kernel class of the last build, as l2g_eri_plan_stats: class code
 * la*1000+lb*100+lc*10+ld (+10000 for FP32 kernels), quartets, primitive
 * quartets and the class's kernel time in ms (HIP events). */
int l2g_fock_class_stats(l2g_fock fock, int cap, int* nclasses, int* class_code, int64_t* nquartets,
                         int64_t* nprim_quartets, float* ms);

int l2g_fock_destroy(l2g_fock fock);

/* ---- Fock-build request server ------------------------------------------ */

/* A server runs Fock builds (as l2g_fock_build) for a stream of requests,
 * each with its own molecule and density.  It has no counterpart in the
 * reference, whose callers loop over builds one at a time on the CPU; the
 * reference-side equivalent of one request is compute_shellpairs +
 * compute_schwarz_ints + compute_2body_fock
 * (tests/hartree-fock/hartree-fock++.cc:1231-1772).
 *
 * Background admission workers build the pair list, Schwarz factors and
 * screened quartet list of several submitted requests in parallel, and
 * admit them in submission order, each as soon as a slot and device memory
 * are free; the request's
 * shell-pair data, density and J/K accumulators then stay resident on the
 * device until it completes.  A launch thread runs rounds of the Fock
 * kernels: each round takes up to round_quartets quartets, split across all
 * admitted requests, and launches each kernel class once over the quartets
 * of every request in it (continuous batching).  A request that arrives
 * while others run joins the next round after its admission. */
typedef struct l2g_server_s* l2g_server;

typedef struct {
  int device;                  /* HIP device */
  double hbm_bytes;            /* device memory for admitted requests' data; 0 = 16 GiB */
  int max_active;              /* requests admitted at once; 0 = 16 */
  int64_t round_quartets;      /* quartets per launch round across requests; 0 = 1 << 20 */
  int screening;               /* primitive screening; 0 = L2G_SCREEN_SCHWARZINF */
  double pair_threshold;       /* significant-pair threshold; 0 = 1e-12 */
  double max_engine_precision; /* shell-pair precision; 0 = epsilon / 1e10 */
  int deterministic;           /* 1: fixed-point J/K accumulation (as l2g_fock_set_deterministic),
                                  so a request's result has the same bits alone and in any mix */
} l2g_server_config;

#define L2G_REQ_QUEUED 0   /* submitted, not yet admitted */
#define L2G_REQ_ACTIVE 1   /* admitted: data resident, quartets in launch rounds */
#define L2G_REQ_DONE 2     /* J, K, G ready for l2g_request_result */
#define L2G_REQ_FAILED 3   /* l2g_request_wait returns the error */

int l2g_server_create(const l2g_server_config* config, l2g_server* server);

/* Submits one Fock build: a basis (as l2g_basis_create, coefficients
 * renormalized), its density D (nbf x nbf row-major) and the Fock precision
 * (as l2g_fock_build); fp32_threshold as l2g_fock_set_fp32_threshold.  All
 * inputs are copied; returns at once with the request id. */
int l2g_request_submit(l2g_server server, int nshell, const int* l, const int* pure, const int* nprim,
                       const double* alpha, const double* coeff, const double* origin, const double* D,
                       double precision, double fp32_threshold, int64_t* id);

/* State of a request (L2G_REQ_*). */
int l2g_request_state(l2g_server server, int64_t id, int* state);

/* Blocks until the request is done (returns 0) or failed (returns its error). */
int l2g_request_wait(l2g_server server, int64_t id);

/* Copies the results of a done request (J, K, G: nbf x nbf row-major, any
 * may be NULL) and its times in ms (ms[0] submit to admission, ms[1] submit
 * to done; may be NULL), then releases the request. */
int l2g_request_result(l2g_server server, int64_t id, double* J, double* K, double* G, double* ms);

/* Counters: launch rounds run, the most requests that shared one round,
 * work items launched (a quartet, or one primitive range of a deep quartet
 * that was split across lanes). */
int l2g_server_stats(l2g_server server, int64_t* rounds, int64_t* max_requests_per_round, int64_t* nquartets);

/* Waits for every submitted request to finish, then stops the threads and
 * frees the server (results not yet collected are dropped). */
int l2g_server_destroy(l2g_server server);

/* ---- single quartets: the per-call path of Engine::compute ------------- */

/* A resident evaluator of one shell quartet per call, for callers that loop
 * over quartets as libint's do (Engine::compute, engine.h:747-749;
 * compute2<coulomb, xx_xx, 0>, :787-791; the driver loop of
 * tests/hartree-fock/hartree-fock++.cc:1628-1743).  Creation allocates the
 * device buffers and pinned host staging for shells of up to max_nprim
 * primitives and l <= max_l; a call then makes one host-to-device copy, one
 * kernel launch per operator pass and one device-to-host copy, and waits
 * for them.  Not thread-safe (one per thread, as libint's Engine). */
typedef struct l2g_quartet_engine_s* l2g_quartet_engine;
int l2g_quartet_engine_create(l2g_context ctx, int max_nprim, int max_l, l2g_quartet_engine* engine);
int l2g_quartet_engine_destroy(l2g_quartet_engine engine);

/* Integrals of the quartet (s0 s1|s2 s3) in caller order, libint's layout
 * (n0 n1 n2 n3 doubles into out).  Shells as l2g_basis_create with nshell 4
 * (l[4], pure[4], nprim[4], alpha and coeff concatenated, origin[12]).
 * ln_scr = NULL: the shell pairs are built here with `screening`, Original
 * or Conservative (as compute2 without ShellPair arguments,
 * engine.impl.h:1212-1216).  Otherwise the caller's precomputed shell pairs
 * (compute2 with spbra/spket, engine.impl.h:1249-1285; ShellPair::init,
 * shell.h:1137-1328): ln_scr holds every primitive pair's ln_scr of the bra
 * pair (s0, s1), n0 n1 values p0-major, then of the ket pair (s2, s3), n2 n3
 * values; -INFINITY marks a primitive pair that is not in the ShellPair's
 * list; nonsph (optional) the same for nonsph_screen_fac.  Any screening
 * method is then allowed.  oper/omega as l2g_eri_plan_set_operator;
 * scale = prescale_by, uniform_cart as l2g_eri_plan_create.  *nonzero
 * (optional) receives 0 when every primitive quartet was screened out
 * (results()[0] == nullptr, engine.impl.h:1781-1784). */
int l2g_quartet_compute(l2g_quartet_engine engine, const int* l, const int* pure, const int* nprim,
                        const double* alpha, const double* coeff, const double* origin, const double* ln_scr,
                        const double* nonsph, double precision, int screening, double scale, int uniform_cart,
                        int oper, double omega, double* out, int8_t* nonzero);

/* The same evaluation for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in libint: it is l2g_quartet_compute without an engine.  It
 * calls no hipMalloc, hipFree or hipHostMalloc: its device memory is a
 * fixed array of the library (about 9.5 MB per GPU, held until the process
 * ends) that l2g_init sets up for its device, and its host staging is
 * static per thread.  Limits: l <= 3 and at most L2G_NOALLOC_MAX_NPRIM
 * primitives per shell; beyond them the call fails with an error naming
 * the limit.  Calls on one GPU share its array and are serialised by the
 * library; the array keeps the last call's inputs and integrals.  The
 * arguments and results are those of l2g_quartet_compute. */
#define L2G_NOALLOC_MAX_NPRIM 16
int l2g_quartet_compute_noalloc(l2g_context ctx, const int* l, const int* pure, const int* nprim,
                                const double* alpha, const double* coeff, const double* origin, const double* ln_scr,
                                const double* nonsph, double precision, int screening, double scale, int uniform_cart,
                                int oper, double omega, double* out, int8_t* nonzero);

/* One-shot convenience: plan, execute, download, destroy.  offsets and
 * nonzero may be NULL. */
int l2g_eri_compute(l2g_basis basis, int64_t nquartets, const int32_t* quartets, double precision, int screening,
                    double* host_out, int64_t* offsets, int8_t* nonzero);

#ifdef __cplusplus
}
#endif
#endif /* L2G_H */
