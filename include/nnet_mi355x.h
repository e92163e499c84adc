/* C-ABI of libnnet_mi355x.so: the MI355X (gfx950) implementation of SPH-EXA-nnet's batched
 * implicit nuclear-network solve.
 *
 * It replaces the device half of
 *   template<bool use_gpu, class Data, typename Float, class nseFunction>
 *   void nnet::parallel::computeNuclearReactions(Data& n, size_t firstIndex, size_t lastIndex,
 *       Float hydro_dt, Float previous_dt, const nnet::ReactionList& reactions,
 *       const nnet::ComputeReactionRatesFunctor<Float>& construct_rates_BE,
 *       const nnet::EosFunctor<Float>& eos, bool use_drhodt, nseFunction jumpToNse)
 *                                     (include/nnet/parallel/parallel_nuclear_net.hpp:86-172)
 * i.e. the reference's nuclear_net.cuh moveToGpu (:61-72) + cudaComputeNuclearReactions
 * (parallel/CUDA/parallel_nuclear_net.cu:116-163).  Plain pointers and sizes only; all particle
 * arrays are caller-owned device memory in the reference's structure-of-arrays layout.
 *
 * Every entry point returns 0 on success or a negative NNET_E* code; nnet_last_error() gives the
 * message.  Nothing throws across this boundary.
 */
#ifndef NNET_MI355X_H
#define NNET_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define NNET_OK 0
#define NNET_E_ARG (-1)       /* invalid argument (null pointer, size, unknown kind) */
#define NNET_E_HIP (-2)       /* HIP runtime error (allocation, launch, copy) */
#define NNET_E_TABLE (-3)     /* Helmholtz table missing or malformed */
#define NNET_E_NETWORK (-4)   /* reaction list inconsistent with the rate parameterization */
#define NNET_E_STATE (-5)     /* context not ready for this call (e.g. no Helmholtz table loaded) */

/* rate parameterizations (construct_rates_BE functor types):
 * nnet::net14/net86/net87::ComputeReactionRatesFunctor<double> */
#define NNET_RATES_NET14 14
#define NNET_RATES_NET86 86
#define NNET_RATES_NET87 87

/* EOS functor types: nnet::eos::IdealGasFunctor<double>, nnet::eos::HelmholtzFunctor<double> */
#define NNET_EOS_IDEAL 0
#define NNET_EOS_HELMHOLTZ 1

/* arithmetic of the Newton linear solve */
#define NNET_PRECISION_FP64 0  /* plain FP64 everywhere (parity path) */
#define NNET_PRECISION_MIXED 1 /* FP32 LU on MFMA + FP64 iterative refinement (<= 3 sweeps on
                                  net14, 1 on net86/net87); a particle with an iteration that does
                                  not converge is integrated by the FP64 path instead */

/* per-particle status written to nnet_batch.status */
#define NNET_STATUS_OK 0      /* integrated to hydro_dt */
#define NNET_STATUS_SKIPPED 1 /* rho <= 1e5 or T <= 1e8: untouched (parallel_nuclear_net_impl.cpp:83) */
#define NNET_STATUS_STALLED 2 /* the reference controller would replay a rejected step forever
                                 (nuclear_net.hpp:744-757, 899-904); state left at the last
                                 accepted substep */
#define NNET_STATUS_ZERO_DT 3 /* substep of length 0 (the reference throws, nuclear_net.hpp:533-548) */
#define NNET_STATUS_MAX_ITER 4 /* Newton-iteration cap reached (guard; the reference has none) */

typedef struct nnet_ctx nnet_ctx;
typedef struct nnet_network nnet_network;

/* Context on one HIP device.  Replaces the reference's per-call cudaMalloc/cudaMemcpy of
 * tables and functors (helmholtz.cpp:174-218, net87.cpp:89-99): constant tables are uploaded
 * once here.  helm_table_path may be NULL (ideal gas only) or the path of a binary Helmholtz
 * table as written by the package's table generator (sph-exa-nnet_amd/helm/helm_table_gen). */
int nnet_ctx_create(nnet_ctx** ctx, int device, const char* helm_table_path);
void nnet_ctx_destroy(nnet_ctx* ctx);
const char* nnet_last_error(const nnet_ctx* ctx);

/* Load (or replace) the Helmholtz table: 21 arrays of 541x201 doubles in the reference order
 * f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt, dpdf, dpdfd, dpdft, dpdfdt, ef, efd, eft, efdt,
 * xf, xfd, xft, xfdt, each [density][temperature] (helmholtz.cpp:100-137), as a binary file of
 * exactly that content or as the reference's text table (see nnet_helm_table_read). */
int nnet_ctx_load_helmholtz(nnet_ctx* ctx, const char* path);

/* Read a Helmholtz table file into host memory `out` (21 x 541 x 201 doubles in the layout
 * above), without a device: the binary layout, or the reference's text table -- for each
 * temperature j and density i the nine f values, then the dpdf, ef and xf blocks of four,
 * whitespace-separated, optionally wrapped in STR( ... ) -- exactly as readCPUTable parses it
 * (helmholtz.cpp:89-137).  nnet_ctx_load_helmholtz accepts the same two formats.  Returns
 * NNET_E_TABLE for a missing, short or malformed file. */
int nnet_helm_table_read(const char* path, double* out);

/* Network = reaction list + rate parameterization.  The list uses the reference's flat
 * ReactionList layout (nuclear_net.hpp:168-213): reaction r has reactants
 * [reactant_begin[r], product_begin[r]) and products [product_begin[r], reactant_begin[r+1])
 * of (species_id, count).  reactant_begin has n_reactions+1 entries.  The order must be the
 * rate functor's push order (as in net14::reactionList etc.). */
int nnet_network_create(nnet_ctx* ctx, int rates_kind, int n_species, int n_reactions,
                        const int* reactant_begin, const int* product_begin,
                        const int* species_id, const int* count, nnet_network** net);
/* The library's own copy of net14::reactionList / net86::reactionList / net87::reactionList. */
int nnet_network_builtin(nnet_ctx* ctx, int rates_kind, nnet_network** net);
void nnet_network_destroy(nnet_network* net);
/* Copy out a network's reaction list (host buffers sized by the caller; NULL to query sizes). */
int nnet_network_reactions(const nnet_network* net, int* n_reactions, int* n_entries,
                           int* reactant_begin, int* product_begin, int* species_id, int* count);

/* One batch of particles sharing a network and an EOS. */
typedef struct nnet_batch {
    const nnet_network* network;
    int eos;              /* NNET_EOS_* */
    double mu;            /* IdealGasFunctor(mu) */
    const double* Z;      /* HelmholtzFunctor(Z, z_dim): host array of z_dim charges */
    int z_dim;
    int precision;        /* NNET_PRECISION_* */
    size_t n;             /* particles [0, n) of the arrays below */
    const double* rho;    /* device, n */
    double* rho_m1;       /* device, n (read/written only if use_drhodt) */
    double* const* Y;     /* host array of n_species device pointers, each n molar abundances */
    double* temp;         /* device, n (in/out) */
    double* dt;           /* device, n: nuclear timestep, persistent solver state (in/out) */
    uint8_t* status;      /* device, n, optional (NULL) */
    uint32_t* substeps;   /* device, n, optional: accepted substeps this call */
    uint32_t* newton;     /* device, n, optional: Newton iterations this call */
    uint32_t* fp64_solves; /* device, n, optional: mixed precision only, Newton iterations solved
                              in FP64.  A particle with an iteration whose FP32 factors do not
                              refine to FP64 accuracy is integrated again from the start of the
                              step by the FP64 pass: then fp64_solves = newton; else 0. */
    const uint32_t* cost_hint; /* device, n, optional: expected relative cost per particle (e.g. the
                              previous call's `newton` counts; the same array may be passed as
                              `newton`).  Particles are then dispatched in descending buckets of
                              the hint, so the longest integrations start first.  Results do not
                              depend on it. */
    const uint32_t* fp64_hint; /* device, n, optional, mixed precision only: a particle with a
                              nonzero entry skips the FP32 attempt and is integrated by the FP64
                              pass (its result is then the FP64 path's).  Passing the previous
                              call's fp64_solves (the same array may be passed as fp64_solves)
                              keeps the particles whose FP32 factors did not refine on FP64.
                              Honoured in calls with one non-empty batch outside HIP graph
                              capture; ignored otherwise. */
} nnet_batch;

/* use_drhodt of nnet_compute: how rho_m1 (the previous hydro step's density) enters */
#define NNET_DRHODT_OFF 0       /* not used (drho/dt = 0) */
#define NNET_DRHODT_CPU 1       /* the reference CPU path (parallel_nuclear_net_impl.cpp:89-92):
                                   rho_m1[i] (if nonzero) is overwritten by (rho - rho_m1)/previous_dt
                                   and the integration still uses drho/dt = 0 */
#define NNET_DRHODT_INTEGRATE 2 /* the reference GPU branch (parallel_nuclear_net.cu:79-81):
                                   drho/dt = (rho - rho_m1)/previous_dt (0 where rho_m1 is 0), used
                                   in the integration, rho(t) = rho - drho/dt (hydro_dt - elapsed)
                                   (nuclear_net.hpp:859); rho_m1 is left unchanged */

/* computeNuclearReactions for nb batches (different networks may share one call).
 * Stream-ordered on `stream` (hipStream_t, NULL = default stream); returns after enqueueing.
 * Semantics follow the reference CPU path (parallel_nuclear_net_impl.cpp:63-103): particles with
 * rho <= 1e5 or T <= 1e8 are skipped; use_drhodt is one of NNET_DRHODT_* (as in the reference,
 * the division by previous_dt happens only where rho_m1 is nonzero). */
int nnet_compute(nnet_ctx* ctx, const nnet_batch* batches, int nb, double hydro_dt,
                 double previous_dt, int use_drhodt, void* stream);

/* nnet_compute without any allocation: the same computation, with the library's work-queue
 * counters, ordering bins, dispatch order and redo lists placed in a fixed device scratch array
 * instead of buffers the library grows with hipMalloc.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The scratch is 64 KiB + 16 MiB per GPU, held from the first context on that
 * GPU until the process ends; a call may hold at most 119 batches and at most 4,194,304 list
 * entries -- one per particle of each batch with a cost_hint (or an honoured fp64_hint), plus one
 * per particle of each mixed-precision batch -- and is refused with NNET_E_ARG beyond that, before
 * anything is enqueued.  Every call on a GPU shares that GPU's scratch: calls that overlap on one
 * GPU (other streams, other threads, other contexts) race.  Same streams, capture rules and
 * results as nnet_compute. */
int nnet_compute_noalloc(nnet_ctx* ctx, const nnet_batch* batches, int nb, double hydro_dt,
                         double previous_dt, int use_drhodt, void* stream);

/* Cap on Newton iterations per particle per call (default 10000; 0 = no cap).  The reference
 * controller can also cycle between two rejected timesteps for thousands of attempts before (or
 * instead of) accepting one; the cap bounds that work.  A capped particle is left at its last
 * accepted substep with status NNET_STATUS_MAX_ITER. */
int nnet_ctx_set_max_iterations(nnet_ctx* ctx, long max_iter);

/* Particles a solve launch integrates concurrently on this context's device: the workgroups of
 * the persistent kernel for (rates_kind, eos, precision) that stay resident (occupancy per CU x
 * CUs) times the particles each integrates at once.  No counterpart in the reference (whose
 * launch is one thread per particle, parallel_nuclear_net.cu:144-156); for tools that model or
 * tune the launch. */
int nnet_solve_slots(nnet_ctx* ctx, int rates_kind, int eos, int precision, int* slots);

/* Particles per workgroup of the solve launches for (rates_kind, eos, precision): net14 (one
 * wavefront per particle) 1, 2 or 4 (default 1), net86 / net87 (four wavefronts per particle)
 * 1.  Results do not depend on it.  Replaces the reference's compile-time launch shape
 * CUDA_NUM_THREAD_PER_BLOCK_NNET, particles per thread block
 * there (parallel/CUDA/parallel_nuclear_net.cuh:51-64); the autotuner
 * (sph-exa-nnet_amd/autotune.py) measures the choices and writes them to its table. */
int nnet_ctx_set_particles_per_workgroup(nnet_ctx* ctx, int rates_kind, int eos, int precision, int ppw);
int nnet_ctx_get_particles_per_workgroup(nnet_ctx* ctx, int rates_kind, int eos, int precision, int* ppw);

/* Calls whose batches belong to several networks (the reference makes one call per network,
 * parallel_nuclear_net.hpp:86-90): fused = 1 (default) runs them as one persistent launch whose
 * workgroups drain the batches' work queues in turn (net87, net86, then net14 at four particles
 * per workgroup) when the batches share the EOS and the precision, number at most 4 and hold at
 * most 192 species and 192 charges in all; otherwise, and with fused = 0, each batch is its own
 * launch on a side stream.  Results do not depend on it.  Chosen by the autotuner. */
int nnet_ctx_set_fused_launch(nnet_ctx* ctx, int fused);
int nnet_ctx_get_fused_launch(nnet_ctx* ctx, int* fused);

/* The launch parameters above as one record, and the autotuner's table that holds them: a text
 * file of lines "ppw <rates_kind> <eos> <precision> <particles per workgroup>" and
 * "fused <0|1>" ('#' starts a comment; settings not listed keep their defaults).
 * sph-exa-nnet_amd/autotune.py measures the choices on the device and writes the table;
 * nnet_ctx_create loads the table named by the environment variable NNET_MI355X_TUNING, else
 * sph-exa-nnet_amd/data/tuning_gfx950.txt next to the library's lib/ directory when it exists
 * (a malformed table fails the creation with NNET_E_ARG).  No counterpart in the reference,
 * whose launch shape is a compile-time constant (parallel/CUDA/parallel_nuclear_net.cuh:51-64). */
typedef struct nnet_tuning {
    int ppw[3][2][2]; /* particles per workgroup, [net14, net86, net87][eos][precision] */
    int fused;        /* see nnet_ctx_set_fused_launch */
} nnet_tuning;
/* Host only (no device): parse a table into `out`, starting from the defaults. */
int nnet_tuning_read(const char* path, nnet_tuning* out);
int nnet_ctx_load_tuning(nnet_ctx* ctx, const char* path);
int nnet_ctx_get_tuning(nnet_ctx* ctx, nnet_tuning* out);

/* Blocking variant of nnet_compute (drop-in for the reference's synchronous launch). */
int nnet_compute_sync(nnet_ctx* ctx, const nnet_batch* batches, int nb, double hydro_dt,
                      double previous_dt, int use_drhodt);

/* Post-step EOS refresh for the hydro: replaces the device half of
 *   template<bool use_gpu, class Data, class Vector>
 *   void nnet::parallel::computeHelmEOS(Data& n, size_t firstIndex, size_t lastIndex, const Vector& Z)
 *                                     (include/nnet/parallel/parallel_nuclear_net.hpp:191-261)
 * i.e. cudaComputeHelmholtz (parallel/CUDA/parallel_nuclear_net.cu:185-226), with the semantics
 * of the CPU path computeHelmholtzImpl (parallel_nuclear_net_impl.cpp:116-141): for each particle
 * i in [0, n), abar_ = sum_j Y[j][i] and zbar_ = sum_j Y[j][i] Z[j] (index order), and the
 * Helmholtz EOS (helmholtz.cpp:340-965) at (temp[i], rho[i]) gives u (specific internal energy),
 * cv, p, c (sound speed) and dP/dT.  Z: host array of `dimension` charges (1 <= dimension <= 128);
 * Y: host array of `dimension` device pointers, each n abundances; temp, rho: device, n; each
 * output (device, n) may be NULL to skip it.  Needs a loaded Helmholtz table.  Stream-ordered. */
int nnet_helm_eos(nnet_ctx* ctx, size_t n, int dimension, const double* Z, const double* temp,
                  const double* rho, const double* const* Y, double* u, double* cv, double* p,
                  double* c, double* dpdT, void* stream);
/* Blocking variant of nnet_helm_eos (the reference's GPU branch synchronises,
 * parallel_nuclear_net.hpp:234). */
int nnet_helm_eos_sync(nnet_ctx* ctx, size_t n, int dimension, const double* Z, const double* temp,
                       const double* rho, const double* const* Y, double* u, double* cv, double* p,
                       double* c, double* dpdT);

#ifdef __cplusplus
}
#endif
#endif
