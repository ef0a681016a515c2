/*
 * ljgpu.h — C ABI of the MI355X (gfx950) Lennard-Jones force library.
 *
 * The hot path of kohnakagawa/lj_gpu: p += F(q) * dt over a Verlet neighbour
 * list, with F the Lennard-Jones force (epsilon = sigma = 1, no shift),
 * positions q fixed during a call. Every entry point below names the
 * reference interface it replaces (paths relative to the reference checkout).
 *
 * Conventions
 *   - Every function returns an int status: LJGPU_OK (0) or a negative
 *     LJGPU_ERR_* code. Nothing aborts and nothing throws across this ABI;
 *     ljgpu_last_error() returns a message for the calling thread.
 *   - Particle counts are int32 (as in the reference), list lengths int64.
 *   - Positions/momenta: AoS is double[N][4] (x, y, z, pad: the reference's
 *     double4 / cpu_ref q[N][4]); SoA is three double[N] arrays (the
 *     reference's q[X][N], q[Y][N], q[Z][N]).
 *   - box > 0 selects a periodic cubic box of that side (minimum image on
 *     every pair; positions must lie in [0, box)); box <= 0 is the
 *     reference's open box.
 *   - Momenta are accumulated into, never overwritten (the reference kernels'
 *     p += F*dt semantics).
 *   - A context owns one HIP stream (default priority). Calls on one context
 *     are not thread-safe; use one context per host thread.
 */
#ifndef LJGPU_H
#define LJGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define LJGPU_OK 0
#define LJGPU_ERR_INVALID_ARGUMENT (-1)
#define LJGPU_ERR_OUT_OF_MEMORY (-2)
#define LJGPU_ERR_HIP (-3)
#define LJGPU_ERR_BAD_LIST (-4)
#define LJGPU_ERR_NO_LIST (-5)
#define LJGPU_ERR_CAPACITY (-6)
#define LJGPU_ERR_NO_DEVICE (-7)
#define LJGPU_ERR_IO (-8)
#define LJGPU_ERR_UNSUPPORTED (-9)

/* ---- enums ------------------------------------------------------------ */
/* List kinds: the reference's action-reaction half list (i < j, built with
 * EN_ACTION_REACTION, cuda/force_cuda.cu:102-120; cpu_ref/force_aos.cpp:60-74)
 * and its full list (every ordered pair i != j, cuda/force_cuda.cu:133-136). */
#define LJGPU_LIST_HALF 0
#define LJGPU_LIST_FULL 1

/* Where caller buffers live. */
#define LJGPU_MEM_HOST 0
#define LJGPU_MEM_DEVICE 1

/* Force precision: FP64 throughout, or the mixed kernels. */
#define LJGPU_PREC_F64 0
#define LJGPU_PREC_MIXED 1

/* Per-particle accumulator of the mixed-precision kernels. */
#define LJGPU_ACCUM_F64 0
#define LJGPU_ACCUM_KAHAN_F32 1

/* How a half list is evaluated on the device. */
#define LJGPU_HALF_AUTO 0       /* library choice (autotuned)               */
#define LJGPU_HALF_ATOMIC 1     /* scatter to j with FP64 atomics, like the
                                   reference's *_with_aar kernels
                                   (cuda/kernel.cuh:238-469); summation
                                   order is not deterministic               */
#define LJGPU_HALF_EXPAND 2     /* expand to the full list once at
                                   set-list time; deterministic             */

/* Order of the particles in LDS tiles (ljgpu_set_tile_order). */
#define LJGPU_TILE_ORDER_INDEX 0    /* blocks of consecutive indices        */
#define LJGPU_TILE_ORDER_SPATIAL 1  /* blocks of nearby particles (default) */

typedef struct ljgpu_ctx ljgpu_ctx;

typedef struct ljgpu_list_info {
  int32_t particle_number;
  int32_t kind;            /* LJGPU_LIST_HALF or LJGPU_LIST_FULL (as given) */
  int64_t n_entries;       /* entries of the list as given                  */
  int64_t n_pairs;         /* unique pairs (half: n_entries; full: /2)      */
  int32_t max_partners;    /* longest row                                   */
  int32_t min_partners;    /* shortest row                                  */
  int32_t half_mode;       /* LJGPU_HALF_ATOMIC / LJGPU_HALF_EXPAND in use  */
  int32_t lanes_per_particle; /* kernel configuration in use                */
  int32_t unroll;
  int32_t block_size;
  int32_t tile_particles;  /* LDS tiling in use: particles per block (0: off) */
  int32_t tile_blocks;     /* number of LDS tiles                            */
  int32_t tile_max_stage;  /* most positions one tile stages in LDS          */
} ljgpu_list_info;

/* ---- context, errors, device memory ----------------------------------- */
/* Replaces the reference's global cuda_ptr<T> buffers and implicit default
 * stream (cuda/cuda_ptr.cuh:10-104, cuda/force_cuda.cu:24-30,265-289). */
int ljgpu_device_count(int* count);
int ljgpu_create(int device, ljgpu_ctx** out);
int ljgpu_destroy(ljgpu_ctx* ctx);
int ljgpu_sync(ljgpu_ctx* ctx);
const char* ljgpu_last_error(void);
const char* ljgpu_version(void);

/* Plain device memory (cuda_ptr::allocate / host2dev / dev2host / set_val,
 * cuda/cuda_ptr.cuh:37-87). Copies are synchronous with respect to the host
 * and ordered on the context stream. */
int ljgpu_malloc(ljgpu_ctx* ctx, size_t bytes, void** out);
int ljgpu_free(ljgpu_ctx* ctx, void* ptr);
int ljgpu_memcpy_htod(ljgpu_ctx* ctx, void* dst, const void* src, size_t bytes);
int ljgpu_memcpy_dtoh(ljgpu_ctx* ctx, void* dst, const void* src, size_t bytes);
int ljgpu_memcpy_dtod(ljgpu_ctx* ctx, void* dst, const void* src, size_t bytes);
int ljgpu_memset(ljgpu_ctx* ctx, void* dst, int value, size_t bytes);

/* Stream timer: HIP events recorded on the context stream around the work
 * issued between start and stop (the reference's gettimeofday pair around
 * measure()'s loop, cuda/force_cuda.cu:330-341). */
int ljgpu_timer_start(ljgpu_ctx* ctx);
int ljgpu_timer_stop(ljgpu_ctx* ctx, float* elapsed_ms);

/* ---- neighbour list --------------------------------------------------- */
/* Install a CSR neighbour list: number_of_partners[N], pointer[N] (row
 * start), sorted_list[n_entries] — the arrays makepair()/sortpair() build
 * (cuda/force_cuda.cu:122-163; cpu_ref/force_aos.cpp:100-140) and every
 * force_kernel_* reads (cuda/kernel.cuh:5-13). The list is validated as the
 * reference's check_loadedpair() does (cuda/force_cuda.cu:183-201, plus
 * j != i, row ends inside the list and half-list j > i) and copied; device
 * inputs are copied too, so the caller keeps ownership. Internal layouts
 * (expanded full list for LJGPU_HALF_EXPAND) are derived here, once. */
int ljgpu_set_list_csr(ljgpu_ctx* ctx, int32_t particle_number,
                       const int32_t* number_of_partners, const int32_t* pointer,
                       const int32_t* sorted_list, int64_t n_entries, int kind, int mem);

/* Install the reference's ELL ("transposed") list: entry k of row i at
 * transposed_list[i + k*N], rows zero-padded to `width`
 * (make_transposed_pairlist, cuda/force_cuda.cu:229-240; read by
 * force_kernel_memopt2 and relatives, cuda/kernel.cuh:102-236,471-819). */
int ljgpu_set_list_transposed(ljgpu_ctx* ctx, int32_t particle_number,
                              const int32_t* number_of_partners,
                              const int32_t* transposed_list, int32_t width, int kind,
                              int mem);

int ljgpu_get_list_info(ljgpu_ctx* ctx, ljgpu_list_info* info);

/* Copy the evaluated list back as CSR (caller buffers on the host): the list
 * as installed, with rows stored in order, or for a half list in
 * LJGPU_HALF_EXPAND mode its expanded full list. */
int ljgpu_get_list_csr(ljgpu_ctx* ctx, int32_t* number_of_partners, int32_t* pointer,
                       int32_t* sorted_list, int64_t capacity);

/* Build the Verlet list on the device from device positions and install it
 * (no host round trip): the reference's makepair() + sortpair()
 * (cpu_ref/force_aos.cpp:100-140; full list cuda/force_cuda.cu:122-163) as a
 * cell list, with the identical pair set and CSR order (predicate
 * r2 = (dx*dx + dy*dy) + dz*dz < search_length^2 evaluated unfused, rows in
 * ascending i, partners ascending; minimum image for box > 0, positions in
 * [0, box)). q is AoS double[N][4] (16-byte aligned) or three SoA arrays.
 * A half list in LJGPU_HALF_EXPAND mode is built directly as its full list.
 * Every derived layout (tiles, ELL) is built on the device as well. */
int ljgpu_build_list(ljgpu_ctx* ctx, const double* q, int32_t particle_number, double box,
                     double search_length, int kind);
int ljgpu_build_list_soa(ljgpu_ctx* ctx, const double* qx, const double* qy, const double* qz,
                         int32_t particle_number, double box, double search_length, int kind);
/* The same from packed double[N][3] positions (8-byte aligned), the
 * reference's double3 buffers (cuda/force_cuda.cu:26-27). */
int ljgpu_build_list_aos3(ljgpu_ctx* ctx, const double* q, int32_t particle_number, double box,
                          double search_length, int kind);

/* Batched replicas: n_replicas independent systems stored back to back,
 * replica r holding particles [offsets[r], offsets[r+1]) in a cubic box of
 * side boxes[r] (periodic for boxes[r] > 0, open otherwise); host arrays,
 * offsets[0] = 0. Afterwards the list (ljgpu_set_list_* with
 * particle_number = offsets[n_replicas], every pair inside one replica, or
 * ljgpu_build_list with box = 0, built per replica) and every force call
 * (box = 0) cover all replicas in one launch of the tiled kernel: its tiles
 * never straddle two replicas and each carries its replica's box. Replaces
 * the installed list; n_replicas = 0 returns to a single system. No
 * reference counterpart (the reference runs one system). */
int ljgpu_set_replicas(ljgpu_ctx* ctx, int32_t n_replicas, const int32_t* offsets,
                       const double* boxes);

/* Half-list evaluation mode (LJGPU_HALF_*); takes effect on the next
 * ljgpu_set_list_* call. */
int ljgpu_set_half_mode(ljgpu_ctx* ctx, int half_mode);

/* Kernel configuration: lanes cooperating on one i-particle (1 selects the
 * thread-per-particle ELL kernel; 4, 8, 16, 32 or 64 the CSR
 * wavefront-segment kernel), the workgroup size (64..1024, multiple of 64)
 * and the list entries each lane keeps in flight per pipelined batch
 * (unroll 1, 2 or 4). 0 keeps the current value. The reference exposes only
 * the block size (argv[1], cuda/force_cuda.cu:377-383). */
int ljgpu_set_kernel_config(ljgpu_ctx* ctx, int lanes_per_particle, int block_size, int unroll);

/* LDS tiling of full lists (on by default, 64 particles per tile): the list
 * is regrouped once per list into blocks of particles whose partners'
 * positions are staged in LDS, and each entry is re-expressed as a 16-bit
 * LDS offset. The particles of a block are consecutive in the order chosen
 * by ljgpu_set_tile_order. 0 turns tiling off (global-gather
 * kernels). The tile actually used is capped at 1024 / lanes_per_particle
 * particles (one workgroup per tile); ELL lists and half lists evaluated
 * with atomics are never tiled. stage_offsets = 1 also stages each tile's
 * entry offsets in LDS (more LDS per workgroup); 0 streams them from HBM.
 * resident_per_cu > 0 launches a persistent grid of that many workgroups per
 * compute unit, each looping over tiles; 0 launches one workgroup per tile.
 * No reference counterpart: it plays the role of the reference's derived
 * transposed list (cuda/force_cuda.cu:229-240), a layout built once from the
 * CSR for the device. */
int ljgpu_set_tiling(ljgpu_ctx* ctx, int32_t tile_particles, int stage_offsets,
                     int32_t resident_per_cu);

/* Copy the LDS tiling in use back to the host, in the layout
 * ljgpu_make_tiles produces (derived on the device for the current kernel
 * configuration; *n_blocks = 0 when the kernels run untiled). Pass null
 * outputs to query the sizes first; local needs one value per entry of the
 * evaluated list. The list of a periodic ljgpu_md_run is tiled without the
 * bank-class row order, but that run leaves no list installed. Inspection
 * only; the tiling's counterpart in the reference is the derived transposed
 * list (cuda/force_cuda.cu:229-240). */
int ljgpu_get_tiles(ljgpu_ctx* ctx, int32_t* desc, int64_t desc_capacity, int32_t* stage,
                    int64_t stage_capacity, uint16_t* local, int64_t local_capacity,
                    int32_t* n_blocks, int64_t* n_stage);

/* Order of the particles in tiles. LJGPU_TILE_ORDER_SPATIAL (the default)
 * sorts each periodic system into columns along x, about 0.75 *
 * cbrt(tile_particles / density) wide, and each column into x bins of about
 * one particle, ordered by (column, bin, index); then, within each block of
 * tile_particles consecutive slots, by row length (longest first, ties kept
 * in order), so that a wavefront's rows need similar numbers of batches. A
 * block of consecutive particles in that order is a compact group that
 * stages fewer partner positions than a block of consecutive indices. The order needs positions:
 * ljgpu_build_list takes it from its own, a list installed with
 * ljgpu_set_list_* from the first force evaluation (index order until
 * then). Open boxes, and bins of more than 256 particles, keep index order.
 * Sums run over each row in its own order either way, so the order changes
 * results only by the rounding of the periodic unwrap. No reference
 * counterpart (a device layout choice). */
int ljgpu_set_tile_order(ljgpu_ctx* ctx, int order);

/* The order of the tiling in use (*order = LJGPU_TILE_ORDER_INDEX or
 * _SPATIAL) and, when member is not null, the particle in each slot
 * (particle_number values). Derives the tiling first, like ljgpu_get_tiles;
 * *order = -1 when the kernels run untiled. Inspection only. */
int ljgpu_get_tile_order(ljgpu_ctx* ctx, int* order, int32_t* member, int64_t capacity);

/* Autotuner. For each candidate configuration, one untimed call of `reps`
 * force evaluations (the production call shape: layouts, slot copies and
 * graphs in place), then 5 timed calls of `reps` evaluations; the median
 * is kept. Candidates: the tiled kernel with 4, 8 or 16 lanes per particle,
 * unroll 2 or 4, tiles of 16, 32 or 64 particles and 1 or 2 launch chains,
 * one persistent grid, the untiled wavefront-segment kernel and the
 * thread-per-particle ELL kernel, over device AoS positions q.
 * The caller's momenta are not touched (a scratch buffer is used). The
 * fastest configuration is installed and returned in *best; the whole table
 * stays available through ljgpu_query_autotune. No reference counterpart (the
 * reference takes only its block size from argv, cuda/force_cuda.cu:377-383). */
typedef struct ljgpu_tune_entry {
  int32_t lanes_per_particle;
  int32_t unroll;
  int32_t tile_particles;   /* requested tile size (0: untiled kernels)   */
  int32_t resident_per_cu;  /* persistent grid (0: one workgroup per tile) */
  int32_t tile_used;        /* tile size the kernel ran with               */
  float us_per_step;        /* GPU time per force evaluation (median call)  */
  int32_t chains;           /* concurrent launch chains (ljgpu_set_chains)  */
  float us_min;             /* the fastest timed call                       */
} ljgpu_tune_entry;

int ljgpu_autotune(ljgpu_ctx* ctx, const double* q, double box, double cl2, int precision,
                   int accum, int32_t reps, ljgpu_tune_entry* best);
/* The table of the last ljgpu_autotune call: *count entries, up to
 * `capacity` copied (pass null results to query the count). */
int ljgpu_query_autotune(ljgpu_ctx* ctx, ljgpu_tune_entry* results, int32_t capacity,
                         int32_t* count);

/* ---- forces ------------------------------------------------------------ */
/* p += F(q) * dt over the installed list, repeated `steps` times (the
 * reference's measure() loop of 100 launches, cuda/force_cuda.cu:319-342;
 * cpu_ref/force_aos.cpp:592-601). q and p are device pointers. FP64
 * throughout: the pair term of cpu_ref/force_aos.cpp:176-189 and
 * cuda/kernel.cuh:22-31. Asynchronous: call ljgpu_sync before reading p. */
int ljgpu_force_aos_f64(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                        double box, int32_t steps);
int ljgpu_force_soa_f64(ljgpu_ctx* ctx, const double* qx, const double* qy, const double* qz,
                        double* px, double* py, double* pz, double dt, double cl2,
                        double box, int32_t steps);

/* The same over packed double[N][3] arrays (x, y, z; 24 B per particle,
 * 8-byte aligned): the reference's double3 instantiation of every kernel
 * (cuda/force_cuda.cu:26-27 buffers, :354-375 MEASURE_FOR_ALLTYPES). Same
 * arithmetic and results as the double4 calls. A call of 16 or more steps
 * over spatially ordered tiles runs on internal 4-double copies in slot
 * order, like the double4 calls, and writes p back in the caller's rows. */
int ljgpu_force_aos3_f64(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                         double box, int32_t steps);

/* Mixed precision: pair distances and r2 in FP64, r2 rounded once to FP32,
 * r^-6 / r^-14 and the force scalar in FP32, accumulation in FP64 or
 * Kahan-compensated FP32 (accum = LJGPU_ACCUM_*), momentum update in FP64.
 * No reference counterpart (the reference allocates float3/float4 buffers,
 * cuda/force_cuda.cu:24-25, but never evaluates forces with them). */
int ljgpu_force_aos_mixed(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                          double box, int32_t steps, int accum);
int ljgpu_force_soa_mixed(ljgpu_ctx* ctx, const double* qx, const double* qy,
                          const double* qz, double* px, double* py, double* pz, double dt,
                          double cl2, double box, int32_t steps, int accum);
int ljgpu_force_aos3_mixed(ljgpu_ctx* ctx, const double* q, double* p, double dt, double cl2,
                           double box, int32_t steps, int accum);

/* One launch with exactly the reference kernel's argument list
 * (force_kernel_*(q, p, particle_number, dt, CL2, sorted_list,
 * number_of_partners, pointer), cuda/kernel.cuh:5-13), all device pointers,
 * AoS double4, open box. kind selects full-list (p[i] only) or half-list
 * (atomic scatter) semantics. No list is installed or validated: every index
 * read is clamped into range, so a malformed list gives wrong numbers, never
 * an out-of-bounds access. */
int ljgpu_force_kernel_csr_f64(ljgpu_ctx* ctx, const double* q, double* p,
                               int32_t particle_number, double dt, double cl2,
                               const int32_t* sorted_list, const int32_t* number_of_partners,
                               const int32_t* pointer, int64_t n_entries, int kind);

/* ---- molecular dynamics around the force path ----------------------------- */
/* No reference counterpart: the reference has no integrator and no energy,
 * it repeats force evaluations on fixed positions (measure()'s loop,
 * cuda/force_cuda.cu:331-335). These entry points wrap that same evaluation
 * (p += F(q) * dt, unit masses, so p is the velocity) in a velocity-Verlet
 * loop with a periodic list rebuild on the device. */
typedef struct ljgpu_md_config {
  double dt;              /* time step                                          */
  double cutoff;          /* force cutoff: pairs with r2 > cutoff^2 are skipped  */
  double skin;            /* Verlet skin: lists hold pairs within cutoff + skin  */
  int32_t rebuild_every;  /* rebuild the list every this many steps; 0: whenever
                             a particle has moved more than skin/2 since the
                             last build (checked every step)                   */
  int32_t precision;      /* LJGPU_PREC_F64 or LJGPU_PREC_MIXED (forces; the
                             integration is FP64 either way)                    */
  int32_t accum;          /* LJGPU_ACCUM_* for mixed precision                   */
  int32_t sample_every;   /* energy sample interval (0: first and last step)     */
} ljgpu_md_config;

typedef struct ljgpu_md_stats {
  int32_t steps;
  int32_t rebuilds;        /* list builds, including the initial one           */
  int32_t samples;         /* energy samples taken                              */
  int32_t stale_lists;     /* rebuilds at which some particle had moved more
                              than skin/2 since the previous build (the list may
                              have missed pairs: rebuild more often)           */
  double potential0, kinetic0;  /* energies at the first sample (step 0); the
                                   potential is shifted: each pair inside the
                                   cutoff counts U(r) - U(cutoff), so that
                                   U + K is the energy truncated forces
                                   conserve                                   */
  double potential1, kinetic1;  /* energies at the last sample (final step)    */
  double max_abs_drift;    /* max |E(t) - E(0)| over the samples, E = U + K    */
  double max_displacement; /* largest displacement found at a list check        */
  int64_t pair_entries;    /* list entries evaluated by force kicks            */
  float ms_total;          /* GPU time of the run                              */
  float ms_rebuild;        /* part of it spent building lists                  */
} ljgpu_md_stats;

/* Run `steps` velocity-Verlet steps in place on device AoS q and p (N x 4,
 * 16-byte aligned): positions wrapped into [0, box) for box > 0, the full
 * list rebuilt on the device (ljgpu_build_list, search length cutoff + skin)
 * at the start and every rebuild_every steps, potential and kinetic energy
 * sampled at step 0, every sample_every steps and at the end. `samples`
 * (optional, host) receives up to sample_capacity rows (step, U, K, U + K)
 * with U the shifted potential (see ljgpu_md_stats).
 * In a periodic box the run works on copies of q and p in a spatial order of
 * the particles, re-sorted at every rebuild, so that index-order tiles over
 * them are compact. It writes them back to q and p in the caller's order on
 * return. Its lists index that order, so none stays installed. With
 * box <= 0 the context's list is replaced by the run's last list. */
int ljgpu_md_run(ljgpu_ctx* ctx, double* q, double* p, int32_t particle_number, double box,
                 const ljgpu_md_config* config, int32_t steps, ljgpu_md_stats* stats,
                 double* samples, int32_t sample_capacity);

/* Potential energy sum of 4 (r^-12 - r^-6) over the installed list's pairs
 * with r2 <= cl2 (the force kernels' cutoff rule, no shift), kinetic energy
 * sum of |p|^2 / 2, device AoS q and p, and (optional) the number of pairs
 * inside the cutoff. Over an installed LDS tiling (the force kernels' layout,
 * used as it stands) the potential runs on the tiles, otherwise over the
 * list's rows. Synchronous. */
int ljgpu_energy(ljgpu_ctx* ctx, const double* q, const double* p, double box, double cl2,
                 double* potential, double* kinetic, int64_t* pairs_in_cutoff);

/* ljgpu_energy without any device allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, checks and results as ljgpu_energy (the same doubles: the
 * partials are added in the same order). Its partial sums live in a static
 * per-GPU array of 3 * 1024 + 3 doubles (24.6 KB, held by every GPU the
 * library runs on until the process ends) instead of context memory. The
 * grids are capped at 1,024 workgroups that stride over the rows or tiles,
 * so no particle count or list length needs more scratch than that. All calls on
 * one GPU share the array: calls that overlap on one GPU (from several
 * threads, or contexts on different streams) race and must be serialised by
 * the caller. The list must already be installed (installing one
 * allocates). Synchronous. */
int ljgpu_energy_noalloc(ljgpu_ctx* ctx, const double* q, const double* p, double box,
                         double cl2, double* potential, double* kinetic,
                         int64_t* pairs_in_cutoff);

/* Maxwell-Boltzmann momenta at `temperature` (unit masses): each component
 * N(0, temperature) from std::mt19937(seed) in particle order x, y, z, then
 * the mean momentum removed. Host AoS p[N][4] (pad set to 0). */
int ljgpu_maxwell_boltzmann(int32_t particle_number, double temperature, uint32_t seed,
                            double* p);

/* Run repeated force evaluations (steps >= 4) and the MD driver's plain steps
 * as replays of captured HIP graphs (1, the default) or as individual
 * launches (0). A force call of `steps` launches replays graphs of 64, 16 and
 * 4 launches (longest first) and launches the remaining 0-3 singly; each
 * graph is captured once per argument set and kept in a cache of 12 (least
 * recently used evicted). */
int ljgpu_set_graphs(ljgpu_ctx* ctx, int enable);

/* Concurrent chains of a many-step call over LDS tiles (1 to 4, default 2).
 * With c > 1, a call of 4 or more steps over at least 64 * c tiles splits
 * the tiles into c contiguous parts and runs each part's steps as its own
 * chain of launches on its own stream of the context, forked from and joined
 * back into the context stream within the call (launched directly, not
 * replayed from graphs). The parts update disjoint particles, so the momenta
 * are bitwise those of one chain; one chain's launch ramp and tail overlap
 * the others' work. No counterpart in the reference (which launches on the
 * default stream). */
int ljgpu_set_chains(ljgpu_ctx* ctx, int chains);

/* Graph cache state: graphs captured over the context's life and graphs
 * cached now. No counterpart in the reference (which has no graphs); lets a
 * caller check that repeated calls replay instead of re-capturing. */
int ljgpu_graph_stats(ljgpu_ctx* ctx, int64_t* captures, int32_t* cached);

/* ---- host-side setup (CPU only, no device needed) ----------------------- */
/* The reference's init(): FCC lattice in an open box of side box_length,
 * jitter U(0, 0.1) from std::mt19937(2), AoS double[N][4].
 * cpu_ref/force_aos.cpp:43-51,142-169; cuda/force_cuda.cu:47-94. */
int ljgpu_lattice_reference(double density, double box_length, double* q, int32_t capacity,
                            int32_t* particle_number);

/* Periodic FCC lattice of 4*cells^3 particles, same lattice constant and
 * particle order as init(), side L = cells * (4/density)^(1/3), jitter
 * U(jitter_lo, jitter_hi) from std::mt19937(seed), wrapped into [0, L).
 * No reference counterpart (the reference has no periodic box). */
int ljgpu_lattice_periodic(int32_t cells, double density, uint32_t seed, double jitter_lo,
                           double jitter_hi, double* q, int32_t capacity, double* box);

/* Verlet list with the reference's exact pair set and order: predicate
 * r2 < search_length^2 with r2 = (dx*dx + dy*dy) + dz*dz unfused, rows in
 * ascending i, partners ascending (makepair + sortpair,
 * cpu_ref/force_aos.cpp:100-140; full list cuda/force_cuda.cu:122-163).
 * Built with a cell list in O(N) instead of the reference's O(N^2) scan.
 * Writes at most `capacity` entries; *n_entries receives the true count
 * (LJGPU_ERR_CAPACITY if it exceeds capacity). */
int ljgpu_makepair_host(int32_t particle_number, const double* q, double box,
                        double search_length, int kind, int32_t* number_of_partners,
                        int32_t* pointer, int32_t* sorted_list, int64_t capacity,
                        int64_t* n_entries);

/* check_loadedpair() (cuda/force_cuda.cu:183-201) made strict: row counts,
 * row starts and ends inside the list, 0 <= j < N, j != i, and j > i for a
 * half list. */
int ljgpu_validate_csr(int32_t particle_number, const int32_t* number_of_partners,
                       const int32_t* pointer, const int32_t* sorted_list, int64_t n_entries,
                       int kind);

/* random_shfl(): shuffle every row with one std::mt19937(seed) stream (seed
 * 10 in the reference), cuda/force_cuda.cu:255-263. */
int ljgpu_random_shfl(int32_t particle_number, const int32_t* number_of_partners,
                      const int32_t* pointer, int32_t* sorted_list, uint32_t seed);

/* make_transposed_pairlist() (cuda/force_cuda.cu:229-240): ELL with stride N,
 * zero padded to width (>= the longest row). */
int ljgpu_make_transposed(int32_t particle_number, const int32_t* number_of_partners,
                          const int32_t* pointer, const int32_t* sorted_list, int32_t width,
                          int32_t* transposed_list);

/* The LDS tiling ljgpu_set_tiling installs, built on the host from a CSR
 * list whose rows are stored in order (pointer[i+1] = pointer[i] +
 * number_of_partners[i], as sortpair produces): the tiled counterpart of
 * make_transposed_pairlist, cuda/force_cuda.cu:229-240. desc receives 8 ints
 * per block (first particle, particle count, staging offset, staging count,
 * first list entry, entry count, replica (0 here), middle particle = first
 * particle + count / 2), stage the staged particle indices,
 * local (n_entries values, row by row as in sorted_list) the LDS byte
 * offset 24*rank of each entry's j within its block's staging area; max_stage
 * the largest staging count. Within a row of at most 256 entries the offsets
 * are permuted so that the partner reads of one LDS instruction spread over
 * the banks: entry k of the row at position r of its block takes a partner
 * of rank class (4*(r & 7) + k) mod 32, or of the next class up that the row
 * still holds, entries of one class keeping their list order (longer rows
 * keep the list order). Pass null outputs to query the sizes first. */
int ljgpu_make_tiles(int32_t particle_number, const int32_t* number_of_partners,
                     const int32_t* pointer, const int32_t* sorted_list, int64_t n_entries,
                     int32_t tile_particles, int32_t* desc, int64_t desc_capacity,
                     int32_t* stage, int64_t stage_capacity, uint16_t* local,
                     int32_t* n_blocks, int64_t* n_stage, int32_t* max_stage);

/* As ljgpu_make_tiles, with the particles taken in the order `member` (a
 * permutation of 0..particle_number-1, one particle per slot; blocks cover
 * consecutive slots): desc's first particle and first entry become the
 * block's first slot and its first entry in the rows stored in slot order,
 * each block's staged particles are sorted by slot instead of index, and
 * local follows the rows in slot order. member = null is ljgpu_make_tiles. No
 * reference counterpart: the host model of ljgpu_get_tile_order's tiling. */
int ljgpu_make_tiles_ordered(int32_t particle_number, const int32_t* number_of_partners,
                             const int32_t* pointer, const int32_t* sorted_list,
                             int64_t n_entries, int32_t tile_particles, const int32_t* member,
                             int32_t* desc, int64_t desc_capacity, int32_t* stage,
                             int64_t stage_capacity, uint16_t* local, int32_t* n_blocks,
                             int64_t* n_stage, int32_t* max_stage);

/* As ljgpu_make_tiles_ordered, with positions q (AoS double[N][4]; box > 0
 * periodic, else open) that put each row of at most 256 entries in cutoff
 * classes before the bank-class order: first the entries whose pair lies
 * within the cutoff, r2 = (dx*dx + dy*dy) + dz*dz <= cl2 over the threshold
 * minimum image of d = q_j - q_i, then the others, each part in list order
 * and then bank-class ordered with its classes continuing from the part's
 * first entry. The tiles a force evaluation derives from its positions take
 * these classes; the force kernels then skip the force arithmetic of batches
 * whose entries are all beyond the cutoff at that step. q = null is
 * ljgpu_make_tiles_ordered. No reference counterpart: the host model of the
 * device tiling. */
int ljgpu_make_tiles_cut(int32_t particle_number, const int32_t* number_of_partners,
                         const int32_t* pointer, const int32_t* sorted_list, int64_t n_entries,
                         int32_t tile_particles, const int32_t* member, const double* q,
                         double box, double cl2, int32_t* desc, int64_t desc_capacity,
                         int32_t* stage, int64_t stage_capacity, uint16_t* local,
                         int32_t* n_blocks, int64_t* n_stage, int32_t* max_stage);

/* ---- pair-list caches (host files, no device needed) ---------------------- */
/* pair.dat, the binary cache of the CPU programs (savepair / loadpair,
 * cpu_ref/force_aos.cpp:603-620; cpu_ref/force_soa.cpp:360-376): int32
 * number_of_pairs, int32 number_of_partners[n_capacity], int32
 * i_particles[pair_capacity], int32 j_particles[pair_capacity], unused tails
 * zero. n_capacity / pair_capacity are the reference's compile-time array
 * sizes; 0 selects its values (400000 and 30 * n_capacity). The writer takes
 * a half list and writes its pairs in CSR storage order (for a list from
 * makepair + sortpair: makepair's order), byte-identical to savepair(). The
 * reader rebuilds the CSR as loadpair() + sortpair() do; pass a null
 * sorted_list to query *n_entries first. */
int ljgpu_pair_dat_write(const char* path, int32_t particle_number,
                         const int32_t* number_of_partners, const int32_t* pointer,
                         const int32_t* sorted_list, int64_t n_entries, int32_t n_capacity,
                         int64_t pair_capacity);
int ljgpu_pair_dat_read(const char* path, int32_t particle_number, int32_t n_capacity,
                        int64_t pair_capacity, int32_t* number_of_partners, int32_t* pointer,
                        int32_t* sorted_list, int64_t capacity, int64_t* n_entries);

/* .cache_pair_half.dat / .cache_pair_all.dat, the GPU harness's text cache
 * (makepaircache / loadpair, cuda/force_cuda.cu:165-227): "N E", then N lines
 * "number_of_partners pointer", then E lines "sorted_list". The reader refuses
 * a cache for another particle number, as the reference does, and validates
 * the list (check_loadedpair, strict form) as `kind`. Pass a null sorted_list
 * to query *n_entries (number_of_partners and pointer are filled). */
int ljgpu_pair_cache_write(const char* path, int32_t particle_number,
                           const int32_t* number_of_partners, const int32_t* pointer,
                           const int32_t* sorted_list, int64_t n_entries);
int ljgpu_pair_cache_read(const char* path, int32_t particle_number, int kind,
                          int32_t* number_of_partners, int32_t* pointer, int32_t* sorted_list,
                          int64_t capacity, int64_t* n_entries);

#ifdef __cplusplus
}
#endif

#endif /* LJGPU_H */
