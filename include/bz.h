/*
 * bz.h -- C ABI of libbz.so, the MI355X (gfx950) electron kinetic step.
 *
 * Every entry point takes plain device pointers and sizes, returns 0 on
 * success and a negative code on failure (text via bz_last_error()), and
 * enqueues its work on the stream held by the context.  No entry point
 * allocates device memory except bz_ctx_create / bz_ctx_destroy.
 *
 * Particle state is structure-of-arrays: x (cm), vx, vy, vz (m/s) and the
 * weight w, each a contiguous array of the context precision (float for
 * BZ_F32, double for BZ_F64).  The reference keeps AoS rows
 * data[:,0:6] = (x,y,z,vx,vy,vz) and tosum[:,0:4] = (w, ion flag, 0, energy)
 * (1bin/boltzmann.py:37-44); y and z are never read by any stage and never
 * written to disk there (kernels/data_write_1D.py:71-75), so they are not
 * stored, and the energy is recomputed from the velocity where it is needed.
 *
 * The live particle count, the step counter (which keys the random stream)
 * and the simulated time live in device memory inside the context (the
 * "clock"), so a whole step can be enqueued or captured in a hipGraph with
 * no host synchronisation.  bz_clock_set / bz_clock_get move them to and
 * from the host.
 */
#ifndef BZ_H_
#define BZ_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bz_ctx* bz_ctx_t;

enum { BZ_F32 = 4, BZ_F64 = 8 };
/* collision model */
enum {
  BZ_XS_FIT = 0,   /* reference argon fits, direct P = 1-exp(-nu dt)
                      (kernels/electron_kernel_numba.py:66-195) */
  BZ_XS_TABLE = 1, /* tabulated sigma(E), 3 channels, null collision */
  BZ_XS_FLIGHT = 2, /* tabulated sigma(E), exact free-flight null collision:
                       any number of collisions per step (Philox stream only) */
  BZ_XS_DIRECT = 3  /* tabulated sigma(E), the reference's direct collision
                       probability P = 1 - exp(-nu_t(E) dt) per particle
                       (electron_kernel_numba.py:186-212): accurate while the
                       particle's own nu_t dt is small, whatever nu_max dt is */
};
/* random stream */
enum {
  BZ_RNG_PHILOX = 0, /* counter-based Philox4x32-10 keyed by (seed, step, index) */
  BZ_RNG_STREAM = 1  /* caller-supplied uniforms R[7][stride], gas_v[3][stride]
                        (the reference's cp_R_vec / gas_v arrays) */
};

/* order of the two updates of the push */
enum {
  BZ_PUSH_REF = 0,     /* x += 100 dt v, then v -= dt q/m E (electron_kernel_numba.py:257-264):
                          both use the state at the start of the step (explicit Euler) */
  BZ_PUSH_LEAPFROG = 1 /* v -= dt q/m E(x), then x += 100 dt v with the new v
                          (kick-drift leapfrog, v staggered by half a step) */
};

/* counters accumulated by the kernels (uint64 each) */
enum {
  BZ_CNT_BIRTHS = 0,
  BZ_CNT_ABSORBED = 1,
  BZ_CNT_ELASTIC = 2,
  BZ_CNT_EXCITATION = 3,
  BZ_CNT_IONIZATION = 4,
  BZ_CNT_NULL = 5,
  BZ_CNT_OVERFLOW = 6,  /* particles dropped because capacity was reached */
  BZ_CNT_NUMAX = 7,     /* null-collision bound exceeded (table mode) */
  BZ_CNT_SPLIT = 8,     /* resampling: segments split */
  BZ_CNT_MERGE = 9,     /* resampling: segments merged */
  BZ_CNT_RESAMPLE = 10, /* resampling passes that fired */
  BZ_CNT_PSTEPS = 11,   /* particle-steps pushed (sum of input counts) */
  BZ_CNT_COPIED = 12,   /* resampling: particles copied unchanged (copy segments) */
  BZ_CNT_SORTED = 13,   /* slots ordered by the binning sorts (sum of sort sizes) */
  BZ_CNT_DEP_RANGE = 14, /* deposit blocks whose fixed-point node sums neared their range */
  BZ_NCOUNTERS = 16
};

typedef struct {
  void* x;
  void* vx;
  void* vy;
  void* vz;
  void* w;
} bz_state_t;

/* Per-run constants of the push / collision step.  Derived constants are
 * computed once on the host so that the device and the CPU check use the
 * same values. */
typedef struct {
  double dt;        /* s */
  double xfac;      /* 100*dt: cm moved per (m/s)          (electron_kernel_numba.py:258) */
  double afac;      /* dt*q_e/m_e                           (electron_kernel_numba.py:264) */
  double kev;       /* 0.5*m_e*j_ev: eV per (m/s)^2         (electron_kernel_numba.py:63) */
  double jm;        /* j_ev*m_e                             (electron_kernel_numba.py:226) */
  double fel;       /* sqrt(1 - 2 m_e/m_n), elastic speed factor (electron_kernel_numba.py:206) */
  double L;         /* domain length, cm */
  double dx;        /* cell size, cm */
  double half_dx;   /* 0.5*dx */
  double L_dx;      /* L+dx                                 (post_processing_1D.py:44) */
  double dep_M;     /* bnf*n_nodes, fine bins of the deposit (boltzmann.py:67) */
  double nn;        /* fit mode: gas density multiplying sigma, m^-3 */
  double e_ion;     /* ionization threshold, eV */
  double e_exc;     /* excitation threshold, eV (table mode) */
  double p_null;    /* table mode: 1-exp(-nu_max dt) */
  double pc;        /* table mode: p_null/nu_max */
  double nu_max;    /* table mode: null-collision frequency bound, 1/s */
  double gas_sigma; /* std of each gas-velocity component, m/s (0: cold gas) */
  double e_field0;  /* 0-D mode: uniform E, V/cm (used when E pointer is NULL) */
  uint64_t seed;
  int32_t dim;      /* 0: 0-D (no advection, no walls), 1: 1-D bounded */
  int32_t xs_mode;  /* BZ_XS_* */
  int32_t rng_mode; /* BZ_RNG_* */
  int32_t n_cells;  /* N: length of E */
  int32_t n_nodes;  /* N+1: deposit nodes */
  int32_t bnf;      /* fine bins per node of the deposit */
  int64_t r_stride; /* BZ_RNG_STREAM: column stride of R and gas_v */
  int32_t push_order; /* BZ_PUSH_*: order of the position and velocity updates */
  int32_t ref_flags;  /* BZ_REF_*: reference behaviours reproduced on request */
} bz_params_t;

/* Reference behaviours that are off by default (bz_params_t.ref_flags).
 * BZ_REF_DEAD_IONIZER: a primary that ionizes and leaves the domain in the
 *   same step adds 1 (its unscaled ionization flag) to the ionization moment
 *   of node 0: the reference zeroes its weight and parks it at x = 1e-8
 *   (kernels/electron_kernel_numba.py:231,267-270), scales the flag by the
 *   weight only where w > 0 and reduces every row
 *   (kernels/post_processing_1D.py:48-55).
 * BZ_REF_ENERGY_ROUNDTRIP: binning uses the energy after (e*w)/w, the value
 *   the reference's tosum[:,3] holds when it bins (post_processing_1D.py:53,62),
 *   in the state precision. */
enum { BZ_REF_DEAD_IONIZER = 1, BZ_REF_ENERGY_ROUNDTRIP = 2 };

/* Tabulated sigma(E)*N_gas, three channels (0 elastic, 1 excitation,
 * 2 ionization), piecewise linear on n nodes.  Device pointers in the
 * context precision.  hint[] maps the top bits of the IEEE energy to the
 * lowest candidate segment (built by the host). */
typedef struct {
  const void* energy;   /* [n] eV, increasing */
  const void* sig;      /* [3][n] sigma*N, 1/m */
  const void* slope;    /* [3][n] d(sigma*N)/dE per segment */
  const int32_t* hint;  /* [n_hint] */
  int32_t n;
  int32_t n_hint;
  int32_t hint_base;    /* bucket number of hint[0] */
  int32_t hint_bits;    /* mantissa bits used by the bucket */
} bz_xsec_t;

/* Field / fluid constants (FP64), 1-D mode. */
typedef struct {
  double dx;        /* cm */
  double dt;        /* s, fluid step (= kinetic step) */
  double eps;       /* permittivity in e/(V cm) (boltzmann.py:56) */
  double v0;        /* RF amplitude, V */
  double freq;      /* Hz */
  double mu_i, D_i, D_s; /* ion mobility / diffusion, metastable diffusion (boltzmann.py:50-55) */
  int32_t n;        /* N cells (N+1 nodes, N-1 interior unknowns) */
  int32_t fluid_on; /* 0: heavies frozen (0-D test problems) */
  int32_t poisson_on;
  int32_t skip_moments; /* 1: leave ni / Te / g2 untouched (stand-alone solves) */
  int32_t ion_scheme;   /* 0: the reference's explicit central fluxes, one step
                           (heavies_kernel_1D.py:25-61); 1: Scharfetter-Gummel
                           fluxes, sub-cycled to the explicit stability bound,
                           with the zero-inflow wall condition of the reference
                           restated for those fluxes */
  int32_t max_substeps; /* scheme 1: cap on the sub-steps of one field step */
} bz_field_params_t;

/* Resampling policy (ES-WPNC merge / split). */
typedef struct {
  double split_below;   /* split when 0 < count < split_below*target (0.5) */
  double merge_at;      /* merge when count >= merge_at*target (3.0) */
  double merge_keep;    /* merged segment keeps int(merge_keep*target) (0.6) */
  double trigger_at;    /* pass fires when some segment count >= trigger_at*target */
  double min_wt;        /* split only particles with w > min_wt (0.1) */
  double split_to;      /* split draws int(split_to*target) - count clones (2) */
  int32_t ref_trigger;  /* 1: reference trigger (last bin's max only, quirk) */
  int32_t max_ppc;      /* >0: merge uses at most max_ppc particles (reference clip) */
  int32_t ref_rescale;  /* 1: reference merge rescale sqrt(T/T_hat) from total energy;
                           0: central moments (conserves energy with drift) */
  int32_t pad;
  /* The resampling draws: NULL, the Philox stream; otherwise a device array
   * of FP64 uniforms in [0,1), U[key*u_stride + m] being draw m of segment
   * `key` (a merge's U_m of (U_m + m)/K, a split's m-th categorical draw), the
   * numbers the reference takes from cp.random.rand / cp.random.choice
   * (1bin/boltzmann.py:814,847).  u_stride must exceed every segment's draws. */
  const double* uniforms;
  int64_t u_stride;
} bz_resample_policy_t;

const char* bz_last_error(void);
int bz_version(void);

/* Context: one device, one stream, workspace sized for `capacity` particles
 * per state buffer and up to `max_nodes` deposit nodes / `max_keys` sort keys. */
int bz_ctx_create(bz_ctx_t* out, int device, int precision, int64_t capacity,
                  int32_t max_nodes, int32_t max_keys);
int bz_ctx_destroy(bz_ctx_t ctx);
int bz_set_stream(bz_ctx_t ctx, void* hip_stream);
/* Tile shape of the fused push (0..4: threads x particles per thread
 * 256x4, 512x2, 256x2, 1024x1, 1024x2), chosen by the autotuner; no
 * counterpart in the reference (its launch shape is fixed at 1024 x 64,
 * 1bin/boltzmann.py:139-140). */
int bz_set_push_shape(bz_ctx_t ctx, int shape);
/* Grid of the 1-D push: at most `blocks_per_cu` blocks per CU (1..8), or
 * as many as fit (0, default).  Fewer blocks spread the per-block set-up
 * (table staging, histogram flush) over more particles; chosen by the
 * autotuner.  No counterpart in the reference. */
int bz_set_push_grid(bz_ctx_t ctx, int blocks_per_cu);
/* Candidate queue of the 1-D push (tabulated sigma, Philox stream, cold
 * gas): a particle whose uniform is at or above the collision-probability
 * bound of its energy bucket is pushed at once; the others are queued per
 * wave and collided 64 at a time.  The results are the same either way.
 * 0 (default): when a wave walks at least 16 tile items (a kernel variant
 * with the queue is launched when the capacity allows that many), 1: always,
 * -1: never.  No counterpart in the reference. */
int bz_set_push_queue(bz_ctx_t ctx, int mode);
/* Binning keys written by the 1-D push: with nb in [1,8] (nb-1 thresholds,
 * host pointer) every 1-D bz_push_mcc also writes the bz_bin_sort key of
 * each slot it leaves (holes and births included), and the next bz_bin_sort
 * of that state with the same nb, thresholds, dx and energy factor sorts
 * those keys instead of recomputing them from the state.  Any other call
 * that changes a state or the sort workspace (clock set, compact,
 * resample, swarm calls, graph capture or replay) drops them.  A caller
 * that writes the state arrays itself between the push and the sort calls
 * this function again first.  nb = 0 turns it off (default).  No
 * counterpart in the reference. */
int bz_set_push_keys(bz_ctx_t ctx, int32_t nb, const double* thresholds);
/* Width of the keys the 1-D push writes (bz_set_push_keys): on = 1 writes
 * them as 16-bit words, half the bytes the push writes and the in-place
 * resampling (bz_resample_inplace) reads three times (key histogram, member
 * count, member extraction); the next bz_resample_inplace of that state
 * reads them as such.  bz_bin_sort, which sorts 32-bit keys, recomputes its
 * keys from the state after a push that wrote 16-bit ones.  Keys above 65535
 * (n_nodes * nb >= 65535) fail the push.  on = 0 (default): 32-bit keys.
 * No counterpart in the reference. */
int bz_set_push_keys16(bz_ctx_t ctx, int32_t on);
/* Tile schedule of the fused push: 0 static tile order under a cooperative
 * launch (default), 1 tiles claimed from a ticket counter by running blocks.
 * No counterpart in the reference. */
int bz_set_push_schedule(bz_ctx_t ctx, int sched);
int bz_synchronize(bz_ctx_t ctx);

/* Device clock: live count, step, time; counters (BZ_NCOUNTERS uint64). */
int bz_clock_set(bz_ctx_t ctx, int64_t n, uint64_t step, double t);
int bz_clock_get(bz_ctx_t ctx, int64_t* n, uint64_t* step, double* t, uint64_t* counters);
int bz_counters_reset(bz_ctx_t ctx);
/* step += 1, t += dt, on the stream */
int bz_clock_advance(bz_ctx_t ctx, double dt);

/* Kernel timing with HIP events recorded on the context stream around each
 * launch of the fused push kernel (k_push_inplace, or k_push for a single
 * 0-D swarm) and of the binning / resampling kernels.
 * bz_timing_enable(ctx, 1) clears and starts, 0 stops; bz_timing_get
 * synchronises and returns the summed milliseconds and launch counts of
 * [0] push, [1] sort, [2] resample.  No counterpart in the reference. */
int bz_timing_enable(bz_ctx_t ctx, int on);
int bz_timing_get(bz_ctx_t ctx, double* ms3, int64_t* launches3);
/* The same for the first n (<= 8) timed kinds: [0] push, [1] sort,
 * [2] resample, [3] the resampling copy kernel alone (k_res_copy), [4] each
 * radix downsweep pass (k_downsweep), and single kernels of the in-place
 * resampling: [5] its merges and splits (k_res_segments), [6] its key
 * histogram (k_ip_hist), [7] the moves of the hole filling (k_fill_move).
 * No counterpart in the reference. */
int bz_timing_get_n(bz_ctx_t ctx, int n, double* ms, int64_t* launches);

/* Fused electron step: MCC collision (elastic / excitation / ionization with
 * a secondary), push in E(x), absorbing walls and node deposit of
 * (sum w, sum w*ionized, sum w*energy).  The slot count is read from and
 * written to the clock.
 *   1-D (dim 1), in place, as the reference's electron_kernel leaves its
 *     arrays: `in` slots [0, n) are updated where they lie, an absorbed
 *     particle's slot keeps w = 0 (a hole, skipped by every stage), and the
 *     secondaries are appended at [n, n + births) in parent order.  `out` is
 *     scratch (births are staged there) and its contents are clobbered.
 *     bz_compact removes the holes (the reference's post_processing
 *     reshuffle, kernels/post_processing_1D.py:26-38).
 *   0-D (dim 0), in -> out: survivors in input order, each secondary
 *     directly after its parent.
 *   E          : [n_cells] field, V/cm, context precision (NULL in 0-D: e_field0)
 *   moments    : [3][n_nodes] FP64 output (overwritten)
 * Replaces electron_kernel (kernels/electron_kernel_numba.py:291-303) plus the
 * deposit of post_processing (kernels/post_processing_1D.py:40-68,
 * kernels/reduction/cuda_reduction/kernels-new.cu:56-132). */
int bz_push_mcc(bz_ctx_t ctx, const bz_state_t* in, const bz_state_t* out,
                const void* E, const bz_params_t* p, const bz_xsec_t* xs,
                const void* R, const void* gas_v, double* moments);

/* Stable compaction: the particles of `in` with w > 0, in slot order, to
 * `out`; the clock's count becomes the live count.  Replaces
 * reshuffle_data_atomics (kernels/post_processing_1D.py:26-38). */
int bz_compact(bz_ctx_t ctx, const bz_state_t* in, const bz_state_t* out);

/* Compaction in place by hole filling: with L the live count (w > 0), the
 * k-th slot below L that holds no particle (in slot order) receives the k-th
 * particle at or above L (in slot order); live particles below L keep their
 * slots and the count becomes L.  Also replaces reshuffle_data_atomics
 * (kernels/post_processing_1D.py:26-38), whose atomics leave the order of
 * the compacted particles open; this moves only the particles above L. */
int bz_compact_inplace(bz_ctx_t ctx, const bz_state_t* s);

/* bin_particles on the reference's own arrays (1bin/kernels/
 * bin_particles.py:7-31, 3bin :7-78, 5bin :7-140): tosum rows of
 * tosum_cols doubles (w in column 0, the energy in eV in column 3, as
 * post_processing leaves it), curr_xbins int64 cells, rows 0..nc-1.  A row
 * with w > 0 and 0 <= cell <= n_cells is listed under its cell and energy
 * bin, the number of thresholds_host[0..nb-2] at or below its energy (a NaN
 * energy is listed nowhere when nb > 1, as the reference's `<` chain).
 * lists[b] is the bin's int64 (max_ppc[b], n_cells+1) array, filled with
 * row indices in ascending order up to max_ppc[b] per cell; write_inds[b]
 * (int64, n_cells+1) receives the full counts.  Bins run from the lowest
 * energy up (the reference names them LE, ME..., HE).  lists, write_inds
 * and max_ppc are host arrays of nb entries; the arrays they point to are
 * device memory.  Resets the context's slot count to nc. */
int bz_bin_lists(bz_ctx_t ctx, const double* tosum, int64_t tosum_cols,
                 const int64_t* curr_xbins, int64_t nc, int32_t n_cells, int32_t nb,
                 const double* thresholds_host, int64_t* const* lists,
                 int64_t* const* write_inds, const int64_t* max_ppc);

/* Collision counters of the 1-D push, the reference's collct deposit
 * (kernels/electron_kernel_numba.py:199,232 set collct[i,0] = w for an
 * elastic collision and collct[i,2] = 1 for an ionization;
 * kernels/post_processing_1D.py:56 reduces them to counter_g0..g3 at the
 * particles' post-push positions, an absorbed particle at x = 1e-8).  With
 * the counters on, every 1-D push adds per node: [0] sum w of the elastic
 * collisions, [1] sum w of the excitations (no column in the reference's
 * 1bin kernel), [2] the number of ionizations; at the primary's node after
 * the push, node 0 when it was absorbed.  The reference never resets its
 * collct rows and never writes the counters (quirk Q10); here they cover
 * the pushes since the last read with reset.  bz_collision_counts copies
 * them to out[3][n_nodes] (device memory) and clears them when reset != 0. */
int bz_set_collision_counts(bz_ctx_t ctx, int32_t on);
int bz_collision_counts(bz_ctx_t ctx, double* out, int32_t n_nodes, int32_t reset);

/* The collision uniforms the Philox stream draws for particles 0..n-1 at
 * `step`, written as R[c*stride + i] (columns 0,1,2,3,5 used, 4 and 6 zero)
 * in the context precision.  The reference refreshes cp_R_vec with fresh
 * uniforms before each push (kernels/electron_kernel_numba.py:300). */
int bz_fill_uniforms(bz_ctx_t ctx, void* R, int64_t stride, int64_t n, uint64_t seed,
                     uint64_t step);

/* Deposit only (no push): node moments of the current state, same layout
 * as bz_push_mcc's output.  Replaces the initial post_processing call
 * (1bin/boltzmann.py:363-371). */
int bz_deposit(bz_ctx_t ctx, const bz_state_t* s, const bz_params_t* p, double* moments);

/* Field step, FP64, on device: ni += ni_src; Te = 2/3 nrg/ne; g2 += ni_src;
 * explicit drift-diffusion of ions / metastables; t += dt; tridiagonal
 * Poisson solve at t; E (FP64) and E_t (context precision) written.
 * Replaces boltzmann.py:643-647 (host moment update),
 * kernels/heavies_kernel_1D.py:5-64 and kernels/EF_kernel_1D.py:8-44.
 * chol: [2][N-1] inverse diagonal and lower band of the Cholesky factor. */
int bz_field_step(bz_ctx_t ctx, const bz_field_params_t* fp, const double* moments,
                  double* ni, double* ns, double* Te, double* g2, double* V,
                  double* E, void* E_t, const double* chol, int32_t advance_time);

/* Histogram of r value columns into coarse rows, the reference's native
 * reduction with the same arguments: for i < ptcl, fine bin b = int(x[i]*M),
 * coarse row k = b / (M/RM); dw[r*k + j] += v[r*i + j].  dw is [M][r] FP64;
 * rows 0..RM-1 receive the coarse sums (all other rows are zeroed), summed in
 * a fixed order.  Replaces cuda_red.calls (kernels/reduction/cuda_reduction/
 * kernels-new.cu:76-132, reda :56-73, redcheck :8-50) behind
 * reduction.reduction(r, p, M, RM, dw, dv, dx, ptcl)
 * (kernels/reduction/cuda_reduction/__init__.py:11-15). */
int bz_reduction(bz_ctx_t ctx, int32_t r, int32_t M, int32_t RM, double* dw, const double* dv,
                 const double* dx, int64_t ptcl);
/* The same reduction with no context and no allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, serving loops, harnesses that forbid hipMalloc).  Its
 * scratch is a static 8 MiB device array per GPU, so RM*r may be at most
 * (2^20 - 64)/2 = 524256 (sizes above are rejected with an error); calls that
 * overlap on one GPU share that array and race.  Work is enqueued on
 * hip_stream (NULL: the default stream) of the current device; the result
 * equals bz_reduction's bit for bit. */
int bz_reduction_static(int32_t r, int32_t M, int32_t RM, double* dw, const double* dv,
                        const double* dx, int64_t ptcl, void* hip_stream);

/* Energy / cell binning and stable radix sort by key = cell*nb + ebin,
 * cell = int((x + dx/2)/dx) (1bin/boltzmann.py:658), n_keys = n_nodes*nb <= 65534.
 * thresholds: [nb-1] eV, ascending (bin b holds thr[b-1] <= e < thr[b]).
 * Particles with w <= 0 get the extra key n_keys and sort last.
 * counts[n_keys+1], offsets[n_keys+2] (offsets[n_keys] = live particles),
 * perm[n] (sorted position -> index); equal keys keep ascending index order.
 * Replaces bin_particles (1bin/kernels/bin_particles.py:7-31,
 * 3bin/kernels/bin_particles.py:39-78, 5bin/kernels/bin_particles.py:86-140). */
int bz_bin_sort(bz_ctx_t ctx, const bz_state_t* s, const bz_params_t* p,
                const double* thresholds_host, int32_t nb,
                int32_t* counts, int32_t* offsets, uint32_t* perm);

/* The counts of bz_bin_sort alone (counts[n_keys+1], same keys), without the
 * sort: what the resample trigger reads (1bin/boltzmann.py:894-916 calls
 * bin_particles for its counts); a triggered pass then calls bz_bin_sort. */
int bz_bin_count(bz_ctx_t ctx, const bz_state_t* s, const bz_params_t* p,
                 const double* thresholds_host, int32_t nb, int32_t* counts);

/* ES-WPNC resampling of the state sorted by bz_bin_sort: per (cell, bin)
 * segment split or merge according to `pol`, written in key order to out.
 * targets: [n_keys] int32 device array.  Count taken from / written to the
 * clock.  Replaces _check_for_resampling + _do_resampling
 * (1bin/boltzmann.py:799-916, 3bin/boltzmann.py:813-1113,
 * 5bin/boltzmann.py:843-1334). */
int bz_resample(bz_ctx_t ctx, const bz_state_t* in, const bz_state_t* out,
                const bz_params_t* p, const int32_t* counts, const int32_t* offsets,
                const uint32_t* perm, const int32_t* targets, int32_t n_keys,
                const bz_resample_policy_t* pol);

/* ES-WPNC resampling in place, in the reference's own layout
 * (1bin/boltzmann.py:806-877): the slots of segments that are neither split
 * nor merged are not touched; a split particle keeps its slot with weight
 * w/(c+1) and its c clones are appended; a merged segment's members become
 * holes (w = 0) and its K particles (weight W/K, rescaled velocities) are
 * appended.  The appended particles follow the slots in use in key order
 * (cell, then energy bin), a split particle's clones in slot order.
 * Selections, weights and velocities are those of bz_resample (segment
 * members in ascending slot order).  The holes stay until bz_compact, the
 * reference's compaction after resampling (1bin/boltzmann.py:879-890).
 * Binning as bz_bin_sort (thresholds_host [nb-1]); the keys the last 1-D
 * push wrote for this state are used when it wrote them (bz_set_push_keys).
 * Replaces _check_for_resampling + _do_resampling (1bin/boltzmann.py:799-916,
 * 3bin/boltzmann.py:813-1113, 5bin/boltzmann.py:843-1334). */
int bz_resample_inplace(bz_ctx_t ctx, const bz_state_t* s, const bz_params_t* p,
                        const double* thresholds_host, int32_t nb, const int32_t* targets,
                        int32_t n_keys, const bz_resample_policy_t* pol);

/* ---- Batched 0-D swarms (no counterpart in the reference: its 0-D test
 * problems are single runs made by hand-editing the driver, README.md:7-8) ----
 * All instances live in one SoA state.  A 0-D particle has no position, so
 * x holds its instance number k (0 <= k < n_inst, exact in either
 * precision).  inst_E: [n_inst] field (V/cm) in the context precision.
 * bz_swarm_step advances every slot in place (velocities only: 0-D has no
 * walls, x and w are kept) and appends the secondaries after the slots in
 * use, in parent order, each in its parent's instance; `out` is staging for
 * them.  The step then advances the clock (step += 1, t += p->dt).
 * inst_off is neither read nor written by the step (it may be NULL).
 * inst_moments: [n_inst][8] FP64 of this step: sum w, sum w e, sum w vx,
 * weights of elastic / excitation / ionization events, births, sum w e^2.
 * inst_acc (optional): running sums of inst_moments when accumulate != 0. */
int bz_swarm_step(bz_ctx_t ctx, const bz_state_t* in, const bz_state_t* out,
                  const bz_params_t* p, const bz_xsec_t* xs, int32_t n_inst, int64_t* inst_off,
                  const void* inst_E, double* inst_moments, double* inst_acc,
                  int32_t accumulate);
/* Population control of every instance towards targets[k] with the ES-WPNC
 * merge / split of bz_resample: the slots of `in` are sorted by instance
 * (stable), each instance is merged, split or copied into `out`, and
 * inst_off [n_inst+1] (int64 device array) receives the new layout: instance
 * k owns [inst_off[k], inst_off[k+1]) of `out`.  counts_ws [n_inst+1] and
 * offsets_ws [n_inst+2] are int32 device scratch. */
int bz_swarm_control(bz_ctx_t ctx, const bz_state_t* in, const bz_state_t* out,
                     const bz_params_t* p, int32_t n_inst, int64_t* inst_off,
                     const int32_t* targets, int32_t* counts_ws, int32_t* offsets_ws,
                     const bz_resample_policy_t* pol);
/* EEDF per instance: sum of weights in n_bins log-spaced energy bins on
 * [e_lo, e_hi] eV (out [n_inst][n_bins] FP64).  Reads the grouped layout
 * inst_off that bz_swarm_control wrote for s. */
int bz_swarm_eedf(bz_ctx_t ctx, const bz_state_t* s, int32_t n_inst, const int64_t* inst_off,
                  int32_t n_bins, double e_lo, double e_hi, double kev, double* out);

/* ---- hipGraph capture of enqueued stages (no counterpart in the reference).
 * Everything between begin and end on the context stream (which must be a
 * created stream) becomes one graph; counts, step and time are read from the
 * device clock, so replays advance the run.  Pushes inside a capture use the
 * ticket schedule. */
int bz_graph_begin(bz_ctx_t ctx);
int bz_graph_end(bz_ctx_t ctx, int32_t* id);
int bz_graph_launch(bz_ctx_t ctx, int32_t id);
int bz_graph_destroy(bz_ctx_t ctx, int32_t id);

/* ---- Learned closure (no counterpart in the reference) ----------------
 * A perceptron 3 -> 64 -> 64 -> 64 -> 3 with SiLU hidden activations maps
 * [log10 E/N (Td), mean energy (eV), log10 n_e (m^-3)] to
 * [log10 k_ion (m^3/s), log10 k_exc (m^3/s), log10 mu*N (1/(V m s))].
 * Parameters are flat, row-major W[out][in]:
 *   W1[64][3] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3].
 * Every pass reads BF16 weights (round-to-nearest-even copy of an FP32
 * master) and accumulates in FP32.  Inputs / outputs are row-major [n][3]
 * FP32 in physical units; the affine normalisation is part of the network. */
#define BZ_MLP_IN 3
#define BZ_MLP_HIDDEN 64
#define BZ_MLP_OUT 3
#define BZ_MLP_NPARAM 8771

typedef struct {
  const uint16_t* w;      /* BF16 parameters on the device, BZ_MLP_NPARAM */
  double in_shift[3];     /* x_n = (x - in_shift) * in_scale */
  double in_scale[3];
  double out_shift[3];    /* y = y_n * out_scale + out_shift */
  double out_scale[3];
  double in_lo[3];        /* input domain of the training set: bz_closure_source */
  double in_hi[3];        /* clamps its features to it (no clamp where lo >= hi) */
} bz_mlp_t;

typedef struct {
  float lr, beta1, beta2, eps;
  float bias1, bias2;     /* 1 - beta1^t, 1 - beta2^t of this update */
} bz_adam_t;

/* y[n][3] = net(x[n][3]). */
int bz_mlp_forward(bz_ctx_t ctx, const bz_mlp_t* net, const float* x, float* y, int64_t n);
/* Floats of scratch that bz_mlp_grad needs for a batch of n samples. */
int64_t bz_mlp_workspace_floats(int64_t n);
/* Gradient of L = (1/2n) sum |y_n - t_n|^2 in normalised output units
 * (t_n = (t - out_shift) / out_scale) w.r.t. the BF16 parameters, reduced
 * in a fixed order (deterministic).  grad [BZ_MLP_NPARAM] FP32, loss [1]. */
int bz_mlp_grad(bz_ctx_t ctx, const bz_mlp_t* net, const float* x, const float* t, int64_t n,
                float* ws, float* grad, float* loss);
/* Adam update of the FP32 master and its moments; refreshes the BF16 copy. */
int bz_mlp_adam(bz_ctx_t ctx, const bz_adam_t* opt, float* master, float* m, float* v,
                const float* grad, uint16_t* w);
/* BF16 copy of an FP32 master (round to nearest even). */
int bz_mlp_round(bz_ctx_t ctx, const float* master, uint16_t* w);

/* Closure source: ionization per node produced in one sub-step of length dt
 * by the learned rate, src_i = k_ion(E/N_i, <e>_i, n_e,i) N_gas ne_i dt in
 * the moments' density units.  E/N_i uses the node field (mean |E| of the
 * adjacent cells, E in V/cm), <e>_i = nrg_i / ne_i and n_e,i = 1e6 ne_i
 * (cm^-3 -> m^-3), all from moments [3][N+1] (the kinetic deposit layout);
 * each feature is clamped to the network's input domain [in_lo, in_hi], so
 * sheath fields beyond the trained range get the rate at its edge.
 * src [N+1] is written; acc [N+1] (optional) accumulates it. */
int bz_closure_source(bz_ctx_t ctx, const bz_mlp_t* net, int32_t n_cells, const double* E,
                      const double* moments, double gas_density_m3, double dt, double* src,
                      double* acc);
/* Field sub-step driven by a closure source: ni += src, g2 += src and the
 * electron density ne (moments row 0) += src (pair creation), then the
 * heavies and the Poisson solve of bz_field_step, t += dt.  The kinetic
 * moment update is not applied. */
int bz_field_substep(bz_ctx_t ctx, const bz_field_params_t* fp, double* moments,
                     const double* src, double* ni, double* ns, double* Te, double* g2,
                     double* V, double* E, void* E_t, const double* chol);

/* Kinetic update of a closure-driven run.  The sub-steps since the last
 * update added the learned ionization acc [N+1] to ni; the kinetic step
 * counted nsrc (moments row 1) over the same interval: ni += nsrc - acc,
 * g2 alike, Te = 2/3 nrg/ne, and the tally's running sums (FP64, fixed
 * order) tally[0..5] += sum nsrc, sum acc, |sum nsrc - sum acc|, 1, sum ne,
 * sum nrg.  acc is cleared. */
int bz_closure_commit(bz_ctx_t ctx, int32_t n_nodes, const double* moments, double* acc,
                      double* ni, double* g2, double* Te, double* tally);

/* Host helper: Philox4x32-10 block (for tests and stream generation). */
void bz_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif /* BZ_H_ */
