// Host-side launch environment shared by the libbz translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <string.h>

#include "../../include/bz.h"
#include "bz_device.h"

namespace bz {

// candidate ring of one wave of the queued push: 5 state words x 128 entries
constexpr int kPushQueueWords = 5 * 128;
// waves per CU the queue buffer is sized for (at most 2048 threads per CU)
constexpr int kPushQueueWavesPerCu = 32;

struct LaunchEnv {
  hipStream_t stream;
  int precision;
  int num_cu;
  int64_t capacity;
  Clock* clock;
  uint64_t* status;          // tile status words
  int64_t status_words;
  double* partials;          // per-block partial histograms
  int64_t partials_doubles;
  double* dep_l2;            // level-2 rows of the partial reduction
  int64_t dep_l2_doubles;
  uint32_t* dep_cols_done;   // per output column: level-2 rows written (reset after use)
  int dep_cols;
  void* sort_ws;             // sort / resample workspace
  int64_t sort_ws_bytes;
  int32_t max_nodes;
  int32_t max_keys;
  int push_shape;            // tile shape picked by the autotuner
  int push_sched;            // 0 static tiles + cooperative launch, 1 ticket counter
  int push_bpc;              // 1-D push: blocks per CU (0: as many as fit)
  int push_queue;            // 1-D push candidate queue: 0 auto, 1 always, -1 never
  unsigned long long* dep_m1_lo;  // 1-D push: global fixed-point ionization sums [max_nodes]
  uint32_t* dep_m1_hi;            //   and their carry words (zero between pushes)
  // 1-D push collision counters (bz_set_collision_counts): fixed-point sums
  // [2][max_nodes] + carry words and the ionization counts [max_nodes],
  // folded after every push into the FP64 accumulator coll_acc [3][max_nodes]
  int coll_on;
  unsigned long long* coll_lo;
  uint32_t* coll_hi;
  unsigned long long* coll_cnt;
  double* coll_acc;
  // queued push: per-wave rings of candidate states, push_q_waves rings of
  // kPushQueueWords words (in the state type) each
  void* push_q;
  int64_t push_q_waves;
  int push_keys_nb;          // 1-D push writes binning keys for nb bins (0: off)
  int push_keys16;           // ... as 16-bit words (bz_set_push_keys16)
  double push_keys_thr[8];   // their energy thresholds
  // optional kernel timing: event pairs per timed kernel class
  struct Timing* timing;
};

// 0 push, 1 sort, 2 resample (whole stages); 3 the resampling copy and
// 4 each radix downsweep (single kernels inside stages 2 and 1); single
// kernels of the in-place resampling: 5 merges and splits (k_res_segments),
// 6 the key histogram (k_ip_hist); 7 the hole filling's moves (k_fill_move)
constexpr int kTimedKinds = 8;
struct Timing {
  int on;
  int cap;
  int used[kTimedKinds];
  int64_t launches[kTimedKinds];
  hipEvent_t* start[kTimedKinds];
  hipEvent_t* stop[kTimedKinds];
};
// record start/stop around a launch when timing is on
void timing_begin(const LaunchEnv& env, int kind);
void timing_end(const LaunchEnv& env, int kind);

int fail(const char* fmt, ...);
int check_launch(const char* what);

int launch_push(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                const void* E, const bz_params_t* p, const bz_xsec_t* xs, const void* R,
                const void* G, double* moments, int shape);
int launch_deposit(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                   double* moments);
// the weight bound of a state into the clock (the deposit's fixed-point scale)
int launch_dep_wmax(const LaunchEnv& env, const bz_state_t* s);
int launch_compact(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out);
// the reference's bin_particles on its tosum rows and curr_xbins
int launch_bin_lists(const LaunchEnv& env, const double* tosum, int64_t cols,
                     const int64_t* curr_xbins, int64_t nc, int n_cells, int nb,
                     const double* thr, int64_t* const* lists, int64_t* const* write_inds,
                     const int64_t* max_ppc);
// the 1-D collision counters accumulated since the last reset, to out[3][nn]
int launch_collision_read(const LaunchEnv& env, double* out, int nn, int reset);
// in place: the holes below the live count filled from the live slots above it
int launch_fill_holes(const LaunchEnv& env, const bz_state_t* s);
// sum the push's per-block partial rows (m doubles each) into out[m];
// commit != 0 also sums the counter rows and commits the clock
// sums the per-block partial rows into out (and adds them to acc when given)
int launch_dep_reduce(const LaunchEnv& env, int nblocks, int m, double* out, int commit,
                      int64_t cap, double* acc = nullptr);
// widest partial row any reduction of a context with these sizes produces
int64_t dep_max_row(int32_t max_nodes, int32_t max_keys);
int launch_reduction(const LaunchEnv& env, int r, int M, int RM, double* dw, const double* dv,
                     const double* dx, int64_t ptcl);
int launch_field(const LaunchEnv& env, const bz_field_params_t* fp, const double* moments,
                 double* ni, double* ns, double* Te, double* g2, double* V, double* E,
                 void* E_t, const double* chol, int advance, const double* src = nullptr);
int launch_closure_source(const LaunchEnv& env, const bz_mlp_t* net, int N, const double* E,
                          const double* mom, double gas_m3, double dt, double* src, double* acc);
int launch_closure_commit(const LaunchEnv& env, int nn, const double* mom, double* acc, double* ni,
                          double* g2, double* Te, double* tally);
// per-key counts of the binning keys, no sort (resample trigger input)
int launch_bin_count(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                     const double* thr, int nb, int32_t* counts);
int launch_bin_sort(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                    const double* thr, int nb, int32_t* counts, int32_t* offsets,
                    uint32_t* perm, bool keys_ready = false);
// batched swarms: stable sort of the slots by instance (held in x)
int launch_group_sort(const LaunchEnv& env, const bz_state_t* s, int n_inst, int32_t* counts,
                      int32_t* offsets, uint32_t* perm, const uint32_t** skeys);
int launch_resample(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                    const bz_params_t* p, const int32_t* counts, const int32_t* offsets,
                    const uint32_t* perm, const int32_t* targets, int32_t n_keys,
                    const bz_resample_policy_t* pol);
size_t sort_workspace_bytes(int64_t capacity, int32_t max_keys);
// in-place resampling (reference layout) of a state whose slot keys are in
// the sort input (written by the push, or by launch_slot_keys), as 32-bit or
// (keys16) 16-bit words
int launch_resample_inplace(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                            const void* keys, int keys16, int nb, const int32_t* targets,
                            int32_t n_keys, const bz_resample_policy_t* pol);
int launch_slot_keys(const LaunchEnv& env, const bz_state_t* s, const bz_params_t* p,
                     const double* thr, int nb);
// per-block partial sums of the push: node histograms (max_nodes) or
// per-instance moments (batched swarms, up to max_keys instances)
int64_t push_partials_doubles(int num_cu, int32_t max_nodes, int32_t max_keys);

}  // namespace bz

namespace bz {
// batched swarms in place (x = instance), births staged in out
int launch_push_swarm(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                      const bz_params_t* p, const bz_xsec_t* xs, int n_inst, const void* inst_E,
                      double* inst_moments, double* acc, int shape);
// resampling kernels on a state already grouped by key (no sort): counts,
// offsets (int32, n_keys+1 / n_keys+2) and keys-per-position built by the caller
int launch_resample_grouped(const LaunchEnv& env, const bz_state_t* in, const bz_state_t* out,
                            const bz_params_t* p, const int32_t* counts, const int32_t* offsets,
                            const uint32_t* perm, const uint32_t* skeys, const int32_t* targets,
                            int32_t n_keys, const bz_resample_policy_t* pol, int32_t** plan_out);
int64_t mlp_workspace_floats(int64_t n);
int launch_mlp_forward(const LaunchEnv& env, const bz_mlp_t* net, const float* x, float* y,
                       int64_t n);
int launch_mlp_grad(const LaunchEnv& env, const bz_mlp_t* net, const float* x, const float* t,
                    int64_t n, float* ws, float* grad, float* loss);
int launch_mlp_adam(const LaunchEnv& env, const bz_adam_t* opt, float* master, float* m, float* v,
                    const float* grad, uint16_t* wbf);
int launch_mlp_round(const LaunchEnv& env, const float* master, uint16_t* wbf);
void* sort_input_keys(const LaunchEnv& env);     // key per slot, input of the sort
BinKeys bin_keys(const bz_params_t* p, const double* thr, int nb);
void* sort_scratch_keys(const LaunchEnv& env);   // keys-per-position buffer
void* sort_scratch_perm(const LaunchEnv& env);   // position -> index buffer
}  // namespace bz
