/* chatcarlo_hip.h -- C-ABI of the MI355X (gfx950) photon-transport library.
 *
 * The reference has no FFI: its hot path is the Python API
 *   run_transport(scene, n_histories, seed, batch_size, dose_grid, grid_resolution_cm,
 *                 n_workers, track_uncertainty)        chatcarlo/transport.py:358-361
 *   transport_photons(pos, dirv, energy, geometry, rng, grid, recorder, tally_rng,
 *                     fluorescence_enabled)             chatcarlo/transport.py:98-103
 *   VoxelGrid / accumulate_track_length / end_batch     chatcarlo/tally.py:26-94, 248-293
 *   Geometry.material_at / next_boundary                chatcarlo/geometry.py:205-231
 * The Python package chatcarlo_amd re-exports those names and calls the entry
 * points below through ctypes.  Conventions: every function returns an int
 * status (CC_OK = 0, negative = error class); cc_last_error() gives the message
 * of the last failure on the calling thread.  Host buffers are caller-owned
 * (pointer + length); device buffers are owned by the opaque context.  One
 * host thread per context; all work of a context runs on its own HIP stream.
 */
#ifndef CHATCARLO_HIP_H
#define CHATCARLO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CC_OK = 0,
  CC_ERR_ARG = -1,     /* invalid argument (sizes, ids, ranges) -> ValueError */
  CC_ERR_HIP = -2,     /* HIP runtime failure */
  CC_ERR_STATE = -3,   /* call order (no scene / tables / grid bound) */
  CC_ERR_DEVICE = -4   /* a kernel raised a device error bit (see cc_run_out.error_bits) */
};

/* device error bits reported in cc_run_out.error_bits */
enum {
  CC_DEV_ENERGY_RANGE = 1,   /* energy outside [1,150] keV (chatcarlo/materials.py:234-237) */
  CC_DEV_ITER_CAP = 2,       /* rejection loop hit its trial cap */
  CC_DEV_Q_RANGE = 4,        /* Rayleigh x_max beyond the form-factor table (physics.py:333-337) */
  CC_DEV_STEP_CAP = 8,       /* boundary-crossing cap of one flight hit */
  CC_DEV_ROUND_CAP = 16      /* round cap of one batch hit */
};

enum { CC_PRIM_BOX = 0, CC_PRIM_SPHERE = 1, CC_PRIM_CYLINDER = 2 };
enum { CC_MAX_PRIMS = 64, CC_MAX_MAT = 32 };

/* One analytic primitive (48 bytes).  box: p[0..2] = lo, p[3..5] = hi;
 * sphere: p[0..2] = centre, p[3] = radius; cylinder: p[0..2] = centre,
 * p[3] = radius, p[4] = half height, axis 0/1/2 = x/y/z.
 * Replaces the scene geometry list of chatcarlo/geometry.py:121-164. */
typedef struct {
  int32_t kind, axis, material, pad;
  float p[8];
} cc_prim;

typedef struct {
  float lo[3], hi[3];        /* world box (chatcarlo/geometry.py:160-164) */
  int32_t background;        /* material id outside every primitive */
  int32_t n_prim;
} cc_world;

/* Physics data pack (host pointers, float32 unless noted).  Layouts:
 * sig/slope [n_mat][5][n_e] parts photo, compton, rayleigh, mu_en, total;
 * edge_* [n_mat][max_edges](...); sel per material (3, n_e, n_elem) at sel_off[m];
 * fl [n_el][32] = edge, omega, 1/dlnE, ln edge, lines[8], cdf[8], k_share[12] (K-shell share of
 *   photoabsorption on nodes ln edge + j dlnE, linear in ln E between them);
 * h10 [4][n_h10]: energies, values, ln-ratios of energies, ln-ratios of values.
 * Replaces chatcarlo/materials.py:202-496 and dose_coefficients.py:26-42. */
typedef struct {
  int32_t n_mat, n_e, n_el, n_q, max_edges, n_h10, sel_len, elem_len;
  const float* density;
  const float* e_node;
  const float* dl_node;
  const float* sig;
  const float* slope;
  const int32_t* edge_of_interval;
  const float* edge_e;
  const float* edge_dl;
  const float* edge_sig;
  const float* edge_slope;
  const int32_t* mat_elem_off;   /* n_mat + 1 */
  const int32_t* mat_elem;
  const int32_t* sel_off;        /* n_mat */
  const float* sel;
  const float* sq_over_z;
  const float* q_s;
  const float* ray_x;
  const float* ray_a;
  const float* fl;
  const float* h10;
} cc_tables;

enum { CC_FIELD_RECT = 0, CC_FIELD_CONE = 1, CC_FIELD_PARALLEL = 2 };

/* Source (chatcarlo/source.py:67-210, spectrum.py:73-82). */
typedef struct {
  int32_t field, n_spec, n_angles, rotation;   /* rotation: 0 none, 1 continuous, 2 discrete */
  double pos[3], dir[3], u[3], v[3];
  double sid, width, height, cos_half;
  double iso[3], axis[3];
  double scan_length;
  const double* spec_e;      /* n_spec energies keV */
  const double* spec_cdf;    /* n_spec cumulative probabilities, last = 1 */
  /* Heel effect (chatcarlo/source.py:43-65,128-149,195-207); heel_bins = 0: none.  The heel
   * axis has coefficients (heel_hu, heel_hv) in the beam basis (u, v), so it turns with a rotating
   * tube.  A photon's heel coordinate on the SID plane is s = a heel_hu + b heel_hv (rect field,
   * a and b its offsets) or sid tan(theta) (cos(phi) heel_hu + sin(phi) heel_hv) (cone field).
   * Field points are drawn until one is kept, with probability heel_rel_flu interpolated at s
   * over the bin centres of [-heel_span, heel_span] (heel_bins equal bins).  The photon then
   * draws its energy from spec_e with the CDF of its bin, the bin clamped to [heel_nz_lo,
   * heel_nz_hi], the bins of non-zero fluence. */
  int32_t heel_bins, heel_nz_lo, heel_nz_hi, heel_pad;
  double heel_hu, heel_hv, heel_span;
  const double* heel_rel_flu;  /* heel_bins relative fluences in [0, 1] */
  const double* heel_cdf;      /* heel_bins rows of n_spec cumulative probabilities on spec_e, each ending at 1 */
} cc_source;

enum {
  CC_RUN_FLUORESCENCE = 1,     /* physics.fluorescence (chatcarlo/transport.py:388) */
  CC_RUN_GRID = 2,             /* score the bound grid */
  CC_RUN_TRACK_UNCERTAINTY = 4, /* batch moments (chatcarlo/tally.py:59-94) */
  CC_RUN_TIMING = 8,            /* per-kernel HIP-event timing into cc_run_out.kernel_ms */
  CC_RUN_REGION = 16,           /* H*(10) track-length score in the region set by cc_set_region */
  CC_RUN_RECORD = 32,           /* contribution records (analog runs only, needs CC_RUN_REGION) */
  CC_RUN_IMPORTANCE = 64,       /* weight windows from the learned importance map */
  CC_RUN_SOURCE_BIAS = 128,     /* source cells drawn by learned importance (cc_set_importance) */
  CC_RUN_TRACE = 256,           /* cc_transport: log every flight segment (cc_trace_alloc, cc_trace_read) */
  CC_RUN_PER_BATCH = 512,       /* cc_run: one batch after the other instead of streaming them (same results) */
  CC_RUN_PERSISTENT = 1024      /* cc_run, streamed: one persistent history-based launch per group of
                                   batches instead of event-based rounds (refill, flight, interactions) in a
                                   particle pool whose last photons one persistent launch carries to their
                                   ends; same counters, tallies up to FP64 summation order */
};

typedef struct {
  uint64_t n_histories, n_absorbed, n_escaped, scatter_sum, n_fluorescence, n_batches;
  uint64_t n_events;            /* flight segments (boundary-to-boundary steps) */
  uint64_t n_flights;           /* photons processed by the flight kernel */
  uint64_t n_voxel_visits;      /* voxel crossings scored by the DDA */
  uint32_t error_bits, rounds_max;
  double edep_keV[CC_MAX_MAT];  /* collision estimator per material */
  double edep_sum2[CC_MAX_MAT]; /* sum over batches of S_b^2 / n_b */
  double kernel_ms[8];          /* source, flight, photo, compton, rayleigh, end_batch, other, interact
                                   (photo + Compton + Rayleigh in one launch: streamed runs; other:
                                   the drain of streamed runs, one launch per group of batches) */
  uint64_t kernel_launches[8];  /* launches timed per kernel kind */
  /* region tally (CC_RUN_REGION), pSv per history summed: per-history sum and sum of squares,
   * and the batch form sum over batches of S_b^2 / n_b */
  double region_sum, region_sum2, region_batch_sum2;
  uint64_t n_split;             /* copies made by splitting */
  uint64_t n_roulette;          /* photons ended by Russian roulette */
  uint64_t n_records;           /* contribution records held after the run */
  double kernel_busy_ms[8];     /* per kernel kind of kernel_ms: the time during which at least one
                                   of its launches ran (the union of their intervals).  Streamed runs
                                   of two lanes overlap launches of one kind, so this is below
                                   kernel_ms; batch-by-batch runs: equal to kernel_ms */
} cc_run_out;

/* Learned importance map: f(x) = MLP 7 -> 64 -> 64 -> 1 (BF16 MFMA, FP32 accumulate).
 * Flat parameters: W1 [16][64] (rows 0..6 features, row 7 bias, rows 8..15 zero),
 * W2 [64][64], b2 [64], w3 [64], b3. */
enum { CC_MLP_IN = 16, CC_MLP_HIDDEN = 64, CC_MLP_PARAMS = 16 * 64 + 64 * 64 + 64 + 64 + 1 };
enum { CC_RECORD_FLOATS = 12 };   /* x y z dx dy dz E score_before target slot batch kind */
/* Two maps of the same shape: 0 for photons about to fly (flight starts, the source),
 * 1 for photons about to interact (collision point, incoming direction and energy). */
enum { CC_MAP_FLIGHT = 0, CC_MAP_COLLISION = 1 };

typedef struct {
  uint32_t steps, batch;        /* Adam steps, records per step */
  float lr, beta1, beta2, eps;
  float yscale;                 /* target scale; <= 0: n_records / sum of targets */
  float lr_end;                 /* cosine decay of the step size from lr to lr_end */
  uint64_t seed;                /* minibatch permutation */
  uint32_t target_power;        /* 1: fit the mean of the target; 2: its second moment */
  uint32_t pad;
} cc_train_cfg;

typedef struct {
  double loss_first, loss_last; /* mean Poisson deviance of the first and last 1% of steps */
  double ms;                    /* device time of the training loop */
  float yscale;
  uint32_t n_records, n_positive;
  uint32_t pad;
} cc_train_out;

typedef struct {
  float cw;                     /* target weight = clamp(cw / exp(f), w_lo, w_hi); <= 0: mean source importance */
  float w_lo, w_hi;
  int32_t split_max;            /* at most this many copies per split */
  float pool_factor;            /* particle pool = batch * pool_factor */
  int32_t source_bias;          /* 1: build the source cells for CC_RUN_SOURCE_BIAS */
  float bias_mix;               /* defensive share of the analog cell probabilities */
  int32_t samples_per_cell;     /* source photons per cell for the mean importance */
  int32_t use_post;             /* window at flight starts (map 0; source photons only with use_pre) */
  int32_t use_pre;              /* split at collisions before the outcome is drawn, roulette the outcomes */
  float split_gain;             /* collision split count gain (<= 0: 1) */
  uint32_t pad;
  uint64_t seed;
} cc_importance_cfg;

const char* cc_last_error(void);
int cc_version(int32_t* major, int32_t* minor);

int cc_ctx_create(int32_t device, void** ctx);
int cc_ctx_destroy(void* ctx);
int cc_upload_tables(void* ctx, const cc_tables* t);
int cc_upload_scene(void* ctx, const cc_prim* prims, int32_t n_prim, const cc_world* world);

/* Grid: origin (3), shape (nx, ny, nz), cubic voxel size; flags bit 1 = track_uncertainty.
 * Replaces VoxelGrid.from_bbox / __post_init__ (chatcarlo/tally.py:42-57). */
int cc_bind_grid(void* ctx, const double* origin, const int32_t* shape, double voxel_cm, int32_t flags);
/* Batch-tally copies used by the next cc_bind_grid (power of two, 1..64; default 4). Workgroup b
 * adds into copy b mod copies, and workgroups are dealt to the MI355X's 8 XCDs round-robin: with 4,
 * two XCDs share a copy, so atomics on the hot voxels of a beam are contended by two XCDs, not 8.
 * No counterpart in the reference (its NumPy tally has no concurrent writers); results do not
 * depend on it beyond FP64 summation order. */
int cc_set_grid_copies(void* ctx, int32_t copies);
/* Calibration of the tally path's bound, with no counterpart in the reference: the rate of
 * no-return FP64 atomic adds (the tally adds) at uniformly random addresses of an n_slots-double
 * array, n_ops adds in one launch (n_ops a multiple 1..4096 of 2^21), as adds per second. On the
 * MI355X these atomics are performed beyond the L2 caches, and this rate bounds the voxel tallies.
 * sum (may be NULL) returns the array's total after the timed launch and a warm-up launch of the
 * same size: 2 n_ops rounded down to whole per-thread counts when every add landed. */
int cc_atomic_rate(void* ctx, uint64_t n_slots, uint64_t n_ops, double* ops_per_s, double* sum);
int cc_unbind_grid(void* ctx);
int cc_reset_grid(void* ctx);
/* Copy the totals and moments back (any pointer may be NULL). */
int cc_read_grid(void* ctx, double* kerma_keV, double* h10_pSv_cm3, double* kerma_sum2,
                 double* h10_sum2, int32_t* n_batches_hit, int64_t* n_batches, int64_t* n_histories);
/* Add caller arrays into the device totals (resume / merge). */
int cc_load_grid(void* ctx, const double* kerma_keV, const double* h10_pSv_cm3, const double* kerma_sum2,
                 const double* h10_sum2, const int32_t* n_batches_hit, int64_t n_batches, int64_t n_histories);

/* Full run: histories [first_history, first_history + n) in batches of `batch`;
 * Philox keyed by seed.  Replaces run_transport's serial branch
 * (chatcarlo/transport.py:257-312, 383-396). */
int cc_run(void* ctx, const cc_source* src, uint64_t n_histories, uint64_t seed, uint64_t first_history,
           uint32_t batch, uint32_t flags, cc_run_out* out);

/* Caller-supplied photons, updated in place (chatcarlo/transport.py:98-233).
 * pos, dir: n x 3; energy: n.  Outputs per photon: n_scatter, absorbed, escaped. */
int cc_transport(void* ctx, int64_t n, double* pos, double* dir, double* energy, uint64_t seed,
                 uint64_t first_history, uint32_t flags, int32_t* n_scatter, uint8_t* absorbed,
                 uint8_t* escaped, cc_run_out* out);

/* cc_transport without device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results and errors as cc_transport, for at most
 * CC_NOALLOC_MAX_PHOTONS photons per call (CC_ERR_ARG beyond) and without CC_RUN_TRACE.  The
 * photon records, queues and staging live in one static device array of 10.9 MB (10.4 MiB) that each
 * GPU holds until the process ends; the context's tables, scene and grid are used as bound.  All
 * calls on one GPU share that array, so calls that overlap on one GPU (from several threads or
 * contexts) race; calls in sequence are independent. */
enum { CC_NOALLOC_MAX_PHOTONS = 1 << 16 };
int cc_transport_noalloc(void* ctx, int64_t n, double* pos, double* dir, double* energy, uint64_t seed,
                         uint64_t first_history, uint32_t flags, int32_t* n_scatter, uint8_t* absorbed,
                         uint8_t* escaped, cc_run_out* out);

/* Geometry queries on the GPU (chatcarlo/geometry.py:205-231). */
int cc_locate(void* ctx, int64_t n, const float* pts, int32_t* material);
int cc_test_boundary(void* ctx, int64_t n, const float* o, const float* d, float eps, float* t,
                     int32_t* escape);

/* Uniform-fed sub-kernels: the production device functions fed caller uniforms. */
/* out per photon: mu, p_photo, p_compt, photo, compton, rayleigh, mu_en, mu_total, h10 */
int cc_test_xs(void* ctx, int64_t n, const int32_t* mat, const float* e, float* out9);
int cc_test_compton(void* ctx, int64_t n, const int32_t* mat, const float* e, const float* u,
                    int32_t k_trials, float* eps, float* cos_t, int32_t* elem, int32_t* trials);
int cc_test_rayleigh(void* ctx, int64_t n, const int32_t* mat, const float* e, const float* u,
                     int32_t k_trials, float* cos_t, int32_t* elem, int32_t* trials);
int cc_test_fluor(void* ctx, int64_t n, const int32_t* mat, const float* e, const float* u4,
                  int32_t* elem, int32_t* emit, float* e_line);
int cc_test_rotate(int64_t n, const float* d, const float* cos_t, const float* u_phi, float* out);
int cc_test_isotropic(int64_t n, const float* u1, const float* u2, float* out);
/* Score segments (o, d, length, weights per cm) into the bound grid's batch buffers and fold them. */
int cc_test_dda(void* ctx, int64_t n, const float* o, const float* d, const float* len,
                const double* w_kerma, const double* w_h10);
int cc_test_philox(int64_t n, const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4);
/* Learned importance (no counterpart in the reference; its transport is analog only).
 * Scoring region: sphere of H*(10) track-length scoring; radius <= 0 clears it.  The MLP
 * feature frame is centred on the region with length scale scale_cm (<= 0: 100 cm). */
int cc_set_region(void* ctx, const double* center, double radius_cm, double scale_cm);
/* Contribution records of analog runs (half the capacity per kind) and the probability
 * that a photon is recorded at a flight start or a collision. */
int cc_records_alloc(void* ctx, uint32_t capacity, float keep_prob);
int cc_records_read(void* ctx, int32_t kind, uint32_t first, uint32_t n, float* out, uint32_t* n_held);
int cc_records_clear(void* ctx);
int cc_mlp_set_params(void* ctx, int32_t map, const float* w);   /* CC_MLP_PARAMS floats; resets Adam */
int cc_mlp_get_params(void* ctx, int32_t map, float* w);
/* Feature rows (n x 16) of states in the frame of the scoring region (without one: the
 * world-box centre, half the largest extent). */
int cc_mlp_features(void* ctx, int64_t n, const float* pos, const float* dir, const float* energy, float* feat);
int cc_mlp_forward(void* ctx, int32_t map, int64_t n, const float* feat, float* out);
/* Gradient of sum_r exp(f_r) - yscale y_r f_r over the rows, and that loss. */
int cc_mlp_grad(void* ctx, int32_t map, int64_t n, const float* feat, const float* y, float yscale, float* grad,
                double* loss);
/* Adam on the held records of the map's kind; loss_curve (steps doubles) may be NULL. */
int cc_mlp_train(void* ctx, int32_t map, const cc_train_cfg* cfg, cc_train_out* out, double* loss_curve);
/* Weight-window parameters and, when cfg->source_bias or cfg->cw <= 0, the source cells of src.
 * cell_prob (512 floats) may be NULL; cw_out receives the cw in use. */
int cc_set_importance(void* ctx, const cc_importance_cfg* cfg, const cc_source* src, float* cw_out,
                      float* cell_prob);

/* Segment trace of cc_transport (chatcarlo/trajectory.py:15-43 TrajectoryRecorder, filled at
 * chatcarlo/transport.py:219-229): with CC_RUN_TRACE, every flight segment of every photon is logged,
 * from its start to the boundary crossed, the interaction point or the world exit.  Logging draws no
 * random numbers: results are those of the same call without it.  The records are in no particular
 * order; sorting by (photon, flight, step) gives each photon's segments in time order. */
enum { CC_EV_BOUNDARY = 0, CC_EV_ESCAPE = 1, CC_EV_PHOTOELECTRIC = 2, CC_EV_FLUORESCENCE = 3,
       CC_EV_COMPTON = 4, CC_EV_RAYLEIGH = 5 };
typedef struct {
  uint32_t photon;              /* index of the photon in the cc_transport call */
  int32_t flight;               /* interactions the photon had before this segment */
  int32_t step;                 /* segment index within its flight */
  int32_t material;             /* material id of the segment */
  float start[3], end[3];       /* cm */
  float energy_keV;             /* photon energy along the segment */
  int32_t event;                /* CC_EV_*: what ends the segment */
} cc_trace_record;
/* Allocates room for `capacity` records (replacing an earlier buffer). */
int cc_trace_alloc(void* ctx, uint32_t capacity);
/* Copies min(records logged, capacity, max_records) records of the last traced call into out
 * (may be NULL) and sets *n_total to the records logged; n_total > capacity means records were
 * dropped and the call should be repeated with a larger buffer. */
int cc_trace_read(void* ctx, uint64_t max_records, cc_trace_record* out, uint64_t* n_total);

/* Batched runs (no counterpart in the reference, which runs one scene per call): many scenes in
 * one particle pool.  Scene s's histories 0 .. n_s - 1 draw from the streams keyed by (seed_s, h),
 * so each scene's tallies are those of cc_run(seed_s, first 0) up to FP64 summation order.  A
 * device cursor hands out the histories, in one of two forms: (persistent = 0) to free slots of a
 * particle pool, with two event-based transport rounds and the status copy captured in one hipGraph
 * and replayed until the queue drains; (persistent = 1) to the lanes of one persistent
 * history-based kernel, captured as a one-node hipGraph and launched once, each wave claiming
 * histories from the cursor as its lanes go idle.  Both give the same event counters; tallies and
 * deposits agree up to FP64 summation order.  Tallies: kerma and H*(10) totals per scene and
 * per-material deposits (no batch moments). */
typedef struct {
  int32_t block;                /* threads per block of the transport kernels (64, 128, 192 or 256) */
  int32_t rounds_per_sync;      /* rounds between host checks of the queue (even; event-based form) */
  uint32_t pool;                /* photons in flight (event-based form) */
  uint32_t tally_copies;        /* copies of the scene tallies (0 or a power of two <= 16; 0 = 1): workgroup
                                   b adds into copy b mod copies, folded once the queue has drained */
  int32_t persistent;           /* 1: one persistent launch; 0: replayed event-based rounds */
  int32_t epoch;                /* persistent form: wave iterations between flushes of a workgroup's
                                   LDS tally cache into the grid (1 .. 65536) */
} cc_batch_tuning;

typedef struct {
  uint64_t n_histories, n_absorbed, n_escaped, scatter_sum, n_fluorescence;
  uint64_t n_events, n_flights, n_voxel_visits;
  uint64_t rounds, graph_launches;
  double device_ms;             /* from the first replay to the drained queue */
  uint32_t error_bits, pad;
} cc_batch_out;

/* prims of all scenes concatenated, scene s owning [prim_off[s], prim_off[s+1]); grid_origin and
 * grid_shape are n_scenes x 3. */
int cc_batch_upload(void* ctx, int32_t n_scenes, const cc_prim* prims, const int32_t* prim_off,
                    const cc_world* worlds, const cc_source* sources, const uint64_t* n_histories,
                    const uint64_t* seeds, const double* grid_origin, const int32_t* grid_shape,
                    const double* voxel_cm);
/* Runs ceil(fraction n_s) histories of every scene (fraction in (0, 1]); tallies start from zero. */
int cc_batch_run(void* ctx, const cc_batch_tuning* tune, double fraction, uint32_t flags, cc_batch_out* out);
int cc_batch_read(void* ctx, int32_t scene, double* kerma_keV, double* h10_pSv_cm3, double* edep_keV);
/* Times launch shapes on a fraction of the histories and returns the fastest, three timing runs:
 * the persistent form at block 512 and 256 (epoch 2) and the event-based form at (block 128,
 * rounds_per_sync 4, pool 2^19), a later one taken only when it beats the best so far by 5 %;
 * tally copies: the largest power
 * of two up to 16 whose copies fit in half the free device memory.  table (7 doubles per row:
 * persistent, block, epoch, rounds_per_sync, pool, tally_copies, histories/s; one row per run)
 * may be NULL.  The scene tallies are cleared afterwards. */
int cc_batch_autotune(void* ctx, double fraction, uint32_t flags, cc_batch_tuning* best, double* table,
                      int32_t max_rows, int32_t* n_rows);

/* Source sampling for histories [first, first + n): pos, dir (n x 3), energy (n). */
int cc_test_source(void* ctx, const cc_source* src, int64_t n, uint64_t seed, uint64_t first_history,
                   float* pos, float* dir, float* energy);

#ifdef __cplusplus
}
#endif
#endif
