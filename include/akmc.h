/*
 * libakmc -- MI355X-native vacancy-hop AKMC engine, C ABI.
 *
 * Plain pointers and sizes only.  The engine owns device memory; the caller
 * owns every host buffer it passes.  Every call returns an int status
 * (AKMC_OK or a negative AKMC_ERR_*); akmc_last_error() gives the message of
 * the last failure on the calling thread.  Nothing in the library calls
 * exit() or abort().  One engine may be used by one host thread at a time;
 * distinct engines may be created and used from different threads.
 *
 * Site encoding used by akmc_upload/akmc_download: one uint8 per lattice site,
 * index x + 2*nx*(y + ny*z) with x doubled (x even = corner site, x odd = body
 * site), bits: 1 = V, 2 = Cu, 4 = Mn, 8 = Ni, none = Fe -- membership in the
 * reference's four occupancy hash containers (src/lattice/lattices_list.h:
 * 329-333); a site reads as the first present of V, Cu, Mn, Ni, else Fe.
 */
#ifndef AKMC_H
#define AKMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AKMC_OK 0
#define AKMC_ERR_INVALID (-1)   /* bad argument or state */
#define AKMC_ERR_HIP (-2)       /* a HIP runtime call failed */
#define AKMC_ERR_CAPACITY (-3)  /* a sector/replica exceeds a kernel capacity */
#define AKMC_ERR_NO_EVENT (-4)  /* u * sum >= sum: the reference asserts here (kmc.cpp:245,393) */
#define AKMC_ERR_CONFIG (-5)    /* config.yaml could not be parsed */
#define AKMC_ERR_BUDGET (-6)    /* HBM admission budget exceeded */

/* Rate precision. */
#define AKMC_PREC_FP64 0  /* bit-exact with the reference CPU build */
#define AKMC_PREC_FP32 1  /* FP32 bond energies and exp, FP64 sums, search and dt */

/* Why akmc_run returned. */
#define AKMC_STOP_TIME 1    /* the time loop finished: ceil(physics_time / T) steps */
#define AKMC_STOP_EVENTS 2  /* the event cap was reached (mid-sector stop) */

/* LatticeTypes::lat_type values (reference src/lattice/lattice_types.h:20-36)
 * used for the alloy list. */
#define AKMC_TYPE_FE 0x0001
#define AKMC_TYPE_CU 0x0002
#define AKMC_TYPE_NI 0x0004
#define AKMC_TYPE_MN 0x0008

#define AKMC_MAX_ALLOY 32

/* Simulation parameters: the fields of the reference's config.yaml that the
 * vacancy-hop path reads (profiles/config/config_values.h:15-70), plus the
 * sector time threshold that the reference hard-codes to 3e-7 s
 * (src/simulation.h:73-74). */
typedef struct akmc_params {
  int64_t box[3];            /* box.size, unit cells (x not doubled) */
  double lattice_const;      /* box.lattice_const */
  double cutoff_radius;      /* box.cutoff_radius; ghost width = ceil(cutoff / a0) */
  int32_t n_alloy;           /* create.random.alloy, in document order */
  int32_t alloy_type[AKMC_MAX_ALLOY]; /* lat_type value of each alloy key */
  int64_t alloy_ratio[AKMC_MAX_ALLOY];
  int64_t va_count;          /* create.random.va_count */
  int64_t seed_create_types; /* seeds.* (0 = std::random_device, as the reference) */
  int64_t seed_create_vacancy;
  int64_t seed_event_selection; /* parsed but unused, as in the reference */
  int64_t seed_time_inc;
  double temperature;        /* simulation.temperature, K */
  double physics_time;       /* simulation.physics_time, s */
  double attempt_freq;       /* simulation.attempt_freq, 1/s */
  double sector_time;        /* sublattice threshold T per sector cycle, s (default 3e-7) */
  int32_t precision;         /* AKMC_PREC_* */
  int32_t reserved;
} akmc_params;

typedef struct akmc_stats {
  int64_t events;            /* events performed since creation */
  int64_t step;              /* current time step */
  int32_t sector_pos;        /* ring position 0..7 of the current sector (ring 0,7,2,5,3,4,1,6) */
  int32_t in_sector;         /* 1 if a sector is in progress (event-cap stop) */
  double evolution[8];       /* evolution time of each sector id */
  double sector_threshold;   /* threshold of the sector in progress */
  double sector_time;        /* partial time of the sector in progress */
  int32_t stop;              /* AKMC_STOP_* of the last akmc_run */
  int32_t sum_errors;        /* events whose running sums disagreed with a serial refold (akmc_set_check_sums) */
  double seconds;            /* wall time of the last akmc_run */
  double kernel_seconds;     /* device time of the event-loop launches of the last akmc_run */
  int64_t launches;          /* event-loop launches of the last akmc_run */
  int64_t n_eval;            /* vacancy evaluations since creation (55-site stencil gathers) */
  int64_t n_fold;            /* vacancies whose running sum was folded from their 8 rates */
  int64_t n_shift;           /* vacancies whose running sum was shifted by an exact delta */
} akmc_stats;

typedef struct akmc_engine akmc_engine;

/* Defaults of the reference example config (examples/config.yaml). */
int akmc_params_default(akmc_params* p);

/* Reads the reference's config.yaml schema (examples/config.yaml; parser
 * profiles/config/config_parsing.cpp:54-230).  Replaces
 * ConfigParsing::newInstance/parse. */
int akmc_params_from_yaml(const char* path, akmc_params* p);

/* Creates an engine on HIP device `device`.  Replaces initialize_gpu and
 * gpu_prepare (reference gpu/gpu_simulate.h:58-60,
 * gpu/gpu_simulate.hip:1764-1831). */
int akmc_create(const akmc_params* p, int device, akmc_engine** out);

/* Initial random occupancy exactly as creation::createRandom
 * (reference src/creation.cpp:96-164), then uploads it.  Replaces the host
 * hash construction and the initInfo gather (gpu_simulate.hip:1874-2040). */
int akmc_init_random(akmc_engine* e);

/* The same initial occupancy written to a host array of 2*nx*ny*nz sites,
 * without a device. */
int akmc_random_sites(const akmc_params* p, uint8_t* sites, int64_t n);

/* Site arrays (see the encoding above); n must equal akmc_num_sites(). */
int64_t akmc_num_sites(const akmc_engine* e);
int akmc_upload(akmc_engine* e, const uint8_t* sites, int64_t n);
int akmc_download(akmc_engine* e, uint8_t* sites, int64_t n);

/* Runs the sublattice time loop on the device until the loop ends or
 * max_events more events have been performed.  Replaces
 * firstSectLocalToGpu, first_calculate_GPU, calculate_GPU,
 * selectAndPerformEventGPU, calculateNextAdvRegion and sector_final
 * (gpu/gpu_simulate.h:31-54) together with SubLattice::startTimeLoop
 * (src/algorithms/sl/sublattice.inl:25-190).  stats may be NULL. */
int akmc_run(akmc_engine* e, int64_t max_events, akmc_stats* stats);

int akmc_get_stats(akmc_engine* e, akmc_stats* stats);

/* Parity hook: evaluates every vacancy of sector `sector` (sector id 0..7,
 * bit0 x high, bit1 y high, bit2 z high) as ABVIModel::calcRates does
 * (models/abvi/kmc.cpp:68-115, including its side effect on stale species
 * tags) and returns their site indices in sweep order, their 8 rates each
 * and the sector total.  *n receives the vacancy count (may exceed cap). */
int akmc_sector_rates(akmc_engine* e, int sector, int64_t cap, int64_t* sites, double* rates, double* total,
                      int64_t* n);

/* Sector id (0..7: bit0 x high, bit1 y high, bit2 z high) of each of n real
 * site indices (x + 2 box[0] (y + box[1] z)) of a box of box[0] x box[1] x
 * box[2] cells, as the event loops assign sites to sectors: the low half of
 * an axis of n cells holds floor(n/2) cells.  Host only.  Replaces the
 * lookup of local_sector_region in the colored domain
 * (src/algorithms/sl/sublattice.inl:195; split of src/simulation.cpp:91-110,
 * per-sector bounds of add_ghost, gpu/gpu_simulate.hip:2093-3086). */
int akmc_site_sectors(const int64_t box[3], const int64_t* sites, int64_t n, int8_t* sectors);

/* Event log: when enabled, akmc_run records every event's (from, to) site
 * indices, for trajectory comparison.  capacity 0 disables it. */
int akmc_set_event_log(akmc_engine* e, int64_t capacity);
int akmc_get_event_log(akmc_engine* e, int64_t* from, int64_t* to, int64_t cap, int64_t* n);

/* Events per random-number chunk (default and maximum 2^20).  A launch of
 * the event loop consumes at most one chunk, so this sets the launch
 * granularity.  Only before the first akmc_run. */
int akmc_set_rng_chunk(akmc_engine* e, int64_t events);

/* Self-check of the running sums: when on, after every event the event loop
 * refolds the current sector's rates serially in sweep order (the order of
 * ABVIModel::calcRates, models/abvi/kmc.cpp:103) and compares every running
 * sum and the total with the ones it keeps incrementally; disagreements are
 * counted in akmc_stats.sum_errors.  A test aid with no counterpart in the
 * reference: it costs a full sector fold per event. */
int akmc_set_check_sums(akmc_engine* e, int on);

/* How the event loop selects while its helper wavefront still repairs the
 * running sums behind the last changed block: 0 (the default) from the
 * exact prefix and the old sums shifted by the block's change when their
 * error bound decides the slot, waiting for the exact sums otherwise; 1
 * always waits; 2 uses the estimate on events of odd index only.  Every mode
 * selects the same events (the estimate is used only where it provably
 * agrees); a test aid with no counterpart in the reference, which selects
 * from sums it refolds in full at every event. */
int akmc_set_select_mode(akmc_engine* e, int mode);

/* Events [first, first + n) of the two random streams akmc_run consumes, one
 * pair per event: u = r::random() of a default-seeded mt19937
 * (src/utils/random/random.cpp:35-41) and l = -log(r), r drawn by
 * SubLattice's uniform_real_distribution on mt19937(seed_time_inc)
 * (src/algorithms/sl/sublattice.inl:118-120). */
int akmc_rng_streams(int64_t seed_time_inc, int64_t first, int64_t n, double* u, double* l);

/* The exp of the rate kernel over an array, on the host and on a device:
 * the table-driven algorithm that reproduces the C library's exp bit for bit
 * (the reference's arrhenius, models/abvi/rate/rates_solver.cpp:25). */
int akmc_exp_host(const double* x, double* y, int64_t n);
int akmc_exp_device(int device, const double* x, double* y, int64_t n);

/* Initial occupancy of pipe creation (creation::createFromPile,
 * src/creation.cpp:184-230; selected by create.pipe in config.yaml,
 * profiles/config/config_parsing.cpp:171-188): the atoms of the random
 * creation's draw (seeds.create_types; p->va_count is ignored), then a
 * vacancy at every "x y z" triple of pipe_file (x doubled, real coordinates;
 * triples outside the box are skipped).  The site becomes a vacancy with no
 * other tag, as the reference writes its type.  A file that cannot be opened
 * is an error (the reference reads it silently as empty).  n_vacancies (may
 * be NULL) receives the vacancy count.  Upload the result with akmc_upload. */
int akmc_pipe_sites(const akmc_params* p, const char* pipe_file, uint8_t* sites, int64_t n, int64_t* n_vacancies);

/* create.pipe.input_box of config.yaml when pipe creation is its create
 * option (profiles/config/config_parsing.cpp:150-188), else AKMC_ERR_CONFIG. */
int akmc_pipe_path_from_yaml(const char* yaml_path, char* out, int64_t cap);

/* Self-interstitial (dumbbell) rates of the abvi model on a typed lattice.
 * types: one LatticeTypes value per real site (uint16: V 0, Fe 1, Cu 2, Ni 4,
 * Mn 8, dumbbell XY = X << 8 | Y with X <= Y; src/lattice/lattice_types.h:
 * 20-36), index x + 2 box[0] (y + box[1] z), periodic box of at least 3
 * cells per axis.  sites: the n_itl dumbbell sites, strictly ascending (the
 * order of ItlList::mp, models/abvi/defect/itl_list.h); orient: their
 * orientations, -6 .. 6 without 0 (models/abvi/orientation.h:12-30).
 * Outputs avail[n_itl] (Itl::beforeRatesUpdate's avail_trans_dir,
 * itl_list.cpp:9-29) and rates[8 n_itl] (Itl::updateRates with
 * ItlRatesSolver::rate as the callback, itl_list.cpp:31-50,
 * rate/itl_rates_solver.cpp:16-80, rate/bonds/energy_dumbbell.cpp:7-206):
 * every mask first, then every rate.  kernel_seconds (may be NULL): device
 * time of the two kernels.  The reference has no caller of this model
 * (models/abvi/kmc.cpp:393-395), and its lattice no longer holds dumbbell
 * types; this is the model's rate step for callers that do. */
int akmc_itl_rates(const int64_t box[3], const uint16_t* types, int64_t n_itl, const int64_t* sites,
                   const int8_t* orient, double attempt_freq, double temperature, int device, uint8_t* avail,
                   double* rates, double* kernel_seconds);

/* akmc_itl_rates for callers that cannot allocate device memory (code
 * between pre-allocated buffers in a serving loop, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same inputs, outputs and
 * results as akmc_itl_rates, but d_types, d_sites, d_orient, d_avail and
 * d_rates are device pointers on `device`, the work runs on `stream` (a
 * hipStream_t; NULL is the null stream) and the call waits for it.  The
 * working arrays live in a static device array of the library, sized for at
 * most AKMC_ITL_NOALLOC_MAX dumbbells (11.5 MiB per device, held until the
 * process ends); more dumbbells return AKMC_ERR_CAPACITY.  Inputs are checked
 * on the device; an invalid input returns AKMC_ERR_INVALID and writes no
 * output.  The call allocates nothing.  Every call on a device shares that
 * device's array and leaves its contents behind: calls that overlap on one
 * device (from several threads or streams) race and must be serialised by
 * the caller. */
#define AKMC_ITL_NOALLOC_MAX 16384
int akmc_itl_rates_noalloc(const int64_t box[3], const uint16_t* d_types, int64_t n_itl, const int64_t* d_sites,
                           const int8_t* d_orient, double attempt_freq, double temperature, int device, void* stream,
                           uint8_t* d_avail, double* d_rates);

/* Waits for all work on a device (benchmark bracketing). */
int akmc_sync(int device);

/* Copies `bytes` between two device buffers `iters` times with a plain
 * streaming kernel and returns the average seconds per copy: a transfer of
 * known size for calibrating the profiler's HBM byte counters.  No
 * counterpart in the reference. */
int akmc_copy_bench(int device, int64_t bytes, int iters, double* seconds);

/* Frees the engine.  Replaces gpu_final (gpu/gpu_simulate.h:56). */
void akmc_destroy(akmc_engine* e);

/* ------------------------------------------------------------------------
 * Replica engine: batches of independent lattices.
 *
 * The reference runs one lattice per MPI rank and has no batched form; these
 * entry points stand where a driver would otherwise loop over simulation::
 * simulate (src/simulation.h:61-85) once per replica.  Each replica follows
 * the reference's physics (stencil, bond counting, rate, perform rule,
 * sublattice time loop) with its own xoshiro256** stream (seeded from
 * seed_time_inc) and FP64 block sums, so trajectories are reproducible per
 * replica but not the reference's bit for bit; the single-lattice engine
 * above is the bit-exact path.  A replica's trajectory depends only on its
 * parameters, initial lattice and seed -- not on the batch it runs in, the
 * slicing of runs or the kernel tuning.
 * ---------------------------------------------------------------------- */

/* Initial occupancy of a replica. */
#define AKMC_INIT_DEVICE 0     /* counter-based draw on the device: each site independently takes a type with
                                  probability ratio / sum(ratio); then va_count uniform site draws become
                                  vacancies (Cu tag dropped, as creation.cpp:131-164) */
#define AKMC_INIT_REFERENCE 1  /* creation::createRandom on the host, bit-exact with akmc_random_sites */

/* Replica status. */
#define AKMC_REP_RUNNING 0
#define AKMC_REP_TIME_DONE 1    /* ceil(physics_time / T) time steps finished */
#define AKMC_REP_TARGET_DONE 2  /* target event count reached */

typedef struct akmc_batch_options {
  int32_t precision;        /* AKMC_PREC_FP32 (default) or AKMC_PREC_FP64 */
  int32_t init;             /* AKMC_INIT_* */
  int32_t error_stats;      /* 1: FP32 rates are also evaluated in FP64 and their error recorded */
  int32_t use_graph;        /* 1: launches replay a captured hipGraph */
  int32_t lds_vcap;         /* tuning: replicas with at most this many vacancies keep their sector in
                               LDS (0 = autotune, -1 = none) */
  int32_t waves_per_group;  /* tuning: replicas per workgroup, 1, 2 or 4 (0 = autotune) */
  int32_t order;            /* tuning: 0 replica order, 1 largest first (-1 = autotune) */
  int32_t interleave;       /* tuning: replica waves per SIMD, 1, 2 or 4 (0 = 1); above 1 the workgroups
                               hold 4 * interleave waves and waves_per_group is ignored */
  int64_t slice_events;     /* events per replica per launch (0 = default 1024) */
} akmc_batch_options;

typedef struct akmc_batch_stats {
  int64_t events;           /* events of all replicas in the last akmc_batch_run */
  double seconds;           /* its wall time */
  double kernel_seconds;    /* device time of its launches */
  int64_t launches;         /* slice launches (graph replays) */
  int32_t running;          /* replicas still running afterwards */
  int32_t failed;           /* replicas stopped by an error */
} akmc_batch_stats;

typedef struct akmc_replica_stats {
  int64_t events;
  int64_t step;
  int32_t status;           /* AKMC_REP_*, or a negative error */
  int32_t sector_pos;
  double evolution[8];
  double sector_time;
  int64_t nvac;
  int64_t n_eval;           /* vacancy evaluations */
  int64_t n_rebuild;        /* sector starts */
  int64_t n_jump;           /* sector visits with a zero total rate */
  double err_max;           /* FP32 rate error vs FP64 (error_stats): max relative error */
  double err_mean;
  int64_t err_n;            /* rates compared */
  int64_t err_tiny;         /* FP64 rates below FLT_MIN, excluded */
  int64_t n_skip;           /* time steps skipped: every occupied sector already past its threshold */
  double busy_seconds;      /* device time of the work items that advanced this replica */
} akmc_replica_stats;

typedef struct akmc_batch_tuning {
  int32_t lds_vcap;
  int32_t waves_per_group;
  int32_t order;
  int32_t interleave;       /* replica waves per SIMD */
  int64_t slice_events;
  double events_per_second; /* best trial */
  int32_t trials;
  int32_t reserved2;
} akmc_batch_tuning;

typedef struct akmc_batch akmc_batch;

int akmc_batch_options_default(akmc_batch_options* o);

/* Creates n replicas on `device`, one per params entry (box, alloy, va_count,
 * seeds, temperature, attempt_freq, sector_time, physics_time; the
 * precision field is ignored in favour of o->precision). */
int akmc_batch_create(const akmc_params* replicas, int32_t n, const akmc_batch_options* o, int device,
                      akmc_batch** out);

/* Advances every running replica by up to events_per_replica events (fewer
 * when its time loop ends). */
int akmc_batch_run(akmc_batch* b, int64_t events_per_replica, akmc_batch_stats* stats);

/* One launch in which every running replica advances until a deadline of
 * `seconds` after the launch starts (at an event boundary; the events per
 * replica then differ).  Trajectories are those of akmc_batch_run. */
int akmc_batch_run_for(akmc_batch* b, double seconds, akmc_batch_stats* stats);

/* Searches the tuning space (LDS capacity, replicas per workgroup, interleave
 * width = replica waves per SIMD, work order) with one deadline launch of
 * real work per setting (the replicas
 * advance; their trajectories do not depend on the tuning), keeps the
 * setting with the most events per device second, then brings every running
 * replica to the same event count. */
int akmc_batch_tune(akmc_batch* b, double trial_seconds, akmc_batch_tuning* chosen);
int akmc_batch_get_tuning(akmc_batch* b, akmc_batch_tuning* t);
/* Every setting the last akmc_batch_tune tried, with its events per device
 * second (events_per_second; trials = 1 each): *n is the number of settings,
 * of which the first min(cap, *n) are written to out. */
int akmc_batch_tune_trials(akmc_batch* b, akmc_batch_tuning* out, int32_t cap, int32_t* n);

int32_t akmc_batch_size(const akmc_batch* b);
int64_t akmc_batch_num_sites(const akmc_batch* b, int32_t replica);
int akmc_batch_replica_stats(akmc_batch* b, int32_t replica, akmc_replica_stats* s);
int akmc_batch_download(akmc_batch* b, int32_t replica, uint8_t* sites, int64_t n);
/* Replaces a replica's lattice (vacancies <= its va_count) and restarts its
 * current sector. */
int akmc_batch_upload(akmc_batch* b, int32_t replica, const uint8_t* sites, int64_t n);

/* Parity hook: the rates of every vacancy of one sector of one replica, as
 * a sector start computes them (with the stale-tag side effect), in
 * vacancy-list order, and the FP64 sum of their per-vacancy totals. */
int akmc_batch_sector_rates(akmc_batch* b, int32_t replica, int sector, int64_t cap, int64_t* sites, double* rates,
                            double* total, int64_t* n);

void akmc_batch_destroy(akmc_batch* b);

/* Solute statistics of a lattice (host only, any site array of the encoding
 * above, box in unit cells): Cu atoms (sites reading as Cu) grouped into
 * clusters by 1nn connectivity; hist[k] = clusters of k atoms (the last bin
 * collects hist_cap - 1 and more); n_isolated = the reference's isolated-Cu
 * count (profiles/m_event_hook.cpp:19-55: sites in the Cu set without a 1nn
 * site reading as Cu). */
int akmc_cu_clusters(const uint8_t* sites, const int64_t box[3], int64_t hist_cap, int64_t* hist,
                     int64_t* n_clusters, int64_t* n_cu, int64_t* n_isolated);

/* ---------------------------------------------------------------- outputs
 * The reference's step hook (MEventHook::onStepFinished,
 * profiles/m_event_hook.cpp:17-70) logs the isolated-Cu count every 100 steps
 * (akmc_cu_clusters) and dumps the lattice every output.dump.interval steps.
 * akmc_run_steps stops the time loop at a step boundary so a caller can run
 * those hooks at the same points. */
#define AKMC_STOP_STEPS 3   /* akmc_run_steps reached its step count */

typedef struct akmc_output {
  int64_t dump_interval;     /* output.dump.interval (0: no dumps) */
  int64_t logs_interval;     /* output.logs.interval */
  char dump_file_path[256];  /* output.dump.file_path, "{}" = the step */
  char logs_file[256];       /* output.logs.logs_file */
} akmc_output;

/* The optional `output` section of config.yaml (ConfigParsing::parseOutput,
 * profiles/config/config_parsing.cpp:216-229); all zero when absent. */
int akmc_output_from_yaml(const char* path, akmc_output* o);

/* fmt::format(dump_file_path, step) for patterns whose replacement fields are
 * all "{}" (m_event_hook.cpp:62). */
int akmc_dump_path(const char* pattern, int64_t step, char* out, int64_t cap);

/* Writes a lattice dump in the reference's format (LatticeDump::dump,
 * profiles/io/lattice_dump.cpp:14-63): the site count, an empty line, then
 * "Type\tx\ty\tz" per site in z, y, x order, coordinates of the
 * ghost-extended grid (ghost width `ghost` = ceil(cutoff / a0)), body sites at
 * +0.5.  Type is the readout (stale tags are not written). */
int akmc_write_dump(const uint8_t* sites, const int64_t box[3], int64_t ghost, const char* path);

/* Reads a dump as the reference's XYZ reader does (XYZReader,
 * tools/xyz_reader.cpp:15-80; particle, tools/particle.h): the box from the
 * coordinate ranges, the ghost width from the smallest x.  sites == NULL
 * returns the box only; otherwise cap must hold 2*box[0]*box[1]*box[2]. */
int akmc_read_dump(const char* path, int64_t box[3], int64_t* ghost, uint8_t* sites, int64_t cap);

/* Runs the event loop like akmc_run, but ends it when `steps` more time steps
 * are complete (stop = AKMC_STOP_STEPS), or earlier at the end of the time
 * loop or after max_events events. */
int akmc_run_steps(akmc_engine* e, int64_t steps, int64_t max_events, akmc_stats* stats);

/* Checkpoint and restart of a single-lattice engine (the reference's restart
 * hook is empty, profiles/pkmc.cpp:183-184): the parameters, the sites with
 * their stale tags, the sector-loop state (clocks, step, sector in progress,
 * event count) and the random stream positions.  A loaded engine continues
 * the same trajectory bit for bit, also from a stop inside a sector. */
int akmc_checkpoint_save(akmc_engine* e, const char* path);
int akmc_checkpoint_load(const char* path, int device, akmc_engine** out);

/* The FP32 rate table the replica engine uses at temperature T (K) and
 * attempt frequency nu: the rate of barrier index m (Ea = m * q / 2 eV) is
 * tab[256 + ((m - m_min) >> 8)] * tab[(m - m_min) & 255] in FP32.  Writes
 * 256 + nhi floats (cap must allow it), the energy quantum q and the bond
 * tables in units of q (e1q, e2q: 25 each, row-major V,Fe,Cu,Ni,Mn; e0q2: 5,
 * 2*e0/q).  Host only; a verification hook. */
int akmc_fp32_rate_table(double temperature, double nu, float* tab, int32_t cap, int32_t* nhi, int32_t* m_min,
                         double* q, float* e1q, float* e2q, float* e0q2);

/* ------------------------------------------------------------------------
 * Serving: a request loop over the replica engine.  Jobs (one lattice each)
 * are admitted in arrival order while the HBM budget and the slot count
 * allow, stay device-resident while they run, and advance together: every
 * launch carries all live jobs, and jobs that arrive are added at the next
 * launch (continuous batching).  A job's result does not depend on the jobs
 * it ran beside.  No counterpart in the reference.
 * ---------------------------------------------------------------------- */

#define AKMC_JOB_PENDING 0
#define AKMC_JOB_RUNNING 1
#define AKMC_JOB_DONE 2
#define AKMC_JOB_FAILED 3

typedef struct akmc_server_options {
  int64_t hbm_budget;       /* bytes of job state admitted at once (default 250e9) */
  int32_t max_live;         /* jobs on the device at once (default 4096) */
  int32_t precision;        /* AKMC_PREC_FP32 (default) or AKMC_PREC_FP64 */
  int32_t lds_vcap;         /* as akmc_batch_options (0: 256) */
  int32_t waves_per_group;  /* (0: 2) */
  int64_t slice_events;     /* events per job per launch (default 4096) */
  double slice_seconds;     /* launch time budget (default 2e-3; 0: none) */
  int32_t keep_results;     /* 1: copy each finished lattice to host memory */
  int32_t reserved;
} akmc_server_options;

typedef struct akmc_job_info {
  int32_t status;           /* AKMC_JOB_* */
  int32_t rep_status;       /* AKMC_REP_* of the replica */
  int64_t events;
  int64_t bytes;            /* HBM bytes the job holds while live */
  double t_submit;          /* server clock, seconds */
  double t_admit;
  double t_done;
} akmc_job_info;

typedef struct akmc_server_stats {
  int64_t submitted, admitted, done, failed;
  int64_t launches;
  int64_t events;           /* events of all jobs so far */
  int64_t live;             /* jobs on the device now */
  int64_t bytes_live;       /* their HBM bytes */
  int64_t bytes_peak;
  double busy_seconds;      /* server time spent in launches */
} akmc_server_stats;

typedef struct akmc_server akmc_server;

int akmc_server_options_default(akmc_server_options* o);
/* Starts the request loop (a host thread) on `device`. */
int akmc_server_create(int device, const akmc_server_options* o, akmc_server** out);
/* Queues a job that stops after target_events events (or at the end of its
 * time loop).  Fails with AKMC_ERR_BUDGET if the job alone exceeds the
 * budget.  Thread-safe. */
int akmc_server_submit(akmc_server* s, const akmc_params* p, int64_t target_events, int64_t* job);
/* Queues a job that also carries self-interstitial dumbbells (the abvi
 * model's vacancy + dumbbell event loop, akmc_mix_*, FP64 rates, device
 * streams): the lattice of p drawn on the device as for akmc_server_submit
 * (alloy, va_count vacancies), plus n_itl dumbbells at distinct atom sites,
 * each of the site's atom and an atom of type itl_atom (AKMC_TYPE_FE for
 * self-interstitials), with random orientations; ef110 as akmc_mix_params.
 * p->seed_event_selection seeds the recombination stand-in draws.  The job
 * shares the request loop, the HBM budget and the arrival order with the
 * vacancy-only jobs; its steps are batched with the other dumbbell jobs in
 * one launch, beside the vacancy jobs' launch. */
int akmc_server_submit_itl(akmc_server* s, const akmc_params* p, int64_t n_itl, int32_t itl_atom, double ef110,
                           int64_t target_events, int64_t* job);
int akmc_server_job_info(akmc_server* s, int64_t job, akmc_job_info* info);
/* The finished lattice of a job (keep_results). */
int akmc_server_result(akmc_server* s, int64_t job, uint8_t* sites, int64_t n);
/* The finished typed lattice of a dumbbell job (keep_results). */
int akmc_server_result_typed(akmc_server* s, int64_t job, uint16_t* types, int64_t n);
/* Waits until no job is pending or running (timeout_s <= 0: no limit). */
int akmc_server_wait_idle(akmc_server* s, double timeout_s);
int akmc_server_get_stats(akmc_server* s, akmc_server_stats* st);
double akmc_server_clock(akmc_server* s);
void akmc_server_destroy(akmc_server* s);

/* ------------------------------------------------------------------------
 * Vacancies and self-interstitial dumbbells: the abvi model's event loop on a
 * typed lattice (one LatticeTypes value per site, as akmc_itl_rates takes).
 *
 * The reference has the pieces but no wired loop: the dumbbell branch of
 * ABVIModel::select is commented out (models/abvi/kmc.cpp:251-268), perform
 * asserts on every event but a vacancy hop (:393-395), and the dumbbell
 * transition, recombination and defect generation are described in comments
 * (:310-355).  These entry points wire them as the oracle restates them
 * (oracle/refcpu.cpp, refmix): the sublattice time loop and selection of
 * the vacancy engine, with a sector's vacancies and dumbbells interleaved in
 * sweep order (8 slots each: a vacancy's direction rates, a dumbbell's
 * Itl::updateRates slots), dumbbell transitions by orientation::tranAtom and
 * orientation::trans (orientation.cc:238-252), recombination of a vacancy
 * and a dumbbell at 1nn or 2nn after every transition (rec::RecList,
 * recombine.cpp, with an mt19937 of seed_recombine in place of std::rand),
 * and defect generation at defect_gen_rate (env.h:23; 0 = none, as the
 * reference configures it): a sector site becomes a vacancy and its atom
 * joins one of the site's 12 third neighbours as a dumbbell.  Each handle holds any number of lattices; a run
 * advances all of them in one launch, one wavefront per lattice.
 * ---------------------------------------------------------------------- */
typedef struct akmc_mix_params {
  int64_t box[3];            /* cells; every edge >= 8 */
  double temperature;        /* K */
  double attempt_freq;       /* 1/s */
  double sector_time;        /* sublattice threshold per sector cycle, s */
  double physics_time;       /* s; the time loop runs ceil(physics_time / sector_time) steps */
  double ef110;              /* <110> dumbbell formation energy in pure Fe, eV (env.h:22; default 3.53) */
  double defect_gen_rate;    /* Frenkel-pair generation rate per sector, 1/s (env.h:23; default 0) */
  int64_t seed_time_inc;     /* time stream, as akmc_params */
  int64_t seed_recombine;    /* stand-in stream for std::rand() and the generation draws */
  int32_t device_rng;        /* 0: the reference's mt19937 streams from the host (bit-exact with the oracle);
                                1: xoshiro256** streams on the device (reproducible, no host work) */
  int32_t reserved;
  int64_t vac_capacity;      /* list capacities (0: the initial counts, + 4096 when generating) */
  int64_t itl_capacity;
} akmc_mix_params;

typedef struct akmc_mix_stats {
  int64_t events;
  int64_t step;
  int32_t sector_pos;
  int32_t in_sector;
  double evolution[8];
  double sector_time;
  int32_t stop;              /* AKMC_STOP_TIME or AKMC_STOP_EVENTS */
  int32_t nvac;              /* vacancies now */
  int32_t nitl;              /* dumbbells now */
  int32_t reserved;
  int64_t n_recombined;      /* vacancy-dumbbell pairs recombined */
  int64_t n_generated;       /* Frenkel pairs generated */
  int64_t n_itl_events;      /* dumbbell transitions */
  int64_t n_eval;            /* rate records evaluated */
  double seconds;            /* wall time of the handle's last akmc_mix_run */
  double kernel_seconds;     /* device time of its launches */
  int64_t launches;
  int64_t n_skipped_steps;   /* time steps passed in one go: every occupied sector's clock was ahead */
} akmc_mix_stats;

typedef struct akmc_mix akmc_mix;

int akmc_mix_params_default(akmc_mix_params* p);

/* Creates n lattices on `device`: types[i] holds 2 box[0] box[1] box[2]
 * LatticeTypes values; itl_sites[i] / itl_orient[i] list its n_itl[i]
 * dumbbells in ascending site order with orientations in [-6, 6] \ {0}. */
int akmc_mix_create(int32_t n, const akmc_mix_params* p, const uint16_t* const* types, const int64_t* n_itl,
                    const int64_t* const* itl_sites, const int8_t* const* itl_orient, int device, akmc_mix** out);

/* Advances every lattice by up to max_events events (fewer when its time loop
 * ends); *total_events (may be NULL) receives the events of all lattices. */
int akmc_mix_run(akmc_mix* m, int64_t max_events, int64_t* total_events);

int akmc_mix_stats_of(akmc_mix* m, int32_t lattice, akmc_mix_stats* stats);

/* The lattice's types (n_sites of them, may be NULL) and its dumbbell list
 * (up to cap entries; *n_itl receives the count). */
int akmc_mix_download(akmc_mix* m, int32_t lattice, uint16_t* types, int64_t n_sites, int64_t cap,
                      int64_t* itl_sites, int8_t* itl_orient, int64_t* n_itl);

/* Parity hook: calcRates of one sector: its vacancies (kind 0) and dumbbells
 * (kind 1) in sweep order with their 8 slots, and the total (with
 * defect_gen_rate). */
int akmc_mix_sector_rates(akmc_mix* m, int32_t lattice, int sector, int64_t cap, int64_t* sites, int32_t* kinds,
                          double* rates, double* total, int64_t* n);

/* Event log of every lattice (capacity events each; 0 disables): from site,
 * to site (-1 for a generation that changed nothing) and kind (0 vacancy
 * hop, 1 dumbbell transition, 2 defect generation). */
int akmc_mix_set_event_log(akmc_mix* m, int64_t capacity);
int akmc_mix_get_event_log(akmc_mix* m, int32_t lattice, int64_t* from, int64_t* to, int32_t* kind, int64_t cap,
                           int64_t* n);

void akmc_mix_destroy(akmc_mix* m);

/* Initial typed lattice: creation::createRandom's occupancy
 * (akmc_random_sites; a vacancy's leftover Ni/Mn tag is dropped), then
 * placeDumbbells of the reference's serial driver
 * (profile_serial/main.cpp:137-161) on an mt19937 of `seed`: for each kind
 * k, itl_counts[k] draws of a site (mt() % sites); a single atom there
 * becomes the dumbbell of itself and an atom of itl_types[k] with
 * orientation::random() (orientation.h:38-42).  The dumbbells come back in
 * ascending site order (up to cap; *n_itl receives the count). */
int akmc_mix_random_lattice(const akmc_params* p, const int32_t* itl_types, const int64_t* itl_counts,
                            int32_t n_kinds, int64_t seed, uint16_t* types, int64_t n_sites, int64_t cap,
                            int64_t* itl_sites, int8_t* itl_orient, int64_t* n_itl);

const char* akmc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* AKMC_H */
