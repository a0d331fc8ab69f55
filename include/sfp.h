/*
 * sfp.h -- C ABI of libsfp_hip.so, the MI355X (gfx950) implementation of the
 * sfparticles per-particle step.
 *
 * Three groups of entry points:
 *
 *  1. Per-operation kernels, one per callable of the reference's three
 *     backend seams, on caller-owned DEVICE arrays (struct-of-arrays, FP64,
 *     bool arrays as uint8):
 *       - backend module  sfparticles/particles.py:14-25 (cpu.py / gpu.py)
 *       - QED model       sfparticles/particles.py:8-12  (qed/ rejection_sampling.py, optical_depth.py)
 *       - field kernel    sfparticles/fields.py:26-46 via particles.py:299-323
 *  2. Species stores: library-owned device buffers with capacity, the
 *     device-resident counterpart of a Particles object's attributes
 *     (particles.py:98-146, 476-516).
 *  3. The fused simulation: the step loop of sfparticles/simulation.py:102-144
 *     as one fused HIP kernel per species and step (fields, chi, QED sampling,
 *     hard-photon cut, recoil, stable append of new particles, push).
 *
 * Every function returns SFP_OK (0) or an error code; sfp_last_error() gives
 * a message.  Pointers named *_dev are device pointers; all work is issued on
 * the HIP stream passed in (NULL = the default stream) and is asynchronous
 * unless stated otherwise.
 */
#ifndef SFP_H
#define SFP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFP_OK 0
#define SFP_E_HIP 1      /* a HIP runtime call failed */
#define SFP_E_ARG 2      /* invalid argument */
#define SFP_E_CAPACITY 3 /* a store ran out of capacity inside a step */
#define SFP_E_STATE 4    /* invalid object state */

typedef void* sfp_stream_t; /* hipStream_t */

/* analytic field term kinds (sfparticles/fields.py:69-271) */
#define SFP_FIELD_STATIC 0
#define SFP_FIELD_SIMPLE 1
#define SFP_FIELD_GAUSSIAN 2
#define SFP_FIELD_FOCUSED 3

/* One field term with the user-level arguments of the reference builder:
 *  STATIC   (fields.py:268)  p = Ex, Ey, Ez, Bx, By, Bz
 *  SIMPLE   (fields.py:69)   p = a0, w0, ctau, t_order, direction, x0, wavelength, pol_angle, cep
 *  GAUSSIAN (fields.py:120)  p = a0, w0, ctau, direction, x0, l0, pol_angle, cep
 *  FOCUSED  (fields.py:173)  p = a0, w0, ctau, direction, xfoc, x0, l0, pol_angle, cep
 * A field is a sum of terms, accumulated left to right (fields.py:52-61). */
typedef struct {
  int32_t kind;
  int32_t pad;
  double p[12];
} sfp_field;

/* Rejection-sampling tables, HOST pointers (qed/rejection_sampling_tables.py:14-26) */
typedef struct {
  const double* int_Ai; /* [z_N] */
  const double* Aip;    /* [z_N] */
  int64_t z_N;
  double z_min, z_max, z_delta;
  /* Ai(z) on the same grid, for the spin-resolved emission rate; may be NULL
   * (no counterpart in the reference, whose emission ignores spin) */
  const double* Ai;
} sfp_rs_tables;

/* Optical-depth tables, HOST pointers (qed/optical_depth_tables.py:14-32) */
typedef struct {
  const double* photon_total; /* [chi_N] */
  const double* pair_total;   /* [chi_N] */
  const double* photon_cdf;   /* [chi_N][delta_N] */
  const double* pair_cdf;     /* [chi_N][delta_N] */
  int64_t chi_N;
  double log_chi_min, log_chi_max, log_chi_delta;
  int64_t delta_N;
  double log_delta_min, log_delta_max, log_delta_delta;
  const double* chi_grid;   /* [chi_N]   10**(log_chi_min + i*log_chi_delta) */
  const double* delta_grid; /* [delta_N] 10**(log_delta_min + j*log_delta_delta) */
  double chi_min;           /* 10**log_chi_min */
  double chi_max;           /* 10**log_chi_max */
} sfp_od_tables;

const char* sfp_last_error(void);
int sfp_version(void);
/* number of compute units of the current device and its name (name may be NULL) */
int sfp_device_info(int32_t* num_cus, char* name, int32_t name_len);

/* Device memory helpers for callers without their own HIP allocator (the
 * Python host layer uses these; no counterpart in the reference, whose arrays
 * are numpy/cupy).  Copies are synchronous; sfp_free waits for the device's
 * queued work before it releases the buffer. */
int sfp_malloc(void** dev, int64_t bytes);
int sfp_free(void* dev);
int sfp_memcpy_htod(void* dst_dev, const void* src_host, int64_t bytes);
int sfp_memcpy_dtoh(void* dst_host, const void* src_dev, int64_t bytes);
int sfp_synchronize(void);
/* Free and total device memory in bytes (for sizing stores up front). */
int sfp_mem_info(int64_t* free_bytes, int64_t* total_bytes);

/* Upload QED tables into library-owned device memory; either pointer may be NULL. */
int sfp_tables_create(const sfp_rs_tables* rs, const sfp_od_tables* od, void** tables);
int sfp_tables_destroy(void* tables);

/* ---------------------------------------------------------------------------
 * 1. Per-operation kernels
 * ------------------------------------------------------------------------- */

/* seam 3: fields.py:26-46 (Fields.field_func), evaluated for [0, n) skipping pruned */
int sfp_eval_field(const sfp_field* terms, int32_t nterms, const double* x_dev,
                   const double* y_dev, const double* z_dev, double t, int64_t n,
                   const uint8_t* pruned_dev, double* Ex_dev, double* Ey_dev, double* Ez_dev,
                   double* Bx_dev, double* By_dev, double* Bz_dev, sfp_stream_t stream);

/* seam 1: cpu.py:5-19 */
int sfp_push_position(double* x_dev, double* y_dev, double* z_dev, const double* ux_dev,
                      const double* uy_dev, const double* uz_dev, const double* inv_gamma_dev,
                      int64_t n, const uint8_t* pruned_dev, double dt, sfp_stream_t stream);
/* cpu.py:22-31 */
int sfp_boris(double* ux_dev, double* uy_dev, double* uz_dev, double* inv_gamma_dev,
              const double* Ex_dev, const double* Ey_dev, const double* Ez_dev,
              const double* Bx_dev, const double* By_dev, const double* Bz_dev, double q,
              int64_t n, const uint8_t* pruned_dev, double dt, sfp_stream_t stream);
/* cpu.py:33-42 */
int sfp_boris_tbmt(double* ux_dev, double* uy_dev, double* uz_dev, double* inv_gamma_dev,
                   double* sx_dev, double* sy_dev, double* sz_dev, const double* Ex_dev,
                   const double* Ey_dev, const double* Ez_dev, const double* Bx_dev,
                   const double* By_dev, const double* Bz_dev, double q, double ae, int64_t n,
                   const uint8_t* pruned_dev, double dt, sfp_stream_t stream);
/* cpu.py:44-51 */
int sfp_LL_push(double* ux_dev, double* uy_dev, double* uz_dev, double* inv_gamma_dev,
                const double* chi_dev, int64_t n, const uint8_t* pruned_dev, double dt,
                sfp_stream_t stream);
/* cpu.py:53-60 */
int sfp_CLL_push(double* ux_dev, double* uy_dev, double* uz_dev, double* inv_gamma_dev,
                 const double* chi_dev, int64_t n, const uint8_t* pruned_dev, double dt,
                 sfp_stream_t stream);
/* cpu.py:62-75 */
int sfp_update_chi(const double* Ex_dev, const double* Ey_dev, const double* Ez_dev,
                   const double* Bx_dev, const double* By_dev, const double* Bz_dev,
                   const double* ux_dev, const double* uy_dev, const double* uz_dev,
                   const double* inv_gamma_dev, double* chi_dev, int64_t n,
                   const uint8_t* pruned_dev, sfp_stream_t stream);
/* cpu.py:227-232 */
int sfp_pick_hard_photon(uint8_t* event_dev, const double* delta_dev,
                         const double* inv_gamma_dev, double threshold, int64_t n,
                         sfp_stream_t stream);
/* cpu.py:78-88 */
int sfp_photon_recoil(double* ux_dev, double* uy_dev, double* uz_dev, double* inv_gamma_dev,
                      const uint8_t* event_dev, const double* delta_dev, int64_t n,
                      const uint8_t* pruned_dev, sfp_stream_t stream);
/* cpu.py:90-118; additionally writes id next_id+k and zeroes chi/tau of the
 * new slot when those pointers are non-NULL */
int sfp_create_photon(const double* x_src, const double* y_src, const double* z_src,
                      const double* ux_src, const double* uy_src, const double* uz_src,
                      double* x_dst, double* y_dst, double* z_dst, double* ux_dst,
                      double* uy_dst, double* uz_dst, double* inv_gamma_dst,
                      uint8_t* photon_pruned_dst, const int64_t* event_index_dev,
                      const double* photon_delta_dev, int64_t n_buffered, int64_t n_photon,
                      uint64_t* id_dst, uint64_t next_id, double* chi_dst, double* tau_dst,
                      sfp_stream_t stream);
/* cpu.py:120-191; same additions, plus spin of the new lepton set to 0 */
int sfp_create_pair(const double* x_src, const double* y_src, const double* z_src,
                    const double* ux_src, const double* uy_src, const double* uz_src,
                    uint8_t* photon_pruned, double* x_dst, double* y_dst, double* z_dst,
                    double* ux_dst, double* uy_dst, double* uz_dst, double* inv_gamma_dst,
                    uint8_t* pair_pruned_dst, const int64_t* event_index_dev,
                    const double* pair_delta_dev, int64_t n_buffered, int64_t n_pair,
                    int32_t inverse_delta, uint64_t* id_dst, uint64_t next_id, double* chi_dst,
                    double* tau_dst, double* sx_dst, double* sy_dst, double* sz_dst,
                    sfp_stream_t stream);
/* cpu.py:193-216: ascending indices of the true flags into event_index_dev
 * (room for n entries); the count goes to *count_host (synchronous). */
int sfp_find_event_index(const uint8_t* event_dev, int64_t n, int64_t* event_index_dev,
                         int64_t* count_host, sfp_stream_t stream);
/* cpu.py:218-225: number of true flags in [0, n) (synchronous) */
int sfp_bool_sum(const uint8_t* flags_dev, int64_t n, int64_t* count_host, sfp_stream_t stream);

/* No counterpart in the reference: sfp_find_event_index for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same indices into
 * event_index_dev (room for n entries); the count goes to the DEVICE pointer
 * count_dev, and the call is asynchronous (no synchronisation, no host
 * copy).  Its scratch (block counts) is a static device array of 2^16
 * int64 (512 KiB) that each GPU holds until the process ends, so n is at
 * most 2^26 (SFP_E_ARG beyond).  Calls that overlap on one GPU share that
 * array and race; calls in order on one stream do not.  It allocates
 * nothing: no hipMalloc, hipFree or allocating library routine. */
#define SFP_NOALLOC_MAX_FLAGS (1ll << 26)
int sfp_find_event_index_noalloc(const uint8_t* event_dev, int64_t n, int64_t* event_index_dev,
                                 int64_t* count_dev, sfp_stream_t stream);

/* 1/gamma from the momentum, [0, n) skipping pruned: 1/sqrt(1+u^2) for
 * leptons (photon == 0), 1/|u| for photons (particles.py:118-123).  No
 * counterpart in the reference, which stores inv_gamma; the fused step keeps
 * it in registers and this rebuilds the array on request. */
int sfp_update_inv_gamma(const double* ux_dev, const double* uy_dev, const double* uz_dev,
                         double* inv_gamma_dev, int64_t n, const uint8_t* pruned_dev,
                         int32_t photon, sfp_stream_t stream);

/* seam 2: qed/rejection_sampling.py:68-85 and :87-101.  Draws are
 * Philox4x32-10 keyed by seed with counter (id, step, tag, slot). */
int sfp_photon_from_rejection_sampling(const double* inv_gamma_dev, const double* chi_dev,
                                       double dt, int64_t n, const uint8_t* pruned_dev,
                                       uint8_t* event_dev, double* delta_dev,
                                       const uint64_t* id_dev, uint64_t seed, uint64_t step,
                                       uint32_t tag, uint32_t slot, void* tables,
                                       sfp_stream_t stream);
int sfp_pair_from_rejection_sampling(const double* inv_gamma_dev, const double* chi_dev,
                                     double dt, int64_t n, const uint8_t* pruned_dev,
                                     uint8_t* event_dev, double* delta_dev, const uint64_t* id_dev,
                                     uint64_t seed, uint64_t step, uint32_t tag, uint32_t slot,
                                     void* tables, sfp_stream_t stream);
/* qed/optical_depth.py:15-62 and :64-110 */
int sfp_update_tau_e(double* tau_dev, const double* inv_gamma_dev, const double* chi_dev,
                     double dt, int64_t n, const uint8_t* pruned_dev, uint8_t* event_dev,
                     double* delta_dev, const uint64_t* id_dev, uint64_t seed, uint64_t step,
                     uint32_t tag, uint32_t slot, void* tables, sfp_stream_t stream);
int sfp_update_tau_gamma(double* tau_dev, const double* inv_gamma_dev, const double* chi_dev,
                         double dt, int64_t n, const uint8_t* pruned_dev, uint8_t* event_dev,
                         double* delta_dev, const uint64_t* id_dev, uint64_t seed, uint64_t step,
                         uint32_t tag, uint32_t slot, void* tables, sfp_stream_t stream);

/* The FP32 envelope bounds the fused step culls with, for n points: without
 * velocities (vx_dev == NULL) out[i] bounds |E| + c|B| at (x, y, z)[i] and
 * time t; with velocities it bounds |E| + c|B| at (x, y, z)[i] + v[i] s and
 * time t + s for every s in [0, T] (a photon's straight path).  +inf where no
 * bound is available.  No counterpart in the reference; exposed so that the
 * bounds can be checked against the exact fields. */
int sfp_field_bound_eval(const sfp_field* terms, int32_t nterms, const double* x_dev,
                         const double* y_dev, const double* z_dev, double t, const double* vx_dev,
                         const double* vy_dev, const double* vz_dev, double T, int64_t n,
                         float* out_dev, sfp_stream_t stream);

/* Elementary functions of the device math library on n device values, for
 * checking that the device and the host compute them identically (no
 * counterpart in the reference).  fn: 0 exp, 1 log, 2 sin, 3 cos, 4 atan,
 * 5 pow(x, y), 6 exp10, 7 log10, 8 sqrt, 9 x / y. */
int sfp_math_eval(int32_t fn, const double* x_dev, const double* y_dev, double* out_dev,
                  int64_t n, sfp_stream_t stream);

/* ---------------------------------------------------------------------------
 * 2. Species stores (device SoA with capacity)
 * ------------------------------------------------------------------------- */
enum sfp_attr {
  SFP_X = 0, SFP_Y, SFP_Z, SFP_UX, SFP_UY, SFP_UZ, /* float64 */
  SFP_ID,                                          /* uint64 */
  SFP_PRUNED,                                      /* uint8 */
  SFP_TAU, SFP_SX, SFP_SY, SFP_SZ,                 /* float64, optional */
  SFP_INV_GAMMA, SFP_CHI,                          /* float64, auxiliary */
  SFP_EX, SFP_EY, SFP_EZ, SFP_BX, SFP_BY, SFP_BZ,  /* float64, auxiliary */
  SFP_EVENT,                                       /* uint8, auxiliary */
  SFP_DELTA,                                       /* float64, auxiliary */
  SFP_SHOT,                                        /* uint32: ensemble shot index, optional */
  SFP_NATTR
};
#define SFP_STORE_TAU 1  /* optical-depth attribute */
#define SFP_STORE_SPIN 2 /* sx, sy, sz */

int sfp_store_create(int32_t flags, int64_t capacity, void** store);
int sfp_store_destroy(void* store);
/* grow to at least `capacity` (content kept; new slots pruned, events false) */
int sfp_store_reserve(void* store, int64_t capacity);
int sfp_store_capacity(void* store, int64_t* capacity);
/* device pointer of an attribute; auxiliary attributes are allocated on first use */
int sfp_store_ptr(void* store, int32_t attr, void** dev_ptr);
int sfp_store_set_count(void* store, int64_t n_buffered, uint64_t next_id);
/* synchronous */
int sfp_store_get_count(void* store, int64_t* n_buffered, uint64_t* next_id);
/* synchronous host<->device copies of elements [offset, offset+n) */
int sfp_store_upload(void* store, int32_t attr, const void* host, int64_t offset, int64_t n);
int sfp_store_download(void* store, int32_t attr, void* host, int64_t offset, int64_t n);
/* particles.py:503-516: stable compaction of every allocated attribute over
 * the non-pruned particles; the new count is returned (synchronous) */
int sfp_store_prune(void* store, int64_t* n_kept);
/* Diagnostics (no counterpart in the reference): per-shot counts and energy
 * spectra of the live particles.  out[nshots][nbins + 1]: column 0 the count
 * of the shot (particles without a SFP_SHOT attribute count as shot 0; larger
 * indices as the last shot), columns 1..nbins the counts in equal bins of
 * log10(energy) over [log10_min, log10_max), energy |u| for photon != 0 and
 * sqrt(1 + u^2) otherwise.  Synchronous. */
int sfp_store_histogram(void* store, int32_t photon, int32_t nshots, int32_t nbins,
                        double log10_min, double log10_max, int64_t* out);

/* ---------------------------------------------------------------------------
 * 3. Fused simulation (simulation.py:12-161)
 * ------------------------------------------------------------------------- */
#define SFP_RR_NONE 0
#define SFP_RR_PHOTON 1
#define SFP_RR_LL 2
#define SFP_RR_CLL 3

#define SFP_QED_RS 0 /* per-step probability, rejection sampling */
#define SFP_QED_OD 1 /* optical depth */

typedef struct {
  void* store;
  double q;        /* charge in C (particles.py:79) */
  double m;        /* mass in units of m_e; 0 for photons */
  double ae;       /* anomalous moment (spin species) */
  int32_t rr;      /* SFP_RR_* */
  int32_t push;    /* particles.py:82 */
  int32_t spin;    /* SpinParticles */
  int32_t photon_dst; /* index of the photon species this lepton emits into, or -1 */
  int32_t pair_e_dst; /* photon: index of the electron species, or -1 */
  int32_t pair_p_dst; /* photon: index of the positron species, or -1 */
  int32_t pad[2];
} sfp_species_desc;

/* Arithmetic of the fused step.  FP64: the reference's precision, results
 * bit-identical to the CPU restatement.  MIXED: fields, pushers and QED
 * sampling in FP32, positions, time and chi in FP64 (no counterpart in the
 * reference; its error against FP64 is measured per observable). */
#define SFP_PRECISION_FP64 0
#define SFP_PRECISION_MIXED 1

typedef struct {
  double dt;
  double photon_threshold; /* simulation.py:19 */
  uint64_t seed;
  int32_t qed_model;       /* SFP_QED_* */
  int32_t precision;       /* SFP_PRECISION_* */
  /* 1: photon emission of spin-carrying leptons under rejection sampling uses
   * the spin-resolved rate (needs the Ai table).  No counterpart in the
   * reference (0, its behaviour: emission ignores spin, cpu.py:78-118). */
  int32_t spin_emission;
  int32_t pad;
} sfp_sim_params;

int sfp_sim_create(const sfp_species_desc* species, int32_t nspecies, const sfp_field* terms,
                   int32_t nterms, void* tables, const sfp_sim_params* params, void** sim);
int sfp_sim_destroy(void* sim);
/* Advance nsteps steps from time *t (updated: t += dt per step, as
 * simulation.py:143) and step index step0 (the Philox counter).  Counts stay
 * on the device; the host synchronises every 64 steps.  A step whose appends
 * would overflow a store halts on the device before changing anything; the
 * store grows and the step is replayed from the halted kernel.
 * QED photons proven cold sleep, with the position of the step they fell
 * asleep at, and the sleep goes on into the next call when it starts at the
 * step this one ended and no store was changed from the host in between
 * (otherwise every photon is visited at its first step).  The sfp_store_*
 * calls that read or change particle data (download, upload, ptr, prune,
 * reserve, a new count) and sfp_sim_destroy first bring sleeping photons to
 * the last step run, so what they return is current; a device pointer taken
 * before the run may show older positions for them. */
#define SFP_RUN_TIMING 1 /* record HIP events around each species kernel */
/* Run each full batch of 64 steps as one replay of a captured HIP graph
 * (captured at the first such batch, again after the kernels or grids in use
 * change); ignored with SFP_RUN_TIMING.  No counterpart in the reference. */
#define SFP_RUN_GRAPH 2
/* Online launch-shape tuner: a batch whose steps alternate between candidate
 * persistent grids of each kernel, timed with events, picks the fastest; the
 * tuner runs such a batch again when the work of a kernel changes by more
 * than a factor of four.  Results do not depend on the grid.  No counterpart
 * in the reference. */
#define SFP_RUN_AUTOTUNE 4
int sfp_sim_run(void* sim, double* t, uint64_t step0, int64_t nsteps, int32_t flags);
/* Pin the persistent grid (workgroups of 256 threads) of one species' sweep
 * (phase 0), QED (1) or emission (2) kernel; 0 returns it to the tuner, whose default
 * is full occupancy.  No counterpart in the reference. */
int sfp_sim_set_grid(void* sim, int32_t species, int32_t phase, int32_t grid);
/* Pin the grid and the most rows of 256 particles per tile (1..32 for the
 * sweep kernels, 1..16 for the QED kernels, clamped to the kernel's own
 * most; 0: the kernel's most) of one kernel; grid 0 (with rows 0) returns
 * both to the tuner, which tries, per grid, the kernel's most rows, half and
 * a quarter of it.  Results do not depend on the shape.  No counterpart in
 * the reference. */
int sfp_sim_set_shape(void* sim, int32_t species, int32_t phase, int32_t grid, int32_t rows);
/* Grid in use, workgroup size, most tile rows (0 for a kernel without row
 * tiles; rows may be NULL) and number of candidate shapes of a kernel (all 0
 * for a species without that kernel). */
int sfp_sim_launch_shape(void* sim, int32_t species, int32_t phase, int32_t* grid,
                         int32_t* tile, int32_t* rows, int32_t* ncand);
/* Accumulated time (ms) and launch count of one species' step kernel since
 * the last reset (requires SFP_RUN_TIMING; synchronous).  phase 0 is the
 * sweep kernel (envelope bound, push of particles that cannot undergo a QED
 * event, QED list), phase 1 the QED kernel over that list, phase 2 the
 * emission kernel of leptons with the learned photon-energy sampler (photon
 * energy of the emission candidates, append, recoil and push). */
int sfp_sim_kernel_time(void* sim, int32_t species, int32_t phase, double* ms,
                        int64_t* launches);
int sfp_sim_reset_timing(void* sim);
/* With SFP_RUN_TIMING, time only the steps whose index `every` divides
 * (default 1: every step); the two events around a launch cost about 2% of
 * the step time of a 1e7-electron collision, so a sample (every coprime to
 * the photon list's 16-step rebuild period samples every phase of it) keeps
 * the timed run closer to an untimed one.  No counterpart in the reference
 * (instrumentation). */
int sfp_sim_set_timing_every(void* sim, int64_t every);
/* With SFP_RUN_TIMING, time one kernel -- species, phase (0 sweep, 1 QED,
 * 2 emission) -- at every step besides the steps sfp_sim_set_timing_every
 * selects for all kernels (species < 0: none): the mean launch time of the
 * kernel a caller prices over a whole run, at two events per step instead
 * of two per launch.  No counterpart in the reference (instrumentation). */
int sfp_sim_set_timing_kernel(void* sim, int32_t species, int32_t phase);
/* The stable ranks of appended particles come from a decoupled look-back
 * over per-tile flag words tagged with a 29-bit launch epoch.  The library
 * zeroes the flags and epochs between batches every `steps` steps (default
 * 2^28, at most 2^28), so that an epoch never wraps onto a stale flag; a
 * smaller interval only exercises that clearing (tests).  No counterpart in
 * the reference. */
int sfp_sim_set_flag_clear_every(void* sim, int64_t steps);
/* Shared QED launches (on by default): two consecutive lepton species with
 * the same QED kernels and photon store -- electrons and positrons -- run in
 * one launch of the lepton QED kernel (and of the emission kernel), the
 * second species' photons appended after all of the first's, as with one
 * launch per species (on == 0), whose results are bit-identical.  The second
 * species' QED launch shapes are then unused.  No counterpart in the
 * reference (simulation.py:126-137 steps one species after the other). */
int sfp_sim_set_shared_launches(void* sim, int32_t on);
/* Ensemble of independent shots (no counterpart in the reference): shot s has
 * its own field, the nterms[s] terms starting at terms + sum_{j<s} nterms[j].
 * Every particle of a species with a SFP_SHOT attribute moves in the field of
 * its shot, and particles it creates inherit the shot.  Draws stay keyed by
 * particle id, so shots are independent streams. */
int sfp_sim_set_shots(void* sim, const sfp_field* terms, const int32_t* nterms, int32_t nshots);

/* Device-side trajectory recorder (the callback pattern of
 * example/trajectory.py:62-70 without a host round trip per step): after
 * every `every`-th step counted from step0, x, y, z, ux, uy, uz of particles
 * [0, n) of a lepton species are stored on the device ([max_rows][n][6],
 * NaN for slots past the species' count) with the time after that step.
 * n = 0 removes the recorder.  Runs as one small kernel at the end of each
 * step, also inside graph replays. */
int sfp_sim_set_recorder(void* sim, int32_t species, int64_t n, int64_t every, int64_t max_rows,
                         uint64_t step0);
/* rows recorded so far (synchronous); t_host [rows], rows_host [rows][n][6]
 * (either may be NULL to query the count) */
int sfp_sim_get_records(void* sim, double* t_host, double* rows_host, int64_t* rows);

/* Use the learned photon-energy sampler (sfp_mlp_*) for the photon emission
 * of the leptons, or stop using it (mlp NULL).  Mixed precision only; needs
 * the optical-depth tables.  The network must outlive its use by the
 * simulation.  No counterpart in the reference. */
int sfp_sim_set_sampler(void* sim, void* mlp);
/* Use the learned spin-resolved emission rate (a network from
 * sfp_mlp_create_rate) in place of the rate table for the photon emission of
 * spin-carrying leptons, or stop using it (NULL).  Mixed precision,
 * rejection sampling and spin_emission only; bins of log10 chi flagged for
 * fallback, and chi outside the network's range, use the table.  No
 * counterpart in the reference. */
int sfp_sim_set_rate_net(void* sim, void* mlp);

/* Counters of one species since the last reset (synchronous), stats[15]:
 * [0] particle-steps (live particles after each step, the reference's
 *     ns/particle denominator, simulation.py:150-151)
 * [1] particles read by the sweep kernel   [2] particles appended by the species
 * [3] current buffered count               [4] particles on QED lists (sum over steps)
 * [5] host synchronisations of the run     [6] replays after a capacity halt
 * [7] current store capacity
 * [8] photon-steps inside multi-step pushes (the sweep reads only a 4-byte
 *     wake step for these)                 [9] multi-step pushes started
 * [10] batches run as a graph replay       [11] graph captures
 * [12] tuning batches (10-12: of the simulation, not the species)
 * [13] emission candidates handed to the emission kernel (sum over steps)
 * [14] run calls that restarted the photon sleep clock (every photon visited
 *      at their first step): the first one, and those that do not continue
 *      the previous call or follow a change of a store from the host */
int sfp_sim_stats(void* sim, int32_t species, int64_t* stats);
/* The same counters and more, n of them (at most SFP_SIM_NSTATS):
 * [0]-[14] as sfp_sim_stats, [1] counting particles read by the QED kernel
 *      on dense steps
 * [15] dense steps: steps on which the QED kernel visited every lepton by
 *      index, the sweep standing down (taken when the last step's culling
 *      bound listed at least 7/8 of the particles; results do not change)
 * [16] photons: steps on which conversions found no room in the pair
 *      stores and were made after the stores grew (deferred conversions;
 *      results do not change)
 * [17] run calls that kept the photon visit lists and window caches of the
 *      call before (no rebuild at their first step: they continue it, no
 *      store changed from the host and no other simulation stepped the
 *      stores in between; results do not change) */
#define SFP_SIM_NSTATS 18
int sfp_sim_stats_n(void* sim, int32_t species, int64_t* stats, int32_t n);
/* Count and next id of one species as the last sfp_sim_run left them (read
 * by the run at its end: no device access; the same values
 * sfp_store_get_count returns until the store changes again).  No
 * counterpart in the reference. */
int sfp_sim_get_counts(void* sim, int32_t species, int64_t* n, uint64_t* next_id);
/* Diagnostics of the exact QED evaluations of one species since the last
 * reset (synchronous; no counterpart in the reference, whose behaviour in
 * these cases is undefined or unreported):
 *  counts[0] rejection-sampling evaluations whose acceptance probability p
 *            exceeded 1 (a step samples at most one event, as in
 *            qed/rejection_sampling.py:78-79,96-97, so such steps
 *            under-sample)
 *  counts[1] photon-emission evaluations with z beyond the table's end (rate
 *            taken as 0; the reference reads past the table,
 *            qed/rejection_sampling_tables.py:40-51)
 *  counts[2] optical-depth evaluations with chi above the table (clamped to
 *            its top node; the reference reads past it,
 *            qed/optical_depth_tables.py:39-40)
 *  counts[3] evaluations with chi NaN (E + v x B cancelled below rounding;
 *            no event)
 *  *p_max    the largest p evaluated (learned sampler: the step's event
 *            probability W dtau).
 * Particles the culling bound proves event-free are not evaluated, and a
 * photon ruled out by its own draw has p < 1/2. */
int sfp_sim_qed_diag(void* sim, int32_t species, int64_t* counts, double* p_max);
int sfp_sim_reset_stats(void* sim);

/* ---------------------------------------------------------------------------
 * 4. Learned QED models (no counterpart in the reference)
 *
 * Photon-energy sampler: an MLP (log10 chi, r) -> log delta,
 * 2 -> 64 -> 64 -> 64 -> 1 with SiLU, BF16 matrix-core products with FP32
 * accumulation, trained on the GPU on samples of the optical-depth
 * photon-emission inverse CDF (the table of qed/optical_depth_tables.py:59-101,
 * which it replaces in the mixed-precision step and which stays its fallback
 * per bin of log10 chi).
 * Parameters: SFP_MLP_NPARAM floats, W1[64][2] b1 W2[64][64] b2 W3[64][64] b3
 * W4[64] b4; norm[4] = {lc_min, lc_max, mu, sd}: inputs (log10 chi - centre) /
 * half-width and 2 r - 1, output (log delta - mu) / sd.
 *
 * Spin-resolved emission rate: the same network with three inputs
 * (log10 chi, r1, sigma = s.zeta) -> log(f + 1e-6), f = 3 r1^2 w(chi, r1^3,
 * sigma) / (alpha m c^2 / hbar) the rejection-sampling acceptance density of
 * qed/rejection_sampling.py:68-85 with the spin-resolved rate;
 * SFP_MLP_NPARAM_RATE parameters, W1[64][3]; inputs as above plus
 * sigma unscaled.  Created by sfp_mlp_create_rate; every sfp_mlp_* call below
 * takes either kind (x arrays have 2 or 3 columns), except sfp_mlp_sample
 * (sampler) and sfp_mlp_rate (rate).
 * ------------------------------------------------------------------------- */
#define SFP_MLP_NPARAM 8577
#define SFP_MLP_NPARAM_RATE 8641
int sfp_mlp_create(void** mlp);
int sfp_mlp_create_rate(void** mlp);
/* number of inputs (2 or 3) and of parameters of a network */
int sfp_mlp_shape(void* mlp, int32_t* nin, int32_t* nparam);
int sfp_mlp_destroy(void* mlp);
int sfp_mlp_set_params(void* mlp, const float* params, const float* norm);
int sfp_mlp_get_params(void* mlp, float* params, float* norm);
/* bit i set: bin i of 64 equal bins of log10 chi over [lc_min, lc_max] uses the table */
int sfp_mlp_set_fallback(void* mlp, uint64_t mask);
/* n samples: x_dev [n][2] = (log10 chi ~ U(lc_min, lc_max), r ~ U(0,1)) from
 * Philox(seed), y_dev [n] = log delta of the table's inverse CDF; for the rate
 * network x_dev [n][3] adds sigma ~ U(-1, 1) and y_dev is log f of the
 * spin-resolved rate table (synchronous) */
int sfp_mlp_make_data(void* mlp, void* tables, int64_t n, uint64_t seed, float* x_dev,
                      float* y_dev);
/* `steps` Adam steps on batches of `batch` samples (a multiple of 32) taken in
 * order from (x_dev, y_dev), cosine-decayed learning rate from lr; the mean
 * squared error of each step goes to loss_curve_host[steps] (may be NULL) */
int sfp_mlp_train(void* mlp, const float* x_dev, const float* y_dev, int64_t n, int32_t steps,
                  int32_t batch, float lr, float* loss_curve_host);
/* gradient of the mean squared error over n samples (n a multiple of 32)
 * from the training kernels, to the host (synchronous) */
int sfp_mlp_grad(void* mlp, const float* x_dev, const float* y_dev, int64_t n, float* grad_host,
                 float* loss_host);
/* normalised network output for raw inputs x_dev [n][2] */
int sfp_mlp_forward(void* mlp, const float* x_dev, float* out_dev, int64_t n);
/* delta for raw inputs x_dev [n][2] by the network (method 0) or the table
 * inverse CDF (method 1); *ms = kernel time (synchronous) */
int sfp_mlp_sample(void* mlp, void* tables, const float* x_dev, int64_t n, int32_t method,
                   double* delta_dev, float* ms);
/* f = 3 r1^2 w(chi, r1^3, sigma) / (alpha m c^2 / hbar) for raw inputs x_dev
 * [n][3] by the rate network (method 0, FP32) or the spin-resolved rate table
 * (method 1, FP64); *ms = kernel time (synchronous) */
int sfp_mlp_rate(void* mlp, void* tables, const float* x_dev, int64_t n, int32_t method,
                 double* f_dev, float* ms);
/* device copy of the network and of its fallback bins, for the simulation */
int sfp_mlp_device(void* mlp, const void** dev, const void** bins);

/* dst[i] = src[i] for n doubles: a plain streaming copy kernel used to
 * calibrate HBM byte counters (no counterpart in the reference) */
int sfp_copy_f64(const double* src_dev, double* dst_dev, int64_t n, sfp_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SFP_H */
