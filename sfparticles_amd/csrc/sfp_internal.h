// sfp_internal.h -- declarations shared by the translation units of libsfp_hip.so.
#ifndef SFP_INTERNAL_H
#define SFP_INTERNAL_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sfp_device.h"
#include "sfp_device32.h"

// errors: set the thread's last-error message and return the code
int sfp_set_err(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
#define set_err sfp_set_err

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err(SFP_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                              \
  } while (0)

#define CHK(expr)          \
  do {                     \
    int rc_ = (expr);      \
    if (rc_) return rc_;   \
  } while (0)

static inline hipStream_t S(sfp_stream_t s) { return (hipStream_t)s; }

#define KLOOP(n) for (int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ip < (n); ip += (int64_t)gridDim.x * blockDim.x)

static constexpr int OPB = 256;

static inline int op_grid(int64_t n) {
  int64_t g = (n + OPB - 1) / OPB;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

__device__ inline int wave_lane() { return (int)__lane_id(); }

// Device memory pointers held in structs the kernels read (the step plan):
// in the device pass they carry the global address space, so that accesses
// through them compile to global loads and stores.  A pointer read from
// memory is otherwise generic, and generic (flat) accesses also count against
// the LDS counter, so every LDS table lookup would wait for the state loads
// in flight.  The host pass sees plain pointers (same size and layout).
#ifdef __HIP_DEVICE_COMPILE__
#define SFP_G __attribute__((address_space(1)))
#define SFP_L __attribute__((address_space(3)))  // LDS, for pointers to __shared__ arrays
#else
#define SFP_G
#define SFP_L
#endif
// a device allocation as the plan holds it (host code)
template <class T>
static inline SFP_G T* gptr(T* p) {
  return (SFP_G T*)p;
}

// QED tables in device memory
struct Tables {
  sfp_rs_node* rs_nodes = nullptr;  // [z_N-1]
  float4* rs_nodes32 = nullptr;     // [z_N-1] the same in FP32 (sfp_rs_node32)
  // decreasing envelopes of the table rates over brackets j.. (see sfp_tables_create)
  double2* rs_env = nullptr;        // [z_N-1] {E1, E2}
  // spin-resolved emission: Ai per bracket {value, slope}, in FP64 and FP32,
  // and the envelopes with the spin term added to E2
  double2* rs_ai = nullptr;
  float2* rs_ai32 = nullptr;
  double2* rs_env_spin = nullptr;
  int64_t z_N = 0;
  double z_min = 0, z_max = 0, z_delta = 0;
  bool has_rs = false;
  double* od_mem = nullptr;
  sfp_od_dev od{};
  double2* od_photon_nodes = nullptr;  // [chi_N-1] {value, slope}
  double2* od_pair_nodes = nullptr;
  bool has_od = false;
};

// per-species device counters
struct Counts {
  int64_t n;          // buffered particles (live + pruned holes)
  int64_t next_id;    // id of the next appended particle
  int64_t n_old;      // n at the start of the current step
  int64_t overflow;   // appends that did not fit (per-operation path)
  int64_t psteps;     // particle-steps advanced (live after each step)
  int64_t processed;  // particles read by this species' sweep kernel
  int64_t spawned;    // particles this species appended (per destination)
  int64_t n_hot;      // particles on this step's QED list
  int64_t hot_total;  // sum of n_hot over steps
  int64_t slept;      // particle-steps skipped inside a multi-step photon push
  int64_t sleeps;     // multi-step photon pushes started
  int64_t list_end;   // photons: n_old when the visit list was last built or extended
  int64_t asleep_n;   // photons: live photons inside a multi-step push
  int64_t n_emit;     // emission candidates handed to the emission kernel this step
  int64_t emit_total; // sum of n_emit over steps
  int64_t cpar;       // photons: which of the two window caches holds this window
  // leptons, dense steps: 1 when this step's QED kernel visits every particle
  // itself (no sweep list); per step parity, the particles a step's bound
  // listed, which decides the next step's mode; dense steps run
  int64_t dense;
  int64_t dense_listed[2];
  int64_t dense_steps;
  // exact QED evaluations (diagnostics): rejection-sampling probabilities
  // above 1, photon-emission z beyond the table, optical-depth chi above the
  // table, NaN chi; the largest probability as the bits of a positive float
  int64_t p_over1;
  int64_t z_clamp;
  int64_t chi_clamp;
  int64_t chi_nan;
  int64_t p_max_bits;
  // photons: conversions of this step that found no room in the pair stores
  // (left unconverted, marked PRUNED_DEFER, converted by the deferred-append
  // kernel once the host has grown the stores), the first list position of
  // one as ~position (0: none), and the recoveries that did so
  int64_t defer_n;
  uint64_t defer_kneg;
  int64_t deferrals;
};

// pruned byte of a photon whose conversion waits for room in the pair stores
static constexpr uint8_t PRUNED_DEFER = 2;

extern const size_t kAttrSize[SFP_NATTR];

struct Store {
  int32_t flags = 0;
  int64_t cap = 0;
  // bumped by every host-side change of the contents (upload, prune, a new
  // count, a writable pointer handed out): a simulation keeps photon sleep
  // state across run calls only while it stays the same
  uint64_t version = 0;
  // run calls that stepped the store (any simulation): a simulation keeps
  // its photon visit lists across its own consecutive run calls only
  uint64_t runs = 0;
  void* p[SFP_NATTR] = {};
  Counts* cnt = nullptr;  // device
  // A simulation whose sleeping photons hold the positions of the steps they
  // fell asleep at: settle(sim) brings them to the last step run.  Every host
  // access to the store calls it first (store_settle).
  int (*settle)(void* sim) = nullptr;
  void* settle_sim = nullptr;
};

static inline int store_settle(Store* s) {
  if (!s->settle) return 0;
  int (*f)(void*) = s->settle;
  void* sim = s->settle_sim;
  s->settle = nullptr;
  s->settle_sim = nullptr;
  return f(sim);
}

int sfp_compile_fields(const sfp_field* terms, int32_t nterms, sfp_fieldset* fs);
int sfp_compact_indices(const uint8_t* f, int64_t n, int invert, int64_t* out, int64_t* count_host,
                        hipStream_t st);

// one optical-depth update (optical_depth.py:42-62); returns the event flag
__device__ inline bool od_update(double* tau, double ig, double chi, double dt, uint64_t seed,
                                 uint64_t id, uint64_t step, uint32_t tag, uint32_t slot,
                                 const double2* nodes, const double* cdf, const sfp_od_dev& od,
                                 double* delta) {
  // total rate (optical_depth_tables.py:46-56) from {value, slope} nodes
  double rate;
  {
    double log_chi = sfp_log10(chi);
    if (log_chi > od.log_chi_max) log_chi = od.log_chi_max;
    if (log_chi < od.log_chi_min) {
      rate = 0.0;
    } else {
      double fidx = (log_chi - od.log_chi_min) / od.log_chi_delta;
      int64_t idx = (int64_t)sfp_floor(fidx);
      if (idx > od.chi_N - 2) idx = od.chi_N - 2;
      double left = od.log_chi_min + (double)idx * od.log_chi_delta;
      double2 nd = nodes[idx];
      rate = nd.x + nd.y * (log_chi - left);
    }
  }
  double dtau = dt * ig;
  double t = *tau;
  if (t == 0.0) {
    sfp_u32x4 d = sfp_draw4(seed, id, step, tag, slot + 1);
    t = -sfp_log(1 - sfp_u01(d.v[0], d.v[1]));
  }
  t -= rate * dtau;
  bool ev = false;
  if (t < 0) {
    sfp_u32x4 d = sfp_draw4(seed, id, step, tag, slot + 2);
    t = -sfp_log(1 - sfp_u01(d.v[0], d.v[1]));
    ev = true;
    *delta = sfp_od_inverse_cdf(cdf, &od, chi, sfp_u01(d.v[2], d.v[3]));
  } else {
    *delta = 0.0;
  }
  *tau = t;
  return ev;
}

#endif  // SFP_INTERNAL_H
