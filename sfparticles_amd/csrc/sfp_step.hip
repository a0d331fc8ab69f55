// sfp_step.hip -- libsfp_hip.so, part 2: the fused simulation step
// (sfparticles/simulation.py:102-144).
//
// Every species advances with two persistent kernels per step:
//
//  sweep  -- one pass over the species' store.  For each live particle an
//            upper bound of chi from the field envelopes (no carrier phase,
//            no Gouy trigonometry) decides whether a QED event is possible at
//            all this step.  Particles that provably cannot emit (or
//            convert) are pushed right away; the others are appended, in
//            ascending index order, to the species' QED list.
//  hot    -- over the QED list only: field at t, chi, Philox draw, table
//            rate, hard-photon cut; the stable rank of every event comes from
//            a wave64 ballot + workgroup scan + decoupled look-back across
//            tiles; new particles are appended at (count + rank), already
//            pushed (the reference pushes particles in the step that creates
//            them); then recoil and push.
//
// The culling is exact: a culled particle's exact QED evaluation could only
// have produced "no event", so results are bit-identical to evaluating every
// particle (the CPU oracle does exactly that).
//
// Capacity: counts stay on the device.  Before appending, the hot kernel
// checks that every destination has room for one new particle per listed
// source particle; if not, it raises a device-side halt flag and every later
// kernel returns at entry without touching state.  The host notices the flag
// at its next synchronisation, grows the destination store and replays the
// step from the halted kernel.  No worst-case reservation and no per-step
// host synchronisation are needed.

#include <cstring>

#include "sfp_internal.h"
#include "sfp_mlp.h"

static constexpr int TILE = 256;
// Photons fall asleep and wake on steps that are multiples of WIN (counted
// from the start of a run); the photon visit list is rebuilt on those steps
// and kept (and extended) in between.
static constexpr uint32_t WIN = 16;
static constexpr int WAVES = TILE / 64;
// kernels of a species and step: 0 sweep, 1 QED, 2 emission (leptons with the
// learned photon-energy sampler)
static constexpr int NPH = 3;

// An emission candidate of a lepton whose photon energy is sampled by the
// emission kernel: particle index, whether the network gives delta (else the
// table value `delta`), the network's normalised inputs, chi and the table's
// uniform (for the hard tail and fallback bins).
struct EmitRec {
  uint32_t i;
  uint32_t net;
  float in0, in1;
  double chi, u_tab, delta;
  // the candidate's state as the mixed step uses it (FP64 position, FP32
  // momentum and spin), so that the emission kernel reads records in order
  // instead of gathering scattered store lines
  double x, y, z;
  float ux, uy, uz;
  float sx, sy, sz;
};
#define SFP_MAX_SPECIES 8

enum { KIND_RS = 0, KIND_OD = 1 };

struct SpeciesDev {
  SFP_G double *x, *y, *z, *ux, *uy, *uz;
  SFP_G uint64_t* id;
  SFP_G uint8_t* pruned;
  SFP_G double* tau;
  SFP_G double *sx, *sy, *sz;
  SFP_G uint32_t* hot;   // this step's QED list (ascending particle indices)
  SFP_G uint32_t* wake;  // photons: step of the next visit (multi-step push), or null
  SFP_G uint32_t* since;  // QED photons asleep: the step whose position the store holds
  SFP_G uint32_t* shot;  // ensemble: the shot each particle belongs to (selects its field), or null
  SFP_G EmitRec* emit;   // leptons with the learned sampler: this step's emission candidates
  // photons: state of the first `ccap` entries of the visit list, by list
  // position, for the window the list is kept; two sets (win_cache), this
  // window's and the last one's (see k_photon)
  SFP_G double* cache;
  int64_t ccap;
  SFP_G uint32_t* slot;  // photons: position in the visit list it was last put on
  SFP_G uint32_t* oldk;  // photons, per list entry at a rebuild: position in the last window's cache, or none
  SFP_G Counts* cnt;
  int64_t cap;
  double q, ae;
  int32_t rr, push, spin, photon, tag, qed;
  int32_t photon_dst, pair_e_dst, pair_p_dst, pad;
};

// One set of the photon window cache: x, y, z, ux, uy, uz, the wake step and
// the id of the first n list entries, as 8 rows of n words (set `par` of the
// two).  The id is the key of a photon's draws: from the cache it loads
// coalesced, from the store it is a gather of one word per 64-byte line
// wherever awake photons lie sparse among sleeping ones.
static constexpr int WIN_WORDS = 8;
struct WinCache {
  SFP_G double* b;
  int64_t n;
  __device__ SFP_G double& x(int64_t k) const { return b[k]; }
  __device__ SFP_G double& y(int64_t k) const { return b[n + k]; }
  __device__ SFP_G double& z(int64_t k) const { return b[2 * n + k]; }
  __device__ SFP_G double& ux(int64_t k) const { return b[3 * n + k]; }
  __device__ SFP_G double& uy(int64_t k) const { return b[4 * n + k]; }
  __device__ SFP_G double& uz(int64_t k) const { return b[5 * n + k]; }
  __device__ SFP_G uint32_t& wake(int64_t k) const { return *(SFP_G uint32_t*)(b + 6 * n + k); }
  __device__ SFP_G uint64_t& id(int64_t k) const { return *(SFP_G uint64_t*)(b + 7 * n + k); }
};
__device__ inline WinCache win_cache(const SpeciesDev& sp, uint32_t par) {
  return WinCache{sp.cache + (int64_t)(par & 1u) * WIN_WORDS * sp.ccap, sp.ccap};
}

struct ScanState {
  uint32_t tile_ctr;
  uint32_t done_ctr;
  uint32_t epoch;
  uint32_t pad;
};

struct Halt {
  int32_t flag;
  int32_t species;
  int64_t step;
};

// Device-side trajectory recorder: after every `every`-th step from step0,
// x, y, z, ux, uy, uz of particles [0, n) of one species go to row r of buf
// ([max_rows][n][6]) and the time to t[r].
struct Recorder {
  SFP_G double* buf;
  SFP_G double* t;
  int64_t n, every, max_rows;
  uint64_t step0;
  int32_t species, on;
};

struct Plan {
  SpeciesDev sp[SFP_MAX_SPECIES];
  SFP_G ScanState* scan[SFP_MAX_SPECIES][NPH];  // [species][0 sweep list, 1 event ranks, 2 emissions]
  SFP_G uint64_t* flags[SFP_MAX_SPECIES][NPH];
  sfp_fieldset fields;
  const SFP_G sfp_rs_node* rs_nodes;
  const SFP_G sfp_rs_node32* rs_nodes32;  // the same nodes in FP32 (mixed precision)
  const SFP_G double2* rs_env;  // rate envelopes (Tables::rs_env)
  const SFP_G double2* rs_ai;        // spin-resolved emission: Ai nodes, FP64 and FP32,
  const SFP_G float2* rs_ai32;       // and the envelopes with the spin term
  const SFP_G double2* rs_env_spin;
  int32_t spin_emission;       // leptons with spin use the spin-resolved rate (RS)
  double rate_unit;       // alpha m c^2 / hbar
  int64_t z_N;
  double z_min, z_max, z_delta;
  const SFP_G double2* od_photon_nodes;
  const SFP_G double2* od_pair_nodes;
  sfp_od_dev od;
  double threshold;
  double chi_coef;  // e hbar / (m^2 c^3)
  double zmax15;    // z_max^1.5: rate is 0 unless delta/(1-delta) <= zmax15 chi (photon)
  uint64_t seed;
  SFP_G Halt* halt;
  const SFP_G sfp_fieldset* shot_fields;  // ensemble: one field set per shot, or null
  uint32_t nshots;
  SFP_G const sfp_mlp_dev* net;        // learned photon-energy sampler, or null
  SFP_G const sfp_mlp_bins* net_bins;  // its per-bin table fallback
  SFP_G const sfp_mlp_dev* rate_net;   // learned spin-resolved emission rate, or null
  SFP_G const sfp_mlp_bins* rate_bins;
  Recorder rec;
  int32_t nspecies;
  int32_t qed;
};

struct StepArgs {
  double t, t_half, dt, cdt_half, cdt, ll_factor;
  uint64_t step;
  uint32_t rel, rel_end;  // step index within this run call, and the run's length
  // step on the photon sleep clock (steps since the simulation's wake origin):
  // QED photons sleep on it, across run calls that continue each other;
  // fresh: the first step of a run call whose photon visit lists are rebuilt
  // (the lists and window caches of the last call are not kept)
  uint32_t wrel, fresh;
};

// --------------------------------------------------------- look-back scan --
// Flag words carry their own payload and no other data is exchanged between
// workgroups inside a launch, so relaxed agent-scope atomics suffice: they
// are coherent across the XCDs' L2s without the cache write-back/invalidate
// that acquire/release ordering costs on every publish.
__device__ inline uint64_t ld_flag(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_flag(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// flag word: [epoch:29][status:2][value:33]; status 1 = aggregate, 2 = inclusive
// prefix.  Values are counts of at most a store's capacity (< 2^32).  The
// epoch advances once per launch of a scanning kernel and is never 0; the
// host clears the flags and epochs long before it could wrap (sim_clear_flags),
// so a flag of an earlier launch can never carry the current epoch.
static constexpr int FLAG_VAL_BITS = 33;
static constexpr int FLAG_EPOCH_SHIFT = FLAG_VAL_BITS + 2;
static constexpr uint32_t FLAG_EPOCH_MASK = (1u << (64 - FLAG_EPOCH_SHIFT)) - 1u;
static constexpr uint64_t FLAG_VAL_MASK = (1ull << FLAG_VAL_BITS) - 1ull;

__device__ inline uint64_t flag_word(uint32_t epoch, uint32_t status, uint64_t v) {
  return ((uint64_t)epoch << FLAG_EPOCH_SHIFT) | ((uint64_t)status << FLAG_VAL_BITS) |
         (v & FLAG_VAL_MASK);
}

// Decoupled look-back (one full wave): publish the tile aggregate, accumulate
// predecessors until an inclusive prefix is found, publish the inclusive
// prefix and return the exclusive prefix (uniform across the wave).
__device__ uint64_t lookback(uint64_t* flags, int64_t tile, uint64_t agg, uint32_t epoch) {
  int lane = wave_lane();
  if (tile == 0) {
    if (lane == 0) st_flag(&flags[0], flag_word(epoch, 2, agg));
    return 0;
  }
  if (lane == 0) st_flag(&flags[tile], flag_word(epoch, 1, agg));
  uint64_t prefix = 0;
  int64_t j = tile - 1;
  while (true) {
    int64_t idx = j - lane;
    uint64_t w;
    uint32_t st;
    while (true) {
      w = (idx < 0) ? flag_word(epoch, 2, 0) : ld_flag(&flags[idx]);
      st = ((uint32_t)(w >> FLAG_EPOCH_SHIFT) == epoch) ? (uint32_t)((w >> FLAG_VAL_BITS) & 3u) : 0u;
      if (__all(st != 0)) break;
      __builtin_amdgcn_s_sleep(1);
    }
    uint64_t pm = __ballot(st == 2);
    uint64_t v = w & FLAG_VAL_MASK;
    int stop = pm ? (int)__builtin_ctzll(pm) : 63;
    uint64_t contrib = (lane <= stop) ? v : 0;
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
    prefix += contrib;
    if (pm) break;
    j -= 64;
  }
  if (lane == 0) st_flag(&flags[tile], flag_word(epoch, 2, agg + prefix));
  return prefix;
}

// A workgroup-uniform value read from LDS, marked uniform for the compiler
// (scalar registers): what is derived from it -- a species index, its
// store's pointers -- then stays in scalar registers and scalar loads.
__device__ inline int64_t wg_uniform(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// A table pointer of the step plan (generic in the table structs the CPU
// restatement shares) in the global address space: loads through it are
// global loads, which unlike generic (flat) ones leave the LDS counter alone.
__device__ inline SFP_G const double* gtab(const double* p) { return (SFP_G const double*)p; }

// The total of a look-back chain of ntiles tiles: its last tile's inclusive
// prefix, waited for until that tile has published it in this epoch.  Safe
// to wait on from a workgroup that has no tile of the chain left to do: every
// tile of the chain is then held by a running workgroup, whose own look-back
// waits only on tiles before it.
__device__ uint64_t chain_total(const uint64_t* flags, int64_t ntiles, uint32_t epoch) {
  if (ntiles <= 0) return 0;
  while (true) {
    const uint64_t w = ld_flag(&flags[ntiles - 1]);
    if ((uint32_t)(w >> FLAG_EPOCH_SHIFT) == epoch && ((w >> FLAG_VAL_BITS) & 3u) == 2u)
      return w & FLAG_VAL_MASK;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Stable rank of `pred` among all lanes of all tiles processed so far:
// exclusive prefix over (tile, wave, lane).  Called by every thread.

__device__ inline int64_t tile_rank(bool pred, int64_t tile, uint64_t* flags, uint32_t epoch,
                                    uint32_t* sh_wcnt, uint64_t* sh_pref) {
  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  uint64_t m = __ballot(pred);
  uint32_t lrank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) sh_wcnt[wv] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t woff = 0, agg = 0;
  for (int k = 0; k < WAVES; ++k) {
    uint32_t c = sh_wcnt[k];
    if (k < wv) woff += c;
    agg += c;
  }
  if (wv == 0) {
    uint64_t pre = lookback(flags, tile, agg, epoch);
    if (lane == 0) *sh_pref = pre;
  }
  __syncthreads();
  return (int64_t)(*sh_pref) + woff + lrank;
}

// Rows of 256 list entries per tile of the QED kernels.  A tile's events are
// ranked with one look-back after all of its rows and then appended from a
// workgroup queue; an event leaves its particle untouched in the row pass and
// is recomputed from that unchanged state when it is appended, so a
// look-back (a round trip through the device-coherent flags, several
// microseconds) is paid once per tile, not once per row.  Short lists take
// fewer rows per tile, so that every workgroup of the grid has a tile.  The
// launch passes the most rows a tile may take (1..QROWS, a choice of the
// launch-shape tuner; results do not depend on it).
static constexpr int QROWS = 16;
static_assert(QROWS * TILE <= 65536, "event queues hold 16-bit tile positions");

__device__ inline int qed_rows(int64_t n_list, int rcap) {
  // the fewest rounds of tiles (one tile per workgroup per round) that
  // rcap-row tiles need, then the fewest rows that fill that many rounds: a
  // list of 1.5 rounds of 16-row tiles would leave half the workgroups idle in
  // the second; in 12-row tiles it fills two
  rcap = rcap < 1 ? 1 : (rcap > QROWS ? QROWS : rcap);
  const int64_t per_round = (int64_t)gridDim.x * TILE;
  const int64_t m = (n_list + per_round * rcap - 1) / (per_round * rcap);
  if (m < 1) return 1;
  const int64_t r = (n_list + per_round * m - 1) / (per_round * m);
  return r < 1 ? 1 : (r > rcap ? rcap : (int)r);
}

// After the rows of a tile have left their event ballots in mask[row][wave]:
// the tile's event count goes through the look-back (exclusive prefix to
// *sh_pref), and the tile positions (row * TILE + thread) of its events go to
// q in (row, wave, lane) order, i.e. list order.  Returns the event count.
// Called by every thread; q and *sh_pref are visible on return.
__device__ inline uint32_t tile_events(const uint64_t (*mask)[WAVES], int rows, uint64_t* flags,
                                       int64_t tile, uint32_t epoch, uint16_t* q,
                                       uint64_t* sh_pref) {
  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  __syncthreads();
  if (wv == 0) {
    uint64_t c = 0;
    for (int k = lane; k < rows * WAVES; k += 64) c += (uint64_t)__popcll((&mask[0][0])[k]);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    const uint64_t pre = lookback(flags, tile, c, epoch);
    if (lane == 0) *sh_pref = pre;
  }
  const uint64_t lt = (1ull << lane) - 1ull;
  uint32_t off = 0;
  for (int row = 0; row < rows; ++row) {
    uint32_t before = 0, total = 0;
    for (int k = 0; k < WAVES; ++k) {
      const uint32_t c = (uint32_t)__popcll(mask[row][k]);
      if (k < wv) before += c;
      total += c;
    }
    const uint64_t m = mask[row][wv];
    if ((m >> lane) & 1ull)
      q[off + before + (uint32_t)__popcll(m & lt)] = (uint16_t)((uint32_t)(row * TILE) + threadIdx.x);
    off += total;
  }
  __syncthreads();
  return off;
}

// The field set of particle i: its shot's in an ensemble (indices at or above
// the number of shots take the last shot), else the common one.
// The field set of particle i: its shot's in an ensemble (SH), else the
// simulation's, the same for every lane (its constants load into scalar
// registers once per wave; per lane they took a vector load each, and the
// lepton QED kernel ran 17 % slower).
template <bool SH>
__device__ inline const sfp_fieldset* fields_of(const Plan& P, const SpeciesDev& sp, int64_t i,
                                               uint32_t& shot) {
  if (SH) {
    shot = min(sp.shot[i], P.nshots - 1u);
    return &P.shot_fields[shot];
  }
  shot = 0;
  return &P.fields;
}

// -------------------------------------------------------- particle push --
// Push of a lepton from the first position half-step to the second
// (simulation.py:128-137 with particles.py:244-296); ig is 1/gamma on entry.
template <int FM>
__device__ inline void lepton_push(const Plan& P, const sfp_fieldset* FS, const SpeciesDev& sp,
                                   const StepArgs& a,
                                   double& x, double& y, double& z, double& ux, double& uy,
                                   double& uz, double& sx, double& sy, double& sz, double ig) {
  sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
  double F[6];
  sfp_field_eval<FM>(FS, x, y, z, a.t_half, F);
  if (sp.spin) {
    sfp_boris_tbmt(&ux, &uy, &uz, &ig, &sx, &sy, &sz, F[0], F[1], F[2], F[3], F[4], F[5], sp.q,
                   sp.ae, a.dt);
  } else {
    sfp_boris(&ux, &uy, &uz, &ig, F[0], F[1], F[2], F[3], F[4], F[5], sp.q, a.dt);
  }
  if (sp.rr == SFP_RR_LL || (sp.rr == SFP_RR_CLL && !sp.spin)) {
    double chi = sfp_chi(F[0], F[1], F[2], F[3], F[4], F[5], ux, uy, uz, ig);
    double factor = (sp.rr == SFP_RR_CLL) ? a.ll_factor * sfp_cll_g(chi) : a.ll_factor;
    sfp_ll_apply(&ux, &uy, &uz, &ig, chi, factor);
  }
  sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
}

// The same push in mixed precision: positions accumulate in FP64 from the
// FP32 momentum, field and pusher run in FP32, chi for LL/CLL in FP64.
template <int FM>
__device__ inline void lepton_push32(const Plan& P, const sfp_fieldset* FS, const SpeciesDev& sp,
                                     const StepArgs& a,
                                     double& x, double& y, double& z, float& ux, float& uy,
                                     float& uz, float& sx, float& sy, float& sz, float ig) {
  double h = a.cdt_half * (double)ig;
  x += h * (double)ux;
  y += h * (double)uy;
  z += h * (double)uz;
  float F[6];
  sfp32_field_eval<FM>(FS, x, y, z, a.t_half, F);
  const float q = (float)sp.q, dt = (float)a.dt;
  if (sp.spin) {
    ig = sfp32_boris_tbmt(ux, uy, uz, sx, sy, sz, F, q, (float)sp.ae, dt);
  } else {
    ig = sfp32_boris(ux, uy, uz, F, q, dt);
  }
  if (sp.rr == SFP_RR_LL || (sp.rr == SFP_RR_CLL && !sp.spin)) {
    const float chi = (float)sfp32_chi(F, ux, uy, uz);
    float factor = (float)a.ll_factor;
    if (sp.rr == SFP_RR_CLL) factor *= sfp32_cll_g(chi);
    ig = sfp32_ll(ux, uy, uz, ig, chi, factor);
  }
  h = a.cdt_half * (double)ig;
  x += h * (double)ux;
  y += h * (double)uy;
  z += h * (double)uz;
}

// Append one new lepton (pair member) at slot d of species `sp`, pushed.
template <int FM, bool MIX = false>
__device__ inline void spawn_lepton(const Plan& P, const sfp_fieldset* FS, uint32_t shot,
                                    const SpeciesDev& sp, const StepArgs& a, int64_t d,
                                    uint64_t id, double x, double y, double z, double ux,
                                    double uy, double uz) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (sp.push) {
    if (MIX) {
      float fx = (float)ux, fy = (float)uy, fz = (float)uz, s0 = 0.f, s1 = 0.f, s2 = 0.f;
      lepton_push32<FM>(P, FS, sp, a, x, y, z, fx, fy, fz, s0, s1, s2,
                        __frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz));
      ux = fx; uy = fy; uz = fz;
    } else {
      double ig = 1.0 / sfp_sqrt(1.0 + ux * ux + uy * uy + uz * uz);
      lepton_push<FM>(P, FS, sp, a, x, y, z, ux, uy, uz, sx, sy, sz, ig);
    }
  }
  sp.x[d] = x; sp.y[d] = y; sp.z[d] = z;
  sp.ux[d] = ux; sp.uy[d] = uy; sp.uz[d] = uz;
  sp.id[d] = id;
  sp.pruned[d] = 0;
  if (sp.tau) sp.tau[d] = 0.0;
  if (sp.spin) { sp.sx[d] = 0.0; sp.sy[d] = 0.0; sp.sz[d] = 0.0; }
  if (sp.shot) sp.shot[d] = shot;
}

// Append one new photon at slot d of species `sp`, pushed.
__device__ inline void spawn_photon(uint32_t shot, const SpeciesDev& sp, const StepArgs& a,
                                    int64_t d, uint64_t id,
                                    double x, double y, double z, double ux, double uy,
                                    double uz) {
  if (sp.push) {
    double ig = sfp_inv_gamma_photon(ux, uy, uz);
    sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
    sfp_push_position(&x, &y, &z, ux, uy, uz, ig, a.cdt_half);
  }
  sp.x[d] = x; sp.y[d] = y; sp.z[d] = z;
  sp.ux[d] = ux; sp.uy[d] = uy; sp.uz[d] = uz;
  sp.id[d] = id;
  sp.pruned[d] = 0;
  if (sp.tau) sp.tau[d] = 0.0;
  if (sp.shot) sp.shot[d] = shot;
}

// the float at or above d (FP32 copies of bound constants stay bounds)
__device__ inline float f32_up(double d) {
  float f = (float)d;
  if ((double)f < d) {
    const uint32_t b = __float_as_uint(f);
    f = f > 0.0f ? __uint_as_float(b + 1u) : (f < 0.0f ? __uint_as_float(b - 1u) : 0x1p-149f);
  }
  return f;
}

// Can a QED event happen this step?  chi_b >= chi (bound with a small margin).  Rejection sampling: an event needs r2 < p with r2 >= 2^-53 (the
// smallest uniform), p = 3 r1^2 w dtau (photon) or w dtau (pair), where the
// table rate w = (alpha m c^2/hbar) |int Ai(z) + (2/z +- chi' sqrt z) Ai'(z)|
// vanishes for z > z_max and otherwise is bounded by the envelopes at the
// smallest reachable z:
//   pair:   z = (1/(d(1-d) chi))^(2/3) >= (4/chi)^(2/3),      chi' = chi
//   photon: z = (d/((1-d) chi))^(2/3) >= (d_min/chi)^(2/3),   chi' = d chi <= chi
// with d_min = threshold / gamma (softer photons are dropped by the hard-photon
// cut, cpu.py:227-232, and leave no trace).  If the bound on p, doubled, is
// below 2^-53 no event is possible.  Optical depth: skipped below chi_min
// (optical_depth.py:43), so only that test is exact.
// Rejection sampling: the bound on p itself (< 0: no bound, evaluate; 0: no
// event possible).
__device__ inline double rs_pbound(const Plan& P, bool photon, float chi_b, double ig, double dt,
                                   bool spin, const float2* env_lds = nullptr) {
  float num;
  if (photon) {
    num = 4.0f;
  } else {
    num = (float)(P.threshold * ig) * 0.999f;
    if (!(num > 0.0f)) return -1.0;
  }
  if (!(chi_b >= 0.0f)) return -1.0;                    // NaN bound: evaluate exactly
  if ((float)P.zmax15 * chi_b < num) return 0.0;        // z > z_max for every reachable d: w = 0
  // lower bound of the smallest reachable z, rounded down
  float q = num / chi_b;
  float zlo = exp2f(log2f(q) * (2.0f / 3.0f)) * 0.999f;
  float fj = (zlo - (float)P.z_min) * (float)(1.0 / P.z_delta);
  if (!(fj >= 1.0f)) return -1.0;                       // first bracket: 2/z unbounded
  int64_t j = (int64_t)fj;
  if (j > P.z_N - 2) return 0.0;
  // from the workgroup's LDS copy where it has one (in FP32, each value
  // rounded up: still a bound, at half the LDS): a global load issued after
  // the next row's state would wait for those loads (they complete in order)
  // and undo the software pipeline
  double2 e;
  if (env_lds) {
    const float2 f = env_lds[j];
    e = make_double2((double)f.x, (double)f.y);
  } else {
    e = spin ? P.rs_env_spin[j] : P.rs_env[j];
  }
  return P.rate_unit * (e.x + (double)chi_b * e.y) * dt * ig * (photon ? 1.0 : 3.0);
}

__device__ inline bool qed_possible(const Plan& P, bool photon, float chi_b, double ig,
                                    double dt, bool spin = false,
                                    const float2* env_lds = nullptr) {
  if (P.qed == SFP_QED_OD) return !(chi_b < (float)P.od.chi_min * 0.999f);
  const double pb = rs_pbound(P, photon, chi_b, ig, dt, spin, env_lds);
  return pb < 0.0 || !(2.0 * pb < 0x1p-53);
}

// ------------------------------------------------------------ exit protocol --
// The last workgroup to finish resets the tile counter for the next launch
// and runs `last` (counts, epoch).
// Everything the last workgroup reads was produced by atomics (counters,
// look-back flags) and what it writes is consumed by later launches, so no
// memory fence is needed.
// Only the first nwg workgroups of the grid take part (see participants).
template <typename F>
__device__ inline void kernel_exit(ScanState* scan, uint32_t nwg, F&& last) {
  if (threadIdx.x == 0) {
    uint32_t prev = atomicAdd(&scan->done_ctr, 1u);
    if (prev == nwg - 1) {
      last();
      scan->tile_ctr = 0;
      scan->done_ctr = 0;
    }
  }
}

// Workgroups of a persistent launch that take part when it has `units` of
// work: no more than the units (at least one, which runs the exit protocol);
// the others return at once, before touching any shared counter, so a short
// list costs a few atomics instead of three per workgroup of the grid.
__device__ inline uint32_t participants(int64_t units) {
  return (uint32_t)(units < 1 ? 1 : (units < (int64_t)gridDim.x ? units : (int64_t)gridDim.x));
}

// Workgroup sums of N per-thread counters; every thread gets the totals.
// sh holds N * WAVES entries.
template <int N>
__device__ inline void block_sum(uint32_t (&v)[N], uint32_t* sh, uint64_t (&tot)[N]) {
  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  for (int off = 32; off > 0; off >>= 1)
    for (int j = 0; j < N; ++j) v[j] += __shfl_xor(v[j], off);
  __syncthreads();
  if (lane == 0)
    for (int j = 0; j < N; ++j) sh[j * WAVES + wv] = v[j];
  __syncthreads();
  for (int j = 0; j < N; ++j) {
    uint64_t s = 0;
    for (int k = 0; k < WAVES; ++k) s += sh[j * WAVES + k];
    tot[j] = s;
  }
}

__device__ inline void count_add(int64_t* dst, uint64_t v) {
  if (v) atomicAdd((unsigned long long*)dst, (unsigned long long)v);
}

__device__ inline uint32_t next_epoch(const ScanState* scan) {
  uint32_t e = (scan->epoch + 1u) & FLAG_EPOCH_MASK;
  return e == 0 ? 1u : e;
}

// --------------------------------------------------- multi-step photon push --
// The photon visit list is rebuilt (and photons fall asleep and wake) on the
// steps of the sleep clock that are multiples of WIN, and on the first step of
// a run call that does not continue the last one (or follows a change of the
// stores): a call that does keeps the last call's lists and window caches.
__device__ inline bool list_rebuild(const StepArgs& a) {
  return (a.wrel % WIN) == 0 || a.fresh != 0;
}

// A QED photon asleep until this step holds the position of the step it fell
// asleep at (since): the two position increments of every step it slept are
// added now -- the same sequential FP64 additions, in closed form, that the
// steps one at a time would have made (cpu.py:5-19).
__device__ inline void photon_wake_position(const SpeciesDev& sp, const StepArgs& a, int64_t i,
                                            double ig, double& x, double& y, double& z,
                                            double ux, double uy, double uz) {
  const uint32_t s0 = sp.since[i];
  if (s0 >= a.wrel) return;
  const uint32_t n = 2u * (a.wrel - s0);
  x = sfp_repeat_add(x, a.cdt_half * ig * ux, n);
  y = sfp_repeat_add(y, a.cdt_half * ig * uy, n);
  z = sfp_repeat_add(z, a.cdt_half * ig * uz, n);
}

// A photon moves on a straight line at c and its momentum never changes, so
// while no QED event is possible its push is the same two position
// increments every step.  Number of steps, from this one on, for which the
// envelope bound along the photon's own path proves that no event is possible
// (1 if only this step is proven).  Windows end at the next multiple of
// 32768, 4096, 512, 64 or WIN steps of the sleep clock (the longest proven
// one), so that neighbouring photons -- one wave's lanes -- wake on the same
// steps, and may last past the end of the run.  Photons of a species without
// QED sleep to the end of the run.
template <int FM>
__device__ inline uint32_t sleep_steps(const Plan& P, const sfp_fieldset* FS, const StepArgs& a,
                                       bool qed, double x,
                                       double y, double z, double ux, double uy, double uz,
                                       double ig) {
  if (!qed) {  // it only moves: to the end of the run
    const uint32_t rem = a.rel_end - a.rel;
    return rem <= 1 ? 1 : rem;
  }
  // QED photons, on the sleep clock: a sleep may last past the end of the run
  const double vx = SFP_C * ig * ux, vy = SFP_C * ig * uy, vz = SFP_C * ig * uz;
  const uint32_t q[5] = {32768u, 4096u, 512u, 64u, WIN};
  for (int k = 0; k < 5; ++k) {
    const uint32_t end = (a.wrel / q[k] + 1u) * q[k];
    if (end < a.wrel + 2u) continue;
    const uint32_t d = end - a.wrel;
    float S = sfp_field_bound_f32_path<FM>(FS, x, y, z, a.t, vx, vy, vz,
                                           a.dt * (double)d * 1.000001);
    float chi_b = (float)P.chi_coef * S * sfp32_div(1.0f, (float)ig) * 1.001f;
    if (!qed_possible(P, true, chi_b, ig, a.dt)) return d;
  }
  return 1;
}

// A photon under rejection sampling converts at a step only if its draw r2
// for that step (keyed by its id and the step) lies below the conversion
// probability, and the probability of every step of a window is below the
// bound pb of the envelope along the photon's straight path over the window.
// A photon whose own draws for the window's steps all lie at or above 2 pb
// (the margin the exact front's prefilter uses) therefore does not convert
// in the window: it sleeps through it as a photon out of the field does,
// bit-identically.  Steps of the window from this one on (this step's draw
// has been tested by the caller), 1 if no window is proven.
template <int FM, bool MIX>
__device__ inline uint32_t sleep_steps_drawn(const Plan& P, const sfp_fieldset* FS,
                                             const StepArgs& a, double x, double y, double z,
                                             double ux, double uy, double uz, double ig,
                                             uint64_t pid, uint32_t tag, const float2* env_lds) {
  const double vx = SFP_C * ig * ux, vy = SFP_C * ig * uy, vz = SFP_C * ig * uz;
  const uint32_t q[2] = {64u, WIN};
  for (int k = 0; k < 2; ++k) {
    const uint32_t end = (a.wrel / q[k] + 1u) * q[k];
    if (end < a.wrel + 2u) continue;
    const uint32_t d = end - a.wrel;
    float S = sfp_field_bound_f32_path<FM>(FS, x, y, z, a.t, vx, vy, vz,
                                           a.dt * (double)d * 1.000001);
    float chi_b = (float)P.chi_coef * S * sfp32_div(1.0f, (float)ig) * 1.001f;
    const double pb = rs_pbound(P, true, chi_b, ig, a.dt, false, env_lds);
    const double lim = 2.0 * pb;
    // no bound, or one a window's draws would likely fall under
    if (pb < 0.0 || !(lim < 0x1p-16)) continue;
    bool ok = true;
    for (uint32_t j = 1; j < d && ok; ++j) {
      const sfp_u32x4 dr = sfp_draw4(P.seed, pid, a.step + j, tag, 0u);
      const double r2 = MIX ? (double)sfp32_u01(dr.v[2]) : sfp_u01(dr.v[2], dr.v[3]);
      ok = !(r2 < lim);
    }
    if (ok) return d;
  }
  return 1;
}

// ------------------------------------------------------------ sweep kernel --
// A tile is up to ROWS rows of 256 consecutive particles (thread t handles particle
// t of every row, so loads stay coalesced); the QED list keeps ascending
// index order, i.e. (row, wave, lane) order, with one look-back per tile.
//
// Photons of a species without QED (this kernel; photons with QED have their
// own kernels, below): a photon whose wake step lies ahead is skipped (it was
// pushed through that step already); a photon is pushed to the end of the
// run at once -- the exact result of the 2n sequential FP64 additions of the
// same increment that n single steps make (sfp_repeat_add), so positions are
// bit-identical -- and sleeps until then, so every position is current when
// the run returns.
//
// Occupancy floor: 5 waves per SIMD in mixed precision (96 VGPRs), 3 in FP64
// (168 VGPRs, a few spilled); each ran faster than the compiler's own choice
// (104 and 203 VGPRs), and 6 or 4 waves ran slower.
static constexpr int ROWS = 32;

// sel: the species, and SEL_PROLOGUE when this launch opens the step: its
// first workgroup then takes the step's snapshot of the counts (what
// k_step_prologue does otherwise, a launch and a gap less per step), and its
// workgroups read their species' count itself, which nothing has changed yet
// in this step.
static constexpr int SEL_PROLOGUE = 0x100;
__device__ inline void step_snapshot(const Plan& P, const StepArgs& a);

// One species' sweep; opens: the launch opens the step (the count is read
// from n), snap: this call takes the step's snapshot.
template <bool PHOTON, bool QED, int FM, bool MIX, bool SH>
__device__ inline void sweep_species(const Plan& P, const StepArgs& a, int s, bool opens, bool snap,
                                     int rmax) {
  const SpeciesDev& sp = P.sp[s];
  __shared__ int64_t sh_tile;
  __shared__ uint32_t sh_w[4 * WAVES];
  __shared__ uint64_t sh_mask[ROWS][WAVES];
  __shared__ uint64_t sh_pref;

  const int64_t n_old = opens ? sp.cnt->n : sp.cnt->n_old;
  if (snap && blockIdx.x == 0) step_snapshot(P, a);
  if (!PHOTON && QED) {
    // Dense step: when the last step's bound listed at least 7/8 of the
    // particles, this kernel stands down and the QED kernel visits every
    // particle by index -- bound, exact QED where the bound allows an event,
    // push -- reading the store once instead of here and again through the
    // list.  A particle the bound culls is pushed there exactly as here, so
    // results do not change.  The list length reported is the store's
    // count: the room the QED kernel asks for, and what the capacity
    // recovery grows the destination by.  Every workgroup reads the same
    // counts (written by earlier launches) and takes the same decision.
    const uint32_t par = (uint32_t)(a.step & 1u);
    const int64_t prev = sp.cnt->dense_listed[par ^ 1u];
    if (sp.push && n_old > 0 && prev * 8 >= n_old * 7) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        sp.cnt->n_hot = n_old;
        sp.cnt->hot_total += n_old;
        sp.cnt->dense = 1;
        sp.cnt->dense_listed[par] = 0;  // the QED kernel adds this step's listed particles
        sp.cnt->dense_steps += 1;
      }
      return;
    }
  }
  // Rows per tile: rmax (at most ROWS, the launch shape's choice) once the
  // store fills a tile per workgroup of the grid; fewer (a power of two) for
  // shorter stores, which then still spread over many workgroups.  The list
  // order does not depend on the tiling.
  int rows = rmax < 1 ? 1 : (rmax > ROWS ? ROWS : rmax);
  {
    const int64_t per = (n_old + (int64_t)TILE * gridDim.x - 1) / ((int64_t)TILE * gridDim.x);
    if (per < rows) {
      rows = 1;
      while (rows < per) rows <<= 1;
    }
  }
  const int64_t tspan = (int64_t)TILE * rows;
  const int64_t ntiles = (n_old + tspan - 1) / tspan;
  const uint32_t nwg = participants(ntiles);
  if (blockIdx.x >= nwg) return;
  ScanState* scan = P.scan[s][0];
  const uint32_t epoch = QED ? next_epoch(scan) : 0u;
  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  // alive, processed, slept, sleeps
  uint32_t cnt[4] = {0, 0, 0, 0};
  uint32_t& alive = cnt[0];
  uint32_t& proc = cnt[1];
  uint32_t* const wake = PHOTON ? sp.wake : nullptr;

  while (true) {
    if (threadIdx.x == 0) sh_tile = (int64_t)atomicAdd(&scan->tile_ctr, 1u);
    __syncthreads();
    const int64_t tile = sh_tile;
    __syncthreads();
    if (tile >= ntiles) break;
    const int64_t base = tile * tspan;
    uint32_t hotbits = 0;
    // photons: the wake steps of all rows of the tile, loaded together (one
    // dependent load per row would leave the sweep latency-bound)
    uint32_t asleep = 0;
    if (PHOTON && wake) {
      uint32_t w[ROWS];
#pragma unroll
      for (int row = 0; row < ROWS; ++row) {
        const int64_t i = base + (int64_t)row * TILE + threadIdx.x;
        w[row] = (row < rows && i < n_old) ? wake[i] : 0u;
      }
#pragma unroll
      for (int row = 0; row < ROWS; ++row) asleep |= (uint32_t)(w[row] > a.rel) << row;
      // a sleeping photon is live (only live photons sleep) and already pushed
      const uint32_t ns = (uint32_t)__popc(asleep);
      alive += ns;
      cnt[2] += ns;
    }

    // leptons: software pipeline over the rows, the state of the next row
    // loads while this row computes
    uint8_t npr = 1;
    double nx = 0, ny = 0, nz = 0, nux = 0, nuy = 0, nuz = 0;
    if (!PHOTON && base + threadIdx.x < n_old) {
      const int64_t i0 = base + threadIdx.x;
      npr = sp.pruned[i0];
      nx = sp.x[i0]; ny = sp.y[i0]; nz = sp.z[i0];
      nux = sp.ux[i0]; nuy = sp.uy[i0]; nuz = sp.uz[i0];
    }
#pragma unroll 1
    for (int row = 0; row < rows; ++row) {
      const int64_t i = base + (int64_t)row * TILE + threadIdx.x;
      if (PHOTON && wake && !__any(i < n_old && !((asleep >> row) & 1u))) {
        // the whole wave sleeps (or lies past the end) in this row
        if (QED && lane == 0) sh_mask[row][wv] = 0;
        continue;
      }
      bool active = false;
      double x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 0, sx = 0, sy = 0, sz = 0;
      if (!PHOTON) {
        if (i < n_old) {
          x = nx; y = ny; z = nz;
          ux = nux; uy = nuy; uz = nuz;
          active = npr == 0;
        }
        if (row + 1 < rows && i + TILE < n_old) {
          const int64_t in = i + TILE;
          npr = sp.pruned[in];
          nx = sp.x[in]; ny = sp.y[in]; nz = sp.z[in];
          nux = sp.ux[in]; nuy = sp.uy[in]; nuz = sp.uz[in];
        }
      } else if (i < n_old && !((asleep >> row) & 1u)) {
        // the pruned flag and the state are loaded together
        const uint8_t pr = sp.pruned[i];
        x = sp.x[i]; y = sp.y[i]; z = sp.z[i];
        ux = sp.ux[i]; uy = sp.uy[i]; uz = sp.uz[i];
        active = pr == 0;
      }
      double ig = 1.0;
      if (active) {
        if (MIX && !PHOTON) {  // mixed precision: the momentum is an FP32 quantity
          const float fx = (float)ux, fy = (float)uy, fz = (float)uz;
          ig = (double)__frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz);
        } else {
          ig = PHOTON ? sfp_inv_gamma_photon(ux, uy, uz) : sfp_inv_gamma_lepton(ux, uy, uz);
        }
        proc += 1;
      }
      uint32_t shot = 0;
      const sfp_fieldset* FS = active ? fields_of<SH>(P, sp, i, shot) : &P.fields;
      bool hot = false;
      if (QED && active && sp.push) {
        // chi <= (e hbar/(m^2 c^3)) gamma (|E| + c|B|) (gamma >= |u|); the FP32
        // bound and this product carry a small margin for their own rounding
        float S = sfp_field_bound_f32<FM>(FS, x, y, z, a.t);
        float chi_b = (float)P.chi_coef * S * sfp32_div(1.0f, (float)ig) * 1.001f;
        hot = qed_possible(P, PHOTON, chi_b, ig, a.dt, !PHOTON && P.spin_emission && sp.spin);
      }
      if (QED) {
        uint64_t m = __ballot(hot);
        if (lane == 0) sh_mask[row][wv] = m;
        if (hot) hotbits |= 1u << row;
      }
      if (active && !hot) {
        if (PHOTON) {
          uint32_t nst = 1;
          // multi-step decisions are taken on every WIN-th step of the run,
          // so that the lanes of a wave take them together
          if (wake)
            nst = !sp.push ? a.rel_end - a.rel
                           : ((a.rel % WIN) == 0 ? sleep_steps<FM>(P, FS, a, QED, x, y, z, ux, uy, uz, ig)
                                                : 1u);
          if (sp.push) {
            // cpu.py:5-19 twice per step: x += (c dt/2 / |u|) ux, the same
            // increment each time
            const double hx = a.cdt_half * ig * ux, hy = a.cdt_half * ig * uy,
                         hz = a.cdt_half * ig * uz;
            if (nst == 1) {
              x += hx; x += hx;
              y += hy; y += hy;
              z += hz; z += hz;
            } else {  // the same 2 nst additions, in closed form
              x = sfp_repeat_add(x, hx, 2 * nst);
              y = sfp_repeat_add(y, hy, 2 * nst);
              z = sfp_repeat_add(z, hz, 2 * nst);
            }
            sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
          }
          if (nst > 1) {
            wake[i] = a.rel + nst;
            cnt[3] += 1;
          }
        } else if (sp.push) {
          if (sp.spin) { sx = sp.sx[i]; sy = sp.sy[i]; sz = sp.sz[i]; }
          if (MIX) {
            float fx = (float)ux, fy = (float)uy, fz = (float)uz;
            float s0 = (float)sx, s1 = (float)sy, s2 = (float)sz;
            lepton_push32<FM>(P, FS, sp, a, x, y, z, fx, fy, fz, s0, s1, s2, (float)ig);
            ux = fx; uy = fy; uz = fz;
            sx = s0; sy = s1; sz = s2;
          } else {
            lepton_push<FM>(P, FS, sp, a, x, y, z, ux, uy, uz, sx, sy, sz, ig);
          }
          sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
          sp.ux[i] = ux; sp.uy[i] = uy; sp.uz[i] = uz;
          if (sp.spin) { sp.sx[i] = sx; sp.sy[i] = sy; sp.sz[i] = sz; }
        }
        alive += 1;
      }
    }

    if (QED) {
      __syncthreads();
      if (wv == 0) {
        uint64_t c = 0;
        for (int k = lane; k < rows * WAVES; k += 64) c += (uint64_t)__popcll((&sh_mask[0][0])[k]);
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
        uint64_t pre = lookback(P.flags[s][0], tile, c, epoch);
        if (lane == 0) sh_pref = pre;
      }
      __syncthreads();
      uint64_t off = sh_pref;
      const uint64_t lt = (1ull << lane) - 1ull;
      for (int row = 0; row < rows; ++row) {
        uint32_t before = 0, total = 0;
        for (int k = 0; k < WAVES; ++k) {
          uint32_t c = (uint32_t)__popcll(sh_mask[row][k]);
          if (k < wv) before += c;
          total += c;
        }
        if (hotbits & (1u << row)) {
          const int64_t i = base + (int64_t)row * TILE + threadIdx.x;
          const uint64_t r = off + before + (uint32_t)__popcll(sh_mask[row][wv] & lt);
          sp.hot[r] = (uint32_t)i;
        }
        off += total;
      }
      __syncthreads();  // sh_mask is rewritten by the next tile
    }
  }

  uint64_t tot[4];
  block_sum(cnt, sh_w, tot);
  if (threadIdx.x == 0) {
    count_add(&sp.cnt->psteps, tot[0]);
    count_add(&sp.cnt->processed, tot[1]);
    count_add(&sp.cnt->slept, tot[2]);
    count_add(&sp.cnt->sleeps, tot[3]);
  }
  kernel_exit(scan, nwg, [&]() {
    if (QED) {
      uint64_t total = ntiles > 0 ? (ld_flag(&P.flags[s][0][ntiles - 1]) & FLAG_VAL_MASK) : 0;
      sp.cnt->n_hot = (int64_t)total;
      sp.cnt->hot_total += (int64_t)total;
      if (!PHOTON) {
        sp.cnt->dense = 0;
        sp.cnt->dense_listed[a.step & 1u] = (int64_t)total;
      }
      scan->epoch = epoch;
    } else {
      sp.cnt->n_hot = 0;
    }
  });
}

template <bool PHOTON, bool QED, int FM, bool MIX, bool SH>
__global__ void __launch_bounds__(TILE, MIX ? 5 : 3) k_sweep(const Plan* __restrict__ plan, int sel, const StepArgs* __restrict__ ap,
        int rmax) {
  const StepArgs a = *ap;
  const Plan& P = *plan;
  if (P.halt->flag) return;
  const bool opens = (sel & SEL_PROLOGUE) != 0;
  sweep_species<PHOTON, QED, FM, MIX, SH>(P, a, sel & 0xff, opens, opens, rmax);
}

// The sweeps of two lepton species with QED that share their QED launches
// (sel: the first species in bits 0-3, the second plus one in bits 4-7), one
// after the other in one launch: a launch and its gap less per step.  The two
// lists and scans are separate, so the second sweep waits on nothing of the
// first; a workgroup takes the second species' tiles once it has left the
// first's (workgroups past the first's participants at once), so every tile
// a look-back waits on belongs to a running workgroup.
template <int FM, bool MIX, bool SH>
__global__ void __launch_bounds__(TILE, MIX ? 5 : 3) k_sweep_pair(const Plan* __restrict__ plan, int sel, const StepArgs* __restrict__ ap,
        int rmax) {
  const StepArgs a = *ap;
  const Plan& P = *plan;
  if (P.halt->flag) return;
  const bool opens = (sel & SEL_PROLOGUE) != 0;
  sweep_species<false, true, FM, MIX, SH>(P, a, sel & 0xf, opens, opens, rmax);
  __syncthreads();  // the workgroup's LDS is reused
  sweep_species<false, true, FM, MIX, SH>(P, a, ((sel >> 4) & 0xf) - 1, opens, false, rmax);
}

// ------------------------------------------------------------- QED front --
// Room for one append per listed particle in every destination, or raise the
// halt flag (every workgroup reads the same counts and takes the same
// decision).  Returns false when the kernel must stop.
__device__ inline bool room_or_halt(const Plan& P, int s, const StepArgs& a, const SpeciesDev* dA,
                                    const SpeciesDev* dB, int64_t n_list) {
  const int64_t baseA = dA ? dA->cnt->n : 0, baseB = dB ? dB->cnt->n : 0;
  if ((dA && baseA + n_list > dA->cap) || (dB && baseB + n_list > dB->cap)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      P.halt->species = s;
      P.halt->step = (int64_t)a.step;
      P.halt->flag = 1;
    }
    return false;
  }
  return true;
}

// Diagnostics of the exact QED evaluations a workgroup makes, counted in LDS
// (one LDS atomic per wave and occurrence; no register held across the
// launch, no contention on a global counter) and added to the species'
// counters at the end of the launch (qdiag_flush): rejection-sampling
// probabilities above 1 (one step samples at most one event, as the
// reference's qed/rejection_sampling.py:78-79,96-97 do), photon-emission z
// beyond the table (rate 0 here; the reference reads past the table's end,
// qed/rejection_sampling_tables.py:40-51), optical-depth chi above the table
// (clamped to its top node; the reference reads past it,
// qed/optical_depth_tables.py:39-40), NaN chi (no event), the largest
// probability (float bits), and the evaluations in which an event was
// possible at all (p >= 2^-54, or chi >= chi_min under optical depth: the
// lepton QED kernel's dense steps decide the next step's mode on it).
enum { QD_P1 = 0, QD_ZC, QD_CC, QD_NAN, QD_PMAX, QD_LIVE, QD_N };
// The counters live in LDS and are typed so (SFP_L): through a generic
// pointer the read below compiled to a flat load, whose wait also waited for
// every global load in flight -- the next row's state included.
struct QDiag {
  SFP_L uint32_t* sh;  // QD_N LDS words of the workgroup
  __device__ static void tally(SFP_L uint32_t* w, bool cond) {
    const uint64_t m = __ballot(cond);
    if (m && wave_lane() == __builtin_ctzll(m))
      __hip_atomic_fetch_add(w, (uint32_t)__popcll(m), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ void prob(double p) {
    tally(&sh[QD_P1], p > 1.0);
    tally(&sh[QD_LIVE], p >= 0x1p-54);
    const float f = (float)p;
    // positive floats order as bits; the atomic only where f exceeds the
    // maximum read so far (a stale read is smaller, never larger): after the
    // first rows of a workgroup it is rare, where one per lane on the same
    // LDS word would serialise
    if (f > 0.0f &&
        __float_as_uint(f) > __hip_atomic_load(&sh[QD_PMAX], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP))
      __hip_atomic_fetch_max(&sh[QD_PMAX], __float_as_uint(f), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // photon emission: z = (delta / ((1 - delta) chi))^(2/3) > z_max
  __device__ void zcheck(const Plan& P, double chi, double delta) {
    tally(&sh[QD_ZC], delta > P.zmax15 * chi * (1.0 - delta));
  }
  __device__ void nan_chi(double chi) { tally(&sh[QD_NAN], chi != chi); }
  // optical depth, chi at or above chi_min: an event is possible
  __device__ void od(bool over) {
    tally(&sh[QD_CC], over);
    tally(&sh[QD_LIVE], true);
  }
  __device__ void clamp(bool over) { tally(&sh[QD_CC], over); }
};

// Zero the workgroup's counters (before the first barrier of the launch).
__device__ inline void qdiag_init(uint32_t* sh) {
  if (threadIdx.x < QD_N) sh[threadIdx.x] = 0;
}

// Add the workgroup's counters to the species' (after all its evaluations);
// returns the evaluations with a possible event.  Called by every thread.
__device__ inline uint32_t qdiag_flush(Counts* c, uint32_t* sh) {
  __syncthreads();
  if (threadIdx.x == 0) {
    count_add(&c->p_over1, sh[QD_P1]);
    count_add(&c->z_clamp, sh[QD_ZC]);
    count_add(&c->chi_clamp, sh[QD_CC]);
    count_add(&c->chi_nan, sh[QD_NAN]);
    if (sh[QD_PMAX])
      atomicMax((unsigned long long*)&c->p_max_bits, (unsigned long long)sh[QD_PMAX]);
  }
  return sh[QD_LIVE];
}

// Exact QED decision of one particle at time t (simulation.py:103-114):
// field, chi, then the rejection-sampling draw (qed/rejection_sampling.py:68-101)
// or the optical-depth update (qed/optical_depth.py:15-110), then the
// hard-photon cut for radiating leptons (cpu.py:227-232).
template <bool PHOTON, int KIND, int FM>
__device__ inline bool qed_event(const Plan& P, const sfp_fieldset* FS, const SpeciesDev& sp,
                                 const StepArgs& a,
                                 const sfp_rs_node* lds_rs, const double2* lds_od, double x,
                                 double y, double z, double ux, double uy, double uz, double ig,
                                 uint64_t pid, double& tau, double& delta, double sx = 0.0,
                                 double sy = 0.0, double sz = 0.0, QDiag* dg = nullptr) {
  bool event = false;
  delta = 0.0;
  double F[6];
  sfp_field_eval<FM>(FS, x, y, z, a.t, F);
  double chi = sfp_chi(F[0], F[1], F[2], F[3], F[4], F[5], ux, uy, uz, ig);
  if (dg) dg->nan_chi(chi);
  if (KIND == KIND_RS) {
    if (chi != 0.0) {
      sfp_u32x4 d = sfp_draw4(P.seed, pid, a.step, (uint32_t)sp.tag, 0u);
      double r1 = sfp_u01(d.v[0], d.v[1]);
      double r2 = sfp_u01(d.v[2], d.v[3]);
      double dtau = a.dt * ig;
      if (!PHOTON) {
        double r1_3 = sfp_powi(r1, 3);
        double w;
        if (P.spin_emission && sp.spin) {
          const double sig = sfp_spin_sigma(sp.q, ux, uy, uz, ig, F, sx, sy, sz);
          w = sfp_rs_rate_spin(lds_rs, P.rs_ai, P.z_N, P.z_min, P.z_max, P.z_delta, chi, r1_3, sig);
        } else {
          w = sfp_rs_rate(lds_rs, P.z_N, P.z_min, P.z_max, P.z_delta, chi, r1_3, 1);
        }
        double p = 3 * sfp_powi(r1, 2) * w * dtau;
        if (dg) {
          dg->prob(p);
          dg->zcheck(P, chi, r1_3);
        }
        if (r2 < p) { event = true; delta = r1_3; }
      } else {
        double p = sfp_rs_rate(lds_rs, P.z_N, P.z_min, P.z_max, P.z_delta, chi, r1, 0) * dtau;
        if (dg) dg->prob(p);
        if (r2 < p) { event = true; delta = r1; }
      }
    }
  } else {
    // below chi_min (or NaN): skipped, tau unchanged
    if (chi >= P.od.chi_min) {
      if (dg) dg->od(chi > P.od.chi_max);
      event = od_update(&tau, ig, chi, a.dt, P.seed, pid, a.step, (uint32_t)sp.tag, 0u, lds_od,
                        PHOTON ? gtab(P.od.pair_cdf) : gtab(P.od.photon_cdf), P.od, &delta);
    }
  }
  if (!PHOTON && event) {
    if (delta / ig < P.threshold) event = false;
  }
  return event;
}

// One optical-depth update in mixed precision (optical_depth.py:42-62): the
// total rate and the draws in FP32, tau accumulated in FP64; the photon energy
// of an event from the FP64 inverse CDF (events are rare).
__device__ inline bool od_update32(double* tau, float ig, double chi, float dt, uint64_t seed,
                                   uint64_t id, uint64_t step, uint32_t tag, const double2* nodes,
                                   const double* cdf, const sfp_od_dev& od, double* delta) {
  float lc = __log10f((float)chi);
  const float lmax = (float)od.log_chi_max, lmin = (float)od.log_chi_min;
  if (lc > lmax) lc = lmax;
  float rate = 0.0f;
  if (lc >= lmin) {
    const float ld = (float)od.log_chi_delta;
    int idx = (int)((lc - lmin) / ld);
    if (idx > (int)od.chi_N - 2) idx = (int)od.chi_N - 2;
    const double2 nd = nodes[idx];
    rate = (float)nd.x + (float)nd.y * (lc - (lmin + (float)idx * ld));
  }
  double t = *tau;
  if (t == 0.0) {
    sfp_u32x4 d = sfp_draw4(seed, id, step, tag, 1u);
    t = (double)(-__logf(1.0f - sfp32_u01(d.v[0])));
  }
  t -= (double)(rate * (dt * ig));
  bool ev = false;
  if (t < 0) {
    sfp_u32x4 d = sfp_draw4(seed, id, step, tag, 2u);
    t = (double)(-__logf(1.0f - sfp32_u01(d.v[0])));
    ev = true;
    *delta = sfp_od_inverse_cdf(cdf, &od, chi, sfp_u01(d.v[2], d.v[3]));
  } else {
    *delta = 0.0;
  }
  *tau = t;
  return ev;
}

// The QED decision in mixed precision: field in FP32, chi in FP64, the
// rejection-sampling rate from FP32 tables with 24-bit uniforms (the leading
// bits of the same Philox words as the FP64 step), optical depth as above.
template <bool PHOTON, int KIND, int FM>
__device__ inline bool qed_event32(const Plan& P, const sfp_fieldset* FS, const SpeciesDev& sp,
                                   const StepArgs& a,
                                   const sfp_rs_node32* lds_rs, const double2* lds_od, double x,
                                   double y, double z, float ux, float uy, float uz, float ig,
                                   uint64_t pid, double& tau, double& delta, float sx = 0.0f,
                                   float sy = 0.0f, float sz = 0.0f, QDiag* dg = nullptr) {
  bool event = false;
  delta = 0.0;
  float F[6];
  sfp32_field_eval<FM>(FS, x, y, z, a.t, F);
  const double chi = PHOTON ? sfp32_chi_photon(F, ux, uy, uz) : sfp32_chi(F, ux, uy, uz);
  if (dg) dg->nan_chi(chi);
  if (KIND == KIND_RS) {
    if (chi != 0.0) {
      sfp_u32x4 d = sfp_draw4(P.seed, pid, a.step, (uint32_t)sp.tag, 0u);
      const float r1 = sfp32_u01(d.v[0]), r2 = sfp32_u01(d.v[2]);
      const float dtau = (float)a.dt * ig;
      const float chif = (float)chi;
      const float zmin = (float)P.z_min, zmax = (float)P.z_max, zd = (float)P.z_delta;
      const float izd = (float)(1.0 / P.z_delta);
      if (!PHOTON) {
        const float r1_3 = r1 * r1 * r1;
        float w;
        if (P.spin_emission && sp.spin) {
          const float sig = sfp32_spin_sigma(sp.q, ux, uy, uz, ig, F, sx, sy, sz);
          w = sfp32_rs_rate_spin(lds_rs, P.rs_ai32, (int)P.z_N, zmin, zmax, zd, izd, chif, r1_3, sig);
        } else {
          w = sfp32_rs_rate(lds_rs, (int)P.z_N, zmin, zmax, zd, izd, chif, r1_3, true);
        }
        const float p = 3.0f * (r1 * r1) * w * dtau;
        if (dg) {
          dg->prob(p);
          dg->zcheck(P, chi, r1_3);
        }
        if (r2 < p) { event = true; delta = r1_3; }
      } else {
        const float p =
            sfp32_rs_rate(lds_rs, (int)P.z_N, zmin, zmax, zd, izd, chif, r1, false) * dtau;
        if (dg) dg->prob(p);
        if (r2 < p) { event = true; delta = r1; }
      }
    }
  } else {
    if (chi >= P.od.chi_min) {
      if (dg) dg->od(chi > P.od.chi_max);
      event = od_update32(&tau, ig, chi, (float)a.dt, P.seed, pid, a.step, (uint32_t)sp.tag,
                          lds_od, PHOTON ? gtab(P.od.pair_cdf) : gtab(P.od.photon_cdf), P.od, &delta);
    }
  }
  if (!PHOTON && event) {
    if (sfp32_div((float)delta, ig) < (float)P.threshold) event = false;
  }
  return event;
}

// The lepton QED decision of the mixed-precision step with the learned
// sampler, which is trained on the inverse of the emission's cumulative rate
// measured from delta = 0 (sfp_od_inverse_cdf_from0).  Optical depth: as
// qed_event32, with delta of an emission from the network, its uniform mapped
// onto that scale.  Rejection sampling: for chi inside the network's range
// the emission is sampled by composition -- an event with probability
// W(chi) dtau, W the total rate (qed/optical_depth_tables.py:46-56, the
// integral of the rejection-sampling rate over delta), then delta from the
// inverse CDF, here the network -- which draws the same (event, delta)
// distribution as rejection sampling; below the range the rejection sampling
// of qed_event32.
// Where the network is used, `net` is set with its normalised inputs and
// delta is filled in by the wave afterwards; bins flagged for fallback take
// the table's inverse CDF.  The hard-photon cut is left to the caller.
template <int KIND, int FM>
__device__ inline bool qed_event32_net(const Plan& P, const sfp_fieldset* FS,
                                       const SpeciesDev& sp, const StepArgs& a,
                                       const sfp_rs_node32* lds_rs, const double2* lds_od,
                                       double x, double y, double z, float ux, float uy, float uz,
                                       float ig, uint64_t pid, double& tau, double& delta_out,
                                       bool& net_out, float& in0_out, float& in1_out,
                                       double& chi_out, double& u_table_out, QDiag* dg = nullptr) {
  // The outputs are written once, at the one exit: stores to two of them at
  // the ends of two branches were merged into one store to a selected
  // address, which kept both of the caller's variables on the stack.
  bool event = false, net = false;
  double delta = 0.0, u_table = 0.0;
  float in0 = 0.0f, in1 = 0.0f;
  float F[6];
  sfp32_field_eval<FM>(FS, x, y, z, a.t, F);
  const double chi = sfp32_chi(F, ux, uy, uz);
  if (dg) dg->nan_chi(chi);
  const sfp_mlp_bins& nb = *P.net_bins;
  do {
    float u = 0.0f;     // uniform that selects delta, on the network's from-zero scale
    double od_u = 0.0;  // optical depth: the uniform of its own inverse CDF
    if (KIND == KIND_RS) {
      if (chi == 0.0) break;
      sfp_u32x4 d = sfp_draw4(P.seed, pid, a.step, (uint32_t)sp.tag, 0u);
      const float r1 = sfp32_u01(d.v[0]), r2 = sfp32_u01(d.v[2]);
      const float dtau = (float)a.dt * ig;
      if (!(chi >= P.od.chi_min)) {
        const float r1_3 = r1 * r1 * r1;
        const float p = 3.0f * (r1 * r1) *
                        sfp32_rs_rate(lds_rs, (int)P.z_N, (float)P.z_min, (float)P.z_max,
                                      (float)P.z_delta, (float)(1.0 / P.z_delta), (float)chi, r1_3,
                                      true) *
                        dtau;
        if (dg) {
          dg->prob(p);
          dg->zcheck(P, chi, r1_3);
        }
        if (r2 < p) { event = true; delta = r1_3; }
        break;
      }
      float lc = __log10f((float)chi);
      const float lmax = (float)P.od.log_chi_max, lmin = (float)P.od.log_chi_min;
      if (dg) dg->clamp(lc > lmax);
      if (lc > lmax) lc = lmax;
      const float ld = (float)P.od.log_chi_delta;
      int idx = (int)((lc - lmin) / ld);
      if (idx > (int)P.od.chi_N - 2) idx = (int)P.od.chi_N - 2;
      const double2 nd = lds_od[idx];
      const float W = (float)nd.x + (float)nd.y * (lc - (lmin + (float)idx * ld));
      if (dg) dg->prob(W * dtau);  // composition: the event probability of the step
      if (!(r2 < W * dtau)) break;
      event = true;
      u = r1;
      u_table = (double)r1;
    } else {
      if (!(chi >= P.od.chi_min)) break;
      if (dg) dg->od(chi > P.od.chi_max);
      event = od_update32(&tau, ig, chi, (float)a.dt, P.seed, pid, a.step, (uint32_t)sp.tag,
                          lds_od, gtab(P.od.photon_cdf), P.od, &delta);
      if (!event) break;
      sfp_u32x4 d = sfp_draw4(P.seed, pid, a.step, (uint32_t)sp.tag, 2u);
      // the optical-depth inverse CDF maps its uniform onto [ymin, ymax] of the
      // cumulative rate; on the network's from-zero scale that is
      // (ymin + u (ymax - ymin)) / ymax
      double ymin, ymax;
      sfp_od_row_ends(gtab(P.od.photon_cdf), &P.od, chi, &ymin, &ymax);
      const double ud = sfp_u01(d.v[2], d.v[3]);
      u = (float)((ymin + ud * (ymax - ymin)) / ymax);
      od_u = ud;
      u_table = ud;
    }
    // delta: the network, or the table in a fallback bin / outside the range
    const float lc = __log10f((float)chi);
    const float fb = (lc - nb.lc_min) / (nb.lc_max - nb.lc_min) * (float)SFP_MLP_BINS;
    const int bin = fb < 0.0f ? 0 : (fb >= (float)SFP_MLP_BINS ? SFP_MLP_BINS - 1 : (int)fb);
    if (lc >= nb.lc_min && lc <= nb.lc_max && !((nb.fallback >> bin) & 1ull)) {
      net = true;
      in0 = sfp_mlp_x0(P.net, lc);
      in1 = 2.0f * u - 1.0f;
    } else if (KIND == KIND_RS) {
      delta = sfp_od_sample_from0(gtab(P.od.photon_cdf), &P.od, chi, (double)u);
    } else {
      delta = sfp_od_inverse_cdf(gtab(P.od.photon_cdf), &P.od, chi, od_u);
    }
  } while (false);
  delta_out = delta;
  net_out = net;
  in0_out = in0;
  in1_out = in1;
  chi_out = chi;
  u_table_out = u_table;
  return event;
}

// The spin-resolved lepton emission decision of the mixed-precision step with
// the learned rate (rejection sampling, qed/rejection_sampling.py:68-85): the
// draws and sigma as in qed_event32; where log10 chi lies in the network's
// range and its bin is not flagged for fallback, `net` is set with the
// normalised inputs and the acceptance density comes from the wave's network
// evaluation afterwards; elsewhere p is the table's.  Returns r2, and p where
// the table decides.
template <int FM>
__device__ inline void qed_rate32_prep(const Plan& P, const sfp_fieldset* FS, const SpeciesDev& sp,
                                       const StepArgs& a, const sfp_rs_node32* lds_rs, double x,
                                       double y, double z, float ux, float uy, float uz, float ig,
                                       uint64_t pid, float sx, float sy, float sz, bool& net,
                                       float& in0, float& in1, float& in2, float& r1_3, float& r2,
                                       float& dtau, float& p_table, QDiag* dg = nullptr) {
  net = false;
  p_table = 0.0f;
  r2 = 1.0f;
  float F[6];
  sfp32_field_eval<FM>(FS, x, y, z, a.t, F);
  const double chi = sfp32_chi(F, ux, uy, uz);
  if (dg) dg->nan_chi(chi);
  if (chi == 0.0) return;
  sfp_u32x4 d = sfp_draw4(P.seed, pid, a.step, (uint32_t)sp.tag, 0u);
  const float r1 = sfp32_u01(d.v[0]);
  r2 = sfp32_u01(d.v[2]);
  r1_3 = r1 * r1 * r1;
  dtau = (float)a.dt * ig;
  if (dg) dg->zcheck(P, chi, r1_3);
  const float sig = sfp32_spin_sigma(sp.q, ux, uy, uz, ig, F, sx, sy, sz);
  const sfp_mlp_bins& nb = *P.rate_bins;
  const float lc = __log10f((float)chi);
  const float fb = (lc - nb.lc_min) / (nb.lc_max - nb.lc_min) * (float)SFP_MLP_BINS;
  const int bin = fb < 0.0f ? 0 : (fb >= (float)SFP_MLP_BINS ? SFP_MLP_BINS - 1 : (int)fb);
  if (lc >= nb.lc_min && lc <= nb.lc_max && !((nb.fallback >> bin) & 1ull)) {
    net = true;
    in0 = sfp_mlp_x0(P.rate_net, lc);
    in1 = 2.0f * r1 - 1.0f;
    in2 = sig;
  } else {
    p_table = 3.0f * (r1 * r1) *
              sfp32_rs_rate_spin(lds_rs, P.rs_ai32, (int)P.z_N, (float)P.z_min, (float)P.z_max,
                                 (float)P.z_delta, (float)(1.0 / P.z_delta), (float)chi, r1_3, sig) *
              dtau;
  }
}

// Normalised output of the rate network for the lanes with `net` set,
// sixteen at a time on the matrix cores.  Called by all lanes of the wave.
__device__ inline float net_rate_wave(const Plan& P, bool net, float in0, float in1, float in2) {
  const sfp_mlp_dev* M = P.rate_net;
  const int lane = wave_lane(), c = lane & 15;
  uint64_t m = __ballot(net);
  const uint64_t below = m & ((1ull << lane) - 1ull);
  int slot = (int)__popcll(below);
  float out = 0.0f;
  while (m) {
    const int cnt = __popcll(m) < 16 ? (int)__popcll(m) : 16;
    uint64_t mm = m;
    const int cc = c < cnt ? c : 0;
    for (int j = 0; j < cc; ++j) mm &= mm - 1ull;
    const int src = (int)__builtin_ctzll(mm);
    const float a0 = __shfl(in0, src), a1 = __shfl(in1, src), a2 = __shfl(in2, src);
    const float o = sfp_mlp_eval16_3(M, a0, a1, a2);
    const float res = __shfl(o, slot >= 0 && slot < 16 ? slot : 0);
    if (net && slot >= 0 && slot < 16) out = res;
    for (int j = 0; j < cnt; ++j) m &= m - 1ull;
    slot -= 16;
  }
  return out;
}

// delta = exp(network(in0, in1)) for the lanes with `net` set, sixteen at a
// time on the matrix cores.  Called by all lanes of the wave.
__device__ inline void net_delta_wave(const Plan& P, const __bf16* W2p, const __bf16* W3p,
                                      bool net, float in0, float in1, double& delta) {
  const sfp_mlp_dev* M = P.net;
  const int lane = wave_lane(), c = lane & 15;
  uint64_t m = __ballot(net);
  const uint64_t below = m & ((1ull << lane) - 1ull);
  int slot = (int)__popcll(below);
  while (m) {
    const int cnt = __popcll(m) < 16 ? (int)__popcll(m) : 16;
    uint64_t mm = m;
    const int cc = c < cnt ? c : 0;
    for (int j = 0; j < cc; ++j) mm &= mm - 1ull;
    const int src = (int)__builtin_ctzll(mm);
    const float a0 = __shfl(in0, src), a1 = __shfl(in1, src);
    const float o = sfp_mlp_eval16<SFP_W_LDS>(M, W2p, W3p, a0, a1);  // weights staged in LDS
    const float res = __shfl(o, slot >= 0 && slot < 16 ? slot : 0);
    if (net && slot >= 0 && slot < 16) delta = sfp_mlp_delta(M, res);
    for (int j = 0; j < cnt; ++j) m &= m - 1ull;
    slot -= 16;
  }
}

template <int KIND, bool PHOTON, bool MIX = false>
__device__ inline void stage_tables(const Plan& P, void* lds_rs, double2* lds_od) {
  if (KIND == KIND_RS) {
    int64_t nn = P.z_N - 1;
    if (MIX) {
      sfp_rs_node32* dst = (sfp_rs_node32*)lds_rs;
      for (int64_t k = threadIdx.x; k < nn && k < 1023; k += TILE) dst[k] = P.rs_nodes32[k];
    } else {
      sfp_rs_node* dst = (sfp_rs_node*)lds_rs;
      for (int64_t k = threadIdx.x; k < nn && k < 1023; k += TILE) dst[k] = P.rs_nodes[k];
    }
  } else {
    const double2* src = PHOTON ? P.od_pair_nodes : P.od_photon_nodes;
    int64_t nn = P.od.chi_N - 1;
    for (int64_t k = threadIdx.x; k < nn && k < 255; k += TILE) lds_od[k] = src[k];
  }
  __syncthreads();
}

// Counts of the destinations after the launch (run by its last workgroup).
// room: the appends the destinations had room for (the photon kernel); the
// ranks past it were deferred, and the launch halts the step so that the
// host grows the destinations and runs the deferred-append kernel.
__device__ inline void finish_appends(const Plan& P, int s, int ph, int64_t ntiles,
                                      const SpeciesDev& sp, const SpeciesDev* dA,
                                      const SpeciesDev* dB, int64_t baseA, int64_t baseB,
                                      int64_t idA, int64_t idB, int64_t room = INT64_MAX,
                                      uint64_t step = 0) {
  uint64_t total = ntiles > 0 ? (ld_flag(&P.flags[s][ph][ntiles - 1]) & FLAG_VAL_MASK) : 0;
  if ((int64_t)total > room) {
    sp.cnt->defer_n = (int64_t)total - room;
    total = (uint64_t)room;
    P.halt->species = s;
    P.halt->step = (int64_t)step;
    P.halt->flag = 1;
  }
  if (total) atomicAdd((unsigned long long*)&sp.cnt->spawned, (unsigned long long)total);
  if (dA) {
    dA->cnt->n = baseA + (int64_t)total;
    dA->cnt->next_id = idA + (int64_t)total;
    atomicAdd((unsigned long long*)&dA->cnt->psteps, (unsigned long long)total);
  }
  if (dB) {
    dB->cnt->n = baseB + (int64_t)total;
    dB->cnt->next_id = idB + (int64_t)total;
    atomicAdd((unsigned long long*)&dB->cnt->psteps, (unsigned long long)total);
  }
}

// ------------------------------------------------------- lepton QED kernel --
// Over the sweep's QED list: exact QED front, stable rank of every emission
// (wave64 ballot + workgroup scan + decoupled look-back), the new photon
// appended already pushed (the reference pushes particles in the step that
// creates them), then recoil (cpu.py:78-88) and push (simulation.py:128-137).
// NETK: 0 tables, 1 the learned photon-energy sampler, 2 the learned
// spin-resolved emission rate (mixed precision, rejection sampling)
//
// Occupancy: the compiler's own register count (about 160 VGPRs, 3 waves
// per SIMD in mixed precision) ran faster than a cap at 128 VGPRs (4 waves)
// or an explicit 3-wave floor, which schedules differently; the NETK == 2
// instances keep the 4-wave cap, which ran faster for them.  In FP64 with
// general fields (focused pulses) the compiler's choice spills into AGPRs and
// runs 1 wave per SIMD; a 2-wave floor (256 VGPRs, a few spills to scratch)
// ran faster (a 1e6-electron focused-pulse collision: 3.2e10 -> 3.5e10
// particle-steps/s).  The LDS of a workgroup (37-50 KB) allows at most 4
// workgroups per CU.
//
// One launch serves one lepton species or two (`sel`: the first in bits 0-3,
// the second plus one in bits 4-7, 0 for none) that share the kernel and the
// photon store: the tiles of the first species' list, then those of the
// second's, come from one counter; each species ranks its events in a
// look-back chain of its own, and the second appends its photons after all
// of the first's (it waits for the first chain's total before its first
// append), so the store's order is that of two launches in species order.
// The second species' tables and launch gap are saved: it is the positrons
// of a collision, a few thousand particles whose own launch was a chain of
// memory latencies (about 20 us) with the GPU otherwise idle.
template <int KIND, int FM, bool MIX, int NETK, bool SH>
__global__ void __launch_bounds__(TILE, (MIX && NETK == 2) ? 4 : ((!MIX && FM == SFP_FM_ANY) ? 2 : 1)) k_hot(const Plan* __restrict__ plan, int sel, const StepArgs* __restrict__ ap,
      int rmax) {
  const StepArgs a = *ap;
  const Plan& P = *plan;
  if (P.halt->flag) return;
  using Node = typename std::conditional<MIX, sfp_rs_node32, sfp_rs_node>::type;

  __shared__ Node lds_rs[KIND == KIND_RS ? 1023 : 1];
  constexpr bool NET = NETK == 1;
  __shared__ double2 lds_od[(KIND == KIND_OD || NET) ? 255 : 1];
  __shared__ int64_t sh_tile;
  __shared__ int64_t sh_off;
  __shared__ uint32_t sh_w[4 * WAVES];
  __shared__ uint64_t sh_pref;
  __shared__ uint64_t sh_mask[NETK == 2 ? 1 : QROWS][WAVES];
  __shared__ uint16_t sh_q[NETK == 2 ? 1 : QROWS * TILE];  // tile positions (< 2^16)
  __shared__ uint32_t sh_qd[2][QD_N];

  const int nseg = (sel >> 4) ? 2 : 1;
  const int sid0 = sel & 15, sid1 = nseg == 2 ? (sel >> 4) - 1 : sid0;
  // per species: tiles of up to rmax (<= QROWS) rows, or of one row for the
  // rate network
  int64_t n_hot_g[2] = {0, 0}, ntiles_g[2] = {0, 0};
  int rows_g[2] = {1, 1};
  uint32_t epoch_g[2] = {0, 0};
  for (int g = 0; g < nseg; ++g) {
    const int sg = g ? sid1 : sid0;
    n_hot_g[g] = P.sp[sg].cnt->n_hot;
    rows_g[g] = NETK == 2 ? 1 : qed_rows(n_hot_g[g], rmax);
    const int64_t sp_g = (int64_t)rows_g[g] * TILE;
    ntiles_g[g] = (n_hot_g[g] + sp_g - 1) / sp_g;
    epoch_g[g] = next_epoch(P.scan[sg][1]);
  }
  const int64_t ntiles_all = ntiles_g[0] + ntiles_g[1];
  // workgroups without work leave before reading the destination counts,
  // which the exit of this launch changes
  const uint32_t nwg = participants(ntiles_all);
  if (blockIdx.x >= nwg) return;
  const SpeciesDev* dA = P.sp[sid0].photon_dst >= 0 ? &P.sp[P.sp[sid0].photon_dst] : nullptr;
  if (!room_or_halt(P, sid0, a, dA, nullptr, n_hot_g[0] + n_hot_g[1])) return;
  const int64_t base0 = dA ? dA->cnt->n : 0;
  const int64_t id0 = dA ? dA->cnt->next_id : 0;
  ScanState* scan = P.scan[sid0][1];

  // per species: the workgroup's pushed and processed counts (its tiles of
  // the first species all come before those of the second)
  uint32_t alive = 0, proc = 0, alive0 = 0, proc0 = 0;
  bool second = false;
  // the first species' appends, where the second's start (waited for once)
  int64_t off = 0;
  bool have_off = false;
  qdiag_init(&sh_qd[0][0]);
  qdiag_init(&sh_qd[1][0]);
  bool staged = false;
  if constexpr (NETK != 2) {
  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  while (true) {
    if (threadIdx.x == 0) sh_tile = (int64_t)atomicAdd(&scan->tile_ctr, 1u);
    __syncthreads();
    const int64_t vt = wg_uniform(sh_tile);
    __syncthreads();
    if (vt >= ntiles_all) break;
    const int g = vt >= ntiles_g[0] ? 1 : 0;
    if (g && !second) {
      alive0 = alive; proc0 = proc;
      alive = 0; proc = 0;
      second = true;
    }
    const int s = g ? sid1 : sid0;
    const SpeciesDev& sp = P.sp[s];
    const int64_t tile = vt - (g ? ntiles_g[0] : 0);
    const int64_t n_hot = g ? n_hot_g[1] : n_hot_g[0];
    const int rows = g ? rows_g[1] : rows_g[0];
    const int64_t SPAN = (int64_t)rows * TILE;
    const uint32_t epoch = g ? epoch_g[1] : epoch_g[0];
    QDiag dg{(SFP_L uint32_t*)&sh_qd[g][0]};
    // dense step (set by this step's sweep): the list is the identity over
    // the store, holes included, and every particle gets the exact evaluation
    // (a particle the sweep's bound would have culled gives "no event" there,
    // and is pushed as the sweep pushes it)
    const bool dense = sp.cnt->dense != 0;
    // tables go to LDS once the workgroup has work (a short list leaves most
    // of the persistent grid without a tile)
    if (!staged) {
      stage_tables<KIND, false, MIX>(P, lds_rs, lds_od);
      if (NET && KIND == KIND_RS) {  // the total photon rate for composition sampling
        const int64_t nn = P.od.chi_N - 1;
        for (int64_t k = threadIdx.x; k < nn && k < 255; k += TILE) lds_od[k] = P.od_photon_nodes[k];
        __syncthreads();
      }

      staged = true;
    }

    // row pass: the QED decision of every listed lepton; the ones without an
    // event (emission, or with the learned sampler an emission candidate)
    // are pushed and stored, the others are left untouched for the event pass
    // software pipeline over the rows: the state of the next row and the
    // list index of the one after are loaded while this row computes
    const int64_t k0 = tile * SPAN + threadIdx.x;
    uint32_t i_cur = k0 < n_hot ? (dense ? (uint32_t)k0 : sp.hot[k0]) : 0u;
    uint32_t i_next = (rows > 1 && k0 + TILE < n_hot)
                          ? (dense ? (uint32_t)(k0 + TILE) : sp.hot[k0 + TILE]) : 0u;
    double nx = 0, ny = 0, nz = 0, nux = 0, nuy = 0, nuz = 0;
    uint64_t npid = 0;
    uint8_t npr = 0;
    if (k0 < n_hot) {
      nx = sp.x[i_cur]; ny = sp.y[i_cur]; nz = sp.z[i_cur];
      nux = sp.ux[i_cur]; nuy = sp.uy[i_cur]; nuz = sp.uz[i_cur];
      npid = sp.id[i_cur];
      if (dense) npr = sp.pruned[i_cur];
    }
#pragma unroll 1
    for (int row = 0; row < rows; ++row) {
      const int64_t k = tile * SPAN + (int64_t)row * TILE + threadIdx.x;
      const bool valid = k < n_hot && npr == 0;
      const int64_t i = valid ? (int64_t)i_cur : 0;
      double x = nx, y = ny, z = nz, ux = nux, uy = nuy, uz = nuz, tau = 0, sx = 0, sy = 0, sz = 0;
      uint64_t pid = npid;
      // this row's own loads go first: loads complete in order, so one issued
      // after the next row's would wait for those
      if (valid) {
        if (KIND == KIND_OD) tau = sp.tau[i];
        if (sp.spin) { sx = sp.sx[i]; sy = sp.sy[i]; sz = sp.sz[i]; }
      }
      // the list index of the row after next loads into a temporary that
      // becomes i_next only at the end of this row: taken at once, its wait
      // (the load issued last) was a wait for every load in flight, the next
      // row's state included, and the rows did not overlap at all
      uint32_t i_after = 0u;
      if (row + 1 < rows && k + TILE < n_hot) {
        i_cur = i_next;
        nx = sp.x[i_cur]; ny = sp.y[i_cur]; nz = sp.z[i_cur];
        nux = sp.ux[i_cur]; nuy = sp.uy[i_cur]; nuz = sp.uz[i_cur];
        npid = sp.id[i_cur];
        npr = dense ? sp.pruned[i_cur] : (uint8_t)0;
        if (row + 2 < rows && k + 2 * TILE < n_hot)
          i_after = dense ? (uint32_t)(k + 2 * TILE) : sp.hot[k + 2 * TILE];
      }
      uint32_t shot = 0;
      const sfp_fieldset* FS = valid ? fields_of<SH>(P, sp, i, shot) : &P.fields;
      float fx = (float)ux, fy = (float)uy, fz = (float)uz;
      float igf = 1.0f;
      double ig = 1.0;
      if (valid) {
        if (MIX) {
          igf = __frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz);
        } else {
          ig = sfp_inv_gamma_lepton(ux, uy, uz);
        }
      }
      const bool need = valid;
      if (dense && valid) proc += 1;
      bool ev = false;
      double delta = 0.0;
      if (need) {
        if (NET) {
          bool net = false;
          float in0 = 0.0f, in1 = 0.0f;
          double chi_n = 0.0, u_tab = 0.0;
          ev = qed_event32_net<KIND, FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, lds_od, x, y, z,
                                         fx, fy, fz, igf, pid, tau, delta, net, in0, in1, chi_n,
                                         u_tab, &dg);
        } else if (MIX) {
          ev = qed_event32<false, KIND, FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, lds_od, x, y,
                                            z, fx, fy, fz, igf, pid, tau, delta, (float)sx,
                                            (float)sy, (float)sz, &dg);
        } else {
          ev = qed_event<false, KIND, FM>(P, FS, sp, a, (const sfp_rs_node*)lds_rs, lds_od, x, y, z,
                                          ux, uy, uz, ig, pid, tau, delta, sx, sy, sz, &dg);
        }
        if (!NET && !dA) ev = false;
      }
      const uint64_t m = __ballot(ev);
      if (lane == 0) sh_mask[row][wv] = m;
      if (valid && !ev) {
        if (MIX) {
          float s0 = (float)sx, s1 = (float)sy, s2 = (float)sz;
          lepton_push32<FM>(P, FS, sp, a, x, y, z, fx, fy, fz, s0, s1, s2, igf);
          ux = fx; uy = fy; uz = fz;
          sx = s0; sy = s1; sz = s2;
        } else {
          lepton_push<FM>(P, FS, sp, a, x, y, z, ux, uy, uz, sx, sy, sz, ig);
        }
        sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
        sp.ux[i] = ux; sp.uy[i] = uy; sp.uz[i] = uz;
        if (sp.spin) { sp.sx[i] = sx; sp.sy[i] = sy; sp.sz[i] = sz; }
        if (KIND == KIND_OD && need) sp.tau[i] = tau;
        alive += 1;
      }
      i_next = i_after;
    }

    // event pass, in list order, each decision recomputed from the lepton's
    // unchanged state.  Tables: the photon appended already pushed (the
    // reference pushes particles in the step that creates them), then recoil
    // (cpu.py:78-88) and push (simulation.py:128-137).  Learned sampler: the
    // candidate goes to the emission kernel, which samples its energy on the
    // matrix cores in full waves, applies the hard-photon cut, appends the
    // photon and pushes the lepton (queue order = index order, so the
    // appends keep their order).
    const uint32_t nev = tile_events(sh_mask, rows, P.flags[s][1], tile, epoch, sh_q, &sh_pref);
    if (!NET && g && !have_off) {  // the second species appends after the first
      if (threadIdx.x == 0) sh_off = (int64_t)chain_total(P.flags[sid0][1], ntiles_g[0], epoch_g[0]);
      __syncthreads();
      off = wg_uniform(sh_off);
      have_off = true;
    }
    const int64_t baseA = base0 + off, idA = id0 + off;
    for (uint32_t e = threadIdx.x; e < nev; e += TILE) {
      const int64_t k = tile * SPAN + (int64_t)sh_q[e];
      const int64_t rank = (int64_t)sh_pref + (int64_t)e;
      const int64_t i = dense ? k : (int64_t)sp.hot[k];
      double x = sp.x[i], y = sp.y[i], z = sp.z[i];
      double ux = sp.ux[i], uy = sp.uy[i], uz = sp.uz[i];
      double tau = KIND == KIND_OD ? sp.tau[i] : 0.0, sx = 0, sy = 0, sz = 0;
      if (sp.spin) { sx = sp.sx[i]; sy = sp.sy[i]; sz = sp.sz[i]; }
      const uint64_t pid = sp.id[i];
      uint32_t shot = 0;
      const sfp_fieldset* FS = fields_of<SH>(P, sp, i, shot);
      float fx = (float)ux, fy = (float)uy, fz = (float)uz;
      float igf = 1.0f;
      double ig = 1.0;
      if (MIX) {
        igf = __frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz);
      } else {
        ig = sfp_inv_gamma_lepton(ux, uy, uz);
      }
      double delta = 0.0;
      if (NET) {
        bool net = false;
        float in0 = 0.0f, in1 = 0.0f;
        double chi_n = 0.0, u_tab = 0.0;
        (void)qed_event32_net<KIND, FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, lds_od, x, y, z,
                                        fx, fy, fz, igf, pid, tau, delta, net, in0, in1, chi_n,
                                        u_tab);
        EmitRec r;
        r.i = (uint32_t)i;
        r.net = net ? 1u : 0u;
        r.in0 = in0;
        r.in1 = in1;
        r.chi = chi_n;
        r.u_tab = u_tab;
        r.delta = delta;
        r.x = x; r.y = y; r.z = z;
        r.ux = fx; r.uy = fy; r.uz = fz;
        r.sx = (float)sx; r.sy = (float)sy; r.sz = (float)sz;
        sp.emit[rank] = r;
        if (KIND == KIND_OD) sp.tau[i] = tau;
        continue;
      }
      if (MIX) {
        (void)qed_event32<false, KIND, FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, lds_od, x, y,
                                           z, fx, fy, fz, igf, pid, tau, delta, (float)sx,
                                           (float)sy, (float)sz);
        const float df = (float)delta;
        spawn_photon(shot, *dA, a, baseA + rank, (uint64_t)(idA + rank), x, y, z, df * fx, df * fy,
                     df * fz);
        if (sp.rr == SFP_RR_PHOTON) {
          const float f = 1.0f - (float)delta;
          fx *= f;
          fy *= f;
          fz *= f;
          igf = __frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz);
        }
        float s0 = (float)sx, s1 = (float)sy, s2 = (float)sz;
        lepton_push32<FM>(P, FS, sp, a, x, y, z, fx, fy, fz, s0, s1, s2, igf);
        ux = fx; uy = fy; uz = fz;
        sx = s0; sy = s1; sz = s2;
      } else {
        (void)qed_event<false, KIND, FM>(P, FS, sp, a, (const sfp_rs_node*)lds_rs, lds_od, x, y, z,
                                         ux, uy, uz, ig, pid, tau, delta, sx, sy, sz);
        spawn_photon(shot, *dA, a, baseA + rank, (uint64_t)(idA + rank), x, y, z, delta * ux,
                     delta * uy, delta * uz);
        if (sp.rr == SFP_RR_PHOTON) {
          ux *= 1 - delta;
          uy *= 1 - delta;
          uz *= 1 - delta;
          ig = sfp_inv_gamma_lepton(ux, uy, uz);
        }
        lepton_push<FM>(P, FS, sp, a, x, y, z, ux, uy, uz, sx, sy, sz, ig);
      }
      sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
      sp.ux[i] = ux; sp.uy[i] = uy; sp.uz[i] = uz;
      if (sp.spin) { sp.sx[i] = sx; sp.sy[i] = sy; sp.sz[i] = sz; }
      if (KIND == KIND_OD) sp.tau[i] = tau;
      alive += 1;
    }
  }
  } else {
  while (true) {
    if (threadIdx.x == 0) sh_tile = (int64_t)atomicAdd(&scan->tile_ctr, 1u);
    __syncthreads();
    const int64_t vt = wg_uniform(sh_tile);
    __syncthreads();
    if (vt >= ntiles_all) break;
    const int g = vt >= ntiles_g[0] ? 1 : 0;
    if (g && !second) {
      alive0 = alive; proc0 = proc;
      alive = 0; proc = 0;
      second = true;
    }
    const int s = g ? sid1 : sid0;
    const SpeciesDev& sp = P.sp[s];
    const int64_t tile = vt - (g ? ntiles_g[0] : 0);
    const int64_t n_hot = g ? n_hot_g[1] : n_hot_g[0];
    const uint32_t epoch = g ? epoch_g[1] : epoch_g[0];
    QDiag dg{(SFP_L uint32_t*)&sh_qd[g][0]};
    const bool dense = sp.cnt->dense != 0;
    // tables go to LDS once the workgroup has work (a short list leaves most
    // of the persistent grid without a tile)
    if (!staged) {
      stage_tables<KIND, false, MIX>(P, lds_rs, lds_od);
      if (NET && KIND == KIND_RS) {  // the total photon rate for composition sampling
        const int64_t nn = P.od.chi_N - 1;
        for (int64_t k = threadIdx.x; k < nn && k < 255; k += TILE) lds_od[k] = P.od_photon_nodes[k];
        __syncthreads();
      }

      staged = true;
    }

    const int64_t k = tile * TILE + threadIdx.x;
    int64_t i = 0;
    bool valid = false;
    if (k < n_hot) {
      i = dense ? k : (int64_t)sp.hot[k];
      valid = !dense || sp.pruned[i] == 0;
    }
    double x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 0, tau = 0, sx = 0, sy = 0, sz = 0;
    uint64_t pid = 0;
    if (valid) {
      x = sp.x[i]; y = sp.y[i]; z = sp.z[i];
      ux = sp.ux[i]; uy = sp.uy[i]; uz = sp.uz[i];
      if (KIND == KIND_OD) tau = sp.tau[i];
      if (sp.spin) { sx = sp.sx[i]; sy = sp.sy[i]; sz = sp.sz[i]; }
      pid = sp.id[i];
    }
    uint32_t shot = 0;
    const sfp_fieldset* FS = valid ? fields_of<SH>(P, sp, i, shot) : &P.fields;
    float fx = (float)ux, fy = (float)uy, fz = (float)uz;
    float igf = 1.0f;
    double ig = 1.0;
    if (valid) {
      if (MIX) {
        igf = __frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz);
      } else {
        ig = sfp_inv_gamma_lepton(ux, uy, uz);
      }
    }
    const bool need = valid;
    if (dense && valid) proc += 1;
    bool event = false;
    double delta = 0.0;
    if (NET) {
      // An emission candidate goes to the emission kernel, which samples its
      // energy on the matrix cores in full waves, applies the hard-photon cut,
      // appends the photon and pushes the lepton (queue order = index order,
      // so the appends keep their order); the others are pushed here.
      bool net = false, cand = false;
      float in0 = 0.0f, in1 = 0.0f;
      double chi_n = 0.0, u_tab = 0.0;
      if (need)
        cand = qed_event32_net<KIND, FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, lds_od, x, y, z,
                                         fx, fy, fz, igf, pid, tau, delta, net, in0, in1, chi_n,
                                         u_tab, &dg);
      const int64_t q = tile_rank(cand, tile, P.flags[s][1], epoch, sh_w, &sh_pref);
      if (cand) {
        EmitRec r;
        r.i = (uint32_t)i;
        r.net = net ? 1u : 0u;
        r.in0 = in0;
        r.in1 = in1;
        r.chi = chi_n;
        r.u_tab = u_tab;
        r.delta = delta;
        r.x = x; r.y = y; r.z = z;
        r.ux = fx; r.uy = fy; r.uz = fz;
        r.sx = (float)sx; r.sy = (float)sy; r.sz = (float)sz;
        sp.emit[q] = r;
        if (KIND == KIND_OD) sp.tau[i] = tau;
      } else if (valid) {
        float s0 = (float)sx, s1 = (float)sy, s2 = (float)sz;
        lepton_push32<FM>(P, FS, sp, a, x, y, z, fx, fy, fz, s0, s1, s2, igf);
        sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
        sp.ux[i] = fx; sp.uy[i] = fy; sp.uz[i] = fz;
        if (sp.spin) { sp.sx[i] = s0; sp.sy[i] = s1; sp.sz[i] = s2; }
        if (KIND == KIND_OD && need) sp.tau[i] = tau;
        alive += 1;
      }
      continue;
    } else if (NETK == 2) {
      bool net = false;
      float in0 = 0.0f, in1 = 0.0f, in2 = 0.0f, r1_3 = 0.0f, r2 = 1.0f, dtau = 0.0f, p = 0.0f;
      if (need)
        qed_rate32_prep<FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, x, y, z, fx, fy, fz, igf,
                            pid, (float)sx, (float)sy, (float)sz, net, in0, in1, in2, r1_3, r2,
                            dtau, p, &dg);
      const float o = net_rate_wave(P, net, in0, in1, in2);
      if (net)
        p = sfp_rate_from_net(P.rate_net, o) * (float)P.rate_unit * dtau;
      if (need) dg.prob(p);
      if (valid && r2 < p) {
        event = true;
        delta = r1_3;
        if (sfp32_div((float)delta, igf) < (float)P.threshold) event = false;  // cpu.py:227-232
      }
    } else if (need) {
      if (MIX) {
        event = qed_event32<false, KIND, FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, lds_od, x, y,
                                             z, fx, fy, fz, igf, pid, tau, delta, (float)sx,
                                             (float)sy, (float)sz, &dg);
      } else {
        event = qed_event<false, KIND, FM>(P, FS, sp, a, (const sfp_rs_node*)lds_rs, lds_od, x, y, z,
                                           ux, uy, uz, ig, pid, tau, delta, sx, sy, sz, &dg);
      }
    }
    if (!dA) event = false;

    int64_t rank = tile_rank(event, tile, P.flags[s][1], epoch, sh_w, &sh_pref);
    if (g && !have_off) {  // the second species appends after the first
      if (threadIdx.x == 0) sh_off = (int64_t)chain_total(P.flags[sid0][1], ntiles_g[0], epoch_g[0]);
      __syncthreads();
      off = wg_uniform(sh_off);
      have_off = true;
    }
    const int64_t baseA = base0 + off, idA = id0 + off;
    if (event) {
      if (MIX) {
        const float df = (float)delta;
        spawn_photon(shot, *dA, a, baseA + rank, (uint64_t)(idA + rank), x, y, z, df * fx, df * fy,
                     df * fz);
      } else {
        spawn_photon(shot, *dA, a, baseA + rank, (uint64_t)(idA + rank), x, y, z, delta * ux,
                     delta * uy, delta * uz);
      }
    }

    if (valid) {
      if (MIX) {
        if (event && sp.rr == SFP_RR_PHOTON) {
          const float f = 1.0f - (float)delta;
          fx *= f;
          fy *= f;
          fz *= f;
          igf = __frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz);
        }
        float s0 = (float)sx, s1 = (float)sy, s2 = (float)sz;
        lepton_push32<FM>(P, FS, sp, a, x, y, z, fx, fy, fz, s0, s1, s2, igf);
        ux = fx; uy = fy; uz = fz;
        sx = s0; sy = s1; sz = s2;
      } else {
        if (event && sp.rr == SFP_RR_PHOTON) {
          ux *= 1 - delta;
          uy *= 1 - delta;
          uz *= 1 - delta;
          ig = sfp_inv_gamma_lepton(ux, uy, uz);
        }
        lepton_push<FM>(P, FS, sp, a, x, y, z, ux, uy, uz, sx, sy, sz, ig);
      }
      sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
      sp.ux[i] = ux; sp.uy[i] = uy; sp.uz[i] = uz;
      if (sp.spin) { sp.sx[i] = sx; sp.sy[i] = sy; sp.sz[i] = sz; }
      if (KIND == KIND_OD && need) sp.tau[i] = tau;
      alive += 1;
    }
  }
  }

  if (!second) {
    alive0 = alive; proc0 = proc;
    alive = 0; proc = 0;
  }
  uint32_t cnt[4] = {alive0, proc0, alive, proc};
  uint64_t tot[4];
  block_sum(cnt, sh_w, tot);
  for (int g = 0; g < nseg; ++g) {
    const SpeciesDev& sp = P.sp[g ? sid1 : sid0];
    const uint32_t live = qdiag_flush(sp.cnt, &sh_qd[g][0]);
    if (threadIdx.x == 0) {
      count_add(&sp.cnt->psteps, tot[2 * g]);
      count_add(&sp.cnt->processed, tot[2 * g + 1]);
      // dense step: the particles with a possible event decide the next
      // step's mode (the sweep counts its list otherwise)
      if (sp.cnt->dense != 0) count_add(&sp.cnt->dense_listed[a.step & 1u], live);
    }
  }
  kernel_exit(scan, nwg, [&]() {
    int64_t done = 0;  // the first species' appends
    for (int g = 0; g < nseg; ++g) {
      const int s = g ? sid1 : sid0;
      const SpeciesDev& sp = P.sp[s];
      const int64_t nt = ntiles_g[g];
      const uint64_t c = nt > 0 ? (ld_flag(&P.flags[s][1][nt - 1]) & FLAG_VAL_MASK) : 0;
      if (NET) {
        sp.cnt->n_emit = (int64_t)c;
        sp.cnt->emit_total += (int64_t)c;
      } else {
        finish_appends(P, s, 1, nt, sp, dA, nullptr, base0 + done, 0, id0 + done, 0);
      }
      done += (int64_t)c;
      P.scan[s][1]->epoch = epoch_g[g];
    }
  });
}

// --------------------------------------------------- lepton emission kernel --
// Mixed precision with the learned photon-energy sampler: over this step's
// emission candidates (from k_hot, in particle index order), delta from the
// network (whole waves of candidates on the matrix cores, weights in LDS) or
// the table, the hard tail from the table, the hard-photon cut (cpu.py:
// 227-232), the stable rank and append of the photon, then recoil
// (cpu.py:78-88) and push of the lepton.  The room for the appends was
// checked by k_hot (one per listed particle).
template <int KIND, int FM, bool SH>
__global__ void __launch_bounds__(TILE) k_emit(const Plan* __restrict__ plan, int sel,
                                               const StepArgs* __restrict__ ap, int rmax) {
  const StepArgs a = *ap;
  const Plan& P = *plan;
  if (P.halt->flag) return;
  __shared__ uint4 lds_w[2 * 4096 * 2 / 16];
  __shared__ int64_t sh_tile;
  __shared__ int64_t sh_off;
  __shared__ uint32_t sh_w[2 * WAVES];
  __shared__ uint64_t sh_pref;
  // one species or two, as in k_hot: the second's candidates follow the
  // first's from one tile counter, and its photons follow the first's
  const int nseg = (sel >> 4) ? 2 : 1;
  const int sid0 = sel & 15, sid1 = nseg == 2 ? (sel >> 4) - 1 : sid0;
  int64_t n_c_g[2] = {0, 0}, ntiles_g[2] = {0, 0};
  uint32_t epoch_g[2] = {0, 0};
  for (int g = 0; g < nseg; ++g) {
    const int sg = g ? sid1 : sid0;
    n_c_g[g] = P.sp[sg].cnt->n_emit;
    ntiles_g[g] = (n_c_g[g] + TILE - 1) / TILE;
    epoch_g[g] = next_epoch(P.scan[sg][2]);
  }
  const int64_t ntiles_all = ntiles_g[0] + ntiles_g[1];
  // workgroups without work leave before reading the destination counts,
  // which the exit of this launch changes
  const uint32_t nwg = participants(ntiles_all);
  if (blockIdx.x >= nwg) return;
  const SpeciesDev* dA = P.sp[sid0].photon_dst >= 0 ? &P.sp[P.sp[sid0].photon_dst] : nullptr;
  const int64_t base0 = dA ? dA->cnt->n : 0;
  const int64_t id0 = dA ? dA->cnt->next_id : 0;
  ScanState* scan = P.scan[sid0][2];
  uint32_t alive = 0, alive0 = 0;
  bool second = false;
  int64_t off = 0;  // the first species' appends (waited for once)
  bool have_off = false;
  bool staged = false;
  while (true) {
    if (threadIdx.x == 0) sh_tile = (int64_t)atomicAdd(&scan->tile_ctr, 1u);
    __syncthreads();
    const int64_t vt = wg_uniform(sh_tile);
    __syncthreads();
    if (vt >= ntiles_all) break;
    const int g = vt >= ntiles_g[0] ? 1 : 0;
    if (g && !second) {
      alive0 = alive;
      alive = 0;
      second = true;
    }
    const int s = g ? sid1 : sid0;
    const SpeciesDev& sp = P.sp[s];
    const int64_t tile = vt - (g ? ntiles_g[0] : 0);
    const int64_t n_c = g ? n_c_g[1] : n_c_g[0];
    const uint32_t epoch = g ? epoch_g[1] : epoch_g[0];
    if (!staged) {
      const uint4* w2 = (const uint4*)P.net->W2p;
      const uint4* w3 = (const uint4*)P.net->W3p;
      constexpr int NQ = 4096 * 2 / 16;
      for (int k = threadIdx.x; k < NQ; k += TILE) {
        lds_w[k] = w2[k];
        lds_w[NQ + k] = w3[k];
      }
      __syncthreads();
      staged = true;
    }
    const int64_t k = tile * TILE + threadIdx.x;
    const bool valid = k < n_c;
    EmitRec r;
    r.i = 0; r.net = 0; r.in0 = 0.0f; r.in1 = 0.0f; r.chi = 0.0; r.u_tab = 0.0; r.delta = 0.0;
    r.x = 0.0; r.y = 0.0; r.z = 0.0;
    r.ux = 0.0f; r.uy = 0.0f; r.uz = 0.0f; r.sx = 0.0f; r.sy = 0.0f; r.sz = 0.0f;
    if (valid) r = sp.emit[k];
    const int64_t i = (int64_t)r.i;
    double x = r.x, y = r.y, z = r.z;
    uint32_t shot = 0;
    const sfp_fieldset* FS = valid ? fields_of<SH>(P, sp, i, shot) : &P.fields;
    float fx = r.ux, fy = r.uy, fz = r.uz;
    float igf = valid ? __frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz) : 1.0f;
    double delta = r.delta;
    const bool net = valid && r.net;
    net_delta_wave(P, (const __bf16*)lds_w, (const __bf16*)lds_w + 4096, net, r.in0, r.in1, delta);
    if (net && delta > SFP_MLP_DELTA_CUT)  // the hard tail from the table, same uniform
      delta = (KIND == KIND_RS) ? sfp_od_sample_from0(gtab(P.od.photon_cdf), &P.od, r.chi, r.u_tab)
                                : sfp_od_inverse_cdf(gtab(P.od.photon_cdf), &P.od, r.chi, r.u_tab);
    // hard-photon cut (cpu.py:227-232)
    bool event = valid && !(sfp32_div((float)delta, igf) < (float)P.threshold);
    if (!dA) event = false;
    const int64_t rank = tile_rank(event, tile, P.flags[s][2], epoch, sh_w, &sh_pref);
    if (g && !have_off) {  // the second species appends after the first
      if (threadIdx.x == 0) sh_off = (int64_t)chain_total(P.flags[sid0][2], ntiles_g[0], epoch_g[0]);
      __syncthreads();
      off = wg_uniform(sh_off);
      have_off = true;
    }
    if (event) {
      const float df = (float)delta;
      spawn_photon(shot, *dA, a, base0 + off + rank, (uint64_t)(id0 + off + rank), x, y, z,
                   df * fx, df * fy, df * fz);
    }
    if (valid) {
      if (event && sp.rr == SFP_RR_PHOTON) {
        const float f = 1.0f - (float)delta;
        fx *= f;
        fy *= f;
        fz *= f;
        igf = __frsqrt_rn(1.0f + fx * fx + fy * fy + fz * fz);
      }
      float s0 = r.sx, s1 = r.sy, s2 = r.sz;
      lepton_push32<FM>(P, FS, sp, a, x, y, z, fx, fy, fz, s0, s1, s2, igf);
      sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
      sp.ux[i] = fx; sp.uy[i] = fy; sp.uz[i] = fz;
      if (sp.spin) { sp.sx[i] = s0; sp.sy[i] = s1; sp.sz[i] = s2; }
      alive += 1;
    }
  }
  if (!second) {
    alive0 = alive;
    alive = 0;
  }
  uint32_t cnt[2] = {alive0, alive};
  uint64_t tot[2];
  block_sum(cnt, sh_w, tot);
  if (threadIdx.x == 0)
    for (int g = 0; g < nseg; ++g) count_add(&P.sp[g ? sid1 : sid0].cnt->psteps, tot[g]);
  kernel_exit(scan, nwg, [&]() {
    int64_t done = 0;  // the first species' appends
    for (int g = 0; g < nseg; ++g) {
      const int s = g ? sid1 : sid0;
      const int64_t nt = ntiles_g[g];
      finish_appends(P, s, 2, nt, P.sp[s], dA, nullptr, base0 + done, 0, id0 + done, 0);
      done += nt > 0 ? (int64_t)(ld_flag(&P.flags[s][2][nt - 1]) & FLAG_VAL_MASK) : 0;
      P.scan[s][2]->epoch = epoch_g[g];
    }
  });
}

// ---------------------------------------------------------- photon kernels --
// Photons with QED advance in two kernels per step:
//  k_awake  -- streams the 4-byte wake steps of the whole store and lists, in
//              ascending index order, the photons to visit this step (wake
//              step reached); photons inside a multi-step push are counted
//              as live and skipped.
//  k_photon -- one listed photon per thread: the envelope bound decides
//              whether a pair is possible; if not, the photon is pushed one
//              step or -- on rebuild steps -- put to sleep for as many steps
//              as the bound along its straight path allows (its position
//              advanced in closed form, sfp_repeat_add, when it wakes); on
//              rebuild steps a photon for which a pair is possible but whose
//              own draws for the next 64 or 16 steps all lie above twice the
//              bound's probability sleeps as well (sleep_steps_drawn); if
//              so, the exact Breit-Wheeler front runs, events are ranked in
//              list order (= index order) and the pairs are appended already
//              pushed; the converted photon becomes a hole and is never
//              visited again.
static constexpr uint32_t WAKE_NEVER = 0xffffffffu;

// Wake steps per thread in k_awake: eight 16-byte loads (a tile of 8192
// words per look-back).  The wake array is allocated in whole tiles
// (WAKE_PAD words), so full-tile loads stay inside it.
static constexpr int AW = 32;
static constexpr int64_t WAKE_PAD = (int64_t)TILE * AW;

__global__ void __launch_bounds__(TILE) k_awake(const Plan* __restrict__ plan, int s, const StepArgs* __restrict__ ap,
        int rmax) {
  const StepArgs a = *ap;
  const Plan& P = *plan;
  if (P.halt->flag) return;
  const SpeciesDev& sp = P.sp[s];
  __shared__ int64_t sh_tile;
  __shared__ uint32_t sh_w[2 * WAVES];
  __shared__ uint64_t sh_pref;
  __shared__ uint32_t sh_list[AW * TILE];
  __shared__ uint8_t sh_cont[AW * TILE];
  const int64_t n_old = sp.cnt->n_old;
  ScanState* scan = P.scan[s][0];
  if (!list_rebuild(a)) {
    // Photons fall asleep and wake only on rebuild steps: in between, the
    // list of the previous step stays valid (holes it holds are skipped) and
    // is extended by the photons appended since, whose indices are larger.
    const int64_t from = sp.cnt->list_end, n_list = sp.cnt->n_hot;
    const uint32_t nwg = participants((n_old - from + TILE - 1) / TILE);
    if (blockIdx.x >= nwg) return;
    const WinCache C = win_cache(sp, (uint32_t)sp.cnt->cpar);
    for (int64_t j = from + (int64_t)blockIdx.x * TILE + threadIdx.x; j < n_old;
         j += (int64_t)nwg * TILE) {
      const int64_t k = n_list + (j - from);
      sp.hot[k] = (uint32_t)j;
      sp.slot[j] = (uint32_t)k;
      if (k < sp.ccap) {  // into the window cache (the new photons are contiguous)
        C.x(k) = sp.x[j]; C.y(k) = sp.y[j]; C.z(k) = sp.z[j];
        C.ux(k) = sp.ux[j]; C.uy(k) = sp.uy[j]; C.uz(k) = sp.uz[j];
        C.wake(k) = sp.pruned[j] ? WAKE_NEVER : sp.wake[j];
        C.id(k) = sp.id[j];
      }
    }
    kernel_exit(scan, nwg, [&]() {
      Counts* c = sp.cnt;
      c->n_hot = n_list + (n_old - from);
      c->list_end = n_old;
      c->psteps += c->asleep_n;
      c->slept += c->asleep_n;
    });
    return;
  }
  const int64_t ntiles = (n_old + WAKE_PAD - 1) / WAKE_PAD;
  const uint32_t nwg = participants(ntiles);
  if (blockIdx.x >= nwg) return;
  const uint32_t epoch = next_epoch(scan);
  const int64_t prev_end = sp.cnt->list_end;  // the last window's list covered [0, prev_end)
  // the last window's cache set and this one's (k_photon sees the new one
  // after this launch flips the parity)
  const uint32_t par0 = (uint32_t)sp.cnt->cpar;
  const WinCache Co = win_cache(sp, par0), Cn = win_cache(sp, par0 ^ 1u);
  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  uint32_t cnt[2] = {0, 0};  // asleep (live), listed

  while (true) {
    if (threadIdx.x == 0) sh_tile = (int64_t)atomicAdd(&scan->tile_ctr, 1u);
    __syncthreads();
    const int64_t tile = sh_tile;
    __syncthreads();
    if (tile >= ntiles) break;
    // thread t owns words [i0, i0 + AW): list order = (thread, word) = index order
    const int64_t i0 = tile * WAKE_PAD + (int64_t)threadIdx.x * AW;
    const uint4* src = (const uint4*)(sp.wake + i0);
    uint32_t w[AW];
#pragma unroll
    for (int q = 0; q < AW / 4; ++q) {
      const uint4 v = src[q];
      w[4 * q] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
    }
    uint32_t up = 0, cont = 0;  // listed; listed and awake through the last window
#pragma unroll
    for (int j = 0; j < AW; ++j) {
      const bool in = i0 + j < n_old;
      up |= (uint32_t)(in && w[j] <= a.wrel) << j;
      cont |= (uint32_t)(in && w[j] < a.wrel) << j;
      cnt[0] += (uint32_t)(in && w[j] > a.wrel && w[j] != WAKE_NEVER);
    }
    const uint32_t c = (uint32_t)__popc(up);
    cnt[1] += c;
    // exclusive prefix of c over the workgroup, then over the tiles
    uint32_t incl = c;
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t v = __shfl_up(incl, d, 64);
      if (lane >= d) incl += v;
    }
    if (lane == 63) sh_w[wv] = incl;
    __syncthreads();
    uint32_t woff = 0, agg = 0;
    for (int k = 0; k < WAVES; ++k) {
      const uint32_t t = sh_w[k];
      if (k < wv) woff += t;
      agg += t;
    }
    if (wv == 0) {
      uint64_t pre = lookback(P.flags[s][0], tile, agg, epoch);
      if (lane == 0) sh_pref = pre;
    }
    __syncthreads();
    // the tile's listed photons go to LDS in list order (with a flag: awake
    // through the last window, i.e. on its list and not just woken) ...
    uint32_t lq = woff + incl - c;
    for (uint32_t m = up; m; m &= m - 1u) {
      const int j = __builtin_ctz(m);
      const int64_t i = i0 + j;
      sh_list[lq] = (uint32_t)i;
      sh_cont[lq++] = (uint8_t)(a.fresh == 0 && ((cont >> j) & 1u) && i < prev_end);
    }
    __syncthreads();
    // ... and are written out one per thread (coalesced).  A photon awake
    // through the last window has its state in that window's cache, at the
    // position it had there; it moves to this window's cache at its new
    // position (both lists are in index order, so the reads are nearly
    // sequential), or to the store if the new position lies past the cache.
    for (uint32_t e = threadIdx.x; e < agg; e += TILE) {
      const int64_t q = (int64_t)sh_pref + e;
      const int64_t i = (int64_t)sh_list[e];
      sp.hot[q] = (uint32_t)i;
      uint32_t ok = WAKE_NEVER;
      if (sh_cont[e]) {
        const uint32_t sl = sp.slot[i];
        if ((int64_t)sl < sp.ccap) {
          if (q < sp.ccap) {
            Cn.x(q) = Co.x(sl); Cn.y(q) = Co.y(sl); Cn.z(q) = Co.z(sl);
            Cn.ux(q) = Co.ux(sl); Cn.uy(q) = Co.uy(sl); Cn.uz(q) = Co.uz(sl);
            Cn.id(q) = Co.id(sl);
            ok = sl;
          } else {
            sp.x[i] = Co.x(sl); sp.y[i] = Co.y(sl); sp.z[i] = Co.z(sl);
          }
        }
      }
      sp.oldk[q] = ok;
      sp.slot[i] = (uint32_t)q;
    }
    __syncthreads();  // sh_w and sh_pref are rewritten by the next tile
  }

  uint64_t tot[2];
  block_sum(cnt, sh_w, tot);
  if (threadIdx.x == 0) {
    count_add(&sp.cnt->psteps, tot[0]);
    count_add(&sp.cnt->slept, tot[0]);
    count_add(&sp.cnt->asleep_n, tot[0]);  // zeroed by this step's prologue
  }
  kernel_exit(scan, nwg, [&]() {
    uint64_t total = ntiles > 0 ? (ld_flag(&P.flags[s][0][ntiles - 1]) & FLAG_VAL_MASK) : 0;
    sp.cnt->n_hot = (int64_t)total;
    sp.cnt->list_end = n_old;
    sp.cnt->cpar ^= 1;  // the new window's cache is the other set
    scan->epoch = epoch;
  });
}

// One conversion of the photon at list position k, recomputed from its
// unchanged state: the pair at the photon's position, electron with delta of
// its momentum, positron with 1 - delta (cpu.py:142-191), at store positions
// posA / posB with ids idA / idB; the photon becomes a hole (cpu.py:189-190),
// not pushed.
template <int KIND, int FM, bool MIX, bool SH, typename Node>
__device__ inline void photon_convert(const Plan& P, const SpeciesDev& sp, const StepArgs& a,
                                      const WinCache& C, const Node* lds_rs,
                                      const double2* lds_od, bool rebuild, int64_t k,
                                      const SpeciesDev* dA, const SpeciesDev* dB, int64_t posA,
                                      int64_t posB, uint64_t idA, uint64_t idB) {
  const int64_t i = (int64_t)sp.hot[k];
  const bool cached = k < sp.ccap;
  double x, y, z, ux, uy, uz;
  bool woke = false;
  const bool moved = rebuild && cached && sp.oldk[k] != WAKE_NEVER;
  if (cached && (!rebuild || moved)) {
    x = C.x(k); y = C.y(k); z = C.z(k);
    ux = C.ux(k); uy = C.uy(k); uz = C.uz(k);
  } else {
    x = sp.x[i]; y = sp.y[i]; z = sp.z[i];
    ux = sp.ux[i]; uy = sp.uy[i]; uz = sp.uz[i];
    woke = rebuild && sp.wake[i] == a.wrel;
  }
  const double ig = sfp_inv_gamma_photon(ux, uy, uz);
  if (woke && sp.push) photon_wake_position(sp, a, i, ig, x, y, z, ux, uy, uz);
  uint32_t shot = 0;
  const sfp_fieldset* FS = fields_of<SH>(P, sp, i, shot);
  double tau = KIND == KIND_OD ? sp.tau[i] : 0.0;
  double delta = 0.0;
  if (MIX) {
    (void)qed_event32<true, KIND, FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, lds_od, x, y, z,
                                      (float)ux, (float)uy, (float)uz, (float)ig, sp.id[i], tau,
                                      delta);
  } else {
    (void)qed_event<true, KIND, FM>(P, FS, sp, a, (const sfp_rs_node*)lds_rs, lds_od, x, y, z, ux,
                                    uy, uz, ig, sp.id[i], tau, delta);
  }
  const double dp = 1.0 - delta;
  if (dA)
    spawn_lepton<FM, MIX>(P, FS, shot, *dA, a, posA, idA, x, y, z, delta * ux, delta * uy,
                          delta * uz);
  if (dB)
    spawn_lepton<FM, MIX>(P, FS, shot, *dB, a, posB, idB, x, y, z, dp * ux, dp * uy, dp * uz);
  sp.pruned[i] = 1;
  sp.wake[i] = WAKE_NEVER;
  if (cached) C.wake(k) = WAKE_NEVER;
}

// Occupancy: 3 waves per SIMD in mixed precision (168 VGPRs; the compiler's
// own choice for some instances is 2), which ran faster than a cap at 128
// VGPRs (4 waves, spilling) and, for optical depth, than the compiler's 2
// waves (a 4e6-electron collision: photon kernel 3.90 -> 3.65 s); 2 waves in
// FP64 (a 1e6-electron focused-pulse collision: 0.26 -> 0.22 s per 3000
// steps against 1 wave).
template <int KIND, int FM, bool MIX, bool SH>
__global__ void __launch_bounds__(TILE, MIX ? 4 : 2) k_photon(const Plan* __restrict__ plan, int s, const StepArgs* __restrict__ ap,
         int rmax) {
  const StepArgs a = *ap;
  const Plan& P = *plan;
  if (P.halt->flag) return;
  const SpeciesDev& sp = P.sp[s];
  using Node = typename std::conditional<MIX, sfp_rs_node32, sfp_rs_node>::type;

  __shared__ Node lds_rs[KIND == KIND_RS ? 1023 : 1];
  __shared__ double2 lds_od[KIND == KIND_OD ? 255 : 1];
  __shared__ float2 lds_env[KIND == KIND_RS ? 1023 : 1];  // rate envelopes of the bound, rounded up
  __shared__ int64_t sh_tile;
  __shared__ uint32_t sh_w[4 * WAVES];
  __shared__ uint64_t sh_pref;
  __shared__ uint64_t sh_mask[QROWS][WAVES];
  __shared__ uint16_t sh_q[QROWS * TILE];  // tile positions (< 2^16)
  __shared__ uint32_t sh_qd[QD_N];

  const int64_t n_list = sp.cnt->n_hot;
  const int rows = qed_rows(n_list, rmax);
  const int64_t SPAN = (int64_t)rows * TILE;
  const int64_t ntiles = (n_list + SPAN - 1) / SPAN;
  // workgroups without work leave before reading the destination counts,
  // which the exit of this launch changes
  const uint32_t nwg = participants(ntiles);
  if (blockIdx.x >= nwg) return;
  const SpeciesDev* dA = sp.pair_e_dst >= 0 ? &P.sp[sp.pair_e_dst] : nullptr;
  const SpeciesDev* dB = sp.pair_p_dst >= 0 ? &P.sp[sp.pair_p_dst] : nullptr;
  const int64_t baseA = dA ? dA->cnt->n : 0, baseB = dB ? dB->cnt->n : 0;
  const int64_t idA = dA ? dA->cnt->next_id : 0, idB = dB ? dB->cnt->next_id : 0;
  // Room in the pair stores: the conversions ranked past it are deferred
  // (the photon is left as it was, marked PRUNED_DEFER) to k_defer, which the
  // host runs after growing the stores by what the step actually converted.
  // The stores thus need room for the conversions, not for one per listed
  // photon (a billion listed photons in a large ensemble).
  int64_t room = INT64_MAX;
  if (dA) room = dA->cap - baseA;
  if (dB && dB->cap - baseB < room) room = dB->cap - baseB;
  room = room < 0 ? 0 : room;
  ScanState* scan = P.scan[s][1];
  const uint32_t epoch = next_epoch(scan);
  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  // Window cache: the list is rebuilt every WIN-th step and kept in between,
  // so the state of its first ccap entries lives in arrays indexed by list
  // position for the window (coalesced).  A rebuild takes a photon that was
  // awake through the last window from that window's cache (the other set,
  // at the position the list kernel found, in the same order: nearly
  // sequential) and gathers the others from the store; appended photons are
  // copied in by the list kernel.  The store gets a photon's position when it
  // falls asleep and, for every cached photon, at the end of the run.
  const bool rebuild = list_rebuild(a);
  const bool last = a.rel + 1 == a.rel_end;
  const WinCache C = win_cache(sp, (uint32_t)sp.cnt->cpar);

  uint32_t cnt[4] = {0, 0, 0, 0};  // alive, processed (live visited), past the bound (id read), multi-step pushes
  qdiag_init(sh_qd);
  QDiag dg{(SFP_L uint32_t*)sh_qd};
  bool staged = false;
  while (true) {
    if (threadIdx.x == 0) sh_tile = (int64_t)atomicAdd(&scan->tile_ctr, 1u);
    __syncthreads();
    const int64_t tile = sh_tile;
    __syncthreads();
    if (tile >= ntiles) break;
    if (!staged) {  // tables go to LDS once the workgroup has work
      stage_tables<KIND, true, MIX>(P, lds_rs, lds_od);
      if (KIND == KIND_RS) {
        for (int64_t k = threadIdx.x; k < P.z_N - 1 && k < 1023; k += TILE) {
          const double2 e = P.rs_env[k];
          lds_env[k] = make_float2(f32_up(e.x), f32_up(e.y));
        }
        __syncthreads();
      }
      staged = true;
    }

    // the list index of the next row is loaded while this row computes
    const int64_t k0 = tile * SPAN + threadIdx.x;
    uint32_t i_next = k0 < n_list ? sp.hot[k0] : 0u;
#pragma unroll 1
    for (int row = 0; row < rows; ++row) {
      const int64_t k = tile * SPAN + (int64_t)row * TILE + threadIdx.x;
      bool live = false;
      const int64_t i = k < n_list ? (int64_t)i_next : 0;
      i_next = (row + 1 < rows && k + TILE < n_list) ? sp.hot[k + TILE] : 0u;
      // the id (the key of the photon's draws; from the window cache where
      // the state comes from it) loads with the state, not after the bound:
      // most visited photons lie past it and draw
      uint64_t pid = 0;
      double x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 0, tau = 0;
      const bool cached = k < sp.ccap;
      bool woke = false;  // asleep until this step
      bool stale = false;  // the store holds an older position
      if (k < n_list) {
        // a list kept from an earlier step holds photons that have fallen
        // asleep since (counted by the list kernel) and holes: skipped
        // at a rebuild, a photon awake through the last window was moved
        // into this window's cache by the list kernel (oldk set)
        const bool moved = rebuild && cached && sp.oldk[k] != WAKE_NEVER;
        if (cached && (!rebuild || moved)) {
          // the cached state loads with the wake step, not after it (an
          // entry asleep since the rebuild is read and dropped)
          const uint32_t wk = C.wake(k);
          const double cx = C.x(k), cy = C.y(k), cz = C.z(k);
          const double cux = C.ux(k), cuy = C.uy(k), cuz = C.uz(k);
          const uint64_t cid = C.id(k);
          if (moved || wk <= a.wrel) {
            x = cx; y = cy; z = cz;
            ux = cux; uy = cuy; uz = cuz;
            pid = cid;
            live = true;
            stale = moved;
          }
        } else {
          // the id loads with the wake step (most visited photons draw)
          const uint32_t wk = sp.wake[i];
          const uint64_t sid = sp.id[i];
          pid = sid;
          if (wk <= a.wrel) {
            const uint8_t pr = sp.pruned[i];
            x = sp.x[i]; y = sp.y[i]; z = sp.z[i];
            ux = sp.ux[i]; uy = sp.uy[i]; uz = sp.uz[i];
            live = pr == 0;
            woke = live && rebuild && wk == a.wrel;
            if (!live) {
              sp.wake[i] = WAKE_NEVER;  // a hole: never visited again
              if (cached) C.wake(k) = WAKE_NEVER;
            }
          } else if (cached) {
            C.wake(k) = WAKE_NEVER;  // asleep at the rebuild: not visited in this window
          }
        }
      }
      const double ig = live ? sfp_inv_gamma_photon(ux, uy, uz) : 1.0;
      if (woke && sp.push) photon_wake_position(sp, a, i, ig, x, y, z, ux, uy, uz);
      uint32_t shot = 0;
      const sfp_fieldset* FS = live ? fields_of<SH>(P, sp, i, shot) : &P.fields;
      bool hot = false;
      double pb = -1.0;
      if (live && sp.push) {
        float S = sfp_field_bound_f32<FM>(FS, x, y, z, a.t);
        float chi_b = (float)P.chi_coef * S * sfp32_div(1.0f, (float)ig) * 1.001f;
        if (KIND == KIND_RS) {
          pb = rs_pbound(P, true, chi_b, ig, a.dt, false, lds_env);
          hot = pb < 0.0 || !(2.0 * pb < 0x1p-53);
        } else {
          hot = qed_possible(P, true, chi_b, ig, a.dt);
        }
      }
      bool event = false;
      double delta = 0.0;
      // Rejection sampling: a pair needs r2 < p <= pb.  The photon's own r2
      // (the same Philox words as the exact front) at or above 2 pb rules
      // it out without the exact field, chi and rate.
      bool exact = hot;
      if (KIND == KIND_RS && hot && pb >= 0.0) {
        const sfp_u32x4 d = sfp_draw4(P.seed, pid, a.step, (uint32_t)sp.tag, 0u);
        const double r2 = MIX ? (double)sfp32_u01(d.v[2]) : sfp_u01(d.v[2], d.v[3]);
        exact = r2 < 2.0 * pb;
      }
      if (exact) {
        if (KIND == KIND_OD) tau = sp.tau[i];
        if (MIX) {
          event = qed_event32<true, KIND, FM>(P, FS, sp, a, (const sfp_rs_node32*)lds_rs, lds_od, x,
                                              y, z, (float)ux, (float)uy, (float)uz, (float)ig,
                                              pid, tau, delta, 0.0f, 0.0f, 0.0f, &dg);
        } else {
          event = qed_event<true, KIND, FM>(P, FS, sp, a, (const sfp_rs_node*)lds_rs, lds_od, x, y,
                                            z, ux, uy, uz, ig, pid, tau, delta, 0.0, 0.0, 0.0,
                                            &dg);
        }
      }
      if (!dA && !dB) event = false;
      const uint64_t m = __ballot(event);
      if (lane == 0) sh_mask[row][wv] = m;

      if (live) {
        cnt[1] += 1;
        cnt[2] += hot ? 1u : 0u;
      }
      if (live && !event) {
        uint32_t nst = 1;
        if (!sp.push) {
          nst = a.rel_end - a.rel;
        } else if (!hot && rebuild) {
          nst = sleep_steps<FM>(P, FS, a, true, x, y, z, ux, uy, uz, ig);
        } else if (KIND == KIND_RS && rebuild && pb >= 0.0) {
          nst = sleep_steps_drawn<FM, MIX>(P, FS, a, x, y, z, ux, uy, uz, ig, pid,
                                           (uint32_t)sp.tag, lds_env);
        }
        if (sp.push) {
          if (nst == 1) {
            // cpu.py:5-19 twice per step: x += (c dt/2 / |u|) ux, the same
            // increment each time
            const double hx = a.cdt_half * ig * ux, hy = a.cdt_half * ig * uy,
                         hz = a.cdt_half * ig * uz;
            x += hx; x += hx;
            y += hy; y += hy;
            z += hz; z += hz;
            if (!cached || last) {
              sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
            }
            if (cached) {
              C.x(k) = x; C.y(k) = y; C.z(k) = z;
              if (rebuild) {
                // a photon moved from the last window's cache has its
                // momentum there already (moved by the list kernel)
                if (!stale) {
                  C.ux(k) = ux; C.uy(k) = uy; C.uz(k) = uz;
                  C.id(k) = pid;
                }
                C.wake(k) = 0u;
              }
            }
          } else if (woke || stale) {
            // falls asleep: the store keeps this step's position (the
            // increments are added when it wakes, or at the end of the run)
            sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
          }
        }
        if (hot && KIND == KIND_OD) sp.tau[i] = tau;
        if (nst > 1) {
          sp.wake[i] = a.wrel + nst;
          sp.since[i] = a.wrel;
          if (cached) C.wake(k) = a.wrel + nst;
          cnt[3] += 1;
        }
        cnt[0] += 1;
      }
    }

    // the tile's conversions in list order (photon_convert); those ranked
    // past the room wait for k_defer, the photon left as it was
    const uint32_t nev = tile_events(sh_mask, rows, P.flags[s][1], tile, epoch, sh_q, &sh_pref);
    for (uint32_t e = threadIdx.x; e < nev; e += TILE) {
      const int64_t k = tile * SPAN + (int64_t)sh_q[e];
      const int64_t rank = (int64_t)sh_pref + (int64_t)e;
      if (rank >= room) {
        sp.pruned[sp.hot[k]] = PRUNED_DEFER;
        // the tile's first deferred conversion records its list position
        if (rank == (room > (int64_t)sh_pref ? room : (int64_t)sh_pref))
          atomicMax((unsigned long long*)&sp.cnt->defer_kneg, (unsigned long long)~(uint64_t)k);
        continue;
      }
      photon_convert<KIND, FM, MIX, SH>(P, sp, a, C, lds_rs, lds_od, rebuild, k, dA, dB, baseA + rank,
                                    baseB + rank, (uint64_t)(idA + rank), (uint64_t)(idB + rank));
    }
  }

  uint64_t tot[4];
  block_sum(cnt, sh_w, tot);
  if (threadIdx.x == 0) {
    count_add(&sp.cnt->psteps, tot[0]);
    count_add(&sp.cnt->processed, tot[1]);
    count_add(&sp.cnt->hot_total, tot[2]);
    count_add(&sp.cnt->sleeps, tot[3]);
    count_add(&sp.cnt->asleep_n, tot[3]);
  }
  (void)qdiag_flush(sp.cnt, sh_qd);
  kernel_exit(scan, nwg, [&]() {
    finish_appends(P, s, 1, ntiles, sp, dA, dB, baseA, baseB, idA, idB, room, a.step);
    scan->epoch = epoch;
  });
}

// Deferred conversions of a step (after k_photon halted it for want of room
// and the host grew the pair stores by what it deferred): the photons marked
// PRUNED_DEFER, from the first one's list position on, ranked in list order
// behind the conversions k_photon appended -- the positions and ids they
// would have had -- and converted from their unchanged state as k_photon
// converts.  The step then continues with the next kernel.
template <int KIND, int FM, bool MIX, bool SH>
__global__ void __launch_bounds__(TILE) k_defer(const Plan* __restrict__ plan, int s,
                                                const StepArgs* __restrict__ ap) {
  const StepArgs a = *ap;
  const Plan& P = *plan;
  if (P.halt->flag) return;
  const SpeciesDev& sp = P.sp[s];
  using Node = typename std::conditional<MIX, sfp_rs_node32, sfp_rs_node>::type;
  __shared__ Node lds_rs[KIND == KIND_RS ? 1023 : 1];
  __shared__ double2 lds_od[KIND == KIND_OD ? 255 : 1];
  __shared__ int64_t sh_tile;
  __shared__ uint64_t sh_pref;
  __shared__ uint64_t sh_mask[1][WAVES];
  __shared__ uint16_t sh_q[TILE];

  const int64_t n_list = sp.cnt->n_hot;
  const int64_t n_def = sp.cnt->defer_n;
  const uint64_t kneg = sp.cnt->defer_kneg;
  const int64_t k0 = kneg ? (int64_t)~kneg : n_list;
  const int64_t ntiles = k0 < n_list ? (n_list - k0 + TILE - 1) / TILE : 0;
  const uint32_t nwg = participants(ntiles);
  if (blockIdx.x >= nwg) return;
  const SpeciesDev* dA = sp.pair_e_dst >= 0 ? &P.sp[sp.pair_e_dst] : nullptr;
  const SpeciesDev* dB = sp.pair_p_dst >= 0 ? &P.sp[sp.pair_p_dst] : nullptr;
  if (!room_or_halt(P, s, a, dA, dB, n_def)) return;
  const int64_t baseA = dA ? dA->cnt->n : 0, baseB = dB ? dB->cnt->n : 0;
  const int64_t idA = dA ? dA->cnt->next_id : 0, idB = dB ? dB->cnt->next_id : 0;
  ScanState* scan = P.scan[s][2];
  const uint32_t epoch = next_epoch(scan);
  const int lane = wave_lane();
  const int wv = threadIdx.x / 64;
  const bool rebuild = list_rebuild(a);
  const WinCache C = win_cache(sp, (uint32_t)sp.cnt->cpar);
  bool staged = false;
  while (true) {
    if (threadIdx.x == 0) sh_tile = (int64_t)atomicAdd(&scan->tile_ctr, 1u);
    __syncthreads();
    const int64_t tile = sh_tile;
    __syncthreads();
    if (tile >= ntiles) break;
    if (!staged) {
      stage_tables<KIND, true, MIX>(P, lds_rs, lds_od);
      staged = true;
    }
    const int64_t k = k0 + tile * TILE + threadIdx.x;
    const bool mark = k < n_list && sp.pruned[sp.hot[k]] == PRUNED_DEFER;
    const uint64_t m = __ballot(mark);
    if (lane == 0) sh_mask[0][wv] = m;
    const uint32_t nev = tile_events(sh_mask, 1, P.flags[s][2], tile, epoch, sh_q, &sh_pref);
    for (uint32_t e = threadIdx.x; e < nev; e += TILE) {
      const int64_t kk = k0 + tile * TILE + (int64_t)sh_q[e];
      const int64_t rank = (int64_t)sh_pref + (int64_t)e;
      photon_convert<KIND, FM, MIX, SH>(P, sp, a, C, lds_rs, lds_od, rebuild, kk, dA, dB,
                                    baseA + rank, baseB + rank, (uint64_t)(idA + rank),
                                    (uint64_t)(idB + rank));
    }
  }
  kernel_exit(scan, nwg, [&]() {
    finish_appends(P, s, 2, ntiles, sp, dA, dB, baseA, baseB, idA, idB);
    sp.cnt->defer_n = 0;
    sp.cnt->defer_kneg = 0;
    sp.cnt->deferrals += 1;
    scan->epoch = epoch;
  });
}

// prologue of a step: snapshot the counts the species kernels iterate over
// (and, on the steps where photons fall asleep and wake, restart the count of
// sleeping photons that the list kernel rebuilds)
__device__ inline void step_snapshot(const Plan& P, const StepArgs& a) {
  const int s = threadIdx.x;
  if (s < P.nspecies) {
    Counts* c = P.sp[s].cnt;
    c->n_old = c->n;
    if (list_rebuild(a)) c->asleep_n = 0;
  }
}
__global__ void k_step_prologue(const Plan* plan, const StepArgs* __restrict__ ap) {
  const StepArgs a = *ap;
  if (plan->halt->flag) return;
  step_snapshot(*plan, a);
}

// end of a step: the recorder's row when this step is one it records
// (particles past the species' count are recorded as NaN)
__global__ void k_record(const Plan* __restrict__ plan, const StepArgs* __restrict__ ap) {
  const StepArgs a = *ap;
  const Plan& P = *plan;
  if (P.halt->flag || !P.rec.on) return;
  const Recorder& R = P.rec;
  if (a.step + 1 <= R.step0) return;
  const uint64_t k1 = a.step + 1 - R.step0;
  if (k1 % (uint64_t)R.every) return;
  const int64_t r = (int64_t)(k1 / (uint64_t)R.every) - 1;
  if (r >= R.max_rows) return;
  const SpeciesDev& sp = P.sp[R.species];
  const int64_t n = sp.cnt->n;
  const double nan = __builtin_nan("");
  double* row = R.buf + r * R.n * 6;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < R.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const bool in = i < n;
    row[6 * i + 0] = in ? sp.x[i] : nan;
    row[6 * i + 1] = in ? sp.y[i] : nan;
    row[6 * i + 2] = in ? sp.z[i] : nan;
    row[6 * i + 3] = in ? sp.ux[i] : nan;
    row[6 * i + 4] = in ? sp.uy[i] : nan;
    row[6 * i + 5] = in ? sp.uz[i] : nan;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) R.t[r] = a.t + a.dt;
}

// end of a run: QED photons asleep past its last step get the position of
// that step (the store must be current when the run returns), and keep
// sleeping from it
__device__ inline bool catchup_one(uint32_t w, uint32_t& s0, uint8_t pr, double& x, double& y,
                                   double& z, double ux, double uy, double uz, uint32_t wend,
                                   double cdt_half) {
  if (w == WAKE_NEVER || w < wend || s0 >= wend || pr) return false;
  const double ig = sfp_inv_gamma_photon(ux, uy, uz);
  const uint32_t m = 2u * (wend - s0);
  x = sfp_repeat_add(x, cdt_half * ig * ux, m);
  y = sfp_repeat_add(y, cdt_half * ig * uy, m);
  z = sfp_repeat_add(z, cdt_half * ig * uz, m);
  s0 = wend;
  return true;
}

// One photon's catch-up when it lags by at most SFP_REPEAT_LOOP additions
// (one by one) or each coordinate's additions stay inside its binade
// (sfp_repeat_add_binade): 0 nothing to do, 1 done, 2 the photon needs the
// general repeated addition (a binade edge, a tie, a coordinate near 0).
__device__ inline int catchup_binade(uint32_t w, uint32_t& s0, uint8_t pr, double& x, double& y,
                                     double& z, double ux, double uy, double uz, uint32_t wend,
                                     double cdt_half) {
  if (w == WAKE_NEVER || w < wend || s0 >= wend || pr) return 0;
  const double ig = sfp_inv_gamma_photon(ux, uy, uz);
  const uint32_t m = 2u * (wend - s0);
  if (m <= SFP_REPEAT_LOOP) {  // a short lag: the additions one by one
    const double hx = cdt_half * ig * ux, hy = cdt_half * ig * uy, hz = cdt_half * ig * uz;
    for (uint32_t k = 0; k < m; ++k) {
      x = x + hx;
      y = y + hy;
      z = z + hz;
    }
    s0 = wend;
    return 1;
  }
  double x1, y1, z1;
  if (!sfp_repeat_add_binade(x, cdt_half * ig * ux, m, &x1) ||
      !sfp_repeat_add_binade(y, cdt_half * ig * uy, m, &y1) ||
      !sfp_repeat_add_binade(z, cdt_half * ig * uz, m, &z1))
    return 2;
  x = x1; y = y1; z = z1;
  s0 = wend;
  return 1;
}

// Two consecutive photons per thread, every load issued up front (the wake
// and since steps, the pruned flags, position and momentum, as 8- and
// 16-byte vectors, so that each load instruction of a wave reads one
// contiguous stretch); the state of photons that need no update is read and
// dropped (most photons sleep at the end of a run).  A pair with an update
// writes its positions and since steps back whole, as vectors.  Lanes try the
// one-binade closed form only, which keeps the kernel at the speed of its
// memory traffic; the photons that need the general repeated addition (a
// binade edge, a tie, a coordinate near 0: a few per cent) are marked in a
// bit mask, two 64-bit ballots per wave and iteration (the wave's first
// and second photons of its 64 pairs), written by the wave's first lane --
// no atomics, and every entry below the count is rewritten by each
// catch-up, so none is cleared -- and k_catchup_slow advances them.  (With
// the general loop inlined as the fallback, the few lanes that took it held
// whole waves in it and the kernel ran at half the speed; a list with one
// atomic per wave on a shared counter was slower still.)  The masks live in
// the visit list's old-position array (oldk: written and read only on the
// steps that rebuild the list, so free between run calls, while the list
// itself is kept for a call that continues this one), whose 4 bytes per
// photon of capacity (in whole multiples of 64 entries, so 16-byte aligned)
// hold the masks' 1/8 byte.  All arrays are allocated for the store's
// capacity, and a full pair lies below the count.
__global__ void __launch_bounds__(256) k_catchup(const Plan* __restrict__ plan, int s,
                                                 uint32_t wend, double cdt_half) {
  const SpeciesDev& sp = plan->sp[s];
  const int64_t n = sp.cnt->n;
  const int64_t npair = (n + 1) / 2;
  uint4* masks = (uint4*)sp.oldk;
  // g of a wave's first lane is a multiple of 64 (workgroups of 256, a
  // stride of whole workgroups): the wave covers pairs g .. g + 63
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < npair;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i0 = 2 * g;
    int r0 = 0, r1 = 0;
    if (i0 + 2 <= n) {
      const uint2 w = *(const uint2*)(sp.wake + i0);
      uint2 s0 = *(const uint2*)(sp.since + i0);
      const uchar2 pr = *(const uchar2*)(sp.pruned + i0);
      double2 x = *(const double2*)(sp.x + i0);
      double2 y = *(const double2*)(sp.y + i0);
      double2 z = *(const double2*)(sp.z + i0);
      const double2 a = *(const double2*)(sp.ux + i0);
      const double2 b = *(const double2*)(sp.uy + i0);
      const double2 c = *(const double2*)(sp.uz + i0);
      r0 = catchup_binade(w.x, s0.x, pr.x, x.x, y.x, z.x, a.x, b.x, c.x, wend, cdt_half);
      r1 = catchup_binade(w.y, s0.y, pr.y, x.y, y.y, z.y, a.y, b.y, c.y, wend, cdt_half);
      if (r0 == 1 || r1 == 1) {
        *(double2*)(sp.x + i0) = x;
        *(double2*)(sp.y + i0) = y;
        *(double2*)(sp.z + i0) = z;
        *(uint2*)(sp.since + i0) = s0;
      }
    }  // else: the last photon of an odd count, left to k_catchup_slow
    const uint64_t m0 = __ballot(r0 == 2), m1 = __ballot(r1 == 2);
    if (wave_lane() == 0)
      masks[g >> 6] = make_uint4((uint32_t)m0, (uint32_t)(m0 >> 32), (uint32_t)m1,
                                 (uint32_t)(m1 >> 32));
  }
}

__device__ inline void catchup_general(const SpeciesDev& sp, int64_t i, uint32_t wend,
                                       double cdt_half) {
  uint32_t s1 = sp.since[i];
  double x = sp.x[i], y = sp.y[i], z = sp.z[i];
  if (catchup_one(sp.wake[i], s1, sp.pruned[i], x, y, z, sp.ux[i], sp.uy[i], sp.uz[i], wend,
                  cdt_half)) {
    sp.x[i] = x; sp.y[i] = y; sp.z[i] = z;
    sp.since[i] = s1;
  }
}

// The photons k_catchup marked, with the general repeated addition, and the
// last photon of an odd count.  A wave reads 64 mask entries at once (lane l
// entry jb + l), then takes the entries with a mark one after the other with
// all its lanes, lane l the pair 64 j + l as k_catchup's lane l had it: after
// a long run most photons are marked (one thread per entry walked 128 of them
// in turn: 59 ms for 2.5e8 photons), after a short one almost none (a wave
// per entry read the 1.2e6 entries of 1.5e8 photons one by one: 69 us).
__global__ void __launch_bounds__(256) k_catchup_slow(const Plan* __restrict__ plan, int s,
                                                      uint32_t wend, double cdt_half) {
  const SpeciesDev& sp = plan->sp[s];
  const int64_t n = sp.cnt->n;
  const int64_t nm = ((n + 1) / 2 + 63) / 64;
  const uint4* masks = (const uint4*)sp.oldk;
  const int lane = wave_lane();
  const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / 64;
  for (int64_t jb = 64 * w0; jb < nm; jb += 64 * nw) {
    const uint4 mk = jb + lane < nm ? masks[jb + lane] : make_uint4(0u, 0u, 0u, 0u);
    uint64_t todo = __ballot((mk.x | mk.y | mk.z | mk.w) != 0u);
    while (todo) {
      const int l = (int)__builtin_ctzll(todo);
      todo &= todo - 1ull;
      const uint64_t m0 = (uint64_t)(uint32_t)__shfl((int)mk.x, l) |
                          ((uint64_t)(uint32_t)__shfl((int)mk.y, l) << 32);
      const uint64_t m1 = (uint64_t)(uint32_t)__shfl((int)mk.z, l) |
                          ((uint64_t)(uint32_t)__shfl((int)mk.w, l) << 32);
      const int64_t j = jb + l;
      if ((m0 >> lane) & 1ull) catchup_general(sp, 2 * (64 * j + lane), wend, cdt_half);
      if ((m1 >> lane) & 1ull) catchup_general(sp, 2 * (64 * j + lane) + 1, wend, cdt_half);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) catchup_general(sp, n - 1, wend, cdt_half);
}

// ------------------------------------------------------------- dispatch ----
typedef void (*kfn)(const Plan*, int, const StepArgs*, int);

template <int FM, bool MIX, bool SH>
static kfn sweep_fn(bool photon, bool qed, bool pair) {
  if (pair) return k_sweep_pair<FM, MIX, SH>;
  if (photon) return qed ? k_awake : k_sweep<true, false, FM, false, false>;
  return qed ? k_sweep<false, true, FM, MIX, SH> : k_sweep<false, false, FM, MIX, SH>;
}

template <int FM, bool MIX, bool SH>
static kfn hot_fn(bool photon, int kind, int netk) {
  if (photon)
    return kind == KIND_OD ? k_photon<KIND_OD, FM, MIX, SH> : k_photon<KIND_RS, FM, MIX, SH>;
  if (MIX && netk == 2 && kind == KIND_RS) return k_hot<KIND_RS, FM, MIX, (MIX ? 2 : 0), SH>;
  if (MIX && netk == 1)
    return kind == KIND_OD ? k_hot<KIND_OD, FM, MIX, (MIX ? 1 : 0), SH>
                           : k_hot<KIND_RS, FM, MIX, (MIX ? 1 : 0), SH>;
  return kind == KIND_OD ? k_hot<KIND_OD, FM, MIX, 0, SH> : k_hot<KIND_RS, FM, MIX, 0, SH>;
}

template <bool MIX, bool SH>
static kfn pick_sweep_p(int fm, bool photon, bool qed, bool pair) {
  if (fm == SFP_FM_GAUSSIAN) return sweep_fn<SFP_FM_GAUSSIAN, MIX, SH>(photon, qed, pair);
  if (fm == SFP_FM_PLANE) return sweep_fn<SFP_FM_PLANE, MIX, SH>(photon, qed, pair);
  return sweep_fn<SFP_FM_ANY, MIX, SH>(photon, qed, pair);
}

template <bool MIX, bool SH>
static kfn pick_hot_p(int fm, bool photon, int kind, int netk) {
  if (fm == SFP_FM_GAUSSIAN) return hot_fn<SFP_FM_GAUSSIAN, MIX, SH>(photon, kind, netk);
  if (fm == SFP_FM_PLANE) return hot_fn<SFP_FM_PLANE, MIX, SH>(photon, kind, netk);
  return hot_fn<SFP_FM_ANY, MIX, SH>(photon, kind, netk);
}

// sh: an ensemble (a field set per shot); pair: the two-species lepton sweep
static kfn pick_sweep(int fm, bool photon, bool qed, bool mixed, bool sh, bool pair = false) {
  if (sh)
    return mixed ? pick_sweep_p<true, true>(fm, photon, qed, pair)
                 : pick_sweep_p<false, true>(fm, photon, qed, pair);
  return mixed ? pick_sweep_p<true, false>(fm, photon, qed, pair)
               : pick_sweep_p<false, false>(fm, photon, qed, pair);
}

static kfn pick_hot(int fm, bool photon, int kind, bool mixed, int netk, bool sh) {
  if (sh)
    return mixed ? pick_hot_p<true, true>(fm, photon, kind, netk)
                 : pick_hot_p<false, true>(fm, photon, kind, 0);
  return mixed ? pick_hot_p<true, false>(fm, photon, kind, netk)
               : pick_hot_p<false, false>(fm, photon, kind, 0);
}

typedef void (*kfn_defer)(const Plan*, int, const StepArgs*);

template <int FM, bool SH>
static kfn_defer defer_fn(int kind, bool mixed) {
  if (mixed) return kind == KIND_OD ? k_defer<KIND_OD, FM, true, SH> : k_defer<KIND_RS, FM, true, SH>;
  return kind == KIND_OD ? k_defer<KIND_OD, FM, false, SH> : k_defer<KIND_RS, FM, false, SH>;
}

template <bool SH>
static kfn_defer pick_defer_p(int fm, int kind, bool mixed) {
  if (fm == SFP_FM_GAUSSIAN) return defer_fn<SFP_FM_GAUSSIAN, SH>(kind, mixed);
  if (fm == SFP_FM_PLANE) return defer_fn<SFP_FM_PLANE, SH>(kind, mixed);
  return defer_fn<SFP_FM_ANY, SH>(kind, mixed);
}

static kfn_defer pick_defer(int fm, int kind, bool mixed, bool sh) {
  return sh ? pick_defer_p<true>(fm, kind, mixed) : pick_defer_p<false>(fm, kind, mixed);
}

template <bool SH>
static kfn pick_emit_p(int fm, int kind) {
  if (fm == SFP_FM_GAUSSIAN)
    return kind == KIND_OD ? k_emit<KIND_OD, SFP_FM_GAUSSIAN, SH> : k_emit<KIND_RS, SFP_FM_GAUSSIAN, SH>;
  if (fm == SFP_FM_PLANE)
    return kind == KIND_OD ? k_emit<KIND_OD, SFP_FM_PLANE, SH> : k_emit<KIND_RS, SFP_FM_PLANE, SH>;
  return kind == KIND_OD ? k_emit<KIND_OD, SFP_FM_ANY, SH> : k_emit<KIND_RS, SFP_FM_ANY, SH>;
}

static kfn pick_emit(int fm, int kind, bool sh) {
  return sh ? pick_emit_p<true>(fm, kind) : pick_emit_p<false>(fm, kind);
}

// ------------------------------------------------------------- host side ---
static constexpr int MAX_CAND = 16;
static constexpr int64_t SYNC_EVERY = 64;

// events around one kernel launch: for the caller's timing and/or for the
// launch-shape tuner (candidate index, or -1)
struct EvRec {
  hipEvent_t e0, e1;
  int cand;
  bool timed;
};

// Launch shape of a persistent kernel: workgroups, and the most rows of 256
// particles a tile may take (kernels without row tiles ignore it)
struct Shape {
  int grid, rows;
};

struct Sim {
  int nsp = 0;
  std::vector<sfp_species_desc> desc;
  std::vector<Store*> stores;
  std::vector<int> qed;            // species has a QED process
  kfn sweep[SFP_MAX_SPECIES] = {};
  kfn sweep2[SFP_MAX_SPECIES] = {};  // leptons with QED: the sweep of this species and the next
  kfn hot[SFP_MAX_SPECIES] = {};
  kfn emit[SFP_MAX_SPECIES] = {};
  kfn_defer defer[SFP_MAX_SPECIES] = {};  // photons: deferred conversions
  // two lepton species in one QED launch (k_hot, k_emit): fuse[s] is the
  // species launched with s (-1: none), inner[s] marks a species whose QED
  // kernels run in the launch of the species before it; fuse_on: allowed
  int fuse[SFP_MAX_SPECIES] = {-1, -1, -1, -1, -1, -1, -1, -1};
  bool inner[SFP_MAX_SPECIES] = {};
  bool fuse_on = true;
  // launch shapes: candidate shapes of each (species, phase) kernel, the one
  // in use, and a grid and tile rows the caller pinned (0: tuned)
  std::vector<Shape> cand[SFP_MAX_SPECIES][NPH];
  int choice[SFP_MAX_SPECIES][NPH] = {};
  int pinned[SFP_MAX_SPECIES][NPH] = {};
  int pinned_rows[SFP_MAX_SPECIES][NPH] = {};
  double tune_ms[SFP_MAX_SPECIES][NPH][MAX_CAND] = {};
  int tune_cnt[SFP_MAX_SPECIES][NPH][MAX_CAND] = {};
  int64_t tune_n[SFP_MAX_SPECIES][NPH] = {};  // work of the kernel when last tuned (+1)
  bool tuned = false;
  int64_t tune_batches = 0;
  // the step's stream, the arguments of one batch of steps (device slots,
  // pinned staging) and the batch captured as a graph
  hipStream_t stream = nullptr;
  StepArgs* args_dev = nullptr;
  StepArgs* args_host = nullptr;
  hipGraphExec_t graph = nullptr;
  bool graph_stale = true;
  int64_t graph_batches = 0, captures = 0;
  std::vector<int64_t> flags_cap[NPH];
  std::vector<int64_t> hot_cap;
  std::vector<uint64_t*> flags[NPH];
  std::vector<uint32_t*> hot_mem;
  std::vector<EmitRec*> emit_mem;   // leptons with the learned sampler
  std::vector<int64_t> emit_cap;
  std::vector<uint32_t*> wake_mem;  // photon species: multi-step push bookkeeping
  std::vector<uint32_t*> since_mem;  // QED photon species: step of the stored position
  // the sleep clock: QED photons keep sleeping across run calls while each
  // call continues the previous one and no store was changed from the host
  uint64_t wake_origin = 0, wake_end = 0;
  bool wake_valid = false;
  int64_t clock_restarts = 0;
  // the photon visit lists and window caches of the last run call hold for a
  // call that continues it (list_kept, with the stores' run counts at its end)
  bool list_kept = false;
  std::vector<uint64_t> list_runs;
  int64_t list_keeps = 0;
  // sleeping photons hold the positions of their sleep start until settled
  // (by the next run that continues this one, or before any host access)
  bool lag = false;
  uint32_t lag_wend = 0;
  std::vector<uint64_t> wake_version;
  std::vector<int64_t> wake_cap;
  std::vector<char*> cache_mem;     // photon species: window cache of the visit list
  std::vector<int64_t> cache_cap;
  ScanState* scan_mem = nullptr;
  Halt* halt = nullptr;            // device
  Halt* halt_host = nullptr;       // pinned
  Counts* counts_host = nullptr;   // pinned [nsp]
  Plan plan{};
  Plan* plan_dev = nullptr;
  Tables* tables = nullptr;
  sfp_sim_params params{};
  int fm = SFP_FM_ANY;
  bool timing = false;
  int64_t timing_every = 1;         // with timing: the steps whose index it divides
  int time_s = -1, time_ph = -1;    // with timing: a kernel timed at every step as well
  std::vector<EvRec> ev_pending[SFP_MAX_SPECIES][NPH];
  double ms_acc[SFP_MAX_SPECIES][NPH] = {};
  int64_t launches[SFP_MAX_SPECIES][NPH] = {};
  std::vector<hipEvent_t> ev_pool;
  int64_t syncs = 0, replays = 0;
  int cus = 0;                       // compute units of the device
  bool net = false;                  // leptons use the learned photon-energy sampler
  uint64_t rec_end = 0;              // recorder: one past the last step run
  // steps run since the look-back flags and epochs were last cleared; each
  // step advances an epoch by at most one
  int64_t steps_since_clear = 0;
  int64_t clear_every = (int64_t)1 << 28;  // well inside the 29-bit epoch
  bool rate_net = false;             // spin leptons use the learned emission rate
  sfp_fieldset* shot_fields = nullptr;  // device, one per shot
  int32_t nshots = 0;
};

// The counts of every species and the halt record, copied behind the step's
// kernels into pinned host memory with one wait (a synchronous copy each
// took a host round trip of its own)
static int sim_read_counts(Sim* sim) {
  for (int s = 0; s < sim->nsp; ++s)
    HIPCHK(hipMemcpyAsync(&sim->counts_host[s], sim->stores[s]->cnt, sizeof(Counts),
                          hipMemcpyDeviceToHost, sim->stream));
  HIPCHK(hipMemcpyAsync(sim->halt_host, sim->halt, sizeof(Halt), hipMemcpyDeviceToHost,
                        sim->stream));
  HIPCHK(hipStreamSynchronize(sim->stream));
  sim->syncs++;
  return SFP_OK;
}

static int sim_refresh_plan(Sim* sim) {
  Plan& P = sim->plan;
  for (int s = 0; s < sim->nsp; ++s) {
    Store* st = sim->stores[s];
    SpeciesDev& d = P.sp[s];
    d.x = gptr((double*)st->p[SFP_X]);
    d.y = gptr((double*)st->p[SFP_Y]);
    d.z = gptr((double*)st->p[SFP_Z]);
    d.ux = gptr((double*)st->p[SFP_UX]);
    d.uy = gptr((double*)st->p[SFP_UY]);
    d.uz = gptr((double*)st->p[SFP_UZ]);
    d.id = gptr((uint64_t*)st->p[SFP_ID]);
    d.pruned = gptr((uint8_t*)st->p[SFP_PRUNED]);
    d.tau = gptr((double*)st->p[SFP_TAU]);
    d.sx = gptr((double*)st->p[SFP_SX]);
    d.sy = gptr((double*)st->p[SFP_SY]);
    d.sz = gptr((double*)st->p[SFP_SZ]);
    d.cnt = gptr(st->cnt);
    d.cap = st->cap;
    // shots take effect once their fields are set
    d.shot = gptr(sim->shot_fields ? (uint32_t*)st->p[SFP_SHOT] : nullptr);
    if (st->cap > (int64_t)UINT32_MAX)
      return set_err(SFP_E_ARG, "species %d: capacity above 2^32 particles", s);
    for (int ph = 0; ph < NPH; ++ph) {
      int64_t need = st->cap / TILE + 2;
      if (need > sim->flags_cap[ph][s]) {
        if (sim->flags[ph][s]) HIPCHK(hipFree(sim->flags[ph][s]));
        int64_t want = need + need / 2;
        HIPCHK(hipMalloc(&sim->flags[ph][s], sizeof(uint64_t) * want));
        HIPCHK(hipMemset(sim->flags[ph][s], 0, sizeof(uint64_t) * want));
        sim->flags_cap[ph][s] = want;
      }
      P.flags[s][ph] = gptr(sim->flags[ph][s]);
      P.scan[s][ph] = gptr(sim->scan_mem + NPH * s + ph);
    }
    if (sim->qed[s] && st->cap > sim->hot_cap[s]) {
      // photons keep their visit list from step to step: grown inside a run,
      // the list keeps its entries; the photons' old cache positions (oldk)
      // follow the list in the same block; whole multiples of 64 entries, so
      // that the catch-up's masks (16 bytes per 128 photons, in oldk) are
      // 16-byte aligned and fit a small store's
      int64_t want = (std::max<int64_t>(st->cap, 64) + 63) & ~(int64_t)63;
      uint32_t* h = nullptr;
      HIPCHK(hipMalloc(&h, sizeof(uint32_t) * 2 * want));
      if (sim->hot_mem[s]) {
        const int64_t hc = sim->hot_cap[s];
        HIPCHK(hipMemcpy(h, sim->hot_mem[s], sizeof(uint32_t) * hc, hipMemcpyDeviceToDevice));
        HIPCHK(hipMemcpy(h + want, sim->hot_mem[s] + hc, sizeof(uint32_t) * hc,
                         hipMemcpyDeviceToDevice));
        HIPCHK(hipFree(sim->hot_mem[s]));
      }
      sim->hot_mem[s] = h;
      sim->hot_cap[s] = want;
    }
    d.hot = gptr(sim->hot_mem[s]);
    d.oldk = gptr(sim->hot_mem[s] ? sim->hot_mem[s] + sim->hot_cap[s] : nullptr);
    if (sim->emit[s] && st->cap > sim->emit_cap[s]) {
      // one step's emission candidates: written by k_hot, read by k_emit
      if (sim->emit_mem[s]) HIPCHK(hipFree(sim->emit_mem[s]));
      sim->emit_mem[s] = nullptr;
      const int64_t want = std::max<int64_t>(st->cap, 1);
      HIPCHK(hipMalloc(&sim->emit_mem[s], sizeof(EmitRec) * want));
      sim->emit_cap[s] = want;
    }
    d.emit = gptr(sim->emit[s] ? sim->emit_mem[s] : nullptr);
    if (d.photon && st->cap > sim->wake_cap[s]) {
      // grown inside a run: sleeping photons keep their wake steps; whole
      // tiles of k_awake, so that its vector loads stay inside the array
      int64_t want = (st->cap + WAKE_PAD - 1) / WAKE_PAD * WAKE_PAD;
      uint32_t* w = nullptr;
      HIPCHK(hipMalloc(&w, sizeof(uint32_t) * want));
      HIPCHK(hipMemset(w, 0, sizeof(uint32_t) * want));
      if (sim->wake_mem[s]) {
        HIPCHK(hipMemcpy(w, sim->wake_mem[s], sizeof(uint32_t) * sim->wake_cap[s],
                         hipMemcpyDeviceToDevice));
        HIPCHK(hipFree(sim->wake_mem[s]));
      }
      // since and slot share one block: [since | slot]
      uint32_t* sn = nullptr;
      HIPCHK(hipMalloc(&sn, sizeof(uint32_t) * 2 * want));
      HIPCHK(hipMemset(sn, 0, sizeof(uint32_t) * 2 * want));
      if (sim->since_mem[s]) {
        const int64_t wc = sim->wake_cap[s];
        HIPCHK(hipMemcpy(sn, sim->since_mem[s], sizeof(uint32_t) * wc, hipMemcpyDeviceToDevice));
        HIPCHK(hipMemcpy(sn + want, sim->since_mem[s] + wc, sizeof(uint32_t) * wc,
                         hipMemcpyDeviceToDevice));
        HIPCHK(hipFree(sim->since_mem[s]));
      }
      sim->wake_mem[s] = w;
      sim->since_mem[s] = sn;
      sim->wake_cap[s] = want;
    }
    d.wake = gptr(sim->wake_mem[s]);
    d.since = gptr(sim->since_mem[s]);
    d.slot = gptr(sim->since_mem[s] ? sim->since_mem[s] + sim->wake_cap[s] : nullptr);
    {
      const int64_t cc = sim->cache_cap[s];
      double* c = (double*)sim->cache_mem[s];
      d.cache = gptr(c);
      d.ccap = c ? cc : 0;
    }
  }
  HIPCHK(hipMemcpy(sim->plan_dev, &P, sizeof(Plan), hipMemcpyHostToDevice));
  return SFP_OK;
}

// Kernel of every (species, phase) for the field mode, precision and sampler,
// with a persistent grid of (workgroups per CU at full occupancy) x CUs.
static int sim_pick_kernels(Sim* sim) {
  if (!sim->cus) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    sim->cus = prop.multiProcessorCount;
  }
  const bool mixed = sim->params.precision == SFP_PRECISION_MIXED;
  const int kind = sim->params.qed_model == SFP_QED_OD ? KIND_OD : KIND_RS;
  const bool sh = sim->shot_fields != nullptr;
  for (int s = 0; s < sim->nsp; ++s) {
    const bool photon = sim->plan.sp[s].photon;
    sim->sweep[s] = pick_sweep(sim->fm, photon, sim->qed[s], mixed, sh);
    sim->sweep2[s] = (!photon && sim->qed[s]) ? pick_sweep(sim->fm, false, true, mixed, sh, true)
                                              : nullptr;
    // learned models: the rate network for spin-carrying leptons under
    // spin-resolved emission, else the photon-energy sampler
    int netk = 0;
    if (!photon && sim->rate_net && sim->plan.spin_emission && sim->plan.sp[s].spin &&
        kind == KIND_RS)
      netk = 2;
    else if (!photon && sim->net)
      netk = 1;
    sim->hot[s] = sim->qed[s] ? pick_hot(sim->fm, photon, kind, mixed, netk, sh) : nullptr;
    sim->emit[s] = (sim->qed[s] && mixed && netk == 1) ? pick_emit(sim->fm, kind, sh) : nullptr;
    sim->defer[s] = (sim->qed[s] && photon) ? pick_defer(sim->fm, kind, mixed, sh) : nullptr;
    for (int ph = 0; ph < NPH; ++ph) {
      kfn fn = ph == 0 ? sim->sweep[s] : (ph == 1 ? sim->hot[s] : sim->emit[s]);
      std::vector<Shape>& c = sim->cand[s][ph];
      c.clear();
      sim->choice[s][ph] = 0;
      if (!fn) continue;
      // most tile rows: the sweep kernels (not the photon wake-list kernel)
      // and the QED kernels with row tiles (not the rate network's one-row
      // tiles); 0 where the kernel has none
      const bool awake = photon && sim->qed[s];
      const int rmax = ph == 0 ? (awake ? 0 : ROWS) : (ph == 1 && netk != 2 ? QROWS : 0);
      if (sim->pinned[s][ph] > 0) {
        const int r = sim->pinned_rows[s][ph];
        c.push_back({sim->pinned[s][ph], rmax == 0 ? 0 : (r > 0 ? std::min(r, rmax) : rmax)});
        continue;
      }
      // Tiles are taken from an atomic counter, so the grid only sets how
      // many workgroups compete for them: full occupancy first (the default),
      // then fewer per CU (shorter look-back chains and less contention when
      // the work is small); and per grid, the kernel's most rows per tile,
      // then half and a quarter of it (more tiles, so fewer idle workgroups
      // in the last round, at one look-back per tile).  Results depend on
      // neither.
      int per_cu = 0;
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, TILE, 0));
      for (int k = std::max(per_cu, 1);; k /= 2) {
        for (int r = rmax, q = 0; q < 3 && (int)c.size() < MAX_CAND; r /= 2, ++q) {
          c.push_back({k * sim->cus, r});
          if (r <= 1) break;
        }
        if (k == 1 || (int)c.size() >= MAX_CAND) break;
      }
    }
  }
  // Two consecutive lepton species with the same QED kernels and photon store
  // share their QED launches (the electrons and positrons of a collision):
  // the second's launch shapes are unused
  for (int s = 0; s < SFP_MAX_SPECIES; ++s) {
    sim->fuse[s] = -1;
    sim->inner[s] = false;
  }
  if (sim->fuse_on)
    for (int s = 0; s + 1 < sim->nsp; ++s) {
      const int b = s + 1;
      if (sim->inner[s] || sim->plan.sp[s].photon || sim->plan.sp[b].photon || !sim->hot[s] ||
          sim->hot[b] != sim->hot[s] || sim->emit[b] != sim->emit[s] ||
          sim->desc[b].photon_dst != sim->desc[s].photon_dst)
        continue;
      sim->fuse[s] = b;
      sim->inner[b] = true;
      sim->cand[b][0].clear();
      sim->cand[b][1].clear();
      sim->cand[b][2].clear();
    }
  sim->tuned = false;
  sim->graph_stale = true;
  return SFP_OK;
}

static Shape sim_shape(const Sim* sim, int s, int ph) { return sim->cand[s][ph][sim->choice[s][ph]]; }

extern "C" int sfp_sim_create(const sfp_species_desc* species, int32_t nspecies,
                              const sfp_field* terms, int32_t nterms, void* tables,
                              const sfp_sim_params* params, void** out) {
  if (nspecies < 1 || nspecies > SFP_MAX_SPECIES)
    return set_err(SFP_E_ARG, "between 1 and %d species are supported", SFP_MAX_SPECIES);
  if (params->qed_model != SFP_QED_RS && params->qed_model != SFP_QED_OD)
    return set_err(SFP_E_ARG, "unknown QED model %d", params->qed_model);
  if (params->precision != SFP_PRECISION_FP64 && params->precision != SFP_PRECISION_MIXED)
    return set_err(SFP_E_ARG, "unknown precision %d", params->precision);
  Sim* sim = new Sim();
  sim->nsp = nspecies;
  sim->params = *params;
  sim->tables = (Tables*)tables;
  sim->desc.assign(species, species + nspecies);
  Plan& P = sim->plan;
  memset(&P, 0, sizeof(P));
  CHK(sfp_compile_fields(terms, nterms, &P.fields));
  sim->fm = sfp_field_mode(&P.fields);
  P.nspecies = nspecies;
  P.seed = params->seed;
  P.threshold = params->photon_threshold;
  P.qed = params->qed_model;
  P.chi_coef = SFP_E * SFP_HBAR / ((SFP_ME * SFP_ME) * (SFP_C * SFP_C * SFP_C));
  bool any_qed = false;
  for (int s = 0; s < nspecies; ++s) {
    const sfp_species_desc& d = species[s];
    if (!d.store) return set_err(SFP_E_ARG, "species %d has no store", s);
    Store* st = (Store*)d.store;
    sim->stores.push_back(st);
    bool photon = d.m == 0.0;
    for (int dst : {d.photon_dst, d.pair_e_dst, d.pair_p_dst})
      if (dst >= nspecies || dst == s)
        return set_err(SFP_E_ARG, "species %d: destination %d invalid", s, dst);
    if (!photon && (d.pair_e_dst >= 0 || d.pair_p_dst >= 0))
      return set_err(SFP_E_ARG, "species %d: only photons create pairs", s);
    if (photon && d.photon_dst >= 0)
      return set_err(SFP_E_ARG, "species %d: photons do not radiate photons", s);
    bool qed = photon ? (d.pair_e_dst >= 0 || d.pair_p_dst >= 0) : (d.photon_dst >= 0);
    if (qed && params->qed_model == SFP_QED_OD && !st->p[SFP_TAU])
      return set_err(SFP_E_ARG, "species %d uses the optical-depth model but has no tau", s);
    if (!photon && d.spin && !st->p[SFP_SX])
      return set_err(SFP_E_ARG, "species %d is a spin species without spin storage", s);
    sim->qed.push_back(qed ? 1 : 0);
    any_qed |= qed;
    SpeciesDev& sd = P.sp[s];
    sd.q = d.q;
    sd.ae = d.ae;
    sd.rr = d.rr;
    sd.push = d.push;
    sd.spin = photon ? 0 : d.spin;
    sd.photon = photon;
    sd.tag = s;
    sd.qed = qed;
    sd.photon_dst = d.photon_dst;
    sd.pair_e_dst = d.pair_e_dst;
    sd.pair_p_dst = d.pair_p_dst;
  }
  if (any_qed && params->qed_model == SFP_QED_RS) {
    if (!sim->tables || !sim->tables->has_rs)
      return set_err(SFP_E_ARG, "rejection-sampling tables required");
    if (sim->tables->z_N - 1 > 1023)
      return set_err(SFP_E_ARG, "rejection-sampling table larger than 1024 nodes");
    P.rs_nodes = gptr(sim->tables->rs_nodes);
    P.rs_nodes32 = gptr((const sfp_rs_node32*)sim->tables->rs_nodes32);
    P.rs_env = gptr(sim->tables->rs_env);
    P.rs_ai = gptr(sim->tables->rs_ai);
    P.rs_ai32 = gptr(sim->tables->rs_ai32);
    P.rs_env_spin = gptr(sim->tables->rs_env_spin);
    P.rate_unit = SFP_ALPHA * SFP_ME * (SFP_C * SFP_C) / SFP_HBAR;
    P.z_N = sim->tables->z_N;
    P.z_min = sim->tables->z_min;
    P.z_max = sim->tables->z_max;
    P.z_delta = sim->tables->z_delta;
    P.zmax15 = sim->tables->z_max * sqrt(sim->tables->z_max);
  }
  if (params->spin_emission) {
    if (params->qed_model != SFP_QED_RS)
      return set_err(SFP_E_ARG, "spin-resolved emission is defined for rejection sampling");
    if (!sim->tables || !sim->tables->rs_ai)
      return set_err(SFP_E_ARG, "spin-resolved emission needs the Ai table");
    P.spin_emission = 1;
  }
  if (any_qed && params->qed_model == SFP_QED_OD) {
    if (!sim->tables || !sim->tables->has_od)
      return set_err(SFP_E_ARG, "optical-depth tables required");
    if (sim->tables->od.chi_N - 1 > 255)
      return set_err(SFP_E_ARG, "optical-depth table larger than 256 chi nodes");
  }
  if (sim->tables && sim->tables->has_od && sim->tables->od.chi_N - 1 <= 255) {
    // (also the learned sampler's total rate and fallback under rejection sampling)
    P.od = sim->tables->od;
    P.od_photon_nodes = gptr(sim->tables->od_photon_nodes);
    P.od_pair_nodes = gptr(sim->tables->od_pair_nodes);
  }
  HIPCHK(hipMalloc(&sim->plan_dev, sizeof(Plan)));
  // default priority; a blocking stream, so the synchronous copies of the
  // host side order with it
  HIPCHK(hipStreamCreate(&sim->stream));
  HIPCHK(hipMalloc(&sim->args_dev, sizeof(StepArgs) * SYNC_EVERY));
  HIPCHK(hipHostMalloc((void**)&sim->args_host, sizeof(StepArgs) * SYNC_EVERY, 0));
  HIPCHK(hipMalloc(&sim->scan_mem, sizeof(ScanState) * NPH * nspecies));
  HIPCHK(hipMemset(sim->scan_mem, 0, sizeof(ScanState) * NPH * nspecies));
  HIPCHK(hipMalloc(&sim->halt, sizeof(Halt)));
  HIPCHK(hipMemset(sim->halt, 0, sizeof(Halt)));
  HIPCHK(hipHostMalloc((void**)&sim->halt_host, sizeof(Halt), 0));
  HIPCHK(hipHostMalloc((void**)&sim->counts_host, sizeof(Counts) * nspecies, 0));
  P.halt = gptr(sim->halt);
  for (int ph = 0; ph < NPH; ++ph) {
    sim->flags[ph].assign(nspecies, nullptr);
    sim->flags_cap[ph].assign(nspecies, 0);
  }
  sim->hot_mem.assign(nspecies, nullptr);
  sim->emit_mem.assign(nspecies, nullptr);
  sim->emit_cap.assign(nspecies, 0);
  sim->hot_cap.assign(nspecies, 0);
  sim->wake_mem.assign(nspecies, nullptr);
  sim->since_mem.assign(nspecies, nullptr);
  sim->wake_version.assign(nspecies, 0);
  sim->list_runs.assign(nspecies, 0);
  sim->wake_cap.assign(nspecies, 0);
  sim->cache_mem.assign(nspecies, nullptr);
  sim->cache_cap.assign(nspecies, 0);
  CHK(sim_pick_kernels(sim));
  CHK(sim_refresh_plan(sim));
  *out = sim;
  return SFP_OK;
}

extern "C" int sfp_sim_set_sampler(void* p, void* mlp) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  if (sim->params.precision != SFP_PRECISION_MIXED)
    return set_err(SFP_E_ARG, "the learned sampler runs in the mixed-precision step");
  if (mlp && sim->plan.spin_emission)
    for (int s = 0; s < sim->nsp; ++s)
      if (sim->plan.sp[s].spin && sim->desc[s].photon_dst >= 0)
        return set_err(SFP_E_ARG, "the learned sampler draws the unpolarised spectrum: not with "
                                  "spin-resolved emission");
  if (!sim->tables || !sim->tables->has_od || !sim->plan.od_photon_nodes)
    return set_err(SFP_E_ARG, "the learned sampler needs the optical-depth tables");
  Plan& P = sim->plan;
  if (mlp) {
    int32_t nin = 0, np = 0;
    CHK(sfp_mlp_shape(mlp, &nin, &np));
    if (nin != 2) return set_err(SFP_E_ARG, "not a photon-energy sampler");
    const void *dev = nullptr, *bins = nullptr;
    CHK(sfp_mlp_device(mlp, &dev, &bins));
    P.net = gptr((const sfp_mlp_dev*)dev);
    P.net_bins = gptr((const sfp_mlp_bins*)bins);
  } else {
    P.net = nullptr;
    P.net_bins = nullptr;
  }
  sim->net = mlp != nullptr;
  CHK(sim_pick_kernels(sim));
  CHK(sim_refresh_plan(sim));  // the emission queue of the kernels picked
  return SFP_OK;
}

extern "C" int sfp_sim_set_rate_net(void* p, void* mlp) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  Plan& P = sim->plan;
  if (mlp) {
    if (sim->params.precision != SFP_PRECISION_MIXED)
      return set_err(SFP_E_ARG, "the learned rate runs in the mixed-precision step");
    if (!P.spin_emission || sim->params.qed_model != SFP_QED_RS)
      return set_err(SFP_E_ARG, "the learned rate needs spin-resolved emission under rejection "
                                "sampling");
    int32_t nin = 0, np = 0;
    CHK(sfp_mlp_shape(mlp, &nin, &np));
    if (nin != 3) return set_err(SFP_E_ARG, "not an emission-rate network");
    const void *dev = nullptr, *bins = nullptr;
    CHK(sfp_mlp_device(mlp, &dev, &bins));
    P.rate_net = gptr((const sfp_mlp_dev*)dev);
    P.rate_bins = gptr((const sfp_mlp_bins*)bins);
  } else {
    P.rate_net = nullptr;
    P.rate_bins = nullptr;
  }
  sim->rate_net = mlp != nullptr;
  CHK(sim_pick_kernels(sim));
  CHK(sim_refresh_plan(sim));
  return SFP_OK;
}

extern "C" int sfp_sim_set_shots(void* p, const sfp_field* terms, const int32_t* nterms,
                                 int32_t nshots) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  if (nshots < 1 || nshots > (1 << 20)) return set_err(SFP_E_ARG, "between 1 and 2^20 shots");
  sim->wake_valid = false;  // sleeps were proven against the old fields
  std::vector<sfp_fieldset> fs(nshots);
  int mode = -1;
  int64_t off = 0;
  for (int32_t k = 0; k < nshots; ++k) {
    CHK(sfp_compile_fields(terms + off, nterms[k], &fs[k]));
    off += nterms[k];
    const int m = sfp_field_mode(&fs[k]);
    mode = (mode < 0 || mode == m) ? m : SFP_FM_ANY;
  }
  if (sim->shot_fields) HIPCHK(hipFree(sim->shot_fields));
  HIPCHK(hipMalloc(&sim->shot_fields, sizeof(sfp_fieldset) * nshots));
  HIPCHK(hipMemcpy(sim->shot_fields, fs.data(), sizeof(sfp_fieldset) * nshots,
                   hipMemcpyHostToDevice));
  sim->nshots = nshots;
  sim->plan.shot_fields = gptr(sim->shot_fields);
  sim->plan.nshots = (uint32_t)nshots;
  // one kernel instance for all shots: the common mode, or the general one
  sim->fm = mode == SFP_FM_ANY ? SFP_FM_ANY : (sfp_field_mode(&sim->plan.fields) == mode ? mode
                                                                                   : SFP_FM_ANY);
  CHK(sim_pick_kernels(sim));
  CHK(sim_refresh_plan(sim));
  return SFP_OK;
}

static int sim_settle(Sim* sim);

extern "C" int sfp_sim_destroy(void* p) {
  Sim* sim = (Sim*)p;
  if (!sim) return SFP_OK;
  (void)sim_settle(sim);  // the stores outlive the simulation: positions current
  (void)hipDeviceSynchronize();
  for (int ph = 0; ph < NPH; ++ph)
    for (auto f : sim->flags[ph])
      if (f) (void)hipFree(f);
  for (auto h : sim->hot_mem)
    if (h) (void)hipFree(h);
  for (auto q : sim->emit_mem)
    if (q) (void)hipFree(q);
  for (auto w : sim->wake_mem)
    if (w) (void)hipFree(w);
  for (auto w : sim->since_mem)
    if (w) (void)hipFree(w);
  for (auto c : sim->cache_mem)
    if (c) (void)hipFree(c);
  if (sim->scan_mem) (void)hipFree(sim->scan_mem);
  if (sim->plan_dev) (void)hipFree(sim->plan_dev);
  if (sim->halt) (void)hipFree(sim->halt);
  if (sim->halt_host) (void)hipHostFree(sim->halt_host);
  if (sim->shot_fields) (void)hipFree(sim->shot_fields);
  if (sim->counts_host) (void)hipHostFree(sim->counts_host);
  if (sim->graph) (void)hipGraphExecDestroy(sim->graph);
  if (sim->args_dev) (void)hipFree(sim->args_dev);
  if (sim->args_host) (void)hipHostFree(sim->args_host);
  if (sim->plan.rec.buf) (void)hipFree(sim->plan.rec.buf);
  if (sim->plan.rec.t) (void)hipFree(sim->plan.rec.t);
  if (sim->stream) (void)hipStreamDestroy(sim->stream);
  for (int s = 0; s < SFP_MAX_SPECIES; ++s)
    for (int ph = 0; ph < NPH; ++ph)
      for (auto& e : sim->ev_pending[s][ph]) {
        (void)hipEventDestroy(e.e0);
        (void)hipEventDestroy(e.e1);
      }
  for (auto e : sim->ev_pool) (void)hipEventDestroy(e);
  delete sim;
  return SFP_OK;
}

static hipEvent_t sim_event(Sim* sim) {
  if (!sim->ev_pool.empty()) {
    hipEvent_t e = sim->ev_pool.back();
    sim->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Launch the kernel of (species s, phase ph) for the step in argument slot j.
// tune >= 0: a tuning launch of the batch's step `tune`, run with candidate
// shape tune mod candidates and timed for the tuner -- except on the photon
// list's rebuild steps (every WIN-th step and a fresh run's first), which cost
// several times a step in between: the few of them in a batch would fall to
// some candidates only, so the tuner compares the steps in between.
static int launch(Sim* sim, int s, int ph, int j, int tune, bool opens = false) {
  const bool pair = sim->fuse[s] >= 0;  // the next species' kernels run in this species' launches
  kfn fn = ph == 0 ? (pair ? sim->sweep2[s] : sim->sweep[s]) : (ph == 1 ? sim->hot[s] : sim->emit[s]);
  const std::vector<Shape>& cand = sim->cand[s][ph];
  const int c = tune >= 0 ? tune % (int)cand.size() : sim->choice[s][ph];
  const StepArgs& ah = sim->args_host[j];
  if (tune >= 0 && (ah.wrel % WIN == 0 || ah.fresh)) tune = -2;  // run, not timed
  const bool timed = sim->timing && (ah.step % (uint64_t)sim->timing_every == 0 ||
                                     (s == sim->time_s && ph == sim->time_ph));
  const bool ev = timed || tune >= 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ev) {
    e0 = sim_event(sim);
    e1 = sim_event(sim);
    if (e0) HIPCHK(hipEventRecord(e0, sim->stream));
  }
  // the kernels of two species take the second in bits 4-7 (plus one); a
  // sweep that opens the step, SEL_PROLOGUE
  const int sel = (pair ? (s | ((sim->fuse[s] + 1) << 4)) : s) | (opens ? SEL_PROLOGUE : 0);
  hipLaunchKernelGGL(fn, dim3(cand[c].grid), dim3(TILE), 0, sim->stream, (const Plan*)sim->plan_dev,
                     sel, (const StepArgs*)(sim->args_dev + j), cand[c].rows);
  HIPCHK(hipGetLastError());
  if (ev && e0 && e1) {
    HIPCHK(hipEventRecord(e1, sim->stream));
    sim->ev_pending[s][ph].push_back({e0, e1, tune >= 0 ? c : -1, timed});
  }
  return SFP_OK;
}

// Launch the kernels of the step in slot j from (species s0, phase ph0) on.
static int launch_step(Sim* sim, int j, int s0, int ph0, int tune) {
  // the step's count snapshot: in the first sweep, unless the first species
  // is a QED photon species (its list kernel adds to a count the snapshot
  // resets)
  const bool opens = s0 == 0 && ph0 == 0;
  const bool merged = opens && !(sim->plan.sp[0].photon && sim->qed[0]);
  if (opens && !merged) {
    hipLaunchKernelGGL(k_step_prologue, dim3(1), dim3(64), 0, sim->stream,
                       (const Plan*)sim->plan_dev, (const StepArgs*)(sim->args_dev + j));
    HIPCHK(hipGetLastError());
  }
  for (int s = s0; s < sim->nsp; ++s) {
    if (sim->inner[s]) continue;  // launched with the species before it
    int first = (s == s0) ? ph0 : 0;
    // with a partner (fuse), one launch sweeps both species
    if (first <= 0) CHK(launch(sim, s, 0, j, tune, merged && s == 0));
    if (first <= 1 && sim->hot[s]) CHK(launch(sim, s, 1, j, tune));
    if (first <= 2 && sim->emit[s]) CHK(launch(sim, s, 2, j, tune));
  }
  if (sim->plan.rec.on) {
    const int g = (int)std::min<int64_t>((sim->plan.rec.n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_record, dim3(std::max(g, 1)), dim3(256), 0, sim->stream,
                       (const Plan*)sim->plan_dev, (const StepArgs*)(sim->args_dev + j));
    HIPCHK(hipGetLastError());
  }
  return SFP_OK;
}

// Capture one batch of SYNC_EVERY steps (slots 0..SYNC_EVERY-1) as a graph.
// The kernels read their step from the slots and the plan from device
// memory, so the graph stays valid while the kernels and grids in use do.
static int sim_capture(Sim* sim) {
  if (sim->graph) {
    HIPCHK(hipGraphExecDestroy(sim->graph));
    sim->graph = nullptr;
  }
  HIPCHK(hipStreamBeginCapture(sim->stream, hipStreamCaptureModeThreadLocal));
  int rc = SFP_OK;
  for (int j = 0; j < SYNC_EVERY && rc == SFP_OK; ++j) rc = launch_step(sim, j, 0, 0, -1);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(sim->stream, &g);
  if (rc != SFP_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_err(SFP_E_HIP, "graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&sim->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_err(SFP_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
  sim->graph_stale = false;
  sim->captures++;
  return SFP_OK;
}

// After a halt: grow the destinations of the halted species by the room the
// halted kernel asks for -- one append per listed particle (the lepton QED
// kernel halts before it starts), or the conversions the photon kernel
// deferred (it ran through and halted at its end: *deferred, the caller runs
// k_defer) -- and clear the flag; the caller replays.
static int sim_recover(Sim* sim, int s, bool* deferred) {
  const sfp_species_desc& d = sim->desc[s];
  const Counts& c = sim->counts_host[s];
  *deferred = c.defer_n > 0;
  // a shared QED launch asks for room for both species' lists
  const int64_t extra = *deferred ? c.defer_n
                                  : c.n_hot + (sim->fuse[s] >= 0 ? sim->counts_host[sim->fuse[s]].n_hot : 0);
  for (int dst : {d.photon_dst, d.pair_e_dst, d.pair_p_dst}) {
    if (dst < 0) continue;
    Store* st = sim->stores[dst];
    const int64_t need = sim->counts_host[dst].n + extra;
    if (need > st->cap) {
      // that room and a quarter more, so that growing lists do not halt
      // every step; where HBM cannot hold that, a sixteenth more
      int64_t want = need + need / 4 + 4096;
      if (sfp_store_reserve(st, want) != SFP_OK) {
        (void)hipGetLastError();
        want = need + need / 16 + 4096;
        CHK(sfp_store_reserve(st, want));
      }
    }
  }
  Halt clear{};
  HIPCHK(hipMemcpy(sim->halt, &clear, sizeof(Halt), hipMemcpyHostToDevice));
  CHK(sim_refresh_plan(sim));
  sim->replays++;
  return SFP_OK;
}

// Wait for the stream and fold the recorded kernel times into the caller's
// timing and the tuner's sums.
static int sim_collect_events(Sim* sim) {
  HIPCHK(hipStreamSynchronize(sim->stream));
  for (int s = 0; s < sim->nsp; ++s)
    for (int ph = 0; ph < NPH; ++ph) {
      for (auto& e : sim->ev_pending[s][ph]) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e.e0, e.e1));
        if (e.timed) {
          sim->ms_acc[s][ph] += ms;
          sim->launches[s][ph] += 1;
        }
        if (e.cand >= 0) {
          sim->tune_ms[s][ph][e.cand] += ms;
          sim->tune_cnt[s][ph][e.cand] += 1;
        }
        sim->ev_pool.push_back(e.e0);
        sim->ev_pool.push_back(e.e1);
      }
      sim->ev_pending[s][ph].clear();
    }
  return SFP_OK;
}

// Work of a kernel as the tuner sees it: particles swept, or listed.
static int64_t sim_work(const Sim* sim, int s, int ph) {
  const Counts& c = sim->counts_host[s];
  return 1 + (ph == 0 ? c.n : (ph == 1 ? c.n_hot : c.n_emit));
}

// The tuner runs a batch again when the work of any kernel has changed by
// more than a factor of four since it last did.
static bool sim_needs_tune(const Sim* sim) {
  if (!sim->tuned) return true;
  for (int s = 0; s < sim->nsp; ++s)
    for (int ph = 0; ph < NPH; ++ph) {
      if (sim->cand[s][ph].size() < 2) continue;
      const int64_t w = sim_work(sim, s, ph), w0 = sim->tune_n[s][ph];
      if (w > 4 * w0 || w0 > 4 * w) return true;
    }
  return false;
}

// After a tuning batch (steps interleaved over the candidates): keep the
// candidate with the least time per launch if it beats the default shape
// (candidate 0: full occupancy, the kernel's most rows) by more than 3%,
// else the default.  A candidate has 5-30 timed launches per batch, whose
// mean varies by a few percent from batch to batch; without the margin the
// tuner followed that noise to shapes that ran the whole step 1-3% slower
// (a 1e7-electron collision, 128 steps about its centre).
static int sim_tune_pick(Sim* sim) {
  CHK(sim_collect_events(sim));
  for (int s = 0; s < sim->nsp; ++s)
    for (int ph = 0; ph < NPH; ++ph) {
      const int nc = (int)sim->cand[s][ph].size();
      int best = sim->choice[s][ph];
      double best_ms = 0.0;
      bool any = false;
      for (int c = 0; c < nc; ++c) {
        if (!sim->tune_cnt[s][ph][c]) continue;
        const double m = sim->tune_ms[s][ph][c] / sim->tune_cnt[s][ph][c];
        if (!any || m < best_ms) {
          best = c;
          best_ms = m;
          any = true;
        }
      }
      if (any && best != 0 && sim->tune_cnt[s][ph][0] &&
          best_ms > 0.97 * sim->tune_ms[s][ph][0] / sim->tune_cnt[s][ph][0])
        best = 0;
      if (best != sim->choice[s][ph]) sim->graph_stale = true;
      sim->choice[s][ph] = best;
      sim->tune_n[s][ph] = sim_work(sim, s, ph);
      for (int c = 0; c < MAX_CAND; ++c) {
        sim->tune_ms[s][ph][c] = 0.0;
        sim->tune_cnt[s][ph][c] = 0;
      }
    }
  sim->tuned = true;
  sim->tune_batches++;
  return SFP_OK;
}

// QED photons asleep past the last step run get that step's position
// (k_catchup); their sleep goes on from it
static int sim_settle(Sim* sim) {
  if (!sim->lag) return SFP_OK;
  sim->lag = false;
  const double cdt_half = SFP_C * (0.5 * sim->params.dt);
  for (int s = 0; s < sim->nsp; ++s) {
    if (!sim->wake_mem[s] || !sim->qed[s] || !sim->plan.sp[s].push) continue;
    hipLaunchKernelGGL(k_catchup, dim3(sim->cus * 8), dim3(256), 0, sim->stream,
                       (const Plan*)sim->plan_dev, s, sim->lag_wend, cdt_half);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_catchup_slow, dim3(sim->cus * 8), dim3(256), 0, sim->stream,
                       (const Plan*)sim->plan_dev, s, sim->lag_wend, cdt_half);
    HIPCHK(hipGetLastError());
    Store* st = sim->stores[s];
    if (st->settle_sim == sim) {
      st->settle = nullptr;
      st->settle_sim = nullptr;
    }
  }
  HIPCHK(hipStreamSynchronize(sim->stream));
  return SFP_OK;
}

static int sim_settle_cb(void* p) { return sim_settle((Sim*)p); }

// Zero every look-back flag and scan state (epochs back to 0) between
// batches, with nothing in flight (a synchronous fill, never inside a
// captured graph), before any epoch could wrap onto a stale flag.
static int sim_clear_flags(Sim* sim) {
  HIPCHK(hipStreamSynchronize(sim->stream));
  for (int s = 0; s < sim->nsp; ++s)
    for (int ph = 0; ph < NPH; ++ph)
      if (sim->flags[ph][s])
        HIPCHK(hipMemset(sim->flags[ph][s], 0, sizeof(uint64_t) * sim->flags_cap[ph][s]));
  HIPCHK(hipMemset(sim->scan_mem, 0, sizeof(ScanState) * NPH * sim->nsp));
  sim->steps_since_clear = 0;
  return SFP_OK;
}

extern "C" int sfp_sim_set_flag_clear_every(void* p, int64_t steps) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  if (steps < 1 || steps > ((int64_t)1 << 28))
    return set_err(SFP_E_ARG, "flag clear interval must be in [1, 2^28] steps");
  sim->clear_every = steps;
  return SFP_OK;
}

extern "C" int sfp_sim_set_shared_launches(void* p, int32_t on) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  sim->fuse_on = on != 0;
  return sim_pick_kernels(sim);
}

extern "C" int sfp_sim_run(void* p, double* t, uint64_t step0, int64_t nsteps, int32_t flags) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  sim->timing = (flags & SFP_RUN_TIMING) != 0;
  const bool use_graph = (flags & SFP_RUN_GRAPH) != 0 && !sim->timing;
  const bool autotune = (flags & SFP_RUN_AUTOTUNE) != 0;
  // stores may have been resized or refilled by the caller since the last run
  bool dirty = false;
  for (int s = 0; s < sim->nsp; ++s)
    if (sim->plan.sp[s].cap != sim->stores[s]->cap || sim->plan.sp[s].x != sim->stores[s]->p[SFP_X])
      dirty = true;
  // the sleep clock goes on if this run continues the last one and no store
  // changed in between (it restarts, waking every photon, otherwise)
  bool keep = sim->wake_valid && !dirty && step0 == sim->wake_end &&
              step0 + (uint64_t)std::max<int64_t>(nsteps, 0) - sim->wake_origin < (1ull << 31);
  for (int s = 0; s < sim->nsp; ++s)
    if (sim->stores[s]->version != sim->wake_version[s]) keep = false;
  sim->wake_valid = false;
  // the visit lists go on as well when no other simulation stepped the
  // stores since, every QED photon species moves (photons that do not sleep
  // to the end of each call) and no window cache is reallocated below
  bool keep_list = keep && sim->list_kept;
  for (int s = 0; s < sim->nsp; ++s) {
    if (sim->stores[s]->runs != sim->list_runs[s]) keep_list = false;
    if (sim->plan.sp[s].photon && sim->qed[s] && !sim->plan.sp[s].push) keep_list = false;
  }
  // ... and only where a rebuild at the call's first step could not pay for
  // itself.  A rebuild drops from the lists the photons appended since the
  // last one that can sleep; that pays when half a window of steps or more
  // follows it before the next rebuild of the clock (at the collision centre
  // of the 1e7-electron benchmark a rebuild at a call's first step cost
  // 0.62 ms and made the 14 steps after it 1.05 ms faster), so such a call
  // rebuilds, while shorter ones (a step or a few per host callback) keep
  // the lists.  Results are the same either way.
  if (keep_list && nsteps >= (int64_t)(WIN / 2) &&
      WIN - (step0 - sim->wake_origin) % WIN >= WIN / 2)
    keep_list = false;
  sim->list_kept = false;
  if (!keep) CHK(sim_settle(sim));  // the restarted clock needs current positions
  // From here the run tracks every photon itself: this simulation's store
  // callbacks (installed at the end of the last run) come off, so that a
  // store grown during the run does not apply the old run's catch-up to
  // photons that have moved since; another simulation's pending catch-up on a
  // shared store runs now.
  for (int s = 0; s < sim->nsp; ++s) {
    Store* st = sim->stores[s];
    if (st->settle_sim == sim) {
      st->settle = nullptr;
      st->settle_sim = nullptr;
    } else if (st->settle) {
      CHK(store_settle(st));
    }
  }
  sim->lag = false;
  if (nsteps < 0 || nsteps >= (int64_t)1 << 31)
    return set_err(SFP_E_ARG, "nsteps must be in [0, 2^31) per run call");
  // window cache of the photon visit lists: a quarter of the store (longer
  // lists keep their tail in the store), and no more than a quarter of the
  // HBM free at sizing (a store of billions of photons leaves no room for a
  // cache of a quarter of it); sized at the start of a run only, so that it
  // stays valid while stores grow inside the run
  for (int s = 0; s < sim->nsp; ++s) {
    if (!sim->plan.sp[s].photon || !sim->qed[s]) continue;
    const int64_t want = sim->stores[s]->cap / 4;
    if (want > sim->cache_cap[s]) {
      // two sets (this window's and the last one's) of WIN_WORDS words per entry
      const int64_t per = 2 * WIN_WORDS * (int64_t)sizeof(double);
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      const int64_t limit = (int64_t)((fr + (size_t)(sim->cache_cap[s] * per)) / 4 / (size_t)per);
      const int64_t cc = std::min<int64_t>(want + want / 2, limit);
      // grown only by a quarter or more (the limit moves with free memory)
      if (cc > sim->cache_cap[s] + sim->cache_cap[s] / 4) {
        if (sim->cache_mem[s]) HIPCHK(hipFree(sim->cache_mem[s]));
        sim->cache_mem[s] = nullptr;
        sim->cache_cap[s] = 0;
        HIPCHK(hipMalloc(&sim->cache_mem[s], (size_t)cc * (size_t)per));
        sim->cache_cap[s] = cc;
        dirty = true;
        keep_list = false;
      }
    }
  }
  if (dirty) CHK(sim_refresh_plan(sim));
  // Photons without QED sleep to the end of each run, so all of them are
  // visited at its first step; QED photons too when the sleep clock restarts
  if (!keep) {
    sim->wake_origin = step0;
    sim->clock_restarts++;
  }
  for (int s = 0; s < sim->nsp; ++s) {
    if (!sim->wake_mem[s] || (keep && sim->qed[s])) continue;
    HIPCHK(hipMemsetAsync(sim->wake_mem[s], 0, sizeof(uint32_t) * sim->wake_cap[s], sim->stream));
    HIPCHK(hipMemsetAsync(sim->since_mem[s], 0, sizeof(uint32_t) * sim->wake_cap[s], sim->stream));
  }
  if (autotune) CHK(sim_read_counts(sim));
  const double dt = sim->params.dt;
  int64_t k = 0;
  double tt = *t;
  while (k < nsteps) {
    const int64_t batch = std::min<int64_t>(SYNC_EVERY, nsteps - k);
    if (sim->steps_since_clear + batch > sim->clear_every) CHK(sim_clear_flags(sim));
    sim->steps_since_clear += batch;
    // the previous batch has been waited for: the staging slots are free
    for (int64_t j = 0; j < batch; ++j) {
      StepArgs& a = sim->args_host[j];
      a.t = tt;
      a.t_half = tt + 0.5 * dt;
      a.dt = dt;
      a.cdt_half = SFP_C * (0.5 * dt);
      a.cdt = SFP_C * dt;
      a.ll_factor = sfp_ll_factor(dt);
      a.step = step0 + (uint64_t)(k + j);
      a.rel = (uint32_t)(k + j);
      a.rel_end = (uint32_t)nsteps;
      a.wrel = (uint32_t)(a.step - sim->wake_origin);
      a.fresh = (k + j == 0 && !keep_list) ? 1u : 0u;
      tt = tt + dt;
    }
    const bool tuning = autotune && sim_needs_tune(sim);
    const bool replay = !tuning && use_graph && batch == SYNC_EVERY;
    // a capture starts on an idle stream, before this batch's arguments are
    // queued (the graph reads them from the slots when it runs)
    if (replay && (sim->graph_stale || !sim->graph)) {
      HIPCHK(hipStreamSynchronize(sim->stream));
      CHK(sim_capture(sim));
    }
    HIPCHK(hipMemcpyAsync(sim->args_dev, sim->args_host, sizeof(StepArgs) * batch,
                          hipMemcpyHostToDevice, sim->stream));
    if (tuning) {
      for (int64_t j = 0; j < batch; ++j) CHK(launch_step(sim, (int)j, 0, 0, (int)j));
    } else if (replay) {
      HIPCHK(hipGraphLaunch(sim->graph, sim->stream));
      sim->graph_batches++;
    } else {
      for (int64_t j = 0; j < batch; ++j) CHK(launch_step(sim, (int)j, 0, 0, -1));
    }
    CHK(sim_read_counts(sim));
    // replay from a halted kernel until the batch completes; a halt that
    // comes back at the same kernel and step means the recovery did not make
    // the room the kernel asks for: an error, not another replay
    int32_t last_s = -1;
    int64_t last_h = -1;
    while (sim->halt_host->flag) {
      int s = sim->halt_host->species;
      int64_t h = sim->halt_host->step - (int64_t)(step0 + k);
      if (h < 0 || h >= batch) return set_err(SFP_E_STATE, "halt outside the launched steps");
      if (s == last_s && h == last_h)
        return set_err(SFP_E_CAPACITY, "species %d halted twice at step %lld: no room after growth",
                       s, (long long)sim->halt_host->step);
      last_s = s;
      last_h = h;
      bool deferred = false;
      CHK(sim_recover(sim, s, &deferred));
      if (deferred) {
        // the photon kernel ran through; its deferred conversions, then the
        // rest of the step
        hipLaunchKernelGGL(sim->defer[s], dim3(sim->cus * 4), dim3(TILE), 0, sim->stream,
                           (const Plan*)sim->plan_dev, s, (const StepArgs*)(sim->args_dev + h));
        HIPCHK(hipGetLastError());
        CHK(launch_step(sim, (int)h, s, 2, -1));
      } else {
        CHK(launch_step(sim, (int)h, s, 1, -1));
      }
      for (int64_t j = h + 1; j < batch; ++j) CHK(launch_step(sim, (int)j, 0, 0, -1));
      CHK(sim_read_counts(sim));
    }
    if (tuning) CHK(sim_tune_pick(sim));
    k += batch;
  }
  *t = tt;
  sim->rec_end = std::max<uint64_t>(sim->rec_end, step0 + (uint64_t)nsteps);
  // sleeping QED photons lag behind the run's last step until settled: the
  // stores call back before the host touches them
  if (nsteps > 0) {
    sim->lag_wend = (uint32_t)(step0 + (uint64_t)nsteps - sim->wake_origin);
    for (int s = 0; s < sim->nsp; ++s) {
      if (!sim->wake_mem[s] || !sim->qed[s] || !sim->plan.sp[s].push) continue;
      Store* st = sim->stores[s];
      if (st->settle && st->settle_sim != sim) CHK(store_settle(st));
      st->settle = sim_settle_cb;
      st->settle_sim = sim;
      sim->lag = true;
    }
  }
  sim->wake_end = step0 + (uint64_t)nsteps;
  for (int s = 0; s < sim->nsp; ++s) sim->wake_version[s] = sim->stores[s]->version;
  sim->wake_valid = true;
  if (nsteps > 0) {
    if (keep_list) sim->list_keeps++;
    for (int s = 0; s < sim->nsp; ++s) sim->list_runs[s] = ++sim->stores[s]->runs;
  }
  sim->list_kept = nsteps > 0 || keep_list;
  return SFP_OK;
}

static int sim_collect_timing(Sim* sim) { return sim_collect_events(sim); }

extern "C" int sfp_sim_set_recorder(void* p, int32_t species, int64_t n, int64_t every,
                                    int64_t max_rows, uint64_t step0) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp) return set_err(SFP_E_ARG, "bad species");
  if (sim->plan.sp[species].photon)
    return set_err(SFP_E_ARG, "the recorder takes a lepton species (photon state is cached "
                              "between the photon list rebuilds)");
  if (n > 0 && (every < 1 || max_rows < 1))
    return set_err(SFP_E_ARG, "every and max_rows must be >= 1");
  Recorder& R = sim->plan.rec;
  HIPCHK(hipDeviceSynchronize());
  if (R.buf) HIPCHK(hipFree(R.buf));
  if (R.t) HIPCHK(hipFree(R.t));
  R = Recorder{};
  if (n > 0) {
    HIPCHK(hipMalloc(&R.buf, sizeof(double) * 6 * (size_t)n * (size_t)max_rows));
    HIPCHK(hipMalloc(&R.t, sizeof(double) * (size_t)max_rows));
    R.n = n;
    R.every = every;
    R.max_rows = max_rows;
    R.step0 = step0;
    R.species = species;
    R.on = 1;
  }
  sim->rec_end = step0;
  sim->graph_stale = true;
  HIPCHK(hipMemcpy(sim->plan_dev, &sim->plan, sizeof(Plan), hipMemcpyHostToDevice));
  return SFP_OK;
}

extern "C" int sfp_sim_get_records(void* p, double* t_host, double* rows_host, int64_t* rows) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  const Recorder& R = sim->plan.rec;
  int64_t nr = 0;
  if (R.on && sim->rec_end > R.step0)
    nr = std::min<int64_t>(R.max_rows, (int64_t)((sim->rec_end - R.step0) / (uint64_t)R.every));
  *rows = nr;
  if (nr > 0) {
    HIPCHK(hipDeviceSynchronize());
    if (t_host) HIPCHK(hipMemcpy(t_host, R.t, sizeof(double) * nr, hipMemcpyDeviceToHost));
    if (rows_host)
      HIPCHK(hipMemcpy(rows_host, R.buf, sizeof(double) * 6 * (size_t)R.n * nr,
                       hipMemcpyDeviceToHost));
  }
  return SFP_OK;
}

extern "C" int sfp_sim_set_grid(void* p, int32_t species, int32_t phase, int32_t grid) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp || phase < 0 || phase >= NPH || grid < 0)
    return set_err(SFP_E_ARG, "bad species/phase/grid");
  sim->pinned[species][phase] = grid;
  sim->pinned_rows[species][phase] = 0;
  return sim_pick_kernels(sim);
}

extern "C" int sfp_sim_set_shape(void* p, int32_t species, int32_t phase, int32_t grid,
                                 int32_t rows) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp || phase < 0 || phase >= NPH || grid < 0)
    return set_err(SFP_E_ARG, "bad species/phase/grid");
  if (rows < 0 || rows > ROWS)
    return set_err(SFP_E_ARG, "tile rows %d outside 0..%d", (int)rows, ROWS);
  if (rows > 0 && grid == 0) return set_err(SFP_E_ARG, "tile rows pinned without a grid");
  sim->pinned[species][phase] = grid;
  sim->pinned_rows[species][phase] = rows;
  return sim_pick_kernels(sim);
}

extern "C" int sfp_sim_launch_shape(void* p, int32_t species, int32_t phase, int32_t* grid,
                                    int32_t* tile, int32_t* rows, int32_t* ncand) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp || phase < 0 || phase >= NPH)
    return set_err(SFP_E_ARG, "bad species/phase");
  const bool has = !sim->cand[species][phase].empty();
  const Shape sh = has ? sim_shape(sim, species, phase) : Shape{0, 0};
  *grid = sh.grid;
  *tile = has ? TILE : 0;
  if (rows) *rows = sh.rows;
  *ncand = (int32_t)sim->cand[species][phase].size();
  return SFP_OK;
}

extern "C" int sfp_sim_kernel_time(void* p, int32_t species, int32_t phase, double* ms,
                                   int64_t* launches) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp || phase < 0 || phase >= NPH)
    return set_err(SFP_E_ARG, "bad species/phase");
  CHK(sim_collect_timing(sim));
  *ms = sim->ms_acc[species][phase];
  *launches = sim->launches[species][phase];
  return SFP_OK;
}

extern "C" int sfp_sim_set_timing_every(void* p, int64_t every) {
  Sim* sim = (Sim*)p;
  if (every < 1) return set_err(SFP_E_ARG, "timing_every must be >= 1, got %lld", (long long)every);
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  sim->timing_every = every;
  return SFP_OK;
}

extern "C" int sfp_sim_get_counts(void* p, int32_t species, int64_t* n, uint64_t* next_id) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  if (species < 0 || species >= sim->nsp) return set_err(SFP_E_ARG, "no species %d", species);
  if (n) *n = sim->counts_host[species].n;
  if (next_id) *next_id = (uint64_t)sim->counts_host[species].next_id;
  return SFP_OK;
}

extern "C" int sfp_sim_set_timing_kernel(void* p, int32_t species, int32_t phase) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  if (species < 0) {
    sim->time_s = sim->time_ph = -1;
    return SFP_OK;
  }
  if (species >= sim->nsp || phase < 0 || phase >= NPH)
    return set_err(SFP_E_ARG, "no kernel (species %d, phase %d)", species, phase);
  sim->time_s = species;
  sim->time_ph = phase;
  return SFP_OK;
}

extern "C" int sfp_sim_reset_timing(void* p) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  CHK(sim_collect_timing(sim));
  for (int s = 0; s < sim->nsp; ++s)
    for (int ph = 0; ph < NPH; ++ph) {
      sim->ms_acc[s][ph] = 0.0;
      sim->launches[s][ph] = 0;
    }
  return SFP_OK;
}

extern "C" int sfp_sim_stats(void* p, int32_t species, int64_t* stats) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp) return set_err(SFP_E_ARG, "bad species");
  CHK(sim_read_counts(sim));
  const Counts& c = sim->counts_host[species];
  stats[0] = c.psteps;
  stats[1] = c.processed;
  stats[2] = c.spawned;
  stats[3] = c.n;
  stats[4] = c.hot_total;
  stats[5] = sim->syncs;
  stats[6] = sim->replays;
  stats[7] = sim->stores[species]->cap;
  stats[8] = c.slept;
  stats[9] = c.sleeps;
  stats[10] = sim->graph_batches;
  stats[11] = sim->captures;
  stats[12] = sim->tune_batches;
  stats[13] = c.emit_total;
  stats[14] = sim->clock_restarts;
  return SFP_OK;
}

extern "C" int sfp_sim_stats_n(void* p, int32_t species, int64_t* stats, int32_t n) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp) return set_err(SFP_E_ARG, "bad species");
  if (n < 0) return set_err(SFP_E_ARG, "n must be >= 0");
  int64_t v[SFP_SIM_NSTATS];
  CHK(sfp_sim_stats(p, species, v));
  v[15] = sim->counts_host[species].dense_steps;
  v[16] = sim->counts_host[species].deferrals;
  v[17] = sim->list_keeps;
  for (int32_t k = 0; k < n && k < SFP_SIM_NSTATS; ++k) stats[k] = v[k];
  return SFP_OK;
}

extern "C" int sfp_sim_qed_diag(void* p, int32_t species, int64_t* counts, double* p_max) {
  Sim* sim = (Sim*)p;
  if (!sim || species < 0 || species >= sim->nsp) return set_err(SFP_E_ARG, "bad species");
  CHK(sim_read_counts(sim));
  const Counts& c = sim->counts_host[species];
  counts[0] = c.p_over1;
  counts[1] = c.z_clamp;
  counts[2] = c.chi_clamp;
  counts[3] = c.chi_nan;
  uint32_t bits = (uint32_t)c.p_max_bits;
  float f;
  memcpy(&f, &bits, sizeof(f));
  *p_max = (double)f;
  return SFP_OK;
}

extern "C" int sfp_sim_reset_stats(void* p) {
  Sim* sim = (Sim*)p;
  if (!sim) return set_err(SFP_E_ARG, "null simulation");
  CHK(sim_read_counts(sim));
  for (int s = 0; s < sim->nsp; ++s) {
    Counts c = sim->counts_host[s];
    c.psteps = c.processed = c.spawned = c.hot_total = c.slept = c.sleeps = c.emit_total = 0;
    c.dense_steps = c.p_over1 = c.z_clamp = c.chi_clamp = c.chi_nan = c.p_max_bits = 0;
    c.deferrals = 0;
    HIPCHK(hipMemcpy(sim->stores[s]->cnt, &c, sizeof(c), hipMemcpyHostToDevice));
  }
  sim->syncs = 0;
  sim->replays = 0;
  sim->graph_batches = sim->captures = sim->tune_batches = 0;
  sim->clock_restarts = 0;
  sim->list_keeps = 0;
  return SFP_OK;
}
