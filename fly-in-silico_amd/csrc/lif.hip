// libfis: MI355X (gfx950) kernels for the whole-brain LIF step.
//
// Semantics: Engine._core + Engine._poisson of lab/flybrain/sim.py:848-953
// and the event SpMV of lab/flybrain/ev_spmv.cu:60-79.  Layout and launch
// shapes are this library's own:
//
//  * state is trial-minor, idx = neuron * bstride + trial, so one wave of 64
//    lanes is 64 trials of one neuron (bstride >= 64) or 64/bstride neurons
//    with all their trials; a spike ballot is then directly the trial mask
//    of an event, and the scatter reads a CSR row once for every trial in
//    that mask;
//  * the delay ring is int32 synapse counts, 40 slots (two delays): the
//    update of step t reads and clears slot (t+1) mod 40, which holds the
//    input computed from the spikes of step t-20 (reference: rec written at
//    step u, read at u+19); the spikes of step t are added into slot
//    (t+21) mod 40.  Since an input needs 20 steps to arrive, W <= 20 steps
//    can run back to back with v, g and the refractory counter in registers,
//    and the slots a window reads are never the slots its spikes go to;
//  * integer atomics make the sum exact and order-free; wScale is applied
//    once per read, after the sum (sim.py:906-918);
//  * every float32 operation of the reference is kept as one rounding in
//    the same order (build with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>
#include <mutex>

#include "../../include/fis.h"

// Event staging of fis_window_noalloc, for callers that cannot allocate:
// one array per device, held until the process ends, shared by every call
// on that device (calls that overlap on one GPU race).  Every word a call
// reads from it is written earlier in the same call, so nothing in it has
// to start at zero.
static __device__ __align__(256) unsigned char scratch[FIS_SCRATCH_BYTES];

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kDelay = 20;  // steps from a spike to the input it makes

// ------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // one 32 x 32 -> 64-bit multiply gives both halves
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

__device__ __forceinline__ uint32_t pick(const U4& x, uint32_t j) {
  return j == 0 ? x.x : (j == 1 ? x.y : (j == 2 ? x.z : x.w));
}

// Bernoulli(p) from one 32-bit word: u = (x >> 8) * 2^-24 (exact), u < p.
__device__ __forceinline__ uint32_t bern(uint32_t x, float p) {
  const float u = (float)(x >> 8) * (1.0f / 16777216.0f);
  return u < p ? 1u : 0u;
}

// Kick bits of one neuron-trial for steps t0 .. t0+W-1 (W <= WMAX):
// kick(word) gives step k's bit from its 32-bit Philox word, called for
// k = 0 .. W-1 in order.  Counter (t>>2, neuron, trial_id, 0), key = 64-bit
// seed, word t & 3.  The window's Philox blocks are independent of each
// other and computed side by side (one block's latency, not six).
template <int WMAX, class K>
__device__ __forceinline__ uint32_t window_kicks(int64_t t0, int W,
                                                 uint32_t neuron, uint32_t tid,
                                                 uint64_t seed, K kick) {
  constexpr int NC = (WMAX + 3) / 4 + 1;  // blocks a window can touch
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const int64_t c0 = t0 >> 2;
  const int off = (int)(t0 & 3);        // word of step t0 in block c0
  const int nc = (off + W + 3) >> 2;    // blocks this window touches
  // The first NC - 1 blocks unconditionally (a full window needs them all;
  // a shorter one ignores the surplus words): one basic block, in which
  // the compiler interleaves their independent rounds.  Each behind a
  // test of its own, they ran one dependent chain after another.
  U4 x[NC];
#pragma unroll
  for (int c = 0; c < NC - 1; ++c)
    x[c] = philox(U4{(uint32_t)(c0 + c), neuron, tid, 0u}, k0, k1);
  x[NC - 1] = U4{0u, 0u, 0u, 0u};
  if (nc == NC)  // a window that starts inside a block touches one more
    x[NC - 1] = philox(U4{(uint32_t)(c0 + NC - 1), neuron, tid, 0u}, k0, k1);
  uint32_t bits = 0;
#pragma unroll
  for (int i = 0; i < 4 * NC; ++i) {  // word i of the blocks = step i - off
    const int k = i - off;
    if (k >= 0 && k < W) bits |= kick(pick(x[i >> 2], (uint32_t)(i & 3))) << k;
  }
  return bits;
}

// ... at a probability that may change per step (p_next() gives the
// probability of each step in turn)
template <int WMAX, class P>
__device__ __forceinline__ uint32_t window_bits(int64_t t0, int W,
                                                uint32_t neuron, uint32_t tid,
                                                uint64_t seed, P p_next) {
  return window_kicks<WMAX>(t0, W, neuron, tid, seed, [&](uint32_t x) {
    const float p = p_next();
    return p > 0.0f ? bern(x, p) : 0u;
  });
}

// ... at one probability p for the whole window.  bern's test u < p, with
// u = m * 2^-24 for the integer m = x >> 8 < 2^24, is m < p * 2^24 (a power
// of two scaling: exact), i.e. m < ceil(p * 2^24), capped at 2^24 where
// every m passes: one integer compare per step, the same bits.
template <int WMAX>
__device__ __forceinline__ uint32_t window_bits_fixed(int64_t t0, int W,
                                                      uint32_t neuron,
                                                      uint32_t tid,
                                                      uint64_t seed, float p) {
  if (!(p > 0.0f)) return 0u;  // bern is never called (nor true) then
  const float q = p * 16777216.0f;
  const uint32_t thr = q >= 16777216.0f ? 16777216u : (uint32_t)ceilf(q);
  return window_kicks<WMAX>(t0, W, neuron, tid, seed, [thr](uint32_t x) {
    return (x >> 8) < thr ? 1u : 0u;
  });
}

// The drive table's probability of one group, step by step from trial step
// t: rate[min(t / seg_steps, n_seg - 1)][group] * prob_scale.  One division
// per window; the segment then advances by counting.
struct TableRate {
  const float* rate;
  int n_groups, n_seg, seg_steps, group;
  float ps;
  int64_t seg;
  int rem;
  float p;
  __device__ TableRate(const float* rate_, int n_groups_, int n_seg_,
                       int seg_steps_, int group_, float ps_, int64_t t)
      : rate(rate_), n_groups(n_groups_), n_seg(n_seg_),
        seg_steps(seg_steps_), group(group_), ps(ps_) {
    seg = t / seg_steps;
    rem = (int)(t - seg * seg_steps);
    p = at(seg);
  }
  // whether the next W steps all take the current probability
  __device__ bool holds(int W) const {
    return rem + W <= seg_steps || seg >= n_seg - 1;
  }
  __device__ float at(int64_t sg) const {
    sg = sg < 0 ? 0 : (sg < n_seg ? sg : n_seg - 1);
    return rate[sg * n_groups + group] * ps;
  }
  __device__ float operator()() {
    const float now = p;
    if (++rem == seg_steps) {
      rem = 0;
      p = at(++seg);
    }
    return now;
  }
};

// The current step: the clock, plus the steps a deferred run of windows
// has launched since the clock was last advanced.  Read as a scalar load
// (through the constant address space): no kernel that reads the clock
// writes it (k_tick and the scatter launches advance it and read it not),
// and a launch sees what earlier launches left.  As a vector load its
// round trip stood at the start of every tile.
__device__ __forceinline__ int64_t now(const fis_sim& s) {
  return *(const __attribute__((address_space(4))) int64_t*)(uintptr_t)
             s.clock +
         s.clock_ahead;
}

__device__ __forceinline__ int64_t origin_of(const fis_sim& s, int b) {
  return s.t_origin ? s.t_origin[b] : 0;
}

// first global step at which trial b's spikes are no longer counted
__device__ __forceinline__ int64_t end_of(const fis_sim& s, int b) {
  return s.t_end ? s.t_end[b] : INT64_MAX;
}

__device__ __forceinline__ int n_waves_of(const fis_sim& s) {
  return (s.n * s.bstride + kWave - 1) >> 6;
}

// Pull road: a kick, a reset or a flag change of neuron-trial idx ends the
// calm of its (neuron, tile part), which would otherwise let the next
// window skip it unread.  Every writer stores the same zero.
__device__ __forceinline__ void unsettle(const fis_sim& s, int64_t idx) {
  if (!s.calm) return;
  const int tile = s.pull_tile ? s.pull_tile : FIS_PULL_TILE;
  const int tpn = s.bstride < tile ? s.bstride : tile;
  const int64_t i = idx / s.bstride;
  const int b = (int)(idx - i * s.bstride);
  uint8_t* c = s.calm + i * (s.bstride / tpn) + b / tpn;
  if (*c) *c = 0;
}

// ------------------------------------------------------------- drive

__global__ __launch_bounds__(kBlock) void k_drive_rates(fis_sim s,
                                                        const float* rates,
                                                        int W) {
  const int nb = s.n * s.bstride;
  const int idx = (int)blockIdx.x * kBlock + (int)threadIdx.x;
  if (idx >= nb) return;
  const int i = idx / s.bstride;
  const int b = idx - i * s.bstride;
  uint32_t bits = 0;
  if (b < s.batch) {
    const float p = rates[(int64_t)b * s.n + i] * s.prob_scale;
    bits = window_bits_fixed<FIS_MAX_WINDOW>(now(s) - origin_of(s, b), W,
                                             (uint32_t)i, s.trial_ids[b],
                                             s.seeds[b], p);
  }
  s.sbits[idx] = bits;
  if (bits) unsettle(s, idx);
}

// Eight waves per SIMD: the compiler's choice (66 VGPRs) left room for
// seven; at 64 nothing spills (55.5 instead of 56.7 us per window at
// B = 256).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_drive_table(
    fis_sim s, const fis_drive_entry* entries, int n_entries,
    const float* rate_hz, int n_groups, int n_seg, int seg_steps, int W) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= n_entries) return;
  const fis_drive_entry e = entries[j];
  // an entry past the state names no neuron-trial: nothing to drive (the
  // host checks the table; this keeps a bad entry from writing anywhere)
  if (e.idx >= (uint32_t)(s.n * s.bstride)) return;
  const int i = (int)e.idx / s.bstride;
  const int b = (int)e.idx - i * s.bstride;
  const int64_t t = now(s) - origin_of(s, b);
  const TableRate rate(rate_hz, n_groups, n_seg, seg_steps, e.group,
                       s.prob_scale, t);
  // a window inside one rate segment (the usual case): one threshold
  const uint32_t bits =
      rate.holds(W)
          ? window_bits_fixed<FIS_MAX_WINDOW>(t, W, (uint32_t)i,
                                              s.trial_ids[b], s.seeds[b],
                                              rate.p)
          : window_bits<FIS_MAX_WINDOW>(t, W, (uint32_t)i, s.trial_ids[b],
                                        s.seeds[b], rate);
  // entries of neighbouring neuron-trials may sit in other workgroups: set
  // the bits with an atomic so no two workgroups store into one line
  if (bits) {
    atomicOr(&s.sbits[e.idx], bits);
    unsettle(s, e.idx);
  }
}

__global__ __launch_bounds__(kBlock) void k_drive_tape(fis_sim s,
                                                       const int64_t* tape_ptr,
                                                       const int32_t* tape_idx,
                                                       int64_t n_tape_steps) {
  // blockIdx.y = window step k
  const int64_t t = now(s) + blockIdx.y;
  if (t >= n_tape_steps) return;
  const int64_t beg = tape_ptr[t], end = tape_ptr[t + 1];
  for (int64_t e = beg + (int64_t)blockIdx.x * kBlock + threadIdx.x; e < end;
       e += (int64_t)gridDim.x * kBlock) {
    atomicOr(&s.sbits[tape_idx[e]], 1u << blockIdx.y);
    unsettle(s, tape_idx[e]);
  }
}

// ------------------------------------------------------------- update

// Lane masks.  A ballot lands in a scalar register pair; it is used here
// only by scalar control flow (is any lane set?).  Every mask that is
// stored to memory is built in vector registers instead, by OR-reduction
// across lanes, and positions come from a vector prefix scan.  Measured on
// gfx950: a ballot copied into a vector register for a store
// (v_mov_b64 v, s[pair]) intermittently arrived with lanes 32..63 cleared,
// while scalar tests of the same ballot were always right.
__device__ __forceinline__ bool any_lane(bool p) {
  uint64_t m = __ballot(p);
  asm volatile("s_mov_b64 %0, %1" : "=s"(m) : "s"(m));
  return m != 0;
}

// x, held in a vector register the compiler cannot see through: a load
// whose address depends on it is a vector load, waited for only where its
// value is used (a uniform address would be a scalar load, which the next
// wait for LDS or another scalar load also waits for)
__device__ __forceinline__ int opaque_vgpr(int x) {
  int y;
  asm("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
  return y;
}

// OR of `x` over aligned groups of `width` lanes (width a power of two <= 64);
// every lane of a group receives the group's value.
__device__ __forceinline__ uint64_t group_or(uint64_t x, int width) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  for (int o = 1; o < width; o <<= 1) {
    lo |= (uint32_t)__shfl_xor((int)lo, o);
    hi |= (uint32_t)__shfl_xor((int)hi, o);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Exclusive prefix count of `p` over the wave, and the wave total.
__device__ __forceinline__ int prefix_count(bool p, int lane, int& total) {
  const int x = p ? 1 : 0;
  int inc = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(inc, o);
    inc += (lane >= o) ? y : 0;
  }
  total = __shfl(inc, kWave - 1);
  return inc - x;
}

// Spike compaction of one step: the raster row (optional) and the events
// of transmitted spikes, appended to this wave's segment.  `pre` and `tb`
// are the lane's neuron and trial (computed once per thread); `live` is
// whether the wave holds neuron-trials at all (uniform over the wave when
// bstride >= 64, the only case that reads it).  Stores are
// made by every lane (identical values, or a sink slot for lanes with
// nothing to write), so no branch narrows the wave around them.
__device__ __forceinline__ void emit(const fis_sim& s, bool live, bool fired,
                                     bool sends, int lane, int pre, int tb,
                                     int wave, int nblk, int rrow, int slot,
                                     int k, fis_event* evs, int& nev) {
  const int bs = s.bstride;
  if (s.raster) {
    const uint64_t all = group_or(fired ? (1ull << lane) : 0ull, kWave);
    s.raster[((int64_t)rrow * nblk + (wave >> 2)) * FIS_LINE_WORDS +
             (wave & 3)] = all;
  }
  const bool any = any_lane(sends);
  if (bs >= kWave) {
    // one wave = one neuron x 64 trials: at most one event per step, kept
    // at position k of the wave's segment (empty mask when no spike), so
    // the scatter can walk the window step by step
    fis_event e;
    e.pre = pre;
    e.slot = (uint16_t)slot;
    e.tb = (uint16_t)(tb - lane);  // trial of lane 0
    e.mask = any ? group_or(sends ? (1ull << lane) : 0ull, kWave) : 0ull;
    // every lane writes the same record; a wave past the last neuron of the
    // last workgroup has no segment (the buffer holds n_waves segments)
    if (live) evs[k] = e;
    return;
  }
  if (!any) return;
  {    // 64/bs neurons per wave; the first lane of each neuron leads
    const uint64_t sub = group_or(sends ? (1ull << tb) : 0ull, bs);
    const bool lead = tb == 0 && sub != 0;
    int total;
    const int pos = prefix_count(lead, lane, total);
    // lanes that do not lead write the segment's last slot (a sink)
    const int at = lead ? nev + pos : s.ev_cap - 1;
    fis_event e;
    e.pre = pre;
    e.slot = (uint16_t)slot;
    e.tb = 0;
    e.mask = sub;
    evs[at] = e;
    nev += total;
  }
}

// Registers of one neuron-trial over a window.  The reference's mask
// factors (exactly 0.0 or 1.0, sim.py:393-410) become lane flags here; see
// lif_step for why that leaves every value unchanged.
struct Cell {
  float v, g;
  int r;     // refractory counter (saturating at FIS_ST_RMAX)
  int R;     // gate threshold; INT_MAX when the input mask is off
  bool sp;   // spiked at the previous step
  bool fire_on, held;
  float v_hold;
};

__device__ __forceinline__ void load_cell(const fis_sim& s, int idx, float& v,
                                          float& g, uint32_t& st,
                                          uint32_t& fl, uint32_t& sb,
                                          float& vh) {
  v = s.v[idx];
  g = s.g[idx];
  st = s.st[idx];
  fl = s.fl[idx];
  sb = s.sbits[idx];
  if ((fl & FIS_HOLD) && s.vhold) vh = s.vhold[idx];
  if (sb) s.sbits[idx] = 0u;  // the window's kicks are consumed
}

// The window's kicks of a table-driven neuron-trial (fused drive): the
// bits k_drive_table would set for it.
// What the fused drive of one neuron-trial reads: its group (< 0: not
// driven) and its trial's Philox identity and origin.  Read at the start
// of the update, all at once, so that they arrive with the state.
struct DriveId {
  int group;
  uint32_t trial_id;
  uint64_t seed;
  int64_t origin;
};

__device__ __forceinline__ DriveId drive_id(const fis_sim& s, int idx,
                                            int trial) {
  DriveId d{-1, 0u, 0ull, 0};
  if (s.drive_group || s.drive_dense) {
    d.group = s.drive_group ? (int)s.drive_group[idx] : 0;
    d.trial_id = s.trial_ids[trial];
    d.seed = s.seeds[trial];
    d.origin = origin_of(s, trial);
  }
  return d;
}

// `fixed`: the table has one segment, and `p` is this lane's probability
// (its group's rate times prob_scale, what TableRate gives at every step).
template <int WMAX>
__device__ __forceinline__ uint32_t fused_kicks(const fis_sim& s,
                                                const DriveId& d, bool fixed,
                                                float p, int neuron,
                                                int64_t t0, int W) {
  if (d.group < 0) return 0u;
  const int64_t t = t0 - d.origin;
  if (fixed)
    return window_bits_fixed<WMAX>(t, W, (uint32_t)neuron, d.trial_id,
                                   d.seed, p);
  return window_bits<WMAX>(t, W, (uint32_t)neuron, d.trial_id, d.seed,
                           TableRate(s.drive_rate, s.drive_n_groups,
                                     s.drive_n_seg, s.drive_seg_steps,
                                     d.group, s.prob_scale, t));
}

__device__ __forceinline__ Cell make_cell(const fis_sim& s, uint32_t fl,
                                          float vh, uint32_t st, float v,
                                          float g) {
  Cell c;
  c.v = v;
  c.g = g;
  c.fire_on = !(fl & FIS_FIRE_OFF);
  c.held = (fl & FIS_HOLD) != 0;
  c.v_hold = vh;
  c.R = (fl & FIS_IN_OFF) ? INT_MAX
                          : ((fl & FIS_NOREFRAC) ? 0 : s.refrac_steps);
  c.r = (int)(st & FIS_ST_RMAX);
  c.sp = (st & FIS_ST_SPIKE) != 0;
  return c;
}

// The two values the reference's stimulus term can take,
// w_scale * (bern * gain) with bern 1.0 or 0.0 (sim.py:859-861), each
// computed with the reference's two roundings.
struct Stim {
  float on, off;
};

__device__ __forceinline__ Stim make_stim(const fis_sim& s) {
  return Stim{s.w_scale * (1.0f * s.gain), s.w_scale * (0.0f * s.gain)};
}

// One step of Engine._core for one neuron-trial (sim.py:920-953), given the
// exact integer synaptic input I read from the delay line and the Poisson
// kick of the step.  Every operation whose operands are not a 0/1 mask is
// one float32 rounding in the reference's order.  A product with a mask
// factor that is exactly 1.0 or 0.0, and the sum of a finite x with a zero,
// are x, x or a zero: the reference's
//   delayed * gate, * m_in;  v * m_hold + v_hold;  spikes * m_fire;
//   v - (v - v_reset) * spikes;  g_new - g_new * spikes
// are therefore selects here, equal in value (a zero may differ in sign,
// which no later +, -, * or comparison can turn into a different number).
// Returns whether the neuron-trial fired (the masked spike).
__device__ __forceinline__ bool lif_step(const fis_sim& s, Cell& c, int32_t I,
                                         bool kick, const Stim& stim) {
  // refractory counter: r = (r + 1) * (1 - s), saturating; gate = r >= R
  // (it saturates at FIS_ST_RMAX only when stored: within a window it
  // stays below FIS_ST_RMAX + FIS_MAX_WINDOW, and the gate compares it with
  // refrac_steps <= 126 or INT_MAX)
  c.r = c.sp ? 0 : c.r + 1;

  // conductance with the delayed input (sim.py:927-929):
  // g * a_syn + (w_scale * I) * m_in * gate
  const float rec = s.w_scale * (float)I;
  const float gd = c.g * s.a_syn;
  const float g_new = (c.r >= c.R) ? gd + rec : gd;

  // membrane with the OLD conductance (sim.py:937-939)
  float vv = c.v + (kick ? stim.on : stim.off);
  const float t1 = vv - s.v_rest;
  const float t2 = c.g - t1;
  const float t3 = s.a_mem * t2;
  vv = vv + t3;
  vv = c.held ? c.v_hold : vv;  // hold before threshold (sim.py:944)
  const bool fire = c.fire_on && vv > s.v_th;
  // reset by the masked spike
  const float t4 = vv - s.v_reset;
  c.v = fire ? vv - t4 : vv;
  c.g = fire ? 0.0f : g_new;
  c.sp = fire;
  return fire;
}

__device__ __forceinline__ void store_cell(const fis_sim& s, const Cell& c,
                                           int idx, int pre, uint32_t nspk) {
  s.v[idx] = c.v;
  s.g[idx] = c.g;
  s.st[idx] = (uint8_t)((c.sp ? FIS_ST_SPIKE : 0u) |
                        (uint32_t)min(c.r, (int)FIS_ST_RMAX));
  if (s.spikes_out) s.spikes_out[idx] = c.sp ? 1.0f : 0.0f;
  if (nspk) {
    // an atomic add returns nothing to wait for: a read-modify-write here
    // would put one more memory round trip at the end of every window
    if (s.counts) atomicAdd(&s.counts[idx], nspk);
    if (s.pop_counts) atomicAdd(&s.pop_counts[pre], nspk);
  }
}

// W steps of lif_step with the inputs in registers; returns the steps at
// which the neuron-trial fired (bit k = step k).  A full window runs as one
// unrolled chain with no per-step test of W.
template <int WMAX>
__device__ __forceinline__ uint32_t run_window(const fis_sim& s, Cell& c,
                                               const int32_t (&in)[WMAX],
                                               uint32_t kicks, Stim stim,
                                               int W) {
  // the two stimulus values stay materialised in registers (the compiler
  // would otherwise redo w_scale * (...) at every step)
  asm volatile("" : "+v"(stim.on), "+v"(stim.off));
  uint32_t fired = 0;
  if (W == WMAX) {
#pragma unroll
    for (int k = 0; k < WMAX; ++k)
      fired |= lif_step(s, c, in[k], (kicks >> k) & 1u, stim) ? (1u << k)
                                                                : 0u;
  } else {
#pragma unroll
    for (int k = 0; k < WMAX; ++k)
      if (k < W)
        fired |= lif_step(s, c, in[k], (kicks >> k) & 1u, stim) ? (1u << k)
                                                                  : 0u;
  }
  return fired;
}

// A full window of lif_step for a wave whose neuron-trials carry no flag and
// no Poisson kick in the window: the fire mask is on, nothing is held, the
// gate threshold is refrac_steps, and the stimulus term is the zero
// w_scale * (0.0 * gain), whose sum with the finite v is v (value-equal; the
// add is left out).  The refractory counter is kept as an offset: after the
// update of step k it is k + b, so the gate r >= R is b >= R - k (a scalar
// per step) and a spike at step k sets b = -(k + 1) for step k + 1.  Equal
// in every value to run_window with those inputs; returns the steps at
// which the neuron-trial fired (bit k = step k).
template <int WMAX>
__device__ __forceinline__ uint32_t run_window_plain(const fis_sim& s, Cell& c,
                                                     const int32_t (&in)[WMAX]) {
  int b = c.sp ? 0 : c.r + 1;  // counter after the update of step 0
  float v = c.v, g = c.g;
  bool fire = false;
  uint32_t fired = 0;
#pragma unroll
  for (int k = 0; k < WMAX; ++k) {
    const float rec = s.w_scale * (float)in[k];
    const float gd = g * s.a_syn;
    const float g_new = (b >= s.refrac_steps - k) ? gd + rec : gd;
    const float t1 = v - s.v_rest;
    const float t2 = g - t1;
    const float t3 = s.a_mem * t2;
    const float vv = v + t3;
    fire = vv > s.v_th;
    const float t4 = vv - s.v_reset;
    v = fire ? vv - t4 : vv;
    g = fire ? 0.0f : g_new;
    fired |= fire ? (1u << k) : 0u;
    if (k + 1 < WMAX) b = fire ? -(k + 1) : b;
  }
  c.v = v;
  c.g = g;
  c.sp = fire;
  c.r = (WMAX - 1) + b;  // the counter after the last update
  return fired;
}

// A plain wave (as run_window_plain) whose every input gate is open at the
// first step -- no spike at the previous step, the refractory counter at
// least refrac_steps - 1 -- keeps it open for the whole window unless the
// lane fires.  Then, without a spike, every step is g_new = g * a_syn +
// w_scale * I and the membrane step with no reset: run here without the
// gate, the reset and the counter (9 operations per step instead of 17).
// Returns true, with the cell advanced, when no lane of the wave fired;
// false, with the cell untouched, when one did (the caller then runs
// run_window_plain from the same state).  Equal in every value to
// run_window_plain where it returns true.  (The same speculation for waves
// with a gate still closed, the gate kept, measured slower: those waves
// belong to neurons spiking now, and most fire again in the window.)
template <int WMAX>
__device__ __forceinline__ bool run_window_open(const fis_sim& s, Cell& c,
                                                const int32_t (&in)[WMAX]) {
  float v = c.v, g = c.g;
  bool fire = false;
#pragma unroll
  for (int k = 0; k < WMAX; ++k) {
    const float rec = s.w_scale * (float)in[k];
    const float gd = g * s.a_syn;
    const float t1 = v - s.v_rest;
    const float t2 = g - t1;
    const float t3 = s.a_mem * t2;
    v = v + t3;
    g = gd + rec;
    fire = fire || v > s.v_th;
  }
  if (any_lane(fire)) return false;
  c.v = v;
  c.g = g;
  c.r += WMAX;  // the counter's WMAX increments; sp stays false
  return true;
}

// Whether a neuron-trial of a plain wave (no flag, no kick, a full window)
// sits at a fixed point of the step for the whole window: no synaptic
// input at any of its WMAX steps, no spike at the previous step, g = +0.0
// and v unchanged by the leak step v + a_mem * (g - (v - v_rest)) without
// reaching threshold.  Then every step of the window maps (v, g) to
// itself: the gated input w_scale * 0 = +0.0 leaves g_new = g * a_syn + 0
// at +0.0 whether the gate is open or not, and the membrane term is the
// one just evaluated; only the refractory counter moves, by one per step.
// Exactly the values run_window_plain would compute, without the chain.
template <int WMAX>
__device__ __forceinline__ bool quiet_cell(const fis_sim& s, const Cell& c,
                                           const int32_t (&in)[WMAX]) {
  int32_t any = 0;
#pragma unroll
  for (int k = 0; k < WMAX; ++k) any |= in[k];
  const float t1 = c.v - s.v_rest;
  const float t2 = c.g - t1;
  const float t3 = s.a_mem * t2;
  const float vv = c.v + t3;
  return any == 0 && !c.sp && __float_as_uint(c.g) == 0u &&
         __float_as_uint(vv) == __float_as_uint(c.v) && !(vv > s.v_th);
}

// The store of a neuron-trial that stayed at its fixed point for a window:
// v and g are unchanged, only the refractory counter moved (no spike, so
// no count, and the last step's spike output is 0).
__device__ __forceinline__ void store_quiet(const fis_sim& s, const Cell& c,
                                            int idx, uint32_t st_old) {
  const uint32_t st = (uint32_t)min(c.r, (int)FIS_ST_RMAX);
  if (st != st_old) s.st[idx] = (uint8_t)st;
  if (s.spikes_out) s.spikes_out[idx] = 0.0f;
}

// A full window of lif_step for a wave with no fire mask off and no hold
// (the push road's common case, also with driven neurons): as
// run_window_plain, with the Poisson kick term kept and the gate threshold
// per lane (refrac_steps, 0 without refractoriness, INT_MAX with the input
// masked: R - k does not overflow, and b never reaches it).  Equal in
// every value to run_window; returns the steps fired (bit k = step k).
template <int WMAX>
__device__ __forceinline__ uint32_t run_window_kick(const fis_sim& s, Cell& c,
                                                    const int32_t (&in)[WMAX],
                                                    uint32_t kicks, Stim stim) {
  asm volatile("" : "+v"(stim.on), "+v"(stim.off));
  int b = c.sp ? 0 : c.r + 1;  // counter after the update of step 0
  float v = c.v, g = c.g;
  bool fire = false;
  uint32_t fired = 0;
#pragma unroll
  for (int k = 0; k < WMAX; ++k) {
    const float rec = s.w_scale * (float)in[k];
    const float gd = g * s.a_syn;
    const float g_new = (b >= c.R - k) ? gd + rec : gd;
    float vv = v + (((kicks >> k) & 1u) ? stim.on : stim.off);
    const float t1 = vv - s.v_rest;
    const float t2 = g - t1;
    const float t3 = s.a_mem * t2;
    vv = vv + t3;
    fire = vv > s.v_th;
    const float t4 = vv - s.v_reset;
    v = fire ? vv - t4 : vv;
    g = fire ? 0.0f : g_new;
    fired |= fire ? (1u << k) : 0u;
    if (k + 1 < WMAX) b = fire ? -(k + 1) : b;
  }
  c.v = v;
  c.g = g;
  c.sp = fire;
  c.r = (WMAX - 1) + b;  // the counter after the last update
  return fired;
}

// One event's CSR row, read once and added into every trial of its mask.
__device__ __forceinline__ void scatter_event(const fis_sim& s,
                                              const fis_event& ev, int lane) {
  const int64_t nb = (int64_t)s.n * s.bstride;  // ring slot stride
  const int bs = s.bstride;
  const int64_t beg = s.offsets[ev.pre], end = s.offsets[ev.pre + 1];
  int32_t* base = s.ring + (int64_t)ev.slot * nb + ev.tb;
  for (int64_t k = beg + lane; k < end; k += kWave) {
    const int32_t post = s.cols[k];
    const int32_t val = s.vals[k];
    int32_t* p = base + (int64_t)post * bs;
    uint64_t m = ev.mask;
    while (m) {
      const int j = __ffsll((long long)m) - 1;
      m &= m - 1ull;
      __hip_atomic_fetch_add(p + j, val, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// The CSR rows of the neurons whose lead lane is set, each added into
// the trials of its `sub` mask in ring slot `slot`, one row after another
// with the whole wave on each (lead, pre and sub are per lane; the loop is
// uniform).
__device__ __forceinline__ void scatter_leads(const fis_sim& s, bool lead,
                                              int pre, uint64_t sub, int slot,
                                              int lane) {
  uint64_t leads = __ballot(lead);
  asm volatile("s_mov_b64 %0, %1" : "=s"(leads) : "s"(leads));
  while (leads) {
    const int l = __ffsll((long long)leads) - 1;
    leads &= leads - 1ull;
    fis_event ev;
    ev.pre = __builtin_amdgcn_readlane(pre, l);
    ev.slot = (uint16_t)slot;
    ev.tb = 0;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)sub, l);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(sub >> 32), l);
    ev.mask = ((uint64_t)hi << 32) | lo;
    scatter_event(s, ev, lane);
  }
}

// One trial per lane (bstride 1): the CSR row of every lane with a set
// bit in `sent` (bit k: the lane's neuron sent at step t0 + k), read once
// by the whole wave, added into ring slot (t0 + k + 21) mod 40 for each
// such step (integer sums: the order of the adds does not matter).
__device__ __forceinline__ void scatter_row_steps(const fis_sim& s,
                                                  uint32_t sent, int pre,
                                                  int64_t t0, int lane) {
  const int64_t nb = (int64_t)s.n;  // ring slot stride at bstride 1
  const int slot0 = (int)((t0 + kDelay + 1) % FIS_RING_LEN);
  uint64_t leads = __ballot(sent != 0u);
  asm volatile("s_mov_b64 %0, %1" : "=s"(leads) : "s"(leads));
  while (leads) {
    const int l = __ffsll((long long)leads) - 1;
    leads &= leads - 1ull;
    const int pl = __builtin_amdgcn_readlane(pre, l);
    const uint32_t steps = (uint32_t)__builtin_amdgcn_readlane((int)sent, l);
    const int64_t beg = s.offsets[pl], end = s.offsets[pl + 1];
    for (int64_t e = beg + lane; e < end; e += kWave) {
      const int32_t post = s.cols[e];
      const int32_t val = s.vals[e];
      for (uint32_t x = steps; x; x &= x - 1u) {  // uniform over the wave
        int sl = slot0 + __ffs(x) - 1;
        sl = sl >= FIS_RING_LEN ? sl - FIS_RING_LEN : sl;
        __hip_atomic_fetch_add(s.ring + (int64_t)sl * nb + post, val,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// A kernel-argument address (as an integer) as a pointer into the
// constant kernel-argument segment, so that loads through it are scalar.
__device__ __forceinline__ const __attribute__((address_space(4))) void*
kernarg_ptr(uint64_t a) {
  return (const __attribute__((address_space(4))) void*)a;
}

// The kernel's fis_sim argument (the first in the kernel-argument segment)
// read again, through an address the compiler cannot see through: its
// fields are then scalar loads at their use, not values kept from the
// kernel's start.  Kept across a long kernel instead, they overflow the
// scalar registers into vector-register lanes, and every spill and reload
// is one more vector instruction.
__device__ __forceinline__ const fis_sim& args_again() {
  uint64_t ka = (uint64_t)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(ka));
  return *(const fis_sim*)kernarg_ptr(ka);
}

// W steps per neuron-trial.  WMAX is the unrolled bound (1 or 20).
// Index arithmetic is 32-bit: n * bstride < 2^31 is checked on the host.
template <int WMAX, bool FUSED>
__global__ __launch_bounds__(kBlock) void k_update(fis_sim s, int W) {
  const int bs = s.bstride;
  const int nb = s.n * bs;
  const int idx = (int)blockIdx.x * kBlock + (int)threadIdx.x;
  const int lane = (int)threadIdx.x & (kWave - 1);
  const int wave = idx >> 6;
  const int nblk = (int)gridDim.x;
  const bool act = idx < nb;
  const int64_t t0 = now(s);
  const int pre = idx / bs;       // neuron of this lane
  const int tb = idx - pre * bs;  // trial of this lane
  const DriveId drive = act ? drive_id(s, idx, tb) : DriveId{-1, 0u, 0ull, 0};
  // A drive table of one segment and at most 64 rates is read by every
  // wave, one rate per lane, alongside the lanes' groups: a driven lane's
  // rate is then a lane shuffle rather than a second dependent load, and
  // no segment is computed (uniform over the grid).
  const bool lane_table = s.drive_group && s.drive_n_seg == 1 &&
                          s.drive_n_groups <= kWave;
  float lane_rate = 0.0f;
  if (lane_table && lane < s.drive_n_groups) lane_rate = s.drive_rate[lane];

  // the window's ring slots, all read (and the non-zero ones cleared)
  // before the recurrence starts: nothing in it waits for memory, and the
  // loads are in flight while the drive's Philox blocks are computed
  int32_t* rp = s.ring + idx;
  const int slot0 = (int)((t0 + 1) % FIS_RING_LEN);
  int32_t in[WMAX];
  {
    int sl = slot0;
#pragma unroll
    for (int k = 0; k < WMAX; ++k) {
      in[k] = (k < W && act) ? rp[(int64_t)sl * nb] : 0;
      sl = (sl + 1 == FIS_RING_LEN) ? 0 : sl + 1;
    }
  }
  // the step masks k < W are formed again where used (scalar compares),
  // not kept from the loads above in scalar registers that would spill
  asm volatile("" : "+s"(W));
  float v = 0.0f, g = 0.0f, vh = 0.0f;
  uint32_t st = 0, fl = FIS_FIRE_OFF, sb = 0;
  if (act) load_cell(s, idx, v, g, st, fl, sb, vh);
  float p_fixed = 0.0f;
  if (lane_table)  // every lane takes part in the shuffle
    p_fixed = __shfl(lane_rate, drive.group < 0 ? 0 : drive.group) *
              s.prob_scale;
  // dense rates: the lane's own probability, as k_drive_rates takes it
  const bool dense = s.drive_dense != nullptr;
  if (dense)
    p_fixed = (act && tb < s.batch)
                  ? s.drive_dense[(int64_t)tb * s.n + pre] * s.prob_scale
                  : 0.0f;
  sb |= fused_kicks<WMAX>(s, drive, lane_table || dense, p_fixed, pre, t0,
                          W);

  Cell c = make_cell(s, fl, vh, st, v, g);
  const Stim stim = make_stim(s);
  const bool transmit = act && !(fl & FIS_OUT_OFF);
  const int64_t t_stop = act ? end_of(s, tb) : 0;
  {
    int sl = slot0;
#pragma unroll
    for (int k = 0; k < WMAX; ++k) {
      if (in[k] != 0) rp[(int64_t)sl * nb] = 0;
      sl = (sl + 1 == FIS_RING_LEN) ? 0 : sl + 1;
    }
  }

  // bit k: this lane fired at step k.  Waves with no fire mask off and no
  // hold in a full window take the kick recurrence.
  const bool kick_path =
      W == WMAX &&
      !any_lane(act && (fl & (FIS_FIRE_OFF | FIS_HOLD)) != 0u);
  // a wave whose every neuron-trial sits at a fixed point of the step with
  // no kick skips the recurrence (quiet_cell; the gate threshold and the
  // transmit flag do not matter without input or spikes)
  const bool quiet =
      kick_path && !any_lane(act && !(sb == 0u && quiet_cell(s, c, in)));
  uint32_t fired_bits = 0u;
  if (quiet)
    c.r += WMAX;
  else if (act)
    fired_bits = kick_path ? run_window_kick<WMAX>(s, c, in, sb, stim)
                           : run_window<WMAX>(s, c, in, sb, stim, W);
  const int64_t left = t_stop - t0;  // steps whose spikes still count
  const int kstop = left <= 0 ? 0 : (left >= W ? W : (int)left);
  const uint32_t nspk = __popc(fired_bits & ((1u << kstop) - 1u));

  // the stores and spikes out read the arguments again
  {
  const fis_sim& s = args_again();
  if constexpr (FUSED) {
    // below 64 trials the wave adds its own spikes into the ring slots no
    // step of this window reads (no event records, no scatter launch)
    if (act) {
      if (quiet)
        store_quiet(s, c, idx, st);
      else
        store_cell(s, c, idx, pre, nspk);
    }
    const uint32_t sent = transmit ? fired_bits : 0u;
    const bool any_sent = any_lane(sent != 0u);
    if (s.raster || any_sent) {
      int rrow = s.raster ? (int)(t0 % s.raster_len) : 0;
      int slot = (int)((t0 + kDelay + 1) % FIS_RING_LEN);
#pragma unroll 1
      for (int k = 0; k < W; ++k) {
        if (s.raster) {
          const uint64_t all =
              group_or(((fired_bits >> k) & 1u) ? (1ull << lane) : 0ull, kWave);
          s.raster[((int64_t)rrow * nblk + (wave >> 2)) * FIS_LINE_WORDS +
                   (wave & 3)] = all;
          rrow = (rrow + 1 == s.raster_len) ? 0 : rrow + 1;
        }
        const bool snd = (sent >> k) & 1u;
        if (bs > 1 && any_sent && any_lane(snd)) {
          // the trials of each neuron (bs lanes from trial 0) that sent
          const uint64_t sub = group_or(snd ? (1ull << tb) : 0ull, bs);
          scatter_leads(s, tb == 0 && sub != 0ull, pre, sub, slot, lane);
        }
        slot = (slot + 1 == FIS_RING_LEN) ? 0 : slot + 1;
      }
      // one trial per lane: each neuron that sent reads its CSR row once
      // and adds it into the slot of every step it sent at (a neuron
      // spiking twice in the window walks its row once, not twice)
      if (bs == 1 && any_sent) scatter_row_steps(s, sent, pre, t0, lane);
    }
    return;  // the clock moves on in k_tick, after every workgroup read it
  }

  // spikes out, step by step: raster rows, and events (below 64 trials
  // only steps with spikes append any, so a wave that stayed silent for the
  // whole window skips this)
  fis_event* evs = s.events + (int64_t)wave * s.ev_cap;
  int nev = 0;  // wave-uniform
  if (s.raster || bs >= kWave || any_lane(fired_bits != 0u)) {
    int rrow = s.raster ? (int)(t0 % s.raster_len) : 0;
    // the spikes of step t0 + k are summed into slot (t0 + k + 21) mod 40,
    // which step t0 + k + 20 reads
    int slot = (int)((t0 + kDelay + 1) % FIS_RING_LEN);
#pragma unroll 1
    for (int k = 0; k < W; ++k) {
      const bool fired = (fired_bits >> k) & 1u;
      emit(s, act, fired, fired && transmit, lane, pre, tb, wave, nblk, rrow,
           slot, k, evs, nev);
      rrow = (rrow + 1 == s.raster_len) ? 0 : rrow + 1;
      slot = (slot + 1 == FIS_RING_LEN) ? 0 : slot + 1;
    }
  }

  if (act) {
    if (quiet)
      store_quiet(s, c, idx, st);
    else
      store_cell(s, c, idx, pre, nspk);
  }
  if (bs < kWave && lane == 0 && act)
    s.ev_count[(int64_t)blockIdx.x * FIS_LINE_INTS + (wave & 3)] = nev;
  }
}

// ------------------------------------------------------------- scatter

// bstride < 64: one wave per update-wave segment (grid-stride) over the
// compacted events of the window.
__global__ __launch_bounds__(kBlock) void k_scatter(fis_sim s, int W) {
  // the window is complete: the next drive and update start W steps later
  // (no kernel of this launch reads the clock)
  if (blockIdx.x == 0 && threadIdx.x == 0 && !s.defer_tick) *s.clock += W;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nw = n_waves_of(s);
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t w = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6; w < nw;
       w += stride) {
    const int cnt = __builtin_amdgcn_readfirstlane(
        s.ev_count[(w >> 2) * FIS_LINE_INTS + (w & 3)]);
    if (cnt == 0) continue;
    const fis_event* evs = s.events + w * s.ev_cap;
    for (int e = 0; e < cnt; ++e) scatter_event(s, evs[e], lane);
  }
}

// bstride >= 64: step-major.  blockIdx.y is the window step; workgroups are
// dispatched x-fastest, so at any time the atomics of the whole device fall
// into one or two ring slots (142 MB each at 138,639 x 256), a working set
// the 256 MB memory-side cache can hold.
__global__ __launch_bounds__(kBlock) void k_scatter_steps(fis_sim s, int W) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && !s.defer_tick)
    *s.clock += W;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nw = n_waves_of(s);
  const int k = blockIdx.y;
  const int64_t stride = (int64_t)gridDim.x * (kBlock / kWave);
  for (int64_t w = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6; w < nw;
       w += stride) {
    const fis_event ev = s.events[w * s.ev_cap + k];
    if (ev.mask == 0) continue;
    scatter_event(s, ev, lane);
  }
}

// Current spikes (st bit 7, transmitted) into one ring slot.
__global__ __launch_bounds__(kBlock) void k_seed_ring(fis_sim s, int slot) {
  const int nb = s.n * s.bstride;
  const int idx = (int)blockIdx.x * kBlock + (int)threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const bool hot = idx < nb && (s.st[idx] & FIS_ST_SPIKE) &&
                   !(s.fl[idx] & FIS_OUT_OFF);
  uint64_t bal = __ballot(hot);
  const int base_idx = idx - lane;
  while (bal) {
    const int j = __ffsll((long long)bal) - 1;
    bal &= bal - 1ull;
    const int src = base_idx + j;
    const int pre = src / s.bstride;
    const int b = src - pre * s.bstride;
    int32_t* base = s.ring + (int64_t)slot * nb + b;
    for (int64_t k = s.offsets[pre] + lane; k < s.offsets[pre + 1]; k += kWave)
      __hip_atomic_fetch_add(base + (int64_t)s.cols[k] * s.bstride,
                             (int32_t)s.vals[k], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// One wave per listed neuron-trial whose transmit flag changed after its
// last spike: add (flag now clear) or subtract (flag now set) its row.
__global__ __launch_bounds__(kBlock) void k_resend(fis_sim s,
                                                   const int32_t* idx,
                                                   int n_idx, int slot) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = (int)((blockIdx.x * kBlock + threadIdx.x) >> 6);
  if (w >= n_idx) return;
  const int src = idx[w];
  if (src < 0 || src >= s.n * s.bstride) return;  // not a neuron-trial
  if (!(s.st[src] & FIS_ST_SPIKE)) return;
  const int32_t sign = (s.fl[src] & FIS_OUT_OFF) ? -1 : 1;
  const int pre = src / s.bstride;
  const int b = src - pre * s.bstride;
  const int64_t nb = (int64_t)s.n * s.bstride;
  int32_t* base = s.ring + (int64_t)slot * nb + b;
  for (int64_t k = s.offsets[pre] + lane; k < s.offsets[pre + 1]; k += kWave)
    __hip_atomic_fetch_add(base + (int64_t)s.cols[k] * s.bstride,
                           sign * (int32_t)s.vals[k], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------- pull road
//
// One launch per window.  A workgroup ("tile") is TILE neuron-trials: P =
// TILE / bstride whole neurons when TILE >= bstride, else one neuron's
// trials split into bstride / TILE parts (bstride a power of two, 64..1024).
// Records are laid out [pre][slot][word]: one neuron's 40 steps x bstride
// bits (320 B at 64 trials) belong to the workgroups that own the neuron,
// and the 20 steps of one delay horizon sit in a few lines for the gather.
//
//  1. gather: the tile's in-edges (W CSR by post) are walked once per
//     64-trial word; for each in-edge whose presynaptic neuron had a spike
//     in the tile's trials at some step t0-20+j of the window's delay
//     horizon (act word of the neuron and part), the record word of that
//     step is read and the synapse count is added into LDS
//     acc[j][post][trial] for every set trial bit -- exact int32,
//     order-free;
//  2. update: W steps of lif_step per neuron-trial with I = acc[j][...],
//     state in registers; each wave (one neuron x 64 trials) writes its
//     spike record words of steps t0..t0+W-1;
//  3. the tile's act words (one per neuron and part) get the window's bits
//     (device atomics: the act words of several tiles share lines).
// Records and act bits of steps t0-20.. (read) and t0.. (written) are
// disjoint slots mod 40 for W <= 20, so no workgroup reads what another
// writes in the same launch.

constexpr int kRec = FIS_REC_SLOTS;
constexpr uint64_t kMask40 = (1ull << kRec) - 1;

// rotate the low 40 bits of x left by r (0 <= r < 40)
__device__ __forceinline__ uint64_t rot40(uint64_t x, int r) {
  x &= kMask40;
  return r == 0 ? x : (((x << r) | (x >> (kRec - r))) & kMask40);
}

// acc[b * stride] += val for every set bit b of the 64-trial word x
template <int stride>
__device__ __forceinline__ void add_bits(int32_t* acc, uint64_t x,
                                         int32_t val) {
  while (x) {
    const int b = __ffsll((long long)x) - 1;
    x &= x - 1ull;
    atomicAdd(acc + b * stride, val);
  }
}

// Record lists: 16 bytes per (neuron, tile part, record slot); byte 0 the
// number of listed trials (<= FIS_REC_LIST_MAX) or FIS_REC_LIST_WORDS, bytes
// 1.. the trials' numbers within the part.
constexpr uint32_t kListMax = FIS_REC_LIST_MAX;

// acc[t * stride] += val for the first `cnt` trials t of list l (one LDS add
// per listed trial; the wave runs as many rounds as its longest list)
template <int stride>
__device__ __forceinline__ void add_list(int32_t* acc, const uint4& l,
                                         uint32_t cnt, int32_t val) {
#pragma unroll
  for (int i = 0; i < (int)kListMax; ++i) {
    if (!any_lane((uint32_t)i < cnt)) break;
    if ((uint32_t)i < cnt) {
      const int byte = i + 1;
      const uint32_t d = byte < 4 ? l.x : (byte < 8 ? l.y : (byte < 12 ? l.z : l.w));
      atomicAdd(acc + ((d >> ((byte & 3) * 8)) & 0xFFu) * stride, val);
    }
  }
}

// prepend byte b to the 16-byte list l (bytes move up by one)
__device__ __forceinline__ void list_push(uint4& l, uint32_t b) {
  l.w = (l.w << 8) | (l.z >> 24);
  l.z = (l.z << 8) | (l.y >> 24);
  l.y = (l.y << 8) | (l.x >> 24);
  l.x = (l.x << 8) | b;
}

// One active (in-edge, step) pair of a wave's 64 in-edges, dealt to the
// lane that takes pair number pi: lane L of the 64 owns pairs
// excl(L) .. excl(L) + popc(m(L)) - 1, one per set step bit of m(L).  All
// 64 lanes must take part (the owner is found by a forward permute and a
// DPP max-scan, its fields read by lane shuffles).
struct Pair {
  int src;      // presynaptic neuron
  int32_t val;  // synapse count
  int pos;      // position of the postsynaptic neuron in the tile
  int step;     // step of the horizon (0 .. W-1)
};

// Inclusive max-scan over the 64 lanes of a wave, x >= 0 (DPP row shifts
// within rows of 16, then the row broadcasts of lanes 15 and 31).
__device__ __forceinline__ int wave_max_scan(int x) {
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false));
  return x;
}

// Inclusive sum-scan over the 64 lanes of a wave (the same DPP steps as
// wave_max_scan; lanes outside the shifted rows add zero).
__device__ __forceinline__ int wave_add_scan(int x) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);
  return x;
}

// The owner of pair `base + lane` (base uniform): the last lane L with
// excl(L) <= base + lane.  The owner of pair `base` is the last lane whose
// run starts at or before it (a ballot); every lane with pairs whose run
// starts inside the round pushes its number to the lane of its start
// (forward permute: the starts of lanes with pairs are distinct), and a
// max-scan carries each start to the lanes after it.  No shuffle chain.
__device__ __forceinline__ int round_owner(int base, int lane, int excl,
                                           uint32_t m) {
  const int carry = 63 - __builtin_clzll(__ballot(excl <= base));
  const int d = excl - base;
  const bool sends = m != 0u && d > 0 && d < kWave;
  int got = __builtin_amdgcn_ds_permute((sends ? d : 0) << 2,
                                        sends ? lane + 1 : 0);
  if (lane == 0) got = carry + 1;  // only lanes without a start sent here
  return wave_max_scan(got) - 1;
}

// Positions of the set bits of a 5-bit value c, lowest first, 3 bits
// each: entry c of the table deal_pair reads (lane c holds entry c).
struct BitPos {
  uint32_t v[32];
};
constexpr BitPos make_bit_pos() {
  BitPos t{};
  for (uint32_t c = 0; c < 32; ++c) {
    uint32_t e = 0, n = 0;
    for (uint32_t b = 0; b < 5; ++b)
      if ((c >> b) & 1u) e |= b << (3 * n++);
    t.v[c] = e;
  }
  return t;
}
__device__ const BitPos kBitPos = make_bit_pos();

// `bitpos`: this lane's entry of kBitPos (entry lane mod 32).  Masks of at
// most 20 steps (four 5-bit chunks).
__device__ __forceinline__ Pair deal_pair(int pi, int excl, uint32_t m,
                                          int src, int32_t val, int pos,
                                          uint32_t bitpos) {
  const int lane = __lane_id();
  const int o = round_owner(pi - lane, lane, excl, m);
  const int eo = __shfl(excl, o);
  Pair p;
  const uint32_t mo = (uint32_t)__shfl((int)m, o);
  p.src = __shfl(src, o);
  p.val = __shfl(val, o);
  p.pos = __shfl(pos, o);
  // the step: the k-th set bit of the owner's mask.  The 5-bit chunk that
  // holds it by the chunks' running counts, then its place in the chunk
  // from the table (a lane permute: lane c holds chunk value c's entry).
  const int k = pi - eo;
  const int n0 = __popc(mo & 31u);
  const int n1 = n0 + __popc((mo >> 5) & 31u);
  const int n2 = n1 + __popc((mo >> 10) & 31u);
  int below = 0, sh = 0;
  if (k >= n0) { below = n0; sh = 5; }
  if (k >= n1) { below = n1; sh = 10; }
  if (k >= n2) { below = n2; sh = 15; }
  const uint32_t chunk = (mo >> sh) & 31u;
  const uint32_t ent =
      (uint32_t)__builtin_amdgcn_ds_bpermute((int)(chunk << 2), (int)bitpos);
  p.step = sh + (int)((ent >> (3 * (k - below))) & 7u);
  return p;
}

// Waves of a one-neuron tile that read the same in-edges and share their
// rounds of pairs.  Measured at 256 trials, window ms (drive + pull): 1
// (each wave its own in-edges) 1.014; 4 (all waves) 0.975; 2 0.898 --
// pairs of waves balance the typical 100-odd in-edges and still give each
// wave two rounds to overlap.
#ifndef FIS_PULL_SHARE
#define FIS_PULL_SHARE 2
#endif
constexpr int kPullShare = FIS_PULL_SHARE;
// Rounds of 64 dealt pairs a wave keeps in flight at once (owner searches
// and record loads of the rounds overlap).  Measured at 138,639 neurons
// (ms per window, drive included): 1 round 0.874 (B = 256) and 0.162
// (B = 64), 2 rounds 0.878 / 0.172, 3 0.894 / 0.191, 4 1.09 / 0.213.
#ifndef FIS_PULL_ROUNDS
#define FIS_PULL_ROUNDS 1
#endif
constexpr int kRounds = FIS_PULL_ROUNDS;
// ... for tiles that read the record lists (parts of 128 and 256 trials),
// where a round waits for its lists: 2 rounds 32.7 us per step against
// 33.3 with 1 (256 trials, sensory drive, graph-replayed windows)
#ifndef FIS_PULL_ROUNDS_LISTS
#define FIS_PULL_ROUNDS_LISTS 2
#endif
constexpr int kRoundsLists = FIS_PULL_ROUNDS_LISTS;

// Neuron-trials per workgroup (fis_sim.pull_tile; 0 = FIS_PULL_TILE).  256
// lanes (4 waves, 20 KB of accumulators) measured fastest at 256 trials:
// smaller workgroups wait less at the barrier between gather and update for
// their slowest wave, and 8 of them fit a CU; 128 and 512 are compiled too
// and picked by measurement where they win.
__host__ __device__ inline int pull_tile_of(const fis_sim& s) {
  return s.pull_tile ? s.pull_tile : FIS_PULL_TILE;
}

// trials of one neuron in one tile, and tiles per neuron (act words per
// neuron), for a trial stride and tile of the pull road
__host__ __device__ inline int pull_trials_per_part(int bs, int tile) {
  return bs < tile ? bs : tile;
}
__host__ __device__ inline int pull_parts(int bs, int tile) {
  return bs / pull_trials_per_part(bs, tile);
}

// Gated pull road, first launch of a window: one lane per tile (the tile
// numbering of k_pull), kPlanTiles tiles per thread.  A tile whose every
// (neuron, part) is calm and to which no input is due in this window (see
// k_pull) has its window done here: the act bits of the window's steps are
// cleared, and its raster rows; nothing else of it is read or written.
// Its record words and lists of those steps keep stale contents that
// nothing reads: every reader reads them only under a set act bit, and the
// one writer that sets act bits outside a window, the m_out re-send
// (k_pull_resend), sets them only for a neuron-trial that spiked at the
// last step, which no trial of a calm part did.  Every other tile is
// appended to the window's work list, which the k_pull launch that follows
// walks -- so the many calm tiles of a sparse network cost a lane each, not
// a workgroup each.
//
// work[]: [0], [1] the list lengths of the two parities, [2] the parity of
// the next window, [3] that of the current one, [4 ..] the list.  A window
// of parity p appends to work[p] (one atomic per workgroup: appends of
// every wave to one counter serialise); its k_pull reads work[3] and
// work[p], flips work[2] and zeroes work[1 - p] for the next window -- no
// counter is read and reset in one launch.
constexpr int kPlanTiles = 4;

template <int BS, int TILE>
__global__ __launch_bounds__(kBlock) void k_pull_plan(fis_sim s, int W,
                                                      int ring_steps,
                                                      int n_tiles) {
  constexpr int bs = BS;
  constexpr int tpn = TILE < BS ? TILE : BS;  // trials of a neuron per tile
  constexpr int P = TILE / tpn;               // neurons per tile
  constexpr int parts = BS / tpn;             // tiles per neuron
  constexpr int words = BS >> 6;              // 64-trial words per neuron
  constexpr int wpt = tpn >> 6;               // of them in one tile
  __shared__ int wave_n[kBlock / kWave];      // busy tiles per wave
  __shared__ int wave_at[kBlock / kWave];     // where each wave's go
  const int tid = (int)threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int first = (int)blockIdx.x * kBlock * kPlanTiles + tid;
  const int64_t t0 = now(s);
  const int par = s.work[2];
  if (blockIdx.x == 0 && tid == 0) s.work[3] = par;
  auto neuron_at = [&](int p) { return s.order ? s.order[p] : p; };
  uint32_t busy = 0;  // bit k: tile first + k * kBlock goes on the list
#pragma unroll
  for (int k = 0; k < kPlanTiles; ++k) {
    const int tile = first + k * kBlock;
    if (tile >= n_tiles) continue;
    const int p0 = (tile / parts) * P, part = tile % parts;
    const int np = min(P, s.n - p0);
    int pre[P];
    int64_t latest = -1;
    bool calm = ring_steps == 0 && s.calm != nullptr;
    uint64_t act_old[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
      pre[q] = neuron_at(p0 + min(q, np - 1));
      act_old[q] = 0;
      if (q < np) {
        const int64_t u = (int64_t)pre[q] * parts + part;
        const int64_t d = s.due[u];
        latest = d > latest ? d : latest;
        if (s.calm && !s.calm[u]) calm = false;
        act_old[q] = s.act[u];
      }
    }
    if (latest >= t0 || !calm) {
      busy |= 1u << k;
      continue;
    }
    const int wbase = (int)(t0 % kRec);  // slot of step t0
    const uint64_t clear = ~rot40((1ull << W) - 1, wbase);
#pragma unroll
    for (int q = 0; q < P; ++q) {
      if (q >= np) continue;
      const uint64_t nw = act_old[q] & clear;
      if (nw != act_old[q]) s.act[(int64_t)pre[q] * parts + part] = nw;
    }
    if (s.raster) {  // the window's raster rows of the tile's live words
      const int w0 = part * wpt;
      const int nblk = (s.n * bs + kBlock - 1) / kBlock;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (q >= np) continue;
#pragma unroll
        for (int j = 0; j < wpt; ++j) {
          if (s.word_end && !(t0 < s.word_end[w0 + j])) continue;
          const int wave = pre[q] * words + w0 + j;
          int rrow = (int)(t0 % s.raster_len);
          for (int r = 0; r < W; ++r) {
            s.raster[((int64_t)rrow * nblk + (wave >> 2)) * FIS_LINE_WORDS +
                     (wave & 3)] = 0ull;
            rrow = rrow + 1 == s.raster_len ? 0 : rrow + 1;
          }
        }
      }
    }
  }
  // this thread's busy tiles go to list positions base + (busy tiles of
  // the waves before it) + (of the lanes before it): a prefix of the
  // counts (<= kPlanTiles) by a ballot per bit, then one atomic per
  // workgroup
  const int c = __popc(busy);
  int excl = 0, total = 0;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const uint64_t mb = __ballot((c >> b) & 1);
    excl += (int)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(mb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mb, 0u))
            << b;
    total += __popcll(mb) << b;
  }
  const int wv = tid >> 6;
  if (lane == 0) wave_n[wv] = total;
  __syncthreads();
  if (tid == 0) {
    int sum = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) {
      wave_at[w] = sum;
      sum += wave_n[w];
    }
    const int base = sum ? atomicAdd(s.work + par, sum) : 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) wave_at[w] += base;
  }
  __syncthreads();
  int at = wave_at[wv] + excl;
#pragma unroll
  for (int k = 0; k < kPlanTiles; ++k)
    if ((busy >> k) & 1u) {
      // (at most n_tiles entries: the counter starts the window at zero)
      if (at < n_tiles) s.work[4 + at] = first + k * kBlock;
      ++at;
    }
}

// Workgroups of a gated window (which walk its work list), in waves: twice
// the 8,192 waves of k_pull the chip holds at once (256 CUs x 32), so that
// workgroups that finish early are followed by others.  Measured at 256
// trials, 138,639 neurons, plain drive: 2.19 us per step against 2.26 with
// one resident grid.
#ifndef FIS_PULL_LIST_WAVES
#define FIS_PULL_LIST_WAVES 16384
#endif
constexpr int kPullListWaves = FIS_PULL_LIST_WAVES;

// One tile's window of k_pull (below): `acc` is the workgroup's LDS.
template <int WMAX, int BS, int TILE, bool GATE>
__device__ __forceinline__ void pull_tile(const fis_sim& s, int32_t* acc,
                                          int tile, int W, int ring_steps,
                                          int tid) {
  constexpr int bs = BS;                        // trial stride (== s.bstride)
  constexpr int tpn = TILE < BS ? TILE : BS;    // trials of a neuron per tile
  constexpr int P = TILE / tpn;                 // neurons per tile
  constexpr int parts = BS / tpn;               // tiles per neuron
  constexpr int words = BS >> 6;                // 64-trial words per neuron
  constexpr int wpt = tpn >> 6;                 // of them in one tile
  // parts of 128 and 256 trials read the record lists (trial numbers fit a
  // byte; below 128 trials a list is no shorter than the one word)
  constexpr bool LISTS = wpt >= 2 && tpn <= 256;
  // rounds in flight (the gated kernel's work-list loop leaves no
  // registers for a second round: it would spill)
  constexpr int NR = LISTS && !GATE ? kRoundsLists : kRounds;
  const bool lists = LISTS && s.rec_ids != nullptr;
  const int lane = tid & (kWave - 1);
  const int pl = tid / tpn;                     // position in the tile
  const int tl = tid - pl * tpn;                // trial within the part
  const int64_t t0 = now(s);
  const int rbase = (int)((t0 + kRec - 20) % kRec);  // slot of step t0-20
  const int wbase = (int)(t0 % kRec);                // slot of step t0
  const int pstride = kRec * words;                  // words per neuron
  const uint64_t wmask = (1ull << W) - 1;
  const int nb = s.n * bs;
  const int nblk = (nb + kBlock - 1) / kBlock;  // raster rows: 256-lane groups
  // neuron at position p of the order (a null order is the identity)
  auto neuron_at = [&](int p) { return s.order ? s.order[p] : p; };
  // one neuron per tile: the waves form groups of SG that read the same 64
  // in-edges and share their rounds of pairs (balanced however few the
  // edges); several neurons per tile: each wave takes its own 64 in-edges
  const int wv = tid >> 6;                     // wave of the tile
  constexpr int NW = TILE / kWave;             // waves per tile
  constexpr int SG = P == 1 ? kPullShare : 1;  // waves per group
  constexpr int NG = NW / SG;                  // groups per tile
  const int grp = wv / SG, mem = wv % SG;
  constexpr int kEdgeStep = NG * kWave;
  constexpr int kRoundGap = SG * kWave;

  // What a tile needs from memory before its gather: loaded ahead.
  struct Tile {
    int p0, part, w0, np;
    bool gather;
    uint32_t calm_old;    // bit q: neuron q of the tile was calm
    uint32_t live;        // bit j: the tile's word j is still simulated
    uint64_t act_old;     // act word of neuron `tid` (tid < np)
    int pre, idx, tb;     // this lane's neuron, neuron-trial, trial
    float v, g;
    uint32_t st, fl, sb;
    int off[P + 1];       // in-edge offsets
    int e1;               // end of the tile's in-edges
    int src;              // this lane's first in-edge
    int32_t val;
    uint64_t a;           // and its act word
  };
  // the tile, whether it is calm, and its act words
  auto head = [&](int tile) {
    Tile T;
    T.p0 = (tile / parts) * P;   // first position of the tile
    T.part = tile % parts;       // which part of its neuron(s)
    T.w0 = T.part * wpt;         // first word of the tile
    T.np = min(P, s.n - T.p0);
    // input gating: skip the in-edge walk when no input sent to the tile's
    // neurons (in its trials) arrives at or after t0 (uniform)
    T.gather = true;
    if (GATE) {
      int64_t latest = -1;
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (q < T.np) {
          const int64_t d =
              s.due[(int64_t)neuron_at(T.p0 + q) * parts + T.part];
          latest = d > latest ? d : latest;
        }
      T.gather = latest >= t0;
    }
    // this tile alone writes the act word of each of its neurons (for its
    // part); read now, updated at the end with a plain store if it changes
    // (other tiles read only the bits of the delay horizon, which the
    // update keeps).  Every lane loads one (a vector load, its own
    // address), so that nothing waits for it before the end.
    T.act_old = s.act[(int64_t)neuron_at(T.p0 + opaque_vgpr(min(tid, T.np - 1))) *
                          parts + T.part];
    // which neurons of the tile were calm (a tile with no input due whose
    // every neuron was calm had its window done by k_pull_plan and is not
    // on the work list)
    T.calm_old = 0;
    if (GATE && s.calm) {
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (q < T.np && s.calm[(int64_t)neuron_at(T.p0 + q) * parts + T.part])
          T.calm_old |= 1u << q;
    }
    T.tb = (T.w0 << 6) + tl;
    return T;
  };
  // the rest of what the window reads before its gather: the loads that
  // depend on nothing else are issued first, and nothing waits for them
  // before its result is needed
  auto load = [&](Tile& T) {
    T.pre = pl < T.np ? neuron_at(T.p0 + pl) : 0;
    T.idx = T.pre * bs + T.tb;
    T.live = (1u << wpt) - 1u;
    if (s.word_end) {
      int64_t we[wpt];
#pragma unroll
      for (int j = 0; j < wpt; ++j) we[j] = s.word_end[T.w0 + j];
      T.live = 0;
#pragma unroll
      for (int j = 0; j < wpt; ++j)
        if (t0 < we[j]) T.live |= 1u << j;
    }
    T.v = 0.0f;
    T.g = 0.0f;
    T.st = 0;
    T.fl = FIS_FIRE_OFF;
    T.sb = 0;
    T.src = 0;
    T.val = 0;
    T.a = 0;
#pragma unroll
    for (int q = 0; q <= P; ++q) T.off[q] = 0x7fffffff;
    if (pl < T.np) {  // in bounds whether or not the trial's word is live
      T.v = s.v[T.idx];
      T.g = s.g[T.idx];
      T.st = s.st[T.idx];
      T.fl = s.fl[T.idx];
      T.sb = s.sbits[T.idx];
    }
    // the tile's in-edge offsets: uniform, held in scalar registers
#pragma unroll
    for (int q = 0; q <= P; ++q)
      T.off[q] = q <= T.np ? s.in_offsets[T.p0 + q] : 0x7fffffff;
    T.e1 = T.off[0];
#pragma unroll
    for (int q = 1; q <= P; ++q)
      if (q <= T.np) T.e1 = T.off[q];
  };
  // the lane's first in-edge, then its act word
  auto first_edge = [&](Tile& T) {
    const int e = T.off[0] + lane * NG + grp;
    if (T.gather && e < T.e1) {
      T.src = s.in_cols[e];
      T.val = s.in_vals[e];
    }
  };
  auto first_act = [&](Tile& T) {
    const int e = T.off[0] + lane * NG + grp;
    if (T.gather && e < T.e1)
      T.a = s.act[T.src * parts + T.part];
  };

  // the tile's window
  auto run = [&](Tile& T) {
    const int p0 = T.p0, part = T.part, w0 = T.w0, np = T.np;
    const int pre = T.pre, idx = T.idx, tb = T.tb;
    const int wave = idx >> 6;
    const uint32_t live_words = T.live;
    // uniform over a wave: a neuron of the tile, in a live 64-trial word
    const bool act = pl < np && (!s.word_end || t0 < s.word_end[tb >> 6]);
    // the lane's entry of the table deal_pair reads (loaded during the
    // accumulator clear)
    static_assert(WMAX <= 20, "deal_pair reads step masks of 20 bits");
    const uint32_t bitpos = kBitPos.v[lane & 31];

    // the accumulators cleared and the waves met while the first in-edges
    // are on their way (the barrier's wait for the slowest wave hides in
    // that latency); then their act words.  The zeros are made here (kept
    // from the kernel's start instead, the work-list loop spilled them to
    // scratch and read them back for every tile).
    int z = 0;
    asm volatile("" : "+v"(z));
    const int4 z4 = make_int4(z, z, z, z);
#pragma unroll
    for (int j = 0; j < WMAX; j += 4) *(int4*)(acc + tid * WMAX + j) = z4;
    __syncthreads();
    first_act(T);

    // 1. gather.  The tile's in-edges are read 64 at a time, one per lane,
    // with their presynaptic act words (the next 64 are fetched meanwhile;
    // the first 64 before the window starts).  The active (in-edge, step)
    // pairs of those 64 -- a step of the horizon at which the presynaptic
    // neuron spiked in some trial of the tile -- are then dealt out 64 at a
    // time, one per lane: the lane reads that step's list (or record words)
    // of the tile's trials and adds the synapse count at each listed trial.
    // A lane thus carries a few trials of one spike of one neuron, and the
    // walk over them is as long as the longest such list of the round (a
    // handful), not as long as the densest of 20 steps of the in-edge.
    if (T.gather) {
      const int e1 = T.e1;
      // this lane's in-edge: the groups take the tile's in-edges in turn,
      // lane by lane, so that they get as many as one another
      int e = T.off[0] + lane * NG + grp;
      int src = T.src;
      int32_t val = T.val;
      uint64_t a = T.a;
      const int rot = (kRec - rbase) % kRec;
      // uniform over the wave: its first in-edge is below e1
      while (e - lane * NG < e1) {
        const int en = e + kEdgeStep;
        int src_n = 0;
        int32_t val_n = 0;
        if (en < e1) {
          src_n = s.in_cols[en];
          val_n = s.in_vals[en];
        }
        int pq = 0;  // position of the in-edge's postsynaptic neuron
#pragma unroll
        for (int q = 1; q < P; ++q) pq += e >= T.off[q] ? 1 : 0;
        const uint32_t m = (e < e1 && live_words)
                               ? (uint32_t)(rot40(a, rot) & wmask) : 0u;
        const uint64_t a_n = en < e1 ? s.act[src_n * parts + part] : 0ull;
        // exclusive prefix of the lanes' active steps, and their total (a
        // DPP sum-scan; no round trip through LDS)
        const int pc = __popc(m);
        const int incl = wave_add_scan(pc);
        const int excl = incl - pc;
        const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
        // rounds of 64 pairs, NR at a time (their owner searches and
        // record loads overlap); shared batches hand the rounds to the
        // waves in turn
        for (int base = mem * kWave; base < total;
             base += NR * kRoundGap) {
          Pair pr[NR];
#pragma unroll
          for (int r = 0; r < NR; ++r)
            pr[r] = deal_pair(base + r * kRoundGap + lane, excl, m, src, val,
                              pq, bitpos);
          if constexpr (LISTS) {
            // the step's list of the part: 16 bytes per pair; a lane whose
            // list overflowed (or was marked) reads the record words
            uint4 li[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              const bool ok = base + r * kRoundGap + lane < total;
              int sl = rbase + pr[r].step;
              sl = sl >= kRec ? sl - kRec : sl;
              li[r] = make_uint4(ok ? FIS_REC_LIST_WORDS : 0u, 0u, 0u, 0u);
              if (ok && lists)
                li[r] = *(const uint4*)(s.rec_ids +
                                        (((int64_t)pr[r].src * parts + part) *
                                             kRec + sl) * 16);
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              int32_t* at = acc + pr[r].pos * tpn * WMAX + pr[r].step;
              // the synapse count arrives by a lane shuffle: wait for it
              // here, once, so that the LDS adds below do not each wait for
              // the previous one
              int32_t cv = pr[r].val;
              asm volatile("s_waitcnt lgkmcnt(0)" : "+v"(cv));
              const uint32_t cnt = li[r].x & 0xFFu;
              const bool fb = cnt > kListMax;
              if (any_lane(fb)) {  // the record words of those lanes
                uint64_t xw[wpt];
#pragma unroll
                for (int q = 0; q < wpt; ++q) xw[q] = 0ull;
                if (fb) {
                  int sl = rbase + pr[r].step;
                  sl = sl >= kRec ? sl - kRec : sl;
                  const uint64_t* rw =
                      s.rec + (int64_t)pr[r].src * pstride + sl * words + w0;
#pragma unroll
                  for (int q = 0; q < wpt; q += 2) {
                    const ulonglong2 w2 = *(const ulonglong2*)(rw + q);
                    xw[q] = (live_words >> q) & 1u ? w2.x : 0ull;
                    xw[q + 1] = (live_words >> (q + 1)) & 1u ? w2.y : 0ull;
                  }
                }
#pragma unroll
                for (int q = 0; q < wpt; ++q)
                  add_bits<WMAX>(at + q * kWave * WMAX, xw[q], cv);
              }
              add_list<WMAX>(at, li[r], fb ? 0u : cnt, cv);
            }
          } else {
            uint64_t xw[NR][wpt];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              const bool ok = base + r * kRoundGap + lane < total;
              int sl = rbase + pr[r].step;
              sl = sl >= kRec ? sl - kRec : sl;
              const uint64_t* rw =
                  s.rec + (int64_t)pr[r].src * pstride + sl * words + w0;
#pragma unroll
              for (int q = 0; q < wpt; ++q) xw[r][q] = 0ull;
              if (ok) {
                // the tile's words of the step: 8 * wpt contiguous bytes,
                // 16-byte aligned from two words up (one vector load per 16
                // bytes)
                if constexpr (wpt >= 2) {
#pragma unroll
                  for (int q = 0; q < wpt; q += 2) {
                    const ulonglong2 w2 = *(const ulonglong2*)(rw + q);
                    xw[r][q] = w2.x;
                    xw[r][q + 1] = w2.y;
                  }
                } else {
                  xw[r][0] = rw[0];
                }
              }
              if (live_words != (1u << wpt) - 1u) {  // idle 64-trial words
#pragma unroll
                for (int q = 0; q < wpt; ++q)
                  if (!((live_words >> q) & 1u)) xw[r][q] = 0ull;
              }
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) {
              int32_t* at = acc + pr[r].pos * tpn * WMAX + pr[r].step;
              int32_t cv = pr[r].val;
              asm volatile("s_waitcnt lgkmcnt(0)" : "+v"(cv));
#pragma unroll
              for (int q = 0; q < wpt; ++q)
                add_bits<WMAX>(at + q * kWave * WMAX, xw[r][q], cv);
            }
          }
        }
        e = en;
        src = src_n;
        val = val_n;
        a = a_n;
      }
    }
    __syncthreads();

    // from here on the arguments are read again (`s` is the kernel argument
    // in both callers)
    const fis_sim& s = args_again();

    // 2. update
    float v = T.v, g = T.g, vh = 0.0f;
    uint32_t st = T.st, fl = T.fl, sb = T.sb;
    if (!act) {  // a lane past the tile's neurons or in a word that has ended
      v = 0.0f;
      g = 0.0f;
      st = 0;
      fl = FIS_FIRE_OFF;
      sb = 0;
    }
    if (sb) s.sbits[idx] = 0u;  // the window's kicks are consumed
    // the clamp value is read only where the flag says HOLD (by now fl has
    // long arrived; 4 B per neuron-trial and window not read otherwise)
    if (act && (fl & FIS_HOLD) && s.vhold) vh = s.vhold[idx];
    Cell c = make_cell(s, fl, vh, st, v, g);
    const Stim stim = make_stim(s);
    const bool transmit = act && !(fl & FIS_OUT_OFF);
    // steps of this window whose spikes still count for the trial
    const int64_t left = act ? end_of(s, tb) - t0 : 0;
    const int kstop = left <= 0 ? 0 : (left >= W ? W : (int)left);

    // the window's inputs, all read before the recurrence starts (nothing
    // in it waits for memory)
    int32_t in[WMAX];
#pragma unroll
    for (int k = 0; k < WMAX; ++k) in[k] = acc[tid * WMAX + k];  // 0 past W
    if (ring_steps > 0 && act) {  // input carried over from a restore
      int32_t* rp = s.ring + idx;
      int rslot = (int)((t0 + 1) % FIS_RING_LEN);
#pragma unroll
      for (int k = 0; k < WMAX; ++k) {
        if (k < W && k < ring_steps) {
          const int32_t x = rp[(int64_t)rslot * nb];
          if (x != 0) {
            in[k] += x;
            rp[(int64_t)rslot * nb] = 0;
          }
        }
        rslot = (rslot + 1 == FIS_RING_LEN) ? 0 : rslot + 1;
      }
    }

    // bit k: this lane fired at step k.  Waves with no flag, no kick and no
    // trial end in a full window (most of them) take the plain recurrence;
    // of those, waves whose every lane sits at a fixed point of the step
    // skip it (quiet_cell).
    const bool plain =
        W == WMAX && !any_lane(!act || fl != 0u || sb != 0u || kstop != W);
    // (a nonzero conductance or a spike anywhere in the wave rules quiet
    // out before the 20 inputs are looked at)
    const bool quiet =
        plain && !any_lane(c.sp || __float_as_uint(c.g) != 0u) &&
        !any_lane(!quiet_cell(s, c, in));
    // of the rest, waves with every input gate open try the window without
    // gate and reset first (most waves: most neurons spike in no trial)
    const bool open =
        plain && !quiet &&
        !any_lane(c.sp || c.r + 1 < (int)s.refrac_steps);
    uint32_t fired_bits = 0u;
    if (quiet)
      c.r += WMAX;  // the counter's WMAX increments; v, g and sp stay
    else if (open && run_window_open<WMAX>(s, c, in))
      ;  // no lane fired: fired_bits stay 0
    else if (plain)
      fired_bits = run_window_plain<WMAX>(s, c, in);
    // waves with a flag but neither fire-off nor hold (driven neurons
    // without refractoriness, silenced outputs): the kick recurrence
    else if (act && W == WMAX &&
             !any_lane((fl & (FIS_FIRE_OFF | FIS_HOLD)) != 0u))
      fired_bits = run_window_kick<WMAX>(s, c, in, sb, stim);
    else if (act)
      fired_bits = run_window<WMAX>(s, c, in, sb, stim, W);
    const uint32_t nspk = __popc(fired_bits & ((1u << kstop) - 1u));
    if (act) {
      if (quiet)
        store_quiet(s, c, idx, st);
      else
        store_cell(s, c, idx, pre, nspk);
    }
    // calm after this window: quiet with the refractory counter saturated
    // (the next quiet window stores nothing then); kept in the wave's cell
    // 4 WMAX + 1 below
    const bool wave_calm = GATE && s.calm && quiet && act &&
                           !any_lane(c.r < (int)FIS_ST_RMAX);

    // The wave's record words of the window (and raster rows): lane k < W
    // owns step k's.  This wave's 64 x WMAX cells are its own and already
    // read; cells 4k .. 4k+3 collect step k's two words (LDS atomics of the
    // lanes that fired, few), cell 4 WMAX the steps at which the wave sent
    // anything.
    {
      const uint32_t sent = transmit ? fired_bits : 0u;
      int32_t* mine = acc + (tid & ~(kWave - 1)) * WMAX;  // this wave's cells
      uint64_t* cell = (uint64_t*)(mine + 4 * (lane < W ? lane : 0));
      if (lane < W) {  // (zeros made here, as for the accumulators)
        uint64_t z = 0ull;
        asm volatile("" : "+v"(z));
        cell[0] = z;
        cell[1] = z;
      }
      if (lane == W) mine[4 * WMAX] = 0;
      if (lane == 0) mine[4 * WMAX + 1] = wave_calm ? 1 : 0;
      __builtin_amdgcn_wave_barrier();
      const uint64_t me = 1ull << lane;
      for (uint32_t x = sent; x; x &= x - 1u)
        atomicOr((uint64_t*)(mine + 4 * (__ffs(x) - 1)), me);
      if (s.raster)
        for (uint32_t x = fired_bits; x; x &= x - 1u)
          atomicOr((uint64_t*)(mine + 4 * (__ffs(x) - 1)) + 1, me);
      __builtin_amdgcn_wave_barrier();
      if (lane < W && cell[0] != 0ull) atomicOr(mine + 4 * WMAX, 1 << lane);
    }
    __syncthreads();

    // The record words: step k's word of this wave is stored when the
    // lane's neuron sent anything at step k in this part (its act bit is
    // then set, and readers read every word of the part at that step), and
    // always for the window's last step (an m_out re-send may set an act
    // bit there later).  Other slots keep stale words that no reader reads:
    // their act bits are clear.
    uint32_t part_sent = 0;  // steps at which the lane's neuron sent (part)
#pragma unroll
    for (int j = 0; j < wpt; ++j)
      part_sent |= (uint32_t)acc[(pl * tpn + j * kWave) * WMAX + 4 * WMAX];
    if (lane < W && act) {
      // `act` is uniform over a wave (n * bstride is a multiple of 64): the
      // waves past the last neuron of the last tile store nothing
      const uint64_t* cell =
          (const uint64_t*)(acc + (tid & ~(kWave - 1)) * WMAX + 4 * lane);
      if (((part_sent >> lane) & 1u) || lane == W - 1) {
        const int slot =
            wbase + lane < kRec ? wbase + lane : wbase + lane - kRec;
        s.rec[(int64_t)pre * pstride + slot * words + (tb >> 6)] = cell[0];
      }
      if (s.raster) {
        int rrow = (int)(t0 % s.raster_len) + lane;  // raster_len >= 20
        rrow = rrow >= s.raster_len ? rrow - s.raster_len : rrow;
        s.raster[((int64_t)rrow * nblk + (wave >> 2)) * FIS_LINE_WORDS +
                 (wave & 3)] = cell[1];
      }
    }
    // The record lists, stored with the words: lane k of the first wave of
    // each neuron of the tile lists the trials of its part that sent at
    // step k (the words of the part's waves, all in LDS by now).  Waves of
    // ended words sent nothing, so a list holds live trials only.
    if (lists && tl < kWave && pl < np && lane < W &&
        (((part_sent >> lane) & 1u) || lane == W - 1)) {
      uint4 l = make_uint4(0u, 0u, 0u, 0u);
      uint32_t cnt = 0;
#pragma unroll
      for (int q = 0; q < wpt; ++q) {
        uint64_t x =
            *(const uint64_t*)(acc + (pl * tpn + q * kWave) * WMAX + 4 * lane);
        while (x) {
          list_push(l, (uint32_t)(q * kWave + __ffsll((long long)x) - 1));
          x &= x - 1ull;
          ++cnt;
        }
      }
      if (cnt <= kListMax)
        list_push(l, cnt);
      else
        l = make_uint4(FIS_REC_LIST_WORDS, 0u, 0u, 0u);
      const int slot =
          wbase + lane < kRec ? wbase + lane : wbase + lane - kRec;
      *(uint4*)(s.rec_ids +
                (((int64_t)pre * parts + part) * kRec + slot) * 16) = l;
    }

    // 3. act bits of the window's steps: clear, then set
    if (tid < np) {
      uint32_t bits32 = 0;
      for (int q = 0; q < wpt; ++q)
        bits32 |= (uint32_t)acc[(tid * tpn + q * kWave) * WMAX + 4 * WMAX];
      const uint64_t nw = (T.act_old & ~rot40(wmask, wbase)) |
                          rot40((uint64_t)bits32, wbase);
      if (nw != T.act_old)
        s.act[(int64_t)neuron_at(p0 + tid) * parts + part] = nw;
      if (GATE && s.calm) {  // every wave of the neuron's part ended calm
        uint32_t now_calm = 1u;
        for (int q = 0; q < wpt; ++q)
          now_calm &=
              (uint32_t)acc[(tid * tpn + q * kWave) * WMAX + 4 * WMAX + 1];
        if (now_calm != ((T.calm_old >> tid) & 1u))
          s.calm[(int64_t)neuron_at(p0 + tid) * parts + part] =
              (uint8_t)now_calm;
      }
    }

    // 4. input gating: a neuron of the tile that sent anything (in this
    // part) stamps its targets with the last step its spikes can arrive
    // at: an atomic max that returns nothing (stamps only grow), so that
    // a hub's thousands of targets cost a few rounds of target loads, each
    // followed by stores nothing waits for.
    if (GATE) {
      const uint64_t stamp = (uint64_t)(t0 + W - 1 + kDelay);
      constexpr int U = 4;  // targets per thread loaded at once
      for (int q = 0; q < np; ++q) {
        uint32_t sent = 0;
#pragma unroll
        for (int j = 0; j < wpt; ++j)
          sent |= (uint32_t)acc[(q * tpn + j * kWave) * WMAX + 4 * WMAX];
        if (!sent) continue;  // uniform: every thread reads the same cells
        const int src = neuron_at(p0 + q);
        const int64_t b1 = s.offsets[src + 1];
        for (int64_t k = s.offsets[src] + tid; k < b1; k += U * TILE) {
          int32_t c[U];
#pragma unroll
          for (int u = 0; u < U; ++u)
            c[u] = k + u * TILE < b1 ? s.cols[k + u * TILE] : -1;
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (c[u] >= 0)
              __hip_atomic_fetch_max(
                  (unsigned long long*)(s.due + (int64_t)c[u] * parts + part),
                  (unsigned long long)stamp, __ATOMIC_RELAXED,
                  __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
  };

  Tile T = head(tile);
  load(T);
  first_edge(T);
  run(T);
}

// Scalar registers capped at 80: measured on MI355X, a workgroup of this
// kernel is resident one fewer per CU than the compiler's occupancy
// estimate whenever the kernel uses more than ~80 (96 SGPRs: 7 of 8
// workgroups, 1.465 ms per window; capped at 80: 8, 1.417 ms).
template <int WMAX, int BS, int TILE, bool GATE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_num_sgpr(80)))
__attribute__((amdgpu_waves_per_eu(8, 8)))
void k_pull(fis_sim s, int W, int ring_steps, int n_tiles, int per_xcd) {
  // acc[trial][step]: a lane's WMAX cells are contiguous (cleared and read
  // back with 16-byte LDS accesses); nothing else lives in LDS
  static_assert(WMAX % 4 == 0, "acc rows are cleared 4 cells at a time");
  __shared__ __align__(16) int32_t acc[TILE * WMAX];
  if constexpr (!GATE) {
    // XCD-aware tile order: workgroup b is dispatched to XCD b mod 8, and
    // XCD x takes the consecutive tiles [x * per_xcd, (x + 1) * per_xcd) of
    // the neuron order, so one XCD's L2 holds the spike records its
    // neurons' presynaptic partners share.  (A workgroup that took two
    // tiles in turn, loading the second's state and first in-edges during
    // the first, measured slower: 0.80 against 0.75 ms per window at 256
    // trials, 64 VGPRs and spills.)
    const int tile = (int)(blockIdx.x & 7u) * per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= n_tiles) return;  // whole workgroup, before any barrier
    pull_tile<WMAX, BS, TILE, GATE>(s, acc, tile, W, ring_steps,
                                    (int)threadIdx.x);
  } else {
    // the window's work list (k_pull_plan), in 8 runs of consecutive
    // entries: run x is walked by the workgroups of XCD x, each taking
    // every (grid / 8)-th entry
    // (both counters and the parity read at once: one round trip)
    const int4 head4 = *(const int4*)s.work;
    const int par = head4.w;
    const int cnt = min(par ? head4.y : head4.x, n_tiles);
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the next window's counter
      s.work[2] = 1 - par;
      s.work[1 - par] = 0;
    }
    const int nk = (int)(gridDim.x >> 3);
    const int per = (cnt + 7) >> 3;
    const int lo = (int)(blockIdx.x & 7u) * per;
    const int hi = min(cnt, lo + per);
    int i = lo + (int)(blockIdx.x >> 3);
    int next = i < hi ? s.work[4 + i] : 0;
    for (; i < hi; i += nk) {
      const int tile = next;
      // the entry after this one, read while this tile runs
      next = i + nk < hi ? s.work[4 + i + nk] : 0;
      // Each tile starts from values the compiler cannot see through: the
      // arguments read again from the kernel-argument segment (scalar
      // loads), the thread number and the window's sizes.  Nothing of one
      // tile's window is then hoisted out of the loop and held in
      // registers across it (it would spill).
      uint64_t ka = (uint64_t)__builtin_amdgcn_kernarg_segment_ptr();
      int w = W, rs = ring_steps, tid = (int)threadIdx.x;
      asm volatile("" : "+s"(ka), "+s"(w), "+s"(rs), "+v"(tid));
      const fis_sim& sk = *(const fis_sim*)kernarg_ptr(ka);
      pull_tile<WMAX, BS, TILE, GATE>(sk, acc, tile, w, rs, tid);
      __syncthreads();  // the next tile clears the accumulators
    }
  }
}

// clock += W after a window (one thread; every workgroup of the window has
// read the clock by then), or after a deferred run of windows
__global__ void k_tick(int64_t* clock, int W) { *clock += W; }

// Re-mask the last step's record of each listed neuron-trial with its
// current transmit flag (one thread per entry; atomics, since entries of
// one record word may come from different workgroups).
__global__ __launch_bounds__(kBlock) void k_pull_resend(fis_sim s,
                                                        const int32_t* idx,
                                                        int n_idx) {
  const int e = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (e >= n_idx) return;
  const int src = idx[e];
  if (src < 0 || src >= s.n * s.bstride) return;
  if (!(s.st[src] & FIS_ST_SPIKE)) return;
  const int pre = src / s.bstride;
  const int b = src - pre * s.bstride;
  const int words = s.bstride >> 6;
  const int slot = (int)((now(s) + kRec - 1) % kRec);
  uint64_t* w = s.rec + ((int64_t)pre * kRec + slot) * words + (b >> 6);
  const uint64_t bit = 1ull << (b & 63);
  const int tile = pull_tile_of(s);
  const int parts = pull_parts(s.bstride, tile);
  const int part = b / pull_trials_per_part(s.bstride, tile);
  // the part's list of that step no longer matches its words: readers are
  // sent to the words (every writer stores the same byte)
  if (s.rec_ids)
    s.rec_ids[(((int64_t)pre * parts + part) * kRec + slot) * 16] =
        (uint8_t)FIS_REC_LIST_WORDS;
  if (s.fl[src] & FIS_OUT_OFF) {
    __hip_atomic_fetch_and(w, ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_fetch_or(w, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(s.act + (int64_t)pre * parts + part, 1ull << slot,
                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s.due) {  // the spike of step now - 1 now reaches its targets
      const int64_t stamp = now(s) - 1 + kDelay;
      for (int64_t k = s.offsets[pre]; k < s.offsets[pre + 1]; ++k)
        __hip_atomic_fetch_max((unsigned long long*)(s.due +
                                   (int64_t)s.cols[k] * parts + part),
                               (unsigned long long)stamp, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// The input in flight in the records, added into a push-road ring: one
// wave per (neuron, step u) of the last 20 steps (blockIdx.y = u - clock
// + 20), rows by presynaptic neuron.
__global__ __launch_bounds__(kBlock) void k_pull_to_ring(fis_sim s,
                                                         int32_t* ring) {
  const int lane = threadIdx.x & (kWave - 1);
  const int pre = (int)((blockIdx.x * kBlock + threadIdx.x) >> 6);
  if (pre >= s.n) return;
  const int64_t u = now(s) - 20 + (int64_t)blockIdx.y;
  if (u < 0) return;
  const int slot = (int)(u % kRec);
  const int words = s.bstride >> 6;
  const int64_t nb = (int64_t)s.n * s.bstride;
  int32_t* base = ring + (int64_t)((u + kDelay + 1) % FIS_RING_LEN) * nb;
  const int parts = pull_parts(s.bstride, pull_tile_of(s));
  const int wpp = words / parts;  // words per part
  for (int w = 0; w < words; ++w) {
    if (!((s.act[(int64_t)pre * parts + w / wpp] >> slot) & 1ull)) continue;
    const uint64_t m = s.rec[((int64_t)pre * kRec + slot) * words + w];
    if (m == 0) continue;
    for (int64_t k = s.offsets[pre] + lane; k < s.offsets[pre + 1];
         k += kWave) {
      int32_t* p = base + (int64_t)s.cols[k] * s.bstride + (w << 6);
      const int32_t val = s.vals[k];
      uint64_t mm = m;
      while (mm) {
        const int j = __ffsll((long long)mm) - 1;
        mm &= mm - 1ull;
        __hip_atomic_fetch_add(p + j, val, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// ------------------------------------------------------------- slots

// One thread per neuron; it resets every listed trial of its neuron, so a
// workgroup writes only the rows of its own 256 neurons (whole lines).
__global__ __launch_bounds__(kBlock) void k_reset_trials(fis_sim s,
                                                         const int32_t* trials,
                                                         int n_trials,
                                                         float v0) {
  const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (i >= s.n) return;
  const int bs = s.bstride;
  const int64_t nb = (int64_t)s.n * bs;
  const int words = bs >> 6;
  for (int q = 0; q < n_trials; ++q) {
    const int b = trials[q];
    if (b < 0 || b >= s.batch) continue;
    const int idx = i * bs + b;
    unsettle(s, idx);
    s.v[idx] = v0;
    s.g[idx] = 0.0f;
    s.st[idx] = (uint8_t)min(s.refrac_steps, (int)FIS_ST_RMAX);
    s.fl[idx] = 0;
    s.sbits[idx] = 0u;
    if (s.vhold) s.vhold[idx] = 0.0f;
    if (s.counts) s.counts[idx] = 0u;
    if (s.spikes_out) s.spikes_out[idx] = 0.0f;
    for (int k = 0; k < FIS_RING_LEN; ++k) s.ring[k * nb + idx] = 0;
    if (s.rec && words > 0) {
      uint64_t* r = s.rec + (int64_t)i * kRec * words + (b >> 6);
      const uint64_t keep = ~(1ull << (b & 63));
      for (int k = 0; k < kRec; ++k) r[k * words] &= keep;
      if (s.rec_ids) {  // the lists of the trial's part: read the words
        const int tile = pull_tile_of(s);
        const int parts = pull_parts(bs, tile);
        const int part = b / pull_trials_per_part(bs, tile);
        uint8_t* li = s.rec_ids + ((int64_t)i * parts + part) * kRec * 16;
        for (int k = 0; k < kRec; ++k) li[k * 16] = (uint8_t)FIS_REC_LIST_WORDS;
      }
    }
  }
}

// Byte-wise flag set through the containing 32-bit word (entries of one
// word may come from different workgroups).
__global__ __launch_bounds__(kBlock) void k_flag_pairs(fis_sim s,
                                                       const int32_t* idx,
                                                       int n, uint32_t bits) {
  const int e = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (e >= n) return;
  const int x = idx[e];
  if (x < 0 || x >= s.n * s.bstride) return;
  const int sh = (x & 3) * 8;
  unsettle(s, x);
  atomicOr((uint32_t*)(s.fl + (x & ~3)), (bits & 0xFFu) << sh);
  if (bits & FIS_NOREFRAC)
    atomicAnd((uint32_t*)(s.st + (x & ~3)), ~(0xFFu << sh));
}

// ------------------------------------------------------------- ev_spmv twin

// out[b, post] += sum_pre W[post, pre] * sp[b, pre] * mo[b, pre].  Each lane
// tests one presynaptic neuron; the wave then walks every hot row together.
template <typename I>
__global__ __launch_bounds__(kBlock) void k_ev_spmv(
    const I* __restrict__ crow, const I* __restrict__ col,
    const float* __restrict__ val, const float* __restrict__ sp,
    const float* __restrict__ mo, float* __restrict__ out, int n) {
  const int64_t off = (int64_t)blockIdx.y * n;
  const int64_t pre = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  float x = 0.0f;
  if (pre < n) x = sp[off + pre] * mo[off + pre];
  uint64_t bal = __ballot(x != 0.0f);
  const int64_t base_pre = pre - lane;
  while (bal) {
    const int j = __ffsll((long long)bal) - 1;
    bal &= bal - 1ull;
    const int64_t p = base_pre + j;
    const float sj = __shfl(x, j);
    const I a = crow[p], e = crow[p + 1];
    for (I k = a + (I)lane; k < e; k += (I)kWave)
      unsafeAtomicAdd(out + off + (int64_t)col[k], val[k] * sj);
  }
}

bool valid_stride(int32_t bs) {
  if (bs <= 0) return false;
  if (bs < kWave) return (bs & (bs - 1)) == 0;
  return bs % kWave == 0;
}

// events of a full window plus one sink slot per wave segment, rounded up
// to whole 128-byte lines
int min_ev_cap(int32_t bs) {
  const int need = FIS_MAX_WINDOW * (bs >= kWave ? 1 : kWave / bs) + 1;
  return (need + FIS_EVENT_ALIGN - 1) / FIS_EVENT_ALIGN * FIS_EVENT_ALIGN;
}

// Everything check_sim checks except the event staging (events, ev_count,
// ev_cap), which fis_window_noalloc brings itself.
int check_sim_state(const fis_sim* s) {
  if (!s || s->n <= 0 || s->batch <= 0 || !valid_stride(s->bstride) ||
      s->batch > s->bstride || s->bstride > 65535 + 1)
    return (int)hipErrorInvalidValue;
  if (!s->offsets || !s->cols || !s->vals || !s->v || !s->g || !s->st ||
      !s->fl || !s->ring || !s->sbits || !s->clock || !s->seeds ||
      !s->trial_ids)
    return (int)hipErrorInvalidValue;
  if (s->raster && s->raster_len < FIS_MAX_WINDOW)
    return (int)hipErrorInvalidValue;
  // the counter saturates at FIS_ST_RMAX = 127 between windows
  if (s->refrac_steps < 0 || s->refrac_steps >= FIS_ST_RMAX)
    return (int)hipErrorInvalidValue;
  if (s->drive_group && (!s->drive_rate || s->drive_n_groups <= 0 ||
                         s->drive_n_seg <= 0 || s->drive_seg_steps <= 0))
    return (int)hipErrorInvalidValue;
  if (s->drive_group && s->drive_dense) return (int)hipErrorInvalidValue;
  if ((int64_t)s->n * s->bstride >= ((int64_t)1 << 31) - kBlock)
    return (int)hipErrorInvalidValue;
  return 0;
}

int check_sim(const fis_sim* s) {
  if (int rc = check_sim_state(s)) return rc;
  if (s->ev_cap < min_ev_cap(s->bstride) || s->ev_cap % FIS_EVENT_ALIGN != 0)
    return (int)hipErrorInvalidValue;
  if (!s->events || !s->ev_count) return (int)hipErrorInvalidValue;
  return 0;
}

int check_pull(const fis_sim* s) {
  if (int rc = check_sim(s)) return rc;
  const int bs = s->bstride;
  if (bs < kWave || bs > 1024 || (bs & (bs - 1)) != 0)
    return (int)hipErrorInvalidValue;
  if (!s->in_offsets || !s->in_cols || !s->in_vals || !s->rec || !s->act)
    return (int)hipErrorInvalidValue;
  if (s->due && (!s->work || ((uintptr_t)s->work & 15u) != 0))
    return (int)hipErrorInvalidValue;
  if (s->pull_tile != 0 && s->pull_tile != 128 && s->pull_tile != 256 &&
      s->pull_tile != 512)
    return (int)hipErrorInvalidValue;
  // the pull road draws a drive table in its own launch: the window kernel
  // is bound by vector-instruction issue, and Philox on the few driven
  // lanes of a wave would cost it more than a separate, fully occupied
  // drive launch
  if (s->drive_group || s->drive_dense) return (int)hipErrorInvalidValue;
  return 0;
}

unsigned grid_for(int64_t items) {
  return (unsigned)((items + kBlock - 1) / kBlock);
}

// W update steps of the push road, spikes into the event segments of
// s.events / s.ev_count (checked by the caller)
void launch_update(const fis_sim& s, int W, hipStream_t stream) {
  const int64_t nb = (int64_t)s.n * s.bstride;
  if (W == 1)
    hipLaunchKernelGGL((k_update<1, false>), dim3(grid_for(nb)), dim3(kBlock),
                       0, stream, s, W);
  else
    hipLaunchKernelGGL((k_update<FIS_MAX_WINDOW, false>), dim3(grid_for(nb)),
                       dim3(kBlock), 0, stream, s, W);
}

// the window's events into the int32 ring, then clock += W (unless
// defer_tick)
void launch_scatter(const fis_sim& s, int W, hipStream_t stream) {
  const int64_t nw = ((int64_t)s.n * s.bstride + kWave - 1) / kWave;
  int64_t blocks = (nw + (kBlock / kWave) - 1) / (kBlock / kWave);
  if (s.bstride >= kWave) {
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_scatter_steps, dim3((unsigned)blocks, (unsigned)W),
                       dim3(kBlock), 0, stream, s, W);
  } else {
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)blocks), dim3(kBlock), 0,
                       stream, s, W);
  }
}

// ---- the static scratch of fis_window_noalloc
//
// Device address of `scratch` on each device, looked up once per device
// with that device current (every GPU holds its own copy of the array at
// its own address; the host-side name of the array is not a device
// address and is never passed to a kernel).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices];

int scratch_address(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch[dev] = (unsigned char*)p;
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" {

int fis_abi_version(void) { return FIS_ABI_VERSION; }

int ev_spmv_abi(void) { return 3; }

int64_t fis_struct_layout(int what) {
  switch (what) {
    case 0: return (int64_t)sizeof(fis_sim);
    case 1: return (int64_t)offsetof(fis_sim, offsets);
    case 2: return (int64_t)offsetof(fis_sim, w_scale);
    case 3: return (int64_t)offsetof(fis_sim, refrac_steps);
    case 4: return (int64_t)offsetof(fis_sim, raster_len);
    case 5: return (int64_t)sizeof(fis_event);
    case 6: return (int64_t)sizeof(fis_drive_entry);
    case 7: return (int64_t)offsetof(fis_sim, in_offsets);
    case 8: return (int64_t)offsetof(fis_sim, t_origin);
    case 9: return (int64_t)offsetof(fis_sim, pull_tile);
    case 10: return (int64_t)offsetof(fis_sim, clock_ahead);
    case 11: return (int64_t)offsetof(fis_sim, drive_dense);
    case 12: return (int64_t)offsetof(fis_sim, due);
    case 13: return (int64_t)offsetof(fis_sim, calm);
    case 14: return (int64_t)offsetof(fis_sim, rec_ids);
    case 15: return (int64_t)offsetof(fis_sim, work);
    default: return -1;
  }
}

int64_t fis_n_blocks(int32_t n, int32_t bstride) {
  const int64_t nb = (int64_t)n * bstride;
  return (nb + kBlock - 1) / kBlock;
}

int64_t fis_n_waves(int32_t n, int32_t bstride) {
  const int64_t nb = (int64_t)n * bstride;
  return (nb + kWave - 1) / kWave;
}

int ev_spmv_i32(const int* crow, const int* col, const float* val,
                const float* sp, const float* mo, float* out, int n, int batch,
                int threads, void* stream) {
  (void)threads;
  if (n <= 0 || batch <= 0) return (int)hipErrorInvalidValue;
  dim3 grid(grid_for(n), (unsigned)batch);
  hipLaunchKernelGGL(k_ev_spmv<int>, grid, dim3(kBlock), 0,
                     (hipStream_t)stream, crow, col, val, sp, mo, out, n);
  return (int)hipGetLastError();
}

int ev_spmv_i64(const long long* crow, const long long* col,
                const float* val, const float* sp, const float* mo, float* out,
                int n, int batch, int threads, void* stream) {
  (void)threads;
  if (n <= 0 || batch <= 0) return (int)hipErrorInvalidValue;
  dim3 grid(grid_for(n), (unsigned)batch);
  hipLaunchKernelGGL(k_ev_spmv<long long>, grid, dim3(kBlock), 0,
                     (hipStream_t)stream, crow, col, val, sp, mo, out, n);
  return (int)hipGetLastError();
}

int fis_drive_rates(const fis_sim* s, const float* rates, int W,
                    void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (!rates || W < 1 || W > FIS_MAX_WINDOW) return (int)hipErrorInvalidValue;
  const int64_t nb = (int64_t)s->n * s->bstride;
  hipLaunchKernelGGL(k_drive_rates, dim3(grid_for(nb)), dim3(kBlock), 0,
                     (hipStream_t)stream, *s, rates, W);
  return (int)hipGetLastError();
}

int fis_drive_table(const fis_sim* s, const fis_drive_entry* entries,
                    int n_entries, const float* rate_hz, int n_groups,
                    int n_seg, int seg_steps, int W, void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (W < 1 || W > FIS_MAX_WINDOW || n_entries < 0 || n_groups <= 0 ||
      n_seg <= 0 || seg_steps <= 0 || !rate_hz ||
      (n_entries > 0 && !entries))
    return (int)hipErrorInvalidValue;
  if (n_entries == 0) return 0;
  hipLaunchKernelGGL(k_drive_table, dim3(grid_for(n_entries)), dim3(kBlock), 0,
                     (hipStream_t)stream, *s, entries, n_entries, rate_hz,
                     n_groups, n_seg, seg_steps, W);
  return (int)hipGetLastError();
}

int fis_drive_tape(const fis_sim* s, const int64_t* tape_ptr,
                   const int32_t* tape_idx, int64_t n_tape_steps, int W,
                   void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (W < 1 || W > FIS_MAX_WINDOW || !tape_ptr || n_tape_steps < 0)
    return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_drive_tape, dim3(8, (unsigned)W), dim3(kBlock), 0,
                     (hipStream_t)stream, *s, tape_ptr, tape_idx,
                     n_tape_steps);
  return (int)hipGetLastError();
}

int fis_update(const fis_sim* s, int W, void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (W < 1 || W > FIS_MAX_WINDOW) return (int)hipErrorInvalidValue;
  launch_update(*s, W, (hipStream_t)stream);
  return (int)hipGetLastError();
}
int fis_tick(const fis_sim* s, int steps, void* stream) {
  if (!s || !s->clock || steps < 0) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, (hipStream_t)stream,
                     s->clock, steps);
  return (int)hipGetLastError();
}

int fis_update_scatter(const fis_sim* s, int W, void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (W < 1 || W > FIS_MAX_WINDOW || s->bstride >= kWave)
    return (int)hipErrorInvalidValue;
  const int64_t nb = (int64_t)s->n * s->bstride;
  if (W == 1)
    hipLaunchKernelGGL((k_update<1, true>), dim3(grid_for(nb)), dim3(kBlock),
                       0, (hipStream_t)stream, *s, W);
  else
    hipLaunchKernelGGL((k_update<FIS_MAX_WINDOW, true>), dim3(grid_for(nb)),
                       dim3(kBlock), 0, (hipStream_t)stream, *s, W);
  if (!s->defer_tick)
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, (hipStream_t)stream,
                       s->clock, W);
  return (int)hipGetLastError();
}

int fis_scatter(const fis_sim* s, int W, void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (W < 1 || W > FIS_MAX_WINDOW) return (int)hipErrorInvalidValue;
  launch_scatter(*s, W, (hipStream_t)stream);
  return (int)hipGetLastError();
}

int fis_window_noalloc(const fis_sim* s, int W, void* stream) {
  if (int rc = check_sim_state(s)) return rc;
  if (W < 1 || W > FIS_MAX_WINDOW) return (int)hipErrorInvalidValue;
  // the staging this window needs: the update's event segments, then one
  // 128-byte line of event counts per update workgroup
  fis_sim t = *s;
  t.ev_cap = min_ev_cap(s->bstride);
  const int64_t ev_bytes =
      (fis_n_waves(s->n, s->bstride) * t.ev_cap * (int64_t)sizeof(fis_event) +
       255) / 256 * 256;
  const int64_t cnt_bytes =
      fis_n_blocks(s->n, s->bstride) * FIS_LINE_INTS * (int64_t)sizeof(int32_t);
  if (ev_bytes + cnt_bytes > (int64_t)FIS_SCRATCH_BYTES)
    return (int)hipErrorInvalidValue;  // larger than the static scratch
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  t.events = (fis_event*)base;
  t.ev_count = (int32_t*)(base + ev_bytes);
  launch_update(t, W, (hipStream_t)stream);
  launch_scatter(t, W, (hipStream_t)stream);
  return (int)hipGetLastError();
}

int fis_seed_ring(const fis_sim* s, int slot, void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (slot < 0 || slot >= FIS_RING_LEN) return (int)hipErrorInvalidValue;
  const int64_t nb = (int64_t)s->n * s->bstride;
  hipLaunchKernelGGL(k_seed_ring, dim3(grid_for(nb)), dim3(kBlock), 0,
                     (hipStream_t)stream, *s, slot);
  return (int)hipGetLastError();
}

int fis_resend(const fis_sim* s, const int32_t* idx, int n_idx, int slot,
               void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (slot < 0 || slot >= FIS_RING_LEN || n_idx < 0 || (n_idx > 0 && !idx))
    return (int)hipErrorInvalidValue;
  if (n_idx == 0) return 0;
  const int64_t threads = (int64_t)n_idx * kWave;
  hipLaunchKernelGGL(k_resend, dim3(grid_for(threads)), dim3(kBlock), 0,
                     (hipStream_t)stream, *s, idx, n_idx, slot);
  return (int)hipGetLastError();
}

int fis_pull_window(const fis_sim* s, int W, int ring_steps, void* stream) {
  if (int rc = check_pull(s)) return rc;
  if (W < 1 || W > FIS_MAX_WINDOW || ring_steps < 0)
    return (int)hipErrorInvalidValue;
  const int tile = pull_tile_of(*s);
  const bool gate = s->due != nullptr;
  auto kern = k_pull<FIS_MAX_WINDOW, 256, 256, false>;
  auto plan = k_pull_plan<256, 256>;
  switch (s->bstride * 4 + tile / 128) {  // tile / 128 is 1, 2 or 4
#define FIS_PULL_CASE(B, T)                                    \
  case B * 4 + T / 128:                                        \
    kern = gate ? k_pull<FIS_MAX_WINDOW, B, T, true>           \
                : k_pull<FIS_MAX_WINDOW, B, T, false>;         \
    plan = k_pull_plan<B, T>;                                  \
    break;
    FIS_PULL_CASE(64, 128) FIS_PULL_CASE(64, 256) FIS_PULL_CASE(64, 512)
    FIS_PULL_CASE(128, 128) FIS_PULL_CASE(128, 256) FIS_PULL_CASE(128, 512)
    FIS_PULL_CASE(256, 128) FIS_PULL_CASE(256, 256) FIS_PULL_CASE(256, 512)
    FIS_PULL_CASE(512, 128) FIS_PULL_CASE(512, 256) FIS_PULL_CASE(512, 512)
    FIS_PULL_CASE(1024, 128) FIS_PULL_CASE(1024, 256)
    FIS_PULL_CASE(1024, 512)
#undef FIS_PULL_CASE
    default: return (int)hipErrorInvalidValue;
  }
  // tiles: P neurons each, or `parts` tiles per neuron
  const int P = tile >= s->bstride ? tile / s->bstride : 1;
  const int parts = tile >= s->bstride ? 1 : s->bstride / tile;
  const int n_tiles = ((s->n + P - 1) / P) * parts;
  // whole neurons per XCD: the parts of one neuron write the same record
  // lines, and share one L2
  const int per_xcd = ((n_tiles + 7) / 8 + parts - 1) / parts * parts;
  if (gate) {
    // the work list first, then kPullListWaves waves' worth of
    // workgroups (a multiple of 8) walk it
    hipLaunchKernelGGL(plan, dim3(grid_for((n_tiles + kPlanTiles - 1) /
                                           kPlanTiles)),
                       dim3(kBlock), 0, (hipStream_t)stream, *s, W,
                       ring_steps, n_tiles);
    const int most = kPullListWaves * kWave / tile;
    const int up8 = (n_tiles + 7) / 8 * 8;
    const int grid = most < up8 ? most : up8;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(tile), 0,
                       (hipStream_t)stream, *s, W, ring_steps, n_tiles,
                       per_xcd);
  } else {
    hipLaunchKernelGGL(kern, dim3((unsigned)(8 * per_xcd)), dim3(tile), 0,
                       (hipStream_t)stream, *s, W, ring_steps, n_tiles,
                       per_xcd);
  }
  if (!s->defer_tick)
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, (hipStream_t)stream,
                       s->clock, W);
  return (int)hipGetLastError();
}

int fis_pull_resend(const fis_sim* s, const int32_t* idx, int n_idx,
                    void* stream) {
  if (int rc = check_pull(s)) return rc;
  if (n_idx < 0 || (n_idx > 0 && !idx)) return (int)hipErrorInvalidValue;
  if (n_idx == 0) return 0;
  hipLaunchKernelGGL(k_pull_resend, dim3(grid_for(n_idx)), dim3(kBlock), 0,
                     (hipStream_t)stream, *s, idx, n_idx);
  return (int)hipGetLastError();
}

int fis_pull_to_ring(const fis_sim* s, int32_t* ring, void* stream) {
  if (int rc = check_pull(s)) return rc;
  if (!ring) return (int)hipErrorInvalidValue;
  const int64_t threads = (int64_t)s->n * kWave;
  hipLaunchKernelGGL(k_pull_to_ring, dim3(grid_for(threads), 20), dim3(kBlock),
                     0, (hipStream_t)stream, *s, ring);
  return (int)hipGetLastError();
}

int fis_reset_trials(const fis_sim* s, const int32_t* trials, int n_trials,
                     float v0, void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (n_trials < 0 || (n_trials > 0 && !trials))
    return (int)hipErrorInvalidValue;
  if (n_trials == 0) return 0;
  hipLaunchKernelGGL(k_reset_trials, dim3(grid_for(s->n)), dim3(kBlock), 0,
                     (hipStream_t)stream, *s, trials, n_trials, v0);
  return (int)hipGetLastError();
}

int fis_flag_pairs(const fis_sim* s, const int32_t* idx, int n, uint32_t bits,
                   void* stream) {
  if (int rc = check_sim(s)) return rc;
  if (n < 0 || (n > 0 && !idx) || bits > 0xFFu)
    return (int)hipErrorInvalidValue;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_flag_pairs, dim3(grid_for(n)), dim3(kBlock), 0,
                     (hipStream_t)stream, *s, idx, n, bits);
  return (int)hipGetLastError();
}

}  // extern "C"
