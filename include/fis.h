/*
 * fis.h -- C ABI of the MI355X LIF engine library (libfis.so).
 *
 * Every entry point takes raw device pointers and sizes, takes the HIP
 * stream as `void*` (0 = the null stream), only enqueues work (no host
 * synchronisation, so every call can be captured into a hipGraph) and
 * returns a hipError_t as int (0 = success).  The library never allocates:
 * the caller owns every buffer.
 *
 * Two groups of entry points:
 *
 *  1. ev_spmv_i32 / ev_spmv_i64 / ev_spmv_abi -- the reference's event SpMV
 *     C ABI under its own names and signatures, lab/flybrain/ev_spmv.cu:
 *     114-137, which lab/flybrain/evspmv.py:886-903 binds by name through
 *     ctypes (so this library loads where the reference's DLL did):
 *         out[b, post] += sum_pre W[post, pre] * sp[b, pre] * mo[b, pre]
 *     over a CSR of W^T (row = presynaptic neuron), float atomics.
 *
 *  2. The fused engine (replaces Engine._core + Engine._poisson,
 *     lab/flybrain/sim.py:848-953, and the csrmv / event roads it calls):
 *     per window of W <= FIS_MAX_WINDOW steps
 *         fis_drive_*   Poisson kicks for the window into `sbits`
 *         fis_update    W steps of the membrane / conductance / refractory
 *                       update per neuron-trial, spike compaction into
 *                       per-wave event segments
 *         fis_scatter   event-driven scatter of the window's spikes into
 *                       the int32 delay ring (exact integer sums)
 */
#ifndef FIS_H
#define FIS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FIS_ABI_VERSION 15
#define FIS_MAX_WINDOW 20   /* steps per window: delay horizon 19 + 1     */
#define FIS_RING_LEN 40     /* int32 ring slots, two delays: the spikes of
                               step u are summed into slot (u + 21) mod 40,
                               which step u + 20 reads (slot (t+1) mod 40
                               for step t); a window's reads and adds never
                               share a slot                                */

/* Line rule.  Within one kernel, outputs written by different workgroups
 * never share a 128-byte line (on MI355X such plain stores from different
 * XCDs were measured to overwrite each other).  Hence the per-block layouts
 * below: a workgroup is FIS_BLOCK_WAVES waves, and per-wave records of one
 * block are packed into a line of their own. */
#define FIS_BLOCK_WAVES 4   /* 256 threads per workgroup                  */
#define FIS_LINE_WORDS 16   /* 128 bytes = 16 x 8-byte words              */
#define FIS_LINE_INTS 32    /* 128 bytes = 32 x 4-byte ints               */
#define FIS_EVENT_ALIGN 8   /* 8 events = 128 bytes                       */
#define FIS_REC_SLOTS 40    /* pull road: spike records of steps t mod 40 */
#define FIS_REC_LIST_MAX 15     /* trials a record list holds */
#define FIS_REC_LIST_WORDS 0xFF /* list count byte: read the record words */
#define FIS_PULL_TILE 256   /* pull road: neuron-trials per workgroup:
                               256 / bstride neurons, or a quarter (512:
                               half) of one neuron's trials               */

/* flag bits of `fl`, one byte per neuron-trial (sim.py:393-410, :866-900) */
#define FIS_FIRE_OFF 0x01u  /* m_fire == 0: spike output suppressed        */
#define FIS_OUT_OFF 0x02u   /* m_out  == 0: spike not transmitted          */
#define FIS_IN_OFF 0x04u    /* m_in   == 0: delayed input not received     */
#define FIS_HOLD 0x08u      /* m_hold == 0: v replaced by vhold            */
#define FIS_NOREFRAC 0x10u  /* refractory period 0 (driven neuron)         */

/* `st`, one byte per neuron-trial: bit 7 = spike of the last step, bits
 * 0..6 = steps since the last spike, saturating at 127 (the refractory
 * gate only compares it with refrac_steps <= 126). */
#define FIS_ST_SPIKE 0x80u
#define FIS_ST_RMAX 127

/* One spike event: neuron `pre` spiked at the step whose input lands in
 * ring slot `slot`, in the trials tb + j for every set bit j of `mask`. */
typedef struct {
  int32_t pre;
  uint16_t slot;
  uint16_t tb;
  uint64_t mask;
} fis_event;

/*
 * Everything the engine kernels touch.  Neuron-trial index
 * idx = neuron * bstride + trial; bstride is 1, 2, 4, ..., 64 or a multiple
 * of 64 (trials >= batch are padding with FIS_FIRE_OFF set).
 */
typedef struct {
  int32_t n;         /* neurons                                        */
  int32_t batch;     /* trials in use                                  */
  int32_t bstride;   /* trial stride (padded batch)                    */
  int32_t ev_cap;    /* event slots per update wave: a multiple of 8,
                        >= 20 * max(1, 64 / bstride) + 1 (the last slot
                        of each wave's segment is a sink)               */
  /* connectome by presynaptic neuron (W^T CSR) */
  const int64_t* offsets; /* [n+1]                                     */
  const int32_t* cols;    /* [E] postsynaptic neuron                   */
  const int16_t* vals;    /* [E] signed synapse count                  */
  /* state, [n * bstride] each */
  float* v;
  float* g;
  uint8_t* st;
  uint8_t* fl;            /* written only by fis_reset_trials and
                             fis_flag_pairs                            */
  float* vhold;           /* nullable; read where FIS_HOLD is set      */
  int32_t* ring;          /* [FIS_RING_LEN][n * bstride]               */
  uint32_t* sbits;        /* [n * bstride] kick bits of the window     */
  fis_event* events;      /* [n_waves * ev_cap]                        */
  int32_t* ev_count;      /* [n_blocks][32]: count of wave w at
                             [w / 4][w % 4]                            */
  int64_t* clock;         /* [1] first step of the next window (plus
                             clock_ahead, below)                       */
  /* per-trial Philox key and stream id, [bstride] */
  const uint64_t* seeds;
  const uint32_t* trial_ids;
  /* float32 constants, already rounded the way the reference's tensor
   * ops round Python scalars (sim.py:361-383) */
  float w_scale;     /* wScale                       0.275           */
  float a_syn;       /* 1 - DT/tauSyn                0.98            */
  float a_mem;       /* DT/tauMem                    0.005           */
  float v_rest;      /* vRest                        -52             */
  float v_reset;     /* vReset                       -52             */
  float v_th;        /* vThreshold                   -45             */
  float gain;        /* scalePoisson                 250             */
  float prob_scale;  /* DT/1000                      1e-4            */
  int32_t refrac_steps; /* round(tRefrac/DT)         22              */
  /* optional outputs (nullable) */
  float* spikes_out;    /* [n * bstride] float 0/1 of the window's last
                           step, same layout as v                      */
  uint32_t* counts;     /* [n * bstride] spike count per neuron-trial  */
  uint32_t* pop_counts; /* [n] spike count per neuron, all trials      */
  uint64_t* raster;     /* [raster_len][n_blocks][16]: spikes of wave w
                           at step t are bit j (lane j) of word
                           [t mod raster_len][w / 4][w % 4]             */
  int32_t raster_len;   /* >= FIS_MAX_WINDOW                           */
  /* pull road only (nullable otherwise): the connectome by postsynaptic
   * neuron (W CSR, the reference's own W[post, pre] orientation) and the
   * spike records of the last 40 steps.  rec word
   * [pre][t mod 40][w] bit j = trial 64 w + j of `pre` transmitted a spike
   * at step t; act[pre][k] bit (t mod 40) = some trial of `pre` among
   * trials k * pull_tile .. (k + 1) * pull_tile - 1 did.                  */
  const int32_t* in_offsets; /* [n+1], by position (see `order`)       */
  const int32_t* in_cols;    /* [E] presynaptic neuron                 */
  const int16_t* in_vals;    /* [E] signed synapse count               */
  uint64_t* rec;             /* [n][40][bstride / 64]                  */
  uint64_t* act;             /* [n][max(1, bstride / pull_tile)]       */
  /* the order in which workgroups take the neurons: position k holds
   * neuron order[k] (a permutation of 0..n-1 that moves aligned units of
   * max(1, 256 / bstride) neurons, e.g. sorted by neuropil group, so that
   * neurons sharing presynaptic partners are updated by the same XCD);
   * in_offsets / in_cols / in_vals list the in-edges by position.       */
  const int32_t* order;      /* [n]                                    */
  /* per-trial clock (nullable = 0 and never): trial b's own step is
   * clock - t_origin[b] -- the Philox counter and the drive table's rate
   * segment use it, so a trial's drive does not depend on when its slot
   * started -- and its spikes are counted only at global steps
   * < t_end[b].  [bstride] each. */
  const int64_t* t_origin;
  const int64_t* t_end;
  /* pull road (nullable = every word live): [bstride / 64] the largest
   * t_end of each 64-trial word.  A word whose word_end is <= the first
   * step of a window holds only finished or never-started trials and is
   * skipped by that window: no input gathered, no state updated, no
   * records written.  Its stale records can only reach trials of the same
   * word, which fis_reset_trials clears when a slot of it restarts. */
  const int64_t* word_end;
  /* fused drive table (nullable; push road only -- fis_pull_window refuses
   * it): drive_group[idx] >= 0 makes fis_update draw the window's Philox
   * kicks of neuron-trial idx itself, at rate
   * drive_rate[min(t / drive_seg_steps, drive_n_seg - 1)][drive_group[idx]]
   * with t the trial's own step -- the same bits fis_drive_table would
   * set, without its launch.  Kicks already in sbits are kept.  */
  const int16_t* drive_group;  /* [n * bstride], -1 = not table-driven */
  const float* drive_rate;     /* [drive_n_seg][drive_n_groups] Hz     */
  int32_t drive_n_groups;
  int32_t drive_n_seg;
  int32_t drive_seg_steps;
  /* pull road: neuron-trials per workgroup, 128, 256 or 512 (0: the
   * default FIS_PULL_TILE).  Fixed for a simulation's life: it sets the
   * layout of `act` ([n][max(1, bstride / pull_tile)] words).  Chosen by
   * measurement (flybrain_amd.autotune). */
  int32_t pull_tile;
  /* Deferred clock, for a run of windows captured into one graph.  Every
   * kernel takes *clock + clock_ahead as the current step; with
   * defer_tick nonzero the window entry points leave *clock alone, and
   * the caller advances it once with fis_tick after the run (a
   * one-thread launch per window cost ~3 us of a 12 us window at B = 1).
   * Both 0 outside such a run. */
  int32_t clock_ahead;
  int32_t defer_tick;
  /* Push road: dense rates [batch][n] in Hz, held over the window, whose
   * kicks the update launches draw themselves -- the bits fis_drive_rates
   * would set, without its launch (Engine.step's rates, sim.py:745-771).
   * Nullable; not together with drive_group; fis_pull_window refuses it. */
  const float* drive_dense;
  /* Pull road input gating (nullable = off): [n][max(1, bstride /
   * pull_tile)] int64, per neuron and tile part the latest step at which
   * input sent to it may arrive.  A window that sends a spike of a part
   * sets its targets' entries (W^T CSR) to the window's last step + 20; a
   * window whose tile's entries are all below its first step has no
   * synaptic input and skips the in-edge walk.  Zero at the start; fixed
   * for a simulation's life (on from its first window, or never). */
  int64_t* due;
  /* Pull road, with `due` (nullable = off): [n][max(1, bstride /
   * pull_tile)] bytes, nonzero where every neuron-trial of the (neuron,
   * tile part) sat at a fixed point of the step through the last window
   * (no input, no kick, no spike, g = +0, v unchanged by the leak step,
   * refractory counter saturated, no flag).  Such a part with no input due
   * and no kick in the next window skips it without reading its state:
   * the window would change nothing but its record and act words.  Set
   * and cleared by fis_pull_window; the drive entry points,
   * fis_reset_trials and fis_flag_pairs clear the entries they touch.  A
   * caller that writes v, g, st, fl or vhold itself zeroes it. */
  uint8_t* calm;
  /* Pull road, nullable: the spike records again as short lists,
   * [n][max(1, bstride / pull_tile)][FIS_REC_SLOTS][16] bytes, one per
   * (neuron, tile part, record slot): byte 0 the number of trials of the
   * part that sent at that step (0..15), or FIS_REC_LIST_WORDS (0xFF):
   * read the record words instead; bytes 1.. their trial numbers within
   * the part, in any order.  Valid exactly where the record words are.
   * Written by fis_pull_window with the words; fis_pull_resend and
   * fis_reset_trials mark what they change 0xFF.  Read in place of the
   * words by parts of 128 and 256 trials (16 bytes per active pair
   * instead of 8 per 64 trials, and one LDS add per listed trial).  Zero
   * (empty lists, matching zero words) at the start; a caller that writes
   * rec itself zeroes the lists with it or fills them with 0xFF. */
  uint8_t* rec_ids;
  /* Pull road, required with `due`: [4 + tiles] int32, zero at the start,
   * tiles = ceil(n / max(1, pull_tile / bstride)) * max(1, bstride /
   * pull_tile).  A gated window first lists the tiles it has to run (the
   * others -- calm, no input due -- are done by that first launch), then
   * walks the list: work[0..3] are the list's counters and parities (kept
   * by the library from window to window), work[4 ..] the list.  16-byte
   * aligned. */
  int32_t* work;
} fis_sim;

/* Drive table entry: neuron-trial `idx` is driven at rate
 * rate_hz[segment][group] with segment = min(t / seg_steps, n_seg - 1). */
typedef struct {
  uint32_t idx;
  uint32_t group;
} fis_drive_entry;

int fis_abi_version(void);

/* ---- the reference's event-SpMV ABI (lab/flybrain/ev_spmv.cu:114-137) --
 * Same names, argument lists and return convention (a hipError_t as int,
 * 0 = success; the reference returns cudaGetLastError()), bound by
 * lab/flybrain/evspmv.py:886-903 with argtypes [c_void_p] * 6 + [c_int] * 3
 * + [c_void_p].  The caller zeroes `out` first (evspmv.py:1131); the kernel
 * adds into it with float atomics (order-free under the exactness gate of
 * evspmv.py:957-994).  n <= 0 or batch <= 0 returns hipErrorInvalidValue
 * without a launch. */

/* replaces ev_spmv_i32 (ev_spmv.cu:114-121): int32 W^T CSR.  `threads` is
 * accepted for the signature; the launch shape is the library's own. */
int ev_spmv_i32(const int* crow, const int* col, const float* val,
                const float* sp, const float* mo, float* out, int n, int batch,
                int threads, void* stream);
/* replaces ev_spmv_i64 (ev_spmv.cu:123-130): int64 W^T CSR. */
int ev_spmv_i64(const long long* crow, const long long* col,
                const float* val, const float* sp, const float* mo, float* out,
                int n, int batch, int threads, void* stream);
/* replaces ev_spmv_abi (ev_spmv.cu:136-137): returns 3, the number
 * evspmv.py:888-894 checks before binding the two entry points above. */
int ev_spmv_abi(void);

/* ---- the fused engine ----------------------------------------------- */

/* Poisson kicks for a window of W steps from dense rates [batch][n] in Hz,
 * constant over the window (Engine._poisson, sim.py:848-861, fed by the
 * `rates` argument of Engine.step, sim.py:745-771). */
int fis_drive_rates(const fis_sim* s, const float* rates, int W, void* stream);

/* Poisson kicks from a drive table: entries [n_entries], rate table
 * rate_hz[n_seg][n_groups] (piecewise-constant schedule, seg_steps steps
 * per segment). */
int fis_drive_table(const fis_sim* s, const fis_drive_entry* entries,
                    int n_entries, const float* rate_hz, int n_groups,
                    int n_seg, int seg_steps, int W, void* stream);

/* Kicks from a recorded tape: for step t the neuron-trial indices
 * tape_idx[tape_ptr[t] .. tape_ptr[t+1]), t in [0, n_tape_steps). */
int fis_drive_tape(const fis_sim* s, const int64_t* tape_ptr,
                   const int32_t* tape_idx, int64_t n_tape_steps, int W,
                   void* stream);

/* W steps of the neuron update (Engine._core minus the SpMV,
 * sim.py:920-953) starting at step *clock, spike compaction into events. */
int fis_update(const fis_sim* s, int W, void* stream);

/* Scatter of the events of the window just updated (W steps) into the
 * int32 ring (the SpMV of sim.py:906-918 / ev_spmv.cu:60-79, integer,
 * exact), then clock += W (unless defer_tick). */
int fis_scatter(const fis_sim* s, int W, void* stream);

/* fis_update + fis_scatter without the event records, bstride < 64 only:
 * each wave adds its window's spikes straight into the ring slots that no
 * step of the window reads (Engine._core + the SpMV, sim.py:906-953); a
 * one-thread launch then advances the clock by W (unless defer_tick). */
int fis_update_scatter(const fis_sim* s, int W, void* stream);

/* Static scratch of fis_window_noalloc: 64 MiB per device. */
#define FIS_SCRATCH_BYTES (64u << 20)

/* fis_update + fis_scatter as one call whose event staging (the per-wave
 * event segments and counts that fis_update writes and fis_scatter reads)
 * lives in a static device array of FIS_SCRATCH_BYTES that the library
 * holds on each GPU it runs on, instead of s->events / s->ev_count (which
 * may be NULL here; s->ev_cap is ignored).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it calls no allocator.  Push road, any trial stride.  The
 * staging needs n_waves * ev_cap * 16 + n_blocks * 128 bytes (ev_cap =
 * 20 * max(1, 64 / bstride) + 1 rounded up to 8): at most about 10 M
 * neuron-trials from 64 trials up, 2.9 M at 16, 0.2 M at 1 (the
 * 138,639-neuron network fits at B = 1, 16 and 64); a larger simulation
 * returns hipErrorInvalidValue without launching.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race.  Same results as fis_update then fis_scatter. */
int fis_window_noalloc(const fis_sim* s, int W, void* stream);

/* *clock += steps (one thread): closes a run of windows launched with
 * defer_tick set.  No counterpart in the reference, whose step counter is
 * a host integer (sim.py:745-771). */
int fis_tick(const fis_sim* s, int steps, void* stream);

/* Scatter the current spike state (st bit 7, not FIS_OUT_OFF) into ring
 * slot `slot` -- used after a restore to rebuild the input in flight. */
int fis_seed_ring(const fis_sim* s, int slot, void* stream);

/* After FIS_OUT_OFF changed on the neuron-trials idx[0..n_idx) between two
 * steps: the reference masks the PREVIOUS step's spikes with the CURRENT
 * m_out (rec = W @ (spikes * m_out), sim.py:910-918, masks written between
 * steps by set_masks, sim.py:523-571), while this engine scattered them at
 * the end of that step with the old flag.  For each listed neuron-trial
 * whose last step spiked (st bit 7), its row is added into ring slot `slot`
 * (flag now clear) or subtracted (flag now set).  fl must already hold the
 * new flags; `slot` is the slot that step's spikes went to:
 * (clock - 1 + 21) mod FIS_RING_LEN. */
int fis_resend(const fis_sim* s, const int32_t* idx, int n_idx, int slot,
               void* stream);

/* ---- pull road: one fused launch per window (bstride a power of two,
 * 64 .. 1024).  Each workgroup owns pull_tile neuron-trials: pull_tile /
 * bstride neurons x all trials, or one neuron x pull_tile of its trials
 * when bstride > pull_tile.  It sums the input of its neurons for the W steps
 * t0 .. t0+W-1 from the spike records of steps t0-20 .. t0+W-21 over their
 * in-edges into LDS (exact int32), then runs the W update steps of
 * Engine._core (sim.py:920-953) from LDS and writes the window's own
 * spike records; finally clock += W (unless defer_tick).  The int32 ring
 * is not used, except
 * that the first `ring_steps` steps of the window also consume (and clear)
 * ring slot (t+1) mod 40, which holds input carried over from a restore. */
int fis_pull_window(const fis_sim* s, int W, int ring_steps, void* stream);

/* Pull-road counterpart of fis_resend: after FIS_OUT_OFF changed on the
 * listed neuron-trials, re-mask the record of the last step (clock - 1)
 * with the new flags (sim.py:910-918 masks the previous step's spikes). */
int fis_pull_resend(const fis_sim* s, const int32_t* idx, int n_idx,
                    void* stream);

/* Add the input still in flight in the pull road's records (the spikes of
 * steps clock-20 .. clock-1) into `ring` ([FIS_RING_LEN][n * bstride],
 * the push road's layout: spikes of step u land in slot (u+21) mod 40),
 * using the W^T CSR -- for state views and snapshots. */
int fis_pull_to_ring(const fis_sim* s, int32_t* ring, void* stream);

/* ---- slots (continuous batching) --------------------------------------
 * Reset trials trials[0..n_trials) to the state Engine.__init__ gives
 * (sim.py:371-410): v = v0, g = 0, refractory counter = refrac_steps, no
 * spike, all flags clear, no kicks, counts 0, and nothing in flight for
 * them (int32 ring column, and on the pull road their record bits).  One
 * thread per neuron walks the listed trials, so every workgroup writes
 * whole lines of its own neurons. */
int fis_reset_trials(const fis_sim* s, const int32_t* trials, int n_trials,
                     float v0, void* stream);

/* OR flag bits into fl[idx[k]] for k < n (byte atomics on the containing
 * word); with FIS_NOREFRAC the refractory counter of those neuron-trials is
 * set to 0 as well (Engine.__init__: refrac = refrac_steps = 0 for driven
 * neurons, sim.py:371-388).  fl and st must be allocated to a multiple of
 * 4 bytes. */
int fis_flag_pairs(const fis_sim* s, const int32_t* idx, int n,
                   uint32_t bits, void* stream);

/* sizeof / offsetof of the structs above (0: sizeof fis_sim, 1: offsets,
 * 2: w_scale, 3: refrac_steps, 4: raster_len, 5: sizeof fis_event,
 * 6: sizeof fis_drive_entry, 7: in_offsets, 8: t_origin, 9: pull_tile,
 * 10: clock_ahead, 11: drive_dense, 12: due, 13: calm, 14: rec_ids,
 * 15: work), so a binding can check its mirror. */
int64_t fis_struct_layout(int what);

/* Number of update waves (event segments) for a given n and bstride. */
int64_t fis_n_waves(int32_t n, int32_t bstride);
/* Number of update workgroups (FIS_BLOCK_WAVES waves each). */
int64_t fis_n_blocks(int32_t n, int32_t bstride);

#ifdef __cplusplus
}
#endif

#endif /* FIS_H */
