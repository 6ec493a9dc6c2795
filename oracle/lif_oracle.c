/*
 * The oracle: a plain-C restatement of the reference LIF step.
 *
 * TEST INFRASTRUCTURE ONLY.  Tests, the smoke check and the CPU-baseline leg
 * of the benchmark load this file's library as the checker; the product path
 * never calls it.
 *
 * Every function restates one piece of lab/flybrain/sim.py and keeps its
 * float32 operation order exactly (compile with -ffp-contract=off and no
 * fast-math: each line below is one IEEE-754 rounding, as each torch op is
 * one kernel with one rounding):
 *
 *   oracle_step      -- Engine._core, lab/flybrain/sim.py:863-953, with the
 *                       event scatter of ev_spmv.cu:60-79 done in exact
 *                       int64 arithmetic (equal to the float32 csrmv under
 *                       the exactness precondition, evspmv.py:957-994)
 *   oracle_stim      -- Engine._poisson, lab/flybrain/sim.py:848-861, with
 *                       the Bernoulli draw taken from a tape or from this
 *                       library's Philox4x32-10 stream instead of torch's
 *                       generator
 *   oracle_philox    -- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11,
 *                       "Parallel random numbers: as easy as 1, 2, 3")
 *   oracle_run_counts -- oracle_step looped over a whole run of one trial
 *                       fed a sparse tape (spike counts only; for long
 *                       statistical comparisons)
 *
 * Arrays are trial-major [B][N] like the reference's tensors; the delay
 * ring is [B][L][N] like Engine.delay_buffer (sim.py:385-386).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ Philox */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void oracle_philox(const uint32_t ctr_in[4], const uint32_t key_in[2],
                   uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* One Bernoulli outcome of the engine's drive stream.
 * Counter (t >> 2, neuron, trial_id, 0), key = the trial's 64-bit seed,
 * word t & 3; uniform u = (x >> 8) * 2^-24, outcome u < p. */
int oracle_draw(uint64_t seed, uint32_t trial_id, uint32_t neuron, uint32_t t,
                float p) {
  uint32_t ctr[4] = {t >> 2, neuron, trial_id, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t x[4];
  oracle_philox(ctr, key, x);
  float u = (float)(x[t & 3u] >> 8) * (1.0f / 16777216.0f);
  return u < p;
}

/* ------------------------------------------------------------ drive */

/* stim[b][i] = w_scale * (bern * gain), bern from rates (Hz) with the
 * Philox stream: p = rates * prob_scale in float32 (sim.py:859-861). */
void oracle_stim(int n, int batch, const float* rates, const uint64_t* seeds,
                 const uint32_t* trial_ids, uint32_t t, float prob_scale,
                 float gain, float w_scale, float* stim) {
  for (int b = 0; b < batch; ++b) {
    for (int i = 0; i < n; ++i) {
      size_t k = (size_t)b * n + i;
      float p = rates[k] * prob_scale;
      float bern = 0.0f;
      if (p > 0.0f)
        bern = (float)oracle_draw(seeds[b], trial_ids[b], (uint32_t)i, t, p);
      float poisson = bern * gain;
      stim[k] = w_scale * poisson;
    }
  }
}

/* stim from a tape: bits[b][i] in {0,1} given as bytes. */
void oracle_stim_tape(int n, int batch, const uint8_t* bits, float gain,
                      float w_scale, float* stim) {
  for (size_t k = 0; k < (size_t)batch * n; ++k) {
    float poisson = (float)bits[k] * gain;
    stim[k] = w_scale * poisson;
  }
}

/* ------------------------------------------------------------ step */

typedef struct {
  float w_scale;  /* 0.275 mV per synapse           */
  float a_syn;    /* float32(1 - DT/tauSyn) = 0.98  */
  float a_mem;    /* float32(DT/tauMem)     = 0.005 */
  float v_rest;   /* -52 */
  float v_reset;  /* -52 */
  float v_th;     /* -45 */
} oracle_params;

/*
 * One step for all trials, in the reference's operation order.
 * refrac_steps is [N] like Engine.refrac_steps, or [B][N] when
 * rs_per_trial is set (trials with their own driven sets, each trial then
 * being an independent reference run with its own no_refractory).
 * Masks may be NULL (all ones / v_hold zero).  rec_out (may be NULL)
 * receives the synaptic input written into the ring this step, and
 * icount_out (may be NULL) the same as the exact integer sum.
 * acc is caller scratch of batch*n int64.
 * Returns 0, or -1 if an integer input sum reached 2^24 (the reference's
 * float32 sum would no longer be exact and parity is undefined).
 */
int oracle_step(int n, int batch, const int64_t* offsets, const int32_t* cols,
                const int16_t* vals, const oracle_params* prm, float* v,
                float* g, float* spikes, float* refrac, float* ring, int L,
                int ptr, const float* refrac_steps, int rs_per_trial,
                const float* m_fire,
                const float* m_out, const float* m_in, const float* m_hold,
                const float* v_hold, const float* stim, float* rec_out,
                int32_t* icount_out, int64_t* acc) {
  int rc = 0;
  memset(acc, 0, sizeof(int64_t) * (size_t)batch * n);
  /* recurrent = w_scale * (W @ (spikes * m_out)) -- sim.py:910-918 */
  for (int b = 0; b < batch; ++b) {
    const float* s = spikes + (size_t)b * n;
    const float* mo = m_out ? m_out + (size_t)b * n : 0;
    int64_t* a = acc + (size_t)b * n;
    for (int pre = 0; pre < n; ++pre) {
      float x = s[pre] * (mo ? mo[pre] : 1.0f);
      if (x == 0.0f) continue;
      /* x is exactly 1.0 here: spikes and masks are 0/1 */
      for (int64_t k = offsets[pre]; k < offsets[pre + 1]; ++k)
        a[cols[k]] += (int64_t)vals[k];
    }
  }
  for (int b = 0; b < batch; ++b) {
    for (int i = 0; i < n; ++i) {
      size_t k = (size_t)b * n + i;
      int64_t I = acc[k];
      if (I >= (1 << 24) || I <= -(1 << 24)) rc = -1;
      float recurrent = prm->w_scale * (float)I;
      if (rec_out) rec_out[k] = recurrent;
      if (icount_out) icount_out[k] = (int32_t)I;

      /* refractory counter: reset on spike, otherwise count up */
      float r = refrac[k] + 1.0f;
      float one_minus_s = 1.0f - spikes[k];
      r = r * one_minus_s;
      refrac[k] = r;
      float gate = (r >= refrac_steps[rs_per_trial ? k : (size_t)i]) ? 1.0f
                                                                     : 0.0f;

      /* delayed input, read-then-write of ring slot ptr */
      float* slot = ring + ((size_t)b * L + ptr) * n;
      float delayed = slot[i] * (m_in ? m_in[k] : 1.0f);
      float gd = g[k] * prm->a_syn;
      float dg = delayed * gate;
      float g_new = gd + dg;
      slot[i] = recurrent;

      /* membrane: uses the OLD conductance (sim.py:932-937) */
      float vv = v[k] + stim[k];
      float t1 = vv - prm->v_rest;
      float t2 = g[k] - t1;
      float t3 = prm->a_mem * t2;
      vv = vv + t3;
      /* hold before threshold (sim.py:942) */
      vv = vv * (m_hold ? m_hold[k] : 1.0f);
      vv = vv + (v_hold ? v_hold[k] : 0.0f);
      float sp = (vv > prm->v_th) ? 1.0f : 0.0f;
      sp = sp * (m_fire ? m_fire[k] : 1.0f);
      /* reset by the masked spike: v - (v - v_reset) * s (sim.py:950) */
      float t4 = vv - prm->v_reset;
      float t5 = t4 * sp;
      vv = vv - t5;
      v[k] = vv;
      spikes[k] = sp;
      float gs = g_new * sp;
      g[k] = g_new - gs;
    }
  }
  return rc;
}

/* ------------------------------------------------------------ helpers */

/* The 24-bit uniforms (x >> 8) the drive compares with p, for neurons
 * n0 .. n0+n-1 and steps t0 .. t0+steps-1 of one trial:
 * out[(t - t0) * n + (i - n0)].  For statistical checks of the stream. */
void oracle_uniform_grid(uint64_t seed, uint32_t trial_id, uint32_t n0,
                         int n, uint32_t t0, int steps, uint32_t* out) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int k = 0; k < steps; ++k) {
    uint32_t t = t0 + (uint32_t)k;
    for (int i = 0; i < n; ++i) {
      uint32_t ctr[4] = {t >> 2, n0 + (uint32_t)i, trial_id, 0u};
      uint32_t x[4];
      oracle_philox(ctr, key, x);
      out[(size_t)k * n + i] = x[t & 3u] >> 8;
    }
  }
}

/* One trial (B = 1) from rest (v = v0, refrac = refrac_steps, nothing in
 * flight) for `steps` steps, kicked by a sparse tape: at step t the
 * neurons tape_idx[tape_ptr[t] .. tape_ptr[t+1]) get the Poisson stim
 * w_scale * (1 * gain) (sim.py:859-861), the others w_scale * (0 * gain).
 * counts[i] (int64) receives the spikes of neuron i.  Returns 0, -1 if an
 * input sum reached 2^24, -2 if out of memory. */
int oracle_run_counts(int n, int steps, const int64_t* offsets,
                      const int32_t* cols, const int16_t* vals,
                      const oracle_params* prm, const float* refrac_steps,
                      float v0, const int64_t* tape_ptr,
                      const int32_t* tape_idx, float gain, int L,
                      int64_t* counts) {
  size_t nn = (size_t)n;
  float* v = malloc(sizeof(float) * nn);
  float* g = calloc(nn, sizeof(float));
  float* sp = calloc(nn, sizeof(float));
  float* rf = malloc(sizeof(float) * nn);
  float* ring = calloc(nn * (size_t)L, sizeof(float));
  float* stim = malloc(sizeof(float) * nn);
  int64_t* acc = malloc(sizeof(int64_t) * nn);
  int rc = 0;
  if (!v || !g || !sp || !rf || !ring || !stim || !acc) {
    rc = -2;
    goto done;
  }
  const float off = prm->w_scale * (0.0f * gain);
  const float on = prm->w_scale * (1.0f * gain);
  for (size_t i = 0; i < nn; ++i) {
    v[i] = v0;
    rf[i] = refrac_steps[i];
    stim[i] = off;
    counts[i] = 0;
  }
  for (int t = 0; t < steps; ++t) {
    for (int64_t e = tape_ptr[t]; e < tape_ptr[t + 1]; ++e)
      stim[tape_idx[e]] = on;
    if (oracle_step(n, 1, offsets, cols, vals, prm, v, g, sp, rf, ring, L,
                    t % L, refrac_steps, 0, 0, 0, 0, 0, 0, stim, 0, 0, acc))
      rc = -1;
    for (int64_t e = tape_ptr[t]; e < tape_ptr[t + 1]; ++e)
      stim[tape_idx[e]] = off;
    for (size_t i = 0; i < nn; ++i) counts[i] += sp[i] != 0.0f;
  }
done:
  free(v);
  free(g);
  free(sp);
  free(rf);
  free(ring);
  free(stim);
  free(acc);
  return rc;
}

int oracle_abi(void) { return 2; }
