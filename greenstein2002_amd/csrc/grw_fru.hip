// grw_fru.hip -- gfx950 kernels and the C ABI (include/grw_fru.h) of the
// release-unit Monte Carlo.
//
// Layout in HBM (ld = nfru rounded up to 64):
//   conc  double [5][ld]   CaJSR, CaSS1..4                      (SoA, coalesced)
//   chan  uint32 [4][ld]   packed LCC/gate/5 RyR/Ito2 per cleft  (grw_model.h)
//   mtidx int32  [ld]      stream position, see MtLane
//   nlog  int32  [ld]      words saved in the undo log since the last checkpoint
//   mt    uint32 [n][640]  one MT19937 row per FRU (624 words used)
//   mtlog uint32 [n][640]  undo log: first pre-checkpoint value of each rewritten word
//   *_h                    checkpoint of conc / chan / mtidx
//
// One lane advances one FRU through the whole sub-interval (simfru_local.c:127-277)
// with every continuous and discrete state in registers; each lane draws from
// its own stream, so lanes never wait on each other except at the flux
// reduction, which is fused into the tail of the same launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/grw_fru.h"
#include "grw_model.h"

#pragma clang fp contract(off)

using namespace grw;

namespace {

// ---------------------------------------------------------------------------
// MT19937 stream of one lane.
//
// The reference regenerates all 624 words when its index reaches 624
// (mt19937_local.c:121-141).  A lane here instead regenerates each word at the
// moment it is drawn: word p of the next generation depends on words p and p+1
// of the current one and on word p+397 (current, p < 227) or p-227 (already
// regenerated, p >= 227), and for p = 623 on the regenerated word 0.  Computing
// the words in draw order therefore yields the reference's sequence exactly,
// with a fixed cost per draw instead of a 624-word burst that would stall the
// other 63 lanes of the wavefront.
//
// idx in [0, 624): the row holds one full generation and the next draw reads
// word idx as is (state freshly loaded from the reference's format).
// idx in [624, 1248): lazy regime, p = idx - 624; words [0, p) belong to the
// new generation, [p, 624) to the previous one.
//
// Rollback: the first overwrite of each word after a checkpoint stores the old
// value in the undo log (nlog counts them, at most 624), so a checkpoint costs
// nothing for the MT rows and a rollback restores only the rewritten words.
struct MtLane {
  uint32_t* row;
  uint32_t* log;
  int idx;
  int nlog;

  __device__ __forceinline__ static uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  __device__ __forceinline__ uint32_t next() {
    if (idx < kMtN) return temper(row[idx++]);
    const int p = idx - kMtN;
    const uint32_t cur = row[p];
    const uint32_t nxt = row[p < kMtN - 1 ? p + 1 : 0];
    const uint32_t far = row[p < kMtN - 397 ? p + 397 : p - (kMtN - 397)];
    const uint32_t y = (cur & 0x80000000u) | (nxt & 0x7fffffffu);
    const uint32_t w = far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    if (nlog < kMtN) {
      log[p] = cur;
      ++nlog;
    }
    row[p] = w;
    idx = (p == kMtN - 1) ? kMtN : idx + 1;
    return temper(w);
  }
  // MersenneTwisterOne_local: y / 2^32 in [0, 1)
  __device__ __forceinline__ double uniform() { return ((double)next()) / 4294967296.0; }

  // The 32 tempered words of one micro-step (MersenneTwister_fast(..., 32, ...),
  // mt19937_local.c:97-154).  In the lazy regime a word depends only on words at
  // least 227 positions older (or, for position 623, on word 0 of the same
  // generation, 623 older), never on another word of the same 32-word window, so
  // all 65 loads are issued before any store and the memory latency is paid
  // once per step instead of once per draw.
  __device__ __forceinline__ void block32(uint32_t out[32]) {
    if (idx >= kMtN) {
      const int p = idx - kMtN;
      uint32_t cur[33], far[32];
#pragma unroll
      for (int k = 0; k < 33; ++k) {
        int q = p + k;
        q -= (q >= kMtN) ? kMtN : 0;
        cur[k] = row[q];
      }
#pragma unroll
      for (int k = 0; k < 32; ++k) {
        int q = p + k;
        q -= (q >= kMtN) ? kMtN : 0;
        far[k] = row[q < kMtN - 397 ? q + 397 : q - (kMtN - 397)];
      }
#pragma unroll
      for (int k = 0; k < 32; ++k) {
        const uint32_t y = (cur[k] & 0x80000000u) | (cur[k + 1] & 0x7fffffffu);
        far[k] = far[k] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        out[k] = temper(far[k]);
      }
#pragma unroll
      for (int k = 0; k < 32; ++k) {
        int q = p + k;
        q -= (q >= kMtN) ? kMtN : 0;
        if (nlog < kMtN) {
          log[q] = cur[k];
          ++nlog;
        }
        row[q] = far[k];
      }
      idx += 32;
      idx -= (idx >= 2 * kMtN) ? kMtN : 0;
    } else if (idx + 32 <= kMtN) {
#pragma unroll
      for (int k = 0; k < 32; ++k) out[k] = temper(row[idx + k]);
      idx += 32;
    } else {  // crossing from a loaded full generation into the lazy regime
#pragma unroll 1
      for (int k = 0; k < 32; ++k) {
        const uint32_t v = next();
#pragma unroll
        for (int j = 0; j < 31; ++j) out[j] = out[j + 1];
        out[31] = v;
      }
    }
  }
};

struct SimArgs {
  double t0, t1;
  double dep0[3], depf[3];
  double fdep[3];   // (V, Cai, CaNSR) used for the flux conversion
  double scale;
  double k23c, gcf; // scaled k23cf, gammacf
  int nfru, ld;
};

struct Dev {
  double* conc;
  uint32_t* chan;
  int32_t* mtidx;
  int32_t* nlog;
  uint32_t* mt;
  uint32_t* mtlog;
  double* conc_h;
  uint32_t* chan_h;
  int32_t* mtidx_h;
  double* part;              // [grid][16] block partial sums
  unsigned int* ticket;      // last-block election counter (self-resetting)
  double* out;               // [32]: 0-3 fluxes, 4-8 flux sums, 16-26 avg sums
  unsigned long long* nsteps;
  int* err;
};

// Categorical transition draw, simfru_local.c:217-227 / :234-245: stay with
// probability 1 - dt*k_tot, else walk the cumulative exit probabilities with '>'.
template <class R>
__device__ __forceinline__ int pick(double u, double dt, const Exits<R>& e) {
  double acc = 1.0 - dt * (double)e.tot;
  if (!(u > acc)) return 0;
  int j = 0;
  acc += dt * (double)e.r[0];
  while (u > acc && j + 1 < e.n) {
    ++j;
    acc += dt * (double)e.r[j];
  }
  return e.d[j];
}

// ---------------------------------------------------------------------------
// Deterministic reduction of NV doubles per lane: xor-butterfly inside the
// wavefront, waves in order through LDS, block partials in fixed slots, and the
// last block to finish (ticket) sums the slots in a fixed order.  Returns true
// in thread 0 of the last block, with the totals in tot[].
template <int NV>
__device__ bool reduce_grid(double v[NV], const Dev& d, int slot_off, double tot[NV]) {
  __shared__ double wsum[4][NV];
  __shared__ bool last;
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) wsum[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; ++k) {
      double s = wsum[0][k];
      for (int w = 1; w < nw; ++w) s += wsum[w][k];
      __hip_atomic_store(&d.part[(size_t)blockIdx.x * 16 + slot_off + k], s, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    const unsigned t = atomicAdd(d.ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  // every thread of the last block sums a strided subset of the slots
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      s[k] += __hip_atomic_load(&d.part[(size_t)b * 16 + slot_off + k], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += __shfl_xor(s[k], off);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) wsum[wave][k] = s[k];
  __syncthreads();
  if (threadIdx.x != 0) return false;
  for (int k = 0; k < NV; ++k) {
    double a = wsum[0][k];
    for (int w = 1; w < nw; ++w) a += wsum[w][k];
    tot[k] = a;
  }
  *d.ticket = 0u;  // ready for the next launch
  return true;
}

// Flux contributions of one FRU (calc_fru_flux_local.c:45-70).
__device__ __forceinline__ void flux_terms(const double conc[5], const uint32_t ch[4], double v[5]) {
  v[0] = ((conc[1] + conc[2]) + conc[3]) + conc[4];
  v[1] = conc[0];
  v[2] = 0.0;
  v[3] = 0.0;
  v[4] = 0.0;
  for (int c = 0; c < 4; ++c) {
    if (ch_lcc_open(ch[c])) {
      v[2] += conc[c + 1];
      v[3] += 1.0;
    }
    if (ch_ito2(ch[c]) == 2) v[4] += 1.0;
  }
}

// ---------------------------------------------------------------------------
// simfru_local.c:32-278 for every FRU over one RK sub-interval, then (REDUCE)
// the flux sums and their conversion (mpi_master.c:401-468).
template <class R, bool SIM, bool REDUCE>
__global__ void __launch_bounds__(256) k_simfru(const SimArgs a, const Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  unsigned steps = 0;
  if (i < a.nfru) {
    double conc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) conc[k] = d.conc[k * a.ld + i];
    uint32_t ch[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) ch[c] = d.chan[c * a.ld + i];
    if (SIM) {
      MtLane m{d.mt + (size_t)i * kMtStride, d.mtlog + (size_t)i * kMtStride, d.mtidx[i], d.nlog[i]};
      const double span0 = a.depf[0] - a.dep0[0];
      const double span1 = a.depf[1] - a.dep0[1];
      const double span2 = a.depf[2] - a.dep0[2];
      const double tint = a.t1 - a.t0;
      double dV = a.dep0[0], dCai = a.dep0[1], dNSR = a.dep0[2];
      bool lopen[4];
      int nopen[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        lopen[c] = ch_lcc_open(ch[c]);
        nopen[c] = ch_ryr_open(ch[c]);
      }
      const R k23c = (R)a.k23c, gcf = (R)a.gcf;
      double t = a.t0;
      bool bad = false;
      while (t < a.t1) {
        // ---- transition rates at the step start (fru_rates_local.c:126-777)
        const R V = (R)dV;
        const R alpha = R(alphacf) * xexp(R(0.012) * (V - R(35.0)));
        const R beta = R(betacf) * xexp(R(-0.05) * (V - R(35.0)));
        const R yinf = R(yCa_frac) / (R(1.0) + xexp((V + R(12.5)) / R(5.0))) + (R(1.0) - R(yCa_frac));
        const R tau = R(60.0) + R(340.0) / (R(1.0) + xexp((V + R(30.0)) / R(12.0)));
        const R vr_open = (R(1.0) - yinf) / tau;
        const R vr_closed = yinf / tau;
        R cold[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) cold[c] = (R)conc[c + 1];
        // Clefts are processed in order 0..3 by always working on slot 0 and
        // rotating the per-cleft registers, so the loop stays rolled (small code,
        // static register indices).
        R maxr = R(0.0);
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
          const R cc = cold[0];
          uint32_t w = ch[0];
          Exits<R> e;
          if (!lcc_exits(ch_lcc(w), alpha, beta, gcf * cc, e)) bad = true;
          maxr = rmax(maxr, e.tot);
          maxr = rmax(maxr, ch_y(w) == 2 ? vr_open : vr_closed);
          maxr = rmax(maxr, ch_ito2(w) == 1 ? R(kfIto2) * cc : R(kbIto2));
          const RyrK<R> k = ryr_k(cc, k23c);
#pragma unroll 1
          for (int r = 0; r < kNryr; ++r) {
            int st = ryr_collapse(ch_ryr(w, r), k.regime);
            if (ryr_needs_draw(st, k.regime)) st = ryr_split(st, m.uniform(), k);
            w = ch_set_ryr(w, r, st);
            Exits<R> er;
            ryr_exits(st, k, er);
            maxr = rmax(maxr, er.tot);
          }
          ch[0] = ch[1]; ch[1] = ch[2]; ch[2] = ch[3]; ch[3] = w;
          cold[0] = cold[1]; cold[1] = cold[2]; cold[2] = cold[3]; cold[3] = cc;
        }
        if (bad) break;
        // ---- step length (simfru_local.c:135-146)
        double dt = 0.1 / (double)maxr;
        dt = rmin(dt, 10.001e-3);
        double cmax = 0.e-3;
#pragma unroll
        for (int c = 0; c < 4; ++c) cmax = rmax(cmax, conc[c + 1]);
        if (cmax > 1.5e-3) dt = rmin(dt, 5.001e-3);
        dt = rmin(dt, a.t1 - t);
        // the step's 32 uniforms (simfru_local.c:211): loads issued here, their
        // latency hidden behind the Heun arithmetic below
        uint32_t draws[32];
        m.block32(draws);
        // ---- Heun step of the local Ca (simfru_local.c:149-173)
        R cr[5], dd[5];
        double k1[5], y1[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) cr[k] = (R)conc[k];
        fcn_fru<R>(cr, (R)dV, (R)dCai, (R)dNSR, lopen, nopen, dd);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          k1[k] = dt * (double)dd[k];
          y1[k] = conc[k] + k1[k];
        }
        t = t + dt;
        dV = a.dep0[0] + (t - a.t0) / tint * span0;
        dCai = a.dep0[1] + (t - a.t0) / tint * span1;
        dNSR = a.dep0[2] + (t - a.t0) / tint * span2;
#pragma unroll
        for (int k = 0; k < 5; ++k) cr[k] = (R)y1[k];
        fcn_fru<R>(cr, (R)dV, (R)dCai, (R)dNSR, lopen, nopen, dd);
#pragma unroll
        for (int k = 0; k < 5; ++k) conc[k] = conc[k] + (k1[k] + dt * (double)dd[k]) / 2.0;
        // ---- transitions, 8 uniforms per cleft (simfru_local.c:215-276); same
        // rotation as above, the cleft's uniforms are always draws[0..7]
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
          const R cc = cold[0];
          uint32_t w = ch[0];
          Exits<R> e;
          double u = ((double)draws[0]) / 4294967296.0;
          lcc_exits(ch_lcc(w), alpha, beta, gcf * cc, e);
          int nd = pick<R>(u, dt, e);
          if (nd) w = ch_set_lcc(w, nd);
          u = ((double)draws[1]) / 4294967296.0;
          if (u >= 1.0 - dt * (double)(ch_y(w) == 2 ? vr_open : vr_closed)) w = ch_flip_y(w);
          const RyrK<R> k = ryr_k(cc, k23c);
          bool moved = false;
#pragma unroll
          for (int r = 0; r < kNryr; ++r) {
            u = ((double)draws[2 + r]) / 4294967296.0;
            ryr_exits(ch_ryr(w, r), k, e);
            nd = pick<R>(u, dt, e);
            if (nd) {
              w = ch_set_ryr(w, r, nd);
              moved = true;
            }
          }
          u = ((double)draws[7]) / 4294967296.0;
          if (u >= 1.0 - dt * (double)(ch_ito2(w) == 1 ? R(kfIto2) * cc : R(kbIto2)))
            w = ch_flip_ito2(w);
          const bool lo = ch_lcc_open(w);
          const int no = moved ? ch_ryr_open(w) : nopen[0];
          ch[0] = ch[1]; ch[1] = ch[2]; ch[2] = ch[3]; ch[3] = w;
          cold[0] = cold[1]; cold[1] = cold[2]; cold[2] = cold[3]; cold[3] = cc;
          lopen[0] = lopen[1]; lopen[1] = lopen[2]; lopen[2] = lopen[3]; lopen[3] = lo;
          nopen[0] = nopen[1]; nopen[1] = nopen[2]; nopen[2] = nopen[3]; nopen[3] = no;
#pragma unroll
          for (int j = 0; j < 24; ++j) draws[j] = draws[j + 8];
        }
        ++steps;
      }
      if (bad) atomicOr(d.err, 1);
#pragma unroll
      for (int k = 0; k < 5; ++k) d.conc[k * a.ld + i] = conc[k];
#pragma unroll
      for (int c = 0; c < 4; ++c) d.chan[c * a.ld + i] = ch[c];
      d.mtidx[i] = m.idx;
      d.nlog[i] = m.nlog;
    }
    if (REDUCE) flux_terms(conc, ch, acc);
  }
  if (SIM) {
    // FRU-step count for the throughput figures
    unsigned long long s = steps;
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(d.nsteps, s);
  }
  if (REDUCE) {
    double tot[5];
    if (reduce_grid<5>(acc, d, 0, tot)) {
      double o4[4];
      flux_currents(tot, a.fdep[0], a.fdep[1], a.fdep[2], a.nfru, a.scale, o4);
      for (int k = 0; k < 4; ++k) d.out[k] = o4[k];
      for (int k = 0; k < 5; ++k) d.out[4 + k] = tot[k];
    }
  }
}

// Average-output sums of calc_fru_avg_local.c:135-176 (tree order).
__global__ void __launch_bounds__(256) k_avg(const SimArgs a, const Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[11];
#pragma unroll
  for (int k = 0; k < 11; ++k) v[k] = 0.0;
  if (i < a.nfru) {
    const double cj = d.conc[i];
    v[1] = cj;
    v[6] = cj * (1.0 + CSQNtot / (KmCSQN + cj));
    for (int c = 0; c < 4; ++c) {
      const double s = d.conc[(c + 1) * a.ld + i];
      const uint32_t w = d.chan[c * a.ld + i];
      v[0] += s;
      v[5] += s * (1.0 + BSRtot / (KBSR + s) + BSLtot / (KBSL + s));
      int nopen = 0, nready = 0;
      for (int r = 0; r < kNryr; ++r) {
        const int st = ch_ryr(w, r);
        nopen += (st == 3 || st == 4 || st == 7);
        nready += (st == 1 || st == 2);
      }
      v[2] += JRyRmax * ((double)nopen) * (cj - s);
      v[3] += (double)nopen;
      v[4] += (double)nready;
      v[7] += (ch_lcc(w) <= 6) ? 1.0 : 0.0;
      v[8] += (ch_y(w) == 2) ? 1.0 : 0.0;
      v[9] += ch_lcc_open(w) ? 1.0 : 0.0;
      v[10] += (ch_ito2(w) == 2) ? 1.0 : 0.0;
    }
  }
  double tot[11];
  if (reduce_grid<11>(v, d, 5, tot))
    for (int k = 0; k < 11; ++k) d.out[16 + k] = tot[k];
}

// Reference-order sums (calc_fru_flux_local.c:45-70 and calc_fru_avg_local.c:135-176)
// in one thread, for bit-parity runs; also applies the flux conversion.
__global__ void k_seq_sums(const SimArgs a, const Dev d) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double ica = 0.0, oca = 0.0;
  int nito2 = 0;
  for (int i = 0; i < a.nfru; ++i)
    for (int c = 0; c < 4; ++c) {
      const uint32_t w = d.chan[c * a.ld + i];
      if (ch_lcc_open(w)) {
        ica = ica + d.conc[(c + 1) * a.ld + i];
        oca = oca + 1;
      }
      if (ch_ito2(w) == 2) nito2 = nito2 + 1;
    }
  double sss = 0.0, sjsr = 0.0;
  for (int i = 0; i < a.nfru; ++i) {
    sjsr += d.conc[i];
    for (int c = 0; c < 4; ++c) sss += d.conc[(c + 1) * a.ld + i];
  }
  double tot[5] = {sss, sjsr, ica, oca, (double)nito2};
  double o4[4];
  flux_currents(tot, a.fdep[0], a.fdep[1], a.fdep[2], a.nfru, a.scale, o4);
  for (int k = 0; k < 4; ++k) d.out[k] = o4[k];
  for (int k = 0; k < 5; ++k) d.out[4 + k] = tot[k];
  // average sums
  double CaSS = 0.0, CaJSR = 0.0, totSS = 0.0, totJSR = 0.0, JRyR = 0.0, PRyR = 0.0;
  double PNorm = 0.0, PnotV = 0.0, PL = 0.0, PIto2 = 0.0;
  int nready = 0;
  for (int i = 0; i < a.nfru; ++i) {
    const double cj = d.conc[i];
    CaJSR = CaJSR + cj;
    totJSR = totJSR + cj * (1.0 + CSQNtot / (KmCSQN + cj));
    for (int c = 0; c < 4; ++c) {
      const double s = d.conc[(c + 1) * a.ld + i];
      const uint32_t w = d.chan[c * a.ld + i];
      CaSS = CaSS + s;
      totSS = totSS + s * (1.0 + BSRtot / (KBSR + s) + BSLtot / (KBSL + s));
      int nopen = 0;
      for (int r = 0; r < kNryr; ++r) {
        const int st = ch_ryr(w, r);
        if (st == 3 || st == 4 || st == 7) nopen = nopen + 1;
        if (st == 1 || st == 2) nready = nready + 1;
      }
      JRyR = JRyR + JRyRmax * ((double)nopen) * (cj - s);
      PRyR = PRyR + ((double)nopen);
      if (ch_lcc(w) <= 6) PNorm = PNorm + 1.0;
      if (ch_y(w) == 2) PnotV = PnotV + 1.0;
      if (ch_lcc_open(w)) PL = PL + 1.0;
      if (ch_ito2(w) == 2) PIto2 = PIto2 + 1.0;
    }
  }
  const double avg[11] = {CaSS, CaJSR, JRyR, PRyR, (double)nready, totSS, totJSR, PNorm, PnotV, PL, PIto2};
  for (int k = 0; k < 11; ++k) d.out[16 + k] = avg[k];
}

// Checkpoint (mpi_slave.c:192-226): the MT rows are covered by the undo log.
__global__ void k_save(const SimArgs a, const Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nfru) return;
  for (int k = 0; k < 5; ++k) d.conc_h[k * a.ld + i] = d.conc[k * a.ld + i];
  for (int c = 0; c < 4; ++c) d.chan_h[c * a.ld + i] = d.chan[c * a.ld + i];
  d.mtidx_h[i] = d.mtidx[i];
  d.nlog[i] = 0;
}

// Rollback (mpi_slave.c:150-184): restore the words rewritten since the
// checkpoint, which are consecutive from the checkpoint's lazy position.
__global__ void k_resume(const SimArgs a, const Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nfru) return;
  for (int k = 0; k < 5; ++k) d.conc[k * a.ld + i] = d.conc_h[k * a.ld + i];
  for (int c = 0; c < 4; ++c) d.chan[c * a.ld + i] = d.chan_h[c * a.ld + i];
  const int idx0 = d.mtidx_h[i];
  const int p0 = idx0 >= kMtN ? idx0 - kMtN : 0;
  const int n = min(d.nlog[i], kMtN);
  uint32_t* row = d.mt + (size_t)i * kMtStride;
  const uint32_t* log = d.mtlog + (size_t)i * kMtStride;
  for (int k = 0; k < n; ++k) {
    int q = p0 + k;
    if (q >= kMtN) q -= kMtN;
    row[q] = log[q];
  }
  d.mtidx[i] = idx0;
  d.nlog[i] = 0;
}

// initialize_ran.c:23-36 + sgrnd_local (mt19937_local.c:75-92) per FRU.
__global__ void k_seed(const SimArgs a, const Dev d, uint32_t a0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nfru) return;
  const long long j = i;
  const long long v = (long long)(int)a0 + 23 * (j * (j + 1) / 2) + 4093 * j * j + 1;
  uint32_t x = (uint32_t)(unsigned long long)v;
  uint32_t* row = d.mt + (size_t)i * kMtStride;
  row[0] = x;
  for (uint32_t k = 1; k < kMtN; ++k) {
    x = 1812433253u * (x ^ (x >> 30)) + k;
    row[k] = x;
  }
  d.mtidx[i] = kMtN;  // the first draw regenerates word 0
  d.nlog[i] = 0;
}

// ---------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(GRW_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

struct grw_ctx {
  grw_config cfg;
  int n = 0, ld = 0, block = 64, grid = 0;
  Dev d{};
  hipStream_t stream = nullptr;
  double* host_out = nullptr;  // pinned [32]
  grw_stats stats{};
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void*> allocs;
};

namespace {

SimArgs base_args(const grw_ctx* c) {
  SimArgs a{};
  a.scale = c->cfg.nfru_scale;
  a.k23c = c->cfg.k23_scale != 1.0 ? k23cf * c->cfg.k23_scale : k23cf;
  a.gcf = c->cfg.gamma_scale != 1.0 ? gammacf * c->cfg.gamma_scale : gammacf;
  a.nfru = c->n;
  a.ld = c->ld;
  return a;
}

template <class T>
int dalloc(grw_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
  HIPCHK(hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T)));
  c->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

int check_err(grw_ctx* c) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, c->d.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (e) return fail(GRW_ESTATE, "kernel met an unknown channel state");
  return 0;
}

template <class R, bool SIM, bool REDUCE>
int launch_sim(grw_ctx* c, const SimArgs& a) {
  if (c->timing) HIPCHK(hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL((k_simfru<R, SIM, REDUCE>), dim3(c->grid), dim3(c->block), 0, c->stream, a, c->d);
  HIPCHK(hipGetLastError());
  if (c->timing) {
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.kernel_ms += ms;
  }
  return 0;
}

int run_sim(grw_ctx* c, const SimArgs& a, bool reduce) {
  int rc;
  const bool mixed = c->cfg.precision == GRW_MIXED;
  const bool tree = reduce && !c->cfg.sequential_sums;
  if (mixed)
    rc = tree ? launch_sim<float, true, true>(c, a) : launch_sim<float, true, false>(c, a);
  else
    rc = tree ? launch_sim<double, true, true>(c, a) : launch_sim<double, true, false>(c, a);
  if (rc) return rc;
  if (reduce && c->cfg.sequential_sums) {
    hipLaunchKernelGGL(k_seq_sums, dim3(1), dim3(1), 0, c->stream, a, c->d);
    HIPCHK(hipGetLastError());
  }
  c->stats.launches++;
  c->stats.simfru_calls += (uint64_t)c->n;
  return 0;
}

// Sub-interval bounds the kernel can finish: finite, ordered, at most 1 s long.
int check_interval(double t0, double t1, const double* dep0, const double* depf) {
  if (!std::isfinite(t0) || !std::isfinite(t1) || t1 < t0 || t1 - t0 > 1000.0)
    return fail(GRW_EINVAL, "sub-interval must be finite with 0 <= t1 - t0 <= 1000 ms");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(dep0[k]) || !std::isfinite(depf[k]))
      return fail(GRW_EINVAL, "non-finite global state (V, Cai, CaNSR)");
  return 0;
}

int fetch_out(grw_ctx* c, int off, int count, double* dst) {
  HIPCHK(hipMemcpyAsync(c->host_out, c->d.out, 32 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < count; ++k) dst[k] = c->host_out[off + k];
  return 0;
}

int flux_now(grw_ctx* c, const double dep[3]) {
  SimArgs a = base_args(c);
  for (int k = 0; k < 3; ++k) a.fdep[k] = dep ? dep[k] : 0.0;
  if (c->cfg.sequential_sums) {
    hipLaunchKernelGGL(k_seq_sums, dim3(1), dim3(1), 0, c->stream, a, c->d);
    HIPCHK(hipGetLastError());
    return 0;
  }
  return launch_sim<double, false, true>(c, a);
}

}  // namespace

// ===========================================================================
extern "C" {

const char* grw_last_error(void) { return g_err.c_str(); }

void grw_config_default(grw_config* cfg) {
  std::memset(cfg, 0, sizeof *cfg);
  cfg->nfru = 12500;
  cfg->precision = GRW_FP64;
  cfg->nfru_scale = 1.0;
  cfg->block = 0;
  cfg->sequential_sums = 0;
  cfg->k23_scale = 1.0;
  cfg->gamma_scale = 1.0;
  cfg->device = 0;
}

int grw_create(const grw_config* cfg, grw_ctx** out) {
  if (!cfg || !out) return fail(GRW_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->nfru <= 0 || cfg->nfru > (1 << 26)) return fail(GRW_EINVAL, "nfru out of range");
  if (cfg->precision != GRW_FP64 && cfg->precision != GRW_MIXED)
    return fail(GRW_EINVAL, "precision must be GRW_FP64 or GRW_MIXED");
  const int blk = cfg->block ? cfg->block : 64;
  if (blk % 64 || blk < 64 || blk > 256) return fail(GRW_EINVAL, "block must be 64, 128, 192 or 256");
  HIPCHK(hipSetDevice(cfg->device));
  grw_ctx* c = new grw_ctx();
  c->cfg = *cfg;
  c->n = cfg->nfru;
  c->ld = (cfg->nfru + 63) / 64 * 64;
  c->block = blk;
  c->grid = (c->n + blk - 1) / blk;
  int rc = 0;
  const size_t ld = c->ld, n = c->n;
  if ((rc = dalloc(c, &c->d.conc, 5 * ld)) || (rc = dalloc(c, &c->d.chan, 4 * ld)) ||
      (rc = dalloc(c, &c->d.mtidx, ld)) || (rc = dalloc(c, &c->d.nlog, ld)) ||
      (rc = dalloc(c, &c->d.mt, n * kMtStride)) || (rc = dalloc(c, &c->d.mtlog, n * kMtStride)) ||
      (rc = dalloc(c, &c->d.conc_h, 5 * ld)) || (rc = dalloc(c, &c->d.chan_h, 4 * ld)) ||
      (rc = dalloc(c, &c->d.mtidx_h, ld)) || (rc = dalloc(c, &c->d.part, (size_t)c->grid * 16)) ||
      (rc = dalloc(c, &c->d.ticket, 1)) || (rc = dalloc(c, &c->d.out, 32)) ||
      (rc = dalloc(c, &c->d.nsteps, 1)) || (rc = dalloc(c, &c->d.err, 1))) {
    grw_destroy(c);
    return rc;
  }
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_out, 32 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    grw_destroy(c);
    return fail(GRW_EDEVICE, std::string("stream/event setup failed: ") + hipGetErrorString(e));
  }
  *out = c;
  return 0;
}

int grw_destroy(grw_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->host_out) (void)hipHostFree(c->host_out);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int grw_load_state(grw_ctx* c, const double* fru, const int32_t* ltype, const int32_t* ryr,
                   const int32_t* ito2, const int32_t* mti, const uint32_t* mt) {
  if (!c || !fru || !ltype || !ryr || !ito2 || !mti || !mt) return fail(GRW_EINVAL, "null argument");
  const int n = c->n, ld = c->ld;
  std::vector<double> conc(5 * (size_t)ld, 0.0);
  std::vector<uint32_t> chan(4 * (size_t)ld, 0u);
  std::vector<int32_t> idx(ld, kMtN);
  std::vector<uint32_t> rows((size_t)n * kMtStride, 0u);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 5; ++k) conc[(size_t)k * ld + i] = fru[5 * (size_t)i + k];
    for (int cl = 0; cl < 4; ++cl) {
      const int lcc = ltype[8 * (size_t)i + 2 * cl], y = ltype[8 * (size_t)i + 2 * cl + 1];
      const int32_t* rr = ryr + 20 * (size_t)i + 5 * cl;
      const int it = ito2[4 * (size_t)i + cl];
      if (lcc < 1 || lcc > 12 || (y != 1 && y != 2) || (it != 1 && it != 2))
        return fail(GRW_EINVAL, "invalid LCC / gate / Ito2 state at FRU " + std::to_string(i));
      for (int r = 0; r < 5; ++r)
        if (rr[r] < 1 || rr[r] > 8) return fail(GRW_EINVAL, "invalid RyR state at FRU " + std::to_string(i));
      chan[(size_t)cl * ld + i] = ch_pack(lcc, y, rr, it);
    }
    if (mti[i] < 0 || mti[i] > kMtN) return fail(GRW_EINVAL, "mti out of range at FRU " + std::to_string(i));
    idx[i] = mti[i];  // mti == 624 is lazy position 0 of the next generation
    std::memcpy(&rows[(size_t)i * kMtStride], mt + (size_t)i * kMtN, kMtN * sizeof(uint32_t));
  }
  HIPCHK(hipMemcpyAsync(c->d.conc, conc.data(), conc.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d.chan, chan.data(), chan.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d.mtidx, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->d.nlog, 0, (size_t)ld * sizeof(int32_t), c->stream));
  HIPCHK(hipMemcpyAsync(c->d.mt, rows.data(), rows.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return grw_save_state(c);
}

int grw_get_state(grw_ctx* c, double* fru, int32_t* ltype, int32_t* ryr, int32_t* ito2,
                  int32_t* mti, uint32_t* mt) {
  if (!c) return fail(GRW_EINVAL, "null context");
  const int n = c->n, ld = c->ld;
  std::vector<double> conc(5 * (size_t)ld);
  std::vector<uint32_t> chan(4 * (size_t)ld);
  std::vector<int32_t> idx(ld);
  HIPCHK(hipMemcpyAsync(conc.data(), c->d.conc, conc.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(chan.data(), c->d.chan, chan.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(idx.data(), c->d.mtidx, idx.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  std::vector<uint32_t> rows;
  if (mt || mti) {
    rows.resize((size_t)n * kMtStride);
    HIPCHK(hipMemcpyAsync(rows.data(), c->d.mt, rows.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    if (fru)
      for (int k = 0; k < 5; ++k) fru[5 * (size_t)i + k] = conc[(size_t)k * ld + i];
    for (int cl = 0; cl < 4; ++cl) {
      const uint32_t w = chan[(size_t)cl * ld + i];
      if (ltype) {
        ltype[8 * (size_t)i + 2 * cl] = ch_lcc(w);
        ltype[8 * (size_t)i + 2 * cl + 1] = ch_y(w);
      }
      if (ryr)
        for (int r = 0; r < 5; ++r) ryr[20 * (size_t)i + 5 * cl + r] = ch_ryr(w, r);
      if (ito2) ito2[4 * (size_t)i + cl] = ch_ito2(w);
    }
    if (mt || mti) {
      // convert the lazy row back to the reference's fully regenerated form
      uint32_t* row = &rows[(size_t)i * kMtStride];
      int m = idx[i];
      if (m >= kMtN) {
        const int p = m - kMtN;
        if (p == 0) {
          m = kMtN;
        } else {
          for (int q = p; q < kMtN; ++q) {
            const uint32_t y = (row[q] & 0x80000000u) | (row[q < kMtN - 1 ? q + 1 : 0] & 0x7fffffffu);
            row[q] = row[q < kMtN - 397 ? q + 397 : q - (kMtN - 397)] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
          }
          m = p;
        }
      }
      if (mti) mti[i] = m;
      if (mt) std::memcpy(mt + (size_t)i * kMtN, row, kMtN * sizeof(uint32_t));
    }
  }
  return 0;
}

int grw_seed(grw_ctx* c, uint32_t a0) {
  if (!c) return fail(GRW_EINVAL, "null context");
  SimArgs a = base_args(c);
  hipLaunchKernelGGL(k_seed, dim3(c->grid), dim3(c->block), 0, c->stream, a, c->d, a0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return grw_save_state(c);
}

int grw_simfru_local(grw_ctx* c, double t0, double t1, const double dep0[3], const double depf[3]) {
  if (!c || !dep0 || !depf) return fail(GRW_EINVAL, "null argument");
  if (int rc0 = check_interval(t0, t1, dep0, depf)) return rc0;
  SimArgs a = base_args(c);
  a.t0 = t0;
  a.t1 = t1;
  for (int k = 0; k < 3; ++k) {
    a.dep0[k] = dep0[k];
    a.depf[k] = depf[k];
  }
  int rc = run_sim(c, a, false);
  if (rc) return rc;
  return check_err(c);
}

int grw_distrib_simfru(grw_ctx* c, double t0, double t1, const double dep0[3], const double depf[3],
                       double out4[4]) {
  if (!c || !dep0 || !depf || !out4) return fail(GRW_EINVAL, "null argument");
  if (int rc0 = check_interval(t0, t1, dep0, depf)) return rc0;
  SimArgs a = base_args(c);
  a.t0 = t0;
  a.t1 = t1;
  for (int k = 0; k < 3; ++k) {
    a.dep0[k] = dep0[k];
    a.depf[k] = depf[k];
    a.fdep[k] = depf[k];  // mpi_master.c:401-407 converts with the end estimate
  }
  int rc = run_sim(c, a, true);
  if (rc) return rc;
  if ((rc = fetch_out(c, 0, 4, out4))) return rc;
  return check_err(c);
}

int grw_calc_flux(grw_ctx* c, const double dep[3], double out4[4]) {
  if (!c || !dep || !out4) return fail(GRW_EINVAL, "null argument");
  int rc = flux_now(c, dep);
  if (rc) return rc;
  return fetch_out(c, 0, 4, out4);
}

int grw_calc_fru_flux_local(grw_ctx* c, double num5[5]) {
  if (!c || !num5) return fail(GRW_EINVAL, "null argument");
  const double dep[3] = {-80.0, 1e-4, 1.0};
  int rc = flux_now(c, dep);
  if (rc) return rc;
  return fetch_out(c, 4, 5, num5);
}

int grw_calc_fru_avg_local(grw_ctx* c, double num11[11]) {
  if (!c || !num11) return fail(GRW_EINVAL, "null argument");
  SimArgs a = base_args(c);
  if (c->cfg.sequential_sums)
    hipLaunchKernelGGL(k_seq_sums, dim3(1), dim3(1), 0, c->stream, a, c->d);
  else
    hipLaunchKernelGGL(k_avg, dim3(c->grid), dim3(c->block), 0, c->stream, a, c->d);
  HIPCHK(hipGetLastError());
  return fetch_out(c, 16, 11, num11);
}

int grw_calc_avg(grw_ctx* c, const double state[37], double other[12]) {
  if (!c || !state || !other) return fail(GRW_EINVAL, "null argument");
  double num[11];
  int rc = grw_calc_fru_avg_local(c, num);
  if (rc) return rc;
  avg_outputs(num, state, c->n, c->cfg.nfru_scale, other);
  return 0;
}

int grw_save_state(grw_ctx* c) {
  if (!c) return fail(GRW_EINVAL, "null context");
  SimArgs a = base_args(c);
  hipLaunchKernelGGL(k_save, dim3(c->grid), dim3(c->block), 0, c->stream, a, c->d);
  HIPCHK(hipGetLastError());
  return 0;
}

int grw_resume_state(grw_ctx* c) {
  if (!c) return fail(GRW_EINVAL, "null context");
  SimArgs a = base_args(c);
  hipLaunchKernelGGL(k_resume, dim3(c->grid), dim3(c->block), 0, c->stream, a, c->d);
  HIPCHK(hipGetLastError());
  return 0;
}

int grw_get_stats(grw_ctx* c, grw_stats* st) {
  if (!c || !st) return fail(GRW_EINVAL, "null argument");
  unsigned long long s = 0;
  HIPCHK(hipMemcpyAsync(&s, c->d.nsteps, sizeof s, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *st = c->stats;
  st->fru_steps = s;
  return 0;
}

int grw_reset_stats(grw_ctx* c) {
  if (!c) return fail(GRW_EINVAL, "null context");
  HIPCHK(hipMemsetAsync(c->d.nsteps, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->stats = grw_stats{};
  return 0;
}

int grw_set_timing(grw_ctx* c, int on) {
  if (!c) return fail(GRW_EINVAL, "null context");
  c->timing = on != 0;
  return 0;
}

}  // extern "C"
