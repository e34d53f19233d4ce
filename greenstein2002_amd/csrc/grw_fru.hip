// grw_fru.hip -- gfx950 kernels and the C ABI (include/grw_fru.h) of the
// release-unit Monte Carlo.
//
// Layout in HBM (ld = nfru rounded up to 64):
//   conc  double [5][ld]   CaJSR, CaSS1..4                      (SoA, coalesced)
//   chan  uint32 [4][ld]   packed LCC/gate/5 RyR/Ito2 per cleft  (grw_model.h)
//   mtidx int32  [ld]      stream position, see MtLane
//   nreg  int32  [ld]      words regenerated since the last checkpoint
//   mt    uint32 [n][640]  one MT19937 row per FRU (624 words used, 16-byte aligned)
//   *_h                    checkpoint of conc / chan / mtidx, and mtw_h: the
//                          first word the next regeneration overwrites
//
// One lane advances one FRU through the whole sub-interval (simfru_local.c:127-277)
// with every continuous and discrete state in registers; each lane draws from
// its own stream, so lanes never wait on each other except at the flux
// reduction, which is fused into the tail of the same launch.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/grw_fru.h"
#include "../../include/grw_batch.h"
#include "grw_global.h"
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
// Rollback: a checkpoint keeps the stream position and the one word the next
// regeneration overwrites first; nreg counts the words regenerated since.  The
// regeneration is invertible (mt_rollback), so a rollback recomputes the
// overwritten words from the row itself and nothing is logged per draw.
struct MtLane {
  uint32_t* row;
  int idx;
  int nreg;

  // (y ^ ((y << s) & B) as one three-input bit operation, truth table 0x78 =
  // src0 ^ (src1 & src2), instead of an and with a literal and an xor)
  __device__ __forceinline__ static uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y = __builtin_amdgcn_bitop3_b32(y, y << 7, 0x9d2c5680u, 0x78);
    y = __builtin_amdgcn_bitop3_b32(y, y << 15, 0xefc60000u, 0x78);
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
    ++nreg;
    row[p] = w;
    if (p < kMtMirror) row[p + kMtN] = w;  // the copy of words 0..7 (kMtMirror)
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
        row[q] = far[k];
        if (q < kMtMirror) row[q + kMtN] = far[k];  // the copy of words 0..7
      }
      nreg += 32;
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

// Micro-steps one FRU may take in one sub-interval before the launch gives up
// with GRW_ESTATE (the fastest exit rate, ~6e5/ms, needs ~6e4 steps per 0.1 ms).
constexpr int kStepCap = 2000000;

struct SimArgs {
  double t0, t1;
  double dep0[3], depf[3];
  double fdep[3];   // (V, Cai, CaNSR) used for the flux conversion
  double scale;
  double k23c, gcf; // scaled k23cf, gammacf
  int nfru, ld;
  int f0, b0;       // the cell's first FRU in the arrays and first block of the grid
  int ncells;
  int pro;          // before the sub-interval: 1 checkpoint, 2 roll back each FRU (first RK stage)
  int avg;          // also reduce the average sums of the FRU state at the end
  int sorted;       // quads take their FRU from d.perm and record its step count in d.cost
};

// One cell of a batch: its FRUs are [f0, f0 + nfru) of the shared arrays (f0 a
// multiple of 64); the Monte Carlo grid (64 FRUs per block) gives it blocks
// [sb0, sb0 + snb), the one-thread-per-FRU grid blocks [bb0, bb0 + bnb).
struct CellInfo {
  int nfru, f0;
  int sb0, snb;
  int bb0, bnb;
  double scale, k23c, gcf;
  unsigned seed;  // master seed of the cell's MT19937 streams
};

struct Dev {
  double* conc;
  uint32_t* chan;
  int32_t* mtidx;
  int32_t* nreg;
  uint32_t* mt;
  double* conc_h;
  uint32_t* chan_h;
  int32_t* mtidx_h;
  uint32_t* mtw_h;
  double* part;              // [grid][kSlot] block partial sums
  double* part2;             // [grid][kSlot] block partials collected by polling (kSlotEmpty when not yet written)
  unsigned int* ticket;      // last-block election counters, one per cell (self-resetting)
  unsigned int* qhead;       // next work item of a queue-layout Monte Carlo launch
  double* out;               // [32]: 0-3 fluxes, 4-8 flux sums, 16-26 avg sums
  unsigned long long* nsteps;        // [0]: FRU-steps; [kLaneUse]: FRU slots x passes of the wavefronts
  int* err;
  unsigned int* maxsteps;           // largest FRU-step count of the current launch
  unsigned long long* critsteps;    // sum over launches of maxsteps
  unsigned long long* tailticks;    // wall-clock ticks spent in the serial launch tails
  const CellInfo* cells;            // [ncells]
  const int* smap;                  // Monte Carlo grid block -> cell
  const int* bmap;                  // one-thread-per-FRU grid block -> cell
  int32_t* perm;                    // [ld]: FRU (cell-local) of each quad slot of the cell, when sorted
  uint32_t* cost;                   // [ld]: micro-steps of each FRU in its last sub-interval, when sorted
  int32_t* bkey;                    // [grid]: cost class of each Monte Carlo block's costliest FRU, when sorted
  int32_t* border;                  // [grid]: Monte Carlo block run by each dispatched block, when sorted
};

// Categorical transition draw, simfru_local.c:217-227 / :234-245: stay with
// probability 1 - dt*k_tot, else walk the cumulative exit probabilities with '>'.
template <class R>
__device__ __forceinline__ int pick(double u, double dt, const Exits<R>& e) {
  // the walk's three partial sums in the reference's order, then selects from
  // the last exit back to "stay" (no branches: lanes disagree on the exit)
  const double acc0 = 1.0 - dt * (double)e.tot;
  const double acc1 = acc0 + dt * (double)e.r[0];
  const double acc2 = acc1 + dt * (double)e.r[1];
  int out = e.d[2];
  out = (!(u > acc2) || e.n == 2) ? e.d[1] : out;
  out = (!(u > acc1) || e.n == 1) ? e.d[0] : out;
  out = !(u > acc0) ? 0 : out;
  return out;
}

// ---------------------------------------------------------------------------
// Deterministic reduction of NV doubles per lane: xor-butterfly inside the
// wavefront, waves in order through LDS, block partials in fixed slots, and the
// last block to finish (ticket) sums the slots in a fixed order.  Returns true
// in thread 0 of the last block, with the totals in tot[].  Block b owns
// d.part[b * kSlot, (b + 1) * kSlot); slot_off + NV must not exceed kSlot.
constexpr int kSlot = 32;
// The lane-use counter's place after the FRU-step counter: a 128-byte line of
// its own, so that the two atomics every wavefront ends with do not meet on one line.
constexpr int kLaneUse = 16;

// The wavefront's xor butterfly v += v[lane ^ k] for k = 32, 16, 8, 4, 2, 1 in
// that order, with each step's partner value moved by lane permutes instead of
// LDS.  k = 32, 16: the permlane swaps leave the two registers of a lane
// holding v_i and v_(i^k) in some order, and their sum is the same either way;
// k = 8, 4: DPP row rotations by 8 and 4, whose source lane is i ^ 8, or i ^ 4
// or i ^ 12 (after the first three steps lanes i ^ 4 and i ^ 12 hold the same
// value); k = 2, 1: quad permutes.  Every step adds the same two numbers as
// the plain butterfly, so the result is the same in every lane, bit for bit.
// (MAX: the larger of the two instead of the sum, for values that are not NaN.)
template <bool MAX = false>
__device__ __forceinline__ double swap_sum(double v, bool rows16) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)b, hi = (unsigned)((unsigned long long)b >> 32);
  unsigned alo, clo, ahi, chi;
  if (rows16) {
    const auto pl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto ph = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    alo = pl[0]; clo = pl[1]; ahi = ph[0]; chi = ph[1];
  } else {
    const auto pl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto ph = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    alo = pl[0]; clo = pl[1]; ahi = ph[0]; chi = ph[1];
  }
  const double a = __longlong_as_double((long long)(((unsigned long long)ahi << 32) | alo));
  const double c = __longlong_as_double((long long)(((unsigned long long)chi << 32) | clo));
  return MAX ? rmax(a, c) : a + c;
}

template <int CTRL, bool MAX = false>
__device__ __forceinline__ double dpp_add(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  const double w = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
  return MAX ? rmax(v, w) : v + w;
}

// Sum and maximum of a 32-bit value over the wavefront, by the same lane
// permutes (integers: the order is immaterial).
template <bool MAX>
__device__ __forceinline__ unsigned wave_u32(unsigned v) {
  const auto p32 = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  v = MAX ? max((unsigned)p32[0], (unsigned)p32[1]) : (unsigned)p32[0] + (unsigned)p32[1];
  const auto p16 = __builtin_amdgcn_permlane16_swap(v, v, false, false);
  v = MAX ? max((unsigned)p16[0], (unsigned)p16[1]) : (unsigned)p16[0] + (unsigned)p16[1];
  unsigned w = (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);
  v = MAX ? max(v, w) : v + w;
  w = (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, false);
  v = MAX ? max(v, w) : v + w;
  w = (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
  v = MAX ? max(v, w) : v + w;
  w = (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
  return MAX ? max(v, w) : v + w;
}

template <bool MAX = false>
__device__ __forceinline__ double wave_sum_xor(double v) {
  v = swap_sum<MAX>(v, false);  // k = 32
  v = swap_sum<MAX>(v, true);   // k = 16
  v = dpp_add<0x128, MAX>(v);   // row_ror:8, k = 8
  v = dpp_add<0x124, MAX>(v);   // row_ror:4, k = 4
  v = dpp_add<0x4E, MAX>(v);    // quad_perm [2,3,0,1], k = 2
  v = dpp_add<0xB1, MAX>(v);    // quad_perm [1,0,3,2], k = 1
  return v;
}

// blk < 0: a block without Monte Carlo work that only counts as an arrival
// (extra: the number of those among the ticket's arrivals).
template <int NV>
__device__ __forceinline__ bool reduce_range(double v[NV], const Dev& d, unsigned int* ticket, int blk, int b0,
                             int nb, int slot_off, double* tot_sh, int extra = 0) {
  static_assert(NV <= kSlot, "block partials overflow their slot row");
  __shared__ double wsum[4][NV];
  __shared__ bool last;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum_xor(v[k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) wsum[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    // The partials go out as agent-scope atomic stores, which write through to
    // the point where every XCD sees them, and the wait for their completion
    // orders them before the ticket.  A device fence here would write back and
    // invalidate this XCD's whole L2, evicting the MT rows of the blocks still
    // running on it; the last block reads the partials with agent-scope atomic
    // loads and needs no acquire fence either (nothing else it reads was
    // written in this launch).
    if (blk >= 0)
      for (int k = 0; k < NV; ++k) {
        double s = wsum[0][k];
        for (int w = 1; w < nw; ++w) s += wsum[w][k];
        __hip_atomic_store(&d.part[(size_t)blk * kSlot + slot_off + k], s, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
    // s_waitcnt is an intrinsic without memory semantics: the compiler-only
    // fences (they emit nothing) keep the compiler from moving the partial
    // stores past it or the ticket before it; the wait itself orders them on
    // the hardware.
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == (unsigned)(nb + extra) - 1u);
  }
  __syncthreads();
  if (!last) return false;
  // every thread of the last block sums a strided subset of the slots (fixed order)
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b0 + nb; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      s[k] += __hip_atomic_load(&d.part[(size_t)b * kSlot + slot_off + k], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = wave_sum_xor(s[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) wsum[wave][k] = s[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; ++k) {
      double a = wsum[0][k];
      for (int w = 1; w < nw; ++w) a += wsum[w][k];
      tot_sh[k] = a;
    }
    *ticket = 0u;  // ready for the next launch
  }
  __syncthreads();
  return true;
}

template <int NV>
__device__ __forceinline__ bool reduce_grid_block(double v[NV], const Dev& d, int slot_off, double* tot_sh) {
  return reduce_range<NV>(v, d, d.ticket, blockIdx.x, 0, gridDim.x, slot_off, tot_sh);
}

// Thread 0 of the last block gets the totals in tot[] (returns true there only).
template <int NV>
__device__ __forceinline__ bool reduce_grid(double v[NV], const Dev& d, int slot_off, double tot[NV]) {
  __shared__ double tsh[NV];
  if (!reduce_grid_block<NV>(v, d, slot_off, tsh)) return false;
  if (threadIdx.x != 0) return false;
  for (int k = 0; k < NV; ++k) tot[k] = tsh[k];
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
// Allocation-free flux sums (grw_calc_fru_flux_local_noalloc): the block
// partials and the last-block ticket of the fixed-order reduction live in one
// file-scope device array instead of memory a context allocates.  Blocks of
// 256 threads, one FRU per thread, as the context path of
// grw_calc_fru_flux_local launches them, so the sums are the same bit for bit.
constexpr int kNoAllocBlocks = 4096;  // at most 4096 x 256 = 1,048,576 FRUs
constexpr int kNoAllocMaxFru = kNoAllocBlocks * 256;
struct NoAllocScratch {
  double part[kNoAllocBlocks * kSlot];  // block partial sums, kSlot doubles per block (256 B)
  unsigned int ticket;                  // arrivals; zero at load, reset by every call's last block
};
static __device__ __align__(256) NoAllocScratch scratch;

// calc_fru_flux_local.c:28-77 over the reference's arrays (fru [n][5],
// ltype [n][4][2], ito2 [n][4]) in device memory; num5 (device) gets the sums.
__global__ void __launch_bounds__(256) k_flux_sums_ref(int n, const int32_t* ltype, const int32_t* ito2,
                                                       const double* fru, double* num5, NoAllocScratch* sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    double conc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) conc[k] = fru[(size_t)5 * i + k];
    uint32_t ch[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {  // the fields flux_terms reads: LCC state, gate, Ito2
      const int lcc = ltype[(size_t)8 * i + 2 * c], y = ltype[(size_t)8 * i + 2 * c + 1];
      const int it = ito2[(size_t)4 * i + c];
      ch[c] = ((uint32_t)lcc & 15u) | ((uint32_t)(y - 1) & 1u) << 4 | ((uint32_t)(it - 1) & 1u) << 20;
    }
    flux_terms(conc, ch, acc);
  }
  Dev d{};
  d.part = sc->part;
  __shared__ double tot[5];
  if (reduce_range<5>(acc, d, &sc->ticket, blockIdx.x, 0, gridDim.x, 0, tot) && threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) num5[k] = tot[k];
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
      MtLane m{d.mt + (size_t)i * kMtStride, d.mtidx[i], d.nreg[i]};
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
      int nguard = 0;
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
        if (!(t + dt > t) || ++nguard > kStepCap) {  // non-finite rates or runaway
          bad = true;
          break;
        }
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
      d.nreg[i] = m.nreg;
    }
    if (REDUCE) flux_terms(conc, ch, acc);
  }
  if (SIM) {
    // FRU-step count for the throughput figures
    unsigned long long s = steps;
    unsigned m = steps;
    for (int off = 32; off >= 1; off >>= 1) {
      s += __shfl_xor(s, off);
      m = max(m, (unsigned)__shfl_xor((int)m, off));
    }
    if ((threadIdx.x & 63) == 0 && s) {
      atomicAdd(d.nsteps, s);
      atomicAdd(d.nsteps + kLaneUse, 64ull * m);  // one FRU per lane
    }
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

// ---------------------------------------------------------------------------
// Four lanes per FRU: lane c of a quad owns cleft c (its LCC, gate, five RyRs,
// Ito2 and CaSS); CaJSR, time, step length and the MT position are kept
// identically in all four lanes.  Cross-cleft terms (the JRyR sum, the ring
// diffusion, the largest exit rate, the count of merged-state uniforms) move
// through DPP quad permutes, and every sum keeps the reference's order, so the
// arithmetic is operation-for-operation that of simfru_local.c / fcn_fru.c.

// DPP quad_perm controls: broadcast lane k, next / previous lane, xor 1 / 2.
constexpr int QB0 = 0x00, QB1 = 0x55, QB2 = 0xAA, QB3 = 0xFF;
constexpr int QNEXT = 0x39, QPREV = 0x93, QX1 = 0xB1, QX2 = 0x4E;

// mov_dpp leaves a lane whose source lane is invalid undefined; every lane of
// a quad reads a lane of its own quad and quads are active as a whole, so all
// lanes are written.  (update_dpp with the value itself as the "old" operand
// makes the compiler copy the value before every permute; the row rotations
// of the reductions below read valid lanes of a full wavefront likewise.)
template <int CTRL>
__device__ __forceinline__ int qperm(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float qperm(float v) {
  return __int_as_float(qperm<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double qperm(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = qperm<CTRL>((int)(unsigned)(b & 0xffffffffll));
  const int hi = qperm<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// LCC exits without per-state branches (fru_rates_local.c:176-334).  For mode
// normal (states 1-5, k = s-1) the exits are back k*beta, forward (4-k)*alpha
// (fL from C4) and mode switch 2^k*gamma; for mode Ca (states 7-11, k = s-7)
// back k*beta', forward (4-k)*alpha' = 2(4-k)*alpha (fL' from CCa4) and mode
// switch omega/bL^k; states 6 and 12 have the single exit gL / gL'.  Multiplying
// by the exact small integers gives the reference's products bit for bit, and
// the total keeps its summation order ((back + fwd) + mode, but
// (back + mode) + fwd for CCa4).
// The mode-Ca switch rates omega / bL^k come from the caller (a table lookup);
// the mode-normal ones 2^k gamma are exact scalings (ldexp).  For k = 0 the
// product k * beta is +0, the reference's absent back exit.
template <class R>
__device__ __forceinline__ void lcc_arith(int s, R alpha, R beta, R bp, R gam, R om, R& back, R& fwd,
                                          R& mode, R& tot) {
  const bool ca = (s >= 7) & (s <= 11);
  const int k = ca ? s - 7 : s - 1;
  back = (R)k * (ca ? bp : beta);
  const R fwd_n = (k < 4) ? (R)(4 - k) * alpha : R(fL);
  const R fwd_c = (k < 4) ? (R)(2 * (4 - k)) * alpha : R(fLprime);
  fwd = ca ? fwd_c : fwd_n;
  mode = ca ? om : ldexp(gam, k & 7);
  tot = (ca && k == 4) ? (back + mode) + fwd : (back + fwd) + mode;
  if (s == 6) tot = R(gL);
  if (s == 12) tot = R(gLprime);
}

// The exit list of an LCC that leaves its state, in the reference's order.
template <class R>
__device__ __forceinline__ int lcc_move(int s, double u, double dt, R back, R fwd, R mode, R tot) {
  Exits<R> e;
  e.tot = tot;
  if (s == 6 || s == 12) {
    e.r[0] = tot; e.d[0] = (s == 6) ? 5 : 11; e.n = 1;
  } else if (s == 1) {
    e.r[0] = fwd; e.d[0] = 2; e.r[1] = mode; e.d[1] = 7; e.n = 2;
  } else if (s <= 5) {
    e.r[0] = back; e.d[0] = s - 1; e.r[1] = fwd; e.d[1] = s + 1; e.r[2] = mode; e.d[2] = s + 6; e.n = 3;
  } else if (s == 7) {
    e.r[0] = mode; e.d[0] = 1; e.r[1] = fwd; e.d[1] = 8; e.n = 2;
  } else {
    e.r[0] = back; e.d[0] = s - 1; e.r[1] = mode; e.d[1] = s - 6; e.r[2] = fwd; e.d[2] = s + 1; e.n = 3;
  }
  return pick<R>(u, dt, e);
}

// ryr_exits (grw_model.h) by selects, for a RyR that moves in the four-lane
// kernel: the exit lists of the eight states with ryr_exits' expressions, the
// merged states' splitting factors a7, a8 taken from the step's rates (they
// are the same quotients), chosen by s.
// tot is the state's total, the same sum ryr_exits forms (looked up by the caller).
template <class R>
__device__ __forceinline__ void ryr_exits_sel(int s, const RyrK<R>& k, R a7, R a8, R tot, Exits<R>& e) {
  const int g = k.regime;
  const R a72 = (R(1.0) - a7) * R(k32), a73 = a7 * R(k45);  // state 7 (ryr_exits case 7)
  const R a82 = a8 * R(k52), a83 = a8 * k.k54;               // state 8
  R r0 = a82, r1 = a83, r2 = R(0.0);
  int d0 = 2, d1 = (g == 3) ? 7 : 4, d2 = 0, n = 2;
  const bool s1 = s == 1, s2 = s == 2, s3 = s == 3, s4 = s == 4, s5 = s == 5, s6 = s == 6, s7 = s == 7;
  r0 = s7 ? a72 : r0; r1 = s7 ? a73 : r1; d1 = s7 ? 8 : d1;
  r0 = s6 ? R(k65) : r0; d0 = s6 ? 5 : d0; n = s6 ? 1 : n;
  r0 = s5 ? R(k52) : r0; r1 = s5 ? k.k54 : r1;
  r2 = s5 ? k.k56 : r2; d1 = s5 ? 4 : d1; d2 = s5 ? 6 : d2; n = s5 ? 3 : n;
  r0 = s4 ? R(k43) : r0; r1 = s4 ? R(k45) : r1;
  d0 = s4 ? 3 : d0; d1 = s4 ? ((g == 1) ? 5 : 8) : d1;
  r0 = s3 ? R(k32) : r0; r1 = s3 ? k.k34 : r1; d1 = s3 ? 4 : d1;
  r0 = s2 ? R(k21) : r0; r1 = s2 ? k.k23 : r1;
  r2 = s2 ? k.k25 : r2; d0 = s2 ? 1 : d0; d1 = s2 ? ((g == 3) ? 7 : 3) : d1;
  d2 = s2 ? ((g == 1) ? 5 : 8) : d2; n = s2 ? 3 : n;
  r0 = s1 ? k.k12 : r0; n = s1 ? 1 : n;
  e.tot = tot;
  e.r[0] = r0; e.r[1] = r1; e.r[2] = r2;
  e.d[0] = (int8_t)d0; e.d[1] = (int8_t)d1; e.d[2] = (int8_t)d2;
  e.n = (int8_t)n;
}

// ryr_collapse and ryr_needs_draw of the raw states 1..8 in one regime, four
// bits per state (bits 0-2: the resolved state - 1; bit 3: the resolved state
// needs a uniform), so that a lane resolves each RyR with bit-field extracts
// instead of compares and selects on its state.
constexpr uint32_t ryr_resolve_bits(int regime) {
  uint32_t v = 0;
  for (int s = 1; s <= 8; ++s) {
    const int t = ryr_collapse(s, regime);
    v |= (uint32_t)((t - 1) | (ryr_needs_draw(t, regime) ? 8 : 0)) << (4 * (s - 1));
  }
  return v;
}
static_assert(((ryr_resolve_bits(3) >> 8) & 15u) == 6u, "regime 3: state 3 collapses to 7");
static_assert(((ryr_resolve_bits(1) >> 24) & 15u) == 14u, "regime 1: state 7 stays and needs a uniform");
static_assert(((ryr_resolve_bits(2) >> 16) & 15u) == 7u, "regime 2: state 5 collapses to 8, no uniform");
static_assert(((ryr_resolve_bits(2) >> 24) & 15u) == 14u, "regime 2: state 7 needs a uniform");
static_assert(((ryr_resolve_bits(3) >> 4) & 15u) == 1u, "regime 3: state 2 stays");

// Row position of absolute stream position g in [0, 1872): positions below 624
// are words of a loaded full generation, read as is; the others are lazy.
__device__ __forceinline__ int mt_row_pos(int g) {
  int q = (g >= kMtN) ? g - kMtN : g;
  return q - ((q >= kMtN) ? kMtN : 0);
}

// One regeneration step of the twist (mt19937_local.c:121-141 for one word).
__device__ __forceinline__ uint32_t mt_twist(uint32_t cur, uint32_t nxt, uint32_t far) {
  const uint32_t y = (cur & 0x80000000u) | (nxt & 0x7fffffffu);
  return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

// MT words of one lane for absolute stream positions g0 .. g0+N-1 (only the
// first n are produced), with dword accesses; used for the merged-state draws
// and after them.  All loads precede all stores: no word of a micro-step
// depends on another word of the same micro-step.
template <int N>
__device__ __forceinline__ void mt_run(uint32_t* row, int g0, int n, uint32_t out[N]) {
  uint32_t cur[N], nxt[N], far[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int g = g0 + k;
    const int qq = mt_row_pos(g);
    cur[k] = 0u; nxt[k] = 0u; far[k] = 0u;
    if (k < n) {
      cur[k] = row[qq];
      if (g >= kMtN) {
        nxt[k] = row[qq < kMtN - 1 ? qq + 1 : 0];
        far[k] = row[qq < kMtN - 397 ? qq + 397 : qq - (kMtN - 397)];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int g = g0 + k;
    if (k < n) {
      if (g >= kMtN) {
        const uint32_t wv = mt_twist(cur[k], nxt[k], far[k]);
        const int pos = mt_row_pos(g);
        row[pos] = wv;
        if (pos < kMtMirror) row[pos + kMtN] = wv;  // the copy of words 0..7 (kMtMirror)
        out[k] = MtLane::temper(wv);
      } else {
        out[k] = MtLane::temper(cur[k]);
      }
    }
  }
}

// The eight words of one cleft's draws: row positions q..q+8 (the draws and the
// word after them, at most 631) and the words they are xored with,
// q+397..q+404 mod 624 (at most 630), both for q < 227 and q >= 227; dword
// loads at fixed offsets from two bases.
struct MtWin8 {
  uint32_t cur[9], far[8];
};

__device__ __forceinline__ void mt_fetch8(const uint32_t* row, int g0, MtWin8& w) {
  const int q = mt_row_pos(g0);
  int qf = q + 397;
  qf -= (qf >= kMtN) ? kMtN : 0;
  const uint32_t* const pc = row + q;
  const uint32_t* const pf = row + qf;
#pragma unroll
  for (int k = 0; k < 9; ++k) w.cur[k] = pc[k];
#pragma unroll
  for (int k = 0; k < 8; ++k) w.far[k] = pf[k];
}

// Regeneration and stores for stream positions g0 .. g0+7 from a fetched window.
// Once a row is past its loaded generation (g0 >= 624, every window but the
// first ~20 after a load) all eight words are regenerated; otherwise a mask
// keeps the loaded word for a position of the loaded generation (g < 624),
// which is stored back unchanged (cur[k] is the word at row position g).  The
// two cases are a branch because a wavefront's lanes almost never disagree on
// it.  Row positions q + k past 623 are the
// copies of words 0..7, so the stores go to q + k and, for a window at either
// end of the row (rare), also to the other copy.
__device__ __forceinline__ void mt_finish8(uint32_t* row, int g0, const MtWin8& w, uint32_t out[8]) {
  const int q = mt_row_pos(g0);
  uint32_t* const p = row + q;
  uint32_t wv[8];
  if (g0 >= kMtN) {  // every word regenerated: all windows after a row's first generation
#pragma unroll
    for (int k = 0; k < 8; ++k) wv[k] = mt_twist(w.cur[k], w.cur[k + 1], w.far[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t m = 0u - (uint32_t)(g0 + k >= kMtN);
      wv[k] = (mt_twist(w.cur[k], w.cur[k + 1], w.far[k]) & m) | (w.cur[k] & ~m);
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    p[k] = wv[k];
    out[k] = MtLane::temper(wv[k]);
  }
  if (q < kMtMirror || q > kMtN - kMtMirror) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int pos = q + k;
      if (pos < kMtMirror) row[pos + kMtN] = wv[k];
      if (pos >= kMtN) row[pos - kMtN] = wv[k];
    }
  }
}

// Undo the nreg regenerations made since a checkpoint at stream position idx_h,
// whose first overwritten word was w_first.  A regenerated word w at row
// position q came from w = far ^ (y >> 1) ^ (y & 1 ? 0x9908b0df : 0), with
// y = (old[q] & 0x80000000) | (next & 0x7fffffff) and far the word at
// (q + 397) mod 624 as it was then.  The magic constant's top bit makes the
// inversion unique: t = w ^ far gives y & 1 = t >> 31, then y >> 1.  So the
// regeneration at q yields old[q]'s top bit and the low 31 bits of the word
// after q, and old[q] = top bit from q | low bits from the regeneration before
// it.  Walking from the newest word back, every far word is what the row holds
// at that moment (words 397 later are already restored, words 227 earlier not
// yet touched), for any count, including more than one generation.  Chunks
// of 32 positions keep the loads of a chunk independent of its stores.
__device__ __forceinline__ uint32_t mt_untwist(const uint32_t* row, int q) {
  const uint32_t t = row[q] ^ row[q < kMtN - 397 ? q + 397 : q - (kMtN - 397)];
  const uint32_t b = t >> 31;
  const uint32_t yh = t ^ (b ? 0x9908b0dfu : 0u);
  return (yh << 1) | b;
}

__device__ __forceinline__ void mt_rollback(uint32_t* row, int idx_h, int nreg, uint32_t w_first) {
  if (nreg <= 0) return;
  const int q0 = idx_h >= kMtN ? idx_h - kMtN : 0;  // row position of the first regenerated word
  int hi = nreg - 1;                                // offsets from q0, restored newest first
  while (hi >= 1) {
    const int lo = hi >= 32 ? hi - 31 : 1;
    uint32_t y[33];
#pragma unroll
    for (int k = 0; k < 33; ++k) {
      const int o = lo - 1 + k;
      y[k] = (o <= hi) ? mt_untwist(row, (q0 + o) % kMtN) : 0u;
    }
#pragma unroll
    for (int k = 1; k < 33; ++k) {
      const int o = lo - 1 + k;
      if (o <= hi) {
        const int q = (q0 + o) % kMtN;
        const uint32_t v = (y[k] & 0x80000000u) | (y[k - 1] & 0x7fffffffu);
        row[q] = v;
        if (q < kMtMirror) row[q + kMtN] = v;  // the copy of words 0..7
      }
    }
    hi = lo - 1;
  }
  row[q0] = w_first;
  if (q0 < kMtMirror) row[q0 + kMtN] = w_first;
}

// Row word that the first regeneration after a checkpoint at idx overwrites.
__device__ __forceinline__ int mt_first_pos(int idx) { return idx >= kMtN ? idx - kMtN : 0; }

__device__ __forceinline__ double unif(uint32_t y) { return ((double)y) / 4294967296.0; }

// The terms of fcn_fru.c that depend on the membrane potential alone: the
// LCC flux factor exp(2VF/RT) and PCa 4 V F^2/RT / (exp(2VF/RT) - 1) / (2 VSS F)
// (fcn_fru.c:60-75), with the reference's operations.  The second evaluation of
// a Heun step and the first of the next step share V, so a step computes them
// once.
template <class R>
struct Ghk {
  R e, a2;
  bool zero;  // |V| < 1e-6: the limit form
};

template <class R>
__device__ __forceinline__ Ghk<R> ghk_terms(R V) {
  const R JDHconstant = rdiv(R(1.0), R(2.0) * R(VSS) * R(Faraday) * R(1000.0));
  const R VF_over_RT = rdiv(V, divk(RT_over_F));
  const R VFsq_over_RT = (R(1000.0) * R(Faraday)) * VF_over_RT;
  Ghk<R> g;
  g.e = xexp(R(2.0) * VF_over_RT);
  g.zero = fabs(V) < R(1.e-6);
  g.a2 = rdiv(R(PCa) * R(4.0) * VFsq_over_RT, g.e - R(1.0)) * JDHconstant;  // unused when zero
  return g;
}

// fcn_fru.c:23-137 for one cleft of the quad: returns d(CaJSR) (same in all
// four lanes) and d(CaSS) of this lane's cleft; g holds the V terms.
template <class R>
__device__ __forceinline__ void fcn_fru4(R CaJSR, R s, const Ghk<R>& g, R Cai, R CaNSR, bool lo, int no,
                                         R& dj, R& ds) {
  const R JDHconstant = rdiv(R(1.0), R(2.0) * R(VSS) * R(Faraday) * R(1000.0));
  const R Jtr = rdiv(CaNSR - CaJSR, divk(tautr));
  const R JRyR = R(JRyRmax) * ((R)no) * (CaJSR - s);
  const R JRyRtot = qperm<QB0>(JRyR) + qperm<QB1>(JRyR) + qperm<QB2>(JRyR) + qperm<QB3>(JRyR);
  const R Jxfer = rdiv(s - Cai, divk(tauxfer));
  const R Jss = rdiv(s - qperm<QNEXT>(s), divk(tauss2ss));
  const R Jss_prev = qperm<QPREV>(Jss);
  const R open = lo ? R(1.0) : R(0.0);
  const R JDHPR = g.zero ? -R(PCa) * R(2.0) * R(1000.0) * R(Faraday) * (s - R(Cao) * R(0.341)) * open * JDHconstant
                         : -(s * g.e - R(Cao) * R(0.341)) * g.a2 * open;
  const R a1s = rdiv(R(BSRtot) * R(KBSR), (s + R(KBSR)) * (s + R(KBSR)));
  const R a2s = rdiv(R(BSLtot) * R(KBSL), (s + R(KBSL)) * (s + R(KBSL)));
  const R bSS = rdiv(R(1.0), R(1.0) + a1s + a2s);
  const R a1j = rdiv(R(CSQNtot) * R(KmCSQN), (CaJSR + R(KmCSQN)) * (CaJSR + R(KmCSQN)));
  const R bJSR = rdiv(R(1.0), R(1.0) + a1j);
  dj = bJSR * (Jtr - rdiv(R(VSS), R(VJSR)) * JRyRtot);
  ds = bSS * (JDHPR + JRyR - Jxfer - Jss + Jss_prev);
}

// A value moved into vector registers: the Monte Carlo loop is short of scalar
// registers (its 64-bit constants are re-materialised or spilled every step),
// so its uniform loop invariants live in VGPRs instead.
template <class T>
__device__ __forceinline__ T in_vgpr(T x) {
  asm volatile("" : "+v"(x));
  return x;
}
template <bool V, class T>
__device__ __forceinline__ T vinv(T x) {
  if constexpr (V) return in_vgpr(x);
  else return x;
}

// One sub-interval of every FRU with four lanes per FRU; gid is the lane's
// index in its cell's four-lane grid (quad slot x 4 + cleft).  Returns this
// lane's FRU-step count (nonzero only in lane 0 of a quad) and its flux terms
// in acc.  NT is the block's thread count (the width of the lane-private LDS
// columns); VINV keeps the launch's uniform values in vector registers.
// One quad's FRU through a sub-interval (four lanes, one per cleft; CaJSR,
// time, step length and MT position are kept identically in the four): the
// state the micro-step loop carries.
template <class R>
struct QuadFru {
  uint32_t* row;
  double cj, css, t, dV, dCai, dNSR;
  uint32_t w;
  int idx, nreg, nopen, bad, nguard;
  bool lopen;
  unsigned steps;
  Ghk<R> gv;
  MtWin8 win;
};

// What every FRU of a cell shares in one launch: the sub-interval, the
// interpolation end points of (V, Cai, CaNSR), the cell's rate scales, and the
// lane's cleft and LDS column.
template <class R>
struct QuadStage {
  int c, tl;
  double t0, t1, o0, o1, o2, span0, span1, span2;
  DivK kint;
  R k23c, gcf;
};

template <class R, bool VINV>
__device__ __forceinline__ QuadStage<R> quad_stage(const SimArgs& a, int c, int tl) {
  QuadStage<R> k;
  k.c = c;
  k.tl = tl;
  k.t0 = vinv<VINV>(a.t0);
  k.t1 = vinv<VINV>(a.t1);
  k.o0 = vinv<VINV>(a.dep0[0]);
  k.o1 = vinv<VINV>(a.dep0[1]);
  k.o2 = vinv<VINV>(a.dep0[2]);
  k.span0 = vinv<VINV>(a.depf[0] - a.dep0[0]);
  k.span1 = vinv<VINV>(a.depf[1] - a.dep0[1]);
  k.span2 = vinv<VINV>(a.depf[2] - a.dep0[2]);
  const double tint = a.t1 - a.t0;
  k.kint = DivK{vinv<VINV>(tint), vinv<VINV>(1.0 / tint)};  // the sub-interval length, fixed for the launch
  k.k23c = vinv<VINV>((R)a.k23c);
  k.gcf = vinv<VINV>((R)a.gcf);
  return k;
}

// The lane's constant LDS entries: the RyR totals of states 4 and 6, and
// omega / bL^k of the mode-Ca LCC states CCa0..CCa4 (fru_rates_local.c:70-76
// constant expressions), one column per lane.
template <class R, int NT>
__device__ __forceinline__ void quad_tables(R (*rtot)[NT], R (*lom)[NT], int tl) {
  rtot[3][tl] = R(k43 + k45);
  rtot[5][tl] = R(k65);
  lom[0][tl] = R(omega);
  lom[1][tl] = R(omega / bL);
  lom[2][tl] = R(omega / bL2);
  lom[3][tl] = R(omega / bL3);
  lom[4][tl] = R(omega / bL4);
}

// The loop state of an FRU at the start of its sub-interval, its
// concentrations, channel word and MT position given.
template <class R>
__device__ __forceinline__ void quad_init(const QuadStage<R>& k, QuadFru<R>& s) {
  s.dV = k.o0;
  s.dCai = k.o1;
  s.dNSR = k.o2;
  s.lopen = ch_lcc_open(s.w);
  s.nopen = ch_ryr_open(s.w);
  s.t = k.t0;
  s.bad = 0;
  s.nguard = 0;
  s.steps = 0;
}

// One micro-step of the quad's FRU (simfru_local.c:127-277), for the refill
// layout: the four-lane kernel's micro-step (simfru4_gid's loop body, with
// the same operations in the same order) on the quad's state in a QuadFru.
// Returns false, with s.bad set, on an unknown channel state, non-finite
// rates or a runaway.
template <class R, int NT, bool LATE>
__device__ __forceinline__ bool quad_step(const QuadStage<R>& k, QuadFru<R>& s, R (*rtot)[NT], R (*lom)[NT]) {
  const int c = k.c, tl = k.tl;
  uint32_t* const row = s.row;
  uint32_t w = s.w;
  int idx = s.idx;
  const double t1 = k.t1;
  if (LATE) mt_fetch8(row, idx + 8 * c, s.win);  // this step's window, fetched at its start
  // ---- rates at the step start (fru_rates_local.c:126-777)
  const R V = (R)s.dV;
  const R vm35 = V - R(35.0);
  const R e0 = R(0.012) * vm35, e1 = R(-0.05) * vm35;
  const R e2 = rdiv(V + R(12.5), divk(5.0)), e3 = rdiv(V + R(30.0), divk(12.0));
  // One exp per lane.  Independent selects: a chained conditional on a
  // lane-varying value becomes a branch tree, and since the lanes of a
  // wavefront disagree every arm of it runs, each behind its own exec-mask
  // handling.
  R earg = e3;
  earg = (c == 2) ? e2 : earg;
  earg = (c == 1) ? e1 : earg;
  earg = (c == 0) ? e0 : earg;
  const R ex = xexp(earg);
  const R alpha = R(alphacf) * qperm<QB0>(ex);
  const R beta = R(betacf) * qperm<QB1>(ex);
  const R yinf = rdiv(R(yCa_frac), R(1.0) + qperm<QB2>(ex)) + (R(1.0) - R(yCa_frac));
  const R tau = R(60.0) + rdiv(R(340.0), R(1.0) + qperm<QB3>(ex));
  const R bp = rdiv(beta, divk(bL * 1.0));
  const R cc = (R)s.css;
  const int ls = ch_lcc(w);
  int bad = s.bad | (ls < 1 || ls > 12);
  R lb, lf, lm, lt;
  const int lk = ls - 7;  // CCa index; other states read entry 0 and ignore it
  lcc_arith<R>(ls, alpha, beta, bp, k.gcf * cc, lom[(lk >= 0 && lk <= 4) ? lk : 0][tl], lb, lf, lm, lt);
  const R vr = rdiv(ch_y(w) == 2 ? (R(1.0) - yinf) : yinf, tau);
  const R ir = (ch_ito2(w) == 1) ? R(kfIto2) * cc : R(kbIto2);
  R maxr = vmax(vmax(vmax(R(0.0), lt), vr), ir);
  const RyrK<R> rk = ryr_k(cc, k.k23c);
  int st[kNryr];
  int nd = 0;
  unsigned need = 0;
  uint32_t rbits = ryr_resolve_bits(3);  // independent selects, as for earg
  rbits = (rk.regime == 2) ? ryr_resolve_bits(2) : rbits;
  rbits = (rk.regime == 1) ? ryr_resolve_bits(1) : rbits;
#pragma unroll
  for (int r = 0; r < kNryr; ++r) {
    const uint32_t b = (rbits >> (4u * ((w >> (5 + 3 * r)) & 7u))) & 15u;  // raw state ch_ryr(w, r)
    st[r] = (int)(b & 7u) + 1;
    nd += (int)(b >> 3);
    need |= (b >> 3) << r;
  }
  // merged-state uniforms come first, cleft by cleft (fru_rates_local.c:452,476,581)
  const int n0 = qperm<QB0>(nd), n1 = qperm<QB1>(nd), n2 = qperm<QB2>(nd), n3 = qperm<QB3>(nd);
  const int E = n0 + n1 + n2 + n3;
  const int lazy0 = idx > kMtN ? idx : kMtN;
  // totals of the resolved RyR states (fru_rates_local.c:390-727); they
  // depend on CaSS alone, so they are computed ahead of the merged draws
  const R T2 = (R(k21) + rk.k23 + rk.k25);
  const R T3 = (R(k32) + rk.k34);
  const R T5 = (R(k52) + rk.k54 + rk.k56);
  const R a7 = rdiv(rk.k34, R(k43) + rk.k34);
  const R T7 = ((R(1.0) - a7) * R(k32)) + (a7 * R(k45));
  const R a8 = rdiv(R(k65), rk.k56 + R(k65));
  const R T8 = (a8 * R(k52)) + (a8 * rk.k54);
  rtot[0][tl] = rk.k12;
  rtot[1][tl] = T2;
  rtot[2][tl] = T3;
  rtot[4][tl] = T5;
  rtot[6][tl] = T7;
  rtot[7][tl] = T8;
  if (nd) {
    const int Ec = (c > 0 ? n0 : 0) + (c > 1 ? n1 : 0) + (c > 2 ? n2 : 0);
    uint32_t vw[kNryr];
    mt_run<kNryr>(row, idx + Ec, nd, vw);
#pragma unroll
    for (int r = 0; r < kNryr; ++r) {
      if ((need >> r) & 1u) {
        st[r] = ryr_split(st[r], unif(vw[0]), rk);
#pragma unroll
        for (int j = 0; j < kNryr - 1; ++j) vw[j] = vw[j + 1];
      }
    }
  }
  R tr[kNryr];
#pragma unroll
  for (int r = 0; r < kNryr; ++r) {
    w = ch_set_ryr(w, r, st[r]);
    tr[r] = rtot[st[r] - 1][tl];  // total of the resolved state (fru_rates_local.c:390-727)
    maxr = vmax_nc(maxr, tr[r]);
  }
  maxr = vmax_nc(maxr, qperm<QX1>(maxr));
  maxr = vmax_nc(maxr, qperm<QX2>(maxr));
  bad |= qperm<QX1>(bad);
  bad |= qperm<QX2>(bad);
  // ---- step length (simfru_local.c:135-146)
  double dt = 0.1 / (double)maxr;
  dt = vmin(dt, 10.001e-3);
  double cm = vmax_nc(s.css, qperm<QX1>(s.css));
  cm = vmax_nc(cm, qperm<QX2>(cm));
  const double cmax = vmax(0.e-3, cm);
  if (cmax > 1.5e-3) dt = vmin(dt, 5.001e-3);
  dt = vmin(dt, t1 - s.t);
  // an unknown channel state, non-finite rates or a runaway (quad-uniform)
  if (bad || !(s.t + dt > s.t) || ++s.nguard > kStepCap) {
    s.w = w;
    s.bad = 1;
    return false;
  }
  // ---- this cleft's 8 of the step's 32 uniforms (simfru_local.c:211), from
  // the words fetched during the previous step; when merged-state draws
  // moved the stream (E > 0, rare) the window is fetched again at the new
  // position first, so that both cases run the same code and the common
  // one waits for nothing but its window
  uint32_t dw[8];
  if (E != 0) mt_fetch8(row, idx + E + 8 * c, s.win);
  mt_finish8(row, idx + E + 8 * c, s.win, dw);
  {
    const int idn = idx + E + 32;
    const int adv = idn - lazy0;
    s.nreg += (adv > 0 ? adv : 0);
    idx = (idn >= 2 * kMtN) ? idn - kMtN : idn;
  }
  // Next step's words, fetched now so their latency hides behind this
  // step's Heun step and transitions and the next step's rates.  The words
  // used are positions idx .. idx+31 and those they regenerate from:
  // positions idx+1 .. idx+32 (not yet regenerated) and words at least 195
  // positions older than idx, none of which this step's stores (positions
  // idx-32 .. idx-1, in either copy) or any store before their use touches.
  if (!LATE) mt_fetch8(row, idx + 8 * c, s.win);
  // ---- transition decisions of this cleft (simfru_local.c:215-276); they
  // use only the step-start state, so they are taken ahead of the Heun step
  // in the same straight-line code and the two interleave
  const double u0 = unif(dw[0]);
  const bool mv_lcc = u0 > 1.0 - dt * (double)lt;
  const bool mv_y = unif(dw[1]) >= 1.0 - dt * (double)vr;
  const bool mv_ito = unif(dw[7]) >= 1.0 - dt * (double)ir;
  double ur[kNryr];
  bool mv_r[kNryr];
  bool moved = false;
#pragma unroll
  for (int r = 0; r < kNryr; ++r) {
    ur[r] = unif(dw[2 + r]);
    mv_r[r] = ur[r] > 1.0 - dt * (double)tr[r];
    moved |= mv_r[r];
  }
  // ---- Heun step (simfru_local.c:149-173)
  R djs, dss;
  fcn_fru4<R>((R)s.cj, cc, s.gv, (R)s.dCai, (R)s.dNSR, s.lopen, s.nopen, djs, dss);
  const double k1j = dt * (double)djs, k1s = dt * (double)dss;
  const double y1j = s.cj + k1j, y1s = s.css + k1s;
  s.t = s.t + dt;
  const double frac = rdiv(s.t - k.t0, k.kint);  // (t - t0) / tint, correctly rounded
  s.dV = k.o0 + frac * k.span0;
  s.dCai = k.o1 + frac * k.span1;
  s.dNSR = k.o2 + frac * k.span2;
  s.gv = ghk_terms<R>((R)s.dV);  // also the next step's first evaluation
  fcn_fru4<R>((R)y1j, (R)y1s, s.gv, (R)s.dCai, (R)s.dNSR, s.lopen, s.nopen, djs, dss);
  s.cj = s.cj + (k1j + dt * (double)djs) / 2.0;
  s.css = s.css + (k1s + dt * (double)dss) / 2.0;
  // ---- the transitions taken
  if (mv_lcc) w = ch_set_lcc(w, lcc_move<R>(ls, u0, dt, lb, lf, lm, lt));
  if (mv_y) w = ch_flip_y(w);
  // the RyRs that move, lowest first, one exit walk per round: lanes moving
  // different RyRs share a round instead of running one body per RyR index
  // (the moves touch disjoint fields of w, so the order is immaterial)
  unsigned mm = 0;
#pragma unroll
  for (int r = 0; r < kNryr; ++r) mm |= mv_r[r] ? (1u << r) : 0u;
  while (mm) {
    const int r = __builtin_ctz(mm);
    mm &= mm - 1u;
    int sr = st[0];
    double ru = ur[0];
#pragma unroll
    for (int q = 1; q < kNryr; ++q) {
      sr = (r == q) ? st[q] : sr;
      ru = (r == q) ? ur[q] : ru;
    }
    Exits<R> e;
    ryr_exits_sel(sr, rk, a7, a8, rtot[sr - 1][tl], e);
    w = ch_set_ryr(w, r, pick<R>(ru, dt, e));
  }
  if (mv_ito) w = ch_flip_ito2(w);
  s.lopen = ch_lcc_open(w);
  if (moved) s.nopen = ch_ryr_open(w);
  s.steps += (c == 0);
  s.w = w;
  s.idx = idx;
  s.bad = bad;
  return true;
}

// The flux-sum terms of one cleft of a finished FRU (calc_fru_flux_local.c,
// and calc_fru_avg_local.c:135-176 when avg is set; the CaJSR ones in lane 0),
// as simfru4_gid forms them at the end of its loop.
__device__ __forceinline__ void quad_acc(double css, double cj, uint32_t w, bool lopen, int nopen, int c, bool avg,
                                         double acc[17]) {
  acc[0] = css;
  acc[1] = (c == 0) ? cj : 0.0;
  acc[2] = lopen ? css : 0.0;
  acc[3] = lopen ? 1.0 : 0.0;
  acc[4] = (ch_ito2(w) == 2) ? 1.0 : 0.0;
  acc[5] = JRyRmax * ((double)nopen) * (cj - css);  // this cleft's release flux (calc_fru_avg_local.c)
  if (avg) {
    int nr = 0;
#pragma unroll
    for (int r = 0; r < kNryr; ++r) {
      const int st = ch_ryr(w, r);
      nr += (st == 1 || st == 2);
    }
    acc[6] = css;
    acc[7] = (c == 0) ? cj : 0.0;
    acc[8] = JRyRmax * ((double)nopen) * (cj - css);
    acc[9] = (double)nopen;
    acc[10] = (double)nr;
    acc[11] = css * (1.0 + BSRtot / (KBSR + css) + BSLtot / (KBSL + css));
    acc[12] = (c == 0) ? cj * (1.0 + CSQNtot / (KmCSQN + cj)) : 0.0;
    acc[13] = (ch_lcc(w) <= 6) ? 1.0 : 0.0;
    acc[14] = (ch_y(w) == 2) ? 1.0 : 0.0;
    acc[15] = lopen ? 1.0 : 0.0;
    acc[16] = (ch_ito2(w) == 2) ? 1.0 : 0.0;
  }
}

// Store the FRU's state at the end of its sub-interval (as simfru4_gid does).
template <class R>
__device__ __forceinline__ void quad_end(const SimArgs& a, const Dev& d, int f, int c, const QuadFru<R>& s) {
  if (s.bad && c == 0) atomicOr(d.err, 1);
  d.conc[(c + 1) * a.ld + f] = s.css;
  d.chan[c * a.ld + f] = s.w;
  if (c == 0) {
    d.conc[f] = s.cj;
    d.mtidx[f] = s.idx;
    d.nreg[f] = s.nreg;
    if (a.sorted && a.t1 > a.t0) d.cost[f] = s.steps;
  }
}

template <class R, int NT = 256, bool VINV = true, bool LATE = false>
__device__ __forceinline__ unsigned simfru4_gid(const SimArgs& a, const Dev& d, double acc[17], int gid) {
  const int slot = gid >> 2;
  const int c = gid & 3;
  unsigned steps = 0;
  if (slot < a.nfru) {  // quads are active or inactive as a whole
    const int f = a.f0 + (a.sorted ? d.perm[a.f0 + slot] : slot);
    uint32_t* const row = d.mt + (size_t)f * kMtStride;
    double cj, css;
    uint32_t w;
    int idx, nreg;
    if (a.pro == 2) {
      // rollback (mpi_slave.c:150-184): start from the checkpoint; lane 0
      // restores the MT words, the fence makes them visible to its quad
      cj = d.conc_h[f];
      css = d.conc_h[(c + 1) * a.ld + f];
      w = d.chan_h[c * a.ld + f];
      idx = d.mtidx_h[f];
      nreg = 0;
      if (c == 0) mt_rollback(row, idx, d.nreg[f], d.mtw_h[f]);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    } else {
      cj = d.conc[f];
      css = d.conc[(c + 1) * a.ld + f];
      w = d.chan[c * a.ld + f];
      idx = d.mtidx[f];
      nreg = a.pro ? 0 : d.nreg[f];
      if (a.pro == 1) {  // checkpoint (mpi_slave.c:192-226)
        d.conc_h[(c + 1) * a.ld + f] = css;
        d.chan_h[c * a.ld + f] = w;
        if (c == 0) {
          d.conc_h[f] = cj;
          d.mtidx_h[f] = idx;
          d.mtw_h[f] = row[mt_first_pos(idx)];
        }
      }
    }
    const double t0 = vinv<VINV>(a.t0), t1 = vinv<VINV>(a.t1);
    const double o0 = vinv<VINV>(a.dep0[0]), o1 = vinv<VINV>(a.dep0[1]), o2 = vinv<VINV>(a.dep0[2]);
    const double span0 = vinv<VINV>(a.depf[0] - a.dep0[0]);
    const double span1 = vinv<VINV>(a.depf[1] - a.dep0[1]);
    const double span2 = vinv<VINV>(a.depf[2] - a.dep0[2]);
    const double tint = a.t1 - a.t0;
    const DivK kint{vinv<VINV>(tint), vinv<VINV>(1.0 / tint)};  // the sub-interval length, fixed for the launch
    double dV = o0, dCai = o1, dNSR = o2;
    bool lopen = ch_lcc_open(w);
    int nopen = ch_ryr_open(w);
    const R k23c = vinv<VINV>((R)a.k23c), gcf = vinv<VINV>((R)a.gcf);
    double t = t0;
    int bad = 0;
    int nguard = 0;
    // Each lane's totals of the eight RyR states, one LDS column per lane, so
    // that a RyR's total is one indexed read instead of a chain of selects
    // over its state.  A lane reads only what it wrote itself, in program
    // order, so no barrier is needed.  States 4 and 6 have constant totals.
    __shared__ R rtot[8][NT];
    const int tl = threadIdx.x;
    rtot[3][tl] = R(k43 + k45);
    rtot[5][tl] = R(k65);
    // omega / bL^k of the mode-Ca LCC states CCa0..CCa4, likewise one column
    // per lane (fru_rates_local.c:70-76 constant expressions)
    __shared__ R lom[5][NT];
    lom[0][tl] = R(omega);
    lom[1][tl] = R(omega / bL);
    lom[2][tl] = R(omega / bL2);
    lom[3][tl] = R(omega / bL3);
    lom[4][tl] = R(omega / bL4);
    MtWin8 win;  // words of the next step's draws, fetched one step ahead
    // Complete the FRU-state loads here: otherwise the loop header merges them
    // with the prefetch below and the first use of the state in every step
    // waits for all outstanding loads, prefetch included.
    __builtin_amdgcn_s_waitcnt(0);
    if (!LATE && t < t1) mt_fetch8(row, idx + 8 * c, win);
    Ghk<R> gv = ghk_terms<R>((R)dV);
    // The refill layout runs the same micro-step through quad_step (its
    // results are held to these bit for bit by the layout parity tests); this
    // kernel keeps the loop inline, which compiles the single-cell launch 1%
    // faster than the same code through quad_step.
    while (t < t1) {
      if (LATE) mt_fetch8(row, idx + 8 * c, win);  // this step's window, fetched at its start
      // ---- rates at the step start (fru_rates_local.c:126-777)
      const R V = (R)dV;
      const R vm35 = V - R(35.0);
      const R e0 = R(0.012) * vm35, e1 = R(-0.05) * vm35;
      const R e2 = rdiv(V + R(12.5), divk(5.0)), e3 = rdiv(V + R(30.0), divk(12.0));
      // One exp per lane.  Independent selects: a chained conditional on a
      // lane-varying value becomes a branch tree, and since the lanes of a
      // wavefront disagree every arm of it runs, each behind its own exec-mask
      // handling.
      R earg = e3;
      earg = (c == 2) ? e2 : earg;
      earg = (c == 1) ? e1 : earg;
      earg = (c == 0) ? e0 : earg;
      const R ex = xexp(earg);
      const R alpha = R(alphacf) * qperm<QB0>(ex);
      const R beta = R(betacf) * qperm<QB1>(ex);
      const R yinf = rdiv(R(yCa_frac), R(1.0) + qperm<QB2>(ex)) + (R(1.0) - R(yCa_frac));
      const R tau = R(60.0) + rdiv(R(340.0), R(1.0) + qperm<QB3>(ex));
      const R bp = rdiv(beta, divk(bL * 1.0));
      const R cc = (R)css;
      const int ls = ch_lcc(w);
      bad |= (ls < 1 || ls > 12);
      R lb, lf, lm, lt;
      const int lk = ls - 7;  // CCa index; other states read entry 0 and ignore it
      lcc_arith<R>(ls, alpha, beta, bp, gcf * cc, lom[(lk >= 0 && lk <= 4) ? lk : 0][tl], lb, lf, lm, lt);
      const R vr = rdiv(ch_y(w) == 2 ? (R(1.0) - yinf) : yinf, tau);
      const R ir = (ch_ito2(w) == 1) ? R(kfIto2) * cc : R(kbIto2);
      R maxr = vmax(vmax(vmax(R(0.0), lt), vr), ir);
      const RyrK<R> k = ryr_k(cc, k23c);
      int st[kNryr];
      int nd = 0;
      unsigned need = 0;
      uint32_t rbits = ryr_resolve_bits(3);  // independent selects, as for earg
      rbits = (k.regime == 2) ? ryr_resolve_bits(2) : rbits;
      rbits = (k.regime == 1) ? ryr_resolve_bits(1) : rbits;
#pragma unroll
      for (int r = 0; r < kNryr; ++r) {
        const uint32_t b = (rbits >> (4u * ((w >> (5 + 3 * r)) & 7u))) & 15u;  // raw state ch_ryr(w, r)
        st[r] = (int)(b & 7u) + 1;
        nd += (int)(b >> 3);
        need |= (b >> 3) << r;
      }
      // merged-state uniforms come first, cleft by cleft (fru_rates_local.c:452,476,581)
      const int n0 = qperm<QB0>(nd), n1 = qperm<QB1>(nd), n2 = qperm<QB2>(nd), n3 = qperm<QB3>(nd);
      const int E = n0 + n1 + n2 + n3;
      const int lazy0 = idx > kMtN ? idx : kMtN;
      // totals of the resolved RyR states (fru_rates_local.c:390-727); they
      // depend on CaSS alone, so they are computed ahead of the merged draws
      const R T2 = (R(k21) + k.k23 + k.k25);
      const R T3 = (R(k32) + k.k34);
      const R T5 = (R(k52) + k.k54 + k.k56);
      const R a7 = rdiv(k.k34, R(k43) + k.k34);
      const R T7 = ((R(1.0) - a7) * R(k32)) + (a7 * R(k45));
      const R a8 = rdiv(R(k65), k.k56 + R(k65));
      const R T8 = (a8 * R(k52)) + (a8 * k.k54);
      rtot[0][tl] = k.k12;
      rtot[1][tl] = T2;
      rtot[2][tl] = T3;
      rtot[4][tl] = T5;
      rtot[6][tl] = T7;
      rtot[7][tl] = T8;
      if (nd) {
        const int Ec = (c > 0 ? n0 : 0) + (c > 1 ? n1 : 0) + (c > 2 ? n2 : 0);
        uint32_t vw[kNryr];
        mt_run<kNryr>(row, idx + Ec, nd, vw);
#pragma unroll
        for (int r = 0; r < kNryr; ++r) {
          if ((need >> r) & 1u) {
            st[r] = ryr_split(st[r], unif(vw[0]), k);
#pragma unroll
            for (int j = 0; j < kNryr - 1; ++j) vw[j] = vw[j + 1];
          }
        }
      }
      R tr[kNryr];
#pragma unroll
      for (int r = 0; r < kNryr; ++r) {
        w = ch_set_ryr(w, r, st[r]);
        tr[r] = rtot[st[r] - 1][tl];  // total of the resolved state (fru_rates_local.c:390-727)
        maxr = vmax_nc(maxr, tr[r]);
      }
      maxr = vmax_nc(maxr, qperm<QX1>(maxr));
      maxr = vmax_nc(maxr, qperm<QX2>(maxr));
      bad |= qperm<QX1>(bad);
      bad |= qperm<QX2>(bad);
      // ---- step length (simfru_local.c:135-146)
      double dt = 0.1 / (double)maxr;
      dt = vmin(dt, 10.001e-3);
      double cm = vmax_nc(css, qperm<QX1>(css));
      cm = vmax_nc(cm, qperm<QX2>(cm));
      const double cmax = vmax(0.e-3, cm);
      if (cmax > 1.5e-3) dt = vmin(dt, 5.001e-3);
      dt = vmin(dt, t1 - t);
      // an unknown channel state, non-finite rates or a runaway (quad-uniform)
      if (bad || !(t + dt > t) || ++nguard > kStepCap) {
        bad = 1;
        break;
      }
      // ---- this cleft's 8 of the step's 32 uniforms (simfru_local.c:211), from
      // the words fetched during the previous step; when merged-state draws
      // moved the stream (E > 0, rare) the window is fetched again at the new
      // position first, so that both cases run the same code and the common
      // one waits for nothing but its window
      uint32_t dw[8];
      if (E != 0) mt_fetch8(row, idx + E + 8 * c, win);
      mt_finish8(row, idx + E + 8 * c, win, dw);
      {
        const int idn = idx + E + 32;
        const int adv = idn - lazy0;
        nreg += (adv > 0 ? adv : 0);
        idx = (idn >= 2 * kMtN) ? idn - kMtN : idn;
      }
      // Next step's words, fetched now so their latency hides behind this
      // step's Heun step and transitions and the next step's rates.  The words
      // used are positions idx .. idx+31 and those they regenerate from:
      // positions idx+1 .. idx+32 (not yet regenerated) and words at least 195
      // positions older than idx, none of which this step's stores (positions
      // idx-32 .. idx-1, in either copy) or any store before their use touches.
      if (!LATE) mt_fetch8(row, idx + 8 * c, win);
      // ---- transition decisions of this cleft (simfru_local.c:215-276); they
      // use only the step-start state, so they are taken ahead of the Heun step
      // in the same straight-line code and the two interleave
      const double u0 = unif(dw[0]);
      const bool mv_lcc = u0 > 1.0 - dt * (double)lt;
      const bool mv_y = unif(dw[1]) >= 1.0 - dt * (double)vr;
      const bool mv_ito = unif(dw[7]) >= 1.0 - dt * (double)ir;
      double ur[kNryr];
      bool mv_r[kNryr];
      bool moved = false;
#pragma unroll
      for (int r = 0; r < kNryr; ++r) {
        ur[r] = unif(dw[2 + r]);
        mv_r[r] = ur[r] > 1.0 - dt * (double)tr[r];
        moved |= mv_r[r];
      }
      // ---- Heun step (simfru_local.c:149-173)
      R djs, dss;
      fcn_fru4<R>((R)cj, cc, gv, (R)dCai, (R)dNSR, lopen, nopen, djs, dss);
      const double k1j = dt * (double)djs, k1s = dt * (double)dss;
      const double y1j = cj + k1j, y1s = css + k1s;
      t = t + dt;
      const double frac = rdiv(t - t0, kint);  // (t - t0) / tint, correctly rounded
      dV = o0 + frac * span0;
      dCai = o1 + frac * span1;
      dNSR = o2 + frac * span2;
      gv = ghk_terms<R>((R)dV);  // also the next step's first evaluation
      fcn_fru4<R>((R)y1j, (R)y1s, gv, (R)dCai, (R)dNSR, lopen, nopen, djs, dss);
      cj = cj + (k1j + dt * (double)djs) / 2.0;
      css = css + (k1s + dt * (double)dss) / 2.0;
      // ---- the transitions taken
      if (mv_lcc) w = ch_set_lcc(w, lcc_move<R>(ls, u0, dt, lb, lf, lm, lt));
      if (mv_y) w = ch_flip_y(w);
      // the RyRs that move, lowest first, one exit walk per round: lanes moving
      // different RyRs share a round instead of running one body per RyR index
      // (the moves touch disjoint fields of w, so the order is immaterial)
      unsigned mm = 0;
#pragma unroll
      for (int r = 0; r < kNryr; ++r) mm |= mv_r[r] ? (1u << r) : 0u;
      while (mm) {
        const int r = __builtin_ctz(mm);
        mm &= mm - 1u;
        int sr = st[0];
        double ru = ur[0];
#pragma unroll
        for (int q = 1; q < kNryr; ++q) {
          sr = (r == q) ? st[q] : sr;
          ru = (r == q) ? ur[q] : ru;
        }
        Exits<R> e;
        ryr_exits_sel(sr, k, a7, a8, rtot[sr - 1][tl], e);
        w = ch_set_ryr(w, r, pick<R>(ru, dt, e));
      }
      if (mv_ito) w = ch_flip_ito2(w);
      lopen = ch_lcc_open(w);
      if (moved) nopen = ch_ryr_open(w);
      steps += (c == 0);
    }
    if (bad && c == 0) atomicOr(d.err, 1);
    d.conc[(c + 1) * a.ld + f] = css;
    d.chan[c * a.ld + f] = w;
    if (c == 0) {
      d.conc[f] = cj;
      d.mtidx[f] = idx;
      d.nreg[f] = nreg;
      if (a.sorted && a.t1 > a.t0) d.cost[f] = steps;
    }
    acc[0] = css;
    acc[1] = (c == 0) ? cj : 0.0;
    acc[2] = lopen ? css : 0.0;
    acc[3] = lopen ? 1.0 : 0.0;
    acc[4] = (ch_ito2(w) == 2) ? 1.0 : 0.0;
    acc[5] = JRyRmax * ((double)nopen) * (cj - css);  // this cleft's release flux (calc_fru_avg_local.c)
    if (a.avg) {  // calc_fru_avg_local.c:135-176, this cleft's terms (CaJSR ones in lane 0)
      int nr = 0;
#pragma unroll
      for (int r = 0; r < kNryr; ++r) {
        const int st = ch_ryr(w, r);
        nr += (st == 1 || st == 2);
      }
      acc[6] = css;
      acc[7] = (c == 0) ? cj : 0.0;
      acc[8] = JRyRmax * ((double)nopen) * (cj - css);
      acc[9] = (double)nopen;
      acc[10] = (double)nr;
      acc[11] = css * (1.0 + BSRtot / (KBSR + css) + BSLtot / (KBSL + css));
      acc[12] = (c == 0) ? cj * (1.0 + CSQNtot / (KmCSQN + cj)) : 0.0;
      acc[13] = (ch_lcc(w) <= 6) ? 1.0 : 0.0;
      acc[14] = (ch_y(w) == 2) ? 1.0 : 0.0;
      acc[15] = lopen ? 1.0 : 0.0;
      acc[16] = (ch_ito2(w) == 2) ? 1.0 : 0.0;
    }
  }
  return steps;
}

template <class R>
__device__ __forceinline__ unsigned simfru4_body(const SimArgs& a, const Dev& d, double acc[17], int blk) {
  return simfru4_gid<R>(a, d, acc, (blk - a.b0) * (int)blockDim.x + (int)threadIdx.x);
}

// ---------------------------------------------------------------------------
// Wave-specialised sub-interval (single cells).  A single cell is bound by the
// micro-steps of its slowest FRU, each issued by one wavefront at one
// instruction per issue slot; splitting a micro-step over the lanes of that
// wavefront would not shorten it (lanes that take different work serialise
// under exec masks), but splitting it over the four wavefronts of a block
// does: each runs on its own SIMD.  Here a block holds 16 FRUs, and lane l of
// every wavefront is cleft l & 3 of FRU l >> 2; the wavefronts take roles:
//   L  the LCC, its voltage gate and Ito2 (rates, draws, moves; MT words 0, 1, 7)
//   RA RyRs 0-1 (rates, draws, moves; MT words 2, 3)
//   RB RyRs 2-4 (MT words 4, 5, 6)
//   H  the Heun step of the Ca concentrations (fcn_fru twice)
// and exchange through LDS twice per micro-step: after the rates (the largest
// exit rate, so that every wavefront takes the same step length) and after the
// step (the concentrations and the channel fields each role changed).  Every
// value is computed with the operations of simfru4_body, so the results are
// the same bit for bit; the flux sums are grouped by reduce_ws into the
// partials of the 64-FRU blocks of the four-lane kernel.
constexpr int kWsL = 0, kWsRA = 1, kWsRB = 2, kWsH = 3;
constexpr int kWsFru = 16;  // FRUs per block
constexpr uint32_t kWsMaskL = 15u | (1u << 4) | (1u << 20);        // LCC, gate, Ito2
constexpr uint32_t kWsMaskRA = (7u << 5) | (7u << 8);              // RyR 0, 1
constexpr uint32_t kWsMaskRB = (7u << 11) | (7u << 14) | (7u << 17);  // RyR 2, 3, 4

// The RyR regime of ryr_k (fru_rates_local.c:384,505,622) alone.
template <class R>
__device__ __forceinline__ int ryr_regime(R ca) {
  const R cm = rmin(ca, R(threshMAXCa));
  return (cm < R(threshCa56to8)) ? 1 : (cm < R(threshCa34to7)) ? 2 : 3;
}

// One role's words of a cleft's eight: positions given as offsets into the
// cleft's 8-word window (K words; the twist of word k also reads word k + 1).
template <int K>
struct MtPart {
  uint32_t cur[K], nxt[K], far[K];
  uint32_t wv[K];  // the words after regeneration
};
template <int... O>
__device__ __forceinline__ void mt_fetch_part(const uint32_t* row, int g0, MtPart<sizeof...(O)>& w) {
  constexpr int off[] = {O...};
  const int q = mt_row_pos(g0);
  int qf = q + 397;
  qf -= (qf >= kMtN) ? kMtN : 0;
  const uint32_t* const pc = row + q;
  const uint32_t* const pf = row + qf;
#pragma unroll
  for (int k = 0; k < (int)sizeof...(O); ++k) {
    w.cur[k] = pc[off[k]];
    w.nxt[k] = pc[off[k] + 1];
    w.far[k] = pf[off[k]];
  }
}
// The regenerated words (kept for mt_store_part) and their tempered draws.
template <int... O>
__device__ __forceinline__ void mt_twist_part(int g0, MtPart<sizeof...(O)>& w, uint32_t* out) {
  constexpr int off[] = {O...};
#pragma unroll
  for (int k = 0; k < (int)sizeof...(O); ++k) {
    const uint32_t m = 0u - (uint32_t)(g0 + off[k] >= kMtN);
    w.wv[k] = (mt_twist(w.cur[k], w.nxt[k], w.far[k]) & m) | (w.cur[k] & ~m);
    out[k] = MtLane::temper(w.wv[k]);
  }
}
template <int... O>
__device__ __forceinline__ void mt_store_part(uint32_t* row, int g0, const MtPart<sizeof...(O)>& w) {
  constexpr int off[] = {O...};
  const int q = mt_row_pos(g0);
  uint32_t* const p = row + q;
#pragma unroll
  for (int k = 0; k < (int)sizeof...(O); ++k) p[off[k]] = w.wv[k];
  if (q < kMtMirror || q > kMtN - kMtMirror) {
#pragma unroll
    for (int k = 0; k < (int)sizeof...(O); ++k) {
      const int pos = q + off[k];
      if (pos < kMtMirror) row[pos + kMtN] = w.wv[k];
      if (pos >= kMtN) row[pos - kMtN] = w.wv[k];
    }
  }
}

// A wavefront's exec-uniform role (readfirstlane: the compiler then branches
// on a scalar instead of masking lanes).
__device__ __forceinline__ int ws_role() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

template <class R>
__device__ __forceinline__ unsigned simfru_ws(const SimArgs& a, const Dev& d, double acc[17], int blk) {
  const int role = ws_role();
  const int lane = threadIdx.x & 63;
  const int slot = (blk - 4 * a.b0) * kWsFru + (lane >> 2);
  const int c = lane & 3;
  const bool on = slot < a.nfru;  // a FRU's four lanes are on or off together, in every wavefront
  const int f = a.f0 + (on ? slot : 0);
  uint32_t* const row = d.mt + (size_t)f * kMtStride;
  // the exchange, one entry per lane (the same lane in every wavefront)
  __shared__ double x_max[3][64];
  __shared__ int x_bad[3][64];
  __shared__ double x_cj[64], x_css[64];
  __shared__ uint32_t x_w[3][64];
  __shared__ uint32_t x_res[64];  // RyR fields resolved by merged-state draws (rare)
  __shared__ R rtot[8][256];      // lane-private RyR totals (roles RA, RB), as in simfru4_body
  __shared__ R lom[5][256];       // lane-private omega / bL^k (role L)
  unsigned steps = 0;
  double cj = 0.0, css = 0.0;
  uint32_t w = 0u;
  int idx = 0, nreg = 0;
  if (on) {
    if (a.pro == 2) {  // rollback from the checkpoint (mpi_slave.c:150-184)
      cj = d.conc_h[f];
      css = d.conc_h[(c + 1) * a.ld + f];
      w = d.chan_h[c * a.ld + f];
      idx = d.mtidx_h[f];
      if (role == kWsL && c == 0) mt_rollback(row, idx, d.nreg[f], d.mtw_h[f]);
    } else {
      cj = d.conc[f];
      css = d.conc[(c + 1) * a.ld + f];
      w = d.chan[c * a.ld + f];
      idx = d.mtidx[f];
      nreg = a.pro ? 0 : d.nreg[f];
      if (a.pro == 1 && role == kWsH) {  // checkpoint (mpi_slave.c:192-226)
        d.conc_h[(c + 1) * a.ld + f] = css;
        d.chan_h[c * a.ld + f] = w;
        if (c == 0) {
          d.conc_h[f] = cj;
          d.mtidx_h[f] = idx;
          d.mtw_h[f] = row[mt_first_pos(idx)];
        }
      }
    }
  }
  __syncthreads();  // the rolled-back row before any wavefront reads it
  const double t0 = in_vgpr(a.t0), t1 = in_vgpr(a.t1);
  const double o0 = in_vgpr(a.dep0[0]), o1 = in_vgpr(a.dep0[1]), o2 = in_vgpr(a.dep0[2]);
  const double span0 = in_vgpr(a.depf[0] - a.dep0[0]);
  const double span1 = in_vgpr(a.depf[1] - a.dep0[1]);
  const double span2 = in_vgpr(a.depf[2] - a.dep0[2]);
  const double tint = a.t1 - a.t0;
  const DivK kint{in_vgpr(tint), in_vgpr(1.0 / tint)};
  double dV = o0, dCai = o1, dNSR = o2;
  bool lopen = ch_lcc_open(w);
  int nopen = ch_ryr_open(w);
  const R k23c = in_vgpr((R)a.k23c), gcf = in_vgpr((R)a.gcf);
  double t = on ? t0 : t1;
  int bad = 0, nguard = 0;
  const int tl = threadIdx.x;
  if (role == kWsRA || role == kWsRB) {
    rtot[3][tl] = R(k43 + k45);
    rtot[5][tl] = R(k65);
  }
  if (role == kWsL) {
    lom[0][tl] = R(omega);
    lom[1][tl] = R(omega / bL);
    lom[2][tl] = R(omega / bL2);
    lom[3][tl] = R(omega / bL3);
    lom[4][tl] = R(omega / bL4);
  }
  // the role's words of the next step, fetched a step ahead
  // (L: words 0, 1, 7 of the cleft; RA: 2, 3; RB: 4, 5, 6)
  MtPart<3> winL{};
  MtPart<2> winA{};
  MtPart<3> winB{};
  if (t < t1) {
    if (role == kWsL) mt_fetch_part<0, 1, 7>(row, idx + 8 * c, winL);
    if (role == kWsRA) mt_fetch_part<2, 3>(row, idx + 8 * c, winA);
    if (role == kWsRB) mt_fetch_part<4, 5, 6>(row, idx + 8 * c, winB);
  }
  // A role's twists read words that another role regenerates in the same step
  // (word k + 1 of the cleft, or the next cleft's word 0), so no role stores
  // its words before barrier 1, by which every role has used its window.  The
  // stores after it touch none of the words that the windows of the next step,
  // fetched at the same time, read (positions idx .. idx+32 and words at least
  // 195 positions older than idx).
  Ghk<R> gv{};
  if (role == kWsH) gv = ghk_terms<R>((R)dV);
  while (t < t1) {
    const R V = (R)dV;
    const R cc = (R)css;
    // ---- merged-state draws (fru_rates_local.c:452,476,581): every role finds
    // the counts from the step-start state it holds, so all agree on whether
    // any lane draws (the condition below is the same in every wavefront)
    const int regime = ryr_regime<R>(cc);
    uint32_t rbits = ryr_resolve_bits(3);
    rbits = (regime == 2) ? ryr_resolve_bits(2) : rbits;
    rbits = (regime == 1) ? ryr_resolve_bits(1) : rbits;
    int st[kNryr];
    int nd = 0;
    unsigned need = 0;
#pragma unroll
    for (int r = 0; r < kNryr; ++r) {
      const uint32_t b = (rbits >> (4u * ((w >> (5 + 3 * r)) & 7u))) & 15u;
      st[r] = (int)(b & 7u) + 1;
      nd += (int)(b >> 3);
      need |= (b >> 3) << r;
    }
    const int n0 = qperm<QB0>(nd), n1 = qperm<QB1>(nd), n2 = qperm<QB2>(nd), n3 = qperm<QB3>(nd);
    const int E = n0 + n1 + n2 + n3;
    const int lazy0 = idx > kMtN ? idx : kMtN;
    RyrK<R> k{};
    R T7 = R(0.0), a7 = R(0.0), a8 = R(0.0);
    if (role == kWsRA || role == kWsRB) {
      k = ryr_k(cc, k23c);
      const R T2 = (R(k21) + k.k23 + k.k25);
      const R T3 = (R(k32) + k.k34);
      const R T5 = (R(k52) + k.k54 + k.k56);
      a7 = rdiv(k.k34, R(k43) + k.k34);
      T7 = ((R(1.0) - a7) * R(k32)) + (a7 * R(k45));
      a8 = rdiv(R(k65), k.k56 + R(k65));
      const R T8 = (a8 * R(k52)) + (a8 * k.k54);
      rtot[0][tl] = k.k12;
      rtot[1][tl] = T2;
      rtot[2][tl] = T3;
      rtot[4][tl] = T5;
      rtot[6][tl] = T7;
      rtot[7][tl] = T8;
    }
    if (__any(nd != 0)) {  // rare: role RA draws for the cleft
      if (role == kWsRA && nd) {
        const RyrK<R> kk = k;
        const int Ec = (c > 0 ? n0 : 0) + (c > 1 ? n1 : 0) + (c > 2 ? n2 : 0);
        uint32_t vw[kNryr];
        mt_run<kNryr>(row, idx + Ec, nd, vw);
#pragma unroll
        for (int r = 0; r < kNryr; ++r) {
          if ((need >> r) & 1u) {
            st[r] = ryr_split(st[r], unif(vw[0]), kk);
#pragma unroll
            for (int j = 0; j < kNryr - 1; ++j) vw[j] = vw[j + 1];
          }
        }
        uint32_t res = 0;
#pragma unroll
        for (int r = 0; r < kNryr; ++r) res |= (uint32_t)(st[r] - 1) << (5 + 3 * r);
        x_res[lane] = res;
      }
      __syncthreads();  // the drawn states, and the words stored for them
      if (nd) {
#pragma unroll
        for (int r = 0; r < kNryr; ++r) st[r] = (int)((x_res[lane] >> (5 + 3 * r)) & 7u) + 1;
      }
      // the step's own words follow the merged-state draws
      if (E != 0) {
        if (role == kWsL) mt_fetch_part<0, 1, 7>(row, idx + E + 8 * c, winL);
        if (role == kWsRA) mt_fetch_part<2, 3>(row, idx + E + 8 * c, winA);
        if (role == kWsRB) mt_fetch_part<4, 5, 6>(row, idx + E + 8 * c, winB);
      }
    }
#pragma unroll
    for (int r = 0; r < kNryr; ++r) w = ch_set_ryr(w, r, st[r]);
    // ---- each role's rates and words of the step
    R lb = R(0.0), lf = R(0.0), lm = R(0.0), lt = R(0.0), vr = R(0.0), ir = R(0.0);
    int ls = 1;
    R tr[kNryr] = {};
    uint32_t dw[3] = {0u, 0u, 0u};
    R djs1 = R(0.0), dss1 = R(0.0);
    R mpart = R(0.0);
    if (role == kWsL) {
      const R vm35 = V - R(35.0);
      const R e0 = R(0.012) * vm35, e1 = R(-0.05) * vm35;
      const R e2 = rdiv(V + R(12.5), divk(5.0)), e3 = rdiv(V + R(30.0), divk(12.0));
      R earg = e3;
      earg = (c == 2) ? e2 : earg;
      earg = (c == 1) ? e1 : earg;
      earg = (c == 0) ? e0 : earg;
      const R ex = xexp(earg);
      const R alpha = R(alphacf) * qperm<QB0>(ex);
      const R beta = R(betacf) * qperm<QB1>(ex);
      const R yinf = rdiv(R(yCa_frac), R(1.0) + qperm<QB2>(ex)) + (R(1.0) - R(yCa_frac));
      const R tau = R(60.0) + rdiv(R(340.0), R(1.0) + qperm<QB3>(ex));
      const R bp = rdiv(beta, divk(bL * 1.0));
      ls = ch_lcc(w);
      bad |= (ls < 1 || ls > 12);
      const int lk = ls - 7;
      lcc_arith<R>(ls, alpha, beta, bp, gcf * cc, lom[(lk >= 0 && lk <= 4) ? lk : 0][tl], lb, lf, lm, lt);
      vr = rdiv(ch_y(w) == 2 ? (R(1.0) - yinf) : yinf, tau);
      ir = (ch_ito2(w) == 1) ? R(kfIto2) * cc : R(kbIto2);
      mpart = vmax(vmax(vmax(R(0.0), lt), vr), ir);
      mt_twist_part<0, 1, 7>(idx + E + 8 * c, winL, dw);
    } else if (role == kWsRA) {
      tr[0] = rtot[st[0] - 1][tl];
      tr[1] = rtot[st[1] - 1][tl];
      mpart = vmax(tr[0], tr[1]);
      mt_twist_part<2, 3>(idx + E + 8 * c, winA, dw);
    } else if (role == kWsRB) {
      tr[2] = rtot[st[2] - 1][tl];
      tr[3] = rtot[st[3] - 1][tl];
      tr[4] = rtot[st[4] - 1][tl];
      mpart = vmax(vmax(tr[2], tr[3]), tr[4]);
      mt_twist_part<4, 5, 6>(idx + E + 8 * c, winB, dw);
    } else {  // H: the first evaluation of the Heun step needs no step length
      fcn_fru4<R>((R)cj, cc, gv, (R)dCai, (R)dNSR, lopen, nopen, djs1, dss1);
    }
    if (role != kWsH) {
      x_max[role][lane] = (double)mpart;
      x_bad[role][lane] = bad;
    }
    __syncthreads();  // barrier 1: the rates of every role
    // ---- step length (simfru_local.c:135-146), the same in every wavefront
    R maxr = vmax(vmax((R)x_max[kWsL][lane], (R)x_max[kWsRA][lane]), (R)x_max[kWsRB][lane]);
    bad = x_bad[kWsL][lane] | x_bad[kWsRA][lane] | x_bad[kWsRB][lane];
    maxr = vmax(maxr, qperm<QX1>(maxr));
    maxr = vmax(maxr, qperm<QX2>(maxr));
    bad |= qperm<QX1>(bad);
    bad |= qperm<QX2>(bad);
    double dt = 0.1 / (double)maxr;
    dt = vmin(dt, 10.001e-3);
    double cm = vmax(css, qperm<QX1>(css));
    cm = vmax(cm, qperm<QX2>(cm));
    const double cmax = vmax(0.e-3, cm);
    if (cmax > 1.5e-3) dt = vmin(dt, 5.001e-3);
    dt = vmin(dt, t1 - t);
    if (bad || !(t + dt > t) || ++nguard > kStepCap) {
      bad = 1;
      break;
    }
    const int gs = idx + E + 8 * c;  // this step's words
    {
      const int idn = idx + E + 32;
      const int adv = idn - lazy0;
      nreg += (adv > 0 ? adv : 0);
      idx = (idn >= 2 * kMtN) ? idn - kMtN : idn;
    }
    // ---- decisions, moves, the Heun step; the next step's words fetched early
    double cj_n = cj, css_n = css;
    if (role == kWsL) {
      mt_fetch_part<0, 1, 7>(row, idx + 8 * c, winL);
      mt_store_part<0, 1, 7>(row, gs, winL);
      const double u0 = unif(dw[0]);
      const bool mv_lcc = u0 > 1.0 - dt * (double)lt;
      const bool mv_y = unif(dw[1]) >= 1.0 - dt * (double)vr;
      const bool mv_ito = unif(dw[2]) >= 1.0 - dt * (double)ir;
      if (mv_lcc) w = ch_set_lcc(w, lcc_move<R>(ls, u0, dt, lb, lf, lm, lt));
      if (mv_y) w = ch_flip_y(w);
      if (mv_ito) w = ch_flip_ito2(w);
    } else if (role == kWsRA || role == kWsRB) {
      if (role == kWsRA) {
        mt_fetch_part<2, 3>(row, idx + 8 * c, winA);
        mt_store_part<2, 3>(row, gs, winA);
      } else {
        mt_fetch_part<4, 5, 6>(row, idx + 8 * c, winB);
        mt_store_part<4, 5, 6>(row, gs, winB);
      }
      const int r0 = role == kWsRA ? 0 : 2, nr = role == kWsRA ? 2 : 3;
      double ur[3];
      unsigned mm = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        ur[q] = unif(dw[q]);
        const R trq = role == kWsRA ? (q == 0 ? tr[0] : tr[1]) : tr[2 + q];
        if (q < nr && ur[q] > 1.0 - dt * (double)trq) mm |= 1u << q;
      }
      while (mm) {
        const int q = __builtin_ctz(mm);
        mm &= mm - 1u;
        const int r = r0 + q;
        int sr = st[0];
        double ru = ur[0];
#pragma unroll
        for (int qq = 1; qq < kNryr; ++qq) sr = (r == qq) ? st[qq] : sr;
#pragma unroll
        for (int qq = 1; qq < 3; ++qq) ru = (q == qq) ? ur[qq] : ru;
        Exits<R> e;
        ryr_exits_sel(sr, k, a7, a8, rtot[sr - 1][tl], e);
        w = ch_set_ryr(w, r, pick<R>(ru, dt, e));
      }
    }
    const double tn = t + dt;
    const double frac = rdiv(tn - t0, kint);
    if (role == kWsH) {
      const double k1j = dt * (double)djs1, k1s = dt * (double)dss1;
      const double y1j = cj + k1j, y1s = css + k1s;
      const double dVn = o0 + frac * span0, dCain = o1 + frac * span1, dNSRn = o2 + frac * span2;
      gv = ghk_terms<R>((R)dVn);
      R djs, dss;
      fcn_fru4<R>((R)y1j, (R)y1s, gv, (R)dCain, (R)dNSRn, lopen, nopen, djs, dss);
      cj_n = cj + (k1j + dt * (double)djs) / 2.0;
      css_n = css + (k1s + dt * (double)dss) / 2.0;
      x_cj[lane] = cj_n;
      x_css[lane] = css_n;
      steps += (c == 0);
    } else {
      x_w[role][lane] = w;
    }
    t = tn;
    dV = o0 + frac * span0;
    dCai = o1 + frac * span1;
    dNSR = o2 + frac * span2;
    __syncthreads();  // barrier 2: the new state of every role
    cj = x_cj[lane];
    css = x_css[lane];
    w = (x_w[kWsL][lane] & kWsMaskL) | (x_w[kWsRA][lane] & kWsMaskRA) | (x_w[kWsRB][lane] & kWsMaskRB);
    lopen = ch_lcc_open(w);
    nopen = ch_ryr_open(w);
  }
  if (on) {
    if (bad && role == kWsH && c == 0) atomicOr(d.err, 1);
    if (role == kWsH) {
      d.conc[(c + 1) * a.ld + f] = css;
      d.chan[c * a.ld + f] = w;
      if (c == 0) d.conc[f] = cj;
      acc[0] = css;
      acc[1] = (c == 0) ? cj : 0.0;
      acc[2] = lopen ? css : 0.0;
      acc[3] = lopen ? 1.0 : 0.0;
      acc[4] = (ch_ito2(w) == 2) ? 1.0 : 0.0;
      acc[5] = JRyRmax * ((double)nopen) * (cj - css);
      if (a.avg) {
        int nr = 0;
#pragma unroll
        for (int r = 0; r < kNryr; ++r) {
          const int s = ch_ryr(w, r);
          nr += (s == 1 || s == 2);
        }
        acc[6] = css;
        acc[7] = (c == 0) ? cj : 0.0;
        acc[8] = JRyRmax * ((double)nopen) * (cj - css);
        acc[9] = (double)nopen;
        acc[10] = (double)nr;
        acc[11] = css * (1.0 + BSRtot / (KBSR + css) + BSLtot / (KBSL + css));
        acc[12] = (c == 0) ? cj * (1.0 + CSQNtot / (KmCSQN + cj)) : 0.0;
        acc[13] = (ch_lcc(w) <= 6) ? 1.0 : 0.0;
        acc[14] = (ch_y(w) == 2) ? 1.0 : 0.0;
        acc[15] = lopen ? 1.0 : 0.0;
        acc[16] = (ch_ito2(w) == 2) ? 1.0 : 0.0;
      }
    }
    if (role == kWsL && c == 0) {
      d.mtidx[f] = idx;
      d.nreg[f] = nreg;
    }
  }
  return steps;
}

// The flux sums of a cell from the partials of its four-lane blocks' wavefronts
// (slot 4 b + w for wavefront w of block b): each block's four added in
// wavefront order, then the blocks in the order of reduce_range's last block
// -- the same numbers as the four-lane kernel's block partials give.  The
// whole block calls it; the totals are in tot_sh afterwards.
template <int NV>
__device__ __forceinline__ void sum_wave_slots(const Dev& d, int b0, int nb, double* tot_sh) {
  __shared__ double wsum[4][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  for (int m = (int)threadIdx.x; m < nb; m += blockDim.x) {
    const size_t g = (size_t)(4 * (b0 + m)) * kSlot;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double p0 = __hip_atomic_load(&d.part[g + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const double p1 = __hip_atomic_load(&d.part[g + kSlot + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const double p2 = __hip_atomic_load(&d.part[g + 2 * kSlot + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const double p3 = __hip_atomic_load(&d.part[g + 3 * kSlot + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s[k] += ((p0 + p1) + p2) + p3;
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = wave_sum_xor(s[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) wsum[wave][k] = s[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; ++k) {
      double a = wsum[0][k];
      for (int w = 1; w < nw; ++w) a += wsum[w][k];
      tot_sh[k] = a;
    }
  }
  __syncthreads();
}

// The flux sums of a wave-specialised launch: the H wavefront's butterfly of
// each block's 16 FRUs into the block's slot, the last block (ticket) adding
// four consecutive slots in the order the four-lane kernel's block adds its
// four wavefronts, then those in that kernel's order -- the same numbers.
template <int NV>
__device__ __forceinline__ bool reduce_ws(double v[NV], const Dev& d, unsigned int* ticket, int blk, int b0, int nb,
                                          double* tot_sh, int extra = 0) {
  static_assert(NV <= kSlot, "block partials overflow their slot row");
  __shared__ bool last;
  const int lane = threadIdx.x & 63;
  // A block's arrival at the ticket follows the completed stores of all its
  // wavefronts (the fcn block's stage terms come from wavefronts 0 and 1).
  __syncthreads();
  if (ws_role() == kWsH) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum_xor(v[k]);
    if (lane == 0) {
      if (blk >= 0)
        for (int k = 0; k < NV; ++k)
          __hip_atomic_store(&d.part[(size_t)blk * kSlot + k], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = (t == (unsigned)(4 * nb + extra) - 1u);
    }
  }
  __syncthreads();
  if (!last) return false;
  sum_wave_slots<NV>(d, b0, nb, tot_sh);
  if (threadIdx.x == 0) *ticket = 0u;
  __syncthreads();
  return true;
}

// (a lane's count is at most kStepCap + 1, so the wavefront's sum fits 32 bits)
static_assert((unsigned long long)(kStepCap + 1) * 64ull < (1ull << 32), "step counts overflow");
__device__ __forceinline__ void count_steps(const Dev& d, unsigned steps) {
  const unsigned s = wave_u32<false>(steps);
  const unsigned m = wave_u32<true>(steps);
  if ((threadIdx.x & 63) == 0 && s) {
    atomicAdd(d.nsteps, (unsigned long long)s);
    atomicMax(d.maxsteps, m);
    atomicAdd(d.nsteps + kLaneUse, 16ull * m);  // a wavefront of four-lane quads holds 16 FRU slots
  }
}

// After the grid's last block has seen every block's atomics: fold the
// launch's largest FRU-step count into the running sum.
__device__ __forceinline__ void close_launch_stats(const Dev& d) {
  const unsigned m = atomicExch(d.maxsteps, 0u);
  atomicAdd(d.critsteps, (unsigned long long)m);
}

// ---- Partials collected by polling (single cells).  The ticket election
// (reduce_range) costs the launch's critical path three device-scope round
// trips after the last wavefront's loop: the wait for the partial stores to
// complete, the ticket's atomic, and the last block's loads of the partials
// (and a fourth for the stage terms the fcn block left in the control block).
// Here the Monte Carlo blocks store their partials without waiting, into slots
// that hold a signalling-NaN marker until written (no arithmetic produces
// one), and the cell's fcn block -- which runs the tail anyway -- loads the
// slots until no marker is left, re-arms them for the next launch and keeps
// its stage terms in LDS.  The blocks never wait for each other but the fcn
// block does, so the grid must hold it beside at least one Monte Carlo block
// (it does: one block per CU at most is the fcn block's, of a single cell);
// its wait is bounded and ends the run with an error if a slot never fills.
constexpr unsigned long long kSlotEmpty = 0x7ff4deadbeef0001ull;  // a signalling NaN
constexpr int kSlotSteps = kSlot - 2, kSlotMaxSteps = kSlot - 1;  // the block's FRU-step sum and maximum
constexpr int kPollSpin = 1 << 20;  // polls of one slot (each a device-scope round trip) before giving up

__global__ void k_arm_slots(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = __longlong_as_double((long long)kSlotEmpty);
}

// A Monte Carlo block's partials: the wavefront butterflies and the waves in
// order, as reduce_range forms them, and its FRU-step sum and maximum.
template <int NV>
__device__ __forceinline__ void publish_partials(double v[NV], unsigned steps, const Dev& d, int blk) {
  static_assert(NV <= kSlotSteps, "block partials overflow their slot row");
  __shared__ double wsum[4][NV];
  __shared__ unsigned wst[4][2];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum_xor(v[k]);
  const unsigned ss = wave_u32<false>(steps), sm = wave_u32<true>(steps);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) wsum[wave][k] = v[k];
    wst[wave][0] = ss;
    wst[wave][1] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* const slot = d.part2 + (size_t)blk * kSlot;
    for (int k = 0; k < NV; ++k) {
      double s = wsum[0][k];
      for (int w = 1; w < nw; ++w) s += wsum[w][k];
      __hip_atomic_store(slot + k, s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned tsum = 0u, tmax = 0u;
    for (int w = 0; w < nw; ++w) {
      tsum += wst[w][0];
      tmax = max(tmax, wst[w][1]);
    }
    __hip_atomic_store(slot + kSlotSteps, (double)tsum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(slot + kSlotMaxSteps, (double)tmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__device__ __forceinline__ double await_slot(double* p, bool& ok) {
  for (int n = 0; ok && n < kPollSpin; ++n) {
    const double v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((unsigned long long)__double_as_longlong(v) != kSlotEmpty) return v;
    __builtin_amdgcn_s_sleep(2);
  }
  ok = false;
  return 0.0;
}

// The cell's fcn block: the partials of blocks b0 .. b0+nb-1 in reduce_range's
// order (each thread its strided subset, the wavefront butterfly, the waves in
// order: the same numbers), the slots re-armed behind it, and the launch's
// step statistics.  Returns false (slots not all written in time) only in
// thread 0.
template <int NV>
__device__ __forceinline__ bool collect_partials(const Dev& d, int b0, int nb, double* tot_sh) {
  __shared__ double wsum[4][NV];
  __shared__ unsigned long long wst[4][2];
  __shared__ int bad_sh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if (threadIdx.x == 0) bad_sh = 0;
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  unsigned long long ssum = 0ull, smax = 0ull;
  bool ok = true;
  const double empty = __longlong_as_double((long long)kSlotEmpty);
  for (int b = b0 + (int)threadIdx.x; b < b0 + nb; b += blockDim.x) {
    double* const slot = d.part2 + (size_t)b * kSlot;
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += await_slot(slot + k, ok);
    ssum += (unsigned long long)await_slot(slot + kSlotSteps, ok);
    smax = max(smax, (unsigned long long)await_slot(slot + kSlotMaxSteps, ok));
#pragma unroll
    for (int k = 0; k < NV; ++k) __hip_atomic_store(slot + k, empty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(slot + kSlotSteps, empty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(slot + kSlotMaxSteps, empty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();  // bad_sh cleared
  if (!ok) bad_sh = 1;
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = wave_sum_xor(s[k]);
  // step statistics: 64-bit sums and maxima over the wavefront by lane shuffles
  for (int o = 32; o > 0; o >>= 1) {
    ssum += __shfl_xor(ssum, o);
    smax = max(smax, __shfl_xor(smax, o));
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) wsum[wave][k] = s[k];
    wst[wave][0] = ssum;
    wst[wave][1] = smax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; ++k) {
      double a = wsum[0][k];
      for (int w = 1; w < nw; ++w) a += wsum[w][k];
      tot_sh[k] = a;
    }
    unsigned long long ts = 0ull, tm = 0ull;
    for (int w = 0; w < nw; ++w) {
      ts += wst[w][0];
      tm = max(tm, wst[w][1]);
    }
    atomicAdd(d.nsteps, ts);
    atomicAdd(d.critsteps, tm);
  }
  __syncthreads();
  return bad_sh == 0;
}

template <class R, bool REDUCE>
__global__ void __launch_bounds__(256) k_simfru4(const SimArgs a, const Dev d) {
  double acc[17] = {};
  count_steps(d, simfru4_body<R>(a, d, acc, blockIdx.x));
  if (REDUCE) {
    double tot[5];
    if (reduce_grid<5>(acc, d, 0, tot)) {
      close_launch_stats(d);
      double o4[4];
      flux_currents(tot, a.fdep[0], a.fdep[1], a.fdep[2], a.nfru, a.scale, o4);
      for (int q = 0; q < 4; ++q) d.out[q] = o4[q];
      for (int q = 0; q < 5; ++q) d.out[4 + q] = tot[q];
    }
  }
}

// Checkpoint of FRU i (mpi_slave.c:192-226): concentrations, channels, the
// stream position and the one MT word the next regeneration overwrites.
__device__ __forceinline__ void fru_checkpoint(const SimArgs& a, const Dev& d, int i) {
  for (int k = 0; k < 5; ++k) d.conc_h[k * a.ld + i] = d.conc[k * a.ld + i];
  for (int c = 0; c < 4; ++c) d.chan_h[c * a.ld + i] = d.chan[c * a.ld + i];
  const int idx = d.mtidx[i];
  d.mtidx_h[i] = idx;
  d.mtw_h[i] = d.mt[(size_t)i * kMtStride + mt_first_pos(idx)];
  d.nreg[i] = 0;
}

// Rollback of FRU i to its checkpoint (mpi_slave.c:150-184).
__device__ __forceinline__ void fru_rollback(const SimArgs& a, const Dev& d, int i) {
  for (int k = 0; k < 5; ++k) d.conc[k * a.ld + i] = d.conc_h[k * a.ld + i];
  for (int c = 0; c < 4; ++c) d.chan[c * a.ld + i] = d.chan_h[c * a.ld + i];
  const int idx = d.mtidx_h[i];
  mt_rollback(d.mt + (size_t)i * kMtStride, idx, d.nreg[i], d.mtw_h[i]);
  d.mtidx[i] = idx;
  d.nreg[i] = 0;
}

// ===========================================================================
// The global integrator on the device.
//
// rk_pd54m.c:29-345 and the caf.c:150-251 output loop run in the tails of the
// Monte Carlo launches: the last block of every sub-interval kernel (after the
// flux reduction) evaluates fcn for that RK stage and prepares the next
// sub-interval, and the last stage decides accept / reject.  The host only
// enqueues attempts (one begin kernel and five stage kernels each) and
// synchronises once per batch; launches after the run is complete return at
// once.  One RkCtl per cell holds the integrator's state.
// Per-beat statistics kept by the integrator from its accepted steps (the
// reference writes only 1 ms samples, too coarse for APD90 or peak ICaL).
// Beat b of a paced cell starts at the stimulus t_s = shift + b * period; its
// accepted steps are those ending in [t_s, t_s + period).  Record layout:
//   0 t_s, 1 V at the t_s output sample (rest), 2 peak V, 3 its time - t_s,
//   4 APD90: time - t_s at which V, after the peak, first falls through
//     V_peak - 0.9 (V_peak - V_rest) (linear between accepted steps; NaN if it
//     does not within the beat, or if the peak is under 20 mV above rest),
//   5 peak Cai, 6 its time - t_s, 7 Cai at the t_s sample, 8 peak (most
//   negative) ICa of the FRU flux conversion, 9 CaNSR at the t_s sample (SR
//   load), 10 mean CaJSR at the t_s sample (JSR load, calc_fru_avg output 1),
//   11 accepted steps in the beat, 12 Ca entered through the LCCs and 13 Ca
//   released by the RyRs over the beat (mM of myoplasm, trapezoids over the
//   accepted steps, the first one from the step before t_s), 14 peak release
//   flux (mM/ms).  Under a voltage-clamp protocol the beats are its periods
//   (iv_shift / iv_period, vpp_shift / vpp_period or vclamp_shift /
//   vclamp_period, the first of these protocols that is on).
constexpr int kBeatStats = 15;
// The part of a stage's fcn that does not depend on the release units'
// fluxes, as the stage's fcn block leaves it for the tail: F[0..36] (the
// assembled components excepted), the FcnAsm terms, V and Istim; and the flux
// conversion's exp at the stage's end potential.
constexpr int kPreAsm = 37, kPreV = kPreAsm + kFcnAsm, kPreIstim = kPreV + 1, kPreE1 = kPreV + 2,
              kPre = kPreV + 3;

struct RkCtl {
  GrwProtocol prot;
  double errw[kNglobal];
  double state[kNglobal];
  double y[kNglobal];
  double F1[kNglobal];
  double k[6][kNglobal];
  double cur[kNcur];
  double J[4];
  double dep0[3], depf[3];
  double start, h, oh, tcall, tstep, ta, tb;
  int stage, success, keepc, resume;
  int pro;  // per-FRU work of the next first stage: 1 checkpoint, 2 rollback
  int active, ms_done, ms_target, ms_base, avg_pending, avg_row;  // rows count from ms_base
  int log_n, log_cap, err;
  long long attempts, rejects, forced;
  double* out_states;  // [ms_target - ms_base][37] or null
  double* out_cur;     // [ms_target - ms_base][24] or null
  double* out_other;   // [ms_target - ms_base][12] or null
  double* out_log;     // [log_cap][4]: accepted steps (t, V, Cai, ICa)
  double* out_beats;   // [beat_cap][kBeatStats] or null
  int beat_cap, beat_n, beat_cur, pend_row;
  double bt[kBeatStats];
  double prev_t, prev_V, prev_jin, prev_jrel;
  double pend_t, pend_V, pend_Cai, pend_NSR, pend_JSR;  // the last output sample
  double pre[kPre];  // this stage's fcn without the assembly, from the launch's fcn block
};

// The beat-tracking fields of a control block in registers, for beat_update:
// loaded together instead of one dependent load after another (beat_load),
// the changed ones written back afterwards (beat_store).
struct BeatLocal {
  double* out_beats;
  double shift, period;  // the beats' origin and length (pacing or clamp)
  int beat_cap, beat_n, beat_cur;
  double bt[kBeatStats];
  double prev_t, prev_V, prev_jin, prev_jrel;
  double pend_t, pend_V, pend_Cai, pend_NSR, pend_JSR;
};

__device__ __forceinline__ BeatLocal beat_load(const RkCtl* c) {
  BeatLocal b;
  b.out_beats = c->out_beats;
  // beats are the pacing periods, or the periods of the clamp protocol in force
  const GrwProtocol& pr = c->prot;
  b.shift = pr.ivclamp ? pr.iv_shift : pr.vppclamp ? pr.vpp_shift : pr.vclamp ? pr.vclamp_shift : pr.shift;
  b.period = pr.ivclamp ? pr.iv_period : pr.vppclamp ? pr.vpp_period : pr.vclamp ? pr.vclamp_period : pr.period;
  b.beat_cap = c->beat_cap;
  b.beat_n = c->beat_n;
  b.beat_cur = c->beat_cur;
  for (int q = 0; q < kBeatStats; ++q) b.bt[q] = c->bt[q];
  b.prev_t = c->prev_t;
  b.prev_V = c->prev_V;
  b.prev_jin = c->prev_jin;
  b.prev_jrel = c->prev_jrel;
  b.pend_t = c->pend_t;
  b.pend_V = c->pend_V;
  b.pend_Cai = c->pend_Cai;
  b.pend_NSR = c->pend_NSR;
  b.pend_JSR = c->pend_JSR;
  return b;
}

__device__ __forceinline__ void beat_store(RkCtl* c, const BeatLocal& b) {
  c->beat_n = b.beat_n;
  c->beat_cur = b.beat_cur;
  for (int q = 0; q < kBeatStats; ++q) c->bt[q] = b.bt[q];
  c->prev_t = b.prev_t;
  c->prev_V = b.prev_V;
  c->prev_jin = b.prev_jin;
  c->prev_jrel = b.prev_jrel;
}

template <class B>
__device__ __forceinline__ void beat_flush(B* c) {
  if (c->beat_cur >= 0 && c->beat_n < c->beat_cap) {
    for (int q = 0; q < kBeatStats; ++q) c->out_beats[(size_t)c->beat_n * kBeatStats + q] = c->bt[q];
    c->beat_n++;
  }
  c->beat_cur = -1;
}

// The output sample at the beat's stimulus time supplies its rest values; it
// may be taken before or after the beat's first accepted step.
template <class B>
__device__ __forceinline__ void beat_take_sample(B* c) {
  if (c->beat_cur >= 0 && c->pend_t == c->bt[0]) {
    c->bt[1] = c->pend_V;
    c->bt[7] = c->pend_Cai;
    c->bt[9] = c->pend_NSR;
    c->bt[10] = c->pend_JSR;
  }
}

__device__ __forceinline__ void beat_update(BeatLocal* c, double t, double V, double Cai, double ICa, double jin,
                                            double jrel) {
  if (!c->out_beats) return;
  const double shift = c->shift;
  const double period = c->period;
  const int b = (t >= shift) ? (int)floor((t - shift) / period) : -1;
  const bool seg = c->prev_t == c->prev_t;  // a previous accepted step exists
  const double din = seg ? 0.5 * (c->prev_jin + jin) * (t - c->prev_t) : 0.0;
  const double drel = seg ? 0.5 * (c->prev_jrel + jrel) * (t - c->prev_t) : 0.0;
  if (b != c->beat_cur) {
    beat_flush(c);
    c->beat_cur = b;
    if (b >= 0) {
      double* bt = c->bt;
      bt[0] = shift + b * period;
      bt[1] = bt[4] = bt[7] = bt[9] = bt[10] = __builtin_nan("");
      bt[2] = V;
      bt[3] = t - bt[0];
      bt[5] = Cai;
      bt[6] = t - bt[0];
      bt[8] = ICa;
      bt[11] = 1.0;
      bt[12] = din;
      bt[13] = drel;
      bt[14] = jrel;
      beat_take_sample(c);
    }
  } else if (b >= 0) {
    double* bt = c->bt;
    if (V > bt[2]) {
      bt[2] = V;
      bt[3] = t - bt[0];
    }
    if (Cai > bt[5]) {
      bt[5] = Cai;
      bt[6] = t - bt[0];
    }
    if (ICa < bt[8]) bt[8] = ICa;
    bt[11] += 1.0;
    bt[12] += din;
    bt[13] += drel;
    if (jrel > bt[14]) bt[14] = jrel;
    if (bt[4] != bt[4] && bt[1] == bt[1] && bt[2] - bt[1] > 20.0 && t - bt[0] > bt[3]) {
      const double thr = bt[2] - 0.9 * (bt[2] - bt[1]);
      if (c->prev_V >= thr && V < thr)
        bt[4] = c->prev_t + (thr - c->prev_V) * (t - c->prev_t) / (V - c->prev_V) - bt[0];
    }
  }
  c->prev_t = t;
  c->prev_V = V;
  c->prev_jin = jin;
  c->prev_jrel = jrel;
}

// Stage s (1..5) sub-interval bounds, rk_pd54m.c:124,145-146,166-167,187-188,208-209.
__device__ __forceinline__ void stage_bounds(int s, double start, double h, double& ta, double& tb) {
  // (divisions by the constants through their reciprocals: rdiv(x, divk(c)))
  const double h5 = rdiv(h, divk(5.0)), h310 = rdiv(h * 3.0, divk(10.0));
  const double h35 = rdiv(h * 3.0, divk(5.0)), h23 = rdiv(h * 2.0, divk(3.0));
  switch (s) {
    case 1: ta = start; tb = start + h5; break;
    case 2: ta = start + h5; tb = start + h310; break;
    case 3: ta = start + h310; tb = start + h35; break;
    case 4: ta = start + h35; tb = start + h23; break;
    default: ta = start + h23; tb = start + h; break;
  }
}

// fcn (fcn.c:29-1072) for the integrator tails, by a whole block of 256
// threads.  First the transcendental calls, one per lane and one call kind per
// wavefront side by side (a wavefront runs the arms its lanes take one after
// the other): the exps on wavefront 0, which also evaluates fcn_voltage for
// the groups; the logs on wavefront 1, then there the IK1 factor's exp, whose
// argument needs EK; the pows on wavefront 2; wavefront 3 is free for the
// caller's work.  Then the independent groups of fcn_finish, balanced over
// the four wavefronts: the channel currents and xKs (thread 0), the two Ito1
// lattices side by side (threads 64 and 65), the INa gates and the HERG chain
// (thread 128), the Ca handling and the pumps (thread 192); thread 0
// assembles the derivatives.  The values are those of the one-thread fcn.
// Call from every thread; J and cur are read / written by thread 0 only, st is
// read by all and written by thread 0.
// fcn_block without the assembly, which alone needs the release units' fluxes:
// F except the assembled components, and in thread 0 the currents c, the Ca
// terms ca, the membrane potential V and the stimulus Istim; *mset as
// fcn_gates leaves it.  Ends behind a barrier.
__device__ __forceinline__ void fcn_block_groups(double t, const double* st, double* F, const GrwProtocol& prot,
                                                 FcnTrans* q, FcnCurrents& c, FcnCa& ca, double& V,
                                                 double& Istim, double* mset) {
  const int j = threadIdx.x;
  // the protocol fields in registers from the start: read through the
  // reference after the caller's stores to its control block, they would be
  // loaded again behind every barrier
  const GrwProtocol p = prot;
  if (j < 64) {
    double Vv, Is;
    fcn_voltage(t, st, p, Vv, Is);
    if (j == 0) {
      q->V = Vv;
      q->Istim = Is;
    }
    if (j < kFcnExp) {
      const double V = fcn_prep_V(t, st, p);
      q->ex[j] = exp(fcn_ex_arg(j, V, rdiv(V, divk(RT_over_F))));
    }
  } else if (j < 64 + kFcnLog) {
    double Vv, Is;
    fcn_voltage(t, st, p, Vv, Is);
    const double lg = log(fcn_lg_arg(j - 64, st));
    q->lg[j - 64] = lg;
    if (j == 65) q->k1x = exp(fcn_k1_arg(Vv, lg));
  } else if (j >= 128 && j < 128 + kFcnPow) {
    double b, e;
    fcn_pw_arg(j - 128, st, b, e);
    q->pw[j - 128] = pow(b, e);
  }
  __syncthreads();
  // fcn_finish with its four independent groups on the block's four wavefronts
  __shared__ FcnCa ca_sh;
  __shared__ double pump_sh[4];
  V = 0.0;
  Istim = 0.0;
  if (j == 0) {
    V = q->V;
    Istim = q->Istim;
    fcn_currents(st, V, *q, q->k1x, c);
    fcn_xks(st, V, *q, F);
  } else if (j == 64 || j == 65) {
    fcn_ito_lattice(st, *q, j - 64, F);
  } else if (j == 128) {
    fcn_gates(st, q->V, *q, F, mset);
  } else if (j == 192) {
    fcn_ca(st, p, *q, F, ca_sh);
    FcnCurrents cp;
    fcn_pumps(st, q->V, *q, cp);
    pump_sh[0] = cp.INaK;
    pump_sh[1] = cp.INaCa;
    pump_sh[2] = cp.ICab;
    pump_sh[3] = cp.IpCa;
  }
  __syncthreads();
  if (j == 0) {
    c.INaK = pump_sh[0];
    c.INaCa = pump_sh[1];
    c.ICab = pump_sh[2];
    c.IpCa = pump_sh[3];
    ca = ca_sh;
  }
}

__device__ __forceinline__ void fcn_block(double t, double* st, double* F, double* cur, bool keepc, const double* J,
                          const GrwProtocol& prot, FcnTrans* q) {
  __shared__ double mset_sh;
  FcnCurrents c;
  FcnCa ca;
  double V, Istim;
  const GrwProtocol p = prot;  // (loaded before the barriers)
  fcn_block_groups(t, st, F, p, q, c, ca, V, Istim, &mset_sh);
  if (threadIdx.x == 0) {
    if (mset_sh >= 0.0) st[1] = mset_sh;
    fcn_assemble(st, V, Istim, c, ca, F, cur, keepc, J, p);
  }
  __syncthreads();
}

// The start of an RK attempt, in the cell's last block (rk_pd54m.c:93-138):
// the averages of a finished output sample (av: the 11 sums of
// calc_fru_avg_local.c over the FRU state), on success fcn at the attempt start
// with the fluxes of the FRU sums fl, then k0, the first stage's end point and
// its bounds.  Call from every thread of the block, after the control block's
// flags are final.
__device__ __forceinline__ void attempt_start(RkCtl* ctl, const CellInfo& ci, const double* fl,
                                              const double* av, double* st, double* F1, double* Jsh,
                                              double* cursh, FcnTrans* fq) {
  const int j = threadIdx.x;
  const int active = ctl->active, avgp = ctl->avg_pending;
  const int resume = ctl->resume, success = ctl->success;
  const bool eval = active && !resume && success;
  const bool keepc = ctl->keepc != 0;
  const double tstart = ctl->start;
  if (j < kNglobal) st[j] = ctl->state[j];
  __syncthreads();
  if (j == 0) {
    if (avgp) {
      double other[12];
      avg_outputs(av, st, ci.nfru, ci.scale, other);
      if (ctl->out_other)
        for (int q = 0; q < 12; ++q) ctl->out_other[(size_t)(ctl->avg_row - ctl->ms_base) * 12 + q] = other[q];
      if (ctl->avg_row == ctl->pend_row) {
        ctl->pend_JSR = other[1];
        beat_take_sample(ctl);
      }
      ctl->avg_pending = 0;
    }
  } else if (j == 192 && eval) {
    // beside the averages on thread 0 and fcn_block's transcendentals on
    // wavefronts 0-2; fcn_block reads J only after its first barrier
    flux_currents(fl, st[0], st[6], st[7], ci.nfru, ci.scale, Jsh);
  }
  if (eval) fcn_block(tstart, st, F1, cursh, keepc, Jsh, ctl->prot, fq);
  else __syncthreads();
  if (j == 0 && active) {
    if (resume) {
      ctl->resume = 0;
    } else if (success) {
      if (keepc)
        for (int q = 0; q < kNcur; ++q) ctl->cur[q] = cursh[q];
      ctl->keepc = 0;
      for (int q = 0; q < 4; ++q) ctl->J[q] = Jsh[q];
    }
    ctl->attempts++;
  }
  __syncthreads();
  if (!active) return;
  const double h = ctl->h, start = ctl->start;
  if (j < kNglobal) {
    if (success && !resume) {
      ctl->F1[j] = F1[j];
      ctl->state[j] = st[j];  // fcn may have set mNa / V (fcn.c:663-669, :811-814)
    } else {
      F1[j] = ctl->F1[j];
    }
    const double k0 = h * F1[j];
    ctl->k[0][j] = k0;
    ctl->y[j] = st[j] + rdiv(k0, divk(5.0));
  }
  __syncthreads();
  if (j == 0) {
    ctl->dep0[0] = st[0]; ctl->dep0[1] = st[6]; ctl->dep0[2] = st[7];
    ctl->depf[0] = ctl->y[0]; ctl->depf[1] = ctl->y[6]; ctl->depf[2] = ctl->y[7];
    double ta, tb;
    stage_bounds(1, start, h, ta, tb);
    ctl->ta = ta;
    ctl->tb = tb;
    ctl->stage = 1;
  }
}

// The Monte Carlo block a dispatched block runs.  A batch dispatches its blocks
// costliest first (k_order_blocks), so that the long ones do not start last
// and leave the launch waiting on them; a block's FRUs, its partial-sum slot
// and so every result are those of the block it runs, whatever the order.
__device__ __forceinline__ int mc_block(const SimArgs& base, const Dev& d, int first) {
  const int i = (int)blockIdx.x - first;
  return base.sorted ? d.border[i] : i;
}

// Attempt start of every cell at the beginning of a run (later attempts start
// in the tail of the previous attempt's last stage): a sub-interval of length
// zero through the stage kernel's per-FRU code, which rolls every FRU back or
// checkpoints it and sums the FRU state with the lanes, blocks and order of a
// stage tail, so that a run continued in pieces adds the same numbers as one
// unbroken run; then attempt_start in the cell's last block.
template <class R, bool WS>
__global__ void __launch_bounds__(256) k_dev_begin(const SimArgs base, const Dev d, RkCtl* ctls) {
  const int blk = mc_block(base, d, 0);
  const int cell = (WS || base.ncells == 1) ? 0 : d.smap[blk];  // (one cell: no dependent load)
  const CellInfo ci = d.cells[cell];
  RkCtl* ctl = ctls + cell;
  const int active = ctl->active, avgp = ctl->avg_pending;
  if (!active && !avgp) return;
  SimArgs a = base;
  a.t0 = a.t1 = ctl->start;
  a.dep0[0] = a.depf[0] = ctl->state[0];
  a.dep0[1] = a.depf[1] = ctl->state[6];
  a.dep0[2] = a.depf[2] = ctl->state[7];
  a.nfru = ci.nfru;
  a.f0 = ci.f0;
  a.b0 = ci.sb0;
  a.k23c = ci.k23c;
  a.gcf = ci.gcf;
  a.pro = !active ? 0 : ctl->resume ? 2 : ctl->success ? 1 : 0;  // mpi_slave.c:150-226
  a.avg = avgp;
  double acc[17] = {};
  __shared__ double tot[17];
  if (WS) {
    (void)simfru_ws<R>(a, d, acc, blk);
    if (!reduce_ws<17>(acc, d, d.ticket + cell, blk, ci.sb0, ci.snb, tot)) return;
  } else {
    (void)simfru4_body<R>(a, d, acc, blk);
    if (!reduce_range<17>(acc, d, d.ticket + cell, blk, ci.sb0, ci.snb, 0, tot)) return;
  }
  __shared__ double st[kNglobal], F1[kNglobal], Jsh[4], cursh[kNcur];
  __shared__ FcnTrans fq;
  if (threadIdx.x == 0) ctl->pro = 0;  // this launch did the per-FRU work
  attempt_start(ctl, ci, tot, tot + 6, st, F1, Jsh, cursh, &fq);
}

// One RK stage of every cell: the Monte Carlo sub-interval of its FRUs (four
// lanes per FRU, 64 FRUs per block, blocks map to cells through d.smap), the
// flux reduction, and in the cell's last block fcn for this stage and then
// the next stage or the step decision (rk_pd54m.c:124-333) and the caf.c
// output loop (caf.c:150-251).
template <class R, bool WS, bool TAIL = false, bool POLL = false>
__device__ __forceinline__ void dev_stage(const SimArgs& base, const Dev& d, RkCtl* ctls) {
  // blocks [0, ncells): each cell's fcn block, dispatched first; then the
  // Monte Carlo blocks
  const bool fblock = TAIL || (int)blockIdx.x < base.ncells;
  const int blk = fblock ? -1 : mc_block(base, d, base.ncells);
  const int cell = fblock ? (int)blockIdx.x : (WS || base.ncells == 1) ? 0 : d.smap[blk];  // (one cell: no dependent load)
  const CellInfo ci = d.cells[cell];
  RkCtl* ctl = ctls + cell;
  if (!ctl->active) return;
  SimArgs a = base;
  a.t0 = ctl->ta;
  a.t1 = ctl->tb;
  for (int q = 0; q < 3; ++q) {
    a.dep0[q] = ctl->dep0[q];
    a.depf[q] = ctl->depf[q];
  }
  a.nfru = ci.nfru;
  a.f0 = ci.f0;
  a.b0 = ci.sb0;
  a.k23c = ci.k23c;
  a.gcf = ci.gcf;
  const int s0 = ctl->stage;
  a.pro = (s0 == 1) ? ctl->pro : 0;
  a.avg = (s0 == 5);  // the attempt's end state: averages for an output sample
  double acc[17] = {};
  const int j = threadIdx.x;
  __shared__ double st[kNglobal], yv[kNglobal], F[kNglobal], kk[6][kNglobal];
  __shared__ FcnTrans fq;
  __shared__ double Jsh[4], pre[kPre];
  if (fblock) {
    // This stage's fcn at (t1, y) but for the assembly, the one part that
    // needs the fluxes of this launch's sub-interval (the stage's y is known
    // before it): computed beside the Monte Carlo blocks and left in the
    // control block for the tail, with agent-scope stores completed before
    // this block's arrival at the ticket.  (fcn_gates' reset of the m gate
    // below -90 mV applies to the stage's y, which nothing reads afterwards.)
    if (s0 >= 1 && s0 <= 5) {
      if (j < kNglobal) yv[j] = ctl->y[j];
      __syncthreads();
      __shared__ double mset_sh;
      FcnCurrents c;
      FcnCa ca;
      double V, Istim;
      fcn_block_groups(a.t1, yv, F, ctl->prot, &fq, c, ca, V, Istim, &mset_sh);
      if (POLL) {  // this block runs the tail: the terms stay in LDS (F in place)
        if (j == 0) {
          const FcnAsm m = fcn_asm_pre(c, ca, Istim, ctl->prot.chf != 0);
          const double mv[kFcnAsm] = {m.dNai, m.dKi, m.a3ca, m.mca, m.nsr, m.beta_i, m.Jup,
                                      m.Jtrpn, m.INa, m.IKr, m.IKs, m.p2, m.a3v};
          for (int q = 0; q < kFcnAsm; ++q) pre[kPreAsm + q] = mv[q];
          pre[kPreV] = V;
          pre[kPreIstim] = Istim;
        }
        if (j == 64) pre[kPreE1] = flux_exp(a.depf[0]);
      } else {
        double* pre_g = ctl->pre;
        if (j < kNglobal) __hip_atomic_store(pre_g + j, F[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (j == 0) {
          const FcnAsm m = fcn_asm_pre(c, ca, Istim, ctl->prot.chf != 0);
          const double mv[kFcnAsm] = {m.dNai, m.dKi, m.a3ca, m.mca, m.nsr, m.beta_i, m.Jup,
                                      m.Jtrpn, m.INa, m.IKr, m.IKs, m.p2, m.a3v};
          for (int q = 0; q < kFcnAsm; ++q)
            __hip_atomic_store(pre_g + kPreAsm + q, mv[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(pre_g + kPreV, V, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(pre_g + kPreIstim, Istim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (j == 64)
          __hip_atomic_store(pre_g + kPreE1, flux_exp(a.depf[0]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // complete the stores before this block arrives at the cell's ticket;
        // the compiler-only fences keep them on this side of the wait
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
      }
    }
  } else if (POLL) {  // publish this block's partials and leave: the fcn block collects them
    const unsigned steps = simfru4_body<R>(a, d, acc, blk);
    if (a.avg) publish_partials<17>(acc, steps, d, blk);
    else publish_partials<6>(acc, steps, d, blk);
    return;
  } else {
    count_steps(d, WS ? simfru_ws<R>(a, d, acc, blk) : simfru4_body<R>(a, d, acc, blk));
  }
  // the control block's state vectors for the tail, loaded by every block
  // before the reduction so that the last block does not wait for them after it
  double st_j = 0.0, y_j = 0.0, k_j[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, errw_j = 0.0;
  if (j < kNglobal && s0 >= 1 && s0 <= 5) {
    st_j = ctl->state[j];
    y_j = ctl->y[j];
    if (s0 == 5) errw_j = ctl->errw[j];
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (q < s0) k_j[q] = ctl->k[q][j];
  }
  __shared__ double tot[17];
  // the eleven average sums (tot[6..16]) exist only in the attempt's last stage;
  // the fcn block is one more arrival
  if (TAIL) {  // the Monte Carlo work ran in the queue kernel before this launch
    __syncthreads();  // the fcn terms' stores are complete in every wavefront
    if (a.avg)
      sum_wave_slots<17>(d, ci.sb0, ci.snb, tot);
    else
      sum_wave_slots<6>(d, ci.sb0, ci.snb, tot);
  } else if (POLL) {  // the fcn block: the Monte Carlo blocks' partials, as they arrive
    const bool okp = a.avg ? collect_partials<17>(d, ci.sb0, ci.snb, tot) : collect_partials<6>(d, ci.sb0, ci.snb, tot);
    if (!okp) {
      if (j == 0) {
        ctl->err = 1;
        ctl->active = 0;
      }
      return;
    }
  } else if (WS ? (a.avg ? !reduce_ws<17>(acc, d, d.ticket + cell, blk, ci.sb0, ci.snb, tot, 1)
                         : !reduce_ws<6>(acc, d, d.ticket + cell, blk, ci.sb0, ci.snb, tot, 1))
                : (a.avg ? !reduce_range<17>(acc, d, d.ticket + cell, blk, ci.sb0, ci.snb, 0, tot, 1)
                         : !reduce_range<6>(acc, d, d.ticket + cell, blk, ci.sb0, ci.snb, 0, tot, 1))) {
    return;
  }
  const unsigned long long tail0 = wall_clock64();
  // ---- last block of the cell: fcn of this stage and the next stage / the step decision
  __shared__ int decision;
  const int s = s0;
  if (s < 1 || s > 5) {  // launches out of sequence: stop the run
    if (j == 0) {
      ctl->err = 1;
      ctl->active = 0;
    }
    return;
  }
  const double h = ctl->h, start = ctl->start;
  // the protocol and the output step's end in registers before any barrier:
  // read through ctl after this tail's stores, they would be loaded again
  const GrwProtocol pl = ctl->prot;
  const double tstep = ctl->tstep;
  if (j < kNglobal) {
    st[j] = st_j;
    yv[j] = y_j;
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (q < s) kk[q][j] = k_j[q];
  }
  // the fcn block's part of this stage's fcn (agent-scope loads: written in
  // this launch), the flux conversion (mpi_master.c:401-468) on wavefront 3
  // and the launch's step statistics (a returning atomic) on wavefront 2
  if (!POLL && j < kPreE1) {
    const double v = __hip_atomic_load(ctl->pre + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (j < kNglobal) F[j] = v;
    else pre[j] = v;
  }
  if (!POLL && j == 128 && base.ncells == 1) close_launch_stats(d);
  if (j == 192) {
    const double e1 =
        POLL ? pre[kPreE1] : __hip_atomic_load(ctl->pre + kPreE1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    flux_currents_e(tot, a.depf[0], e1, a.depf[1], a.depf[2], ci.nfru, ci.scale, Jsh);
  }
  __syncthreads();
  if (j == 0) {  // fcn's assembly with this launch's fluxes
    const double* m = pre + kPreAsm;
    const FcnAsm a{m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11], m[12]};
    (void)fcn_asm_fin(a, yv, pre[kPreV], pre[kPreIstim], Jsh, clamped(pl), F);
    for (int q = 0; q < 4; ++q) ctl->J[q] = Jsh[q];
  }
  __syncthreads();
  if (j < kNglobal) {
    kk[s][j] = h * F[j];
    ctl->k[s][j] = kk[s][j];
  }
  if (s < 5) {  // next stage's y (rk_pd54m.c:138,159,180,201)
    if (j < kNglobal) {
      double y;
      if (s == 1) y = st[j] + (3.0 / 40.0 * kk[0][j] + 9.0 / 40.0 * kk[1][j]);
      else if (s == 2) y = st[j] + (3.0 / 10.0 * kk[0][j] - 9.0 / 10.0 * kk[1][j] + 6.0 / 5.0 * kk[2][j]);
      else if (s == 3)
        y = st[j] + 226.0 / 729.0 * kk[0][j] - 25.0 / 27.0 * kk[1][j] + 880.0 / 729.0 * kk[2][j] +
            55.0 / 729.0 * kk[3][j];
      else
        y = st[j] - 181.0 / 270.0 * kk[0][j] + 5.0 / 2.0 * kk[1][j] - 266.0 / 297.0 * kk[2][j] -
            91.0 / 27.0 * kk[3][j] + 189.0 / 55.0 * kk[4][j];
      yv[j] = y;
      ctl->y[j] = y;
    }
    __syncthreads();
    if (j == 0) {
      ctl->dep0[0] = st[0]; ctl->dep0[1] = st[6]; ctl->dep0[2] = st[7];
      ctl->depf[0] = yv[0]; ctl->depf[1] = yv[6]; ctl->depf[2] = yv[7];
      double ta, tb;
      stage_bounds(s + 1, start, h, ta, tb);
      ctl->ta = ta;
      ctl->tb = tb;
      ctl->stage = s + 1;
      if (s == 1) ctl->pro = 0;  // every FRU did its checkpoint / rollback
      atomicAdd(d.tailticks, wall_clock64() - tail0);
    }
    return;
  }
  // ---- truncation error and step control (rk_pd54m.c:225-333)
  // (thread 0 loads the beat tracking and the step log's position now, for
  // the accepted step's bookkeeping below)
  BeatLocal bl;
  int log_n = 0, log_cap = 0;
  double* out_log = nullptr;
  if (j == 0) {
    bl = beat_load(ctl);
    log_n = ctl->log_n;
    log_cap = ctl->log_cap;
    out_log = ctl->out_log;
  }
  const double huge_number = 1e99;
  __shared__ double tts_sh[2];
  double evj = 0.0;
  if (j < kNglobal) {
    const double y4 = st[j] + 31.0 / 540.0 * kk[0][j] + 190.0 / 297.0 * kk[2][j] - 145.0 / 108.0 * kk[3][j] +
                      351.0 / 220.0 * kk[4][j] + 1.0 / 20.0 * kk[5][j];
    const double ym = st[j] + 19.0 / 216.0 * kk[0][j] + 1000.0 / 2079.0 * kk[2][j] -
                      125.0 / 216.0 * kk[3][j] + 81.0 / 88.0 * kk[4][j] + 5.0 / 56.0 * kk[5][j];
    evj = fabs((ym - y4) * 0.2 * errw_j);
    yv[j] = y4;
  }
  // The error norm (rk_pd54m.c:225-237): the largest component, or
  // huge_number once any component is not below it (NaN included), which is
  // the sequential loop's result whatever the order; lanes past the state hold
  // 0, which changes no maximum of non-negative values.
  double err = 0.0;
  if (j < 64) err = wave_sum_xor<true>((evj < huge_number) ? evj : huge_number);
  // the stimulus alignment's remainders for either outcome (rk_pd54m.c:279,
  // :317), on wavefront 1 beside the norm
  if (j == 64) {
    tts_sh[0] = pl.shift - fmod(start + h, pl.period);
    tts_sh[1] = pl.shift - fmod(start, pl.period);
  }
  __syncthreads();
  if (j == 0) {
    const GrwProtocol& p = pl;
    const bool accept = err < p.tolrk;
    double hn = h;
    if (accept || h <= p.step_min) {
      decision = 1;
      const double ns = start + h;
      ctl->start = ns;
      ctl->success = 1;
      if (!accept) ctl->forced++;
      if (ns >= tstep) {
        decision = 2;  // this rk54pd call is complete
      } else if (accept) {
        hn = rmin(0.85 * h * pow(p.tolrk / err, 0.2), 4.0 * h);
        ctl->oh = hn;
        double tts = tts_sh[0];
        if (tts < 0.0) tts = tts + p.period;
        if (tts < hn) hn = rmax(p.step_min, tts - 0.5 * p.step_min);
      } else {
        hn = p.step_min;
      }
      if (!accept) {
        ctl->oh = hn;
        hn = rmin(p.step_max, rmax(p.step_min, hn));
        hn = rmin(hn, tstep - ns);
      }
    } else {
      decision = 0;
      ctl->rejects++;
      hn = 0.85 * h * pow(p.tolrk / err, 0.2);
      double tts = tts_sh[1];
      if (tts < 0.0) tts = tts + p.period;
      if (tts < hn) hn = rmax(p.step_min, tts - 0.5 * p.step_min);
      ctl->success = 0;
      ctl->resume = 1;
      hn = rmin(p.step_max, rmax(p.step_min, hn));
      hn = rmin(hn, tstep - start);
    }
    ctl->h = hn;
    ctl->stage = 0;
  }
  __syncthreads();
  if (decision != 0) {
  if (j < kNglobal) {
    double y = yv[j];
    if (fabs(y) < 1.e-10) y = 0.0;
    yv[j] = y;
    ctl->state[j] = y;
  }
  __syncthreads();
  if (j == 0) {
    const double tnew = start + h;  // the accepted step's end, ctl->start now
    if (log_n < log_cap) {
      double* L = out_log + 4 * (size_t)log_n;
      L[0] = tnew; L[1] = yv[0]; L[2] = yv[6]; L[3] = Jsh[2];
      ctl->log_n = log_n + 1;
    }
    // Ca entering through the LCCs and released by the RyRs, both per litre of
    // myoplasm (fcn.c:1058 and the scaling of send_calc_fru_avg, mpi_master.c:178)
    const double jin = rdiv(Jsh[2] * Acap, divk(-2.0 * Vmyo * Faraday * 1000.0));
    const double jrel = rdiv(tot[5] * ci.scale * VSS, divk(Vmyo));
    beat_update(&bl, tnew, yv[0], yv[6], Jsh[2], jin, jrel);
    beat_store(ctl, bl);
  }
  if (decision == 2) {
  // ---- the caf.c loop: record this call's outputs, set up the next call
  const int row = ctl->ms_done;
  const size_t orow = (size_t)(row - ctl->ms_base);
  if (ctl->out_states && j < kNglobal) ctl->out_states[orow * kNglobal + j] = yv[j];
  if (ctl->out_cur && j < kNcur) ctl->out_cur[orow * kNcur + j] = ctl->cur[j];
  __syncthreads();
  if (j == 0) {
    const GrwProtocol& p = pl;
    ctl->pend_t = tstep;
    ctl->pend_row = row;
    ctl->pend_V = yv[0];
    ctl->pend_Cai = yv[6];
    ctl->pend_NSR = yv[7];
    ctl->pend_JSR = __builtin_nan("");
    beat_take_sample(ctl);
    ctl->ms_done = row + 1;
    ctl->avg_pending = 1;
    ctl->avg_row = row;
    // the next call starts on the output grid (caf.c:150) with the step of
    // rk_pd54m.c:71-74; a later run continues from here
    const double tc = tstep;
    ctl->tcall = tc;
    ctl->tstep = tc + p.time_step;
    double hn = rmin(ctl->oh, ctl->tstep - tc);
    hn = rmin(p.step_max, hn);
    ctl->h = hn;
    ctl->oh = hn;
    ctl->start = tc;
    ctl->success = 1;
    ctl->keepc = 1;
    if (row + 1 >= ctl->ms_target) ctl->active = 0;
  }
  }
  }
  // ---- the next attempt starts here (rk_pd54m.c:93-138): its first stage
  // checkpoints every FRU after an accepted step and rolls it back after a
  // rejected one; the FRU sums of this launch are those of the state it starts from
  if (j == 0) ctl->pro = (decision == 0) ? 2 : 1;
  __syncthreads();
  __shared__ double F1[kNglobal], cursh[kNcur];
  attempt_start(ctl, ci, tot, tot + 6, st, F1, Jsh, cursh, &fq);
  if (j == 0) atomicAdd(d.tailticks, wall_clock64() - tail0);
}

template <class R, bool WS>
__global__ void __launch_bounds__(256) k_dev_stage(const SimArgs base, const Dev d, RkCtl* ctls) {
  dev_stage<R, WS>(base, d, ctls);
}

// The four-lane stage of a single cell whose fcn block collects the Monte
// Carlo blocks' partials by polling (publish_partials / collect_partials) and
// runs the tail: the same numbers, two to three fewer device-scope round trips
// on the launch's critical path.
template <class R>
__global__ void __launch_bounds__(256) k_dev_stage_poll(const SimArgs base, const Dev d, RkCtl* ctls) {
  dev_stage<R, false, false, true>(base, d, ctls);
}

// Queue layout (switched off: grw_set_layout refuses GRW_LAYOUT_QUEUE, since
// its one device run did not finish), second kernel of a stage: one block per cell runs the cell's
// fcn terms, adds the wavefront partials the queue kernel left and runs the
// integrator stage (the tail of k_dev_stage, unchanged).  Resets the queue for
// the next launch.
template <class R>
__global__ void __launch_bounds__(256) k_dev_tail(const SimArgs base, const Dev d, RkCtl* ctls) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *d.qhead = 0u;
  dev_stage<R, false, true>(base, d, ctls);
}

// Queue layout (switched off, see k_dev_tail), first kernel of a stage: the Monte Carlo sub-interval of every
// active cell, one work item per wavefront's worth (16 FRUs, wavefront w of a
// four-lane block), taken from a counter in dispatch order (the batch's
// costliest blocks first).  A wavefront that finishes its item takes the next
// one instead of waiting for the other wavefronts of its block; the grid is
// what the GPU holds at once.  Each item's partials go to slot 4 b + w, which
// k_dev_tail adds in the four-lane kernel's order: the same numbers.  No
// wavefront waits for another: the launch ends when the counter runs past the
// last item.
template <class R>
__global__ void __launch_bounds__(256) k_mc_queue(const SimArgs base, const Dev d, RkCtl* ctls, int nitems) {
  const int lane = threadIdx.x & 63;
  while (true) {
    unsigned item = 0;
    if (lane == 0) item = atomicAdd(d.qhead, 1u);
    item = __builtin_amdgcn_readfirstlane(__shfl(item, 0));
    if (item >= (unsigned)nitems) break;
    const int i = (int)(item >> 2), w = (int)(item & 3u);
    const int blk = base.sorted ? d.border[i] : i;
    const int cell = base.ncells == 1 ? 0 : d.smap[blk];
    const CellInfo ci = d.cells[cell];
    const RkCtl* ctl = ctls + cell;
    if (!ctl->active) continue;
    SimArgs a = base;
    a.t0 = ctl->ta;
    a.t1 = ctl->tb;
    for (int q = 0; q < 3; ++q) {
      a.dep0[q] = ctl->dep0[q];
      a.depf[q] = ctl->depf[q];
    }
    a.nfru = ci.nfru;
    a.f0 = ci.f0;
    a.b0 = ci.sb0;
    a.k23c = ci.k23c;
    a.gcf = ci.gcf;
    const int s0 = ctl->stage;
    a.pro = (s0 == 1) ? ctl->pro : 0;
    a.avg = (s0 == 5);
    double acc[17] = {};
    count_steps(d, simfru4_gid<R>(a, d, acc, (blk - a.b0) * 256 + w * 64 + lane));
    double* const slot = d.part + (size_t)(4 * blk + w) * kSlot;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double v = wave_sum_xor(acc[k]);
      if (lane == 0) slot[k] = v;
    }
    if (a.avg) {
#pragma unroll
      for (int k = 6; k < 17; ++k) {
        const double v = wave_sum_xor(acc[k]);
        if (lane == 0) slot[k] = v;
      }
    }
  }
}

// Single-wavefront blocks, first kernel of a stage: the Monte Carlo
// sub-interval of every active cell with one wavefront per block -- block i
// runs wavefront i & 3 of four-lane block i >> 2 (in the batch's dispatch
// order, costliest first), 16 FRUs.  A 256-thread block holds the registers of
// its four wavefronts until the slowest is done; here a wavefront's registers
// return to its SIMD as soon as its own FRUs are, and the next block starts
// there.  The partials of wavefront w of four-lane block b go to slot 4 b + w;
// k_dev_tail (one block per cell, after this kernel) adds them in the
// four-lane kernel's order, evaluates the stage's fcn and runs the integrator
// stage: the same numbers as GRW_LAYOUT_QUAD.  No block waits for another.
template <class R, bool VINV = true, bool LATE = false>
__device__ __forceinline__ void mc_wave(const SimArgs& base, const Dev& d, const RkCtl* ctls) {
  const int i = (int)blockIdx.x;
  const int blk = base.sorted ? d.border[i >> 2] : (i >> 2);
  const int w = i & 3;
  const int lane = threadIdx.x;
  const int cell = base.ncells == 1 ? 0 : d.smap[blk];
  const CellInfo ci = d.cells[cell];
  const RkCtl* ctl = ctls + cell;
  if (!ctl->active) return;
  SimArgs a = base;
  a.t0 = ctl->ta;
  a.t1 = ctl->tb;
  for (int q = 0; q < 3; ++q) {
    a.dep0[q] = ctl->dep0[q];
    a.depf[q] = ctl->depf[q];
  }
  a.nfru = ci.nfru;
  a.f0 = ci.f0;
  a.b0 = ci.sb0;
  a.k23c = ci.k23c;
  a.gcf = ci.gcf;
  const int s0 = ctl->stage;
  a.pro = (s0 == 1) ? ctl->pro : 0;
  a.avg = (s0 == 5);
  double acc[17] = {};
  count_steps(d, simfru4_gid<R, 64, VINV, LATE>(a, d, acc, (blk - a.b0) * 256 + w * 64 + lane));
  double* const slot = d.part + (size_t)(4 * blk + w) * kSlot;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double v = wave_sum_xor(acc[k]);
    if (lane == 0) slot[k] = v;
  }
  if (a.avg) {
#pragma unroll
    for (int k = 6; k < 17; ++k) {
      const double v = wave_sum_xor(acc[k]);
      if (lane == 0) slot[k] = v;
    }
  }
}

template <class R>
__global__ void __launch_bounds__(64) k_mc_wave(const SimArgs base, const Dev d, RkCtl* ctls) {
  mc_wave<R>(base, d, ctls);
}

// The same with its registers capped so that W wavefronts fit on a SIMD.
template <class R, int W>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) k_mc_wave_occ(const SimArgs base,
                                                                                       const Dev d, RkCtl* ctls) {
  mc_wave<R, false, true>(base, d, ctls);
}

// Refill blocks (GRW_LAYOUT_REFILL): one wavefront per four-lane
// block, whose 16 quads work through the block's 64 FRUs and take the next one
// as soon as their own is done.  In the one-wavefront layout a wavefront runs
// as long as the slowest of its 16 FRUs, and an FRU's micro-step count in the
// next sub-interval is poorly predicted by its last one (a channel opening
// raises CaSS and with it the rates), so most FRU lanes of a batch wait.  Here
// a quad finishing its FRU stores it and takes the next from the block's
// states, staged in LDS at the start: the block's slots are handed out
// costliest first (they ascend by cost class), in quad order among the quads
// free in the same micro-step.  Each FRU's arithmetic is the
// four-lane kernel's; the flux sums are formed afterwards from the final
// states kept in LDS, lane for lane as the one-wavefront layout's four
// wavefronts form them, into the same slots, so k_dev_tail adds the same
// numbers.  The loop ends when every quad is out of FRUs: each pass either
// advances every working quad's time (a micro-step, or a stop for a bad state
// or after kStepCap steps) or hands out slots, of which there are at most 64.
template <class R, bool LATE>
__device__ __forceinline__ void mc_refill(const SimArgs& base, const Dev& d, const RkCtl* ctls) {
  const int i = (int)blockIdx.x;
  const int blk = base.sorted ? d.border[i] : i;
  const int lane = threadIdx.x;
  const int q = lane >> 2, c = lane & 3;
  const int cell = base.ncells == 1 ? 0 : d.smap[blk];
  const CellInfo ci = d.cells[cell];
  const RkCtl* ctl = ctls + cell;
  if (!ctl->active) return;
  SimArgs a = base;
  a.t0 = ctl->ta;
  a.t1 = ctl->tb;
  for (int r = 0; r < 3; ++r) {
    a.dep0[r] = ctl->dep0[r];
    a.depf[r] = ctl->depf[r];
  }
  a.nfru = ci.nfru;
  a.f0 = ci.f0;
  a.b0 = ci.sb0;
  a.k23c = ci.k23c;
  a.gcf = ci.gcf;
  const int s0 = ctl->stage;
  a.pro = (s0 == 1) ? ctl->pro : 0;
  a.avg = (s0 == 5);
  const int first = (blk - a.b0) * 64;        // the block's first quad slot in the cell
  const int nslot = min(64, a.nfru - first);  // FRUs of the block
  const QuadStage<R> k = quad_stage<R, !LATE>(a, c, lane);
  __shared__ R rtot[8][64];
  __shared__ R lom[5][64];
  quad_tables<R, 64>(rtot, lom, lane);
  // Every slot's state, staged in LDS at the start so that a quad taking a
  // slot waits for no memory round trip (lane l loads slot l, with the
  // checkpoint or the rollback of mpi_slave.c:150-226); slot s's clefts at
  // entries 4 s + c.  A finished slot's entries then hold its final state,
  // entry 4 s + c as lane 4 (s mod 16) + c of the one-wavefront layout's
  // wavefront s / 16 holds it.
  __shared__ double scj[64], scss[256];
  __shared__ uint32_t sw[256];
  __shared__ int sidx[64], snreg[64], sf[64], sopen[256];  // sopen: lopen | nopen << 1 at the end
  if (lane < nslot) {
    const int fl = a.f0 + (a.sorted ? d.perm[a.f0 + first + lane] : first + lane);
    uint32_t* const row = d.mt + (size_t)fl * kMtStride;
    double cjl, cs[4];
    uint32_t wl[4];
    int idxl, nregl;
    if (a.pro == 2) {  // rollback (mpi_slave.c:150-184): from the checkpoint, MT words restored
      cjl = d.conc_h[fl];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cs[r] = d.conc_h[(r + 1) * a.ld + fl];
        wl[r] = d.chan_h[r * a.ld + fl];
      }
      idxl = d.mtidx_h[fl];
      nregl = 0;
      mt_rollback(row, idxl, d.nreg[fl], d.mtw_h[fl]);
    } else {
      cjl = d.conc[fl];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cs[r] = d.conc[(r + 1) * a.ld + fl];
        wl[r] = d.chan[r * a.ld + fl];
      }
      idxl = d.mtidx[fl];
      nregl = a.pro ? 0 : d.nreg[fl];
      if (a.pro == 1) {  // checkpoint (mpi_slave.c:192-226)
        d.conc_h[fl] = cjl;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          d.conc_h[(r + 1) * a.ld + fl] = cs[r];
          d.chan_h[r * a.ld + fl] = wl[r];
        }
        d.mtidx_h[fl] = idxl;
        d.mtw_h[fl] = row[mt_first_pos(idxl)];
      }
    }
    scj[lane] = cjl;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      scss[4 * lane + r] = cs[r];
      sw[4 * lane + r] = wl[r];
    }
    sidx[lane] = idxl;
    snreg[lane] = nregl;
    sf[lane] = fl;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // restored MT words seen by every quad
  __syncthreads();
  const Ghk<R> gv0 = ghk_terms<R>((R)k.o0);  // every FRU's first flux factors (V at the sub-interval start)
  QuadFru<R> s;
  int sl = nslot - 1 - q;  // this quad's slot (costliest first), < 0: none
  bool live = sl >= 0;
  int f = 0;
  auto take = [&](int slot) {
    f = sf[slot];
    s.row = d.mt + (size_t)f * kMtStride;
    s.cj = scj[slot];
    s.css = scss[4 * slot + c];
    s.w = sw[4 * slot + c];
    s.idx = sidx[slot];
    s.nreg = snreg[slot];
    quad_init<R>(k, s);
    s.gv = gv0;
    if (!LATE && s.t < k.t1) mt_fetch8(s.row, s.idx + 8 * c, s.win);
  };
  if (live) take(sl);
  int next = 16;  // slots handed out (wave-uniform; may run past nslot)
  unsigned stot = 0u, smax = 0u, passes = 0u;
  for (;;) {
    const bool work = live && s.t < k.t1 && !s.bad;
    passes += __ballot(work) != 0ull ? 1u : 0u;
    if (work) (void)quad_step<R, 64, LATE>(k, s, rtot, lom);
    if (live && !(s.t < k.t1 && !s.bad)) {  // this quad's FRU is done
      quad_end<R>(a, d, f, c, s);
      const int e = 4 * sl + c;
      scss[e] = s.css;
      sw[e] = s.w;
      sopen[e] = (s.lopen ? 1 : 0) | (s.nopen << 1);
      if (c == 0) {
        scj[sl] = s.cj;
        stot += s.steps;
        smax = max(smax, s.steps);
      }
      live = false;
    }
    const unsigned long long freeq = __ballot(!live && c == 0);  // lane 0 of each free quad
    if (next < nslot && freeq != 0ull) {
      const int ns = next + __popcll(freeq & ((1ull << (4 * q)) - 1ull));
      if (!live && ns < nslot) {
        sl = nslot - 1 - ns;
        live = true;
        take(sl);
      }
      next += __popcll(freeq);
    }
    if (__ballot(live) == 0ull) break;
  }
  __syncthreads();  // the final states in LDS, written by other quads
#pragma unroll 1
  for (int w = 0; w < 4; ++w) {
    const int e = 64 * w + lane;
    double acc[17] = {};
    if ((e >> 2) < nslot) {
      const int o = sopen[e];
      quad_acc(scss[e], scj[e >> 2], sw[e], (o & 1) != 0, o >> 1, c, a.avg, acc);
    }
    double* const slot = d.part + (size_t)(4 * blk + w) * kSlot;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      const double v = wave_sum_xor(acc[r]);
      if (lane == 0) slot[r] = v;
    }
    if (a.avg) {
#pragma unroll
      for (int r = 6; r < 17; ++r) {
        const double v = wave_sum_xor(acc[r]);
        if (lane == 0) slot[r] = v;
      }
    }
  }
  const unsigned ss = wave_u32<false>(stot), sm = wave_u32<true>(smax);
  if (lane == 0 && ss) {
    atomicAdd(d.nsteps, (unsigned long long)ss);
    atomicMax(d.maxsteps, sm);
    atomicAdd(d.nsteps + kLaneUse, 16ull * passes);  // 16 FRU slots per pass of the wavefront
  }
}

template <class R>
__global__ void __launch_bounds__(64) k_mc_refill(const SimArgs base, const Dev d, RkCtl* ctls) {
  mc_refill<R, false>(base, d, ctls);
}

// The same with its registers capped so that three wavefronts fit on a SIMD,
// each step's MT window fetched at the step's start.
template <class R>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_mc_refill3(const SimArgs base,
                                                                                      const Dev d, RkCtl* ctls) {
  mc_refill<R, true>(base, d, ctls);
}

// The four-lane stage with its registers capped so that W wavefronts fit on a
// SIMD (512 / W VGPRs): more wavefronts to hide latency when a batch has more
// of them than the chip holds at once, against spills.  Same arithmetic, same
// results.
template <class R, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) k_dev_stage_occ(const SimArgs base, const Dev d,
                                                                                          RkCtl* ctls) {
  dev_stage<R, false>(base, d, ctls);
}

// Regrouping of a batch cell's FRUs between RK attempts.  A wavefront runs as
// long as the slowest of its 16 FRUs, and a FRU that took many micro-steps in
// one sub-interval (high CaSS, open RyRs) keeps taking many for a while; so
// before each attempt every active cell's quad slots are refilled with its
// FRUs ordered by the step count of their last sub-interval, in kSortBins
// bins, FRU index order within a bin -- a stable counting sort, one block per
// cell, the same result every run.  Every per-FRU quantity is keyed by the FRU
// index, so only the order of the flux and average sums changes; a cell that
// has finished its run keeps its order, so a run continued in pieces sees the
// same orders as one unbroken run.
constexpr int kSortBins = 16;

__device__ __forceinline__ int cost_bin(uint32_t s) {
  constexpr uint32_t edge[kSortBins - 1] = {2, 3, 4, 6, 9, 14, 22, 34, 52, 80, 120, 180, 270, 400, 600};
  int b = 0;
#pragma unroll
  for (int k = 0; k < kSortBins - 1; ++k) b += (s >= edge[k]);
  return b;
}

// Each block's key for the dispatch order: the cost class of its last quad
// slot, the costliest of the block since the slots ascend by class; -1 for the
// blocks of a cell that has finished.
__global__ void __launch_bounds__(256) k_sort_cells(const Dev d, const RkCtl* ctls) {
  const int cell = blockIdx.x;
  const CellInfo ci = d.cells[cell];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (!ctls[cell].active) {
    for (int b = t; b < ci.snb; b += blockDim.x) d.bkey[ci.sb0 + b] = -1;
    return;
  }
  const int per = (ci.nfru + 255) / 256;
  const int lo = min(ci.nfru, t * per), hi = min(ci.nfru, lo + per);
  int cnt[kSortBins];
#pragma unroll
  for (int b = 0; b < kSortBins; ++b) cnt[b] = 0;
  for (int i = lo; i < hi; ++i) {
    const int bin = cost_bin(d.cost[ci.f0 + i]);
#pragma unroll
    for (int b = 0; b < kSortBins; ++b) cnt[b] += (bin == b);
  }
  // slot of this thread's first FRU of each bin: earlier bins, then earlier
  // threads of the same bin (wavefront scan, then the waves in order)
  __shared__ int wtot[kSortBins][4];
  __shared__ int binbase[kSortBins];
  int pos[kSortBins];
#pragma unroll
  for (int b = 0; b < kSortBins; ++b) {
    int x = cnt[b];
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    pos[b] = x - cnt[b];
    if (lane == 63) wtot[b][wave] = x;
  }
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int b = 0; b < kSortBins; ++b) {
      binbase[b] = acc;
      for (int w = 0; w < 4; ++w) acc += wtot[b][w];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < kSortBins; ++b) {
    int off = binbase[b];
    for (int w = 0; w < wave; ++w) off += wtot[b][w];
    pos[b] += off;
  }
  for (int i = lo; i < hi; ++i) {
    const int bin = cost_bin(d.cost[ci.f0 + i]);
    int p = 0;
#pragma unroll
    for (int b = 0; b < kSortBins; ++b) {
      p = (bin == b) ? pos[b] : p;
      pos[b] += (bin == b);
    }
    d.perm[ci.f0 + p] = i;
    if ((p & 63) == 63 || p == ci.nfru - 1) d.bkey[ci.sb0 + (p >> 6)] = bin;
  }
}

// The dispatch order of a batch's Monte Carlo blocks: by key, highest first,
// block index order within a key (a stable counting sort in one block of 1024
// threads, each over a contiguous range of blocks).  Longest-first list
// scheduling: the blocks that take the most micro-steps start in the first
// wave of the launch and the short ones fill the compute units behind them.
constexpr int kOrderThreads = 1024;
constexpr int kOrderRanks = kSortBins + 1;  // the keys kSortBins-1 .. 0, then -1

__global__ void __launch_bounds__(kOrderThreads) k_order_blocks(const Dev d, int nblk) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  constexpr int nw = kOrderThreads / 64;
  const int per = (nblk + kOrderThreads - 1) / kOrderThreads;
  const int lo = min(nblk, t * per), hi = min(nblk, lo + per);
  int cnt[kOrderRanks];
#pragma unroll
  for (int r = 0; r < kOrderRanks; ++r) cnt[r] = 0;
  for (int i = lo; i < hi; ++i) {
    const int r = kSortBins - 1 - d.bkey[i];
#pragma unroll
    for (int q = 0; q < kOrderRanks; ++q) cnt[q] += (r == q);
  }
  __shared__ int wtot[kOrderRanks][nw];
  __shared__ int rbase[kOrderRanks];
  int pos[kOrderRanks];
#pragma unroll
  for (int r = 0; r < kOrderRanks; ++r) {
    int x = cnt[r];
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    pos[r] = x - cnt[r];
    if (lane == 63) wtot[r][wave] = x;
  }
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int r = 0; r < kOrderRanks; ++r) {
      rbase[r] = acc;
      for (int w = 0; w < nw; ++w) acc += wtot[r][w];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kOrderRanks; ++r) {
    int off = rbase[r];
    for (int w = 0; w < wave; ++w) off += wtot[r][w];
    pos[r] += off;
  }
  for (int i = lo; i < hi; ++i) {
    const int r = kSortBins - 1 - d.bkey[i];
    int p = 0;
#pragma unroll
    for (int q = 0; q < kOrderRanks; ++q) {
      p = (r == q) ? pos[q] : p;
      pos[q] += (r == q);
    }
    d.border[p] = i;
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

__global__ void k_save(const SimArgs a, const Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.nfru) fru_checkpoint(a, d, i);
}

__global__ void k_resume(const SimArgs a, const Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.nfru) fru_rollback(a, d, i);
}


// One thread per FRU of every cell (blocks map to cells through d.bmap).
__device__ __forceinline__ int cell_fru(const Dev& d, int& cell) {
  cell = d.bmap[blockIdx.x];
  const CellInfo ci = d.cells[cell];
  const int il = (blockIdx.x - ci.bb0) * blockDim.x + threadIdx.x;
  return il < ci.nfru ? il : -1;
}

// Every cell starts from the same initial FRU set of nic records, FRU j taking
// record j mod nic (initialize_state.c:38-41,62-65).
__global__ void k_load_ic(const SimArgs a, const Dev d, const double* icc, const uint32_t* ich, int nic) {
  int cell;
  const int il = cell_fru(d, cell);
  if (il < 0) return;
  const int f = d.cells[cell].f0 + il, r = il % nic;
  for (int k = 0; k < 5; ++k) d.conc[k * a.ld + f] = icc[k * nic + r];
  for (int c = 0; c < 4; ++c) d.chan[c * a.ld + f] = ich[c * nic + r];
}

// initialize_ran.c:23-36 + sgrnd_local (mt19937_local.c:75-92) for every FRU
// j of every cell, with the cell's master seed.
__global__ void k_seed_cells(const SimArgs a, const Dev d) {
  int cell;
  const int il = cell_fru(d, cell);
  if (il < 0) return;
  const CellInfo ci = d.cells[cell];
  const long long j = il;
  const long long v = (long long)(int)ci.seed + 23 * (j * (j + 1) / 2) + 4093 * j * j + 1;
  uint32_t x = (uint32_t)(unsigned long long)v;
  const int f = ci.f0 + il;
  uint32_t* row = d.mt + (size_t)f * kMtStride;
  row[0] = x;
  row[kMtN] = x;
  for (uint32_t k = 1; k < kMtN; ++k) {
    x = 1812433253u * (x ^ (x >> 30)) + k;
    row[k] = x;
    if (k < (uint32_t)kMtMirror) row[k + kMtN] = x;  // the copy of words 0..7
  }
  d.mtidx[f] = kMtN;  // the first draw regenerates word 0
  d.nreg[f] = 0;
}

// Checkpoint of every FRU slot of the arrays (padding slots included).
__global__ void k_save_all(const SimArgs a, const Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.ld) fru_checkpoint(a, d, i);
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
  int grid_sim = 0;  // grid of the Monte Carlo kernel (4 lanes per FRU: 4x the FRU grid)
  Dev d{};
  hipStream_t stream = nullptr;
  double* host_out = nullptr;  // pinned [32]
  grw_stats stats{};
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> evpool;  // per-launch timing of device-driven runs
  bool graphs = false;             // device-driven runs replay captured attempts
  bool ws = false;                 // device integrator: wave-specialised Monte Carlo kernel (single cells)
  int occ = 0;                     // 0: the stage kernel as compiled; 3, 4: registers capped for 3 / 4 waves per SIMD
  bool queue = false;              // device integrator: Monte Carlo work queue + per-cell tail kernel
  bool wave1 = false;              // device integrator: single-wavefront Monte Carlo blocks + per-cell tail kernel
  bool refill = false;             // device integrator: refill blocks (quads take the block's next FRU) + tail kernel
  bool poll = false;               // device integrator (single cells): the fcn block collects the partials by polling
  int qgrid = 0;                   // blocks of the queue kernel: as many as the GPU holds at once
  hipGraphExec_t graph = nullptr;  // kGraphAttempts RK attempts, captured on first use
  SimArgs graph_base{};            // the launch arguments the graph holds
  Dev graph_dev{};
  std::vector<void*> allocs;
  // cells: one for a context made by grw_create, many for grw_batch_create
  int ncells = 1;
  bool batch = false;
  std::vector<CellInfo> cells;
  std::vector<double> t0s;      // each cell's start time (grw_cell_spec.t0)
  int grid_b = 0, grid_s = 0;   // integrator grids: one thread per FRU / 64 FRUs per block
  RkCtl* dctl = nullptr;        // [ncells] integrator control blocks, persistent across runs
  std::vector<RkCtl> hctl;      // host copies
  double* dbeats = nullptr;     // [ncells][beat_cap][kBeatStats]
  int beat_cap = 0;
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
  if (e) return fail(GRW_ESTATE, "kernel met an unknown channel state, a non-finite rate or a runaway step count");
  return 0;
}

// Timed launches go through hipExtLaunchKernelGGL, whose start / stop events
// take the dispatch's own begin and end timestamps: the kernel's execution
// time, without the gap to the launch before it (what a profiler's kernel
// trace reports), rather than the span between two event packets.
template <class R, bool SIM, bool REDUCE>
int launch_sim(grw_ctx* c, const SimArgs& a) {
  hipEvent_t e0 = c->timing ? c->ev0 : nullptr, e1 = c->timing ? c->ev1 : nullptr;
  if (SIM && c->cfg.lanes_per_fru == 4)
    hipExtLaunchKernelGGL((k_simfru4<R, REDUCE>), dim3(c->grid_sim), dim3(c->block), 0, c->stream, e0, e1, 0, a,
                          c->d);
  else
    hipExtLaunchKernelGGL((k_simfru<R, SIM, REDUCE>), dim3(c->grid), dim3(c->block), 0, c->stream, e0, e1, 0, a,
                          c->d);
  HIPCHK(hipGetLastError());
  if (c->timing) {
    HIPCHK(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.kernel_ms += ms;
    c->stats.timed_launches++;
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

// Device arrays for the FRUs of all cells (c->ld slots, c->grid_s Monte Carlo blocks).
int alloc_fru_state(grw_ctx* c) {
  int rc = 0;
  const size_t ld = c->ld;
  const size_t nslots = std::max<size_t>({(size_t)c->grid_sim, (size_t)4 * c->grid_s, (size_t)c->grid_b,
                                          (size_t)c->grid});  // (4 x grid_s: the wave-specialised layout)
  if ((rc = dalloc(c, &c->d.conc, 5 * ld)) || (rc = dalloc(c, &c->d.chan, 4 * ld)) ||
      (rc = dalloc(c, &c->d.mtidx, ld)) || (rc = dalloc(c, &c->d.nreg, ld)) ||
      (rc = dalloc(c, &c->d.mt, ld * kMtStride)) || (rc = dalloc(c, &c->d.mtw_h, ld)) ||
      (rc = dalloc(c, &c->d.conc_h, 5 * ld)) || (rc = dalloc(c, &c->d.chan_h, 4 * ld)) ||
      (rc = dalloc(c, &c->d.mtidx_h, ld)) || (rc = dalloc(c, &c->d.part, nslots * kSlot)) ||
      (rc = dalloc(c, &c->d.part2, (size_t)c->grid_s * kSlot)) ||
      (rc = dalloc(c, &c->d.ticket, (size_t)c->ncells)) || (rc = dalloc(c, &c->d.qhead, 1)) || (rc = dalloc(c, &c->d.out, 32)) ||
      (rc = dalloc(c, &c->d.nsteps, kLaneUse + 1)) || (rc = dalloc(c, &c->d.err, 1)) ||
      (rc = dalloc(c, &c->d.maxsteps, 1)) || (rc = dalloc(c, &c->d.critsteps, 1)) ||
      (rc = dalloc(c, &c->d.tailticks, 1)) || (rc = dalloc(c, &c->d.perm, ld)) ||
      (rc = dalloc(c, &c->d.cost, ld)) || (rc = dalloc(c, &c->d.bkey, (size_t)c->grid_s)) ||
      (rc = dalloc(c, &c->d.border, (size_t)c->grid_s)))
    return rc;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_out, 32 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) return fail(GRW_EDEVICE, std::string("stream/event setup failed: ") + hipGetErrorString(e));
  return 0;
}

// Cell placement: cell k's FRUs start at a multiple of 64; block maps for the
// two integrator grids; control blocks and beat records.  Sets c->ld, grids.
int place_cells(grw_ctx* c, std::vector<CellInfo>& cells) {
  int f0 = 0, sb = 0, bb = 0;
  for (CellInfo& ci : cells) {
    ci.f0 = f0;
    ci.sb0 = sb;
    ci.snb = (ci.nfru + 63) / 64;
    ci.bb0 = bb;
    ci.bnb = (ci.nfru + 255) / 256;
    f0 += ci.snb * 64;
    sb += ci.snb;
    bb += ci.bnb;
  }
  c->cells = cells;
  c->ncells = (int)cells.size();
  c->ld = f0;
  c->grid_s = sb;
  c->grid_b = bb;
  return 0;
}

int upload_cells(grw_ctx* c) {
  const int nc = c->ncells;
  std::vector<int> smap(c->grid_s), bmap(c->grid_b);
  for (int k = 0; k < nc; ++k) {
    const CellInfo& ci = c->cells[k];
    for (int b = 0; b < ci.snb; ++b) smap[ci.sb0 + b] = k;
    for (int b = 0; b < ci.bnb; ++b) bmap[ci.bb0 + b] = k;
  }
  CellInfo* dc = nullptr;
  int* ds = nullptr;
  int* db = nullptr;
  int rc = 0;
  if ((rc = dalloc(c, &dc, (size_t)nc)) || (rc = dalloc(c, &ds, (size_t)c->grid_s)) ||
      (rc = dalloc(c, &db, (size_t)c->grid_b)) || (rc = dalloc(c, &c->dctl, (size_t)nc)) ||
      (rc = dalloc(c, &c->dbeats, (size_t)nc * (size_t)c->beat_cap * kBeatStats)))
    return rc;
  HIPCHK(hipMemcpy(dc, c->cells.data(), sizeof(CellInfo) * nc, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ds, smap.data(), sizeof(int) * smap.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(db, bmap.data(), sizeof(int) * bmap.size(), hipMemcpyHostToDevice));
  c->d.cells = dc;
  c->d.smap = ds;
  c->d.bmap = db;
  c->hctl.assign(nc, RkCtl{});
  return 0;
}

int refresh_cells(grw_ctx* c) {
  HIPCHK(hipMemcpy((void*)c->d.cells, c->cells.data(), sizeof(CellInfo) * c->ncells, hipMemcpyHostToDevice));
  return 0;
}

// Forget the integrator's clock and beat tracking (a new initial state).
void reset_tracking(grw_ctx* c) {
  for (int k = 0; k < c->ncells; ++k) {
    RkCtl& h = c->hctl[k];
    h.ms_done = 0;
    h.beat_cap = c->beat_cap;
    h.beat_n = 0;
    h.beat_cur = -1;
    h.pend_row = -1;
    h.prev_t = h.prev_V = h.prev_jin = h.prev_jrel = NAN;
    h.pend_t = h.pend_V = h.pend_Cai = h.pend_NSR = h.pend_JSR = NAN;
    for (int q = 0; q < kBeatStats; ++q) h.bt[q] = NAN;
    h.out_beats = c->dbeats + (size_t)k * c->beat_cap * kBeatStats;
  }
}

constexpr int kTimeEvery = 8;  // the device integrator times one RK attempt in eight

// Between RK attempts of a batch: regroup every active cell's FRUs by cost,
// then order the Monte Carlo blocks for dispatch.
void sort_and_order(grw_ctx* c) {
  hipLaunchKernelGGL(k_sort_cells, dim3(c->ncells), dim3(256), 0, c->stream, c->d, c->dctl);
  hipLaunchKernelGGL(k_order_blocks, dim3(1), dim3(kOrderThreads), 0, c->stream, c->d, c->grid_s);
}

// The FRU order of every cell back to the identity, step counts to zero (a
// new initial state).
int reset_order(grw_ctx* c) {
  std::vector<int32_t> perm((size_t)c->ld, 0);
  for (const CellInfo& ci : c->cells)
    for (int i = 0; i < ci.nfru; ++i) perm[(size_t)ci.f0 + i] = i;
  HIPCHK(hipMemcpy(c->d.perm, perm.data(), perm.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->d.cost, 0, (size_t)c->ld * sizeof(uint32_t)));
  std::vector<int32_t> order((size_t)c->grid_s);
  for (int b = 0; b < c->grid_s; ++b) order[(size_t)b] = b;
  HIPCHK(hipMemcpy(c->d.border, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return 0;
}

// Monte Carlo blocks of a device-integrator launch: 64 FRUs per block with the
// four-lane kernel, 16 with the wave-specialised one (four blocks per 64).
int mc_grid(const grw_ctx* c) { return c->ws ? 4 * c->grid_s : c->grid_s; }

void launch_begin(grw_ctx* c, const SimArgs& base, bool mixed) {
  const dim3 grid(mc_grid(c));
  if (c->ws) {
    if (mixed) hipLaunchKernelGGL((k_dev_begin<float, true>), grid, dim3(256), 0, c->stream, base, c->d, c->dctl);
    else hipLaunchKernelGGL((k_dev_begin<double, true>), grid, dim3(256), 0, c->stream, base, c->d, c->dctl);
  } else {
    if (mixed) hipLaunchKernelGGL((k_dev_begin<float, false>), grid, dim3(256), 0, c->stream, base, c->d, c->dctl);
    else hipLaunchKernelGGL((k_dev_begin<double, false>), grid, dim3(256), 0, c->stream, base, c->d, c->dctl);
  }
}

// One RK attempt's launches: five k_dev_stage, then a batch's regrouping.
// ek (may be null) holds ten events, the start and stop of each of the five
// launches (hipExtLaunchKernelGGL: the dispatch's own begin / end timestamps).
void enqueue_attempt(grw_ctx* c, const SimArgs& base, bool mixed, const hipEvent_t* ek) {
  const dim3 grid(c->ncells + mc_grid(c));
  for (int s5 = 0; s5 < 5; ++s5) {
    hipEvent_t e0 = ek ? ek[2 * s5] : nullptr, e1 = ek ? ek[2 * s5 + 1] : nullptr;
    if (c->poll) {
      if (mixed)
        hipExtLaunchKernelGGL(k_dev_stage_poll<float>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d, c->dctl);
      else
        hipExtLaunchKernelGGL(k_dev_stage_poll<double>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d,
                              c->dctl);
    } else if (c->refill) {
      if (mixed) {
        if (c->occ == 3)
          hipExtLaunchKernelGGL(k_mc_refill3<float>, dim3(c->grid_s), dim3(64), 0, c->stream, e0, nullptr, 0, base,
                                c->d, c->dctl);
        else
          hipExtLaunchKernelGGL(k_mc_refill<float>, dim3(c->grid_s), dim3(64), 0, c->stream, e0, nullptr, 0, base,
                                c->d, c->dctl);
        hipExtLaunchKernelGGL(k_dev_tail<float>, dim3(c->ncells), dim3(256), 0, c->stream, nullptr, e1, 0, base, c->d,
                              c->dctl);
      } else {
        if (c->occ == 3)
          hipExtLaunchKernelGGL(k_mc_refill3<double>, dim3(c->grid_s), dim3(64), 0, c->stream, e0, nullptr, 0, base,
                                c->d, c->dctl);
        else
          hipExtLaunchKernelGGL(k_mc_refill<double>, dim3(c->grid_s), dim3(64), 0, c->stream, e0, nullptr, 0, base,
                                c->d, c->dctl);
        hipExtLaunchKernelGGL(k_dev_tail<double>, dim3(c->ncells), dim3(256), 0, c->stream, nullptr, e1, 0, base,
                              c->d, c->dctl);
      }
    } else if (c->wave1) {
      const dim3 wgrid(4 * c->grid_s);
      if (mixed) {
        if (c->occ == 3)
          hipExtLaunchKernelGGL((k_mc_wave_occ<float, 3>), wgrid, dim3(64), 0, c->stream, e0, nullptr, 0, base, c->d,
                                c->dctl);
        else if (c->occ == 4)
          hipExtLaunchKernelGGL((k_mc_wave_occ<float, 4>), wgrid, dim3(64), 0, c->stream, e0, nullptr, 0, base, c->d,
                                c->dctl);
        else
          hipExtLaunchKernelGGL(k_mc_wave<float>, wgrid, dim3(64), 0, c->stream, e0, nullptr, 0, base, c->d, c->dctl);
        hipExtLaunchKernelGGL(k_dev_tail<float>, dim3(c->ncells), dim3(256), 0, c->stream, nullptr, e1, 0, base, c->d,
                              c->dctl);
      } else {
        if (c->occ == 3)
          hipExtLaunchKernelGGL((k_mc_wave_occ<double, 3>), wgrid, dim3(64), 0, c->stream, e0, nullptr, 0, base, c->d,
                                c->dctl);
        else if (c->occ == 4)
          hipExtLaunchKernelGGL((k_mc_wave_occ<double, 4>), wgrid, dim3(64), 0, c->stream, e0, nullptr, 0, base, c->d,
                                c->dctl);
        else
          hipExtLaunchKernelGGL(k_mc_wave<double>, wgrid, dim3(64), 0, c->stream, e0, nullptr, 0, base, c->d, c->dctl);
        hipExtLaunchKernelGGL(k_dev_tail<double>, dim3(c->ncells), dim3(256), 0, c->stream, nullptr, e1, 0, base,
                              c->d, c->dctl);
      }
    } else if (c->queue) {
      const int nitems = 4 * c->grid_s;
      if (mixed) {
        hipExtLaunchKernelGGL(k_mc_queue<float>, dim3(c->qgrid), dim3(256), 0, c->stream, e0, nullptr, 0, base, c->d,
                              c->dctl, nitems);
        hipExtLaunchKernelGGL(k_dev_tail<float>, dim3(c->ncells), dim3(256), 0, c->stream, nullptr, e1, 0, base, c->d,
                              c->dctl);
      } else {
        hipExtLaunchKernelGGL(k_mc_queue<double>, dim3(c->qgrid), dim3(256), 0, c->stream, e0, nullptr, 0, base,
                              c->d, c->dctl, nitems);
        hipExtLaunchKernelGGL(k_dev_tail<double>, dim3(c->ncells), dim3(256), 0, c->stream, nullptr, e1, 0, base,
                              c->d, c->dctl);
      }
    } else if (c->occ == 3 || c->occ == 4) {
      if (c->occ == 3) {
        if (mixed)
          hipExtLaunchKernelGGL(k_dev_stage_occ<float, 3>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d, c->dctl);
        else
          hipExtLaunchKernelGGL(k_dev_stage_occ<double, 3>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d, c->dctl);
      } else {
        if (mixed)
          hipExtLaunchKernelGGL(k_dev_stage_occ<float, 4>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d, c->dctl);
        else
          hipExtLaunchKernelGGL(k_dev_stage_occ<double, 4>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d, c->dctl);
      }
    } else if (c->ws) {
      if (mixed)
        hipExtLaunchKernelGGL(k_dev_stage<float, true>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d, c->dctl);
      else
        hipExtLaunchKernelGGL(k_dev_stage<double, true>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d, c->dctl);
    } else {
      if (mixed)
        hipExtLaunchKernelGGL(k_dev_stage<float, false>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d, c->dctl);
      else
        hipExtLaunchKernelGGL(k_dev_stage<double, false>, grid, dim3(256), 0, c->stream, e0, e1, 0, base, c->d,
                              c->dctl);
    }
  }
  if (base.sorted) sort_and_order(c);
}

// Attempts per graph replay: with the attempt that run_device launches
// directly (the one it may time), a group of kTimeEvery.
constexpr int kGraphAttempts = kTimeEvery - 1;

// Capture kGraphAttempts attempts into c->graph, once per context: every
// attempt's launches are the same kernels with the same arguments (the
// control blocks on the device decide what each one does), so one graph
// serves every run.  Kernel launches only: nothing is cleared inside it.
int ensure_graph(grw_ctx* c, const SimArgs& base, bool mixed) {
  if (c->graph && std::memcmp(&c->graph_base, &base, sizeof base) == 0 &&
      std::memcmp(&c->graph_dev, &c->d, sizeof c->d) == 0)
    return 0;
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  for (int k = 0; k < kGraphAttempts; ++k) enqueue_attempt(c, base, mixed, nullptr);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e == hipSuccess) e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->graph = nullptr;
    return fail(GRW_EDEVICE, std::string("attempt graph capture failed: ") + hipGetErrorString(e));
  }
  c->graph_base = base;
  c->graph_dev = c->d;
  return 0;
}

// Enqueue RK attempts (five k_dev_stage each, after one k_dev_begin) until every active cell
// of c->hctl has reached its ms_target; the host synchronises once per batch of
// attempts sized to ~25 simulated ms of the slowest cell, and launches after a
// cell is done return at once for it.
int run_device(grw_ctx* c) {
  const int nc = c->ncells;
  HIPCHK(hipMemcpyAsync(c->dctl, c->hctl.data(), sizeof(RkCtl) * nc, hipMemcpyHostToDevice, c->stream));
  if (c->poll) {  // every polled slot empty (a run stopped by an error may have left some written)
    const int n = c->grid_s * kSlot;
    hipLaunchKernelGGL(k_arm_slots, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d.part2, n);
    HIPCHK(hipGetLastError());
  }
  SimArgs base = base_args(c);
  base.ncells = nc;
  base.sorted = c->batch ? 1 : 0;  // batch cells regroup their FRUs between attempts
  const bool mixed = c->cfg.precision == GRW_MIXED;
  std::vector<long long> att_prev(nc, 0);
  std::vector<int> ms_prev(nc);
  std::vector<double> ms_per_att(nc, 1.0 / 40.0);
  for (int k = 0; k < nc; ++k) ms_prev[k] = c->hctl[k].ms_done;
  std::vector<hipEvent_t>& evs = c->evpool;
  bool first = true;
  while (true) {
    int K = 0;
    for (int k = 0; k < nc; ++k) {
      const RkCtl& h = c->hctl[k];
      if (!h.active) continue;
      const int chunk = std::min(h.ms_target - h.ms_done, 25);
      K = std::max(K, (int)std::ceil(chunk / ms_per_att[k] * 1.1) + 2);
    }
    if (K == 0) break;
    K = std::max(4, std::min(K, 4000));
    const bool graph = c->graphs;
    if (graph) {  // whole groups of kTimeEvery attempts (4000 is a multiple of 8)
      K = (K + kTimeEvery - 1) / kTimeEvery * kTimeEvery;
      if (int rc = ensure_graph(c, base, mixed)) return rc;
    }
    // timing: HIP events around the five launches of one attempt in kTimeEvery
    // (events around every launch cost the run itself a few percent)
    const int nsamp = (K + kTimeEvery - 1) / kTimeEvery;
    if (c->timing) {
      while ((int)evs.size() < 10 * nsamp) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) break;
        evs.push_back(ev);
      }
    }
    const bool timed = c->timing && (int)evs.size() >= 10 * nsamp;
    std::vector<char> was_active(nc);
    for (int k = 0; k < nc; ++k) was_active[k] = c->hctl[k].active != 0;
    if (first) {  // the run's first attempt; later ones start in the last stage's tail
      if (mixed)
        launch_begin(c, base, true);
      else
        launch_begin(c, base, false);
      if (base.sorted) sort_and_order(c);
    }
    first = false;
    hipError_t eg = hipSuccess;
    for (int k = 0; k < K; ++k) {
      if (graph && k % kTimeEvery != 0) {  // the group's other attempts in one replay
        if (k % kTimeEvery == 1 && eg == hipSuccess) eg = hipGraphLaunch(c->graph, c->stream);
        continue;
      }
      const bool tk = timed && k % kTimeEvery == 0;
      enqueue_attempt(c, base, mixed, tk ? evs.data() + 10 * (k / kTimeEvery) : nullptr);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = eg;
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->hctl.data(), c->dctl, sizeof(RkCtl) * nc, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(GRW_EDEVICE, std::string("device run failed: ") + hipGetErrorString(e));
    int kerr = 0;
    HIPCHK(hipMemcpy(&kerr, c->d.err, sizeof(int), hipMemcpyDeviceToHost));
    long long eff = 0;
    for (int k = 0; k < nc; ++k) {
      const RkCtl& h = c->hctl[k];
      bool finite = true;
      for (int q = 0; q < kNglobal; ++q) finite = finite && std::isfinite(h.state[q]);
      const std::string who = nc > 1 ? " (cell " + std::to_string(k) + ")" : std::string();
      if (!finite) return fail(GRW_ESTATE, "global state became non-finite" + who);
      if (h.attempts > 20000LL * (long long)(h.ms_target - h.ms_base) + 100000LL)
        return fail(GRW_ESTATE, "integrator made no progress (step size collapsed)" + who);
      if (kerr || h.err)
        return fail(GRW_ESTATE, "kernel met an unknown channel state, a non-finite rate or a runaway step count" + who);
      const long long done = h.attempts - att_prev[k];
      // a cell still active after a batch of at least four attempts' launches
      // started attempts in it (one finishing its last attempt may start none)
      if (was_active[k] && h.active && done == 0)
        return fail(GRW_EDEVICE, "a batch of launches completed no RK attempt" + who);
      eff = std::max(eff, done);
      c->stats.simfru_calls += (uint64_t)(5 * done) * (uint64_t)c->cells[k].nfru;
      att_prev[k] = h.attempts;
      if (h.ms_done > ms_prev[k] && done > 0) ms_per_att[k] = (double)(h.ms_done - ms_prev[k]) / (double)done;
      ms_prev[k] = h.ms_done;
    }
    if (timed) {  // the sampled attempts that ran (later launches of the batch return at once)
      for (long long k = 0; k < std::min<long long>(eff, K); k += kTimeEvery)
        for (int s5 = 0; s5 < 5; ++s5) {
          float ms = 0.f;
          const hipEvent_t* ek = evs.data() + 10 * (k / kTimeEvery);
          if (hipEventElapsedTime(&ms, ek[2 * s5], ek[2 * s5 + 1]) == hipSuccess) {
            c->stats.kernel_ms += ms;
            c->stats.timed_launches++;
          }
        }
    }
    c->stats.launches += (uint64_t)(5 * eff);
  }
  bool avg = false;
  for (const RkCtl& h : c->hctl) avg = avg || h.avg_pending;
  if (avg) {  // averages of the last output sample
    if (mixed)
      launch_begin(c, base, true);
    else
      launch_begin(c, base, false);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->hctl.data(), c->dctl, sizeof(RkCtl) * nc, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(GRW_EDEVICE, std::string("device run failed: ") + hipGetErrorString(e));
  }
  return 0;
}

// The per-FRU ABI calls act on one cell.
int single(const grw_ctx* c) {
  if (!c) return fail(GRW_EINVAL, "null context");
  if (c->ncells != 1) return fail(GRW_EINVAL, "this call acts on a single-cell context (grw_create)");
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
  cfg->lanes_per_fru = 4;
}

int grw_create(const grw_config* cfg, grw_ctx** out) {
  if (!cfg || !out) return fail(GRW_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->nfru <= 0 || cfg->nfru > (1 << 26)) return fail(GRW_EINVAL, "nfru out of range");
  if (cfg->precision != GRW_FP64 && cfg->precision != GRW_MIXED)
    return fail(GRW_EINVAL, "precision must be GRW_FP64 or GRW_MIXED");
  const int blk = cfg->block ? cfg->block : (cfg->lanes_per_fru == 4 ? 256 : 64);
  if (blk % 64 || blk < 64 || blk > 256) return fail(GRW_EINVAL, "block must be 64, 128, 192 or 256");
  if (cfg->lanes_per_fru != 1 && cfg->lanes_per_fru != 4)
    return fail(GRW_EINVAL, "lanes_per_fru must be 1 or 4");
  HIPCHK(hipSetDevice(cfg->device));
  grw_ctx* c = new grw_ctx();
  c->cfg = *cfg;
  c->n = cfg->nfru;
  c->block = blk;
  c->grid = (c->n + blk - 1) / blk;
  c->grid_sim = cfg->lanes_per_fru == 4 ? (4 * c->n + blk - 1) / blk : c->grid;
  std::vector<CellInfo> one(1);
  one[0].nfru = c->n;
  one[0].scale = cfg->nfru_scale;
  one[0].k23c = cfg->k23_scale != 1.0 ? k23cf * cfg->k23_scale : k23cf;
  one[0].gcf = cfg->gamma_scale != 1.0 ? gammacf * cfg->gamma_scale : gammacf;
  one[0].seed = 0;
  place_cells(c, one);
  c->beat_cap = 4096;
  int rc = alloc_fru_state(c);
  if (!rc) rc = upload_cells(c);
  if (rc) {
    grw_destroy(c);
    return rc;
  }
  reset_tracking(c);
  if (int rc2 = reset_order(c)) {
    grw_destroy(c);
    return rc2;
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
  for (hipEvent_t ev : c->evpool) (void)hipEventDestroy(ev);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int grw_load_state(grw_ctx* c, const double* fru, const int32_t* ltype, const int32_t* ryr,
                   const int32_t* ito2, const int32_t* mti, const uint32_t* mt) {
  if (int rc = single(c)) return rc;
  if (!fru || !ltype || !ryr || !ito2 || !mti || !mt) return fail(GRW_EINVAL, "null argument");
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
    std::memcpy(&rows[(size_t)i * kMtStride + kMtN], mt + (size_t)i * kMtN, kMtMirror * sizeof(uint32_t));
  }
  HIPCHK(hipMemcpyAsync(c->d.conc, conc.data(), conc.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d.chan, chan.data(), chan.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d.mtidx, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->d.nreg, 0, (size_t)ld * sizeof(int32_t), c->stream));
  HIPCHK(hipMemcpyAsync(c->d.mt, rows.data(), rows.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  reset_tracking(c);
  if (int rc = reset_order(c)) return rc;
  return grw_save_state(c);
}

int grw_get_state(grw_ctx* c, double* fru, int32_t* ltype, int32_t* ryr, int32_t* ito2,
                  int32_t* mti, uint32_t* mt) {
  if (int rc = single(c)) return rc;
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
  if (int rc = single(c)) return rc;
  c->cells[0].seed = a0;
  if (int rc = refresh_cells(c)) return rc;
  SimArgs a = base_args(c);
  hipLaunchKernelGGL(k_seed_cells, dim3(c->grid_b), dim3(256), 0, c->stream, a, c->d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  reset_tracking(c);
  if (int rc = reset_order(c)) return rc;
  return grw_save_state(c);
}

int grw_simfru_local(grw_ctx* c, double t0, double t1, const double dep0[3], const double depf[3]) {
  if (int rc = single(c)) return rc;
  if (!dep0 || !depf) return fail(GRW_EINVAL, "null argument");
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
  if (int rc = single(c)) return rc;
  if (!dep0 || !depf || !out4) return fail(GRW_EINVAL, "null argument");
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
  if (int rc = single(c)) return rc;
  if (!dep || !out4) return fail(GRW_EINVAL, "null argument");
  int rc = flux_now(c, dep);
  if (rc) return rc;
  return fetch_out(c, 0, 4, out4);
}

int grw_calc_fru_flux_local(grw_ctx* c, double num5[5]) {
  if (int rc = single(c)) return rc;
  if (!num5) return fail(GRW_EINVAL, "null argument");
  const double dep[3] = {-80.0, 1e-4, 1.0};
  int rc = flux_now(c, dep);
  if (rc) return rc;
  return fetch_out(c, 4, 5, num5);
}

int grw_calc_fru_flux_local_noalloc(int nfru, const int32_t* ltype, const int32_t* ito2, const double* fru,
                                    double* num5, void* stream) {
  if (nfru < 0 || nfru > kNoAllocMaxFru)
    return fail(GRW_EINVAL, "grw_calc_fru_flux_local_noalloc: nfru " + std::to_string(nfru) +
                                " outside [0, " + std::to_string(kNoAllocMaxFru) + "], the allocation-free maximum");
  if (!num5 || (nfru > 0 && (!ltype || !ito2 || !fru))) return fail(GRW_EINVAL, "null argument");
  // The scratch array's address on the current device, looked up once per
  // device (each GPU holds its own copy of the array).
  static std::mutex mu;
  static void* ptr[64] = {};
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(GRW_EDEVICE, "device ordinal beyond 63");
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(GRW_EDEVICE, std::string("scratch symbol address: ") + hipGetErrorString(e));
      ptr[dev] = p;
    }
  }
  const int grid = nfru > 0 ? (nfru + 255) / 256 : 1;
  hipLaunchKernelGGL(k_flux_sums_ref, dim3(grid), dim3(256), 0, (hipStream_t)stream, nfru, ltype, ito2, fru, num5,
                     (NoAllocScratch*)ptr[dev]);
  HIPCHK(hipGetLastError());
  return 0;
}

int grw_calc_fru_avg_local(grw_ctx* c, double num11[11]) {
  if (int rc = single(c)) return rc;
  if (!num11) return fail(GRW_EINVAL, "null argument");
  SimArgs a = base_args(c);
  if (c->cfg.sequential_sums)
    hipLaunchKernelGGL(k_seq_sums, dim3(1), dim3(1), 0, c->stream, a, c->d);
  else
    hipLaunchKernelGGL(k_avg, dim3(c->grid), dim3(c->block), 0, c->stream, a, c->d);
  HIPCHK(hipGetLastError());
  return fetch_out(c, 16, 11, num11);
}

int grw_calc_avg(grw_ctx* c, const double state[37], double other[12]) {
  if (int rc = single(c)) return rc;
  if (!state || !other) return fail(GRW_EINVAL, "null argument");
  double num[11];
  int rc = grw_calc_fru_avg_local(c, num);
  if (rc) return rc;
  avg_outputs(num, state, c->n, c->cfg.nfru_scale, other);
  return 0;
}

int grw_save_state(grw_ctx* c) {
  if (int rc = single(c)) return rc;
  SimArgs a = base_args(c);
  hipLaunchKernelGGL(k_save, dim3(c->grid), dim3(c->block), 0, c->stream, a, c->d);
  HIPCHK(hipGetLastError());
  return 0;
}

int grw_resume_state(grw_ctx* c) {
  if (int rc = single(c)) return rc;
  SimArgs a = base_args(c);
  hipLaunchKernelGGL(k_resume, dim3(c->grid), dim3(c->block), 0, c->stream, a, c->d);
  HIPCHK(hipGetLastError());
  return 0;
}

int grw_get_stats(grw_ctx* c, grw_stats* st) {
  if (!c || !st) return fail(GRW_EINVAL, "null argument");
  unsigned long long s = 0, cs = 0, tt = 0, ss = 0;
  HIPCHK(hipMemcpyAsync(&s, c->d.nsteps, sizeof s, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&cs, c->d.critsteps, sizeof cs, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&ss, c->d.nsteps + kLaneUse, sizeof ss, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&tt, c->d.tailticks, sizeof tt, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->cfg.device) != hipSuccess || khz <= 0)
    khz = 100000;
  *st = c->stats;
  st->fru_steps = s;
  st->critical_steps = cs;
  st->slot_steps = ss;
  st->tail_ms = (double)tt / (double)khz;
  return 0;
}

int grw_reset_stats(grw_ctx* c) {
  if (!c) return fail(GRW_EINVAL, "null context");
  HIPCHK(hipMemsetAsync(c->d.nsteps, 0, (kLaneUse + 1) * sizeof(unsigned long long), c->stream));
  HIPCHK(hipMemsetAsync(c->d.critsteps, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(hipMemsetAsync(c->d.tailticks, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->stats = grw_stats{};
  return 0;
}

int grw_set_timing(grw_ctx* c, int on) {
  if (!c) return fail(GRW_EINVAL, "null context");
  c->timing = on != 0;
  return 0;
}

int grw_set_layout(grw_ctx* c, int layout) {
  if (!c) return fail(GRW_EINVAL, "null context");
  if (layout < GRW_LAYOUT_QUAD || layout > GRW_LAYOUT_REFILL_3)
    return fail(GRW_EINVAL, "layout must be one of GRW_LAYOUT_QUAD, _WAVES, _QUAD3, _QUAD4, _QUEUE, _WAVE1, _WAVE1_3, "
                            "_QUADP, _WAVE1_4, _REFILL, _REFILL_3");
  if (layout == GRW_LAYOUT_QUADP && (c->batch || c->ncells != 1))
    return fail(GRW_EINVAL, "the polling layout runs single-cell contexts only");
  if (layout == GRW_LAYOUT_WAVES && (c->batch || c->ncells != 1))
    return fail(GRW_EINVAL, "the wave-specialised layout runs single-cell contexts only");
  // The queue layout is switched off: the one run of it on the device (a
  // 384-FRU cell) did not finish, and no cause has been found in its code.
  if (layout == GRW_LAYOUT_QUEUE)
    return fail(GRW_EINVAL, "the queue layout is switched off (a device run of it did not finish)");
  const bool ws = layout == GRW_LAYOUT_WAVES;
  const bool queue = layout == GRW_LAYOUT_QUEUE;
  const bool wave1 = layout == GRW_LAYOUT_WAVE1 || layout == GRW_LAYOUT_WAVE1_3 || layout == GRW_LAYOUT_WAVE1_4;
  const bool poll = layout == GRW_LAYOUT_QUADP;
  const bool refill = layout == GRW_LAYOUT_REFILL || layout == GRW_LAYOUT_REFILL_3;
  const int occ = (layout == GRW_LAYOUT_QUAD3 || layout == GRW_LAYOUT_WAVE1_3 || layout == GRW_LAYOUT_REFILL_3) ? 3
                  : (layout == GRW_LAYOUT_QUAD4 || layout == GRW_LAYOUT_WAVE1_4) ? 4
                                                                                  : 0;
  if (queue && c->qgrid == 0) {  // as many queue blocks as the GPU holds at once
    int dev = 0, ncu = 0, per = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const bool mixed = c->cfg.precision == GRW_MIXED;
    if (mixed)
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_mc_queue<float>, 256, 0));
    else
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_mc_queue<double>, 256, 0));
    c->qgrid = std::max(1, std::min(c->grid_s, std::max(per, 1) * std::max(ncu, 1)));
  }
  if ((ws != c->ws || occ != c->occ || queue != c->queue || wave1 != c->wave1 || poll != c->poll ||
       refill != c->refill) &&
      c->graph) {
    // the captured launches are another kernel's
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  c->ws = ws;
  c->occ = occ;
  c->queue = queue;
  c->wave1 = wave1;
  c->refill = refill;
  if (poll && !c->poll) {  // every slot empty before the first polled launch
    const int n = c->grid_s * kSlot;
    hipLaunchKernelGGL(k_arm_slots, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d.part2, n);
    HIPCHK(hipGetLastError());
  }
  c->poll = poll;
  return 0;
}

int grw_set_graphs(grw_ctx* c, int on) {
  if (!c) return fail(GRW_EINVAL, "null context");
  c->graphs = on != 0;
  return 0;
}

// ---- device-resident cell run (include/grw_cell.h)
int grw_run_cell_device(grw_ctx* c, const GrwProtocol* p, const double errw[37], double t0,
                        int nsteps, double state[37], double* h, double* states, double* currents,
                        double* other, double* log, int log_cap, int* log_n, GrwRkStats* stt) {
  if (int rc = single(c)) return rc;
  if (!p || !errw || !state || !h || nsteps < 0 || log_cap < 0)
    return fail(GRW_EINVAL, "bad argument to grw_run_cell_device");
  if (nsteps == 0) return 0;
  if (!std::isfinite(t0) || !(p->time_step > 0.0) || !(p->step_min > 0.0) || !(p->step_max >= p->step_min) ||
      !(p->period > 0.0))
    return fail(GRW_EINVAL, "bad protocol time controls");
  const size_t ns = (size_t)nsteps, lc = (size_t)std::max(log_cap, 1);
  double* dbuf = nullptr;
  hipError_t e0 = hipMalloc((void**)&dbuf, sizeof(double) * (ns * (kNglobal + kNcur + 12) + lc * 4));
  if (e0 != hipSuccess) return fail(GRW_EDEVICE, std::string("hipMalloc failed: ") + hipGetErrorString(e0));
  RkCtl& hc = c->hctl[0];  // keeps the output clock and the beat tracking of earlier runs
  hc.prot = *p;
  for (int q = 0; q < kNglobal; ++q) {
    hc.errw[q] = errw[q];
    hc.state[q] = state[q];
  }
  hc.tcall = t0;
  hc.start = t0;
  hc.tstep = t0 + p->time_step;
  double h0 = (*h < hc.tstep - t0) ? *h : hc.tstep - t0;  // rk_pd54m.c:71-74
  h0 = (p->step_max < h0) ? p->step_max : h0;
  hc.h = h0;
  hc.oh = h0;
  hc.stage = 0;
  hc.success = 1;
  hc.keepc = 1;
  hc.resume = 0;
  hc.active = 1;
  hc.avg_pending = 0;
  hc.ms_base = hc.ms_done;
  hc.ms_target = hc.ms_done + nsteps;
  hc.log_n = 0;
  hc.log_cap = log ? log_cap : 0;
  hc.err = 0;
  hc.attempts = hc.rejects = hc.forced = 0;
  hc.out_states = dbuf;
  hc.out_cur = dbuf + ns * kNglobal;
  hc.out_other = hc.out_cur + ns * kNcur;
  hc.out_log = hc.out_other + ns * 12;
  hc.out_beats = c->dbeats;
  hc.beat_cap = c->beat_cap;
  int rc = run_device(c);
  if (!rc) {
    std::vector<double> hb(ns * (kNglobal + kNcur + 12) + (size_t)hc.log_n * 4);
    hipError_t e = hipMemcpy(hb.data(), dbuf, sizeof(double) * hb.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(GRW_EDEVICE, std::string("output download failed: ") + hipGetErrorString(e));
    } else {
      if (states) std::memcpy(states, hb.data(), sizeof(double) * ns * kNglobal);
      if (currents) std::memcpy(currents, hb.data() + ns * kNglobal, sizeof(double) * ns * kNcur);
      if (other) std::memcpy(other, hb.data() + ns * (kNglobal + kNcur), sizeof(double) * ns * 12);
      if (log && log_n) {
        const int add = std::min(hc.log_n, log_cap - *log_n);
        if (add > 0)
          std::memcpy(log + 4 * (size_t)(*log_n), hb.data() + ns * (kNglobal + kNcur + 12),
                      sizeof(double) * 4 * (size_t)add);
        *log_n += std::max(add, 0);
      }
      for (int q = 0; q < kNglobal; ++q) state[q] = hc.state[q];
      *h = hc.oh;
      if (stt) {
        stt->attempts += hc.attempts;
        stt->rejects += hc.rejects;
        stt->forced += hc.forced;
      }
    }
  }
  hc.active = 0;
  hc.out_states = hc.out_cur = hc.out_other = hc.out_log = nullptr;
  hc.log_cap = 0;
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(dbuf);
  return rc;
}

int grw_get_beats(grw_ctx* c, double* out, int cap, int* n, int include_current) {
  return grw_batch_get_beats(c, 0, out, cap, n, include_current);
}

// ---- batches of independent cells (include/grw_batch.h)
int grw_batch_create(const grw_batch_config* cfg, const grw_cell_spec* spec, grw_ctx** out) {
  if (!cfg || !spec || !out) return fail(GRW_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->ncells <= 0 || cfg->ncells > (1 << 16)) return fail(GRW_EINVAL, "ncells out of range");
  if (cfg->precision != GRW_FP64 && cfg->precision != GRW_MIXED)
    return fail(GRW_EINVAL, "precision must be GRW_FP64 or GRW_MIXED");
  if (cfg->beat_cap < 0 || cfg->beat_cap > (1 << 20)) return fail(GRW_EINVAL, "beat_cap out of range");
  std::vector<CellInfo> cells(cfg->ncells);
  long long total = 0;
  for (int k = 0; k < cfg->ncells; ++k) {
    const grw_cell_spec& sp = spec[k];
    if (sp.nfru <= 0 || sp.nfru > (1 << 24)) return fail(GRW_EINVAL, "cell " + std::to_string(k) + ": nfru out of range");
    const GrwProtocol& p = sp.prot;
    if (!(p.time_step > 0.0) || !(p.step_min > 0.0) || !(p.step_max >= p.step_min) || !(p.period > 0.0) ||
        !(p.tolrk > 0.0) || !std::isfinite(p.shift))
      return fail(GRW_EINVAL, "cell " + std::to_string(k) + ": bad protocol time controls");
    if (!(sp.k23_scale > 0.0) || !(sp.gamma_scale > 0.0) || !(sp.nfru_scale >= 0.0) || !(p.serca_scale > 0.0))
      return fail(GRW_EINVAL, "cell " + std::to_string(k) + ": scales must be positive");
    if (!std::isfinite(sp.t0) || (p.vppclamp && !(p.vpp_period > 0.0)) || (p.ivclamp && !(p.iv_period > 0.0)) ||
        (p.vclamp && !(p.vclamp_period > 0.0)))
      return fail(GRW_EINVAL, "cell " + std::to_string(k) + ": bad start time or clamp period");
    cells[k].nfru = sp.nfru;
    cells[k].scale = sp.nfru_scale > 0.0 ? sp.nfru_scale : 12500.0 / (double)sp.nfru;
    cells[k].k23c = sp.k23_scale != 1.0 ? k23cf * sp.k23_scale : k23cf;
    cells[k].gcf = sp.gamma_scale != 1.0 ? gammacf * sp.gamma_scale : gammacf;
    cells[k].seed = sp.seed;
    total += sp.nfru;
  }
  if (total > (1LL << 28)) return fail(GRW_EINVAL, "too many FRUs in the batch");
  HIPCHK(hipSetDevice(cfg->device));
  grw_ctx* c = new grw_ctx();
  grw_config_default(&c->cfg);
  c->cfg.nfru = (int)total;
  c->cfg.precision = cfg->precision;
  c->cfg.device = cfg->device;
  c->batch = true;
  c->n = (int)total;
  c->block = 256;
  place_cells(c, cells);
  c->grid = c->grid_sim = c->grid_s;
  c->beat_cap = cfg->beat_cap;
  int rc = alloc_fru_state(c);
  if (!rc) rc = upload_cells(c);
  if (rc) {
    grw_destroy(c);
    return rc;
  }
  c->t0s.resize(c->ncells);
  for (int k = 0; k < c->ncells; ++k) {
    c->hctl[k].prot = spec[k].prot;
    c->t0s[k] = spec[k].t0;
  }
  reset_tracking(c);
  if (int rc2 = reset_order(c)) {
    grw_destroy(c);
    return rc2;
  }
  *out = c;
  return 0;
}

int grw_batch_load_ic(grw_ctx* c, const double* states, const double errw[37], int nic, const double* fru,
                      const int32_t* ltype, const int32_t* ryr, const int32_t* ito2) {
  if (!c || !states || !errw || !fru || !ltype || !ryr || !ito2) return fail(GRW_EINVAL, "null argument");
  if (!c->batch) return fail(GRW_EINVAL, "grw_batch_load_ic needs a batch context (grw_batch_create)");
  if (nic <= 0) return fail(GRW_EINVAL, "empty initial FRU set");
  std::vector<double> icc(5 * (size_t)nic);
  std::vector<uint32_t> ich(4 * (size_t)nic);
  for (int i = 0; i < nic; ++i) {
    for (int k = 0; k < 5; ++k) icc[(size_t)k * nic + i] = fru[5 * (size_t)i + k];
    for (int cl = 0; cl < 4; ++cl) {
      const int lcc = ltype[8 * (size_t)i + 2 * cl], y = ltype[8 * (size_t)i + 2 * cl + 1];
      const int32_t* rr = ryr + 20 * (size_t)i + 5 * cl;
      const int it = ito2[4 * (size_t)i + cl];
      if (lcc < 1 || lcc > 12 || (y != 1 && y != 2) || (it != 1 && it != 2))
        return fail(GRW_EINVAL, "invalid LCC / gate / Ito2 state at FRU " + std::to_string(i));
      for (int r = 0; r < 5; ++r)
        if (rr[r] < 1 || rr[r] > 8) return fail(GRW_EINVAL, "invalid RyR state at FRU " + std::to_string(i));
      ich[(size_t)cl * nic + i] = ch_pack(lcc, y, rr, it);
    }
  }
  for (int k = 0; k < c->ncells; ++k)
    for (int q = 0; q < kNglobal; ++q)
      if (!std::isfinite(states[(size_t)k * kNglobal + q]))
        return fail(GRW_EINVAL, "non-finite initial state of cell " + std::to_string(k));
  double* dic = nullptr;
  HIPCHK(hipMalloc((void**)&dic, icc.size() * sizeof(double) + ich.size() * sizeof(uint32_t)));
  uint32_t* dich = (uint32_t*)(dic + icc.size());
  hipError_t e = hipMemcpy(dic, icc.data(), icc.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dich, ich.data(), ich.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  SimArgs a = base_args(c);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_load_ic, dim3(c->grid_b), dim3(256), 0, c->stream, a, c->d, dic, dich, nic);
    hipLaunchKernelGGL(k_seed_cells, dim3(c->grid_b), dim3(256), 0, c->stream, a, c->d);
    hipLaunchKernelGGL(k_save_all, dim3((c->ld + 255) / 256), dim3(256), 0, c->stream, a, c->d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dic);
  if (e != hipSuccess) return fail(GRW_EDEVICE, std::string("initial state upload failed: ") + hipGetErrorString(e));
  reset_tracking(c);
  if (int rc = reset_order(c)) return rc;
  for (int k = 0; k < c->ncells; ++k) {
    RkCtl& hc = c->hctl[k];
    const GrwProtocol& p = hc.prot;
    for (int q = 0; q < kNglobal; ++q) {
      hc.errw[q] = errw[q];
      hc.state[q] = states[(size_t)k * kNglobal + q];
    }
    if (p.vclamp) hc.state[0] = p.vclamp_hold;  // initialize_state.c:222-228
    if (p.ivclamp) hc.state[0] = p.iv_hold;
    hc.tcall = hc.start = c->t0s[k];
    hc.tstep = c->t0s[k] + p.time_step;
    hc.h = hc.oh = std::min(p.step_max, p.time_step);
    hc.success = 1;
    hc.keepc = 1;
    hc.resume = 0;
    hc.stage = 0;
    hc.active = 0;
  }
  return 0;
}

int grw_batch_run(grw_ctx* c, const int32_t* nsteps, double* states, GrwRkStats* stt) {
  if (!c || !nsteps) return fail(GRW_EINVAL, "null argument");
  if (!c->batch) return fail(GRW_EINVAL, "grw_batch_run needs a batch context (grw_batch_create)");
  for (int k = 0; k < c->ncells; ++k) {
    if (nsteps[k] < 0) return fail(GRW_EINVAL, "negative step count for cell " + std::to_string(k));
    RkCtl& hc = c->hctl[k];
    hc.active = nsteps[k] > 0;
    hc.ms_base = hc.ms_done;
    hc.ms_target = hc.ms_done + nsteps[k];
    hc.attempts = hc.rejects = hc.forced = 0;
    hc.err = 0;
    hc.log_cap = hc.log_n = 0;
    hc.out_states = hc.out_cur = hc.out_other = hc.out_log = nullptr;
  }
  int rc = run_device(c);
  if (rc) return rc;
  for (int k = 0; k < c->ncells; ++k) {
    const RkCtl& hc = c->hctl[k];
    if (states)
      for (int q = 0; q < kNglobal; ++q) states[(size_t)k * kNglobal + q] = hc.state[q];
    if (stt) {
      stt[k].attempts += hc.attempts;
      stt[k].rejects += hc.rejects;
      stt[k].forced += hc.forced;
    }
  }
  return 0;
}

int grw_batch_get_beats(grw_ctx* c, int cell, double* out, int cap, int* n, int include_current) {
  if (!c || !n || (cap > 0 && !out)) return fail(GRW_EINVAL, "null argument");
  if (cell < 0 || cell >= c->ncells) return fail(GRW_EINVAL, "cell index out of range");
  const RkCtl& hc = c->hctl[cell];
  const int done = std::min(std::min(hc.beat_n, c->beat_cap), std::max(cap, 0));
  if (done > 0)
    HIPCHK(hipMemcpy(out, c->dbeats + (size_t)cell * c->beat_cap * kBeatStats,
                     sizeof(double) * (size_t)done * kBeatStats, hipMemcpyDeviceToHost));
  int m = done;
  if (include_current && hc.beat_cur >= 0 && m < cap) {
    for (int q = 0; q < kBeatStats; ++q) out[(size_t)m * kBeatStats + q] = hc.bt[q];
    ++m;
  }
  *n = m;
  return 0;
}

}  // extern "C"


// ---------------------------------------------------------------------------
// Autotuner: the device integrator's kernel layouts (grw_set_layout) and graph
// replays give the same results bit for bit, so the choice is a matter of speed
// alone, and that depends on the workload (FRU counts, precision, number of
// cells).  The tuner times a short run of the caller's own workload under each
// candidate, every one from the same snapshot of the context's device state,
// restores the snapshot and keeps the fastest.
namespace {

// Every device array a run of the device integrator reads or writes and keeps
// (the per-launch scratch -- partials, tickets, out -- excepted), with the
// host copies of the control blocks and the counters.
struct CtxSnapshot {
  std::vector<std::pair<void*, size_t>> regions;  // device address, bytes
  char* mem = nullptr;
  std::vector<RkCtl> hctl;
  grw_stats stats{};
  ~CtxSnapshot() {
    if (mem) (void)hipFree(mem);
  }
};

int snapshot_take(grw_ctx* c, CtxSnapshot& s) {
  const size_t ld = (size_t)c->ld, nc = (size_t)c->ncells;
  s.regions = {{c->d.conc, 5 * ld * sizeof(double)},
               {c->d.chan, 4 * ld * sizeof(uint32_t)},
               {c->d.mtidx, ld * sizeof(int32_t)},
               {c->d.nreg, ld * sizeof(int32_t)},
               {c->d.mt, ld * kMtStride * sizeof(uint32_t)},
               {c->d.mtw_h, ld * sizeof(uint32_t)},
               {c->d.conc_h, 5 * ld * sizeof(double)},
               {c->d.chan_h, 4 * ld * sizeof(uint32_t)},
               {c->d.mtidx_h, ld * sizeof(int32_t)},
               {c->d.nsteps, (kLaneUse + 1) * sizeof(unsigned long long)},
               {c->d.err, sizeof(int)},
               {c->d.maxsteps, sizeof(unsigned int)},
               {c->d.critsteps, sizeof(unsigned long long)},
               {c->d.tailticks, sizeof(unsigned long long)},
               {c->d.perm, ld * sizeof(int32_t)},
               {c->d.cost, ld * sizeof(uint32_t)},
               {c->d.bkey, (size_t)std::max(c->grid_s, 1) * sizeof(int32_t)},
               {c->d.border, (size_t)std::max(c->grid_s, 1) * sizeof(int32_t)},
               {c->dctl, nc * sizeof(RkCtl)},
               {c->dbeats, nc * (size_t)c->beat_cap * kBeatStats * sizeof(double)}};
  size_t total = 0;
  for (auto& r : s.regions)
    if (r.first) total += (r.second + 255) / 256 * 256;
  HIPCHK(hipMalloc((void**)&s.mem, std::max<size_t>(total, 256)));
  size_t off = 0;
  for (auto& r : s.regions) {
    if (!r.first || !r.second) continue;
    HIPCHK(hipMemcpyAsync(s.mem + off, r.first, r.second, hipMemcpyDeviceToDevice, c->stream));
    off += (r.second + 255) / 256 * 256;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  s.hctl = c->hctl;
  s.stats = c->stats;
  return 0;
}

int snapshot_restore(grw_ctx* c, const CtxSnapshot& s) {
  size_t off = 0;
  for (auto& r : s.regions) {
    if (!r.first || !r.second) continue;
    HIPCHK(hipMemcpyAsync(r.first, s.mem + off, r.second, hipMemcpyDeviceToDevice, c->stream));
    off += (r.second + 255) / 256 * 256;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->hctl = s.hctl;
  c->stats = s.stats;
  return 0;
}

// Candidates in a fixed order, the default first; a warm-up run of one output
// step (graph capture, code loading) precedes each timed run, both from the
// snapshot.
constexpr double kTuneMargin = 0.01;
int autotune(grw_ctx* c, const std::vector<int>& layouts, const std::function<int(int)>& run, int trial_steps,
             GrwTune* out) {
  const bool ws0 = c->ws, g0 = c->graphs;
  const int occ0 = c->occ;
  const int layout0 = c->poll     ? GRW_LAYOUT_QUADP
                      : c->refill ? (occ0 == 3 ? GRW_LAYOUT_REFILL_3 : GRW_LAYOUT_REFILL)
                      : c->queue  ? GRW_LAYOUT_QUEUE
                      : c->wave1  ? (occ0 == 3 ? GRW_LAYOUT_WAVE1_3 : occ0 == 4 ? GRW_LAYOUT_WAVE1_4 : GRW_LAYOUT_WAVE1)
                      : ws0       ? GRW_LAYOUT_WAVES
                      : occ0 == 3 ? GRW_LAYOUT_QUAD3
                      : occ0 == 4 ? GRW_LAYOUT_QUAD4
                                  : GRW_LAYOUT_QUAD;
  CtxSnapshot snap;
  if (int rc = snapshot_take(c, snap)) return rc;
  GrwTune r{};
  double best = 0.0;
  int rc = 0;
  for (int lay : layouts) {
    for (int g = 0; g < 2 && rc == 0; ++g) {
      if (r.ncand >= GRW_TUNE_MAX) break;
      if ((rc = grw_set_layout(c, lay)) || (rc = grw_set_graphs(c, g))) break;
      if ((rc = snapshot_restore(c, snap)) || (rc = run(1)) || (rc = snapshot_restore(c, snap))) break;
      const auto t0 = std::chrono::steady_clock::now();
      if ((rc = run(trial_steps))) break;
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      r.cand_layout[r.ncand] = lay;
      r.cand_graphs[r.ncand] = g;
      r.cand_ms[r.ncand] = ms;
      // a candidate replaces the one kept so far only if it is faster by more
      // than the trials' run-to-run spread, so that ties keep the earlier
      // (default) setting
      if (r.ncand == 0 || ms < best * (1.0 - kTuneMargin)) {
        best = ms;
        r.layout = lay;
        r.graphs = g;
      }
      ++r.ncand;
    }
  }
  const int rc2 = snapshot_restore(c, snap);
  if (rc || rc2 || r.ncand == 0) {  // leave the context as it was
    std::string msg = grw_last_error();
    (void)grw_set_layout(c, layout0);
    (void)grw_set_graphs(c, g0 ? 1 : 0);
    return fail(rc ? rc : rc2 ? rc2 : GRW_EINVAL, "autotune: " + msg);
  }
  if ((rc = grw_set_layout(c, r.layout)) || (rc = grw_set_graphs(c, r.graphs))) return rc;
  if (out) *out = r;
  return 0;
}

}  // namespace

extern "C" {

int grw_autotune_cell(grw_ctx* c, const GrwProtocol* p, const double errw[37], double t0, const double state[37],
                      double h, int trial_steps, GrwTune* out) {
  if (int rc = single(c)) return rc;
  if (!p || !errw || !state || trial_steps < 1 || !(h > 0.0)) return fail(GRW_EINVAL, "bad argument to grw_autotune_cell");
  auto run = [&](int nsteps) {
    double st[kNglobal];
    for (int q = 0; q < kNglobal; ++q) st[q] = state[q];
    double hh = h;
    int nlog = 0;
    return grw_run_cell_device(c, p, errw, t0, nsteps, st, &hh, nullptr, nullptr, nullptr, nullptr, 0, &nlog,
                               nullptr);
  };
  // the polling variant first: it is never slower than the four-lane kernel
  // it derives from (launches shorter by ~2 us), so ties keep it
  return autotune(c, {GRW_LAYOUT_QUADP, GRW_LAYOUT_QUAD, GRW_LAYOUT_WAVES, GRW_LAYOUT_QUAD3, GRW_LAYOUT_WAVE1}, run,
                  trial_steps, out);
}

int grw_autotune_batch(grw_ctx* c, int trial_steps, GrwTune* out) {
  if (!c) return fail(GRW_EINVAL, "null context");
  if (!c->batch) return fail(GRW_EINVAL, "grw_autotune_batch acts on a batch context (grw_batch_create)");
  if (trial_steps < 1) return fail(GRW_EINVAL, "trial_steps must be positive");
  auto run = [&](int nsteps) {
    std::vector<int32_t> ns((size_t)c->ncells, nsteps);
    return grw_batch_run(c, ns.data(), nullptr, nullptr);
  };
  return autotune(c,
                  {GRW_LAYOUT_QUAD, GRW_LAYOUT_QUAD3, GRW_LAYOUT_QUAD4, GRW_LAYOUT_WAVE1, GRW_LAYOUT_WAVE1_3,
                   GRW_LAYOUT_WAVE1_4, GRW_LAYOUT_REFILL, GRW_LAYOUT_REFILL_3},
                  run, trial_steps, out);
}

}  // extern "C"
