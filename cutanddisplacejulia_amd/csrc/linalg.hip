// Blocked right-looking LU with partial pivoting (the dgetrf behind `A\b`, SlipCalculator3D.jl:22) and
// blocked triangular solves, for gfx950.
//
//   panel   : one cooperative launch per 64-column panel.  The panel rows are split over the workgroups, one row
//             per thread, held in registers; each column needs one grid-wide exchange of pivot candidates (none
//             when the panel fits one workgroup), after which every thread applies the rank-1 update to its row.
//   laswp   : the panel's row interchanges applied to the columns outside the panel.
//   trsm    : U12 = L11^-1 A12 with L11 staged in LDS.
//   update  : A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), 128x128 tiles, K
//             streamed through double-buffered LDS.  The product is formed as (U^T L^T)^T so each lane's accumulator holds a column
//             segment of A22, making the read-modify-write of A22 column-contiguous.
// Matrices may carry extra right-hand-side columns (ncols > n): they are swapped, solved and updated with the
// rest, so factoring [A | B] leaves L^-1 P B in place of B.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <type_traits>
#include <utility>
#include <vector>

#include "linalg.hpp"
#include "mixed.hpp"

namespace cg = cooperative_groups;

namespace cdj {

namespace {

constexpr int PT = 256;        // panel threads
constexpr int MAX_P = 1024;    // max panel workgroups
constexpr int MAX_RPB = PT;    // panel rows per workgroup, one per thread
constexpr int LU_OUTER = 256;      // columns per outer block (rank of the trailing update)
constexpr int LU_OUTER_BIG = 512;  // the same for n >= LU_BIG: the trailing matrix is read and written half as often
constexpr int64_t LU_BIG = 32768;  // (measured: 512 is faster at n = 48672, slower at 2485, 15000 and 19602)
inline int64_t lu_outer(int64_t n) { return n >= LU_BIG ? LU_OUTER_BIG : LU_OUTER; }
constexpr int PERM_STRIDE = 1 + 4 * LU_NB;                        // ints of one panel's row permutation
constexpr int PERM_INTS = (LU_OUTER_BIG / LU_NB) * PERM_STRIDE;   // the permutations of one outer block's panels
constexpr int PERM_SETS = 2;  // two blocks' permutations (one block's are still read while the next is factored)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: A operand lane l = A[l % 16][l / 16], B operand lane l = B[l / 16][l % 16].  The accumulator
// layouts differ: f32 item v of lane l is D[4 (l / 16) + v][l % 16]; f64 item v of lane l is D[4 v + l / 16][l % 16].
// row(g, v) gives the D row of item v for lane group g = l / 16.
template <class T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 acc;
  __device__ static acc step(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int v) { return 4 * v + g; }
};
template <> struct Mfma<float> {
  typedef f4 acc;
  __device__ static acc step(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int v) { return 4 * g + v; }
};

// Panel factorization: one cooperative launch per panel of nbw <= 64 columns; the panel's m rows are split over P
// co-resident workgroups, one row per thread.
//
// Rows never move.  Each thread keeps its row in registers, shifted so that x[0] is always the current column j
// (x[i] is column j + i while i < nbw - j), and tracks the row's current position `pos`.  The row interchange of
// step j (position j <-> pivot position) is a relabelling: the pivot row takes position j, the row that held
// position j takes the pivot's old position.  Each step's shift frees x[63], which receives the step's multiplier:
// after t steps multiplier c sits in x[64 - t + c].  The pivot row used by the shift is zero beyond its active
// part, so the shift carries the multipliers along unchanged.  A pivot row stops shifting and keeps its U row in
// x[0..) with its multipliers behind it; every thread writes its finished row to its final position at the end.
// No LDS holds rows or multipliers (about 5 KiB per workgroup in all), so a panel workgroup fits on a CU beside
// other kernels' workgroups.  The arithmetic, the pivot rule (largest |a|, ties to the lowest position, as idamax)
// and hence the result are those of the row-moving algorithm, bit for bit (a zero may change sign).
//
// Per column: each wave finds its candidate with DPP maxima of the 64-bit key bits(|a|) (a non-negative double's
// bit pattern orders like its value) and stages the candidate row in LDS; after one workgroup barrier every thread
// knows the workgroup's candidate.  With one workgroup that is the pivot.  With several, wave 0 publishes the
// candidate and polls the others (below), then a second barrier hands the pivot row to the workgroup.
//
// Grid exchange (P > 1).  Each workgroup owns a slot word per column parity: the key with its low 12 bits replaced
// by c mod 4096 (c = global column), or the top bit set when it has no candidate, and a record of tagged words
// (32-bit payload, 32-bit tag c): the key, the candidate's position and its row.  All are relaxed agent-scope
// atomics: a reader takes a word only when its tag is c, so no fences are needed and a record read can never see
// another column's data.  Wave 0 polls the slots until all carry tag c (this is the grid barrier), picks the
// largest key (equal truncated keys are settled from the records' exact keys, then the lowest position) and reads
// the winner's record in one batch.  A slot or record is rewritten two columns later by the same workgroup, which
// can only happen after every workgroup has finished reading it; slots and records start all ones before a
// factorization.  Waits are bounded (about a second: co-resident workgroups publish within microseconds); on
// overrun the error flag is set and the kernel carries on, so it always exits (the host then reports the failure).
constexpr uint64_t SLOT_TAG = 0xFFFull;
constexpr uint64_t SLOT_NONE = 1ull << 63;
constexpr int XREC = 3 + 2 * LU_NB;              // tagged words per record: key (2), position (1), row (2 per entry)

template <int CTRL>
__device__ __forceinline__ long long dpp64(long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned long long)v, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)((unsigned long long)v >> 32), CTRL, 0xF, 0xF, false);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ long long readlane64(long long v, int l) {
  const int lo = __builtin_amdgcn_readlane((int)(unsigned long long)v, l);
  const int hi = __builtin_amdgcn_readlane((int)((unsigned long long)v >> 32), l);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ long long kmax64(long long a, long long b) { return a > b ? a : b; }
// value of v in lane l (l a compile-time constant), through scalar registers
__device__ __forceinline__ float bcast_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bcast_lane(double v, int l) {
  return __longlong_as_double(readlane64(__double_as_longlong(v), l));
}

// wave-wide maximum, in every lane: DPP within rows of 16 lanes, then the four row maxima
__device__ __forceinline__ long long wave_max(long long k) {
  k = kmax64(k, dpp64<0xB1>(k));   // quad_perm [1,0,3,2]
  k = kmax64(k, dpp64<0x4E>(k));   // quad_perm [2,3,0,1]
  k = kmax64(k, dpp64<0x141>(k));  // row_half_mirror
  k = kmax64(k, dpp64<0x140>(k));  // row_mirror
  return kmax64(kmax64(readlane64(k, 0), readlane64(k, 16)), kmax64(readlane64(k, 32), readlane64(k, 48)));
}

__device__ __forceinline__ void put_tagged(unsigned long long* p, uint32_t tag, uint32_t v) {
  __hip_atomic_store(p, ((unsigned long long)tag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double join_words(unsigned long long lo, unsigned long long hi) {
  return __longlong_as_double((long long)((hi << 32) | (lo & 0xFFFFFFFFull)));
}

// QPL: slots polled per lane (P <= 64 QPL)
// The panel of one matrix, factored by workgroup b of the P workgroups assigned to it.
template <class T, int QPL>
__device__ __forceinline__ void panel_body(T* A, int64_t lda, int64_t k, int64_t m, int nbw, int rpb, int32_t* ipiv,
                                           int* info, unsigned long long* xw, unsigned long long* slots, int* bar_err,
                                           int* perm, const int P, const int b, const T* __restrict__ save) {
  __shared__ T upad[2][LU_NB];               // the pivot row, zero beyond the active part, by column parity
  __shared__ long long wk[2][PT / 64];       // each wave's candidate key (-1: none) and position
  __shared__ int wp[2][PT / 64];
  __shared__ int s_prow[2];
  __shared__ int s_piv[LU_NB];  // pivot positions of this panel (relative to k)
  __shared__ int s_info;        // first column with a zero (or NaN) pivot, as a 1-based global index; 0 if none
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = (int64_t)b * rpb;
  const int64_t r1 = r0 + rpb < m ? r0 + rpb : m;
  const int nr = r1 > r0 ? (int)(r1 - r0) : 0;
  const bool mine = tid < nr;
  int pos = (int)(r0 + tid);  // current position of this thread's row
  bool done = !mine;          // row already taken as a pivot (or no row)
  bool pivot = false;         // this row is U row `pos` (chosen at step pos)
  if (tid == 0) s_info = 0;
  T x[LU_NB];
  {
    // unconditional loads from clamped (in-range) addresses; the clamped values are never used.  save (after the
    // diagonal-pivot fast path found another pivot row): the panel's original rows, m x nbw with leading dimension
    // m, which the fast path copied before changing them in A
    const int64_t rel = mine ? pos : (m > 0 ? m - 1 : 0);
    const T* const src = save ? save : A + k + k * lda;
    const int64_t lds = save ? m : lda;
#pragma unroll
    for (int c = 0; c < LU_NB; ++c) x[c] = src[rel + (c < nbw ? c : nbw - 1) * lds];
  }
  for (int j = 0; j < nbw; ++j) {
    const int par = j & 1;
    const uint64_t tag = (uint64_t)(k + j) & SLOT_TAG;
    const uint32_t tg = (uint32_t)(k + j);
    const int na = nbw - j;  // active columns
    // wave candidate: largest key, ties to the lowest position
    const long long key = done ? -1ll : (long long)__double_as_longlong(fabs((double)x[0]));
    const long long kw = wave_max(key);
    int wl = 0;  // this wave's candidate lane
    if (kw >= 0) {
      unsigned long long msk = __ballot(key == kw);
      wl = __ffsll((long long)msk) - 1;
      if (__popcll(msk) > 1) {
        int bp = __builtin_amdgcn_readlane(pos, wl);
        for (msk &= msk - 1; msk; msk &= msk - 1) {
          const int l2 = __ffsll((long long)msk) - 1;
          const int p2 = __builtin_amdgcn_readlane(pos, l2);
          if (p2 < bp) {
            bp = p2;
            wl = l2;
          }
        }
      }
      if (lane == wl) {
        wk[par][wave] = kw;
        wp[par][wave] = pos;
      }
    } else if (lane == 0) {
      wk[par][wave] = -1;
      wp[par][wave] = 0;
    }
    __syncthreads();
    // workgroup candidate (uniform)
    long long bk = wk[par][0];
    int bpos = wp[par][0], bw = 0;
#pragma unroll
    for (int w = 1; w < PT / 64; ++w) {
      const long long k2 = wk[par][w];
      const int p2 = wp[par][w];
      if (k2 > bk || (k2 == bk && k2 >= 0 && p2 < bpos)) {
        bk = k2;
        bpos = p2;
        bw = w;
      }
    }
    // the candidate's wave stages its row in upad (zero beyond the active part: the entries the candidate lane
    // writes past it are overwritten by the lanes >= na after it, in order, as one wave's LDS accesses are) and,
    // with several workgroups, runs the exchange; only the winning wave's lane writes a row to LDS
    if (wave == bw) {
      if (bk >= 0 && lane == wl) {
#pragma unroll
        for (int c8 = 0; c8 < LU_NB; c8 += 8)
          if (c8 < na) {
#pragma unroll
            for (int i = c8; i < c8 + 8; ++i) upad[par][i] = x[i];
          }
      }
      if (lane >= na) upad[par][lane] = T(0);
      if (P == 1) {
        if (lane == 0) {
          s_piv[j] = bk >= 0 ? bpos : j;
          const double gv = bk >= 0 ? __longlong_as_double(bk) : -1.0;
          if (!(gv > 0.0) && s_info == 0) s_info = (int)(k + j + 1);
        }
      } else {
        unsigned long long* rec = xw + (size_t)(par * MAX_P + b) * XREC;
        if (bk >= 0 && lane < na) {
          const unsigned long long xv = (unsigned long long)__double_as_longlong((double)upad[par][lane]);
          put_tagged(rec + 3 + 2 * lane, tg, (uint32_t)xv);
          put_tagged(rec + 4 + 2 * lane, tg, (uint32_t)(xv >> 32));
        }
        if (lane == 0) {
          put_tagged(rec, tg, (uint32_t)(unsigned long long)bk);
          put_tagged(rec + 1, tg, (uint32_t)((unsigned long long)bk >> 32));
          put_tagged(rec + 2, tg, (uint32_t)bpos);
          const uint64_t word = bk >= 0 ? (((uint64_t)bk & ~SLOT_TAG) | tag) : (SLOT_NONE | tag);
          __hip_atomic_store(&slots[par * MAX_P + b], (unsigned long long)word, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        uint64_t wq[QPL];
        long long it = 0;
        for (;;) {
          bool seen = true;
#pragma unroll
          for (int i = 0; i < QPL; ++i) {
            const int q = lane + 64 * i;
            wq[i] = q < P ? ld_agent(&slots[par * MAX_P + q]) : (SLOT_NONE | tag);
            seen = seen && (wq[i] & SLOT_TAG) == tag;
          }
          if (__all(seen)) break;
          if (++it > (1ll << 20)) {
            if (lane == 0) atomicOr(bar_err, 1);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        // the largest truncated key, ties to the lowest workgroup: one DPP maximum of (key | 1023 - q)
        long long vq = -1;
#pragma unroll
        for (int i = 0; i < QPL; ++i) {
          const int q = lane + 64 * i;
          if (q < P && !(wq[i] & SLOT_NONE)) vq = kmax64(vq, (long long)((wq[i] & ~SLOT_TAG) | (uint64_t)(1023 - q)));
        }
        const long long vmax = wave_max(vq);
        int gw = vmax >= 0 ? 1023 - (int)(vmax & 1023) : -1;
        const uint64_t kmax = vmax >= 0 ? ((uint64_t)vmax & ~SLOT_TAG) : 0;
        double gv = -1.0;
        int gr = -1;
        int ntie = 0;
#pragma unroll
        for (int i = 0; i < QPL; ++i) {
          const int q = lane + 64 * i;
          ntie += __popcll(__ballot(q < P && !(wq[i] & SLOT_NONE) && (wq[i] & ~SLOT_TAG) == kmax));
        }
        if (gw >= 0 && ntie > 1) {  // equal truncated keys: exact keys, then the lowest position
          gw = -1;
#pragma unroll
          for (int i = 0; i < QPL; ++i) {
            unsigned long long msk = __ballot(lane + 64 * i < P && !(wq[i] & SLOT_NONE) && (wq[i] & ~SLOT_TAG) == kmax);
            while (msk) {
              const int q = 64 * i + __ffsll((long long)msk) - 1;
              msk &= msk - 1;
              const unsigned long long* rw = xw + (size_t)(par * MAX_P + q) * XREC;
              unsigned long long w0, w1, w2;
              for (long long t = 0;; ++t) {
                w0 = ld_agent(rw);
                w1 = ld_agent(rw + 1);
                w2 = ld_agent(rw + 2);
                if ((uint32_t)(w0 >> 32) == tg && (uint32_t)(w1 >> 32) == tg && (uint32_t)(w2 >> 32) == tg) break;
                if (t > (1ll << 20)) {
                  if (lane == 0) atomicOr(bar_err, 2);
                  break;
                }
                __builtin_amdgcn_s_sleep(1);
              }
              const double vv = join_words(w0, w1);
              const int rq = (int)(uint32_t)w2;
              if (gw < 0 || vv > gv || (vv == gv && rq < gr)) {
                gv = vv;
                gr = rq;
                gw = q;
              }
            }
          }
        }
        if (gw == b) {  // this workgroup's candidate: its row is in upad already
          if (ntie <= 1) {
            gv = __longlong_as_double(bk);
            gr = bpos;
          }
        } else if (gw >= 0) {
          // the winner's record in one batch: lane i its row entry i, lane 63 also key and position
          const unsigned long long* rw = xw + (size_t)(par * MAX_P + gw) * XREC;
          const bool need = lane < na;
          for (long long t = 0;; ++t) {
            bool ok = true;
            unsigned long long a0 = 0, a1 = 0;
            if (need) {
              a0 = ld_agent(rw + 3 + 2 * lane);
              a1 = ld_agent(rw + 4 + 2 * lane);
              ok = (uint32_t)(a0 >> 32) == tg && (uint32_t)(a1 >> 32) == tg;
            }
            unsigned long long w0 = 0, w1 = 0, w2 = 0;
            if (lane == 63) {
              w0 = ld_agent(rw);
              w1 = ld_agent(rw + 1);
              w2 = ld_agent(rw + 2);
              ok = ok && (uint32_t)(w0 >> 32) == tg && (uint32_t)(w1 >> 32) == tg && (uint32_t)(w2 >> 32) == tg;
            }
            if (__all(ok)) {
              upad[par][lane] = need ? (T)join_words(a0, a1) : T(0);
              if (lane == 63) {
                gv = join_words(w0, w1);
                gr = (int)(uint32_t)w2;
              }
              break;
            }
            if (t > (1ll << 20)) {
              if (lane == 0) atomicOr(bar_err, 2);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          gv = __shfl(gv, 63);
          gr = __shfl(gr, 63);
        }
        if (lane == 0) {
          s_prow[par] = gr >= 0 ? gr : j;
          s_piv[j] = gr >= 0 ? gr : j;
          if (!(gv > 0.0) && s_info == 0) s_info = (int)(k + j + 1);
        }
      }
    }
    __syncthreads();
    const int prow = P == 1 ? (bk >= 0 ? bpos : j) : s_prow[par];
    // the pivot row, lane i holding its column j + i (zero beyond the active part); its entries reach the update
    // as scalar operands (lane reads), not as 64-lane LDS broadcasts.  u is loaded and its lane 0 read by every
    // lane before the divergent update below, which reads the other lanes of the same register.
    const T u = upad[par][lane];
    const T piv = bcast_lane(u, 0);
    // the pivot row becomes U row j; every other active row takes its multiplier and the rank-1 update
    if (!done) {
      if (pos == prow) {
        pos = j;
        done = true;
        pivot = true;
      } else {
        if (pos == j) pos = prow;
        T l = x[0];
        if (piv != T(0)) {
          l = x[0] / piv;
          // u is zero beyond the active part, so x - l * 0 moves the multipliers (and the unused tail) along
          // unchanged; only a row that already holds an infinite or NaN value (then l * 0 is NaN) differs from
          // the row-moving algorithm, in that row's earlier multipliers
          // the first half of the row's entries reach every lane as lane reads into scalar registers (VALU),
          // the second half as LDS broadcasts: the two paths run side by side, each carrying half the row
#pragma unroll
          for (int i0 = 0; i0 < LU_NB / 2; i0 += 8) {
            T ub[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) ub[e] = bcast_lane(u, i0 + e + 1);
#pragma unroll
            for (int e = 0; e < 8; ++e) x[i0 + e] = x[i0 + e + 1] - l * ub[e];
          }
          const T* uh = upad[par];
#pragma unroll
          for (int i = LU_NB / 2; i + 1 < LU_NB; ++i) x[i] = x[i + 1] - l * uh[i + 1];
        } else {
#pragma unroll
          for (int i = 0; i + 1 < LU_NB; ++i) x[i] = x[i + 1];
        }
        x[LU_NB - 1] = l;
      }
    }
  }
  __syncthreads();
  if (mine) {
    // a pivot row (chosen at step pos): U columns pos.. in x[0..), multiplier c < pos in x[64 - pos + c]; any
    // other row: multiplier c in x[64 - nbw + c]
    const int lead = pivot ? LU_NB - pos : LU_NB - nbw;  // register holding multiplier 0
    T* const out = A + (k + pos) + k * lda;
#pragma unroll
    for (int i = 0; i < LU_NB; ++i) {
      const int c = i >= lead ? i - lead : (pivot && i < nbw - pos ? pos + i : -1);
      if (c >= 0) out[c * lda] = x[i];
    }
  }
  if (b == 0 && tid < nbw) ipiv[k + tid] = (int32_t)(k + s_piv[tid]);
  if (b == 0 && tid == 0 && s_info != 0 && *info == 0) *info = s_info;
  // block 0, wave 0: the panel's swaps as one row permutation of the touched rows (rows j < nbw and the pivot
  // rows beyond the panel): perm[0] = count, then (destination, source) pairs, relative to k
  if (b == 0 && wave == 0) {
    __shared__ int touched[2 * LU_NB], cur[2 * LU_NB];
    __shared__ int s_cnt;
    if (lane < nbw) {
      touched[lane] = lane;
      cur[lane] = lane;
    }
    if (lane == 0) s_cnt = nbw;
    __builtin_amdgcn_wave_barrier();
    for (int j = 0; j < nbw; ++j) {
      const int p = s_piv[j];
      int cnt = s_cnt;
      // index of row p among the touched rows (rows < nbw are at their own index)
      int at = p < nbw ? p : -1;
      if (at < 0) {
        const bool hit = lane + nbw < cnt && touched[lane + nbw] == p;
        const unsigned long long mm = __ballot(hit);
        if (mm) at = nbw + __ffsll((long long)mm) - 1;
      }
      if (at < 0) {
        at = cnt;
        if (lane == 0) {
          touched[cnt] = p;
          cur[cnt] = p;
          s_cnt = cnt + 1;
        }
      }
      if (lane == 0 && at != j) {
        const int t = cur[j];
        cur[j] = cur[at];
        cur[at] = t;
      }
      __builtin_amdgcn_wave_barrier();
    }
    const int cnt = s_cnt;
    for (int t = lane; t < cnt; t += 64) {
      perm[1 + 2 * t] = touched[t];
      perm[2 + 2 * t] = cur[t];
    }
    if (lane == 0) perm[0] = cnt;
  }
}

// skip (may be NULL): the diagonal-pivot fast path's words {violation, zero-pivot column}; with no violation the
// fast path has factored the panel already, and the kernel only passes its zero-pivot column on to info.
template <class T, int QPL>
__global__ void __launch_bounds__(PT) lu_panel_kernel(T* A, int64_t lda, int64_t k, int64_t m, int nbw, int rpb,
                                                      int32_t* ipiv, int* info, unsigned long long* xw,
                                                      unsigned long long* slots, int* bar_err, int* perm,
                                                      int* skip, const T* save) {
  const T* src = nullptr;
  if (skip) {
    if (skip[0] == 0) {
      if (blockIdx.x == 0 && threadIdx.x == 0 && skip[1] != 0 && *info == 0) *info = skip[1];
      return;
    }
    if (skip[0] == 1) {  // the fast path changed the rows in A: factor the saved originals, count the fall-back
      src = save;
      if (blockIdx.x == 0 && threadIdx.x == 0) skip[2] += 1;
    }
  }
  panel_body<T, QPL>(A, lda, k, m, nbw, rpb, ipiv, info, xw, slots, bar_err, perm, (int)gridDim.x, (int)blockIdx.x,
                     src);
}

// ---- diagonal-pivot fast path of a panel ----------------------------------------------------------------
// Partial pivoting takes, at column j, the row of largest |a| (ties to the lowest position).  When that is the
// diagonal row at every column of the panel (the collocation matrices of this engine never pivot), the panel's
// factors are those of elimination without interchanges, which needs no grid-wide exchange: one wave factors the
// nbw x nbw diagonal block (fp_diag_kernel), then every row below solves against its U on its own
// (fp_rows_kernel).  Both compare each candidate with the diagonal as the exchange does (64-bit keys bits(|a|),
// ties to the diagonal) and raise fpw[0] when another row would have been taken; both first copy the rows they
// change to `save`.  On a violation lu_panel_kernel factors the panel with the exchange from the rows in `save` (its
// final write covers every entry the fast path changed); otherwise it returns at once.  Each entry receives the exchange panel's operations (the same
// divisions and FMAs in the same order), so the factors are bit-identical either way.
template <class T>
__device__ __forceinline__ long long abs_key(T v) {
  return (long long)__double_as_longlong(fabs((double)v));
}

// fpw: [0] 0 = no violation, 1 = violation (rows saved), 2 = fast path not tried; [1] the fast path's first zero
// pivot (1-based global column, or 0); [2] panels of this factorization that fell back.  After FP_GIVE_UP
// fall-backs a factorization stops trying the fast path (a matrix that pivots would pay for every attempt).
constexpr int FP_GIVE_UP = 2;

// Step J of the diagonal block (J a constant, so that every register index is one): the pivot row's entries reach
// the lanes as lane reads into scalar registers, a group at a time and before the products that use them (the reads
// then issue back to back); the rows below the diagonal divide by the pivot and subtract its multiple.  Each lane
// keeps a running violation flag (a row below holding a larger |a| than the pivot).
template <class T, bool FULL, int J>
__device__ __forceinline__ void fp_step(T (&x)[LU_NB], int lane, int nbw, int64_t k, bool mine, bool& viol,
                                        int& zinfo) {
  if (FULL) nbw = LU_NB;  // a full block: every column test below is a constant
  if (J >= nbw) return;  // uniform
  const T piv = bcast_lane(x[J], J);
  const long long pk = abs_key(piv);
  if (!(__longlong_as_double(pk) > 0.0) && zinfo == 0) zinfo = (int)(k + J + 1);
  const bool below = lane > J && mine;
  T l = x[J];
  if (below) {
    viol = viol || abs_key(x[J]) > pk;
    if (piv != T(0)) l = x[J] / piv;
  }
  constexpr int B = sizeof(T) == 8 ? 4 : 8;  // entries per group (FP64 entries take two scalar registers each)
  if (piv != T(0)) {
#pragma unroll
    for (int c0 = J + 1; c0 < LU_NB; c0 += B) {
      T u[B];
#pragma unroll
      for (int e = 0; e < B; ++e)
        if (c0 + e < LU_NB) u[e] = c0 + e < nbw ? bcast_lane(x[(c0 + e) & (LU_NB - 1)], J) : T(0);
      if (below) {
#pragma unroll
        for (int e = 0; e < B; ++e)
          if (c0 + e < LU_NB && c0 + e < nbw) x[(c0 + e) & (LU_NB - 1)] -= l * u[e];
      }
    }
  }
  if (below) x[J] = l;
}

template <class T, bool FULL, int... J>
__device__ __forceinline__ void fp_steps(T (&x)[LU_NB], int lane, int nbw, int64_t k, bool mine, bool& viol,
                                         int& zinfo, std::integer_sequence<int, J...>) {
  (fp_step<T, FULL, J>(x, lane, nbw, k, mine, viol, zinfo), ...);
}

// One wave factors the diagonal block: lane r holds row k + r, shifted as in panel_body; each step's pivot row is
// staged in LDS by its lane (zero beyond its active part) and read back as broadcasts.  Both forms write ipiv
// (identity), an empty permutation, fpw[0] (violation inside the block) and fpw[1].
template <class T>
__device__ __forceinline__ void fp_diag_lds(T* A, int64_t lda, int64_t k, int nbw, int32_t* ipiv, int* perm,
                                            int* fpw, T* save, int64_t lds) {
  __shared__ T urow[LU_NB];  // the pivot row of the current column, zero beyond its active part
  const int lane = threadIdx.x;
  const bool mine = lane < nbw;
  T x[LU_NB];
  {
    const int64_t row = k + (mine ? lane : 0);
#pragma unroll
    for (int c = 0; c < LU_NB; ++c) x[c] = A[row + (k + (c < nbw ? c : nbw - 1)) * lda];
  }
  if (mine) {
#pragma unroll
    for (int c = 0; c < LU_NB; ++c)
      if (c < nbw) save[lane + c * lds] = x[c];
  }
  bool done = !mine, viol = false;
  int zinfo = 0;
  for (int j = 0; j < nbw; ++j) {
    const int na = nbw - j;
    const long long key = done ? -1ll : abs_key(x[0]);
    const long long kmax = wave_max(key);
    const long long kj = readlane64(key, j);
    viol = viol || kmax > kj;
    if (!(__longlong_as_double(kj) > 0.0) && zinfo == 0) zinfo = (int)(k + j + 1);
    if (lane == j) {
#pragma unroll
      for (int c8 = 0; c8 < LU_NB; c8 += 8)
        if (c8 < na) {
#pragma unroll
          for (int i = c8; i < c8 + 8; ++i) urow[i] = x[i];
        }
    }
    if (lane >= na) urow[lane] = T(0);
    __builtin_amdgcn_wave_barrier();
    const T u = urow[lane];
    const T piv = bcast_lane(u, 0);
    __builtin_amdgcn_wave_barrier();
    if (!done) {
      if (lane == j) {
        done = true;
      } else {
        T l = x[0];
        if (piv != T(0)) {
          l = x[0] / piv;
#pragma unroll
          for (int i = 0; i + 1 < LU_NB; ++i) x[i] = x[i + 1] - l * urow[i + 1];
        } else {
#pragma unroll
          for (int i = 0; i + 1 < LU_NB; ++i) x[i] = x[i + 1];
        }
        x[LU_NB - 1] = l;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (mine) {
    // row `lane` was taken at step lane: U columns lane.. in x[0..), multiplier c < lane in x[64 - lane + c]
    const int lead = LU_NB - lane;
    T* const out = A + (k + lane) + k * lda;
#pragma unroll
    for (int i = 0; i < LU_NB; ++i) {
      const int c = i >= lead ? i - lead : (i < nbw - lane ? lane + i : -1);
      if (c >= 0) out[c * lda] = x[i];
    }
    ipiv[k + lane] = (int32_t)(k + lane);
  }
  if (lane == 0) {
    perm[0] = 0;
    fpw[0] = viol ? 1 : 0;
    fpw[1] = zinfo;
  }
}

// The same for a full block (nbw = 64) with lane r holding row k + r in registers unshifted, x[c] = column k + c,
// the steps unrolled (fp_step) and the pivot row's entries as grouped lane reads.
template <class T>
__device__ __forceinline__ void fp_diag_lanes(T* A, int64_t lda, int64_t k, int nbw, int32_t* ipiv, int* perm,
                                              int* fpw, T* save, int64_t lds) {
  const int lane = threadIdx.x;
  const bool mine = lane < nbw;
  T x[LU_NB];
  {
    const int64_t row = k + (mine ? lane : 0);
#pragma unroll
    for (int c = 0; c < LU_NB; ++c) x[c] = A[row + (k + (c < nbw ? c : nbw - 1)) * lda];
  }
  if (mine) {
#pragma unroll
    for (int c = 0; c < LU_NB; ++c)
      if (c < nbw) save[lane + c * lds] = x[c];
  }
  bool viol = false;
  int zinfo = 0;
  fp_steps<T, true>(x, lane, LU_NB, k, mine, viol, zinfo, std::make_integer_sequence<int, LU_NB>{});
  if (mine) {
    T* const out = A + (k + lane) + k * lda;
#pragma unroll
    for (int c = 0; c < LU_NB; ++c)
      if (c < nbw) out[c * lda] = x[c];
    ipiv[k + lane] = (int32_t)(k + lane);
  }
  const bool any = __any(viol);
  if (lane == 0) {
    perm[0] = 0;
    fpw[0] = any ? 1 : 0;
    fpw[1] = zinfo;
  }
}

// FULL (nbw == 64, every panel but possibly the last): the lane-read form, each kernel its own register
// allocation; a narrower last panel: the LDS-staged form.
template <class T, bool FULL>
__global__ void __launch_bounds__(64) fp_diag_kernel(T* A, int64_t lda, int64_t k, int nbw, int32_t* ipiv, int* perm,
                                                     int* fpw, T* save, int64_t lds) {
  if (fpw[2] >= FP_GIVE_UP) {
    if (threadIdx.x == 0) fpw[0] = 2;
    return;
  }
  if constexpr (FULL)
    fp_diag_lanes<T>(A, lda, k, LU_NB, ipiv, perm, fpw, save, lds);
  else
    fp_diag_lds<T>(A, lda, k, nbw, ipiv, perm, fpw, save, lds);
}

// Rows k + nbw + r (r < nrows): one thread per row, the row in registers, U of the block from LDS.  Column j
// divides the row's entry by u_jj (none when u_jj is 0) and subtracts the multiple of U's row j from the entries
// right of it: the operations the exchange panel applies to a row that is not taken.  The loads of U and of the
// row are issued before the first wait (FP32; in FP64 the row's after U is staged); an FP32 block without a zero
// pivot (fpw[1] == 0, from the diagonal block) takes a loop with no branch on the pivot, which the compiler schedules
// across columns.
template <class T>
__device__ __forceinline__ void fp_rows_body(T* A, int64_t lda, int64_t k, int nbw, int64_t nrows, int* fpw, T* save,
                                             int64_t lds, int64_t blk) {
  __shared__ T U[LU_NB][LU_NB + 1];  // U[j][c]: row j of the block's U, zero left of the diagonal and from nbw on
  __shared__ long long ukey[LU_NB];
  const int tid = threadIdx.x;
  if (fpw[0] == 2) return;
  constexpr int UPT = LU_NB * LU_NB / 256;  // U entries staged per thread
  T ureg[UPT];
#pragma unroll
  for (int i = 0; i < UPT; ++i) {
    const int idx = tid + i * 256, j = idx % LU_NB, c = idx / LU_NB;
    ureg[i] = A[(k + (j < nbw ? j : nbw - 1)) + (k + (c < nbw ? c : nbw - 1)) * lda];  // clamped: always valid
  }
  auto stage = [&] {
#pragma unroll
    for (int i = 0; i < UPT; ++i) {
      const int idx = tid + i * 256, j = idx % LU_NB, c = idx / LU_NB;
      U[j][c] = (c >= j && c < nbw) ? ureg[i] : T(0);
      if (j == c) ukey[j] = c < nbw ? abs_key(ureg[i]) : 0;
    }
  };
  if (sizeof(T) == 8) stage();  // FP64: U leaves registers before the row arrives (168 registers, no spill)
  const int64_t r = blk * 256 + tid;
  const bool mine = r < nrows;
  const int64_t row = k + nbw + (mine ? r : nrows - 1);
  T x[LU_NB];
#pragma unroll
  for (int c = 0; c < LU_NB; ++c) x[c] = A[row + (k + (c < nbw ? c : nbw - 1)) * lda];
  if (sizeof(T) == 4) stage();
  __syncthreads();
  if (mine) {
#pragma unroll
    for (int c = 0; c < LU_NB; ++c)
      if (c < nbw) save[(nbw + r) + c * lds] = x[c];
  }
  bool viol = false;
  // (FP32 only: in FP64 the compiler's schedule of this loop needs more than 168 registers and spills)
  if (sizeof(T) == 4 && fpw[1] == 0) {
#pragma unroll
    for (int j = 0; j < LU_NB; ++j) {
      if (j < nbw) {
        viol = viol || abs_key(x[j]) > ukey[j];
        const T l = x[j] / U[j][j];
#pragma unroll
        for (int c = j + 1; c < LU_NB; ++c) x[c] -= l * U[j][c];
        x[j] = l;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < LU_NB; ++j) {
      if (j < nbw) {
        viol = viol || abs_key(x[j]) > ukey[j];
        const T piv = U[j][j];
        T l = x[j];
        if (piv != T(0)) {
          l = x[j] / piv;
#pragma unroll
          for (int c = j + 1; c < LU_NB; ++c) x[c] -= l * U[j][c];
        }
        x[j] = l;
      }
    }
  }
  if (__any(mine && viol) && (tid & 63) == 0) atomicOr(&fpw[0], 1);
  if (mine) {
#pragma unroll
    for (int c = 0; c < LU_NB; ++c)
      if (c < nbw) A[row + (k + c) * lda] = x[c];
  }
}

// At most 168 registers per lane (three workgroups per CU): a workgroup then fits where one workgroup of a running
// trailing-update product retires, instead of waiting for a whole CU.
template <class T>
__global__ void __launch_bounds__(256, 3) fp_rows_kernel(T* A, int64_t lda, int64_t k, int nbw, int64_t nrows,
                                                         int* fpw, T* save, int64_t lds) {
  fp_rows_body<T>(A, lda, k, nbw, nrows, fpw, save, lds, (int64_t)blockIdx.x);
}

// The fast path for the panels of several matrices at the same column k (the batched factorization): job j owns
// workgroups [b0, b0 + nb) of a launch (nb = 1 for the diagonal blocks).
template <class T>
struct FpJob {
  T* A;
  int64_t lda, m;  // m: the panel's rows, [k, k + m)
  int32_t nbw, b0;
  int32_t* ipiv;
  int* perm;
  int* fpw;
  T* save;
};
constexpr int FP_JOBS = 24;
template <class T>
struct FpBatch {
  int njobs;
  FpJob<T> job[FP_JOBS];
};
template <class T>
__device__ __forceinline__ int fp_job_of(const FpBatch<T>& jb) {
  int j = 0;
  while (j + 1 < jb.njobs && (int)blockIdx.x >= jb.job[j + 1].b0) ++j;
  return j;
}

template <class T, bool FULL>
__global__ void __launch_bounds__(64) fp_diag_batched_kernel(const FpBatch<T> jb, int64_t k) {
  const FpJob<T>& J = jb.job[blockIdx.x];
  if (J.fpw[2] >= FP_GIVE_UP) {
    if (threadIdx.x == 0) J.fpw[0] = 2;
    return;
  }
  if constexpr (FULL)
    fp_diag_lanes<T>(J.A, J.lda, k, LU_NB, J.ipiv, J.perm, J.fpw, J.save, J.m);
  else
    fp_diag_lds<T>(J.A, J.lda, k, J.nbw, J.ipiv, J.perm, J.fpw, J.save, J.m);
}

template <class T>
__global__ void __launch_bounds__(256, 3) fp_rows_batched_kernel(const FpBatch<T> jb, int64_t k) {
  const FpJob<T>& J = jb.job[fp_job_of(jb)];
  fp_rows_body<T>(J.A, J.lda, k, J.nbw, J.m - J.nbw, J.fpw, J.save, J.m, (int64_t)blockIdx.x - J.b0);
}

// Panels of several matrices at the same column k in one cooperative launch: matrix j owns workgroups
// [b0, b0 + P) and exchanges its pivot candidates only among them, through its own workspace.
template <class T>
struct PanelJob {
  T* A;
  int64_t lda, m;
  int32_t nbw, rpb, P, b0;
  int32_t* ipiv;
  int* info;
  unsigned long long* xw;
  unsigned long long* slots;
  int* bar_err;
  int* perm;
  int* skip;      // the fast path's words (NULL: none), as in lu_panel_kernel
  const T* save;  // the fast path's copy of the panel's rows, as in lu_panel_kernel
};
constexpr int PANEL_JOBS = 24;  // matrices per launch (the job table travels in the kernel arguments)
template <class T>
struct PanelBatch {
  int njobs;
  PanelJob<T> job[PANEL_JOBS];
};

template <class T, int QPL>
__global__ void __launch_bounds__(PT) lu_panel_batched_kernel(const PanelBatch<T> jb, int64_t k) {
  int j = 0;
  while (j + 1 < jb.njobs && (int)blockIdx.x >= jb.job[j + 1].b0) ++j;
  const PanelJob<T>& J = jb.job[j];
  const T* src = nullptr;
  if (J.skip) {
    if (J.skip[0] == 0) {  // every workgroup of the job leaves, so none waits for another
      if ((int)blockIdx.x == J.b0 && threadIdx.x == 0 && J.skip[1] != 0 && *J.info == 0) *J.info = J.skip[1];
      return;
    }
    if (J.skip[0] == 1) {
      src = J.save;
      if ((int)blockIdx.x == J.b0 && threadIdx.x == 0) J.skip[2] += 1;
    }
  }
  panel_body<T, QPL>(J.A, J.lda, k, J.m, J.nbw, J.rpb, J.ipiv, J.info, J.xw, J.slots, J.bar_err, J.perm, J.P,
                     (int)blockIdx.x - J.b0, src);
}

// Row interchanges of one panel applied to columns [0, k) and [k + nbw, ncols), as the permutation of the
// touched rows emitted by the panel (perm = {count, (dst, src) pairs} relative to k).  Two columns per block,
// 128 threads each: all reads of a column complete before any write.
template <class T>
__device__ __forceinline__ void laswp_body(T* A, int64_t lda, int64_t ncols, int64_t k, int nbw,
                                           const int* __restrict__ perm, int64_t blk, int64_t nblocks) {
  const int cnt = perm[0];
  if (cnt == 0) return;  // no interchange (the diagonal-pivot fast path held): the workgroup leaves at once
  const int t = threadIdx.x % 128;
  // column pairs blk, blk + nblocks, ... (the same count for every thread of the workgroup)
  for (int64_t c2 = blk; 2 * c2 < ncols; c2 += nblocks) {
    const int64_t c = 2 * c2 + threadIdx.x / 128;
    const bool act = c < ncols && !(c >= k && c < k + nbw) && t < cnt;
    T v = T(0);
    int dst = 0;
    if (act) {
      dst = perm[1 + 2 * t];
      v = A[(k + perm[2 + 2 * t]) + c * lda];
    }
    __syncthreads();
    if (act) A[(k + dst) + c * lda] = v;
  }
}

template <class T>
__global__ void __launch_bounds__(256) laswp_kernel(T* A, int64_t lda, int64_t ncols, int64_t k, int nbw,
                                                    const int* __restrict__ perm) {
  laswp_body<T>(A, lda, ncols, k, nbw, perm, (int64_t)blockIdx.x, (int64_t)gridDim.x);
}

// The interchanges of one panel in the columns [c0, c1) only (the panel's own columns excepted).
template <class T>
__global__ void __launch_bounds__(256) laswp_range_kernel(T* A, int64_t lda, int64_t c0, int64_t c1, int64_t k,
                                                          int nbw, const int* __restrict__ perm) {
  const int64_t c = c0 + (int64_t)blockIdx.x * 2 + threadIdx.x / 128;
  const int t = threadIdx.x % 128;
  const int cnt = perm[0];
  const bool act = c < c1 && !(c >= k && c < k + nbw) && t < cnt;
  T v = T(0);
  int dst = 0;
  if (act) {
    dst = perm[1 + 2 * t];
    v = A[(k + perm[2 + 2 * t]) + c * lda];
  }
  __syncthreads();
  if (act) A[(k + dst) + c * lda] = v;
}

// The interchanges of the np panels of the block starting at column K0 (panel p at column K0 + LU_NB p, its
// permutation at perms + p * PERM_STRIDE), in panel order, in the columns [c0, c1) (outside the block).
template <class T>
__global__ void __launch_bounds__(256) laswp_deferred_kernel(T* A, int64_t lda, int64_t K0, int64_t c0, int64_t c1,
                                                             int np, const int* __restrict__ perms) {
  const int64_t c = c0 + (int64_t)blockIdx.x * 2 + threadIdx.x / 128;
  const int t = threadIdx.x % 128;
  for (int p = 0; p < np; ++p) {
    const int* perm = perms + p * PERM_STRIDE;
    const int64_t k = K0 + (int64_t)p * LU_NB;
    const int cnt = perm[0];
    const bool act = c < c1 && t < cnt;
    T v = T(0);
    int dst = 0;
    if (act) {
      dst = perm[1 + 2 * t];
      v = A[(k + perm[2 + 2 * t]) + c * lda];
    }
    __syncthreads();
    if (act) A[(k + dst) + c * lda] = v;
    __syncthreads();
  }
}

// The row interchanges of the panels at column k of several matrices (job j owns workgroups [b0, ...)).
template <class T>
struct SwapJob {
  T* A;
  int64_t lda, ncols;
  const int* perm;
  int32_t nbw, b0, nb;  // workgroups [b0, b0 + nb) of the launch
};
constexpr int SWAP_BLOCKS = 256;  // workgroups per matrix at most; each walks its share of the column pairs
constexpr int SWAP_JOBS = 48;
template <class T>
struct SwapBatch {
  int njobs;
  SwapJob<T> job[SWAP_JOBS];
};

template <class T>
__global__ void __launch_bounds__(256) laswp_batched_kernel(const SwapBatch<T> jb, int64_t k) {
  int j = 0;
  while (j + 1 < jb.njobs && (int)blockIdx.x >= jb.job[j + 1].b0) ++j;
  const SwapJob<T>& J = jb.job[j];
  laswp_body<T>(J.A, J.lda, J.ncols, k, J.nbw, J.perm, (int64_t)blockIdx.x - J.b0, (int64_t)J.nb);
}

// A[k:k+nbw, c0:c0+ncols] = L11^-1 A[k:k+nbw, c0:...] with L11 unit lower (the panel's diagonal block).  One
// wave per column at a time (cpw columns per wave): lane r holds x[r]; step q broadcasts x[q] and the lanes
// below update with L11 read from LDS (stored transposed, so a step's reads are consecutive).  A column's solve is
// a latency chain of nbw steps, so few columns get one column per wave (more workgroups, each staging L11).
template <class T>
__device__ __forceinline__ void trsm_body(T* A, int64_t lda, int64_t k, int nbw, int64_t c0, int64_t ncols, int cpw,
                                          int64_t blk) {
  __shared__ T LT[LU_NB][LU_NB + 1];  // LT[q][r] = L11[r][q]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int idx = tid; idx < LU_NB * LU_NB; idx += 256) {
    const int r = idx % LU_NB, q = idx / LU_NB;
    LT[q][r] = (r < nbw && q < nbw) ? A[(k + r) + (k + q) * lda] : T(0);
  }
  __syncthreads();
  const int64_t cb = c0 + (blk * 4 + wave) * cpw;
  const int64_t ce = std::min<int64_t>(cb + cpw, c0 + ncols);  // this wave's columns [cb, ce)
  if (nbw == LU_NB) {
    // full block: two columns at a time, their chains interleaved, the next pair's loads in flight during the
    // chains; steps unrolled, x[q] reaching the lanes as a lane read into scalar registers, the lanes' L11 entries
    // from LDS independently of the chain (the same FMAs in the same order as the generic loop below)
    auto ld = [&](int64_t c) { return c < ce ? A[k + lane + c * lda] : T(0); };
    T na0 = ld(cb), na1 = ld(cb + 1);
    for (int64_t c = cb; c < ce; c += 2) {
      T x0 = na0, x1 = na1;
      na0 = ld(c + 2);
      na1 = ld(c + 3);
#pragma unroll
      for (int q = 0; q < LU_NB - 1; ++q) {
        const T lq = LT[q][lane];
        const T x0q = bcast_lane(x0, q), x1q = bcast_lane(x1, q);
        if (lane > q) {
          x0 -= lq * x0q;
          x1 -= lq * x1q;
        }
      }
      A[k + lane + c * lda] = x0;
      if (c + 1 < ce) A[k + lane + (c + 1) * lda] = x1;
    }
    return;
  }
  for (int64_t c = cb; c < ce; ++c) {
    T* col = A + k + c * lda;
    T x = lane < nbw ? col[lane] : T(0);
    for (int q = 0; q < nbw - 1; ++q) {
      const T xq = __shfl(x, q);
      if (lane > q) x -= LT[q][lane] * xq;
    }
    if (lane < nbw) col[lane] = x;
  }
}

template <class T>
__global__ void __launch_bounds__(256) trsm_kernel(T* A, int64_t lda, int64_t k, int nbw, int64_t c0, int64_t ncols,
                                                   int cpw) {
  trsm_body<T>(A, lda, k, nbw, c0, ncols, cpw, (int64_t)blockIdx.x);
}


// C[0:M, 0:N] -= L[0:M, 0:K] * U[0:K, 0:N], column-major with leading dimensions ldl, ldu, ldc, any K.
// 128 x 128 tile per workgroup (4 waves in 2 x 2, 64 x 64 each = 4 x 4 MFMA tiles).  K is streamed in slices of
// GB_K through two LDS buffers: while slice s is multiplied, slice s + 1 goes from registers into the other
// buffer and slice s + 2 is fetched into registers, so each slice costs one workgroup barrier.  Loads are
// branch-free: row and column indices are clamped into the matrix (a clamped row or column only feeds C entries
// outside [0, M) x [0, N), which are not written) and k >= K reads as zero.  At most 256 VGPRs per lane, so two
// workgroups share a CU and one wave's stalls are covered by the other's MFMAs.  The product is formed as
// (U^T L^T)^T so each lane's accumulator items run down a column segment of C.
constexpr int GB_M = 128, GB_N = 128, GB_K = 16;

// Tile of workgroup b in a gx x gy tile grid.  Workgroups are dispatched round-robin over the 8 XCDs, each with
// its own L2: the tiles are renumbered so that every XCD owns one contiguous run of tiles, and the run is walked
// in groups of 8 row tiles per column tile, so the workgroups resident on one XCD at a time share L and U slices
// in that XCD's L2.
__device__ __forceinline__ void xcd_tile(int b, int gx, int gy, int& tx, int& ty) {
  const int T = gx * gy, q = T / 8, r = T % 8, x = b % 8;
  const int t = x * q + (x < r ? x : r) + b / 8;
  constexpr int G = 8;
  const int grp = t / (G * gy), first = grp * G;
  const int gs = gx - first < G ? gx - first : G;
  const int w = t - grp * G * gy;
  tx = first + w % gs;
  ty = w / gs;
}

// vec (decided by the host): L and U are aligned to two elements with even leading dimensions and K is a multiple
// of GB_K, so an interior tile fetches its operands as pairs of elements, half the load instructions of the
// element-wise fetch (the FP64 product is bound by operand load issue: 43.9 -> 52.4 TFLOP/s at 16128^2 x 256).
// The staged LDS contents, and so the results, are the same either way.
template <class T>
__device__ __forceinline__ void gemm_sub_tile(const T* __restrict__ L, int64_t ldl, const T* __restrict__ U,
                                              int64_t ldu, T* __restrict__ C, int64_t ldc, int64_t M, int64_t N,
                                              int64_t K, int tx, int ty, bool vec) {
  __shared__ T Ls[2][GB_K][GB_M + 4];  // Ls[buf][k][i] = L[i0 + i, k0 + k]
  __shared__ T Us[2][GB_K][GB_N + 4];  // Us[buf][k][j] = U[k0 + k, j0 + j]
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)tx * GB_M, j0 = (int64_t)ty * GB_N;
  const int wave = tid / 64, lane = tid % 64;
  const int wr = (wave % 2) * 64, wc = (wave / 2) * 64;
  const int li = lane % 16, lk = lane / 16;
  constexpr int PER = GB_M * GB_K / 256;  // elements of each operand staged per thread (8)
  // L: thread -> row il, k = kl + 2 e (rows contiguous across the wave); U: thread -> k = ku, column ju + 16 e
  // (16 consecutive k of one column per 16 lanes)
  const int il = tid % GB_M, kl = tid / GB_M;
  const int ku = tid % GB_K, ju = tid / GB_K;
  const T* Lrow = L + (i0 + il < M ? i0 + il : M - 1);
  T lr[PER], ur[PER];
  // paired fetch: L rows 2 lp, 2 lp + 1 at k = lkp + 4 e; U k = 2 up, 2 up + 1 of column ujp + 32 e
  vec = vec && i0 + GB_M <= M && j0 + GB_N <= N;
  const int lp = tid % 64, lkp = tid / 64, up = tid % 8, ujp = tid / 8;
  typedef T t2 __attribute__((ext_vector_type(2)));
  auto fetch = [&](int64_t k0) {
    if (vec) {
#pragma unroll
      for (int e = 0; e < PER / 2; ++e) {
        const t2 a = *reinterpret_cast<const t2*>(L + (i0 + 2 * lp) + (k0 + lkp + 4 * e) * ldl);
        const t2 b = *reinterpret_cast<const t2*>(U + (k0 + 2 * up) + (j0 + ujp + 32 * e) * ldu);
        lr[2 * e] = a.x;
        lr[2 * e + 1] = a.y;
        ur[2 * e] = b.x;
        ur[2 * e + 1] = b.y;
      }
      return;
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int64_t kL = k0 + kl + 2 * e;
      const T vl = Lrow[(kL < K ? kL : K - 1) * ldl];
      lr[e] = kL < K ? vl : T(0);
      const int64_t kU = k0 + ku;
      const int64_t jU = j0 + ju + 16 * e;
      const T vu = U[(kU < K ? kU : K - 1) + (jU < N ? jU : N - 1) * ldu];
      ur[e] = kU < K ? vu : T(0);
    }
  };
  auto stash = [&](int buf) {
    if (vec) {
#pragma unroll
      for (int e = 0; e < PER / 2; ++e) {
        Ls[buf][lkp + 4 * e][2 * lp] = lr[2 * e];
        Ls[buf][lkp + 4 * e][2 * lp + 1] = lr[2 * e + 1];
        Us[buf][2 * up][ujp + 32 * e] = ur[2 * e];
        Us[buf][2 * up + 1][ujp + 32 * e] = ur[2 * e + 1];
      }
      return;
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      Ls[buf][kl + 2 * e][il] = lr[e];
      Us[buf][ku][ju + 16 * e] = ur[e];
    }
  };
  typename Mfma<T>::acc acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b2 = 0; b2 < 4; ++b2) acc[a][b2] = typename Mfma<T>::acc{0, 0, 0, 0};
  const int64_t nk = (K + GB_K - 1) / GB_K;
  fetch(0);
  stash(0);
  if (nk > 1) fetch(GB_K);
  __syncthreads();
  for (int64_t s = 0; s < nk; ++s) {
    const int buf = (int)(s & 1);
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 4) {
      T lv[4], uv[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        lv[t] = Ls[buf][kk + lk][wr + 16 * t + li];
        uv[t] = Us[buf][kk + lk][wc + 16 * t + li];
      }
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) acc[tj][ti] = Mfma<T>::step(uv[tj], lv[ti], acc[tj][ti]);
    }
    if (s + 1 < nk) {
      stash(buf ^ 1);
      if (s + 2 < nk) fetch((s + 2) * GB_K);
    }
    __syncthreads();
  }
  // acc[tj][ti][v] = (L U)[row = wr + 16 ti + li, col = wc + 16 tj + Mfma::row(lk, v)]
  // C's read-modify-write, one column tile tj (16 entries per lane) at a time: its loads are issued together, from
  // clamped (always valid) addresses, and the next tile's loads are in flight while this one is stored.  Predicated
  // element by element, each load waited for its own round trip before the store that followed it.
  int64_t rows[4], cols[4][4];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti) rows[ti] = i0 + wr + 16 * ti + li;
#pragma unroll
  for (int tj = 0; tj < 4; ++tj)
#pragma unroll
    for (int v = 0; v < 4; ++v) cols[tj][v] = j0 + wc + 16 * tj + Mfma<T>::row(lk, v);
  auto cload = [&](int tj, T (&cv)[4][4]) {
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = rows[ti] < M ? rows[ti] : M - 1, c = cols[tj][v] < N ? cols[tj][v] : N - 1;
        cv[ti][v] = C[r + c * ldc];
      }
  };
  auto cstore = [&](int tj, const T (&cv)[4][4]) {
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (rows[ti] < M && cols[tj][v] < N) C[rows[ti] + cols[tj][v] * ldc] = cv[ti][v] - acc[tj][ti][v];
  };
  T ca[4][4], cb[4][4];
  cload(0, ca);
  cload(1, cb);
  cstore(0, ca);
  cload(2, ca);
  cstore(1, cb);
  cload(3, cb);
  cstore(2, ca);
  cstore(3, cb);
}

template <class T>
__global__ void __launch_bounds__(256, 2) gemm_sub_kernel(const T* __restrict__ L, int64_t ldl,
                                                          const T* __restrict__ U, int64_t ldu, T* __restrict__ C,
                                                          int64_t ldc, int64_t M, int64_t N, int64_t K, int gx,
                                                          int gy, int vec) {
  int tx, ty;
  xcd_tile(blockIdx.x, gx, gy, tx, ty);
  gemm_sub_tile<T>(L, ldl, U, ldu, C, ldc, M, N, K, tx, ty, vec != 0);
}

// Several independent products in one launch (job j owns workgroups [t0, t0 + gx gy)), for the small per-matrix
// updates of a batched factorization, which alone would occupy a fraction of the CUs.
template <class T>
struct GemmJob {
  const T* L;
  const T* U;
  T* C;
  int64_t ldl, ldu, ldc, M, N, K;
  int32_t gx, gy, t0, vec;  // vec: as gemm_sub_tile's
};
constexpr int BATCH_JOBS = 24;
template <class T>
struct GemmBatch {
  int njobs;
  GemmJob<T> job[BATCH_JOBS];
};

template <class T>
__global__ void __launch_bounds__(256, 2) gemm_sub_batched_kernel(const GemmBatch<T> jb) {
  int j = 0;
  while (j + 1 < jb.njobs && (int)blockIdx.x >= jb.job[j + 1].t0) ++j;
  const GemmJob<T>& J = jb.job[j];
  int tx, ty;
  xcd_tile((int)blockIdx.x - J.t0, J.gx, J.gy, tx, ty);
  gemm_sub_tile<T>(J.L, J.ldl, J.U, J.ldu, J.C, J.ldc, J.M, J.N, J.K, tx, ty, J.vec != 0);
}

template <class T>
struct TrsmJob {
  T* A;
  int64_t lda, k, c0, ncols;
  int32_t nbw, cpw, b0, pad;
};
template <class T>
struct TrsmBatch {
  int njobs;
  TrsmJob<T> job[BATCH_JOBS];
};

template <class T>
__global__ void __launch_bounds__(256) trsm_batched_kernel(const TrsmBatch<T> jb) {
  int j = 0;
  while (j + 1 < jb.njobs && (int)blockIdx.x >= jb.job[j + 1].b0) ++j;
  const TrsmJob<T>& J = jb.job[j];
  trsm_body<T>(J.A, J.lda, J.k, J.nbw, J.c0, J.ncols, J.cpw, (int64_t)blockIdx.x - J.b0);
}

// FP16 variant of the trailing update for FP32 factors: C[0:M, 0:N] -= L[0:M, 0:K] * U[0:K, 0:N] with both
// operands rounded to FP16 and the products accumulated in FP32 on v_mfma_f32_16x16x32_f16.  |L| <= 1 under
// partial pivoting; each column c of U is scaled by a power of two sigma[c] (max |U[:, c]| * sigma in [0.5, 1))
// before rounding, and the accumulated product divided by it, so FP16's range never clips the update.
// The operands are rounded once per update into packed FP16 copies with k contiguous (Lh: row i's K values,
// Uh: column j's K values, K <= LU_OUTER_BIG), so the GEMM streams them with 16-byte loads.
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
constexpr int GH_K = 32;        // k per LDS stage
constexpr int GH_LD = LU_OUTER_BIG; // k stride of the packed copies (the largest outer block)

// Bytes from the column scales to the packed FP16 copies: n floats, rounded up to 256 bytes
inline size_t f16_pack_offset(int64_t n) { return ((sizeof(float) * (size_t)n + 255) / 256) * 256; }
// Layout of lu_factor's workspace: the panels' slots and records, flags, the panels' permutations and the fast
// path's two words; n column scales and the two packed FP16 operand copies of the FP16 trailing update (n x GH_LD
// halves each); then, 256-byte aligned, the fast path's copy of a panel's rows.
inline size_t fast_save_offset(int64_t n) {
  const int64_t nn = n > 0 ? n : 1;
  const size_t head = sizeof(unsigned long long) * 2 * MAX_P * (1 + (size_t)XREC) +
                      sizeof(int) * (2 + PERM_SETS * PERM_INTS + 4) + f16_pack_offset(nn) +
                      2 * sizeof(_Float16) * (size_t)nn * GH_LD + 1024;
  return (head + 255) / 256 * 256;
}

// Lh[i * GH_LD + k] = fp16(L[i, k]) (zero for k >= K): 32 rows per workgroup, read down the columns (lanes over
// rows) and written along the rows (lanes over 8-value groups of k) through an LDS transpose
__global__ void __launch_bounds__(256) pack_l_f16_kernel(const float* __restrict__ L, int64_t ldl, int64_t M,
                                                         int64_t K, _Float16* __restrict__ Lh) {
  __shared__ _Float16 tile[32][GH_LD + 8];
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int KP = (int)((K + GH_K - 1) / GH_K) * GH_K;  // the GEMM reads whole slices: zeros up to KP
  for (int idx = tid; idx < 32 * KP; idx += 256) {
    const int r = idx % 32, k = idx / 32;
    const int64_t i = i0 + r;
    tile[r][k] = (_Float16)((i < M && k < K) ? L[i + k * ldl] : 0.f);
  }
  __syncthreads();
  for (int idx = tid; idx < 32 * (KP / 8); idx += 256) {
    const int r = idx / (KP / 8), kq = idx % (KP / 8);
    const int64_t i = i0 + r;
    if (i >= M) continue;
    h8 v;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = tile[r][8 * kq + t];
    *reinterpret_cast<h8*>(Lh + i * GH_LD + 8 * kq) = v;
  }
}

// sigma[j] = 2^-e with max_k |U[k, j]| in [2^(e-1), 2^e) (1 for a zero column) and Uh[j * GH_LD + k] =
// fp16(sigma[j] U[k, j]); one wave per column, lane l taking k = l, l + 64, ...
__global__ void __launch_bounds__(256) pack_u_f16_kernel(const float* __restrict__ U, int64_t ldu, int64_t N,
                                                         int64_t K, float* __restrict__ sigma,
                                                         _Float16* __restrict__ Uh) {
  const int64_t j = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (j >= N) return;
  float x[GH_LD / 64];
  float mx = 0.f;
#pragma unroll
  for (int t = 0; t < GH_LD / 64; ++t) {
    const int64_t k = lane + 64 * t;
    x[t] = k < K ? U[k + j * ldu] : 0.f;
    mx = fmaxf(mx, fabsf(x[t]));
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  int e = 0;
  if (mx > 0.f) (void)frexpf(mx, &e);
  const float sg = mx > 0.f ? ldexpf(1.0f, -e) : 1.0f;
  if (lane == 0) sigma[j] = sg;
  const int64_t KP = (K + GH_K - 1) / GH_K * GH_K;
#pragma unroll
  for (int t = 0; t < GH_LD / 64; ++t)
    if (lane + 64 * t < KP) Uh[j * GH_LD + lane + 64 * t] = (_Float16)(x[t] * sg);
}

// 128 x 128 tile per workgroup (4 waves in 2 x 2, 64 x 64 each), K in GH_K slices through two LDS buffers (one
// barrier per slice), operands from the packed copies: 4 consecutive lanes read one row's (column's) 32 k of a
// slice as two 16-byte loads each.  One v_mfma_f32_16x16x32_f16 per tile pair and slice: lane l holds k
// 8 (l / 16) .. + 7 of row (column) l % 16 of both operands.  Row and column indices are clamped (the extra rows
// and columns only reach C entries that are not written); the packed copies are zero beyond K.
__global__ void __launch_bounds__(256, 2) gemm_sub_f16_kernel(const _Float16* __restrict__ Lh,
                                                              const _Float16* __restrict__ Uh,
                                                              float* __restrict__ C, int64_t ldc, int64_t M, int64_t N,
                                                              int64_t K, const float* __restrict__ sigma, int gx,
                                                              int gy) {
  static_assert(GH_K == 32, "one 16x16x32 MFMA per slice");
  constexpr int OK = GH_K / 8;  // octets of k per row and slice (4)
  __shared__ h8 Lq[2][GB_M][OK + 1];
  __shared__ h8 Uq[2][GB_N][OK + 1];
  const int tid = threadIdx.x;
  int tx, ty;
  xcd_tile(blockIdx.x, gx, gy, tx, ty);
  const int64_t i0 = (int64_t)tx * GB_M, j0 = (int64_t)ty * GB_N;
  const int wave = tid / 64, lane = tid % 64;
  const int wr = (wave % 2) * 64, wc = (wave / 2) * 64;
  const int li = lane % 16, lk = lane / 16;
  // staging: thread -> row (column) r = tid / 4 + 64 e, octet tid % 4
  const int sp = tid % 4, sr = tid / 4;
  const _Float16* lsrc[2];
  const _Float16* usrc[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int64_t i = i0 + sr + 64 * e, j = j0 + sr + 64 * e;
    lsrc[e] = Lh + (i < M ? i : M - 1) * GH_LD + 8 * sp;
    usrc[e] = Uh + (j < N ? j : N - 1) * GH_LD + 8 * sp;
  }
  h8 lr[2], ur[2];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      lr[e] = *reinterpret_cast<const h8*>(lsrc[e] + k0);
      ur[e] = *reinterpret_cast<const h8*>(usrc[e] + k0);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      Lq[buf][sr + 64 * e][sp] = lr[e];
      Uq[buf][sr + 64 * e][sp] = ur[e];
    }
  };
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b2 = 0; b2 < 4; ++b2) acc[a][b2] = f4{0.f, 0.f, 0.f, 0.f};
  const int64_t nk = (K + GH_K - 1) / GH_K;
  fetch(0);
  stash(0);
  if (nk > 1) fetch(GH_K);
  __syncthreads();
  for (int64_t s = 0; s < nk; ++s) {
    const int buf = (int)(s & 1);
    h8 lv[4], uv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      lv[t] = Lq[buf][wr + 16 * t + li][lk];
      uv[t] = Uq[buf][wc + 16 * t + li][lk];
    }
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
        acc[tj][ti] = __builtin_amdgcn_mfma_f32_16x16x32_f16(uv[tj], lv[ti], acc[tj][ti], 0, 0, 0);
    if (s + 1 < nk) {
      stash(buf ^ 1);
      if (s + 2 < nk) fetch((s + 2) * GH_K);
    }
    __syncthreads();
  }
  // acc[tj][ti][v] = (L sigma U)[row = wr + 16 ti + li, col = wc + 16 tj + 4 lk + v]
  // C's read-modify-write as in gemm_sub_tile: all loads of a column tile issued together from clamped addresses,
  // the next tile's loads in flight while this one is stored.
  int64_t rows[4], cols[4][4];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti) rows[ti] = i0 + wr + 16 * ti + li;
  float sg[4][4];
#pragma unroll
  for (int tj = 0; tj < 4; ++tj)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      cols[tj][v] = j0 + wc + 16 * tj + 4 * lk + v;
      sg[tj][v] = sigma[cols[tj][v] < N ? cols[tj][v] : N - 1];
    }
  auto cload = [&](int tj, float (&cv)[4][4]) {
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = rows[ti] < M ? rows[ti] : M - 1, c = cols[tj][v] < N ? cols[tj][v] : N - 1;
        cv[ti][v] = C[r + c * ldc];
      }
  };
  auto cstore = [&](int tj, const float (&cv)[4][4]) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float s = 1.0f / sg[tj][v];
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
        if (rows[ti] < M && cols[tj][v] < N) C[rows[ti] + cols[tj][v] * ldc] = cv[ti][v] - acc[tj][ti][v] * s;
    }
  };
  float ca[4][4], cb[4][4];
  cload(0, ca);
  cload(1, cb);
  cstore(0, ca);
  cload(2, ca);
  cstore(1, cb);
  cload(3, cb);
  cstore(2, ca);
  cstore(3, cb);
}

// ---- triangular solves with nrhs <= 64 right-hand sides -------------------------------------------------
// Diagonal-block solve of rows [r0, r0+nb) for all right-hand sides: lower (unit) or upper.  One wave per
// right-hand side: lane i holds x[i]; step r broadcasts x[r] and the lanes below (lower) or above (upper) update.
template <class T, bool LOWER>
__global__ void __launch_bounds__(256) trsv_diag_kernel(const T* __restrict__ A, int64_t lda, int64_t r0, int nb,
                                                        T* B, int64_t ldb, int64_t nrhs) {
  __shared__ T D[LU_NB][LU_NB + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int r = idx % nb, c = idx / nb;
    D[r][c] = A[(r0 + r) + (r0 + c) * lda];
  }
  __syncthreads();
  const int64_t q = (int64_t)blockIdx.x * 4 + wave;
  if (q >= nrhs) return;
  T* x = B + q * ldb + r0;
  T xi = lane < nb ? x[lane] : T(0);
  if (LOWER) {
    for (int r = 0; r < nb - 1; ++r) {
      const T xr = __shfl(xi, r);
      if (lane > r && lane < nb) xi -= D[lane][r] * xr;
    }
  } else {
    for (int r = nb - 1; r >= 0; --r) {
      if (lane == r) xi = xi / D[r][r];
      const T xr = __shfl(xi, r);
      if (lane < r) xi -= D[lane][r] * xr;
    }
  }
  if (lane < nb) x[lane] = xi;
}

// B[rows] -= A[rows, c0:c0+nb] * B[c0:c0+nb] for rows in [rb, re).
template <class T>
__global__ void __launch_bounds__(256) trsv_update_kernel(const T* __restrict__ A, int64_t lda, int64_t c0, int nb,
                                                          int64_t rb, int64_t re, T* B, int64_t ldb, int64_t nrhs) {
  const int64_t i = rb + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= re) return;
  for (int64_t q = 0; q < nrhs; ++q) {
    const T* x = B + q * ldb + c0;
    T s = T(0);
    for (int c = 0; c < nb; ++c) s += A[i + (c0 + c) * lda] * x[c];
    B[i + q * ldb] -= s;
  }
}

template <class T>
__global__ void permute_kernel(const int32_t* __restrict__ perm, const T* __restrict__ in, T* __restrict__ out, int64_t n,
                               int64_t nrhs, int64_t ld) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * nrhs) return;
  const int64_t q = t / n, i = t % n;
  out[i + q * ld] = in[perm[i] + q * ld];
}

template <class TA>
__global__ void __launch_bounds__(256) gemv_kernel(const TA* __restrict__ A, int64_t m, int64_t n, int64_t lda,
                                                   const double* __restrict__ x, double* __restrict__ y, double alpha,
                                                   double beta) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int64_t c = 0; c < n; ++c) s += (double)A[i + c * lda] * x[c];
  y[i] = alpha * s + (beta == 0.0 ? 0.0 : beta * y[i]);
}

__global__ void d2f_kernel(const double* a, float* b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void f2d_kernel(const float* a, double* b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
__global__ void d2f_matrix_kernel(const double* A, int64_t lda, float* B, int64_t ldb, int64_t m, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (i < m && c < n) B[i + c * ldb] = (float)A[i + c * lda];
}

unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// K = [[I, A], [A^T, 0]] ((m + k) x (m + k), leading dimension ldk) from A (m x k, leading dimension lda).
__global__ void kkt_kernel(const double* __restrict__ A, int64_t lda, int64_t m, int64_t k, double* __restrict__ K,
                           int64_t ldk) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t j = blockIdx.y;
  if (i >= m + k) return;
  double v;
  if (j < m) v = i < m ? (i == j ? 1.0 : 0.0) : A[j + (i - m) * lda];
  else v = i < m ? A[i + (j - m) * lda] : 0.0;
  K[i + j * ldk] = v;
}

// info = -1 when a panel barrier wait overran (the factors are then invalid)
// A matrix's factorization state before its first panel: info and the wait flags cleared, slots and records all
// ones (tag 4095 / 2^32 - 1 match no column).  A kernel rather than memsets, so the factorization can be captured
// in a graph.
__global__ void lu_init_kernel(int* info, unsigned* bar, unsigned long long* slots, int64_t nslots, int* fpw) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t == 0) {
    *info = 0;
    bar[0] = 0;
    bar[1] = 0;
  }
  if (fpw && t < 4) fpw[t] = 0;
  for (int64_t i = t; i < nslots; i += (int64_t)gridDim.x * blockDim.x) slots[i] = ~0ull;
}

__global__ void barrier_status_kernel(const int* bar_err, int* info) {
  if (*bar_err) *info = -1;
}

// The paired operand fetch of gemm_sub_tile applies (both operands aligned to two elements with even leading
// dimensions, whole k slices).
template <class T>
int gemm_vec_ok(const T* L, int64_t ldl, const T* U, int64_t ldu, int64_t K) {
  return reinterpret_cast<uintptr_t>(L) % (2 * sizeof(T)) == 0 && reinterpret_cast<uintptr_t>(U) % (2 * sizeof(T)) == 0 &&
         ldl % 2 == 0 && ldu % 2 == 0 && K % GB_K == 0;
}

template <class T>
void gemm_sub(const T* L, int64_t ldl, const T* U, int64_t ldu, T* C, int64_t ldc, int64_t M, int64_t N, int64_t K,
              hipStream_t st) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  const int gx = (int)nblk(M, GB_M), gy = (int)nblk(N, GB_N);
  hipLaunchKernelGGL(gemm_sub_kernel<T>, dim3(gx * gy), dim3(256), 0, st, L, ldl, U, ldu, C, ldc, M, N, K, gx, gy,
                     gemm_vec_ok(L, ldl, U, ldu, K));
}

// Rows [r0, r1) of columns [c0, c0 + nc): X = L^-1 X with L the unit-lower block of the sub-panels in [r0, r1)
// (64 rows each): per sub-panel a diagonal-block solve, then the update of the rows below it inside [r0, r1).
template <class T>
void block_trsm(T* A, int64_t lda, int64_t r0, int64_t r1, int64_t c0, int64_t nc, hipStream_t st) {
  if (nc <= 0) return;
  // columns per wave: about nc / 1024, at most 4 (more columns per wave lengthen each wave's chain of solves more
  // than they save in staging L11: 64 x 18000 FP64 in 24 us at 4, 29 us at 16)
  const int cpw = (int)std::min<int64_t>(4, std::max<int64_t>(1, nc / 1024));
  for (int64_t s = r0; s < r1; s += LU_NB) {
    const int nb = (int)std::min<int64_t>(LU_NB, r1 - s);
    hipLaunchKernelGGL(trsm_kernel<T>, dim3(nblk(nc, 4 * cpw)), dim3(256), 0, st, A, lda, s, nb, c0, nc, cpw);
    gemm_sub<T>(A + (s + nb) + s * lda, lda, A + s + c0 * lda, lda, A + (s + nb) + c0 * lda, lda, r1 - s - nb, nc, nb,
                st);
  }
}

// Job lists of the batched factorization, launched BATCH_JOBS products (or solves) at a time.
template <class T>
struct GemmList {
  std::vector<GemmJob<T>> v;
  void add(const T* L, int64_t ldl, const T* U, int64_t ldu, T* C, int64_t ldc, int64_t M, int64_t N, int64_t K) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    v.push_back(GemmJob<T>{L, U, C, ldl, ldu, ldc, M, N, K, (int32_t)nblk(M, GB_M), (int32_t)nblk(N, GB_N), 0,
                           gemm_vec_ok(L, ldl, U, ldu, K)});
  }
  void launch(hipStream_t st) {
    for (size_t a = 0; a < v.size(); a += BATCH_JOBS) {
      GemmBatch<T> b;
      b.njobs = 0;
      int t = 0;
      for (size_t i = a; i < v.size() && b.njobs < BATCH_JOBS; ++i) {
        GemmJob<T> j = v[i];
        j.t0 = t;
        t += j.gx * j.gy;
        b.job[b.njobs++] = j;
      }
      hipLaunchKernelGGL(gemm_sub_batched_kernel<T>, dim3(t), dim3(256), 0, st, b);
    }
    v.clear();
  }
};

template <class T>
struct TrsmList {
  std::vector<TrsmJob<T>> v;
  void add(T* A, int64_t lda, int64_t k, int nbw, int64_t c0, int64_t nc) {
    if (nc <= 0) return;
    // as block_trsm: about nc / 1024 columns per wave, at most 4 (each workgroup stages L11, so a launch over 24
    // meshes of thousands of columns at one column per wave staged it ~48k times)
    const int cpw = (int)std::min<int64_t>(4, std::max<int64_t>(1, nc / 1024));
    v.push_back(TrsmJob<T>{A, lda, k, c0, nc, nbw, cpw, 0, 0});
  }
  void launch(hipStream_t st) {
    for (size_t a = 0; a < v.size(); a += BATCH_JOBS) {
      TrsmBatch<T> b;
      b.njobs = 0;
      int t = 0;
      for (size_t i = a; i < v.size() && b.njobs < BATCH_JOBS; ++i) {
        TrsmJob<T> j = v[i];
        j.b0 = t;
        t += (int)nblk(j.ncols, 4 * j.cpw);
        b.job[b.njobs++] = j;
      }
      hipLaunchKernelGGL(trsm_batched_kernel<T>, dim3(t), dim3(256), 0, st, b);
    }
    v.clear();
  }
};

// The trailing update C -= L U of an outer block, in two steps so that disjoint column ranges of C can be updated
// by different streams from one preparation: trailing_prepare forms what every range reads (for the FP16 update,
// the column scales and the packed FP16 copies of L and of U's N columns, in the workspace after sigma),
// trailing_cols updates columns [j0, j1) of C.
template <class T>
struct Trailing {
  T* L;
  T* U;
  T* C;
  int64_t lda, M, N, K;
  float* sigma;
  bool fp16;
  _Float16* Lh = nullptr;
  _Float16* Uh = nullptr;
};

template <class T>
Trailing<T> trailing_prepare(T* L, T* U, T* C, int64_t lda, int64_t M, int64_t N, int64_t K, float* sigma, bool fp16,
                             hipStream_t st) {
  Trailing<T> t{L, U, C, lda, M, N, K, sigma, fp16 && std::is_same<T, float>::value && M > 0 && N > 0 && K > 0};
  if (t.fp16) {
    // packed FP16 copies after the column scales in the same workspace region, 256-byte aligned
    const uintptr_t pk = (reinterpret_cast<uintptr_t>(sigma) + sizeof(float) * (size_t)(M > N ? M : N) + 255) &
                         ~(uintptr_t)255;
    t.Lh = reinterpret_cast<_Float16*>(pk);
    t.Uh = t.Lh + M * GH_LD;
    hipLaunchKernelGGL(pack_l_f16_kernel, dim3(nblk(M, 32)), dim3(256), 0, st, (const float*)L, lda, M, K, t.Lh);
    hipLaunchKernelGGL(pack_u_f16_kernel, dim3(nblk(N, 4)), dim3(256), 0, st, (const float*)U, lda, N, K, sigma,
                       t.Uh);
  }
  return t;
}

template <class T>
void trailing_cols(const Trailing<T>& t, int64_t j0, int64_t j1, hipStream_t st) {
  const int64_t N = j1 - j0;
  if (t.M <= 0 || N <= 0 || t.K <= 0) return;
  if (!t.fp16) {
    gemm_sub<T>(t.L, t.lda, t.U + j0 * t.lda, t.lda, t.C + j0 * t.lda, t.lda, t.M, N, t.K, st);
    return;
  }
  const int gx = (int)nblk(t.M, GB_M), gy = (int)nblk(N, GB_N);
  hipLaunchKernelGGL(gemm_sub_f16_kernel, dim3(gx * gy), dim3(256), 0, st, (const _Float16*)t.Lh,
                     (const _Float16*)(t.Uh + j0 * GH_LD), reinterpret_cast<float*>(t.C + j0 * t.lda), t.lda, t.M, N,
                     t.K, (const float*)(t.sigma + j0), gx, gy);
}

// The caller's second stream and one factorization's events for the look-ahead (the events are created per call
// and released when it returns; the runtime frees them once the work that records them has completed).
struct LookAhead {
  hipStream_t s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, panels = nullptr;
  hipError_t init(hipStream_t side) {
    s2 = side;
    hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&panels, hipEventDisableTiming);
    return e;
  }
  ~LookAhead() {
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (panels) (void)hipEventDestroy(panels);
  }
};

// Look-ahead is on unless CDJ_LU_LOOKAHEAD=0, the panels' diagonal-pivot fast path unless CDJ_LU_FASTPANEL=0 (read
// per call; the factors are the same either way).
inline bool lookahead_enabled() {
  const char* v = std::getenv("CDJ_LU_LOOKAHEAD");
  return !(v && v[0] == '0');
}
inline bool fastpanel_enabled() {
  const char* v = std::getenv("CDJ_LU_FASTPANEL");
  return !(v && v[0] == '0');
}

}  // namespace

size_t lu_workspace_bytes(int64_t n) {
  // then the fast path's copy of a panel's rows (n x LU_NB entries, FP64 at most)
  return fast_save_offset(n) + sizeof(double) * (size_t)(n > 0 ? n : 1) * LU_NB;
}

// Blocked LU with partial pivoting of the n x n leading block of the n x ncols matrix A (extra columns are
// carried along).  Outer blocks of lu_outer(n) columns are factored as LU_NB-wide panels, each first brought up to
// date with the earlier panels of its block (left-looking inside the block); the rest of the matrix then gets a
// single rank-lu_outer(n) update (right-looking across blocks), so the trailing matrix is read and written once
// per outer block.
//
// Look-ahead: the trailing update of block b is split.  The next block's columns are updated first on the
// caller's stream, whose next work is that block's panels; the rest runs at the same time on a second stream.
// The panels' row interchanges reach only their own block's columns then; the interchanges in every other column
// (among them block b's multipliers, which the second stream's product reads) wait for that product and are
// applied panel by panel afterwards.  Each entry goes through the same operations as without look-ahead, so the
// factors are the same bit for bit.  A panel launched beside the running product is an ordinary launch: its
// workgroups take CUs as the product's workgroups finish (a panel workgroup needs 6.5 KiB of LDS and fits beside
// one product workgroup), and its waits are bounded as always.
template <class T>
hipError_t lu_factor_cols(T* A, int64_t n, int64_t ncols, int64_t lda, int32_t* ipiv, int* info, void* ws,
                          hipStream_t st, bool fp16_update, hipStream_t side) {
  char* w = static_cast<char*>(ws);
  unsigned long long* slots = reinterpret_cast<unsigned long long*>(w);  // the panels' slot words, then records
  unsigned long long* xw = slots + 2 * MAX_P;
  const size_t xbytes = sizeof(unsigned long long) * 2 * MAX_P * (1 + (size_t)XREC);
  w += xbytes;
  unsigned* bar = reinterpret_cast<unsigned*>(w);  // unused word, then the panels' wait-overrun flag
  int* bar_err = reinterpret_cast<int*>(bar + 1);
  int* perm_sets = bar_err + 1;                     // the row permutations of an outer block's panels, two sets
  int* fpw = perm_sets + PERM_SETS * PERM_INTS;       // the fast path's {violation, zero-pivot column}
  float* sigma = reinterpret_cast<float*>(fpw + 4);  // n column scales (FP16 update)
  T* save = reinterpret_cast<T*>(static_cast<char*>(ws) + fast_save_offset(n));  // a panel's rows (n x LU_NB)
  if (ncols > n) fp16_update = false;
  hipError_t e = hipMemsetAsync(info, 0, sizeof(int), st);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(bar, 0, 2 * sizeof(unsigned), st);
  if (e != hipSuccess) return e;
  // all ones: slot tag 4095 and record tag 2^32 - 1 match no column of this factorization
  e = hipMemsetAsync(slots, 0xFF, xbytes, st);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(fpw, 0, 4 * sizeof(int), st);
  if (e != hipSuccess) return e;
  const int64_t ob = lu_outer(n);
  // panel workgroups that fit on the device at once (ordinary launches beside a running product rely on it)
  static const int panel_capacity = [] {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)lu_panel_kernel<T, MAX_P / 64>, PT, 0) !=
        hipSuccess)
      return 0;
    return per_cu * cus;
  }();
  LookAhead la;
  const bool ahead = side != nullptr && side != st && lookahead_enabled() && n > ob && panel_capacity > 0;
  const bool fast = fastpanel_enabled();
  if (ahead && (e = la.init(side)) != hipSuccess) return e;
  bool pending = false;  // the previous block's update of the columns right of this block runs on la.s2
  bool queued = false;   // work was queued on la.s2
  for (int64_t K0 = 0, blk = 0; K0 < n; K0 += ob, ++blk) {
    const int64_t K1 = std::min<int64_t>(n, K0 + ob);
    int* perms = perm_sets + (blk % PERM_SETS) * PERM_INTS;
    int np = 0;
    for (int64_t k = K0; k < K1; k += LU_NB, ++np) {
      const int nbw = (int)std::min<int64_t>(LU_NB, K1 - k);
      int* perm = perms + np * PERM_STRIDE;
      if (k > K0) {
        // this panel's columns: U rows from the block's earlier panels, then their update of rows [k, n)
        block_trsm<T>(A, lda, K0, k, k, nbw, st);
        gemm_sub<T>(A + k + K0 * lda, lda, A + K0 + k * lda, lda, A + k + k * lda, lda, n - k, nbw, k - K0, st);
      }
      const int64_t m = n - k;
      const int P = (int)((m + MAX_RPB - 1) / MAX_RPB);
      int rpb = (int)((m + P - 1) / P);
      if (P > MAX_P) return hipErrorInvalidValue;
      int* skip = nullptr;
      const T* csave = save;
      if (fast) {
        // the diagonal-pivot fast path; the exchange panel below runs only if it found another pivot row
        if (nbw == LU_NB)
          hipLaunchKernelGGL((fp_diag_kernel<T, true>), dim3(1), dim3(LU_NB), 0, st, A, lda, k, nbw, ipiv, perm, fpw,
                             save, m);
        else
          hipLaunchKernelGGL((fp_diag_kernel<T, false>), dim3(1), dim3(LU_NB), 0, st, A, lda, k, nbw, ipiv, perm, fpw,
                             save, m);
        if (m > nbw)
          hipLaunchKernelGGL(fp_rows_kernel<T>, dim3(nblk(m - nbw, 256)), dim3(256), 0, st, A, lda, k, nbw, m - nbw,
                             fpw, save, m);
        skip = fpw;
      }
      void* args[] = {&A,  (void*)&lda, &k,       (void*)&m, (void*)&nbw,  &rpb, &ipiv, &info,
                      &xw, &slots,      &bar_err, &perm,     (void*)&skip, (void*)&csave};
      const void* kern = P <= 128 ? (const void*)lu_panel_kernel<T, 2> : (const void*)lu_panel_kernel<T, MAX_P / 64>;
      if (pending && P <= panel_capacity)
        e = hipLaunchKernel(kern, dim3(P), dim3(PT), args, 0, st);
      else
        e = hipLaunchCooperativeKernel(kern, dim3(P), dim3(PT), args, 0, st);
      if (e != hipSuccess) return e;
      if (pending) {
        if (K1 - K0 > nbw)
          hipLaunchKernelGGL(laswp_range_kernel<T>, dim3(nblk(K1 - K0, 2)), dim3(256), 0, st, A, lda, K0, K1, k, nbw,
                             (const int*)perm);
      } else if (ncols > nbw) {
        hipLaunchKernelGGL(laswp_kernel<T>, dim3(nblk(ncols, 2)), dim3(256), 0, st, A, lda, ncols, k, nbw,
                           (const int*)perm);
      }
    }
    if (pending) {
      // this block's interchanges in the other columns: those left of the block on the second stream, behind its
      // product (nothing on this stream reads them again; this permutation set is rewritten only two blocks on,
      // after a join that follows them), those right of it here once the product is done
      if (K0 > 0) {
        if ((e = hipEventRecord(la.panels, st)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(la.s2, la.panels, 0)) != hipSuccess) return e;
        hipLaunchKernelGGL(laswp_deferred_kernel<T>, dim3(nblk(K0, 2)), dim3(256), 0, la.s2, A, lda, K0, (int64_t)0,
                           K0, np, (const int*)perms);
      }
      if ((e = hipStreamWaitEvent(st, la.join, 0)) != hipSuccess) return e;
      if (ncols > K1)
        hipLaunchKernelGGL(laswp_deferred_kernel<T>, dim3(nblk(ncols - K1, 2)), dim3(256), 0, st, A, lda, K0, K1,
                           ncols, np, (const int*)perms);
      pending = false;
    }
    if (K1 >= ncols) break;
    // U12 of the block, then the rank-(K1 - K0) update of everything right of and below it
    block_trsm<T>(A, lda, K0, K1, K1, ncols - K1, st);
    const Trailing<T> t = trailing_prepare<T>(A + K1 + K0 * lda, A + K0 + K1 * lda, A + K1 + K1 * lda, lda, n - K1,
                                              ncols - K1, K1 - K0, sigma, fp16_update, st);
    const int64_t K2 = std::min<int64_t>(n, K1 + ob);  // the next block's columns are [K1, K2)
    if (ahead && K1 < n && K2 < ncols) {
      if ((e = hipEventRecord(la.fork, st)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(la.s2, la.fork, 0)) != hipSuccess) return e;
      trailing_cols<T>(t, K2 - K1, ncols - K1, la.s2);
      if ((e = hipEventRecord(la.join, la.s2)) != hipSuccess) return e;
      trailing_cols<T>(t, 0, K2 - K1, st);
      pending = true;
      queued = true;
    } else {
      trailing_cols<T>(t, 0, ncols - K1, st);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (queued) {  // everything queued on the second stream
    if ((e = hipEventRecord(la.join, la.s2)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(st, la.join, 0)) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(barrier_status_kernel, dim3(1), dim3(1), 0, st, (const int*)bar_err, info);
  return hipGetLastError();
}

// Co-resident workgroups of the batched panel kernel (what one cooperative launch may hold).  Thread-safe one-time
// set-up.
template <class T>
int lu_batched_capacity() {
  static const int capacity = [] {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t r = hipGetDevice(&dev);
    if (r == hipSuccess) r = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (r == hipSuccess)
      r = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)lu_panel_batched_kernel<T, MAX_P / 64>,
                                                       PT, 0);
    return r == hipSuccess ? per_cu * cus : 0;
  }();
  return capacity;
}

// Several matrices factored together, each exactly as lu_factor_cols would factor it alone when n < LU_BIG (the
// same operations in the same order, so the same factors bit for bit; larger matrices keep LU_OUTER-column blocks
// here), but stepped in lockstep by panel column so that the panels
// of all matrices at column k share cooperative launches: a panel's cost is the latency chain of its 64 columns,
// which one launch now pays once for up to PANEL_JOBS matrices.  The catch-up and trailing updates, triangular
// solves and row interchanges stay per matrix, on the same stream.
template <class T>
hipError_t lu_factor_batched(const LuJob<T>* jobs, int njobs, hipStream_t st, bool fp16_update) {
  if (njobs <= 0) return hipSuccess;
  struct WsView {
    unsigned long long *slots, *xw;
    unsigned* bar;
    int *bar_err, *perm, *fpw;
    float* sigma;
    T* save;
    size_t xbytes;
  };
  std::vector<WsView> wv(njobs);
  int64_t nmax = 0;
  for (int j = 0; j < njobs; ++j) {
    const LuJob<T>& J = jobs[j];
    if (J.n < 0 || J.ncols < J.n || J.lda < J.n) return hipErrorInvalidValue;
    if ((J.n + MAX_RPB - 1) / MAX_RPB > MAX_P) return hipErrorInvalidValue;
    char* w = static_cast<char*>(J.ws);
    WsView& v = wv[j];
    v.slots = reinterpret_cast<unsigned long long*>(w);
    v.xw = v.slots + 2 * MAX_P;
    v.xbytes = sizeof(unsigned long long) * 2 * MAX_P * (1 + (size_t)XREC);
    w += v.xbytes;
    v.bar = reinterpret_cast<unsigned*>(w);
    v.bar_err = reinterpret_cast<int*>(v.bar + 1);
    v.perm = v.bar_err + 1;
    v.fpw = v.perm + PERM_SETS * PERM_INTS;  // the same layout as lu_factor_cols'
    v.sigma = reinterpret_cast<float*>(v.fpw + 4);
    v.save = reinterpret_cast<T*>(static_cast<char*>(J.ws) + fast_save_offset(J.n));
    hipLaunchKernelGGL(lu_init_kernel, dim3(64), dim3(256), 0, st, J.info, v.bar, v.slots,
                       (int64_t)(v.xbytes / sizeof(unsigned long long)), v.fpw);
    nmax = std::max(nmax, J.n);
  }
  const int capacity = lu_batched_capacity<T>();
  if (capacity <= 0) return hipErrorInvalidConfiguration;
  const bool fast = fastpanel_enabled();
  std::vector<int> act;
  GemmList<T> gl;
  TrsmList<T> tl;
  for (int64_t K0 = 0; K0 < nmax; K0 += LU_OUTER) {
    for (int64_t k = K0; k < K0 + LU_OUTER && k < nmax; k += LU_NB) {
      act.clear();
      for (int j = 0; j < njobs; ++j)
        if (jobs[j].n > k) act.push_back(j);
      // this panel's columns of every matrix: U rows from the block's earlier panels (sub-panel by sub-panel, a
      // triangular solve and the update of the rows below it), then their update of rows [k, n)
      auto nbw_of = [&](const LuJob<T>& J) {
        return (int)std::min<int64_t>(LU_NB, std::min<int64_t>(J.n, K0 + LU_OUTER) - k);
      };
      if (k > K0) {
        for (int64_t s = K0; s < k; s += LU_NB) {
          for (int j : act) tl.add(jobs[j].A, jobs[j].lda, s, LU_NB, k, nbw_of(jobs[j]));
          tl.launch(st);
          for (int j : act) {
            const LuJob<T>& J = jobs[j];
            const int64_t ld = J.lda;
            gl.add(J.A + (s + LU_NB) + s * ld, ld, J.A + s + k * ld, ld, J.A + (s + LU_NB) + k * ld, ld, k - s - LU_NB,
                   nbw_of(J), LU_NB);
          }
          gl.launch(st);
        }
        for (int j : act) {
          const LuJob<T>& J = jobs[j];
          const int64_t ld = J.lda;
          gl.add(J.A + k + K0 * ld, ld, J.A + K0 + k * ld, ld, J.A + k + k * ld, ld, J.n - k, nbw_of(J), k - K0);
        }
        gl.launch(st);
      }
      // the diagonal-pivot fast path of every matrix's panel (as in lu_factor_cols), FP_JOBS matrices per launch;
      // the exchange panels below run only for the matrices where it found another pivot row
      if (fast) {
        for (size_t a0 = 0; a0 < act.size(); a0 += FP_JOBS) {
          const size_t a1 = std::min(act.size(), a0 + FP_JOBS);
          FpBatch<T> full, part, all;
          full.njobs = part.njobs = all.njobs = 0;
          for (size_t a = a0; a < a1; ++a) {
            const LuJob<T>& J = jobs[act[a]];
            const WsView& v = wv[act[a]];
            const int nbw = nbw_of(J);
            const FpJob<T> q{J.A, J.lda, J.n - k, nbw, 0, J.ipiv, v.perm, v.fpw, v.save};
            (nbw == LU_NB ? full : part).job[(nbw == LU_NB ? full : part).njobs++] = q;
            all.job[all.njobs++] = q;
          }
          int64_t kk = k;
          if (full.njobs)
            hipLaunchKernelGGL((fp_diag_batched_kernel<T, true>), dim3(full.njobs), dim3(LU_NB), 0, st, full, kk);
          if (part.njobs)
            hipLaunchKernelGGL((fp_diag_batched_kernel<T, false>), dim3(part.njobs), dim3(LU_NB), 0, st, part, kk);
          int blocks = 0;
          for (int q = 0; q < all.njobs; ++q) {
            all.job[q].b0 = blocks;
            blocks += (int)nblk(all.job[q].m - all.job[q].nbw, 256);
          }
          if (blocks > 0) hipLaunchKernelGGL(fp_rows_batched_kernel<T>, dim3(blocks), dim3(256), 0, st, all, kk);
        }
      }
      // the panels, as few launches as co-residency and the job table allow
      size_t a = 0;
      while (a < act.size()) {
        PanelBatch<T> pb;
        pb.njobs = 0;
        int blocks = 0, pmax = 1;
        for (; a < act.size() && pb.njobs < PANEL_JOBS; ++a) {
          const LuJob<T>& J = jobs[act[a]];
          const int64_t m = J.n - k;
          const int P = (int)((m + MAX_RPB - 1) / MAX_RPB);
          if (blocks > 0 && blocks + P > capacity) break;
          const int64_t K1 = std::min<int64_t>(J.n, K0 + LU_OUTER);
          PanelJob<T>& q = pb.job[pb.njobs++];
          q.A = J.A;
          q.lda = J.lda;
          q.m = m;
          q.nbw = (int)std::min<int64_t>(LU_NB, K1 - k);
          q.rpb = (int)((m + P - 1) / P);
          q.P = P;
          q.b0 = blocks;
          q.ipiv = J.ipiv;
          q.info = J.info;
          q.xw = wv[act[a]].xw;
          q.slots = wv[act[a]].slots;
          q.bar_err = wv[act[a]].bar_err;
          q.perm = wv[act[a]].perm;
          q.skip = fast ? wv[act[a]].fpw : nullptr;
          q.save = wv[act[a]].save;
          blocks += P;
          pmax = std::max(pmax, P);
        }
        if (blocks > capacity) return hipErrorCooperativeLaunchTooLarge;
        int64_t kk = k;
        void* args[] = {&pb, &kk};
        const void* kern = pmax <= 128 ? (const void*)lu_panel_batched_kernel<T, 2>
                                       : (const void*)lu_panel_batched_kernel<T, MAX_P / 64>;
        hipError_t e = hipLaunchCooperativeKernel(kern, dim3(blocks), dim3(PT), args, 0, st);
        if (e != hipSuccess) return e;
      }
      // the panels' row interchanges in the other columns, SWAP_JOBS matrices per launch
      {
        SwapBatch<T> sb;
        sb.njobs = 0;
        int blocks = 0;
        for (size_t a2 = 0; a2 <= act.size(); ++a2) {
          if (sb.njobs == SWAP_JOBS || (a2 == act.size() && sb.njobs > 0)) {
            int64_t kk = k;
            hipLaunchKernelGGL(laswp_batched_kernel<T>, dim3(blocks), dim3(256), 0, st, sb, kk);
            sb.njobs = 0;
            blocks = 0;
          }
          if (a2 == act.size()) break;
          const LuJob<T>& J = jobs[act[a2]];
          const int nbw = nbw_of(J);
          if (J.ncols <= nbw) continue;
          const int nb = (int)std::min<int64_t>(SWAP_BLOCKS, nblk(J.ncols, 2));
          sb.job[sb.njobs++] = SwapJob<T>{J.A, J.lda, J.ncols, (const int*)wv[act[a2]].perm, nbw, blocks, nb};
          blocks += nb;
        }
      }
    }
    // U12 of the block (sub-panel by sub-panel), then the rank-(K1 - K0) update of everything right of and below it
    for (int64_t s = K0; s < K0 + LU_OUTER && s < nmax; s += LU_NB) {
      for (int j = 0; j < njobs; ++j) {
        const LuJob<T>& J = jobs[j];
        const int64_t K1 = std::min<int64_t>(J.n, K0 + LU_OUTER);
        if (J.n > K0 && s < K1) tl.add(J.A, J.lda, s, (int)std::min<int64_t>(LU_NB, K1 - s), K1, J.ncols - K1);
      }
      tl.launch(st);
      for (int j = 0; j < njobs; ++j) {
        const LuJob<T>& J = jobs[j];
        const int64_t K1 = std::min<int64_t>(J.n, K0 + LU_OUTER), ld = J.lda;
        if (J.n <= K0 || s >= K1) continue;
        const int64_t nb = std::min<int64_t>(LU_NB, K1 - s);
        gl.add(J.A + (s + nb) + s * ld, ld, J.A + s + K1 * ld, ld, J.A + (s + nb) + K1 * ld, ld, K1 - s - nb,
               J.ncols - K1, nb);
      }
      gl.launch(st);
    }
    for (int j = 0; j < njobs; ++j) {
      const LuJob<T>& J = jobs[j];
      if (J.n <= K0) continue;
      const int64_t K1 = std::min<int64_t>(J.n, K0 + LU_OUTER), ld = J.lda;
      if (fp16_update && J.ncols == J.n) {
        const Trailing<T> t = trailing_prepare<T>(J.A + K1 + K0 * ld, J.A + K0 + K1 * ld, J.A + K1 + K1 * ld, ld,
                                                  J.n - K1, J.ncols - K1, K1 - K0, wv[j].sigma, true, st);
        trailing_cols<T>(t, 0, J.ncols - K1, st);
      } else
        gl.add(J.A + K1 + K0 * ld, ld, J.A + K0 + K1 * ld, ld, J.A + K1 + K1 * ld, ld, J.n - K1, J.ncols - K1, K1 - K0);
    }
    gl.launch(st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int j = 0; j < njobs; ++j)
    hipLaunchKernelGGL(barrier_status_kernel, dim3(1), dim3(1), 0, st, (const int*)wv[j].bar_err, jobs[j].info);
  return hipGetLastError();
}

template <class T>
hipError_t lu_factor(T* A, int64_t n, int64_t lda, int32_t* ipiv, int* info, void* ws, hipStream_t st,
                     bool fp16_update, hipStream_t side) {
  return lu_factor_cols<T>(A, n, n, lda, ipiv, info, ws, st, fp16_update, side);
}

// Permutation of the rows produced by the pivots: row i of P A is row perm[i] of A.
static std::vector<int32_t> perm_from_ipiv(const std::vector<int32_t>& ipiv) {
  const int64_t n = (int64_t)ipiv.size();
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
  for (int64_t i = 0; i < n; ++i) std::swap(perm[i], perm[ipiv[i]]);
  return perm;
}

hipError_t lu_permutation(const int32_t* ipiv, int64_t n, int32_t* perm_dev, hipStream_t st) {
  std::vector<int32_t> h(n);
  hipError_t e = hipMemcpyAsync(h.data(), ipiv, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  std::vector<int32_t> p = perm_from_ipiv(h);
  e = hipMemcpyAsync(perm_dev, p.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

template <class T>
hipError_t tri_solve_lower_unit(const T* LU, int64_t n, int64_t lda, T* B, int64_t nrhs, int64_t ldb, hipStream_t st) {
  for (int64_t r0 = 0; r0 < n; r0 += LU_NB) {
    const int nb = (int)std::min<int64_t>(LU_NB, n - r0);
    hipLaunchKernelGGL((trsv_diag_kernel<T, true>), dim3(nblk(nrhs, 4)), dim3(256), 0, st, LU, lda, r0, nb, B, ldb,
                       nrhs);
    const int64_t rb = r0 + nb;
    if (rb < n)
      hipLaunchKernelGGL(trsv_update_kernel<T>, dim3(nblk(n - rb, 256)), dim3(256), 0, st, LU, lda, r0, nb, rb, n, B,
                         ldb, nrhs);
  }
  return hipGetLastError();
}

template <class T>
hipError_t tri_solve_upper(const T* LU, int64_t n, int64_t lda, T* B, int64_t nrhs, int64_t ldb, hipStream_t st) {
  const int64_t nblocks = (n + LU_NB - 1) / LU_NB;
  for (int64_t bi = nblocks - 1; bi >= 0; --bi) {
    const int64_t r0 = bi * LU_NB;
    const int nb = (int)std::min<int64_t>(LU_NB, n - r0);
    hipLaunchKernelGGL((trsv_diag_kernel<T, false>), dim3(nblk(nrhs, 4)), dim3(256), 0, st, LU, lda, r0, nb, B, ldb,
                       nrhs);
    if (r0 > 0)
      hipLaunchKernelGGL(trsv_update_kernel<T>, dim3(nblk(r0, 256)), dim3(256), 0, st, LU, lda, r0, nb, (int64_t)0, r0,
                         B, ldb, nrhs);
  }
  return hipGetLastError();
}

// U X = B for many right-hand sides: per 64-row block from the bottom, the diagonal-block solve (one wave per
// right-hand side) and the update of the rows above as an MFMA GEMM.
template <class T>
hipError_t tri_solve_upper_many(const T* LU, int64_t n, int64_t lda, T* B, int64_t nrhs, int64_t ldb,
                                hipStream_t st) {
  const int64_t nblocks = (n + LU_NB - 1) / LU_NB;
  for (int64_t bi = nblocks - 1; bi >= 0; --bi) {
    const int64_t r0 = bi * LU_NB;
    const int nb = (int)std::min<int64_t>(LU_NB, n - r0);
    hipLaunchKernelGGL((trsv_diag_kernel<T, false>), dim3(nblk(nrhs, 4)), dim3(256), 0, st, LU, lda, r0, nb, B, ldb,
                       nrhs);
    if (r0 > 0) gemm_sub<T>(LU + r0 * lda, lda, B + r0, ldb, B, ldb, r0, nrhs, nb, st);
  }
  return hipGetLastError();
}

template <class T>
__global__ void set_identity_kernel(T* X, int64_t n, int64_t ldx) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int64_t i = t % n, j = t / n;
  X[i + j * ldx] = i == j ? T(1) : T(0);
}

hipError_t invert(double* AI, int64_t n, int32_t* ipiv, int* info, void* ws, hipStream_t st) {
  hipLaunchKernelGGL(set_identity_kernel<double>, dim3(nblk(n * n, 256)), dim3(256), 0, st, AI + n * n, n, n);
  hipError_t e = lu_factor_cols<double>(AI, n, 2 * n, n, ipiv, info, ws, st, false, nullptr);
  if (e != hipSuccess) return e;
  return tri_solve_upper_many<double>(AI, n, n, AI + n * n, n, n, st);
}

template <class T>
hipError_t lu_solve(const T* LU, int64_t n, int64_t lda, const int32_t* ipiv, T* B, int64_t nrhs, int64_t ldb,
                    hipStream_t st) {
  int32_t* perm = nullptr;
  T* tmp = nullptr;
  hipError_t e = hipMalloc(&perm, sizeof(int32_t) * std::max<int64_t>(n, 1));
  if (e != hipSuccess) return e;
  e = hipMalloc(&tmp, sizeof(T) * std::max<int64_t>(ldb * nrhs, 1));
  if (e == hipSuccess) e = lu_permutation(ipiv, n, perm, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(permute_kernel<T>, dim3(nblk(n * nrhs, 256)), dim3(256), 0, st, perm, (const T*)B, tmp, n, nrhs,
                       ldb);
    e = hipMemcpyAsync(B, tmp, sizeof(T) * ldb * nrhs, hipMemcpyDeviceToDevice, st);
  }
  if (nrhs == 1 && n <= trsv_persistent_max_n<T>()) {
    // one right-hand side: persistent solves (one launch each, the 64-row blocks chained through flags) in place
    // of two launches per block
    int* flags = nullptr;
    const int64_t nf = (n + LU_NB - 1) / LU_NB + 1;  // block flags, then the error word
    if (e == hipSuccess) e = hipMalloc(&flags, sizeof(int) * nf);
    if (e == hipSuccess) e = hipMemsetAsync(flags, 0, sizeof(int) * nf, st);
    TrsvState s{flags, flags + nf - 1, 0, 1ll << 20};
    if (e == hipSuccess) e = trsv_persistent<T>(LU, n, lda, B, true, s, st);
    if (e == hipSuccess) e = trsv_persistent<T>(LU, n, lda, B, false, s, st);
    int herr = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, flags + nf - 1, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && herr != 0) e = hipErrorLaunchTimeOut;  // a block's wait overran
    if (flags) (void)hipFree(flags);
  } else {
    if (e == hipSuccess) e = tri_solve_lower_unit<T>(LU, n, lda, B, nrhs, ldb, st);
    if (e == hipSuccess) e = tri_solve_upper<T>(LU, n, lda, B, nrhs, ldb, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  (void)hipFree(perm);
  if (tmp) (void)hipFree(tmp);
  return e;
}

template <class TA>
void gemv(const TA* A, int64_t m, int64_t n, int64_t lda, const double* x, double* y, double alpha, double beta,
          hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(gemv_kernel<TA>, dim3(nblk(m, 256)), dim3(256), 0, st, A, m, n, lda, x, y, alpha, beta);
}

void build_kkt(const double* A, int64_t lda, int64_t m, int64_t k, double* K, int64_t ldk, hipStream_t st) {
  if (m + k > 0)
    hipLaunchKernelGGL(kkt_kernel, dim3(nblk(m + k, 256), (unsigned)(m + k)), dim3(256), 0, st, A, lda, m, k, K, ldk);
}

void convert_d2f(const double* a, float* b, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(d2f_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, a, b, n);
}
void convert_f2d(const float* a, double* b, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(f2d_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, a, b, n);
}
void convert_matrix_d2f(const double* A, int64_t lda, float* B, int64_t ldb, int64_t m, int64_t n, hipStream_t st) {
  if (m > 0 && n > 0) hipLaunchKernelGGL(d2f_matrix_kernel, dim3(nblk(m, 256), (unsigned)n), dim3(256), 0, st, A, lda, B, ldb, m, n);
}

template int lu_batched_capacity<double>();
template int lu_batched_capacity<float>();
template hipError_t lu_factor_batched<double>(const LuJob<double>*, int, hipStream_t, bool);
template hipError_t lu_factor_batched<float>(const LuJob<float>*, int, hipStream_t, bool);
template hipError_t lu_factor<double>(double*, int64_t, int64_t, int32_t*, int*, void*, hipStream_t, bool,
                                      hipStream_t);
template hipError_t lu_factor<float>(float*, int64_t, int64_t, int32_t*, int*, void*, hipStream_t, bool, hipStream_t);
template hipError_t lu_factor_cols<double>(double*, int64_t, int64_t, int64_t, int32_t*, int*, void*, hipStream_t,
                                           bool, hipStream_t);
template hipError_t lu_factor_cols<float>(float*, int64_t, int64_t, int64_t, int32_t*, int*, void*, hipStream_t, bool,
                                          hipStream_t);
template hipError_t lu_solve<double>(const double*, int64_t, int64_t, const int32_t*, double*, int64_t, int64_t, hipStream_t);
template hipError_t lu_solve<float>(const float*, int64_t, int64_t, const int32_t*, float*, int64_t, int64_t, hipStream_t);
template hipError_t tri_solve_upper<double>(const double*, int64_t, int64_t, double*, int64_t, int64_t, hipStream_t);
template hipError_t tri_solve_upper_many<double>(const double*, int64_t, int64_t, double*, int64_t, int64_t,
                                                 hipStream_t);
template hipError_t tri_solve_upper<float>(const float*, int64_t, int64_t, float*, int64_t, int64_t, hipStream_t);
template hipError_t tri_solve_lower_unit<double>(const double*, int64_t, int64_t, double*, int64_t, int64_t, hipStream_t);
template hipError_t tri_solve_lower_unit<float>(const float*, int64_t, int64_t, float*, int64_t, int64_t, hipStream_t);
template void gemv<double>(const double*, int64_t, int64_t, int64_t, const double*, double*, double, double, hipStream_t);
template void gemv<float>(const float*, int64_t, int64_t, int64_t, const double*, double*, double, double, hipStream_t);

}  // namespace cdj
