// Adjoint lane table of the register-path coil stage (host code, no HIP).
//
// The adjoint of the sampled dim-2 DFT adds each sample of a state into the
// 16 dim-2 rows of a group at its phase-encode row k1.  With one lane per
// row a wave runs as long as its most-sampled row; the table instead deals
// the state's samples, sorted by row, to the lanes in equal runs of cs =
// ceil(samples / lanes) and records per lane how its row segments are
// written: a segment whose row began in an earlier lane is that lane's
// carry, a row that begins in the lane and continues into later lanes is
// finished by the lane (its owner) adding their carries in lane order, and
// every other segment is a whole row written directly.  Empty rows are dealt
// one per lane, to be zeroed.  A state of more than kAdjMaxSamples samples
// takes one table block per chunk of kAdjMaxSamples consecutive row-sorted
// samples: each chunk's rows are summed within the chunk, and the chunks'
// contributions add up in the adjoint's output as the coil stage's sample
// chunks do.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace dis {

constexpr int kAdjLanes = 256;       // lanes of a coil-stage workgroup
constexpr int kAdjMaxSamples = 768;  // samples per table block (3 per lane): one coil-stage chunk
constexpr int kAdjMaxState = 16383;  // samples per state the 14-bit sample field holds

// Entry of a lane (.x/.y/.z): k2 (bits 0-9) | row k1 (bits 10-17) | sample
// within its state (bits 18-31).  Flags of the lane (.w): the first
// segment's row began in an earlier lane (its partial goes to the lane's
// carry slot); the last segment's row is owned and continues into the next
// `ncont` lanes, whose carries the owner adds in lane order; an empty row the
// lane zeroes; per entry, whether it holds a sample, whether a row segment
// closes after it and whether that segment is the lane's carry.
constexpr int kAdjRowShift = 10;
constexpr int kAdjSampleShift = 18;
constexpr int kAdjHead = 1;
constexpr int kAdjDeferred = 2;
constexpr int kAdjNcontShift = 2;   // 10 bits
constexpr int kAdjZrowShift = 12;   // 10 bits
constexpr int kAdjZvalid = 1 << 22;
constexpr int kAdjCloseShift = 23;  // 3 bits, one per entry
constexpr int kAdjCarryShift = 26;  // 3 bits: the segment closing at that entry is the lane's carry
constexpr int kAdjValidShift = 29;  // 3 bits: the entry holds a sample

// Builds the table (CoilArgs::adjtab, blocks of kAdjLanes x int4) and the
// per-state block offsets off[S+1] (CoilArgs::adjoff) from the per-state
// CSR by k1: rp [S][n1+1] row pointers, ent (k2 << 16) | sample within the
// state.  Returns false (no table) when a state has more than kAdjMaxState
// samples or n1 > kAdjLanes or n2 > 1024.
inline bool adjoint_lane_table(int n1, int n2, int S, const int32_t* state_off, const std::vector<int>& rp,
                               const std::vector<int>& ent, std::vector<int>* tab, std::vector<int>* off) {
  if (n1 > kAdjLanes || n2 > 1024) return false;
  off->assign(S + 1, 0);
  for (int s = 0; s < S; ++s) {
    const int nent = state_off[s + 1] - state_off[s];
    if (nent > kAdjMaxState) return false;
    (*off)[s + 1] = (*off)[s] + std::max(1, (nent + kAdjMaxSamples - 1) / kAdjMaxSamples);
  }
  tab->assign((size_t)(*off)[S] * kAdjLanes * 4, 0);
  std::vector<int> rowof(kAdjMaxSamples), r(n1 + 1);
  for (int s = 0; s < S; ++s) {
    const int o0 = state_off[s], ns = state_off[s + 1] - o0;
    const int* rs = rp.data() + (size_t)s * (n1 + 1);
    for (int ch = 0; ch < (*off)[s + 1] - (*off)[s]; ++ch) {
      // the chunk's own CSR: entries [c0, c0 + nent) of the state's row-sorted list
      const int c0 = ch * kAdjMaxSamples, nent = std::min(kAdjMaxSamples, ns - c0);
      for (int k = 0; k <= n1; ++k) r[k] = std::min(std::max(rs[k], c0), c0 + nent) - c0;
      for (int k = 0; k < n1; ++k)
        for (int e = r[k]; e < r[k + 1]; ++e) rowof[e] = k;
      const int cs = std::max(1, (nent + kAdjLanes - 1) / kAdjLanes);
      int* T = tab->data() + (size_t)((*off)[s] + ch) * kAdjLanes * 4;
      // empty rows, dealt one per lane (n1 <= kAdjLanes)
      int z = 0;
      for (int k = 0; k < n1; ++k)
        if (r[k] == r[k + 1]) T[4 * (z++) + 3] |= kAdjZvalid | (k << kAdjZrowShift);
      for (int L = 0; L < kAdjLanes; ++L) {
        const int eb = cs * L, ee = std::min(eb + cs, nent);
        if (eb >= ee) continue;
        int info = T[4 * L + 3];
        for (int e = eb; e < ee; ++e) {
          const int row = rowof[e], v = ent[o0 + c0 + e];
          T[4 * L + (e - eb)] = (v >> 16) | (row << kAdjRowShift) | ((v & 0xffff) << kAdjSampleShift);
          info |= 1 << (kAdjValidShift + (e - eb));
          // a segment closes at the row's last entry in this lane, unless the
          // row is owned here and continues into later lanes (deferred)
          const bool row_ends = e + 1 < ee ? rowof[e + 1] != row : r[row + 1] == ee;
          const bool owned = r[row] >= eb;
          if (row_ends || (e + 1 == ee && !owned)) {
            info |= 1 << (kAdjCloseShift + (e - eb));
            if (!owned) info |= 1 << (kAdjCarryShift + (e - eb));  // only the first segment can be begun earlier
          }
        }
        if (r[rowof[eb]] < eb) info |= kAdjHead;
        const int lrow = rowof[ee - 1];
        if (r[lrow] >= eb && r[lrow + 1] > ee)
          info |= kAdjDeferred | ((((r[lrow + 1] - 1) / cs) - L) << kAdjNcontShift);
        T[4 * L + 3] = info;
      }
    }
  }
  return true;
}

}  // namespace dis
