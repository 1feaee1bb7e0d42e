// Host model of the register coil stage's adjoint as it consumes the lane
// table of disorder_amd/csrc/lanetab.hpp: per lane, the same sequence the
// kernel runs (each entry's contribution added with the restart factor of
// the previous entry's close bit, a closing segment written to the tile row
// or the lane's carry slot, the empty row zeroed; after the barrier the
// deferred owner adds the next ncont lanes' carries in lane order).  Sample
// values are integers, so every row must equal the direct sum of its
// samples exactly, and every row must be written exactly once.
//
//   lanetab_harness TRIALS SEED [plant]
// prints "ok <rows>" and exits 0, or names the first violation and exits 1.
// With "plant", one close bit of one lane per table block is flipped after
// the table is built, and the run must report a violation.  States hold up
// to 2,000 samples, so that they span up to three blocks of 768.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../disorder_amd/csrc/lanetab.hpp"

using namespace dis;

// One table block: `chunk` lists the samples (within the state) of the
// block's chunk; `val` holds every sample value of the state.
static int check_block(const int* T, int n1, const std::vector<int>& k1, const std::vector<int>& k2, int o0,
                       const std::vector<int>& chunk, const std::vector<long long>& val, long long* rows) {
  std::vector<long long> Z(n1, 0), direct(n1, 0), carry(kAdjLanes, 0), acc_end(kAdjLanes, 0);
  std::vector<int> writes(n1, 0), lrow_end(kAdjLanes, 0), seen(val.size(), 0);
  for (int jl : chunk) direct[k1[o0 + jl]] += val[jl];
  for (int L = 0; L < kAdjLanes; ++L) {
    const int info = T[4 * L + 3];
    long long acc = 0;
    int lrow = 0, last = -1;
    for (int i = 0; i < 3; ++i) {
      if (!((info >> (kAdjValidShift + i)) & 1)) continue;
      const int e = T[4 * L + i];
      last = i;
      const int kk2 = e & 1023, row = (e >> kAdjRowShift) & 0xff, jl = (e >> kAdjSampleShift) & 0x3fff;
      if (jl >= (int)val.size() || k1[o0 + jl] != row || k2[o0 + jl] != kk2) return printf("entry %d of lane %d\n", i, L), 1;
      if (seen[jl]++) return printf("sample %d dealt twice\n", jl), 1;
      const long long keep = i > 0 && ((info >> (kAdjCloseShift + i - 1)) & 1) ? 0 : 1;
      acc = acc * keep + val[jl];
      lrow = row;
      if ((info >> (kAdjCloseShift + i)) & 1) {
        if ((info >> (kAdjCarryShift + i)) & 1) {
          carry[L] = acc;
        } else {
          Z[row] = acc;
          ++writes[row];
        }
      }
    }
    if (last >= 0 && bool((info >> (kAdjCloseShift + last)) & 1) == bool(info & kAdjDeferred))
      return printf("lane %d: last segment neither closed nor deferred\n", L), 1;
    if (info & kAdjZvalid) {
      const int zr = (info >> kAdjZrowShift) & 1023;
      if (zr >= n1) return printf("lane %d: empty row %d out of range\n", L, zr), 1;
      Z[zr] = 0;
      ++writes[zr];
    }
    acc_end[L] = acc;
    lrow_end[L] = lrow;
  }
  for (int L = 0; L < kAdjLanes; ++L) {
    const int info = T[4 * L + 3];
    if (!(info & kAdjDeferred)) continue;
    const int nc = (info >> kAdjNcontShift) & 1023;
    if (L + nc >= kAdjLanes) return printf("lane %d: carries beyond the last lane\n", L), 1;
    long long acc = acc_end[L];
    for (int l = 1; l <= nc; ++l) acc += carry[L + l];
    Z[lrow_end[L]] = acc;
    ++writes[lrow_end[L]];
  }
  for (int k = 0; k < n1; ++k) {
    if (writes[k] != 1) return printf("row %d written %d times\n", k, writes[k]), 1;
    if (Z[k] != direct[k]) return printf("row %d: %lld against %lld\n", k, Z[k], direct[k]), 1;
  }
  for (int jl : chunk)
    if (!seen[jl]) return printf("sample %d of the chunk not dealt\n", jl), 1;
  *rows += n1;
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const int trials = atoi(argv[1]);
  const bool plant = argc > 3 && !strcmp(argv[3], "plant");
  std::mt19937 rng((unsigned)atoi(argv[2]));
  long long rows = 0;
  int caught = 0, planted = 0;
  for (int trial = 0; trial < trials; ++trial) {
    const int n1 = 1 + rng() % kAdjLanes, n2 = 1 + rng() % 64, S = 1 + rng() % 3;
    std::vector<int> off(S + 1, 0), k1, k2;
    for (int s = 0; s < S; ++s) {
      // empty, a few samples, exactly 768, any count, or half of them in one row
      const int mode = rng() % 5;
      const int ns = mode == 0 ? 0 : mode == 1 ? 1 + rng() % 10 : mode == 2 ? kAdjMaxSamples * (1 + rng() % 2)
                                                                                  : (int)(rng() % 2001);
      const int hot = rng() % n1;
      for (int j = 0; j < ns; ++j) {
        k1.push_back((mode == 4 && rng() % 2) ? hot : (int)(rng() % n1));
        k2.push_back(rng() % n2);
      }
      off[s + 1] = off[s] + ns;
    }
    // per-state CSR by k1, stable in sample order (as dis_plan_create builds it)
    std::vector<int> rp((size_t)S * (n1 + 1), 0), ent(std::max(1, off[S]));
    for (int s = 0; s < S; ++s) {
      int* r = rp.data() + (size_t)s * (n1 + 1);
      for (int j = off[s]; j < off[s + 1]; ++j) ++r[k1[j] + 1];
      for (int k = 0; k < n1; ++k) r[k + 1] += r[k];
      std::vector<int> fill(r, r + n1);
      for (int j = off[s]; j < off[s + 1]; ++j) ent[off[s] + fill[k1[j]]++] = (k2[j] << 16) | (j - off[s]);
    }
    std::vector<int> tab, toff;
    if (!adjoint_lane_table(n1, n2, S, off.data(), rp, ent, &tab, &toff)) return printf("no table\n"), 1;
    for (int s = 0; s < S; ++s) {
      const int ns = off[s + 1] - off[s];
      std::vector<long long> val(ns), total(n1, 0), direct(n1, 0);
      for (auto& v : val) v = 1 + rng() % 1000003;
      const int nch = toff[s + 1] - toff[s];
      if (nch != std::max(1, (ns + kAdjMaxSamples - 1) / kAdjMaxSamples)) return printf("state %d: %d blocks\n", s, nch), 1;
      for (int ch = 0; ch < nch; ++ch) {
        // the chunk: row-sorted entries [768 ch, 768 ch + 768) of the state's CSR
        std::vector<int> chunk;
        for (int e = ch * kAdjMaxSamples; e < std::min(ns, (ch + 1) * kAdjMaxSamples); ++e)
          chunk.push_back(ent[off[s] + e] & 0xffff);
        int* T = tab.data() + (size_t)(toff[s] + ch) * kAdjLanes * 4;
        bool flipped = false;
        if (plant && !chunk.empty()) {
          const int L = rng() % ((chunk.size() + 2) / 3);  // a lane that holds samples
          T[4 * L + 3] ^= 1 << kAdjCloseShift;
          flipped = true;
          ++planted;
        }
        const int bad = check_block(T, n1, k1, k2, off[s], chunk, val, &rows);
        if (bad && !flipped) return 1;
        if (bad) ++caught;
      }
    }
  }
  if (plant) {
    printf("planted %d caught %d\n", planted, caught);
    return caught == planted ? 0 : 1;
  }
  printf("ok %lld\n", rows);
  return 0;
}
