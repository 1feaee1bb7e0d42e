// CPU harness of the LM update loop (disorder_amd/csrc/lm_host.hpp): the
// library's device solve and trial-energy callbacks are replaced by stubs,
// so scripted energies (NaN and Inf included) drive the same loop the
// library runs.
#include <cstring>
#include <vector>

#include "../../disorder_amd/csrc/lm_host.hpp"

extern "C" {

// steps: [NM x 6] row per state, the undamped step s_0; the stub solve
// returns s_0 * winit / w (the damped step shrinks as w grows, as
// -(winit/w) MH\g does).  trial: [rounds x NM] energies returned by the stub
// trial callback in each round (the last row repeats).  Returns the
// LmStatus; *rounds receives the number of rounds run.
int harness_lm(int NM, double* T, double* w, unsigned char* cT, const double* tL, double winit, const double* EnPrev,
               const double* steps, const double* trial, int trial_rows, int* rounds, int* n_trials,
               double* En_out) {
  int st = dis::lm_check_motion(NM, T, w);
  if (st) return st;
  dis::LmUpdate u;
  u.NM = NM;
  for (int s = 0; s < NM; ++s)
    if (!cT[s]) u.est.push_back(s);
  u.EnPrev.assign(EnPrev, EnPrev + NM);
  int round = 0;
  auto solve = [&](const std::vector<int>& sol, std::vector<double>& out) -> int {
    for (size_t i = 0; i < sol.size(); ++i)
      for (int g = 0; g < 6; ++g) out[i * 6 + g] = steps[(size_t)sol[i] * 6 + g] * winit / w[sol[i]];
    return 0;
  };
  auto tri = [&](const std::vector<int>& todo, const std::vector<double>&, std::vector<double>& En) -> int {
    const int r = round < trial_rows ? round : trial_rows - 1;
    for (int s : todo) En[s] = trial[(size_t)r * NM + s];
    ++round;
    return 0;
  };
  int cb = 0;
  *n_trials = 0;
  st = dis::lm_update(u, T, w, cT, tL, solve, tri, n_trials, &cb);
  *rounds = u.rounds + (st == dis::LM_OK ? 1 : 0);
  for (int s = 0; s < NM && !u.En.empty(); ++s) En_out[s] = u.En[s] < u.EnPrev[s] ? u.En[s] : u.EnPrev[s];
  return st;
}

double harness_restrict(double v) { return dis::lm_restrict_rotation(v); }
}
