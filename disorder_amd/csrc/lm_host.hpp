// Host bookkeeping of one LMsolver call (Reconstruction/LMsolver.m:126-215):
// damped steps, restrictTransform, trial energies, per-state accept/reject
// and the damping/convergence updates.  Free of device code: the library
// drives it with callbacks that solve and evaluate on the GPU, and a CPU test
// drives it with scripted energies.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace dis {

// Rounds of the accept/reject loop before the call gives up.  With finite
// energies a rejected state doubles its damping every round, so w climbs from
// its 1e-8 floor past the 1e10 cut-off (which accepts the state) in at most
// 61 rounds.
constexpr int kLmMaxRounds = 256;

enum LmStatus {
  LM_OK = 0,
  LM_NONFINITE_T = 1,       // a motion parameter or damping weight is NaN or Inf on entry
  LM_NONFINITE_ENERGY = 2,  // the residual energy at the current motion is NaN or Inf
  LM_NO_CONVERGENCE = 3,    // the loop reached kLmMaxRounds
  LM_CALLBACK = 4,          // a solve or trial callback failed (its status is kept)
};

inline const char* lm_status_text(int s) {
  switch (s) {
    case LM_NONFINITE_T: return "LM: motion parameters or damping weights are not finite";
    case LM_NONFINITE_ENERGY: return "LM: residual energy is not finite (x or y holds NaN/Inf)";
    case LM_NO_CONVERGENCE: return "LM: accept/reject loop did not terminate";
    default: return "LM: ok";
  }
}

// restrictTransform (Methods/restrictTransform.m:28-29): rotations into
// [-pi, pi] by repeated 2 pi steps; far-out values are first reduced with
// fmod so the loop is bounded.
inline double lm_restrict_rotation(double v) {
  if (!std::isfinite(v)) return v;
  if (std::fabs(v) > 64 * M_PI) v = std::fmod(v, 2 * M_PI);
  while (v > M_PI) v -= 2 * M_PI;
  while (v < -M_PI) v += 2 * M_PI;
  return v;
}

// State of one LMsolver call for NM motion states; T is column-major
// [NM x 6] (T[s + NM g]).
struct LmUpdate {
  int NM = 0;
  std::vector<int> est;              // states not converged on entry
  std::vector<double> EnPrev, En;    // energies at T and at the trial motion
  std::vector<double> dGEff;         // [NM x 6] row per state: current damped step
  std::vector<unsigned char> flagw;  // state accepted (or converged)
  std::vector<double> Tup;           // trial motion, column-major like T
  int rounds = 0;
};

// Checks on entry: T and the damping weights w finite for every state.
inline int lm_check_motion(int NM, const double* T, const double* w) {
  for (long k = 0; k < 6L * NM; ++k)
    if (!std::isfinite(T[k])) return LM_NONFINITE_T;
  for (int s = 0; w && s < NM; ++s)
    if (!std::isfinite(w[s])) return LM_NONFINITE_T;
  return LM_OK;
}

// The update loop after the partial derivatives (LMsolver.m:126-215).
//   solve(list, steps): damped 6x6 solves for the states of `list`
//     (non-converged, w < 1e10), writing 6 steps per listed state (row
//     order of the list); a non-finite step is replaced by no step (the
//     reference's mldivide returns Inf/NaN on a singular system).
//   trial(list, Tup, En): energies at the trial motion for the listed
//     states, written to En[state].
// Both return 0 or a library status, which is passed back through *cb_rc.
// A trial energy that is not finite counts as a rejection: in the reference
// NaN compares false both ways, so the state would be neither accepted nor
// damped and LMsolver.m:129 would loop forever.
template <class SolveFn, class TrialFn>
int lm_update(LmUpdate& u, double* T, double* w, unsigned char* cT, const double tL[2], SolveFn solve,
              TrialFn trial, int* n_trials, int* cb_rc) {
  const int NM = u.NM;
  for (int s : u.est)
    if (!std::isfinite(u.EnPrev[s])) return LM_NONFINITE_ENERGY;
  u.En = u.EnPrev;
  u.dGEff.assign((size_t)NM * 6, 0.0);
  u.flagw.assign(NM, 0);
  for (int s = 0; s < NM; ++s)
    if (cT[s]) u.flagw[s] = 1;  // En = EnPrev = 0 for converged states
  u.Tup.assign((size_t)NM * 6, 0.0);
  for (u.rounds = 0; u.rounds < kLmMaxRounds; ++u.rounds) {
    std::vector<int> sol;
    for (int s : u.est)
      if (w[s] < 1e10) sol.push_back(s);
    if (!sol.empty()) {
      std::vector<double> steps(sol.size() * 6, 0.0);
      const int rc = solve(sol, steps);
      if (rc) {
        *cb_rc = rc;
        return LM_CALLBACK;
      }
      for (size_t i = 0; i < sol.size(); ++i) {
        bool finite = true;
        for (int g = 0; g < 6; ++g) finite = finite && std::isfinite(steps[i * 6 + g]);
        for (int g = 0; g < 6; ++g) u.dGEff[(size_t)sol[i] * 6 + g] = finite ? steps[i * 6 + g] : 0.0;
      }
    }
    // Tup = restrictTransform(T + dGEff)
    for (int s = 0; s < NM; ++s)
      for (int g = 0; g < 6; ++g) {
        const double v = T[s + (size_t)NM * g] + u.dGEff[(size_t)s * 6 + g];
        u.Tup[s + (size_t)NM * g] = g >= 3 ? lm_restrict_rotation(v) : v;
      }
    // trial energies of the states neither converged nor accepted
    std::vector<int> todo;
    for (int s : u.est)
      if (!u.flagw[s]) todo.push_back(s);
    if (n_trials) *n_trials += (int)todo.size();
    if (!todo.empty()) {
      const int rc = trial(todo, u.Tup, u.En);
      if (rc) {
        *cb_rc = rc;
        return LM_CALLBACK;
      }
    }
    bool all = true;
    for (int s = 0; s < NM; ++s) {
      if (w[s] > 1e10) u.En[s] = u.EnPrev[s];
      if (u.En[s] <= u.EnPrev[s]) u.flagw[s] = 1;
      if (!u.flagw[s]) all = false;
    }
    if (!all) {
      for (int s = 0; s < NM; ++s)
        if (!cT[s] && !u.flagw[s] && !(u.En[s] <= u.EnPrev[s])) w[s] *= 2.0;
      continue;
    }
    for (int s = 0; s < NM; ++s) {
      if (!cT[s]) w[s] /= 1.2;
      if (w[s] < 1e-8) w[s] *= 1.2;
    }
    for (size_t k = 0; k < u.Tup.size(); ++k) T[k] = u.Tup[k];
    for (int s = 0; s < NM; ++s) {
      double tra = 0, rot = 0;
      for (int g = 0; g < 3; ++g) tra = std::max(tra, std::fabs(u.dGEff[(size_t)s * 6 + g]));
      for (int g = 3; g < 6; ++g) rot = std::max(rot, std::fabs(u.dGEff[(size_t)s * 6 + g] * 180.0 / M_PI));
      if (tra < tL[0] && rot < tL[1]) cT[s] = 1;
    }
    return LM_OK;
  }
  return LM_NO_CONVERGENCE;
}

}  // namespace dis
