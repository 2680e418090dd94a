// Opt-in per-launch HIP-event profiler (the only process-wide state of the
// library; guarded by a mutex, off unless wcfd_profile_begin() is called).
// The events of a launch are stamped by the runtime with the kernel's own
// begin and end (hipExtLaunchKernelGGL in WCFD_LAUNCH_B).
//
// Krylov solves keep launching while the host has not yet seen their stop
// (two check groups in flight), and those launches exit at entry.  Each
// launch of a Krylov iteration carries its iteration index; when the solve
// returns, launches at or beyond the iteration count it ran are retired, so
// the totals cover only launches that did an iteration's work.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "common.h"

namespace wcfd {

namespace {
struct Rec {
  int id;
  int iter;     // Krylov iteration of the launch, -1 for any other launch
  bool dead;    // exited at entry: its solve had stopped before this iteration
  hipEvent_t a, b;
};
thread_local int t_iter = -1;
int32_t g_skipped[K_COUNT] = {};
std::mutex g_mu;
std::atomic<bool> g_on{false};
std::vector<hipEvent_t> g_pool;
std::vector<Rec> g_recs;
size_t g_next = 0;
const char* kNames[K_COUNT] = {
    "set_bc", "face_interp", "green_gauss", "mass_flux", "divergence", "assemble_momentum",
    "assemble_pressure", "rua_hbya", "relax_p", "face_p_fc", "grad_p_correct", "inv_diag", "spmv",
    "dot", "cg_init", "cg_1", "cg_2", "bicgstab_init", "bicgstab_1", "bicgstab_2", "bicgstab_3", "fold", "ell_to_csr", "term_weights", "equation",
    "convergence", "cg_xcd"};
}  // namespace

bool prof_active() { return g_on.load(std::memory_order_relaxed); }

bool prof_events(int id, hipEvent_t* start, hipEvent_t* stop) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_on || g_next + 2 > g_pool.size()) return false;
  Rec r{id, t_iter, false, g_pool[g_next], g_pool[g_next + 1]};
  g_next += 2;
  g_recs.push_back(r);
  *start = r.a;
  *stop = r.b;
  return true;
}

void prof_set_iter(int it) { t_iter = it; }

size_t prof_mark() {
  if (!prof_active()) return 0;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_recs.size();
}

void prof_retire(size_t from, int iters_run) {
  if (!prof_active()) return;
  std::lock_guard<std::mutex> lk(g_mu);
  for (size_t i = from; i < g_recs.size(); ++i)
    if (g_recs[i].iter >= iters_run) g_recs[i].dead = true;
}

}  // namespace wcfd

using namespace wcfd;

extern "C" {

int wcfd_kernel_count(void) { return K_COUNT; }

const char* wcfd_kernel_name(int id) { return (id >= 0 && id < K_COUNT) ? kNames[id] : "?"; }

int wcfd_profile_begin(int capacity) {
  if (capacity <= 0) return WCFD_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto e : g_pool) (void)hipEventDestroy(e);
  g_pool.clear();
  g_recs.clear();
  g_next = 0;
  g_pool.resize(2 * static_cast<size_t>(capacity));
  for (auto& e : g_pool) {
    int rc = check_hip(hipEventCreate(&e));
    if (rc) return rc;
  }
  g_recs.reserve(capacity);
  g_on = true;
  return WCFD_OK;
}

int wcfd_profile_end(double* total_ms, int32_t* counts, int n_ids) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = false;
  if (n_ids < K_COUNT || !total_ms || !counts) return WCFD_EINVAL;
  for (int i = 0; i < K_COUNT; ++i) { total_ms[i] = 0.0; counts[i] = 0; g_skipped[i] = 0; }
  for (auto& r : g_recs) {
    if (r.dead) {
      g_skipped[r.id] += 1;
      continue;
    }
    int rc = check_hip(hipEventSynchronize(r.b));
    if (rc) return rc;
    float ms = 0.f;
    rc = check_hip(hipEventElapsedTime(&ms, r.a, r.b));
    if (rc) return rc;
    total_ms[r.id] += ms;
    counts[r.id] += 1;
  }
  for (auto e : g_pool) (void)hipEventDestroy(e);
  g_pool.clear();
  g_recs.clear();
  g_next = 0;
  return WCFD_OK;
}

int wcfd_profile_timeline(int32_t* ids, double* start_ms, double* end_ms, int capacity) {
  if (capacity < 0 || (capacity > 0 && (!ids || !start_ms || !end_ms))) return WCFD_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_recs.empty()) return 0;
  const hipEvent_t base = g_recs.front().a;
  int n = 0;
  for (auto& r : g_recs) {
    if (n >= capacity) break;
    int rc = check_hip(hipEventSynchronize(r.b));
    if (rc) return -rc;
    float a = 0.f, b = 0.f;
    if ((rc = check_hip(hipEventElapsedTime(&a, base, r.a)))) return -rc;
    if ((rc = check_hip(hipEventElapsedTime(&b, base, r.b)))) return -rc;
    ids[n] = r.dead ? -1 - r.id : r.id;
    start_ms[n] = a;
    end_ms[n] = b;
    ++n;
  }
  return n;
}

int wcfd_profile_skipped(int32_t* counts, int n_ids) {
  if (n_ids < K_COUNT || !counts) return WCFD_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  for (int i = 0; i < K_COUNT; ++i) counts[i] = g_skipped[i];
  return WCFD_OK;
}

}  // extern "C"
