// serve.cpp -- request loop over the batched strain-step engine
// (include/vpfft_serve.h).  Host C++ over the C ABI of vpfft.h only.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/vpfft.h"
#include "../../include/vpfft_serve.h"
#include "capi_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Active {
  int req;         // index into the request array
  void* grid;
  uint64_t bytes;  // charged against the budget
  int next;        // next increment
  double voxels;
};

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// the material of a request: gdot0 = 1, its exponent, the reference's Voce
// defaults (MaterialGrid.cpp:322-336)
vpfft_material request_material(int nexp) {
  vpfft_material m;
  std::memset(&m, 0, sizeof m);
  for (int k = 0; k < VPFFT_NSLIP; k++) {
    m.gamma_dot_base[k] = 1.0;
    m.rate_sensitivity[k] = nexp;
    m.tau0[k] = m.tau1[k] = 11.0;
    m.theta0[k] = 430.0;
    m.theta1[k] = 110.0;
    for (int j = 0; j < VPFFT_NSLIP; j++) m.hardening[k * VPFFT_NSLIP + j] = 1.0;
  }
  return m;
}

double percentile(std::vector<double> v, double p) {  // nearest rank
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(p * (double)v.size());
  k = std::max<size_t>(1, std::min(k, v.size()));
  return v[k - 1];
}

}  // namespace

extern "C" int vpfft_serve(const vpfft_serve_config* cfg, const vpfft_request* reqs, int count,
                           vpfft_request_result* res, vpfft_serve_stats* stats) {
  using vpfft::set_error;
  if (!cfg || count < 0 || (count > 0 && (!reqs || !res))) return set_error(VPFFT_E_INVALID, "bad argument");
  if (cfg->precision != VPFFT_FP64 && cfg->precision != VPFFT_MIXED)
    return set_error(VPFFT_E_INVALID, "precision must be VPFFT_FP64 or VPFFT_MIXED");
  if (cfg->max_active < 1 || cfg->hbm_budget_bytes == 0 || cfg->max_iter < 1 || !(cfg->tol >= 0))
    return set_error(VPFFT_E_INVALID, "bad serving configuration");
  std::memset(res, 0, sizeof(vpfft_request_result) * (size_t)count);
  // a request with bad fields is refused alone (its status says why); the
  // others are served
  std::vector<char> refused(count, 0);
  for (int i = 0; i < count; i++) {
    const vpfft_request& q = reqs[i];
    int st = VPFFT_OK;
    if (q.edge < 2 || q.grains < 1 || q.num_increments < 1 || !q.edot || !std::isfinite(q.arrival_s) ||
        !std::isfinite(q.dt))
      st = set_error(VPFFT_E_INVALID, "request %d: bad fields", i);
    for (int k = 0; !st && k < q.num_increments; k++) {
      double ee = 0;
      bool fin = true;
      for (int c = 0; c < 5; c++) {
        fin = fin && std::isfinite(q.edot[5 * k + c]);
        ee += q.edot[5 * k + c] * q.edot[5 * k + c];
      }
      if (!fin) st = set_error(VPFFT_E_NONFINITE, "request %d: non-finite strain rate in increment %d", i, k);
      else if (!(ee > 0)) st = set_error(VPFFT_E_DEGENERATE, "request %d: zero strain rate in increment %d", i, k);
    }
    if (st) {
      res[i].status = st;
      refused[i] = 1;
    }
  }
  const bool fallback = cfg->fp64_fallback && cfg->precision == VPFFT_MIXED;
  std::vector<int> order(count);
  for (int i = 0; i < count; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });

  void* batch = nullptr;
  int rc = vpfft_batch_create(&batch);
  if (rc) return rc;
  if (cfg->tune) vpfft_batch_set_tuning(batch, 1);

  std::vector<Active> active;
  std::vector<double> latencies;
  uint64_t used = 0, max_used = 0;
  int max_active = 0, steps_run = 0, reruns_total = 0;
  long rve_steps = 0;
  double voxel_iters = 0;
  size_t next_pending = 0;
  const Clock::time_point t0 = Clock::now();

  auto finish = [&](size_t a, int status) {
    Active& A = active[a];
    vpfft_request_result& r = res[A.req];
    r.finished_s = seconds_since(t0);
    r.latency_s = r.finished_s - reqs[A.req].arrival_s;
    r.status = status;
    if (status == VPFFT_OK) latencies.push_back(r.latency_s);
    vpfft_grid_destroy(A.grid);
    used -= A.bytes;
    active.erase(active.begin() + (long)a);
  };

  while (next_pending < order.size() || !active.empty()) {
    // admission in arrival order, by slots and HBM budget
    while (next_pending < order.size()) {
      const int qi = order[next_pending];
      const vpfft_request& q = reqs[qi];
      if (refused[qi]) {
        next_pending++;
        continue;
      }
      if (q.arrival_s > seconds_since(t0) || (int)active.size() >= cfg->max_active) break;
      const vpfft_grid_desc desc = {q.edge, q.edge, q.edge, cfg->precision};
      uint64_t bytes = 0;
      if ((rc = vpfft_grid_bytes(&desc, &bytes))) {
        res[qi].status = rc;
        next_pending++;
        continue;
      }
      const long n = (long)q.edge * q.edge * q.edge;
      if (fallback) bytes += sizeof(double) * 32 * (uint64_t)n;
      if (bytes > cfg->hbm_budget_bytes) {  // never fits
        res[qi].status = VPFFT_E_NOMEM;
        next_pending++;
        continue;
      }
      if (used + bytes > cfg->hbm_budget_bytes) break;  // wait for memory to free up
      next_pending++;
      const vpfft_material mat = request_material(q.rate_sensitivity);
      void* g = nullptr;
      if ((rc = vpfft_grid_create(&desc, &mat, &g))) {
        res[qi].status = rc;
        continue;
      }
      std::vector<double> R(9 * (size_t)n);
      rc = vpfft_generate_voronoi(q.edge, q.edge, q.edge, q.grains, q.seed, R.data(), nullptr, nullptr);
      if (!rc) rc = vpfft_grid_upload(g, R.data(), nullptr, 1.0, nullptr);
      if (!rc) rc = vpfft_grid_set_voxel_length(g, 1.0, 1.0, 1.0);
      if (rc) {
        res[qi].status = rc;
        vpfft_grid_destroy(g);
        continue;
      }
      res[qi].admitted_s = seconds_since(t0);
      active.push_back(Active{qi, g, bytes, 0, (double)n});
      used += bytes;
      max_used = std::max(max_used, used);
      max_active = std::max(max_active, (int)active.size());
    }
    if (active.empty()) {
      if (next_pending < order.size()) {
        const double wait = reqs[order[next_pending]].arrival_s - seconds_since(t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
      }
      continue;
    }
    // one increment of every admitted request, in one batched step
    const int m = (int)active.size();
    std::vector<void*> gs(m);
    std::vector<double> ed(5 * (size_t)m), dts(m), avg(5 * (size_t)m);
    std::vector<vpfft_step_info> infos(m);
    for (int j = 0; j < m; j++) {
      const vpfft_request& q = reqs[active[j].req];
      gs[j] = active[j].grid;
      std::memcpy(&ed[5 * j], q.edot + 5 * active[j].next, 5 * sizeof(double));
      dts[j] = q.dt;
      if (fallback && (rc = vpfft_grid_save_state(gs[j]))) break;
    }
    if (!rc)
      rc = vpfft_batch_run_strain_step(batch, gs.data(), m, ed.data(), cfg->max_iter, dts.data(), cfg->tol,
                                       avg.data(), infos.data());
    steps_run++;
    if (rc && rc != VPFFT_E_DEGENERATE && rc != VPFFT_E_NONFINITE) break;  // the engine failed: stop serving
    rc = VPFFT_OK;
    for (int j = 0; j < m; j++) {  // work done, counted per attempt
      res[active[j].req].iterations += infos[j].iterations;
      voxel_iters += active[j].voxels * infos[j].iterations;
    }
    // mixed steps that ended unconverged and stalled, or met a non-finite
    // value: rerun them in FP64 from the saved state
    if (fallback) {
      std::vector<int> redo;
      for (int j = 0; j < m; j++)
        if ((infos[j].status == VPFFT_OK && !infos[j].converged && infos[j].err_ratio_5 >= cfg->stall_ratio) ||
            infos[j].status == VPFFT_E_NONFINITE)
          redo.push_back(j);
      if (!redo.empty()) {
        std::vector<void*> rg;
        std::vector<double> red, rdt;
        for (int j : redo) {
          vpfft_grid_restore_state(gs[j]);
          vpfft_grid_set_precision(gs[j], VPFFT_FP64);
          rg.push_back(gs[j]);
          red.insert(red.end(), &ed[5 * j], &ed[5 * j] + 5);
          rdt.push_back(dts[j]);
        }
        std::vector<double> ravg(5 * redo.size());
        std::vector<vpfft_step_info> rinf(redo.size());
        const int rrc = vpfft_batch_run_strain_step(batch, rg.data(), (int)rg.size(), red.data(), cfg->max_iter,
                                                    rdt.data(), cfg->tol, ravg.data(), rinf.data());
        steps_run++;
        for (size_t k = 0; k < redo.size(); k++) {
          const int j = redo[k];
          vpfft_grid_set_precision(gs[j], VPFFT_MIXED);
          std::memcpy(&avg[5 * j], &ravg[5 * k], 5 * sizeof(double));
          infos[j] = rinf[k];
          res[active[j].req].iterations += rinf[k].iterations;
          voxel_iters += active[j].voxels * rinf[k].iterations;
          if (rrc && rrc != VPFFT_E_DEGENERATE && rrc != VPFFT_E_NONFINITE && infos[j].status == VPFFT_OK)
            infos[j].status = rrc;
          res[active[j].req].fp64_reruns++;
          reruns_total++;
        }
      }
    }
    // record, then retire finished requests (from the back so indices stay valid)
    for (int j = 0; j < m; j++) {
      Active& A = active[j];
      const vpfft_request& q = reqs[A.req];
      vpfft_request_result& r = res[A.req];
      if (q.avg_stress) std::memcpy(q.avg_stress + 5 * A.next, &avg[5 * j], 5 * sizeof(double));
      if (infos[j].status == VPFFT_OK) {
        A.next++;
        r.increments_done = A.next;
        rve_steps++;
      }
    }
    for (int j = m - 1; j >= 0; j--) {
      const vpfft_request& q = reqs[active[j].req];
      if (infos[j].status != VPFFT_OK) finish((size_t)j, infos[j].status);
      else if (active[j].next >= q.num_increments) finish((size_t)j, VPFFT_OK);
    }
  }
  const int loop_rc = rc;
  for (size_t a = active.size(); a > 0; a--) finish(a - 1, loop_rc ? loop_rc : VPFFT_E_STATE);
  vpfft_batch_destroy(batch);
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->wall_s = seconds_since(t0);
    stats->rve_steps = rve_steps;
    stats->rve_steps_per_s = stats->wall_s > 0 ? (double)rve_steps / stats->wall_s : 0.0;
    stats->voxel_iters_per_s = stats->wall_s > 0 ? voxel_iters / stats->wall_s : 0.0;
    stats->p50_latency_s = percentile(latencies, 0.50);
    stats->p99_latency_s = percentile(latencies, 0.99);
    double sum = 0;
    for (double l : latencies) sum += l;
    stats->mean_latency_s = latencies.empty() ? 0.0 : sum / (double)latencies.size();
    stats->max_active_seen = max_active;
    stats->max_bytes_seen = max_used;
    stats->batched_steps = steps_run;
    stats->fp64_reruns = reruns_total;
  }
  return loop_rc;
}
