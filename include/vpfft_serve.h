/* vpfft_serve.h -- a request loop over the batched strain-step engine of
 * vpfft.h (no reference counterpart: the reference runs one grid per job).
 *
 * Requests (one RVE and a sequence of strain increments each) arrive over
 * time.  The loop admits them in arrival order while the HBM budget and the
 * slot count allow, keeps every admitted RVE's fields resident on the GPU
 * between its increments, and advances all admitted RVEs by one increment per
 * batched step, so requests that arrive while others run join the next step
 * (continuous batching).  A mixed-precision step that ends unconverged with
 * its residual stalled (no longer decreasing) is rerun in FP64 from the saved
 * state.  A request's results do not depend on
 * which other requests share its steps. */
#ifndef VPFFT_SERVE_H_
#define VPFFT_SERVE_H_

#include <stdint.h>

#include "vpfft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double arrival_s;    /* arrival time, seconds after the start of the run */
  int edge;            /* cubic grid edge (factors 2, 3, 5, 7) */
  int grains;          /* periodic Voronoi grains (vpfft_generate_voronoi) */
  uint64_t seed;       /* microstructure seed */
  int rate_sensitivity;
  int num_increments;
  const double* edot;  /* num_increments x 5 strain rates */
  double dt;           /* time step of every increment */
  double* avg_stress;  /* out, may be NULL: num_increments x 5, <sigma> of each increment */
} vpfft_request;

typedef struct {
  double admitted_s;   /* when its grid was created */
  double finished_s;   /* when its last increment completed */
  double latency_s;    /* finished - arrival */
  int increments_done;
  int fp64_reruns;     /* increments rerun in FP64 */
  int iterations;      /* fixed-point iterations over all increments, FP64 reruns included */
  int status;          /* VPFFT_OK or the error that ended it */
} vpfft_request_result;

typedef struct {
  uint64_t hbm_budget_bytes;  /* device bytes the admitted grids may hold */
  int max_active;             /* admitted requests at once */
  int precision;              /* VPFFT_MIXED or VPFFT_FP64 */
  int fp64_fallback;          /* 1: rerun a mixed step in FP64 when it ends unconverged and stalled */
  double stall_ratio;         /* stalled: vpfft_step_info.err_ratio_5 >= stall_ratio (0 = always rerun
                                 an unconverged mixed step) */
  int max_iter;
  double tol;
  int tune;                   /* FFT block-shape tuning of the batch */
} vpfft_serve_config;

typedef struct {
  double wall_s;
  long rve_steps;             /* increments completed */
  double rve_steps_per_s;
  double voxel_iters_per_s;
  double p50_latency_s, p99_latency_s, mean_latency_s;
  int max_active_seen;
  uint64_t max_bytes_seen;
  int batched_steps;          /* batched strain steps run */
  int fp64_reruns;
} vpfft_serve_stats;

/* Run the requests (any order; served by arrival time, in real time) to
 * completion.  results: count entries; stats may be NULL.  Returns VPFFT_OK
 * unless the loop itself failed.  A request that cannot be served is marked
 * in its result and skipped, alone: VPFFT_E_NOMEM when it can never fit the
 * budget, VPFFT_E_INVALID for bad fields (edge, grains, increments, arrival
 * time, time step), VPFFT_E_NONFINITE / VPFFT_E_DEGENERATE for a non-finite
 * or zero strain rate in any increment; a request whose step fails
 * (degenerate voxels, non-finite values that an FP64 rerun does not cure)
 * ends with that step's status. */
int vpfft_serve(const vpfft_serve_config* cfg, const vpfft_request* requests, int count,
                vpfft_request_result* results, vpfft_serve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* VPFFT_SERVE_H_ */
