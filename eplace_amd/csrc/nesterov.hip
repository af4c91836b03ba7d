// Device-resident Nesterov optimiser pieces (Optimization/nesterov.hpp,
// optutils.hpp) and the position setter (PlaceDB::setModuleCenter_2D,
// placedb.cpp:253-288).  Compiled with -ffp-contract=off so the element
// update rounds like the reference's VECTOR_3D arithmetic.
#include "ep_internal.h"
#include "reduce.h"

namespace ep {

__global__ void set_positions_kernel(int64_t n, const float2 *__restrict__ xy, const int32_t *__restrict__ idx,
                                     float2 *pos, const float *__restrict__ w, const float *__restrict__ h,
                                     Core k) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int o = idx[t];
  pos[o] = clamp_center(xy[t].x, xy[t].y, w[o], h[o], k);
}

__global__ void copy_float2_kernel(int64_t n, float2 *dst, const float2 *src) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Step length 1/L from the Lipschitz partial sums (optutils.hpp:19-35), the
// momentum a_{k+1} = (1 + sqrt(4 a_k^2 + 1))/2 and (a_k - 1)/a_{k+1}
// (nesterov.hpp:375-415).  Iteration 0 uses step 1.  A non-finite 1/L
// (identical consecutive positions, 0/0) keeps the previous step and is
// counted in bad_numeric.
// Step of an iteration: 1/L from the Lipschitz sums of the combine
// (optutils.hpp:19-35; 1 at iteration 0), or with bb the Barzilai-Borwein
// step of opt_step_bb (nesterov.hpp:245-281): the short step s.y / y.y when
// positive, else min(|s| / |y|, 1/L), s = y_k - y_{k-1}, y = g_k - g_{k-1}.
// Also the Nesterov coefficient a_{k+1} and the controller's gradient norms.
__device__ void step_scalar_block(int nblocks, double *partials, DevScalars *sc, int bb, int lip_slot) {
  if (sc->done) return;
  const double *const src[5] = {partials + lip_slot, partials + lip_slot + 1, partials + kPartialNorm,
                                partials + kPartialNorm + 1, partials + kPartialSY};
  double sum[5];
  block_sum_partials_strided_n(src, nblocks, 2, sum);
  const double num = sum[0], den = sum[1], nw = sum[2], nd = sum[3], sy = sum[4];
  if (threadIdx.x == 0) {
    sc->gwl_norm2 = nw;
    sc->gd_norm2 = nd;
    double step = 1.0;
    if (sc->iter > 0) {
      double L = sqrt(num) / sqrt(den);
      step = 1.0 / L;
      if (bb) {
        const double short_step = sy / num;
        if (short_step > 0.0) {
          step = short_step;
          sc->bb_short += 1;
        } else {
          step = fmin(sqrt(den) / sqrt(num), step);
        }
      }
      if (!isfinite(step) || step <= 0.0) {
        sc->bad_numeric += 1;
        step = sc->step;
      }
    }
    sc->lip_num = num;
    sc->lip_den = den;
    sc->step = step;
    double a = sc->a_k;
    double an = (1.0 + sqrt(4.0 * a * a + 1.0)) / 2.0;
    sc->a_next = an;  // committed by done_check
    sc->coef = (a - 1.0) / an;
  }
}

__global__ void __launch_bounds__(kFinish) step_scalar_kernel(int nblocks, double *partials, DevScalars *sc,
                                                              int bb, int lip_slot) {
  step_scalar_block(nblocks, partials, sc, bb, lip_slot);
}

// Backtracking acceptance (nesterov.hpp:337-364): accept when
// 0.95 * step <= 1/L', L' from the trial evaluation; step := 1/L' either way.
__global__ void __launch_bounds__(kFinish) accept_kernel(int nblocks, double *partials, DevScalars *sc) {
  if (sc->done) return;
  const double *const src[2] = {partials + kPartialLip, partials + kPartialLip + 1};
  double sum[2];
  block_sum_partials_strided_n(src, nblocks, 2, sum);
  const double num = sum[0], den = sum[1];
  if (threadIdx.x == 0) {
    double ns = 1.0 / (sqrt(num) / sqrt(den));
    if (!isfinite(ns) || ns <= 0.0) {
      sc->bad_numeric += 1;
      sc->accepted = 1;
      return;
    }
    int ok = 0.95 * sc->step <= ns;
    sc->accepted = ok;
    if (!ok) {
      sc->backtracks += 1;
      if (sc->trial == 0 && sc->ctrl_active) sc->fallbacks += 1;  // controller step rejected
    }
    sc->trial += 1;
    sc->step = ns;
  }
}

// The move of every object of the stage (move_object).
__global__ void __launch_bounds__(kBlock) nesterov_move_kernel(MoveArgs a, int O, float2 *pos,
                                                               const float *__restrict__ w,
                                                               const float *__restrict__ h, const DevScalars *sc) {
  if (sc->done) return;
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= O || !a.in_stage[o]) return;
  move_object(a, o, pos, w[o], h[o], (float)sc->step, (float)sc->coef);
}

// End of an iteration: commit a_{k+1}, advance the counter, evaluate the
// stop condition for the next iteration (nesterov.hpp:86-149) with the tau
// of the last evaluation, append a log row.
__device__ void done_check_body(DevScalars *sc, int next_iter, double *log, int log_cap, int set_iter,
                                int apply_next) {
  if (sc->done) return;
  // updatePenaltyFactor as the iteration's last trial staged it
  // (reduce_wl_kernel, lambda_update 3 and 4)
  if (apply_next) {
    sc->ref_mult_ema += (::log(sc->ref_mult_next) - sc->ref_mult_ema) * (1.0 / 16.0);
    sc->lambda = sc->lambda_next;
    sc->last_hpwl = sc->hpwl;
  }
  int k = next_iter - 1;
  if (log && k >= 0 && k < log_cap) {
    double *r = log + (size_t)k * EP_LOG_COLS;
    r[EP_LOG_ITER] = k;
    r[EP_LOG_TAU] = sc->overflow;
    r[EP_LOG_LAMBDA] = sc->lambda;
    r[EP_LOG_HPWL] = sc->hpwl;
    r[EP_LOG_STEP] = sc->step;
    r[EP_LOG_INV_GAMMA] = sc->inv_gamma_x;
    r[EP_LOG_EVALS] = sc->evals;
    r[EP_LOG_BACKTRACKS] = sc->backtracks;
    for (int i = 0; i < EP_CTRL_FEATURES; ++i) r[EP_LOG_FEATURES + i] = sc->feat[i];
    r[EP_LOG_STEP_PROP] = sc->step_prop;
    r[EP_LOG_STEP_USED] = sc->step_used;
    r[EP_LOG_LAMBDA_BEFORE] = sc->lambda_before;
    r[EP_LOG_FALLBACKS] = sc->fallbacks;
  }
  if (set_iter) {
    sc->a_k = sc->a_next;
    sc->iter = next_iter;
    // accepted step history for the controller features
    sc->step_hist[2] = sc->step_hist[1];
    sc->step_hist[1] = sc->step_hist[0];
    sc->step_hist[0] = sc->step;
    if (k == 0) sc->step0 = sc->step;
  }
  bool stop;
  if (sc->stage == EP_STAGE_MGP) stop = ((float)sc->overflow < sc->target_overflow) || (next_iter > sc->max_iter);
  else if (sc->stage == EP_STAGE_FILLERONLY) stop = next_iter >= 20;
  else stop = ((float)sc->overflow < 0.07f) || (next_iter > sc->max_iter);
  if (stop) sc->done = 1;
}

__global__ void done_check_kernel(DevScalars *sc, int next_iter, double *log, int log_cap, int set_iter,
                                  int apply_next) {
  done_check_body(sc, next_iter, log, log_cap, set_iter, apply_next);
}

// The end of a backtracking iteration and the step of the next one in one
// launch: the stop check (applying the staged penalty update) by thread 0,
// then, after the barrier, the step from the accepted trial's Lipschitz sums.
__global__ void __launch_bounds__(kFinish) done_step_kernel(DevScalars *sc, int next_iter, double *log, int log_cap,
                                                            int nblocks, double *partials, int lip_slot) {
  if (threadIdx.x == 0) done_check_body(sc, next_iter, log, log_cap, 1, 1);
  __syncthreads();
  step_scalar_block(nblocks, partials, sc, 0, lip_slot);
}


static int blocks_for(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

static Core core_of(const Ctx &c) { return Core{c.core_llx, c.core_lly, c.core_urx, c.core_ury}; }

void launch_set_positions(Ctx &c, const float2 *xy, const int32_t *idx, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(set_positions_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, c.stream, n, xy, idx, c.pos, c.w,
                     c.h, core_of(c));
}

void launch_positions_to_ref(Ctx &c, float2 *dst) {
  hipLaunchKernelGGL(copy_float2_kernel, dim3(blocks_for(c.O)), dim3(kBlock), 0, c.stream, (int64_t)c.O, dst,
                     c.pos);
}

void launch_copy_float2(Ctx &c, float2 *dst, const float2 *src, int64_t n) {
  hipLaunchKernelGGL(copy_float2_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, c.stream, n, dst, src);
}

void launch_step_scalar(Ctx &c, int bb, int lip_slot) {
  hipLaunchKernelGGL(step_scalar_kernel, dim3(1), dim3(kFinish), 0, c.stream, object_blocks(c.O), c.partials,
                     c.sc, bb, lip_slot);
}

void launch_accept(Ctx &c) {
  hipLaunchKernelGGL(accept_kernel, dim3(1), dim3(kFinish), 0, c.stream, object_blocks(c.O), c.partials, c.sc);
}

void launch_nesterov_move_ex(Ctx &c, const float2 *ref, const float2 *grad, const float2 *main_in,
                             float2 *main_out, float2 *ref_unclamped_out, float2 *save_ref, float2 *save_grad) {
  const MoveArgs a{c.in_stage, ref, grad, main_in, main_out, ref_unclamped_out, save_ref, save_grad, core_of(c)};
  hipLaunchKernelGGL(nesterov_move_kernel, dim3(blocks_for(c.O)), dim3(kBlock), 0, c.stream, a, c.O, c.pos, c.w,
                     c.h, c.sc);
}

void launch_done_step(Ctx &c, int next_iter, double *log, int log_cap, int lip_slot) {
  hipLaunchKernelGGL(done_step_kernel, dim3(1), dim3(kFinish), 0, c.stream, c.sc, next_iter, log, log_cap,
                     object_blocks(c.O), c.partials, lip_slot);
}

void launch_nesterov_move_hist(Ctx &c, const float2 *ref, const float2 *grad, const float2 *main_in,
                               float2 *main_out, float2 *ref_unclamped_out, float2 *save_ref, float2 *save_grad) {
  launch_move_hist(c, MoveArgs{c.in_stage, ref, grad, main_in, main_out, ref_unclamped_out, save_ref, save_grad,
                               core_of(c)});
}

void launch_done_check_ex(Ctx &c, int next_iter, double *log, int log_cap, int set_iter, int apply_next) {
  hipLaunchKernelGGL(done_check_kernel, dim3(1), dim3(1), 0, c.stream, c.sc, next_iter, log, log_cap, set_iter,
                     apply_next);
}

// The scalars to pinned coherent host memory, then the sequence number with
// a system-scope release: once the host sees the number, the scalars it
// reads are this point's.
__global__ void __launch_bounds__(64) publish_scalars_kernel(const DevScalars *sc, Ctx::HostScalars *host,
                                                             unsigned int seq) {
  static_assert(sizeof(DevScalars) % 4 == 0, "scalars are copied as 32-bit words");
  const uint32_t *src = reinterpret_cast<const uint32_t *>(sc);
  uint32_t *dst = reinterpret_cast<uint32_t *>(&host->s);
  for (int i = threadIdx.x; i < (int)(sizeof(DevScalars) / 4); i += blockDim.x) dst[i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&host->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_publish_scalars(Ctx &c, unsigned int seq) {
  hipLaunchKernelGGL(publish_scalars_kernel, dim3(1), dim3(64), 0, c.stream, c.sc, c.h_sc, seq);
}

// Holds the stream for `ticks` periods of the constant-rate wall clock, so
// that the host has enqueued a whole evaluation before the device starts it
// when the evaluation's phases are timed with events (otherwise an event can
// be stamped while the device waits for the host to submit the next kernel).
// The loop is also bounded by a count of sleeps, so it ends even if the
// clock did not advance.
__global__ void __launch_bounds__(64) hold_kernel(unsigned long long ticks) {
  const unsigned long long t0 = wall_clock64();
  for (int i = 0; i < 100000 && wall_clock64() - t0 < ticks; ++i) __builtin_amdgcn_s_sleep(8);
}

void launch_hold(Ctx &c, unsigned long long ticks) {
  hipLaunchKernelGGL(hold_kernel, dim3(1), dim3(64), 0, c.stream, ticks);
}



}  // namespace ep
