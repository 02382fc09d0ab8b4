// Many independent systems (replicas of uneven size) stepped together, and
// the autotuner that picks launch shapes and the schedule.  No counterpart
// in the reference, which runs one system per process.
//
// Each replica is an ordinary engine with its own stream and device buffers.
// Schedules:
//   MGPU_SCHED_GRAPH          one hipGraph holds one MD step of every
//                             replica as parallel branches (fork/join
//                             events); a step of the whole set is one
//                             hipGraphLaunch;
//   MGPU_SCHED_STREAM_GRAPHS  each replica replays its own step graph on its
//                             own stream, round robin;
//   MGPU_SCHED_STREAMS        eager launches, round robin over the streams.
// The autotuner first times each engine's tunable launch shapes on its own
// force evaluation (pure: no state changes), then times the schedules over a
// few MD steps and keeps the fastest.
#include "mgpu_internal.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <tuple>

using namespace mgpu;

struct mgpu_batch {
   int device = 0;
   hipStream_t origin = nullptr;
   std::vector<mgpu_engine *> eng;
   hipGraph_t graph = nullptr;
   hipGraphExec_t exec = nullptr;
   std::vector<int> cap_version, cap_precision;
   unsigned long cap_epoch = 0;
   std::vector<hipEvent_t> ev;          // [0] fork, [1] done, then one per engine
   int schedule = MGPU_SCHED_GRAPH;
   std::string err;
};

namespace {

template <class F> int bguard(mgpu_batch *b, F &&f) {
   if (!b) return MGPU_E_ARG;
   try {
      return f();
   } catch (const HipError &h) {
      b->err = std::string("HIP error ") + hipGetErrorString(h.err) + " in " + h.what;
      return MGPU_E_HIP;
   } catch (const StatusError &s) {
      b->err = s.msg;
      return s.code;
   }
}

void drop_graph(mgpu_batch *b) {
   if (b->exec) (void)hipGraphExecDestroy(b->exec);
   if (b->graph) (void)hipGraphDestroy(b->graph);
   b->exec = nullptr;
   b->graph = nullptr;
}

bool graph_stale(const mgpu_batch *b) {
   if (!b->exec || b->cap_epoch != g_alloc_epoch || b->cap_version.size() != b->eng.size()) return true;
   for (size_t r = 0; r < b->eng.size(); r++)
      if (b->cap_version[r] != b->eng[r]->geom_version || b->cap_precision[r] != b->eng[r]->precision) return true;
   return false;
}

void ensure_events(mgpu_batch *b) {
   while (b->ev.size() < b->eng.size() + 2) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->ev.push_back(e);
   }
}

// One step of every replica captured as parallel branches of one graph.
void capture_batch(mgpu_batch *b) {
   drop_graph(b);
   ensure_events(b);
   unsigned long epoch = g_alloc_epoch;
   HIPCHK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
   try {
      HIPCHK(hipEventRecord(b->ev[0], b->origin));
      for (size_t r = 0; r < b->eng.size(); r++) {
         mgpu_engine *e = b->eng[r];
         HIPCHK(hipStreamWaitEvent(e->stream, b->ev[0], 0));
         e->capturing = true;
         md_step_launch(e, false);
         e->capturing = false;
         HIPCHK(hipEventRecord(b->ev[2 + r], e->stream));
         HIPCHK(hipStreamWaitEvent(b->origin, b->ev[2 + r], 0));
      }
   } catch (...) {
      for (auto *e : b->eng) e->capturing = false;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b->origin, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
   }
   HIPCHK(hipStreamEndCapture(b->origin, &b->graph));
   if (epoch != g_alloc_epoch) throw StatusError{MGPU_E_STATE, "allocation during batch graph capture"};
   HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
   b->cap_epoch = g_alloc_epoch;
   b->cap_version.clear();
   b->cap_precision.clear();
   for (auto *e : b->eng) {
      b->cap_version.push_back(e->geom_version);
      b->cap_precision.push_back(e->precision);
   }
}

// Enqueue nsteps of every replica under schedule s (no host sync).
void enqueue_steps(mgpu_batch *b, int nsteps, bool init_h0, int s) {
   for (auto *e : b->eng) {
      md_load_cell(e);
      md_reset_record(e, nsteps);
      zero_async(e, e->d_iflags.p, sizeof(int) * 48);
   }
   if (nsteps <= 0) return;
   for (auto *e : b->eng) md_step_launch(e, init_h0);   // eager first step makes every buffer
   if (nsteps == 1) return;
   if (s == MGPU_SCHED_GRAPH) {
      if (graph_stale(b)) capture_batch(b);
      ensure_events(b);
      for (size_t r = 0; r < b->eng.size(); r++) {
         HIPCHK(hipEventRecord(b->ev[2 + r], b->eng[r]->stream));
         HIPCHK(hipStreamWaitEvent(b->origin, b->ev[2 + r], 0));
      }
      for (int k = 1; k < nsteps; k++) HIPCHK(hipGraphLaunch(b->exec, b->origin));
      HIPCHK(hipEventRecord(b->ev[1], b->origin));
      for (auto *e : b->eng) HIPCHK(hipStreamWaitEvent(e->stream, b->ev[1], 0));
   } else if (s == MGPU_SCHED_STREAM_GRAPHS) {
      for (auto *e : b->eng) ensure_step_graph(e);
      for (int k = 1; k < nsteps; k++)
         for (auto *e : b->eng) HIPCHK(hipGraphLaunch(e->graph_exec, e->stream));
   } else {
      for (int k = 1; k < nsteps; k++)
         for (auto *e : b->eng) md_step_launch(e, false);
   }
}

void finish(mgpu_batch *b, int nsteps, double *trace) {
   std::vector<std::vector<int>> fl(b->eng.size(), std::vector<int>(41));
   for (size_t r = 0; r < b->eng.size(); r++) {
      mgpu_engine *e = b->eng[r];
      if (trace && nsteps > 0)
         HIPCHK(hipMemcpyAsync(trace + (size_t)r * nsteps * md_trace_width(), e->d_trace.p,
                               sizeof(double) * md_trace_width() * nsteps, hipMemcpyDeviceToHost, e->stream));
      HIPCHK(hipMemcpyAsync(fl[r].data(), e->d_iflags.p, sizeof(int) * 41, hipMemcpyDeviceToHost, e->stream));
   }
   for (auto *e : b->eng) HIPCHK(hipStreamSynchronize(e->stream));
   for (size_t r = 0; r < b->eng.size(); r++) {
      try {
         handle_flags(b->eng[r], fl[r].data());
      } catch (const StatusError &s) {
         throw StatusError{s.code, "replica " + std::to_string(r) + ": " + s.msg};
      }
   }
}

// Process-wide memory of tuned shapes: (precision, nrows, lmax, nsites/4096)
// -> (S(k) workgroups, MFMA back-projection partitions, FP64 partitions).
std::map<std::tuple<int, int, int, int>, std::array<int, 4>> g_tuned;

}  // namespace

// Time the tunable launch shapes of e's force evaluation and keep the
// fastest: the S(k) workgroup count and the back-projection row partitions
// of the MFMA Ewald (mixed precision), or the row partitions of the FP64
// back-projection.  Results are identical up to the summation order of
// partial sums.
void mgpu::autotune_engine(mgpu_engine *e, int reps, int *sk_blocks, int *npart) {
   if (e->alpha <= 1e-7) return;
   auto key = std::make_tuple(e->precision, e->nrows, e->lmax, e->nsites / 4096);
   auto it = g_tuned.find(key);
   if (it != g_tuned.end()) {
      e->tune_sk_blocks = it->second[0];
      e->tune_force_parts = it->second[1];
      e->tune_npart = it->second[2];
      e->overlap = it->second[3];
   } else {
      hipEvent_t t0, t1;
      HIPCHK(hipEventCreate(&t0));
      HIPCHK(hipEventCreate(&t1));
      const bool prof = e->prof_on;
      e->prof_on = false;
      int nrep = reps;
      auto time_eval = [&]() {
         e->geom_version++;   // re-plan with the new shape
         eval_forces_device(e, true);
         HIPCHK(hipEventRecord(t0, e->stream));
         for (int k = 0; k < nrep; k++) eval_forces_device(e, true);
         HIPCHK(hipEventRecord(t1, e->stream));
         HIPCHK(hipEventSynchronize(t1));
         float ms = 0;
         HIPCHK(hipEventElapsedTime(&ms, t0, t1));
         return ms;
      };
      // coordinate descent over one shape at a time; workgroup targets in
      // whole rounds of the resident slots
      // two passes over the candidates, each scored by its faster pass, so
      // that a clock ramp or a neighbour's burst does not pick the shape
      auto tune = [&](int &field, const std::vector<int> &cand, int passes = 2) {
         std::vector<float> t(cand.size(), 1e30f);
         for (int pass = 0; pass < passes; pass++)
            for (size_t c = 0; c < cand.size(); c++) {
               field = cand[c];
               t[c] = std::min(t[c], time_eval());
            }
         field = cand[std::min_element(t.begin(), t.end()) - t.begin()];
      };
      if (e->precision == MGPU_MIXED) {
         const int slots = e->ncu;   // one S(k) workgroup per CU (LDS)
         tune(e->tune_sk_blocks, {2 * slots, 3 * slots, 4 * slots, 6 * slots, 8 * slots, 12 * slots});
         tune(e->tune_force_parts, {0, 1, 2, 4});
      } else {
         tune(e->tune_npart, {4, 8, 16, 32});
      }
      // real space beside the Ewald sum, or after it: a difference of a few
      // per cent, so longer trials and more passes than the shapes above
      nrep = std::max(reps, 10);
      tune(e->overlap, {1, 0}, 4);
      e->prof_on = prof;
      (void)hipEventDestroy(t0);
      (void)hipEventDestroy(t1);
      g_tuned[key] = {e->tune_sk_blocks, e->tune_force_parts, e->tune_npart, e->overlap};
   }
   e->geom_version++;
   if (sk_blocks) *sk_blocks = e->precision == MGPU_MIXED ? e->tune_sk_blocks : 0;
   if (npart) *npart = e->precision == MGPU_MIXED ? e->tune_force_parts : e->tune_npart;
}

extern "C" {

int mgpu_batch_create(mgpu_batch **bp, int device) {
   if (!bp) return MGPU_E_ARG;
   *bp = nullptr;
   mgpu_batch *b = new mgpu_batch();
   b->device = device;
   int rc = bguard(b, [&] {
      HIPCHK(hipSetDevice(device));
      HIPCHK(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking));
      return 0;
   });
   if (rc) {
      delete b;
      return rc;
   }
   *bp = b;
   return 0;
}

void mgpu_batch_destroy(mgpu_batch *b) {
   if (!b) return;
   (void)hipSetDevice(b->device);
   if (b->origin) (void)hipStreamSynchronize(b->origin);
   drop_graph(b);
   for (auto ev : b->ev) (void)hipEventDestroy(ev);
   if (b->origin) (void)hipStreamDestroy(b->origin);
   delete b;
}

const char *mgpu_batch_last_error(const mgpu_batch *b) { return b ? b->err.c_str() : "null batch"; }

int mgpu_batch_add(mgpu_batch *b, mgpu_engine *e) {
   return bguard(b, [&] {
      if (!e || !e->have_dyn || e->n_full == 0) {
         b->err = "engine needs mgpu_set_dynamics, mgpu_upload_state and mgpu_upload_momenta";
         return (int)MGPU_E_STATE;
      }
      if (e->device != b->device) {
         b->err = "engine on another device";
         return (int)MGPU_E_ARG;
      }
      if (e->dyn_const_pressure) {
         b->err = "a barostat engine re-plans its cell every step: run it with mgpu_md_steps";
         return (int)MGPU_E_ARG;
      }
      b->eng.push_back(e);
      drop_graph(b);
      return 0;
   });
}

int mgpu_batch_steps(mgpu_batch *b, int nsteps, int init_h0, int schedule, double *trace) {
   return bguard(b, [&] {
      if (nsteps < 0 || b->eng.empty()) {
         b->err = "no engines or nsteps < 0";
         return (int)MGPU_E_ARG;
      }
      HIPCHK(hipSetDevice(b->device));
      const int s = schedule < 0 ? b->schedule : schedule;
      enqueue_steps(b, nsteps, init_h0 != 0, s);
      finish(b, nsteps, trace);
      return 0;
   });
}

int mgpu_batch_autotune(mgpu_batch *b, int trial_steps, int *schedule, double ms_per_step[3]) {
   return bguard(b, [&] {
      if (b->eng.empty() || trial_steps < 2) {
         b->err = "no engines or fewer than 2 trial steps";
         return (int)MGPU_E_ARG;
      }
      HIPCHK(hipSetDevice(b->device));
      for (auto *e : b->eng) autotune_engine(e, 3, nullptr, nullptr);
      // each schedule with the engines' own real/reciprocal stream overlap
      // and with it off (replicas already run side by side)
      std::vector<int> own;
      for (auto *e : b->eng) own.push_back(e->overlap);
      auto set_overlap = [&](bool off) {
         for (size_t r = 0; r < b->eng.size(); r++) {
            const int want = off ? 0 : own[r];
            if (b->eng[r]->overlap != want) {
               b->eng[r]->overlap = want;
               b->eng[r]->geom_version++;   // captured graphs follow the stream layout
            }
         }
         drop_graph(b);
      };
      double t[2][3];
      for (int off = 0; off < 2; off++) {
         set_overlap(off);
         for (int s = 0; s < 3; s++) {
            enqueue_steps(b, 2, false, s);     // capture / warm up outside the timing
            finish(b, 2, nullptr);
            HIPCHK(hipDeviceSynchronize());
            auto t0 = std::chrono::steady_clock::now();
            enqueue_steps(b, trial_steps, false, s);
            finish(b, trial_steps, nullptr);
            auto t1 = std::chrono::steady_clock::now();
            t[off][s] = std::chrono::duration<double, std::milli>(t1 - t0).count() / trial_steps;
         }
      }
      int best = 0;
      for (int k = 1; k < 6; k++)
         if (t[k / 3][k % 3] < t[best / 3][best % 3]) best = k;
      set_overlap(best / 3);
      b->schedule = best % 3;
      if (schedule) *schedule = best % 3;
      if (ms_per_step)
         for (int s = 0; s < 3; s++) ms_per_step[s] = std::min(t[0][s], t[1][s]);
      return 0;
   });
}

int mgpu_autotune(mgpu_engine *e, int reps, int *sk_blocks, int *npart) {
   if (!e) return MGPU_E_ARG;
   try {
      if (!e->have_system || e->n_full == 0) {
         e->err = "mgpu_upload_state not called";
         return MGPU_E_STATE;
      }
      HIPCHK(hipSetDevice(e->device));
      autotune_engine(e, reps < 1 ? 3 : reps, sk_blocks, npart);
      return 0;
   } catch (const HipError &h) {
      e->err = std::string("HIP error ") + hipGetErrorString(h.err) + " in " + h.what;
      return MGPU_E_HIP;
   } catch (const StatusError &s) {
      e->err = s.msg;
      return s.code;
   }
}

}  // extern "C"
