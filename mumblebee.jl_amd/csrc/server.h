// Continuous-batching request loop for conic solves on one GPU.
//
// Requests pass two pools of host threads. Analysis workers take them in arrival order and run
// the host analysis (problem permutation, symbolic factor, equilibration) and the device-byte
// bound; slots (one HIP stream each) take analysed requests in ticket order, wait until the bound
// fits the memory budget beside the running solves, build the device state, solve, copy the
// result out and hand the memory back. A slot therefore holds only device work, and the next
// requests are analysed while the GPU runs the current ones. Requests that arrive while others
// are on the GPU are admitted as slots and memory free up, and their kernels overlap on the
// device with the running solves. The persistent large-front factorization takes a share of the
// CUs per slot.
#pragma once
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gang.h"
#include "host.h"
#include "solver.h"

namespace mb {

struct ServeJob {
  int64_t id = 0;
  const Problem* prob = nullptr;  // owned by the caller until the result is collected
  Settings set;
  bool done = false;
  int status = -1;
  std::string error;
  std::vector<double> p, d, y;
  double info[8] = {0};
  // seconds on one steady clock: submitted, admitted (device memory reserved), finished
  double t_submit = 0, t_admit = 0, t_end = 0;
  double bytes = 0;    // device bytes the solve allocated
  double reserve = 0;  // the admission bound (estimate_device_bytes) held while it ran
  std::unique_ptr<Solver> solver;  // host-analysed, between analysis and admission
};

// gang mode: requests whose factorization takes fewer flops than GANG_CLASS_FLOPS[0] go to the
// first gang, then the next class, and so on
inline const std::vector<double> GANG_CLASS_FLOPS = {1e8, 2e9};

class Server {
 public:
  // nanalysis ≤ 0: one analysis worker per two hardware threads (at least two); gang: the slots
  // are members of gangs, one gang of nslots per size class (their solves advance in lockstep and
  // share every launch), so nslots × classes requests run at a time
  Server(int nslots, double mem_budget_bytes, int dag_grid, int nanalysis = 0, bool gang = false);
  ~Server();  // finishes the queued and running requests, then joins the threads
  int64_t submit(const Problem* prob, const Settings& st);
  // waits up to timeout_s (< 0: no limit); returns the job once done, nullptr on timeout
  std::shared_ptr<ServeJob> wait(int64_t id, double timeout_s);
  void release(int64_t id);  // forget a collected job
  // queued (submitted, not yet admitted), running, completed, device bytes reserved, peak
  // bytes reserved
  void stats(double* out);
  static double now();

 private:
  void analysis_main();
  void slot_main(int slot);
  void finish(const std::shared_ptr<ServeJob>& job, double reserve);
  int nslots_, dag_grid_;
  double budget_, reserved_ = 0, peak_ = 0;
  int64_t next_id_ = 0;
  int running_ = 0;
  int oversized_waiting_ = 0;  // admitted only on an empty device; others wait behind it
  int64_t completed_ = 0;
  size_t lookahead_ = 0;       // analysed-but-not-admitted requests kept at most
  bool stop_ = false;
  int analysing_ = 0;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_ready_, cv_done_, cv_room_;
  std::deque<std::shared_ptr<ServeJob>> queue_;           // submitted, not analysed
  std::map<int64_t, std::shared_ptr<ServeJob>> ready_;    // analysed, by ticket
  std::unordered_map<int64_t, std::shared_ptr<ServeJob>> jobs_;
  std::vector<std::thread> workers_, slots_;
  std::vector<std::unique_ptr<Gang>> gangs_;  // gang mode: one per size class
};

}  // namespace mb
