// Continuous-batching request loop for conic solves on one GPU.
//
// A fixed pool of slots, each a host thread with its own HIP stream, takes requests in arrival
// order. A slot runs the host analysis of its request (symbolic factor, equilibration). It then
// waits until the request's device-memory bound fits the budget, builds the device state, solves,
// copies the result out and frees the memory. Requests that arrive while others are on the GPU are
// admitted as slots and memory free up, and their kernels overlap on the device with the running
// solves. The persistent large-front factorization takes a share of the CUs per slot.
#pragma once
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

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
  double bytes = 0;  // device bytes the solve allocated
};

class Server {
 public:
  Server(int nslots, double mem_budget_bytes, int dag_grid);
  ~Server();  // finishes the queued and running requests, then joins the slots
  int64_t submit(const Problem* prob, const Settings& st);
  // waits up to timeout_s (< 0: no limit); returns the job once done, nullptr on timeout
  std::shared_ptr<ServeJob> wait(int64_t id, double timeout_s);
  void release(int64_t id);  // forget a collected job
  // queued, running, completed, device bytes reserved, peak bytes reserved
  void stats(double* out);
  static double now();

 private:
  void slot_main(int slot);
  int nslots_, dag_grid_;
  double budget_, reserved_ = 0, peak_ = 0;
  int64_t next_id_ = 0;
  int running_ = 0;
  int64_t completed_ = 0;
  bool stop_ = false;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_mem_, cv_done_;
  std::deque<std::shared_ptr<ServeJob>> queue_;
  std::unordered_map<int64_t, std::shared_ptr<ServeJob>> jobs_;
  std::vector<std::thread> slots_;
};

}  // namespace mb
