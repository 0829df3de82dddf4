#include "server.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <stdexcept>

namespace mb {

double Server::now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

Server::Server(int nslots, double mem_budget_bytes, int dag_grid, int nanalysis, bool gang)
    : nslots_(std::max(1, nslots)), dag_grid_(dag_grid), budget_(mem_budget_bytes) {
  if (gang) {
    // one gang per size class (the factorization flops of a request): members of one class take
    // similar work per step, so lockstep costs little, and the classes overlap on the device
    nslots_ = std::min(nslots_, GANG_MAX);
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      ncu = 256;
    for (size_t c = 0; c <= GANG_CLASS_FLOPS.size(); ++c)
      gangs_.push_back(std::make_unique<Gang>(nslots_, ncu, false));
  }
  if (budget_ <= 0.0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) throw std::runtime_error("hipMemGetInfo failed");
    budget_ = 0.9 * (double)free_b;
  }
  if (dag_grid_ <= 0) {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      ncu = 256;
    dag_grid_ = std::max(16, ncu / nslots_);
  }
  if (nanalysis <= 0) nanalysis = std::min(8, std::max(2, (int)std::thread::hardware_concurrency() / 2));
  lookahead_ = (size_t)(2 * nslots_ * std::max<size_t>(1, gangs_.size()) + nanalysis);
  for (int k = 0; k < nanalysis; ++k) workers_.emplace_back(&Server::analysis_main, this);
  const int nthreads = gangs_.empty() ? nslots_ : nslots_ * (int)gangs_.size();
  for (int s = 0; s < nthreads; ++s) slots_.emplace_back(&Server::slot_main, this, s);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  cv_ready_.notify_all();
  cv_room_.notify_all();
  for (auto& t : workers_) t.join();
  for (auto& t : slots_) t.join();
  gangs_.clear();
}

int64_t Server::submit(const Problem* prob, const Settings& st) {
  auto job = std::make_shared<ServeJob>();
  job->prob = prob;
  job->set = st;
  job->set.dag_grid = dag_grid_;
  job->t_submit = now();
  int64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw std::runtime_error("server is shutting down");
    id = job->id = next_id_++;
    jobs_[id] = job;
    queue_.push_back(job);
  }
  cv_work_.notify_one();
  return id;
}

std::shared_ptr<ServeJob> Server::wait(int64_t id, double timeout_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) throw std::invalid_argument("unknown ticket");
  auto job = it->second;
  auto ready = [&] { return job->done; };
  if (timeout_s < 0) cv_done_.wait(lk, ready);
  else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return nullptr;
  return job;
}

void Server::release(int64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  jobs_.erase(id);
}

void Server::stats(double* out) {
  std::lock_guard<std::mutex> lk(mu_);
  out[0] = (double)(queue_.size() + ready_.size() + (size_t)analysing_);
  out[1] = (double)running_;
  out[2] = (double)completed_;
  out[3] = reserved_;
  out[4] = peak_;
}

// Host analysis ahead of admission, in arrival order, at most lookahead_ requests ahead of the
// slots (an analysed request holds its permuted problem and symbolic factor in host memory).
void Server::analysis_main() {
  for (;;) {
    std::shared_ptr<ServeJob> job;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] {
        return (stop_ && queue_.empty()) || (!queue_.empty() && ready_.size() + (size_t)analysing_ < lookahead_);
      });
      if (queue_.empty()) return;  // stop_ with nothing left
      job = queue_.front();
      queue_.pop_front();
      ++analysing_;
    }
    try {
      job->solver = std::make_unique<Solver>(*job->prob, job->set, Solver::Deferred{});
      job->reserve = job->solver->estimate_device_bytes();
    } catch (const std::exception& e) {
      job->solver.reset();
      job->error = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      --analysing_;
      ready_[job->id] = job;
    }
    cv_ready_.notify_all();
  }
}

void Server::finish(const std::shared_ptr<ServeJob>& job, double reserve) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    reserved_ -= reserve;
    if (reserved_ < 1.0) reserved_ = 0.0;
    --running_;
    ++completed_;
    job->t_end = now();
    job->done = true;
  }
  cv_room_.notify_all();
  cv_done_.notify_all();
}

void Server::slot_main(int slot) {
  (void)slot;
  for (;;) {
    std::shared_ptr<ServeJob> job;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_ready_.wait(lk, [&] {
        return !ready_.empty() || (stop_ && queue_.empty() && analysing_ == 0);
      });
      if (ready_.empty()) return;  // stop_ with nothing left
      job = ready_.begin()->second;
      ready_.erase(ready_.begin());
      ++running_;
    }
    cv_work_.notify_all();  // room for the analysis workers to run ahead again
    if (!job->solver) {     // the host analysis failed
      job->status = -1;
      finish(job, 0.0);
      continue;
    }
    const double reserve = job->reserve;
    {
      // Admission: the bound must fit what the running solves leave of the budget. A request
      // larger than the whole budget runs alone: it waits for an empty device, and no other
      // request is admitted while it waits (so it cannot be starved by smaller ones).
      std::unique_lock<std::mutex> lk(mu_);
      if (reserve > budget_) {
        ++oversized_waiting_;
        cv_room_.wait(lk, [&] { return reserved_ == 0.0; });
        --oversized_waiting_;
      } else {
        cv_room_.wait(lk, [&] { return oversized_waiting_ == 0 && reserved_ + reserve <= budget_; });
      }
      reserved_ += reserve;
      peak_ = std::max(peak_, reserved_);
    }
    try {
      Solver& S = *job->solver;
      job->t_admit = now();
      S.device_setup(nullptr, nullptr, nullptr);
      job->bytes = (double)S.device_bytes();
      if (gangs_.empty()) {
        job->status = S.solve();
      } else {
        const double fl = S.front_flops();
        size_t c = 0;
        while (c < GANG_CLASS_FLOPS.size() && fl >= GANG_CLASS_FLOPS[c]) ++c;
        job->status = S.solve_in(*gangs_[c]);
      }
      job->p.resize(S.n());
      job->d.resize(S.n());
      job->y.resize(S.m());
      S.result(job->p.data(), job->d.data(), job->y.data(), job->info);
    } catch (const std::exception& e) {
      job->error = e.what();
      job->status = -1;
    }
    job->solver.reset();  // buffers back to the cache before the reservation is released
    finish(job, reserve);
  }
}

}  // namespace mb
