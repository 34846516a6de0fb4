// wzfh_serve.cpp -- cell-search serving loop with continuous batching.
//
// The reference has no server: its lanes are fixed arm processes started by the lane driver
// (cluster/deploy/cluster_firsthit_probe.sh:100-107). This is a serving form of the same
// cell unit (wzfh_search_cells): requests name a
// problem (an (n, class) handle, set up once and kept device-resident) and a cell.
// One worker thread per problem repeatedly admits whatever requests of its problem have
// arrived -- up to max_batch cells and to the batch that fits its share of free device
// memory -- and runs them as one persistent stream launch plus the completer launches on the
// problem's own HIP stream. Requests that
// arrive while a batch is on the device join the next batch of their problem as soon as
// the current one returns, and batches of different problems run concurrently on their
// streams, so the device work of mixed-n traffic overlaps. Results are identical to
// running each cell alone: a cell's search depends only on the cell (the batched stream
// is exact per cell, see wzfh_stream.hip), which tests/test_gpu_serve.py checks.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "wzfh_internal.h"

using Clock = std::chrono::steady_clock;

struct ServeReq {
  long long ticket;
  int problem, cell;
  Clock::time_point submitted;
};

struct wzfh_server {
  std::vector<wzfh_problem *> problems;
  wzfh_serve_opts opts;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::deque<ServeReq>> pending;   // per problem
  std::vector<long long> mem_slots;             // per problem: cells per batch that fit in memory
  std::deque<wzfh_serve_result> done;
  std::vector<std::thread> workers;
  bool closing = false;
  long long next_ticket = 0;
  int error = WZFH_OK;
  std::string error_msg;       // message of the first failing batch (error strings are per thread)
  Clock::time_point t0 = Clock::now();
};

static void worker_loop(wzfh_server *S, int pidx) {
  wzfh_problem *P = S->problems[pidx];
  while (true) {
    std::vector<ServeReq> batch;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->closing || !S->pending[pidx].empty(); });
      if (S->pending[pidx].empty() && S->closing) return;
      const long long slots = std::max(1ll, std::min((long long)S->opts.max_batch, S->mem_slots[pidx]));
      while (!S->pending[pidx].empty() && (long long)batch.size() < slots) {
        batch.push_back(S->pending[pidx].front());
        S->pending[pidx].pop_front();
      }
    }
    std::vector<int> cells(batch.size());
    for (size_t i = 0; i < batch.size(); i++) cells[i] = batch[i].cell;
    std::vector<wzfh_search_result> res(batch.size());
    const auto started = Clock::now();
    int rc = wzfh_search_cells(P, cells.data(), (int)cells.size(), &S->opts.search, res.data());
    const auto finished = Clock::now();
    const char *msg = rc != WZFH_OK ? wzfh_last_error() : "";
    std::lock_guard<std::mutex> lk(S->mu);
    if (rc != WZFH_OK && rc != WZFH_EINTERNAL && S->error == WZFH_OK) { S->error = rc; S->error_msg = msg; }
    for (size_t i = 0; i < batch.size(); i++) {
      wzfh_serve_result r{};
      r.ticket = batch[i].ticket;
      r.problem = pidx;
      r.cell = batch[i].cell;
      r.status = rc;
      r.batch_size = (int)batch.size();
      r.result = res[i];
      r.submit_s = std::chrono::duration<double>(batch[i].submitted - S->t0).count();
      r.start_s = std::chrono::duration<double>(started - S->t0).count();
      r.finish_s = std::chrono::duration<double>(finished - S->t0).count();
      S->done.push_back(r);
    }
    S->cv_done.notify_all();
  }
}

extern "C" {

int wzfh_serve_open(wzfh_problem **problems, int nproblems, const wzfh_serve_opts *opts, wzfh_server **out) {
  if (!problems || nproblems <= 0 || !opts || !out) return WZFH_EINVAL;
  for (int i = 0; i < nproblems; i++)
    if (!problems[i] || !problems[i]->dev) { wzfh::set_error("serve: problem without a device"); return WZFH_EINVAL; }
  wzfh_server *S = new wzfh_server();
  S->problems.assign(problems, problems + nproblems);
  S->opts = *opts;
  S->pending.resize(nproblems);
  for (int i = 0; i < nproblems; i++) S->mem_slots.push_back(wzfh::serve_cell_slots(*problems[i], nproblems, opts->search));
  for (int i = 0; i < nproblems; i++) S->workers.emplace_back(worker_loop, S, i);
  *out = S;
  return WZFH_OK;
}

int wzfh_serve_submit(wzfh_server *S, int problem, int cell, long long *ticket) {
  if (!S || problem < 0 || problem >= (int)S->problems.size()) return WZFH_EINVAL;
  if (cell < 0 || cell >= wzfh_num_cells(S->problems[problem])) return WZFH_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->closing) return WZFH_EINVAL;
  ServeReq r{S->next_ticket++, problem, cell, Clock::now()};
  if (ticket) *ticket = r.ticket;
  S->pending[problem].push_back(r);
  S->cv_work.notify_all();
  return WZFH_OK;
}

int wzfh_serve_poll(wzfh_server *S, wzfh_serve_result *out, int max, int min_results, int timeout_ms) {
  if (!S || !out || max <= 0) return WZFH_EINVAL;
  std::unique_lock<std::mutex> lk(S->mu);
  const int want = std::max(0, std::min(min_results, max));
  S->cv_done.wait_for(lk, std::chrono::milliseconds(std::max(0, timeout_ms)),
                      [&] { return (int)S->done.size() >= want || S->error != WZFH_OK; });
  int k = 0;
  while (k < max && !S->done.empty()) { out[k++] = S->done.front(); S->done.pop_front(); }
  if (S->error != WZFH_OK && k == 0) { wzfh::set_error("serve: " + S->error_msg); return S->error; }
  return k;
}

int wzfh_serve_close(wzfh_server *S) {
  if (!S) return WZFH_EINVAL;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->closing = true;
    S->cv_work.notify_all();
  }
  for (auto &t : S->workers) t.join();
  const int rc = S->error;
  delete S;
  return rc;
}

}  // extern "C"
