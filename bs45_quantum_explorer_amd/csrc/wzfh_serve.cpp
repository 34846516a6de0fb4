// wzfh_serve.cpp -- cell-search serving loop with continuous, in-flight batching.
//
// The reference has no server: its lanes are fixed arm processes started by the lane driver
// (cluster/deploy/cluster_firsthit_probe.sh:100-107). This is a serving form of the same cell
// unit (wzfh_search_cells): requests name a problem (an (n, class) handle, set up once and kept
// device-resident) and a cell.
//
// Problems whose stream kernels use the same tables (every class of one n under the same
// flags) form a group. Each group has a stream stage: a thread with a device state of its own
// that admits every request of the group that has arrived -- up to max_batch cells and to the
// cells that fit the group's share of the device's free memory, re-read at every admission --
// runs them as ONE persistent stream launch (cells of different classes side by side), picks
// each cell's flattest top_k candidates, hands them to the problems' completion stages and
// immediately admits the next requests. Each problem has a completion stage: a thread running
// streaming completer runs under the cells' profile rows on the problem's own HIP stream; the
// stream batches queued for the problem join the running run between its rounds, and each
// batch's results return as soon as its own candidates are decided. So a request that arrives
// while its group is streaming waits for that one launch only; the completions of earlier
// batches run meanwhile on the other streams, and groups of different n run concurrently. A result equals the same cell searched alone: a cell's stream is exact per
// cell whatever else shares its launch, and its completions depend only on its candidates.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "wzfh_internal.h"

using Clock = std::chrono::steady_clock;

namespace {

struct ServeReq {
  long long ticket;
  int problem, cell;
  Clock::time_point submitted;
};

// One stream launch's share for one problem: its requests, their results (stream part set),
// and their top-K candidates (owner = index into reqs).
struct CompBatch {
  std::vector<ServeReq> reqs;
  std::vector<wzfh_search_result> res;
  wzfh::TopK topk;
  Clock::time_point started;
  int launch_cells = 0;
};

struct Group {
  std::vector<int> members;
  wzfh_problem *sh = nullptr;      // stream handle (device state of the stream stage)
  std::deque<ServeReq> pending;
  bool stream_done = false;        // the stream stage has exited (closing)
};

}  // namespace

struct wzfh_server {
  std::vector<wzfh_problem *> problems;
  int workers = 1;                           // completion runs per problem at once
  std::vector<std::vector<wzfh::DeviceState *>> comp_states;   // per problem: states of workers 1..

  wzfh_serve_opts opts;
  std::mutex mu;
  std::condition_variable cv_work, cv_comp, cv_done;
  std::vector<Group> groups;
  std::vector<int> group_of;                 // per problem
  std::vector<std::deque<CompBatch>> comp_q; // per problem
  std::deque<wzfh_serve_result> done;
  std::vector<std::thread> stream_threads, comp_threads;
  bool closing = false;
  long long next_ticket = 0;
  int error = WZFH_OK;
  std::string error_msg;       // message of the first failure (error strings are per thread)
  Clock::time_point t0 = Clock::now();

  double since(Clock::time_point t) const { return std::chrono::duration<double>(t - t0).count(); }

  // Results of requests that failed before their completions (caller holds mu).
  void fail(const std::vector<ServeReq> &reqs, int rc, const char *msg, Clock::time_point started, int launch) {
    if (error == WZFH_OK) { error = rc; error_msg = msg; }
    const Clock::time_point now = Clock::now();
    for (const ServeReq &q : reqs) {
      wzfh_serve_result r{};
      r.ticket = q.ticket; r.problem = q.problem; r.cell = q.cell; r.status = rc; r.batch_size = launch;
      r.submit_s = since(q.submitted); r.start_s = since(started); r.finish_s = since(now);
      done.push_back(r);
    }
    cv_done.notify_all();
  }
};

static void stream_loop(wzfh_server *S, int g) {
  Group &G = S->groups[g];
  const int K = S->opts.search.top_k;
  while (true) {
    const long long mem = wzfh::serve_cell_slots(*G.sh, (int)S->groups.size(), S->opts.search);
    std::vector<ServeReq> batch;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->closing || !G.pending.empty(); });
      if (G.pending.empty()) {
        G.stream_done = true;
        S->cv_comp.notify_all();
        return;
      }
      const long long slots = std::max(1ll, std::min((long long)S->opts.max_batch, mem));
      while (!G.pending.empty() && (long long)batch.size() < slots) {
        batch.push_back(G.pending.front());
        G.pending.pop_front();
      }
    }
    const int nc = (int)batch.size();
    std::vector<int> cells(nc);
    std::vector<const wzfh_problem *> owner(nc);
    for (int i = 0; i < nc; i++) { cells[i] = batch[i].cell; owner[i] = S->problems[batch[i].problem]; }
    const auto started = Clock::now();
    std::vector<wzfh_cell_stats> st(nc);
    wzfh::StreamOut so;
    // several groups (several n) share the device: a launch of a few cells takes at most 64 blocks
    // per cell, so the groups' launches run side by side (measured on the mixed-n serving sweep:
    // p50 / p99 at 0.2 load 137 / 774 ms against 295 / 4,022 ms with the whole grid each)
    const int cell_blocks = S->groups.size() > 1 ? 64 : 0;
    const int rc = wzfh::run_stream(*G.sh, cells.data(), nc, S->opts.search.stream, st.data(), so, owner.data(),
                                    cell_blocks);
    if (rc != WZFH_OK) {
      const std::string msg = wzfh_last_error();
      std::lock_guard<std::mutex> lk(S->mu);
      S->fail(batch, rc, msg.c_str(), started, nc);
      continue;
    }
    wzfh::TopK t;
    wzfh::select_topk(G.sh->n, so, nc, K, t);
    // split the launch by problem, keeping each cell's candidates in rank order
    std::map<int, CompBatch> parts;
    std::vector<int> local(nc);
    for (int i = 0; i < nc; i++) {
      CompBatch &b = parts[batch[i].problem];
      local[i] = (int)b.reqs.size();
      b.reqs.push_back(batch[i]);
      wzfh_search_result r{};
      r.stream = st[i];
      r.hit_rank = -1; r.hit_idx = -1;
      b.res.push_back(r);
      b.started = started;
      b.launch_cells = nc;
    }
    for (size_t k = 0; k < t.owner.size(); k++) {
      CompBatch &b = parts[batch[t.owner[k]].problem];
      b.topk.cd.push_back(t.cd[2 * k]); b.topk.cd.push_back(t.cd[2 * k + 1]);
      b.topk.owner.push_back(local[t.owner[k]]);
      b.topk.rank.push_back(t.rank[k]); b.topk.sidx.push_back(t.sidx[k]);
    }
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto &kv : parts) S->comp_q[kv.first].push_back(std::move(kv.second));
    S->cv_comp.notify_all();
  }
}

// Completion worker w of problem pidx: worker 0 runs on the problem's own device state, the
// others on states of their own. A worker runs one streaming completer run at a time: the
// batches queued for its problem join the run between its rounds (continuous batching inside the
// completer), and each batch's results go out as soon as its own candidates are decided. A batch
// that arrives while a run waits for its heaviest candidates therefore starts at the next round
// instead of after the whole run.
static void complete_loop(wzfh_server *S, int pidx, int w) {
  wzfh_problem *P = S->problems[pidx];
  wzfh::DeviceState *dev = w > 0 ? S->comp_states[pidx][w - 1] : nullptr;
  const int g = S->group_of[pidx];
  constexpr long long kRunCands = 1 << 20;   // candidates one run takes before it lets a new run start
  while (true) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_comp.wait(lk, [&] { return !S->comp_q[pidx].empty() || S->groups[g].stream_done; });
      if (S->comp_q[pidx].empty()) return;
    }
    struct Live {
      CompBatch b;
      int base = 0, left = 0;
      bool out = false;
    };
    std::deque<Live> live;        // batches of this run, in joining order (stable references)
    std::vector<int> live_of;     // candidate -> batch
    std::vector<wzfh_completion> comp;
    long long taken = 0;
    auto finish = [&](Live &L) {   // every candidate of the batch is decided: its results go out
      wzfh::aggregate_completions(L.b.topk, comp.data() + L.base, L.b.res.data());
      const auto finished = Clock::now();
      std::lock_guard<std::mutex> lk(S->mu);
      for (size_t i = 0; i < L.b.reqs.size(); i++) {
        wzfh_serve_result r{};
        r.ticket = L.b.reqs[i].ticket;
        r.problem = pidx;
        r.cell = L.b.reqs[i].cell;
        r.status = L.b.res[i].internal ? WZFH_EINTERNAL : WZFH_OK;
        r.batch_size = L.b.launch_cells;
        r.result = L.b.res[i];
        r.submit_s = S->since(L.b.reqs[i].submitted);
        r.start_s = S->since(L.b.started);
        r.finish_s = S->since(finished);
        S->done.push_back(r);
      }
      L.out = true;
      S->cv_done.notify_all();
    };
    wzfh::CompleteFeed feed;
    feed.more = [&](std::vector<uint64_t> &cd, std::vector<int> &cell_of) -> int {
      if (taken >= kRunCands) return 0;
      std::vector<CompBatch> got;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        while (!S->comp_q[pidx].empty()) {
          got.push_back(std::move(S->comp_q[pidx].front()));
          S->comp_q[pidx].pop_front();
        }
      }
      int added = 0;
      for (CompBatch &b : got) {
        live.emplace_back();
        Live &L = live.back();
        L.b = std::move(b);
        L.base = (int)cell_of.size();
        L.left = (int)L.b.topk.owner.size();
        const int li = (int)live.size() - 1;
        for (size_t k = 0; k < L.b.topk.owner.size(); k++) {
          cd.push_back(L.b.topk.cd[2 * k]);
          cd.push_back(L.b.topk.cd[2 * k + 1]);
          cell_of.push_back(L.b.reqs[L.b.topk.owner[k]].cell);
          live_of.push_back(li);
          comp.emplace_back();
        }
        added += L.left;
        taken += L.left;
        if (L.left == 0) finish(L);   // no candidate in any of its cells
      }
      return added;
    };
    feed.done = [&](int i, const wzfh_completion &c) {
      comp[i] = c;
      Live &L = live[live_of[i]];
      if (--L.left == 0) finish(L);
    };
    if (dev) wzfh::complete_state_sync_tuning(*P, dev);
    const int rc = wzfh::run_complete_feed(*P, S->opts.search.budget, dev, feed);
    if (rc != WZFH_OK) {
      const std::string msg = wzfh_last_error();
      std::lock_guard<std::mutex> lk(S->mu);
      for (Live &L : live)
        if (!L.out) S->fail(L.b.reqs, rc, msg.c_str(), L.b.started, L.b.launch_cells);
    }
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
  S->comp_q.resize(nproblems);
  S->group_of.assign(nproblems, -1);
  std::map<std::string, int> by_key;
  for (int i = 0; i < nproblems; i++) {
    const std::string k = wzfh::stream_key(*problems[i]);
    auto it = by_key.find(k);
    if (it == by_key.end()) {
      it = by_key.emplace(k, (int)S->groups.size()).first;
      S->groups.emplace_back();
    }
    S->group_of[i] = it->second;
    S->groups[it->second].members.push_back(i);
  }
  // one worker per problem by default: on served n = 44 steps (12 problems) two or three workers
  // were 10-20 % slower than one (more HIP streams than hardware queues, and more completer waves
  // beside the stream launches)
  S->workers = opts->complete_workers > 0 ? std::min(opts->complete_workers, 8) : 1;
  S->comp_states.resize(nproblems);
  auto undo = [&](int rc) {
    for (Group &H : S->groups) if (H.sh) wzfh_destroy(H.sh);
    for (auto &v : S->comp_states) for (wzfh::DeviceState *d : v) wzfh::complete_state_free(d);
    delete S;
    return rc;
  };
  for (Group &G : S->groups) {
    const int rc = wzfh::stream_handle(*problems[G.members[0]], &G.sh);
    if (rc != WZFH_OK) return undo(rc);
  }
  for (int i = 0; i < nproblems; i++)
    for (int w = 1; w < S->workers; w++) {
      wzfh::DeviceState *d = nullptr;
      const int rc = wzfh::complete_state(*problems[i], &d);
      if (rc != WZFH_OK) return undo(rc);
      S->comp_states[i].push_back(d);
    }
  for (int g = 0; g < (int)S->groups.size(); g++) S->stream_threads.emplace_back(stream_loop, S, g);
  for (int i = 0; i < nproblems; i++)
    for (int w = 0; w < S->workers; w++) S->comp_threads.emplace_back(complete_loop, S, i, w);
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
  S->groups[S->group_of[problem]].pending.push_back(r);
  S->cv_work.notify_all();
  return WZFH_OK;
}

int wzfh_serve_submit_batch(wzfh_server *S, int count, const int *problems, const int *cells, long long *tickets) {
  if (!S || count < 0 || (count > 0 && (!problems || !cells))) return WZFH_EINVAL;
  for (int i = 0; i < count; i++) {
    if (problems[i] < 0 || problems[i] >= (int)S->problems.size()) return WZFH_EINVAL;
    if (cells[i] < 0 || cells[i] >= wzfh_num_cells(S->problems[problems[i]])) return WZFH_EINVAL;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->closing) return WZFH_EINVAL;
  const Clock::time_point now = Clock::now();
  for (int i = 0; i < count; i++) {
    ServeReq r{S->next_ticket++, problems[i], cells[i], now};
    if (tickets) tickets[i] = r.ticket;
    S->groups[S->group_of[problems[i]]].pending.push_back(r);
  }
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

int wzfh_serve_kernel_totals(wzfh_server *S, wzfh_kernel_sums *out, int reset) {
  if (!S) return WZFH_EINVAL;
  wzfh_kernel_sums sum{};
  // stream launches run on the groups' own handles, completer launches on the problems'
  std::vector<wzfh_problem *> hs;
  for (Group &G : S->groups) hs.push_back(G.sh);
  hs.insert(hs.end(), S->problems.begin(), S->problems.end());
  auto add = [&](const wzfh_kernel_sums &t) {
    for (int k = 0; k < 2; k++) {
      sum.ms[k] += t.ms[k];
      sum.units[k] += t.units[k];
      sum.launches[k] += t.launches[k];
    }
  };
  for (wzfh_problem *h : hs) {
    wzfh_kernel_sums t{};
    if (int rc = wzfh_kernel_totals(h, &t, reset)) return rc;
    add(t);
  }
  for (auto &v : S->comp_states)   // the extra completion workers' states
    for (wzfh::DeviceState *d : v) {
      wzfh_kernel_sums t{};
      wzfh::state_kernel_totals(d, &t, reset);
      add(t);
    }
  if (out) *out = sum;
  return WZFH_OK;
}

int wzfh_serve_close(wzfh_server *S) {
  if (!S) return WZFH_EINVAL;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->closing = true;
    S->cv_work.notify_all();
    S->cv_comp.notify_all();
  }
  for (auto &t : S->stream_threads) t.join();
  for (auto &t : S->comp_threads) t.join();
  for (Group &G : S->groups) if (G.sh) wzfh_destroy(G.sh);
  for (auto &v : S->comp_states) for (wzfh::DeviceState *d : v) wzfh::complete_state_free(d);
  const int rc = S->error;
  delete S;
  return rc;
}

}  // extern "C"
