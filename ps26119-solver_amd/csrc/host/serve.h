// serve.h — continuous batching of LP solves on one GPU (no counterpart in the reference, whose
// solve() runs one LP per call). Requests of any size arrive at any time; a host thread pool
// prepares each (validation, scaling, ‖Ã‖₂), the serving thread admits prepared requests while
// the HBM byte budget has room, and every round advances all live instances by one chunk of
// r²HPDHG iterations in a single packed block-diagonal launch sequence (psh_server_*), checks
// them, restarts or retires each one on its own schedule, and hands finished ones back to the
// pool for the finish (unscale, KKT on the original model, certified bound, in-process check).
// Per instance the iteration is the reference's r²HPDHG (src/pdhg/r2hpdhg.cpp:11-162) in fp64
// with the restart and PID rules of the single-LP engine; an instance's result does not depend on
// the other instances in the mix.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "contract.h"

struct psh_server;

namespace ps26119::amd {

struct ServerOptions {
  int device = 0;
  std::size_t budget_bytes = std::size_t(8) << 30;  // HBM byte budget of the arena
  int max_instances = 256;                          // live instances at once
  int check_every = 64;                             // K
  int setup_threads = 4;                            // host workers for setup and finish
  bool graphs = true;                               // replay captured chunk graphs (false: eager launches)
  bool timing = false;                              // HIP-event timing of every chunk
};

struct ServedResult {
  Solution solution;
  double submit_time = 0;    // seconds since the server started
  double admit_time = 0;     // request admitted to the device
  double done_time = 0;      // result ready
  double latency() const { return done_time - submit_time; }
};

class Server {
 public:
  explicit Server(const ServerOptions& o);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  bool ok() const { return error_.empty(); }
  const std::string& error() const { return error_; }

  // Queues a solve; returns its id. Options honoured: tolerance, iteration_limit, time_limit
  // (from submission), verbosity, engine_params (restart / PID knobs). Requests the server cannot
  // serve (integer columns, invalid data, too large for the budget) finish at once with a status.
  std::int64_t submit(Model model, const Options& options);
  // Blocks until the request has finished (timeout_s < 0: no timeout); false on timeout.
  bool wait(std::int64_t id, ServedResult& out, double timeout_s = -1.0);
  // Requests queued, being prepared or live on the device.
  int in_flight() const;
  // timed serving (ServerOptions::timing): iterations, their device milliseconds and bytes
  bool kernel_time(long long& iterations, double& total_ms, double& bytes);
  double now() const;

 private:
  struct Request;
  void serve_loop();
  void setup_worker();
  void prepare(Request& r);
  void finish(Request& r);
  void complete(const std::shared_ptr<Request>& r);

  ServerOptions opt_;
  std::string error_;
  psh_server* dev_ = nullptr;
  std::size_t capacity_ = 0;  // arena bytes
  double t0_ = 0;

  mutable std::mutex mu_;
  std::mutex dev_mu_;  // held by the serving thread for each round of device calls
  std::condition_variable cv_work_, cv_done_, cv_serve_;
  std::deque<std::shared_ptr<Request>> to_prepare_, to_finish_, ready_;
  std::map<std::int64_t, std::shared_ptr<Request>> all_;
  std::int64_t next_id_ = 0;
  int live_ = 0;
  bool stop_ = false;
  std::vector<std::thread> workers_;
  std::thread loop_;
};

}  // namespace ps26119::amd
