// Serving loop over the batched solver (gfem_batch_*): requests on resident
// meshes arrive over time, are admitted while request slots and the HBM
// budget allow, and iterate together with the requests already on the GPU
// (continuous batching).  Each request sets term coefficients of its mesh's
// formulation (wavenumber, source position, ...), is assembled with them
// into its own buffers and solved to its tolerance.  Latency is measured
// from arrival to the solution being available on the host.
#pragma once

#include <complex>
#include <cstddef>
#include <string>
#include <utility>
#include <vector>

#include "gfem.h"
#include "gmshfem.hpp"

namespace gmshfem {

namespace serve {

struct Request {
  int mesh = 0;          // index returned by Server::addMesh
  double arrival = 0.;   // seconds after the start of run()
  std::vector<std::pair<unsigned int, function::ScalarFunction>> params;  // (term tag, coefficient)
  double tol = 1e-8;
  int maxIter = 20000;
};

struct Completed {
  size_t request = 0;    // index into the schedule
  double arrival = 0., admitted = 0., done = 0.;  // seconds after the start of run()
  int iterations = 0, cycles = 0;
  double relResidual = 0.;
  bool converged = false;
  std::vector<std::complex<double>> x;  // when kept
  double latency() const { return done - arrival; }
};

struct RunStats {
  double seconds = 0.;   // first arrival to last completion
  int steps = 0;         // batch steps (restart cycles)
  int maxLive = 0;       // most requests iterating together
  size_t peakBytes = 0;  // most HBM held by admitted requests
  size_t maxWaiting = 0; // most requests arrived and not yet admitted
};

class Server {
 public:
  Server(int restart = 50, size_t hbmBytes = size_t(16) << 30, int maxRequests = 256);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  // A pre()-processed formulation; its system is uploaded here.  The
  // formulation must outlive the server.
  int addMesh(problem::Formulation<std::complex<double>> &f);
  size_t requestBytes(int mesh) const;

  // Serve the schedule (sorted by arrival) in real time; returns one record
  // per request, in completion order.
  std::vector<Completed> run(const std::vector<Request> &schedule, bool keepSolutions = false);
  const RunStats &stats() const { return _stats; }
  // Print a line to stderr every `seconds` of a run (0: never): time,
  // requests arrived / admitted / completed, iterating and waiting.
  void setProgress(double seconds) { _progress = seconds; }

 private:
  gfem_batch *_batch = nullptr;
  int _restart;
  size_t _budget;
  std::vector<problem::Formulation<std::complex<double>> *> _meshes;
  RunStats _stats;
  double _progress = 0.;
};

}  // namespace serve

}  // namespace gmshfem
