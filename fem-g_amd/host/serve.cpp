#include "serve.hpp"

#include <cstdio>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <thread>

namespace gmshfem {

namespace serve {

namespace {

void check(gfem_status st, const char *what)
{
  if(st != GFEM_OK) throw common::Exception(std::string(what) + ": " + gfem_ctx_last_error(common::deviceContext()));
}

}  // namespace

Server::Server(int restart, size_t hbmBytes, int maxRequests) : _restart(restart), _budget(hbmBytes)
{
  check(gfem_batch_create(common::deviceContext(), restart, hbmBytes, maxRequests, &_batch), "gfem_batch_create");
}

Server::~Server() { gfem_batch_destroy(_batch); }

int Server::addMesh(problem::Formulation<std::complex<double>> &f)
{
  if(!f.system()) {
    problem::AssembleOptions o;
    o.download = false;
    f.assemble(o);  // uploads the mesh, pattern and maps
    if(!f.lastError().empty()) throw common::Exception("Server::addMesh: " + f.lastError());
  }
  _meshes.push_back(&f);
  return (int)_meshes.size() - 1;
}

size_t Server::requestBytes(int mesh) const
{
  size_t b = 0;
  check(gfem_batch_request_bytes(_meshes.at(mesh)->system(), _restart, &b), "gfem_batch_request_bytes");
  return b;
}

std::vector<Completed> Server::run(const std::vector<Request> &schedule, bool keepSolutions)
{
  using clock = std::chrono::steady_clock;
  _stats = RunStats();
  std::vector<Completed> out;
  out.reserve(schedule.size());
  for(const Request &r : schedule)
    if(r.mesh < 0 || r.mesh >= (int)_meshes.size()) throw common::Exception("Server::run: unknown mesh");
  std::vector<size_t> bytes(schedule.size());
  for(size_t i = 0; i < schedule.size(); ++i) bytes[i] = requestBytes(schedule[i].mesh);
  const auto t0 = clock::now();
  auto now = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
  size_t next = 0;                     // next arrival
  std::deque<size_t> waiting;          // arrived, not admitted
  std::map<int64_t, size_t> running;   // batch id -> request
  std::map<int64_t, double> admitted;
  size_t held = 0, nadmitted = 0;
  double report = _progress;
  while(out.size() < schedule.size()) {
    const double t = now();
    while(next < schedule.size() && schedule[next].arrival <= t) waiting.push_back(next++);
    _stats.maxWaiting = std::max(_stats.maxWaiting, waiting.size());
    if(_progress > 0. && t >= report) {
      std::fprintf(stderr, "serve: %.0f s  arrived %zu  admitted %zu  completed %zu  iterating %zu  waiting %zu\n", t,
                   next, nadmitted, out.size(), running.size(), waiting.size());
      std::fflush(stderr);
      report = t + _progress;
    }
    // admit in arrival order while a slot and memory are free
    while(!waiting.empty()) {
      const size_t ri = waiting.front();
      const Request &r = schedule[ri];
      problem::Formulation<std::complex<double>> &f = *_meshes[r.mesh];
      for(const auto &p : r.params) f.setTermCoefficient(p.first, p.second);
      int64_t id = -1;
      const gfem_status st = gfem_batch_submit(_batch, f.system(), r.tol, r.maxIter, nullptr, &id);
      if(st == GFEM_ERR_BUSY) {
        if(running.empty())
          throw common::Exception("Server::run: a request needs more than the whole HBM budget");
        break;
      }
      check(st, "gfem_batch_submit");
      waiting.pop_front();
      ++nadmitted;
      running[id] = ri;
      admitted[id] = now();
      held += bytes[ri];
      _stats.peakBytes = std::max(_stats.peakBytes, held);
    }
    int live = 0;
    if(!running.empty()) {
      check(gfem_batch_step(_batch, nullptr, &live), "gfem_batch_step");
      ++_stats.steps;
    }
    // collect what ended (also requests that converged at submission)
    int64_t ids[256];
    int cnt = 0;
    do {
      check(gfem_batch_done(_batch, 256, ids, &cnt), "gfem_batch_done");
      for(int k = 0; k < cnt; ++k) {
        const auto it = running.find(ids[k]);
        if(it == running.end()) continue;
        const size_t ri = it->second;
        Completed c;
        c.request = ri;
        c.arrival = schedule[ri].arrival;
        c.admitted = admitted[ids[k]];
        gfem_solve_info info{};
        const int64_t n = (int64_t)_meshes[schedule[ri].mesh]->getNumberOfUnknownDof();
        if(keepSolutions) c.x.resize((size_t)n);
        check(gfem_batch_result(_batch, ids[k], keepSolutions ? c.x.data() : nullptr,
                                keepSolutions ? sizeof(std::complex<double>) * (size_t)n : 0, &info),
              "gfem_batch_result");
        c.done = now();
        c.iterations = info.iterations;
        c.cycles = info.cycles;
        c.relResidual = info.rel_residual;
        c.converged = info.converged != 0;
        held -= bytes[ri];
        running.erase(it);
        admitted.erase(ids[k]);
        out.push_back(std::move(c));
      }
    } while(cnt == 256);
    _stats.maxLive = std::max(_stats.maxLive, live);
    // idle: wait for the next arrival
    if(running.empty() && waiting.empty() && next < schedule.size()) {
      const double dt = schedule[next].arrival - now();
      if(dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(dt));
    }
  }
  if(!out.empty()) {
    double first = schedule.front().arrival, last = 0.;
    for(const auto &c : out) last = std::max(last, c.done);
    _stats.seconds = last - first;
  }
  return out;
}

}  // namespace serve

}  // namespace gmshfem
