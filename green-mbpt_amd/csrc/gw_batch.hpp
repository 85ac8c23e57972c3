// Several independent GW systems (uneven nk / nao / naux) evaluated together: one stream, one
// shared transient workspace sized for the largest system, and the whole multi-system
// evaluation captured once into a hipGraph that every later call replays with a single
// hipGraphLaunch.  Included by gw_hip_kernel.hip (one translation unit).
#pragma once

#include <memory>
#include <vector>

namespace gwhip {

class gw_hip_batch {
public:
  gw_hip_batch(int n, const Dims* dims, int device) : device_(device) {
    if (n <= 0 || !dims) throw status_error(GWHIP_ERR_ARG, "a batch needs at least one system");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
      (void)hipGetLastError();
      throw status_error(GWHIP_ERR_DEVICE, "no HIP device " + std::to_string(device));
    }
    GWHIP_CHECK(hipSetDevice(device_));
    GWHIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    GWHIP_CHECK(hipMalloc(&flag_, sizeof(int)));
    for (int i = 0; i < n; ++i) systems_.push_back(std::make_unique<gw_hip_kernel>(dims[i], device_, stream_));
  }

  ~gw_hip_batch() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    systems_.clear();
    free_ws();
    if (flag_) (void)hipFree(flag_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  int size() const { return (int)systems_.size(); }
  gw_hip_kernel& system(int i) {
    if (i < 0 || i >= size()) throw status_error(GWHIP_ERR_ARG, "system index out of range");
    return *systems_[i];
  }

  // Autotune every system on its own operands (eagerly, before any capture).
  void autotune(const double2* const* G, int precision) {
    if (!G) throw status_error(GWHIP_ERR_ARG, "null G list");
    GWHIP_CHECK(hipSetDevice(device_));
    for (int i = 0; i < size(); ++i)
      if (!systems_[i]->has_inputs() || !G[i])
        throw status_error(GWHIP_ERR_STATE, "system " + std::to_string(i) + ": grid, integrals or G missing");
    if (!precision_valid(precision)) throw status_error(GWHIP_ERR_ARG, "unknown precision");
    prepare(any_split(precision));
    for (auto& s : systems_) s->ensure_orb_buffers();
    for (int i = 0; i < size(); ++i) systems_[i]->autotune(G[i], precision);
    drop_graphs();
  }

  void eval(const double2* const* G, double2* const* S, int precision) {
    if (!G || !S) throw status_error(GWHIP_ERR_ARG, "null G or Sigma list");
    for (int i = 0; i < size(); ++i)
      if (!G[i] || !S[i]) throw status_error(GWHIP_ERR_ARG, "null G or Sigma of system " + std::to_string(i));
    if (!precision_valid(precision)) throw status_error(GWHIP_ERR_ARG, "unknown precision");
    GWHIP_CHECK(hipSetDevice(device_));
    for (int i = 0; i < size(); ++i)
      if (!systems_[i]->has_inputs())
        throw status_error(GWHIP_ERR_STATE, "system " + std::to_string(i) + ": grid or integrals not set");
    prepare(any_split(precision));
    for (auto& s : systems_) {
      s->check_ready(precision);
      s->ensure_orb_buffers();  // before any capture: the GEMM form of the orbital products allocates
    }
    std::vector<const void*> key;
    for (int i = 0; i < size(); ++i) {
      key.push_back(G[i]);
      key.push_back(S[i]);
      // a graph holds the P0 tile set of each system's G (upper tiles only when G is Hermitian)
      key.push_back(systems_[i]->detect_hermitian_G(G[i]) ? (const void*)1 : nullptr);
    }
    const uint64_t gen = gen_sum();
    GraphEntry* e = nullptr;
    for (auto& g : graphs_)
      if (g.precision == precision && g.ptrs == key) e = &g;
    if (e && e->gen != gen) {
      (void)hipGraphExecDestroy(e->exec);
      graphs_.erase(graphs_.begin() + (e - graphs_.data()));
      e = nullptr;
    }
    if (!e) {
      graphs_.push_back(capture(G, S, precision));
      graphs_.back().ptrs = key;
      graphs_.back().gen = gen;
      e = &graphs_.back();
    }
    Event a, b;
    GWHIP_CHECK(hipEventRecord(a, stream_));
    GWHIP_CHECK(hipGraphLaunch(e->exec, stream_));
    GWHIP_CHECK(hipEventRecord(b, stream_));
    GWHIP_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    GWHIP_CHECK(hipEventElapsedTime(&ms, a, b));
    last_ms_ = ms;
    last_nodes_ = e->nodes;
    for (int i = 0; i < size(); ++i)
      if (systems_[i]->dyson_failed())
        throw status_error(GWHIP_ERR_NOT_POSITIVE_DEFINITE,
                           "system " + std::to_string(i) + ": I - P0(i Omega) is not positive definite for some (q, Omega)");
    int flag = 0;
    GWHIP_CHECK(hipMemcpy(&flag, flag_, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) throw status_error(GWHIP_ERR_NONFINITE, "Sigma of some system contains non-finite values");
  }

  double last_ms() const { return last_ms_; }
  int64_t last_nodes() const { return last_nodes_; }
  int64_t captures() const { return captures_; }
  size_t workspace_bytes() const { return ws_bytes_; }

private:
  struct GraphEntry {
    int precision = 0;
    std::vector<const void*> ptrs;
    uint64_t gen = 0;
    hipGraphExec_t exec = nullptr;
    int64_t nodes = 0;
  };

  int device_ = 0;
  hipStream_t stream_ = nullptr;
  int* flag_ = nullptr;
  std::vector<std::unique_ptr<gw_hip_kernel>> systems_;
  std::vector<GraphEntry> graphs_;
  Workspace ws_;
  size_t ws_bytes_ = 0;
  bool ws_ok_ = false, ws_split_ = false;
  std::vector<size_t> ws_slice_sig_;  // slice bytes per tau of every system when sized
  double last_ms_ = 0.0;
  int64_t last_nodes_ = 0, captures_ = 0;

  uint64_t gen_sum() const {
    uint64_t g = 0;
    for (auto& s : systems_) g = g * 1000003ull + s->generation();
    return g;
  }

  void drop_graphs() {
    for (auto& g : graphs_)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
  }

  void free_ws() {
    for (void* p : {(void*)ws_.scratch, (void*)ws_.dyson, (void*)ws_.slices, (void*)ws_.exps})
      if (p) (void)hipFree(p);
    ws_ = Workspace{};
    ws_bytes_ = 0;
    ws_ok_ = false;
  }

  template <typename T>
  T* walloc(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      throw status_error(GWHIP_ERR_OOM, "shared workspace allocation of " + std::to_string(count * sizeof(T)) +
                                            " bytes failed");
    }
    ws_bytes_ += count * sizeof(T);
    return static_cast<T*>(p);
  }

  // Split stores on every system, then a workspace that fits all of them (re)allocated when
  // the systems' needs changed, then bound to each system.
  void prepare(bool split) {
    if (split)
      for (auto& s : systems_) s->prepare_split();
    std::vector<size_t> sig;
    for (auto& s : systems_) sig.push_back(split ? s->slice_bytes_per_t() : 0);
    if (ws_ok_ && (!split || (ws_split_ && sig == ws_slice_sig_))) return;
    drop_graphs();
    free_ws();
    size_t sc_full = 0, sc_min = 0, dy_full = 0, dy_min = 0, sl_full = 0, sl_min = 0, ex_full = 0, ex_min = 0;
    for (auto& s : systems_) {
      const Dims& d = s->dims();
      sc_full = std::max(sc_full, s->scratch_per_t() * (d.nts / 2));  // both stages use half tau grids
      sc_min = std::max(sc_min, s->scratch_per_t());
      dy_full = std::max(dy_full, s->dyson_per_q() * d.nk);
      dy_min = std::max(dy_min, s->dyson_per_q());
      if (split) {
        sl_full = std::max(sl_full, s->slice_bytes_per_t() * (d.nts / 2));
        sl_min = std::max(sl_min, s->slice_bytes_per_t());
        ex_full = std::max(ex_full, s->exps_per_t() * (d.nts / 2));
        ex_min = std::max(ex_min, s->exps_per_t());
      }
    }
    size_t free_b = 0, total_b = 0;
    GWHIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const double budget = 0.8 * (double)free_b;
    const double full = 16.0 * (sc_full + dy_full) + (double)sl_full + 4.0 * ex_full;
    const double f = full <= budget ? 1.0 : budget / full;
    auto scaled = [&](size_t fullv, size_t minv) { return std::max(minv, (size_t)(f * (double)fullv)); };
    ws_.scratch_elems = scaled(sc_full, sc_min);
    ws_.scratch = walloc<double2>(ws_.scratch_elems);
    ws_.dyson_elems = scaled(dy_full, dy_min);
    ws_.dyson = walloc<double2>(ws_.dyson_elems);
    if (split) {
      ws_.slice_bytes = scaled(sl_full, sl_min);
      ws_.slices = walloc<int8_t>(ws_.slice_bytes);
      ws_.exp_elems = scaled(ex_full, ex_min);
      ws_.exps = walloc<int>(ws_.exp_elems);
    }
    for (auto& s : systems_) s->bind_workspace(ws_);
    ws_ok_ = true;
    ws_split_ = split;
    ws_slice_sig_ = sig;
  }

  GraphEntry capture(const double2* const* G, double2* const* S, int precision) {
    for (auto& s : systems_) s->set_timing(false);
    hipGraph_t graph = nullptr;
    GWHIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
    try {
      hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, stream_, flag_, (int64_t)1);
      GWHIP_CHECK(hipGetLastError());
      for (int i = 0; i < size(); ++i) systems_[i]->enqueue_eval(G[i], S[i], precision, flag_);
    } catch (...) {
      (void)hipStreamEndCapture(stream_, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      for (auto& s : systems_) s->set_timing(true);
      throw;
    }
    GWHIP_CHECK(hipStreamEndCapture(stream_, &graph));
    for (auto& s : systems_) s->set_timing(true);
    GraphEntry e;
    e.precision = precision;
    size_t nodes = 0;
    GWHIP_CHECK(hipGraphGetNodes(graph, nullptr, &nodes));
    e.nodes = (int64_t)nodes;
    hipError_t err = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GWHIP_CHECK(err);
    ++captures_;
    return e;
  }
};

}  // namespace gwhip
