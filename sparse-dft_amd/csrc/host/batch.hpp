// Batched k-point application: the H psi / S psi of many plane-wave operators
// that share one FFT grid (the k-points of one crystal, src/scf_core.cpp's
// k-point loop) in one persistent kernel launch, captured in a HIP graph and
// replayed, with a launch shape chosen by a measuring autotuner and cached per
// problem shape.
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../hip/batch.hpp"
#include "device.hpp"

namespace sdft_hip {

class OperatorBatch {
 public:
  // ops: operators of one context, same grid shape, atom-block projectors.
  // columns: columns per operator. The operators must outlive the batch.
  OperatorBatch(Context& ctx, const std::vector<DeviceOperator*>& ops, long columns, Precision prec);
  ~OperatorBatch();
  OperatorBatch(const OperatorBatch&) = delete;
  OperatorBatch& operator=(const OperatorBatch&) = delete;

  // offsets[k] (complex elements) of operator k's n_k x columns block in the
  // concatenated input/output; offsets[ops] = total
  const std::vector<long>& offsets() const { return offsets_; }
  long columns() const { return columns_; }
  // device buffers of the batch input and outputs (concatenated blocks)
  double* input() { return x_.p; }
  const double* output_h() const { return hy_.p; }
  const double* output_s() const { return sy_.p; }
  // one graph replay: H psi and S psi of every block (asynchronous on the
  // context stream)
  void replay();
  // the same launch issued directly, without the graph
  void launch_direct();
  const BatchLaunch& plan() const { return launch_; }
  double tuned_ms() const { return tuned_ms_; }
  bool plan_from_cache() const { return cached_; }
  // measured average replay time over `replays` replays (HIP events)
  double benchmark(int replays);

 private:
  void autotune();
  void capture();
  Context& ctx_;
  long columns_;
  Precision prec_;
  std::vector<long> offsets_;
  BatchLaunch launch_{};
  DevBuf<BatchOpDev> ops_dev_;
  std::vector<BatchOpDev> ops_host_;
  DevBuf<double> x_, hy_, sy_;
  DevBuf<int> inverse_, proj_atom_;
  DevBuf<char> bt_;
  hipGraph_t graph_{};
  hipGraphExec_t exec_{};
  double tuned_ms_ = 0;
  bool cached_ = false;
  std::string key_;
};

}  // namespace sdft_hip
