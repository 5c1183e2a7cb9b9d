// guide_host.h — the guide object (network, optimiser state, training store) and
// the launch helpers the tracer uses.
#pragma once
#include "guide.h"
#include "runtime.h"

struct mupsi_guide {
  mupsi_ctx* ctx = nullptr;
  mupsi_guide_config cfg{};
  mupsi::DeviceBuffers mem;
  float* theta = nullptr;  // FP32 master parameters
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  mupsi::guide::bf16x8* wf = nullptr;  // forward fragments
  mupsi::guide::bf16x8* wt = nullptr;  // transposed fragments
  float* grad_part = nullptr;          // per-block partial gradients
  float* loss_part = nullptr;
  float* grad_out = nullptr;           // reduced gradient (mupsi_guide_gradient)
  int part_blocks = 0;
  int64_t step = 0;
  mupsi_guide_stats stats{};

  mupsi::guide::NetDev net() const { return {theta, wf, wt}; }
  mupsi::guide::EncodeBox box() const {
    mupsi::guide::EncodeBox b;
    for (int a = 0; a < 3; ++a) b.c[a] = cfg.box_center[a];
    b.inv_r = 1.0f / cfg.box_radius;
    return b;
  }
};

// Frees a guide (the context's list is updated by the caller).
void destroy_guide(mupsi_guide* g);
// Detaches g from scene s if it is attached there (defined with the tracer).
void detach_guide(mupsi_scene* s, const mupsi_guide* g);

namespace mupsi {
namespace guide {

// Forward pass over n feature rows (n read from count_dev when it is non-null).
// raw_out: row-major n x 41, or null.  mix_out: mixture parameters (44 floats)
// written at row q_pid[i] (or i when q_pid is null), or null.
void launch_infer(const mupsi_guide* g, const float* feats, const int* count_dev, int count_host,
                  const int* q_pid, float* raw_out, float* mix_out, hipStream_t st);

// `steps` Adam steps on minibatches of `batch` samples drawn uniformly (with
// replacement) from n records (feats n x 32, dirs n x 3, w n).
mupsi_status train_random(mupsi_guide* g, const float* feats, const float* dirs, const float* w,
                          int n, int steps, int batch, hipStream_t st);

// Blocks of a training step on `batch` samples: one per 64 samples, at most the
// number of per-block partials the guide keeps (each block writes one loss partial).
int train_blocks(int batch);

}  // namespace guide
}  // namespace mupsi
