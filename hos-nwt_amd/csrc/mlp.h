// Internal interface of the order-prediction MLP (mlp.hip) used by the learned
// order selection in hosnwt.hip.  The public C-ABI is in include/hosnwt.h.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/hosnwt.h"

namespace hosmlp {

constexpr int kIn = 512;    // input features (8 blocks of 64)
constexpr int kWidth = 256; // hidden width (3 hidden layers)
constexpr int kOut = 64;    // predicted band energies of the next order
constexpr int kLayers = 4;
constexpr int kRowTile = 64;  // batches are padded to multiples of 64 rows

// Per-caller inference workspace (activations of up to max_rows rows), so that
// several contexts can run the shared network on their own streams.
struct Workspace {
  int max_rows = 0;
  __bf16* x = nullptr;      // [max_rows][kIn]
  __bf16* h[3] = {};        // [max_rows][kWidth]
  float* y = nullptr;       // [max_rows][kOut]
};

// Allocates / frees a workspace (hipMalloc; called at set-up, not per step).
void workspace_alloc(Workspace& w, int max_rows);
void workspace_free(Workspace& w);

// y = net(x) for rows (a multiple of kRowTile, <= w.max_rows) on stream st;
// w.x must hold the input rows.  Weights are read-only here.
void forward(const hos_mlp* net, int rows, Workspace& w, hipStream_t st);

}  // namespace hosmlp
