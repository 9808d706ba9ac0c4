// Captured launch sequences: a GraphSegment holds the instantiated graph of
// one sequence and the arguments it was captured with.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <vector>

#include "common.hpp"

namespace crispy {

inline void drop_segment(GraphSegment &g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
  g.key.clear();
}

// Runs enqueue(capturing) on the main stream through the segment's graph:
// captured again when the key differs from the one it was captured with, and
// then either applied to the instantiated graph in place (same launches, new
// arguments: hipGraphExecUpdate) or instantiated anew.  A captured graph must
// hold kernels and event dependencies only (memset nodes have given wrong
// results on replays), which is checked.
template <class Enqueue>
int run_segment(crispy_ctx *ctx, GraphSegment &g, std::vector<uintptr_t> key, const char *what,
                Enqueue &&enqueue) {
  if (!ctx->graphs) return enqueue(false);
  if (!g.exec || key != g.key) {
    g.key.clear();
    CRISPY_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(true);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != CRISPY_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc != CRISPY_OK ? rc : set_err(ctx, CRISPY_ERR_HIP, "capturing %s: %s", what, hipGetErrorString(e));
    }
    size_t nn = 0;
    hipError_t ge = hipGraphGetNodes(graph, nullptr, &nn);
    std::vector<hipGraphNode_t> nodes(nn);
    if (ge == hipSuccess && nn) ge = hipGraphGetNodes(graph, nodes.data(), &nn);
    for (size_t k = 0; ge == hipSuccess && k < nn; ++k) {
      hipGraphNodeType t;
      ge = hipGraphNodeGetType(nodes[k], &t);
      if (ge == hipSuccess && t == hipGraphNodeTypeMemset) {
        (void)hipGraphDestroy(graph);
        return set_err(ctx, CRISPY_ERR_HIP, "capturing %s: the graph holds a memset node", what);
      }
    }
    if (ge != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return set_err(ctx, CRISPY_ERR_HIP, "inspecting the %s graph: %s", what, hipGetErrorString(ge));
    }
    bool updated = false;
    if (g.exec) {
      hipGraphNode_t bad = nullptr;
      hipGraphExecUpdateResult ur = hipGraphExecUpdateError;
      updated = hipGraphExecUpdate(g.exec, graph, &bad, &ur) == hipSuccess && ur == hipGraphExecUpdateSuccess;
      if (!updated) {
        (void)hipGetLastError();
        drop_segment(g);
      }
    }
    const hipError_t ie = updated ? hipSuccess : hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      g.exec = nullptr;
      return set_err(ctx, CRISPY_ERR_HIP, "instantiating the %s graph: %s", what, hipGetErrorString(ie));
    }
    g.key = std::move(key);
  }
  CRISPY_HIP(ctx, hipGraphLaunch(g.exec, ctx->stream));
  return CRISPY_OK;
}

}  // namespace crispy
