// nn helpers (see include/loop_tool/nn.h); same graphs as
// src/frontends/nn.cpp:15-76 of the reference.
#include "loop_tool/nn.h"

#include <unordered_set>

namespace loop_tool {
namespace nn {

using lazy::Constraint;
using lazy::Expr;
using lazy::Symbol;
using lazy::Tensor;

// X[..., spatial...] viewed as X[batch..., out_spatial..., window...] with
// spatial = out*stride + window; multiplied by W and summed over the window
// and every dim X shares with W.
Tensor convolve(Tensor X, Tensor W, std::vector<Symbol> spatial_dims,
                std::vector<Symbol> window_dims, int stride) {
  ASSERT(spatial_dims.size() == window_dims.size())
      << "one window dim per spatial dim";
  std::vector<Symbol> out_spatial;
  std::vector<Constraint> constraints;
  for (size_t i = 0; i < spatial_dims.size(); ++i) {
    Symbol o(spatial_dims[i].name() + "_x_" + window_dims[i].name());
    out_spatial.emplace_back(o);
    constraints.emplace_back(Expr(spatial_dims[i]),
                             Expr(o) * Expr(stride) + Expr(window_dims[i]));
  }
  std::unordered_set<int> w_ids, sp_ids;
  for (const auto &s : W.shape()) w_ids.insert(s.id());
  for (const auto &s : spatial_dims) sp_ids.insert(s.id());
  std::vector<Symbol> batch;
  std::vector<Symbol> reduced = window_dims;
  for (const auto &s : X.shape()) {
    if (w_ids.count(s.id()))
      reduced.emplace_back(s);
    else if (!sp_ids.count(s.id()))
      batch.emplace_back(s);
  }
  batch.insert(batch.end(), out_spatial.begin(), out_spatial.end());
  batch.insert(batch.end(), window_dims.begin(), window_dims.end());
  X = X.to(batch, constraints);
  return (X * W).sum(reduced);
}

// X viewed with a k-wide window per spatial dim (spatial = out*stride + win),
// max-reduced over the windows.
Tensor maxpool(Tensor X, std::vector<Symbol> spatial_dims, int k, int stride) {
  std::vector<Symbol> out_spatial, windows;
  std::vector<Constraint> constraints;
  for (const auto &s : spatial_dims) {
    Symbol o(s.name() + "_p");
    Symbol w(s.name() + "_k");
    out_spatial.emplace_back(o);
    windows.emplace_back(w);
    constraints.emplace_back(Expr(s), Expr(o) * Expr(stride) + Expr(w));
    constraints.emplace_back(Expr::size(Expr(w)), Expr(k));
  }
  std::unordered_set<int> sp_ids;
  for (const auto &s : spatial_dims) sp_ids.insert(s.id());
  std::vector<Symbol> batch;
  for (const auto &s : X.shape())
    if (!sp_ids.count(s.id())) batch.emplace_back(s);
  batch.insert(batch.end(), out_spatial.begin(), out_spatial.end());
  batch.insert(batch.end(), windows.begin(), windows.end());
  X = X.to(batch, constraints);
  return X.max(windows);
}

}  // namespace nn
}  // namespace loop_tool
