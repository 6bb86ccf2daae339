// TensorNetwork on the device executor (interface: reference
// include/tensor/network.hpp; semantics: src/tensor/network.cpp:60-170).
//
// A contraction order is decided on the host first, from metadata only: each
// step's wires and result placement come from the placement logic of
// Tensor::contract (detail::place), and after every step the remaining bonds
// are re-keyed through that placement, as the reference's contractBond does.
// The steps then run as one planned device network per result
// (qtnh_net_plan_steps + qtnh_net_run_device): leaves are read in place from
// their device storage, intermediates take fixed arena slots of the context's
// workspace, and each step is kernel launches on the context's stream, with no
// host round trip, rebcast or allocation between steps.  An identity tensor
// contracted in place only relabels its partner's indices (as in
// Tensor::contract), so it adds no step.
#include <algorithm>
#include <iostream>
#include <map>
#include <stdexcept>

#include "placement.hpp"
#include "status.hpp"
#include "tensor/network.hpp"

namespace qtnh {
  namespace {
    // One tensor of the order being planned: a leaf (a tensor of the network)
    // or the result of a step over two earlier ones.
    struct PNode {
      qtnh::uint leaf_tid = 0;  // leaves: the network's tensor id
      int a = -1, b = -1;       // steps: operand nodes
      int alias = -1;           // in-place identity step: the node whose data this is
      qtnh::tidx_tup dims;
      int ndis = 0;
      qtnh::uint off = 0;
      std::vector<uint16_t> wires, repl1, repl2;
      int nw = 0;
    };
  }

  TensorNetwork::TensorNetwork() = default;

  TensorNetwork::Bond::Bond(std::pair<qtnh::uint, qtnh::uint> tids, std::vector<qtnh::wire> ws)
    : tensor_ids(tids), wires(std::move(ws)) {}

  Tensor* TensorNetwork::tensor(qtnh::uint tid) { return tensors_.at(tid).get(); }
  const TensorNetwork::Bond& TensorNetwork::bond(qtnh::uint bid) { return bonds_.at(bid); }

  std::vector<qtnh::uint> TensorNetwork::tensorIDs() {
    std::vector<qtnh::uint> ids;
    for (auto& kv : tensors_) ids.push_back(kv.first);
    return ids;
  }

  std::vector<qtnh::uint> TensorNetwork::bondIDs() {
    std::vector<qtnh::uint> ids;
    for (auto& kv : bonds_) ids.push_back(kv.first);
    return ids;
  }

  qtnh::tptr TensorNetwork::extract(qtnh::uint tid) {
    qtnh::tptr tp = std::move(tensors_.at(tid));
    tensors_.erase(tid);
    return tp;
  }

  qtnh::uint TensorNetwork::insert(qtnh::tptr tp) {
    tensors_.insert({++tensor_counter, std::move(tp)});
    return tensor_counter;
  }

  qtnh::uint TensorNetwork::addBond(qtnh::uint tid1, qtnh::uint tid2, std::vector<qtnh::wire> ws) {
    bonds_.insert({++bond_counter, Bond({tid1, tid2}, ws)});
    return bond_counter;
  }

  qtnh::uint TensorNetwork::contractBond(qtnh::uint bid) { return run(std::vector<qtnh::uint>{bid}, false, false); }

  qtnh::uint TensorNetwork::contractAll() {
    // Every bond in the map's iteration order, one contraction each.
    std::vector<qtnh::uint> ids;
    for (auto& kv : bonds_) ids.push_back(kv.first);
    return run(ids, false, false);
  }

  qtnh::uint TensorNetwork::contractAll(std::vector<qtnh::uint> bids) { return run(bids, true, true); }

  qtnh::uint TensorNetwork::run(const std::vector<qtnh::uint>& bids, bool merge, bool rebcast) {
    qtnh::uint tid = tensors_.empty() ? 0 : tensors_.begin()->first;
    last_plan_steps_ = 0;
    if (bids.empty()) return tid;

    // ---- plan: the order's steps on metadata; bonds re-keyed as they go.
    std::vector<PNode> nodes;
    std::map<qtnh::uint, int> live;  // tensor id -> node
    auto node_of = [&](qtnh::uint t) {
      auto it = live.find(t);
      if (it != live.end()) return it->second;
      const Tensor& x = *tensors_.at(t);
      PNode n;
      n.leaf_tid = t;
      n.dims = x.totDims();
      n.ndis = (int)x.disDims().size();
      n.off = x.bc().off;
      nodes.push_back(n);
      live[t] = (int)nodes.size() - 1;
      return (int)nodes.size() - 1;
    };
    for (std::size_t i = 0, j = 1; i < bids.size(); i += j) {
      j = 1;
      Bond& b1 = bonds_.at(bids[i]);
      if (merge) {
        // The following bonds that join the same two tensors are one contraction.
        for (; i + j < bids.size(); ++j) {
          const Bond& b2 = bonds_.at(bids[i + j]);
          if (b1.tensor_ids != b2.tensor_ids) break;
          b1.wires.insert(b1.wires.end(), b2.wires.begin(), b2.wires.end());
          bonds_.erase(bids[i + j]);
        }
      }
      const qtnh::uint tid1 = b1.tensor_ids.first, tid2 = b1.tensor_ids.second;
      const int n1 = node_of(tid1), n2 = node_of(tid2);
      if (n1 == n2) throw std::invalid_argument("Bond joins a tensor to itself.");
      // Only a leaf can be symmetric: every contraction result is dense.
      int n_dis_in = -1;
      bool iden = false;
      if (nodes[n2].leaf_tid) {
        const Tensor* t2 = tensors_.at(nodes[n2].leaf_tid).get();
        if (auto* s2 = dynamic_cast<const SymmTensorBase*>(t2)) n_dis_in = (int)s2->nDisIn();
        iden = t2->type() == TT::idenTensor;
      }
      ConParams params(b1.wires);
      const detail::Placement P = detail::place(nodes[n1].dims, nodes[n1].ndis, nodes[n2].dims, nodes[n2].ndis,
                                                n_dis_in, params);
      PNode r;
      r.dims = P.dims;
      r.ndis = P.ndis;
      if (P.in_place && iden) {
        r.alias = n1;  // the state's data, relabelled
        r.off = nodes[n1].off;
      } else {
        r.a = n1;
        r.b = n2;
        r.nw = (int)params.wires.size();
        r.wires = P.wires;
        r.repl1 = P.repl1;
        r.repl2 = P.repl2;
        r.off = P.in_place ? nodes[n1].off : std::min(nodes[n1].off, nodes[n2].off);
      }
      nodes.push_back(r);
      const int n3 = (int)nodes.size() - 1;
      const qtnh::uint tid3 = ++tensor_counter;
      bonds_.erase(bids[i]);
      live.erase(tid1);
      live.erase(tid2);
      live[tid3] = n3;
      for (auto& kv : bonds_) {
        Bond& o = kv.second;
        for (int end = 0; end < 2; ++end) {
          qtnh::uint& t = end ? o.tensor_ids.second : o.tensor_ids.first;
          if (t != tid1 && t != tid2) continue;
          const auto& repl = t == tid1 ? params.dimRepls1 : params.dimRepls2;
          t = tid3;
          for (auto& w : o.wires) {
            auto& ax = end ? w.second : w.first;
            ax = repl.at(ax);
          }
        }
      }
      tid = tid3;
    }

    // ---- run: one device network per surviving result.
    const QTNHEnv& env = tensors_.at(nodes[0].leaf_tid)->bc().env;
    qtnh_ctx_t ctx = env.ctx();
    std::map<qtnh::uint, qtnh::tptr> results;
    std::vector<qtnh::tptr> keep;  // consumed leaves, alive until the stream has read them
    std::map<qtnh::uint, const void*> leaf_ptr;
    auto device_of = [&](qtnh::uint t) -> const void* {
      auto it = leaf_ptr.find(t);
      if (it != leaf_ptr.end()) return it->second;
      std::unique_ptr<DenseTensor> d = Tensor::convert<DenseTensor>(std::move(tensors_.at(t)));
      const void* p = d->totSize() ? d->store().device() : nullptr;
      leaf_ptr[t] = p;
      keep.push_back(qtnh::tptr(d.release()));
      return p;
    };
    auto resolve = [&](int x) {
      while (nodes[x].alias >= 0) x = nodes[x].alias;
      return x;
    };
    for (auto& kv : live) {
      const int root = kv.second;
      if (nodes[root].leaf_tid) continue;  // untouched by the order
      // The root's tree: leaves in first use, steps in plan order.
      std::vector<int> order_leaves, order_steps;
      std::vector<char> in_tree(nodes.size(), 0);
      std::vector<int> stack{resolve(root)};
      while (!stack.empty()) {
        const int x = stack.back();
        stack.pop_back();
        in_tree[x] = 1;
        if (nodes[x].a >= 0) {
          stack.push_back(resolve(nodes[x].a));
          stack.push_back(resolve(nodes[x].b));
        }
      }
      std::map<int, int> net_id;
      for (std::size_t x = 0; x < nodes.size(); ++x) {
        if (!in_tree[x]) continue;
        if (nodes[x].leaf_tid) order_leaves.push_back((int)x);
        else order_steps.push_back((int)x);
      }
      std::sort(order_leaves.begin(), order_leaves.end(), [&](int p, int q) {
        return nodes[p].leaf_tid < nodes[q].leaf_tid;
      });
      qtnh_net_t net = nullptr;
      detail::check(ctx, qtnh_net_create(&net), "network");
      try {
        bool has_empty = false;
        for (int x : order_leaves) has_empty = has_empty || utils::dims_to_size(nodes[x].dims) == 0;
        for (int x : order_leaves) {
          const void* p = device_of(nodes[x].leaf_tid);
          if (has_empty) continue;
          const auto& d = nodes[x].dims;
          std::vector<uint64_t> dims(d.begin(), d.end());
          int id = 0;
          detail::check(ctx, qtnh_net_add_tensor_device(net, (int)dims.size(), dims.data(), nullptr, p, &id),
                        "network leaf");
          net_id[x] = id;
        }
        std::vector<qtnh_net_step_spec> specs;
        for (int x : order_steps) {
          if (has_empty) break;
          const PNode& n = nodes[x];
          qtnh_net_step_spec sp;
          sp.a = net_id.at(resolve(n.a));
          sp.b = net_id.at(resolve(n.b));
          sp.nw = n.nw;
          sp.wires = n.wires.data();
          sp.repl1 = n.repl1.data();
          sp.repl2 = n.repl2.data();
          specs.push_back(sp);
          net_id[x] = (int)order_leaves.size() + (int)specs.size() - 1;
        }
        if (!has_empty) {
          detail::check(ctx, qtnh_net_plan_steps(net, (int)specs.size(), specs.data()), "network plan");
          detail::check(ctx, qtnh_net_set_precision(net, env.precision), "network precision");
        }
        const PNode& rn = nodes[root];
        const std::size_t nC = utils::dims_to_size(rn.dims);
        std::unique_ptr<detail::Store> out;
        if (nC > 0 && !has_empty) {
          out.reset(new detail::Store(env, nC));
          detail::check(ctx, qtnh_net_run_device(ctx, net, out->device_out()), "network run");
        } else {
          // An empty leaf: every sum over its (wired) zero-size index is empty.
          out.reset(new detail::Store(env, std::vector<qtnh::tel>(nC, 0.0)));
        }
        last_plan_steps_ += specs.size();
        auto split = utils::split_dims(rn.dims, (qtnh::tidx_tup_st)rn.ndis);
        const qtnh::uint off = rebcast ? 0 : rn.off;
        results[kv.first] =
            qtnh::tptr(new DenseTensor(env, split.first, split.second, std::move(out), {1, 1, off}));
      } catch (...) {
        qtnh_net_destroy(net);
        throw;
      }
      qtnh_net_destroy(net);
    }
    // The consumed leaves' device memory is released only after the stream has read it.
    if (!keep.empty()) detail::check(ctx, qtnh_sync(ctx), "network sync");
    keep.clear();
    for (auto& n : nodes)
      if (n.leaf_tid && !live.count(n.leaf_tid)) tensors_.erase(n.leaf_tid);
    for (auto& kv : results) tensors_[kv.first] = std::move(kv.second);
    return tid;
  }

  std::size_t TensorNetwork::lastPlanSteps() const noexcept { return last_plan_steps_; }

  // Format of reference src/tensor/network.cpp:172-190.
  void TensorNetwork::print() {
    using namespace qtnh::ops;
    std::cout << "================================================================\n";
    std::cout << "Tensor Network of " << tensors_.size() << " tensors and " << bonds_.size() << " bonds\n";
    std::cout << "----------------------------------------------------------------\n";
    std::cout << "Tensors: \n";
    for (auto& kv : tensors_) std::cout << "ID: " << kv.first << " | Els: " << *kv.second << "\n";
    std::cout << "----------------------------------------------------------------\n";
    std::cout << "Bonds: \n";
    for (auto& kv : bonds_) std::cout << "ID: " << kv.first << " | Bond: " << kv.second << "\n";
    std::cout << "================================================================\n";
  }

  // Format of reference src/tensor/network.cpp:192-203.
  std::ostream& ops::operator<<(std::ostream& out, const TensorNetwork::Bond& o) {
    out << "(" << o.tensor_ids.first << ", " << o.tensor_ids.second << "); {";
    for (std::size_t i = 0; i < o.wires.size(); ++i) {
      out << "(" << o.wires[i].first << ", " << o.wires[i].second << ")";
      if (i + 1 < o.wires.size()) out << ", ";
    }
    return out << "}";
  }
}
