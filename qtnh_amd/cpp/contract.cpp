// Tensor::contract (reference src/tensor/contract.cpp:204-313).
//
// Host side: the same validation, the same default and symmetric index
// placements (computed by qtnh_default_repls / qtnh_symm_repls), the same
// side effects on ConParams.  Device side, by route:
//  * symmetric placement with every input wired and a dense state: in place
//    on the state -- IdenTensor: nothing to do; SwapTensor: axis swap;
//    DiagTensor: diagonal kernel; SymmTensor: gate kernel;
//  * anything else: qtnh_contract_c128 into a new tensor.
// Distributed indices are metadata on one GPU: the result's distributed
// count is the number of open distributed indices of both operands.
#include <algorithm>
#include <stdexcept>

#include "placement.hpp"
#include "status.hpp"
#include "tensor/dense.hpp"
#include "tensor/diag.hpp"
#include "tensor/symm.hpp"

namespace qtnh {
  namespace {
    std::vector<uint16_t> flat_wires(const std::vector<qtnh::wire>& ws) {
      std::vector<uint16_t> f;
      for (auto& w : ws) {
        f.push_back(w.first);
        f.push_back(w.second);
      }
      if (f.empty()) f.push_back(0);
      return f;
    }

    std::vector<uint16_t> u16(const std::vector<qtnh::tidx_tup_st>& v) {
      std::vector<uint16_t> r(v.begin(), v.end());
      if (r.empty()) r.push_back(0);
      return r;
    }

    // A dense tensor holding tp's elements: tp itself when dense, else a new one.
    std::unique_ptr<DenseTensor> as_dense(qtnh::tptr tp) { return Tensor::convert<DenseTensor>(std::move(tp)); }

    // Gate matrix G[in, out] of a symmetric tensor whose inputs are the wired
    // indices `ins` (t2 order) and outputs the remaining indices (t2 order).
    std::vector<qtnh::tel> gate_matrix(const Tensor& g, const std::vector<std::size_t>& ins,
                                       const std::vector<std::size_t>& outs) {
      const auto dims = g.totDims();
      qtnh::tidx_tup di, dout;
      for (auto i : ins) di.push_back(dims[i]);
      for (auto o : outs) dout.push_back(dims[o]);
      const std::size_t K = utils::dims_to_size(di), L = utils::dims_to_size(dout);
      std::vector<qtnh::tel> G(K * L);
      qtnh::tidx_tup t(dims.size());
      for (std::size_t a = 0; a < K; ++a) {
        auto va = utils::i_to_idxs(a, di);
        for (std::size_t m = 0; m < ins.size(); ++m) t[ins[m]] = va[m];
        for (std::size_t b = 0; b < L; ++b) {
          auto vb = utils::i_to_idxs(b, dout);
          for (std::size_t m = 0; m < outs.size(); ++m) t[outs[m]] = vb[m];
          G[a * L + b] = g.at(t);
        }
      }
      return G;
    }
  }

  namespace detail {
    Placement place(const qtnh::tidx_tup& d1, int nd1, const qtnh::tidx_tup& d2, int nd2, int symm_n_dis_in,
                    ConParams& params) {
      const int r1 = (int)d1.size(), r2 = (int)d2.size();
      int n_dis_ws = 0;
      for (auto& w : params.wires) {
        if (d1.at(w.first) != d2.at(w.second)) throw std::invalid_argument("Incompatible contraction dimensions.");
        if (((int)w.first < nd1) != ((int)w.second < nd2))
          throw std::invalid_argument("Attempted to contract distributed and local dimensions.");
        if ((int)w.first < nd1) ++n_dis_ws;
      }
      const int nw = (int)params.wires.size();
      auto byfirst = [](const qtnh::wire& a, const qtnh::wire& b) { return a.first < b.first; };
      auto bysecond = [](const qtnh::wire& a, const qtnh::wire& b) { return a.second < b.second; };

      if (params.useDefRepls) {
        std::vector<uint16_t> o1(std::max(r1, 1)), o2(std::max(r2, 1));
        auto fw = flat_wires(params.wires);
        int st = qtnh_default_repls(r1, nd1, r2, nd2, nw, fw.data(), o1.data(), o2.data());
        if (st != QTNH_OK) throw std::invalid_argument("Invalid contraction wires.");
        params.dimRepls1.assign(o1.begin(), o1.begin() + r1);
        params.dimRepls2.assign(o2.begin(), o2.begin() + r2);
        std::sort(params.wires.begin(), params.wires.end(), byfirst);
        std::sort(params.wires.begin(), params.wires.end(), bysecond);
      }

      Placement P;
      if (symm_n_dis_in >= 0) {
        std::vector<uint16_t> o1(std::max(r1, 1)), o2(std::max(r2, 1));
        auto fw = flat_wires(params.wires);
        int applies = 0;
        int st = qtnh_symm_repls(r1, nd2, r2 - nd2, symm_n_dis_in, nw, fw.data(), o1.data(), o2.data(), &applies);
        if (st != QTNH_OK) throw std::invalid_argument("Invalid contraction wires.");
        if (applies) {
          params.dimRepls1.assign(o1.begin(), o1.begin() + r1);
          params.dimRepls2.assign(o2.begin(), o2.begin() + r2);
          std::sort(params.wires.begin(), params.wires.end(), byfirst);
          std::sort(params.wires.begin(), params.wires.end(), bysecond);
          P.symm = true;
        }
      }
      if ((int)params.dimRepls1.size() != r1 || (int)params.dimRepls2.size() != r2)
        throw std::out_of_range("Index replacements do not match the tensors.");

      // Result dims under the placement; a placement that is not a permutation
      // of the result indices is rejected (the reference indexes past its vectors).
      P.wires = flat_wires(params.wires);
      P.repl1 = u16(params.dimRepls1);
      P.repl2 = u16(params.dimRepls2);
      int rc = 0;
      std::vector<uint64_t> dC(QTNH_MAX_RANK);
      if (qtnh_contract_dims(r1, d1.data(), r2, d2.data(), nw, P.wires.data(), P.repl1.data(), P.repl2.data(), &rc,
                             dC.data()) != QTNH_OK)
        throw std::out_of_range("Invalid index replacements.");
      P.dims.assign(dC.begin(), dC.begin() + rc);
      P.ndis = (nd1 - n_dis_ws) + (nd2 - n_dis_ws);
      P.in_place = P.symm && nw * 2 == r2 && P.dims == d1;
      return P;
    }
  }

  qtnh::tptr Tensor::contract(qtnh::tptr tp1, qtnh::tptr tp2, ConParams& params) {
    const QTNHEnv& env = tp1->bc().env;
    const auto d1 = tp1->totDims(), d2 = tp2->totDims();
    const int r1 = (int)d1.size(), r2 = (int)d2.size();
    SymmTensorBase* s2 = dynamic_cast<SymmTensorBase*>(tp2.get());
    const detail::Placement P = detail::place(d1, (int)tp1->disDims().size(), d2, (int)tp2->disDims().size(),
                                              s2 ? (int)s2->nDisIn() : -1, params);
    const int nw = (int)params.wires.size();
    const qtnh::tidx_tup& dCt = P.dims;
    auto split = utils::split_dims(dCt, (qtnh::tidx_tup_st)P.ndis);

    // In-place route: symmetric placement, exactly the inputs wired, result dims = state dims.
    if (P.in_place) {
      std::vector<std::size_t> ins, outs, axes;
      std::vector<char> wired(r2, 0);
      for (auto& w : params.wires) {  // sorted by second
        ins.push_back(w.second);
        axes.push_back(w.first);
        wired[w.second] = 1;
      }
      for (int i = 0; i < r2; ++i)
        if (!wired[i]) outs.push_back(i);
      const TT t2 = tp2->type();
      std::unique_ptr<DenseTensor> st = as_dense(std::move(tp1));
      if (t2 != TT::idenTensor && st->totSize() > 0) {
        qtnh_ctx_t c = env.ctx();
        std::vector<uint16_t> ax(axes.begin(), axes.end());
        void* p = st->store().device_mut();
        int status = QTNH_OK;
        if (t2 == TT::swapTensor) {
          status = qtnh_swap_c128(c, p, r1, d1.data(), ax.at(0), ax.at(1));
        } else if (t2 == TT::diagTensor) {
          const auto& dg = static_cast<DiagTensor*>(tp2.get())->diagonal();
          status = qtnh_diag_apply_c128(c, p, r1, d1.data(), nw, ax.data(), reinterpret_cast<const double*>(dg.data()));
        } else {
          status = QTNH_EUNSUPPORTED;
        }
        if (status == QTNH_EUNSUPPORTED) {
          auto G = gate_matrix(*tp2, ins, outs);
          status = qtnh_gate_apply_c128(c, p, r1, d1.data(), nw, ax.data(), reinterpret_cast<const double*>(G.data()), 0);
        }
        detail::check(c, status, "gate application");
      }
      // Same data, result split of the reference.
      st->dis_dims_ = split.first;
      st->loc_dims_ = split.second;
      Broadcaster nb(env, (qtnh::uint)st->disSize(), {1, 1, st->bc().off});
      st->bc_ = std::move(nb);
      return qtnh::tptr(st.release());
    }

    // General route.
    const qtnh::uint off = std::min(tp1->bc().off, tp2->bc().off);
    std::unique_ptr<DenseTensor> a = as_dense(std::move(tp1));
    std::unique_ptr<DenseTensor> b = as_dense(std::move(tp2));
    const std::size_t nC = utils::dims_to_size(dCt);
    std::unique_ptr<detail::Store> out(new detail::Store(env, nC));
    if (nC > 0) {
      qtnh_ctx_t c = env.ctx();
      const void* pa = a->store().device();
      const void* pb = b->store().device();
      void* pc = out->device_out();
      double err = 0.0;
      detail::check(c,
                    qtnh_contract_c128(c, pa, r1, d1.data(), pb, r2, d2.data(), nw, P.wires.data(), P.repl1.data(),
                                       P.repl2.data(), pc, env.precision, &err),
                    "contraction");
      env.last_split_error = err;
    }
    return qtnh::tptr(new DenseTensor(env, split.first, split.second, std::move(out), {1, 1, off}));
  }
}
